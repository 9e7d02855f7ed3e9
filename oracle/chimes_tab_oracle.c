/*
 * ORACLE -- test infrastructure only.
 *
 * Plain-C, CPU, FP64 restatement of the reference's tabulated 2- and 3-body
 * interactions (chimesFF built with -DTABULATION, used by the LAMMPS pair
 * style): the table readers read_2B_tab / read_3B_tab (chimesFF.cpp:198-352),
 * the parameter-file entries that name the tables (715-745, 876-918), the
 * local cubic spline of get_tab_2B (1858-1924), the Catmull-Rom tricubic of
 * interpolateTricubic (2335-2450) and compute_2B_tab / compute_3B_tab
 * (1777-1856, 2178-2313).  Same operations in the same order, built with
 * -ffp-contract=off, so it can be compared with the compiled reference bit
 * for bit.
 *
 * Parity pins (tests/test_tab_oracle.py):
 *   - the compiled reference with -DTABULATION (oracle/_ref), call by call,
 *     on the reference's own 2-body table (etc/lmp/tests/example-tabulate/
 *     chimes_scan_2b.type_0.dat) and on a generated 3-body table;
 *   - what the mathematics fixes: the spline reproduces the table at its
 *     knots, and a table of a cubic polynomial is reproduced exactly by the
 *     Catmull-Rom tricubic inside the grid.
 *
 * Deliberate differences from the reference, all on inputs where it exits or
 * reads outside its arrays: a distance at or below the second table point
 * (2-body) and a grid index past the table's last point (3-body) set *err
 * instead of exiting or reading past the table; table files are resolved
 * against the parameter file's directory with a separator (the reference
 * concatenates the directory and the name without one, which only works for
 * a parameter file given without a directory).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "chimes_oracle.h"

#define TLEN 4096

typedef struct {
    int n;
    double *r, *e, *f;
} tab2_t;

typedef struct {
    long rows;
    int n;        /* grid points per axis: (int) cbrt(rows) */
    double dr;    /* 1 / grid spacing */
    double *rij, *rik, *rjk, *e, *fij, *fik, *fjk;
} tab3_t;

struct oc_tables {
    int ntab2, ntab3;
    tab2_t *t2;
    tab3_t *t3;
};

static char g_tab_err[512];
const char *oc_tab_error(void) { return g_tab_err; }

/* split_line (chimesFF.cpp:148-178): '!' and '##' start comments */
static int tab_split(char *line, char **tok, int maxtok)
{
    char *c = strchr(line, '!');
    if (c) *c = 0;
    c = strstr(line, "##");
    if (c) *c = 0;
    int n = 0;
    for (char *s = strtok(line, " \t\r\n"); s && n < maxtok; s = strtok(NULL, " \t\r\n")) tok[n++] = s;
    return n;
}

/* One table file: a count line, then that many rows of cmin or cmax items;
 * cols[0 .. cmax-1] are allocated (count values each) and receive the
 * items of every row (columns a row does not have stay zero). */
static int read_table(const char *path, int cmin, int cmax, double *cols[6], long *rows)
{
    for (int c = 0; c < 6; c++) cols[c] = NULL;
    FILE *fp = fopen(path, "r");
    if (!fp) {
        snprintf(g_tab_err, sizeof g_tab_err, "could not open table file %.400s", path);
        return -1;
    }
    char line[TLEN], *tok[16];
    const long n = fgets(line, sizeof line, fp) ? strtol(line, NULL, 10) : 0;
    if (n <= 0) {
        fclose(fp);
        snprintf(g_tab_err, sizeof g_tab_err, "%.400s: missing or bad row count", path);
        return -1;
    }
    for (int c = 0; c < cmax; c++) cols[c] = (double *)calloc(n, sizeof(double));
    for (long i = 0; i < n; i++) {
        const int k = fgets(line, sizeof line, fp) ? tab_split(line, tok, 16) : 0;
        if (k != cmin && k != cmax) {
            fclose(fp);
            snprintf(g_tab_err, sizeof g_tab_err, "%.400s: row %ld has %d items", path, i + 1, k);
            return -1;
        }
        for (int c = 0; c < k; c++) cols[c][i] = strtod(tok[c], NULL);
    }
    fclose(fp);
    *rows = n;
    return 0;
}

static void table_path(const char *param_path, const char *name, const char *suffix, char *out, size_t len)
{
    const char *slash = strrchr(param_path, '/');
    if (slash)
        snprintf(out, len, "%.*s/%s%s", (int)(slash - param_path), param_path, name, suffix);
    else
        snprintf(out, len, "%s%s", name, suffix);
}

void oc_tab_free(oc_tables *T)
{
    if (!T) return;
    for (int p = 0; p < T->ntab2; p++) {
        free(T->t2[p].r);
        free(T->t2[p].e);
        free(T->t2[p].f);
    }
    for (int t = 0; t < T->ntab3; t++) {
        tab3_t *b = &T->t3[t];
        free(b->rij); free(b->rik); free(b->rjk);
        free(b->e); free(b->fij); free(b->fik); free(b->fjk);
    }
    free(T->t2);
    free(T->t3);
    free(T);
}

/* read_2B_tab (chimesFF.cpp:200-257): energy file rows "r E", force file rows
 * "r F" (its distances are not kept). */
static int load_2b(tab2_t *t, const char *base)
{
    char path[TLEN + 16];
    double *c[6];
    long ne = 0, nf = 0;
    snprintf(path, sizeof path, "%s.energy", base);
    const int re = read_table(path, 2, 2, c, &ne);
    t->r = c[0];
    t->e = c[1];
    if (re) return -1;
    snprintf(path, sizeof path, "%s.force", base);
    const int rf = read_table(path, 2, 2, c, &nf);
    free(c[0]);
    t->f = c[1];
    if (rf) return -1;
    if (nf != ne) {
        snprintf(g_tab_err, sizeof g_tab_err, "%.400s: energy and force tables differ in length", base);
        return -1;
    }
    if (ne < 3) {
        snprintf(g_tab_err, sizeof g_tab_err, "%.400s: a 2-body table needs at least 3 rows", base);
        return -1;
    }
    t->n = (int)ne;
    return 0;
}

/* read_3B_tab (chimesFF.cpp:259-352): energy rows "rij rik rjk E", force rows
 * "rij rik rjk Fij Fik Fjk" (their distances are not kept); the grid size and
 * spacing of interpolateTricubic (2339-2342). */
static int load_3b(tab3_t *t, const char *base)
{
    char path[TLEN + 16];
    double *c[6];
    long ne = 0, nf = 0;
    snprintf(path, sizeof path, "%s.energy", base);
    const int re = read_table(path, 4, 6, c, &ne);
    t->rij = c[0]; t->rik = c[1]; t->rjk = c[2]; t->e = c[3];
    free(c[4]);
    free(c[5]);
    if (re) return -1;
    snprintf(path, sizeof path, "%s.force", base);
    const int rf = read_table(path, 6, 6, c, &nf);
    free(c[0]); free(c[1]); free(c[2]);
    t->fij = c[3]; t->fik = c[4]; t->fjk = c[5];
    if (rf) return -1;
    if (nf != ne) {
        snprintf(g_tab_err, sizeof g_tab_err, "%.400s: energy and force tables differ in length", base);
        return -1;
    }
    t->rows = ne;
    t->n = (int)cbrt((double)ne);
    const long n = t->n;
    if (n < 2 || n * n + n + 1 >= ne) {
        snprintf(g_tab_err, sizeof g_tab_err, "%.400s: a 3-body table needs at least 2 points per axis", base);
        return -1;
    }
    t->dr = 1 / (t->rjk[n * n + n + 1] - t->rjk[0]);
    return 0;
}

/* The parameter-file entries that name tables (chimesFF.cpp:715-745,
 * 876-918): a "PAIRTYPE PARAMS:" line of 7 items names the 2-body table of
 * the next pair (tables in reading order), an "INDEX:" line of 8 items in a
 * "TRIPLETTYPE PARAMS:" block the 3-body table of the next triplet type. */
oc_tables *oc_tab_load(const oc_params *P, const char *param_path)
{
    FILE *fp = fopen(param_path, "r");
    g_tab_err[0] = 0;
    if (!fp) {
        snprintf(g_tab_err, sizeof g_tab_err, "could not open %.400s", param_path);
        return NULL;
    }
    oc_tables *T = (oc_tables *)calloc(1, sizeof *T);
    char line[TLEN], copy[TLEN], *tok[64];
    int in_trip = 0, plain2 = 0, bad = 0;
    while (fgets(line, sizeof line, fp)) {
        if (strstr(line, "ENDFILE")) break;
        strcpy(copy, line);
        if (strstr(line, "PAIRTYPE PARAMS:")) {
            const int k = tab_split(copy, tok, 64);
            if (k == 7) {
                if (plain2) bad = 1;
                T->t2 = (tab2_t *)realloc(T->t2, sizeof(tab2_t) * (T->ntab2 + 1));
                memset(&T->t2[T->ntab2], 0, sizeof(tab2_t));
                char base[TLEN];
                table_path(param_path, tok[6], "", base, sizeof base);
                T->ntab2++;
                if (load_2b(&T->t2[T->ntab2 - 1], base)) goto fail;
            } else {
                if (T->ntab2) bad = 1;
                plain2 = 1;
            }
        }
        if (strstr(line, "TRIPLETTYPE PARAMS:")) in_trip = 1;
        else if (in_trip && strstr(line, "INDEX:")) {
            in_trip = 0;
            const int k = tab_split(copy, tok, 64);
            if (k == 8) {
                T->t3 = (tab3_t *)realloc(T->t3, sizeof(tab3_t) * (T->ntab3 + 1));
                memset(&T->t3[T->ntab3], 0, sizeof(tab3_t));
                char base[TLEN];
                table_path(param_path, tok[7], "", base, sizeof base);
                T->ntab3++;
                if (load_3b(&T->t3[T->ntab3 - 1], base)) goto fail;
            }
        }
    }
    fclose(fp);
    fp = NULL;
    if (bad || (T->ntab2 && T->ntab2 != oc_npairs(P))) {
        snprintf(g_tab_err, sizeof g_tab_err, "all 2-body parameters must be tabulated, or none");
        goto fail;
    }
    return T;
fail:
    if (fp) fclose(fp);
    oc_tab_free(T);
    return NULL;
}

int oc_tab_bodies(const oc_tables *T) { return (T && T->ntab2 ? 1 : 0) | (T && T->ntab3 ? 2 : 0); }

/* get_tab_2B (chimesFF.cpp:1858-1924).  lower_bound gives the first knot not
 * below rij; a knot equal to rij moves one interval left; otherwise the
 * spline piece of [x_i, x_i+1] is evaluated at rij < x_i (dx < 0), as the
 * reference does.  The piece's coefficients are the one-step natural-spline
 * recurrence with c_prev = 0. */
double oc_tab_2b_value(const oc_tables *T, int pair, double rij, int for_energy, int *err)
{
    const tab2_t *t = &T->t2[pair];
    const double *x = t->r;
    const int n = t->n;
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = lo + (hi - lo) / 2;
        if (x[mid] < rij) lo = mid + 1;
        else hi = mid;
    }
    int i = lo;
    if (i == n) return 0.0;
    if (i == 0) {
        *err = 1;
        return 0.0;
    }
    if (rij == x[i] && i > 0) i--;
    const double *y = for_energy ? t->e : t->f;
    if (i >= n - 1) return 0.0;
    if (i <= 0) {
        *err = 1;
        return 0.0;
    }
    double h = x[i + 1] - x[i];
    double alpha = (3 * (y[i + 1] - y[i]) / h) - (3 * (y[i] - y[i - 1]) / (x[i] - x[i - 1]));
    double l = 2 * (x[i + 1] - x[i - 1]) - h;
    double mu = h / l;
    double z = alpha / l;
    double c_prev = 0.0;
    double c = z - mu * c_prev;
    double b = (y[i + 1] - y[i]) / h - h * (c + 2 * c_prev) / 3.0;
    double d = (c - c_prev) / (3.0 * h);
    double a = y[i];
    double dx = rij - x[i];
    return a + b * dx + c * dx * dx + d * dx * dx * dx;
}

/* cubicInterpolate / bicubicInterpolate / tricubicInterpolate
 * (chimesFF.cpp:2317-2333): Catmull-Rom in the innermost index first. */
static double cr(double p0, double p1, double p2, double p3, double x)
{
    return p1 + 0.5 * x * (p2 - p0 + x * (2 * p0 - 5 * p1 + 4 * p2 - p3 + x * (3 * (p1 - p2) + p3 - p0)));
}

static double bicubic(const double *p, double x, double y)
{
    double a[4];
    for (int i = 0; i < 4; i++) a[i] = cr(p[i * 4 + 0], p[i * 4 + 1], p[i * 4 + 2], p[i * 4 + 3], x);
    return cr(a[0], a[1], a[2], a[3], y);
}

static double tricubic(const double *p, double x, double y, double z)
{
    double a[4];
    for (int i = 0; i < 4; i++) a[i] = bicubic(p + 16 * i, x, y);
    return cr(a[0], a[1], a[2], a[3], z);
}

/* interpolateTricubic (chimesFF.cpp:2335-2395): grid index from the first
 * row's distance and the spacing, at least 1; the 4x4x4 stencil around it,
 * zero past the last grid point; fractions from the rows at (i, j, k),
 * (0, j, k) and (0, 0, k); out = energy, then the force scalars of ij, ik, jk. */
void oc_tab_3b_interp(const oc_tables *T, int trip, double rij, double rik, double rjk, double out[4], int *err)
{
    const tab3_t *t = &T->t3[trip];
    const int size_ik = t->n, size_ij = size_ik * size_ik;
    const double dr = t->dr;
    int i = (int)((rij - t->rij[0]) * dr);
    int j = (int)((rik - t->rik[0]) * dr);
    int k = (int)((rjk - t->rjk[0]) * dr);
    i = i > 1 ? i : 1;
    j = j > 1 ? j : 1;
    k = k > 1 ? k : 1;
    if (i >= size_ik || j >= size_ik || k >= size_ik) {
        *err = 1;
        out[0] = out[1] = out[2] = out[3] = 0.0;
        return;
    }
    double v[4][64];
    memset(v, 0, sizeof v);
    for (int di = -1; di <= 2; di++)
        for (int dj = -1; dj <= 2; dj++)
            for (int dk = -1; dk <= 2; dk++) {
                const int ti = i + di, tj = j + dj, tk = k + dk;
                if (ti < size_ik && tj < size_ik && tk < size_ik) {
                    const long index = (long)size_ij * ti + (long)size_ik * tj + tk;
                    const int q = (di + 1) * 16 + (dj + 1) * 4 + (dk + 1);
                    v[0][q] = t->e[index];
                    v[1][q] = t->fij[index];
                    v[2][q] = t->fik[index];
                    v[3][q] = t->fjk[index];
                }
            }
    const double xi = (rij - t->rij[(long)i * size_ij + (long)j * size_ik + k]) * dr;
    const double yi = (rik - t->rik[(long)j * size_ik + k]) * dr;
    const double zi = (rjk - t->rjk[k]) * dr;
    for (int m = 0; m < 4; m++) out[m] = tricubic(v[m], zi, yi, xi);
}

int oc_tab_3b_count(const oc_tables *T) { return T ? T->ntab3 : 0; }

/* compute_2B_tab (chimesFF.cpp:1782-1856): cutoff test, energy and force
 * scalar from the tables, force and stress of the pair; the penalty term is
 * not added (the table is expected to hold it). */
int oc_compute_2b_tab(const oc_params *P, const oc_tables *T, double dx, const double dr[3], const int t[2],
                      double force[6], double stress[6], double *energy)
{
    const int pair = oc_pair_type(P, t[0], t[1]);
    if (dx >= oc_cutoff_2b(P, t[0], t[1])) return 0;
    int err = 0;
    *energy += oc_tab_2b_value(T, pair, dx, 1, &err);
    const double fs = oc_tab_2b_value(T, pair, dx, 0, &err);
    force[0] += fs * dr[0]; force[1] += fs * dr[1]; force[2] += fs * dr[2];
    force[3] -= fs * dr[0]; force[4] -= fs * dr[1]; force[5] -= fs * dr[2];
    stress[0] -= fs * dr[0] * dr[0];
    stress[1] -= fs * dr[0] * dr[1];
    stress[2] -= fs * dr[0] * dr[2];
    stress[3] -= fs * dr[1] * dr[1];
    stress[4] -= fs * dr[1] * dr[2];
    stress[5] -= fs * dr[2] * dr[2];
    return err;
}

static void tab_pair_update(double fs, const double *d, double *fa, double *fb, double *s)
{
    fa[0] += fs * d[0]; fa[1] += fs * d[1]; fa[2] += fs * d[2];
    fb[0] -= fs * d[0]; fb[1] -= fs * d[1]; fb[2] -= fs * d[2];
    s[0] -= fs * (d[0] * d[0]);
    s[1] -= fs * (d[0] * d[1]);
    s[2] -= fs * (d[0] * d[2]);
    s[3] -= fs * (d[1] * d[1]);
    s[4] -= fs * (d[1] * d[2]);
    s[5] -= fs * (d[2] * d[2]);
}

/* compute_3B_tab (chimesFF.cpp:2183-2313): excluded types and the cutoffs of
 * the mapped slots as compute_3B; the table is read at the runtime pair
 * distances (ij, ik, jk) in that order (get_tab_3B's sort is not executed,
 * 2459-2470); force and stress of ij, ik, jk with the distance-tensor
 * products. */
int oc_compute_3b_tab(const oc_params *P, const oc_tables *T, const double dx[3], const double dr[9], const int t[3],
                      double force[9], double stress[6], double *energy)
{
    const int trip = oc_trip_type(P, t);
    if (trip < 0) return 0;
    for (int p = 0; p < 3; p++) {
        double rin, rout;
        oc_cluster_pair_cutoffs(P, 3, t, p, &rin, &rout);
        if (dx[p] >= rout) return 0;
    }
    if (trip >= T->ntab3) return 1;
    int err = 0;
    double v[4];
    oc_tab_3b_interp(T, trip, dx[0], dx[1], dx[2], v, &err);
    *energy += v[0];
    tab_pair_update(v[1], dr + 0, force + 0, force + 3, stress);
    tab_pair_update(v[2], dr + 3, force + 0, force + 6, stress);
    tab_pair_update(v[3], dr + 6, force + 3, force + 6, stress);
    return err;
}
