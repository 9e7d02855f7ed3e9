/*
 * ORACLE -- test infrastructure only.
 *
 * Plain-C, CPU, FP64 restatement of the LAMMPS pair style of the reference
 * (etc/lmp/src/pair_chimes.cpp) on one process: the many-body neighbour
 * lists built from a LAMMPS full neighbour list with ghosts, the 1-/2-/3-/4-
 * body loops of PairCHIMES::compute and the energy/virial tally of
 * Pair::ev_tally_mb (etc/lmp/etc/pair.cpp:1425-1482).  The cluster terms are
 * the oracle's own oc_compute_{1,2,3,4}b.  Only tests/ and bench.py's
 * cpu_baseline leg load it.
 *
 * Inputs mirror LAMMPS's arrays: x[nall][3] (local atoms, then ghosts),
 * ctype[nall] (ChIMES type index of each atom, i.e. chimes_type[type[i]-1]),
 * tag[nall], ilist[inum], numneigh[] and the neighbours of ilist[ii] at
 * nbr[nbr_off[ii] .. nbr_off[ii] + numneigh[ilist[ii]]) (the rows of
 * firstneigh, packed).
 *
 * Parity pins (tests/test_lmp_oracle.py):
 *   - the pressure and the summed per-atom stress of the reference's own
 *     LAMMPS test (etc/lmp/tests/test_suite-ev_tally_mb: 512-atom diamond,
 *     Carbon-2.0 2+3+4-body, step 0);
 *   - on boxes wider than twice the padded cutoffs, where the tag rule and
 *     the serial path's parent rule select the same clusters, the energy,
 *     virial and the ghost-folded forces equal oc_calculate's;
 *   - the per-atom split sums to the global energy and virial, and 3- and
 *     4-mer lists on small boxes match a brute-force enumeration of the
 *     same predicates in the test.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "chimes_oracle.h"

/* get_dist (pair_chimes.cpp:303-312): dr = x_j - x_i, same operation order */
static double lmp_dist(const double *x, int i, int j, double dr[3])
{
    dr[0] = x[3 * j + 0] - x[3 * i + 0];
    dr[1] = x[3 * j + 1] - x[3 * i + 1];
    dr[2] = x[3 * j + 2] - x[3 * i + 2];
    return sqrt(dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2]);
}

typedef struct {
    long n, cap;
    int *v;
    int w;
} ivec;

static void ivec_push(ivec *a, const int *e)
{
    if (a->n == a->cap) {
        a->cap = a->cap ? 2 * a->cap : 1024;
        a->v = (int *)realloc(a->v, sizeof(int) * a->cap * a->w);
    }
    memcpy(a->v + a->n * a->w, e, sizeof(int) * a->w);
    a->n++;
}

/* PairCHIMES::build_mb_neighlists (pair_chimes.cpp:321-478).  trips: 3 ints
 * (i, j, k), quads: 4 ints (i, j, k, l), atom indices, in loop order.
 * Returns nonzero on allocation failure. */
int oc_lmp_build_lists(const oc_params *P, const double *x, const int *tag, int inum, const int *ilist,
                       const int *numneigh, const int *nbr_off, const int *nbr, double skin, long *n3, int **trips,
                       long *n4, int **quads)
{
    ivec t3 = {0, 0, NULL, 3}, t4 = {0, 0, NULL, 4};
    const int o3 = oc_order(P, 3), o4 = oc_order(P, 4);
    *n3 = *n4 = 0;
    *trips = *quads = NULL;
    if (o3 == 0 && o4 == 0) return 0;
    const double pad3 = oc_max_cutoff(P, 3) + skin, pad4 = oc_max_cutoff(P, 4) + skin;
    double dr[3];
    for (int ii = 0; ii < inum; ii++) {
        const int i = ilist[ii], itag = tag[i];
        const int *jl = nbr + nbr_off[ii];
        const int jn = numneigh[i];
        for (int jj = 0; jj < jn; jj++) {
            const int j = jl[jj];
            if (j == i || tag[j] < itag) continue;
            const double dij = lmp_dist(x, i, j, dr);
            if (dij >= pad3 && dij >= pad4) continue;
            for (int kk = 0; kk < jn; kk++) {
                const int k = jl[kk];
                if (k == i || k == j) continue;
                if (tag[k] < itag || tag[k] < tag[j]) continue;
                const double dik = lmp_dist(x, i, k, dr);
                if (dik >= pad3 && dik >= pad4) continue;
                const double djk = lmp_dist(x, j, k, dr);
                if (dij < pad3 && dik < pad3 && djk < pad3) {
                    const int e[3] = {i, j, k};
                    ivec_push(&t3, e);
                }
                if (dij >= pad4 || dik >= pad4 || djk >= pad4) continue;
                if (o4 == 0) continue;
                for (int ll = 0; ll < jn; ll++) {
                    const int l = jl[ll];
                    if (l == i || l == j || l == k) continue;
                    if (tag[l] < itag || tag[l] < tag[j] || tag[l] < tag[k]) continue;
                    if (lmp_dist(x, i, l, dr) >= pad4) continue;
                    if (lmp_dist(x, j, l, dr) >= pad4) continue;
                    if (lmp_dist(x, k, l, dr) >= pad4) continue;
                    const int e[4] = {i, j, k, l};
                    ivec_push(&t4, e);
                }
            }
        }
    }
    *n3 = t3.n;
    *trips = t3.v;
    *n4 = t4.n;
    *quads = t4.v;
    return 0;
}

/* Pair::ev_tally_mb (pair.cpp:1425-1482) for n interaction atoms given by
 * their atom indices; stress in ChIMES order (xx xy xz yy yz zz) is
 * tallied in LAMMPS order (xx yy zz xy xz yz). */
static void ev_tally_mb(int n, const int *atoms, double evdwl, const double s[6], int eflag_global,
                        int eflag_atom, int vflag_global, int vflag_atom, double *eng, double *virial,
                        double *eatom, double *vatom)
{
    if (eflag_global) *eng += evdwl;
    if (eflag_atom)
        for (int a = 0; a < n; a++) eatom[atoms[a]] += evdwl / n;
    if (n < 2) return;
    if (vflag_global) {
        virial[0] += s[0];
        virial[1] += s[3];
        virial[2] += s[5];
        virial[3] += s[1];
        virial[4] += s[2];
        virial[5] += s[4];
    }
    if (vflag_atom)
        for (int a = 0; a < n; a++) {
            double *v = vatom + 6L * atoms[a];
            v[0] += s[0] / n;
            v[1] += s[3] / n;
            v[2] += s[5] / n;
            v[3] += s[1] / n;
            v[4] += s[2] / n;
            v[5] += s[4] / n;
        }
}

static const oc_tables *g_tab;
static long g_tab_errors;

void oc_lmp_use_tables(const oc_tables *T)
{
    g_tab = T;
    g_tab_errors = 0;
}

long oc_lmp_tab_errors(void) { return g_tab_errors; }

/* PairCHIMES::compute (pair_chimes.cpp:480-831) over given many-body lists.
 * f[nall][3], eng, virial[6], eatom[nall], vatom[nall][6] are accumulated
 * onto, as the reference's force and tally arrays. */
void oc_lmp_compute(const oc_params *P, const double *x, const int *ctype, const int *tag, int inum,
                    const int *ilist, const int *numneigh, const int *nbr_off, const int *nbr, long n3,
                    const int *trips, long n4, const int *quads, int eflag_global, int eflag_atom, int vflag_global,
                    int vflag_atom, double *f, double *eng, double *virial, double *eatom, double *vatom)
{
    double s[6] = {0, 0, 0, 0, 0, 0}, fo[12], dist[6], dr[18], e;
    /* 1- and 2-body (pair_chimes.cpp:567-656) */
    for (int ii = 0; ii < inum; ii++) {
        const int i = ilist[ii], itag = tag[i];
        e = 0.0;
        oc_compute_1b(P, ctype[i], &e);
        ev_tally_mb(1, &i, e, s, eflag_global, eflag_atom, vflag_global, vflag_atom, eng, virial, eatom, vatom);
        const int *jl = nbr + nbr_off[ii];
        for (int jj = 0; jj < numneigh[i]; jj++) {
            const int j = jl[jj];
            if (tag[j] <= itag) continue;
            dist[0] = lmp_dist(x, i, j, dr);
            const int t[2] = {ctype[i], ctype[j]};
            memset(fo, 0, sizeof fo);
            memset(s, 0, sizeof s);
            e = 0.0;
            if (oc_tab_bodies(g_tab) & 1) g_tab_errors += oc_compute_2b_tab(P, g_tab, dist[0], dr, t, fo, s, &e) != 0;
            else oc_compute_2b(P, dist[0], dr, t, fo, s, &e);
            for (int d = 0; d < 3; d++) {
                f[3 * i + d] += fo[d];
                f[3 * j + d] += fo[3 + d];
            }
            const int at[2] = {i, j};
            ev_tally_mb(2, at, e, s, eflag_global, eflag_atom, vflag_global, vflag_atom, eng, virial, eatom, vatom);
        }
    }
    /* 3-body (pair_chimes.cpp:673-748): pairs ij, ik, jk */
    if (oc_order(P, 3) > 0)
        for (long c = 0; c < n3; c++) {
            const int *m = trips + 3 * c;
            dist[0] = lmp_dist(x, m[0], m[1], dr + 0);
            dist[1] = lmp_dist(x, m[0], m[2], dr + 3);
            dist[2] = lmp_dist(x, m[1], m[2], dr + 6);
            const int t[3] = {ctype[m[0]], ctype[m[1]], ctype[m[2]]};
            memset(fo, 0, sizeof fo);
            memset(s, 0, sizeof s);
            e = 0.0;
            if (oc_tab_bodies(g_tab) & 2) g_tab_errors += oc_compute_3b_tab(P, g_tab, dist, dr, t, fo, s, &e) != 0;
            else oc_compute_3b(P, dist, dr, t, fo, s, &e);
            for (int a = 0; a < 3; a++)
                for (int d = 0; d < 3; d++) f[3 * m[a] + d] += fo[3 * a + d];
            ev_tally_mb(3, m, e, s, eflag_global, eflag_atom, vflag_global, vflag_atom, eng, virial, eatom, vatom);
        }
    /* 4-body (pair_chimes.cpp:749-829): pairs ij, ik, il, jk, jl, kl */
    if (oc_order(P, 4) > 0)
        for (long c = 0; c < n4; c++) {
            const int *m = quads + 4 * c;
            static const int pa[6] = {0, 0, 0, 1, 1, 2}, pb[6] = {1, 2, 3, 2, 3, 3};
            for (int p = 0; p < 6; p++) dist[p] = lmp_dist(x, m[pa[p]], m[pb[p]], dr + 3 * p);
            const int t[4] = {ctype[m[0]], ctype[m[1]], ctype[m[2]], ctype[m[3]]};
            memset(fo, 0, sizeof fo);
            memset(s, 0, sizeof s);
            e = 0.0;
            oc_compute_4b(P, dist, dr, t, fo, s, &e);
            for (int a = 0; a < 4; a++)
                for (int d = 0; d < 3; d++) f[3 * m[a] + d] += fo[3 * a + d];
            ev_tally_mb(4, m, e, s, eflag_global, eflag_atom, vflag_global, vflag_atom, eng, virial, eatom, vatom);
        }
}

/* The cluster rows of the pair style's FINGERPRINT build: PairCHIMES::compute
 * passes a cluster to the row record only when its members are in order
 * (2-body: i < j by atom index, pair_chimes.cpp:625-626; 3-body: tag[i] <
 * tag[j], tag[i] < tag[k], tag[j] < tag[k], 703-704; 4-body: tag[i] < tag[j] <
 * tag[k] < tag[l], 777-778), and compute_nB records it after its exclusion
 * and cutoff checks (chimesFF.cpp:1691-1703, 1989-2022, 2575-2613): the pair
 * distances in runtime pair order, then the member types.  Rows come in the
 * loops' order.  Returns the number of rows; *rows (np + nb doubles each) is
 * malloc'd (free with oc_lmp_free), NULL when there are none. */
long oc_lmp_fp_rows(const oc_params *P, int nb, const double *x, const int *ctype, const int *tag, int inum,
                    const int *ilist, const int *numneigh, const int *nbr_off, const int *nbr, long n3,
                    const int *trips, long n4, const int *quads, double **rows)
{
    static const int pa[6] = {0, 0, 0, 1, 1, 2}, pb[6] = {1, 2, 3, 2, 3, 3};
    static const int pa3[3] = {0, 0, 1}, pb3[3] = {1, 2, 2};
    const int np = nb * (nb - 1) / 2, w = np + nb;
    long n = 0, cap = 0;
    double *out = NULL, dr[3], d[6], rin, rout;
    int m[4], t[4];
    const long nlist = nb == 2 ? 0 : (nb == 3 ? n3 : n4);
    const int *list = nb == 3 ? trips : quads;
    long total = 0;
    if (nb == 2)
        for (int ii = 0; ii < inum; ii++) total += numneigh[ilist[ii]];
    else
        total = nlist;
    for (long e = 0, ii = 0, jj = 0; e < total; e++) {
        if (nb == 2) {
            while (jj >= numneigh[ilist[ii]]) { ii++; jj = 0; }
            m[0] = ilist[ii];
            m[1] = nbr[nbr_off[ii] + jj];
            jj++;
            if (tag[m[1]] <= tag[m[0]] || !(m[0] < m[1])) continue;
        } else {
            for (int a = 0; a < nb; a++) m[a] = list[e * nb + a];
            if (nb == 3 && !(tag[m[0]] < tag[m[1]] && tag[m[0]] < tag[m[2]] && tag[m[1]] < tag[m[2]])) continue;
            if (nb == 4 && !(tag[m[0]] < tag[m[1]] && tag[m[1]] < tag[m[2]] && tag[m[2]] < tag[m[3]])) continue;
        }
        for (int a = 0; a < nb; a++) t[a] = ctype[m[a]];
        int inside = 1;
        for (int q = 0; q < np && inside; q++) {
            const int a = nb == 3 ? pa3[q] : pa[q], b = nb == 3 ? pb3[q] : pb[q];
            d[q] = lmp_dist(x, m[a], m[b], dr);
            if (oc_cluster_pair_cutoffs(P, nb, t, q, &rin, &rout) != 0) inside = 0;  /* excluded type */
            else if (d[q] >= rout) inside = 0;
        }
        if (!inside) continue;
        if (n == cap) {
            cap = cap ? 2 * cap : 4096;
            out = (double *)realloc(out, sizeof(double) * cap * w);
        }
        for (int q = 0; q < np; q++) out[n * w + q] = d[q];
        for (int a = 0; a < nb; a++) out[n * w + np + a] = (double)t[a];
        n++;
    }
    *rows = out;
    return n;
}

void oc_lmp_free(void *p) { free(p); }
