/*
 * ORACLE -- test infrastructure only (see chimes_oracle.h).
 *
 * CPU FP64 restatement of the ChIMES serial path.  It follows the reference's
 * arithmetic operation by operation (same association, no FMA contraction: the
 * Makefile builds it with -ffp-contract=off) so that it can be compared with
 * the compiled reference bit for bit, and with the HIP product to 1e-10.
 *
 * What pins each part (tests/test_oracle.py):
 *   parser + maps (oc_load)         : compiled reference (oracle/_ref) on every
 *                                     force field in serial_interface/tests,
 *                                     via end-to-end E/F/stress equality;
 *                                     round trip of generated files.
 *   oc_cheby                        : T_n(cos t) = cos(n t), dT/dx = n U_{n-1},
 *                                     Morse endpoints x(r_in)=+1, x(r_out)=-1,
 *                                     finite differences; out-of-range branch
 *                                     continuity at r_in.
 *   oc_fcut                         : closed forms (cubic, Tersoff).
 *   oc_compute_{2,3,4}b             : brute-force long-double term sums,
 *                                     finite-difference forces, virial identity.
 *   system setup + neighbour lists  : golden cluster lists and E/F/stress from
 *                                     the compiled reference (tests/golden):
 *                                     identical sorted lists on every fixture.
 *   oc_calculate                    : reference expected_output fixtures
 *                                     (1e-3), golden %.17g vectors (1e-12).
 */
#include "chimes_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define OC_PI 3.14159265359 /* truncated constant of chimesFF.h:22 */
#define NAMELEN 64
#define MAXTOK 64

struct oc_params {
    int nt;
    char (*atm)[NAMELEN];
    double *mass, *eoff;
    int order[3];
    /* pairs */
    int npairs;
    char (*pa1)[NAMELEN], (*pa2)[NAMELEN];
    double *rin, *rout, *lam;
    int tersoff;
    double fcut_var, pen_dist, pen_scale, smooth;
    int *nc2;
    int **pow2;
    double **c2;
    int *pairmap;             /* [nt*nt] parameter pair index */
    char (*prpr)[NAMELEN];    /* [nt*nt] proper pair name */
    /* triplets */
    int ntrip;
    char (*trip_atm)[3][NAMELEN];
    char (*trip_pr)[3][NAMELEN];
    int *nc3;
    int **pow3;               /* [K*3] */
    double **c3;
    double (*cut3)[2][3];
    int *tripmap;             /* [nt^3] */
    int *trip_pm;             /* [nt^3 * 3] */
    /* quadruplets */
    int nquad;
    char (*quad_atm)[4][NAMELEN];
    char (*quad_pr)[6][NAMELEN];
    int *nc4;
    int **pow4;               /* [K*6] */
    double **c4;
    double (*cut4)[2][6];
    int *quadmap;             /* [nt^4] */
    int *quad_pm;             /* [nt^4 * 6] */
};

static long g_penalty_count, g_oor_count;
long oc_last_penalty_count(void) { return g_penalty_count; }
long oc_last_out_of_range_count(void) { return g_oor_count; }

/* ------------------------------------------------------------------------ */
/* Line reader and tokenizer (chimesFF.cpp:148-197)                          */
/* ------------------------------------------------------------------------ */

typedef struct {
    char **lines;
    int n;
    int pos;
    int err;
} lines_t;

static int read_lines(const char *path, lines_t *L)
{
    FILE *f = fopen(path, "rb");
    if (!f) return -1;
    fseek(f, 0, SEEK_END);
    long sz = ftell(f);
    fseek(f, 0, SEEK_SET);
    char *buf = (char *)malloc(sz + 2);
    if (fread(buf, 1, sz, f) != (size_t)sz) { fclose(f); free(buf); return -1; }
    fclose(f);
    buf[sz] = 0;
    int cap = 1024;
    L->lines = (char **)malloc(sizeof(char *) * cap);
    L->n = 0;
    char *p = buf;
    while (*p) {
        char *e = strchr(p, '\n');
        if (L->n == cap) { cap *= 2; L->lines = (char **)realloc(L->lines, sizeof(char *) * cap); }
        L->lines[L->n++] = p;
        if (!e) break;
        *e = 0;
        p = e + 1;
    }
    L->pos = 0;
    L->err = 0;
    if (L->n == 0) { free(buf); free(L->lines); return -1; }
    return 0;
}

static void free_lines(lines_t *L)
{
    if (L->n) free(L->lines[0]);
    free(L->lines);
}

/* get_next_line: the reference exits when the stream runs out. */
static const char *next_line(lines_t *L)
{
    if (L->pos >= L->n) { L->err = 1; return ""; }
    return L->lines[L->pos++];
}

/* split_line: strip at '!' and "##", then whitespace tokens. */
static int split(const char *line, char tok[][NAMELEN])
{
    char tmp[4096];
    strncpy(tmp, line, sizeof(tmp) - 1);
    tmp[sizeof(tmp) - 1] = 0;
    char *c = strchr(tmp, '!');
    if (c) *c = 0;
    c = strstr(tmp, "##");
    if (c) *c = 0;
    int n = 0;
    char *s = strtok(tmp, " \t\r\n\v\f");
    while (s && n < MAXTOK) {
        strncpy(tok[n], s, NAMELEN - 1);
        tok[n][NAMELEN - 1] = 0;
        n++;
        s = strtok(NULL, " \t\r\n\v\f");
    }
    return n;
}

static int find_name(char (*names)[NAMELEN], int n, const char *s)
{
    for (int i = 0; i < n; i++)
        if (strcmp(names[i], s) == 0) return i;
    return -1;
}

/* get_proper_pair (chimesFF.cpp:68-84) */
static int proper_pair(const oc_params *P, const char *t1, const char *t2)
{
    for (int i = 0; i < P->npairs; i++) {
        if (!strcmp(t1, P->pa1[i]) && !strcmp(t2, P->pa2[i])) return i;
        if (!strcmp(t2, P->pa1[i]) && !strcmp(t1, P->pa2[i])) return i;
    }
    return -1;
}

/* ------------------------------------------------------------------------ */
/* Permutation search (chimesFF.h:438-539)                                   */
/* ------------------------------------------------------------------------ */

static int next_perm(int *a, int n)
{
    int i = n - 2;
    while (i >= 0 && a[i] >= a[i + 1]) i--;
    if (i < 0) {
        for (int l = 0, r = n - 1; l < r; l++, r--) { int t = a[l]; a[l] = a[r]; a[r] = t; }
        return 0;
    }
    int j = n - 1;
    while (a[j] <= a[i]) j--;
    int t = a[i]; a[i] = a[j]; a[j] = t;
    for (int l = i + 1, r = n - 1; l < r; l++, r--) { t = a[l]; a[l] = a[r]; a[r] = t; }
    return 1;
}

/* mapped[j] = index i of the parameter pair that runtime pair j plays.
 * The first permutation (in next_permutation order) of the atoms whose
 * proper pair names equal the parameter block's pair names is used. */
static void atom_pair_mapper(const oc_params *P, int natoms, const int *typ,
                             char (*clu_pairs)[NAMELEN], int *mapped)
{
    int perm[4], tp[6][2], rp[6][2];
    int npairs = natoms * (natoms - 1) / 2;
    for (int i = 0; i < natoms; i++) perm[i] = i;
    int idx = 0;
    for (int i = 0; i < natoms; i++)
        for (int j = i + 1; j < natoms; j++) { tp[idx][0] = i; tp[idx][1] = j; idx++; }
    for (int i = 0; i < npairs; i++) rp[i][0] = rp[i][1] = 0;
    do {
        int match = 1;
        idx = 0;
        for (int i = 0; i < natoms && match; i++)
            for (int j = i + 1; j < natoms; j++) {
                const char *nm = P->prpr[typ[perm[i]] * P->nt + typ[perm[j]]];
                if (strcmp(clu_pairs[idx], nm) != 0) { match = 0; break; }
                idx++;
            }
        if (match) {
            idx = 0;
            for (int i = 0; i < natoms; i++)
                for (int j = i + 1; j < natoms; j++) { rp[idx][0] = perm[i]; rp[idx][1] = perm[j]; idx++; }
            break;
        }
    } while (next_perm(perm, natoms));
    for (int i = 0; i < npairs; i++)
        for (int j = 0; j < npairs; j++)
            if ((rp[i][0] == tp[j][0] && rp[i][1] == tp[j][1]) ||
                (rp[i][0] == tp[j][1] && rp[i][1] == tp[j][0]))
                mapped[j] = i;
}

/* ------------------------------------------------------------------------ */
/* Parameter file (chimesFF.cpp:354-1570)                                    */
/* ------------------------------------------------------------------------ */

#define TOKS char tok[MAXTOK][NAMELEN]

static int pass_orders(oc_params *P, lines_t *L)
{
    TOKS;
    L->pos = 0;
    for (;;) {
        const char *line = next_line(L);
        if (L->err) return -1;
        if (strstr(line, "ENDFILE")) return -1;
        if (strstr(line, "PAIRTYP: CHEBYSHEV")) {
            int n = split(line, tok);
            if (n < 3) return -1;
            P->order[0] = atoi(tok[2]);
            P->order[1] = n >= 4 ? atoi(tok[3]) : 0;
            P->order[2] = n >= 5 ? atoi(tok[4]) : 0;
            return 0;
        }
    }
}

static int pass_general(oc_params *P, lines_t *L)
{
    TOKS;
    int no_pairs = 0;
    L->pos = 0;
    for (;;) {
        const char *line = next_line(L);
        if (L->err) return -1;
        if (strstr(line, "ENDFILE")) break;
        if (strstr(line, "ATOM TYPES:")) {
            split(line, tok);
            P->nt = atoi(tok[2]);
            P->eoff = (double *)calloc(P->nt, sizeof(double));
        }
        if (strstr(line, "# TYPEIDX #")) {
            P->atm = calloc(P->nt, NAMELEN);
            P->mass = (double *)calloc(P->nt, sizeof(double));
            for (int i = 0; i < P->nt; i++) {
                const char *l2 = next_line(L);
                if (L->err) return -1;
                split(l2, tok);
                strcpy(P->atm[i], tok[1]);
                P->mass[i] = strtod(tok[3], NULL);
            }
        }
        if (strstr(line, "ATOM PAIRS:")) {
            split(line, tok);
            no_pairs = atoi(tok[2]);
        }
        if (strstr(line, "# PAIRIDX #")) {
            if (strstr(line, "# USEOVRP #")) continue;
            P->npairs = no_pairs;
            P->pa1 = calloc(no_pairs, NAMELEN);
            P->pa2 = calloc(no_pairs, NAMELEN);
            P->rin = (double *)calloc(no_pairs, sizeof(double));
            P->rout = (double *)calloc(no_pairs, sizeof(double));
            P->lam = (double *)calloc(no_pairs, sizeof(double));
            char style0[NAMELEN] = "";
            for (int i = 0; i < no_pairs; i++) {
                const char *l2 = next_line(L);
                if (L->err) return -1;
                int n = split(l2, tok);
                int sidx, midx;
                if (n == 8) { sidx = 6; midx = 7; }
                else if (n == 7) { sidx = 5; midx = 6; }
                else return -1;
                strcpy(P->pa1[i], tok[1]);
                strcpy(P->pa2[i], tok[2]);
                P->rin[i] = strtod(tok[3], NULL);
                P->rout[i] = strtod(tok[4], NULL);
                if (i == 0) strcpy(style0, tok[sidx]);
                else if (strcmp(style0, tok[sidx])) return -1;
                if (!strcmp(style0, "MORSE")) {
                    if (n > midx) P->lam[i] = strtod(tok[midx], NULL);
                    else return -1;
                }
            }
        }
        if (strstr(line, "FCUT TYPE:")) {
            split(line, tok);
            if (!strcmp(tok[2], "CUBIC")) P->tersoff = 0;
            else if (!strcmp(tok[2], "TERSOFF")) { P->tersoff = 1; P->fcut_var = strtod(tok[3], NULL); }
            else return -1;
        }
        if (strstr(line, "PAIR CHEBYSHEV PENALTY DIST:")) {
            split(line, tok);
            P->pen_dist = strtod(tok[4], NULL);
        }
        if (strstr(line, "PAIR CHEBYSHEV PENALTY SCALING:")) {
            split(line, tok);
            P->pen_scale = strtod(tok[4], NULL);
        }
        if (strstr(line, "NO ENERGY OFFSETS:")) {
            int n = split(line, tok);
            if (atoi(tok[n - 1]) != P->nt) return -1;
            for (int i = 0; i < P->nt; i++) {
                const char *l2 = next_line(L);
                if (L->err) return -1;
                split(l2, tok);
                int k = atoi(tok[2]);
                P->eoff[k - 1] = strtod(tok[3], NULL);
            }
        }
    }
    return 0;
}

static int pass_2b(oc_params *P, lines_t *L)
{
    TOKS;
    L->pos = 0;
    P->nc2 = (int *)calloc(P->npairs, sizeof(int));
    P->pow2 = (int **)calloc(P->npairs, sizeof(int *));
    P->c2 = (double **)calloc(P->npairs, sizeof(double *));
    for (;;) {
        const char *line = next_line(L);
        if (L->err) return -1;
        if (strstr(line, "ENDFILE")) break;
        if (strstr(line, "PAIRTYPE PARAMS:")) {
            split(line, tok);
            int p = atoi(tok[2]);
            if (p < 0 || p >= P->npairs) return -1;
            next_line(L); /* blank line */
            int K = P->order[0];
            P->nc2[p] = K;
            P->pow2[p] = (int *)malloc(sizeof(int) * K);
            P->c2[p] = (double *)malloc(sizeof(double) * K);
            for (int i = 0; i < K; i++) {
                const char *l2 = next_line(L);
                if (L->err) return -1;
                split(l2, tok);
                P->pow2[p][i] = atoi(tok[0]);
                P->c2[p][i] = strtod(tok[1], NULL);
            }
        }
        if (strstr(line, "PAIRMAPS:")) {
            split(line, tok);
            int nm = atoi(tok[1]);
            char (*names)[NAMELEN] = calloc(nm, NAMELEN);
            int *idx = (int *)calloc(nm, sizeof(int));
            for (int i = 0; i < nm; i++) {
                const char *l2 = next_line(L);
                if (L->err) { free(names); free(idx); return -1; }
                split(l2, tok);
                idx[i] = atoi(tok[0]);
                strcpy(names[i], tok[1]);
            }
            int nt = P->nt;
            P->pairmap = (int *)calloc(nt * nt, sizeof(int));
            P->prpr = calloc(nt * nt, NAMELEN);
            for (int i = 0; i < nt; i++)
                for (int j = 0; j < nt; j++) {
                    char nmij[2 * NAMELEN];
                    snprintf(nmij, sizeof nmij, "%s%s", P->atm[i], P->atm[j]);
                    int k = find_name(names, nm, nmij);
                    if (k < 0) { free(names); free(idx); return -1; }
                    P->pairmap[i * nt + j] = idx[k];
                    int pp = proper_pair(P, P->atm[i], P->atm[j]);
                    if (pp < 0) { free(names); free(idx); return -1; }
                    snprintf(P->prpr[i * nt + j], NAMELEN, "%s%s", P->pa1[pp], P->pa2[pp]);
                }
            free(names);
            free(idx);
        }
    }
    return 0;
}

/* Reads a "SPECIAL nB S_MAXIM/S_MINIM" block for body order nb (3 or 4);
 * which = 1 for the outer (S_MAXIM), 0 for the inner (S_MINIM) cutoffs. */
static int special_block(oc_params *P, lines_t *L, char tok[][NAMELEN], int nb, int which,
                         char (*slow)[NAMELEN], int *slow_idx, int nslow)
{
    int npr = nb == 3 ? 3 : 6;
    int ncl = nb == 3 ? P->ntrip : P->nquad;
    if (!strcmp(tok[3], "ALL")) {
        double v = strtod(tok[4], NULL);
        for (int i = 0; i < ncl; i++)
            for (int q = 0; q < npr; q++) {
                if (nb == 3) P->cut3[i][which][q] = v;
                else P->cut4[i][which][q] = v;
            }
        return 0;
    }
    int nent = atoi(tok[4]);
    if (nb == 4 && which == 0) /* chimesFF.cpp:1511-1519 */
        for (int i = 0; i < ncl; i++)
            for (int q = 0; q < 6; q++) P->cut4[i][0][q] = -1.0;
    for (int e = 0; e < nent; e++) {
        char t2[MAXTOK][NAMELEN];
        const char *l2 = next_line(L);
        if (L->err) return -1;
        split(l2, t2);
        int k = find_name(slow, nslow, t2[0]);
        if (k < 0) return -1;
        int ci = slow_idx[k];
        if (ci < 0 || ci >= ncl) return -1;
        int disq[6] = {0, 0, 0, 0, 0, 0};
        for (int q = 0; q < npr; q++) {
            const char *nm = t2[1 + q];
            double v = strtod(t2[1 + npr + q], NULL);
            int hit = -1;
            for (int r = 0; r < npr; r++) {
                const char *cand = nb == 3 ? P->trip_pr[ci][r] : P->quad_pr[ci][r];
                if (!strcmp(cand, nm) && !disq[r]) { disq[r] = 1; hit = r; break; }
            }
            if (hit < 0) return -1;
            if (nb == 3) P->cut3[ci][which][hit] = v;
            else P->cut4[ci][which][hit] = v;
        }
    }
    return 0;
}

/* 3B block (chimesFF.cpp:834-1169) and 4B block (chimesFF.cpp:1173-1567). */
static int pass_nb(oc_params *P, lines_t *L, int nb)
{
    TOKS;
    const int npr = nb == 3 ? 3 : 6;
    const int nat = nb;
    int ncl = 0;
    int nslow = 0;
    char (*slow)[NAMELEN] = NULL;
    int *slow_idx = NULL;
    int nt = P->nt;
    int ntup = nb == 3 ? nt * nt * nt : nt * nt * nt * nt;
    int *fastmap = NULL;

    L->pos = 0;
    for (;;) {
        const char *line = next_line(L);
        if (L->err) goto fail;
        if (strstr(line, "ENDFILE")) break;
        if (nb == 3 && strstr(line, "ATOM PAIR TRIPLETS:")) {
            split(line, tok);
            ncl = P->ntrip = atoi(tok[3]);
            P->trip_atm = calloc(ncl, sizeof(*P->trip_atm));
            P->trip_pr = calloc(ncl, sizeof(*P->trip_pr));
            P->nc3 = (int *)calloc(ncl, sizeof(int));
            P->pow3 = (int **)calloc(ncl, sizeof(int *));
            P->c3 = (double **)calloc(ncl, sizeof(double *));
            P->cut3 = calloc(ncl, sizeof(*P->cut3));
        }
        if (nb == 4 && strstr(line, "ATOM PAIR QUADRUPLETS:")) {
            split(line, tok);
            ncl = P->nquad = atoi(tok[3]);
            P->quad_atm = calloc(ncl, sizeof(*P->quad_atm));
            P->quad_pr = calloc(ncl, sizeof(*P->quad_pr));
            P->nc4 = (int *)calloc(ncl, sizeof(int));
            P->pow4 = (int **)calloc(ncl, sizeof(int *));
            P->c4 = (double **)calloc(ncl, sizeof(double *));
            P->cut4 = calloc(ncl, sizeof(*P->cut4));
        }
        if ((nb == 3 && strstr(line, "TRIPLETTYPE PARAMS:")) ||
            (nb == 4 && strstr(line, "QUADRUPLETYPE PARAMS:"))) {
            const char *l2 = next_line(L);
            if (L->err) goto fail;
            split(l2, tok);
            int ci = atoi(tok[1]);
            if (ci < 0 || ci >= ncl) goto fail;
            for (int a = 0; a < nat; a++) {
                if (nb == 3) strcpy(P->trip_atm[ci][a], tok[3 + a]);
                else strcpy(P->quad_atm[ci][a], tok[3 + a]);
            }
            l2 = next_line(L);
            if (L->err) goto fail;
            int ntk = split(l2, tok);
            for (int q = 0; q < npr; q++) {
                if (nb == 3) strcpy(P->trip_pr[ci][q], tok[1 + q]);
                else strcpy(P->quad_pr[ci][q], tok[1 + q]);
            }
            if (ntk > npr + 1 && strcmp(tok[1 + npr], "EXCLUDED:") != 0) {
                int K = atoi(tok[npr + 4]);
                next_line(L);
                next_line(L);
                int *pw = (int *)malloc(sizeof(int) * (K > 0 ? K : 1) * npr);
                double *cf = (double *)malloc(sizeof(double) * (K > 0 ? K : 1));
                for (int k = 0; k < K; k++) {
                    const char *l3 = next_line(L);
                    if (L->err) { free(pw); free(cf); goto fail; }
                    split(l3, tok);
                    for (int q = 0; q < npr; q++) pw[k * npr + q] = atoi(tok[1 + q]);
                    cf[k] = strtod(tok[npr + 3], NULL);
                }
                if (nb == 3) { P->nc3[ci] = K; P->pow3[ci] = pw; P->c3[ci] = cf; }
                else { P->nc4[ci] = K; P->pow4[ci] = pw; P->c4[ci] = cf; }
            }
        }
        if ((nb == 3 && strstr(line, "TRIPMAPS:")) || (nb == 4 && strstr(line, "QUADMAPS:"))) {
            split(line, tok);
            nslow = atoi(tok[1]);
            slow = calloc(nslow, NAMELEN);
            slow_idx = (int *)calloc(nslow, sizeof(int));
            for (int i = 0; i < nslow; i++) {
                const char *l2 = next_line(L);
                if (L->err) goto fail;
                split(l2, tok);
                slow_idx[i] = atoi(tok[0]);
                strcpy(slow[i], tok[1]);
            }
            fastmap = (int *)calloc(ntup, sizeof(int));
            for (int t = 0; t < ntup; t++) {
                int ty[4];
                int r = t;
                for (int a = nat - 1; a >= 0; a--) { ty[a] = r % nt; r /= nt; }
                char nm[8 * NAMELEN] = "";
                for (int a = 0; a < nat; a++)
                    for (int b = a + 1; b < nat; b++) {
                        int pp = proper_pair(P, P->atm[ty[a]], P->atm[ty[b]]);
                        if (pp < 0) goto fail;
                        strcat(nm, P->pa1[pp]);
                        strcat(nm, P->pa2[pp]);
                    }
                int k = find_name(slow, nslow, nm);
                if (k < 0) goto fail;
                fastmap[t] = slow_idx[k];
            }
        }
    }
    if (!fastmap) goto fail;
    /* default cutoffs from the 2B pair cutoffs (chimesFF.cpp:1014-1039, 1358-1393) */
    for (int i = 0; i < ncl; i++) {
        int at[4];
        for (int a = 0; a < nat; a++) {
            const char *nm = nb == 3 ? P->trip_atm[i][a] : P->quad_atm[i][a];
            at[a] = find_name(P->atm, nt, nm);
            if (at[a] < 0) at[a] = 0; /* reference reads out of range here */
        }
        int q = 0;
        for (int a = 0; a < nat; a++)
            for (int b = a + 1; b < nat; b++) {
                int pt = P->pairmap[at[a] * nt + at[b]];
                if (nb == 3) { P->cut3[i][0][q] = P->rin[pt]; P->cut3[i][1][q] = P->rout[pt]; }
                else { P->cut4[i][0][q] = P->rin[pt]; P->cut4[i][1][q] = P->rout[pt]; }
                q++;
            }
    }
    /* special cutoffs */
    L->pos = 0;
    for (;;) {
        const char *line = next_line(L);
        if (L->err) goto fail;
        if (strstr(line, "ENDFILE")) break;
        const char *kmax = nb == 3 ? "SPECIAL 3B S_MAXIM:" : "SPECIAL 4B S_MAXIM:";
        const char *kmin = nb == 3 ? "SPECIAL 3B S_MINIM:" : "SPECIAL 4B S_MINIM:";
        if (strstr(line, kmax)) {
            split(line, tok);
            if (special_block(P, L, tok, nb, 1, slow, slow_idx, nslow)) goto fail;
        }
        if (strstr(line, kmin)) {
            split(line, tok);
            if (special_block(P, L, tok, nb, 0, slow, slow_idx, nslow)) goto fail;
        }
    }
    /* build_pair_int_{trip,quad}_map (chimesFF.cpp:2958-3080); the map buffer
     * persists across tuples exactly like the reference's pair_map vector. */
    {
        int *pm = (int *)calloc(ntup * npr, sizeof(int));
        int buf[6] = {0, 0, 0, 0, 0, 0};
        for (int t = 0; t < ntup; t++) {
            int ci = fastmap[t];
            if (ci < 0) continue;
            int ty[4];
            int r = t;
            for (int a = nat - 1; a >= 0; a--) { ty[a] = r % nt; r /= nt; }
            atom_pair_mapper(P, nat, ty, nb == 3 ? P->trip_pr[ci] : P->quad_pr[ci], buf);
            for (int q = 0; q < npr; q++) pm[t * npr + q] = buf[q];
        }
        if (nb == 3) { P->tripmap = fastmap; P->trip_pm = pm; }
        else { P->quadmap = fastmap; P->quad_pm = pm; }
    }
    free(slow);
    free(slow_idx);
    return 0;
fail:
    free(slow);
    free(slow_idx);
    free(fastmap);
    return -1;
}

oc_params *oc_load(const char *path)
{
    lines_t L;
    if (read_lines(path, &L)) return NULL;
    oc_params *P = (oc_params *)calloc(1, sizeof(oc_params));
    P->pen_dist = 0.01; /* chimesFF.cpp:110-113 */
    P->pen_scale = 1.0e4;
    P->smooth = 0.01;
    int ok = !pass_orders(P, &L) && !pass_general(P, &L) && P->nt > 0 && P->npairs > 0 &&
             !pass_2b(P, &L) && P->pairmap;
    if (ok && P->order[1] > 0) ok = !pass_nb(P, &L, 3);
    if (ok && P->order[2] > 0) ok = !pass_nb(P, &L, 4);
    free_lines(&L);
    if (!ok) { oc_free(P); return NULL; }
    return P;
}

void oc_free(oc_params *P)
{
    if (!P) return;
    for (int i = 0; i < P->npairs && P->pow2; i++) { free(P->pow2[i]); free(P->c2[i]); }
    for (int i = 0; i < P->ntrip && P->pow3; i++) { free(P->pow3[i]); free(P->c3[i]); }
    for (int i = 0; i < P->nquad && P->pow4; i++) { free(P->pow4[i]); free(P->c4[i]); }
    free(P->atm); free(P->mass); free(P->eoff);
    free(P->pa1); free(P->pa2); free(P->rin); free(P->rout); free(P->lam);
    free(P->nc2); free(P->pow2); free(P->c2); free(P->pairmap); free(P->prpr);
    free(P->trip_atm); free(P->trip_pr); free(P->nc3); free(P->pow3); free(P->c3);
    free(P->cut3); free(P->tripmap); free(P->trip_pm);
    free(P->quad_atm); free(P->quad_pr); free(P->nc4); free(P->pow4); free(P->c4);
    free(P->cut4); free(P->quadmap); free(P->quad_pm);
    free(P);
}

int oc_order(const oc_params *P, int body) { return P->order[body - 2]; }
int oc_ntypes(const oc_params *P) { return P->nt; }
const char *oc_type_name(const oc_params *P, int t) { return P->atm[t]; }

/* max_cutoff_{2,3,4}B (chimesFF.cpp:2890-2943) */
double oc_max_cutoff(const oc_params *P, int body)
{
    if (body == 2) {
        double m = P->rout[0];
        for (int i = 0; i < P->npairs; i++) if (P->rout[i] > m) m = P->rout[i];
        return m;
    }
    if (body == 3) {
        if (P->order[1] == 0) return 0.0;
        double m = P->cut3[0][1][0];
        for (int i = 0; i < P->ntrip; i++)
            for (int q = 0; q < 3; q++) if (P->cut3[i][1][q] > m) m = P->cut3[i][1][q];
        return m;
    }
    if (P->order[2] == 0) return 0.0;
    double m = P->cut4[0][1][0];
    for (int i = 0; i < P->nquad; i++)
        for (int q = 0; q < 6; q++) if (P->cut4[i][1][q] > m) m = P->cut4[i][1][q];
    return m;
}

/* ------------------------------------------------------------------------ */
/* Polynomials, cutoff, penalty                                              */
/* ------------------------------------------------------------------------ */

/* set_cheby_polys (chimesFF.h:542-628) + set_polys_out_of_range
 * (chimesFF.cpp:1572-1613). */
void oc_cheby(double dx, double morse, double rin, double rout, int order, double *Tn,
              double *Tnd)
{
    double x_min = exp(-1 * rin / morse);
    double x_max = exp(-1 * rout / morse);
    double x_avg = 0.5 * (x_max + x_min);
    double x_diff = 0.5 * (x_max - x_min);
    x_diff *= -1.0;
    int oor = 0;
    double dx_orig = dx;
    if (dx < rin) { oor = 1; dx = rin; }
    double exprlen = exp(-1 * dx / morse);
    double x = (exprlen - x_avg) / x_diff;
    double dx_dr = (-exprlen / morse) / x_diff;
    Tn[0] = 1.0;
    Tn[1] = x;
    Tnd[0] = 1.0;
    Tnd[1] = 2.0 * x;
    for (int i = 2; i <= order; i++) {
        Tn[i] = 2.0 * x * Tn[i - 1] - Tn[i - 2];
        Tnd[i] = 2.0 * x * Tnd[i - 1] - Tnd[i - 2];
    }
    for (int i = order; i >= 1; i--) Tnd[i] = i * dx_dr * Tnd[i - 1];
    Tnd[0] = 0.0;
    if (oor) {
        g_oor_count++;
        double damp = exp((dx_orig - rin) / 0.01);
        for (int i = 0; i <= order; i++) {
            Tn[i] += 0.01 * (damp - 1.0) * Tnd[i];
            Tnd[i] *= damp;
        }
    }
}

/* get_fcut (chimesFF.h:347-386) */
void oc_fcut(const oc_params *P, double dx, double rc, double *fcut, double *fcutd)
{
    if (!P->tersoff) {
        double f0 = (1.0 - dx / rc);
        *fcut = pow(f0, 3.0);
        *fcutd = pow(f0, 2.0);
        *fcutd *= -1.0 * 3.0 / rc;
    } else {
        double thr = rc - P->fcut_var * rc;
        if (dx < thr) { *fcut = 1.0; *fcutd = 0.0; }
        else if (dx > rc) { *fcut = 0.0; *fcutd = 0.0; }
        else {
            double f0 = (dx - thr) / (rc - thr) * OC_PI + OC_PI / 2.0;
            double f0d = OC_PI / (rc - thr);
            *fcut = 0.5 + 0.5 * sin(f0);
            *fcutd = 0.5 * cos(f0) * f0d;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Cluster evaluators                                                        */
/* ------------------------------------------------------------------------ */

/* compute_2B (chimesFF.cpp:1657-1775) incl. get_penalty (chimesFF.h:388-419) */
/* chimesFF::compute_1B (chimesFF.cpp:1636-1641) */
void oc_compute_1b(const oc_params *P, int t, double *energy) { *energy += P->eoff[t]; }

double oc_mass(const oc_params *P, int t) { return P->mass[t]; }

/* Inner and outer cutoff of runtime pair p of an nb-body cluster of the
 * given member types: the pair's own cutoffs for nb = 2, else the cluster
 * type's cutoffs at the parameter slot the pair maps to (the cutoffs
 * compute_3B/4B test, chimesFF.cpp:1990-2011, 2560-2600).  Nonzero for an
 * excluded cluster type. */
int oc_cluster_pair_cutoffs(const oc_params *P, int nb, const int *t, int p, double *rin, double *rout)
{
    const int nt = P->nt;
    if (nb == 2) {
        const int pair = P->pairmap[t[0] * nt + t[1]];
        *rin = P->rin[pair];
        *rout = P->rout[pair];
        return 0;
    }
    if (nb == 3) {
        const int tix = (t[0] * nt + t[1]) * nt + t[2];
        const int trip = P->tripmap[tix];
        if (trip < 0) return 1;
        const int m = P->trip_pm[tix * 3 + p];
        *rin = P->cut3[trip][0][m];
        *rout = P->cut3[trip][1][m];
        return 0;
    }
    const int tix = ((t[0] * nt + t[1]) * nt + t[2]) * nt + t[3];
    const int quad = P->quadmap[tix];
    if (quad < 0) return 1;
    const int m = P->quad_pm[tix * 6 + p];
    *rin = P->cut4[quad][0][m];
    *rout = P->cut4[quad][1][m];
    return 0;
}

int oc_npairs(const oc_params *P) { return P->npairs; }

/* Parameter pair index of two atom types (atom_int_pair_map). */
int oc_pair_type(const oc_params *P, int ti, int tj) { return P->pairmap[ti * P->nt + tj]; }

/* Triplet type of three atom types (atom_int_trip_map); negative: excluded. */
int oc_trip_type(const oc_params *P, const int t[3])
{
    return P->tripmap[(t[0] * P->nt + t[1]) * P->nt + t[2]];
}

/* Morse lambda of the pair type of two atom types (chimesFF.cpp:354-1570). */
double oc_morse_lambda(const oc_params *P, int ti, int tj) { return P->lam[P->pairmap[ti * P->nt + tj]]; }

/* Outer 2-body cutoff of a type pair (chimesFF::get_cutoff_2B, chimesFF.cpp:2886-2901). */
double oc_cutoff_2b(const oc_params *P, int ti, int tj) { return P->rout[P->pairmap[ti * P->nt + tj]]; }

void oc_compute_2b(const oc_params *P, double dx, const double dr[3], const int t[2],
                   double force[6], double stress[6], double *energy)
{
    int pair = P->pairmap[t[0] * P->nt + t[1]];
    if (dx >= P->rout[pair]) return;
    double Tn[64], Tnd[64], fc, fcd;
    oc_cheby(dx, P->lam[pair], P->rin[pair], P->rout[pair], P->order[0], Tn, Tnd);
    oc_fcut(P, dx, P->rout[pair], &fc, &fcd);
    double dx_inv = (dx > 0.0) ? 1.0 / dx : 1e20;
    for (int c = 0; c < P->nc2[pair]; c++) {
        double cv = P->c2[pair][c];
        int p = P->pow2[pair][c] + 1;
        *energy += cv * fc * Tn[p];
        double deriv = fc * Tnd[p] + fcd * Tn[p];
        double fs = cv * deriv * dx_inv;
        force[0] += fs * dr[0]; force[1] += fs * dr[1]; force[2] += fs * dr[2];
        force[3] -= fs * dr[0]; force[4] -= fs * dr[1]; force[5] -= fs * dr[2];
        stress[0] -= fs * dr[0] * dr[0];
        stress[1] -= fs * dr[0] * dr[1];
        stress[2] -= fs * dr[0] * dr[2];
        stress[3] -= fs * dr[1] * dr[1];
        stress[4] -= fs * dr[1] * dr[2];
        stress[5] -= fs * dr[2] * dr[2];
    }
    double Epen = 0.0, fs = 1.0, rpen = 0.0;
    if (dx - P->pen_dist < P->rin[pair]) rpen = P->rin[pair] + P->pen_dist - dx;
    if (rpen > 0.0) {
        Epen = rpen * rpen * rpen * P->pen_scale;
        fs = -3.0 * rpen * rpen * P->pen_scale;
        g_penalty_count++;
    }
    if (Epen > 0.0) {
        *energy += Epen;
        fs /= dx;
        force[0] += fs * dr[0]; force[1] += fs * dr[1]; force[2] += fs * dr[2];
        force[3] -= fs * dr[0]; force[4] -= fs * dr[1]; force[5] -= fs * dr[2];
        stress[0] -= fs * dr[0] * dr[0];
        stress[1] -= fs * dr[0] * dr[1];
        stress[2] -= fs * dr[0] * dr[2];
        stress[3] -= fs * dr[1] * dr[1];
        stress[4] -= fs * dr[1] * dr[2];
        stress[5] -= fs * dr[2] * dr[2];
    }
}

/* Force and stress update of one pair inside a 3B/4B cluster; the stress uses
 * the distance tensor product dr_a*dr_b (chimesFF.cpp:1627-1634, 2100-2108). */
static inline void pair_update(double fs, const double *d, double *fa, double *fb, double *s)
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

/* compute_3B (chimesFF.cpp:1944-2176) */
void oc_compute_3b(const oc_params *P, const double dx[3], const double dr[9], const int t[3],
                   double force[9], double stress[6], double *energy)
{
    int nt = P->nt;
    int tix = t[0] * nt * nt + t[1] * nt + t[2];
    int trip = P->tripmap[tix];
    if (trip < 0) return;
    const int *m = P->trip_pm + tix * 3;
    double rc[3], ri[3];
    for (int p = 0; p < 3; p++) {
        rc[p] = P->cut3[trip][1][m[p]];
        ri[p] = P->cut3[trip][0][m[p]];
        if (dx[p] >= rc[p]) return;
    }
    int pt[3] = {P->pairmap[t[0] * nt + t[1]], P->pairmap[t[0] * nt + t[2]],
                 P->pairmap[t[1] * nt + t[2]]};
    int order = P->order[1];
    double T[3][64], Td[3][64], fc[3], fcd[3];
    for (int p = 0; p < 3; p++) oc_cheby(dx[p], P->lam[pt[p]], ri[p], rc[p], order, T[p], Td[p]);
    for (int p = 0; p < 3; p++) oc_fcut(P, dx[p], rc[p], &fc[p], &fcd[p]);
    double fall = fc[0] * fc[1] * fc[2];
    double f2[3] = {fc[1] * fc[2] / dx[0], fc[0] * fc[2] / dx[1], fc[0] * fc[1] / dx[2]};
    const int *pw = P->pow3[trip];
    const double *cf = P->c3[trip];
    for (int c = 0; c < P->nc3[trip]; c++) {
        double coeff = cf[c];
        int a0 = pw[c * 3 + m[0]], a1 = pw[c * 3 + m[1]], a2 = pw[c * 3 + m[2]];
        *energy += coeff * fall * T[0][a0] * T[1][a1] * T[2][a2];
        double d0 = fc[0] * Td[0][a0] + fcd[0] * T[0][a0];
        double d1 = fc[1] * Td[1][a1] + fcd[1] * T[1][a1];
        double d2 = fc[2] * Td[2][a2] + fcd[2] * T[2][a2];
        double fs0 = coeff * d0 * f2[0] * T[1][a1] * T[2][a2];
        double fs1 = coeff * d1 * f2[1] * T[0][a0] * T[2][a2];
        double fs2 = coeff * d2 * f2[2] * T[0][a0] * T[1][a1];
        pair_update(fs0, dr + 0, force + 0, force + 3, stress);
        pair_update(fs1, dr + 3, force + 0, force + 6, stress);
        pair_update(fs2, dr + 6, force + 3, force + 6, stress);
    }
}

/* compute_4B (chimesFF.cpp:2509-2872) */
void oc_compute_4b(const oc_params *P, const double dx[6], const double dr[18], const int t[4],
                   double force[12], double stress[6], double *energy)
{
    int nt = P->nt;
    int tix = ((t[0] * nt + t[1]) * nt + t[2]) * nt + t[3];
    int quad = P->quadmap[tix];
    if (quad < 0) return;
    const int *m = P->quad_pm + tix * 6;
    double rc[6], ri[6];
    for (int p = 0; p < 6; p++) {
        rc[p] = P->cut4[quad][1][m[p]];
        ri[p] = P->cut4[quad][0][m[p]];
        if (dx[p] >= rc[p]) return;
    }
    static const int pa[6] = {0, 0, 0, 1, 1, 2}, pb[6] = {1, 2, 3, 2, 3, 3};
    int pt[6];
    for (int p = 0; p < 6; p++) pt[p] = P->pairmap[t[pa[p]] * nt + t[pb[p]]];
    int order = P->order[2];
    double T[6][64], Td[6][64], fc[6], fcd[6];
    for (int p = 0; p < 6; p++) oc_cheby(dx[p], P->lam[pt[p]], ri[p], rc[p], order, T[p], Td[p]);
    for (int p = 0; p < 6; p++) oc_fcut(P, dx[p], rc[p], &fc[p], &fcd[p]);
    double fall = fc[0] * fc[1] * fc[2] * fc[3] * fc[4] * fc[5];
    double f5[6];
    f5[0] = fc[1] * fc[2] * fc[3] * fc[4] * fc[5] / dx[0];
    f5[1] = fc[0] * fc[2] * fc[3] * fc[4] * fc[5] / dx[1];
    f5[2] = fc[0] * fc[1] * fc[3] * fc[4] * fc[5] / dx[2];
    f5[3] = fc[0] * fc[1] * fc[2] * fc[4] * fc[5] / dx[3];
    f5[4] = fc[0] * fc[1] * fc[2] * fc[3] * fc[5] / dx[4];
    f5[5] = fc[0] * fc[1] * fc[2] * fc[3] * fc[4] / dx[5];
    const int *pw = P->pow4[quad];
    const double *cf = P->c4[quad];
    for (int c = 0; c < P->nc4[quad]; c++) {
        double coeff = cf[c];
        int a[6];
        for (int p = 0; p < 6; p++) a[p] = pw[c * 6 + m[p]];
        double t012 = T[0][a[0]] * T[1][a[1]] * T[2][a[2]];
        double t34 = T[3][a[3]] * T[4][a[4]];
        double t5 = T[5][a[5]];
        *energy += coeff * fall * t012 * t34 * t5;
        double d[6];
        for (int p = 0; p < 6; p++) d[p] = fc[p] * Td[p][a[p]] + fcd[p] * T[p][a[p]];
        double fs0 = coeff * d[0] * f5[0] * T[1][a[1]] * T[2][a[2]] * t34 * t5;
        double fs1 = coeff * d[1] * f5[1] * T[0][a[0]] * T[2][a[2]] * t34 * t5;
        double fs2 = coeff * d[2] * f5[2] * T[0][a[0]] * T[1][a[1]] * t34 * t5;
        double fs3 = coeff * d[3] * f5[3] * t012 * T[4][a[4]] * t5;
        double fs4 = coeff * d[4] * f5[4] * t012 * T[3][a[3]] * t5;
        double fs5 = coeff * d[5] * f5[5] * t012 * t34;
        pair_update(fs0, dr + 0, force + 0, force + 3, stress);
        pair_update(fs1, dr + 3, force + 0, force + 6, stress);
        pair_update(fs2, dr + 6, force + 0, force + 9, stress);
        pair_update(fs3, dr + 9, force + 3, force + 6, stress);
        pair_update(fs4, dr + 12, force + 3, force + 9, stress);
        pair_update(fs5, dr + 15, force + 6, force + 9, stress);
    }
}

/* ------------------------------------------------------------------------ */
/* simulation_system (serial_chimes_interface.{h,cpp})                       */
/* ------------------------------------------------------------------------ */

typedef struct {
    int n_atoms, n_ghost, cap;
    double *x, *y, *z;
    int *parent, *rep_parent, *type;
    double h[9], ih[9];
    double la, lb, lc, alpha, beta, gamma, vol;
    int n_rep, n_layers;
} sysm;

static void sys_free(sysm *S)
{
    free(S->x); free(S->y); free(S->z); free(S->parent); free(S->rep_parent); free(S->type);
}

static void sys_push(sysm *S, double x, double y, double z, int parent, int rep, int type)
{
    if (S->n_ghost == S->cap) {
        S->cap = S->cap ? 2 * S->cap : 1024;
        S->x = (double *)realloc(S->x, sizeof(double) * S->cap);
        S->y = (double *)realloc(S->y, sizeof(double) * S->cap);
        S->z = (double *)realloc(S->z, sizeof(double) * S->cap);
        S->parent = (int *)realloc(S->parent, sizeof(int) * S->cap);
        S->rep_parent = (int *)realloc(S->rep_parent, sizeof(int) * S->cap);
        S->type = (int *)realloc(S->type, sizeof(int) * S->cap);
    }
    int g = S->n_ghost++;
    S->x[g] = x; S->y[g] = y; S->z[g] = z;
    S->parent[g] = parent; S->rep_parent[g] = rep; S->type[g] = type;
}

/* mag_a, a_dot_b, angle_ab (serial_chimes_interface.cpp:20-65) */
static double mag3(double a, double b, double c)
{
    double m = 0;
    m += a * a; m += b * b; m += c * c;
    return sqrt(m);
}
static double dot3(const double *a, const double *b)
{
    double d = 0;
    d += a[0] * b[0]; d += a[1] * b[1]; d += a[2] * b[2];
    return d;
}
static double angle3(const double *a, const double *b)
{
    double ang = dot3(a, b);
    ang /= mag3(a[0], a[1], a[2]);
    ang /= mag3(b[0], b[1], b[2]);
    return acos(ang);
}

/* set_hmat (serial_chimes_interface.cpp:85-113) */
static void set_hmat(const double *ca, const double *cb, const double *cc, double *h, double *ih,
                     int rep)
{
    h[0] = ca[0] * (rep + 1); h[3] = ca[1] * (rep + 1); h[6] = ca[2] * (rep + 1);
    h[1] = cb[0] * (rep + 1); h[4] = cb[1] * (rep + 1); h[7] = cb[2] * (rep + 1);
    h[2] = cc[0] * (rep + 1); h[5] = cc[1] * (rep + 1); h[8] = cc[2] * (rep + 1);
    double det = h[0] * (h[4] * h[8] - h[5] * h[7]) - h[1] * (h[3] * h[8] - h[5] * h[6]) +
                 h[2] * (h[3] * h[7] - h[4] * h[6]);
    double t[9];
    t[0] = (h[4] * h[8] - h[5] * h[7]);
    t[3] = -1 * (h[1] * h[8] - h[2] * h[7]);
    t[6] = (h[1] * h[5] - h[2] * h[4]);
    t[1] = -1 * (h[3] * h[8] - h[5] * h[6]);
    t[4] = (h[0] * h[8] - h[2] * h[6]);
    t[7] = -1 * (h[0] * h[5] - h[2] * h[3]);
    t[2] = (h[3] * h[7] - h[4] * h[6]);
    t[5] = -1 * (h[0] * h[7] - h[1] * h[6]);
    t[8] = (h[0] * h[4] - h[1] * h[3]);
    ih[0] = t[0]; ih[3] = t[1]; ih[6] = t[2];
    ih[1] = t[3]; ih[4] = t[4]; ih[7] = t[5];
    ih[2] = t[6]; ih[5] = t[7]; ih[8] = t[8];
    for (int k = 0; k < 9; k++) ih[k] /= det;
}

static inline void mat_vec(const double *m, double x, double y, double z, double *o)
{
    o[0] = m[0] * x + m[1] * y + m[2] * z;
    o[1] = m[3] * x + m[4] * y + m[5] * z;
    o[2] = m[6] * x + m[7] * y + m[8] * z;
}

static void cell_metrics(sysm *S)
{
    const double *h = S->h;
    S->la = mag3(h[0], h[3], h[6]);
    S->lb = mag3(h[1], h[4], h[7]);
    S->lc = mag3(h[2], h[5], h[8]);
    double A[3] = {h[0], h[3], h[6]}, B[3] = {h[1], h[4], h[7]}, C[3] = {h[2], h[5], h[8]};
    S->alpha = angle3(B, C);
    S->beta = angle3(C, A);
    S->gamma = angle3(A, B);
}

/* simulation_system::init (serial_chimes_interface.cpp:150-341) */
static int sys_init(sysm *S, int natom, const double *x, const double *y, const double *z,
                    const int *typ, const double *ca, const double *cb, const double *cc,
                    double rc2, int small)
{
    memset(S, 0, sizeof(*S));
    for (int a = 0; a < natom; a++) sys_push(S, x[a], y[a], z[a], a, a, typ[a]);
    double la = mag3(ca[0], ca[1], ca[2]), lb = mag3(cb[0], cb[1], cb[2]),
           lc = mag3(cc[0], cc[1], cc[2]);
    double mn = la;
    if (lb < mn) mn = lb;
    if (lc < mn) mn = lc;
    S->n_rep = 0;
    if (small) S->n_rep = (int)(ceil(rc2 / mn) - 1);
    set_hmat(ca, cb, cc, S->h, S->ih, 0);
    int nrepl = natom;
    for (int i = 0; i <= S->n_rep; i++)
        for (int j = 0; j <= S->n_rep; j++)
            for (int k = 0; k <= S->n_rep; k++) {
                if (i == 0 && j == 0 && k == 0) continue;
                for (int a = 0; a < natom; a++) {
                    double t[3];
                    mat_vec(S->ih, S->x[a], S->y[a], S->z[a], t);
                    t[0] += i; t[1] += j; t[2] += k;
                    double o[3];
                    mat_vec(S->h, t[0], t[1], t[2], o);
                    sys_push(S, o[0], o[1], o[2], nrepl, a, typ[a]);
                    nrepl++;
                }
            }
    S->n_atoms = nrepl;
    set_hmat(ca, cb, cc, S->h, S->ih, S->n_rep);
    for (int i = 0; i < S->n_atoms; i++) {
        double t[3];
        mat_vec(S->ih, S->x[i], S->y[i], S->z[i], t);
        t[0] -= floor(t[0]); t[1] -= floor(t[1]); t[2] -= floor(t[2]);
        double o[3];
        mat_vec(S->h, t[0], t[1], t[2], o);
        S->x[i] = o[0]; S->y[i] = o[1]; S->z[i] = o[2];
    }
    cell_metrics(S);
    double v = 1;
    v += 2 * cos(S->alpha) * cos(S->beta) * cos(S->gamma);
    v -= cos(S->alpha) * cos(S->alpha);
    v -= cos(S->beta) * cos(S->beta);
    v -= cos(S->gamma) * cos(S->gamma);
    S->vol = S->la * S->lb * S->lc * sqrt(v);
    return 0;
}

/* simulation_system::reorient (serial_chimes_interface.cpp:342-476) */
static void sys_reorient(sysm *S)
{
    for (int i = 0; i < S->n_atoms; i++) {
        double t[3];
        mat_vec(S->ih, S->x[i], S->y[i], S->z[i], t);
        S->x[i] = t[0]; S->y[i] = t[1]; S->z[i] = t[2];
    }
    double A[3] = {S->h[0], S->h[3], S->h[6]}, B[3] = {S->h[1], S->h[4], S->h[7]},
           C[3] = {S->h[2], S->h[5], S->h[8]};
    double ma = mag3(A[0], A[1], A[2]);
    double u[3] = {0, 0, 0};
    u[0] += A[0] / ma; u[1] += A[1] / ma; u[2] += A[2] / ma;
    double cr[3];
    cr[0] = (u[1] * B[2] - u[2] * B[1]);
    cr[1] = -1 * (u[0] * B[2] - u[2] * B[0]);
    cr[2] = (u[0] * B[1] - u[1] * B[0]);
    double na[3], nb[3], nc[3];
    na[0] = ma; na[1] = 0; na[2] = 0;
    nb[0] = dot3(B, u); nb[1] = mag3(cr[0], cr[1], cr[2]); nb[2] = 0;
    nc[0] = dot3(C, u);
    nc[1] = (dot3(B, C) - nb[0] * nc[0]) / nb[1];
    double mc = mag3(C[0], C[1], C[2]);
    nc[2] = sqrt(mc * mc - nc[0] * nc[0] - nc[1] * nc[1]);
    set_hmat(na, nb, nc, S->h, S->ih, 0);
    for (int i = 0; i < S->n_atoms; i++) {
        double o[3];
        mat_vec(S->h, S->x[i], S->y[i], S->z[i], o);
        S->x[i] = o[0]; S->y[i] = o[1]; S->z[i] = o[2];
    }
    cell_metrics(S);
}

/* simulation_system::build_layered_system (serial_chimes_interface.cpp:477-587) */
static int sys_layers(sysm *S, const oc_params *P, double rc2, double rc3, double rc4)
{
    double lmin = S->la;
    if (S->lb < lmin) lmin = S->lb;
    if (S->lc < lmin) lmin = S->lc;
    S->n_layers = (int)ceil(rc2 * 2.0 / lmin + 1);
    int nl = S->n_layers;
    double ex = S->la * (2 * nl + 1), ey = S->lb * (2 * nl + 1), ez = S->lc * (2 * nl + 1);
    if (rc2 > 0.5 * ex || rc2 > 0.5 * ey || rc2 > 0.5 * ez) return -1;
    if (P->order[1] > 0 && (rc3 > 0.5 * ex || rc3 > 0.5 * ey || rc3 > 0.5 * ez)) return -1;
    if (P->order[2] > 0 && (rc4 > 0.5 * ex || rc4 > 0.5 * ey || rc4 > 0.5 * ez)) return -1;
    S->n_ghost = S->n_atoms;
    for (int i = -nl; i <= nl; i++)
        for (int j = -nl; j <= nl; j++)
            for (int k = -nl; k <= nl; k++) {
                if (i == 0 && j == 0 && k == 0) continue;
                for (int a = 0; a < S->n_atoms; a++) {
                    double t[3];
                    mat_vec(S->ih, S->x[a], S->y[a], S->z[a], t);
                    t[0] += i; t[1] += j; t[2] += k;
                    double o[3];
                    mat_vec(S->h, t[0], t[1], t[2], o);
                    sys_push(S, o[0], o[1], o[2], a, S->rep_parent[a], S->type[a]);
                }
            }
    return 0;
}

/* get_dist (serial_chimes_interface.h:141-208) */
static inline double get_dist(const sysm *S, int i, int j, double *r)
{
    const double *ih = S->ih, *h = S->h;
    double ix = ih[0] * S->x[i] + ih[1] * S->y[i] + ih[2] * S->z[i];
    double iy = ih[3] * S->x[i] + ih[4] * S->y[i] + ih[5] * S->z[i];
    double iz = ih[6] * S->x[i] + ih[7] * S->y[i] + ih[8] * S->z[i];
    double jx = ih[0] * S->x[j] + ih[1] * S->y[j] + ih[2] * S->z[j];
    double jy = ih[3] * S->x[j] + ih[4] * S->y[j] + ih[5] * S->z[j];
    double jz = ih[6] * S->x[j] + ih[7] * S->y[j] + ih[8] * S->z[j];
    double dx = jx - ix, dy = jy - iy, dz = jz - iz;
    r[0] = h[0] * dx + h[1] * dy + h[2] * dz;
    r[1] = h[3] * dx + h[4] * dy + h[5] * dz;
    r[2] = h[6] * dx + h[7] * dy + h[8] * dz;
    return sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
}

typedef struct {
    int n;
    int *start, *cnt, *idx;  /* CSR of the 2B neighbour list over real atoms */
    long n3, cap3;
    int *l3;
    long n4, cap4;
    int *l4;
} nlists;

static void push_int(int **arr, long *n, long *cap, const int *v, int k)
{
    if (*n == *cap) {
        *cap = *cap ? 2 * *cap : 4096;
        *arr = (int *)realloc(*arr, sizeof(int) * k * (*cap));
    }
    memcpy(*arr + (*n) * k, v, sizeof(int) * k);
    (*n)++;
}

/* simulation_system::build_neigh_lists (serial_chimes_interface.cpp:588-864) */
static int build_lists(const sysm *S, const oc_params *P, double rc2, double rc3, double rc4,
                       nlists *NL)
{
    memset(NL, 0, sizeof(*NL));
    double mx[3] = {-1.0e100, -1.0e100, -1.0e100}, mn[3] = {1.0e100, 1.0e100, 1.0e100};
    for (int i = 0; i < S->n_ghost; i++) {
        if (S->x[i] > mx[0]) mx[0] = S->x[i];
        if (S->y[i] > mx[1]) mx[1] = S->y[i];
        if (S->z[i] > mx[2]) mx[2] = S->z[i];
        if (S->x[i] < mn[0]) mn[0] = S->x[i];
        if (S->y[i] < mn[1]) mn[1] = S->y[i];
        if (S->z[i] < mn[2]) mn[2] = S->z[i];
    }
    double sd = rc2;
    if (rc3 > sd) sd = rc3;
    if (rc4 > sd) sd = rc4;
    int nbx = (int)ceil((mx[0] - mn[0]) / sd);
    int nby = (int)ceil((mx[1] - mn[1]) / sd);
    int nbz = (int)ceil((mx[2] - mn[2]) / sd);
    long total = (long)nbx * nby * nbz;
    if (nbx <= 0 || nby <= 0 || nbz <= 0 || total <= 0) return -1;
    int *bcnt = (int *)calloc(total + 1, sizeof(int));
    int *bstart = (int *)calloc(total + 1, sizeof(int));
    int *bin_of = (int *)malloc(sizeof(int) * S->n_ghost);
    for (int i = 0; i < S->n_ghost; i++) {
        int bx = (int)floor((S->x[i] - mn[0]) / sd);
        int by = (int)floor((S->y[i] - mn[1]) / sd);
        int bz = (int)floor((S->z[i] - mn[2]) / sd);
        if (bx < 0 || by < 0 || bz < 0) goto fail;
        long ib = bx + (long)by * nbx + (long)bz * nbx * nby;
        if (ib >= total) goto fail;
        bin_of[i] = (int)ib;
        bcnt[ib]++;
    }
    for (long b = 0; b < total; b++) bstart[b + 1] = bstart[b] + bcnt[b];
    int *bins = (int *)malloc(sizeof(int) * (S->n_ghost + 1));
    {
        int *fill = (int *)calloc(total, sizeof(int));
        for (int i = 0; i < S->n_ghost; i++) {
            int b = bin_of[i];
            bins[bstart[b] + fill[b]++] = i;
        }
        free(fill);
    }
    NL->n = S->n_atoms;
    NL->start = (int *)malloc(sizeof(int) * (S->n_atoms + 1));
    NL->cnt = (int *)calloc(S->n_atoms, sizeof(int));
    long cap = 0, used = 0;
    NL->idx = NULL;
    for (int ai = 0; ai < S->n_atoms; ai++) {
        int bx = (int)floor((S->x[ai] - mn[0]) / sd);
        int by = (int)floor((S->y[ai] - mn[1]) / sd);
        int bz = (int)floor((S->z[ai] - mn[2]) / sd);
        if (bx < 1 || by < 1 || bz < 1) { free(bins); goto fail; }
        NL->start[ai] = (int)used;
        for (int i = bx - 1; i <= bx + 1; i++)
            for (int j = by - 1; j <= by + 1; j++)
                for (int k = bz - 1; k <= bz + 1; k++) {
                    long ib = i + (long)j * nbx + (long)k * nbx * nby;
                    if (ib >= total) { free(bins); goto fail; }
                    for (int q = bstart[ib]; q < bstart[ib + 1]; q++) {
                        int g = bins[q];
                        if (g == ai) continue;
                        double r[3];
                        if (ai <= S->parent[g] && get_dist(S, ai, g, r) < sd) {
                            if (used == cap) {
                                cap = cap ? 2 * cap : 4096;
                                NL->idx = (int *)realloc(NL->idx, sizeof(int) * cap);
                            }
                            NL->idx[used++] = g;
                        }
                    }
                }
        NL->cnt[ai] = (int)(used - NL->start[ai]);
    }
    NL->start[S->n_atoms] = (int)used;
    free(bins);
    free(bcnt); free(bstart); free(bin_of);
    if (P->order[1] == 0 && P->order[2] == 0) return 0;

    for (int i = 0; i < S->n_atoms; i++) {
        const int *nb = NL->idx + NL->start[i];
        int n = NL->cnt[i];
        for (int j = 0; j < n; j++) {
            int v3 = 1, v4 = 1;
            int jj = nb[j];
            double r[3];
            double dij = get_dist(S, i, jj, r);
            if (dij >= rc3) v3 = 0;
            if (dij >= rc4) v4 = 0;
            if (!v3 && !v4) continue;
            for (int k = 0; k < n; k++) {
                int kk = nb[k];
                if (jj == kk) continue;
                if (S->parent[jj] > S->parent[kk]) continue;
                double dik = get_dist(S, i, kk, r);
                if (dik >= rc3) v3 = 0;
                if (dik >= rc4) v4 = 0;
                double djk = get_dist(S, jj, kk, r);
                if (djk >= rc3) v3 = 0;
                if (djk >= rc4) v4 = 0;
                if (!v3 && !v4) {
                    if (dij < rc3) v3 = 1;
                    if (dij < rc4) v4 = 1;
                    continue;
                }
                int t3[3] = {i, jj, kk};
                if (v3) push_int(&NL->l3, &NL->n3, &NL->cap3, t3, 3);
                if (P->order[2] == 0) continue;
                if (!v4) {
                    if (dij < rc4) v4 = 1;
                    continue;
                }
                for (int l = 0; l < n; l++) {
                    int ll = nb[l];
                    if (jj == ll || kk == ll) continue;
                    if (S->parent[jj] > S->parent[ll]) continue;
                    if (S->parent[kk] > S->parent[ll]) continue;
                    if (get_dist(S, i, ll, r) >= rc4) continue;
                    if (get_dist(S, jj, ll, r) >= rc4) continue;
                    if (get_dist(S, kk, ll, r) >= rc4) continue;
                    int t4[4] = {i, jj, kk, ll};
                    if (v4) push_int(&NL->l4, &NL->n4, &NL->cap4, t4, 4);
                }
            }
        }
    }
    return 0;
fail:
    free(bcnt); free(bstart); free(bin_of);
    return -1;
}

static void nl_free(nlists *NL)
{
    free(NL->start); free(NL->cnt); free(NL->idx); free(NL->l3); free(NL->l4);
}

/* Layered index -> canonical key (parent, shift). Ghost g >= n_atoms was
 * created for shift ordinal c = (g - n_atoms) / n_atoms of the lexicographic
 * (i, j, k) in [-n, n]^3 without the origin (serial_chimes_interface.cpp:546-585). */
static void key_of(const sysm *S, int g, int *out)
{
    if (g < S->n_atoms) { out[0] = g; out[1] = out[2] = out[3] = 0; return; }
    int gp = g - S->n_atoms;
    int c = gp / S->n_atoms;
    int a = gp % S->n_atoms;
    int w = 2 * S->n_layers + 1;
    int mid = (w * w * w - 1) / 2;
    int t = c < mid ? c : c + 1;
    out[0] = a;
    out[1] = t / (w * w) - S->n_layers;
    out[2] = (t / w) % w - S->n_layers;
    out[3] = t % w - S->n_layers;
}

static int g_exact_sums = 0;

/* Compensated accumulation for oc_calculate (not a reference behaviour): the
   reference adds every term straight into running double totals
   (chimesFF.cpp:1736-1741, 2103-2168), whose rounding grows with the number of
   terms; with this mode on, the energy and stress terms of each cluster are
   summed from zero and the per-cluster sums added in long double, so the
   totals carry the rounding of the per-term products only. */
void oc_set_exact_sums(int on) { g_exact_sums = on != 0; }

/* serial_chimes_interface::calculate (serial_chimes_interface.cpp:960-1124) */
int oc_calculate(const oc_params *P, int natom, const double *x, const double *y, const double *z,
                 const char *const *types, const double ca[3], const double cb[3],
                 const double cc[3], int small, double *energy, double *fx, double *fy,
                 double *fz, double stress[9], oc_lists *lists, int owner_stride)
{
    g_penalty_count = 0;
    g_oor_count = 0;
    if (owner_stride < 1) owner_stride = 1;
    double rc2 = oc_max_cutoff(P, 2), rc3 = oc_max_cutoff(P, 3), rc4 = oc_max_cutoff(P, 4);
    int *typ = (int *)malloc(sizeof(int) * (natom > 0 ? natom : 1));
    for (int a = 0; a < natom; a++) {
        typ[a] = -1;
        for (int t = 0; t < P->nt; t++)
            if (!strcmp(types[a], P->atm[t])) { typ[a] = t; break; }
        if (typ[a] < 0) { free(typ); return -2; }
    }
    sysm S, N;
    sys_init(&S, natom, x, y, z, typ, ca, cb, cc, rc2, small);
    if (sys_layers(&S, P, rc2, rc3, rc4)) { sys_free(&S); free(typ); return -3; }
    /* run_checks (serial_chimes_interface.cpp:866-905) */
    {
        double mc[3] = {rc2, rc3, rc4};
        int nl = S.n_layers;
        for (int i = 0; i < 3; i++)
            if (mc[i] > 2 * S.la * (2 * nl + 1) || mc[i] > 2 * S.lb * (2 * nl + 1) ||
                mc[i] > 2 * S.lc * (2 * nl + 1)) { sys_free(&S); free(typ); return -4; }
        int bod = 2 + (P->order[1] > 0) + (P->order[2] > 0);
        if (bod > S.n_ghost) { sys_free(&S); free(typ); return -4; }
    }
    sys_init(&N, natom, x, y, z, typ, ca, cb, cc, rc2, small);
    free(typ);
    sys_reorient(&N);
    if (sys_layers(&N, P, rc2, rc3, rc4)) { sys_free(&S); sys_free(&N); return -3; }
    nlists NL;
    if (build_lists(&N, P, rc2, rc3, rc4, &NL)) { sys_free(&S); sys_free(&N); return -5; }

    double sc[6] = {0, 0, 0, 0, 0, 0};
    double E = *energy;
    double f2[6], f3[9], f4[12], r[3];
    /* compensated mode: each cluster's energy and stress terms are summed
       from zero in double, then added into long double totals */
    long double EL = 0.0L, SL[6] = {0, 0, 0, 0, 0, 0};
    double Ec, scc[6];
#define OC_BEGIN() double *sc_ = g_exact_sums ? scc : sc, *E_ = g_exact_sums ? &Ec : &E; \
    if (g_exact_sums) { Ec = 0.0; memset(scc, 0, sizeof scc); }
#define OC_END() if (g_exact_sums) { EL += Ec; for (int k_ = 0; k_ < 6; k_++) SL[k_] += scc[k_]; }
    for (int i = 0; i < S.n_atoms; i++) {
        E += P->eoff[S.type[i]];
        if (i % owner_stride) continue;
        for (int q = NL.start[i]; q < NL.start[i] + NL.cnt[i]; q++) {
            int jj = NL.idx[q];
            double d = get_dist(&S, i, jj, r);
            int t[2] = {S.type[i], S.type[jj]};
            memset(f2, 0, sizeof f2);
            OC_BEGIN();
            oc_compute_2b(P, d, r, t, f2, sc_, E_);
            OC_END();
            int a = S.rep_parent[i], b = S.rep_parent[S.parent[jj]];
            fx[a] += f2[0]; fy[a] += f2[1]; fz[a] += f2[2];
            fx[b] += f2[3]; fy[b] += f2[4]; fz[b] += f2[5];
        }
    }
    if (P->order[1] > 0)
        for (long c = 0; c < NL.n3; c++) {
            const int *m = NL.l3 + 3 * c;
            if (m[0] % owner_stride) continue;
            double dx[3], dr[9];
            dx[0] = get_dist(&S, m[0], m[1], dr + 0);
            dx[1] = get_dist(&S, m[0], m[2], dr + 3);
            dx[2] = get_dist(&S, m[1], m[2], dr + 6);
            int t[3] = {S.type[m[0]], S.type[m[1]], S.type[m[2]]};
            memset(f3, 0, sizeof f3);
            OC_BEGIN();
            oc_compute_3b(P, dx, dr, t, f3, sc_, E_);
            OC_END();
            for (int a = 0; a < 3; a++) {
                int o = S.rep_parent[S.parent[m[a]]];
                fx[o] += f3[3 * a]; fy[o] += f3[3 * a + 1]; fz[o] += f3[3 * a + 2];
            }
        }
    if (P->order[2] > 0)
        for (long c = 0; c < NL.n4; c++) {
            const int *m = NL.l4 + 4 * c;
            if (m[0] % owner_stride) continue;
            static const int pa[6] = {0, 0, 0, 1, 1, 2}, pb[6] = {1, 2, 3, 2, 3, 3};
            double dx[6], dr[18];
            for (int p = 0; p < 6; p++) dx[p] = get_dist(&S, m[pa[p]], m[pb[p]], dr + 3 * p);
            int t[4] = {S.type[m[0]], S.type[m[1]], S.type[m[2]], S.type[m[3]]};
            memset(f4, 0, sizeof f4);
            OC_BEGIN();
            oc_compute_4b(P, dx, dr, t, f4, sc_, E_);
            OC_END();
            for (int a = 0; a < 4; a++) {
                int o = S.rep_parent[S.parent[m[a]]];
                fx[o] += f4[3 * a]; fy[o] += f4[3 * a + 1]; fz[o] += f4[3 * a + 2];
            }
        }
#undef OC_BEGIN
#undef OC_END
    if (g_exact_sums) {
        E = (double)((long double)E + EL);
        for (int k = 0; k < 6; k++) sc[k] = (double)SL[k];
    }
    E /= pow(S.n_rep + 1.0, 3.0);
    *energy = E;
    stress[0] = sc[0]; stress[1] = sc[1]; stress[2] = sc[2];
    stress[3] = sc[1]; stress[4] = sc[3]; stress[5] = sc[4];
    stress[6] = sc[2]; stress[7] = sc[4]; stress[8] = sc[5];
    for (int k = 0; k < 9; k++) stress[k] /= S.vol;

    if (lists) {
        memset(lists, 0, sizeof(*lists));
        lists->n_atoms = N.n_atoms;
        lists->n_layers = N.n_layers;
        lists->n_rep = N.n_rep;
        long np = NL.start[N.n_atoms];
        lists->n2 = np;
        lists->pairs = (int *)malloc(sizeof(int) * 5 * (np > 0 ? np : 1));
        long w = 0;
        for (int i = 0; i < N.n_atoms; i++)
            for (int q = NL.start[i]; q < NL.start[i] + NL.cnt[i]; q++) {
                lists->pairs[5 * w] = i;
                key_of(&N, NL.idx[q], lists->pairs + 5 * w + 1);
                w++;
            }
        lists->n3 = NL.n3;
        lists->trips = (int *)malloc(sizeof(int) * 9 * (NL.n3 > 0 ? NL.n3 : 1));
        for (long c = 0; c < NL.n3; c++) {
            int *o = lists->trips + 9 * c;
            o[0] = NL.l3[3 * c];
            key_of(&N, NL.l3[3 * c + 1], o + 1);
            key_of(&N, NL.l3[3 * c + 2], o + 5);
        }
        lists->n4 = NL.n4;
        lists->quads = (int *)malloc(sizeof(int) * 13 * (NL.n4 > 0 ? NL.n4 : 1));
        for (long c = 0; c < NL.n4; c++) {
            int *o = lists->quads + 13 * c;
            o[0] = NL.l4[4 * c];
            key_of(&N, NL.l4[4 * c + 1], o + 1);
            key_of(&N, NL.l4[4 * c + 2], o + 5);
            key_of(&N, NL.l4[4 * c + 3], o + 9);
        }
    }
    nl_free(&NL);
    sys_free(&S);
    sys_free(&N);
    return 0;
}

void oc_lists_free(oc_lists *L)
{
    if (!L) return;
    free(L->pairs); free(L->trips); free(L->quads);
    L->pairs = L->trips = L->quads = NULL;
}
