/*
 * ORACLE -- test infrastructure only.
 *
 * A plain-C, CPU, FP64 restatement of the ChIMES serial evaluation path of
 * rk-lindsey/chimes_calculator (chimesFF + serial_chimes_interface).  It is the
 * checker for the HIP product and the CPU baseline timed by bench.py.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it;
 * the product library never links or calls it.
 *
 * Every function cites the reference file:line it follows.  Parity pins for
 * each function are listed at the top of chimes_oracle.c.
 */
#ifndef CHIMES_ORACLE_H
#define CHIMES_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oc_params oc_params;

/* Cluster lists in canonical key form.  A member of a cluster is the pair
 * (parent atom index in the (replicated) system, image shift sx, sy, sz), the
 * shift being the ghost-layer offset in cell vectors of the neighbour frame.
 *   pairs   : 5 ints per entry  (i, a, sx, sy, sz)
 *   triplets: 9 ints per entry  (i, a_j, s_j[3], a_k, s_k[3])
 *   quads   : 13 ints per entry (i, a_j, s_j[3], a_k, s_k[3], a_l, s_l[3])
 * Entries are in the order the reference's loops generate them. */
typedef struct {
    int n_atoms;     /* real atoms after replication */
    int n_layers;    /* ghost layers of the neighbour frame */
    int n_rep;       /* replication count (0 = none) */
    long n2, n3, n4;
    int *pairs, *trips, *quads;
} oc_lists;

/* Parse a ChIMES parameter file (chimesFF.cpp:354-1570) and build the
 * triplet/quadruplet pair maps (chimesFF.cpp:2958-3080).  NULL on error. */
oc_params *oc_load(const char *path);
void oc_free(oc_params *P);

int    oc_order(const oc_params *P, int body);          /* body = 2, 3, 4 */
double oc_max_cutoff(const oc_params *P, int body);     /* chimesFF.cpp:2903-2943 */
int    oc_ntypes(const oc_params *P);
const char *oc_type_name(const oc_params *P, int t);

/* Full system evaluation (serial_chimes_interface.cpp:960-1124).
 * energy, fx/fy/fz are accumulated onto (like the reference); stress[9] is
 * overwritten.  If owner_stride > 1 only clusters whose first (owner) atom i
 * satisfies i % owner_stride == 0 are evaluated (a work sample for timing).
 * Returns 0 on success, nonzero on an error the reference would exit() on. */
int oc_calculate(const oc_params *P, int natom, const double *x, const double *y,
                 const double *z, const char *const *types, const double ca[3],
                 const double cb[3], const double cc[3], int small,
                 double *energy, double *fx, double *fy, double *fz,
                 double stress[9], oc_lists *lists, int owner_stride);

void oc_lists_free(oc_lists *L);

/* 1-body energy offset of type t, added onto *energy (chimesFF.cpp:1636-1641). */
void oc_compute_1b(const oc_params *P, int t, double *energy);
/* Mass of type t as read from the parameter file (chimesFF.cpp:354-1570). */
double oc_mass(const oc_params *P, int t);
/* Outer 2-body cutoff of a type pair (chimesFF.cpp:2886-2901). */
double oc_cutoff_2b(const oc_params *P, int ti, int tj);

/* Single-cluster evaluators (chimesFF.cpp:1657-1775, 1944-2176, 2509-2872).
 * Type arguments are type indices.  force / stress are packed as in the
 * reference (force[natoms*3], stress[6]) and are accumulated onto. */
void oc_compute_2b(const oc_params *P, double dx, const double dr[3], const int t[2],
                   double force[6], double stress[6], double *energy);
void oc_compute_3b(const oc_params *P, const double dx[3], const double dr[9],
                   const int t[3], double force[9], double stress[6], double *energy);
void oc_compute_4b(const oc_params *P, const double dx[6], const double dr[18],
                   const int t[4], double force[12], double stress[6], double *energy);

/* Inner/outer cutoff of runtime pair p of a cluster of types t (nonzero for
 * an excluded cluster type) and the Morse lambda of a type pair. */
int oc_cluster_pair_cutoffs(const oc_params *P, int nb, const int *t, int p, double *rin, double *rout);
double oc_morse_lambda(const oc_params *P, int ti, int tj);

/* Cluster fingerprint histograms (chimes_fp_oracle.c, calc_histograms.cpp):
 * dumped rows (nb(nb-1)/2 distances, nb types) -> sorted transformed edges,
 * and the histogram of cluster-cluster distances (same != 0: pairs j > i of
 * one set).  Nonzero on an excluded cluster type or a bin out of range. */
int oc_fp_edges(const oc_params *P, int nb, long n, const double *dist, const int *types, double *edges);
int oc_fp_hist(int np, long n1, const double *e1, long n2, const double *e2, int same, int nbin, double binw,
               double maxd, long long *counts, long long *nsamples);

/* LAMMPS pair style on one process (chimes_lmp_oracle.c): many-body lists
 * from a full neighbour list with ghosts (pair_chimes.cpp:321-478; trips as
 * (i, j, k), quads as (i, j, k, l), free with oc_lmp_free) and the compute
 * loops with ev_tally_mb (pair_chimes.cpp:480-831, pair.cpp:1425-1482). */
int oc_lmp_build_lists(const oc_params *P, const double *x, const int *tag, int inum, const int *ilist,
                       const int *numneigh, const int *nbr_off, const int *nbr, double skin, long *n3, int **trips,
                       long *n4, int **quads);
void oc_lmp_compute(const oc_params *P, const double *x, const int *ctype, const int *tag, int inum,
                    const int *ilist, const int *numneigh, const int *nbr_off, const int *nbr, long n3,
                    const int *trips, long n4, const int *quads, int eflag_global, int eflag_atom, int vflag_global,
                    int vflag_atom, double *f, double *eng, double *virial, double *eatom, double *vatom);
long oc_lmp_fp_rows(const oc_params *P, int nb, const double *x, const int *ctype, const int *tag, int inum,
                    const int *ilist, const int *numneigh, const int *nbr_off, const int *nbr, long n3,
                    const int *trips, long n4, const int *quads, double **rows);
void oc_lmp_free(void *p);

/* Chebyshev polynomials of the Morse-transformed distance and the cutoff
 * function, exposed for unit tests (chimesFF.h:347-386, 542-628,
 * chimesFF.cpp:1572-1613). */
void oc_cheby(double dx, double lambda, double rin, double rout, int order,
              double *Tn, double *Tnd);
void oc_fcut(const oc_params *P, double dx, double rout, double *fcut, double *fcutd);

/* Compensated energy and stress sums in oc_calculate (per-cluster sums added
   in long double); off by default, when the sums follow the reference's order. */
void oc_set_exact_sums(int on);

/* Penalty / out-of-range event counters of the last oc_calculate call. */
long oc_last_penalty_count(void);
long oc_last_out_of_range_count(void);

int oc_npairs(const oc_params *P);
int oc_pair_type(const oc_params *P, int ti, int tj);
int oc_trip_type(const oc_params *P, const int t[3]);

/* Tabulated 2-/3-body interactions of the -DTABULATION build
 * (chimes_tab_oracle.c).  oc_tab_load reads the tables a parameter file
 * names (NULL on error, message in oc_tab_error()); oc_tab_bodies: bit 0 the
 * 2-body, bit 1 the 3-body interactions are tabulated.  The compute functions
 * return nonzero where the reference exits or reads past its table. */
typedef struct oc_tables oc_tables;
oc_tables *oc_tab_load(const oc_params *P, const char *param_path);
void oc_tab_free(oc_tables *T);
const char *oc_tab_error(void);
int oc_tab_bodies(const oc_tables *T);
int oc_tab_3b_count(const oc_tables *T);
double oc_tab_2b_value(const oc_tables *T, int pair, double rij, int for_energy, int *err);
void oc_tab_3b_interp(const oc_tables *T, int trip, double rij, double rik, double rjk, double out[4], int *err);
int oc_compute_2b_tab(const oc_params *P, const oc_tables *T, double dx, const double dr[3], const int t[2],
                      double force[6], double stress[6], double *energy);
int oc_compute_3b_tab(const oc_params *P, const oc_tables *T, const double dx[3], const double dr[9],
                      const int t[3], double force[9], double stress[6], double *energy);

/* The pair style's loops (oc_lmp_compute) with the tabulated interactions
 * of T (NULL: the polynomial ones), as pair_chimes.cpp:620-623, 697-700 of a
 * -DTABULATION build; oc_lmp_tab_errors counts calls that returned an error
 * since the last oc_lmp_use_tables. */
void oc_lmp_use_tables(const oc_tables *T);
long oc_lmp_tab_errors(void);

#ifdef __cplusplus
}
#endif
#endif
