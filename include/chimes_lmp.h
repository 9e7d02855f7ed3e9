/*
 * The LAMMPS pair style chimesFF (etc/lmp/src/pair_chimes.cpp) evaluated on
 * the GPU: what PairCHIMES::compute does with one process's local and ghost
 * atoms, behind plain arrays a pair style can hand over unchanged (x, type,
 * tag, and the full neighbour list with ghosts: inum, ilist, numneigh,
 * firstneigh).  Results are added onto the caller's force and tally arrays,
 * as the reference's loops and Pair::ev_tally_mb (etc/lmp/etc/pair.cpp:
 * 1425-1482) add onto LAMMPS's.  Every function returns 0 on success and a
 * nonzero code on error (message in chimes_gpu_last_error()); none exits.
 */
#ifndef CHIMES_LMP_H
#define CHIMES_LMP_H

#include "chimes_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct chimes_lmp chimes_lmp;

/* PairCHIMES::coeff and set_chimes_type (pair_chimes.cpp:176-236, 833-879):
 * LAMMPS types 1..ntypes are matched to the model's atom types by mass
 * within 1e-3; mass[1..ntypes] as LAMMPS's atom->mass (mass[0] unused).
 * Every type must match: the reference only warns when one does not, after
 * which its type table is shifted for the later types. */
chimes_lmp *chimes_lmp_create(chimes_gpu_model *m, int ntypes, const double *mass);
void chimes_lmp_free(chimes_lmp *p);

/* The pair style as built with -DTABULATION (pair_chimes.cpp:620-623,
 * 697-700; chimesFF.cpp:198-352, 715-745, 876-918, 1777-1924, 2178-2484).
 * on = 1: the 2-body interactions are read from the tables the parameter
 * file names on its "PAIRTYPE PARAMS:" lines of seven items (compute_2B_tab:
 * a local cubic spline per pair type, no penalty term), the 3-body ones from
 * the tables named on "INDEX:" lines of eight items (compute_3B_tab: a
 * Catmull-Rom tricubic at the runtime pair distances); the table files
 * <name>.energy and <name>.force are read now, relative to the parameter
 * file's directory.  A body order whose entries name no table stays
 * polynomial; 4-body is always polynomial.  on = 0: polynomials throughout.
 * Errors: unreadable or malformed tables; at compute, a pair at or below the
 * second point of its table or a triplet past the end of its table (where
 * the reference exits or reads past its arrays). */
int chimes_lmp_set_tabulation(chimes_lmp *p, int on);

/* PairCHIMES::build_mb_neighlists (pair_chimes.cpp:321-478): the 3- and
 * 4-mer lists within the largest 3-/4-body cutoffs plus the neighbour skin,
 * kept on the GPU for the following compute calls. */
int chimes_lmp_build_mb_neighlists(chimes_lmp *p, int nall, const double *x, const int *tag, int inum,
                                   const int *ilist, const int *numneigh, int *const *firstneigh, double skin);

/* PairCHIMES::compute (pair_chimes.cpp:480-831).  x[nall*3] (local atoms,
 * then ghosts), type[nall] (LAMMPS types), tag[nall]; the lists are rebuilt
 * first when ago == 0 (neighbor->ago) or none exist.  Between rebuilds only x
 * is read again: LAMMPS keeps the neighbour list, tags, types and ilist
 * unchanged until it rebuilds the list.  Adds onto f[nall*3]
 * and, as the flags ask, *eng_vdwl, virial[6] (xx yy zz xy xz yz),
 * eatom[nall], vatom[nall*6]; ghost forces are left for LAMMPS's reverse
 * communication. */
int chimes_lmp_compute(chimes_lmp *p, int nall, const double *x, const int *type, const int *tag, int inum,
                       const int *ilist, const int *numneigh, int *const *firstneigh, int ago, double skin,
                       int eflag_global, int eflag_atom, int vflag_global, int vflag_atom, double *f,
                       double *eng_vdwl, double *virial, double *eatom, double *vatom);

/* The many-body lists of the last build as atom indices in the reference's
 * loop order: trips (i, j, k), quads (i, j, k, l); free with chimes_gpu_free. */
int chimes_lmp_mb_lists(chimes_lmp *p, long *n3, int **trips, long *n4, int **quads);

/* The cluster rows of the pair style's FINGERPRINT build for the last
 * compute (pair_chimes.cpp:526-535, 625-640, 703-714, 777-790, recorded by
 * chimesFF.cpp:1694-1703, 2013-2022, 2603-2613), made on the GPU from the
 * device-resident neighbour rows and many-body lists: nb = 2, 3 or 4; a
 * cluster is recorded when its members are in order (2-body: i < j by atom
 * index, among the pairs tag[j] > tag[i]; 3-body: tag[i] < tag[j], tag[i] <
 * tag[k], tag[j] < tag[k]; 4-body: tag[i] < tag[j] < tag[k] < tag[l]) and
 * compute_nB evaluates it (type not excluded, every pair below its slot's
 * outer cutoff).  Each row holds the nb(nb-1)/2 pair distances (pair order
 * 01 02 03 12 13 23) and then the nb ChIMES type indices, as doubles, the
 * values writeClusterDataComp prints; rows come in the loops' order.  *rows
 * is malloc'd (free with chimes_gpu_free); chimes_gpu_fp_edges takes them. */
int chimes_lmp_fp_rows(chimes_lmp *p, int nb, long *nrows, double **rows);

/* Milliseconds of the last compute (HIP events): [0] input copies, [1] list
 * build (0 when the lists were reused), [2] 1- and 2-body, [3] 3-/4-body and
 * the read-back of the results. */
void chimes_lmp_timing(const chimes_lmp *p, double ms[4]);

#ifdef __cplusplus
}
#endif
#endif
