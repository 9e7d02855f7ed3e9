// LAMMPS pair style on the GPU: device arrays of one call and the kernel
// launchers (kernels_lmp.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "chx_tab.h"
#include "chx_types.h"

namespace chx {

// Longest filtered neighbour row (images within the padded many-body
// cutoffs, tag not below the owner's) the list kernel holds in LDS.
constexpr int CHX_LMP_CAND = 512;
// status words of the adapter
enum : int { CHX_LMP_ST_FLAGS = 0, CHX_LMP_ST_MAXCAND = 1, CHX_LMP_ST_WORDS = 4 };
enum : int { CHX_LMP_BAD_CAND = 1, CHX_LMP_BAD_RANGE = 2, CHX_LMP_BAD_TAB2 = 4, CHX_LMP_BAD_TAB3 = 8 };

struct chx_lmp_args {
    int inum, nall;
    const double *x;      // [nall * 3]
    const int *ctype;     // [nall] ChIMES type index
    const int *tag;       // [nall]
    const int *ilist;     // [inum]
    const int *nbr_off;   // [inum + 1] packed rows of firstneigh
    const int *nbr;       // [nbr_off[inum]]
    unsigned long long *Facc;  // [nall * 3 * 2] fixed-point forces
    unsigned long long *Eacc;  // [nall * 2] per-atom energy, or null
    unsigned long long *Vacc;  // [nall * 6 * 2] per-atom virial (LAMMPS order), or null
    unsigned long long *Gacc;  // [nall * 7 * 2] partial sums of the global energy + virial (only the total over all atoms is read), or null
    int *status;               // [CHX_LMP_ST_WORDS]
    int eflag_global, vflag_global;
};

// T: the tables of a -DTABULATION build (null or empty: the polynomials)
void launch_lmp_pair(const chx_lmp_args &A, const chx_model_dev &M, const chx_tabs_dev *T, hipStream_t st);
// mode 0: per-owner counts into cnt3 / cnt4; mode 1: lists written at off3 / off4
void launch_lmp_lists(int mode, const chx_lmp_args &A, double pad3, double pad4, int has4, int *cnt3, int *cnt4,
                      const long long *off3, const long long *off4, int *trips, int *quads, hipStream_t st);
void launch_lmp_clusters(int nb, int D, long n, const int *list, const chx_lmp_args &A, const chx_model_dev &M,
                         const chx_tabs_dev *T, hipStream_t st);
// Fingerprint rows of nb-body clusters (n loop entries: list entries, or for
// nb = 2 the packed neighbour-row entries): mode 0 counts rows per block of
// 256 entries into blk_cnt, mode 1 writes them from blk_off (rows of
// nb(nb-1)/2 distances then nb types).
void launch_lmp_fp_rows(int nb, int mode, long n, const int *list, const chx_lmp_args &A, const chx_model_dev &M,
                        int *blk_cnt, const long long *blk_off, double *rows, hipStream_t st);
// F[nall*3], eatom[nall] (if Eacc), vatom[nall*6] (if Vacc), glob[7] (if
// Gacc; gsum[14] and *done zeroed scratch of the global reduction)
void launch_lmp_out(const chx_lmp_args &A, double *F, double *eatom, double *vatom, double *glob,
                    unsigned long long *gsum, int *done, hipStream_t st);

}  // namespace chx
