// Declarations of the evaluation kernels and their host launchers.
#pragma once
#include <hip/hip_runtime.h>

#include "chx_kernels.h"
#include "chx_types.h"

namespace chx {

// Uniform (per cluster type) arguments of the 3-/4-body kernels.
struct chx_type_args3 {
    int nt, tersoff;
    double fcut_var;
    const int *perm;           // [nt^3 * 3] canonical member order per type tuple
    chx_slot s[3];
};
struct chx_type_args4 {
    int nt, tersoff;
    double fcut_var;
    const int *perm;           // [nt^4 * 4]
    chx_slot s[6];
};

struct chx_red_args {
    const chx_red_model *models;
    const double *rec2E, *rec2F;
    const int *off3;
    const double *rec3E, *rec3F;
    const int *rec3M;
    const int *off4;
    const double *rec4E, *rec4F;
    const long long *rec4M;
};

__global__ void k_eval2(long nent, const int *e_owner, const chx_cfg *cfgs, const chx_atom_s *as,
                        const chx_entry_s *es, const chx_model_dev *models, double *rec2E, double *rec2F,
                        int *cfg_stat);
__global__ void k_reduce(int natoms, const chx_cfg *cfgs, const chx_atom_s *as, const int *nbr_off,
                         const chx_entry_s *es, const unsigned char *e_in3, const chx_red_args ra, double *Fe,
                         double *Fself, double *Eown, double *Sown);
__global__ void k_gather_rep(int natoms, const int *rstart, const int *rev, const double *Fe, const double *Fself,
                             double *Frep);
__global__ void k_gather_in(int natoms_in, int ncfg, const int *cfg_in_off, const chx_cfg *cfgs,
                            const double *Frep, double *Fout);
__global__ void k_cfg_sum(const chx_cfg *cfgs, const double *Eown, const double *Sown, double *Ecfg, double *Vcfg);

bool supported_D3(int D);
bool supported_D4(int D);
void launch_eval3(int D, bool mixed, long begin, long end, hipStream_t st, const int2 *list, const chx_cfg *cfgs,
                  const chx_atom_s *as, const int *nbr_off, const chx_entry_s *es, const chx_type_args3 &ta,
                  const void *Cd, double *recE, double *recF, int *recM, int *cfg_stat);
void launch_eval4(int D, bool mixed, long begin, long end, hipStream_t st, const int2 *list, const chx_cfg *cfgs,
                  const chx_atom_s *as, const int *nbr_off, const chx_entry_s *es, const chx_type_args4 &ta,
                  const void *Cd, double *recE, double *recF, long long *recM, int *cfg_stat);
void launch_clusters(int nb, int D, int n, hipStream_t st, const double *dist, const double *drv, const int *types,
                     const chx_model_dev &M, double *E, double *F, double *S, double *last, int *flags);

}  // namespace chx
