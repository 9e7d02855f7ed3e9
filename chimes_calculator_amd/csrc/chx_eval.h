// Declarations of the evaluation kernels and their host launchers.
#pragma once
#include <hip/hip_runtime.h>

#include "chx_kernels.h"
#include "chx_types.h"

namespace chx {

// Arguments of the 3-/4-body evaluation over one list.
struct chx_list_args {
    long cap;                      // list capacity (slots)
    int G;                         // groups of this body order
    const int *gbase, *gend;       // [G+1], [G] aligned segment of each group
    const chx_group *groups;       // [G]
    const int2 *list;
    const chx_cfg *cfgs;
    const chx_atom_s *as;
    int nb_cap;  // neighbour images per owner slab
    const chx_entry_s *es;
    unsigned long long *Facc;
    unsigned long long *Cacc;      // [ncfg][CHX_CACC_SLOTS][CHX_CACC_STRIDE] energy and virial per configuration
    int *cfg_stat, *status;
};

__global__ void k_fin(int ncfg, int nitems, const int *cfg_item_off, const int *item_cfg, int natoms_in,
                      const int *in_cfg,
                      const chx_cfg *cfgs, const chx_atom_s *as, const chx_model_dev *models,
                      const unsigned long long *Facc,
                      const unsigned long long *Cacc, const int *nbr_cnt,
                      long long *part, int *done,
                      double *Fout, double *Ecfg, double *Vcfg, int *status);

bool supported_D3(int D);
bool supported_D4(int D);
// Cacc: the per-configuration energy and virial accumulators
void launch_eval2(int natoms, int nb_cap, hipStream_t st, const int *nbr_cnt, const chx_cfg *cfgs,
                  const chx_atom_s *as, const chx_entry_s *es, const chx_model_dev *models, unsigned long long *Facc,
                  unsigned long long *Cacc, int *cfg_stat, int *status);
// Vector-unit kernel for nb = 3 or 4 bodies.
void launch_evalv(int nb, int D, int prec, const chx_list_args &la, hipStream_t st);
// nb-body evaluation at precision prec (chx_precision) with the given kernel
// variant (CHX_VAR_*; BF16 always uses the matrix cores); the vector kernel
// (FP32 for the reduced precisions) stands in where the matrix-core kernels
// do not take D.
void launch_eval_body(int nb, int D, int prec, int variant, const chx_list_args &la, hipStream_t st);
// Whether variant has a kernel of its own for nb bodies at precision prec
// (the one-cluster-per-lane matrix-core variant: 3-body FP32-mixed only).
bool variant_applies(int nb, int prec, int variant);
void launch_clusters(int nb, int D, int n, hipStream_t st, const double *dist, const double *drv, const int *types,
                     const chx_model_dev &M, double *E, double *F, double *S, double *last, int *flags);

}  // namespace chx
