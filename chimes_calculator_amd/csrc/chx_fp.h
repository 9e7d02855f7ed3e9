// Cluster fingerprint histogram kernel (kernels_fp.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "chx_types.h"

namespace chx {

// Largest histogram the kernel keeps in LDS (32-bit counters).
constexpr int CHX_FP_MAX_BINS = 8192;

// counts[nbin] (zeroed by the caller) += the histogram of the distances
// between the np-edge clusters e1[n1] and e2[n2] (same != 0: e1 == e2 and
// pairs j > i); *bad set when a distance falls outside the bins.  False if
// np or the sizes are not supported.
bool launch_fp_hist(int np, long n1, const double *e1, long n2, const double *e2, int same, int nbin, double binw,
                    double maxd, unsigned long long *counts, int *bad, hipStream_t st);

// Sorted Morse-transformed edges edges[n*P] of cluster rows dist[n*P],
// types[n*nb] (read_flat_clusters); types already checked by the caller.
void launch_fp_edges(int nb, long n, const double *dist, const int *types, const chx_model_dev &M, double *edges,
                     hipStream_t st);

}  // namespace chx
