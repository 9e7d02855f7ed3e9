// Cluster fingerprint histogram (chimesFF/src/FP/calc_histograms.cpp:351-489,
// gen_flat_hists): the distance between every cluster i of one set and every
// cluster j of another (or every j > i of the same set) is binned.  Clusters
// are their NP sorted, Morse-transformed edge lengths; the distance is the
// Euclidean one (gen_flat_hists's composition weights are always zero: the
// vectors it hands to compute_composition_weight_{2,3,4}b hold only zeros in
// the positions those functions read).
//
// One thread per cluster i of a 256-cluster block; the block walks a span of
// the second set through LDS tiles, so every edge of set 2 is read from HBM
// once per 256 clusters of set 1.  Counts go to an LDS histogram (one per
// wavefront when the bins fit, against same-bin contention) and are added
// into the global 64-bit counts once per block.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "chx_fp.h"

namespace chx {

template <int NP, bool PERWAVE>
__global__ __launch_bounds__(256) void k_fp_hist(long n1, const double *__restrict__ e1, long n2,
                                                 const double *__restrict__ e2, int same, long jspan, int nbin,
                                                 double binw, double rbinw, double maxd, unsigned long long *counts,
                                                 int *bad)
{
    extern __shared__ unsigned int hist[];  // PERWAVE ? 4 * nbin : nbin
    __shared__ double tile[256 * NP];
    const int nh = PERWAVE ? 4 * nbin : nbin;
    for (int b = threadIdx.x; b < nh; b += 256) hist[b] = 0;
    const long ilo = (long)blockIdx.x * 256;
    const long i = ilo + threadIdx.x;
    const bool vi = i < n1;
    double a[NP];
#pragma unroll
    for (int k = 0; k < NP; k++) a[k] = vi ? e1[i * NP + k] : 0.0;
    unsigned int *h = PERWAVE ? hist + (threadIdx.x >> 6) * nbin : hist;
    const long j0 = (long)blockIdx.y * jspan, j1 = std::min(n2, j0 + jspan);
    bool badl = false;
    for (long jb = j0; jb < j1; jb += 256) {
        if (same && jb + 255 <= ilo) continue;  // every j of the tile <= every i of the block
        const int cnt = (int)std::min<long>(256, j1 - jb);
        __syncthreads();
        for (int q = threadIdx.x; q < cnt * NP; q += 256) tile[q] = e2[jb * NP + q];
        __syncthreads();
        if (!vi) continue;
        const int jstart = same ? (int)std::max<long>(0, std::min<long>(cnt, i + 1 - jb)) : 0;
        for (int jj = jstart; jj < cnt; jj++) {
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < NP; k++) {
                const double d = a[k] - tile[jj * NP + k];
                s += d * d;
            }
            // Bin estimate from the hardware reciprocal square root and one
            // Newton step (relative error about 1.5 e^2 for an estimate of
            // relative error e: below 1e-11 for any e up to 2^-20), times
            // 1 / binw.
            // Where it lies more than 1e-7 from a bin edge (|estimate - exact
            // quotient| < 1e-8 for at most 8192 bins) its floor is the exact
            // bin; elsewhere the exact arithmetic below decides.
            int bin = 0;
            bool exact = !(s > 0.0);
            if (!exact) {
                double y = __builtin_amdgcn_rsq(s);
                y = y * (1.5 - 0.5 * s * y * y);
                const double qa = s * y * rbinw;
                const double fl = floor(qa);
                const double fr = qa - fl;
                exact = fr < 1e-7 || fr > 1.0 - 1e-7;
                bin = (int)fl;
            }
            if (exact) {
                // sqrt(s) + |0 - 0|: adding +0 changes nothing
                const double tot = sqrt(s);
                bin = (int)floor(tot / binw);
                if (tot == maxd) bin -= 1;  // get_bin (calc_histograms.cpp:251-259)
            }
            if (bin < 0 || bin >= nbin) {
                badl = true;
                continue;
            }
            atomicAdd(&h[bin], 1u);
        }
    }
    __syncthreads();
    for (int b = threadIdx.x; b < nbin; b += 256) {
        unsigned long long v = hist[b];
        if (PERWAVE) v += (unsigned long long)hist[nbin + b] + hist[2 * nbin + b] + hist[3 * nbin + b];
        if (v) atomicAdd(&counts[b], v);
    }
    if (badl) atomicOr(bad, 1);
}

bool launch_fp_hist(int np, long n1, const double *e1, long n2, const double *e2, int same, int nbin, double binw,
                    double maxd, unsigned long long *counts, int *bad, hipStream_t st)
{
    if (n1 <= 0 || n2 <= 0) return true;
    if (nbin <= 0 || nbin > CHX_FP_MAX_BINS) return false;
    // spans of set 2 per block: enough blocks to fill the device, at most
    // 65535 spans, and far fewer than 2^32 / 256 pairs per block (32-bit LDS
    // counters)
    const long bx = (n1 + 255) / 256;
    long jspan = std::max<long>(2048, (n2 * bx + 4095) / 4096);
    jspan = std::max<long>(jspan, (n2 + 65534) / 65535);
    jspan = std::min<long>(jspan, 1L << 22);
    jspan = (jspan + 255) / 256 * 256;
    const long by = (n2 + jspan - 1) / jspan;
    if (by > 65535 || bx > 0x7fffffffL) return false;
    const bool perwave = nbin <= 1024;
    const size_t shm = sizeof(unsigned int) * (perwave ? 4 * nbin : nbin);
    const dim3 grid((unsigned)bx, (unsigned)by), block(256);
#define X(P)                                                                                                   \
    if (np == P) {                                                                                             \
        if (perwave)                                                                                           \
            hipLaunchKernelGGL((k_fp_hist<P, true>), grid, block, shm, st, n1, e1, n2, e2, same, jspan, nbin, \
                               binw, 1.0 / binw, maxd, counts, bad);                                                       \
        else                                                                                                   \
            hipLaunchKernelGGL((k_fp_hist<P, false>), grid, block, shm, st, n1, e1, n2, e2, same, jspan, nbin, \
                               binw, 1.0 / binw, maxd, counts, bad);                                                       \
        return true;                                                                                           \
    }
    X(1) X(3) X(6)
#undef X
    return false;
}

}  // namespace chx
