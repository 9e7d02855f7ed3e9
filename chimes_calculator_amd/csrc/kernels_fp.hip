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

#include "chx_cluster.h"
#include "chx_fp.h"

namespace chx {

// ---------------------------------------------------------------------------
// read_flat_clusters (calc_histograms.cpp:140-249) on the device: each cluster
// row (runtime pair distances, member types) becomes its NP Morse-transformed
// edge lengths, sorted ascending.  Pair p of the row uses the inner and outer
// cutoffs of its slot in the cluster type (the 2-body pair's own for NB = 2)
// and the Morse lambda of its pair type; transform() (calc_histograms.cpp:
// 126-134) is kept operation by operation (this file is built without FMA
// contraction).  The caller has checked the types (range, no excluded type).
// ---------------------------------------------------------------------------
__device__ __forceinline__ double fp_transform(double rcin, double rcout, double lambda, double rij)
{
    const double x_min = exp(-1 * rcin / lambda);
    const double x_max = exp(-1 * rcout / lambda);
    const double x_avg = 0.5 * (x_max + x_min);
    double x_diff = 0.5 * (x_max - x_min);
    x_diff *= -1.0;
    return (exp(-1 * rij / lambda) - x_avg) / x_diff;
}

template <int NB>
__global__ __launch_bounds__(256) void k_fp_edges(long n, const double *__restrict__ dist,
                                                  const int *__restrict__ types, chx_model_dev M, double *edges)
{
    constexpr int P = NB * (NB - 1) / 2;
    constexpr int pa[6] = {0, 0, 0, 1, 1, 2}, pb[6] = {1, 2, 3, 2, 3, 3};
    constexpr int pa3[3] = {0, 0, 1}, pb3[3] = {1, 2, 2};
    const long c = (long)blockIdx.x * 256 + threadIdx.x;
    if (c >= n) return;
    int ty[NB];
#pragma unroll
    for (int a = 0; a < NB; a++) ty[a] = types[c * NB + a];
    double e[P];
    int tup = 0;
#pragma unroll
    for (int a = 0; a < NB; a++) tup = tup * M.nt + ty[a];
    const int tau = NB == 2 ? 0 : (NB == 3 ? M.tripmap[tup] : M.quadmap[tup]);
#pragma unroll
    for (int p = 0; p < P; p++) {
        const int a = NB == 3 ? pa3[p] : pa[p], b = NB == 3 ? pb3[p] : pb[p];
        const chx_slot &s2 = M.slots2[M.pairmap[ty[a] * M.nt + ty[b]]];
        double rin = s2.rin, rout = s2.rout;
        if constexpr (NB > 2) {
            // the canonical slot whose runtime pair is p
            const int *perm = (NB == 3 ? M.trip_perm : M.quad_perm) + tup * NB;
            const chx_slot *sl = (NB == 3 ? M.slots3 : M.slots4) + tau * P;
            for (int q = 0; q < P; q++)
                if (runtime_pair<NB>(perm, q) == p) {
                    rin = sl[q].rin;
                    rout = sl[q].rout;
                }
        }
        e[p] = fp_transform(rin, rout, s2.lam, dist[c * P + p]);
    }
    // ascending (insertion order of equal values is immaterial)
#pragma unroll
    for (int i = 1; i < P; i++)
#pragma unroll
        for (int j = i; j > 0; j--) {
            const double lo = e[j - 1] < e[j] ? e[j - 1] : e[j], hi = e[j - 1] < e[j] ? e[j] : e[j - 1];
            e[j - 1] = lo;
            e[j] = hi;
        }
#pragma unroll
    for (int p = 0; p < P; p++) edges[c * P + p] = e[p];
}

void launch_fp_edges(int nb, long n, const double *dist, const int *types, const chx_model_dev &M, double *edges,
                     hipStream_t st)
{
    if (n <= 0) return;
    const dim3 grid((unsigned)((n + 255) / 256)), block(256);
    if (nb == 2) hipLaunchKernelGGL(k_fp_edges<2>, grid, block, 0, st, n, dist, types, M, edges);
    else if (nb == 3) hipLaunchKernelGGL(k_fp_edges<3>, grid, block, 0, st, n, dist, types, M, edges);
    else hipLaunchKernelGGL(k_fp_edges<4>, grid, block, 0, st, n, dist, types, M, edges);
}

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
