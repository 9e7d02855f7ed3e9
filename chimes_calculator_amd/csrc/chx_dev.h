// Device helpers shared by the geometry and evaluation kernels: status
// checks and exact fixed-point accumulation.
//
// Fixed point: a contribution v is split as v = (hi + lo * 2^-40) * 2^-20
// with integer hi and lo in [0, 2^40), both taken exactly from the double
// (v * 2^20 and its fractional part are exact operations; only bits below
// 2^-60 are dropped).  Accumulators hold the integer sums of hi and of lo, so
// every total is the exact sum of its contributions, independent of the order
// in which threads add them; converting a total back to double rounds once.
// |v| must stay below 2^42; larger values set CHX_BAD_RANGE.
#pragma once
#include <hip/hip_runtime.h>

#include "chx_types.h"

namespace chx {

#define CHX_FX_LIMIT 0x1p42

// Status flags as one value for the whole workgroup (read once, then shared),
// so that every wave of a block takes the same early exit.
__device__ __forceinline__ bool block_stop(const int *status, int mask)
{
    __shared__ int flags_s;
    if (threadIdx.x == 0) flags_s = __atomic_load_n(&status[CHX_ST_FLAGS], __ATOMIC_RELAXED);
    __syncthreads();
    return (flags_s & mask) != 0;
}

__device__ __forceinline__ void fx_split(double v, long long &hi, long long &lo, bool &bad)
{
    if (!(fabs(v) < CHX_FX_LIMIT)) {
        bad = true;
        hi = 0;
        lo = 0;
        return;
    }
    const double s = v * 0x1p20;
    const double f = floor(s);
    hi = (long long)f;
    lo = (long long)((s - f) * 0x1p40);
}

// acc[0] += hi, acc[1] += lo (integer atomics; zero parts are skipped).
__device__ __forceinline__ void fx_atomic(unsigned long long *acc, double v, bool &bad)
{
    if (v == 0.0) return;
    long long hi, lo;
    fx_split(v, hi, lo, bad);
    if (hi) atomicAdd(acc, (unsigned long long)hi);
    if (lo) atomicAdd(acc + 1, (unsigned long long)lo);
}

// Segmented sums over runs of equal keys in groups of W consecutive lanes
// (W = 16 or 64).  Clusters and pair entries are stored sorted by owner, so
// the owners of a wavefront's lanes come in a few runs: each run's integer
// parts are summed exactly across lanes and only the run's last lane issues
// global atomics.  Must be reached by all lanes of the wavefront.

// first lane (within the group) of this lane's run
template <int W>
__device__ __forceinline__ int fx_run_first(int key)
{
    const int l = threadIdx.x & (W - 1);
    const int kprev = __shfl_up(key, 1, W);
    int seg = (l == 0 || kprev != key) ? l : 0;
#pragma unroll
    for (int d = 1; d < W; d <<= 1) {
        const int t = __shfl_up(seg, d, W);
        if (l >= d && t > seg) seg = t;
    }
    return seg;
}

// inclusive segmented scan of K integers: lane l ends with the sum over
// lanes seg..l of its run
template <int W, int K>
__device__ __forceinline__ void fx_run_scan(long long (&x)[K], int seg)
{
    const int l = threadIdx.x & (W - 1);
#pragma unroll
    for (int d = 1; d < W; d <<= 1) {
        long long t[K];
#pragma unroll
        for (int k = 0; k < K; k++) t[k] = __shfl_up(x[k], d, W);
        if (l - d >= seg)
#pragma unroll
            for (int k = 0; k < K; k++) x[k] += t[k];
    }
}

// Adds the owner's force f into Facc[key] (acc[(3*key + k)*2 + {0,1}]) and
// the energy and virial terms q into Oacc[key] (acc[(7*key + k)*2 + {0,1}]),
// one atomic pair per nonzero component and run.  key < 0 contributes nothing.
template <int W>
__device__ __forceinline__ void fx_owner_runs(unsigned long long *Facc, unsigned long long *Oacc, int key,
                                              const double (&f)[3], const double (&q)[7], bool &bad)
{
    if (__ballot(key >= 0) == 0) return;
    long long hi[10], lo[10];
#pragma unroll
    for (int k = 0; k < 10; k++) {
        hi[k] = 0;
        lo[k] = 0;
        if (key >= 0) fx_split(k < 3 ? f[k] : q[k - 3], hi[k], lo[k], bad);
    }
    const int seg = fx_run_first<W>(key);
    fx_run_scan<W, 10>(hi, seg);
    fx_run_scan<W, 10>(lo, seg);
    const int l = threadIdx.x & (W - 1);
    const int knext = __shfl_down(key, 1, W);
    if (key >= 0 && (l == W - 1 || knext != key)) {
#pragma unroll
        for (int k = 0; k < 10; k++) {
            unsigned long long *a = k < 3 ? Facc + 2L * (3L * key + k) : Oacc + 2L * (7L * key + (k - 3));
            if (hi[k]) atomicAdd(a, (unsigned long long)hi[k]);
            if (lo[k]) atomicAdd(a + 1, (unsigned long long)lo[k]);
        }
    }
}

// Owner terms of tiles in which the four lanes l, l + 16, l + 32, l + 48
// (sub = lane / 16) hold one cluster and the same key: sub 0 adds the owner's
// force f into Facc[key], subs 1, 2, 3 add the energy and virial terms q
// (components 0-2, 3-4 and 5-6) into Oacc[key]; each sub sums over runs of
// equal keys in its 16 lanes, so a lane scans three values, not ten.
__device__ __forceinline__ void fx_owner_runs4(unsigned long long *Facc, unsigned long long *Oacc, int key,
                                               const double (&f)[3], const double (&q)[7], bool &bad)
{
    if (__ballot(key >= 0) == 0) return;
    const int sub = (threadIdx.x >> 4) & 3;
    double v[3];
    v[0] = sub == 0 ? f[0] : (sub == 1 ? q[0] : (sub == 2 ? q[3] : q[5]));
    v[1] = sub == 0 ? f[1] : (sub == 1 ? q[1] : (sub == 2 ? q[4] : q[6]));
    v[2] = sub == 0 ? f[2] : (sub == 1 ? q[2] : 0.0);
    long long hi[3], lo[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        hi[k] = 0;
        lo[k] = 0;
        if (key >= 0) fx_split(v[k], hi[k], lo[k], bad);
    }
    const int seg = fx_run_first<16>(key);
    fx_run_scan<16, 3>(hi, seg);
    fx_run_scan<16, 3>(lo, seg);
    const int l = threadIdx.x & 15;
    const int knext = __shfl_down(key, 1, 16);
    if (key >= 0 && (l == 15 || knext != key)) {
        unsigned long long *a = sub == 0 ? Facc + 6L * key
                                         : Oacc + 14L * key + (sub == 1 ? 0 : (sub == 2 ? 6 : 10));
        const int n = sub < 2 ? 3 : 2;
#pragma unroll
        for (int k = 0; k < 3; k++) {
            if (k >= n) break;
            if (hi[k]) atomicAdd(a + 2 * k, (unsigned long long)hi[k]);
            if (lo[k]) atomicAdd(a + 2 * k + 1, (unsigned long long)lo[k]);
        }
    }
}

// Per-workgroup LDS table that merges the force contributions of a tile by
// atom before they reach global memory: open addressing on the atom index,
// exact integer LDS atomics, then one global atomic pair per component and
// distinct atom.  CAP must exceed the number of contributions per flush.
template <int CAP>
struct fx_table {
    static_assert((CAP & (CAP - 1)) == 0, "table capacity must be a power of two");
    int key[CAP];
    unsigned long long val[CAP][6];  // (hi, lo) per component

    // all threads of the block; a barrier must follow before use
    __device__ __forceinline__ void clear()
    {
        for (int s = threadIdx.x; s < CAP; s += blockDim.x) {
            key[s] = -1;
#pragma unroll
            for (int k = 0; k < 6; k++) val[s][k] = 0ull;
        }
    }
    __device__ __forceinline__ void add(int atom, const double (&v)[3], bool &bad)
    {
        if (atom < 0) return;
        unsigned h = ((unsigned)atom * 2654435761u) >> 8;
        int s = (int)(h & (CAP - 1));
        for (int probe = 0; probe < CAP; probe++) {
            const int prev = atomicCAS(&key[s], -1, atom);
            if (prev == -1 || prev == atom) break;
            s = (s + 1) & (CAP - 1);
        }
#pragma unroll
        for (int k = 0; k < 3; k++) {
            if (v[k] == 0.0) continue;
            long long hi, lo;
            fx_split(v[k], hi, lo, bad);
            if (hi) atomicAdd(&val[s][2 * k], (unsigned long long)hi);
            if (lo) atomicAdd(&val[s][2 * k + 1], (unsigned long long)lo);
        }
    }
    // all threads of the block, after a barrier
    __device__ __forceinline__ void flush(unsigned long long *acc)
    {
        for (int s = threadIdx.x; s < CAP; s += blockDim.x) {
            const int a = key[s];
            if (a < 0) continue;
#pragma unroll
            for (int k = 0; k < 6; k++)
                if (val[s][k]) atomicAdd(acc + 6L * a + k, val[s][k]);
        }
    }
};

// Exact running total of normalised accumulators: lo parts are reduced below
// 2^40 before summing, so up to 2^23 accumulators add without overflow.
struct fx_sum {
    long long hi = 0;
    unsigned long long lo = 0;
    __device__ __forceinline__ void add(const unsigned long long *a)
    {
        hi += (long long)a[0] + (long long)(a[1] >> 40);
        lo += a[1] & 0xffffffffffull;
    }
    __device__ __forceinline__ void add_value(double v, bool &bad)
    {
        long long h, l;
        fx_split(v, h, l, bad);
        hi += h;
        lo += (unsigned long long)l;
    }
    __device__ __forceinline__ double value() const
    {
        const long long h = hi + (long long)(lo >> 40);
        const unsigned long long l = lo & 0xffffffffffull;
        return (double)h * 0x1p-20 + (double)l * 0x1p-60;
    }
};

}  // namespace chx
