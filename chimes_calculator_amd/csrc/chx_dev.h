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

// Adds K per-lane values into the accumulators of each lane's owner
// (acc[(K*owner + k)*2 + {0,1}]); owner < 0 contributes nothing.  Must be
// reached by all 64 lanes.  When every contributing lane has the same owner
// the integer parts are first summed across the wave (exactly), then added
// once.
template <int K>
__device__ __forceinline__ void fx_owner(unsigned long long *acc, int owner, const double (&q)[K], bool &bad)
{
    long long hi[K], lo[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        hi[k] = 0;
        lo[k] = 0;
        if (owner >= 0) fx_split(q[k], hi[k], lo[k], bad);
    }
    const unsigned long long act = __ballot(owner >= 0);
    if (act == 0) return;
    const int o0 = __shfl(owner, __builtin_ctzll(act), 64);
    if (__ballot(owner >= 0 && owner != o0) == 0) {
#pragma unroll
        for (int w = 32; w >= 1; w >>= 1)
#pragma unroll
            for (int k = 0; k < K; k++) {
                hi[k] += __shfl_xor(hi[k], w, 64);
                lo[k] += __shfl_xor(lo[k], w, 64);
            }
        if ((threadIdx.x & 63) == 0)
#pragma unroll
            for (int k = 0; k < K; k++) {
                unsigned long long *a = acc + 2L * ((long)K * o0 + k);
                if (hi[k]) atomicAdd(a, (unsigned long long)hi[k]);
                if (lo[k]) atomicAdd(a + 1, (unsigned long long)lo[k]);
            }
    } else if (owner >= 0) {
#pragma unroll
        for (int k = 0; k < K; k++) {
            unsigned long long *a = acc + 2L * ((long)K * owner + k);
            if (hi[k]) atomicAdd(a, (unsigned long long)hi[k]);
            if (lo[k]) atomicAdd(a + 1, (unsigned long long)lo[k]);
        }
    }
}

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
