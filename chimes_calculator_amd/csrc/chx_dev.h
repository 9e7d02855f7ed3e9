// Device helpers shared by the geometry and evaluation kernels: status
// checks and exact fixed-point accumulation.
//
// Fixed point: a contribution v is split as v = (hi + lo * 2^-40) * 2^-20
// with integers hi and |lo| <= 2^40, both taken exactly from the double
// (only bits below 2^-60 are rounded away).  Accumulators hold the integer
// sums of hi and of lo (lo as a signed 64-bit sum), so every total is the
// exact sum of its contributions, independent of the order in which threads
// add them; converting a total back to double rounds once.  |v| must stay
// below 2^42; larger values set CHX_BAD_RANGE.
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
    const double s = v * 0x1p20;  // exact
    if (fabs(v) < 0x1p31) {
        // |s| < 2^51: adding 1.5 * 2^52 rounds s to an integer held in the
        // low bits of the sum, and the remainder r = s - round(s), |r| <=
        // 1/2, is exact; lo = round(r * 2^40) the same way (one fused
        // operation, r * 2^40 being exact)
        const double M = 0x1.8p52;
        const double t = s + M;
        hi = __double_as_longlong(t) - __double_as_longlong(M);
        const double r = s - (t - M);
        lo = __double_as_longlong(fma(r, 0x1p40, M)) - __double_as_longlong(M);
        return;
    }
    if (!(fabs(v) < CHX_FX_LIMIT)) {
        bad = true;
        hi = 0;
        lo = 0;
        return;
    }
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

// Lane movement within rows of 16 lanes (DPP row_shr:D, zero where the
// source lane is outside the row), and a value of a fixed lane read into a
// scalar register: the scans below use no LDS.
template <int D>
__device__ __forceinline__ int row_shr_i32(int x)
{
    return __builtin_amdgcn_update_dpp(0, x, 0x110 + D, 0xf, 0xf, true);
}
template <int D>
__device__ __forceinline__ long long row_shr_i64(long long x)
{
    const int lo = row_shr_i32<D>((int)(unsigned long long)x);
    const int hi = row_shr_i32<D>((int)((unsigned long long)x >> 32));
    return (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ long long readlane_i64(long long x, int l)
{
    const int lo = __builtin_amdgcn_readlane((int)(unsigned long long)x, l);
    const int hi = __builtin_amdgcn_readlane((int)((unsigned long long)x >> 32), l);
    return (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

// first lane (within the group) of this lane's run: a running maximum of the
// run starts, within rows of 16 lanes, then across rows (W = 64)
template <int W>
__device__ __forceinline__ int fx_run_first(int key)
{
    static_assert(W == 16 || W == 64, "run groups of 16 or 64 lanes");
    const int l = threadIdx.x & (W - 1);
    const int kprev = __shfl_up(key, 1, W);
    int seg = (l == 0 || kprev != key) ? l : 0;
    seg = max(seg, row_shr_i32<1>(seg));
    seg = max(seg, row_shr_i32<2>(seg));
    seg = max(seg, row_shr_i32<4>(seg));
    seg = max(seg, row_shr_i32<8>(seg));
    if constexpr (W == 64) {
        const int r = l >> 4;
        const int s15 = __builtin_amdgcn_readlane(seg, 15);
        const int s31 = max(__builtin_amdgcn_readlane(seg, 31), s15);
        const int s47 = max(__builtin_amdgcn_readlane(seg, 47), s31);
        seg = max(seg, r == 0 ? 0 : (r == 1 ? s15 : (r == 2 ? s31 : s47)));
    }
    return seg;
}

// inclusive segmented scan of K integers: lane l ends with the sum over
// lanes seg..l of its run.  Within rows of 16 lanes by row shifts; then each
// row adds the row totals before it that lie in its run (W = 64).
template <int W, int K>
__device__ __forceinline__ void fx_run_scan(long long (&x)[K], int seg)
{
    const int l = threadIdx.x & (W - 1);
    const int lr = threadIdx.x & 15;
#define CHX_RUN_STEP(D)                                                  \
    _Pragma("unroll") for (int k = 0; k < K; k++)                        \
    {                                                                    \
        const long long t = row_shr_i64<D>(x[k]);                        \
        if (lr >= D && l - D >= seg) x[k] += t;                          \
    }
    CHX_RUN_STEP(1)
    CHX_RUN_STEP(2)
    CHX_RUN_STEP(4)
    CHX_RUN_STEP(8)
#undef CHX_RUN_STEP
    if constexpr (W == 64) {
        const int r = l >> 4;
#pragma unroll
        for (int k = 0; k < K; k++) {
            const long long R0 = readlane_i64(x[k], 15), R1 = readlane_i64(x[k], 31), R2 = readlane_i64(x[k], 47);
            const long long Rm1 = r == 1 ? R0 : (r == 2 ? R1 : R2);  // row r - 1
            const long long Rm2 = r == 2 ? R0 : R1;                   // row r - 2
            if (r >= 1 && seg <= 16 * r - 1) x[k] += Rm1;
            if (r >= 2 && seg <= 16 * r - 17) x[k] += Rm2;
            if (r >= 3 && seg <= 16 * r - 33) x[k] += R0;
        }
    }
}

// Inclusive sum of a double over each group of W = 8 or 16 lanes (row
// shifts, a fixed order): the group's last lane holds its total.
template <int W>
__device__ __forceinline__ double group_sum(double x)
{
    static_assert(W == 8 || W == 16, "groups of 8 or 16 lanes");
    const int l = threadIdx.x & (W - 1);
#define CHX_ROW_ADD(D)                                                                      \
    {                                                                                       \
        const long long b = row_shr_i64<D>(__double_as_longlong(x));                        \
        if (W == 16 || l >= D) x += __longlong_as_double(b);                                \
    }
    CHX_ROW_ADD(1)
    CHX_ROW_ADD(2)
    CHX_ROW_ADD(4)
    if constexpr (W == 16) CHX_ROW_ADD(8)
#undef CHX_ROW_ADD
    return x;
}

// A lane's value at DPP control CTRL (every source lane inside the row).
template <int CTRL>
__device__ __forceinline__ long long dpp_i64(long long x)
{
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned long long)x, CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)((unsigned long long)x >> 32), CTRL, 0xf, 0xf, false);
    return (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

// Wavefront totals of 16 integers per lane, transposed: lane l returns the
// sum over all 64 lanes of x[l & 15].  Within each row of 16 lanes, four
// exchange steps pair every lane with one whose bit 3, 2, 1, 0 differs and
// whose higher bits agree (row rotation by 8, half-row mirror, two quad
// permutations); at each step a lane keeps the half of its slots selected
// by that bit, adds the partner's copy of them and sends the other half, so
// 8 + 4 + 2 + 1 exchanges replace sixteen separate row sums.  The four rows
// are then added by two lane exchanges.  Must be reached by all lanes.
__device__ __forceinline__ long long fx_wave_sums16(long long (&x)[16])
{
    const int i = threadIdx.x & 15;
#define CHX_T_STEP(N, BIT, CTRL)                                           \
    {                                                                      \
        const bool up = (i >> BIT) & 1;                                    \
        _Pragma("unroll") for (int j = 0; j < N; j++)                      \
        {                                                                  \
            const long long keep = up ? x[j + N] : x[j];                   \
            const long long send = up ? x[j] : x[j + N];                   \
            x[j] = keep + dpp_i64<CTRL>(send);                             \
        }                                                                  \
    }
    CHX_T_STEP(8, 3, 0x128)  // row_ror:8 (lane i ^ 8)
    CHX_T_STEP(4, 2, 0x141)  // row_half_mirror (lane i ^ 7)
    CHX_T_STEP(2, 1, 0x4e)   // quad_perm [2, 3, 0, 1] (lane i ^ 2)
    CHX_T_STEP(1, 0, 0xb1)   // quad_perm [1, 0, 3, 2] (lane i ^ 1)
#undef CHX_T_STEP
    long long t = x[0];
    t += __shfl_xor(t, 16, 64);
    t += __shfl_xor(t, 32, 64);
    return t;
}

// Adds the energy and virial terms q of each lane's cluster into the
// accumulator of its configuration, copy `slot` (wave-uniform) of the
// CHX_CACC_SLOTS: Cacc[(key * SLOTS + slot) * STRIDE + 2k + {0,1}] (key < 0:
// nothing).  A wavefront whose clusters all belong to one configuration (the
// common case: configurations hold thousands of clusters) sums the fourteen
// integer parts over its lanes at once (fx_wave_sums16) and its first
// fourteen lanes add one each; otherwise the lanes' keys form runs
// (clusters are stored by owner, owners by configuration) and each run adds
// its sums.  Must be reached by all lanes of the wavefront.
__device__ __forceinline__ unsigned long long *fx_cfg_line(unsigned long long *Cacc, int key, int slot)
{
    return Cacc + ((long)key * CHX_CACC_SLOTS + slot) * CHX_CACC_STRIDE;
}

__device__ __forceinline__ void fx_cfg_sums(unsigned long long *Cacc, int key, int slot, const double (&q)[7],
                                            bool &bad)
{
    const unsigned long long act = __ballot(key >= 0);
    if (act == 0) return;
    const int k0 = __builtin_amdgcn_readlane(key, __ffsll((long long)act) - 1);
    const int l = threadIdx.x & 63;
    if (__ballot(key >= 0 && key != k0) == 0) {
        long long x[16];
#pragma unroll
        for (int k = 0; k < 7; k++) {
            x[2 * k] = x[2 * k + 1] = 0;
            if (key >= 0) fx_split(q[k], x[2 * k], x[2 * k + 1], bad);
        }
        x[14] = x[15] = 0;
        const long long t = fx_wave_sums16(x);
        if (l < 14 && t) atomicAdd(fx_cfg_line(Cacc, k0, slot) + l, (unsigned long long)t);
        return;
    }
    const int seg = fx_run_first<64>(key);
    const int knext = __shfl_down(key, 1, 64);
    const bool last = key >= 0 && (l == 63 || knext != key);
#pragma unroll
    for (int k = 0; k < 7; k++) {
        long long x[2] = {0, 0};
        if (key >= 0) fx_split(q[k], x[0], x[1], bad);
        fx_run_scan<64, 2>(x, seg);
        if (last) {
            unsigned long long *a = fx_cfg_line(Cacc, key, slot) + 2 * k;
            if (x[0]) atomicAdd(a, (unsigned long long)x[0]);
            if (x[1]) atomicAdd(a + 1, (unsigned long long)x[1]);
        }
        __builtin_amdgcn_sched_barrier(0);
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
    // as add, but when every slot is taken by other atoms the contribution
    // goes straight to the global accumulators acc[(3*atom + k)*2 + {0,1}]
    __device__ __forceinline__ void add_or_global(int atom, const double (&v)[3], unsigned long long *acc, bool &bad)
    {
        if (atom < 0) return;
        unsigned h = ((unsigned)atom * 2654435761u) >> 8;
        int s = (int)(h & (CAP - 1));
        bool found = false;
        for (int probe = 0; probe < CAP; probe++) {
            const int prev = atomicCAS(&key[s], -1, atom);
            if (prev == -1 || prev == atom) { found = true; break; }
            s = (s + 1) & (CAP - 1);
        }
#pragma unroll
        for (int k = 0; k < 3; k++) {
            if (v[k] == 0.0) continue;
            if (!found) { fx_atomic(acc + 2L * (3L * atom + k), v[k], bad); continue; }
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

// Exact running total of normalised accumulators: a signed lo sum is split
// into its multiple of 2^40 (moved into hi, arithmetic shift) and a remainder
// in [0, 2^40) before summing, so up to 2^23 accumulators add without
// overflow.
struct fx_sum {
    long long hi = 0;
    long long lo = 0;
    __device__ __forceinline__ void add(const unsigned long long *a)
    {
        hi += (long long)a[0] + ((long long)a[1] >> 40);
        lo += (long long)(a[1] & 0xffffffffffull);
    }
    __device__ __forceinline__ void add_value(double v, bool &bad)
    {
        long long h, l;
        fx_split(v, h, l, bad);
        hi += h;
        lo += l;
    }
    __device__ __forceinline__ double value() const
    {
        const long long h = hi + (lo >> 40);
        const long long l = lo & 0xffffffffffll;
        return (double)h * 0x1p-20 + (double)l * 0x1p-60;
    }
};

}  // namespace chx
