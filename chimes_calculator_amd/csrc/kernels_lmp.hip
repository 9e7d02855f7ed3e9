// LAMMPS pair style chimesFF (etc/lmp/src/pair_chimes.cpp) on the GPU, one
// process's local and ghost atoms per call:
//   k_lmp_pair      1- and 2-body loops of PairCHIMES::compute (567-656)
//   k_lmp_lists     PairCHIMES::build_mb_neighlists (321-478)
//   k_lmp_clusters  3- and 4-body loops (673-829) over those lists
//   k_lmp_out       fixed-point accumulators -> forces and per-atom tallies
//   k_lmp_global    global energy and virial
// Every tally of Pair::ev_tally_mb (etc/lmp/etc/pair.cpp:1425-1482) is added
// in the exact fixed-point form of the batch pipeline, so results do not
// depend on the order of the lists or on the kernel schedule.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "chx_cluster.h"
#include "chx_dev.h"
#include "chx_lmp.h"
#include "chx_tab.h"

namespace chx {

// get_dist (pair_chimes.cpp:303-312): dr = x_b - x_a, same operation order
__device__ __forceinline__ double lmp_dist(const double *a, const double *b, double (&dr)[3])
{
    dr[0] = b[0] - a[0];
    dr[1] = b[1] - a[1];
    dr[2] = b[2] - a[2];
    return sqrt(dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2]);
}

// ChIMES stress (xx xy xz yy yz zz) -> LAMMPS virial order (xx yy zz xy xz yz)
__device__ __forceinline__ void lmp_order(const double (&s)[6], double (&v)[6])
{
    v[0] = s[0]; v[1] = s[3]; v[2] = s[5];
    v[3] = s[1]; v[4] = s[2]; v[5] = s[4];
}

// ---------------------------------------------------------------------------
// 1- and 2-body: sixteen lanes per owner ilist[ii]; lane l takes row entries
// l, l + 16, ...  Pairs with tag[j] > tag[i] only (pair_chimes.cpp:603).  The
// owner's force and tallies are summed over its lanes in double in that fixed
// order and added exactly once; the partner's are added per pair.  TAB: the
// pair's energy and force scalar come from its table (compute_2B_tab,
// chimesFF.cpp:1782-1856: cutoff test, table lookups, no penalty term).
// ---------------------------------------------------------------------------
template <bool TAB>
__global__ __launch_bounds__(256) void k_lmp_pair(chx_lmp_args A, chx_model_dev M, chx_tabs_dev T)
{
    const int sub = threadIdx.x & 15;
    const int ii = (int)(blockIdx.x * 16 + (threadIdx.x >> 4));
    if (ii >= A.inum) return;  // whole groups of sixteen leave together
    const int i = A.ilist[ii];
    const int itag = A.tag[i], ti = A.ctype[i];
    const double xi[3] = {A.x[3L * i], A.x[3L * i + 1], A.x[3L * i + 2]};
    // owner force, owner energy share, owner virial share, global energy, global virial
    double fi[3] = {0, 0, 0}, ea = 0.0, va[6] = {0, 0, 0, 0, 0, 0}, ge = 0.0, gv[6] = {0, 0, 0, 0, 0, 0};
    bool bad = false, tab_err = false;
    int flags = 0;
    const int beg = A.nbr_off[ii], end = A.nbr_off[ii + 1];
    for (int q = beg + sub; q < end; q += 16) {
        const int j = A.nbr[q];
        if (A.tag[j] <= itag) continue;
        const double xj[3] = {A.x[3L * j], A.x[3L * j + 1], A.x[3L * j + 2]};
        double dr[3];
        const double r = lmp_dist(xi, xj, dr);
        const int pt = M.pairmap[ti * M.nt + A.ctype[j]];
        double E, fs;
        if constexpr (TAB) {
            if (r >= M.slots2[pt].rout) continue;
            tab2_lookup(T, pt, r, E, fs, tab_err);
        } else {
            eval_pair(r, M.slots2[pt], M.c2 + (long)pt * (M.order[0] + 1), M.order[0], M.tersoff, M.fcut_var,
                      M.pen_dist, M.pen_scale, E, fs, flags);
            if (r >= M.slots2[pt].rout) continue;
        }
        const double F[3] = {fs * dr[0], fs * dr[1], fs * dr[2]};
        const double s[6] = {-(fs * (dr[0] * dr[0])), -(fs * (dr[0] * dr[1])), -(fs * (dr[0] * dr[2])),
                             -(fs * (dr[1] * dr[1])), -(fs * (dr[1] * dr[2])), -(fs * (dr[2] * dr[2]))};
        double v[6];
        lmp_order(s, v);
        for (int k = 0; k < 3; k++) {
            fi[k] += F[k];
            fx_atomic(A.Facc + 2L * (3L * j + k), -F[k], bad);
        }
        ge += E;
        ea += E / 2;
        if (A.Eacc) fx_atomic(A.Eacc + 2L * j, E / 2, bad);
        for (int k = 0; k < 6; k++) {
            gv[k] += v[k];
            va[k] += v[k] / 2;
            if (A.Vacc) fx_atomic(A.Vacc + 2L * (6L * j + k), v[k] / 2, bad);
        }
    }
#pragma unroll
    for (int w = 8; w >= 1; w >>= 1) {
        for (int k = 0; k < 3; k++) fi[k] += __shfl_xor(fi[k], w, 64);
        ea += __shfl_xor(ea, w, 64);
        ge += __shfl_xor(ge, w, 64);
        for (int k = 0; k < 6; k++) {
            va[k] += __shfl_xor(va[k], w, 64);
            gv[k] += __shfl_xor(gv[k], w, 64);
        }
    }
    if (sub == 0) {
        const double e1 = M.eoff[ti];  // compute_1B, tallied for one atom
        for (int k = 0; k < 3; k++) fx_atomic(A.Facc + 2L * (3L * i + k), fi[k], bad);
        if (A.Eacc) {
            fx_atomic(A.Eacc + 2L * i, e1, bad);
            fx_atomic(A.Eacc + 2L * i, ea, bad);
        }
        if (A.Vacc)
            for (int k = 0; k < 6; k++) fx_atomic(A.Vacc + 2L * (6L * i + k), va[k], bad);
        if (A.Gacc) {
            if (A.eflag_global) {
                fx_atomic(A.Gacc + 2L * (7L * i), e1, bad);
                fx_atomic(A.Gacc + 2L * (7L * i), ge, bad);
            }
            if (A.vflag_global)
                for (int k = 0; k < 6; k++) fx_atomic(A.Gacc + 2L * (7L * i + 1 + k), gv[k], bad);
        }
    }
    if (bad) atomicOr(&A.status[CHX_LMP_ST_FLAGS], CHX_LMP_BAD_RANGE);
    if (tab_err) atomicOr(&A.status[CHX_LMP_ST_FLAGS], CHX_LMP_BAD_TAB2);
}

// ---------------------------------------------------------------------------
// Many-body lists, one wavefront per owner ilist[ii].  The owner's row is
// filtered to J = {j : j != i, tag[j] >= tag[i], d_ij below either padded
// cutoff}, kept in row order in LDS; every k and l of the reference's loops
// lies in J.  Triplets come from the (j, k) pairs of J x J in row-major order,
// quadruplets from each such pair within the 4-body cutoff followed by every
// l of J, so ballot compaction writes them in the reference's loop order.
// MODE 0 counts per owner, MODE 1 writes at the scanned offsets.
// ---------------------------------------------------------------------------
template <int MODE>
__global__ __launch_bounds__(64) void k_lmp_lists(chx_lmp_args A, double pad3, double pad4, int has4, int *cnt3,
                                                  int *cnt4, const long long *off3, const long long *off4,
                                                  int *trips, int *quads)
{
    __shared__ int cj[CHX_LMP_CAND], ctag[CHX_LMP_CAND];
    __shared__ double cx[CHX_LMP_CAND][3], cd[CHX_LMP_CAND];
    const int ii = blockIdx.x, lane = threadIdx.x;
    const unsigned long long below = lane ? (~0ull >> (64 - lane)) : 0ull;
    const int i = A.ilist[ii], itag = A.tag[i];
    const double xi[3] = {A.x[3L * i], A.x[3L * i + 1], A.x[3L * i + 2]};
    const int beg = A.nbr_off[ii], n = A.nbr_off[ii + 1] - beg;
    int m = 0;
    for (int b0 = 0; b0 < n; b0 += 64) {
        bool v = false;
        int j = 0;
        double xj[3] = {0, 0, 0}, d = 0.0;
        if (b0 + lane < n) {
            j = A.nbr[beg + b0 + lane];
            if (j != i && A.tag[j] >= itag) {
                xj[0] = A.x[3L * j]; xj[1] = A.x[3L * j + 1]; xj[2] = A.x[3L * j + 2];
                double dr[3];
                d = lmp_dist(xi, xj, dr);
                v = !(d >= pad3 && d >= pad4);
            }
        }
        const unsigned long long bal = __ballot(v);
        const int r = m + __popcll(bal & below);
        if (v && r < CHX_LMP_CAND) {
            cj[r] = j;
            ctag[r] = A.tag[j];
            cx[r][0] = xj[0]; cx[r][1] = xj[1]; cx[r][2] = xj[2];
            cd[r] = d;
        }
        m += __popcll(bal);
    }
    if (m > CHX_LMP_CAND) {
        if (lane == 0) {
            atomicOr(&A.status[CHX_LMP_ST_FLAGS], CHX_LMP_BAD_CAND);
            atomicMax(&A.status[CHX_LMP_ST_MAXCAND], m);
        }
        return;
    }
    __syncthreads();
    long long c3 = 0, c4 = 0;
    const long long o3 = MODE ? off3[ii] : 0, o4 = MODE ? off4[ii] : 0;
    const int total = m * m;
    for (int p0 = 0; p0 < total; p0 += 64) {
        const int p = p0 + lane;
        bool v3 = false, v4 = false;
        int a = 0, b = 0;
        if (p < total) {
            a = p / m;
            b = p - a * m;
            if (cj[a] != cj[b] && ctag[b] >= ctag[a]) {
                double dr[3];
                const double dab = lmp_dist(cx[a], cx[b], dr);
                const double da = cd[a], db = cd[b];
                v3 = da < pad3 && db < pad3 && dab < pad3;
                v4 = has4 && da < pad4 && db < pad4 && dab < pad4;
            }
        }
        const unsigned long long b3 = __ballot(v3);
        if (MODE && v3) {
            int *t = trips + 3 * (o3 + c3 + __popcll(b3 & below));
            t[0] = i; t[1] = cj[a]; t[2] = cj[b];
        }
        c3 += __popcll(b3);
        unsigned long long b4 = __ballot(v4);
        while (b4) {
            const int t = __ffsll((long long)b4) - 1;
            b4 &= b4 - 1;
            const int pt = p0 + t, at = pt / m, bt = pt - at * m;
            const int ja = cj[at], jb = cj[bt], tb = ctag[bt];
            for (int q0 = 0; q0 < m; q0 += 64) {
                const int q = q0 + lane;
                bool v = false;
                if (q < m) {
                    const int l = cj[q];
                    if (l != ja && l != jb && ctag[q] >= tb && cd[q] < pad4) {
                        double dr[3];
                        v = lmp_dist(cx[at], cx[q], dr) < pad4 && lmp_dist(cx[bt], cx[q], dr) < pad4;
                    }
                }
                const unsigned long long bl = __ballot(v);
                if (MODE && v) {
                    int *u = quads + 4 * (o4 + c4 + __popcll(bl & below));
                    u[0] = i; u[1] = ja; u[2] = jb; u[3] = cj[q];
                }
                c4 += __popcll(bl);
            }
        }
    }
    if (!MODE && lane == 0) {
        cnt3[ii] = (int)c3;
        cnt4[ii] = (int)c4;
    }
}

// ---------------------------------------------------------------------------
// 3- and 4-body: one thread per listed cluster.  Pair distances and vectors
// as the reference forms them (ij, ik, jk / ij, ik, il, jk, jl, kl; x_b - x_a),
// the terms of compute_3B / compute_4B, then the forces of every member and
// the tallies of ev_tally_mb (energy and virial shared equally by the
// members; the global part kept per owner atom).
// ---------------------------------------------------------------------------
// One value of the owner atom: summed over the wavefront's runs of equal keys
// (the lists are sorted by owner) and added by each run's last lane at a.
__device__ __forceinline__ void lmp_run_add(int key, int seg, bool last, double v, unsigned long long *a, bool &bad)
{
    long long x[2] = {0, 0};
    if (key >= 0) fx_split(v, x[0], x[1], bad);
    fx_run_scan<64, 2>(x, seg);
    if (last) {
        if (x[0]) atomicAdd(a, (unsigned long long)x[0]);
        if (x[1]) atomicAdd(a + 1, (unsigned long long)x[1]);
    }
    __builtin_amdgcn_sched_barrier(0);
}

// compute_3B_tab (chimesFF.cpp:2183-2313): exclusion and the cutoffs of the
// mapped slots as compute_3B, then the table at the runtime pair distances
// (ij, ik, jk; get_tab_3B's sort is not executed, 2459-2470); forces and
// stress of ij, ik, jk with the distance-tensor products, in that order.
__device__ __forceinline__ void tab3_cluster(const chx_model_dev &M, const chx_tabs_dev &T, const int (&ty)[3],
                                             const double (&dist)[3], const double (&dr)[3][3], double &E,
                                             double (&fo)[3][3], double (&S)[6], bool &evaluated, bool &err)
{
    E = 0.0;
    for (int a = 0; a < 3; a++) fo[a][0] = fo[a][1] = fo[a][2] = 0.0;
    for (int k = 0; k < 6; k++) S[k] = 0.0;
    evaluated = false;
    const int tau = M.tripmap[(ty[0] * M.nt + ty[1]) * M.nt + ty[2]];
    if (tau < 0 || !cluster_inside<3>(M, ty, dist)) return;
    if (tau >= T.ntab3) {
        err = true;
        return;
    }
    double v[4];
    tab3_lookup(T, tau, dist[0], dist[1], dist[2], v, err);
    evaluated = true;
    E = v[0];
    const int ma[3] = {0, 0, 1}, mb[3] = {1, 2, 2};
#pragma unroll
    for (int q = 0; q < 3; q++) {
        const double f = v[1 + q];
        const double *d = dr[q];
        for (int k = 0; k < 3; k++) {
            fo[ma[q]][k] += f * d[k];
            fo[mb[q]][k] -= f * d[k];
        }
        S[0] -= f * (d[0] * d[0]);
        S[1] -= f * (d[0] * d[1]);
        S[2] -= f * (d[0] * d[2]);
        S[3] -= f * (d[1] * d[1]);
        S[4] -= f * (d[1] * d[2]);
        S[5] -= f * (d[2] * d[2]);
    }
}

template <int NB, int D, bool TAB>
__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(2))) void k_lmp_clusters(long n, const int *__restrict__ list, chx_lmp_args A,
                                                      chx_model_dev M, chx_tabs_dev T)
{
    // the members' forces of the workgroup's clusters, merged per atom
    __shared__ fx_table<512> tab;  // >= 128 clusters x NB members
    tab.clear();
    __syncthreads();
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = c < n;
    constexpr int P = NB * (NB - 1) / 2;
    int mem[NB], ty[NB];
    double xm[NB][3];
#pragma unroll
    for (int a = 0; a < NB; a++) {
        mem[a] = valid ? list[c * NB + a] : 0;
        ty[a] = A.ctype[mem[a]];
        for (int k = 0; k < 3; k++) xm[a][k] = A.x[3L * mem[a] + k];
    }
    double dist[P], dr[P][3];
    {
        int q = 0;
#pragma unroll
        for (int a = 0; a < NB; a++)
#pragma unroll
            for (int b = a + 1; b < NB; b++, q++) dist[q] = lmp_dist(xm[a], xm[b], dr[q]);
    }
    double Ec = 0.0, fo[NB][3], So[6];
    int flags = 0;
    bool evaluated = false, tab_err = false;
    if constexpr (TAB) {
        if (valid) tab3_cluster(M, T, ty, dist, dr, Ec, fo, So, evaluated, tab_err);
    } else {
        if (valid) explicit_cluster<NB, D>(M, ty, dist, dr, Ec, fo, So, flags, evaluated, nullptr);
    }
    if (tab_err) atomicOr(&A.status[CHX_LMP_ST_FLAGS], CHX_LMP_BAD_TAB3);
    bool bad = false;
    double v[6];
    lmp_order(So, v);
    // forces of every member, the owner's included, through the LDS table
#pragma unroll
    for (int a = 0; a < NB; a++) tab.add(evaluated ? mem[a] : -1, fo[a], bad);
    // per-atom tallies (ev_tally_mb's equal shares): members 1.. by exact
    // atomics, the owner by runs of equal owners over the wavefront
    if (A.Eacc || A.Vacc) {
        if (evaluated)
#pragma unroll
            for (int a = 1; a < NB; a++) {
                if (A.Eacc) fx_atomic(A.Eacc + 2L * mem[a], Ec / NB, bad);
                if (A.Vacc)
                    for (int k = 0; k < 6; k++) fx_atomic(A.Vacc + 2L * (6L * mem[a] + k), v[k] / NB, bad);
            }
        const int key = evaluated ? mem[0] : -1;
        if (__ballot(key >= 0)) {
            const int seg = fx_run_first<64>(key);
            const int l = threadIdx.x & 63;
            const int knext = __shfl_down(key, 1, 64);
            const bool last = key >= 0 && (l == 63 || knext != key);
            const long o = key >= 0 ? key : 0;
            if (A.Eacc) lmp_run_add(key, seg, last, Ec / NB, A.Eacc + 2L * o, bad);
            if (A.Vacc)
                for (int k = 0; k < 6; k++) lmp_run_add(key, seg, last, v[k] / NB, A.Vacc + 2L * (6L * o + k), bad);
        }
    }
    // global energy and virial: only their sum over all atoms' words is
    // read, so the wavefront's totals (fourteen integer parts summed at
    // once) are added to the words of its first evaluated cluster's owner
    if (A.Gacc) {
        const unsigned long long on = __ballot(evaluated);
        if (on) {
            const int k0 = __builtin_amdgcn_readlane(mem[0], __ffsll((long long)on) - 1);
            long long x[16];
#pragma unroll
            for (int k = 0; k < 16; k++) x[k] = 0;
            if (evaluated) {
                if (A.eflag_global) fx_split(Ec, x[0], x[1], bad);
                if (A.vflag_global)
#pragma unroll
                    for (int k = 0; k < 6; k++) fx_split(v[k], x[2 + 2 * k], x[3 + 2 * k], bad);
            }
            const long long t = fx_wave_sums16(x);
            const int l = threadIdx.x & 63;
            if (l < 14 && t) atomicAdd(A.Gacc + 14L * k0 + l, (unsigned long long)t);
        }
    }
    __syncthreads();
    tab.flush(A.Facc);
    if (bad) atomicOr(&A.status[CHX_LMP_ST_FLAGS], CHX_LMP_BAD_RANGE);
}

// ---------------------------------------------------------------------------
// Accumulators -> doubles: one thread per atom; block 0 also sums the global
// energy and virial over the owner atoms (integer sums, exact).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_lmp_out(chx_lmp_args A, double *F, double *eatom, double *vatom)
{
    const long a = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= A.nall) return;
    for (int k = 0; k < 3; k++) {
        fx_sum t;
        t.add(A.Facc + 2 * (3 * a + k));
        F[3 * a + k] = t.value();
    }
    if (A.Eacc) {
        fx_sum t;
        t.add(A.Eacc + 2 * a);
        eatom[a] = t.value();
    }
    if (A.Vacc)
        for (int k = 0; k < 6; k++) {
            fx_sum t;
            t.add(A.Vacc + 2 * (6 * a + k));
            vatom[6 * a + k] = t.value();
        }
}

// global energy and virial: each block sums a strided share of the owner
// atoms' words (integer sums, exact) and adds them into gsum[7][2]; the
// last block to finish converts them
__global__ __launch_bounds__(256) void k_lmp_global(chx_lmp_args A, unsigned long long *gsum, int *done, double *glob)
{
    __shared__ long long sh[256], sl[256];
    __shared__ bool is_last;
    const long stride = (long)gridDim.x * 256;
    for (int k = 0; k < 7; k++) {
        long long hi = 0, lo = 0;
        for (long a = (long)blockIdx.x * 256 + threadIdx.x; a < A.nall; a += stride) {
            const unsigned long long *g = A.Gacc + 2 * (7 * a + k);
            hi += (long long)g[0] + ((long long)g[1] >> 40);
            lo += (long long)(g[1] & 0xffffffffffull);
        }
        sh[threadIdx.x] = hi;
        sl[threadIdx.x] = lo;
        __syncthreads();
        for (int w = 128; w >= 1; w >>= 1) {
            if ((int)threadIdx.x < w) {
                sh[threadIdx.x] += sh[threadIdx.x + w];
                sl[threadIdx.x] += sl[threadIdx.x + w];
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            if (sh[0]) atomicAdd(gsum + 2 * k, (unsigned long long)sh[0]);
            if (sl[0]) atomicAdd(gsum + 2 * k + 1, (unsigned long long)sl[0]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        __threadfence();
        is_last = atomicAdd(done, 1) == (int)gridDim.x - 1;
    }
    __syncthreads();
    if (is_last && threadIdx.x < 7) {
        __threadfence();
        const unsigned long long h = __atomic_load_n(gsum + 2 * threadIdx.x, __ATOMIC_RELAXED);
        const unsigned long long l = __atomic_load_n(gsum + 2 * threadIdx.x + 1, __ATOMIC_RELAXED);
        fx_sum t;
        t.hi = (long long)h + ((long long)l >> 40);
        t.lo = (long long)(l & 0xffffffffffull);
        glob[threadIdx.x] = t.value();
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// Cluster rows of the FINGERPRINT build (pair_chimes.cpp:526-535, 625-640,
// 703-714, 777-790, recorded by chimesFF.cpp:1694-1703, 2013-2022,
// 2603-2613): a cluster of the pair style's loops is recorded when its
// members are in order (2-body: i < j by atom index, on pairs with tag[j] >
// tag[i]; 3-body: tag[i] < tag[j], tag[i] < tag[k], tag[j] < tag[k]; 4-body:
// tag[i] < tag[j] < tag[k] < tag[l]) and compute_nB evaluates it (type not
// excluded, every pair below its slot's outer cutoff).  Row: the pair
// distances in runtime order (get_dist's arithmetic), then the member types
// as doubles.  One thread per loop entry (2-body: per packed neighbour-row
// entry), 256 entries per block: MODE 0 counts each block's rows, MODE 1
// writes them from the block's offset in entry order, so rows come in the
// loops' order.
// ---------------------------------------------------------------------------
template <int NB, int MODE>
__global__ __launch_bounds__(256) void k_lmp_fp_rows(long n, const int *__restrict__ list, chx_lmp_args A,
                                                     chx_model_dev M, int *blk_cnt, const long long *blk_off,
                                                     double *rows)
{
    constexpr int P = NB * (NB - 1) / 2;
    constexpr int pa[6] = {0, 0, 0, 1, 1, 2}, pb[6] = {1, 2, 3, 2, 3, 3};
    constexpr int pa3[3] = {0, 0, 1}, pb3[3] = {1, 2, 2};
    __shared__ int wsum[4];
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    bool keep = false;
    int ty[NB];
    double dist[P];
    if (e < n) {
        int m[NB];
        if constexpr (NB == 2) {
            int lo = 0, hi = A.inum - 1;  // owner: the last row starting at or before e
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (A.nbr_off[mid] <= e) lo = mid; else hi = mid - 1;
            }
            m[0] = A.ilist[lo];
            m[1] = A.nbr[e];
            keep = A.tag[m[1]] > A.tag[m[0]] && m[0] < m[1];
        } else {
#pragma unroll
            for (int a = 0; a < NB; a++) m[a] = list[e * NB + a];
            const int *t = A.tag;
            if constexpr (NB == 3) keep = t[m[0]] < t[m[1]] && t[m[0]] < t[m[2]] && t[m[1]] < t[m[2]];
            else keep = t[m[0]] < t[m[1]] && t[m[1]] < t[m[2]] && t[m[2]] < t[m[3]];
        }
        if (keep) {
#pragma unroll
            for (int a = 0; a < NB; a++) ty[a] = A.ctype[m[a]];
#pragma unroll
            for (int q = 0; q < P; q++) {
                const int a = NB == 3 ? pa3[q] : pa[q], b = NB == 3 ? pb3[q] : pb[q];
                double dr[3];
                dist[q] = lmp_dist(A.x + 3L * m[a], A.x + 3L * m[b], dr);
            }
            keep = cluster_inside<NB>(M, ty, dist);
        }
    }
    const unsigned long long bal = __ballot(keep);
    const int before = __popcll(bal & (lane ? (~0ull >> (64 - lane)) : 0ull));
    if (lane == 0) wsum[wv] = __popcll(bal);
    __syncthreads();
    int base = 0, total = 0;
#pragma unroll
    for (int w = 0; w < 4; w++) {
        base += w < wv ? wsum[w] : 0;
        total += wsum[w];
    }
    if (MODE == 0) {
        if (threadIdx.x == 0) blk_cnt[blockIdx.x] = total;
        return;
    }
    if (!keep) return;
    double *row = rows + (blk_off[blockIdx.x] + base + before) * (P + NB);
#pragma unroll
    for (int q = 0; q < P; q++) row[q] = dist[q];
#pragma unroll
    for (int a = 0; a < NB; a++) row[P + a] = (double)ty[a];
}

void launch_lmp_fp_rows(int nb, int mode, long n, const int *list, const chx_lmp_args &A, const chx_model_dev &M,
                        int *blk_cnt, const long long *blk_off, double *rows, hipStream_t st)
{
    if (n <= 0) return;
    const dim3 grid((unsigned)((n + 255) / 256)), block(256);
#define L(NB_, MODE_) \
    hipLaunchKernelGGL((k_lmp_fp_rows<NB_, MODE_>), grid, block, 0, st, n, list, A, M, blk_cnt, blk_off, rows)
    if (nb == 2) { if (mode) L(2, 1); else L(2, 0); }
    else if (nb == 3) { if (mode) L(3, 1); else L(3, 0); }
    else { if (mode) L(4, 1); else L(4, 0); }
#undef L
}

void launch_lmp_pair(const chx_lmp_args &A, const chx_model_dev &M, const chx_tabs_dev *T, hipStream_t st)
{
    if (A.inum <= 0) return;
    const dim3 grid((unsigned)((A.inum + 15) / 16)), block(256);
    if (T && T->ntab2) hipLaunchKernelGGL(k_lmp_pair<true>, grid, block, 0, st, A, M, *T);
    else hipLaunchKernelGGL(k_lmp_pair<false>, grid, block, 0, st, A, M, chx_tabs_dev{});
}

void launch_lmp_lists(int mode, const chx_lmp_args &A, double pad3, double pad4, int has4, int *cnt3, int *cnt4,
                      const long long *off3, const long long *off4, int *trips, int *quads, hipStream_t st)
{
    if (A.inum <= 0) return;
    if (mode == 0)
        hipLaunchKernelGGL(k_lmp_lists<0>, dim3((unsigned)A.inum), dim3(64), 0, st, A, pad3, pad4, has4, cnt3,
                           cnt4, off3, off4, trips, quads);
    else
        hipLaunchKernelGGL(k_lmp_lists<1>, dim3((unsigned)A.inum), dim3(64), 0, st, A, pad3, pad4, has4, cnt3,
                           cnt4, off3, off4, trips, quads);
}

void launch_lmp_clusters(int nb, int D, long n, const int *list, const chx_lmp_args &A, const chx_model_dev &M,
                         const chx_tabs_dev *T, hipStream_t st)
{
    if (n <= 0) return;
    const dim3 grid((unsigned)((n + 127) / 128)), block(128);
    if (nb == 3 && T && T->ntab3) {
        hipLaunchKernelGGL((k_lmp_clusters<3, 0, true>), grid, block, 0, st, n, list, A, M, *T);
        return;
    }
#define X(d)                                                                                         \
    if (D == d) {                                                                                    \
        hipLaunchKernelGGL((k_lmp_clusters<3, d, false>), grid, block, 0, st, n, list, A, M, chx_tabs_dev{}); \
        return;                                                                                      \
    }
    if (nb == 3) { CHX_D3_CASES(X) }
#undef X
#define X(d)                                                                                         \
    if (D == d) {                                                                                    \
        hipLaunchKernelGGL((k_lmp_clusters<4, d, false>), grid, block, 0, st, n, list, A, M, chx_tabs_dev{}); \
        return;                                                                                      \
    }
    if (nb == 4) { CHX_D4_CASES(X) }
#undef X
}

void launch_lmp_out(const chx_lmp_args &A, double *F, double *eatom, double *vatom, double *glob,
                    unsigned long long *gsum, int *done, hipStream_t st)
{
    if (A.nall > 0)
        hipLaunchKernelGGL(k_lmp_out, dim3((unsigned)((A.nall + 255) / 256)), dim3(256), 0, st, A, F, eatom, vatom);
    if (A.Gacc) {
        const unsigned nb = (unsigned)std::min<long>(256, std::max<long>(1, ((long)A.nall + 1023) / 1024));
        hipLaunchKernelGGL(k_lmp_global, dim3(nb), dim3(256), 0, st, A, gsum, done, glob);
    }
}

}  // namespace chx
