// Evaluation kernels: 2-, 3- and 4-body cluster terms with exact fixed-point
// accumulation of forces (per replicated atom) and of energy and virial (per
// owner atom), the final per-atom / per-configuration totals, and the
// explicit-cluster kernel behind the per-cluster API.
//
// 3-/4-body lists are laid out in segments of one (model, cluster type) group
// starting on multiples of CHX_LIST_ALIGN, so every 64-cluster tile uses one
// coefficient table.  The 4-body contraction for dense dimensions D <= 4 runs
// on the matrix cores: with X = T_01 (x) T_02 (x) T_03 and Y = T_12 (x) T_13 (x)
// T_23 per cluster, the energy form is S = X^T C Y for the coefficient matrix
// C[(a,b,c)][(d,e,f)], and the gradients take X with one factor replaced by
// its derivative (three more products) or Y likewise (elementwise).  One
// wavefront evaluates 16 clusters per step as Z_v = C^T X_v (v = energy form
// and the three X gradients) with 16x16x4 MFMA, then folds Z_v against Y.
//
// Compiled with -ffp-contract=off: the cluster geometry (dr = h (v_b - v_a))
// reproduces the reference's get_dist bit for bit; the polynomial and
// coefficient arithmetic uses explicit fma().
#include <hip/hip_runtime.h>

#include "chx_dev.h"
#include "chx_engine.h"
#include "chx_eval.h"
#include "eval_math.h"
#include "chx_cluster.h"

#include <algorithm>

namespace chx {

typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef float flt4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void mv3(const double *m, const double *d, double *o)
{
    o[0] = m[0] * d[0] + m[1] * d[1] + m[2] * d[2];
    o[1] = m[3] * d[0] + m[4] * d[1] + m[5] * d[2];
    o[2] = m[6] * d[0] + m[7] * d[1] + m[8] * d[2];
}

// dr = h (b - a) and its length (serial_chimes_interface.h:176-208)
__device__ __forceinline__ double pair_vec(const double *h, const double *a, const double *b, double *dr)
{
    const double d[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    mv3(h, d, dr);
    return sqrt(dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2]);
}

__device__ __forceinline__ void note_flags(int *cfg_stat, int cfg, int flags)
{
    if (!flags) return;
    atomicMax(&cfg_stat[4 * cfg + 0], (flags & FLAG_BELOW_RIN) ? 2 : ((flags & FLAG_PENALTY) ? 1 : 0));
    if (flags & FLAG_PENALTY) atomicAdd(&cfg_stat[4 * cfg + 1], 1);
    if (flags & FLAG_OUT_OF_RANGE) atomicAdd(&cfg_stat[4 * cfg + 2], 1);
}

// Group of list slot t: the last group whose aligned base is <= t.
__device__ __forceinline__ int find_group(const int *__restrict__ gbase, int G, long t)
{
    int lo = 0, hi = G - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (gbase[mid] <= t) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// ---------------------------------------------------------------------------
// 2-body: LPO lanes per owner atom over its image slab (64 / LPO owners per
// wavefront; OWN owners per workgroup, a small workgroup for batches too
// small to fill the device with the large one).  F_owner += f dr, F_image -=
// f dr, virial -= f dr dr (compute_2B, chimesFF.cpp:1657-1775).  The owner's
// force, energy and virial are summed in double in a fixed order (lane l of
// the LPO takes images l, l + LPO, ... of the slab, which is sorted by image
// key, and row shifts add the lanes), so the sum depends only on the owner's
// own images, never on the batch around it; the force is then added exactly
// per owner, the energy and virial exactly per configuration.  Image forces
// are merged in the workgroup's LDS table.
// ---------------------------------------------------------------------------
template <int OWN, int LPO>
__global__ __launch_bounds__(LPO * OWN) __attribute__((amdgpu_waves_per_eu(3))) void k_eval2(int natoms, int nb_cap, const int *__restrict__ nbr_cnt,
                                               const chx_cfg *__restrict__ cfgs, const chx_atom_s *__restrict__ as,
                                               const chx_entry_s *__restrict__ es,
                                               const chx_model_dev *__restrict__ models, unsigned long long *Facc,
                                               unsigned long long *Cacc, int *cfg_stat, int *status)
{
    __shared__ fx_table<LPO * OWN <= 64 ? 256 : 512> tab;  // image side of the workgroup's owners
    // the pair tables (slots, coefficients, pair map) of the model of the
    // workgroup's first owner, staged in LDS when they are small; lanes whose
    // owner uses another model read them from memory
    constexpr int SP = 12, SC = 320;
    __shared__ chx_slot sl_s[SP];
    __shared__ double c2_s[SC];
    __shared__ int pm_s[SP * 2];
    // status read overlapped with the table clear; one barrier for both
    __shared__ int flags_s, m_s;
    if (threadIdx.x == 0) {
        flags_s = __atomic_load_n(&status[CHX_ST_FLAGS], __ATOMIC_RELAXED);
        const int i0 = (int)(blockIdx.x * OWN);
        int m = -1;
        if (i0 < natoms) {
            m = cfgs[as[i0].cfg].model;
            const chx_model_dev &M0 = models[m];
            if (M0.npairs > SP || M0.npairs * (M0.order[0] + 1) > SC || M0.nt * M0.nt > 2 * SP) m = -1;
        }
        m_s = m;
    }
    tab.clear();
    __syncthreads();
    if (flags_s & CHX_STOP) return;
    const int ms = m_s;
    if (ms >= 0) {
        const chx_model_dev &M0 = models[ms];
        const int nc = M0.npairs * (M0.order[0] + 1);
        for (int t = threadIdx.x; t < nc; t += LPO * OWN) c2_s[t] = M0.c2[t];
        for (int t = threadIdx.x; t < M0.npairs; t += LPO * OWN) sl_s[t] = M0.slots2[t];
        for (int t = threadIdx.x; t < M0.nt * M0.nt; t += LPO * OWN) pm_s[t] = M0.pairmap[t];
    }
    __syncthreads();
    const int sub = threadIdx.x & (LPO - 1);
    const int i = (int)(blockIdx.x * OWN + threadIdx.x / LPO);
    double acc[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};  // owner force, energy, virial
    bool bad = false;
    int ckey = -1;  // configuration of the owner, on its last lane
    if (i < natoms) {
        const int n = nbr_cnt[i];
        const chx_atom_s ai = as[i];
        const chx_cfg &C = cfgs[ai.cfg];
        const chx_model_dev &M = models[C.model];
        int flags = 0;
        // the owner's images, with the pair tables at pm / sl / c2
        auto images = [&](const int *pm, const chx_slot *sl, const double *c2) {
            // one image: energy, force scalar and pair vector
            auto pair = [&](const chx_entry_s &g, double &E, double &fs, double (&dr)[3]) {
                const double r = pair_vec(C.h, ai.u, g.v, dr);
                const int pt = pm[ai.type * M.nt + g.type];
                eval_pair(r, sl[pt], c2 + (long)pt * (M.order[0] + 1), M.order[0], M.tersoff, M.fcut_var,
                          M.pen_dist, M.pen_scale, E, fs, flags);
            };
            // its contributions, added in slab order
            auto add = [&](const chx_entry_s &g, double E, double fs, const double (&dr)[3]) {
                const double F0[3] = {fs * dr[0], fs * dr[1], fs * dr[2]};
                const double F1[3] = {-F0[0], -F0[1], -F0[2]};
                tab.add_or_global(g.parent, F1, Facc, bad);
                acc[0] += F0[0];
                acc[1] += F0[1];
                acc[2] += F0[2];
                acc[3] += E;
                acc[4] -= fs * (dr[0] * dr[0]);
                acc[5] -= fs * (dr[0] * dr[1]);
                acc[6] -= fs * (dr[0] * dr[2]);
                acc[7] -= fs * (dr[1] * dr[1]);
                acc[8] -= fs * (dr[1] * dr[2]);
                acc[9] -= fs * (dr[2] * dr[2]);
            };
            // two images per pass, evaluated side by side (independent
            // latency chains) and added in slab order, j then j + LPO, so
            // the sums are those of one image per pass
            const chx_entry_s *slab = es + (long)i * nb_cap;
            for (int j = sub; j < n; j += 2 * LPO) {
                const bool two = j + LPO < n;
                const chx_entry_s g0 = slab[j], g1 = slab[two ? j + LPO : j];
                double E0, fs0, dr0[3], E1, fs1, dr1[3];
                pair(g0, E0, fs0, dr0);
                pair(g1, E1, fs1, dr1);
                add(g0, E0, fs0, dr0);
                if (two) add(g1, E1, fs1, dr1);
            }
        };
        if (C.model == ms) images(pm_s, sl_s, c2_s);
        else images(M.pairmap, M.slots2, M.c2);
        note_flags(cfg_stat, ai.cfg, flags);
        if (sub == LPO - 1) ckey = ai.cfg;
    }
    // the LPO lanes of each owner summed in double, in a fixed order
    // (inclusive sums by row shifts: the owner's last lane holds the total)
#pragma unroll
    for (int k = 0; k < 10; k++) acc[k] = group_sum<LPO>(acc[k]);
    if (sub == LPO - 1 && i < natoms)
#pragma unroll
        for (int k = 0; k < 3; k++) fx_atomic(Facc + 2L * (3L * i + k), acc[k], bad);
    // energy and virial per configuration, exact over the wavefront's owners
    double q7[7];
#pragma unroll
    for (int k = 0; k < 7; k++) q7[k] = acc[3 + k];
    fx_cfg_sums(Cacc, ckey, blockIdx.x & (CHX_CACC_SLOTS - 1), q7, bad);
    __syncthreads();
    tab.flush(Facc);
    if (bad) atomicOr(&status[CHX_ST_FLAGS], CHX_BAD_RANGE);
}

// Member positions, types and replicated atom indices of a cluster record.
template <int NB>
__device__ __forceinline__ void load_members(const int2 rec, const chx_atom_s *__restrict__ as,
                                             int nb_cap, const chx_entry_s *__restrict__ es,
                                             double (&p)[NB][3], int (&t)[NB], int (&par)[NB], int &cfg)
{
    const int i = rec.x;
    const chx_atom_s ai = as[i];
    cfg = ai.cfg;
    p[0][0] = ai.u[0]; p[0][1] = ai.u[1]; p[0][2] = ai.u[2];
    t[0] = ai.type;
    par[0] = i;
    const long off = (long)i * nb_cap - 1;  // owner i's slab
#pragma unroll
    for (int m = 1; m < NB; m++) {
        const int s = (rec.y >> (10 * (m - 1))) & 1023;
        const chx_entry_s g = es[off + s];
        p[m][0] = g.v[0]; p[m][1] = g.v[1]; p[m][2] = g.v[2];
        t[m] = g.type;
        par[m] = g.parent;
    }
}

// Reorders the members into parameter-file atom order: canonical member a is
// runtime member perm[a] (the first matching permutation, chimesFF.h:480-526).
template <int NB>
__device__ __forceinline__ void canonical(const int *__restrict__ perm_tab, int tup, double (&p)[NB][3],
                                          int (&par)[NB], int (&pm)[NB])
{
#pragma unroll
    for (int a = 0; a < NB; a++) pm[a] = perm_tab[tup * NB + a];
    double q[NB][3];
    int c[NB];
#pragma unroll
    for (int a = 0; a < NB; a++) {
        c[a] = par[0];
        q[a][0] = p[0][0]; q[a][1] = p[0][1]; q[a][2] = p[0][2];
#pragma unroll
        for (int b = 1; b < NB; b++) {
            const bool s = pm[a] == b;
            c[a] = s ? par[b] : c[a];
            q[a][0] = s ? p[b][0] : q[a][0];
            q[a][1] = s ? p[b][1] : q[a][1];
            q[a][2] = s ? p[b][2] : q[a][2];
        }
    }
#pragma unroll
    for (int a = 0; a < NB; a++) {
        par[a] = c[a];
        p[a][0] = q[a][0]; p[a][1] = q[a][1]; p[a][2] = q[a][2];
    }
}

template <int NB>
struct pairs_of;
template <>
struct pairs_of<3> {
    static constexpr int P = 3;
    static constexpr int a[3] = {0, 0, 1};
    static constexpr int b[3] = {1, 2, 2};
};
template <>
struct pairs_of<4> {
    static constexpr int P = 6;
    static constexpr int a[6] = {0, 0, 0, 1, 1, 2};
    static constexpr int b[6] = {1, 2, 3, 2, 3, 3};
};

// Cluster geometry in canonical order: members, pair vectors and distances,
// and whether every pair is inside its cutoff; canonical member a is runtime
// member pm[a] (runtime member 0 is the owner).
template <int NB>
__device__ __forceinline__ bool cluster_geometry(const chx_group &grp, const int2 rec,
                                                 const chx_cfg *__restrict__ cfgs,
                                                 const chx_atom_s *__restrict__ as, int nb_cap,
                                                 const chx_entry_s *__restrict__ es, double (&p)[NB][3],
                                                 int (&par)[NB], double (&r)[pairs_of<NB>::P], int &cfg,
                                                 int (&pm)[NB], double (&dr)[pairs_of<NB>::P][3])
{
    using PQ = pairs_of<NB>;
    int t[NB];
    load_members<NB>(rec, as, nb_cap, es, p, t, par, cfg);
    int tup = 0;
#pragma unroll
    for (int a = 0; a < NB; a++) tup = tup * grp.nt + t[a];
    canonical<NB>(grp.perm, tup, p, par, pm);
    const chx_cfg &C = cfgs[cfg];
    bool in = true;
#pragma unroll
    for (int q = 0; q < PQ::P; q++) {
        r[q] = pair_vec(C.h, p[PQ::a[q]], p[PQ::b[q]], dr[q]);
        in = in && (r[q] < grp.s[q].rout);
    }
    return in;
}

// The same without the pair vectors.
template <int NB>
__device__ __forceinline__ bool cluster_geometry(const chx_group &grp, const int2 rec,
                                                 const chx_cfg *__restrict__ cfgs,
                                                 const chx_atom_s *__restrict__ as, int nb_cap,
                                                 const chx_entry_s *__restrict__ es, double (&p)[NB][3],
                                                 int (&par)[NB], double (&r)[pairs_of<NB>::P], int &cfg,
                                                 int (&pm)[NB])
{
    double dr[pairs_of<NB>::P][3];
    return cluster_geometry<NB>(grp, rec, cfgs, as, nb_cap, es, p, par, r, cfg, pm, dr);
}

// Forces of one cluster from its pair force scalars: F_a += fs dr, F_b -= fs
// dr for pair (a, b) of canonical members; virial -= fs dr dr into the q7
// owner accumulator slots 1..6.
template <int NB>
__device__ __forceinline__ void cluster_forces(const double (&drq)[pairs_of<NB>::P][3],
                                               const double (&fs)[pairs_of<NB>::P], double (&F)[NB][3],
                                               double (&q7)[7])
{
    using PQ = pairs_of<NB>;
#pragma unroll
    for (int m = 0; m < NB; m++) F[m][0] = F[m][1] = F[m][2] = 0.0;
#pragma unroll
    for (int q = 0; q < PQ::P; q++) {
        const double (&dr)[3] = drq[q];
        const double f = fs[q];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            F[PQ::a[q]][k] += f * dr[k];
            F[PQ::b[q]][k] -= f * dr[k];
        }
        q7[1] -= f * (dr[0] * dr[0]);
        q7[2] -= f * (dr[0] * dr[1]);
        q7[3] -= f * (dr[0] * dr[2]);
        q7[4] -= f * (dr[1] * dr[1]);
        q7[5] -= f * (dr[1] * dr[2]);
        q7[6] -= f * (dr[2] * dr[2]);
    }
}

// ---------------------------------------------------------------------------
// 3- and 4-body on the vector units: one thread per cluster, one wavefront per
// 64-cluster tile of one group (any supported D); the workgroup's member
// forces are merged per atom in an LDS table before they reach memory.
// Occupancy targets, measured: the FP64 4-body kernels for D <= 3 are held to
// two wavefronts per SIMD (left alone the compiler gives them one; at two
// they need no scratch); to three: the FP64 3-body kernels for D <= 8, the
// FP32 3-body kernels for D <= 10 and the FP32 4-body kernels for D <= 4 (at
// most 84 bytes of scratch, faster end to end); to four: the FP32 3-body
// kernels for D <= 7 (at most 128 registers, no scratch; 4% faster than
// three).
// ---------------------------------------------------------------------------
template <int NB, int D, typename R>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(NB == 3 && D <= 7 && sizeof(R) == 4 ? 4 : NB == 4 && D <= 3 && sizeof(R) == 8 ? 2 : ((NB == 3 && D <= (sizeof(R) == 8 ? 8 : 10)) || (NB == 4 && D <= (sizeof(R) == 8 ? 3 : 4)) ? 3 : 1)))) void k_evalv(int G, const int *__restrict__ gbase, const int *__restrict__ gend,
                                                const chx_group *__restrict__ groups, int total_word,
                                                const int2 *__restrict__ list, const chx_cfg *__restrict__ cfgs,
                                                const chx_atom_s *__restrict__ as, int nb_cap,
                                                const chx_entry_s *__restrict__ es, unsigned long long *Facc,
                                                unsigned long long *Cacc, int *cfg_stat, int *status)
{
    using PQ = pairs_of<NB>;
    constexpr int P = PQ::P;
    // the workgroup's member forces merged per atom; the table holds the
    // distinct atoms of a typical tile (a few hundred) with room to spare,
    // and anything past it goes straight to memory, so the workgroup stays
    // small enough for five per CU
    __shared__ fx_table<512> tab;
    // status read overlapped with the table clear; one barrier for both
    __shared__ int flags_s;
    if (threadIdx.x == 0) flags_s = __atomic_load_n(&status[CHX_ST_FLAGS], __ATOMIC_RELAXED);
    tab.clear();
    __syncthreads();
    if (flags_s & CHX_STOP) return;
    const int lane = threadIdx.x & 63;
    // wave index made explicitly uniform, so the group, its slots and its
    // coefficients are read through the scalar unit
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const long t0 = ((long)blockIdx.x * 4 + wave) * 64;
    bool wave_on = t0 < status[total_word];
    const int g = wave_on ? find_group(gbase, G, t0) : 0;
    const chx_group &grp = groups[g];
    wave_on = wave_on && grp.type >= 0 && grp.D == D;
    const long c = t0 + lane;
    int ckey = -1;
    double q7[7] = {0, 0, 0, 0, 0, 0, 0};
    double Fr[NB][3];   // force on each member (canonical order)
    int key[NB];        // replicated atom of each member, -1 none
#pragma unroll
    for (int m = 0; m < NB; m++) { key[m] = -1; Fr[m][0] = Fr[m][1] = Fr[m][2] = 0.0; }
    bool bad = false;
    if (wave_on && c < gend[g]) {
        const int2 rec = list[c];
        double p[NB][3], r[P], dr[P][3];
        int par[NB], cfg, pm[NB];
        if (cluster_geometry<NB>(grp, rec, cfgs, as, nb_cap, es, p, par, r, cfg, pm, dr)) {
            R T[P][D], dT[P][D];
            double f[P], fd[P];
            int flags = 0;
#pragma unroll
            for (int q = 0; q < P; q++) {
                cheby<D, R>(r[q], grp.s[q], T[q], dT[q], flags);
                fcut(r[q], grp.s[q], grp.tersoff, grp.fcut_var, f[q], fd[q]);
            }
            const R *Cd = sizeof(R) == 8 ? (const R *)grp.C : (const R *)grp.Cf;
            double S, Gq[P], E, fs[P];
            if constexpr (NB == 3) contract3<D, R, double>(Cd, T, dT, S, Gq);
            else contract4<D, R, double>(Cd, T, dT, S, Gq);
            cluster_scalars<P, double>(r, f, fd, S, Gq, E, fs);
            q7[0] = E;
            double F[NB][3];
            cluster_forces<NB>(dr, fs, F, q7);
#pragma unroll
            for (int a = 0; a < NB; a++) {
                key[a] = par[a];
                Fr[a][0] = F[a][0]; Fr[a][1] = F[a][1]; Fr[a][2] = F[a][2];
            }
            ckey = cfg;
            note_flags(cfg_stat, cfg, flags);
        }
    }
    // every member's force through the LDS table (the owner's too: its
    // lanes' LDS atomics cost fewer instructions than six wavefront scans);
    // energy and virial go to the configuration
#pragma unroll
    for (int m = 0; m < NB; m++) tab.add_or_global(key[m], Fr[m], Facc, bad);
    fx_cfg_sums(Cacc, ckey, blockIdx.x & (CHX_CACC_SLOTS - 1), q7, bad);
    __syncthreads();
    tab.flush(Facc);
    if (bad) atomicOr(&status[CHX_ST_FLAGS], CHX_BAD_RANGE);
}

// ---------------------------------------------------------------------------
// 4-body on the matrix cores (D <= 4).  Workgroup = 4 wavefronts and one
// 64-cluster tile of one group (16 clusters per wavefront); the group's
// coefficient matrix is staged in LDS.  Per wavefront: (1) the four lanes of
// a cluster compute its geometry, lane sub the polynomials of pairs sub and
// sub + 4, staged in LDS; (2) MFMA contraction; (3) lane sub accumulates the
// force on member sub and the virial of pairs sub and sub + 4.
//
// Matrix image (built on the host): Cm[k][m], k = (a*D + b)*4 + c over pairs
// 01, 02, 03 and m = (d*D + e)*4 + f over pairs 12, 13, 23, the innermost
// index padded to 4 and m to a multiple of 16; padding is zero.
//
// Lane l of a wavefront: cluster j = l % 16 of the wavefront's 16, sub = l / 16.
// MFMA 16x16x4 operands: A[m][k] = Cm[4s + sub][16 rb + j] (LDS), B[k][n] =
// X_v[cluster j][4s + sub], i.e. c = sub and (a, b) = (s / D, s % D).  Output
// row m of lane l, register v: FP64 m = sub + 4v (f = sub, (d, e) from
// 4 rb + v); FP32 m = 4 sub + v (f = v, (d, e) from 4 rb + sub).
// ---------------------------------------------------------------------------
template <typename R>
struct mfma_acc;
template <>
struct mfma_acc<double> {
    typedef dbl4 type;
    static constexpr bool f64 = true;
    __device__ __forceinline__ static dbl4 op(double a, double b, dbl4 c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};
template <>
struct mfma_acc<float> {
    typedef flt4 type;
    static constexpr bool f64 = false;
    __device__ __forceinline__ static flt4 op(float a, float b, flt4 c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
};

// LDS layout of a KP x MP matrix image read as A operands (16 consecutive
// columns by each group of 16 lanes, four consecutive rows per wavefront):
// rows whose stride is a multiple of 64 elements get their 16-column blocks
// XOR-swizzled by (row & 3); other strides are padded to 16 or 48 mod 64, so
// the four rows of one read land in different banks.
template <int KP, int MP>
struct lds_mat {
    static constexpr bool swz = MP % 64 == 0;
    static constexpr int LD = swz ? MP : ((MP % 64 == 16 || MP % 64 == 48) ? MP : MP + (64 + 16 - MP % 64) % 64);
    static constexpr int size = KP * LD;
    __device__ __forceinline__ static int at(int k, int m) { return swz ? k * MP + (m ^ ((k & 3) << 4)) : k * LD + m; }
};

// Square P x P image (P a multiple of 16) read both by rows and by columns:
// in a 16-lane group either the column (16 consecutive m, one k) or the row
// (16 consecutive k, one m) varies, and XOR-ing m's low four bits with k's
// keeps both patterns on distinct banks; for P a multiple of 64 the k & 3
// part also moves the four rows of one MFMA k-group to different 16-column
// blocks.
template <int P>
struct lds_sq {
    static_assert(P % 16 == 0, "square image size must be a multiple of 16");
    static constexpr int size = P * P;
    __device__ __forceinline__ static int at(int k, int m)
    {
        return k * P + (m ^ ((P % 64 == 0 ? (k & 3) << 4 : 0) ^ (k & 15)));
    }
};

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

// Round-to-nearest-even FP32 -> BF16 bits.
__device__ __forceinline__ unsigned short bf16_bits(float f)
{
    const unsigned u = __float_as_uint(f);
    return (unsigned short)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

// Eight FP32 values as one BF16 MFMA operand (k = 8 (lane / 16) + q).
__device__ __forceinline__ bf16x8 pack_bf16(const float (&x)[8])
{
    bf16x8 r;
#pragma unroll
    for (int q = 0; q < 8; q++) r[q] = __builtin_bit_cast(__bf16, bf16_bits(x[q]));
    return r;
}

// Split of eight FP32 values into BF16 high parts and BF16 residuals
// (x ~ hi + lo to about 16 significant bits).
__device__ __forceinline__ void split_bf16(const float (&x)[8], bf16x8 &hi, bf16x8 &lo)
{
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const unsigned short h = bf16_bits(x[q]);
        hi[q] = __builtin_bit_cast(__bf16, h);
        lo[q] = __builtin_bit_cast(__bf16, bf16_bits(x[q] - __uint_as_float((unsigned)h << 16)));
    }
}

__device__ __forceinline__ flt4 mfma_bf16(const bf16x8 &a, const bf16x8 &b, flt4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// LDS layout of a transposed BF16 image [MP][K32]: rows of K32 + 8 halves
// (16 B of padding, so the 16 columns read by a lane group fall in different
// banks), 16-byte aligned for the 8-value operand reads.
template <int MP, int K32>
struct lds_matb {
    static constexpr int LD = K32 + 8;
    static constexpr int size = MP * LD;
};

template <int D>
struct mfma_shape {
    static constexpr int KP = D * D * 4;                // rows of Cm (k)
    static constexpr int MP = (D * D * 4 + 15) / 16 * 16;  // columns of Cm (m)
    static constexpr int KS = D * D;                    // k steps of 4
    static constexpr int RB = MP / 16;                  // 16-row output blocks
};

// A block's share of a coefficient image on its way to LDS: every thread
// issues all of its loads first and stores them later, so that the loads are
// in flight together and overlap the cluster geometry.
template <typename R, int N, int NT, int NSRC = N>
struct mat_stage {
    static constexpr int PER = (N + NT - 1) / NT;
    R v[PER];
    // src holds the first NSRC of the N elements; the rest are zeros
    __device__ __forceinline__ void load(const R *__restrict__ src)
    {
#pragma unroll
        for (int u = 0; u < PER; u++) {
            const int x = (int)threadIdx.x + u * NT;
            v[u] = x < NSRC ? src[x] : (R)0;
        }
    }
    template <class AT>
    __device__ __forceinline__ void store(AT at)
    {
#pragma unroll
        for (int u = 0; u < PER; u++) {
            const int x = (int)threadIdx.x + u * NT;
            if (x < N) at(x) = v[u];
        }
    }
};

template <int D, typename R, int WPB, int BF>
__global__ __launch_bounds__(64 * WPB, 2) void k_eval4m(int G, const int *__restrict__ gbase,
                                                    const int *__restrict__ gend,
                                                    const chx_group *__restrict__ groups,
                                                    const int2 *__restrict__ list, const chx_cfg *__restrict__ cfgs,
                                                    const chx_atom_s *__restrict__ as,
                                                    int nb_cap,
                                                    const chx_entry_s *__restrict__ es, unsigned long long *Facc,
                                                    unsigned long long *Cacc, int *cfg_stat, int *status)
{
    using SH = mfma_shape<D>;
    using MA = mfma_acc<R>;
    typedef typename MA::type acc_t;
    using LQ = lds_sq<SH::MP>;
    constexpr int K32 = (SH::KP + 31) / 32 * 32;
    using LB = lds_matb<SH::MP, K32>;
    constexpr int TS = 6 * D + 1;     // per-cluster stride of the staged polynomials (odd: conflict-free)
    // matrix image in LDS: R[MP][MP] (rows k >= KP zero, swizzled), or BF16
    // transposed [MP][K32 + 8]; BF = 1: BF16 image; BF = 3: BF16 high and
    // residual images (three-product split)
    __shared__ __attribute__((aligned(16))) unsigned short ClRaw[BF ? (BF == 3 ? 2 : 1) * LB::size
                                                                     : (int)(LQ::size * sizeof(R) / 2)];
    R *Cl = (R *)ClRaw;
    // staged polynomials; after the contraction the same memory holds the force table
    union stage_t {
        struct { R T[WPB][16 * TS], dT[WPB][16 * TS]; } p;
        fx_table<256> tab;
    };
    __shared__ stage_t st;
    R (&Ts)[WPB][16 * TS] = st.p.T;
    R (&dTs)[WPB][16 * TS] = st.p.dT;
    __shared__ double Gm[WPB][16][6][3];  // r, fcut, fcut' per pair
    __shared__ double Pm[WPB][16][4][3];  // member positions (evaluation-frame fractional)
    if (block_stop(status, CHX_STOP)) return;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int j = lane & 15, sub = lane >> 4;
    const long nbt = status[CHX_ST_N4] / (16 * WPB);
    {
        const long bt = blockIdx.x;
        if (bt >= nbt) return;
        const long t0 = bt * (16 * WPB);
        const int g = find_group(gbase, G, t0);
        const chx_group &grp = groups[g];
        if (grp.type < 0 || grp.D != D) return;
        // rows KP .. MP - 1 of the square image read as zeros (src holds KP x MP)
        mat_stage<R, BF ? 1 : SH::MP * SH::MP, 64 * WPB, BF ? 1 : SH::KP * SH::MP> cm;
        if constexpr (BF) {
            for (int x = threadIdx.x; x < (BF == 3 ? 2 : 1) * SH::MP * K32; x += 64 * WPB) {
                const int img = x / (SH::MP * K32), y = x % (SH::MP * K32);
                ClRaw[img * LB::size + (y / K32) * LB::LD + y % K32] = grp.Cmb[x];
            }
        } else {
            cm.load(sizeof(R) == 8 ? (const R *)grp.Cm : (const R *)grp.Cmf);
        }
        // ---- stage: geometry (lane sub loads runtime member sub and computes
        // pairs sub and sub + 4), polynomials of those pairs
        const long c = t0 + wv * 16 + j;
        bool act = c < gend[g];
        int2 rec = make_int2(-1, 0);
        int cfg = 0, flags = 0;
        int a_s = 0, key_s = -1;  // canonical index and replicated atom of runtime member sub
        {
            double xr[3] = {0.0, 0.0, 0.0};
            int ty = 0;
            if (act) {
                rec = list[c];
                const chx_atom_s ai = as[rec.x];
                cfg = ai.cfg;
                if (sub == 0) {
                    xr[0] = ai.u[0]; xr[1] = ai.u[1]; xr[2] = ai.u[2];
                    ty = ai.type;
                    key_s = rec.x;
                } else {
                    const int slot = (rec.y >> (10 * (sub - 1))) & 1023;
                    const chx_entry_s e = es[(long)rec.x * nb_cap - 1 + slot];
                    xr[0] = e.v[0]; xr[1] = e.v[1]; xr[2] = e.v[2];
                    ty = e.type;
                    key_s = e.parent;
                }
            }
            if constexpr (!BF) cm.store([&](int x) -> R & { return Cl[LQ::at(x / SH::MP, x % SH::MP)]; });
            // the cluster's type tuple from its four lanes, and the canonical
            // slot of this lane's member (the first matching permutation)
            int tup = 0;
#pragma unroll
            for (int m = 0; m < 4; m++) tup = tup * grp.nt + __shfl(ty, j + 16 * m, 64);
            if (act)
#pragma unroll
                for (int a = 0; a < 4; a++)
                    if (grp.perm[tup * 4 + a] == sub) a_s = a;
            if (act) {
                Pm[wv][j][a_s][0] = xr[0];
                Pm[wv][j][a_s][1] = xr[1];
                Pm[wv][j][a_s][2] = xr[2];
            }
            __syncthreads();
            // pairs sub and sub + 4 in canonical order; the cluster is in
            // when all six of its pairs are inside their cutoffs
            const double *h = cfgs[cfg].h;
            double rp[2] = {0.0, 0.0};
            bool in = true;
#pragma unroll
            for (int u = 0; u < 2; u++) {
                const int q = sub + 4 * u;
                if (q < 6 && act) {
                    double dr[3];
                    rp[u] = pair_vec(h, Pm[wv][j][pairs_of<4>::a[q]], Pm[wv][j][pairs_of<4>::b[q]], dr);
                    in = in && rp[u] < grp.s[q].rout;
                }
            }
            const unsigned long long okm = __ballot(in) >> j;
            act = act && (okm & 0x0001000100010001ull) == 0x0001000100010001ull;
            for (int q = sub; q < 6; q += 4) {
                const double rq = q == sub ? rp[0] : rp[1];
                R Tq[D], dTq[D];
                double fq = 0.0, fdq = 0.0;
                if (act) {
                    const chx_slot sl = grp.s[q];
                    cheby<D, R>(rq, sl, Tq, dTq, flags);
                    fcut(rq, sl, grp.tersoff, grp.fcut_var, fq, fdq);
                } else {
#pragma unroll
                    for (int x = 0; x < D; x++) { Tq[x] = (R)0; dTq[x] = (R)0; }
                }
#pragma unroll
                for (int x = 0; x < D; x++) {
                    Ts[wv][j * TS + q * D + x] = Tq[x];
                    dTs[wv][j * TS + q * D + x] = dTq[x];
                }
                Gm[wv][j][q][0] = rq;
                Gm[wv][j][q][1] = fq;
                Gm[wv][j][q][2] = fdq;
            }
        }
        __syncthreads();
        // ---- contraction on the matrix cores
        const R *Tj = &Ts[wv][j * TS], *dTj = &dTs[wv][j * TS];
        R T0[D], dT0[D], T1[D], dT1[D], T3[D], dT3[D], T4[D], dT4[D], T5[D], dT5[D];
#pragma unroll
        for (int x = 0; x < D; x++) {
            T0[x] = Tj[0 * D + x]; dT0[x] = dTj[0 * D + x];
            T1[x] = Tj[1 * D + x]; dT1[x] = dTj[1 * D + x];
            T3[x] = Tj[3 * D + x]; dT3[x] = dTj[3 * D + x];
            T4[x] = Tj[4 * D + x]; dT4[x] = dTj[4 * D + x];
            T5[x] = Tj[5 * D + x]; dT5[x] = dTj[5 * D + x];
        }
        // lane-selected factors: c = sub (pair 03), and f = sub (pair 23) in the FP64 layout
        const R t2 = sub < D ? Tj[2 * D + sub] : (R)0, dt2 = sub < D ? dTj[2 * D + sub] : (R)0;
        const R t5 = sub < D ? Tj[5 * D + sub] : (R)0, dt5 = sub < D ? dTj[5 * D + sub] : (R)0;
        R S = 0, Gs[6] = {0, 0, 0, 0, 0, 0};
        if constexpr (!BF) {
            // Two products per 16-row block of the square image C[abc][def]:
            // Z = C^T X (rows def) gives the energy and the gradients of pairs
            // 12, 13, 23 against Y and its derivatives; W = C Y (rows abc)
            // gives those of pairs 01, 02, 03 against the derivatives of X.
            R T2[D], dT2[D];
#pragma unroll
            for (int x = 0; x < D; x++) { T2[x] = Tj[2 * D + x]; dT2[x] = dTj[2 * D + x]; }
#pragma unroll
            for (int rb = 0; rb < SH::RB; rb++) {
                acc_t z = {0, 0, 0, 0}, w = {0, 0, 0, 0};
#pragma unroll
                for (int s = 0; s < SH::KS; s++) {
                    const int a = s / D, b = s % D;
                    const R Az = Cl[LQ::at(4 * s + sub, 16 * rb + j)];
                    const R Aw = Cl[LQ::at(16 * rb + j, 4 * s + sub)];
                    z = MA::op(Az, (T0[a] * T1[b]) * t2, z);
                    w = MA::op(Aw, (T3[a] * T4[b]) * t5, w);
                }
                if (MA::f64) {
                    // output row 16 rb + sub + 4 v: (d, e) or (a, b) = 4 rb + v, f or c = sub
#pragma unroll
                    for (int v = 0; v < 4; v++) {
                        const int de = 4 * rb + v;
                        if (de >= D * D) continue;
                        const int d = de / D, e = de % D;
                        const R t34 = T3[d] * T4[e];
                        S += z[v] * (t34 * t5);
                        Gs[3] += z[v] * ((dT3[d] * T4[e]) * t5);
                        Gs[4] += z[v] * ((T3[d] * dT4[e]) * t5);
                        Gs[5] += z[v] * (t34 * dt5);
                        Gs[0] += w[v] * ((dT0[d] * T1[e]) * t2);
                        Gs[1] += w[v] * ((T0[d] * dT1[e]) * t2);
                        Gs[2] += w[v] * ((T0[d] * T1[e]) * dt2);
                    }
                } else {
                    // output row 16 rb + 4 sub + v: (d, e) or (a, b) = 4 rb + sub, f or c = v
                    R q34 = 0, q3 = 0, q4 = 0, p01 = 0, p0 = 0, p1 = 0;
#pragma unroll
                    for (int u = 0; u < 4; u++) {
                        const int de = 4 * rb + u;
                        if (de >= D * D) continue;
                        const int d = de / D, e = de % D;
                        if (sub == u) {
                            q34 = T3[d] * T4[e];
                            q3 = dT3[d] * T4[e];
                            q4 = T3[d] * dT4[e];
                            p01 = T0[d] * T1[e];
                            p0 = dT0[d] * T1[e];
                            p1 = T0[d] * dT1[e];
                        }
                    }
#pragma unroll
                    for (int v = 0; v < 4; v++) {
                        if (v >= D) continue;
                        S += z[v] * (q34 * T5[v]);
                        Gs[3] += z[v] * (q3 * T5[v]);
                        Gs[4] += z[v] * (q4 * T5[v]);
                        Gs[5] += z[v] * (q34 * dT5[v]);
                        Gs[0] += w[v] * (p0 * T2[v]);
                        Gs[1] += w[v] * (p1 * T2[v]);
                        Gs[2] += w[v] * (p01 * dT2[v]);
                    }
                }
                // keep the operand loads of the next row block out of this one
                __builtin_amdgcn_sched_barrier(0);
            }
        } else {
#pragma unroll
            for (int rb = 0; rb < SH::RB; rb++) {
                acc_t z0 = {0, 0, 0, 0}, z1 = {0, 0, 0, 0}, z2 = {0, 0, 0, 0}, z3 = {0, 0, 0, 0};
                // k = 32 s + 8 sub + q over (a, b, c) = (k / 4D, (k / 4) % D, k % 4)
#pragma unroll
                for (int s = 0; s < K32 / 32; s++) {
                    const bf16x8 A = *(const bf16x8 *)&ClRaw[(16 * rb + j) * LB::LD + 32 * s + 8 * sub];
                    float x[8], y0[8], y1[8], y2[8];
#pragma unroll
                    for (int q = 0; q < 8; q++) {
                        const int k = 32 * s + 8 * sub + q;
                        const int a = k / (4 * D), b = (k / 4) % D, c = q & 3;
                        const bool ok = k < SH::KP && c < D;
                        const float t0 = ok ? (float)Tj[a] : 0.f, e0 = ok ? (float)dTj[a] : 0.f;
                        const float t1 = ok ? (float)Tj[D + b] : 0.f, e1 = ok ? (float)dTj[D + b] : 0.f;
                        const float tc = ok ? (float)Tj[2 * D + c] : 0.f, ec = ok ? (float)dTj[2 * D + c] : 0.f;
                        x[q] = (t0 * t1) * tc;
                        y0[q] = (e0 * t1) * tc;
                        y1[q] = (t0 * e1) * tc;
                        y2[q] = (t0 * t1) * ec;
                    }
                    if (BF == 1) {
                        z0 = mfma_bf16(A, pack_bf16(x), z0);
                        z1 = mfma_bf16(A, pack_bf16(y0), z1);
                        z2 = mfma_bf16(A, pack_bf16(y1), z2);
                        z3 = mfma_bf16(A, pack_bf16(y2), z3);
                    } else {
                        const bf16x8 Al =
                            *(const bf16x8 *)&ClRaw[LB::size + (16 * rb + j) * LB::LD + 32 * s + 8 * sub];
                        bf16x8 h, l;
                        split_bf16(x, h, l);
                        z0 = mfma_bf16(A, h, z0); z0 = mfma_bf16(A, l, z0); z0 = mfma_bf16(Al, h, z0);
                        split_bf16(y0, h, l);
                        z1 = mfma_bf16(A, h, z1); z1 = mfma_bf16(A, l, z1); z1 = mfma_bf16(Al, h, z1);
                        split_bf16(y1, h, l);
                        z2 = mfma_bf16(A, h, z2); z2 = mfma_bf16(A, l, z2); z2 = mfma_bf16(Al, h, z2);
                        split_bf16(y2, h, l);
                        z3 = mfma_bf16(A, h, z3); z3 = mfma_bf16(A, l, z3); z3 = mfma_bf16(Al, h, z3);
                    }
                }
                // fold against Y (FP32 output layout: row 16 rb + 4 sub + v)
                R q34 = 0, q3 = 0, q4 = 0;
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const int de = 4 * rb + u;
                    if (de >= D * D) continue;
                    const int d = de / D, e = de % D;
                    if (sub == u) {
                        q34 = T3[d] * T4[e];
                        q3 = dT3[d] * T4[e];
                        q4 = T3[d] * dT4[e];
                    }
                }
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    if (v >= D) continue;
                    const R y = q34 * T5[v];
                    S += z0[v] * y;
                    Gs[0] += z1[v] * y;
                    Gs[1] += z2[v] * y;
                    Gs[2] += z3[v] * y;
                    Gs[3] += z0[v] * (q3 * T5[v]);
                    Gs[4] += z0[v] * (q4 * T5[v]);
                    Gs[5] += z0[v] * (q34 * dT5[v]);
                }
                __builtin_amdgcn_sched_barrier(0);
            }
        }
        // sum over the four lanes of each cluster
#pragma unroll
        for (int w = 16; w <= 32; w <<= 1) {
            S += __shfl_xor(S, w, 64);
#pragma unroll
            for (int q = 0; q < 6; q++) Gs[q] += __shfl_xor(Gs[q], w, 64);
            flags |= __shfl_xor(flags, w, 64);
        }
        // ---- forces (lane sub: runtime member sub, virial of pairs sub and sub + 4), energy, virial
        __syncthreads();
        st.tab.clear();
        __syncthreads();
        double q7[7] = {0, 0, 0, 0, 0, 0, 0};
        double F[3] = {0.0, 0.0, 0.0};
        bool bad = false;
        if (act) {
            double r[6], f[6], fd[6], E, fs[6], Gd[6];
#pragma unroll
            for (int q = 0; q < 6; q++) {
                r[q] = Gm[wv][j][q][0];
                f[q] = Gm[wv][j][q][1];
                fd[q] = Gm[wv][j][q][2];
                Gd[q] = (double)Gs[q];
            }
            cluster_scalars<6, double>(r, f, fd, (double)S, Gd, E, fs);
            const double *h = cfgs[cfg].h;
#pragma unroll
            for (int q = 0; q < 6; q++) {
                const int pa = pairs_of<4>::a[q], pb = pairs_of<4>::b[q];
                const bool mine_f = pa == a_s || pb == a_s;
                const bool mine_v = (q & 3) == sub;
                if (!mine_f && !mine_v) continue;
                double dr[3];
                pair_vec(h, Pm[wv][j][pa], Pm[wv][j][pb], dr);
                const double fq = fs[q];
                if (mine_f) {
                    const double sg = pa == a_s ? 1.0 : -1.0;
                    F[0] += sg * (fq * dr[0]);
                    F[1] += sg * (fq * dr[1]);
                    F[2] += sg * (fq * dr[2]);
                }
                if (mine_v) {
                    q7[1] -= fq * (dr[0] * dr[0]);
                    q7[2] -= fq * (dr[0] * dr[1]);
                    q7[3] -= fq * (dr[0] * dr[2]);
                    q7[4] -= fq * (dr[1] * dr[1]);
                    q7[5] -= fq * (dr[1] * dr[2]);
                    q7[6] -= fq * (dr[2] * dr[2]);
                }
            }
            if (sub == 0) {
                q7[0] = E;
                note_flags(cfg_stat, cfg, flags);
            }
        }
        // energy and virial of each cluster summed over its four lanes and
        // added to its configuration from lane sub 0; the members' forces
        // (lane sub: runtime member sub) through the LDS table
#pragma unroll
        for (int w = 16; w <= 32; w <<= 1)
#pragma unroll
            for (int k = 0; k < 7; k++) q7[k] += __shfl_xor(q7[k], w, 64);
        st.tab.add(act ? key_s : -1, F, bad);
        fx_cfg_sums(Cacc, (act && sub == 0) ? cfg : -1, blockIdx.x & (CHX_CACC_SLOTS - 1), q7, bad);
        __syncthreads();
        st.tab.flush(Facc);
        if (bad) atomicOr(&status[CHX_ST_FLAGS], CHX_BAD_RANGE);
    }
}

// ---------------------------------------------------------------------------
// 3-body on the matrix cores (D <= 16): Z_v[(b,c)][cluster] = sum_a
// Cm3[a][(b,c)] X_v[cluster][a] with X_0 = T_01 and X_1 = dT_01, folded
// against T_02[b] T_12[c] and its derivatives.  Matrix image Cm3[a][16 b + c]
// (a padded to a multiple of 4, c to 16, zero padding).  Same tiling as the
// 4-body kernel: lane (j = l % 16, sub = l / 16); output row m = 16 b + c with
// c = sub + 4v (FP64) or 4 sub + v (FP32).
// ---------------------------------------------------------------------------
template <int D>
struct mfma3_shape {
    static constexpr int KP = (D + 3) / 4 * 4;  // rows of Cm3 (a)
    static constexpr int MP = 16 * D;           // columns of Cm3 (b, c)
    static constexpr int KS = KP / 4;
    static constexpr int RB = D;
};

template <int D, typename R, int WPB, int BF>
__global__ __launch_bounds__(64 * WPB, 2) void k_eval3m(int G, const int *__restrict__ gbase,
                                                    const int *__restrict__ gend,
                                                    const chx_group *__restrict__ groups,
                                                    const int2 *__restrict__ list, const chx_cfg *__restrict__ cfgs,
                                                    const chx_atom_s *__restrict__ as,
                                                    int nb_cap,
                                                    const chx_entry_s *__restrict__ es, unsigned long long *Facc,
                                                    unsigned long long *Cacc, int *cfg_stat, int *status)
{
    using SH = mfma3_shape<D>;
    using MA = mfma_acc<R>;
    typedef typename MA::type acc_t;
    using LM = lds_mat<SH::KP, SH::MP>;
    constexpr int K32 = (SH::KP + 31) / 32 * 32;
    using LB = lds_matb<SH::MP, K32>;
    // staged polynomials per cluster: pair 01 [KP], pair 02 [D], pair 12 [16]; odd stride
    constexpr int O1 = SH::KP, O2 = SH::KP + D, TS = (SH::KP + D + 16) | 1;
    // BF = 1: BF16 image; BF = 3: BF16 high and residual images (three-product split)
    __shared__ __attribute__((aligned(16))) unsigned short ClRaw[BF ? (BF == 3 ? 2 : 1) * LB::size
                                                                     : (int)(LM::size * sizeof(R) / 2)];
    R *Cl = (R *)ClRaw;
    // staged polynomials; after the contraction the same memory holds the force table
    union stage_t {
        struct { R T[WPB][16 * TS], dT[WPB][16 * TS]; } p;
        fx_table<256> tab;
    };
    __shared__ stage_t st;
    R (&Ts)[WPB][16 * TS] = st.p.T;
    R (&dTs)[WPB][16 * TS] = st.p.dT;
    __shared__ double Gm[WPB][16][3][3];  // r, fcut, fcut' per pair
    __shared__ double Pm[WPB][16][3][3];  // member positions
    if (block_stop(status, CHX_STOP)) return;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int j = lane & 15, sub = lane >> 4;
    const long nbt = status[CHX_ST_N3] / (16 * WPB);
    const long bt = blockIdx.x;
    if (bt >= nbt) return;
    const long t0 = bt * (16 * WPB);
    const int g = find_group(gbase, G, t0);
    const chx_group &grp = groups[g];
    if (grp.type < 0 || grp.D != D) return;
    mat_stage<R, BF ? 1 : SH::KP * SH::MP, 64 * WPB> cm;
    if constexpr (BF) {
        for (int x = threadIdx.x; x < (BF == 3 ? 2 : 1) * SH::MP * K32; x += 64 * WPB) {
            const int img = x / (SH::MP * K32), y = x % (SH::MP * K32);
            ClRaw[img * LB::size + (y / K32) * LB::LD + y % K32] = grp.Cmb[x];
        }
    } else {
        cm.load(sizeof(R) == 8 ? (const R *)grp.Cm : (const R *)grp.Cmf);
    }
    // ---- stage: geometry (all lanes of a cluster), lane sub < 3 the polynomials of pair sub
    const long c = t0 + wv * 16 + j;
    bool act = c < gend[g];
    int2 rec = make_int2(-1, 0);
    int cfg = 0, flags = 0;
    int a_s = 0, key_s = -1;  // canonical index and replicated atom of runtime member sub
    {
        double p[3][3], r[3];
        int par[3], perm[3];
        if (act) {
            rec = list[c];
            act = cluster_geometry<3>(grp, rec, cfgs, as, nb_cap, es, p, par, r, cfg, perm);
        }
        if constexpr (!BF) cm.store([&](int x) -> R & { return Cl[LM::at(x / SH::MP, x % SH::MP)]; });
        if (act)
#pragma unroll
            for (int a = 0; a < 3; a++)
                if (perm[a] == sub) { a_s = a; key_s = par[a]; }
        if (sub < 3) {
            double pm[3] = {p[0][0], p[0][1], p[0][2]};
            double rq = r[0];
#pragma unroll
            for (int m = 1; m < 3; m++)
                if (sub == m) { pm[0] = p[m][0]; pm[1] = p[m][1]; pm[2] = p[m][2]; rq = r[m]; }
            Pm[wv][j][sub][0] = pm[0];
            Pm[wv][j][sub][1] = pm[1];
            Pm[wv][j][sub][2] = pm[2];
            R Tq[D], dTq[D];
            double fq = 0.0, fdq = 0.0;
            if (act) {
                const chx_slot sl = grp.s[sub];
                cheby<D, R>(rq, sl, Tq, dTq, flags);
                fcut(rq, sl, grp.tersoff, grp.fcut_var, fq, fdq);
            } else {
#pragma unroll
                for (int x = 0; x < D; x++) { Tq[x] = (R)0; dTq[x] = (R)0; }
            }
            const int o = sub == 0 ? 0 : (sub == 1 ? O1 : O2);
            const int w = sub == 0 ? SH::KP : (sub == 1 ? D : 16);
#pragma unroll
            for (int x = 0; x < 16; x++)
                if (x < w) {
                    Ts[wv][j * TS + o + x] = x < D ? Tq[x < D ? x : 0] : (R)0;
                    dTs[wv][j * TS + o + x] = x < D ? dTq[x < D ? x : 0] : (R)0;
                }
            Gm[wv][j][sub][0] = rq;
            Gm[wv][j][sub][1] = fq;
            Gm[wv][j][sub][2] = fdq;
        }
    }
    __syncthreads();
    // ---- contraction on the matrix cores
    const R *Tj = &Ts[wv][j * TS], *dTj = &dTs[wv][j * TS];
    R t2s[4], dt2s[4];
#pragma unroll
    for (int v = 0; v < 4; v++) {
        const int cc = MA::f64 ? sub + 4 * v : 4 * sub + v;
        t2s[v] = Tj[O2 + cc];
        dt2s[v] = dTj[O2 + cc];
    }
    R S = 0, Gs[3] = {0, 0, 0};
#pragma unroll
    for (int rb = 0; rb < SH::RB; rb++) {
        acc_t z0 = {0, 0, 0, 0}, z1 = {0, 0, 0, 0};
        if constexpr (BF) {
#pragma unroll
            for (int s = 0; s < K32 / 32; s++) {
                const bf16x8 A = *(const bf16x8 *)&ClRaw[(16 * rb + j) * LB::LD + 32 * s + 8 * sub];
                float x[8], y[8];
#pragma unroll
                for (int q = 0; q < 8; q++) {
                    const int a = 32 * s + 8 * sub + q;
                    x[q] = a < D ? (float)Tj[a < D ? a : 0] : 0.f;
                    y[q] = a < D ? (float)dTj[a < D ? a : 0] : 0.f;
                }
                if (BF == 1) {
                    z0 = mfma_bf16(A, pack_bf16(x), z0);
                    z1 = mfma_bf16(A, pack_bf16(y), z1);
                } else {
                    const bf16x8 Al = *(const bf16x8 *)&ClRaw[LB::size + (16 * rb + j) * LB::LD + 32 * s + 8 * sub];
                    bf16x8 h, l;
                    split_bf16(x, h, l);
                    z0 = mfma_bf16(A, h, z0); z0 = mfma_bf16(A, l, z0); z0 = mfma_bf16(Al, h, z0);
                    split_bf16(y, h, l);
                    z1 = mfma_bf16(A, h, z1); z1 = mfma_bf16(A, l, z1); z1 = mfma_bf16(Al, h, z1);
                }
            }
        } else {
#pragma unroll
            for (int s = 0; s < SH::KS; s++) {
                const R A = Cl[LM::at(4 * s + sub, 16 * rb + j)];
                z0 = MA::op(A, Tj[4 * s + sub], z0);
                z1 = MA::op(A, dTj[4 * s + sub], z1);
            }
        }
        const R t1 = Tj[O1 + rb], dt1 = dTj[O1 + rb];
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const R y = t1 * t2s[v];
            S += z0[v] * y;
            Gs[0] += z1[v] * y;
            Gs[1] += z0[v] * (dt1 * t2s[v]);
            Gs[2] += z0[v] * (t1 * dt2s[v]);
        }
    }
#pragma unroll
    for (int w = 16; w <= 32; w <<= 1) {
        S += __shfl_xor(S, w, 64);
#pragma unroll
        for (int q = 0; q < 3; q++) Gs[q] += __shfl_xor(Gs[q], w, 64);
        flags |= __shfl_xor(flags, w, 64);
    }
    // ---- forces (lane sub < 3: runtime member sub and the virial of pair sub), energy, virial
    __syncthreads();
    st.tab.clear();
    __syncthreads();
    double q7[7] = {0, 0, 0, 0, 0, 0, 0};
    double F[3] = {0.0, 0.0, 0.0};
    bool bad = false;
    if (act && sub < 3) {
        double r[3], f[3], fd[3], E, fs[3], Gd[3];
#pragma unroll
        for (int q = 0; q < 3; q++) {
            r[q] = Gm[wv][j][q][0];
            f[q] = Gm[wv][j][q][1];
            fd[q] = Gm[wv][j][q][2];
            Gd[q] = (double)Gs[q];
        }
        cluster_scalars<3, double>(r, f, fd, (double)S, Gd, E, fs);
        const double *h = cfgs[cfg].h;
#pragma unroll
        for (int q = 0; q < 3; q++) {
            const int pa = pairs_of<3>::a[q], pb = pairs_of<3>::b[q];
            const bool mine_f = pa == a_s || pb == a_s;
            const bool mine_v = q == sub;
            double dr[3];
            pair_vec(h, Pm[wv][j][pa], Pm[wv][j][pb], dr);
            const double fq = fs[q];
            if (mine_f) {
                const double sg = pa == a_s ? 1.0 : -1.0;
                F[0] += sg * (fq * dr[0]);
                F[1] += sg * (fq * dr[1]);
                F[2] += sg * (fq * dr[2]);
            }
            if (mine_v) {
                q7[1] -= fq * (dr[0] * dr[0]);
                q7[2] -= fq * (dr[0] * dr[1]);
                q7[3] -= fq * (dr[0] * dr[2]);
                q7[4] -= fq * (dr[1] * dr[1]);
                q7[5] -= fq * (dr[1] * dr[2]);
                q7[6] -= fq * (dr[2] * dr[2]);
            }
        }
        if (sub == 0) {
            q7[0] = E;
            note_flags(cfg_stat, cfg, flags);
        }
    }
    // as in the 4-body kernel: per-cluster energy and virial over the lanes
    // to the configuration, the members' forces through the LDS table
#pragma unroll
    for (int w = 16; w <= 32; w <<= 1)
#pragma unroll
        for (int k = 0; k < 7; k++) q7[k] += __shfl_xor(q7[k], w, 64);
    st.tab.add((act && sub < 3) ? key_s : -1, F, bad);
    fx_cfg_sums(Cacc, (act && sub == 0) ? cfg : -1, blockIdx.x & (CHX_CACC_SLOTS - 1), q7, bad);
    __syncthreads();
    st.tab.flush(Facc);
    if (bad) atomicOr(&status[CHX_ST_FLAGS], CHX_BAD_RANGE);
}

// ---------------------------------------------------------------------------
// 3-body with one cluster per lane, as k_evalv, and the contraction over the
// first pair's index on the matrix cores (FP32 forms, D <= CHX_MFMA3V_MAX_D).
// Each lane stages its cluster's polynomials in LDS; a wavefront's 64
// clusters are then contracted as four 16-cluster MFMA tiles:
//   Z_p[(b, c)][n] = sum_a C[a][b][c] X_p[a][n],  X_0 = T_01, X_1 = dT_01,
// rows (b, c) = b CP + c (c padded to CP, a multiple of 4), the coefficient
// operands held in registers for the whole wavefront.  Output lane l holds
// rows 16 rb + 4 (l / 16) + v, v = 0..3: one b and four consecutive c, so
//   E   += T_02[b] sum_v Z_0 T_12[c],   G_02 += dT_02[b] sum_v Z_0 T_12[c],
//   G_12 += T_02[b] sum_v Z_0 dT_12[c], G_01 += T_02[b] sum_v Z_1 T_12[c]
// (the four-term sums in FP32, the rest in FP64), and the four lanes of a
// cluster column are added by two exchanges.  Every cluster's sums follow
// the same fixed order wherever it sits in a tile, so results keep the
// batch-independence of the other kernels.  Scalars, forces and
// accumulation are those of k_evalv.
// ---------------------------------------------------------------------------
template <int D>
struct mfma3v_shape {
    static constexpr int CP = (D + 3) / 4 * 4;       // padded c (and a) range
    static constexpr int RB = (D * CP + 15) / 16;    // 16-row blocks over (b, c)
    static constexpr int KS = CP / 4;                // k steps of 4 over a
    static constexpr int STR = 6 * CP + 1;           // per-cluster LDS stride (odd: conflict-free columns)
};

template <int D>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_eval3v(
    int G, const int *__restrict__ gbase, const int *__restrict__ gend, const chx_group *__restrict__ groups,
    int total_word, const int2 *__restrict__ list, const chx_cfg *__restrict__ cfgs,
    const chx_atom_s *__restrict__ as, int nb_cap, const chx_entry_s *__restrict__ es, unsigned long long *Facc,
    unsigned long long *Cacc, int *cfg_stat, int *status)
{
    using SH = mfma3v_shape<D>;
    constexpr int NB = 3, P = 3, CP = SH::CP, STR = SH::STR;
    // the staged polynomials (T_01, dT_01, T_02, dT_02, T_12, dT_12, each
    // padded with zeros to CP) and, after the contraction, the force table
    union shared_t {
        float T[4][64 * STR];
        fx_table<1024> tab;
    };
    __shared__ shared_t sh;
    __shared__ int flags_s;
    if (threadIdx.x == 0) flags_s = __atomic_load_n(&status[CHX_ST_FLAGS], __ATOMIC_RELAXED);
    __syncthreads();
    if (flags_s & CHX_STOP) return;
    const int lane = threadIdx.x & 63, col = lane & 15, r4 = lane >> 4;
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const long t0 = ((long)blockIdx.x * 4 + wave) * 64;
    bool wave_on = t0 < status[total_word];
    const int g = wave_on ? find_group(gbase, G, t0) : 0;
    const chx_group &grp = groups[g];
    wave_on = wave_on && grp.type >= 0 && grp.D == D;
    const long c = t0 + lane;
    float *Tw = sh.T[wave];
    // ---- geometry and polynomials, one cluster per lane
    bool act = false;
    int2 rec = make_int2(-1, 0);
    double p[NB][3], r[P], f[P], fd[P], dr[P][3];
    int par[NB], cfg = 0, pm[NB], flags = 0;
    {
        float T[P][D], dT[P][D];
#pragma unroll
        for (int q = 0; q < P; q++)
#pragma unroll
            for (int x = 0; x < D; x++) { T[q][x] = 0.f; dT[q][x] = 0.f; }
        if (wave_on && c < gend[g]) {
            rec = list[c];
            act = cluster_geometry<NB>(grp, rec, cfgs, as, nb_cap, es, p, par, r, cfg, pm, dr);
            if (act)
#pragma unroll
                for (int q = 0; q < P; q++) {
                    cheby<D, float>(r[q], grp.s[q], T[q], dT[q], flags);
                    fcut(r[q], grp.s[q], grp.tersoff, grp.fcut_var, f[q], fd[q]);
                }
        }
        float *Tl = Tw + lane * STR;
#pragma unroll
        for (int q = 0; q < P; q++)
#pragma unroll
            for (int x = 0; x < CP; x++) {
                Tl[(2 * q) * CP + x] = x < D ? T[q][x] : 0.f;
                Tl[(2 * q + 1) * CP + x] = x < D ? dT[q][x] : 0.f;
            }
    }
    __syncthreads();
    // ---- contraction on the matrix cores (whole wavefronts)
    double S = 0.0, Gq[P] = {0.0, 0.0, 0.0};
    if (wave_on) {
        // coefficient operands: A[row][a] = C[a][b][c], row = 16 rb + col = b CP + c, a = 4 s + r4
        const float *Cf = grp.Cf;
        float A[SH::RB][SH::KS];
#pragma unroll
        for (int rb = 0; rb < SH::RB; rb++)
#pragma unroll
            for (int s = 0; s < SH::KS; s++) {
                const int row = 16 * rb + col, a = 4 * s + r4;
                const int b = row / CP, cc = row % CP;
                A[rb][s] = (a < D && b < D && cc < D) ? Cf[(a * D + b) * D + cc] : 0.f;
            }
        // per sub-tile q: energy and the three gradients of the column's
        // cluster, partial over this lane's rows
        double part[4][4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const float *Tn = Tw + (16 * q + col) * STR;  // the cluster of this lane's column
            float x0[SH::KS], x1[SH::KS];
#pragma unroll
            for (int s = 0; s < SH::KS; s++) {
                x0[s] = Tn[4 * s + r4];
                x1[s] = Tn[CP + 4 * s + r4];
            }
            double e = 0.0, g01 = 0.0, g02 = 0.0, g12 = 0.0;
#pragma unroll
            for (int rb = 0; rb < SH::RB; rb++) {
                flt4 z0 = {0.f, 0.f, 0.f, 0.f}, z1 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int s = 0; s < SH::KS; s++) {
                    z0 = __builtin_amdgcn_mfma_f32_16x16x4f32(A[rb][s], x0[s], z0, 0, 0, 0);
                    z1 = __builtin_amdgcn_mfma_f32_16x16x4f32(A[rb][s], x1[s], z1, 0, 0, 0);
                }
                // rows 16 rb + 4 r4 + v: b = row0 / CP, c = row0 % CP + v
                const int row0 = 16 * rb + 4 * r4, b = row0 / CP, c0 = row0 % CP;
                if (b < D) {
                    float u0 = z0[0] * Tn[4 * CP + c0], u1 = z0[0] * Tn[5 * CP + c0], w = z1[0] * Tn[4 * CP + c0];
#pragma unroll
                    for (int v = 1; v < 4; v++) {
                        const float t12 = Tn[4 * CP + c0 + v], d12 = Tn[5 * CP + c0 + v];
                        u0 = fmaf(z0[v], t12, u0);
                        u1 = fmaf(z0[v], d12, u1);
                        w = fmaf(z1[v], t12, w);
                    }
                    const double t02 = Tn[2 * CP + b], d02 = Tn[3 * CP + b];
                    e = fma((double)u0, t02, e);
                    g02 = fma((double)u0, d02, g02);
                    g12 = fma((double)u1, t02, g12);
                    g01 = fma((double)w, t02, g01);
                }
            }
            part[q][0] = e;
            part[q][1] = g01;
            part[q][2] = g02;
            part[q][3] = g12;
        }
        // the four lanes of each column added, sub-tile q ending on the lane
        // with r4 = q: exchange across lane halves (r4 ^ 2) keeping the
        // sub-tiles of this lane's half, then across rows (r4 ^ 1)
        const bool hi2 = (r4 & 2) != 0, hi1 = (r4 & 1) != 0;
        double h[2][4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
#pragma unroll
            for (int j = 0; j < 2; j++) {
                const double keep = hi2 ? part[2 + j][k] : part[j][k];
                const double send = hi2 ? part[j][k] : part[2 + j][k];
                h[j][k] = keep + __shfl_xor(send, 32, 64);
            }
            const double keep = hi1 ? h[1][k] : h[0][k];
            const double send = hi1 ? h[0][k] : h[1][k];
            const double t = keep + __shfl_xor(send, 16, 64);
            if (k == 0) S = t;
            else Gq[k - 1] = t;
        }
    }
    __syncthreads();
    sh.tab.clear();
    __syncthreads();
    // ---- scalars, forces and accumulation (as k_evalv)
    int ckey = -1;
    double q7[7] = {0, 0, 0, 0, 0, 0, 0};
    double Fr[NB][3];
    int key[NB];
#pragma unroll
    for (int m = 0; m < NB; m++) { key[m] = -1; Fr[m][0] = Fr[m][1] = Fr[m][2] = 0.0; }
    bool bad = false;
    if (act) {
        double E, fs[P];
        cluster_scalars<P, double>(r, f, fd, S, Gq, E, fs);
        q7[0] = E;
        double F[NB][3];
        cluster_forces<NB>(dr, fs, F, q7);
#pragma unroll
        for (int a = 0; a < NB; a++) {
            key[a] = par[a];
            Fr[a][0] = F[a][0]; Fr[a][1] = F[a][1]; Fr[a][2] = F[a][2];
        }
        ckey = cfg;
        note_flags(cfg_stat, cfg, flags);
    }
#pragma unroll
    for (int m = 0; m < NB; m++) sh.tab.add(key[m], Fr[m], bad);
    fx_cfg_sums(Cacc, ckey, blockIdx.x & (CHX_CACC_SLOTS - 1), q7, bad);
    __syncthreads();
    sh.tab.flush(Facc);
    if (bad) atomicOr(&status[CHX_ST_FLAGS], CHX_BAD_RANGE);
}

// ---------------------------------------------------------------------------
// Totals, one launch.  Blocks [0, nitems): energy and virial, one chunk of
// CHX_FIN_CHUNK owners of one configuration each (item; cfg_item_off[c] is
// the first item of c), exact partial sums normalised to part[item][7][2] =
// (hi, lo < 2^40) plus the 1-body offsets (compute_1B, chimesFF.cpp:1636-
// 1641); the last block of a configuration to finish (done[c], zeroed before
// the run) adds its items' partials and writes the totals.  Blocks from
// nitems on: force on input atom a, the exact sum over its replicas' force
// accumulators.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_fin(int ncfg, int nitems, const int *__restrict__ cfg_item_off,
                                              const int *__restrict__ item_cfg, int natoms_in,
                                              const int *__restrict__ in_cfg,
                                              const chx_cfg *__restrict__ cfgs, const chx_atom_s *__restrict__ as,
                                              const chx_model_dev *__restrict__ models,
                                              const unsigned long long *__restrict__ Facc,
                                              const unsigned long long *__restrict__ Cacc,
                                              const int *__restrict__ nbr_cnt, long long *part, int *done,
                                              double *Fout, double *Ecfg, double *Vcfg, int *status)
{
    __shared__ long long wsum[4][16];  // per wavefront: (hi, lo) of the seven sums
    __shared__ int last, nimg;
    if ((int)blockIdx.x >= nitems) {
        const int a = ((int)blockIdx.x - nitems) * 256 + threadIdx.x;
        if (a >= natoms_in) return;
        const chx_cfg &C = cfgs[in_cfg[a]];
        const int la = a - C.atom_in_off;
        const int ncopy = (C.n_rep + 1) * (C.n_rep + 1) * (C.n_rep + 1);
        fx_sum f[3];
        for (int rho = 0; rho < ncopy; rho++) {
            const long r = C.atom_off + (long)rho * C.natoms_in + la;
#pragma unroll
            for (int k = 0; k < 3; k++) f[k].add(Facc + 2L * (3L * r + k));
        }
#pragma unroll
        for (int k = 0; k < 3; k++) Fout[3L * a + k] = f[k].value();
        return;
    }
    const int item = blockIdx.x;
    const int c = item_cfg[item];
    const chx_cfg &C = cfgs[c];
    const chx_model_dev &M = models[C.model];
    const long r0 = C.atom_off + (long)(item - cfg_item_off[c]) * CHX_FIN_CHUNK;
    const long r1 = r0 + CHX_FIN_CHUNK < (long)C.atom_off + C.natoms ? r0 + CHX_FIN_CHUNK
                                                                      : (long)C.atom_off + C.natoms;
    fx_sum acc[7];
    bool bad = false;
    int ni = 0;  // neighbour images of these owners (reported count)
    if (threadIdx.x == 0) nimg = 0;
    for (long r = r0 + threadIdx.x; r < r1; r += 256) {
        acc[0].add_value(M.eoff[as[r].type], bad);
        ni += nbr_cnt[r];
    }
    // the configuration's own sums (clusters added per configuration, one
    // copy per thread of the configuration's first item)
    if ((int)threadIdx.x < CHX_CACC_SLOTS && item == cfg_item_off[c]) {
        const unsigned long long *a = Cacc + ((long)c * CHX_CACC_SLOTS + threadIdx.x) * CHX_CACC_STRIDE;
#pragma unroll
        for (int k = 0; k < 7; k++) acc[k].add(a + 2 * k);
    }
    __syncthreads();
    if (ni) atomicAdd(&nimg, ni);
    // the seven exact sums over the workgroup: each wavefront's by the
    // transposed exchanges (lane l holds slot l of its wavefront), then the
    // four wavefronts' through LDS
    {
        long long x[16];
#pragma unroll
        for (int k = 0; k < 7; k++) {
            x[2 * k] = acc[k].hi + (acc[k].lo >> 40);
            x[2 * k + 1] = acc[k].lo & 0xffffffffffll;
        }
        x[14] = x[15] = 0;
        const long long t = fx_wave_sums16(x);
        if ((threadIdx.x & 63) < 16) wsum[threadIdx.x >> 6][threadIdx.x & 15] = t;
    }
    __syncthreads();
    if (bad) atomicOr(&status[CHX_ST_FLAGS], CHX_BAD_RANGE);
    if (threadIdx.x == 0 && nimg) atomicAdd(&status[CHX_ST_NENT], nimg);
    const int nit = cfg_item_off[c + 1] - cfg_item_off[c];
    if (threadIdx.x < 7) {
        const int k = threadIdx.x;
        long long hs = 0, ls = 0;
#pragma unroll
        for (int w = 0; w < 4; w++) {
            hs += wsum[w][2 * k];
            ls += wsum[w][2 * k + 1];
        }
        const long long h = hs + (ls >> 40);
        const long long l = ls & 0xffffffffffll;
        if (nit == 1) {
            fx_sum t;
            t.hi = h;
            t.lo = l;
            if (k == 0) Ecfg[c] = t.value(); else Vcfg[6L * c + k - 1] = t.value();
        } else {
            __atomic_store_n(&part[14L * item + 2 * k], h, __ATOMIC_RELAXED);
            __atomic_store_n(&part[14L * item + 2 * k + 1], l, __ATOMIC_RELAXED);
        }
    }
    if (nit == 1) return;
    // publish this item's partials, then count it; the last item of the
    // configuration sums all of them
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) last = atomicAdd(&done[c], 1) == nit - 1;
    __syncthreads();
    if (!last) return;
    __threadfence();
    if (threadIdx.x < 7) {
        const int k = threadIdx.x;
        fx_sum t;
        for (int it = cfg_item_off[c]; it < cfg_item_off[c + 1]; it++) {
            const unsigned long long v[2] = {
                (unsigned long long)__atomic_load_n(&part[14L * it + 2 * k], __ATOMIC_RELAXED),
                (unsigned long long)__atomic_load_n(&part[14L * it + 2 * k + 1], __ATOMIC_RELAXED)};
            t.add(v);
        }
        if (k == 0) Ecfg[c] = t.value(); else Vcfg[6L * c + k - 1] = t.value();
    }
}

// ---------------------------------------------------------------------------
// Explicit clusters (per-cluster API): one thread per cluster of NB bodies
// given by runtime-order pair distances and vectors (x_b - x_a).  Besides E,
// forces and the packed stress it returns, for 3/4 bodies, the force scalars
// of the last term of the type's list (compute_3B/4B's force_scalar_in,
// chimesFF.cpp:2171-2173, 2864-2869) and, for 2 bodies, the penalty scalar
// (chimesFF.cpp:1745-1774).
// ---------------------------------------------------------------------------
template <int NB, int D>
__global__ __launch_bounds__(128) void k_clusters(int n, const double *__restrict__ dist, const double *__restrict__ drv,
                           const int *__restrict__ types, chx_model_dev M, double *E, double *F, double *Sout,
                           double *last, int *flags_out)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    constexpr int P = NB * (NB - 1) / 2;
    int ty[NB];
    double rq[P], dr[P][3];
#pragma unroll
    for (int a = 0; a < NB; a++) ty[a] = types[(long)c * NB + a];
#pragma unroll
    for (int p = 0; p < P; p++) {
        rq[p] = dist[(long)c * P + p];
#pragma unroll
        for (int k = 0; k < 3; k++) dr[p][k] = drv[((long)c * P + p) * 3 + k];
    }
    double Ec, fo[NB][3], So[6];
    int flags = 0;
    bool evaluated;
    explicit_cluster<NB, D>(M, ty, rq, dr, Ec, fo, So, flags, evaluated, last + (long)c * P);
    E[c] = Ec;
    for (int a = 0; a < NB; a++)
        for (int k = 0; k < 3; k++) F[((long)c * NB + a) * 3 + k] = fo[a][k];
    for (int k = 0; k < 6; k++) Sout[(long)c * 6 + k] = So[k];
    flags_out[c] = flags | (evaluated ? 8 : 0);
}

// ---------------------------------------------------------------------------
// Host-side launchers: dispatch on the dense dimension D and precision.
// ---------------------------------------------------------------------------
#define CHX_D4M_CASES(X) X(2) X(3) X(4)
#define CHX_D3V_CASES(X) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10)

bool supported_D3(int D) { return D >= 2 && D <= 13; }
bool supported_D4(int D) { return D >= 2 && D <= 6; }

void launch_eval2(int natoms, int nb_cap, hipStream_t st, const int *nbr_cnt, const chx_cfg *cfgs,
                  const chx_atom_s *as, const chx_entry_s *es, const chx_model_dev *models, unsigned long long *Facc,
                  unsigned long long *Cacc, int *cfg_stat, int *status)
{
    if (natoms <= 0) return;
    // results do not depend on the shape: per-owner sums run over that owner's
    // sixteen lanes and image forces are exact sums
    constexpr int LPO = CHX_EVAL2_LANES;
    if (natoms < 16384)
        hipLaunchKernelGGL((k_eval2<64 / LPO, LPO>), dim3((unsigned)((natoms + 64 / LPO - 1) / (64 / LPO))), dim3(64), 0,
                           st, natoms, nb_cap, nbr_cnt, cfgs, as, es, models, Facc, Cacc, cfg_stat, status);
    else
        hipLaunchKernelGGL((k_eval2<256 / LPO, LPO>), dim3((unsigned)((natoms + 256 / LPO - 1) / (256 / LPO))), dim3(256),
                           0, st, natoms, nb_cap, nbr_cnt, cfgs, as, es, models, Facc, Cacc, cfg_stat, status);
}

void launch_evalv(int nb, int D, int prec, const chx_list_args &la, hipStream_t st)
{
    if (la.cap <= 0 || la.G <= 0) return;
    const bool f32 = prec != CHX_FP64;
    const dim3 grid((unsigned)((la.cap + 255) / 256)), block(256);
    const int word = nb == 3 ? CHX_ST_N3 : CHX_ST_N4;
#define X(d)                                                                                                   \
    if (D == d) {                                                                                              \
        if (f32)                                                                                               \
            hipLaunchKernelGGL((k_evalv<NB_, d, float>), grid, block, 0, st, la.G, la.gbase, la.gend, la.groups, \
                               word, la.list, la.cfgs, la.as, la.nb_cap, la.es, la.Facc, la.Cacc, la.cfg_stat,  \
                               la.status);                                                                     \
        else                                                                                                   \
            hipLaunchKernelGGL((k_evalv<NB_, d, double>), grid, block, 0, st, la.G, la.gbase, la.gend,          \
                               la.groups, word, la.list, la.cfgs, la.as, la.nb_cap, la.es, la.Facc, la.Cacc,  \
                               la.cfg_stat, la.status);                                                        \
        return;                                                                                                \
    }
    if (nb == 3) {
#define NB_ 3
        CHX_D3_CASES(X)
#undef NB_
    } else {
#define NB_ 4
        CHX_D4_CASES(X)
#undef NB_
    }
#undef X
}

// Matrix-core kernels: FP64 (prec 0), FP32 (1), BF16 operands (2) or split
// BF16 operands, three products each (3); BF16 forms accumulate in FP32.
#define CHX_LAUNCH_M(KER, d)                                                                                   \
    {                                                                                                          \
        if (prec == CHX_FP64)                                                                                  \
            hipLaunchKernelGGL((KER<d, double, WPB, 0>), grid, block, 0, st, la.G, la.gbase, la.gend,           \
                               la.groups, la.list, la.cfgs, la.as, la.nb_cap, la.es, la.Facc, la.Cacc,        \
                               la.cfg_stat, la.status);                                                        \
        else if (prec == CHX_MIXED)                                                                            \
            hipLaunchKernelGGL((KER<d, float, WPB, 0>), grid, block, 0, st, la.G, la.gbase, la.gend,            \
                               la.groups, la.list, la.cfgs, la.as, la.nb_cap, la.es, la.Facc, la.Cacc,        \
                               la.cfg_stat, la.status);                                                        \
        else if (prec == CHX_BF16)                                                                             \
            hipLaunchKernelGGL((KER<d, float, WPB, 1>), grid, block, 0, st, la.G, la.gbase, la.gend,            \
                               la.groups, la.list, la.cfgs, la.as, la.nb_cap, la.es, la.Facc, la.Cacc,        \
                               la.cfg_stat, la.status);                                                        \
        else                                                                                                   \
            hipLaunchKernelGGL((KER<d, float, WPB, 3>), grid, block, 0, st, la.G, la.gbase, la.gend,            \
                               la.groups, la.list, la.cfgs, la.as, la.nb_cap, la.es, la.Facc, la.Cacc,        \
                               la.cfg_stat, la.status);                                                        \
        return true;                                                                                           \
    }

template <int WPB>
bool launch_eval3m(int D, int prec, const chx_list_args &la, hipStream_t st)
{
    if (D > CHX_MFMA3_MAX_D) return false;
    if (la.cap <= 0 || la.G <= 0) return true;
    const dim3 grid((unsigned)std::max<long>(1, la.cap / (16 * WPB))), block(64 * WPB);
#define X(d) \
    if (D == d) CHX_LAUNCH_M(k_eval3m, d)
    CHX_D3_CASES(X)
#undef X
    return false;
}

template <int WPB>
bool launch_eval4m(int D, int prec, const chx_list_args &la, hipStream_t st)
{
    if (D > CHX_MFMA_MAX_D) return false;
    if (la.cap <= 0 || la.G <= 0) return true;
    const dim3 grid((unsigned)std::max<long>(1, la.cap / (16 * WPB))), block(64 * WPB);
#define X(d) \
    if (D == d) CHX_LAUNCH_M(k_eval4m, d)
    CHX_D4M_CASES(X)
#undef X
    return false;
}

bool launch_eval3v(int D, const chx_list_args &la, hipStream_t st)
{
    if (D > CHX_MFMA3V_MAX_D) return false;
    if (la.cap <= 0 || la.G <= 0) return true;
    const dim3 grid((unsigned)((la.cap + 255) / 256)), block(256);
#define X(d)                                                                                                   \
    if (D == d) {                                                                                              \
        hipLaunchKernelGGL((k_eval3v<d>), grid, block, 0, st, la.G, la.gbase, la.gend, la.groups, CHX_ST_N3,     \
                           la.list, la.cfgs, la.as, la.nb_cap, la.es, la.Facc, la.Cacc, la.cfg_stat, la.status); \
        return true;                                                                                           \
    }
    CHX_D3V_CASES(X)
#undef X
    return false;
}

bool variant_applies(int nb, int prec, int variant)
{
    return variant != CHX_VAR_MFMAV || (nb == 3 && prec == CHX_MIXED);
}

void launch_eval_body(int nb, int D, int prec, int variant, const chx_list_args &la, hipStream_t st)
{
    bool done = false;
    if (variant == CHX_VAR_MFMAV && nb == 3 && prec == CHX_MIXED)
        done = launch_eval3v(D, la, st);
    else if (variant == CHX_VAR_MFMA4 || (variant >= CHX_VAR_VECTOR && prec >= CHX_BF16))
        done = nb == 3 ? launch_eval3m<4>(D, prec, la, st) : launch_eval4m<4>(D, prec, la, st);
    else if (variant == CHX_VAR_MFMA2)
        done = nb == 3 ? launch_eval3m<2>(D, prec, la, st) : launch_eval4m<2>(D, prec, la, st);
    if (!done) launch_evalv(nb, D, prec, la, st);
}

void launch_clusters(int nb, int D, int n, hipStream_t st, const double *dist, const double *drv, const int *types,
                     const chx_model_dev &M, double *E, double *F, double *S, double *last, int *flags)
{
    if (n <= 0) return;
    const dim3 grid((unsigned)((n + 127) / 128)), block(128);
    if (nb == 2) {
        hipLaunchKernelGGL((k_clusters<2, 2>), grid, block, 0, st, n, dist, drv, types, M, E, F, S, last, flags);
        return;
    }
#define X(d)                                                                                                 \
    if (D == d) {                                                                                            \
        hipLaunchKernelGGL((k_clusters<3, d>), grid, block, 0, st, n, dist, drv, types, M, E, F, S, last, flags); \
        return;                                                                                              \
    }
    if (nb == 3) { CHX_D3_CASES(X) }
#undef X
#define X(d)                                                                                                 \
    if (D == d) {                                                                                            \
        hipLaunchKernelGGL((k_clusters<4, d>), grid, block, 0, st, n, dist, drv, types, M, E, F, S, last, flags); \
        return;                                                                                              \
    }
    if (nb == 4) { CHX_D4_CASES(X) }
#undef X
}

}  // namespace chx
