// Evaluation kernels: 2-, 3- and 4-body cluster terms, the deterministic
// owner-ordered reduction of pair force scalars into forces and virial, and
// the explicit-cluster kernel behind the per-cluster API.
//
// Compiled with -ffp-contract=off: the cluster geometry (dr = h (v_b - v_a))
// reproduces the reference's get_dist bit for bit; the polynomial and
// coefficient arithmetic uses explicit fma().
#include <hip/hip_runtime.h>

#include "chx_eval.h"
#include "eval_math.h"

namespace chx {

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

// ---------------------------------------------------------------------------
// 2-body: one thread per neighbour entry.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_eval2(long nent, const int *__restrict__ e_owner,
                                                const chx_cfg *__restrict__ cfgs, const chx_atom_s *__restrict__ as,
                                                const chx_entry_s *__restrict__ es,
                                                const chx_model_dev *__restrict__ models, double *rec2E,
                                                double *rec2F, int *cfg_stat)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nent) return;
    const int i = e_owner[e];
    const chx_atom_s ai = as[i];
    const chx_entry_s g = es[e];
    const chx_cfg &C = cfgs[ai.cfg];
    const chx_model_dev &M = models[C.model];
    double dr[3];
    const double r = pair_vec(C.h, ai.u, g.v, dr);
    const int pt = M.pairmap[ai.type * M.nt + g.type];
    double E, fs;
    int flags = 0;
    eval_pair(r, M.slots2[pt], M.c2 + (long)pt * (M.order[0] + 1), M.order[0], M.tersoff, M.fcut_var, M.pen_dist,
              M.pen_scale, E, fs, flags);
    rec2E[e] = E;
    rec2F[e] = fs;
    note_flags(cfg_stat, ai.cfg, flags);
}

// Member positions and types of a cluster record.
template <int NB>
__device__ __forceinline__ void load_members(const int2 rec, const chx_atom_s *__restrict__ as,
                                             const int *__restrict__ nbr_off, const chx_entry_s *__restrict__ es,
                                             double (&p)[NB][3], int (&t)[NB], int (&code)[NB], int &cfg)
{
    const int i = rec.x;
    const chx_atom_s ai = as[i];
    cfg = ai.cfg;
    p[0][0] = ai.u[0]; p[0][1] = ai.u[1]; p[0][2] = ai.u[2];
    t[0] = ai.type;
    code[0] = 0;
    const int off = nbr_off[i] - 1;
#pragma unroll
    for (int m = 1; m < NB; m++) {
        const int s = (rec.y >> (10 * (m - 1))) & 1023;
        code[m] = s;
        const chx_entry_s g = es[off + s];
        p[m][0] = g.v[0]; p[m][1] = g.v[1]; p[m][2] = g.v[2];
        t[m] = g.type;
    }
}

// Reorders the members into parameter-file atom order: canonical member a is
// runtime member perm[a] (the first matching permutation, chimesFF.h:480-526).
template <int NB>
__device__ __forceinline__ void canonical(const int *__restrict__ perm_tab, int tup, double (&p)[NB][3],
                                          int (&code)[NB])
{
    int pm[NB];
#pragma unroll
    for (int a = 0; a < NB; a++) pm[a] = perm_tab[tup * NB + a];
    double q[NB][3];
    int c[NB];
#pragma unroll
    for (int a = 0; a < NB; a++) {
        c[a] = code[0];
        q[a][0] = p[0][0]; q[a][1] = p[0][1]; q[a][2] = p[0][2];
#pragma unroll
        for (int b = 1; b < NB; b++) {
            const bool s = pm[a] == b;
            c[a] = s ? code[b] : c[a];
            q[a][0] = s ? p[b][0] : q[a][0];
            q[a][1] = s ? p[b][1] : q[a][1];
            q[a][2] = s ? p[b][2] : q[a][2];
        }
    }
#pragma unroll
    for (int a = 0; a < NB; a++) {
        code[a] = c[a];
        p[a][0] = q[a][0]; p[a][1] = q[a][1]; p[a][2] = q[a][2];
    }
}

// ---------------------------------------------------------------------------
// 3-body: one thread per triplet of one cluster type (uniform coefficients).
// ---------------------------------------------------------------------------
template <int D, typename R>
__global__ __launch_bounds__(256) void k_eval3(long begin, long end, const int2 *__restrict__ list,
                                                const chx_cfg *__restrict__ cfgs, const chx_atom_s *__restrict__ as,
                                                const int *__restrict__ nbr_off, const chx_entry_s *__restrict__ es,
                                                chx_type_args3 ta, const R *__restrict__ Cd, double *recE,
                                                double *recF, int *recM, int *cfg_stat)
{
    const long c = begin + (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= end) return;
    double p[3][3];
    int t[3], code[3], cfg;
    load_members<3>(list[c], as, nbr_off, es, p, t, code, cfg);
    const int tup = (t[0] * ta.nt + t[1]) * ta.nt + t[2];
    canonical<3>(ta.perm, tup, p, code);
    const chx_cfg &C = cfgs[cfg];
    double r[3], dr[3];
    r[0] = pair_vec(C.h, p[0], p[1], dr);
    r[1] = pair_vec(C.h, p[0], p[2], dr);
    r[2] = pair_vec(C.h, p[1], p[2], dr);
    recM[c] = code[0] | (code[1] << 10) | (code[2] << 20);
    bool in = true;
#pragma unroll
    for (int q = 0; q < 3; q++) in = in && (r[q] < ta.s[q].rout);
    if (!in) {
        recE[c] = 0.0;
        recF[3 * c] = 0.0; recF[3 * c + 1] = 0.0; recF[3 * c + 2] = 0.0;
        return;
    }
    R T[3][D], dT[3][D];
    double f[3], fd[3];
    int flags = 0;
#pragma unroll
    for (int q = 0; q < 3; q++) {
        cheby<D, R>(r[q], ta.s[q], T[q], dT[q], flags);
        fcut(r[q], ta.s[q].rout, ta.tersoff, ta.fcut_var, f[q], fd[q]);
    }
    double S, G[3], E, fs[3];
    contract3<D, R, double>(Cd, T, dT, S, G);
    cluster_scalars<3, double>(r, f, fd, S, G, E, fs);
    recE[c] = E;
    recF[3 * c] = fs[0]; recF[3 * c + 1] = fs[1]; recF[3 * c + 2] = fs[2];
    note_flags(cfg_stat, cfg, flags);
}

// ---------------------------------------------------------------------------
// 4-body: one thread per quadruplet of one cluster type.
// ---------------------------------------------------------------------------
template <int D, typename R>
__global__ __launch_bounds__(256) void k_eval4(long begin, long end, const int2 *__restrict__ list,
                                                const chx_cfg *__restrict__ cfgs, const chx_atom_s *__restrict__ as,
                                                const int *__restrict__ nbr_off, const chx_entry_s *__restrict__ es,
                                                chx_type_args4 ta, const R *__restrict__ Cd, double *recE,
                                                double *recF, long long *recM, int *cfg_stat)
{
    const long c = begin + (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= end) return;
    double p[4][3];
    int t[4], code[4], cfg;
    load_members<4>(list[c], as, nbr_off, es, p, t, code, cfg);
    const int tup = ((t[0] * ta.nt + t[1]) * ta.nt + t[2]) * ta.nt + t[3];
    canonical<4>(ta.perm, tup, p, code);
    const chx_cfg &C = cfgs[cfg];
    double r[6], dr[3];
    r[0] = pair_vec(C.h, p[0], p[1], dr);
    r[1] = pair_vec(C.h, p[0], p[2], dr);
    r[2] = pair_vec(C.h, p[0], p[3], dr);
    r[3] = pair_vec(C.h, p[1], p[2], dr);
    r[4] = pair_vec(C.h, p[1], p[3], dr);
    r[5] = pair_vec(C.h, p[2], p[3], dr);
    recM[c] = (long long)code[0] | ((long long)code[1] << 10) | ((long long)code[2] << 20) |
              ((long long)code[3] << 30);
    bool in = true;
#pragma unroll
    for (int q = 0; q < 6; q++) in = in && (r[q] < ta.s[q].rout);
    if (!in) {
        recE[c] = 0.0;
#pragma unroll
        for (int q = 0; q < 6; q++) recF[6 * c + q] = 0.0;
        return;
    }
    R T[6][D], dT[6][D];
    double f[6], fd[6];
    int flags = 0;
#pragma unroll
    for (int q = 0; q < 6; q++) {
        cheby<D, R>(r[q], ta.s[q], T[q], dT[q], flags);
        fcut(r[q], ta.s[q].rout, ta.tersoff, ta.fcut_var, f[q], fd[q]);
    }
    double S, G[6], E, fs[6];
    contract4<D, R, double>(Cd, T, dT, S, G);
    cluster_scalars<6, double>(r, f, fd, S, G, E, fs);
    recE[c] = E;
#pragma unroll
    for (int q = 0; q < 6; q++) recF[6 * c + q] = fs[q];
    note_flags(cfg_stat, cfg, flags);
}

// ---------------------------------------------------------------------------
// Owner reduction: one 64-lane workgroup per owner atom.  The pair force
// scalars of every cluster the owner enumerated are summed into cells (s, t)
// of its local slot pairs (slot 0 = owner, slot s = neighbour s-1) in list
// order, each cell by one fixed lane; forces F_s += cell * h (v_t - v_s) and
// the virial -cell * dr (x) dr follow per slot and per cell.  Everything is
// summed in a fixed order, so results do not depend on scheduling.
// ---------------------------------------------------------------------------
#define RTILE 4096

__device__ __forceinline__ int tri(int s, int t)
{
    const int lo = s < t ? s : t, hi = s < t ? t : s;
    return hi * (hi - 1) / 2 + lo;
}

__device__ __forceinline__ void scan_chunk(const int *cidx, const double *cval, int np, int t0, int lane,
                                           double *cells)
{
    for (int p = 0; p < np; p++) {
        const int c = cidx[p] - t0;
        if (c >= 0 && c < RTILE && (c & 63) == lane) cells[c] += cval[p];
    }
}

__global__ __launch_bounds__(64) void k_reduce(int natoms, const chx_cfg *__restrict__ cfgs,
                                                const chx_atom_s *__restrict__ as, const int *__restrict__ nbr_off,
                                                const chx_entry_s *__restrict__ es,
                                                const unsigned char *__restrict__ e_in3, const chx_red_args ra,
                                                double *Fe, double *Fself, double *Eown, double *Sown)
{
    __shared__ double cells[RTILE];
    __shared__ double pos[CHX_MAX_NB + 1][3];
    __shared__ short s2m[CHX_MAX_NB + 1];     // slot -> many-body index (0 = owner), -1 if none
    __shared__ short m2s[CHX_MAX_N3 + 1];     // many-body index -> slot
    __shared__ int cidx[64 * 6];
    __shared__ double cval[64 * 6];
    __shared__ int nmb;
    const int i = blockIdx.x;
    const int lane = threadIdx.x;
    const int off = nbr_off[i];
    const int nb = nbr_off[i + 1] - off;
    const chx_atom_s ai = as[i];
    const chx_cfg &C = cfgs[ai.cfg];
    const chx_red_model &M = ra.models[C.model];
    const long N = natoms;
    for (int s = lane; s <= nb; s += 64) {
        if (s == 0) { pos[0][0] = ai.u[0]; pos[0][1] = ai.u[1]; pos[0][2] = ai.u[2]; }
        else {
            const chx_entry_s g = es[off + s - 1];
            pos[s][0] = g.v[0]; pos[s][1] = g.v[1]; pos[s][2] = g.v[2];
        }
    }
    // many-body slot set: owner + images flagged by the enumeration, in slot order
    if (lane == 0) { nmb = 1; s2m[0] = 0; m2s[0] = 0; }
    __syncthreads();
    for (int base = 1; base <= nb; base += 64) {
        const int s = base + lane;
        const bool in = s <= nb && e_in3[off + s - 1];
        const unsigned long long b = __ballot(in);
        const unsigned long long below = lane ? (~0ull >> (64 - lane)) : 0ull;
        const int m0 = nmb;
        if (s <= nb) {
            const int m = in ? m0 + __popcll(b & below) : -1;
            s2m[s] = (short)m;
            if (in) m2s[m] = (short)s;
        }
        __syncthreads();
        if (lane == 0) nmb = m0 + __popcll(b);
        __syncthreads();
    }
    const int nm = nmb;
    // 2-body: direct, slot s gets -f dr, the owner +f dr
    double F2[CHX_MAX_NB / 64][3];
#pragma unroll
    for (int q = 0; q < CHX_MAX_NB / 64; q++) F2[q][0] = F2[q][1] = F2[q][2] = 0.0;
    double F0[3] = {0.0, 0.0, 0.0};
    double Sv[6] = {0, 0, 0, 0, 0, 0};
    double Eacc = 0.0;
#pragma unroll
    for (int q = 0; q < CHX_MAX_NB / 64; q++) {
        const int s = 1 + lane + 64 * q;
        if (s > nb) continue;
        const double f = ra.rec2F[off + s - 1];
        Eacc += ra.rec2E[off + s - 1];
        if (f == 0.0) continue;
        double dr[3];
        pair_vec(C.h, pos[0], pos[s], dr);
        F0[0] += f * dr[0]; F0[1] += f * dr[1]; F0[2] += f * dr[2];
        F2[q][0] = -f * dr[0]; F2[q][1] = -f * dr[1]; F2[q][2] = -f * dr[2];
        Sv[0] -= f * (dr[0] * dr[0]);
        Sv[1] -= f * (dr[0] * dr[1]);
        Sv[2] -= f * (dr[0] * dr[2]);
        Sv[3] -= f * (dr[1] * dr[1]);
        Sv[4] -= f * (dr[1] * dr[2]);
        Sv[5] -= f * (dr[2] * dr[2]);
    }
    // many-body: pair cells over the many-body slot set
    double Fm[(CHX_MAX_N3 + 64) / 64][3];
#pragma unroll
    for (int q = 0; q < (CHX_MAX_N3 + 64) / 64; q++) Fm[q][0] = Fm[q][1] = Fm[q][2] = 0.0;
    const int ncell = nm * (nm - 1) / 2;
    const bool has_mb = (ra.off3 && M.ntrip) || (ra.off4 && M.nquad);
    for (int t0 = 0; has_mb && t0 < ncell; t0 += RTILE) {
        for (int c = lane; c < RTILE; c += 64) cells[c] = 0.0;
        __syncthreads();
        if (ra.off3) {
            for (int g = M.g3; g < M.g3 + M.ntrip; g++) {
                const long b = ra.off3[g * N + i], e = ra.off3[g * N + i + 1];
                for (long cb = b; cb < e; cb += 64) {
                    const long c = cb + lane;
                    if (c < e) {
                        const int m = ra.rec3M[c];
                        const int s0 = s2m[m & 1023], s1 = s2m[(m >> 10) & 1023], s2 = s2m[(m >> 20) & 1023];
                        cidx[3 * lane] = tri(s0, s1);
                        cidx[3 * lane + 1] = tri(s0, s2);
                        cidx[3 * lane + 2] = tri(s1, s2);
                        cval[3 * lane] = ra.rec3F[3 * c];
                        cval[3 * lane + 1] = ra.rec3F[3 * c + 1];
                        cval[3 * lane + 2] = ra.rec3F[3 * c + 2];
                        if (t0 == 0) Eacc += ra.rec3E[c];
                    } else {
                        cidx[3 * lane] = cidx[3 * lane + 1] = cidx[3 * lane + 2] = -1;
                    }
                    __syncthreads();
                    scan_chunk(cidx, cval, 3 * 64, t0, lane, cells);
                    __syncthreads();
                }
            }
        }
        if (ra.off4) {
            for (int g = M.g4; g < M.g4 + M.nquad; g++) {
                const long b = ra.off4[g * N + i], e = ra.off4[g * N + i + 1];
                for (long cb = b; cb < e; cb += 64) {
                    const long c = cb + lane;
                    if (c < e) {
                        const long long m = ra.rec4M[c];
                        const int s[4] = {s2m[m & 1023], s2m[(m >> 10) & 1023], s2m[(m >> 20) & 1023],
                                          s2m[(m >> 30) & 1023]};
                        cidx[6 * lane] = tri(s[0], s[1]);
                        cidx[6 * lane + 1] = tri(s[0], s[2]);
                        cidx[6 * lane + 2] = tri(s[0], s[3]);
                        cidx[6 * lane + 3] = tri(s[1], s[2]);
                        cidx[6 * lane + 4] = tri(s[1], s[3]);
                        cidx[6 * lane + 5] = tri(s[2], s[3]);
#pragma unroll
                        for (int q = 0; q < 6; q++) cval[6 * lane + q] = ra.rec4F[6 * c + q];
                        if (t0 == 0) Eacc += ra.rec4E[c];
                    } else {
#pragma unroll
                        for (int q = 0; q < 6; q++) cidx[6 * lane + q] = -1;
                    }
                    __syncthreads();
                    scan_chunk(cidx, cval, 6 * 64, t0, lane, cells);
                    __syncthreads();
                }
            }
        }
        // forces on the many-body slots and the virial from this tile
#pragma unroll
        for (int q = 0; q < (CHX_MAX_N3 + 64) / 64; q++) {
            const int a = lane + 64 * q;
            if (a >= nm) continue;
            const int sa = m2s[a];
            for (int b = 0; b < nm; b++) {
                if (b == a) continue;
                const int c = tri(a, b) - t0;
                if (c < 0 || c >= RTILE) continue;
                const double v = cells[c];
                if (v == 0.0) continue;
                double dr[3];
                pair_vec(C.h, pos[sa], pos[m2s[b]], dr);
                Fm[q][0] += v * dr[0];
                Fm[q][1] += v * dr[1];
                Fm[q][2] += v * dr[2];
                if (a < b) {
                    Sv[0] -= v * (dr[0] * dr[0]);
                    Sv[1] -= v * (dr[0] * dr[1]);
                    Sv[2] -= v * (dr[0] * dr[2]);
                    Sv[3] -= v * (dr[1] * dr[1]);
                    Sv[4] -= v * (dr[1] * dr[2]);
                    Sv[5] -= v * (dr[2] * dr[2]);
                }
            }
        }
        __syncthreads();
    }
    // many-body forces to LDS (reusing the cell tile), then per-slot totals
    double *fm = cells;
    for (int a = lane; a < 3 * nm && a < RTILE; a += 64) fm[a] = 0.0;
    __syncthreads();
#pragma unroll
    for (int q = 0; q < (CHX_MAX_N3 + 64) / 64; q++) {
        const int a = lane + 64 * q;
        if (a >= nm || !has_mb) continue;
        fm[3 * a] = Fm[q][0]; fm[3 * a + 1] = Fm[q][1]; fm[3 * a + 2] = Fm[q][2];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < CHX_MAX_NB / 64; q++) {
        const int s = 1 + lane + 64 * q;
        if (s > nb) continue;
        const int m = s2m[s];
        double f[3] = {F2[q][0], F2[q][1], F2[q][2]};
        if (m > 0) { f[0] += fm[3 * m]; f[1] += fm[3 * m + 1]; f[2] += fm[3 * m + 2]; }
        double *dst = Fe + 3L * (off + s - 1);
        dst[0] = f[0]; dst[1] = f[1]; dst[2] = f[2];
    }
    // owner force, energy and virial: fixed-pattern wave reductions
    double red[10] = {Eacc, Sv[0], Sv[1], Sv[2], Sv[3], Sv[4], Sv[5], F0[0], F0[1], F0[2]};
#pragma unroll
    for (int w = 32; w >= 1; w >>= 1)
#pragma unroll
        for (int k = 0; k < 10; k++) red[k] += __shfl_xor(red[k], w, 64);
    if (lane == 0) {
        Eown[i] = red[0] + M.eoff[ai.type];
#pragma unroll
        for (int k = 0; k < 6; k++) Sown[6L * i + k] = red[1 + k];
        Fself[3L * i] = red[7] + fm[0];
        Fself[3L * i + 1] = red[8] + fm[1];
        Fself[3L * i + 2] = red[9] + fm[2];
    }
}

// Force on replicated atom r: its own owner force plus every entry whose
// parent is r, in entry order.
__global__ void k_gather_rep(int natoms, const int *__restrict__ rstart, const int *__restrict__ rev,
                             const double *__restrict__ Fe, const double *__restrict__ Fself, double *Frep)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= natoms) return;
    double f[3] = {Fself[3L * r], Fself[3L * r + 1], Fself[3L * r + 2]};
    for (int p = rstart[r]; p < rstart[r + 1]; p++) {
        const long e = rev[p];
        f[0] += Fe[3 * e]; f[1] += Fe[3 * e + 1]; f[2] += Fe[3 * e + 2];
    }
    Frep[3L * r] = f[0]; Frep[3L * r + 1] = f[1]; Frep[3L * r + 2] = f[2];
}

// Force on input atom a: sum over its replicas in replica order.
__global__ void k_gather_in(int natoms_in, int ncfg, const int *__restrict__ cfg_in_off,
                            const chx_cfg *__restrict__ cfgs, const double *__restrict__ Frep, double *Fout)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= natoms_in) return;
    int lo = 0, hi = ncfg - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (cfg_in_off[mid] <= a) lo = mid; else hi = mid - 1;
    }
    const chx_cfg &C = cfgs[lo];
    const int la = a - C.atom_in_off;
    const int ncopy = (C.n_rep + 1) * (C.n_rep + 1) * (C.n_rep + 1);
    double f[3] = {0.0, 0.0, 0.0};
    for (int rho = 0; rho < ncopy; rho++) {
        const long r = C.atom_off + (long)rho * C.natoms_in + la;
        f[0] += Frep[3 * r]; f[1] += Frep[3 * r + 1]; f[2] += Frep[3 * r + 2];
    }
    Fout[3L * a] = f[0]; Fout[3L * a + 1] = f[1]; Fout[3L * a + 2] = f[2];
}

// Energy and virial of each configuration: fixed-order block reduction.
__global__ __launch_bounds__(256) void k_cfg_sum(const chx_cfg *__restrict__ cfgs, const double *__restrict__ Eown,
                                                  const double *__restrict__ Sown, double *Ecfg, double *Vcfg)
{
    __shared__ double sm[7][256];
    const chx_cfg &C = cfgs[blockIdx.x];
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int r = C.atom_off + threadIdx.x; r < C.atom_off + C.natoms; r += 256) {
        acc[0] += Eown[r];
#pragma unroll
        for (int k = 0; k < 6; k++) acc[1 + k] += Sown[6L * r + k];
    }
#pragma unroll
    for (int k = 0; k < 7; k++) sm[k][threadIdx.x] = acc[k];
    __syncthreads();
    for (int w = 128; w >= 1; w >>= 1) {
        if ((int)threadIdx.x < w)
#pragma unroll
            for (int k = 0; k < 7; k++) sm[k][threadIdx.x] += sm[k][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        Ecfg[blockIdx.x] = sm[0][0];
#pragma unroll
        for (int k = 0; k < 6; k++) Vcfg[6L * blockIdx.x + k] = sm[1 + k][0];
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
__global__ void k_clusters(int n, const double *__restrict__ dist, const double *__restrict__ drv,
                           const int *__restrict__ types, chx_model_dev M, double *E, double *F, double *Sout,
                           double *last, int *flags_out)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    constexpr int P = NB * (NB - 1) / 2;
    int pa[P], pb[P];
    {
        int q = 0;
        for (int a = 0; a < NB; a++)
            for (int b = a + 1; b < NB; b++) { pa[q] = a; pb[q] = b; q++; }
    }
    const int *ty = types + (long)c * NB;
    double Ec = 0.0;
    double fs_rt[P];
    for (int p = 0; p < P; p++) fs_rt[p] = 0.0;
    int flags = 0;
    bool evaluated = false;
    if constexpr (NB == 2) {
        const int pt = M.pairmap[ty[0] * M.nt + ty[1]];
        double fs;
        const double r = dist[c];
        eval_pair(r, M.slots2[pt], M.c2 + (long)pt * (M.order[0] + 1), M.order[0], M.tersoff, M.fcut_var,
                  M.pen_dist, M.pen_scale, Ec, fs, flags);
        fs_rt[0] = fs;
        evaluated = r < M.slots2[pt].rout;
        if (evaluated) {
            double ls = 1.0;
            if (r - M.pen_dist < M.slots2[pt].rin) {
                const double rp = M.slots2[pt].rin + M.pen_dist - r;
                if (rp > 0.0) ls = (-3.0 * rp * rp * M.pen_scale) / r;
            }
            last[c] = ls;
        }
    } else {
        int tup = 0;
        for (int a = 0; a < NB; a++) tup = tup * M.nt + ty[a];
        const int tau = NB == 3 ? M.tripmap[tup] : M.quadmap[tup];
        if (tau >= 0) {
            const int *perm = (NB == 3 ? M.trip_perm : M.quad_perm) + tup * NB;
            const chx_slot *sl = (NB == 3 ? M.slots3 : M.slots4) + tau * P;
            int rt[P];
            double rq[P];
            bool in = true;
            for (int q = 0; q < P; q++) {
                const int a = perm[pa[q]], b = perm[pb[q]];
                const int lo = a < b ? a : b, hi = a < b ? b : a;
                int idx = 0;
                for (int x = 0; x < lo; x++) idx += NB - 1 - x;
                idx += hi - lo - 1;
                rt[q] = idx;
                rq[q] = dist[(long)c * P + idx];
                in = in && rq[q] < sl[q].rout;
            }
            if (in) {
                evaluated = true;
                double T[P][D], dT[P][D], f[P], fd[P];
                for (int q = 0; q < P; q++) {
                    cheby<D, double>(rq[q], sl[q], T[q], dT[q], flags);
                    fcut(rq[q], sl[q].rout, M.tersoff, M.fcut_var, f[q], fd[q]);
                }
                double S, G[P], fs[P];
                if constexpr (NB == 3) contract3<D, double, double>(M.C3 + (long)tau * D * D * D, T, dT, S, G);
                else contract4<D, double, double>(M.C4 + (long)tau * D * D * D * D * D * D, T, dT, S, G);
                cluster_scalars<P, double>(rq, f, fd, S, G, Ec, fs);
                for (int q = 0; q < P; q++) fs_rt[rt[q]] = fs[q];
                // last term of the list, in runtime pair order
                const int *lp = (NB == 3 ? M.last3 : M.last4) + tau * P;
                const double lc = (NB == 3 ? M.lastc3 : M.lastc4)[tau];
                for (int q = 0; q < P; q++) {
                    double v = lc * (f[q] * dT[q][lp[q]] + fd[q] * T[q][lp[q]]);
                    double other = 1.0;
                    for (int k = 0; k < P; k++)
                        if (k != q) other *= f[k];
                    v *= other / rq[q];
                    for (int k = 0; k < P; k++)
                        if (k != q) v *= T[k][lp[k]];
                    last[(long)c * P + rt[q]] = v;
                }
            }
        }
    }
    double fo[NB][3];
    for (int a = 0; a < NB; a++) fo[a][0] = fo[a][1] = fo[a][2] = 0.0;
    double So[6] = {0, 0, 0, 0, 0, 0};
    for (int p = 0; p < P; p++) {
        const double fs = fs_rt[p];
        const double *d = drv + ((long)c * P + p) * 3;
        fo[pa[p]][0] += fs * d[0]; fo[pa[p]][1] += fs * d[1]; fo[pa[p]][2] += fs * d[2];
        fo[pb[p]][0] -= fs * d[0]; fo[pb[p]][1] -= fs * d[1]; fo[pb[p]][2] -= fs * d[2];
        So[0] -= fs * (d[0] * d[0]);
        So[1] -= fs * (d[0] * d[1]);
        So[2] -= fs * (d[0] * d[2]);
        So[3] -= fs * (d[1] * d[1]);
        So[4] -= fs * (d[1] * d[2]);
        So[5] -= fs * (d[2] * d[2]);
    }
    E[c] = Ec;
    for (int a = 0; a < NB; a++)
        for (int k = 0; k < 3; k++) F[((long)c * NB + a) * 3 + k] = fo[a][k];
    for (int k = 0; k < 6; k++) Sout[(long)c * 6 + k] = So[k];
    flags_out[c] = flags | (evaluated ? 8 : 0);
}

// ---------------------------------------------------------------------------
// Host-side launchers: dispatch on the dense dimension D and precision.
// ---------------------------------------------------------------------------
#define CHX_D3_CASES(X) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13)
#define CHX_D4_CASES(X) X(2) X(3) X(4) X(5) X(6)

bool supported_D3(int D) { return D >= 2 && D <= 13; }
bool supported_D4(int D) { return D >= 2 && D <= 6; }

void launch_eval3(int D, bool mixed, long begin, long end, hipStream_t st, const int2 *list, const chx_cfg *cfgs,
                  const chx_atom_s *as, const int *nbr_off, const chx_entry_s *es, const chx_type_args3 &ta,
                  const void *Cd, double *recE, double *recF, int *recM, int *cfg_stat)
{
    const long n = end - begin;
    if (n <= 0) return;
    const dim3 grid((unsigned)((n + 255) / 256)), block(256);
#define X(d)                                                                                                    \
    if (D == d) {                                                                                               \
        if (mixed)                                                                                              \
            hipLaunchKernelGGL((k_eval3<d, float>), grid, block, 0, st, begin, end, list, cfgs, as, nbr_off, es, \
                               ta, (const float *)Cd, recE, recF, recM, cfg_stat);                              \
        else                                                                                                    \
            hipLaunchKernelGGL((k_eval3<d, double>), grid, block, 0, st, begin, end, list, cfgs, as, nbr_off,    \
                               es, ta, (const double *)Cd, recE, recF, recM, cfg_stat);                         \
        return;                                                                                                 \
    }
    CHX_D3_CASES(X)
#undef X
}

void launch_eval4(int D, bool mixed, long begin, long end, hipStream_t st, const int2 *list, const chx_cfg *cfgs,
                  const chx_atom_s *as, const int *nbr_off, const chx_entry_s *es, const chx_type_args4 &ta,
                  const void *Cd, double *recE, double *recF, long long *recM, int *cfg_stat)
{
    const long n = end - begin;
    if (n <= 0) return;
    const dim3 grid((unsigned)((n + 255) / 256)), block(256);
#define X(d)                                                                                                    \
    if (D == d) {                                                                                               \
        if (mixed)                                                                                              \
            hipLaunchKernelGGL((k_eval4<d, float>), grid, block, 0, st, begin, end, list, cfgs, as, nbr_off, es, \
                               ta, (const float *)Cd, recE, recF, recM, cfg_stat);                              \
        else                                                                                                    \
            hipLaunchKernelGGL((k_eval4<d, double>), grid, block, 0, st, begin, end, list, cfgs, as, nbr_off,    \
                               es, ta, (const double *)Cd, recE, recF, recM, cfg_stat);                         \
        return;                                                                                                 \
    }
    CHX_D4_CASES(X)
#undef X
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

