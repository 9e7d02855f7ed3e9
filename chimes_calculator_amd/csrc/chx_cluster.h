// One explicit cluster of NB = 2, 3 or 4 bodies evaluated by one thread, for
// callers that bring their own clusters: the per-cluster API (compute_2B /
// compute_3B / compute_4B, chimesFF.cpp:1657-1775, 1944-2176, 2509-2872) and
// the LAMMPS pair style (pair_chimes.cpp:567-829).
#pragma once
#include "eval_math.h"

// Dense dimensions D = order + 1 instantiated for the 3- and 4-body kernels.
#define CHX_D3_CASES(X) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13)
#define CHX_D4_CASES(X) X(2) X(3) X(4) X(5) X(6)

namespace chx {

// a[idx] of a small register array, by selects (no indexed private memory)
template <int P>
__device__ __forceinline__ double pick(const double (&a)[P], int idx)
{
    double v = a[0];
#pragma unroll
    for (int t = 1; t < P; t++) v = idx == t ? a[t] : v;
    return v;
}

// Canonical pair slot q of a cluster type and the runtime pair rt it takes:
// canonical member a is runtime member perm[a] (first-match permutation,
// chimesFF.h:480-526), runtime pairs in the order 01 02 03 12 13 23.
template <int NB>
__device__ __forceinline__ int runtime_pair(const int *perm, int q)
{
    constexpr int pa[6] = {0, 0, 0, 1, 1, 2}, pb[6] = {1, 2, 3, 2, 3, 3};
    constexpr int pa3[3] = {0, 0, 1}, pb3[3] = {1, 2, 2};
    const int a = perm[NB == 3 ? pa3[q] : pa[q]], b = perm[NB == 3 ? pb3[q] : pb[q]];
    const int lo = a < b ? a : b, hi = a < b ? b : a;
    int idx = 0;
    for (int x = 0; x < lo; x++) idx += NB - 1 - x;
    return idx + hi - lo - 1;
}

// Whether compute_nB evaluates a cluster of types ty with runtime pair
// distances dist: its type is not excluded and every pair is below the outer
// cutoff of its slot (chimesFF.cpp:1691-1692, 1989-2012, 2575-2600).
template <int NB>
__device__ __forceinline__ bool cluster_inside(const chx_model_dev &M, const int (&ty)[NB],
                                               const double (&dist)[NB * (NB - 1) / 2])
{
    constexpr int P = NB * (NB - 1) / 2;
    if constexpr (NB == 2) {
        return dist[0] < M.slots2[M.pairmap[ty[0] * M.nt + ty[1]]].rout;
    } else {
        int tup = 0;
        for (int a = 0; a < NB; a++) tup = tup * M.nt + ty[a];
        const int tau = NB == 3 ? M.tripmap[tup] : M.quadmap[tup];
        if (tau < 0) return false;
        const int *perm = (NB == 3 ? M.trip_perm : M.quad_perm) + tup * NB;
        const chx_slot *sl = (NB == 3 ? M.slots3 : M.slots4) + tau * P;
        bool in = true;
        for (int q = 0; q < P; q++) {
            const int rt = runtime_pair<NB>(perm, q);
            double r = dist[0];
            for (int t = 1; t < P; t++) r = rt == t ? dist[t] : r;
            in = in && r < sl[q].rout;
        }
        return in;
    }
}

// dist[P] / drv[P][3]: pair distances and vectors x_b - x_a in runtime pair
// order (01 02 03 12 13 23), ty[NB] the members' type indices.  Returns the
// energy Ec, member forces fo, the stress So (xx xy xz yy yz zz, -sum f dr dr),
// whether the cluster was inside every cutoff, and, when last is not null,
// for 2 bodies the penalty scalar (chimesFF.cpp:1745-1774) and for 3/4 bodies
// the force scalars of the last term of the type's list in runtime pair
// order (compute_3B/4B's force_scalar_in, chimesFF.cpp:2171-2173, 2864-2869).
template <int NB, int D>
__device__ __forceinline__ void explicit_cluster(const chx_model_dev &M, const int (&ty)[NB],
                                                 const double (&dist)[NB * (NB - 1) / 2],
                                                 const double (&drv)[NB * (NB - 1) / 2][3], double &Ec,
                                                 double (&fo)[NB][3], double (&So)[6], int &flags, bool &evaluated,
                                                 double *last)
{
    constexpr int P = NB * (NB - 1) / 2;
    int pa[P], pb[P];
    {
        int q = 0;
        for (int a = 0; a < NB; a++)
            for (int b = a + 1; b < NB; b++) { pa[q] = a; pb[q] = b; q++; }
    }
    Ec = 0.0;
    evaluated = false;
    double fs_rt[P];
#pragma unroll
    for (int p = 0; p < P; p++) fs_rt[p] = 0.0;
    if constexpr (NB == 2) {
        const int pt = M.pairmap[ty[0] * M.nt + ty[1]];
        double fs;
        const double r = dist[0];
        eval_pair(r, M.slots2[pt], M.c2 + (long)pt * (M.order[0] + 1), M.order[0], M.tersoff, M.fcut_var,
                  M.pen_dist, M.pen_scale, Ec, fs, flags);
        fs_rt[0] = fs;
        evaluated = r < M.slots2[pt].rout;
        if (evaluated && last) {
            double ls = 1.0;
            if (r - M.pen_dist < M.slots2[pt].rin) {
                const double rp = M.slots2[pt].rin + M.pen_dist - r;
                if (rp > 0.0) ls = (-3.0 * rp * rp * M.pen_scale) / r;
            }
            last[0] = ls;
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
                rq[q] = pick<P>(dist, idx);
                in = in && rq[q] < sl[q].rout;
            }
            if (in) {
                evaluated = true;
                double T[P][D], dT[P][D], f[P], fd[P];
                for (int q = 0; q < P; q++) {
                    cheby<D, double>(rq[q], sl[q], T[q], dT[q], flags);
                    fcut(rq[q], sl[q], M.tersoff, M.fcut_var, f[q], fd[q]);
                }
                double S, G[P], fs[P];
                // when every active lane has the same cluster type (always for
                // one-element models) the coefficient address is wave-uniform
                // and the contraction reads it through scalar loads
                const int tu = __builtin_amdgcn_readfirstlane(tau);
                const long csz = NB == 3 ? (long)D * D * D : (long)D * D * D * D * D * D;
                if (__ballot(tau != tu) == 0) {
                    if constexpr (NB == 3) contract3<D, double, double>(M.C3 + tu * csz, T, dT, S, G);
                    else contract4<D, double, double>(M.C4 + tu * csz, T, dT, S, G);
                } else {
                    if constexpr (NB == 3) contract3<D, double, double>(M.C3 + tau * csz, T, dT, S, G);
                    else contract4<D, double, double>(M.C4 + tau * csz, T, dT, S, G);
                }
                cluster_scalars<P, double>(rq, f, fd, S, G, Ec, fs);
#pragma unroll
                for (int q = 0; q < P; q++)
#pragma unroll
                    for (int t = 0; t < P; t++)
                        if (rt[q] == t) fs_rt[t] = fs[q];
                if (last) {
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
                        last[rt[q]] = v;
                    }
                }
            }
        }
    }
    for (int a = 0; a < NB; a++) fo[a][0] = fo[a][1] = fo[a][2] = 0.0;
    for (int k = 0; k < 6; k++) So[k] = 0.0;
#pragma unroll
    for (int p = 0; p < P; p++) {
        const double fs = fs_rt[p];
        const double *d = drv[p];
        fo[pa[p]][0] += fs * d[0]; fo[pa[p]][1] += fs * d[1]; fo[pa[p]][2] += fs * d[2];
        fo[pb[p]][0] -= fs * d[0]; fo[pb[p]][1] -= fs * d[1]; fo[pb[p]][2] -= fs * d[2];
        So[0] -= fs * (d[0] * d[0]);
        So[1] -= fs * (d[0] * d[1]);
        So[2] -= fs * (d[0] * d[2]);
        So[3] -= fs * (d[1] * d[1]);
        So[4] -= fs * (d[1] * d[2]);
        So[5] -= fs * (d[2] * d[2]);
    }
}

}  // namespace chx
