// Device math of one ChIMES cluster: Morse-transformed Chebyshev polynomials,
// smooth cutoffs, the 2-body sum with its penalty, and the 3-/4-body
// coefficient contraction.
//
// Semantics follow chimesFF (chimesFF/src/chimesFF.h:347-628,
// chimesFF/src/chimesFF.cpp:1572-1613, 1657-1775, 1944-2176, 2509-2872).
// The 3-/4-body term lists are held as dense coefficient tensors C[a_0..a_P-1]
// in parameter pair order and contracted one pair index at a time, carrying
// the energy form S = sum C prod_q T_q[a_q] and its derivatives
// G_q = sum C dT_q[a_q] prod_{r!=q} T_r[a_r] together; the reference's per-term
// force scalars sum to (prod_{r!=q} f_r / r_q) (f_q G_q + f'_q S).
#pragma once
#include <hip/hip_runtime.h>

#include "chx_types.h"

#define CHX_PI 3.14159265359          // truncated constant of chimesFF.h:22
#define CHX_SMOOTH 0.01               // inner_smooth_distance, chimesFF.cpp:113

namespace chx {

enum : int { FLAG_PENALTY = 1, FLAG_BELOW_RIN = 2, FLAG_OUT_OF_RANGE = 4 };

// Chebyshev polynomials T_0..T_{D-1} of the Morse variable and dT/dr
// (set_cheby_polys + set_polys_out_of_range).  R is the recurrence type.
template <int D, typename R>
__device__ __forceinline__ void cheby(double r, const chx_slot &s, R (&T)[D], R (&dT)[D], int &flags)
{
    double dx = r;
    const bool oor = dx < s.rin;
    if (oor) dx = s.rin;
    const double ex = exp(-dx * s.ilam);
    const double xd = (ex - s.xavg) * s.ixdiff;
    const double dxdr_d = (-ex * s.ilam) * s.ixdiff;
    const R x = (R)xd, dxdr = (R)dxdr_d, two_x = (R)(2.0 * xd);
    R U[D];
    T[0] = (R)1;
    U[0] = (R)1;
    if (D > 1) { T[1] = x; U[1] = two_x; }
#pragma unroll
    for (int i = 2; i < D; i++) {
        T[i] = two_x * T[i - 1] - T[i - 2];
        U[i] = two_x * U[i - 1] - U[i - 2];
    }
    dT[0] = (R)0;
#pragma unroll
    for (int i = 1; i < D; i++) dT[i] = (R)i * dxdr * U[i - 1];
    if (oor) {
        flags |= FLAG_OUT_OF_RANGE;
        const double damp = exp((r - s.rin) / CHX_SMOOTH);
        const R k = (R)(CHX_SMOOTH * (damp - 1.0)), dm = (R)damp;
#pragma unroll
        for (int i = 0; i < D; i++) {
            T[i] += k * dT[i];
            dT[i] *= dm;
        }
    }
}

// get_fcut: cubic (1 - r/rc)^3 or Tersoff's sine ramp, rc = the slot's outer
// cutoff; divisions by rc and by the ramp width are products with the
// slot's reciprocals.
__device__ __forceinline__ void fcut(double r, const chx_slot &s, int tersoff, double fvar, double &f, double &fd)
{
    const double rc = s.rout;
    if (!tersoff) {
        const double f0 = 1.0 - r * s.irout;
        f = f0 * f0 * f0;
        fd = f0 * f0 * (-3.0 * s.irout);
    } else {
        const double thr = rc - fvar * rc;
        if (r < thr) { f = 1.0; fd = 0.0; }
        else if (r > rc) { f = 0.0; fd = 0.0; }
        else {
            const double a = (r - thr) * s.iramp * CHX_PI + CHX_PI / 2.0;
            double sa, ca;
            sincos(a, &sa, &ca);
            f = 0.5 + 0.5 * sa;
            fd = 0.5 * ca * (CHX_PI * s.iramp);
        }
    }
}

// 2-body pair (compute_2B incl. get_penalty).  c2 points at the dense
// coefficients of the pair type by Chebyshev degree 0..order.  Returns the
// energy and the force scalar f with F_i += f dr, F_j -= f dr.
__device__ __forceinline__ void eval_pair(double r, const chx_slot &s, const double *__restrict__ c2, int order,
                                          int tersoff, double fvar, double pen_dist, double pen_scale,
                                          double &E, double &fs, int &flags)
{
    E = 0.0;
    fs = 0.0;
    if (r >= s.rout) return;
    double dx = r;
    const bool oor = dx < s.rin;
    if (oor) dx = s.rin;
    const double ex = exp(-dx * s.ilam);
    const double x = (ex - s.xavg) * s.ixdiff;
    const double dxdr = (-ex * s.ilam) * s.ixdiff;
    const double two_x = 2.0 * x;
    double Tm2 = 1.0, Tm1 = x, Um2 = 1.0, Um1 = two_x;
    double sT = c2[1] * x, sdT = c2[1] * dxdr;
    for (int n = 2; n <= order; n++) {
        const double Tn = two_x * Tm1 - Tm2;
        const double dTn = (double)n * dxdr * Um1;
        const double Un = two_x * Um1 - Um2;
        const double c = c2[n];
        sT = fma(c, Tn, sT);
        sdT = fma(c, dTn, sdT);
        Tm2 = Tm1; Tm1 = Tn;
        Um2 = Um1; Um1 = Un;
    }
    if (oor) {
        flags |= FLAG_OUT_OF_RANGE;
        const double damp = exp((r - s.rin) / CHX_SMOOTH);
        sT += CHX_SMOOTH * (damp - 1.0) * sdT;
        sdT *= damp;
    }
    double f, fd;
    fcut(r, s, tersoff, fvar, f, fd);
    const double rinv = (r > 0.0) ? 1.0 / r : 1e20;
    E = f * sT;
    fs = (f * sdT + fd * sT) * rinv;
    if (r - pen_dist < s.rin) {
        flags |= (r < s.rin) ? (FLAG_PENALTY | FLAG_BELOW_RIN) : FLAG_PENALTY;
        const double rp = s.rin + pen_dist - r;
        if (rp > 0.0) {
            E += rp * rp * rp * pen_scale;
            fs += (-3.0 * rp * rp * pen_scale) / r;
        }
    }
}

// ---------------------------------------------------------------------------
// Dense contraction of the 3-body form (pairs q = 0, 1, 2).
// ---------------------------------------------------------------------------
// Coefficient tensors are read-only for a whole launch: reading them through
// the constant address space lets wave-uniform reads use scalar loads.
template <typename R>
using const_ptr = const __attribute__((address_space(4))) R *;

// Index 0 of every pair is T_0 = 1 with dT_0 = 0 (cheby sets them exactly,
// damping included), so its products are plain copies and its derivative
// products vanish; the unrolled levels skip them (the same values: x * 1 and
// fma(x, 0, y) are exact).
template <int D, typename R, typename A>
__device__ __forceinline__ void contract3(const R *__restrict__ C, const R (&T)[3][D], const R (&dT)[3][D],
                                          A &S, A (&G)[3])
{
    A s0 = 0, g00 = 0, g01 = 0, g02 = 0;
    if constexpr (D <= 7) {
        // unrolled: every T[0] factor is a register of its own, the second
        // pair's factors are converted once, and the a0 = 0 level is a copy
        A t1[D], d1[D];
#pragma unroll
        for (int u = 0; u < D; u++) { t1[u] = (A)T[1][u]; d1[u] = (A)dT[1][u]; }
        auto level = [&](int a0, A t0, A dt0, bool first) {
            A s1 = 0, g11 = 0, g12 = 0;
#pragma unroll
            for (int a1 = 0; a1 < D; a1++) {
                const_ptr<R> c = (const_ptr<R>)C + (a0 * D + a1) * D;
                R z = c[0], z2 = 0;
#pragma unroll
                for (int a2 = 1; a2 < D; a2++) {
                    z = fma(c[a2], T[2][a2], z);
                    z2 = fma(c[a2], dT[2][a2], z2);
                }
                if (a1 == 0) {
                    s1 = (A)z;
                    g12 = (A)z2;
                } else {
                    s1 = fma(t1[a1], (A)z, s1);
                    g11 = fma(d1[a1], (A)z, g11);
                    g12 = fma(t1[a1], (A)z2, g12);
                }
            }
            if (first) {
                s0 = s1;
                g01 = g11;
                g02 = g12;
            } else {
                s0 = fma(t0, s1, s0);
                g00 = fma(dt0, s1, g00);
                g01 = fma(t0, g11, g01);
                g02 = fma(t0, g12, g02);
            }
        };
#pragma unroll
        for (int a0 = 0; a0 < D; a0++) level(a0, (A)T[0][a0], (A)dT[0][a0], a0 == 0);
    } else {
        // larger D: a loop over a0 (code size and registers), T[0] and dT[0]
        // consumed from the front of rotating copies, since indexing them by
        // a0 would place them in scratch
        R t0v[D], d0v[D];
#pragma unroll
        for (int u = 0; u < D; u++) { t0v[u] = T[0][u]; d0v[u] = dT[0][u]; }
#pragma unroll 1
        for (int a0 = 0; a0 < D; a0++) {
            A s1 = 0, g11 = 0, g12 = 0;
#pragma unroll
            for (int a1 = 0; a1 < D; a1++) {
                const_ptr<R> c = (const_ptr<R>)C + (a0 * D + a1) * D;
                R z = c[0], z2 = 0;
#pragma unroll
                for (int a2 = 1; a2 < D; a2++) {
                    z = fma(c[a2], T[2][a2], z);
                    z2 = fma(c[a2], dT[2][a2], z2);
                }
                if (a1 == 0) {
                    s1 = (A)z;
                    g12 = (A)z2;
                } else {
                    s1 = fma((A)T[1][a1], (A)z, s1);
                    g11 = fma((A)dT[1][a1], (A)z, g11);
                    g12 = fma((A)T[1][a1], (A)z2, g12);
                }
            }
            const A t0 = (A)t0v[0], dt0 = (A)d0v[0];
            s0 = fma(t0, s1, s0);
            g00 = fma(dt0, s1, g00);
            g01 = fma(t0, g11, g01);
            g02 = fma(t0, g12, g02);
#pragma unroll
            for (int u = 0; u + 1 < D; u++) { t0v[u] = t0v[u + 1]; d0v[u] = d0v[u + 1]; }
        }
    }
    S = s0;
    G[0] = g00; G[1] = g01; G[2] = g02;
}

// ---------------------------------------------------------------------------
// Dense contraction of the 4-body form (pairs q = 0..5 = ij ik il jk jl kl).
// ---------------------------------------------------------------------------
template <int D, typename R, typename A>
__device__ __forceinline__ void contract4(const R *__restrict__ C, const R (&T)[6][D], const R (&dT)[6][D],
                                          A &S, A (&G)[6])
{
    A s0 = 0, h0[6] = {0, 0, 0, 0, 0, 0};
    // the (a2, a3, a4, a5) sums of one (a0, a1): s2 and the gradients of
    // pairs 2..5 (h2)
    auto inner = [&](int a0, int a1, R &s2, R (&h2)[4]) {
        s2 = 0;
        h2[0] = h2[1] = h2[2] = h2[3] = 0;
#pragma unroll
        for (int a2 = 0; a2 < D; a2++) {
            R s3 = 0, h3[3] = {0, 0, 0};
#pragma unroll
            for (int a3 = 0; a3 < D; a3++) {
                R s4 = 0, h4[2] = {0, 0};
#pragma unroll
                for (int a4 = 0; a4 < D; a4++) {
                    const_ptr<R> c = (const_ptr<R>)C + ((((a0 * D + a1) * D + a2) * D + a3) * D + a4) * D;
                    R z = c[0], z5 = 0;
#pragma unroll
                    for (int a5 = 1; a5 < D; a5++) {
                        z = fma(c[a5], T[5][a5], z);
                        z5 = fma(c[a5], dT[5][a5], z5);
                    }
                    if (a4 == 0) {
                        s4 = z;
                        h4[1] = z5;
                    } else {
                        s4 = fma(T[4][a4], z, s4);
                        h4[0] = fma(dT[4][a4], z, h4[0]);
                        h4[1] = fma(T[4][a4], z5, h4[1]);
                    }
                }
                if (a3 == 0) {
                    s3 = s4;
                    h3[1] = h4[0];
                    h3[2] = h4[1];
                } else {
                    s3 = fma(T[3][a3], s4, s3);
                    h3[0] = fma(dT[3][a3], s4, h3[0]);
                    h3[1] = fma(T[3][a3], h4[0], h3[1]);
                    h3[2] = fma(T[3][a3], h4[1], h3[2]);
                }
            }
            if (a2 == 0) {
                s2 = s3;
#pragma unroll
                for (int q = 0; q < 3; q++) h2[q + 1] = h3[q];
            } else {
                s2 = fma(T[2][a2], s3, s2);
                h2[0] = fma(dT[2][a2], s3, h2[0]);
#pragma unroll
                for (int q = 0; q < 3; q++) h2[q + 1] = fma(T[2][a2], h3[q], h2[q + 1]);
            }
        }
    };
    // s1 / h1 += the a1 term (first: T[1][0] = 1, dT[1][0] = 0)
    auto add1 = [&](A t1, A dt1, bool first, R s2, const R (&h2)[4], A &s1, A (&h1)[5]) {
        if (first) {
            s1 = (A)s2;
#pragma unroll
            for (int q = 0; q < 4; q++) h1[q + 1] = (A)h2[q];
        } else {
            s1 = fma(t1, (A)s2, s1);
            h1[0] = fma(dt1, (A)s2, h1[0]);
#pragma unroll
            for (int q = 0; q < 4; q++) h1[q + 1] = fma(t1, (A)h2[q], h1[q + 1]);
        }
    };
    // s0 / h0 += the a0 term (first: T[0][0] = 1, dT[0][0] = 0)
    auto add0 = [&](A t0, A dt0, bool first, A s1, const A (&h1)[5]) {
        if (first) {
            s0 = s1;
#pragma unroll
            for (int q = 0; q < 5; q++) h0[q + 1] = h1[q];
        } else {
            s0 = fma(t0, s1, s0);
            h0[0] = fma(dt0, s1, h0[0]);
#pragma unroll
            for (int q = 0; q < 5; q++) h0[q + 1] = fma(t0, h1[q], h0[q + 1]);
        }
    };
    // loops over a0 and a1 (code size and registers), the T[0] and T[1]
    // factors consumed from the front of rotating copies (indexing them by a
    // loop index would place them in scratch)
    R t0v[D], d0v[D];
#pragma unroll
    for (int u = 0; u < D; u++) { t0v[u] = T[0][u]; d0v[u] = dT[0][u]; }
#pragma unroll 1
    for (int a0 = 0; a0 < D; a0++) {
        A s1 = 0, h1[5] = {0, 0, 0, 0, 0};
        R t1v[D], d1v[D];
#pragma unroll
        for (int u = 0; u < D; u++) { t1v[u] = T[1][u]; d1v[u] = dT[1][u]; }
#pragma unroll 1
        for (int a1 = 0; a1 < D; a1++) {
            R s2, h2[4];
            inner(a0, a1, s2, h2);
            add1((A)t1v[0], (A)d1v[0], false, s2, h2, s1, h1);
#pragma unroll
            for (int u = 0; u + 1 < D; u++) { t1v[u] = t1v[u + 1]; d1v[u] = d1v[u + 1]; }
        }
        add0((A)t0v[0], (A)d0v[0], false, s1, h1);
#pragma unroll
        for (int u = 0; u + 1 < D; u++) { t0v[u] = t0v[u + 1]; d0v[u] = d0v[u + 1]; }
    }
    S = s0;
#pragma unroll
    for (int q = 0; q < 6; q++) G[q] = h0[q];
}

// Force scalars of a P-pair cluster from S, G and the cutoffs.
template <int P, typename A>
__device__ __forceinline__ void cluster_scalars(const double (&r)[P], const double (&f)[P], const double (&fd)[P],
                                                A S, const A (&G)[P], double &E, double (&fs)[P])
{
    double fall = f[0];
#pragma unroll
    for (int q = 1; q < P; q++) fall *= f[q];
    E = fall * (double)S;
#pragma unroll
    for (int q = 0; q < P; q++) {
        double other = 1.0;
#pragma unroll
        for (int k = 0; k < P; k++)
            if (k != q) other *= f[k];
        fs[q] = other / r[q] * (f[q] * (double)G[q] + fd[q] * (double)S);
    }
}

}  // namespace chx
