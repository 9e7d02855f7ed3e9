// Tabulated 2- and 3-body interactions of the LAMMPS pair style built with
// -DTABULATION (chimesFF.cpp:198-352, 1777-1924, 2178-2484): device layout of
// the tables and the two lookups.
//
// Every lookup does the reference's floating-point operations in the
// reference's order (the library is built without contraction), so a table
// value equals the reference's bit for bit; the only departures are inputs
// where the reference exits or reads past its arrays, which set *err here.
#pragma once
#include <hip/hip_runtime.h>

namespace chx {

struct chx_tabs_dev {
    int ntab2, ntab3;
    // 2-body, table p (parameter pair p): rows off2[p] .. off2[p + 1] of r2, e2, f2
    const int *off2;
    const double *r2, *e2, *f2;
    // 3-body, table t (triplet type t): n3[t] points per axis, 1 / spacing
    // dr3[t]; the rij column (n^3 rows) at rij3 + o3[t], the first n^2 rik and
    // the first n rjk at rik3 + o3ik[t] and rjk3 + o3jk[t], and per row the
    // energy and the ij, ik, jk force scalars at v3 + 4 (o3[t] + row)
    const int *n3;
    const double *dr3;
    const long *o3, *o3ik, *o3jk;
    const double *rij3, *rik3, *rjk3, *v3;
};

// get_tab_2B's interval search (chimesFF.cpp:1861-1891): index of the spline
// piece for rij, or -1 past the table (value 0), -2 where the reference exits
// (rij at or below the first knot, or a knot match that leaves index 0).
__device__ __forceinline__ int tab2_piece(const double *__restrict__ x, int n, double rij)
{
    int lo = 0, hi = n;  // lower_bound: first knot not below rij
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (x[mid] < rij) lo = mid + 1;
        else hi = mid;
    }
    int i = lo;
    if (i == n) return -1;
    if (i == 0) return -2;
    if (rij == x[i]) i--;
    if (i >= n - 1) return -1;
    if (i <= 0) return -2;
    return i;
}

// The local spline of piece i at rij (chimesFF.cpp:1903-1923): one step of
// the natural-spline recurrence from c_prev = 0, evaluated at rij - x_i
// (negative when rij lies left of x_i, as the reference's search allows).
__device__ __forceinline__ double tab2_spline(const double *__restrict__ x, const double *__restrict__ y, int i,
                                              double rij)
{
    const double xm = x[i - 1], x0 = x[i], x1 = x[i + 1];
    const double ym = y[i - 1], y0 = y[i], y1 = y[i + 1];
    const double h = x1 - x0;
    const double alpha = (3 * (y1 - y0) / h) - (3 * (y0 - ym) / (x0 - xm));
    const double l = 2 * (x1 - xm) - h;
    const double mu = h / l;
    const double z = alpha / l;
    const double c_prev = 0.0;
    const double c = z - mu * c_prev;
    const double b = (y1 - y0) / h - h * (c + 2 * c_prev) / 3.0;
    const double d = (c - c_prev) / (3.0 * h);
    const double dx = rij - x0;
    return y0 + b * dx + c * dx * dx + d * dx * dx * dx;
}

// Energy and force scalar of pair type p at distance r; err set where the
// reference exits.
__device__ __forceinline__ void tab2_lookup(const chx_tabs_dev &T, int p, double r, double &E, double &fs, bool &err)
{
    const int o = T.off2[p], n = T.off2[p + 1] - o;
    const double *x = T.r2 + o;
    const int i = tab2_piece(x, n, r);
    E = 0.0;
    fs = 0.0;
    if (i == -2) err = true;
    if (i < 0) return;
    E = tab2_spline(x, T.e2 + o, i, r);
    fs = tab2_spline(x, T.f2 + o, i, r);
}

// cubicInterpolate (chimesFF.cpp:2317-2320), Catmull-Rom through p1, p2
__device__ __forceinline__ double tab_cr(double p0, double p1, double p2, double p3, double x)
{
    return p1 + 0.5 * x * (p2 - p0 + x * (2 * p0 - 5 * p1 + 4 * p2 - p3 + x * (3 * (p1 - p2) + p3 - p0)));
}

// interpolateTricubic (chimesFF.cpp:2335-2395) of triplet table t at the
// runtime pair distances: out = energy and the ij, ik, jk force scalars.  The
// 4 x 4 x 4 stencil is reduced as it is read, innermost (rjk) first, in the
// order of tricubicInterpolate / bicubicInterpolate, so the four channels
// need 32 registers rather than 256.  err: a grid index past the table.
__device__ __forceinline__ void tab3_lookup(const chx_tabs_dev &T, int t, double rij, double rik, double rjk,
                                            double (&out)[4], bool &err)
{
    const int n = T.n3[t];
    const long nn = (long)n * n;
    const double dr = T.dr3[t];
    const double *cij = T.rij3 + T.o3[t], *cik = T.rik3 + T.o3ik[t], *cjk = T.rjk3 + T.o3jk[t];
    int i = (int)((rij - cij[0]) * dr);
    int j = (int)((rik - cik[0]) * dr);
    int k = (int)((rjk - cjk[0]) * dr);
    i = i > 1 ? i : 1;
    j = j > 1 ? j : 1;
    k = k > 1 ? k : 1;
    out[0] = out[1] = out[2] = out[3] = 0.0;
    if (i >= n || j >= n || k >= n) {
        err = true;
        return;
    }
    const double xi = (rij - cij[i * nn + (long)j * n + k]) * dr;
    const double yi = (rik - cik[(long)j * n + k]) * dr;
    const double zi = (rjk - cjk[k]) * dr;
    const double4 *v = (const double4 *)(T.v3 + 4 * T.o3[t]);
    double a_i[4][4];  // per channel, the bicubic value of each di
    for (int di = 0; di < 4; di++) {
        double a_j[4][4];  // per channel, the cubic over dk of each dj
        for (int dj = 0; dj < 4; dj++) {
            double p[4][4];  // channel, dk
            for (int dk = 0; dk < 4; dk++) {
                const int ti = i + di - 1, tj = j + dj - 1, tk = k + dk - 1;
                double4 w = make_double4(0.0, 0.0, 0.0, 0.0);
                if (ti < n && tj < n && tk < n) w = v[ti * nn + (long)tj * n + tk];
                p[0][dk] = w.x;
                p[1][dk] = w.y;
                p[2][dk] = w.z;
                p[3][dk] = w.w;
            }
#pragma unroll
            for (int c = 0; c < 4; c++) a_j[c][dj] = tab_cr(p[c][0], p[c][1], p[c][2], p[c][3], zi);
        }
#pragma unroll
        for (int c = 0; c < 4; c++) a_i[c][di] = tab_cr(a_j[c][0], a_j[c][1], a_j[c][2], a_j[c][3], yi);
    }
#pragma unroll
    for (int c = 0; c < 4; c++) out[c] = tab_cr(a_i[c][0], a_i[c][1], a_i[c][2], a_i[c][3], xi);
}

}  // namespace chx
