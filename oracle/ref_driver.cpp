// ORACLE -- test infrastructure only.
//
// A thin C-callable driver around the reference's own serial_chimes_interface
// and chimesFF classes, compiled from the sources where they lie under the
// reference checkout (see oracle/Makefile).  It exposes E/F/stress and the
// reference's cluster lists in canonical key form so the C restatement
// (chimes_oracle.c) and the HIP product can be pinned against the reference
// itself.  Built into oracle/_ref/ only; never shipped with the product.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

// The reference keeps its neighbour lists private; the checker reads them.
#define private public
#include "serial_chimes_interface.h"
#undef private

namespace {

struct quiet {
    std::streambuf *old;
    std::ofstream sink;
    quiet() : sink("/dev/null") { old = std::cout.rdbuf(sink.rdbuf()); }
    ~quiet() { std::cout.rdbuf(old); }
};

struct ref_handle {
    serial_chimes_interface *ifc;
};

// Layered index -> (parent, shift) for the reference's ghost ordering
// (serial_chimes_interface.cpp:546-585).
void key_of(const simulation_system &S, int g, int *out)
{
    if (g < S.n_atoms) { out[0] = g; out[1] = out[2] = out[3] = 0; return; }
    int gp = g - S.n_atoms;
    int c = gp / S.n_atoms;
    int a = gp % S.n_atoms;
    int w = 2 * S.n_layers + 1;
    int mid = (w * w * w - 1) / 2;
    int t = c < mid ? c : c + 1;
    out[0] = a;
    out[1] = t / (w * w) - S.n_layers;
    out[2] = (t / w) % w - S.n_layers;
    out[3] = t % w - S.n_layers;
}

}  // namespace

extern "C" {

void *ref_open(const char *paramfile, int small)
{
    quiet q;
    ref_handle *h = new ref_handle;
    h->ifc = new serial_chimes_interface(small != 0);
    h->ifc->init_chimesFF(paramfile, 0);
    return h;
}

void ref_close(void *p)
{
    ref_handle *h = (ref_handle *)p;
    delete h->ifc;
    delete h;
}

// Returns the list sizes through n2/n3/n4 and the malloc'd key arrays
// (5 / 9 / 13 ints per entry) through the pointer arguments, when non-null.
int ref_calculate(void *p, int natom, const double *x, const double *y, const double *z,
                  const char *const *types, const double *ca, const double *cb,
                  const double *cc, double *energy, double *fx, double *fy, double *fz,
                  double *stress, long *n2, int **pairs, long *n3, int **trips, long *n4,
                  int **quads, int *meta)
{
    quiet q;
    ref_handle *h = (ref_handle *)p;
    std::vector<double> X(x, x + natom), Y(y, y + natom), Z(z, z + natom);
    std::vector<double> A(ca, ca + 3), B(cb, cb + 3), C(cc, cc + 3);
    std::vector<std::string> T(natom);
    for (int i = 0; i < natom; i++) T[i] = types[i];
    std::vector<std::vector<double> > F(natom, std::vector<double>(3));
    for (int i = 0; i < natom; i++) { F[i][0] = fx[i]; F[i][1] = fy[i]; F[i][2] = fz[i]; }
    std::vector<double> S(stress, stress + 9);
    h->ifc->calculate(X, Y, Z, A, B, C, T, *energy, F, S);
    for (int i = 0; i < natom; i++) { fx[i] = F[i][0]; fy[i] = F[i][1]; fz[i] = F[i][2]; }
    for (int k = 0; k < 9; k++) stress[k] = S[k];

    const simulation_system &N = h->ifc->neigh;
    if (meta) { meta[0] = N.n_atoms; meta[1] = N.n_layers; meta[2] = N.n_replicates; }
    if (pairs) {
        long cnt = 0;
        for (int i = 0; i < N.n_atoms; i++) cnt += h->ifc->neighlist_2b[i].size();
        *n2 = cnt;
        int *o = (int *)malloc(sizeof(int) * 5 * (cnt ? cnt : 1));
        long w = 0;
        for (int i = 0; i < N.n_atoms; i++)
            for (size_t k = 0; k < h->ifc->neighlist_2b[i].size(); k++) {
                o[5 * w] = i;
                key_of(N, h->ifc->neighlist_2b[i][k], o + 5 * w + 1);
                w++;
            }
        *pairs = o;
    }
    if (trips) {
        long cnt = h->ifc->neighlist_3b.size();
        *n3 = cnt;
        int *o = (int *)malloc(sizeof(int) * 9 * (cnt ? cnt : 1));
        for (long c = 0; c < cnt; c++) {
            const std::vector<int> &m = h->ifc->neighlist_3b[c];
            o[9 * c] = m[0];
            key_of(N, m[1], o + 9 * c + 1);
            key_of(N, m[2], o + 9 * c + 5);
        }
        *trips = o;
    }
    if (quads) {
        long cnt = h->ifc->neighlist_4b.size();
        *n4 = cnt;
        int *o = (int *)malloc(sizeof(int) * 13 * (cnt ? cnt : 1));
        for (long c = 0; c < cnt; c++) {
            const std::vector<int> &m = h->ifc->neighlist_4b[c];
            o[13 * c] = m[0];
            key_of(N, m[1], o + 13 * c + 1);
            key_of(N, m[2], o + 13 * c + 5);
            key_of(N, m[3], o + 13 * c + 9);
        }
        *quads = o;
    }
    return 0;
}

void ref_free(void *p) { free(p); }

// Single-cluster calls on the chimesFF base object.
void ref_compute_2b(void *p, double dx, const double *dr, const int *t, double *force,
                    double *stress, double *energy)
{
    quiet q;
    ref_handle *h = (ref_handle *)p;
    std::vector<double> d(dr, dr + 3), f(6, 0.0), s(stress, stress + 6);
    std::vector<int> ty(t, t + 2);
    chimes2BTmp tmp(h->ifc->poly_orders[0]);
    h->ifc->compute_2B(dx, d, ty, f, s, *energy, tmp);
    for (int k = 0; k < 6; k++) { force[k] += f[k]; stress[k] = s[k]; }
}

void ref_compute_3b(void *p, const double *dx, const double *dr, const int *t, double *force,
                    double *stress, double *energy)
{
    quiet q;
    ref_handle *h = (ref_handle *)p;
    std::vector<double> x(dx, dx + 3), d(dr, dr + 9), f(9, 0.0), s(stress, stress + 6);
    std::vector<int> ty(t, t + 3);
    chimes3BTmp tmp(h->ifc->poly_orders[1]);
    h->ifc->compute_3B(x, d, ty, f, s, *energy, tmp);
    for (int k = 0; k < 9; k++) force[k] += f[k];
    for (int k = 0; k < 6; k++) stress[k] = s[k];
}

void ref_compute_4b(void *p, const double *dx, const double *dr, const int *t, double *force,
                    double *stress, double *energy)
{
    quiet q;
    ref_handle *h = (ref_handle *)p;
    std::vector<double> x(dx, dx + 6), d(dr, dr + 18), f(12, 0.0), s(stress, stress + 6);
    std::vector<int> ty(t, t + 4);
    chimes4BTmp tmp(h->ifc->poly_orders[2]);
    h->ifc->compute_4B(x, d, ty, f, s, *energy, tmp);
    for (int k = 0; k < 12; k++) force[k] += f[k];
    for (int k = 0; k < 6; k++) stress[k] = s[k];
}

}  // extern "C"
