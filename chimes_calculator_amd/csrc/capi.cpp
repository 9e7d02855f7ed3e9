// The reference's two C ABIs over the GPU implementation:
//   chimescalc_serial_C (serial_interface/api/chimescalc_serial_C.cpp:22-124)
//   chimescalc_C        (chimesFF/api/chimescalc_C.cpp:22-318)
#include <algorithm>
#include <iostream>
#include <string>
#include <vector>

#include "chimesFF.h"
#include "errors.h"
#include "chimescalc_C.h"
#include "chimescalc_serial_C.h"
#include "serial_chimes_interface.h"

namespace {

serial_chimes_interface *g_serial = nullptr;  // set by set_chimes_serial
chimesFF *g_ff = nullptr;                     // set by set_chimes

serial_chimes_interface &serial_default()
{
    static serial_chimes_interface inst;
    return inst;
}

chimesFF &ff_default()
{
    static chimesFF inst;
    return inst;
}

void check_small(int small)
{
    if (small != 0 && small != 1) {
        std::cout << "ERROR: Small must be set to 0 (false) or 1 (true)" << std::endl;
        chx::fatal_error("Received: " + std::to_string(small));
    }
}

int type_of(chimesFF *ff, const char *name)
{
    auto it = std::find(ff->atmtyps.begin(), ff->atmtyps.end(), std::string(name));
    if (it == ff->atmtyps.end()) {
        chx::fatal_error(std::string("ERROR: input atom type not in parameter file: ") + name);
    }
    return (int)(it - ff->atmtyps.begin());
}

}  // namespace

extern "C" {

// ---------------------------------------------------------------------------
// chimescalc_serial_C
// ---------------------------------------------------------------------------

void *chimes_open_instance() { return (void *)new serial_chimes_interface; }

void chimes_close_instance(void *handle) { delete (serial_chimes_interface *)handle; }

void set_chimes_serial(int small)
{
    check_small(small);
    g_serial = &serial_default();
    set_chimes_serial_instance(g_serial, small);
}

void set_chimes_serial_instance(void *handle, int small)
{
    check_small(small);
    ((serial_chimes_interface *)handle)->allow_replication = small != 0;
}

void init_chimes_serial(char *param_file, int *rank)
{
    if (!g_serial) {
        chx::fatal_error("ERROR: set_chimes_serial must be called before init_chimes_serial");
    }
    g_serial->init_chimesFF(param_file, *rank);
}

void init_chimes_serial_instance(void *handle, char *param_file, int rank)
{
    (void)rank;
    ((serial_chimes_interface *)handle)->init_chimesFF(param_file, 0);
}

void calculate_chimes(int natom, double *xc, double *yc, double *zc, char *atom_types[], double ca[3],
                      double cb[3], double cc[3], double *energy, double fx[], double fy[], double fz[],
                      double stress[9])
{
    if (!g_serial) {
        chx::fatal_error("ERROR: set_chimes_serial must be called before calculate_chimes");
    }
    calculate_chimes_instance(g_serial, natom, xc, yc, zc, atom_types, ca, cb, cc, energy, fx, fy, fz, stress);
}

void calculate_chimes_instance(void *handle, int natom, double *xc, double *yc, double *zc, char *atom_types[],
                               double ca[3], double cb[3], double cc[3], double *energy, double fx[], double fy[],
                               double fz[], double stress[9])
{
    serial_chimes_interface *s = (serial_chimes_interface *)handle;
    std::vector<double> x(xc, xc + natom), y(yc, yc + natom), z(zc, zc + natom);
    std::vector<std::vector<double> > f(natom, std::vector<double>(3));
    std::vector<std::string> t(natom);
    for (int i = 0; i < natom; i++) {
        f[i][0] = fx[i];
        f[i][1] = fy[i];
        f[i][2] = fz[i];
        t[i] = atom_types[i];
    }
    std::vector<double> st(stress, stress + 9);
    std::vector<double> a(ca, ca + 3), b(cb, cb + 3), c(cc, cc + 3);
    s->calculate(x, y, z, a, b, c, t, *energy, f, st);
    for (int i = 0; i < natom; i++) {
        fx[i] = f[i][0];
        fy[i] = f[i][1];
        fz[i] = f[i][2];
    }
    for (int k = 0; k < 9; k++) stress[k] = st[k];
}

// ---------------------------------------------------------------------------
// chimescalc_C
// ---------------------------------------------------------------------------

static chimesFF *ff()
{
    if (!g_ff) {
        chx::fatal_error("ERROR: set_chimes must be called first");
    }
    return g_ff;
}

double get_chimes_max_2b_cutoff() { return ff()->max_cutoff_2B(); }
double get_chimes_max_3b_cutoff() { return ff()->max_cutoff_3B(); }
double get_chimes_max_4b_cutoff() { return ff()->max_cutoff_4B(); }

void set_chimes() { g_ff = &ff_default(); }

int get_chimes_2b_order() { return ff()->poly_orders[0]; }
int get_chimes_3b_order() { return ff()->poly_orders.size() > 1 ? ff()->poly_orders[1] : 0; }
int get_chimes_4b_order() { return ff()->poly_orders.size() > 2 ? ff()->poly_orders[2] : 0; }

void init_chimes(int rank) { ff()->init(rank); }

void chimes_read_params(char *param_file)
{
    ff()->read_parameters(param_file);
    ff()->build_pair_int_trip_map();
    ff()->build_pair_int_quad_map();
}

void chimes_build_pair_int_trip_map() { ff()->build_pair_int_trip_map(); }
void chimes_build_pair_int_quad_map() { ff()->build_pair_int_quad_map(); }

void chimes_compute_2b_props(double rij, double dr[3], char *atype2b[2], double force[2][3], double stress[9],
                             double *epot)
{
    chimesFF *f = ff();
    std::vector<double> d(dr, dr + 3), F(2 * CHDIM, 0.0), S(9, 0.0);
    std::vector<int> t = {type_of(f, atype2b[0]), type_of(f, atype2b[1])};
    chimes2BTmp tmp(f->poly_orders[0]);
    f->compute_2B(rij, d, t, F, S, *epot, tmp);
    for (int a = 0; a < 2; a++)
        for (int k = 0; k < 3; k++) force[a][k] += F[a * CHDIM + k];
    for (int k = 0; k < 9; k++) stress[k] += S[k];
}

void chimes_compute_3b_props(double dr_3b[3], double dist_3b[3][3], char *atype3b[3], double f3b[3][3],
                             double stress[9], double *epot)
{
    chimesFF *f = ff();
    std::vector<double> dx(dr_3b, dr_3b + 3), d(9), F(3 * CHDIM, 0.0), S(9, 0.0);
    for (int p = 0; p < 3; p++)
        for (int k = 0; k < 3; k++) d[p * CHDIM + k] = dist_3b[p][k];
    std::vector<int> t = {type_of(f, atype3b[0]), type_of(f, atype3b[1]), type_of(f, atype3b[2])};
    chimes3BTmp tmp(f->poly_orders[1]);
    f->compute_3B(dx, d, t, F, S, *epot, tmp);
    for (int a = 0; a < 3; a++)
        for (int k = 0; k < 3; k++) f3b[a][k] += F[a * CHDIM + k];
    for (int k = 0; k < 9; k++) stress[k] += S[k];
}

void chimes_compute_4b_props(double dr_4b[6], double dist_4b[6][3], char *atype4b[4], double f4b[4][3],
                             double stress[9], double *epot)
{
    chimesFF *f = ff();
    std::vector<double> dx(dr_4b, dr_4b + 6), d(18), F(4 * CHDIM, 0.0), S(stress, stress + 9);
    for (int p = 0; p < 6; p++)
        for (int k = 0; k < 3; k++) d[p * CHDIM + k] = dist_4b[p][k];
    std::vector<int> t = {type_of(f, atype4b[0]), type_of(f, atype4b[1]), type_of(f, atype4b[2]),
                          type_of(f, atype4b[3])};
    chimes4BTmp tmp(f->poly_orders[2]);
    f->compute_4B(dx, d, t, F, S, *epot, tmp);
    for (int a = 0; a < 4; a++)
        for (int k = 0; k < 3; k++) f4b[a][k] += F[a * CHDIM + k];
    for (int k = 0; k < 9; k++) stress[k] = S[k];
}

void chimes_compute_2b_props_fromf90(double *rij, double dr[3], char *type1, char *type2, double force[2][3],
                                     double stress[9], double *epot)
{
    char *t[2] = {type1, type2};
    chimes_compute_2b_props(*rij, dr, t, force, stress, epot);
}

void chimes_compute_3b_props_fromf90(double dr_3b[3], double dist_3b[3][3], char *type1, char *type2, char *type3,
                                     double f3b[3][3], double stress[9], double *epot)
{
    char *t[3] = {type1, type2, type3};
    chimes_compute_3b_props(dr_3b, dist_3b, t, f3b, stress, epot);
}

void chimes_compute_4b_props_fromf90(double dr_4b[6], double dist_4b[6][3], char *type1, char *type2, char *type3,
                                     char *type4, double f4b[4][3], double stress[9], double *epot)
{
    char *t[4] = {type1, type2, type3, type4};
    chimes_compute_4b_props(dr_4b, dist_4b, t, f4b, stress, epot);
}

}  // extern "C"
