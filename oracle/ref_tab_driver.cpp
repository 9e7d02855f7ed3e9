// ORACLE -- test infrastructure only.
//
// A thin C-callable driver around the reference's own chimesFF class compiled
// with -DTABULATION (see oracle/Makefile): one call of compute_2B_tab /
// compute_3B_tab, or of get_tab_2B / interpolateTricubic, with the caller's
// inputs.  It pins the tabulation restatement (chimes_tab_oracle.c) to the
// reference itself.  Built into oracle/_ref/ only; never shipped.
#include <unistd.h>

#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "chimesFF.h"

namespace {

struct quiet {
    std::streambuf *old;
    std::ofstream sink;
    quiet() : sink("/dev/null") { old = std::cout.rdbuf(sink.rdbuf()); }
    ~quiet() { std::cout.rdbuf(old); }
};

}  // namespace

extern "C" {

// The reference prefixes a table's name with the parameter file's directory
// without a separator, so the file is read from inside its directory under
// its bare name (the form the reference's own example uses).
void *reft_open(const char *dir, const char *file)
{
    quiet q;
    char cwd[4096];
    if (!getcwd(cwd, sizeof cwd)) return nullptr;
    if (chdir(dir) != 0) return nullptr;
    chimesFF *ff = new chimesFF;
    ff->init(1);
    ff->read_parameters(file);
    ff->build_pair_int_trip_map();
    if (chdir(cwd) != 0) {
        delete ff;
        return nullptr;
    }
    return ff;
}

void reft_close(void *h) { delete (chimesFF *)h; }

int reft_flags(void *h)
{
    chimesFF *ff = (chimesFF *)h;
    return (ff->tabulate_2B ? 1 : 0) | (ff->tabulate_3B ? 2 : 0);
}

double reft_get_tab_2b(void *h, int pair, double r, int for_energy)
{
    return ((chimesFF *)h)->get_tab_2B(pair, r, for_energy != 0);
}

void reft_interp_3b(void *h, int trip, double rij, double rik, double rjk, double *out4)
{
    chimesFF *ff = (chimesFF *)h;
    const std::vector<double> v = ff->interpolateTricubic(trip, rij, rik, rjk, ff->tab_e_3B[trip], ff->tab_f_ij_3B[trip],
                                                          ff->tab_f_ik_3B[trip], ff->tab_f_jk_3B[trip]);
    for (int k = 0; k < 4; k++) out4[k] = v[k];
}

void reft_compute_2b(void *h, double dx, const double *dr, const int *types, double *force6, double *stress6,
                     double *energy)
{
    quiet q;
    chimesFF *ff = (chimesFF *)h;
    std::vector<double> d(dr, dr + 3), f(6, 0.0), s(6, 0.0);
    std::vector<int> t(types, types + 2);
    chimes2BTmp tmp(1);
    double e = 0.0;
    ff->compute_2B_tab(dx, d, t, f, s, e, tmp);
    memcpy(force6, f.data(), sizeof(double) * 6);
    memcpy(stress6, s.data(), sizeof(double) * 6);
    *energy = e;
}

void reft_compute_3b(void *h, const double *dx, const double *dr, const int *types, double *force9, double *stress6,
                     double *energy)
{
    chimesFF *ff = (chimesFF *)h;
    std::vector<double> x(dx, dx + 3), d(dr, dr + 9), f(9, 0.0), s(6, 0.0);
    std::vector<int> t(types, types + 3);
    chimes3BTmp tmp(1);
    double e = 0.0;
    ff->compute_3B_tab(x, d, t, f, s, e, tmp);
    memcpy(force9, f.data(), sizeof(double) * 9);
    memcpy(stress6, s.data(), sizeof(double) * 6);
    *energy = e;
}

}  // extern "C"
