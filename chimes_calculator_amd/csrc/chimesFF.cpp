// chimesFF and serial_chimes_interface (C++ API) on top of the HIP engine.
//
// Errors print a "chimesFF: ERROR" / "ERROR" line and exit, as the reference
// does (chimesFF/src/chimesFF.cpp:363-371, serial_interface/src/
// serial_chimes_interface.cpp:142-146); the GPU path has no CPU fallback: a
// missing device is such an error.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <memory>
#include <mutex>

#include "chimesFF.h"
#include "host_model.h"
#include "errors.h"
#include "serial_chimes_interface.h"

struct chimesFF_impl {
    std::unique_ptr<chx::DeviceModel> dm;
};

namespace {

[[noreturn]] void fatal(const std::string &msg) { chx::fatal_error("chimesFF: ERROR: " + msg); }

bool quiet() { return getenv("CHIMES_QUIET") != nullptr; }

}  // namespace

chimesFF::chimesFF() : rank(0), natmtyps(0), impl_(new chimesFF_impl), badness(0) {}

chimesFF::~chimesFF() { delete impl_; }

void chimesFF::init(int mpi_rank)
{
    rank = mpi_rank;
    if (rank == 0 && !quiet()) std::cout << "chimesFF: ChIMES calculator (MI355X/HIP)" << std::endl;
}

void chimesFF::read_parameters(string paramfile)
{
    std::string err;
    chx::ParamSet P;
    if (rank == 0 && !quiet()) std::cout << "chimesFF: Reading parameters from file: " << paramfile << std::endl;
    if (!chx::load_params(paramfile, P, err)) fatal(err);
    impl_->dm = chx::make_device_model(P, chx::default_device(), err);
    if (!impl_->dm) fatal(err);
    natmtyps = P.natmtyps;
    poly_orders.assign(P.order, P.order + 3);
    atmtyps = P.atmtyps;
    masses = P.masses;
    if (rank == 0 && !quiet())
        std::cout << "chimesFF: Using respective 2, 3, and 4-body orders of: " << P.order[0] << " " << P.order[1]
                  << " " << P.order[2] << std::endl;
}

void chimesFF::require_model(const char *who)
{
    if (!impl_->dm) fatal(std::string(who) + " called before read_parameters");
}

void chimesFF::compute_1B(const int typ_idx, double &energy)
{
    require_model("compute_1B");
    energy += impl_->dm->P.energy_offsets[typ_idx];
}

void chimesFF::cluster_call(int nb, const double *dx, const double *dr, const int *typ, double *force,
                            double *stress, double &energy, double *last)
{
    require_model("compute");
    const int P = nb * (nb - 1) / 2;
    double E = 0.0, F[12], S[6];
    int flags = 0;
    chx::DeviceModel &dm = *impl_->dm;
    {
        std::lock_guard<std::mutex> lk(dm.eh->mu);
        if (chx_eval_clusters(dm.eh->eng, dm.m, nb, 1, dx, dr, typ, &E, F, S, last, &flags))
            fatal(chx_last_error(dm.eh->eng));
    }
    if (!(flags & 8)) return;  // cluster outside its cutoffs: nothing is touched
    energy += E;
    for (int k = 0; k < 3 * nb; k++) force[k] += F[k];
    for (int k = 0; k < 6; k++) stress[k] += S[k];
    if (flags & 3) badness = std::max(badness, (flags & 2) ? 2 : 1);
    (void)P;
}

void chimesFF::compute_2B(const double dx, const vector<double> &dr, const vector<int> typ_idxs,
                          vector<double> &force, vector<double> &stress, double &energy, chimes2BTmp &tmp)
{
    double dummy = 0.0;
    compute_2B(dx, dr, typ_idxs, force, stress, energy, tmp, dummy);
}

void chimesFF::compute_2B(const double dx, const vector<double> &dr, const vector<int> typ_idxs,
                          vector<double> &force, vector<double> &stress, double &energy, chimes2BTmp &,
                          double &force_scalar_in)
{
    double d = dx, last = force_scalar_in;
    cluster_call(2, &d, dr.data(), typ_idxs.data(), force.data(), stress.data(), energy, &last);
    force_scalar_in = last;
}

void chimesFF::compute_3B(const vector<double> &dx, const vector<double> &dr, const vector<int> &typ_idxs,
                          vector<double> &force, vector<double> &stress, double &energy, chimes3BTmp &tmp)
{
    vector<double> dummy(3, 0.0);
    compute_3B(dx, dr, typ_idxs, force, stress, energy, tmp, dummy);
}

void chimesFF::compute_3B(const vector<double> &dx, const vector<double> &dr, const vector<int> &typ_idxs,
                          vector<double> &force, vector<double> &stress, double &energy, chimes3BTmp &,
                          vector<double> &force_scalar_in)
{
    if (force_scalar_in.size() < 3) force_scalar_in.resize(3);
    cluster_call(3, dx.data(), dr.data(), typ_idxs.data(), force.data(), stress.data(), energy,
                 force_scalar_in.data());
}

void chimesFF::compute_4B(const vector<double> &dx, const vector<double> &dr, const vector<int> &typ_idxs,
                          vector<double> &force, vector<double> &stress, double &energy, chimes4BTmp &tmp)
{
    vector<double> dummy(6, 0.0);
    compute_4B(dx, dr, typ_idxs, force, stress, energy, tmp, dummy);
}

void chimesFF::compute_4B(const vector<double> &dx, const vector<double> &dr, const vector<int> &typ_idxs,
                          vector<double> &force, vector<double> &stress, double &energy, chimes4BTmp &,
                          vector<double> &force_scalar_in)
{
    if (force_scalar_in.size() < 6) force_scalar_in.resize(6);
    cluster_call(4, dx.data(), dr.data(), typ_idxs.data(), force.data(), stress.data(), energy,
                 force_scalar_in.data());
}

void chimesFF::compute_clusters(int nb, int n, const double *dx, const double *dr, const int *typ_idxs,
                                double *energy, double *force, double *stress)
{
    require_model("compute_clusters");
    if (n <= 0) return;
    const int P = nb * (nb - 1) / 2;
    std::vector<double> last((size_t)n * P, 0.0);
    std::vector<int> flags(n);
    chx::DeviceModel &dm = *impl_->dm;
    std::lock_guard<std::mutex> lk(dm.eh->mu);
    if (chx_eval_clusters(dm.eh->eng, dm.m, nb, n, dx, dr, typ_idxs, energy, force, stress, last.data(),
                          flags.data()))
        fatal(chx_last_error(dm.eh->eng));
}

void chimesFF::get_cutoff_2B(vector<vector<double> > &cutoff_2b)
{
    require_model("get_cutoff_2B");
    const chx::ParamSet &P = impl_->dm->P;
    cutoff_2b.assign(P.npairs(), vector<double>());
    for (int p = 0; p < P.npairs(); p++) cutoff_2b[p] = {P.pair_rin[p], P.pair_rout[p]};
}

double chimesFF::max_cutoff_2B(bool silent)
{
    require_model("max_cutoff_2B");
    const double m = impl_->dm->P.max_cutoff(2);
    if (rank == 0 && !silent && !quiet()) std::cout << "chimesFF: \tSetting 2-body max cutoff to: " << m << std::endl;
    return m;
}

double chimesFF::max_cutoff_3B(bool silent)
{
    require_model("max_cutoff_3B");
    const double m = impl_->dm->P.max_cutoff(3);
    if (rank == 0 && !silent && !quiet() && poly_orders[1] > 0)
        std::cout << "chimesFF: \tSetting 3-body max cutoff to: " << m << std::endl;
    return m;
}

double chimesFF::max_cutoff_4B(bool silent)
{
    require_model("max_cutoff_4B");
    const double m = impl_->dm->P.max_cutoff(4);
    if (rank == 0 && !silent && !quiet() && poly_orders[2] > 0)
        std::cout << "chimesFF: \tSetting 4-body max cutoff to: " << m << std::endl;
    return m;
}

void chimesFF::set_atomtypes(vector<string> &type_list)
{
    type_list.resize(natmtyps);
    for (int i = 0; i < natmtyps; i++) type_list[i] = atmtyps[i];
}

int chimesFF::get_atom_pair_index(int pair_id)
{
    require_model("get_atom_pair_index");
    return impl_->dm->P.pairmap_slow.at(pair_id);
}

// The permutation tables are built while the parameters are read; these are
// kept for callers that invoke them explicitly (chimesFF.cpp:2958-3080).
void chimesFF::build_pair_int_trip_map() {}
void chimesFF::build_pair_int_quad_map() {}

int chimesFF::get_badness() { return badness; }
void chimesFF::reset_badness() { badness = 0; }

// ---------------------------------------------------------------------------
// serial_chimes_interface
// ---------------------------------------------------------------------------

serial_chimes_interface::serial_chimes_interface(bool small) : allow_replication(small) {}

serial_chimes_interface::~serial_chimes_interface() {}

void serial_chimes_interface::init_chimesFF(string chimesFF_paramfile, int rank_in)
{
    init(rank_in);
    read_parameters(chimesFF_paramfile);
}

void serial_chimes_interface::calculate(vector<double> &x_in, vector<double> &y_in, vector<double> &z_in,
                                        vector<double> &cella_in, vector<double> &cellb_in,
                                        vector<double> &cellc_in, vector<string> &atmtyps_in, double &energy,
                                        vector<vector<double> > &force, vector<double> &stress)
{
    require_model("calculate");
    const int n = (int)x_in.size();
    if ((int)y_in.size() != n) fatal("x and y coordinate vector lengths do not match!");
    if ((int)z_in.size() != n) fatal("x and z coordinate vector lengths do not match!");
    chx::DeviceModel &dm = *impl_->dm;
    // the reference truncates the caller's type list to n (serial_chimes_interface.cpp:187)
    if ((int)atmtyps_in.size() > n) atmtyps_in.erase(atmtyps_in.begin() + n, atmtyps_in.end());
    std::vector<double> pos(3 * (size_t)n);
    std::vector<int> typ(n);
    for (int a = 0; a < n; a++) {
        pos[3 * a] = x_in[a];
        pos[3 * a + 1] = y_in[a];
        pos[3 * a + 2] = z_in[a];
        typ[a] = chx::type_index(dm.P, a < (int)atmtyps_in.size() ? atmtyps_in[a] : std::string());
        if (typ[a] < 0)
            fatal("Couldn't assign an atom type index for (index/type) " + std::to_string(a) + " " +
                  (a < (int)atmtyps_in.size() ? atmtyps_in[a] : std::string()));
    }
    double cells[9] = {cella_in[0], cella_in[1], cella_in[2], cellb_in[0], cellb_in[1],
                       cellb_in[2], cellc_in[0], cellc_in[1], cellc_in[2]};
    int model = 0;
    chx_batch b{};
    b.nconf = 1;
    b.natoms = &n;
    b.pos = pos.data();
    b.types = typ.data();
    b.cells = cells;
    b.model = &model;
    b.small = allow_replication ? 1 : 0;
    b.precision = CHX_FP64;
    std::vector<double> F(3 * (size_t)n);
    double E = 0.0, V[6], vol = 0.0;
    int nrep = 0, bad = 0;
    long long counts[4];
    chx_out o{};
    o.energy = &E;
    o.forces = F.data();
    o.virial = V;
    o.volume = &vol;
    o.n_rep = &nrep;
    o.counts = counts;
    o.badness = &bad;
    {
        std::lock_guard<std::mutex> lk(dm.eh->mu);
        chx_model *m = dm.m;
        if (chx_evaluate(dm.eh->eng, &m, 1, &b, &o)) fatal(chx_last_error(dm.eh->eng));
    }
    if (bad > badness) badness = bad;
    if (counts[1] && !quiet())
        std::cout << "Warning: " << counts[1] << " intermolecular distance(s) less than the inner cutoff were found"
                  << std::endl;
    if (counts[0] && !quiet())
        std::cout << "chimesFF: Added the penalty potential for " << counts[0] << " pair(s)" << std::endl;
    // replication appends the replicas' types to the caller's list
    // (serial_chimes_interface.cpp:257)
    const int ncopy = (nrep + 1) * (nrep + 1) * (nrep + 1);
    for (int c = 1; c < ncopy; c++)
        for (int a = 0; a < n; a++) atmtyps_in.push_back(atmtyps_in[a]);
    energy = (energy + E) / pow(nrep + 1.0, 3.0);
    for (int a = 0; a < n; a++)
        for (int k = 0; k < 3; k++) force[a][k] += F[3 * a + k];
    const double s6[6] = {V[0], V[1], V[2], V[3], V[4], V[5]};
    stress.resize(9);
    stress[0] = s6[0]; stress[1] = s6[1]; stress[2] = s6[2];
    stress[3] = s6[1]; stress[4] = s6[3]; stress[5] = s6[4];
    stress[6] = s6[2]; stress[7] = s6[4]; stress[8] = s6[5];
    for (int k = 0; k < 9; k++) stress[k] /= vol;
}
