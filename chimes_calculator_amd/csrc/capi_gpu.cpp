// chimes_gpu.h: batched, mixed-precision and device-buffer entry points.
#include "chimes_gpu.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "host_model.h"


struct chimes_gpu_plan {
    chx_plan *p = nullptr;
    std::vector<chimes_gpu_model *> models;  // for the type checks of each run
    std::vector<int> model_of, natoms_of;
    int nconf = 0;
    long natoms = 0;
    std::vector<int> nrep;
    std::vector<double> vol, V;
    int out_dev = 0;
    ~chimes_gpu_plan()
    {
        if (p) chx_plan_destroy(p);
    }
};

namespace {

thread_local std::string g_err;

int fail(const std::string &m)
{
    g_err = m;
    return 1;
}

// Checks that all models live on one engine; returns it.
chx::EngineHandle *common_engine(chimes_gpu_model *const *models, int nmodels, std::vector<chx_model *> &ms)
{
    if (nmodels <= 0 || !models) return nullptr;
    chx::EngineHandle *eh = models[0]->dm->eh;
    ms.resize(nmodels);
    for (int k = 0; k < nmodels; k++) {
        if (!models[k] || models[k]->dm->eh != eh) return nullptr;
        ms[k] = models[k]->dm->m;
    }
    return eh;
}

// Energy / (R+1)^3 and the 9-component stress (virial / volume) of each
// configuration from the engine's raw outputs.
void finish_host(int nconf, const int *nrep, const double *vol, const double *V, double *energy, double *stress)
{
    for (int c = 0; c < nconf; c++) {
        energy[c] /= pow(nrep[c] + 1.0, 3.0);
        const double *v = &V[6 * c];
        double *s = stress + 9 * c;
        s[0] = v[0]; s[1] = v[1]; s[2] = v[2];
        s[3] = v[1]; s[4] = v[3]; s[5] = v[4];
        s[6] = v[2]; s[7] = v[4]; s[8] = v[5];
        for (int k = 0; k < 9; k++) s[k] /= vol[c];
    }
}

int check_types(chimes_gpu_model *const *models, int nmodels, int nconf, const int *model_of, const int *natoms,
                const int *types)
{
    for (int c = 0; c < nconf; c++)
        if (model_of[c] < 0 || model_of[c] >= nmodels) return fail("model index out of range");
    if (!types) return 0;
    long a = 0;
    for (int c = 0; c < nconf; c++) {
        const int nt = models[model_of[c]]->dm->P.natmtyps;
        for (int k = 0; k < natoms[c]; k++, a++)
            if (types[a] < 0 || types[a] >= nt) return fail("atom type index out of range");
    }
    return 0;
}

void sort_rows(int *a, long n, int w)
{
    std::vector<std::vector<int>> rows(n);
    for (long i = 0; i < n; i++) rows[i].assign(a + i * w, a + (i + 1) * w);
    std::sort(rows.begin(), rows.end());
    for (long i = 0; i < n; i++) std::memcpy(a + i * w, rows[i].data(), sizeof(int) * w);
}

}  // namespace

extern "C" {

const char *chimes_gpu_last_error(void) { return g_err.c_str(); }

chimes_gpu_model *chimes_gpu_model_load(const char *param_file, int device)
{
    chx::ParamSet P;
    std::string err;
    if (!chx::load_params(param_file, P, err)) { fail(err); return nullptr; }
    std::unique_ptr<chx::DeviceModel> dm = chx::make_device_model(P, device, err);
    if (!dm) { fail(err); return nullptr; }
    chimes_gpu_model *m = new chimes_gpu_model;
    m->dm = std::move(dm);
    return m;
}

void chimes_gpu_model_free(chimes_gpu_model *m) { delete m; }

int chimes_gpu_model_ntypes(const chimes_gpu_model *m) { return m->dm->P.natmtyps; }

int chimes_gpu_model_type_index(const chimes_gpu_model *m, const char *name)
{
    return chx::type_index(m->dm->P, name);
}

int chimes_gpu_model_order(const chimes_gpu_model *m, int body)
{
    return (body >= 2 && body <= 4) ? m->dm->P.order[body - 2] : 0;
}

double chimes_gpu_model_max_cutoff(const chimes_gpu_model *m, int body)
{
    return (body >= 2 && body <= 4) ? m->dm->P.max_cutoff(body) : 0.0;
}

int chimes_gpu_eval_batch(chimes_gpu_model *const *models, int nmodels, int nconf, const int *model_of,
                          const int *natoms, const double *pos, const int *types, const double *cells, int small,
                          int precision, double *energy, double *forces, double *stress)
{
    std::vector<chx_model *> ms;
    chx::EngineHandle *eh = common_engine(models, nmodels, ms);
    if (!eh) return fail("models must be loaded on the same device");
    if (check_types(models, nmodels, nconf, model_of, natoms, types)) return 1;
    chx_batch b{};
    b.nconf = nconf;
    b.natoms = natoms;
    b.pos = pos;
    b.types = types;
    b.cells = cells;
    b.model = model_of;
    b.small = small;
    b.precision = precision;
    std::vector<double> V(6 * (size_t)nconf), vol(nconf);
    std::vector<int> nrep(nconf);
    chx_out o{};
    o.energy = energy;
    o.forces = forces;
    o.virial = V.data();
    o.volume = vol.data();
    o.n_rep = nrep.data();
    {
        std::lock_guard<std::mutex> lk(eh->mu);
        if (chx_evaluate(eh->eng, ms.data(), nmodels, &b, &o)) return fail(chx_last_error(eh->eng));
    }
    finish_host(nconf, nrep.data(), vol.data(), V.data(), energy, stress);
    return 0;
}

int chimes_gpu_eval_batch_device(chimes_gpu_model *const *models, int nmodels, int nconf, const int *model_of,
                                 const int *natoms, const double *pos_dev, const int *types_dev,
                                 const double *cells, int small, int precision, double *energy_dev,
                                 double *forces_dev, double *virial_dev)
{
    std::vector<chx_model *> ms;
    chx::EngineHandle *eh = common_engine(models, nmodels, ms);
    if (!eh) return fail("models must be loaded on the same device");
    chx_batch b{};
    b.nconf = nconf;
    b.natoms = natoms;
    b.pos = pos_dev;
    b.types = types_dev;
    b.cells = cells;
    b.model = model_of;
    b.small = small;
    b.pos_dev = 1;
    b.precision = precision;
    chx_out o{};
    o.energy = energy_dev;
    o.forces = forces_dev;
    o.virial = virial_dev;
    o.out_dev = 1;
    std::lock_guard<std::mutex> lk(eh->mu);
    if (chx_evaluate(eh->eng, ms.data(), nmodels, &b, &o)) return fail(chx_last_error(eh->eng));
    return 0;
}

chimes_gpu_plan *chimes_gpu_plan_create(chimes_gpu_model *const *models, int nmodels, int nconf,
                                        const int *model_of, const int *natoms, const double *pos, const int *types,
                                        const double *cells, int small, int precision, int tune, double *energy,
                                        double *forces, double *stress)
{
    std::vector<chx_model *> ms;
    chx::EngineHandle *eh = common_engine(models, nmodels, ms);
    if (!eh) { fail("models must be loaded on the same device"); return nullptr; }
    if (check_types(models, nmodels, nconf, model_of, natoms, types)) return nullptr;
    std::unique_ptr<chimes_gpu_plan> P(new chimes_gpu_plan);
    P->nconf = nconf;
    P->models.assign(models, models + nmodels);
    P->model_of.assign(model_of, model_of + nconf);
    P->natoms_of.assign(natoms, natoms + nconf);
    P->nrep.resize(nconf);
    P->vol.resize(nconf);
    P->V.resize(6 * (size_t)nconf);
    chx_batch b{};
    b.nconf = nconf;
    b.natoms = natoms;
    b.pos = pos;
    b.types = types;
    b.cells = cells;
    b.model = model_of;
    b.small = small;
    b.precision = precision;
    chx_out o{};
    o.energy = energy;
    o.forces = forces;
    o.virial = P->V.data();
    o.volume = P->vol.data();
    o.n_rep = P->nrep.data();
    char err[512];
    P->p = chx_plan_create(models[0]->dm->device, ms.data(), nmodels, &b, tune, &o, err, sizeof(err));
    if (!P->p) { fail(err); return nullptr; }
    finish_host(nconf, P->nrep.data(), P->vol.data(), P->V.data(), energy, stress);
    return P.release();
}

int chimes_gpu_plan_run(chimes_gpu_plan *P, const double *pos, const int *types, const double *cells,
                        double *energy, double *forces, double *stress)
{
    if (!P || P->out_dev) return fail("not a host-buffer plan");
    if (check_types(P->models.data(), (int)P->models.size(), P->nconf, P->model_of.data(), P->natoms_of.data(), types))
        return 1;
    chx_out o{};
    o.energy = energy;
    o.forces = forces;
    o.virial = P->V.data();
    o.volume = P->vol.data();
    o.n_rep = P->nrep.data();
    if (chx_plan_run(P->p, pos, types, cells, &o)) return fail(chx_plan_error(P->p));
    finish_host(P->nconf, P->nrep.data(), P->vol.data(), P->V.data(), energy, stress);
    return 0;
}

chimes_gpu_plan *chimes_gpu_plan_create_device(chimes_gpu_model *const *models, int nmodels, int nconf,
                                               const int *model_of, const int *natoms, const double *pos_dev,
                                               const int *types_dev, const double *cells, int small, int precision,
                                               int tune, double *energy_dev, double *forces_dev,
                                               double *virial_dev)
{
    std::vector<chx_model *> ms;
    chx::EngineHandle *eh = common_engine(models, nmodels, ms);
    if (!eh) { fail("models must be loaded on the same device"); return nullptr; }
    if (check_types(models, nmodels, nconf, model_of, natoms, nullptr)) return nullptr;
    std::unique_ptr<chimes_gpu_plan> P(new chimes_gpu_plan);
    P->nconf = nconf;
    P->out_dev = 1;
    chx_batch b{};
    b.nconf = nconf;
    b.natoms = natoms;
    b.pos = pos_dev;
    b.types = types_dev;
    b.cells = cells;
    b.model = model_of;
    b.small = small;
    b.pos_dev = 1;
    b.precision = precision;
    chx_out o{};
    o.energy = energy_dev;
    o.forces = forces_dev;
    o.virial = virial_dev;
    o.out_dev = 1;
    char err[512];
    P->p = chx_plan_create(models[0]->dm->device, ms.data(), nmodels, &b, tune, &o, err, sizeof(err));
    if (!P->p) { fail(err); return nullptr; }
    return P.release();
}

int chimes_gpu_plan_run_device(chimes_gpu_plan *P, const double *cells, double *energy_dev, double *forces_dev,
                               double *virial_dev)
{
    if (!P || !P->out_dev) return fail("not a device-buffer plan");
    chx_out o{};
    o.energy = energy_dev;
    o.forces = forces_dev;
    o.virial = virial_dev;
    o.out_dev = 1;
    if (chx_plan_run(P->p, nullptr, nullptr, cells, &o)) return fail(chx_plan_error(P->p));
    return 0;
}

void chimes_gpu_plan_info(const chimes_gpu_plan *P, int variants[2], double tune_ms[8], long long counts[3])
{
    chx_plan_info(P->p, variants, tune_ms, counts);
}

void chimes_gpu_plan_free(chimes_gpu_plan *P) { delete P; }

int chimes_gpu_plan_set_variants(chimes_gpu_plan *P, int var3, int var4)
{
    if (!P) return fail("no plan");
    if (chx_plan_set_variants(P->p, var3, var4)) return fail(chx_plan_error(P->p));
    return 0;
}

int chimes_gpu_set_variants(int device, int var3, int var4)
{
    std::string err;
    chx::EngineHandle *eh = chx::engine_for_device(device, err);
    if (!eh) return fail(err);
    std::lock_guard<std::mutex> lk(eh->mu);
    if (chx_set_variants(eh->eng, var3, var4)) return fail(chx_last_error(eh->eng));
    return 0;
}

int chimes_gpu_get_variants(int device, int variants[2])
{
    std::string err;
    chx::EngineHandle *eh = chx::engine_for_device(device, err);
    if (!eh) return fail(err);
    std::lock_guard<std::mutex> lk(eh->mu);
    chx_get_variants(eh->eng, variants);
    return 0;
}

void chimes_gpu_last_timing(const chimes_gpu_model *m, double ms[4])
{
    chx_last_timing(m->dm->eh->eng, ms);
}

void chimes_gpu_last_counts(const chimes_gpu_model *m, long long counts[3])
{
    chx_last_counts(m->dm->eh->eng, counts);
}

int chimes_gpu_compute_clusters(chimes_gpu_model *m, int nb, int n, const double *dist, const double *dr,
                                const int *types, double *energy, double *force, double *stress)
{
    if (nb < 2 || nb > 4) return fail("cluster size must be 2, 3 or 4");
    const int P = nb * (nb - 1) / 2;
    const int nt = m->dm->P.natmtyps;
    for (long k = 0; k < (long)n * nb; k++)
        if (types[k] < 0 || types[k] >= nt) return fail("atom type index out of range");
    std::vector<double> last((size_t)n * P, 0.0);
    std::vector<int> flags(n);
    chx::EngineHandle *eh = m->dm->eh;
    std::lock_guard<std::mutex> lk(eh->mu);
    if (chx_eval_clusters(eh->eng, m->dm->m, nb, n, dist, dr, types, energy, force, stress, last.data(),
                          flags.data()))
        return fail(chx_last_error(eh->eng));
    return 0;
}

int chimes_gpu_compute_clusters_static(chimes_gpu_model *m, int nb, int n, const double *dist, const double *dr,
                                       const int *types, double *energy, double *force, double *stress)
{
    if (!m) return fail("no model");
    if (nb < 2 || nb > 4) return fail("cluster size must be 2, 3 or 4");
    if (n < 0 || n > CHIMES_GPU_STATIC_MAX_CLUSTERS) return fail("cluster count out of range");
    const int nt = m->dm->P.natmtyps;
    for (long k = 0; k < (long)n * nb; k++)
        if (types[k] < 0 || types[k] >= nt) return fail("atom type index out of range");
    chx::EngineHandle *eh = m->dm->eh;
    std::lock_guard<std::mutex> lk(eh->mu);
    if (chx_eval_clusters_static(eh->eng, m->dm->m, nb, n, dist, dr, types, energy, force, stress))
        return fail(chx_last_error(eh->eng));
    return 0;
}

int chimes_gpu_cluster_lists(chimes_gpu_model *m, int conf, long *n2, int **pairs, long *n3, int **trips,
                             long *n4, int **quads)
{
    chx::EngineHandle *eh = m->dm->eh;
    std::lock_guard<std::mutex> lk(eh->mu);
    if (chx_export_lists(eh->eng, conf, n2, pairs, n3, trips, n4, quads)) return fail(chx_last_error(eh->eng));
    sort_rows(*pairs, *n2, 5);
    sort_rows(*trips, *n3, 9);
    sort_rows(*quads, *n4, 13);
    return 0;
}

void chimes_gpu_free(void *p) { chx_free(p); }

}  // extern "C"

// ---------------------------------------------------------------------------
// chimes_lmp.h: the LAMMPS pair style
// ---------------------------------------------------------------------------
#include "chimes_lmp.h"
#include "tables.h"

struct chimes_lmp {
    chx_lmp *p = nullptr;
    const chx::ParamSet *P = nullptr;  // the model's parameter set (table names)
    ~chimes_lmp()
    {
        if (p) chx_lmp_destroy(p);
    }
};

extern "C" {

chimes_lmp *chimes_lmp_create(chimes_gpu_model *m, int ntypes, const double *mass)
{
    if (!m || ntypes <= 0 || !mass) {
        fail("model, atom types and masses are required");
        return nullptr;
    }
    const std::vector<double> &cm = m->dm->P.masses;
    std::vector<int> ct(ntypes, -1);
    for (int t = 1; t <= ntypes; t++) {
        for (size_t j = 0; j < cm.size(); j++)
            if (std::fabs(mass[t] - cm[j]) < 1e-3) {
                ct[t - 1] = (int)j;
                break;
            }
        if (ct[t - 1] < 0) {
            fail("LAMMPS atom type " + std::to_string(t) + " matches no ChIMES atom type by mass");
            return nullptr;
        }
    }
    char err[256] = {0};
    chx_lmp *p = chx_lmp_create(m->dm->m, ntypes, ct.data(), err, sizeof err);
    if (!p) {
        fail(err);
        return nullptr;
    }
    chimes_lmp *h = new chimes_lmp();
    h->p = p;
    h->P = &m->dm->P;
    return h;
}

int chimes_lmp_set_tabulation(chimes_lmp *p, int on)
{
    if (!p) return fail("no pair style");
    chx::TableSet T;
    if (on) {
        std::string err;
        if (!chx::load_tables(*p->P, T, err)) return fail(err);
    }
    if (chx_lmp_set_tables(p->p, &T)) return fail(chx_lmp_error(p->p));
    return 0;
}

void chimes_lmp_free(chimes_lmp *p) { delete p; }

int chimes_lmp_build_mb_neighlists(chimes_lmp *p, int nall, const double *x, const int *tag, int inum,
                                   const int *ilist, const int *numneigh, int *const *firstneigh, double skin)
{
    if (!p) return fail("no handle");
    if (chx_lmp_build_lists(p->p, nall, x, tag, inum, ilist, numneigh, (const int *const *)firstneigh, skin))
        return fail(chx_lmp_error(p->p));
    return 0;
}

int chimes_lmp_compute(chimes_lmp *p, int nall, const double *x, const int *type, const int *tag, int inum,
                       const int *ilist, const int *numneigh, int *const *firstneigh, int ago, double skin,
                       int eflag_global, int eflag_atom, int vflag_global, int vflag_atom, double *f,
                       double *eng_vdwl, double *virial, double *eatom, double *vatom)
{
    if (!p) return fail("no handle");
    if (chx_lmp_compute(p->p, nall, x, type, tag, inum, ilist, numneigh, (const int *const *)firstneigh, ago, skin,
                        eflag_global, eflag_atom, vflag_global, vflag_atom, f, eng_vdwl, virial, eatom, vatom))
        return fail(chx_lmp_error(p->p));
    return 0;
}

int chimes_lmp_mb_lists(chimes_lmp *p, long *n3, int **trips, long *n4, int **quads)
{
    if (!p) return fail("no handle");
    if (chx_lmp_lists(p->p, n3, trips, n4, quads)) return fail(chx_lmp_error(p->p));
    return 0;
}

int chimes_lmp_fp_rows(chimes_lmp *p, int nb, long *nrows, double **rows)
{
    if (!p) return fail("no pair style");
    if (chx_lmp_fp_rows(p->p, nb, nrows, rows)) return fail(chx_lmp_error(p->p));
    return 0;
}

void chimes_lmp_timing(const chimes_lmp *p, double ms[4])
{
    if (p) chx_lmp_timing(p->p, ms);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Cluster fingerprints (chimesFF/src/FP/calc_histograms.cpp)
// ---------------------------------------------------------------------------
extern "C" {

int chimes_gpu_fp_edges(chimes_gpu_model *m, int nb, long n, const double *dist, const int *types, double *edges)
{
    if (!m || nb < 2 || nb > 4 || n < 0) return fail("cluster size must be 2, 3 or 4");
    const chx::ParamSet &P = m->dm->P;
    const int nt = P.natmtyps;
    for (long c = 0; c < n; c++) {
        const int *t = types + c * nb;
        for (int a = 0; a < nb; a++)
            if (t[a] < 0 || t[a] >= nt) return fail("atom type index out of range");
        int tuple = 0;
        for (int a = 0; a < nb; a++) tuple = tuple * nt + t[a];
        if ((nb == 3 && P.tripmap[tuple] < 0) || (nb == 4 && P.quadmap[tuple] < 0))
            return fail("cluster of an excluded type");
    }
    chx::EngineHandle *eh = m->dm->eh;
    std::lock_guard<std::mutex> lk(eh->mu);
    if (chx_fp_edges(eh->eng, m->dm->m, nb, n, dist, types, edges)) return fail(chx_last_error(eh->eng));
    return 0;
}

int chimes_gpu_fp_histogram(chimes_gpu_model *m, int nb, long n1, const double *edges1, long n2,
                            const double *edges2, int same, int nbin, double binw, double maxd, long long *counts,
                            long long *nsamples)
{
    if (!m || nb < 2 || nb > 4) return fail("cluster size must be 2, 3 or 4");
    if (same) {
        n2 = n1;
        edges2 = edges1;
    }
    chx::EngineHandle *eh = m->dm->eh;
    std::lock_guard<std::mutex> lk(eh->mu);
    double ms = 0.0;
    if (chx_fp_histogram(eh->eng, nb * (nb - 1) / 2, n1, edges1, n2, edges2, same, nbin, binw, maxd, counts, &ms))
        return fail(chx_last_error(eh->eng));
    if (nsamples) *nsamples = same ? n1 * (n1 - 1) / 2 : n1 * n2;
    return 0;
}

}  // extern "C"
