// chimes_gpu.h serving entry points: a continuous-batching request loop over
// resident parameter sets.
//
// Requests (one configuration each, any model, any size) enter a FIFO.  Each
// of `nslots` workers owns an engine (its own HIP stream and device buffers,
// kept between batches) and loops: take every pending request that fits the
// batch atom limit and the HBM budget (estimated bytes of the requests in
// flight on all slots), evaluate them as one multi-model batch, deliver the
// results, repeat.  A batch that fails is evaluated again request by request,
// so an error fails only the request that causes it.  Requests that arrive while batches run wait only until a
// slot frees up; batches of different slots overlap on the device.  Every
// result is an exact fixed-point sum of per-cluster terms that do not depend
// on the other members of the batch, so a request gives the same bits alone
// and in any mix.
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "chimes_gpu.h"
#include "host_model.h"


namespace {

using clk = std::chrono::steady_clock;

struct Request {
    long id = 0;
    int model = 0, natoms = 0, small = 0;
    std::vector<double> pos;
    std::vector<int> types;
    double cell[9];
    // sizing inputs of the engine's capacity estimates (per request)
    double nrep_atoms = 0;   // atoms after replication
    double nb_cap = 0;       // neighbour-slab entries per owner
    double l3 = 0, l4 = 0;   // triplet / quadruplet list slots
    double rcap = 0;         // cluster records per owner
    double cells = 0;        // cell-grid cells
    // results: written by the worker before `done` is set under the lock
    double energy = 0.0;
    std::vector<double> forces;
    double stress[9];
    double t_submit = 0, t_start = 0, t_done = 0;  // ms since server creation
    int status = 0;
    std::string err;
    bool done = false;
};

// The engine's capacity estimates for one configuration (the density model
// of guess_capacities in engine.hip, the same ghost-layer and cell-grid
// geometry), from which the device bytes of any batch follow.
void size_request(const chx::DeviceModel &m, Request &r)
{
    const double *a = r.cell, *b = r.cell + 3, *c = r.cell + 6;
    const double vol = fabs(a[0] * (b[1] * c[2] - b[2] * c[1]) - a[1] * (b[0] * c[2] - b[2] * c[0]) +
                            a[2] * (b[0] * c[1] - b[1] * c[0]));
    const double rc2 = m.rc[0], rc3 = m.P.order[1] ? m.rc[1] : 0.0, rc4 = m.P.order[2] ? m.rc[2] : 0.0;
    double lmin = 1e300;
    for (int k = 0; k < 3; k++) {
        const double *v = r.cell + 3 * k;
        lmin = std::min(lmin, sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]));
    }
    const int rep = r.small ? std::max(0, (int)ceil(rc2 / lmin) - 1) : 0;
    const double ncopy = (rep + 1.0) * (rep + 1.0) * (rep + 1.0);
    const double n = (double)r.natoms * ncopy;
    const double rho = n / std::max(vol * ncopy, 1e-9);
    const double pi = 3.141592653589793;
    const double search = std::max(rc2, std::max(rc3, rc4));
    const double nb = 0.5 * rho * 4.0 / 3.0 * pi * search * search * search;
    const double n3 = 0.5 * rho * 4.0 / 3.0 * pi * rc3 * rc3 * rc3;
    const double n4 = 0.5 * rho * 4.0 / 3.0 * pi * rc4 * rc4 * rc4;
    r.nrep_atoms = n;
    r.nb_cap = std::min(512.0, ceil((1.4 * nb + 12) / 8) * 8);
    r.l3 = rc3 > 0 ? n * (0.4 * n3 * n3 + 4) : 0.0;
    r.l4 = rc4 > 0 ? n * (0.08 * n4 * n4 * n4 + 4) : 0.0;
    // record slabs: the engine starts at twice the mean clusters per owner
    // and grows to the largest owner's count; three times the mean bounds
    // the owners of uniform configurations
    const double rmean = (rc3 > 0 ? 0.4 * n3 * n3 + 4 : 0.0) + (rc4 > 0 ? 0.08 * n4 * n4 * n4 + 4 : 0.0);
    r.rcap = rmean > 0 ? 3.0 * rmean + 64 : 0.0;
    // half-search-radius cells over the replicated cell
    const double w = 0.5 * search;
    r.cells = std::max(1.0, vol * ncopy / (w * w * w)) + 64.0;
}

// Device bytes an engine asks for when it evaluates these requests as one
// batch: the work buffers of engine.hip's layout() and size_work() at the
// estimated capacities (neighbour slabs at the batch's largest per-owner
// capacity), the blobs, and a 50% margin for capacities that a retry grows
// past the density estimate (an owner with more images than the estimate).
size_t batch_bytes(const std::vector<Request *> &b, int groups)
{
    double N = 0, Nin = 0, nb_cap = 8, l3 = 0, l4 = 0, cells = 0, rcap = 0;
    for (const Request *r : b) {
        N += r->nrep_atoms;
        Nin += r->natoms;
        nb_cap = std::max(nb_cap, r->nb_cap);
        rcap = std::max(rcap, r->rcap);
        l3 += r->l3;
        l4 += r->l4;
        cells += r->cells;
    }
    const double ncfg = (double)b.size();
    l3 += 64.0 * (groups + 1) + 256;
    l4 += 64.0 * (groups + 1) + 256;
    const double per_atom = 64.0             // atom records (two frames)
                            + 16.0           // cell index, neighbour counts, offsets, queue
                            + 8.0 * groups   // per-(group, owner) cluster counts and their scan
                            + 48.0           // force accumulators
                            + 68.0 * nb_cap  // neighbour slabs
                            + 8.0 * rcap + 4 // cluster record slabs and counts
                            + 8.0;           // scan partials
    const double bytes = N * per_atom + 8.0 * (l3 + l4) + 16.0 * cells + 40.0 * Nin  // blobs: positions, types, forces
                         + (1024.0 + 4096.0) * ncfg + 65536.0;  // + per-configuration accumulator lines
    return (size_t)(1.5 * bytes);
}

}  // namespace

struct chimes_gpu_server {
    std::vector<chimes_gpu_model *> models;
    std::vector<chx_model *> ms;
    int device = 0, precision = 0, nslots = 1;
    int groups = 0;  // cluster-type groups of all models (a batch may hold every one)
    long max_batch_atoms = 0;
    size_t budget = 0;
    clk::time_point t0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request *> pending;
    std::unordered_map<long, std::unique_ptr<Request>> reqs;
    std::deque<long> completed;
    size_t inflight = 0;
    bool stop = false;
    long next_id = 1;
    long long n_req = 0, n_batches = 0, n_atoms = 0, max_batch = 0;
    long long n_failed = 0, n_isolated = 0, n_under = 0;
    double peak_inflight = 0, peak_held = 0, worst_ratio = 0;
    std::vector<size_t> held;  // device bytes held by each slot's engine
    std::vector<chx_engine *> engines;
    std::vector<std::thread> workers;
    std::string err;

    double now_ms() const { return std::chrono::duration<double, std::milli>(clk::now() - t0).count(); }

    // Takes the next batch under the lock: FIFO, same replication flag, within
    // the atom limit and the budget (the estimated bytes of this batch plus
    // those of the batches in flight; a lone request larger than the budget
    // runs when nothing else is in flight).  Returns the batch's bytes.
    size_t take(std::vector<Request *> &b)
    {
        long atoms = 0;
        size_t bytes = 0;
        while (!pending.empty()) {
            Request *r = pending.front();
            if (!b.empty() && (r->small != b[0]->small || atoms + r->natoms > max_batch_atoms)) break;
            b.push_back(r);
            const size_t nb = batch_bytes(b, groups);
            if (inflight + nb > budget && !(inflight == 0 && b.size() == 1)) {
                b.pop_back();
                break;
            }
            pending.pop_front();
            atoms += r->natoms;
            bytes = nb;
        }
        inflight += bytes;
        peak_inflight = std::max(peak_inflight, (double)inflight);
        return bytes;
    }

    // Evaluates requests as one batch; on success writes their results.
    int eval(chx_engine *e, Request *const *b, int n, std::string &msg)
    {
        std::vector<int> natoms(n), model_of(n);
        std::vector<double> pos, cells(9 * (size_t)n);
        std::vector<int> types;
        for (int k = 0; k < n; k++) {
            const Request *r = b[k];
            natoms[k] = r->natoms;
            model_of[k] = r->model;
            pos.insert(pos.end(), r->pos.begin(), r->pos.end());
            types.insert(types.end(), r->types.begin(), r->types.end());
            memcpy(&cells[9 * (size_t)k], r->cell, sizeof(r->cell));
        }
        chx_batch bt{};
        bt.nconf = n;
        bt.natoms = natoms.data();
        bt.pos = pos.data();
        bt.types = types.data();
        bt.cells = cells.data();
        bt.model = model_of.data();
        bt.small = b[0]->small;
        bt.precision = precision;
        std::vector<double> E(n), F(pos.size()), V(6 * (size_t)n), vol(n);
        std::vector<int> nrep(n);
        chx_out o{};
        o.energy = E.data();
        o.forces = F.data();
        o.virial = V.data();
        o.volume = vol.data();
        o.n_rep = nrep.data();
        // size the buffers for this batch alone (allocations are reused)
        chx_engine_reset_capacities(e);
        if (chx_evaluate(e, ms.data(), (int)ms.size(), &bt, &o)) {
            msg = chx_last_error(e);
            return 1;
        }
        size_t off = 0;
        for (int k = 0; k < n; k++) {
            Request *r = b[k];
            r->status = 0;
            r->err.clear();
            r->energy = E[k] / pow(nrep[k] + 1.0, 3.0);
            r->forces.assign(F.begin() + off, F.begin() + off + 3 * (size_t)r->natoms);
            const double *v = &V[6 * (size_t)k];
            const double s9[9] = {v[0], v[1], v[2], v[1], v[3], v[4], v[2], v[4], v[5]};
            for (int q = 0; q < 9; q++) r->stress[q] = s9[q] / vol[k];
            off += 3 * (size_t)r->natoms;
        }
        return 0;
    }

    // One batch.  When the batch fails, each member is evaluated alone, so an
    // error fails only the request that causes it.  Returns the device bytes
    // the batch's evaluations asked for (the largest of them).
    size_t run_batch(chx_engine *e, std::vector<Request *> &b, bool &isolated)
    {
        const double ts = now_ms();
        for (Request *r : b) r->t_start = ts;
        std::string msg;
        size_t hn[2];
        isolated = false;
        const int rc = eval(e, b.data(), (int)b.size(), msg);
        chx_engine_bytes(e, hn);
        size_t need = hn[1];
        if (rc && b.size() == 1) {
            b[0]->status = rc;
            b[0]->err = msg;
        } else if (rc) {
            isolated = true;
            for (Request *r : b) {
                std::string m1;
                if (eval(e, &r, 1, m1)) {
                    r->status = 1;
                    r->err = m1;
                }
                chx_engine_bytes(e, hn);
                need = std::max(need, hn[1]);
            }
        }
        const double td = now_ms();
        for (Request *r : b) r->t_done = td;
        return need;
    }

    void worker(int slot)
    {
        chx_engine *e = engines[slot];
        for (;;) {
            std::vector<Request *> b;
            size_t bytes = 0;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv_work.wait(lk, [&] {
                    if (stop) return true;
                    if (pending.empty()) return false;
                    if (inflight == 0) return true;
                    std::vector<Request *> one{pending.front()};
                    return inflight + batch_bytes(one, groups) <= budget;
                });
                if (stop && pending.empty()) return;
                bytes = take(b);
                if (b.empty()) continue;
            }
            bool isolated = false;
            const size_t need = run_batch(e, b, isolated);
            // device memory held by this slot: released when all slots
            // together hold more than the budget
            size_t hn[2];
            chx_engine_bytes(e, hn);
            bool trim = false;
            {
                std::lock_guard<std::mutex> lk(mu);
                held[slot] = hn[0];
                size_t total = 0;
                for (size_t h : held) total += h;
                peak_held = std::max(peak_held, (double)total);
                trim = total > budget;
                if (trim) held[slot] = 0;
            }
            if (trim) chx_engine_trim(e);
            {
                std::lock_guard<std::mutex> lk(mu);
                long atoms = 0;
                for (Request *r : b) {
                    atoms += r->natoms;
                    if (r->status) n_failed++;
                    completed.push_back(r->id);
                }
                worst_ratio = std::max(worst_ratio, (double)need / std::max<double>(1.0, (double)bytes));
                if (need > bytes) n_under++;
                if (isolated) n_isolated++;
                inflight -= bytes;
                n_req += (long long)b.size();
                n_batches++;
                n_atoms += atoms;
                max_batch = std::max<long long>(max_batch, (long long)b.size());
                // last: from here on result() may release the requests
                for (Request *r : b) r->done = true;
            }
            cv_done.notify_all();
            cv_work.notify_all();
        }
    }
};

namespace {
thread_local std::string s_err;
}

extern "C" {

const char *chimes_gpu_server_last_error(void) { return s_err.c_str(); }

chimes_gpu_server *chimes_gpu_server_create(chimes_gpu_model *const *models, int nmodels, int precision,
                                            int nslots, long max_batch_atoms, double hbm_budget_gb)
{
    if (nmodels <= 0 || !models) { s_err = "no models"; return nullptr; }
    if (nslots < 1 || nslots > 8) { s_err = "nslots must be 1..8"; return nullptr; }
    std::unique_ptr<chimes_gpu_server> s(new chimes_gpu_server);
    s->device = models[0]->dm->device;
    for (int k = 0; k < nmodels; k++) {
        if (!models[k] || models[k]->dm->device != s->device) { s_err = "models must share one device"; return nullptr; }
        s->models.push_back(models[k]);
        s->ms.push_back(models[k]->dm->m);
        const chx::ParamSet &P = models[k]->dm->P;
        if (P.order[1] > 0) s->groups += (int)P.trip.size() + 1;
        if (P.order[2] > 0) s->groups += (int)P.quad.size() + 1;
    }
    s->precision = precision;
    s->nslots = nslots;
    s->max_batch_atoms = max_batch_atoms > 0 ? max_batch_atoms : 65536;
    s->budget = (size_t)(std::max(hbm_budget_gb, 0.01) * 1e9);
    s->t0 = clk::now();
    s->held.assign(nslots, 0);
    for (int k = 0; k < nslots; k++) {
        chx_engine *e = chx_engine_create(s->device);
        if (!e) {
            for (chx_engine *x : s->engines) chx_engine_destroy(x);
            s_err = "cannot create an engine";
            return nullptr;
        }
        s->engines.push_back(e);
    }
    chimes_gpu_server *sp = s.get();
    for (int k = 0; k < nslots; k++) s->workers.emplace_back([sp, k] { sp->worker(k); });
    return s.release();
}

long chimes_gpu_server_submit(chimes_gpu_server *s, int model, int natoms, const double *pos, const int *types,
                              const double *cell, int small)
{
    if (model < 0 || model >= (int)s->models.size()) { s_err = "model index out of range"; return -1; }
    if (natoms <= 0) { s_err = "configuration without atoms"; return -1; }
    const int nt = s->models[model]->dm->P.natmtyps;
    for (int k = 0; k < natoms; k++)
        if (types[k] < 0 || types[k] >= nt) { s_err = "atom type index out of range"; return -1; }
    std::unique_ptr<Request> r(new Request);
    r->model = model;
    r->natoms = natoms;
    r->small = small ? 1 : 0;
    r->pos.assign(pos, pos + 3 * (size_t)natoms);
    r->types.assign(types, types + natoms);
    memcpy(r->cell, cell, sizeof(r->cell));
    size_request(*s->models[model]->dm, *r);
    long id;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        id = s->next_id++;
        r->id = id;
        r->t_submit = s->now_ms();
        s->pending.push_back(r.get());
        s->reqs[id] = std::move(r);
    }
    s->cv_work.notify_one();
    return id;
}

long chimes_gpu_server_wait(chimes_gpu_server *s, long id, double timeout_s)
{
    std::unique_lock<std::mutex> lk(s->mu);
    const auto until = clk::now() + std::chrono::duration_cast<clk::duration>(std::chrono::duration<double>(timeout_s));
    for (;;) {
        if (id > 0) {
            auto it = s->reqs.find(id);
            if (it == s->reqs.end()) { s_err = "unknown request"; return -1; }
            for (auto c = s->completed.begin(); c != s->completed.end(); ++c)
                if (*c == id) {
                    s->completed.erase(c);
                    return id;
                }
        } else if (!s->completed.empty()) {
            const long c = s->completed.front();
            s->completed.pop_front();
            return c;
        }
        if (s->cv_done.wait_until(lk, until) == std::cv_status::timeout) return 0;
    }
}

int chimes_gpu_server_result(chimes_gpu_server *s, long id, double *energy, double *forces, double *stress,
                             double times_ms[3])
{
    std::unique_ptr<Request> r;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->reqs.find(id);
        if (it == s->reqs.end() || !it->second->done) { s_err = "request not complete"; return 1; }
        r = std::move(it->second);
        s->reqs.erase(it);
        for (auto c = s->completed.begin(); c != s->completed.end(); ++c)
            if (*c == id) {
                s->completed.erase(c);
                break;
            }
    }
    if (r->status) { s_err = r->err; return 1; }
    *energy = r->energy;
    memcpy(forces, r->forces.data(), r->forces.size() * sizeof(double));
    memcpy(stress, r->stress, sizeof(r->stress));
    if (times_ms) {
        times_ms[0] = r->t_submit;
        times_ms[1] = r->t_start;
        times_ms[2] = r->t_done;
    }
    return 0;
}

void chimes_gpu_server_stats(chimes_gpu_server *s, long long stats[4])
{
    std::lock_guard<std::mutex> lk(s->mu);
    stats[0] = s->n_req;
    stats[1] = s->n_batches;
    stats[2] = s->n_atoms;
    stats[3] = s->max_batch;
}

void chimes_gpu_server_memory(chimes_gpu_server *s, double out[5])
{
    std::lock_guard<std::mutex> lk(s->mu);
    out[0] = (double)s->budget;
    out[1] = s->peak_inflight;
    out[2] = s->peak_held;
    out[3] = s->worst_ratio;
    out[4] = (double)s->n_under;
}

void chimes_gpu_server_failures(chimes_gpu_server *s, long long out[2])
{
    std::lock_guard<std::mutex> lk(s->mu);
    out[0] = s->n_failed;
    out[1] = s->n_isolated;
}

void chimes_gpu_server_free(chimes_gpu_server *s)
{
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (std::thread &t : s->workers) t.join();
    for (chx_engine *e : s->engines) chx_engine_destroy(e);
    delete s;
}

}  // extern "C"
