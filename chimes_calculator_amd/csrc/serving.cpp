// chimes_gpu.h serving entry points: a continuous-batching request loop over
// resident parameter sets.
//
// Requests (one configuration each, any model, any size) enter a FIFO.  Each
// of `nslots` workers owns an engine (its own HIP stream and device buffers,
// kept between batches) and loops: take every pending request that fits the
// batch atom limit and the HBM budget (estimated bytes of the requests in
// flight on all slots), evaluate them as one multi-model batch, deliver the
// results, repeat.  Requests that arrive while batches run wait only until a
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
    size_t est_bytes = 0;
    // results
    double energy = 0.0;
    std::vector<double> forces;
    double stress[9];
    double t_submit = 0, t_start = 0, t_done = 0;  // ms since server creation
    int status = 0;
    std::string err;
};

// Device bytes a request needs in a batch, from the density-based list sizes
// of its configuration (neighbour entries, triplet and quadruplet records,
// per-atom state and accumulators).
size_t estimate_bytes(const chx::DeviceModel &m, int natoms, const double *cell, int small)
{
    const double *a = cell, *b = cell + 3, *c = cell + 6;
    const double vol = fabs(a[0] * (b[1] * c[2] - b[2] * c[1]) - a[1] * (b[0] * c[2] - b[2] * c[0]) +
                            a[2] * (b[0] * c[1] - b[1] * c[0]));
    const double rc2 = m.rc[0], rc3 = m.P.order[1] ? m.rc[1] : 0.0, rc4 = m.P.order[2] ? m.rc[2] : 0.0;
    double lmin = 1e300;
    for (int k = 0; k < 3; k++) {
        const double *v = cell + 3 * k;
        lmin = std::min(lmin, sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]));
    }
    const int rep = small ? std::max(0, (int)ceil(rc2 / lmin) - 1) : 0;
    const double n = (double)natoms * (rep + 1) * (rep + 1) * (rep + 1);
    const double rho = natoms / std::max(vol, 1e-9);
    const double pi = 3.141592653589793;
    const double search = std::max(rc2, std::max(rc3, rc4));
    const double nb = 0.5 * rho * 4.0 / 3.0 * pi * search * search * search;
    const double n3 = 0.5 * rho * 4.0 / 3.0 * pi * rc3 * rc3 * rc3;
    const double n4 = 0.5 * rho * 4.0 / 3.0 * pi * rc4 * rc4 * rc4;
    const double per_atom = (1.3 * nb + 8) * 72.0 + (0.4 * n3 * n3 + 4) * 8.0 + (0.08 * n4 * n4 * n4 + 4) * 8.0 + 400.0;
    return (size_t)(n * per_atom) + 65536;
}

}  // namespace

struct chimes_gpu_server {
    std::vector<chimes_gpu_model *> models;
    std::vector<chx_model *> ms;
    int device = 0, precision = 0, nslots = 1;
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
    std::vector<chx_engine *> engines;
    std::vector<std::thread> workers;
    std::string err;

    double now_ms() const { return std::chrono::duration<double, std::milli>(clk::now() - t0).count(); }

    // Takes the next batch under the lock: FIFO, same replication flag, within
    // the atom limit and the budget (a lone request larger than the budget
    // runs when nothing else is in flight).
    std::vector<Request *> take()
    {
        std::vector<Request *> b;
        long atoms = 0;
        size_t bytes = 0;
        while (!pending.empty()) {
            Request *r = pending.front();
            if (!b.empty() && (r->small != b[0]->small || atoms + r->natoms > max_batch_atoms)) break;
            if (inflight + bytes + r->est_bytes > budget && !(inflight == 0 && b.empty())) break;
            pending.pop_front();
            b.push_back(r);
            atoms += r->natoms;
            bytes += r->est_bytes;
        }
        inflight += bytes;
        return b;
    }

    void run_batch(chx_engine *e, std::vector<Request *> &b)
    {
        const int n = (int)b.size();
        std::vector<int> natoms(n), model_of(n);
        std::vector<double> pos, cells(9 * (size_t)n);
        std::vector<int> types;
        const double ts = now_ms();
        for (int k = 0; k < n; k++) {
            Request *r = b[k];
            r->t_start = ts;
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
        const int rc = chx_evaluate(e, ms.data(), (int)ms.size(), &bt, &o);
        const std::string msg = rc ? chx_last_error(e) : "";
        const double td = now_ms();
        size_t off = 0;
        for (int k = 0; k < n; k++) {
            Request *r = b[k];
            r->t_done = td;
            r->status = rc;
            r->err = msg;
            if (!rc) {
                r->energy = E[k] / pow(nrep[k] + 1.0, 3.0);
                r->forces.assign(F.begin() + off, F.begin() + off + 3 * (size_t)r->natoms);
                const double *v = &V[6 * (size_t)k];
                const double s9[9] = {v[0], v[1], v[2], v[1], v[3], v[4], v[2], v[4], v[5]};
                for (int q = 0; q < 9; q++) r->stress[q] = s9[q] / vol[k];
            }
            off += 3 * (size_t)r->natoms;
        }
    }

    void worker(int slot)
    {
        chx_engine *e = engines[slot];
        for (;;) {
            std::vector<Request *> b;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv_work.wait(lk, [&] {
                    if (stop) return true;
                    if (pending.empty()) return false;
                    return inflight == 0 || inflight + pending.front()->est_bytes <= budget;
                });
                if (stop && pending.empty()) return;
                b = take();
                if (b.empty()) continue;
            }
            run_batch(e, b);
            {
                std::lock_guard<std::mutex> lk(mu);
                size_t bytes = 0;
                long atoms = 0;
                for (Request *r : b) {
                    bytes += r->est_bytes;
                    atoms += r->natoms;
                    completed.push_back(r->id);
                }
                inflight -= bytes;
                n_req += (long long)b.size();
                n_batches++;
                n_atoms += atoms;
                max_batch = std::max<long long>(max_batch, (long long)b.size());
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
    }
    s->precision = precision;
    s->nslots = nslots;
    s->max_batch_atoms = max_batch_atoms > 0 ? max_batch_atoms : 65536;
    s->budget = (size_t)(std::max(hbm_budget_gb, 0.01) * 1e9);
    s->t0 = clk::now();
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
    r->est_bytes = estimate_bytes(*s->models[model]->dm, natoms, cell, r->small);
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
        if (it == s->reqs.end() || it->second->t_done == 0) { s_err = "request not complete"; return 1; }
        r = std::move(it->second);
        s->reqs.erase(it);
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
