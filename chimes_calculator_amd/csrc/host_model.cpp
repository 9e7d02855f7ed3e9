// Dense device tables from a parameter set (host).
#include "host_model.h"

#include <cmath>
#include <cstdlib>
#include <map>

namespace chx {

namespace {

std::mutex g_mu;
std::map<int, EngineHandle *> g_engines;

void slot_consts(double rin, double rout, double lam, double *o)
{
    // x range of the Morse variable (chimesFF.h:553-559)
    const double x_min = exp(-1 * rin / lam);
    const double x_max = exp(-1 * rout / lam);
    const double x_avg = 0.5 * (x_max + x_min);
    double x_diff = 0.5 * (x_max - x_min);
    x_diff *= -1.0;
    o[0] = rin;
    o[1] = rout;
    o[2] = lam;
    o[3] = x_avg;
    o[4] = x_diff;
}

struct DescStorage {
    std::vector<double> eoff, pair_slots, c2, trip_slots, C3, quad_slots, C4, lastc3, lastc4;
    std::vector<int> pairmap, tripmap, trip_perm, quadmap, quad_perm, last3, last4;
};

// Builds the tables of one body order (3 or 4); returns false with err set.
bool cluster_tables(const ParamSet &P, int nb, int &D, std::vector<int> &map, std::vector<int> &perm,
                    std::vector<double> &slots, std::vector<double> &C, std::vector<int> &last,
                    std::vector<double> &lastc, std::string &err)
{
    const std::vector<ClusterType> &T = nb == 3 ? P.trip : P.quad;
    const std::vector<int> &fmap = nb == 3 ? P.tripmap : P.quadmap;
    const std::vector<int> &fperm = nb == 3 ? P.trip_perm : P.quad_perm;
    const std::vector<int> &fok = nb == 3 ? P.trip_perm_ok : P.quad_perm_ok;
    const int npr = nb == 3 ? 3 : 6;
    const int nt = P.natmtyps;
    int maxp = 1;
    for (const ClusterType &c : T)
        for (int v : c.powers) maxp = std::max(maxp, v);
    D = maxp + 1;
    size_t cs = 1;
    for (int q = 0; q < npr; q++) cs *= (size_t)D;
    C.assign(T.size() * cs, 0.0);
    last.assign(T.size() * npr, 0);
    lastc.assign(T.size(), 0.0);
    for (size_t t = 0; t < T.size(); t++) {
        const ClusterType &c = T[t];
        const size_t K = c.coef.size();
        for (size_t k = 0; k < K; k++) {
            size_t idx = 0;
            for (int q = 0; q < npr; q++) idx = idx * D + c.powers[k * npr + q];
            C[t * cs + idx] += c.coef[k];
        }
        if (K) {
            for (int q = 0; q < npr; q++) last[t * npr + q] = c.powers[(K - 1) * npr + q];
            lastc[t] = c.coef[K - 1];
        }
    }
    map = fmap;
    perm = fperm;
    // per (type, slot): cutoffs and the Morse lambda of the slot's pair type
    slots.assign(T.size() * npr * 5, 0.0);
    std::vector<int> have(T.size(), 0);
    std::vector<double> lam(T.size() * npr, 0.0);
    static const int pa3[3] = {0, 0, 1}, pb3[3] = {1, 2, 2};
    static const int pa4[6] = {0, 0, 0, 1, 1, 2}, pb4[6] = {1, 2, 3, 2, 3, 3};
    const int *pa = nb == 3 ? pa3 : pa4, *pb = nb == 3 ? pb3 : pb4;
    for (size_t tup = 0; tup < fmap.size(); tup++) {
        const int tau = fmap[tup];
        if (tau < 0) continue;
        if (!fok[tup]) {
            err = "no atom permutation matches the parameter pair order of cluster type " + std::to_string(tau);
            return false;
        }
        int ty[4], r = (int)tup;
        for (int a = nb - 1; a >= 0; a--) { ty[a] = r % nt; r /= nt; }
        const int *pm = &fperm[tup * nb];
        for (int q = 0; q < npr; q++) {
            const int x = pm[pa[q]], y = pm[pb[q]];
            const int lo = x < y ? x : y, hi = x < y ? y : x;
            const double l = P.morse[P.pairmap[ty[lo] * nt + ty[hi]]];
            if (!have[tau]) lam[tau * npr + q] = l;
            else if (lam[tau * npr + q] != l) {
                err = "inconsistent pair types within cluster type " + std::to_string(tau);
                return false;
            }
        }
        have[tau] = 1;
    }
    for (size_t t = 0; t < T.size(); t++)
        for (int q = 0; q < npr; q++) {
            const double l = have[t] ? lam[t * npr + q] : 1.0;
            slot_consts(T[t].rin[q], T[t].rout[q], l, &slots[(t * npr + q) * 5]);
        }
    return true;
}

}  // namespace

int default_device()
{
    const char *s = getenv("CHIMES_GPU_DEVICE");
    return s ? atoi(s) : 0;
}

EngineHandle *engine_for_device(int device, std::string &err)
{
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_engines.find(device);
    if (it != g_engines.end()) return it->second;
    chx_engine *e = chx_engine_create(device);
    if (!e) {
        err = "no usable HIP device " + std::to_string(device) + " (the ChIMES GPU library needs an AMD GPU)";
        return nullptr;
    }
    EngineHandle *h = new EngineHandle;
    h->eng = e;
    g_engines[device] = h;
    return h;
}

DeviceModel::~DeviceModel()
{
    if (m) chx_model_destroy(m);
}

std::unique_ptr<DeviceModel> make_device_model(const ParamSet &P, int device, std::string &err)
{
    EngineHandle *eh = engine_for_device(device, err);
    if (!eh) return nullptr;
    DescStorage S;
    chx_model_desc d{};
    const int nt = P.natmtyps;
    d.nt = nt;
    for (int k = 0; k < 3; k++) d.order[k] = P.order[k];
    d.npairs = P.npairs();
    d.tersoff = P.tersoff ? 1 : 0;
    d.fcut_var = P.fcut_var;
    d.pen_dist = P.penalty_dist;
    d.pen_scale = P.penalty_scale;
    S.eoff = P.energy_offsets;
    S.pairmap = P.pairmap;
    S.pair_slots.resize(d.npairs * 5);
    const int o2 = P.order[0];
    S.c2.assign((size_t)d.npairs * (o2 + 1), 0.0);
    for (int p = 0; p < d.npairs; p++) {
        slot_consts(P.pair_rin[p], P.pair_rout[p], P.morse[p], &S.pair_slots[5 * p]);
        for (size_t k = 0; k < P.pow2[p].size(); k++) S.c2[(size_t)p * (o2 + 1) + P.pow2[p][k] + 1] += P.c2[p][k];
    }
    for (int b = 0; b < 3; b++) d.rc[b] = P.max_cutoff(b + 2);
    if (P.order[1] > 0) {
        if (!cluster_tables(P, 3, d.D3, S.tripmap, S.trip_perm, S.trip_slots, S.C3, S.last3, S.lastc3, err))
            return nullptr;
        d.ntrip = (int)P.trip.size();
    }
    if (P.order[2] > 0) {
        if (!cluster_tables(P, 4, d.D4, S.quadmap, S.quad_perm, S.quad_slots, S.C4, S.last4, S.lastc4, err))
            return nullptr;
        d.nquad = (int)P.quad.size();
    }
    d.eoff = S.eoff.data();
    d.pairmap = S.pairmap.data();
    d.pair_slots = S.pair_slots.data();
    d.c2 = S.c2.data();
    d.tripmap = S.tripmap.data();
    d.trip_perm = S.trip_perm.data();
    d.trip_slots = S.trip_slots.data();
    d.C3 = S.C3.data();
    d.last3 = S.last3.data();
    d.lastc3 = S.lastc3.data();
    d.quadmap = S.quadmap.data();
    d.quad_perm = S.quad_perm.data();
    d.quad_slots = S.quad_slots.data();
    d.C4 = S.C4.data();
    d.last4 = S.last4.data();
    d.lastc4 = S.lastc4.data();
    std::unique_ptr<DeviceModel> dm(new DeviceModel);
    dm->P = P;
    dm->eh = eh;
    dm->device = device;
    {
        std::lock_guard<std::mutex> lk(eh->mu);
        dm->m = chx_model_create(eh->eng, &d);
        if (!dm->m) {
            err = chx_last_error(eh->eng);
            return nullptr;
        }
    }
    for (int b = 0; b < 3; b++) dm->rc[b] = d.rc[b];
    return dm;
}

}  // namespace chx
