// HIP engine: device-resident parameter tables, per-configuration frame
// set-up and the evaluation pipeline behind the internal C ABI (chx_engine.h).
//
// Pipeline of one batch (all on one stream):
//   k_prep        replicate, wrap, evaluation and neighbour frames, cell index
//   cells         count, scan, fill, sort (fixed atom order per cell)
//   k_nbr         image neighbour lists (count, scan, fill)
//   k_enum        sort each owner's images, 3-/4-body lists grouped by
//                 (model, cluster type) so evaluation waves share coefficients
//   reverse lists entries grouped by parent atom
//   k_eval2/3/4   per-cluster energies and pair force scalars
//   k_reduce      owner-ordered pair-cell sums -> forces, virial, energy
//   gathers       forces per replica and per input atom, energy and virial
//                 per configuration
// Host frame set-up follows serial_chimes_interface.cpp:150-476 operation by
// operation (this file is built with -ffp-contract=off).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "chx_engine.h"
#include "chx_eval.h"
#include "chx_kernels.h"

using namespace chx;

namespace {

#define HIPCHK(x)                                                                                   \
    do {                                                                                            \
        hipError_t err_ = (x);                                                                      \
        if (err_ != hipSuccess) throw EngineError(std::string(#x) + ": " + hipGetErrorString(err_)); \
    } while (0)

struct EngineError {
    std::string msg;
    explicit EngineError(std::string m) : msg(std::move(m)) {}
};

template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t cap = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p(o.p), cap(o.cap) { o.p = nullptr; o.cap = 0; }
    ~DevBuf() { release(); }
    void ensure(size_t n)
    {
        if (n <= cap && p) return;
        if (p) (void)hipFree(p);
        p = nullptr;
        size_t c = std::max<size_t>(n, 16);
        c += c / 4;
        HIPCHK(hipMalloc((void **)&p, c * sizeof(T)));
        cap = c;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
};

template <typename T>
T *upload(const T *h, size_t n)
{
    T *d = nullptr;
    if (n == 0) return nullptr;
    HIPCHK(hipMalloc((void **)&d, n * sizeof(T)));
    HIPCHK(hipMemcpy(d, h, n * sizeof(T), hipMemcpyHostToDevice));
    return d;
}

unsigned nblk(long n, int b) { return (unsigned)std::max<long>(1, (n + b - 1) / b); }

// ---------------------------------------------------------------------------
// Host frame arithmetic (serial_chimes_interface.cpp:20-113, 342-476)
// ---------------------------------------------------------------------------
double mag3(double a, double b, double c)
{
    double m = 0;
    m += a * a;
    m += b * b;
    m += c * c;
    return sqrt(m);
}

double dot3(const double *a, const double *b)
{
    double d = 0;
    d += a[0] * b[0];
    d += a[1] * b[1];
    d += a[2] * b[2];
    return d;
}

double angle3(const double *a, const double *b)
{
    double ang = dot3(a, b);
    ang /= mag3(a[0], a[1], a[2]);
    ang /= mag3(b[0], b[1], b[2]);
    return acos(ang);
}

void set_hmat(const double *ca, const double *cb, const double *cc, double *h, double *ih, int rep)
{
    h[0] = ca[0] * (rep + 1); h[3] = ca[1] * (rep + 1); h[6] = ca[2] * (rep + 1);
    h[1] = cb[0] * (rep + 1); h[4] = cb[1] * (rep + 1); h[7] = cb[2] * (rep + 1);
    h[2] = cc[0] * (rep + 1); h[5] = cc[1] * (rep + 1); h[8] = cc[2] * (rep + 1);
    const double det = h[0] * (h[4] * h[8] - h[5] * h[7]) - h[1] * (h[3] * h[8] - h[5] * h[6]) +
                       h[2] * (h[3] * h[7] - h[4] * h[6]);
    double t[9];
    t[0] = (h[4] * h[8] - h[5] * h[7]);
    t[3] = -1 * (h[1] * h[8] - h[2] * h[7]);
    t[6] = (h[1] * h[5] - h[2] * h[4]);
    t[1] = -1 * (h[3] * h[8] - h[5] * h[6]);
    t[4] = (h[0] * h[8] - h[2] * h[6]);
    t[7] = -1 * (h[0] * h[5] - h[2] * h[3]);
    t[2] = (h[3] * h[7] - h[4] * h[6]);
    t[5] = -1 * (h[0] * h[7] - h[1] * h[6]);
    t[8] = (h[0] * h[4] - h[1] * h[3]);
    ih[0] = t[0]; ih[3] = t[1]; ih[6] = t[2];
    ih[1] = t[3]; ih[4] = t[4]; ih[7] = t[5];
    ih[2] = t[6]; ih[5] = t[7]; ih[8] = t[8];
    for (int k = 0; k < 9; k++) ih[k] /= det;
}

void col(const double *h, int k, double *v) { v[0] = h[k]; v[1] = h[3 + k]; v[2] = h[6 + k]; }

void reorient(const double *h, double *hn, double *ihn)
{
    double A[3], B[3], C[3];
    col(h, 0, A); col(h, 1, B); col(h, 2, C);
    const double ma = mag3(A[0], A[1], A[2]);
    double u[3] = {0, 0, 0};
    u[0] += A[0] / ma; u[1] += A[1] / ma; u[2] += A[2] / ma;
    double cr[3];
    cr[0] = (u[1] * B[2] - u[2] * B[1]);
    cr[1] = -1 * (u[0] * B[2] - u[2] * B[0]);
    cr[2] = (u[0] * B[1] - u[1] * B[0]);
    double na[3] = {ma, 0, 0};
    double nb[3] = {dot3(B, u), mag3(cr[0], cr[1], cr[2]), 0};
    double nc[3];
    nc[0] = dot3(C, u);
    nc[1] = (dot3(B, C) - nb[0] * nc[0]) / nb[1];
    const double mc = mag3(C[0], C[1], C[2]);
    nc[2] = sqrt(mc * mc - nc[0] * nc[0] - nc[1] * nc[1]);
    set_hmat(na, nb, nc, hn, ihn, 0);
}

void lattice_lengths(const double *h, double *l)
{
    l[0] = mag3(h[0], h[3], h[6]);
    l[1] = mag3(h[1], h[4], h[7]);
    l[2] = mag3(h[2], h[5], h[8]);
}

// ---------------------------------------------------------------------------
// Models
// ---------------------------------------------------------------------------
struct TypeLaunch {
    chx_slot s[6];
    const double *C = nullptr;   // device, FP64
    const float *Cf = nullptr;   // device, FP32 copy for the mixed path
};

}  // namespace

struct chx_model {
    chx_engine *eng = nullptr;
    chx_model_desc d{};
    chx_model_dev dev{};
    std::vector<void *> allocs;
    std::vector<TypeLaunch> t3, t4;
    std::vector<int> tripmap_h, quadmap_h;
    double rc[3];
    ~chx_model()
    {
        for (void *p : allocs) (void)hipFree(p);
    }
    template <typename T>
    const T *up(const T *h, size_t n)
    {
        T *d = upload(h, n);
        if (d) allocs.push_back(d);
        return d;
    }
};

struct chx_engine {
    int device = 0;
    hipStream_t st = nullptr;
    hipEvent_t ev[6];
    std::string err;
    double timing[4] = {0, 0, 0, 0};
    // batch state
    std::vector<chx_cfg> cfg_h;
    std::vector<int> cfg_atom_off_h, cfg_in_off_h;
    int N = 0, Nin = 0, ncfg = 0, ncell = 0;
    long nent = 0;
    int G3 = 0, G4 = 0;
    std::vector<int> g3b, g4b;
    std::vector<int> goff3, goff4;  // group boundaries (G+1)
    std::vector<chx_model *> models;
    // device buffers
    DevBuf<chx_cfg> cfgs;
    DevBuf<int> cfg_atom_off, cfg_in_off;
    DevBuf<double> pos;
    DevBuf<int> types;
    DevBuf<chx_atom_s> as;
    DevBuf<chx_atom_n> an;
    DevBuf<int> cell_cnt, cell_start, cell_cur, cell_atoms;
    DevBuf<int> nbr_cnt, nbr_off, e_owner;
    DevBuf<chx_entry_s> es, es2;
    DevBuf<chx_entry_n> en, en2;
    DevBuf<unsigned char> e_in3;
    DevBuf<int> cnt3, off3, cnt4, off4, status, gtmp;
    DevBuf<int2> list3, list4;
    DevBuf<int> rcnt, rstart, rcur, rev;
    DevBuf<double> rec2E, rec2F, rec3E, rec3F, rec4E, rec4F;
    DevBuf<int> rec3M;
    DevBuf<long long> rec4M;
    DevBuf<double> Fe, Fself, Eown, Sown, Frep, Fout, Ecfg, Vcfg;
    DevBuf<int> cfg_stat;
    DevBuf<chx_enum_model> em;
    DevBuf<chx_red_model> rm;
    DevBuf<chx_model_dev> mdev;
    std::vector<DevBuf<int>> scan_tmp;
    // per-cluster API buffers
    DevBuf<double> cl_in, cl_out;
    DevBuf<int> cl_types, cl_flags;

    void scan(const int *in, int *out, long n, int level = 0)
    {
        if (scan_tmp.size() < 32) scan_tmp.resize(32);  // fixed: references below stay valid
        if (level >= 15) throw EngineError("scan too deep");
        const long nb = (n + 2047) / 2048;
        DevBuf<int> &bsum = scan_tmp[2 * level];
        DevBuf<int> &bscan = scan_tmp[2 * level + 1];
        bsum.ensure(nb + 1);
        bscan.ensure(nb + 2);
        if (n > 0) hipLaunchKernelGGL(k_scan_block, dim3(nblk(n, 2048)), dim3(256), 0, st, in, out, n, bsum.p);
        if (nb > 1) {
            scan(bsum.p, bscan.p, nb, level + 1);
            hipLaunchKernelGGL(k_scan_add, dim3(nblk(n, 256)), dim3(256), 0, st, out, n, bscan.p);
        }
        hipLaunchKernelGGL(k_scan_total, dim3(1), dim3(1), 0, st, out, n, in);
    }
    void zero_i(int *p, long n, int v = 0)
    {
        if (n > 0) hipLaunchKernelGGL(k_fill_i32, dim3(std::min<long>(nblk(n, 256), 4096)), dim3(256), 0, st, p, n, v);
    }
    int read_int(const int *p)
    {
        int v;
        HIPCHK(hipMemcpyAsync(&v, p, sizeof(int), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        return v;
    }
};

extern "C" {

chx_engine *chx_engine_create(int device)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return nullptr;
    if (device < 0 || device >= n) return nullptr;
    if (hipSetDevice(device) != hipSuccess) return nullptr;
    chx_engine *e = new chx_engine;
    e->device = device;
    if (hipStreamCreateWithFlags(&e->st, hipStreamNonBlocking) != hipSuccess) { delete e; return nullptr; }
    for (int k = 0; k < 6; k++) (void)hipEventCreate(&e->ev[k]);
    return e;
}

void chx_engine_destroy(chx_engine *e)
{
    if (!e) return;
    (void)hipSetDevice(e->device);
    (void)hipStreamSynchronize(e->st);
    for (int k = 0; k < 6; k++) (void)hipEventDestroy(e->ev[k]);
    (void)hipStreamDestroy(e->st);
    delete e;
}

const char *chx_last_error(chx_engine *e) { return e ? e->err.c_str() : "no engine"; }

void chx_free(void *p) { free(p); }

void chx_last_counts(chx_engine *e, long long *c3)
{
    c3[0] = e->nent;
    c3[1] = e->goff3.empty() ? 0 : e->goff3.back();
    c3[2] = e->goff4.empty() ? 0 : e->goff4.back();
}

void chx_last_timing(chx_engine *e, double *ms4)
{
    for (int k = 0; k < 4; k++) ms4[k] = e->timing[k];
}

chx_model *chx_model_create(chx_engine *e, const chx_model_desc *d)
{
    try {
        HIPCHK(hipSetDevice(e->device));
        if (d->nt <= 0 || d->nt > CHX_MAX_TYPES) throw EngineError("unsupported number of atom types");
        if (d->ntrip > CHX_MAX_CLUSTER_TYPES || d->nquad > CHX_MAX_CLUSTER_TYPES)
            throw EngineError("too many cluster types");
        if (d->order[1] > 0 && !supported_D3(d->D3)) throw EngineError("unsupported 3-body order");
        if (d->order[2] > 0 && !supported_D4(d->D4)) throw EngineError("unsupported 4-body order");
        chx_model *m = new chx_model;
        m->eng = e;
        m->d = *d;
        const int nt = d->nt;
        chx_model_dev &v = m->dev;
        v.nt = nt;
        for (int k = 0; k < 3; k++) { v.order[k] = d->order[k]; m->rc[k] = d->rc[k]; }
        v.npairs = d->npairs;
        v.tersoff = d->tersoff;
        v.fcut_var = d->fcut_var;
        v.pen_dist = d->pen_dist;
        v.pen_scale = d->pen_scale;
        v.eoff = m->up(d->eoff, nt);
        v.pairmap = m->up(d->pairmap, (size_t)nt * nt);
        std::vector<chx_slot> s2(d->npairs);
        for (int p = 0; p < d->npairs; p++) {
            const double *q = d->pair_slots + 5 * p;
            s2[p] = chx_slot{q[0], q[1], q[2], q[3], q[4]};
        }
        v.slots2 = m->up(s2.data(), s2.size());
        v.c2 = m->up(d->c2, (size_t)d->npairs * (d->order[0] + 1));
        v.ntrip = d->ntrip;
        v.D3 = d->D3;
        v.nquad = d->nquad;
        v.D4 = d->D4;
        if (d->order[1] > 0 && d->ntrip > 0) {
            const size_t nt3 = (size_t)nt * nt * nt, cs = (size_t)d->D3 * d->D3 * d->D3;
            v.tripmap = m->up(d->tripmap, nt3);
            v.trip_perm = m->up(d->trip_perm, nt3 * 3);
            std::vector<chx_slot> s3(d->ntrip * 3);
            for (int k = 0; k < d->ntrip * 3; k++) {
                const double *q = d->trip_slots + 5 * k;
                s3[k] = chx_slot{q[0], q[1], q[2], q[3], q[4]};
            }
            v.slots3 = m->up(s3.data(), s3.size());
            v.C3 = m->up(d->C3, d->ntrip * cs);
            v.last3 = m->up(d->last3, (size_t)d->ntrip * 3);
            v.lastc3 = m->up(d->lastc3, d->ntrip);
            std::vector<float> cf(d->C3, d->C3 + d->ntrip * cs);
            const float *Cf = m->up(cf.data(), cf.size());
            m->t3.resize(d->ntrip);
            for (int t = 0; t < d->ntrip; t++) {
                for (int q = 0; q < 3; q++) m->t3[t].s[q] = s3[3 * t + q];
                m->t3[t].C = v.C3 + t * cs;
                m->t3[t].Cf = Cf + t * cs;
            }
            m->tripmap_h.assign(d->tripmap, d->tripmap + nt3);
        }
        if (d->order[2] > 0 && d->nquad > 0) {
            const size_t nt4 = (size_t)nt * nt * nt * nt;
            size_t cs = 1;
            for (int k = 0; k < 6; k++) cs *= d->D4;
            v.quadmap = m->up(d->quadmap, nt4);
            v.quad_perm = m->up(d->quad_perm, nt4 * 4);
            std::vector<chx_slot> s4(d->nquad * 6);
            for (int k = 0; k < d->nquad * 6; k++) {
                const double *q = d->quad_slots + 5 * k;
                s4[k] = chx_slot{q[0], q[1], q[2], q[3], q[4]};
            }
            v.slots4 = m->up(s4.data(), s4.size());
            v.C4 = m->up(d->C4, d->nquad * cs);
            v.last4 = m->up(d->last4, (size_t)d->nquad * 6);
            v.lastc4 = m->up(d->lastc4, d->nquad);
            std::vector<float> cf(d->C4, d->C4 + d->nquad * cs);
            const float *Cf = m->up(cf.data(), cf.size());
            m->t4.resize(d->nquad);
            for (int t = 0; t < d->nquad; t++) {
                for (int q = 0; q < 6; q++) m->t4[t].s[q] = s4[6 * t + q];
                m->t4[t].C = v.C4 + t * cs;
                m->t4[t].Cf = Cf + t * cs;
            }
            m->quadmap_h.assign(d->quadmap, d->quadmap + nt4);
        }
        return m;
    } catch (const EngineError &x) {
        e->err = x.msg;
        return nullptr;
    }
}

void chx_model_destroy(chx_model *m) { delete m; }

int chx_evaluate(chx_engine *e, chx_model *const *models, int nmodels, const chx_batch *b, chx_out *o)
{
    try {
        HIPCHK(hipSetDevice(e->device));
        hipStream_t st = e->st;
        const int ncfg = b->nconf;
        if (ncfg <= 0) throw EngineError("empty batch");
        e->ncfg = ncfg;
        e->models.assign(models, models + nmodels);
        // ---------------- host frame set-up per configuration --------------
        e->cfg_h.assign(ncfg, chx_cfg{});
        e->cfg_atom_off_h.assign(ncfg + 1, 0);
        e->cfg_in_off_h.assign(ncfg + 1, 0);
        long N = 0, Nin = 0, ncell = 0;
        for (int c = 0; c < ncfg; c++) {
            chx_cfg &C = e->cfg_h[c];
            const int mi = b->model[c];
            if (mi < 0 || mi >= nmodels) throw EngineError("model index out of range");
            const chx_model *M = models[mi];
            const double rc2 = M->rc[0], rc3 = M->rc[1], rc4 = M->rc[2];
            const double *ca = b->cells + 9 * c, *cb = ca + 3, *cc = ca + 6;
            const int nin = b->natoms[c];
            if (nin <= 0) throw EngineError("configuration without atoms");
            double la = mag3(ca[0], ca[1], ca[2]), lb = mag3(cb[0], cb[1], cb[2]), lc = mag3(cc[0], cc[1], cc[2]);
            double mn = la;
            if (lb < mn) mn = lb;
            if (lc < mn) mn = lc;
            C.n_rep = b->small ? (int)(ceil(rc2 / mn) - 1) : 0;
            if (C.n_rep < 0) C.n_rep = 0;
            set_hmat(ca, cb, cc, C.h0, C.ih0, 0);
            set_hmat(ca, cb, cc, C.h, C.ih, C.n_rep);
            double L[3];
            lattice_lengths(C.h, L);
            double A[3], B[3], Cv[3];
            col(C.h, 0, A); col(C.h, 1, B); col(C.h, 2, Cv);
            const double al = angle3(B, Cv), be = angle3(Cv, A), ga = angle3(A, B);
            double vv = 1;
            vv += 2 * cos(al) * cos(be) * cos(ga);
            vv -= cos(al) * cos(al);
            vv -= cos(be) * cos(be);
            vv -= cos(ga) * cos(ga);
            C.vol = L[0] * L[1] * L[2] * sqrt(vv);
            // ghost layers (serial_chimes_interface.cpp:477-534) in both frames
            double lm = L[0];
            if (L[1] < lm) lm = L[1];
            if (L[2] < lm) lm = L[2];
            const int nl_s = (int)ceil(rc2 * 2.0 / lm + 1);
            reorient(C.h, C.hn, C.ihn);
            double Ln[3];
            lattice_lengths(C.hn, Ln);
            double lmn = Ln[0];
            if (Ln[1] < lmn) lmn = Ln[1];
            if (Ln[2] < lmn) lmn = Ln[2];
            C.n_layers = (int)ceil(rc2 * 2.0 / lmn + 1);
            if (C.n_layers != nl_s) throw EngineError("ghost layer counts of the two frames differ");
            const int nl = C.n_layers;
            for (int f = 0; f < 2; f++) {
                const double *LL = f ? Ln : L;
                const double ex = LL[0] * (2 * nl + 1), ey = LL[1] * (2 * nl + 1), ez = LL[2] * (2 * nl + 1);
                if (rc2 > 0.5 * ex || rc2 > 0.5 * ey || rc2 > 0.5 * ez ||
                    (M->d.order[1] > 0 && (rc3 > 0.5 * ex || rc3 > 0.5 * ey || rc3 > 0.5 * ez)) ||
                    (M->d.order[2] > 0 && (rc4 > 0.5 * ex || rc4 > 0.5 * ey || rc4 > 0.5 * ez)))
                    throw EngineError("Maximum cutoff is greater than half at least one box length.");
            }
            {
                const double mc[3] = {rc2, rc3, rc4};
                for (int k = 0; k < 3; k++)
                    if (mc[k] > 2 * L[0] * (2 * nl + 1) || mc[k] > 2 * L[1] * (2 * nl + 1) ||
                        mc[k] > 2 * L[2] * (2 * nl + 1))
                        throw EngineError("Layered system is smaller than 2x the model maximum outer cutoff.");
            }
            C.search = rc2;
            if (rc3 > C.search) C.search = rc3;
            if (rc4 > C.search) C.search = rc4;
            C.rc3 = M->d.order[1] > 0 ? rc3 : 0.0;
            C.rc4 = M->d.order[2] > 0 ? rc4 : 0.0;
            if (C.rc4 > C.rc3 && C.rc3 > 0.0)
                throw EngineError("4-body cutoff above the 3-body cutoff is not supported");
            // cell grid over the cell's fractional coordinates
            double a[3], bb[3], cv[3];
            col(C.hn, 0, a); col(C.hn, 1, bb); col(C.hn, 2, cv);
            auto cross = [](const double *x, const double *y, double *z) {
                z[0] = x[1] * y[2] - x[2] * y[1];
                z[1] = x[2] * y[0] - x[0] * y[2];
                z[2] = x[0] * y[1] - x[1] * y[0];
            };
            double bc[3], ca2[3], ab[3];
            cross(bb, cv, bc); cross(cv, a, ca2); cross(a, bb, ab);
            const double V = fabs(dot3(a, bc));
            const double w[3] = {V / mag3(bc[0], bc[1], bc[2]), V / mag3(ca2[0], ca2[1], ca2[2]),
                                 V / mag3(ab[0], ab[1], ab[2])};
            long cells = 1;
            for (int k = 0; k < 3; k++) {
                int m = (int)floor(w[k] / (C.search * (1.0 + 1e-6)));
                if (m < 1) m = 1;
                if (m > 256) m = 256;
                const double delta = C.search / w[k] * (1.0 + 1e-8) + 1e-12;
                int reach = (int)ceil(delta * m);
                if (reach < 1) reach = 1;
                C.m[k] = m;
                C.reach[k] = reach;
                cells *= m;
            }
            C.natoms_in = nin;
            C.natoms = nin * (C.n_rep + 1) * (C.n_rep + 1) * (C.n_rep + 1);
            C.atom_off = (int)N;
            C.atom_in_off = (int)Nin;
            C.cell_off = (int)ncell;
            C.model = mi;
            e->cfg_atom_off_h[c] = (int)N;
            e->cfg_in_off_h[c] = (int)Nin;
            N += C.natoms;
            Nin += nin;
            ncell += cells;
        }
        e->cfg_atom_off_h[ncfg] = (int)N;
        e->cfg_in_off_h[ncfg] = (int)Nin;
        if (N >= (1L << 30)) throw EngineError("batch too large");
        e->N = (int)N;
        e->Nin = (int)Nin;
        e->ncell = (int)ncell;
        // group bases per model
        e->g3b.assign(nmodels, 0);
        e->g4b.assign(nmodels, 0);
        int G3 = 0, G4 = 0;
        std::vector<chx_enum_model> emh(nmodels);
        std::vector<chx_red_model> rmh(nmodels);
        std::vector<chx_model_dev> mdh(nmodels);
        for (int mi = 0; mi < nmodels; mi++) {
            const chx_model *M = models[mi];
            chx_enum_model &q = emh[mi];
            q.nt = M->d.nt;
            q.has3 = M->d.order[1] > 0 && M->d.ntrip > 0;
            q.has4 = M->d.order[2] > 0 && M->d.nquad > 0;
            q.ntrip = M->d.ntrip;
            q.nquad = M->d.nquad;
            q.g3 = G3;
            q.g4 = G4;
            q.tripmap = M->dev.tripmap;
            q.quadmap = M->dev.quadmap;
            e->g3b[mi] = G3;
            e->g4b[mi] = G4;
            if (q.has3) G3 += M->d.ntrip + 1;
            if (q.has4) G4 += M->d.nquad + 1;
            rmh[mi] = chx_red_model{q.g3, q.has3 ? M->d.ntrip : 0, q.g4, q.has4 ? M->d.nquad : 0, M->dev.eoff};
            mdh[mi] = M->dev;
        }
        e->G3 = G3;
        e->G4 = G4;
        // ---------------- uploads ------------------------------------------
        (void)hipEventRecord(e->ev[0], st);
        e->cfgs.ensure(ncfg);
        e->cfg_atom_off.ensure(ncfg + 1);
        e->cfg_in_off.ensure(ncfg + 1);
        e->em.ensure(nmodels);
        e->rm.ensure(nmodels);
        e->mdev.ensure(nmodels);
        HIPCHK(hipMemcpyAsync(e->cfgs.p, e->cfg_h.data(), ncfg * sizeof(chx_cfg), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(e->cfg_atom_off.p, e->cfg_atom_off_h.data(), (ncfg + 1) * sizeof(int),
                              hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(e->cfg_in_off.p, e->cfg_in_off_h.data(), (ncfg + 1) * sizeof(int),
                              hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(e->em.p, emh.data(), nmodels * sizeof(chx_enum_model), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(e->rm.p, rmh.data(), nmodels * sizeof(chx_red_model), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(e->mdev.p, mdh.data(), nmodels * sizeof(chx_model_dev), hipMemcpyHostToDevice, st));
        const double *pos_d;
        const int *typ_d;
        if (b->pos_dev) {
            pos_d = b->pos;
            typ_d = b->types;
        } else {
            e->pos.ensure(3 * Nin);
            e->types.ensure(Nin);
            HIPCHK(hipMemcpyAsync(e->pos.p, b->pos, 3 * Nin * sizeof(double), hipMemcpyHostToDevice, st));
            HIPCHK(hipMemcpyAsync(e->types.p, b->types, Nin * sizeof(int), hipMemcpyHostToDevice, st));
            pos_d = e->pos.p;
            typ_d = e->types.p;
        }
        // ---------------- atoms and cells ----------------------------------
        e->as.ensure(N);
        e->an.ensure(N);
        e->cell_cnt.ensure(ncell + 1);
        e->cell_start.ensure(ncell + 1);
        e->cell_cur.ensure(ncell + 1);
        e->cell_atoms.ensure(N);
        e->zero_i(e->cell_cnt.p, ncell + 1);
        e->zero_i(e->cell_cur.p, ncell + 1);
        hipLaunchKernelGGL(k_prep, dim3(nblk(N, 256)), dim3(256), 0, st, (int)N, ncfg, e->cfg_atom_off.p, e->cfgs.p,
                           pos_d, typ_d, e->as.p, e->an.p, e->cell_cnt.p);
        e->scan(e->cell_cnt.p, e->cell_start.p, ncell);
        hipLaunchKernelGGL(k_cell_fill, dim3(nblk(N, 256)), dim3(256), 0, st, (int)N, e->an.p, e->cell_start.p,
                           e->cell_cur.p, e->cell_atoms.p);
        hipLaunchKernelGGL(k_segment_sort, dim3(nblk(ncell, 256)), dim3(256), 0, st, (int)ncell, e->cell_start.p,
                           e->cell_atoms.p);
        // ---------------- neighbour lists ----------------------------------
        e->nbr_cnt.ensure(N + 1);
        e->nbr_off.ensure(N + 1);
        hipLaunchKernelGGL((k_nbr<0>), dim3(nblk(N, 128)), dim3(128), 0, st, (int)N, e->cfgs.p, e->as.p, e->an.p,
                           e->cell_start.p, e->cell_atoms.p, e->nbr_cnt.p, (const int *)nullptr,
                           (chx_entry_s *)nullptr, (chx_entry_n *)nullptr, (int *)nullptr);
        e->scan(e->nbr_cnt.p, e->nbr_off.p, N);
        const long nent = e->read_int(e->nbr_off.p + N);
        e->nent = nent;
        e->es.ensure(nent);
        e->en.ensure(nent);
        e->e_owner.ensure(nent);
        hipLaunchKernelGGL((k_nbr<1>), dim3(nblk(N, 128)), dim3(128), 0, st, (int)N, e->cfgs.p, e->as.p, e->an.p,
                           e->cell_start.p, e->cell_atoms.p, e->nbr_cnt.p, e->nbr_off.p, e->es.p, e->en.p,
                           e->e_owner.p);
        // canonical (parent, shift) order of every owner's images
        e->es2.ensure(nent);
        e->en2.ensure(nent);
        e->e_in3.ensure(nent);
        e->status.ensure(4);
        e->zero_i(e->status.p, 4);
        hipLaunchKernelGGL(k_nbr_sort, dim3((unsigned)N), dim3(64), 0, st, (int)N, e->nbr_off.p, e->es.p, e->en.p,
                           e->es2.p, e->en2.p, e->status.p);
        std::swap(e->es.p, e->es2.p);
        std::swap(e->es.cap, e->es2.cap);
        std::swap(e->en.p, e->en2.p);
        std::swap(e->en.cap, e->en2.cap);
        {
            std::vector<int> s0(2);
            HIPCHK(hipMemcpyAsync(s0.data(), e->status.p, 2 * sizeof(int), hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
            if (s0[0])
                throw EngineError("an atom has " + std::to_string(s0[1]) + " neighbour images within the search "
                                  "cutoff; at most " + std::to_string(CHX_MAX_NB) + " are supported");
        }
        // ---------------- cluster enumeration ------------------------------
        const long n3c = (long)G3 * N + 1, n4c = (long)G4 * N + 1;
        e->cnt3.ensure(n3c);
        e->off3.ensure(n3c + 1);
        e->cnt4.ensure(n4c);
        e->off4.ensure(n4c + 1);
        e->zero_i(e->cnt3.p, n3c);
        e->zero_i(e->cnt4.p, n4c);
        hipLaunchKernelGGL((k_enum<0>), dim3((unsigned)N), dim3(64), 0, st, (int)N, e->cfgs.p, e->as.p, e->an.p,
                           e->nbr_off.p, e->es.p, e->en.p, e->em.p, e->cnt3.p, e->cnt4.p, (const int *)nullptr,
                           (const int *)nullptr, (int2 *)nullptr, (int2 *)nullptr, (unsigned char *)nullptr,
                           e->status.p);
        e->scan(e->cnt3.p, e->off3.p, (long)G3 * N);
        e->scan(e->cnt4.p, e->off4.p, (long)G4 * N);
        e->gtmp.ensure(G3 + G4 + 2);
        hipLaunchKernelGGL(k_gather_strided, dim3(1), dim3(128), 0, st, e->off3.p, (long)N, G3, e->gtmp.p);
        hipLaunchKernelGGL(k_gather_strided, dim3(1), dim3(128), 0, st, e->off4.p, (long)N, G4, e->gtmp.p + G3 + 1);
        std::vector<int> gh(G3 + G4 + 2), sth(4);
        HIPCHK(hipMemcpyAsync(gh.data(), e->gtmp.p, (G3 + G4 + 2) * sizeof(int), hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(sth.data(), e->status.p, 4 * sizeof(int), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        if (sth[0])
            throw EngineError("an atom has more than " + std::to_string(CHX_MAX_N3) + " images within the 3-body "
                              "cutoff or more than " + std::to_string(CHX_MAX_N4) + " within the 4-body cutoff");
        e->goff3.assign(gh.begin(), gh.begin() + G3 + 1);
        e->goff4.assign(gh.begin() + G3 + 1, gh.end());
        const long n3 = e->goff3[G3], n4 = e->goff4[G4];
        e->list3.ensure(n3);
        e->list4.ensure(n4);
        hipLaunchKernelGGL((k_enum<1>), dim3((unsigned)N), dim3(64), 0, st, (int)N, e->cfgs.p, e->as.p, e->an.p,
                           e->nbr_off.p, e->es.p, e->en.p, e->em.p, e->cnt3.p, e->cnt4.p, e->off3.p, e->off4.p,
                           e->list3.p, e->list4.p, e->e_in3.p, e->status.p);
        // ---------------- reverse lists ------------------------------------
        e->rcnt.ensure(N + 1);
        e->rstart.ensure(N + 1);
        e->rcur.ensure(N + 1);
        e->rev.ensure(nent);
        e->zero_i(e->rcnt.p, N + 1);
        e->zero_i(e->rcur.p, N + 1);
        hipLaunchKernelGGL(k_rev_count, dim3(nblk(nent, 256)), dim3(256), 0, st, nent, e->es.p, e->rcnt.p);
        e->scan(e->rcnt.p, e->rstart.p, N);
        hipLaunchKernelGGL(k_rev_fill, dim3(nblk(nent, 256)), dim3(256), 0, st, nent, e->es.p, e->rstart.p,
                           e->rcur.p, e->rev.p);
        hipLaunchKernelGGL(k_segment_sort, dim3(nblk(N, 256)), dim3(256), 0, st, (int)N, e->rstart.p, e->rev.p);
        (void)hipEventRecord(e->ev[1], st);
        // ---------------- evaluation ---------------------------------------
        const bool mixed = b->precision == CHX_MIXED;
        e->cfg_stat.ensure(4 * ncfg);
        e->zero_i(e->cfg_stat.p, 4 * ncfg);
        e->rec2E.ensure(nent);
        e->rec2F.ensure(nent);
        hipLaunchKernelGGL(k_eval2, dim3(nblk(nent, 256)), dim3(256), 0, st, nent, e->e_owner.p, e->cfgs.p, e->as.p,
                           e->es.p, e->mdev.p, e->rec2E.p, e->rec2F.p, e->cfg_stat.p);
        e->rec3E.ensure(n3);
        e->rec3F.ensure(3 * n3);
        e->rec3M.ensure(n3);
        e->rec4E.ensure(n4);
        e->rec4F.ensure(6 * n4);
        e->rec4M.ensure(n4);
        (void)hipEventRecord(e->ev[2], st);
        for (int mi = 0; mi < nmodels; mi++) {
            const chx_model *M = models[mi];
            if (!emh[mi].has3) continue;
            for (int t = 0; t < M->d.ntrip; t++) {
                const int g = e->g3b[mi] + t;
                chx_type_args3 ta;
                ta.nt = M->d.nt;
                ta.tersoff = M->d.tersoff;
                ta.fcut_var = M->d.fcut_var;
                ta.perm = M->dev.trip_perm;
                for (int q = 0; q < 3; q++) ta.s[q] = M->t3[t].s[q];
                launch_eval3(M->d.D3, mixed, e->goff3[g], e->goff3[g + 1], st, e->list3.p, e->cfgs.p, e->as.p,
                             e->nbr_off.p, e->es.p, ta, mixed ? (const void *)M->t3[t].Cf : (const void *)M->t3[t].C,
                             e->rec3E.p, e->rec3F.p, e->rec3M.p, e->cfg_stat.p);
            }
        }
        (void)hipEventRecord(e->ev[3], st);
        for (int mi = 0; mi < nmodels; mi++) {
            const chx_model *M = models[mi];
            if (!emh[mi].has4) continue;
            for (int t = 0; t < M->d.nquad; t++) {
                const int g = e->g4b[mi] + t;
                chx_type_args4 ta;
                ta.nt = M->d.nt;
                ta.tersoff = M->d.tersoff;
                ta.fcut_var = M->d.fcut_var;
                ta.perm = M->dev.quad_perm;
                for (int q = 0; q < 6; q++) ta.s[q] = M->t4[t].s[q];
                launch_eval4(M->d.D4, mixed, e->goff4[g], e->goff4[g + 1], st, e->list4.p, e->cfgs.p, e->as.p,
                             e->nbr_off.p, e->es.p, ta, mixed ? (const void *)M->t4[t].Cf : (const void *)M->t4[t].C,
                             e->rec4E.p, e->rec4F.p, e->rec4M.p, e->cfg_stat.p);
            }
        }
        (void)hipEventRecord(e->ev[4], st);
        // ---------------- reduction and gathers ----------------------------
        e->Fe.ensure(3 * nent);
        e->Fself.ensure(3 * N);
        e->Eown.ensure(N);
        e->Sown.ensure(6 * N);
        e->Frep.ensure(3 * N);
        e->Fout.ensure(3 * Nin);
        e->Ecfg.ensure(ncfg);
        e->Vcfg.ensure(6 * ncfg);
        chx_red_args ra;
        ra.models = e->rm.p;
        ra.rec2E = e->rec2E.p;
        ra.rec2F = e->rec2F.p;
        ra.off3 = G3 ? e->off3.p : nullptr;
        ra.rec3E = e->rec3E.p;
        ra.rec3F = e->rec3F.p;
        ra.rec3M = e->rec3M.p;
        ra.off4 = G4 ? e->off4.p : nullptr;
        ra.rec4E = e->rec4E.p;
        ra.rec4F = e->rec4F.p;
        ra.rec4M = e->rec4M.p;
        hipLaunchKernelGGL(k_reduce, dim3((unsigned)N), dim3(64), 0, st, (int)N, e->cfgs.p, e->as.p, e->nbr_off.p,
                           e->es.p, e->e_in3.p, ra, e->Fe.p, e->Fself.p, e->Eown.p, e->Sown.p);
        hipLaunchKernelGGL(k_gather_rep, dim3(nblk(N, 256)), dim3(256), 0, st, (int)N, e->rstart.p, e->rev.p,
                           e->Fe.p, e->Fself.p, e->Frep.p);
        double *Fout = o->out_dev ? o->forces : e->Fout.p;
        double *Ecfg = o->out_dev ? o->energy : e->Ecfg.p;
        double *Vcfg = o->out_dev ? o->virial : e->Vcfg.p;
        hipLaunchKernelGGL(k_gather_in, dim3(nblk(Nin, 256)), dim3(256), 0, st, (int)Nin, ncfg, e->cfg_in_off.p,
                           e->cfgs.p, e->Frep.p, Fout);
        hipLaunchKernelGGL(k_cfg_sum, dim3(ncfg), dim3(256), 0, st, e->cfgs.p, e->Eown.p, e->Sown.p, Ecfg, Vcfg);
        (void)hipEventRecord(e->ev[5], st);
        std::vector<int> stat(4 * ncfg);
        HIPCHK(hipMemcpyAsync(stat.data(), e->cfg_stat.p, 4 * ncfg * sizeof(int), hipMemcpyDeviceToHost, st));
        if (!o->out_dev) {
            HIPCHK(hipMemcpyAsync(o->forces, Fout, 3 * Nin * sizeof(double), hipMemcpyDeviceToHost, st));
            HIPCHK(hipMemcpyAsync(o->energy, Ecfg, ncfg * sizeof(double), hipMemcpyDeviceToHost, st));
            HIPCHK(hipMemcpyAsync(o->virial, Vcfg, 6 * ncfg * sizeof(double), hipMemcpyDeviceToHost, st));
        }
        HIPCHK(hipStreamSynchronize(st));
        HIPCHK(hipGetLastError());
        float ms = 0;
        (void)hipEventElapsedTime(&ms, e->ev[0], e->ev[5]);
        e->timing[0] = ms;
        (void)hipEventElapsedTime(&ms, e->ev[3], e->ev[4]);
        e->timing[1] = ms;
        (void)hipEventElapsedTime(&ms, e->ev[2], e->ev[3]);
        e->timing[2] = ms;
        (void)hipEventElapsedTime(&ms, e->ev[0], e->ev[1]);
        e->timing[3] = ms;
        for (int c = 0; c < ncfg; c++) {
            if (o->volume) o->volume[c] = e->cfg_h[c].vol;
            if (o->n_rep) o->n_rep[c] = e->cfg_h[c].n_rep;
            if (o->badness) o->badness[c] = stat[4 * c];
            if (o->counts) {
                o->counts[4 * c + 0] = stat[4 * c + 1];
                o->counts[4 * c + 1] = stat[4 * c + 2];
                o->counts[4 * c + 2] = 0;
                o->counts[4 * c + 3] = 0;
            }
        }
        e->err.clear();
        return 0;
    } catch (const EngineError &x) {
        e->err = x.msg;
        (void)hipStreamSynchronize(e->st);
        return 1;
    }
}

int chx_export_lists(chx_engine *e, int cfg, long *n2, int **pairs, long *n3, int **trips, long *n4, int **quads)
{
    try {
        HIPCHK(hipSetDevice(e->device));
        if (cfg < 0 || cfg >= e->ncfg) throw EngineError("configuration index out of range");
        const chx_cfg &C = e->cfg_h[cfg];
        const long N = e->N;
        std::vector<int> off(C.natoms + 1);
        HIPCHK(hipMemcpy(off.data(), e->nbr_off.p + C.atom_off, (C.natoms + 1) * sizeof(int), hipMemcpyDeviceToHost));
        const long e0 = off[0], ne = off[C.natoms] - off[0];
        std::vector<chx_entry_s> es(ne);
        std::vector<chx_entry_n> en(ne);
        if (ne) {
            HIPCHK(hipMemcpy(es.data(), e->es.p + e0, ne * sizeof(chx_entry_s), hipMemcpyDeviceToHost));
            HIPCHK(hipMemcpy(en.data(), e->en.p + e0, ne * sizeof(chx_entry_n), hipMemcpyDeviceToHost));
        }
        auto key = [&](int owner_local, int code, int *o) {
            if (code == 0) { o[0] = owner_local; o[1] = o[2] = o[3] = 0; return; }
            const long idx = off[owner_local] - e0 + code - 1;
            o[0] = es[idx].parent - C.atom_off;
            const int s = en[idx].shift;
            o[1] = (s & 255) - 128;
            o[2] = ((s >> 8) & 255) - 128;
            o[3] = ((s >> 16) & 255) - 128;
        };
        *n2 = ne;
        *pairs = (int *)malloc(sizeof(int) * 5 * std::max<long>(ne, 1));
        for (int i = 0; i < C.natoms; i++)
            for (int s = 1; s <= off[i + 1] - off[i]; s++) {
                int *o = *pairs + 5 * (off[i] - e0 + s - 1);
                o[0] = i;
                key(i, s, o + 1);
            }
        for (int body = 3; body <= 4; body++) {
            const int G = body == 3 ? e->G3 : e->G4;
            const chx_model *M = e->models[C.model];
            const bool has = body == 3 ? (M->d.order[1] > 0 && M->d.ntrip > 0) : (M->d.order[2] > 0 && M->d.nquad > 0);
            std::vector<int> out;
            if (has && G > 0) {
                const int gb = body == 3 ? e->g3b[C.model] : e->g4b[C.model];
                const int ng = (body == 3 ? M->d.ntrip : M->d.nquad) + 1;
                const int *offv = body == 3 ? e->off3.p : e->off4.p;
                const int2 *lst = body == 3 ? e->list3.p : e->list4.p;
                for (int g = gb; g < gb + ng; g++) {
                    int r[2];
                    HIPCHK(hipMemcpy(&r[0], offv + (long)g * N + C.atom_off, sizeof(int), hipMemcpyDeviceToHost));
                    HIPCHK(hipMemcpy(&r[1], offv + (long)g * N + C.atom_off + C.natoms, sizeof(int),
                                     hipMemcpyDeviceToHost));
                    std::vector<int2> recs(r[1] - r[0]);
                    if (!recs.empty())
                        HIPCHK(hipMemcpy(recs.data(), lst + r[0], recs.size() * sizeof(int2), hipMemcpyDeviceToHost));
                    for (const int2 &q : recs) {
                        const int il = q.x - C.atom_off;
                        out.push_back(il);
                        for (int m = 0; m < body - 1; m++) {
                            int k[4];
                            key(il, (q.y >> (10 * m)) & 1023, k);
                            out.insert(out.end(), k, k + 4);
                        }
                    }
                }
            }
            const int w = body == 3 ? 9 : 13;
            int *buf = (int *)malloc(sizeof(int) * std::max<size_t>(out.size(), 1));
            if (!out.empty()) memcpy(buf, out.data(), out.size() * sizeof(int));
            if (body == 3) { *n3 = (long)out.size() / w; *trips = buf; }
            else { *n4 = (long)out.size() / w; *quads = buf; }
        }
        return 0;
    } catch (const EngineError &x) {
        e->err = x.msg;
        return 1;
    }
}

int chx_eval_clusters(chx_engine *e, chx_model *m, int nb, int n, const double *dist, const double *dr,
                      const int *types, double *energy, double *force, double *stress, double *last, int *flags)
{
    try {
        HIPCHK(hipSetDevice(e->device));
        if (nb < 2 || nb > 4) throw EngineError("cluster size must be 2, 3 or 4");
        if (n <= 0) return 0;
        if ((nb == 3 && (m->d.order[1] == 0 || !m->dev.C3)) || (nb == 4 && (m->d.order[2] == 0 || !m->dev.C4)))
            throw EngineError("model has no terms of this body order");
        const int P = nb * (nb - 1) / 2;
        hipStream_t st = e->st;
        e->cl_in.ensure((size_t)n * P * 4);
        e->cl_out.ensure((size_t)n * (1 + 3 * nb + 6 + P));
        e->cl_types.ensure((size_t)n * nb);
        e->cl_flags.ensure(n);
        double *d_dist = e->cl_in.p, *d_dr = e->cl_in.p + (size_t)n * P;
        double *oE = e->cl_out.p, *oF = oE + n, *oS = oF + (size_t)n * nb * 3, *oL = oS + (size_t)n * 6;
        HIPCHK(hipMemcpyAsync(d_dist, dist, sizeof(double) * n * P, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(d_dr, dr, sizeof(double) * n * P * 3, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(e->cl_types.p, types, sizeof(int) * n * nb, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(oL, last, sizeof(double) * n * P, hipMemcpyHostToDevice, st));
        launch_clusters(nb, nb == 3 ? m->d.D3 : (nb == 4 ? m->d.D4 : 2), n, st, d_dist, d_dr, e->cl_types.p, m->dev,
                        oE, oF, oS, oL, e->cl_flags.p);
        HIPCHK(hipMemcpyAsync(energy, oE, sizeof(double) * n, hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(force, oF, sizeof(double) * n * nb * 3, hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(stress, oS, sizeof(double) * n * 6, hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(last, oL, sizeof(double) * n * P, hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(flags, e->cl_flags.p, sizeof(int) * n, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        HIPCHK(hipGetLastError());
        return 0;
    } catch (const EngineError &x) {
        e->err = x.msg;
        return 1;
    }
}

}  // extern "C"
