// HIP engine: device-resident parameter tables, per-configuration frame
// set-up and the evaluation pipeline behind the internal C ABI (chx_engine.h).
//
// Pipeline of one batch (one stream; a captured plan runs the 2-body and the
// 4-body evaluation on a second one):
//   k_prep        replicate, wrap, evaluation and neighbour frames, cell index
//   cells         count, scan, fill
//   k_nbr         image neighbour lists sorted by (parent, shift) per owner
//   k_enum        3-/4-body lists in segments of one (model, cluster type),
//                 aligned so evaluation tiles share coefficients
//   k_eval2/3/4   cluster energies, forces and virials, accumulated exactly
//                 in fixed point per atom and per owner
//   k_fin         forces per input atom, energy and virial per configuration
// Buffers are sized by capacity: a run that finds a capacity short sets a
// status bit, skips the rest and is repeated with the size it reported, so a
// run needs one host synchronisation (at the end).
// Host frame set-up follows serial_chimes_interface.cpp:150-476 operation by
// operation (this file is built with -ffp-contract=off).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "chx_engine.h"
#include "chx_eval.h"
#include "chx_kernels.h"
#include "chx_fp.h"
#include "chx_lmp.h"
#include "tables.h"

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

// Persistent host workers: run(f) calls f(k) once for each k in [0, size()),
// k = 0 on the calling thread, and returns when all have returned (f must
// not throw).
class WorkerPool {
public:
    explicit WorkerPool(int n) : n_(n)
    {
        for (int k = 1; k < n; k++) th_.emplace_back([this, k] { loop(k); });
    }
    ~WorkerPool()
    {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        for (std::thread &t : th_) t.join();
    }
    int size() const { return n_; }
    void run(const std::function<void(int)> &f)
    {
        {
            std::lock_guard<std::mutex> lk(mu_);
            job_ = &f;
            left_ = n_ - 1;
            gen_++;
        }
        cv_.notify_all();
        f(0);
        std::unique_lock<std::mutex> lk(mu_);
        done_.wait(lk, [this] { return left_ == 0; });
        job_ = nullptr;
    }

private:
    void loop(int k)
    {
        long seen = 0;
        for (;;) {
            const std::function<void(int)> *f;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
                if (stop_) return;
                seen = gen_;
                f = job_;
            }
            (*f)(k);
            std::lock_guard<std::mutex> lk(mu_);
            if (--left_ == 0) done_.notify_one();
        }
    }
    int n_;
    std::vector<std::thread> th_;
    std::mutex mu_;
    std::condition_variable cv_, done_;
    const std::function<void(int)> *job_ = nullptr;
    int left_ = 0;
    long gen_ = 0;
    bool stop_ = false;
};

template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t cap = 0;
    size_t need = 0;  // elements the last ensure() asked for
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p(o.p), cap(o.cap), need(o.need) { o.p = nullptr; o.cap = 0; o.need = 0; }
    ~DevBuf() { release(); }
    void ensure(size_t n)
    {
        need = n;
        if (n <= cap && p) return;
        if (p) (void)hipFree(p);
        p = nullptr;
        // 25% slack, and at least double a buffer that grows: a stream of
        // growing batches (a server's first minutes) reallocates a few times,
        // not once per new largest batch (each free and allocation waits for
        // the device)
        size_t c = std::max<size_t>(n, 16);
        c += c / 4;
        c = std::max(c, 2 * cap);
        HIPCHK(hipMalloc((void **)&p, c * sizeof(T)));
        cap = c;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        need = 0;
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
    const double *C = nullptr;   // device, FP64 dense coefficients
    const float *Cf = nullptr;   // device, FP32 copy for the mixed path
    const double *Cm = nullptr;  // device, matrix image for the matrix-core kernels
    const float *Cmf = nullptr;
    const unsigned short *Cmb = nullptr;  // device, BF16 images (rounded, residual) transposed [column][row to 32]
};

// Round-to-nearest-even BF16 bits of x.
unsigned short bf16_of(double x)
{
    const float f = (float)x;
    unsigned u;
    memcpy(&u, &f, 4);
    return (unsigned short)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

// BF16 copies of a row-major KP x MP matrix image, transposed to [MP][K32]
// with K32 = KP rounded up to 32 (zero padding): the rounded values, then the
// rounded residuals (value - high part, in FP32).
std::vector<unsigned short> bf16_transposed(const std::vector<double> &m, int KP, int MP)
{
    const int K32 = (KP + 31) / 32 * 32;
    const size_t n = (size_t)MP * K32;
    std::vector<unsigned short> t(2 * n, 0);
    for (int k = 0; k < KP; k++)
        for (int c = 0; c < MP; c++) {
            const float f = (float)m[(size_t)k * MP + c];
            const unsigned short h = bf16_of(f);
            unsigned u = (unsigned)h << 16;
            float hf;
            memcpy(&hf, &u, 4);
            t[(size_t)c * K32 + k] = h;
            t[n + (size_t)c * K32 + k] = bf16_of(f - hf);
        }
    return t;
}

// Matrix image of one 4-body coefficient tensor for the matrix-core kernel:
// Cm[k][m] with k = (a*D + b)*4 + c (pairs 01, 02, 03) and m = (d*D + e)*4 + f
// (pairs 12, 13, 23), innermost indices padded to 4, m padded to a multiple
// of 16, zero padding.
std::vector<double> quad_matrix_image(const double *C, int D)
{
    const int KP = D * D * 4, MP = (D * D * 4 + 15) / 16 * 16;
    std::vector<double> m((size_t)KP * MP, 0.0);
    for (int a = 0; a < D; a++)
        for (int b = 0; b < D; b++)
            for (int c = 0; c < D; c++)
                for (int d = 0; d < D; d++)
                    for (int e = 0; e < D; e++)
                        for (int f = 0; f < D; f++) {
                            const size_t src = (((((size_t)a * D + b) * D + c) * D + d) * D + e) * D + f;
                            const size_t k = (size_t)(a * D + b) * 4 + c, col = (size_t)(d * D + e) * 4 + f;
                            m[k * MP + col] = C[src];
                        }
    return m;
}

// A pair slot from (rin, rout, lambda, xavg, xdiff) and the model's Tersoff
// ramp fraction, with the reciprocals the kernels multiply by.
chx_slot make_slot(const double *q, double fcut_var)
{
    chx_slot s{q[0], q[1], q[2], q[3], q[4], 0, 0, 0, 0};
    s.ilam = 1.0 / s.lam;
    s.ixdiff = 1.0 / s.xdiff;
    s.irout = 1.0 / s.rout;
    // Tersoff ramp: from thr = rout - fcut_var * rout up to rout
    const double thr = s.rout - fcut_var * s.rout;
    const double width = s.rout - thr;
    s.iramp = width != 0.0 ? 1.0 / width : 0.0;
    return s;
}

// Matrix image of one 3-body coefficient tensor: Cm3[a][16 b + c], a padded to
// a multiple of 4, c to 16, zero padding.
std::vector<double> trip_matrix_image(const double *C, int D)
{
    const int KP = (D + 3) / 4 * 4, MP = 16 * D;
    std::vector<double> m((size_t)KP * MP, 0.0);
    for (int a = 0; a < D; a++)
        for (int b = 0; b < D; b++)
            for (int c = 0; c < D; c++) m[(size_t)a * MP + 16 * b + c] = C[((size_t)a * D + b) * D + c];
    return m;
}

}  // namespace

// Scratch of the allocation-free explicit-cluster entry point: inputs and
// outputs of at most CHX_STATIC_MAX_CLUSTERS clusters.  Each device holds its
// own copy for the life of the process; host code uses the address that
// hipGetSymbolAddress returns for the current device (scratch_init, once per
// device at engine creation), never the array's name.
constexpr size_t CHX_SCRATCH_BYTES = (size_t)CHX_STATIC_MAX_CLUSTERS * 416 + 8 * 256;
static __device__ __align__(256) char scratch[CHX_SCRATCH_BYTES];

namespace {
constexpr int CHX_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void *g_scratch[CHX_MAX_DEVICES] = {};
hipError_t g_scratch_rc[CHX_MAX_DEVICES] = {};

// Resolves the scratch array's address on `device` (which must be current).
void scratch_init(int device)
{
    if (device < 0 || device >= CHX_MAX_DEVICES) return;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch[device]) return;
    void *p = nullptr;
    const hipError_t rc = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_scratch_rc[device] = rc;
    if (rc == hipSuccess) g_scratch[device] = p;
}

void *scratch_ptr(int device)
{
    if (device < 0 || device >= CHX_MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return g_scratch[device];
}
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

// Page-locked host buffer (hipHostMalloc), grown on demand.
struct PinnedBuf {
    char *p = nullptr;
    size_t cap = 0;
    PinnedBuf() = default;
    PinnedBuf(const PinnedBuf &) = delete;
    PinnedBuf &operator=(const PinnedBuf &) = delete;
    ~PinnedBuf()
    {
        if (p) (void)hipHostFree(p);
    }
    void ensure(size_t n)
    {
        if (n <= cap && p) return;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        const size_t c = std::max(std::max<size_t>(n + n / 4, 4096), 2 * cap);  // as DevBuf::ensure
        HIPCHK(hipHostMalloc((void **)&p, c, hipHostMallocDefault));
        cap = c;
    }
};

struct chx_engine {
    int device = 0;
    hipStream_t st = nullptr;
    hipEvent_t ev[6];
    std::string err;
    double timing[4] = {0, 0, 0, 0};
    long long counts[3] = {0, 0, 0};
    // batch state
    std::vector<chx_cfg> cfg_h;
    std::vector<int> cfg_atom_off_h, cfg_in_off_h;
    std::vector<int> cfg_item_off_h;   // energy/virial reduction items (owner chunks) per configuration
    std::vector<long> cfg_cells_h;     // cells of each configuration's grid
    int nitems = 0;
    int N = 0, Nin = 0, ncfg = 0, ncell = 0;
    int G3 = 0, G4 = 0;
    std::vector<int> g3b, g4b;
    std::vector<chx_model *> models;
    std::vector<chx_enum_model> emh;
    std::vector<chx_model_dev> mdh;
    std::vector<chx_group> g3h, g4h;
    std::vector<int> D3s, D4s;
    int ngl_max = 1;
    int prec = CHX_FP64;
    // kernel variants of the 3-/4-body evaluation (chosen by the plan autotuner)
    // untuned defaults: the vector-unit 3-body kernel and the 64-cluster
    // matrix-core 4-body kernel (the fastest on the C1-C3 workloads)
    int var3 = CHX_VAR_VECTOR, var4 = CHX_VAR_MFMA4;
    bool capturing = false;  // inside a graph capture: no event records
    // host workers of the frame set-up of large batches (created on first use)
    std::unique_ptr<WorkerPool> pool;
    // second stream of a captured plan, and its fork / join events
    hipStream_t st2 = nullptr;
    hipEvent_t fork_ev[2] = {nullptr, nullptr}, join_ev = nullptr;
    // buffer capacities (entries, list slots); grow on demand
    long l3_cap = 0, l4_cap = 0;
    int nb_cap = 0;  // neighbour images per owner slab
    int nb_cap_run = 0;  // the slab stride of the last run (capacities may change after it)
    int rcap = 0;    // cluster records per owner slab
    long cap_key[4] = {-1, -1, -1, -1};  // batch shape the capacities were measured on
    bool big_on = false;  // launch the crowded-owner enumeration (set once an owner needed it)
    // one host-to-device and one device-to-host copy per run
    PinnedBuf in_h, out_h;
    DevBuf<char> in_d, out_d;
    size_t in_bytes = 0, out_bytes = 0;
    size_t o_stat = 0, o_cstat = 0, o_E = 0, o_V = 0, o_F = 0;  // offsets in the output blob
    // device pointers into the input blob
    chx_cfg *cfgs = nullptr;
    int *cfg_atom_off = nullptr, *cfg_in_off = nullptr, *cfg_item_off = nullptr;
    chx_enum_model *em = nullptr;
    chx_model_dev *mdev = nullptr;
    chx_group *grp3 = nullptr, *grp4 = nullptr;
    chx_clear_range *clear = nullptr;
    int nclear = 0;
    double *pos_in = nullptr;
    int *types_in = nullptr;
    // device pointers into the output blob
    int *status = nullptr, *cfg_stat = nullptr;
    double *E_out = nullptr, *V_out = nullptr, *F_out = nullptr;
    // work buffers
    DevBuf<chx_atom_s> as;
    DevBuf<chx_atom_n> an;
    DevBuf<int> cell_cnt, cell_start, cell_cur, cell_atoms;
    DevBuf<int> nbr_cnt, nbr_off, big;
    DevBuf<int> in_cfg, item_cfg;  // configuration of each input atom and of each totals item
    DevBuf<chx_entry_s> es;
    DevBuf<chx_entry_n> en;
    DevBuf<int> cnt34, off34;   // per (group, owner) counts and their scan: 3-body groups, then 4-body
    DevBuf<int> gb3, ge3, gb4, ge4;
    DevBuf<int2> list3, list4;
    DevBuf<int2> rec;           // per owner slab of rcap cluster records (enumeration -> placement)
    DevBuf<int> nrec;
    DevBuf<unsigned long long> Facc;
    DevBuf<unsigned long long> Cacc;  // [ncfg][CHX_CACC_SLOTS][CHX_CACC_STRIDE] energy and virial per configuration
    DevBuf<long long> part;
    DevBuf<int> done;  // per configuration: reduction items finished
    std::vector<DevBuf<int>> scan_tmp;
    // per-cluster API buffers
    DevBuf<double> cl_in, cl_out;
    DevBuf<int> cl_types, cl_flags;

    // Exclusive scan of n ints (out[n] = total).  With gb.off set (to out),
    // the group bases are computed too: in the scan's own launch when it is a
    // single workgroup, else by a k_group_bases launch after it.
    void scan(const int *in, int *out, long n, int level = 0, chx_bases_args gb = chx_bases_args{})
    {
        if (scan_tmp.size() < 32) scan_tmp.resize(32);  // fixed: references below stay valid
        if (level >= 15) throw EngineError("scan too deep");
        // one workgroup for short arrays, the block sums of long ones included
        if (n <= CHX_SCAN1_MAX) {
            hipLaunchKernelGGL(k_scan_one, dim3(1), dim3(1024), 0, st, in, out, n, gb);
            return;
        }
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
        if (gb.off) hipLaunchKernelGGL(k_group_bases, dim3(1), dim3(64), 0, st, gb);
    }
    void record(int k)
    {
        if (!capturing) (void)hipEventRecord(ev[k], st);
    }
    // Applies f to every device work buffer.
    template <typename F>
    void each_buf(F f)
    {
        f(as); f(an); f(cell_cnt); f(cell_start); f(cell_cur); f(cell_atoms); f(nbr_cnt); f(nbr_off); f(big);
        f(in_cfg); f(item_cfg);
        f(es); f(en); f(cnt34); f(off34); f(gb3); f(ge3); f(gb4); f(ge4); f(list3); f(list4); f(rec); f(nrec);
        f(Facc); f(Cacc);
        f(part); f(done); f(in_d); f(out_d); f(cl_in); f(cl_out); f(cl_types); f(cl_flags);
        for (auto &b : scan_tmp) f(b);
    }
};

extern "C" {

chx_engine *chx_engine_create(int device)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return nullptr;
    if (device < 0 || device >= n) return nullptr;
    if (hipSetDevice(device) != hipSuccess) return nullptr;
    scratch_init(device);
    chx_engine *e = new chx_engine;
    e->device = device;
    if (hipStreamCreateWithFlags(&e->st, hipStreamNonBlocking) != hipSuccess) { delete e; return nullptr; }
    if (hipStreamCreateWithFlags(&e->st2, hipStreamNonBlocking) != hipSuccess) {
        (void)hipStreamDestroy(e->st);
        delete e;
        return nullptr;
    }
    for (int k = 0; k < 6; k++) (void)hipEventCreate(&e->ev[k]);
    for (hipEvent_t *ev : {&e->fork_ev[0], &e->fork_ev[1], &e->join_ev})
        (void)hipEventCreateWithFlags(ev, hipEventDisableTiming);
    return e;
}

void chx_engine_destroy(chx_engine *e)
{
    if (!e) return;
    (void)hipSetDevice(e->device);
    (void)hipStreamSynchronize(e->st);
    (void)hipStreamSynchronize(e->st2);
    for (int k = 0; k < 6; k++) (void)hipEventDestroy(e->ev[k]);
    for (hipEvent_t ev : {e->fork_ev[0], e->fork_ev[1], e->join_ev}) (void)hipEventDestroy(ev);
    (void)hipStreamDestroy(e->st2);
    (void)hipStreamDestroy(e->st);
    delete e;
}

const char *chx_last_error(chx_engine *e) { return e ? e->err.c_str() : "no engine"; }

void chx_free(void *p) { free(p); }

void chx_last_counts(chx_engine *e, long long *c3)
{
    for (int k = 0; k < 3; k++) c3[k] = e->counts[k];
}

void chx_last_timing(chx_engine *e, double *ms4)
{
    for (int k = 0; k < 4; k++) ms4[k] = e->timing[k];
}

void chx_engine_bytes(chx_engine *e, size_t *held_need)
{
    size_t held = 0, need = 0;
    e->each_buf([&](const auto &b) {
        held += b.cap * sizeof(*b.p);
        need += b.need * sizeof(*b.p);
    });
    held_need[0] = held;
    held_need[1] = need;
}

void chx_engine_trim(chx_engine *e)
{
    (void)hipSetDevice(e->device);
    (void)hipStreamSynchronize(e->st);
    e->each_buf([](auto &b) { b.release(); });
    e->nb_cap = e->rcap = 0;
    e->l3_cap = e->l4_cap = 0;
    for (int k = 0; k < 4; k++) e->cap_key[k] = -1;
}

void chx_engine_reset_capacities(chx_engine *e)
{
    e->nb_cap = e->rcap = 0;
    e->l3_cap = e->l4_cap = 0;
    for (int k = 0; k < 4; k++) e->cap_key[k] = -1;
}

int chx_set_variants(chx_engine *e, int var3, int var4)
{
    if (var3 < -1 || var3 >= CHX_NVAR || var4 < -1 || var4 >= CHX_NVAR) {
        e->err = "unknown kernel variant";
        return 1;
    }
    if (var3 >= 0) e->var3 = var3;
    if (var4 >= 0) e->var4 = var4;
    return 0;
}

void chx_get_variants(chx_engine *e, int *v2)
{
    v2[0] = e->var3;
    v2[1] = e->var4;
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
            s2[p] = make_slot(q, d->fcut_var);
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
                s3[k] = make_slot(q, d->fcut_var);
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
                if (d->D3 <= CHX_MFMA3_MAX_D) {
                    const std::vector<double> img = trip_matrix_image(d->C3 + t * cs, d->D3);
                    const std::vector<float> imgf(img.begin(), img.end());
                    m->t3[t].Cm = m->up(img.data(), img.size());
                    m->t3[t].Cmf = m->up(imgf.data(), imgf.size());
                    const int D = d->D3, KP = (D + 3) / 4 * 4, MP = 16 * D;
                    const std::vector<unsigned short> imgb = bf16_transposed(img, KP, MP);
                    m->t3[t].Cmb = m->up(imgb.data(), imgb.size());
                }
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
                s4[k] = make_slot(q, d->fcut_var);
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
                if (d->D4 <= CHX_MFMA_MAX_D) {
                    const std::vector<double> img = quad_matrix_image(d->C4 + t * cs, d->D4);
                    const std::vector<float> imgf(img.begin(), img.end());
                    m->t4[t].Cm = m->up(img.data(), img.size());
                    m->t4[t].Cmf = m->up(imgf.data(), imgf.size());
                    const int D = d->D4, KP = D * D * 4, MP = (D * D * 4 + 15) / 16 * 16;
                    const std::vector<unsigned short> imgb = bf16_transposed(img, KP, MP);
                    m->t4[t].Cmb = m->up(imgb.data(), imgb.size());
                }
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

}  // extern "C"

namespace {

// Host frame set-up of every configuration (serial_chimes_interface.cpp:150-534).
// Frame of configuration c: everything but its offsets in the batch; returns
// its cell count.  It depends only on the configuration's cell, atom count
// and model.
long frame_of(chx_model *const *models, int nmodels, const chx_batch *b, int c, chx_cfg &C)
{
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
    {
        // the cheap neighbour test skips the round trip of an image
        // through Cartesian coordinates: delta bounds what that and the
        // rounding of both tests move a distance by (fractional
        // coordinates within nl + 2 of the origin, B and Bi the row-sum
        // norms of the neighbour frame and its inverse)
        double B = 0.0, Bi = 0.0;
        for (int r = 0; r < 3; r++) {
            B = std::max(B, fabs(C.hn[3 * r]) + fabs(C.hn[3 * r + 1]) + fabs(C.hn[3 * r + 2]));
            Bi = std::max(Bi, fabs(C.ihn[3 * r]) + fabs(C.ihn[3 * r + 1]) + fabs(C.ihn[3 * r + 2]));
        }
        const double eps = 0x1p-52;
        const double delta = B * (nl + 2) * eps * (32.0 * B * Bi + 10.0);
        const double lim = C.search * (1.0 + 1e-12) + 2.0 * delta;
        C.near2 = lim * lim * (1.0 + 1e-12);
        // smallest eigenvalue of the metric hn^T hn (closed form for a
        // symmetric 3 x 3), lowered by margins: |hn du|^2 >= lam_min |du|^2
        // lets the neighbour search skip cells beyond the search radius
        double A[3][3];
        for (int r = 0; r < 3; r++)
            for (int k = 0; k < 3; k++)
                A[r][k] = C.hn[r] * C.hn[k] + C.hn[3 + r] * C.hn[3 + k] + C.hn[6 + r] * C.hn[6 + k];
        const double p1 = A[0][1] * A[0][1] + A[0][2] * A[0][2] + A[1][2] * A[1][2];
        const double q = (A[0][0] + A[1][1] + A[2][2]) / 3.0;
        double emin, emax;
        if (p1 == 0.0) {
            emin = std::min(A[0][0], std::min(A[1][1], A[2][2]));
            emax = std::max(A[0][0], std::max(A[1][1], A[2][2]));
        } else {
            const double p2 = (A[0][0] - q) * (A[0][0] - q) + (A[1][1] - q) * (A[1][1] - q) +
                              (A[2][2] - q) * (A[2][2] - q) + 2.0 * p1;
            const double p = sqrt(p2 / 6.0);
            double Bm[3][3];
            for (int r = 0; r < 3; r++)
                for (int k = 0; k < 3; k++) Bm[r][k] = (A[r][k] - (r == k ? q : 0.0)) / p;
            const double det = Bm[0][0] * (Bm[1][1] * Bm[2][2] - Bm[1][2] * Bm[2][1]) -
                               Bm[0][1] * (Bm[1][0] * Bm[2][2] - Bm[1][2] * Bm[2][0]) +
                               Bm[0][2] * (Bm[1][0] * Bm[2][1] - Bm[1][1] * Bm[2][0]);
            const double rr = std::max(-1.0, std::min(1.0, det / 2.0));
            const double phi = acos(rr) / 3.0;
            emax = q + 2.0 * p * cos(phi);
            emin = q + 2.0 * p * cos(phi + 2.0 * M_PI / 3.0);
        }
        C.lam_min = std::max(0.0, emin - 1e-9 * emax) * (1.0 - 1e-6);
    }
    C.rc3 =M->d.order[1] > 0 ? rc3 : 0.0;
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
        // cells at least half the search radius wide: 5 x 5 x 5 searched
        // cells hold ~2.4x fewer candidates than 3 x 3 x 3 of full width
        int m = (int)floor(2.0 * w[k] / (C.search * (1.0 + 1e-6)));
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
    C.model = mi;
    return cells;
}

// Frames of a batch, then the configurations' offsets.  Large batches share
// the frames out over the engine's workers in fixed chunks; a failing
// configuration reports the error of the first one in batch order either way.
void setup_frames(chx_engine *e, chx_model *const *models, int nmodels, const chx_batch *b)
{
    const int ncfg = b->nconf;
    e->cfg_h.resize(ncfg);
    e->cfg_atom_off_h.resize(ncfg + 1);
    e->cfg_in_off_h.resize(ncfg + 1);
    e->cfg_item_off_h.resize(ncfg + 1);
    std::vector<long> &ncells = e->cfg_cells_h;
    ncells.resize(ncfg);
    // up to eight workers, for batches of 512 configurations or more
    const int hw = std::min(8, (int)std::thread::hardware_concurrency());
    const int nw = ncfg >= 512 && hw > 1 ? hw : 1;
    std::vector<int> bad(nw, ncfg);
    std::vector<std::string> why(nw);
    auto chunk = [&](int k) {
        const int c0 = (int)((long)ncfg * k / nw), c1 = (int)((long)ncfg * (k + 1) / nw);
        for (int c = c0; c < c1; c++) {
            try {
                e->cfg_h[c] = chx_cfg{};
                ncells[c] = frame_of(models, nmodels, b, c, e->cfg_h[c]);
            } catch (const EngineError &x) {
                bad[k] = c;
                why[k] = x.msg;
                return;
            }
        }
    };
    if (nw > 1) {
        if (!e->pool) e->pool.reset(new WorkerPool(nw));
        e->pool->run(chunk);
    } else {
        chunk(0);
    }
    for (int k = 0; k < nw; k++)
        if (bad[k] < ncfg) throw EngineError(why[k]);
    long N = 0, Nin = 0, ncell = 0, items = 0;
    for (int c = 0; c < ncfg; c++) {
        chx_cfg &C = e->cfg_h[c];
        C.atom_off = (int)N;
        C.atom_in_off = (int)Nin;
        C.cell_off = (int)ncell;
        e->cfg_atom_off_h[c] = (int)N;
        e->cfg_in_off_h[c] = (int)Nin;
        e->cfg_item_off_h[c] = (int)items;
        items += std::max<long>(1, (C.natoms + CHX_FIN_CHUNK - 1) / CHX_FIN_CHUNK);
        N += C.natoms;
        Nin += C.natoms_in;
        ncell += ncells[c];
        if (N >= (1L << 28)) throw EngineError("batch too large");
    }
    e->cfg_atom_off_h[ncfg] = (int)N;
    e->cfg_in_off_h[ncfg] = (int)Nin;
    e->cfg_item_off_h[ncfg] = (int)items;
    e->nitems = (int)items;
    e->N = (int)N;
    e->Nin = (int)Nin;
    e->ncell = (int)ncell;
}

// First capacity guesses from the density of each configuration; a run that
// finds them too small reports the sizes it needs and is repeated.
void guess_capacities(chx_engine *e, chx_model *const *models)
{
    const double pi = 3.141592653589793;
    double nbmax = 0, l3 = 0, l4 = 0, recmax = 0;
    for (const chx_cfg &C : e->cfg_h) {
        const chx_model *M = models[C.model];
        const double rho = C.natoms / C.vol;
        // the mean half-list count; the first owners of a configuration see
        // nearly their whole neighbourhood (twice that), plus fluctuation
        const double nb = 0.5 * rho * 4.0 / 3.0 * pi * C.search * C.search * C.search;
        nbmax = std::max(nbmax, 2.0 * nb + 3.0 * sqrt(2.0 * nb) + 12);
        double r = 0;  // clusters of an owner, on average
        if (M->d.order[1] > 0) {
            const double n3 = 0.5 * rho * 4.0 / 3.0 * pi * C.rc3 * C.rc3 * C.rc3;
            l3 += C.natoms * (0.4 * n3 * n3 + 4);
            r += 0.4 * n3 * n3 + 4;
        }
        if (M->d.order[2] > 0) {
            const double n4 = 0.5 * rho * 4.0 / 3.0 * pi * C.rc4 * C.rc4 * C.rc4;
            l4 += C.natoms * (0.08 * n4 * n4 * n4 + 4);
            r += 0.08 * n4 * n4 * n4 + 4;
        }
        recmax = std::max(recmax, 2.0 * r + 32);
    }
    // A batch of the shape of the last one (same atoms, configurations and
    // groups: a plan, or repeated evaluations of one system) keeps the
    // capacities its runs measured; another shape starts from the estimate,
    // so a large batch's capacities are not carried into small ones (their
    // launch grids follow the capacities).
    const long key[4] = {e->N, e->ncfg, e->G3, e->G4};
    bool same = true;
    for (int k = 0; k < 4; k++) same = same && key[k] == e->cap_key[k];
    if (!same) {
        for (int k = 0; k < 4; k++) e->cap_key[k] = key[k];
        e->nb_cap = std::min(CHX_MAX_NB, ((int)nbmax + 7) / 8 * 8);
        e->rcap = ((int)std::min(recmax, 1e6) + 31) / 32 * 32;
        e->l3_cap = (long)l3 + (long)CHX_LIST_ALIGN * (e->G3 + 1);
        e->l4_cap = (long)l4 + (long)CHX_LIST_ALIGN * (e->G4 + 1);
    }
    e->l3_cap = (e->l3_cap + 255) / 256 * 256;
    e->l4_cap = (e->l4_cap + 255) / 256 * 256;
}

// After a complete run: capacities brought down to what the run used plus a
// margin (never below it), so the list kernels' launch grids follow the
// measured cluster counts rather than the density estimate.  A later run
// that needs more reports it and is repeated with the size it needs.
void tighten_capacities(chx_engine *e, double margin)
{
    const int *st = (const int *)(e->out_h.p + e->o_stat);
    const int nb = std::min(CHX_MAX_NB, ((int)(st[CHX_ST_MAXNB] * margin) + 8 + 7) / 8 * 8);
    const long n3 = ((long)(st[CHX_ST_N3] * margin) + 1024 + 255) / 256 * 256;
    const long n4 = ((long)(st[CHX_ST_N4] * margin) + 1024 + 255) / 256 * 256;
    e->nb_cap = std::min(e->nb_cap, std::max(nb, 8));
    const int rc = ((int)(st[CHX_ST_MAXREC] * margin) + 16 + 31) / 32 * 32;
    if (e->G3 + e->G4) e->rcap = std::min(e->rcap, std::max(rc, 32));
    if (e->G3) e->l3_cap = std::min(e->l3_cap, n3);
    if (e->G4) e->l4_cap = std::min(e->l4_cap, n4);
}

// Prepares a batch on the host: frames, groups, model tables and capacity
// guesses (nothing is enqueued).
void prepare(chx_engine *e, chx_model *const *models, int nmodels, const chx_batch *b)
{
    const int ncfg = b->nconf;
    if (ncfg <= 0) throw EngineError("empty batch");
    e->ncfg = ncfg;
    e->models.assign(models, models + nmodels);
    setup_frames(e, models, nmodels, b);
    e->g3b.assign(nmodels, 0);
    e->g4b.assign(nmodels, 0);
    e->emh.assign(nmodels, chx_enum_model{});
    e->mdh.assign(nmodels, chx_model_dev{});
    e->g3h.clear();
    e->g4h.clear();
    e->D3s.clear();
    e->D4s.clear();
    e->ngl_max = 1;
    for (int mi = 0; mi < nmodels; mi++) {
        const chx_model *M = models[mi];
        chx_enum_model &q = e->emh[mi];
        q.nt = M->d.nt;
        q.has3 = M->d.order[1] > 0 && M->d.ntrip > 0;
        q.has4 = M->d.order[2] > 0 && M->d.nquad > 0;
        q.ntrip = M->d.ntrip;
        q.nquad = M->d.nquad;
        q.g3 = (int)e->g3h.size();
        q.g4 = (int)e->g4h.size();
        q.tripmap = M->dev.tripmap;
        q.quadmap = M->dev.quadmap;
        e->g3b[mi] = q.g3;
        e->g4b[mi] = q.g4;
        for (int body = 3; body <= 4; body++) {
            const bool has = body == 3 ? q.has3 : q.has4;
            if (!has) continue;
            const int nty = body == 3 ? M->d.ntrip : M->d.nquad;
            const int D = body == 3 ? M->d.D3 : M->d.D4;
            std::vector<chx_group> &gv = body == 3 ? e->g3h : e->g4h;
            std::vector<int> &Ds = body == 3 ? e->D3s : e->D4s;
            if (std::find(Ds.begin(), Ds.end(), D) == Ds.end()) Ds.push_back(D);
            e->ngl_max = std::max(e->ngl_max, nty + 1);
            for (int t = 0; t <= nty; t++) {
                chx_group g{};
                g.model = mi;
                g.type = t < nty ? t : -1;
                g.nt = M->d.nt;
                g.D = D;
                g.tersoff = M->d.tersoff;
                g.fcut_var = M->d.fcut_var;
                g.perm = body == 3 ? M->dev.trip_perm : M->dev.quad_perm;
                if (t < nty) {
                    const TypeLaunch &tl = body == 3 ? M->t3[t] : M->t4[t];
                    g.C = tl.C;
                    g.Cf = tl.Cf;
                    g.Cm = tl.Cm;
                    g.Cmf = tl.Cmf;
                    g.Cmb = tl.Cmb;
                    for (int k = 0; k < (body == 3 ? 3 : 6); k++) g.s[k] = tl.s[k];
                }
                gv.push_back(g);
            }
        }
        e->mdh[mi] = M->dev;
    }
    e->G3 = (int)e->g3h.size();
    e->G4 = (int)e->g4h.size();
    if (e->G3 >= 2048 || e->G4 >= 2048) throw EngineError("too many (model, cluster type) groups in one batch");
    if (b->precision < CHX_FP64 || b->precision > CHX_BF16X3) throw EngineError("unknown precision");
    e->prec = b->precision;
    guess_capacities(e, models);
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

constexpr int CHX_NCLEAR = 10;  // ranges of the clear list

// Sizes every buffer, lays out the input and output blobs and fills the host
// input blob.  host_pos: positions/types are host arrays copied in the blob;
// host_out: energy/virial/forces come back in the output blob.
void layout(chx_engine *e, const chx_batch *b, bool host_out)
{
    const long N = e->N, Nin = e->Nin, ncell = e->ncell;
    const int ncfg = e->ncfg, nmodels = (int)e->models.size(), G3 = e->G3, G4 = e->G4;
    e->as.ensure(N);
    e->an.ensure(N);
    e->cell_cnt.ensure(ncell + 1);
    e->cell_start.ensure(ncell + 1);
    e->cell_cur.ensure(ncell + 1);
    e->cell_atoms.ensure(N);
    e->nbr_cnt.ensure(N + 1);
    e->big.ensure(N + 1);
    e->in_cfg.ensure(Nin + 1);
    e->item_cfg.ensure(e->nitems + 1);
    e->nbr_off.ensure(N + 1);
    e->cnt34.ensure((long)(G3 + G4) * N + 1);
    e->off34.ensure((long)(G3 + G4) * N + 2);
    e->gb3.ensure(G3 + 1);
    e->ge3.ensure(G3 + 1);
    e->gb4.ensure(G4 + 1);
    e->ge4.ensure(G4 + 1);
    e->Facc.ensure(6 * N);
    e->Cacc.ensure((long)CHX_CACC_SLOTS * CHX_CACC_STRIDE * ncfg);
    e->part.ensure(14L * e->nitems);
    e->done.ensure(ncfg + 1);
    // input blob
    size_t off = 0;
    const size_t o_cfg = off; off = align256(off + ncfg * sizeof(chx_cfg));
    const size_t o_cao = off; off = align256(off + (ncfg + 1) * sizeof(int));
    const size_t o_cio = off; off = align256(off + (ncfg + 1) * sizeof(int));
    const size_t o_cfi = off; off = align256(off + (ncfg + 1) * sizeof(int));
    const size_t o_em = off; off = align256(off + nmodels * sizeof(chx_enum_model));
    const size_t o_md = off; off = align256(off + nmodels * sizeof(chx_model_dev));
    const size_t o_g3 = off; off = align256(off + (G3 + 1) * sizeof(chx_group));
    const size_t o_g4 = off; off = align256(off + (G4 + 1) * sizeof(chx_group));
    const size_t o_cl = off; off = align256(off + CHX_NCLEAR * sizeof(chx_clear_range));
    const bool host_pos = !b->pos_dev;
    const size_t o_pos = off; if (host_pos) off = align256(off + 3 * Nin * sizeof(double));
    const size_t o_typ = off; if (host_pos) off = align256(off + Nin * sizeof(int));
    e->in_bytes = off;
    e->in_h.ensure(off);
    e->in_d.ensure(off);
    char *hb = e->in_h.p, *db = e->in_d.p;
    e->cfgs = (chx_cfg *)(db + o_cfg);
    e->cfg_atom_off = (int *)(db + o_cao);
    e->cfg_in_off = (int *)(db + o_cio);
    e->cfg_item_off = (int *)(db + o_cfi);
    e->em = (chx_enum_model *)(db + o_em);
    e->mdev = (chx_model_dev *)(db + o_md);
    e->grp3 = (chx_group *)(db + o_g3);
    e->grp4 = (chx_group *)(db + o_g4);
    e->clear = (chx_clear_range *)(db + o_cl);
    e->pos_in = host_pos ? (double *)(db + o_pos) : const_cast<double *>(b->pos);
    e->types_in = host_pos ? (int *)(db + o_typ) : const_cast<int *>(b->types);
    // the frames of a large batch are copied by the frame workers in chunks
    if (e->pool && ncfg >= 512) {
        const int nw = e->pool->size();
        e->pool->run([&](int k) {
            const long c0 = (long)ncfg * k / nw, c1 = (long)ncfg * (k + 1) / nw;
            memcpy(hb + o_cfg + c0 * sizeof(chx_cfg), e->cfg_h.data() + c0, (c1 - c0) * sizeof(chx_cfg));
        });
    } else {
        memcpy(hb + o_cfg, e->cfg_h.data(), ncfg * sizeof(chx_cfg));
    }
    memcpy(hb + o_cao, e->cfg_atom_off_h.data(), (ncfg + 1) * sizeof(int));
    memcpy(hb + o_cio, e->cfg_in_off_h.data(), (ncfg + 1) * sizeof(int));
    memcpy(hb + o_cfi, e->cfg_item_off_h.data(), (ncfg + 1) * sizeof(int));
    memcpy(hb + o_em, e->emh.data(), nmodels * sizeof(chx_enum_model));
    memcpy(hb + o_md, e->mdh.data(), nmodels * sizeof(chx_model_dev));
    if (G3) memcpy(hb + o_g3, e->g3h.data(), G3 * sizeof(chx_group));
    if (G4) memcpy(hb + o_g4, e->g4h.data(), G4 * sizeof(chx_group));
    if (host_pos) {
        memcpy(hb + o_pos, b->pos, 3 * Nin * sizeof(double));
        memcpy(hb + o_typ, b->types, Nin * sizeof(int));
    }
    // output blob: status, per-configuration flags[, energy, virial, forces]
    off = 0;
    e->o_stat = off; off = align256(off + CHX_ST_WORDS * sizeof(int));
    e->o_cstat = off; off = align256(off + 4 * ncfg * sizeof(int));
    e->o_E = off; if (host_out) off = align256(off + ncfg * sizeof(double));
    e->o_V = off; if (host_out) off = align256(off + 6 * ncfg * sizeof(double));
    e->o_F = off; if (host_out) off = align256(off + 3 * Nin * sizeof(double));
    e->out_bytes = off;
    e->out_h.ensure(off);
    e->out_d.ensure(off);
    e->status = (int *)(e->out_d.p + e->o_stat);
    e->cfg_stat = (int *)(e->out_d.p + e->o_cstat);
    if (host_out) {
        e->E_out = (double *)(e->out_d.p + e->o_E);
        e->V_out = (double *)(e->out_d.p + e->o_V);
        e->F_out = (double *)(e->out_d.p + e->o_F);
    }
}

// Work buffers sized by the current capacities, and the clear list (the
// ranges one kernel zeroes at the start of a run).
void size_work(chx_engine *e)
{
    if ((long)e->N * e->nb_cap >= (1L << 31)) throw EngineError("batch too large for the neighbour slabs");
    e->es.ensure((long)e->N * e->nb_cap);
    e->en.ensure((long)e->N * e->nb_cap);
    e->list3.ensure(e->l3_cap);
    e->list4.ensure(e->l4_cap);
    const long N = e->N;
    if (e->G3 + e->G4) {
        if (N * e->rcap >= (1L << 40)) throw EngineError("batch too large for the cluster record slabs");
        e->rec.ensure(N * e->rcap);
        e->nrec.ensure(N);
    }
    chx_clear_range cl[CHX_NCLEAR];
    int n = 0;
    cl[n++] = chx_clear_range{e->status, CHX_ST_WORDS};
    cl[n++] = chx_clear_range{e->cfg_stat, 4L * e->ncfg};
    cl[n++] = chx_clear_range{e->cell_cnt.p, (long)e->ncell + 1};
    cl[n++] = chx_clear_range{e->cell_cur.p, (long)e->ncell + 1};
    cl[n++] = chx_clear_range{e->cnt34.p, (long)(e->G3 + e->G4) * N + 1};
    if (e->G3 + e->G4) cl[n++] = chx_clear_range{e->nrec.p, N};
    cl[n++] = chx_clear_range{(int *)e->Facc.p, 12 * N};
    cl[n++] = chx_clear_range{(int *)e->Cacc.p, 2L * CHX_CACC_SLOTS * CHX_CACC_STRIDE * e->ncfg};
    cl[n++] = chx_clear_range{e->done.p, (long)e->ncfg + 1};
    e->nclear = n;
    memcpy(e->in_h.p + ((char *)e->clear - e->in_d.p), cl, sizeof(cl));
}

// One pass of the pipeline on the engine's stream, between the blob copies
// (no host synchronisation).
void run_pipeline(chx_engine *e, double *Fout, double *Ecfg, double *Vcfg)
{
    hipStream_t st = e->st;
    const long N = e->N, Nin = e->Nin, ncell = e->ncell;
    const int ncfg = e->ncfg, G3 = e->G3, G4 = e->G4;
    const int prec = e->prec;
    e->nb_cap_run = e->nb_cap;
    // A captured plan evaluates the 2-body terms beside the enumeration and
    // the 4-body terms beside the 3-body ones, on a second stream (every sum
    // is exact, so the order of the atomic adds cannot change a bit).  Eager
    // runs keep one stream, in the order the kernel-group timing assumes.
    const bool fork = e->capturing;
    hipStream_t s2 = fork ? e->st2 : st;
    auto fork_at = [&](int k) {
        HIPCHK(hipEventRecord(e->fork_ev[k], st));
        HIPCHK(hipStreamWaitEvent(e->st2, e->fork_ev[k], 0));
    };
    auto eval2 = [&] {
        launch_eval2((int)N, e->nb_cap, s2, e->nbr_cnt.p, e->cfgs, e->as.p, e->es.p, e->mdev, e->Facc.p,
                     e->Cacc.p, e->cfg_stat, e->status);
    };
    e->record(0);
    HIPCHK(hipMemcpyAsync(e->in_d.p, e->in_h.p, e->in_bytes, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_clear, dim3(1024), dim3(256), 0, st, e->clear, e->nclear);
    hipLaunchKernelGGL(k_prep, dim3(nblk(N, 256)), dim3(256), 0, st, (int)N, ncfg, e->cfg_atom_off, e->cfgs,
                       e->pos_in, e->types_in, e->as.p, e->an.p, e->cell_cnt.p, e->cfg_item_off, e->in_cfg.p,
                       e->item_cfg.p);
    e->scan(e->cell_cnt.p, e->cell_start.p, ncell);
    hipLaunchKernelGGL(k_cell_fill, dim3(nblk(N, 256)), dim3(256), 0, st, (int)N, e->an.p, e->cell_start.p,
                       e->cell_cur.p, e->cell_atoms.p);
    // neighbour images, sorted per owner
    // small batches: one owner per workgroup (four wavefronts), else one per wavefront
    const bool wide = N <= CHX_ENUM_WIDE_MAX;
    if (wide)
        hipLaunchKernelGGL(k_nbr_wide, dim3((unsigned)N), dim3(256), 0, st, (int)N, e->cfgs, e->as.p, e->an.p,
                           e->cell_start.p, e->cell_atoms.p, e->nbr_cnt.p, e->nbr_off.p, e->es.p, e->en.p,
                           e->nb_cap, e->status);
    else
        hipLaunchKernelGGL(k_nbr, dim3(nblk(N, 4)), dim3(256), 0, st, (int)N, e->cfgs, e->as.p, e->an.p,
                           e->cell_start.p, e->cell_atoms.p, e->nbr_cnt.p, e->nbr_off.p, e->es.p, e->en.p,
                           e->nb_cap, e->status);
    if (fork) {
        fork_at(0);
        eval2();
    }
    // 3-/4-body lists in type-aligned segments
    if (G3 + G4 > 0) {
        chx_enum_args ea{};
        ea.natoms = (int)N;
        ea.cfgs = e->cfgs;
        ea.as = e->as.p;
        ea.an = e->an.p;
        ea.nbr_off = e->nbr_off.p;
        ea.nbr_cnt = e->nbr_cnt.p;
        ea.es = e->es.p;
        ea.en = e->en.p;
        ea.em = e->em;
        ea.cnt3 = e->cnt34.p;
        ea.cnt4 = e->cnt34.p + (long)G3 * N;
        ea.off3 = e->off34.p;
        ea.off4 = e->off34.p + (long)G3 * N;
        ea.gb3 = e->gb3.p;
        ea.gb4 = e->gb4.p;
        ea.G3 = G3;
        ea.G4 = G4;
        ea.list3 = e->list3.p;
        ea.list4 = e->list4.p;
        ea.rec = e->rec.p;
        ea.nrec = e->nrec.p;
        ea.rcap = e->rcap;
        ea.big = e->big.p;
        ea.big_on = e->big_on;
        ea.status = e->status;
        // owners with many images are queued by k_enum for the large-LDS
        // kernel, which is launched once a batch has needed it (until then a
        // queued owner asks for a retry)
        const dim3 gbig(1280);
        // small batches: four wavefronts share an owner's enumeration (latency)
        if (wide) hipLaunchKernelGGL((k_enum<256>), dim3((unsigned)N), dim3(256), 0, st, ea);
        else hipLaunchKernelGGL((k_enum<64>), dim3((unsigned)N), dim3(64), 0, st, ea);
        if (e->big_on) hipLaunchKernelGGL(k_enum_big, gbig, dim3(64), 0, st, ea);
        // one scan over both bodies' counts: segments and positions use
        // differences of the scan only
        chx_bases_args gb{};
        gb.off = e->off34.p;
        gb.N = N;
        gb.G3 = G3;
        gb.G4 = G4;
        gb.cap3 = e->l3_cap;
        gb.cap4 = e->l4_cap;
        gb.base3 = e->gb3.p;
        gb.end3 = e->ge3.p;
        gb.base4 = e->gb4.p;
        gb.end4 = e->ge4.p;
        gb.status = e->status;
        e->scan(e->cnt34.p, e->off34.p, (long)(G3 + G4) * N, 0, gb);
        // every record to its list position
        hipLaunchKernelGGL(k_place, dim3(nblk(N, 4)), dim3(256), 0, st, ea);
    }
    e->record(1);
    // evaluation with exact accumulation
    if (!fork) eval2();
    e->record(2);
    if (fork) fork_at(1);
    chx_list_args la{};
    la.cfgs = e->cfgs;
    la.as = e->as.p;
    la.nb_cap = e->nb_cap;
    la.es = e->es.p;
    la.Facc = e->Facc.p;
    la.Cacc = e->Cacc.p;
    la.cfg_stat = e->cfg_stat;
    la.status = e->status;
    if (G3) {
        chx_list_args l3 = la;
        l3.cap = e->l3_cap;
        l3.G = G3;
        l3.gbase = e->gb3.p;
        l3.gend = e->ge3.p;
        l3.groups = e->grp3;
        l3.list = e->list3.p;
        for (int D : e->D3s) launch_eval_body(3, D, prec, e->var3, l3, st);
    }
    e->record(3);
    if (G4) {
        chx_list_args l4 = la;
        l4.cap = e->l4_cap;
        l4.G = G4;
        l4.gbase = e->gb4.p;
        l4.gend = e->ge4.p;
        l4.groups = e->grp4;
        l4.list = e->list4.p;
        for (int D : e->D4s) launch_eval_body(4, D, prec, e->var4, l4, s2);
    }
    if (fork) {
        HIPCHK(hipEventRecord(e->join_ev, e->st2));
        HIPCHK(hipStreamWaitEvent(st, e->join_ev, 0));
    }
    e->record(4);
    hipLaunchKernelGGL(k_fin, dim3(e->nitems + nblk(Nin, 256)), dim3(256), 0, st, ncfg, e->nitems,
                       e->cfg_item_off, e->item_cfg.p, (int)Nin, e->in_cfg.p, e->cfgs, e->as.p, e->mdev, e->Facc.p,
                       e->Cacc.p, e->nbr_cnt.p, e->part.p, e->done.p, Fout, Ecfg, Vcfg, e->status);
    HIPCHK(hipMemcpyAsync(e->out_h.p, e->out_d.p, e->out_bytes, hipMemcpyDeviceToHost, st));
    e->record(5);
}

// Status of the last run (host copy): throws on errors, grows capacities and
// returns true when the run must be repeated.
bool check_status(chx_engine *e)
{
    const int *st_h = (const int *)(e->out_h.p + e->o_stat);
    const int fl = st_h[CHX_ST_FLAGS];
    if (fl & CHX_BAD_NB)
        throw EngineError("an atom has " + std::to_string(st_h[CHX_ST_MAXNB]) + " neighbour images within the "
                          "search cutoff; at most " + std::to_string(CHX_MAX_NB) + " are supported");
    if (fl & CHX_BAD_N34)
        throw EngineError("an atom has more than " + std::to_string(CHX_MAX_N3) + " images within the 3-body "
                          "cutoff or more than " + std::to_string(CHX_MAX_N4) + " within the 4-body cutoff");
    if (!(fl & CHX_RETRY_ANY)) {
        if (fl & CHX_BAD_RANGE)
            throw EngineError("a cluster energy, force or virial contribution is outside the accumulator range "
                              "(|value| >= 2^42)");
        e->counts[0] = st_h[CHX_ST_NENT];
        e->counts[1] = e->G3 ? st_h[CHX_ST_C3] : 0;
        e->counts[2] = e->G4 ? st_h[CHX_ST_C4] : 0;
        return false;
    }
    if (fl & CHX_RETRY_BIG) e->big_on = true;
    if (fl & CHX_RETRY_ENT)
        e->nb_cap = std::max(e->nb_cap, std::min(CHX_MAX_NB, (st_h[CHX_ST_MAXNB] + 8 + 7) / 8 * 8));
    if (fl & CHX_RETRY_REC) e->rcap = std::max(e->rcap, ((int)(st_h[CHX_ST_MAXREC] * 1.2) + 32 + 31) / 32 * 32);
    if (fl & CHX_RETRY_L3) e->l3_cap = std::max<long>(e->l3_cap, (long)(st_h[CHX_ST_N3] * 1.2) + 4096);
    if (fl & CHX_RETRY_L4) e->l4_cap = std::max<long>(e->l4_cap, (long)(st_h[CHX_ST_N4] * 1.2) + 4096);
    e->l3_cap = (e->l3_cap + 255) / 256 * 256;
    e->l4_cap = (e->l4_cap + 255) / 256 * 256;
    return true;
}

void read_timing(chx_engine *e)
{
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e->ev[0], e->ev[5]);
    e->timing[0] = ms;
    (void)hipEventElapsedTime(&ms, e->ev[3], e->ev[4]);
    e->timing[1] = ms;
    (void)hipEventElapsedTime(&ms, e->ev[2], e->ev[3]);
    e->timing[2] = ms;
    (void)hipEventElapsedTime(&ms, e->ev[0], e->ev[1]);
    e->timing[3] = ms;
}

// Per-configuration results from the output blob into the caller's arrays.
void deliver(chx_engine *e, chx_out *o)
{
    const int ncfg = e->ncfg;
    const int *stat = (const int *)(e->out_h.p + e->o_cstat);
    if (!o->out_dev) {
        memcpy(o->energy, e->out_h.p + e->o_E, ncfg * sizeof(double));
        memcpy(o->virial, e->out_h.p + e->o_V, 6 * ncfg * sizeof(double));
        memcpy(o->forces, e->out_h.p + e->o_F, 3 * (size_t)e->Nin * sizeof(double));
    }
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
}

// Eager evaluation: runs the pipeline until no capacity is short.
void evaluate_eager(chx_engine *e, chx_out *o)
{
    double *Fout = o->out_dev ? o->forces : e->F_out;
    double *Ecfg = o->out_dev ? o->energy : e->E_out;
    double *Vcfg = o->out_dev ? o->virial : e->V_out;
    for (int attempt = 0;; attempt++) {
        size_work(e);
        run_pipeline(e, Fout, Ecfg, Vcfg);
        HIPCHK(hipStreamSynchronize(e->st));
        HIPCHK(hipGetLastError());
        if (!check_status(e)) break;
        if (attempt >= 3) throw EngineError("buffer sizing did not converge");
    }
    read_timing(e);
}

}  // namespace

extern "C" {

int chx_evaluate(chx_engine *e, chx_model *const *models, int nmodels, const chx_batch *b, chx_out *o)
{
    try {
        HIPCHK(hipSetDevice(e->device));
        prepare(e, models, nmodels, b);
        layout(e, b, !o->out_dev);
        evaluate_eager(e, o);
        deliver(e, o);
        tighten_capacities(e, 1.1);  // for the next evaluation of this shape
        e->err.clear();
        return 0;
    } catch (const EngineError &x) {
        e->err = x.msg;
        (void)hipStreamSynchronize(e->st);
        return 1;
    }
}

// ---------------------------------------------------------------------------
// Batch plans: a fixed batch shape evaluated by replaying one captured graph.
// ---------------------------------------------------------------------------
struct chx_plan {
    chx_engine *e = nullptr;
    std::vector<chx_model *> models;
    std::vector<int> natoms, model_of;
    int nconf = 0, small = 0, precision = 0, pos_dev = 0, out_dev = 0;
    const double *pos_dev_p = nullptr;
    const int *types_dev_p = nullptr;
    double *dev_out[3] = {nullptr, nullptr, nullptr};
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    long sig[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // N, Nin, ncell, nb_cap, l3_cap, l4_cap, nitems, rcap of the captured graph
    double tune_ms[2][CHX_NVAR] = {};  // per (body, variant); 0 where a variant was not timed
    std::string err;
    ~chx_plan()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (e) chx_engine_destroy(e);
    }
};

namespace {

void plan_batch(chx_plan *p, const double *pos, const int *types, const double *cells, chx_batch &b)
{
    b.nconf = p->nconf;
    b.natoms = p->natoms.data();
    b.pos = p->pos_dev ? p->pos_dev_p : pos;
    b.types = p->pos_dev ? p->types_dev_p : types;
    b.cells = cells;
    b.model = p->model_of.data();
    b.small = p->small;
    b.pos_dev = p->pos_dev;
    b.precision = p->precision;
}

void plan_release_graph(chx_plan *p)
{
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    p->exec = nullptr;
    p->graph = nullptr;
}

// Captures the pipeline at the current sizes (buffers already grown by an
// eager run, so no allocation happens while capturing).
void plan_capture(chx_plan *p, chx_out *o)
{
    chx_engine *e = p->e;
    plan_release_graph(p);
    double *Fout = o->out_dev ? o->forces : e->F_out;
    double *Ecfg = o->out_dev ? o->energy : e->E_out;
    double *Vcfg = o->out_dev ? o->virial : e->V_out;
    size_work(e);
    e->capturing = true;
    HIPCHK(hipStreamBeginCapture(e->st, hipStreamCaptureModeThreadLocal));
    try {
        run_pipeline(e, Fout, Ecfg, Vcfg);
    } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(e->st, &g);
        if (g) (void)hipGraphDestroy(g);
        e->capturing = false;
        throw;
    }
    e->capturing = false;
    HIPCHK(hipStreamEndCapture(e->st, &p->graph));
    HIPCHK(hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0));
    const long sig[8] = {e->N, e->Nin, e->ncell, (long)e->nb_cap, e->l3_cap, e->l4_cap, e->nitems, (long)e->rcap};
    for (int k = 0; k < 8; k++) p->sig[k] = sig[k];
}

// Results of the last eager run (raw energy, virial and forces), read from
// the output blob or from the caller's device buffers.
struct Snap {
    std::vector<double> E, V, F;
};

void snapshot(chx_engine *e, const chx_out *o, Snap &s)
{
    const int ncfg = e->ncfg;
    const long Nin = e->Nin;
    s.E.resize(ncfg);
    s.V.resize(6 * (size_t)ncfg);
    s.F.resize(3 * (size_t)Nin);
    if (o->out_dev) {
        HIPCHK(hipMemcpyAsync(s.E.data(), o->energy, ncfg * sizeof(double), hipMemcpyDeviceToHost, e->st));
        HIPCHK(hipMemcpyAsync(s.V.data(), o->virial, 6 * ncfg * sizeof(double), hipMemcpyDeviceToHost, e->st));
        HIPCHK(hipMemcpyAsync(s.F.data(), o->forces, 3 * Nin * sizeof(double), hipMemcpyDeviceToHost, e->st));
        HIPCHK(hipStreamSynchronize(e->st));
    } else {
        memcpy(s.E.data(), e->out_h.p + e->o_E, ncfg * sizeof(double));
        memcpy(s.V.data(), e->out_h.p + e->o_V, 6 * ncfg * sizeof(double));
        memcpy(s.F.data(), e->out_h.p + e->o_F, 3 * Nin * sizeof(double));
    }
}

// Whether the results a of a kernel variant agree with the reference results
// r within the bars of precision prec, per configuration (n = its replicated
// atoms; energies and virials are raw sums in kcal/mol, forces kcal/mol/A):
//   FP64   energy 1e-10 relative, forces 1e-10 of the largest + 1e-9,
//          virial 1e-10 of the largest component;
//   MIXED  energy 1e-6 eV/atom, forces 1e-4 eV/A (r is the FP64 result),
//          virial 1e-5 of the largest component;
//   BF16 forms: 1e-6 of each quantity's scale (r is the same form's result
//          with the default variants: the variants differ only in tiling).
bool agrees(const chx_engine *e, const Snap &a, const Snap &r, int prec)
{
    const double kcal_per_ev = 1.0 / 0.0433641;
    for (int c = 0; c < e->ncfg; c++) {
        const double n = e->cfg_h[c].natoms;
        const long f0 = 3L * e->cfg_h[c].atom_in_off, f1 = f0 + 3L * e->cfg_h[c].natoms_in;
        double fmax = 0, vmax = 0, df = 0, dv = 0;
        for (long k = f0; k < f1; k++) {
            fmax = std::max(fmax, fabs(r.F[k]));
            df = std::max(df, fabs(a.F[k] - r.F[k]));
        }
        for (int k = 0; k < 6; k++) {
            vmax = std::max(vmax, fabs(r.V[6 * c + k]));
            dv = std::max(dv, fabs(a.V[6 * c + k] - r.V[6 * c + k]));
        }
        const double de = fabs(a.E[c] - r.E[c]);
        bool ok;
        if (prec == CHX_FP64)
            ok = de <= 1e-10 * fabs(r.E[c]) + 1e-12 * n && df <= 1e-10 * fmax + 1e-9 && dv <= 1e-10 * vmax + 1e-9;
        else if (prec == CHX_MIXED)
            ok = de <= 1e-6 * kcal_per_ev * n && df <= 1e-4 * kcal_per_ev && dv <= 1e-5 * vmax + 1e-6 * n;
        else
            ok = de <= 1e-6 * (fabs(r.E[c]) + n) && df <= 1e-6 * (fmax + 1.0) && dv <= 1e-6 * (vmax + n);
        if (!ok) return false;
    }
    return true;
}

// Times each 3-/4-body kernel variant on the lists of the last eager run, then
// keeps per body order the fastest variant whose results agree with a
// reference evaluation (the default variants; in FP64 for the FP32-mixed
// form) within the precision's bars.  A variant that disagrees is never
// chosen; its timing is stored negated.  Ends with an eager run of the chosen
// variants (results in o).
void plan_tune(chx_plan *p, chx_out *o)
{
    chx_engine *e = p->e;
    hipEvent_t a, z;
    HIPCHK(hipEventCreate(&a));
    HIPCHK(hipEventCreate(&z));
    for (int body = 3; body <= 4; body++) {
        const int G = body == 3 ? e->G3 : e->G4;
        if (!G) continue;
        chx_list_args la{};
        la.cfgs = e->cfgs;
        la.as = e->as.p;
        la.nb_cap = e->nb_cap;
        la.es = e->es.p;
        la.Facc = e->Facc.p;  // accumulators are overwritten by the next real run
        la.Cacc = e->Cacc.p;
        la.cfg_stat = e->cfg_stat;
        la.status = e->status;
        la.cap = body == 3 ? e->l3_cap : e->l4_cap;
        la.G = G;
        la.gbase = body == 3 ? e->gb3.p : e->gb4.p;
        la.gend = body == 3 ? e->ge3.p : e->ge4.p;
        la.groups = body == 3 ? e->grp3 : e->grp4;
        la.list = body == 3 ? e->list3.p : e->list4.p;
        const std::vector<int> &Ds = body == 3 ? e->D3s : e->D4s;
        // the BF16 forms are defined by their matrix-core contraction: no vector variant
        const int nvar = e->prec >= CHX_BF16 ? CHX_VAR_VECTOR : CHX_NVAR;
        for (int var = 0; var < nvar; var++) {
            if (!variant_applies(body, e->prec, var)) continue;
            float best_rep = 1e30f;
            for (int rep = 0; rep < 3; rep++) {
                HIPCHK(hipEventRecord(a, e->st));
                for (int D : Ds) launch_eval_body(body, D, e->prec, var, la, e->st);
                HIPCHK(hipEventRecord(z, e->st));
                HIPCHK(hipEventSynchronize(z));
                float ms = 0;
                HIPCHK(hipEventElapsedTime(&ms, a, z));
                if (rep > 0) best_rep = std::min(best_rep, ms);
            }
            p->tune_ms[body - 3][var] = best_rep;
        }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(z);
    HIPCHK(hipGetLastError());
    // reference results: the default variants, in FP64 for the mixed form
    const int prec = e->prec;
    Snap ref, cand;
    if (prec == CHX_MIXED) e->prec = CHX_FP64;
    try {
        evaluate_eager(e, o);
    } catch (...) {
        e->prec = prec;
        throw;
    }
    e->prec = prec;
    snapshot(e, o, ref);
    for (int body = 3; body <= 4; body++) {
        const int G = body == 3 ? e->G3 : e->G4;
        if (!G) continue;
        int &var = body == 3 ? e->var3 : e->var4;
        const int def = var;
        const int nvar = prec >= CHX_BF16 ? CHX_VAR_VECTOR : CHX_NVAR;
        std::vector<int> order;
        for (int v = 0; v < nvar; v++)
            if (variant_applies(body, prec, v)) order.push_back(v);
        std::stable_sort(order.begin(), order.end(),
                         [&](int x, int y) { return p->tune_ms[body - 3][x] < p->tune_ms[body - 3][y]; });
        int chosen = -1;
        for (int v : order) {
            var = v;
            evaluate_eager(e, o);
            snapshot(e, o, cand);
            if (agrees(e, cand, ref, prec)) {
                chosen = v;
                break;
            }
            p->tune_ms[body - 3][v] = -fabs(p->tune_ms[body - 3][v]);
        }
        var = chosen >= 0 ? chosen : def;
    }
    evaluate_eager(e, o);  // results of the chosen variants
}

}  // namespace

chx_plan *chx_plan_create(int device, chx_model *const *models, int nmodels, const chx_batch *b, int tune,
                          chx_out *o, char *err, int errlen)
{
    chx_plan *p = new chx_plan;
    try {
        p->e = chx_engine_create(device);
        if (!p->e) throw EngineError("cannot create an engine on this device");
        chx_engine *e = p->e;
        HIPCHK(hipSetDevice(e->device));
        p->models.assign(models, models + nmodels);
        p->nconf = b->nconf;
        p->natoms.assign(b->natoms, b->natoms + b->nconf);
        p->model_of.assign(b->model, b->model + b->nconf);
        p->small = b->small;
        p->precision = b->precision;
        p->pos_dev = b->pos_dev;
        p->out_dev = o->out_dev;
        if (p->pos_dev) {
            p->pos_dev_p = b->pos;
            p->types_dev_p = b->types;
        }
        if (p->out_dev) {
            p->dev_out[0] = o->energy;
            p->dev_out[1] = o->forces;
            p->dev_out[2] = o->virial;
        }
        prepare(e, models, nmodels, b);
        layout(e, b, !o->out_dev);
        evaluate_eager(e, o);
        if (tune) plan_tune(p, o);
        deliver(e, o);
        tighten_capacities(e, 1.02);  // the captured grids follow this batch's counts
        plan_capture(p, o);
        return p;
    } catch (const EngineError &x) {
        if (err && errlen > 0) {
            strncpy(err, x.msg.c_str(), errlen - 1);
            err[errlen - 1] = 0;
        }
        delete p;
        return nullptr;
    }
}

int chx_plan_run(chx_plan *p, const double *pos, const int *types, const double *cells, chx_out *o)
{
    chx_engine *e = p->e;
    try {
        HIPCHK(hipSetDevice(e->device));
        if (o->out_dev != p->out_dev) throw EngineError("output location differs from the plan's");
        if (p->out_dev && (o->energy != p->dev_out[0] || o->forces != p->dev_out[1] || o->virial != p->dev_out[2]))
            throw EngineError("device output buffers differ from the plan's");
        chx_batch b{};
        plan_batch(p, pos, types, cells, b);
        prepare(e, p->models.data(), (int)p->models.size(), &b);
        layout(e, &b, !o->out_dev);
        const long sig[8] = {e->N, e->Nin, e->ncell, (long)e->nb_cap, e->l3_cap, e->l4_cap, e->nitems, (long)e->rcap};
        bool same = p->exec != nullptr;
        for (int k = 0; k < 8; k++) same = same && sig[k] == p->sig[k];
        if (same) {
            size_work(e);
            HIPCHK(hipGraphLaunch(p->exec, e->st));
            HIPCHK(hipStreamSynchronize(e->st));
            HIPCHK(hipGetLastError());
            if (!check_status(e)) {
                deliver(e, o);
                e->err.clear();
                return 0;
            }
        }
        // new shape or a capacity ran short: eager run, then capture again
        evaluate_eager(e, o);
        deliver(e, o);
        tighten_capacities(e, 1.02);
        plan_capture(p, o);
        e->err.clear();
        return 0;
    } catch (const EngineError &x) {
        p->err = x.msg;
        (void)hipStreamSynchronize(e->st);
        return 1;
    }
}

void chx_plan_destroy(chx_plan *p) { delete p; }

int chx_plan_set_variants(chx_plan *p, int var3, int var4)
{
    if (chx_set_variants(p->e, var3, var4)) {
        p->err = p->e->err;
        return 1;
    }
    plan_release_graph(p);  // the next run evaluates directly and captures again
    return 0;
}

const char *chx_plan_error(chx_plan *p) { return p ? p->err.c_str() : "no plan"; }

void chx_plan_info(chx_plan *p, int *variants, double *tune_ms, long long *counts)
{
    variants[0] = p->e->var3;
    variants[1] = p->e->var4;
    for (int b = 0; b < 2; b++)
        for (int v = 0; v < CHX_NVAR; v++) tune_ms[CHX_NVAR * b + v] = p->tune_ms[b][v];
    for (int k = 0; k < 3; k++) counts[k] = p->e->counts[k];
}

}  // extern "C"

extern "C" {

int chx_export_lists(chx_engine *e, int cfg, long *n2, int **pairs, long *n3, int **trips, long *n4, int **quads)
{
    try {
        HIPCHK(hipSetDevice(e->device));
        if (cfg < 0 || cfg >= e->ncfg) throw EngineError("configuration index out of range");
        const chx_cfg &C = e->cfg_h[cfg];
        const long N = e->N;
        // the configuration's owners' slabs, and their counts
        std::vector<int> cnt(C.natoms);
        HIPCHK(hipMemcpy(cnt.data(), e->nbr_cnt.p + C.atom_off, C.natoms * sizeof(int), hipMemcpyDeviceToHost));
        const long cap = e->nb_cap_run, e0 = (long)C.atom_off * cap, nslab = (long)C.natoms * cap;
        std::vector<long> off(C.natoms + 1, 0);
        for (int i = 0; i < C.natoms; i++) off[i + 1] = off[i] + cnt[i];
        const long ne = off[C.natoms];
        std::vector<chx_entry_s> es(nslab);
        std::vector<chx_entry_n> en(nslab);
        if (nslab) {
            HIPCHK(hipMemcpy(es.data(), e->es.p + e0, nslab * sizeof(chx_entry_s), hipMemcpyDeviceToHost));
            HIPCHK(hipMemcpy(en.data(), e->en.p + e0, nslab * sizeof(chx_entry_n), hipMemcpyDeviceToHost));
        }
        auto key = [&](int owner_local, int code, int *o) {
            if (code == 0) { o[0] = owner_local; o[1] = o[2] = o[3] = 0; return; }
            const long idx = (long)owner_local * cap + code - 1;
            o[0] = es[idx].parent - C.atom_off;
            const int s = en[idx].shift;
            o[1] = (s & 255) - 128;
            o[2] = ((s >> 8) & 255) - 128;
            o[3] = ((s >> 16) & 255) - 128;
        };
        *n2 = ne;
        *pairs = (int *)malloc(sizeof(int) * 5 * std::max<long>(ne, 1));
        for (int i = 0; i < C.natoms; i++)
            for (int s = 1; s <= cnt[i]; s++) {
                int *o = *pairs + 5 * (off[i] + s - 1);
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
                const int *offv = body == 3 ? e->off34.p : e->off34.p + (long)e->G3 * N;
                const int *gbase = body == 3 ? e->gb3.p : e->gb4.p;
                const int2 *lst = body == 3 ? e->list3.p : e->list4.p;
                for (int g = gb; g < gb + ng; g++) {
                    int r[3], base;
                    HIPCHK(hipMemcpy(&r[0], offv + (long)g * N, sizeof(int), hipMemcpyDeviceToHost));
                    HIPCHK(hipMemcpy(&r[1], offv + (long)g * N + C.atom_off, sizeof(int), hipMemcpyDeviceToHost));
                    HIPCHK(hipMemcpy(&r[2], offv + (long)g * N + C.atom_off + C.natoms, sizeof(int),
                                     hipMemcpyDeviceToHost));
                    HIPCHK(hipMemcpy(&base, gbase + g, sizeof(int), hipMemcpyDeviceToHost));
                    std::vector<int2> recs(r[2] - r[1]);
                    if (!recs.empty())
                        HIPCHK(hipMemcpy(recs.data(), lst + base + (r[1] - r[0]), recs.size() * sizeof(int2),
                                         hipMemcpyDeviceToHost));
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

int chx_eval_clusters_static(chx_engine *e, chx_model *m, int nb, int n, const double *dist, const double *dr,
                             const int *types, double *energy, double *force, double *stress)
{
    try {
        HIPCHK(hipSetDevice(e->device));
        if (nb < 2 || nb > 4) throw EngineError("cluster size must be 2, 3 or 4");
        if (n < 0 || n > CHX_STATIC_MAX_CLUSTERS)
            throw EngineError("at most " + std::to_string(CHX_STATIC_MAX_CLUSTERS) +
                              " clusters per call of the allocation-free entry point");
        if (n == 0) return 0;
        if ((nb == 3 && (m->d.order[1] == 0 || !m->dev.C3)) || (nb == 4 && (m->d.order[2] == 0 || !m->dev.C4)))
            throw EngineError("model has no terms of this body order");
        char *base = (char *)scratch_ptr(e->device);
        if (!base) throw EngineError("the device scratch array has no address on this device");
        const int P = nb * (nb - 1) / 2;
        hipStream_t st = e->st;
        // scratch layout (each part 256-byte aligned): distances, pair vectors,
        // types, energy, forces, stress, last-term scalars, flags
        size_t off = 0;
        auto take = [&](size_t bytes) {
            char *p = base + off;
            off += (bytes + 255) & ~(size_t)255;
            return p;
        };
        double *d_dist = (double *)take(sizeof(double) * n * P);
        double *d_dr = (double *)take(sizeof(double) * n * P * 3);
        int *d_types = (int *)take(sizeof(int) * n * nb);
        double *oE = (double *)take(sizeof(double) * n);
        double *oF = (double *)take(sizeof(double) * n * nb * 3);
        double *oS = (double *)take(sizeof(double) * n * 6);
        double *oL = (double *)take(sizeof(double) * n * P);
        int *oflags = (int *)take(sizeof(int) * n);
        if (off > CHX_SCRATCH_BYTES) throw EngineError("scratch layout exceeds the scratch array");
        HIPCHK(hipMemcpyAsync(d_dist, dist, sizeof(double) * n * P, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(d_dr, dr, sizeof(double) * n * P * 3, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(d_types, types, sizeof(int) * n * nb, hipMemcpyHostToDevice, st));
        // the last-term scalars keep their value for a cluster that is not
        // evaluated: they start at zero, as chimes_gpu_compute_clusters's do
        hipLaunchKernelGGL(k_fill_f64, dim3(nblk((long)n * P, 256)), dim3(256), 0, st, oL, (long)n * P);
        launch_clusters(nb, nb == 3 ? m->d.D3 : (nb == 4 ? m->d.D4 : 2), n, st, d_dist, d_dr, d_types, m->dev, oE,
                        oF, oS, oL, oflags);
        HIPCHK(hipMemcpyAsync(energy, oE, sizeof(double) * n, hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(force, oF, sizeof(double) * n * nb * 3, hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(stress, oS, sizeof(double) * n * 6, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        HIPCHK(hipGetLastError());
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

// ---------------------------------------------------------------------------
// LAMMPS pair style (etc/lmp/src/pair_chimes.cpp) on one process's local and
// ghost atoms.  The neighbour rows (LAMMPS's firstneigh) are packed on the
// host after every index has been checked, so no kernel reads outside its
// arrays; the many-body lists stay on the device between rebuilds.
// ---------------------------------------------------------------------------
struct chx_lmp {
    chx_model *m = nullptr;
    int device = 0;
    hipStream_t st = nullptr;
    hipEvent_t ev[5];
    std::string err;
    std::vector<int> chimes_type;
    std::vector<int> off_h, nbr_h, ctype_h, cnt_h;
    std::vector<long long> scan_h;
    std::vector<double> out_h;
    DevBuf<double> x, out;
    DevBuf<int> ctype, tag, ilist, nbr_off, nbr, cnt, trips, quads, status, fp_cnt;
    DevBuf<long long> off, fp_off;
    DevBuf<double> fp_rows;
    std::vector<int> fp_cnt_h;
    std::vector<long long> fp_off_h;
    int types_nall = -1;  // atoms whose ChIMES types are on the device
    DevBuf<unsigned long long> acc;
    // tables of the tabulated 2-/3-body path (chx_lmp_set_tables)
    DevBuf<char> tab_buf;
    chx_tabs_dev tabs{};
    long n3 = 0, n4 = 0;
    bool have_lists = false;
    int nall = 0, inum = 0;
    int lists_nall = -1, lists_inum = -1;  // shape the lists were built for
    double ms[4] = {0, 0, 0, 0};
    ~chx_lmp()
    {
        if (st) {
            (void)hipStreamSynchronize(st);
            for (hipEvent_t &e : ev) (void)hipEventDestroy(e);
            (void)hipStreamDestroy(st);
        }
    }
    chx_lmp_args args() const
    {
        chx_lmp_args A{};
        A.inum = inum;
        A.nall = nall;
        A.x = x.p;
        A.ctype = ctype.p;
        A.tag = tag.p;
        A.ilist = ilist.p;
        A.nbr_off = nbr_off.p;
        A.nbr = nbr.p;
        A.status = status.p;
        return A;
    }
};

namespace {

// LAMMPS keeps special-bond bits above NEIGHMASK in neighbour indices
constexpr int LMP_NEIGHMASK = 0x3FFFFFFF;

void lmp_upload(chx_lmp *p, int nall, const double *x, const int *type, const int *tag, int inum, const int *ilist,
                const int *numneigh, const int *const *firstneigh)
{
    if (nall <= 0 || inum < 0 || inum > nall) throw EngineError("atom counts out of range");
    if (!x || !tag || (inum > 0 && (!ilist || !numneigh || !firstneigh))) throw EngineError("missing input array");
    // row offsets first (checking owners and counts), then the rows
    p->off_h.assign((size_t)inum + 1, 0);
    long total = 0;
    for (int ii = 0; ii < inum; ii++) {
        const int i = ilist[ii];
        if (i < 0 || i >= nall) throw EngineError("ilist entry out of range");
        const int n = numneigh[i];
        if (n < 0) throw EngineError("negative neighbour count");
        if (n > 0 && !firstneigh[i]) throw EngineError("missing neighbour row");
        total += n;
        if (total >= (1L << 31)) throw EngineError("too many neighbour entries");
        p->off_h[ii + 1] = (int)total;
    }
    p->nbr_h.resize((size_t)total);
    int *dst = p->nbr_h.data();
    bool bad_index = false;
    for (int ii = 0; ii < inum; ii++) {
        const int *row = firstneigh[ilist[ii]];
        const int n = p->off_h[ii + 1] - p->off_h[ii];
        for (int q = 0; q < n; q++) {
            const int j = row[q] & LMP_NEIGHMASK;
            bad_index |= j >= nall;
            *dst++ = j;
        }
    }
    if (bad_index) throw EngineError("neighbour index out of range");
    hipStream_t st = p->st;
    p->nall = nall;
    p->inum = inum;
    p->x.ensure((size_t)nall * 3);
    p->tag.ensure(nall);
    p->ilist.ensure(std::max(inum, 1));
    p->nbr_off.ensure((size_t)inum + 1);
    p->nbr.ensure(std::max<size_t>(p->nbr_h.size(), 1));
    HIPCHK(hipMemcpyAsync(p->x.p, x, sizeof(double) * nall * 3, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(p->tag.p, tag, sizeof(int) * nall, hipMemcpyHostToDevice, st));
    if (inum > 0) HIPCHK(hipMemcpyAsync(p->ilist.p, ilist, sizeof(int) * inum, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(p->nbr_off.p, p->off_h.data(), sizeof(int) * (inum + 1), hipMemcpyHostToDevice, st));
    if (!p->nbr_h.empty())
        HIPCHK(hipMemcpyAsync(p->nbr.p, p->nbr_h.data(), sizeof(int) * p->nbr_h.size(), hipMemcpyHostToDevice, st));
    p->ctype.ensure(nall);
    if (type) {
        const int nt = (int)p->chimes_type.size();
        p->ctype_h.resize(nall);
        for (int a = 0; a < nall; a++) {
            if (type[a] < 1 || type[a] > nt) throw EngineError("atom type out of range");
            p->ctype_h[a] = p->chimes_type[type[a] - 1];
        }
        HIPCHK(hipMemcpyAsync(p->ctype.p, p->ctype_h.data(), sizeof(int) * nall, hipMemcpyHostToDevice, st));
        p->types_nall = nall;
    } else if (p->types_nall != nall) {
        p->types_nall = -1;
    }
    p->status.ensure(CHX_LMP_ST_WORDS);
    HIPCHK(hipMemsetAsync(p->status.p, 0, sizeof(int) * CHX_LMP_ST_WORDS, st));
}

void lmp_check_status(chx_lmp *p, const int *s)
{
    if (s[CHX_LMP_ST_FLAGS] & CHX_LMP_BAD_CAND)
        throw EngineError("an owner has " + std::to_string(s[CHX_LMP_ST_MAXCAND]) +
                          " neighbours within the padded many-body cutoffs (at most " +
                          std::to_string(CHX_LMP_CAND) + ")");
    if (s[CHX_LMP_ST_FLAGS] & CHX_LMP_BAD_RANGE) throw EngineError("a contribution is out of the accumulator range");
    if (s[CHX_LMP_ST_FLAGS] & CHX_LMP_BAD_TAB2)
        throw EngineError("a pair distance lies at or below the second point of its 2-body table");
    if (s[CHX_LMP_ST_FLAGS] & CHX_LMP_BAD_TAB3)
        throw EngineError("a triplet lies outside its 3-body table (or its type has no table)");
}

// build_mb_neighlists (pair_chimes.cpp:321-478) on the uploaded arrays
void lmp_build(chx_lmp *p, double skin)
{
    const chx_model *m = p->m;
    p->n3 = p->n4 = 0;
    p->have_lists = false;  // until this build completes
    p->lists_nall = p->nall;
    p->lists_inum = p->inum;
    const int inum = p->inum;
    if ((m->d.order[1] == 0 && m->d.order[2] == 0) || inum == 0) {
        p->have_lists = true;
        return;
    }
    const double pad3 = m->rc[1] + skin, pad4 = m->rc[2] + skin;
    const int has4 = m->d.order[2] > 0;
    chx_lmp_args A = p->args();
    hipStream_t st = p->st;
    p->cnt.ensure((size_t)2 * inum + CHX_LMP_ST_WORDS);
    launch_lmp_lists(0, A, pad3, pad4, has4, p->cnt.p, p->cnt.p + inum, nullptr, nullptr, nullptr, nullptr, st);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(p->cnt.p + 2 * inum, p->status.p, sizeof(int) * CHX_LMP_ST_WORDS, hipMemcpyDeviceToDevice,
                          st));
    p->cnt_h.resize((size_t)2 * inum + CHX_LMP_ST_WORDS);
    HIPCHK(hipMemcpyAsync(p->cnt_h.data(), p->cnt.p, sizeof(int) * p->cnt_h.size(), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    lmp_check_status(p, p->cnt_h.data() + 2 * inum);
    p->scan_h.resize((size_t)2 * inum);
    long long t3 = 0, t4 = 0;
    for (int ii = 0; ii < inum; ii++) {
        p->scan_h[ii] = t3;
        p->scan_h[inum + ii] = t4;
        t3 += p->cnt_h[ii];
        t4 += p->cnt_h[inum + ii];
    }
    if (t3 * 3 >= (1LL << 31) || t4 * 4 >= (1LL << 31)) throw EngineError("many-body lists too long");
    p->n3 = (long)t3;
    p->n4 = (long)t4;
    if (t3 + t4 == 0) {
        p->have_lists = true;
        return;
    }
    p->off.ensure((size_t)2 * inum);
    p->trips.ensure(std::max<size_t>((size_t)t3 * 3, 3));
    p->quads.ensure(std::max<size_t>((size_t)t4 * 4, 4));
    HIPCHK(hipMemcpyAsync(p->off.p, p->scan_h.data(), sizeof(long long) * 2 * inum, hipMemcpyHostToDevice, st));
    launch_lmp_lists(1, A, pad3, pad4, has4, nullptr, nullptr, p->off.p, p->off.p + inum, p->trips.p, p->quads.p,
                     st);
    HIPCHK(hipGetLastError());
    p->have_lists = true;
}

}  // namespace

extern "C" {

chx_lmp *chx_lmp_create(chx_model *m, int ntypes, const int *chimes_type, char *err, int errlen)
{
    chx_lmp *p = new chx_lmp();
    try {
        if (!m) throw EngineError("no model");
        if (ntypes <= 0 || !chimes_type) throw EngineError("no atom types");
        for (int t = 0; t < ntypes; t++)
            if (chimes_type[t] < 0 || chimes_type[t] >= m->d.nt) throw EngineError("ChIMES type index out of range");
        p->m = m;
        p->device = m->eng->device;
        p->chimes_type.assign(chimes_type, chimes_type + ntypes);
        HIPCHK(hipSetDevice(p->device));
        HIPCHK(hipStreamCreateWithFlags(&p->st, hipStreamNonBlocking));
        for (hipEvent_t &e : p->ev) HIPCHK(hipEventCreate(&e));
        return p;
    } catch (const EngineError &x) {
        if (err && errlen > 0) snprintf(err, errlen, "%s", x.msg.c_str());
        delete p;
        return nullptr;
    }
}

void chx_lmp_destroy(chx_lmp *p)
{
    if (!p) return;
    (void)hipSetDevice(p->device);
    delete p;
}

const char *chx_lmp_error(chx_lmp *p) { return p ? p->err.c_str() : "no handle"; }

extern "C++" int chx_lmp_set_tables(chx_lmp *p, const TableSet *T)
{
    try {
        HIPCHK(hipSetDevice(p->device));
        HIPCHK(hipStreamSynchronize(p->st));
        p->tabs = chx_tabs_dev{};
        p->tab_buf.release();
        if (!T || (T->t2.empty() && T->t3.empty())) return 0;
        const chx_model *m = p->m;
        if (!T->t2.empty() && (int)T->t2.size() != m->d.npairs) throw EngineError("one 2-body table per pair type");
        // one device buffer: offsets and descriptors, then the columns
        std::vector<char> h;
        auto put = [&](const void *src, size_t bytes) {
            const size_t off = align256(h.size());
            h.resize(off + std::max<size_t>(bytes, 1));
            if (bytes) memcpy(h.data() + off, src, bytes);
            return off;
        };
        std::vector<int> off2{0}, n3;
        std::vector<double> r2, e2, f2, dr3, rij3, rik3, rjk3, v3;
        std::vector<long> o3, o3ik, o3jk;
        for (const Table2 &t : T->t2) {
            r2.insert(r2.end(), t.r.begin(), t.r.end());
            e2.insert(e2.end(), t.e.begin(), t.e.end());
            f2.insert(f2.end(), t.f.begin(), t.f.end());
            off2.push_back((int)r2.size());
        }
        for (const Table3 &t : T->t3) {
            const size_t n = (size_t)t.n;
            n3.push_back(t.n);
            dr3.push_back(t.dr);
            o3.push_back((long)rij3.size());
            o3ik.push_back((long)rik3.size());
            o3jk.push_back((long)rjk3.size());
            rij3.insert(rij3.end(), t.rij.begin(), t.rij.begin() + n * n * n);
            rik3.insert(rik3.end(), t.rik.begin(), t.rik.begin() + n * n);
            rjk3.insert(rjk3.end(), t.rjk.begin(), t.rjk.begin() + n);
            v3.insert(v3.end(), t.v.begin(), t.v.begin() + 4 * n * n * n);
        }
        const size_t a_off2 = put(off2.data(), off2.size() * sizeof(int));
        const size_t a_r2 = put(r2.data(), r2.size() * 8), a_e2 = put(e2.data(), e2.size() * 8),
                     a_f2 = put(f2.data(), f2.size() * 8);
        const size_t a_n3 = put(n3.data(), n3.size() * sizeof(int)), a_dr3 = put(dr3.data(), dr3.size() * 8);
        const size_t a_o3 = put(o3.data(), o3.size() * sizeof(long)), a_o3ik = put(o3ik.data(), o3ik.size() * sizeof(long)),
                     a_o3jk = put(o3jk.data(), o3jk.size() * sizeof(long));
        const size_t a_rij = put(rij3.data(), rij3.size() * 8), a_rik = put(rik3.data(), rik3.size() * 8),
                     a_rjk = put(rjk3.data(), rjk3.size() * 8), a_v3 = put(v3.data(), v3.size() * 8);
        p->tab_buf.ensure(h.size());
        HIPCHK(hipMemcpy(p->tab_buf.p, h.data(), h.size(), hipMemcpyHostToDevice));
        char *d = p->tab_buf.p;
        chx_tabs_dev &D = p->tabs;
        D.ntab2 = (int)T->t2.size();
        D.ntab3 = (int)T->t3.size();
        D.off2 = (const int *)(d + a_off2);
        D.r2 = (const double *)(d + a_r2);
        D.e2 = (const double *)(d + a_e2);
        D.f2 = (const double *)(d + a_f2);
        D.n3 = (const int *)(d + a_n3);
        D.dr3 = (const double *)(d + a_dr3);
        D.o3 = (const long *)(d + a_o3);
        D.o3ik = (const long *)(d + a_o3ik);
        D.o3jk = (const long *)(d + a_o3jk);
        D.rij3 = (const double *)(d + a_rij);
        D.rik3 = (const double *)(d + a_rik);
        D.rjk3 = (const double *)(d + a_rjk);
        D.v3 = (const double *)(d + a_v3);
        return 0;
    } catch (const EngineError &e) {
        p->err = e.msg;
        p->tabs = chx_tabs_dev{};
        return 1;
    }
}

int chx_lmp_build_lists(chx_lmp *p, int nall, const double *x, const int *tag, int inum, const int *ilist,
                        const int *numneigh, const int *const *firstneigh, double skin)
{
    try {
        HIPCHK(hipSetDevice(p->device));
        lmp_upload(p, nall, x, nullptr, tag, inum, ilist, numneigh, firstneigh);
        lmp_build(p, skin);
        HIPCHK(hipStreamSynchronize(p->st));
        return 0;
    } catch (const EngineError &e) {
        p->err = e.msg;
        p->have_lists = false;
        return 1;
    }
}

int chx_lmp_compute(chx_lmp *p, int nall, const double *x, const int *type, const int *tag, int inum,
                    const int *ilist, const int *numneigh, const int *const *firstneigh, int ago, double skin,
                    int eflag_global, int eflag_atom, int vflag_global, int vflag_atom, double *f, double *eng,
                    double *virial, double *eatom, double *vatom)
{
    try {
        HIPCHK(hipSetDevice(p->device));
        if (!type || !f) throw EngineError("missing input array");
        if ((eflag_global && !eng) || (vflag_global && !virial) || (eflag_atom && !eatom) ||
            (vflag_atom && !vatom))
            throw EngineError("missing tally array");
        if (nall == 0 && inum == 0) return 0;  // a process without atoms: the loops do nothing
        hipStream_t st = p->st;
        HIPCHK(hipEventRecord(p->ev[0], st));
        // Between rebuilds (ago > 0) LAMMPS keeps its neighbour list, tags,
        // types and ilist unchanged, so only the positions are sent again.
        const bool rebuild = ago == 0 || !p->have_lists || p->lists_inum != inum || p->lists_nall != nall;
        if (rebuild) {
            lmp_upload(p, nall, x, type, tag, inum, ilist, numneigh, firstneigh);
        } else {
            if (!x) throw EngineError("missing input array");
            HIPCHK(hipMemcpyAsync(p->x.p, x, sizeof(double) * nall * 3, hipMemcpyHostToDevice, st));
            HIPCHK(hipMemsetAsync(p->status.p, 0, sizeof(int) * CHX_LMP_ST_WORDS, st));
        }
        HIPCHK(hipEventRecord(p->ev[1], st));
        if (rebuild) lmp_build(p, skin);
        HIPCHK(hipEventRecord(p->ev[2], st));
        const chx_model *m = p->m;
        // accumulators: forces, per-atom energy, per-atom virial, per-owner global
        const size_t nF = (size_t)nall * 6, nE = eflag_atom ? (size_t)nall * 2 : 0,
                     nV = vflag_atom ? (size_t)nall * 12 : 0,
                     nG = (eflag_global || vflag_global) ? (size_t)nall * 14 : 0;
        // plus the global reduction's sums (14 words) and block counter (1)
        p->acc.ensure(nF + nE + nV + nG + 15);
        HIPCHK(hipMemsetAsync(p->acc.p, 0, sizeof(unsigned long long) * (nF + nE + nV + nG + 15), st));
        chx_lmp_args A = p->args();
        A.Facc = p->acc.p;
        A.Eacc = nE ? A.Facc + nF : nullptr;
        A.Vacc = nV ? A.Facc + nF + nE : nullptr;
        A.Gacc = nG ? A.Facc + nF + nE + nV : nullptr;
        A.eflag_global = eflag_global;
        A.vflag_global = vflag_global;
        launch_lmp_pair(A, m->dev, &p->tabs, st);
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(p->ev[3], st));
        if (m->d.order[1] > 0 && p->n3 > 0)
            launch_lmp_clusters(3, m->d.D3, p->n3, p->trips.p, A, m->dev, &p->tabs, st);
        if (m->d.order[2] > 0 && p->n4 > 0)
            launch_lmp_clusters(4, m->d.D4, p->n4, p->quads.p, A, m->dev, &p->tabs, st);
        HIPCHK(hipGetLastError());
        const size_t nout = (size_t)nall * 3 + (nE ? nall : 0) + (nV ? (size_t)nall * 6 : 0) + 7;
        p->out.ensure(nout + CHX_LMP_ST_WORDS);
        double *F = p->out.p, *Ea = F + (size_t)nall * 3, *Va = Ea + (nE ? nall : 0),
               *G = Va + (nV ? (size_t)nall * 6 : 0);
        unsigned long long *gsum = p->acc.p + nF + nE + nV + nG;
        launch_lmp_out(A, F, nE ? Ea : nullptr, nV ? Va : nullptr, G, gsum, (int *)(gsum + 14), st);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(p->out.p + nout, p->status.p, sizeof(int) * CHX_LMP_ST_WORDS, hipMemcpyDeviceToDevice,
                              st));
        HIPCHK(hipEventRecord(p->ev[4], st));
        p->out_h.resize(nout + CHX_LMP_ST_WORDS);
        HIPCHK(hipMemcpyAsync(p->out_h.data(), p->out.p, sizeof(double) * p->out_h.size(), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        int s[CHX_LMP_ST_WORDS];
        memcpy(s, p->out_h.data() + nout, sizeof(s));
        lmp_check_status(p, s);
        const double *o = p->out_h.data();
        for (size_t k = 0; k < (size_t)nall * 3; k++) f[k] += o[k];
        o += (size_t)nall * 3;
        if (nE) {
            for (int a = 0; a < nall; a++) eatom[a] += o[a];
            o += nall;
        }
        if (nV) {
            for (size_t k = 0; k < (size_t)nall * 6; k++) vatom[k] += o[k];
            o += (size_t)nall * 6;
        }
        if (eflag_global) *eng += o[0];
        if (vflag_global)
            for (int k = 0; k < 6; k++) virial[k] += o[1 + k];
        float t[4];
        for (int k = 0; k < 4; k++) HIPCHK(hipEventElapsedTime(&t[k], p->ev[k], p->ev[k + 1]));
        for (int k = 0; k < 4; k++) p->ms[k] = t[k];
        return 0;
    } catch (const EngineError &e) {
        p->err = e.msg;
        return 1;
    }
}

int chx_lmp_fp_rows(chx_lmp *p, int nb, long *nrows, double **rows)
{
    *nrows = 0;
    *rows = nullptr;
    try {
        HIPCHK(hipSetDevice(p->device));
        if (nb < 2 || nb > 4) throw EngineError("cluster size must be 2, 3 or 4");
        if (p->types_nall != p->nall || p->nall <= 0)
            throw EngineError("no compute has run on these atoms: rows follow the last compute");
        const chx_model *m = p->m;
        if ((nb == 3 && m->d.order[1] == 0) || (nb == 4 && m->d.order[2] == 0)) return 0;
        if (nb > 2 && (!p->have_lists || p->lists_nall != p->nall || p->lists_inum != p->inum))
            throw EngineError("no many-body lists for these atoms");
        const long n = nb == 2 ? (long)p->off_h[p->inum] : (nb == 3 ? p->n3 : p->n4);
        if (n == 0) return 0;
        const int *list = nb == 3 ? p->trips.p : (nb == 4 ? p->quads.p : nullptr);
        const int w = nb * (nb - 1) / 2 + nb;
        const long nb_blk = (n + 255) / 256;
        hipStream_t st = p->st;
        const chx_lmp_args A = p->args();
        p->fp_cnt.ensure(nb_blk);
        p->fp_off.ensure(nb_blk);
        launch_lmp_fp_rows(nb, 0, n, list, A, m->dev, p->fp_cnt.p, nullptr, nullptr, st);
        HIPCHK(hipGetLastError());
        p->fp_cnt_h.resize(nb_blk);
        p->fp_off_h.resize(nb_blk);
        HIPCHK(hipMemcpyAsync(p->fp_cnt_h.data(), p->fp_cnt.p, sizeof(int) * nb_blk, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        long long total = 0;
        for (long b = 0; b < nb_blk; b++) {
            p->fp_off_h[b] = total;
            total += p->fp_cnt_h[b];
        }
        if (total == 0) return 0;
        double *out = (double *)malloc(sizeof(double) * total * w);
        if (!out) throw EngineError("out of host memory for the rows");
        p->fp_rows.ensure((size_t)total * w);
        HIPCHK(hipMemcpyAsync(p->fp_off.p, p->fp_off_h.data(), sizeof(long long) * nb_blk, hipMemcpyHostToDevice,
                              st));
        launch_lmp_fp_rows(nb, 1, n, list, A, m->dev, nullptr, p->fp_off.p, p->fp_rows.p, st);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(out, p->fp_rows.p, sizeof(double) * total * w, hipMemcpyDeviceToHost, st));
        if (hipStreamSynchronize(st) != hipSuccess) {
            free(out);
            throw EngineError("row read-back failed");
        }
        *nrows = (long)total;
        *rows = out;
        return 0;
    } catch (const EngineError &x) {
        p->err = x.msg;
        return 1;
    }
}

int chx_lmp_lists(chx_lmp *p, long *n3, int **trips, long *n4, int **quads)
{
    try {
        HIPCHK(hipSetDevice(p->device));
        if (!p->have_lists) throw EngineError("no many-body lists built");
        *n3 = p->n3;
        *n4 = p->n4;
        *trips = (int *)malloc(sizeof(int) * std::max<long>(3 * p->n3, 1));
        *quads = (int *)malloc(sizeof(int) * std::max<long>(4 * p->n4, 1));
        if (p->n3) HIPCHK(hipMemcpy(*trips, p->trips.p, sizeof(int) * 3 * p->n3, hipMemcpyDeviceToHost));
        if (p->n4) HIPCHK(hipMemcpy(*quads, p->quads.p, sizeof(int) * 4 * p->n4, hipMemcpyDeviceToHost));
        return 0;
    } catch (const EngineError &e) {
        p->err = e.msg;
        return 1;
    }
}

void chx_lmp_timing(chx_lmp *p, double *ms4)
{
    for (int k = 0; k < 4; k++) ms4[k] = p->ms[k];
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Cluster fingerprint histogram (chimesFF/src/FP/calc_histograms.cpp:351-489)
// ---------------------------------------------------------------------------
extern "C" int chx_fp_edges(chx_engine *e, chx_model *m, int nb, long n, const double *dist, const int *types,
                            double *edges)
{
    try {
        HIPCHK(hipSetDevice(e->device));
        if (nb < 2 || nb > 4 || n < 0) throw EngineError("cluster size must be 2, 3 or 4");
        if ((nb == 3 && m->d.order[1] == 0) || (nb == 4 && m->d.order[2] == 0))
            throw EngineError("model has no terms of this body order");
        if (n == 0) return 0;
        const int P = nb * (nb - 1) / 2;
        hipStream_t st = e->st;
        e->cl_in.ensure((size_t)n * P * 2);
        e->cl_types.ensure((size_t)n * nb);
        double *d_dist = e->cl_in.p, *d_edges = e->cl_in.p + (size_t)n * P;
        HIPCHK(hipMemcpyAsync(d_dist, dist, sizeof(double) * n * P, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(e->cl_types.p, types, sizeof(int) * n * nb, hipMemcpyHostToDevice, st));
        launch_fp_edges(nb, n, d_dist, e->cl_types.p, m->dev, d_edges, st);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(edges, d_edges, sizeof(double) * n * P, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        return 0;
    } catch (const EngineError &x) {
        e->err = x.msg;
        return 1;
    }
}

extern "C" int chx_fp_histogram(chx_engine *e, int np, long n1, const double *e1, long n2, const double *e2,
                                int same, int nbin, double binw, double maxd, long long *counts, double *ms)
{
    try {
        HIPCHK(hipSetDevice(e->device));
        if (np != 1 && np != 3 && np != 6) throw EngineError("clusters must have 1, 3 or 6 edges");
        if (nbin <= 0 || nbin > CHX_FP_MAX_BINS)
            throw EngineError("number of bins must be in 1.." + std::to_string(CHX_FP_MAX_BINS));
        if (!(binw > 0.0)) throw EngineError("bin width must be positive");
        if (n1 < 0 || n2 < 0 || (same && n1 != n2)) throw EngineError("cluster counts out of range");
        hipStream_t st = e->st;
        DevBuf<double> d1, d2;
        DevBuf<unsigned long long> cnt;
        d1.ensure(std::max<size_t>((size_t)n1 * np, 1));
        if (!same) d2.ensure(std::max<size_t>((size_t)n2 * np, 1));
        cnt.ensure((size_t)nbin + 1);
        const double *g2 = same ? d1.p : d2.p;
        if (n1) HIPCHK(hipMemcpyAsync(d1.p, e1, sizeof(double) * n1 * np, hipMemcpyHostToDevice, st));
        if (!same && n2) HIPCHK(hipMemcpyAsync(d2.p, e2, sizeof(double) * n2 * np, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemsetAsync(cnt.p, 0, sizeof(unsigned long long) * (nbin + 1), st));
        // events of its own: the engine's time the last evaluation
        struct Ev {
            hipEvent_t e = nullptr;
            ~Ev()
            {
                if (e) (void)hipEventDestroy(e);
            }
        } t0, t1;
        HIPCHK(hipEventCreate(&t0.e));
        HIPCHK(hipEventCreate(&t1.e));
        HIPCHK(hipEventRecord(t0.e, st));
        if (!launch_fp_hist(np, n1, d1.p, n2, g2, same, nbin, binw, maxd, cnt.p, (int *)(cnt.p + nbin), st))
            throw EngineError("cluster sets too large for one histogram launch");
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(t1.e, st));
        std::vector<unsigned long long> h((size_t)nbin + 1);
        HIPCHK(hipMemcpyAsync(h.data(), cnt.p, sizeof(unsigned long long) * (nbin + 1), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        if (h[nbin]) throw EngineError("a cluster distance falls outside the histogram");
        for (int b = 0; b < nbin; b++) counts[b] = (long long)h[b];
        float t = 0.f;
        HIPCHK(hipEventElapsedTime(&t, t0.e, t1.e));
        if (ms) *ms = t;
        return 0;
    } catch (const EngineError &x) {
        e->err = x.msg;
        return 1;
    }
}
