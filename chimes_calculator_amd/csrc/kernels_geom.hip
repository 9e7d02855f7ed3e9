// Geometry kernels: atom preparation (replication, wrapping, frames), cell
// lists, sorted neighbour (image) lists, 3-/4-body cluster enumeration into
// type-aligned list segments, and an integer scan.
//
// All arithmetic on positions reproduces the reference's operation order
// (serial_chimes_interface.h:141-208, serial_chimes_interface.cpp:150-341,
// 477-587) and this file is compiled with -ffp-contract=off, so cluster
// membership decided here is bit-identical to the reference's.
#include <hip/hip_runtime.h>

#include "chx_dev.h"
#include "chx_kernels.h"

namespace chx {

__device__ __forceinline__ void mv(const double *m, double x, double y, double z, double *o)
{
    o[0] = m[0] * x + m[1] * y + m[2] * z;
    o[1] = m[3] * x + m[4] * y + m[5] * z;
    o[2] = m[6] * x + m[7] * y + m[8] * z;
}

// Squared length of h * (b - a); its square root is the reference's get_dist
// on fractional coordinates.
__device__ __forceinline__ double frac_dist2(const double *h, const double *a, const double *b)
{
    const double dx = b[0] - a[0], dy = b[1] - a[1], dz = b[2] - a[2];
    double r[3];
    mv(h, dx, dy, dz, r);
    return r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
}

// The reference's test get_dist < c from the squared length s: decided by s
// wherever that is clear and by the square root only within a few ulps of
// c * c.  The square root is correctly rounded, so both give the same answer
// everywhere: a square below lo = c*c (1 - 2^-49) has a rounded root below c,
// one at or above hi = c*c (1 + 2^-49) a rounded root above it.
struct dist_cut {
    double c, lo, hi;
};

__device__ __forceinline__ dist_cut make_cut(double c)
{
    const double c2 = c * c;
    return {c, c2 * (1.0 - 0x1p-49), c2 * (1.0 + 0x1p-49)};
}

__device__ __forceinline__ bool within(double s, const dist_cut &k)
{
    if (s < k.lo) return true;
    if (s >= k.hi) return false;
    return sqrt(s) < k.c;
}

// Unordered pair p of n items as (u, v), u < v, v the larger index: v from a
// single-precision root (within one of the answer) and one correction each way.
__device__ __forceinline__ void pair_of(int p, int &u, int &v)
{
    v = (int)((1.0f + __builtin_amdgcn_sqrtf(1.0f + 8.0f * (float)p)) * 0.5f);
    while (v * (v - 1) / 2 > p) v--;
    while ((v + 1) * v / 2 <= p) v++;
    u = p - v * (v - 1) / 2;
}

__device__ __forceinline__ int find_cfg(const int *off, int n, int r)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= r) lo = mid; else hi = mid - 1;
    }
    return lo;
}

__global__ void k_fill_i32(int *p, long n, int v)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (long)gridDim.x * blockDim.x) p[i] = v;
}

__global__ void k_fill_f64(double *p, long n)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (long)gridDim.x * blockDim.x) p[i] = 0.0;
}

// ---------------------------------------------------------------------------
// Atom preparation: one thread per replicated atom.
// ---------------------------------------------------------------------------
__global__ void k_prep(int natoms, int ncfg, const int *__restrict__ cfg_atom_off, const chx_cfg *__restrict__ cfgs,
                       const double *__restrict__ pos, const int *__restrict__ types, chx_atom_s *as,
                       chx_atom_n *an, int *cell_count, const int *__restrict__ cfg_item_off, int *in_cfg,
                       int *item_cfg)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= natoms) return;
    const int c = find_cfg(cfg_atom_off, ncfg, r);
    const chx_cfg &C = cfgs[c];
    const int l = r - C.atom_off;
    const int rho = l / C.natoms_in;
    const int a = l - rho * C.natoms_in;
    const int ain = C.atom_in_off + a;
    // the configuration of each input atom and of each totals item (one
    // writer each: the first replica, the first owner of the item's chunk)
    if (rho == 0) in_cfg[ain] = c;
    if (l % CHX_FIN_CHUNK == 0) item_cfg[cfg_item_off[c] + l / CHX_FIN_CHUNK] = c;
    double x[3] = {pos[3 * (long)ain], pos[3 * (long)ain + 1], pos[3 * (long)ain + 2]};
    if (rho > 0) {  // replica (i, j, k) of serial_chimes_interface.cpp:243-283
        const int w = C.n_rep + 1;
        const int ri = rho / (w * w), rj = (rho / w) % w, rk = rho % w;
        double t[3];
        mv(C.ih0, x[0], x[1], x[2], t);
        t[0] += ri; t[1] += rj; t[2] += rk;
        mv(C.h0, t[0], t[1], t[2], x);
    }
    double t[3];
    mv(C.ih, x[0], x[1], x[2], t);  // wrap (serial_chimes_interface.cpp:301-320)
    t[0] -= floor(t[0]); t[1] -= floor(t[1]); t[2] -= floor(t[2]);
    double xs[3], us[3], xn[3], un[3];
    mv(C.h, t[0], t[1], t[2], xs);
    mv(C.ih, xs[0], xs[1], xs[2], us);
    mv(C.hn, us[0], us[1], us[2], xn);  // reorient (serial_chimes_interface.cpp:355-412)
    mv(C.ihn, xn[0], xn[1], xn[2], un);
    int cc[3];
    for (int d = 0; d < 3; d++) {
        int v = (int)floor(us[d] * C.m[d]);
        cc[d] = v < 0 ? 0 : (v >= C.m[d] ? C.m[d] - 1 : v);
    }
    const int cell = C.cell_off + cc[0] + C.m[0] * (cc[1] + C.m[1] * cc[2]);
    chx_atom_s s;
    s.u[0] = us[0]; s.u[1] = us[1]; s.u[2] = us[2];
    s.cfg = c;
    s.type = types[ain];
    as[r] = s;
    chx_atom_n n;
    n.u[0] = un[0]; n.u[1] = un[1]; n.u[2] = un[2];
    n.cell = cell;
    n.orig = ain;
    an[r] = n;
    atomicAdd(&cell_count[cell], 1);
}

__global__ void k_cell_fill(int natoms, const chx_atom_n *__restrict__ an, const int *__restrict__ cell_start,
                            int *cursor, int *cell_atoms)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= natoms) return;
    const int c = an[r].cell;
    const int p = atomicAdd(&cursor[c], 1);
    cell_atoms[cell_start[c] + p] = r;
}

// Image of atom a under lattice shift sh in the neighbour frame, and whether
// it lies within the search radius of the owner at ui
// (serial_chimes_interface.cpp:534-587, 700-726).
__device__ __forceinline__ bool nbr_image(const chx_cfg &C, const double *ui, const double *ua, const int *sh,
                                          bool zero, double *vn)
{
    if (zero) {
        vn[0] = ua[0]; vn[1] = ua[1]; vn[2] = ua[2];
    } else {
        double t[3] = {ua[0], ua[1], ua[2]};
        t[0] += sh[0]; t[1] += sh[1]; t[2] += sh[2];
        double xg[3];
        mv(C.hn, t[0], t[1], t[2], xg);
        mv(C.ihn, xg[0], xg[1], xg[2], vn);
    }
    return within(frac_dist2(C.hn, ui, vn), make_cut(C.search));
}

// The cheap first test of the same membership: the image's fractional offset
// taken as ua + sh - ui, without the round trip through Cartesian
// coordinates, and no square root.  C.near2 is the square of the search
// radius plus a bound on what the round trip and the rounding of both tests
// can move a distance by, so every image nbr_image accepts passes here;
// those that pass are confirmed by nbr_image.
__device__ __forceinline__ bool nbr_near(const chx_cfg &C, const double *ui, const double *ua, const int *sh)
{
    const double dx = (ua[0] + sh[0]) - ui[0], dy = (ua[1] + sh[1]) - ui[1], dz = (ua[2] + sh[2]) - ui[2];
    double r[3];
    mv(C.hn, dx, dy, dz, r);
    return r[0] * r[0] + r[1] * r[1] + r[2] * r[2] < C.near2;
}

// floor(u / m) and the remainder in [0, m), for m >= 1 and |u| < 2^20: a
// single-precision estimate (relative error below 2^-22, so off by at most
// one) and one correction, in place of two integer division sequences.
__device__ __forceinline__ int floor_div(int u, int m, int &r)
{
    int q = (int)floorf((float)u * __builtin_amdgcn_rcpf((float)m));
    r = u - q * m;
    if (r < 0) {
        r += m;
        q--;
    } else if (r >= m) {
        r -= m;
        q++;
    }
    return q;
}

// Cell coordinates of a cell of configuration C.
__device__ __forceinline__ void cell_coords(const chx_cfg &C, int cell, int (&ci)[3])
{
    const int y = floor_div(cell - C.cell_off, C.m[0], ci[0]);
    ci[2] = floor_div(y, C.m[1], ci[1]);
}

// Searched cell at offsets (dx, dy, dz) from the owner's cell ci: the wrapped
// cell, the image shift, and whether the shift lies within the ghost layers.
__device__ __forceinline__ bool searched_cell(const chx_cfg &C, const int (&ci)[3], int dx, int dy, int dz,
                                              int &cell, int (&sh)[3])
{
    const int u[3] = {ci[0] + dx, ci[1] + dy, ci[2] + dz};
    int wc[3];
    bool ok = true;
    for (int d = 0; d < 3; d++) {
        sh[d] = floor_div(u[d], C.m[d], wc[d]);
        if (sh[d] > C.n_layers || sh[d] < -C.n_layers) ok = false;
    }
    cell = C.cell_off + wc[0] + C.m[0] * (wc[1] + C.m[1] * wc[2]);
    return ok;
}

__device__ __forceinline__ int pack_shift(const int (&sh)[3])
{
    return (sh[0] + 128) | ((sh[1] + 128) << 8) | ((sh[2] + 128) << 16);
}

__device__ __forceinline__ void unpack_shift(int spk, int (&sh)[3])
{
    sh[0] = (spk & 255) - 128;
    sh[1] = ((spk >> 8) & 255) - 128;
    sh[2] = ((spk >> 16) & 255) - 128;
}

// The searched cells of one owner as one index range of candidates, laid out
// in LDS by one wavefront: per cell the inclusive prefix of the atom counts
// (cpre), the cell_atoms index of its first atom minus its first candidate
// index (cadj), and the packed image shift (cshf); per candidate below
// CHX_NBR_CANDS the cell it lies in (ccell).  Cells whose shift lies outside
// the ghost layers hold no candidates.  Returns the candidate count.
__device__ __forceinline__ int nbr_cells(const chx_cfg &C, const int (&ci)[3], const double *ui, int nc, int lane,
                                         const int *__restrict__ cell_start, int *cpre, int *cadj, int *cshf,
                                         unsigned char *ccell)
{
    const int w0 = 2 * C.reach[0] + 1, w1 = 2 * C.reach[1] + 1;
    const double im[3] = {1.0 / C.m[0], 1.0 / C.m[1], 1.0 / C.m[2]};
    int carry = 0;
    for (int c0 = 0; c0 < nc; c0 += 64) {
        const int c = c0 + lane;
        int cnt = 0, cs = 0, spk = 0;
        if (c < nc) {
            int r0, r1, cell, sh[3];
            const int c2 = floor_div(floor_div(c, w0, r0), w1, r1);
            const int d[3] = {r0 - C.reach[0], r1 - C.reach[1], c2 - C.reach[2]};
            // the cell's fractional box (unwrapped: its image) against the
            // owner: |h (u - u_i)|^2 >= lam_min |u - u_i|^2, so a cell whose
            // nearest point lies beyond the cheap test's radius holds no
            // image that passes it (1e-7 absorbs the rounding of the binning)
            double gap2 = 0.0;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const int cu = ci[k] + d[k];
                const double lo = cu * im[k], hi = (cu + 1) * im[k];
                const double g = fmax(0.0, fmax(lo - ui[k], ui[k] - hi) - 1e-7);
                gap2 += g * g;
            }
            if (!(gap2 * C.lam_min > C.near2) && searched_cell(C, ci, d[0], d[1], d[2], cell, sh)) {
                cs = cell_start[cell];
                cnt = cell_start[cell + 1] - cs;
                spk = pack_shift(sh);
            }
        }
        int v = cnt;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int t = __shfl_up(v, d, 64);
            if (lane >= d) v += t;
        }
        if (c < nc) {
            const int first = carry + v - cnt;
            cpre[c] = carry + v;
            cadj[c] = cs - first;
            cshf[c] = spk;
            for (int k = first; k < first + cnt && k < CHX_NBR_CANDS; k++) ccell[k] = (unsigned char)c;
        }
        carry += __shfl(v, 63, 64);
    }
    return carry;
}

// Candidate t of the layout above: its atom and packed shift.
__device__ __forceinline__ void nbr_cand(int t, int T, int nc, const int *cpre, const int *cadj, const int *cshf,
                                         const unsigned char *ccell, const int *__restrict__ cell_atoms, int &a,
                                         int &spk)
{
    int c;
    if (T <= CHX_NBR_CANDS) {
        c = ccell[t];
    } else {
        int lo = 0, hi = nc - 1;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (cpre[mid] > t) hi = mid; else lo = mid + 1;
        }
        c = lo;
    }
    a = cell_atoms[cadj[c] + t];
    spk = cshf[c];
}

// Key of a listed image: parent in the high bits, packed shift in the low 24.
// Images that fail the exact test are marked with the largest key.
constexpr long long NBR_DROPPED = 0x7fffffffffffffffll;

__device__ __forceinline__ bool nbr_confirm(const chx_cfg &C, const chx_atom_n &ni,
                                            const chx_atom_n *__restrict__ an, long long k)
{
    int sh[3];
    unpack_shift((int)(k & 0xffffff), sh);
    const bool zero = sh[0] == 0 && sh[1] == 0 && sh[2] == 0;
    double vn[3];
    return nbr_image(C, ni.u, an[(int)(k >> 24)].u, sh, zero, vn);
}

// Slab entry of image key k (both frames).
__device__ __forceinline__ void nbr_write(const chx_cfg &C, const chx_atom_n &ni, const chx_atom_s *__restrict__ as,
                                          const chx_atom_n *__restrict__ an, long long k, chx_entry_s *es,
                                          chx_entry_n *en)
{
    const int a = (int)(k >> 24);
    const int spack = (int)(k & 0xffffff);
    int sh[3];
    unpack_shift(spack, sh);
    const bool zero = sh[0] == 0 && sh[1] == 0 && sh[2] == 0;
    chx_entry_n f;
    nbr_image(C, ni.u, an[a].u, sh, zero, f.v);
    f.shift = spack;
    f.pad = 0;
    const chx_atom_s sa = as[a];
    chx_entry_s e;
    if (zero) {
        e.v[0] = sa.u[0]; e.v[1] = sa.u[1]; e.v[2] = sa.u[2];
    } else {
        double t[3] = {sa.u[0], sa.u[1], sa.u[2]};
        t[0] += sh[0]; t[1] += sh[1]; t[2] += sh[2];
        double xg[3];
        mv(C.h, t[0], t[1], t[2], xg);
        mv(C.ih, xg[0], xg[1], xg[2], e.v);
    }
    e.parent = a;
    e.type = sa.type;
    *es = e;
    *en = f;
}

// ---------------------------------------------------------------------------
// Neighbour lists: for real atom i every image g = (a, n) with i <= a,
// g != i and |r_g - r_i| < search in the neighbour frame
// (serial_chimes_interface.cpp:682-731).  One wavefront per owner, four
// owners per workgroup; lanes stride over the atoms of each searched cell.
// Candidates that pass the cheap test are collected as (parent, shift) keys
// in LDS, confirmed by the reference's arithmetic, ranked, and each owner's
// images written sorted by key into its slab, so every list and every later
// traversal is in one fixed order.
// ---------------------------------------------------------------------------
// Owner i's image count and slab: entries of owner i live at i * nb_cap
// (a capacity that grows on demand), so no count pass or scan is needed.
__device__ __forceinline__ void nbr_slab_note(int i, int n, int nb_cap, int *nbr_cnt, int *nbr_off, int *status)
{
    nbr_cnt[i] = n;
    nbr_off[i] = i * nb_cap;
    // the word only grows, so a count not above its current value needs no
    // atomic (one address for all owners)
    if (n > __atomic_load_n(&status[CHX_ST_MAXNB], __ATOMIC_RELAXED)) atomicMax(&status[CHX_ST_MAXNB], n);
    if (n > CHX_MAX_NB) atomicOr(&status[CHX_ST_FLAGS], CHX_BAD_NB);
    else if (n > nb_cap) atomicOr(&status[CHX_ST_FLAGS], CHX_RETRY_ENT);
}

__global__ __launch_bounds__(256) void k_nbr(int natoms, const chx_cfg *__restrict__ cfgs,
                                              const chx_atom_s *__restrict__ as, const chx_atom_n *__restrict__ an,
                                              const int *__restrict__ cell_start, const int *__restrict__ cell_atoms,
                                              int *nbr_cnt, int *nbr_off, chx_entry_s *es, chx_entry_n *en,
                                              int nb_cap, int *status)
{
    __shared__ long long keys[4][CHX_MAX_NB];
    __shared__ int cpre[4][CHX_NBR_CELLS], cadj[4][CHX_NBR_CELLS], cshf[4][CHX_NBR_CELLS];
    __shared__ unsigned char ccell[4][CHX_NBR_CANDS];
    if (block_stop(status, CHX_STOP)) return;
    // the wavefront index as a wave-uniform value: the owner's atom and frame
    // are then read with scalar loads into scalar registers
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int i = blockIdx.x * 4 + wv;
    const bool own = i < natoms;
    const unsigned long long below = lane ? (~0ull >> (64 - lane)) : 0ull;
    int n1 = 0;  // candidates that passed the cheap test
    chx_atom_n ni{};
    const chx_cfg &C = cfgs[own ? as[i].cfg : 0];
    int ci[3] = {0, 0, 0}, nc = 0, T = 0;
    // one candidate per lane: cheap test, keys of those that pass appended
    auto take = [&](bool valid, int a, int spk) {
        bool in = false;
        if (valid) {
            int sh[3];
            unpack_shift(spk, sh);
            const bool zero = sh[0] == 0 && sh[1] == 0 && sh[2] == 0;
            if (!(a < i || (a == i && zero))) in = nbr_near(C, ni.u, an[a].u, sh);
        }
        const unsigned long long b = __ballot(in);
        if (in) {
            const int idx = n1 + __popcll(b & below);
            if (idx < CHX_MAX_NB) keys[wv][idx] = ((long long)a << 24) | spk;
        }
        n1 += __popcll(b);
    };
    if (own) {
        ni = an[i];
        cell_coords(C, ni.cell, ci);
        nc = (2 * C.reach[0] + 1) * (2 * C.reach[1] + 1) * (2 * C.reach[2] + 1);
        if (nc <= CHX_NBR_CELLS) T = nbr_cells(C, ci, ni.u, nc, lane, cell_start, cpre[wv], cadj[wv], cshf[wv], ccell[wv]);
    }
    __syncthreads();
    if (own) {
        if (nc <= CHX_NBR_CELLS) {
            for (int t0 = 0; t0 < T; t0 += 64) {
                const int t = t0 + lane;
                int a = 0, spk = 0;
                if (t < T) nbr_cand(t, T, nc, cpre[wv], cadj[wv], cshf[wv], ccell[wv], cell_atoms, a, spk);
                take(t < T, a, spk);
            }
        } else {
            for (int dz = -C.reach[2]; dz <= C.reach[2]; dz++)
                for (int dy = -C.reach[1]; dy <= C.reach[1]; dy++)
                    for (int dx = -C.reach[0]; dx <= C.reach[0]; dx++) {
                        int cell, sh[3];
                        if (!searched_cell(C, ci, dx, dy, dz, cell, sh)) continue;
                        const int cs = cell_start[cell], ce = cell_start[cell + 1];
                        const int spk = pack_shift(sh);
                        for (int p0 = cs; p0 < ce; p0 += 64) {
                            const int p = p0 + lane;
                            take(p < ce, p < ce ? cell_atoms[p] : 0, spk);
                        }
                    }
        }
    }
    __syncthreads();
    // the reference's test on the keys that passed; failures are marked
    int n = n1;
    if (own && n1 <= CHX_MAX_NB) {
        n = 0;
        for (int j0 = 0; j0 < n1; j0 += 64) {
            const int j = j0 + lane;
            bool in = false;
            if (j < n1) {
                in = nbr_confirm(C, ni, an, keys[wv][j]);
                if (!in) keys[wv][j] = NBR_DROPPED;
            }
            n += __popcll(__ballot(in));
        }
    }
    if (own && lane == 0) nbr_slab_note(i, n, nb_cap, nbr_cnt, nbr_off, status);
    __syncthreads();
    if (!own || n > CHX_MAX_NB || n > nb_cap) return;
    const long off = (long)i * nb_cap;
    for (int j = lane; j < n1; j += 64) {
        const long long k = keys[wv][j];
        if (k == NBR_DROPPED) continue;
        int rk = 0;
        for (int q = 0; q < n1; q++) rk += keys[wv][q] < k;
        nbr_write(C, ni, as, an, k, es + off + rk, en + off + rk);
    }
}

// The same lists for small batches: one owner per workgroup, the searched
// candidates spread over its four wavefronts (latency-bound otherwise).
// Cells beyond the flattened range are searched by wavefront 0 alone.
__global__ __launch_bounds__(256) void k_nbr_wide(int natoms, const chx_cfg *__restrict__ cfgs,
                                                   const chx_atom_s *__restrict__ as,
                                                   const chx_atom_n *__restrict__ an,
                                                   const int *__restrict__ cell_start,
                                                   const int *__restrict__ cell_atoms, int *nbr_cnt, int *nbr_off,
                                                   chx_entry_s *es, chx_entry_n *en, int nb_cap, int *status)
{
    __shared__ long long keys[CHX_MAX_NB];
    __shared__ int cpre[CHX_NBR_CELLS], cadj[CHX_NBR_CELLS], cshf[CHX_NBR_CELLS];
    __shared__ unsigned char ccell[CHX_NBR_CANDS];
    __shared__ int ncount, nexact, ncand;
    if (block_stop(status, CHX_STOP)) return;
    const int i = blockIdx.x;
    if (i >= natoms) return;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const unsigned long long below = lane ? (~0ull >> (64 - lane)) : 0ull;
    const chx_atom_n ni = an[i];
    const chx_cfg &C = cfgs[as[i].cfg];
    int ci[3];
    cell_coords(C, ni.cell, ci);
    const int nc = (2 * C.reach[0] + 1) * (2 * C.reach[1] + 1) * (2 * C.reach[2] + 1);
    // candidates of this wavefront's lanes: cheap test, keys of those that
    // pass appended to the owner's key list
    auto take = [&](bool valid, int a, int spk) {
        bool in = false;
        if (valid) {
            int sh[3];
            unpack_shift(spk, sh);
            const bool zero = sh[0] == 0 && sh[1] == 0 && sh[2] == 0;
            if (!(a < i || (a == i && zero))) in = nbr_near(C, ni.u, an[a].u, sh);
        }
        const unsigned long long b = __ballot(in);
        if (b == 0) return;
        int base = 0;
        if (lane == 0) base = atomicAdd(&ncount, __popcll(b));
        base = __shfl(base, 0, 64);
        if (in) {
            const int idx = base + __popcll(b & below);
            if (idx < CHX_MAX_NB) keys[idx] = ((long long)a << 24) | spk;
        }
    };
    if (tid == 0) {
        ncount = 0;
        nexact = 0;
    }
    if (nc <= CHX_NBR_CELLS && wv == 0) {
        const int T = nbr_cells(C, ci, ni.u, nc, lane, cell_start, cpre, cadj, cshf, ccell);
        if (lane == 0) ncand = T;
    }
    __syncthreads();
    if (nc <= CHX_NBR_CELLS) {
        const int T = ncand;
        for (int t0 = wv * 64; t0 < T; t0 += 256) {
            const int t = t0 + lane;
            int a = 0, spk = 0;
            if (t < T) nbr_cand(t, T, nc, cpre, cadj, cshf, ccell, cell_atoms, a, spk);
            take(t < T, a, spk);
        }
    } else if (wv == 0) {
        for (int dz = -C.reach[2]; dz <= C.reach[2]; dz++)
            for (int dy = -C.reach[1]; dy <= C.reach[1]; dy++)
                for (int dx = -C.reach[0]; dx <= C.reach[0]; dx++) {
                    int cell, sh[3];
                    if (!searched_cell(C, ci, dx, dy, dz, cell, sh)) continue;
                    const int cs = cell_start[cell], ce = cell_start[cell + 1];
                    const int spk = pack_shift(sh);
                    for (int p0 = cs; p0 < ce; p0 += 64) {
                        const int p = p0 + lane;
                        take(p < ce, p < ce ? cell_atoms[p] : 0, spk);
                    }
                }
    }
    __syncthreads();
    // the reference's test on the keys that passed; failures are marked
    const int n1 = ncount;
    if (n1 <= CHX_MAX_NB) {
        for (int j0 = wv * 64; j0 < n1; j0 += 256) {
            const int j = j0 + lane;
            bool in = false;
            if (j < n1) {
                in = nbr_confirm(C, ni, an, keys[j]);
                if (!in) keys[j] = NBR_DROPPED;
            }
            const unsigned long long b = __ballot(in);
            if (lane == 0 && b) atomicAdd(&nexact, __popcll(b));
        }
    }
    __syncthreads();
    const int n = n1 <= CHX_MAX_NB ? nexact : n1;
    if (tid == 0) nbr_slab_note(i, n, nb_cap, nbr_cnt, nbr_off, status);
    if (n > CHX_MAX_NB || n > nb_cap) return;
    const long off = (long)i * nb_cap;
    for (int j = tid; j < n1; j += 256) {
        const long long k = keys[j];
        if (k == NBR_DROPPED) continue;
        int rk = 0;
        for (int q = 0; q < n1; q++) rk += keys[q] < k;
        nbr_write(C, ni, as, an, k, es + off + rk, en + off + rk);
    }
}

// Aligned segment of each (model, cluster type) group in the 3- and 4-body
// lists (thread 0: triplets, thread 1: quadruplets): base[g] = sum over h < g
// of the group sizes rounded up to CHX_LIST_ALIGN, end[g] = base[g] + size of
// g.  off is the exclusive scan of the per-(group, owner) counts laid out
// [G3 + G4][N]; a group's size is a difference of two scan values.
// Thread 0: triplet groups, thread 1: quadruplet groups.
__device__ __forceinline__ void group_bases_of(const int *__restrict__ off, long N, int G3, int G4, long cap3,
                                               long cap4, int *base3, int *end3, int *base4, int *end4, int *status,
                                               int t)
{
    const bool four = t == 1;
    const int G = four ? G4 : G3;
    if (G == 0) return;
    const int *o = four ? off + (long)G3 * N : off;
    int *base = four ? base4 : base3, *end = four ? end4 : end3;
    const long cap = four ? cap4 : cap3;
    const int word = four ? CHX_ST_N4 : CHX_ST_N3;
    status[four ? CHX_ST_C4 : CHX_ST_C3] = o[(long)G * N] - o[0];
    long b = 0;
    for (int g = 0; g < G; g++) {
        const long sz = (long)o[(long)(g + 1) * N] - o[(long)g * N];
        base[g] = (int)b;
        end[g] = (int)(b + sz);
        b += (sz + CHX_LIST_ALIGN - 1) / CHX_LIST_ALIGN * CHX_LIST_ALIGN;
    }
    base[G] = (int)b;
    status[word] = (int)b;
    if (b > cap || b >= (1L << 31) - CHX_LIST_ALIGN) atomicOr(&status[CHX_ST_FLAGS], four ? CHX_RETRY_L4 : CHX_RETRY_L3);
}

__global__ void k_group_bases(chx_bases_args a)
{
    if (blockIdx.x || threadIdx.x > 1) return;
    group_bases_of(a.off, a.N, a.G3, a.G4, a.cap3, a.cap4, a.base3, a.end3, a.base4, a.end4, a.status, threadIdx.x);
}

// ---------------------------------------------------------------------------
// Cluster enumeration over the sorted images, one 64-lane workgroup per owner:
// clusters are counted per (model, cluster type) group and each is recorded
// in the owner's record slab as (packed member slot codes, group | its slot
// among the owner's clusters of that group).  After the scan of the counts
// k_place moves every record to the owner's part of its group segment (in
// any order: every later sum is exact, and exported lists are sorted), so
// the enumeration runs once.  Predicates (serial_chimes_interface.cpp:749-857, with the max 4B
// cutoff <= the max 3B cutoff): j, k in N2(i), j != k, parent(j) <=
// parent(k), all three distances < rc3; quadruplets add l in N2(i), l not in
// {j, k}, parent(j), parent(k) <= parent(l), all six distances < rc4.
// Candidates are the images within rc3 (S3) of the owner; quadruplets use
// adjacency bit masks over the images within rc4 (S4).
//
// The LDS working set is sized for C3 / C4 candidates.  k_enum uses a small
// size (occupancy); an owner with more than CHX_ENUM_SMALL images, or more
// candidates than its working set holds, is queued and handled by k_enum_big
// with the full size.
// ---------------------------------------------------------------------------
template <int C3, int C4>
struct EnumShared {
    static constexpr int EW4 = C4 / 32;
    double vn3[C3][3];
    int idx3[C3];
    int par3[C3];
    int typ3[C3];
    int idx4[C4];      // S4 member -> S3 index
    int s4of[C3];      // S3 member -> S4 index, -1 outside S4
    int gst4[C4];
    unsigned A4[C4][EW4];
    int cnt[CHX_MAX_CLUSTER_TYPES + 1];
    int map[256];                         // type tuple -> group, for nt <= 4
    int wc3[4], wc4[4];                   // per-wavefront candidate counts of one compaction pass
    int n3, n4, nrec;
};

// One cluster of group gg (global index) at its slot: a record in the
// owner's slab; y = 4-body flag (sign bit) | gg << 20 | slot.
__device__ __forceinline__ void enum_record(const chx_enum_args &x, int i, int &nrec_sh, int code, bool four, long gg,
                                            int slot)
{
    const int r = atomicAdd(&nrec_sh, 1);
    if (r < x.rcap)
        x.rec[(long)i * x.rcap + r] = make_int2(code, (four ? (int)0x80000000 : 0) | ((int)gg << 20) | slot);
}

// Owner i to the large-LDS enumeration's queue (nothing of it is written
// yet); a retry when that kernel was not launched.
__device__ __forceinline__ void enum_queue_big(const chx_enum_args &x, int i)
{
    x.big[atomicAdd(&x.status[CHX_ST_NBIG], 1)] = i;
    if (!x.big_on) atomicOr(&x.status[CHX_ST_FLAGS], CHX_RETRY_BIG);
}

// Data that earlier kernels wrote and a launch only reads, at a
// workgroup-uniform index, read through the constant address space: scalar
// loads into scalar registers.
template <typename T>
__device__ __forceinline__ T uniform_load(const T *p, long k)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return ((const __attribute__((address_space(4))) T *)p)[k];
#else
    return p[k];  // the host pass never runs device code
#endif
}

template <int C3, int C4, int NT>
__device__ __forceinline__ void enum_owner(const chx_enum_args &x, int i, EnumShared<C3, C4> &sh)
{
    static_assert(NT == 64 || NT == 256, "one or four wavefronts per owner");
    constexpr int EW4 = EnumShared<C3, C4>::EW4;
    const int lane = threadIdx.x, wl = lane & 63, wv = lane >> 6;
    const int off = uniform_load(x.nbr_off, i);
    const int nb = uniform_load(x.nbr_cnt, i);
    const chx_atom_s ai = uniform_load(x.as, i);
    const chx_atom_n ni = uniform_load(x.an, i);
    const chx_cfg C = uniform_load(x.cfgs, ai.cfg);
    const chx_enum_model M = uniform_load(x.em, C.model);
    const long N = x.natoms;
    const double rc3 = C.rc3, rc4 = C.rc4;
    const bool any3 = M.has3 && rc3 > 0.0, any4 = M.has4 && rc4 > 0.0;
    const double rcS = rc3 > rc4 ? rc3 : rc4;  // many-body candidate radius
    if (!any3 && !any4) return;
    const dist_cut cutS = make_cut(rcS), cut3 = make_cut(rc3), cut4 = make_cut(rc4);
    if (lane == 0) { sh.n3 = 0; sh.n4 = 0; sh.nrec = 0; }
    __syncthreads();
    // compact the images within rc3 / rc4, in sorted order (positions: the
    // running count, the counts of the lower wavefronts, the lower lanes)
    for (int base = 0; base < nb; base += NT) {
        const int j = base + lane;
        bool in3 = false, in4 = false;
        chx_entry_n e;
        if (j < nb) {
            e = x.en[off + j];
            const double d2 = frac_dist2(C.hn, ni.u, e.v);
            in3 = within(d2, cutS);
            in4 = within(d2, cut4);
        }
        const unsigned long long b3 = __ballot(in3), b4 = __ballot(in4);
        const unsigned long long below = wl ? (~0ull >> (64 - wl)) : 0ull;
        int n3b = sh.n3, n4b = sh.n4, t3 = __popcll(b3), t4 = __popcll(b4);
        if constexpr (NT > 64) {
            if (wl == 0) { sh.wc3[wv] = t3; sh.wc4[wv] = t4; }
            __syncthreads();
            t3 = t4 = 0;
            for (int w = 0; w < NT / 64; w++) {
                if (w < wv) { n3b += sh.wc3[w]; n4b += sh.wc4[w]; }
                t3 += sh.wc3[w];
                t4 += sh.wc4[w];
            }
        }
        if (in3) {
            const int a = n3b + __popcll(b3 & below);
            if (a < C3) {
                sh.vn3[a][0] = e.v[0]; sh.vn3[a][1] = e.v[1]; sh.vn3[a][2] = e.v[2];
                sh.idx3[a] = j;
                const chx_entry_s f = x.es[off + j];
                sh.par3[a] = f.parent;
                sh.typ3[a] = f.type;
                if (in4) {
                    const int u = n4b + __popcll(b4 & below);
                    if (u < C4) sh.idx4[u] = a;
                }
            }
        }
        __syncthreads();
        if (lane == 0) { sh.n3 += t3; sh.n4 += t4; }
        __syncthreads();
    }
    const int n3 = sh.n3, n4 = sh.n4;
    if (n3 > C3 || n4 > C4) {
        if (lane == 0) {
            if constexpr (C3 < CHX_MAX_N3 || C4 < CHX_MAX_N4) enum_queue_big(x, i);
            else atomicOr(&x.status[CHX_ST_FLAGS], CHX_BAD_N34);
        }
        return;
    }
    for (int u = lane; u < n3; u += NT) sh.s4of[u] = -1;
    __syncthreads();
    for (int u = lane; u < n4; u += NT) {
        int g = u;
        const int p = sh.par3[sh.idx4[u]];
        while (g > 0 && sh.par3[sh.idx4[g - 1]] == p) g--;
        sh.gst4[u] = g;
        sh.s4of[sh.idx4[u]] = u;
        for (int w = 0; w < EW4; w++) sh.A4[u][w] = 0u;
    }
    for (int g = lane; g <= CHX_MAX_CLUSTER_TYPES; g += NT) sh.cnt[g] = 0;
    __syncthreads();
    // 4-body adjacency: in the 3-body pair pass below when there is one
    // (S4 lies within S3, and the distance is the same), else here
    if (any4 && !any3) {
        const int npair = n4 * (n4 - 1) / 2;
        for (int p = lane; p < npair; p += NT) {
            int u, v;
            pair_of(p, u, v);
            if (within(frac_dist2(C.hn, sh.vn3[sh.idx4[u]], sh.vn3[sh.idx4[v]]), cut4)) {
                atomicOr(&sh.A4[u][v >> 5], 1u << (v & 31));
                atomicOr(&sh.A4[v][u >> 5], 1u << (u & 31));
            }
        }
    }
    __syncthreads();
    const int nt = M.nt;
    const int ti = ai.type;
    // group of a type tuple from LDS when the map is small
    const bool lmap = nt <= 4;
    auto stage = [&](const int *map, int nmap) {
        if (lmap)
            for (int t = lane; t < nmap; t += NT) sh.map[t] = map[t];
        __syncthreads();
    };
    if (any3) {
        const int ngl = M.ntrip + 1;
        const long gbase = M.g3;
        stage(M.tripmap, nt * nt * nt);
        // every unordered pair u < v of S3 once, spread over the lanes: (j, k)
        // = (u, v) always, and (v, u) too when both are images of one atom
        // (parent(j) <= parent(k) in both orders); the distance is symmetric
        const int npair = n3 * (n3 - 1) / 2;
        for (int p = lane; p < npair; p += NT) {
            int u, v;
            pair_of(p, u, v);
            const double d2 = frac_dist2(C.hn, sh.vn3[u], sh.vn3[v]);
            if (any4) {
                const int u4 = sh.s4of[u], v4 = sh.s4of[v];
                if (u4 >= 0 && v4 >= 0 && within(d2, cut4)) {
                    atomicOr(&sh.A4[u4][v4 >> 5], 1u << (v4 & 31));
                    atomicOr(&sh.A4[v4][u4 >> 5], 1u << (u4 & 31));
                }
            }
            if (!within(d2, cut3)) continue;
            const int nrec = sh.par3[u] == sh.par3[v] ? 2 : 1;
            for (int q = 0; q < nrec; q++) {
                const int a = q ? v : u, b = q ? u : v;
                const int t = (ti * nt + sh.typ3[a]) * nt + sh.typ3[b];
                int g = lmap ? sh.map[t] : M.tripmap[t];
                if (g < 0) g = ngl - 1;
                const int slot = atomicAdd(&sh.cnt[g], 1);
                enum_record(x, i, sh.nrec, (sh.idx3[a] + 1) | ((sh.idx3[b] + 1) << 10), false, gbase + g, slot);
            }
        }
        __syncthreads();
        for (int g = lane; g < ngl; g += NT) {
            x.cnt3[(gbase + g) * N + i] = sh.cnt[g];
            sh.cnt[g] = 0;
        }
        __syncthreads();
    }
    if (any4) {
        const int ngl = M.nquad + 1;
        const long gbase = M.g4;
        const int W4 = (n4 + 31) >> 5;
        stage(M.quadmap, nt * nt * nt * nt);
        // every adjacent unordered pair u < v of S4 once, spread over the
        // lanes, as (j, k) = (u, v) and, for two images of one atom, also
        // (v, u); then l over the common neighbours with parent(k) <= parent(l)
        const int npair = n4 * (n4 - 1) / 2;
        for (int p = lane; p < npair; p += NT) {
            int u, v;
            pair_of(p, u, v);
            if (!((sh.A4[u][v >> 5] >> (v & 31)) & 1u)) continue;
            const int nrec = sh.par3[sh.idx4[u]] == sh.par3[sh.idx4[v]] ? 2 : 1;
            for (int q = 0; q < nrec; q++) {
                const int ua = q ? v : u, vb = q ? u : v;
                const int au = sh.idx4[ua], av = sh.idx4[vb];
                const int j = sh.idx3[au], k = sh.idx3[av];
                const int tjk = (ti * nt + sh.typ3[au]) * nt + sh.typ3[av];
                const int gv = sh.gst4[vb];
                for (int y = 0; y < W4; y++) {
                    unsigned ml = sh.A4[ua][y] & sh.A4[vb][y];
                    const int lo2 = y * 32;
                    if (gv > lo2) ml &= (gv - lo2 >= 32) ? 0u : (~0u << (gv - lo2));
                    if (ua >= lo2 && ua < lo2 + 32) ml &= ~(1u << (ua - lo2));
                    if (vb >= lo2 && vb < lo2 + 32) ml &= ~(1u << (vb - lo2));
                    while (ml) {
                        const int wl = lo2 + __ffs(ml) - 1;
                        ml &= ml - 1;
                        const int aw = sh.idx4[wl];
                        const int t = tjk * nt + sh.typ3[aw];
                        int g = lmap ? sh.map[t] : M.quadmap[t];
                        if (g < 0) g = ngl - 1;
                        const int slot = atomicAdd(&sh.cnt[g], 1);
                        enum_record(x, i, sh.nrec, (j + 1) | ((k + 1) << 10) | ((sh.idx3[aw] + 1) << 20), true,
                                    gbase + g, slot);
                    }
                }
            }
        }
        __syncthreads();
        for (int g = lane; g < ngl; g += NT) {
            x.cnt4[(gbase + g) * N + i] = sh.cnt[g];
            sh.cnt[g] = 0;
        }
        __syncthreads();
    }
    // the owner's record count; the largest one (a word that only grows, so a
    // count not above it needs no atomic) sizes the slabs of a retry
    if (lane == 0) {
        const int nr = sh.nrec;
        x.nrec[i] = nr;
        if (nr > __atomic_load_n(&x.status[CHX_ST_MAXREC], __ATOMIC_RELAXED)) atomicMax(&x.status[CHX_ST_MAXREC], nr);
        if (nr > x.rcap) atomicOr(&x.status[CHX_ST_FLAGS], CHX_RETRY_REC);
    }
}

template <int NT>
__global__ __launch_bounds__(NT) void k_enum(chx_enum_args x)
{
    // one wavefront per owner: room for CHX_ENUM_SMALL / 2 images within
    // each many-body cutoff (several times the usual count), so more owners
    // share a CU; an owner with more is queued once they are counted
    constexpr int CS = NT == 64 ? CHX_ENUM_SMALL / 2 : CHX_ENUM_SMALL;
    __shared__ EnumShared<CS, CS> sh;
    if (block_stop(x.status, CHX_STOP)) return;
    const int i = blockIdx.x;
    if (x.nbr_cnt[i] > CHX_ENUM_SMALL) {
        if (threadIdx.x == 0) enum_queue_big(x, i);
        return;
    }
    enum_owner<CS, CS, NT>(x, i, sh);
}

// Owners queued by k_enum, any number of them, over a fixed grid.
__global__ __launch_bounds__(64) void k_enum_big(chx_enum_args x)
{
    __shared__ EnumShared<CHX_MAX_N3, CHX_MAX_N4> sh;
    if (block_stop(x.status, CHX_STOP)) return;
    const int nbig = x.status[CHX_ST_NBIG];
    for (int t = blockIdx.x; t < nbig; t += gridDim.x) {
        enum_owner<CHX_MAX_N3, CHX_MAX_N4, 64>(x, x.big[t], sh);
        __syncthreads();
    }
}

// Records to list positions, one wavefront per owner: a record of group gg at
// its slot goes to the group's base + the owner's offset within the group
// (the scan of the counts) + the slot.  With few groups the owner's position
// in each group is computed once, into LDS.
constexpr int PLACE_GROUPS = 32;

__global__ __launch_bounds__(256) void k_place(chx_enum_args x)
{
    __shared__ int pos[4][2][PLACE_GROUPS];
    if (block_stop(x.status, CHX_STOP)) return;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int i = (int)(blockIdx.x * 4 + wv);
    const int lane = threadIdx.x & 63;
    const long N = x.natoms;
    const bool own = i < x.natoms;
    const bool table = x.G3 <= PLACE_GROUPS && x.G4 <= PLACE_GROUPS;
    if (own && table) {
        if (lane < x.G3) pos[wv][0][lane] = x.gb3[lane] + (x.off3[lane * N + i] - x.off3[lane * N]);
        else if (lane >= 32 && lane - 32 < x.G4)
            pos[wv][1][lane - 32] = x.gb4[lane - 32] + (x.off4[(lane - 32) * N + i] - x.off4[(lane - 32) * N]);
    }
    __syncthreads();
    if (!own) return;
    const int n = uniform_load(x.nrec, i);
    const int2 *rec = x.rec + (long)i * x.rcap;
    for (int r = lane; r < n; r += 64) {
        const int2 e = rec[r];
        const int gg = (e.y >> 20) & 0x7ff;
        const int slot = e.y & 0xfffff;
        const bool four = e.y < 0;
        int p;
        if (table) p = pos[wv][four][gg];
        else if (four) p = x.gb4[gg] + (x.off4[gg * N + i] - x.off4[gg * N]);
        else p = x.gb3[gg] + (x.off3[gg * N + i] - x.off3[gg * N]);
        (four ? x.list4 : x.list3)[p + slot] = make_int2(i, e.x);
    }
}

template __global__ void k_enum<64>(chx_enum_args);
template __global__ void k_enum<256>(chx_enum_args);

// ---------------------------------------------------------------------------
// Exclusive scan of int32 (out[n] = total), 2048 elements per block.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_scan_block(const int *in, int *out, long n, int *bsum)
{
    __shared__ int s[256];
    const long base = (long)blockIdx.x * 2048;
    int v[8], t = 0;
    for (int q = 0; q < 8; q++) {
        const long idx = base + threadIdx.x * 8 + q;
        v[q] = idx < n ? in[idx] : 0;
        t += v[q];
    }
    s[threadIdx.x] = t;
    __syncthreads();
    for (int d = 1; d < 256; d <<= 1) {
        int x = threadIdx.x >= d ? s[threadIdx.x - d] : 0;
        __syncthreads();
        s[threadIdx.x] += x;
        __syncthreads();
    }
    int run = s[threadIdx.x] - t;
    for (int q = 0; q < 8; q++) {
        const long idx = base + threadIdx.x * 8 + q;
        if (idx < n) out[idx] = run;
        run += v[q];
    }
    if (threadIdx.x == 255) bsum[blockIdx.x] = s[255];
}

// Exclusive scan of a short array (n <= CHX_SCAN1_MAX) in one workgroup of
// 1024 threads, each summing a contiguous chunk; out[n] = total.
__global__ __launch_bounds__(1024) void k_scan_one(const int *in, int *out, long n, chx_bases_args gb)
{
    constexpr int PER = CHX_SCAN1_MAX / 1024;
    static_assert(PER % 4 == 0, "segments of whole int4");
    __shared__ int wsum[16];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    // contiguous segments of a multiple of 4 elements, read and written as
    // int4 where the arrays are 16-byte aligned (fewer cache lines per access)
    const int per = (int)(((n + 1023) / 1024 + 3) & ~3L);
    const long b = (long)tid * per;
    const bool vec = ((reinterpret_cast<unsigned long>(in) | reinterpret_cast<unsigned long>(out)) & 15) == 0;
    int x[PER];
    int t = 0;
#pragma unroll
    for (int q = 0; q < PER / 4; q++) {
        const long e0 = b + 4 * q;
        if (vec && 4 * q < per && e0 + 3 < n) {
            const int4 v = *reinterpret_cast<const int4 *>(in + e0);
            x[4 * q] = v.x; x[4 * q + 1] = v.y; x[4 * q + 2] = v.z; x[4 * q + 3] = v.w;
        } else {
#pragma unroll
            for (int r = 0; r < 4; r++) x[4 * q + r] = (4 * q + r < per && e0 + r < n) ? in[e0 + r] : 0;
        }
#pragma unroll
        for (int r = 0; r < 4; r++) t += x[4 * q + r];
    }
    // inclusive scan of the thread totals: within each wavefront, then over
    // the 16 wavefront totals
    int v = t;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int q = __shfl_up(v, d, 64);
        if (lane >= d) v += q;
    }
    if (lane == 63) wsum[w] = v;
    __syncthreads();
    if (w == 0) {
        int q = lane < 16 ? wsum[lane] : 0;
#pragma unroll
        for (int d = 1; d < 16; d <<= 1) {
            const int y = __shfl_up(q, d, 64);
            if (lane >= d) q += y;
        }
        if (lane < 16) wsum[lane] = q;
    }
    __syncthreads();
    int run = v - t + (w ? wsum[w - 1] : 0);
#pragma unroll
    for (int q = 0; q < PER / 4; q++) {
        const long e0 = b + 4 * q;
        int o[4];
#pragma unroll
        for (int r = 0; r < 4; r++) {
            o[r] = run;
            run += x[4 * q + r];
        }
        if (vec && 4 * q < per && e0 + 3 < n) {
            *reinterpret_cast<int4 *>(out + e0) = make_int4(o[0], o[1], o[2], o[3]);
        } else {
#pragma unroll
            for (int r = 0; r < 4; r++)
                if (4 * q + r < per && e0 + r < n) out[e0 + r] = o[r];
        }
    }
    if (tid == 1023) out[n] = wsum[15];
    // the group bases of the 3-/4-body count scan, in the same launch
    if (gb.off == nullptr) return;
    __syncthreads();
    if (tid < 2)
        group_bases_of(out, gb.N, gb.G3, gb.G4, gb.cap3, gb.cap4, gb.base3, gb.end3, gb.base4, gb.end4, gb.status,
                       tid);
}

// Zeroes every listed range (one launch for all counters and accumulators).
__global__ void k_clear(const chx_clear_range *__restrict__ r, int nr)
{
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x, gs = (long)gridDim.x * blockDim.x;
    for (int k = 0; k < nr; k++) {
        int *p = r[k].p;
        const long n = r[k].n;
        for (long i = g; i < n; i += gs) p[i] = 0;
    }
}

__global__ void k_scan_add(int *out, long n, const int *bscan)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < n) out[idx] += bscan[idx / 2048];
}

__global__ void k_scan_total(int *out, long n, const int *in)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) out[n] = (n > 0 ? out[n - 1] + in[n - 1] : 0);
}

}  // namespace chx
