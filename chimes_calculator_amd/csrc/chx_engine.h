// Internal C-ABI layer between the C++ host classes and the HIP engine.
// Plain pointers and sizes only; every array argument is host memory unless
// its name ends in _dev.
#pragma once
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct chx_engine chx_engine;
typedef struct chx_model chx_model;

// A model description built by the host from a parsed parameter file.
typedef struct {
    int nt;
    int order[3];
    int npairs;
    int tersoff;
    double fcut_var, pen_dist, pen_scale;
    const double *eoff;          // [nt]
    const int *pairmap;          // [nt*nt]
    const double *pair_slots;    // [npairs*5]: rin rout lam xavg xdiff
    const double *c2;            // [npairs*(order2+1)] dense by Chebyshev degree
    double rc[3];                // max outer cutoffs per body order
    int ntrip, D3;
    const int *tripmap;          // [nt^3]
    const int *trip_perm;        // [nt^3*3]
    const double *trip_slots;    // [ntrip*3*5]
    const double *C3;            // [ntrip*D3^3]
    const int *last3;            // [ntrip*3] powers of each type's last listed term
    const double *lastc3;        // [ntrip]
    int nquad, D4;
    const int *quadmap;          // [nt^4]
    const int *quad_perm;        // [nt^4*4]
    const double *quad_slots;    // [nquad*6*5]
    const double *C4;            // [nquad*D4^6]
    const int *last4;            // [nquad*6]
    const double *lastc4;        // [nquad]
} chx_model_desc;

// FP64; FP32 recurrences and contractions; BF16 matrix-core contractions with
// FP32 accumulation, with plain BF16 operands or with each operand split into
// a BF16 high part and a BF16 residual (three products); everything else as
// FP32-mixed.
typedef enum { CHX_FP64 = 0, CHX_MIXED = 1, CHX_BF16 = 2, CHX_BF16X3 = 3 } chx_precision;

// A batch of independent configurations.  Positions are xyz-interleaved
// input coordinates; types are indices into the configuration's model.
typedef struct {
    int nconf;
    const int *natoms;           // [nconf]
    const double *pos;           // [sum natoms * 3]   (host, or device if pos_dev)
    const int *types;            // [sum natoms]       (host, or device if pos_dev)
    const double *cells;         // [nconf * 9]: a, b, c cell vectors
    const int *model;            // [nconf] index into the models array
    int small;                   // replicate small cells (serial interface flag)
    int pos_dev;                 // 1: pos/types are device pointers
    int precision;               // chx_precision
} chx_batch;

typedef struct {
    double *energy;              // [nconf] raw sum (before the 1/(R+1)^3 factor)
    double *forces;              // [sum natoms * 3]
    double *virial;              // [nconf * 6]: xx xy xz yy yz zz (before /volume)
    double *volume;              // [nconf]
    int *n_rep;                  // [nconf]
    long long *counts;           // [nconf * 4]: penalty events, out-of-range events, 0, 0
    int *badness;                // [nconf] 0 good, 1 penalty region, 2 below inner cutoff
    int out_dev;                 // 1: energy/forces/virial are device pointers (others stay host)
} chx_out;

chx_engine *chx_engine_create(int device);
void chx_engine_destroy(chx_engine *e);
const char *chx_last_error(chx_engine *e);

chx_model *chx_model_create(chx_engine *e, const chx_model_desc *d);
void chx_model_destroy(chx_model *m);

// Runs the whole pipeline on the engine's stream and waits for it.
// Returns 0 on success, nonzero on error (message in chx_last_error).
int chx_evaluate(chx_engine *e, chx_model *const *models, int nmodels, const chx_batch *b, chx_out *o);

// Cluster lists of configuration cfg of the last evaluate call, in canonical
// key form (see oracle/chimes_oracle.h): 5 / 9 / 13 ints per entry.  Arrays
// are malloc'd; free with chx_free.
int chx_export_lists(chx_engine *e, int cfg, long *n2, int **pairs, long *n3, int **trips, long *n4,
                     int **quads);
void chx_free(void *p);

// Explicit clusters (the per-cluster API): nb = 2, 3 or 4 bodies, n clusters.
// dist[n*P], dr[n*P*3], types[n*nb]; outputs energy[n], force[n*nb*3],
// stress[n*6] (packed xx xy xz yy yz zz), last[n*P] (last-term scalars,
// 2-body: penalty scalar), flags[n].
int chx_eval_clusters(chx_engine *e, chx_model *m, int nb, int n, const double *dist, const double *dr,
                      const int *types, double *energy, double *force, double *stress, double *last,
                      int *flags);

// The same without allocating: inputs and outputs are staged in the
// library's static device scratch array (at most CHX_STATIC_MAX_CLUSTERS
// clusters; its address on each device is resolved when the device's first
// engine is created).  Calls that overlap on one device race on the array.
int chx_eval_clusters_static(chx_engine *e, chx_model *m, int nb, int n, const double *dist, const double *dr,
                             const int *types, double *energy, double *force, double *stress);

// Batch plans: a batch shape (models, model_of, natoms, small, precision and,
// when pos_dev / out_dev are set, the device buffers) fixed at creation.
// Creation evaluates b once (results in o), optionally times the 3-/4-body
// kernel variants on it and keeps the fastest (tune != 0), and captures the
// whole pipeline - input copy, kernels, output copy - in one HIP graph.  Each
// run replays the graph for new positions, types and cells; a run whose
// shape or cluster counts no longer fit the captured sizes is evaluated
// directly and captured again.  A plan owns its engine (stream and buffers).
typedef struct chx_plan chx_plan;
chx_plan *chx_plan_create(int device, chx_model *const *models, int nmodels, const chx_batch *b, int tune,
                          chx_out *o, char *err, int errlen);
int chx_plan_run(chx_plan *p, const double *pos, const int *types, const double *cells, chx_out *o);
void chx_plan_destroy(chx_plan *p);
const char *chx_plan_error(chx_plan *p);
// variants[2]: chosen variant of the 3- and 4-body kernels; tune_ms[8]: the
// timed kernel milliseconds per (body, variant) as [2][CHX_NVAR], negated for a variant whose
// results disagreed with the reference evaluation; counts[3]: as
// chx_last_counts.
void chx_plan_info(chx_plan *p, int *variants, double *tune_ms, long long *counts);

// Cluster counts of the last evaluate call: [0] pair entries, [1] triplets,
// [2] quadruplets (all configurations, excluded types included).
void chx_last_counts(chx_engine *e, long long *c3);

// Device bytes of the engine's work buffers: held_need[0] allocated,
// held_need[1] asked for by the last evaluation (its own need).
void chx_engine_bytes(chx_engine *e, size_t *held_need);
// Frees the work buffers (allocated again on demand) and forgets the grown
// capacities.
void chx_engine_trim(chx_engine *e);
// Forgets the grown capacities: the next evaluation sizes its buffers from its
// own estimates (allocations are kept and reused when large enough).
void chx_engine_reset_capacities(chx_engine *e);

// Kernel variants of the 3- and 4-body terms for the engine's later
// evaluations (CHX_VAR_*: matrix cores with 64- or 32-cluster tiles, vector
// units; -1 keeps the current one).  A matrix-core request for a dense
// dimension above the matrix-core kernels' range, and a vector request of a
// BF16 form, fall back as launch_eval_body does.  Nonzero for an unknown
// variant.
int chx_set_variants(chx_engine *e, int var3, int var4);
void chx_get_variants(chx_engine *e, int *v2);
// The same for a plan; the next run evaluates directly and captures again.
int chx_plan_set_variants(chx_plan *p, int var3, int var4);

// Timing of the last evaluate call (HIP events on the engine's stream, ms):
// [0] whole pipeline, [1] 4-body kernels, [2] 3-body kernels, [3] enumeration.
void chx_last_timing(chx_engine *e, double *ms4);

// Sorted Morse-transformed edges of cluster rows on the device
// (read_flat_clusters, calc_histograms.cpp:140-249): dist[n*P], types[n*nb]
// (checked by the caller) -> edges[n*P] (host arrays).
int chx_fp_edges(chx_engine *e, chx_model *m, int nb, long n, const double *dist, const int *types, double *edges);
// Cluster fingerprint histogram (calc_histograms.cpp:351-489): counts[nbin]
// of the distances between the np-edge clusters e1[n1] and e2[n2] (host
// arrays; same != 0: one set, pairs j > i).  ms: kernel milliseconds.
int chx_fp_histogram(chx_engine *e, int np, long n1, const double *e1, long n2, const double *e2, int same,
                     int nbin, double binw, double maxd, long long *counts, double *ms);

// LAMMPS pair style (etc/lmp/src/pair_chimes.cpp) over one process's local
// and ghost atoms.  chimes_type[t - 1] is the ChIMES type of LAMMPS type t.
// x[nall*3], type[nall] (LAMMPS types), tag[nall], ilist[inum] and the
// neighbour rows firstneigh[i][0 .. numneigh[i]) of the full list with ghosts
// are host arrays.  compute rebuilds the many-body lists when ago == 0 (or
// none exist) and adds onto f[nall*3], *eng, virial[6] (xx yy zz xy xz yz),
// eatom[nall] and vatom[nall*6] as the flags ask.
typedef struct chx_lmp chx_lmp;
chx_lmp *chx_lmp_create(chx_model *m, int ntypes, const int *chimes_type, char *err, int errlen);
void chx_lmp_destroy(chx_lmp *p);
const char *chx_lmp_error(chx_lmp *p);
int chx_lmp_build_lists(chx_lmp *p, int nall, const double *x, const int *tag, int inum, const int *ilist,
                        const int *numneigh, const int *const *firstneigh, double skin);
int chx_lmp_compute(chx_lmp *p, int nall, const double *x, const int *type, const int *tag, int inum,
                    const int *ilist, const int *numneigh, const int *const *firstneigh, int ago, double skin,
                    int eflag_global, int eflag_atom, int vflag_global, int vflag_atom, double *f, double *eng,
                    double *virial, double *eatom, double *vatom);
// Many-body lists of the last build: (i, j, k) and (i, j, k, l) atom indices
// in the reference's loop order; malloc'd, free with chx_free.
int chx_lmp_lists(chx_lmp *p, long *n3, int **trips, long *n4, int **quads);
// Fingerprint rows of the last compute (FINGERPRINT build of pair_chimes.cpp):
// nb(nb-1)/2 pair distances then nb types per row, malloc'd, in the loops'
// order; free with chx_free.
int chx_lmp_fp_rows(chx_lmp *p, int nb, long *nrows, double **rows);
// Milliseconds of the last compute (HIP events): [0] input copies, [1] list
// build, [2] 1-/2-body, [3] 3-/4-body, read-back of tallies.
void chx_lmp_timing(chx_lmp *p, double *ms4);

#ifdef __cplusplus
}

namespace chx {
struct TableSet;
}
// The tables of the tabulated 2-/3-body path (a -DTABULATION build of the
// pair style) for the following compute calls; null or empty: the polynomials.
int chx_lmp_set_tables(chx_lmp *p, const chx::TableSet *T);
#endif
