/*
 * MI355X extensions of the ChIMES evaluator: batched configurations, batched
 * explicit clusters, FP32-mixed precision and device-resident buffers.
 *
 * These entry points have no counterpart in the reference; they sit beside
 * its C ABIs (chimescalc_serial_C.h, chimescalc_C.h) for callers that
 * evaluate many systems at once.  Every function returns 0 on success and a
 * nonzero code on error (message in chimes_gpu_last_error()); none exits.
 */
#ifndef CHIMES_GPU_H
#define CHIMES_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct chimes_gpu_model chimes_gpu_model;

/* Precision of the 3-/4-body terms: FP64; FP32-mixed (FP32 polynomials and
 * contractions, FP64 distances, cutoffs, force scalars and accumulation);
 * BF16 (as FP32-mixed, with BF16 matrix-core contractions accumulated in
 * FP32); BF16X3 (each BF16 operand split into a high part and a residual,
 * three products per term).  The 2-body term is always FP64. */
enum { CHIMES_GPU_FP64 = 0, CHIMES_GPU_MIXED = 1, CHIMES_GPU_BF16 = 2, CHIMES_GPU_BF16X3 = 3 };
/* 3-/4-body kernel variants (chimes_gpu_set_variants, chimes_gpu_plan_info) */
enum {
    CHIMES_GPU_VAR_MFMA4 = 0,  /* matrix cores, 64-cluster tiles */
    CHIMES_GPU_VAR_MFMA2 = 1,  /* matrix cores, 32-cluster tiles */
    CHIMES_GPU_VAR_VECTOR = 2, /* vector units, one cluster per lane */
    CHIMES_GPU_VAR_MFMAV = 3   /* one cluster per lane, 3-body contraction on the matrix cores (FP32-mixed) */
};

/* Parses a ChIMES parameter file and uploads its tables to GPU `device`. */
chimes_gpu_model *chimes_gpu_model_load(const char *param_file, int device);
void chimes_gpu_model_free(chimes_gpu_model *m);
const char *chimes_gpu_last_error(void);

int chimes_gpu_model_ntypes(const chimes_gpu_model *m);
/* Index of an element name in the model's atom types, or -1. */
int chimes_gpu_model_type_index(const chimes_gpu_model *m, const char *name);
int chimes_gpu_model_order(const chimes_gpu_model *m, int body);          /* body = 2, 3, 4 */
double chimes_gpu_model_max_cutoff(const chimes_gpu_model *m, int body);  /* body = 2, 3, 4 */

/* Evaluates nconf independent periodic configurations in one pipeline.
 *   models[nmodels], model_of[nconf]  model of each configuration
 *   natoms[nconf]                      atoms per configuration
 *   pos[3*sum natoms]                  x y z interleaved, configurations back to back
 *   types[sum natoms]                  atom type indices of each configuration's model
 *   cells[9*nconf]                     cell vectors a, b, c
 *   small                              1: replicate cells shorter than the 2B cutoff
 *   precision                          CHIMES_GPU_FP64, _MIXED, _BF16 or _BF16X3
 * Outputs (overwritten): energy[nconf] (kcal/mol, 1B..4B, divided by (R+1)^3
 * when replicated), forces[3*sum natoms] (kcal/mol/A), stress[9*nconf]
 * (virial / volume, kcal/mol/A^3, xx xy xz yx yy yz zx zy zz). */
int chimes_gpu_eval_batch(chimes_gpu_model *const *models, int nmodels, int nconf, const int *model_of,
                          const int *natoms, const double *pos, const int *types, const double *cells, int small,
                          int precision, double *energy, double *forces, double *stress);

/* Same, with pos/types already in device memory and energy/forces/virial
 * written to device memory (virial: 6 per configuration, xx xy xz yy yz zz,
 * not divided by volume; energy not divided by (R+1)^3). */
int chimes_gpu_eval_batch_device(chimes_gpu_model *const *models, int nmodels, int nconf, const int *model_of,
                                 const int *natoms, const double *pos_dev, const int *types_dev,
                                 const double *cells, int small, int precision, double *energy_dev,
                                 double *forces_dev, double *virial_dev);

/* Batch plans: one batch shape (models, model_of, natoms, small, precision)
 * evaluated repeatedly by replaying a single captured HIP graph (input copy,
 * every kernel, output copy) - the form for thousands of uneven
 * configurations evaluated again and again.  Creation evaluates the given
 * inputs once (results in energy/forces/stress, as chimes_gpu_eval_batch) and,
 * when tune != 0, times the 3-/4-body kernel variants (matrix cores with 64-
 * or 32-cluster tiles, or vector units) on this batch and keeps the fastest.
 * chimes_gpu_plan_run takes new positions, types and cells of the same shape;
 * a run whose cluster counts outgrow the captured buffers is evaluated
 * directly and captured again.  A plan owns its GPU stream and buffers; calls
 * on one plan must not overlap. */
typedef struct chimes_gpu_plan chimes_gpu_plan;
chimes_gpu_plan *chimes_gpu_plan_create(chimes_gpu_model *const *models, int nmodels, int nconf,
                                        const int *model_of, const int *natoms, const double *pos, const int *types,
                                        const double *cells, int small, int precision, int tune, double *energy,
                                        double *forces, double *stress);
int chimes_gpu_plan_run(chimes_gpu_plan *p, const double *pos, const int *types, const double *cells,
                        double *energy, double *forces, double *stress);
/* Device-buffer plans: pos_dev/types_dev and the outputs (as in
 * chimes_gpu_eval_batch_device) are fixed at creation; each run re-reads the
 * same input buffers and takes the cells (host) of this run. */
chimes_gpu_plan *chimes_gpu_plan_create_device(chimes_gpu_model *const *models, int nmodels, int nconf,
                                               const int *model_of, const int *natoms, const double *pos_dev,
                                               const int *types_dev, const double *cells, int small, int precision,
                                               int tune, double *energy_dev, double *forces_dev,
                                               double *virial_dev);
int chimes_gpu_plan_run_device(chimes_gpu_plan *p, const double *cells, double *energy_dev, double *forces_dev,
                               double *virial_dev);
/* variants[2]: kernel variant of the 3- and 4-body terms (0 matrix cores,
 * 64-cluster tiles; 1 matrix cores, 32-cluster tiles; 2 vector units; 3 one
 * cluster per lane with the contraction on the matrix cores, 3-body
 * FP32-mixed only, D <= 10);
 * tune_ms[8]: timed milliseconds per (body, variant), tune_ms[4 (body - 3) +
 * variant], 0 if not tuned or not applicable, negated
 * for a variant the tuner rejected: the tuner keeps the fastest variant whose
 * results agree with a reference evaluation of the batch (the default
 * variants, in FP64 for the FP32-mixed form) within the precision's bars;
 * counts[3]: pair images, triplets, quadruplets of the last run. */
void chimes_gpu_plan_info(const chimes_gpu_plan *p, int variants[2], double tune_ms[8], long long counts[3]);
void chimes_gpu_plan_free(chimes_gpu_plan *p);

/* Kernel variants of the 3- and 4-body terms (numbered as in
 * chimes_gpu_plan_info; -1 keeps the current one).  chimes_gpu_set_variants
 * sets them for every later evaluation on `device` that does not go through
 * a plan (chimes_gpu_eval_batch*, calculate_chimes*, the chimesFF classes);
 * the defaults are 2 (vector units) for the 3-body and 0 (matrix cores,
 * 64-cluster tiles) for the 4-body term.  chimes_gpu_plan_set_variants sets
 * them for a plan, whose next run evaluates directly and captures again.  A
 * matrix-core variant for a dense dimension above the matrix-core kernels'
 * range (4-body D > 4) runs on the vector units, and so does variant 3 where
 * it does not apply (4-body, FP64, 3-body D > 10); a vector-unit request of a
 * BF16 form runs on the matrix cores.  Nonzero for an unknown variant. */
int chimes_gpu_set_variants(int device, int var3, int var4);
int chimes_gpu_get_variants(int device, int variants[2]);
int chimes_gpu_plan_set_variants(chimes_gpu_plan *p, int var3, int var4);

/* Serving: a continuous-batching request loop over resident models.  Each of
 * nslots workers owns a GPU stream and device buffers (kept between batches)
 * and repeatedly evaluates every pending request that fits max_batch_atoms
 * and the HBM budget (estimated device bytes of all requests in flight) as one
 * multi-model batch; batches of different slots overlap on the device.
 * Requests are single configurations with host inputs (copied at submit).
 * Every result is bit-identical to evaluating the request alone.  A batch that
 * fails is evaluated again request by request, so an error fails only the
 * request that causes it. */
typedef struct chimes_gpu_server chimes_gpu_server;
chimes_gpu_server *chimes_gpu_server_create(chimes_gpu_model *const *models, int nmodels, int precision,
                                            int nslots, long max_batch_atoms, double hbm_budget_gb);
/* Returns the request id (> 0), or -1 on error (chimes_gpu_server_last_error). */
long chimes_gpu_server_submit(chimes_gpu_server *s, int model, int natoms, const double *pos, const int *types,
                              const double *cell, int small);
/* Waits up to timeout_s for request id (id > 0) or for any request (id == 0)
 * to complete; returns its id, 0 on timeout, -1 on error. */
long chimes_gpu_server_wait(chimes_gpu_server *s, long id, double timeout_s);
/* Results of a completed request (as chimes_gpu_eval_batch for one
 * configuration) and its times in ms since server creation: submitted,
 * started, done.  Releases the request; nonzero if it failed. */
int chimes_gpu_server_result(chimes_gpu_server *s, long id, double *energy, double *forces, double *stress,
                             double times_ms[3]);
/* [0] requests served, [1] batches, [2] atoms, [3] largest batch (requests). */
void chimes_gpu_server_stats(chimes_gpu_server *s, long long stats[4]);
/* Memory accounting in bytes: [0] the budget; [1] the largest estimated
 * device bytes of the batches in flight together; [2] the largest device
 * bytes held by the slots' buffers together (a slot releases its buffers
 * when the slots hold more than the budget); [3] the largest ratio of the
 * bytes a batch's evaluation asked for to its estimate; [4] batches that
 * asked for more than their estimate. */
void chimes_gpu_server_memory(chimes_gpu_server *s, double out[5]);
/* [0] failed requests, [1] batches that failed and were evaluated again
 * request by request (so that an error fails only its own request). */
void chimes_gpu_server_failures(chimes_gpu_server *s, long long out[2]);
const char *chimes_gpu_server_last_error(void);
/* Finishes the pending requests and stops the workers. */
void chimes_gpu_server_free(chimes_gpu_server *s);

/* Milliseconds of the last evaluation on the model's device, from HIP events:
 * [0] pipeline, [1] 4-body kernels, [2] 3-body kernels, [3] lists. */
void chimes_gpu_last_timing(const chimes_gpu_model *m, double ms[4]);

/* Cluster counts of the last evaluation on the model's device: [0] pair
 * images, [1] triplets, [2] quadruplets, summed over configurations. */
void chimes_gpu_last_counts(const chimes_gpu_model *m, long long counts[3]);

/* Batched explicit clusters: n clusters of nb = 2, 3 or 4 atoms given by the
 * nb(nb-1)/2 pair distances dist[n*P], pair vectors dr[n*P*3] (x_b - x_a,
 * pair order 01 02 03 12 13 23) and type indices types[n*nb].  Overwrites
 * energy[n], force[n*nb*3], stress[n*6] (xx xy xz yy yz zz, -sum f dr dr). */
int chimes_gpu_compute_clusters(chimes_gpu_model *m, int nb, int n, const double *dist, const double *dr,
                                const int *types, double *energy, double *force, double *stress);

/* Allocation-free chimes_gpu_compute_clusters.  It has no counterpart in the
 * reference; it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments and results as chimes_gpu_compute_clusters; the clusters'
 * inputs and outputs are staged in a static device array of the library
 * (about 52 MiB, held on each GPU the library runs on until the process
 * ends), so n is at most CHIMES_GPU_STATIC_MAX_CLUSTERS (a larger n is an
 * error) and no call allocates.  All calls on one GPU share that array: calls
 * that overlap on one GPU race. */
#define CHIMES_GPU_STATIC_MAX_CLUSTERS 131072
int chimes_gpu_compute_clusters_static(chimes_gpu_model *m, int nb, int n, const double *dist, const double *dr,
                                       const int *types, double *energy, double *force, double *stress);

/* Cluster lists of configuration `conf` of the last evaluation, as sorted
 * canonical keys: pairs (i, a, sx, sy, sz), triplets (i, a_j, s_j, a_k, s_k),
 * quadruplets (i, a_j, s_j, a_k, s_k, a_l, s_l); free with chimes_gpu_free. */
int chimes_gpu_cluster_lists(chimes_gpu_model *m, int conf, long *n2, int **pairs, long *n3, int **trips,
                             long *n4, int **quads);
void chimes_gpu_free(void *p);

/* Cluster fingerprints (chimesFF/src/FP/calc_histograms.cpp).
 * chimes_gpu_fp_edges: rows of a cluster dump - the nb(nb-1)/2 pair
 * distances dist[n*P] (pair order 01 02 03 12 13 23) and the nb type indices
 * types[n*nb] of each cluster, as chimesFF's FINGERPRINT mode records them
 * (chimesFF.cpp:1694-1703, 2013-2020, 2603-2612; chimes_lmp_fp_rows makes
 * them on the GPU) - become edges[n*P], the sorted Morse-transformed
 * distances with the cutoffs of each pair's slot (read_flat_clusters,
 * calc_histograms.cpp:140-249), computed on the GPU (the exponentials are
 * the device library's: an edge can differ from a host libm's by an ulp).
 * chimes_gpu_fp_histogram: counts[nbin] of the distances between every
 * cluster of edges1[n1] and every cluster of edges2[n2] - or, with same != 0,
 * every pair j > i of edges1 - binned as get_bin does with bin width binw and
 * top value maxd (gen_flat_hists, calc_histograms.cpp:351-489; the reference
 * divides the counts by nsamples when it writes the .hist file).  A distance
 * outside the nbin bins (at most 8192) is an error; the reference leaves the
 * case bin == nbin unchecked. */
int chimes_gpu_fp_edges(chimes_gpu_model *m, int nb, long n, const double *dist, const int *types, double *edges);
int chimes_gpu_fp_histogram(chimes_gpu_model *m, int nb, long n1, const double *edges1, long n2,
                            const double *edges2, int same, int nbin, double binw, double maxd, long long *counts,
                            long long *nsamples);

#ifdef __cplusplus
}
#endif
#endif
