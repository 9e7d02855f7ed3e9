// Plain-old-data layouts shared by the host engine and the HIP kernels.
#pragma once
#include <stdint.h>

#define CHX_MAX_TYPES 8
#define CHX_MAX_NB 512         // neighbour images per owner atom within the search cutoff
#define CHX_MAX_N3 512         // images per owner within the 3-body cutoff
#define CHX_MAX_N4 256         // images per owner within the 4-body cutoff
#define CHX_FIN_CHUNK 2048     // owners per block of the energy/virial reduction
#define CHX_NBR_CELLS 128      // searched cells per owner for the flattened candidate loop
#define CHX_NBR_CANDS 512      // candidates per owner whose cell is tabulated in LDS (binary search above)
#define CHX_ENUM_WIDE_MAX 16384 // batches up to this many atoms enumerate with four wavefronts per owner
#define CHX_ENUM_SMALL 128     // owners with at most this many images use the small-LDS enumeration
#define CHX_EVAL2_LANES 8      // lanes per owner of the 2-body kernel (fixes its summation order: one value for every batch)
#define CHX_MAX_CLUSTER_TYPES 64
#define CHX_LIST_ALIGN 64      // 3-/4-body list segments of one cluster type start on this multiple
#define CHX_MFMA_MAX_D 4       // 4-body dense dimension handled by the matrix-core kernel
#define CHX_MFMA3_MAX_D 13     // 3-body dense dimension handled by the matrix-core kernel
#define CHX_MFMA3V_MAX_D 10    // 3-body dense dimension of the one-cluster-per-lane matrix-core kernel (FP32)
#define CHX_STATIC_MAX_CLUSTERS 131072  // clusters per call of the allocation-free explicit-cluster entry
// Per-configuration energy and virial accumulators: CHX_CACC_SLOTS copies per
// configuration, one 128-byte line each (7 (hi, lo) pairs), picked by
// workgroup, so the wavefronts of one large configuration spread their
// atomics over many lines instead of queueing on one.
#define CHX_CACC_SLOTS 32
#define CHX_CACC_STRIDE 16

// Status words of one evaluation (device int array, read back at the end).
enum {
    CHX_ST_FLAGS = 0,          // error / retry bits below
    CHX_ST_MAXNB = 1,          // largest image count of an owner
    CHX_ST_NENT = 2,           // neighbour entries needed
    CHX_ST_N3 = 3,             // triplet list slots needed (aligned)
    CHX_ST_N4 = 4,             // quadruplet list slots needed (aligned)
    CHX_ST_C3 = 5,             // triplets listed
    CHX_ST_C4 = 6,             // quadruplets listed
    CHX_ST_NBIG = 7,           // owners queued for the large enumeration kernel
    CHX_ST_MAXREC = 8,         // largest cluster record count of an owner
    CHX_ST_WORDS = 9
};
enum {
    CHX_BAD_NB = 1,            // an owner has more than CHX_MAX_NB images
    CHX_BAD_N34 = 2,           // an owner exceeds CHX_MAX_N3 / CHX_MAX_N4
    CHX_RETRY_ENT = 4,         // entry buffer too small
    CHX_RETRY_L3 = 8,          // triplet list buffer too small
    CHX_RETRY_L4 = 16,         // quadruplet list buffer too small
    CHX_BAD_RANGE = 32,        // a contribution outside the fixed-point range
    CHX_RETRY_BIG = 64,        // an owner needs the crowded-owner enumeration, which was not launched
    CHX_RETRY_REC = 128,       // an owner's cluster records exceed the record slab
    CHX_RETRY_ANY = CHX_RETRY_ENT | CHX_RETRY_L3 | CHX_RETRY_L4 | CHX_RETRY_BIG | CHX_RETRY_REC,
    CHX_STOP = CHX_BAD_NB | CHX_BAD_N34 | CHX_RETRY_ANY
};

// A range of ints zeroed by the clear kernel at the start of a run.
struct chx_clear_range {
    int *p;
    long n;
};

// 3-/4-body evaluation kernel variants (chosen per body order by the plan
// autotuner): matrix cores with 4 or 2 wavefronts per workgroup (64 or 32
// clusters per tile), the vector units, or (3-body, FP32-mixed) one cluster
// per lane with the contraction on the matrix cores; where a variant does
// not apply the vector kernel stands in.
enum { CHX_VAR_MFMA4 = 0, CHX_VAR_MFMA2 = 1, CHX_VAR_VECTOR = 2, CHX_VAR_MFMAV = 3, CHX_NVAR = 4 };
#define CHX_SCAN1_MAX 32768    // scans up to this length run in one workgroup

// Per-configuration frame data (one per configuration in a batch).
struct chx_cfg {
    double h[9], ih[9];        // evaluation frame: replicated cell and inverse
    double h0[9], ih0[9];      // input cell and inverse (replication step)
    double hn[9], ihn[9];      // neighbour frame (reoriented cell) and inverse
    double vol;
    double search;             // max(rc2, rc3, rc4)
    double rc3, rc4;           // max 3B / 4B outer cutoffs (list predicates)
    int natoms_in;             // input atoms
    int n_rep;                 // replication count R (R+1)^3 copies
    int natoms;                // replicated atoms
    int n_layers;              // ghost layers of the neighbour frame
    int m[3];                  // cell grid
    int reach[3];              // cells searched on each side
    int atom_off;              // first replicated atom (global index)
    int atom_in_off;           // first input atom (global index)
    int cell_off;              // first cell (global index)
    int model;                 // parameter set index
    double near2;              // squared bound of the cheap neighbour test (see nbr_near)
    double lam_min;            // lower bound of the smallest eigenvalue of hn^T hn (cell pruning)
};

// Replicated atom, evaluation frame: fractional coordinates u = ih * x.
struct chx_atom_s {
    double u[3];
    int cfg;
    int type;
};

// Replicated atom, neighbour frame.
struct chx_atom_n {
    double u[3];
    int cell;
    int orig;                  // global input atom this replica maps to
};

// Neighbour entry of an owner atom: an image (parent atom, lattice shift).
struct chx_entry_s {
    double v[3];               // evaluation-frame fractional coordinates of the image
    int parent;                // global replicated atom index
    int type;
};
struct chx_entry_n {
    double v[3];               // neighbour-frame fractional coordinates of the image
    int shift;                 // packed (sx+128) | (sy+128) << 8 | (sz+128) << 16
    int pad;
};

// Per-slot constants of one cluster type (or pair type).
// Cutoffs and Morse constants of one pair slot, with the reciprocals the
// kernels multiply by (1 / lambda, 1 / xdiff, 1 / rout and, for the Tersoff
// cutoff, 1 / (rout - thr), the width of the ramp that starts at
// thr = rout - fcut_var * rout).
struct chx_slot {
    double rin, rout, lam, xavg, xdiff;
    double ilam, ixdiff, irout, iramp;
};

// One (model, cluster type) group of the 3- or 4-body list.  type < 0 marks
// the group of excluded type tuples (listed, never evaluated).
struct chx_group {
    int model, type, nt, D;
    int tersoff, pad;
    double fcut_var;
    const int *perm;           // canonical member order per type tuple
    const double *C;           // dense coefficients, parameter pair order
    const float *Cf;           // FP32 copy
    const double *Cm;          // 4-body matrix image for the matrix-core kernel (D <= 4)
    const float *Cmf;          // FP32 copy
    const unsigned short *Cmb; // BF16 images (rounded, then residual), transposed: [column][row padded to 32]
    chx_slot s[6];
};

// Device parameter tables of one model; all pointers are device pointers.
struct chx_model_dev {
    int nt;
    int order[3];
    int npairs;
    int tersoff;
    double fcut_var;
    double pen_dist, pen_scale;
    const double *eoff;        // [nt]
    const int *pairmap;        // [nt*nt]
    // 2-body, per pair type p: slots2[p], coefficient k of pair p at c2[p*o2 + k]
    // multiplies T_{pow2[p*o2+k]+1}
    const chx_slot *slots2;
    const int *ncoef2;
    const int *pow2;
    const double *c2;
    // 3-body
    int ntrip, D3;
    const int *tripmap;        // [nt^3] cluster type, -1 excluded
    const int *trip_perm;      // [nt^3 * 3]
    const chx_slot *slots3;    // [ntrip * 3]
    const double *C3;          // [ntrip * D3^3] dense coefficients, parameter slot order
    const int *last3;          // [ntrip * 3] powers of the last listed term
    const double *lastc3;      // [ntrip] its coefficient
    // 4-body
    int nquad, D4;
    const int *quadmap;        // [nt^4]
    const int *quad_perm;      // [nt^4 * 4]
    const chx_slot *slots4;    // [nquad * 6]
    const double *C4;          // [nquad * D4^6]
    const int *last4;          // [nquad * 6]
    const double *lastc4;      // [nquad]
};
