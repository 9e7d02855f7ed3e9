// Kernel declarations shared by the HIP translation units of the engine.
#pragma once
#include <hip/hip_runtime.h>

#include "chx_types.h"

namespace chx {

// Per-model data the enumeration kernel needs.
struct chx_enum_model {
    int nt, has3, has4, ntrip, nquad, g3, g4, pad;
    const int *tripmap;   // [nt^3] cluster type, -1 excluded
    const int *quadmap;   // [nt^4]
};

// Per-model data the owner reduction needs.
struct chx_red_model {
    int g3, ntrip, g4, nquad;
    const double *eoff;
};

__global__ void k_fill_i32(int *p, long n, int v);
__global__ void k_fill_f64(double *p, long n);
__global__ void k_prep(int natoms, int ncfg, const int *cfg_atom_off, const chx_cfg *cfgs, const double *pos,
                       const int *types, chx_atom_s *as, chx_atom_n *an, int *cell_count,
                       const int *cfg_item_off, int *in_cfg, int *item_cfg);
__global__ void k_cell_fill(int natoms, const chx_atom_n *an, const int *cell_start, int *cursor, int *cell_atoms);
__global__ void k_nbr(int natoms, const chx_cfg *cfgs, const chx_atom_s *as, const chx_atom_n *an,
                      const int *cell_start, const int *cell_atoms, int *nbr_cnt, int *nbr_off, chx_entry_s *es,
                      chx_entry_n *en, int nb_cap, int *status);
__global__ void k_nbr_wide(int natoms, const chx_cfg *cfgs, const chx_atom_s *as, const chx_atom_n *an,
                           const int *cell_start, const int *cell_atoms, int *nbr_cnt, int *nbr_off,
                           chx_entry_s *es, chx_entry_n *en, int nb_cap, int *status);
// Aligned group segments from the scan of the per-(group, owner) counts.
struct chx_bases_args {
    const int *off;  // the scan ([G3 + G4][N] + total); null: no bases
    long N;
    int G3, G4;
    long cap3, cap4;
    int *base3, *end3, *base4, *end4;
    int *status;
};
__global__ void k_group_bases(chx_bases_args a);
// arguments of the enumeration kernels
struct chx_enum_args {
    int natoms;
    const chx_cfg *cfgs;
    const chx_atom_s *as;
    const chx_atom_n *an;
    const int *nbr_off;
    const int *nbr_cnt;
    const chx_entry_s *es;
    const chx_entry_n *en;
    const chx_enum_model *em;
    int *cnt3, *cnt4;              // per (group, owner) counts [G][N]
    int2 *rec;                     // per owner slab of rcap cluster records: (members, group | slot)
    int *nrec;                     // [natoms] records of each owner
    int rcap;
    const int *off3, *off4;        // k_place: exclusive scans of the counts
    const int *gb3, *gb4;          // k_place: aligned group bases
    int G3, G4;                    // k_place: triplet and quadruplet groups
    int2 *list3, *list4;           // k_place output
    int *big;                      // [natoms]: owners queued for k_enum_big
    int big_on;                    // k_enum_big runs in this pipeline (else a queued owner asks for a retry)
    int *status;
};
template <int NT>
__global__ void k_enum(chx_enum_args x);
__global__ void k_enum_big(chx_enum_args x);
__global__ void k_place(chx_enum_args x);
__global__ void k_scan_block(const int *in, int *out, long n, int *bsum);
__global__ void k_scan_add(int *out, long n, const int *bscan);
__global__ void k_scan_one(const int *in, int *out, long n, chx_bases_args gb);
__global__ void k_clear(const chx_clear_range *r, int nr);
__global__ void k_scan_total(int *out, long n, const int *in);

}  // namespace chx
