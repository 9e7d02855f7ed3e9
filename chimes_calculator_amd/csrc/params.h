// ChIMES parameter set: host-side model of a parameter file and the
// permutation tables derived from it.
//
// Grammar and semantics follow chimesFF::read_parameters
// (chimesFF/src/chimesFF.cpp:354-1570) and the pair-permutation search of
// chimesFF::build_atom_and_pair_mappers (chimesFF/src/chimesFF.h:438-539).
#pragma once

#include <string>
#include <vector>

namespace chx {

struct ClusterType {
    std::vector<std::string> atoms;  // element of each cluster atom, parameter order
    std::vector<std::string> pairs;  // proper pair name of each cluster pair, parameter order
    bool excluded = false;
    std::vector<int> powers;         // K x npairs, parameter pair order
    std::vector<double> coef;        // K
    double rin[6] = {0, 0, 0, 0, 0, 0};
    double rout[6] = {0, 0, 0, 0, 0, 0};
};

struct ParamSet {
    std::string path;
    int order[3] = {0, 0, 0};           // 2-, 3-, 4-body polynomial orders
    int natmtyps = 0;
    std::vector<std::string> atmtyps;
    std::vector<double> masses;
    std::vector<double> energy_offsets;
    // pairs
    std::vector<std::string> pair_a1, pair_a2;
    std::vector<double> pair_rin, pair_rout, morse;
    std::vector<std::vector<int>> pow2;  // [pair][k]
    std::vector<std::vector<double>> c2;
    std::vector<int> pairmap;            // [ti*nt+tj] -> pair parameter index
    std::vector<int> pairmap_slow;       // PAIRMAPS entries, file order -> pair index
    std::vector<std::string> proper;     // [ti*nt+tj] -> proper pair name
    // smoothing / penalty
    bool tersoff = false;
    double fcut_var = 0.0;
    double penalty_dist = 0.01;          // chimesFF.cpp:110
    double penalty_scale = 1.0e4;        // chimesFF.cpp:111
    // tables named by the parameter file for the tabulated 2-/3-body path of
    // the LAMMPS pair style (chimesFF.cpp:715-745, 876-918), reading order;
    // the polynomial coefficients of those entries are read as well
    std::vector<std::string> tab2_files, tab3_files;
    bool tab2_plain = false;             // a 2-body entry without a table
    bool tab2_mixed = false;             // tabulated and plain 2-body entries
    // many-body
    std::vector<ClusterType> trip, quad;
    std::vector<int> tripmap, quadmap;   // [tuple] -> cluster type (negative: excluded)
    std::vector<int> trip_mapped, quad_mapped;  // [tuple*npairs + p] -> parameter pair slot
    std::vector<int> trip_perm, quad_perm;      // [tuple*natoms + a] -> runtime atom at
                                                // parameter atom position a
    std::vector<int> trip_perm_ok, quad_perm_ok;

    int npairs() const { return (int)pair_a1.size(); }
    double max_cutoff(int body) const;   // chimesFF.cpp:2903-2943
};

// Parses the file; on failure returns false and fills err (the reference
// prints and exits; the C ABI layer decides what to do with the message).
bool load_params(const std::string &path, ParamSet &P, std::string &err);

int type_index(const ParamSet &P, const std::string &name);

}  // namespace chx
