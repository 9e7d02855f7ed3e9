// Tables of the tabulated 2- and 3-body path (the LAMMPS pair style of a
// -DTABULATION build): reading the files a parameter file names.
//
// File formats follow chimesFF::read_2B_tab / read_3B_tab
// (chimesFF/src/chimesFF.cpp:198-352): a row count, then the rows; '!' and
// '##' start comments (split_line, chimesFF.cpp:148-178).
//   <name>.energy  2-body "r E",           3-body "rij rik rjk E"
//   <name>.force   2-body "r F",           3-body "rij rik rjk Fij Fik Fjk"
// Only the energy file's distances are kept.  A 3-body table is a uniform
// grid, rij slowest and rjk fastest, of n = (int) cbrt(rows) points per axis
// whose spacing is taken from the rjk column (chimesFF.cpp:2339-2342).
#pragma once

#include <string>
#include <vector>

#include "params.h"

namespace chx {

struct Table2 {
    std::vector<double> r, e, f;
};

struct Table3 {
    int n = 0;        // points per axis
    double dr = 0.0;  // 1 / spacing, as interpolateTricubic computes it
    std::vector<double> rij, rik, rjk;  // columns of the energy file
    std::vector<double> v;              // per row: E, Fij, Fik, Fjk
};

struct TableSet {
    std::vector<Table2> t2;  // by parameter pair index (reading order)
    std::vector<Table3> t3;  // by triplet type index (reading order)
};

// Reads every table P names, relative to the parameter file's directory.
// Errors (the reference exits or reads past its arrays on most of them): a
// missing or short file, a row of the wrong width, energy and force files of
// different lengths, a 2-body table of fewer than 3 rows, a 3-body table of
// fewer than 2 points per axis, 2-body tables for some pairs only, and a
// 3-body type without a table when others have one.
bool load_tables(const ParamSet &P, TableSet &T, std::string &err);

}  // namespace chx
