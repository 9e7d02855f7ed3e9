// Table files of the tabulated 2-/3-body path (see tables.h).
#include "tables.h"

#include <cmath>
#include <cstdlib>
#include <fstream>
#include <sstream>

namespace chx {

namespace {

struct TabFail {
    std::string msg;
};

// A table file: count line, then count rows of ncol_a or ncol_b items.
std::vector<std::vector<double>> read_rows(const std::string &path, size_t ncol_a, size_t ncol_b)
{
    std::ifstream in(path);
    if (!in) throw TabFail{"could not open table file " + path};
    std::string line;
    if (!std::getline(in, line)) throw TabFail{path + ": missing row count"};
    const long n = std::strtol(line.c_str(), nullptr, 10);
    if (n <= 0) throw TabFail{path + ": bad row count"};
    std::vector<std::vector<double>> rows((size_t)n);
    for (long i = 0; i < n; i++) {
        if (!std::getline(in, line)) throw TabFail{path + ": fewer rows than its count"};
        size_t c = line.find('!');
        if (c != std::string::npos) line.erase(c);
        c = line.find("##");
        if (c != std::string::npos) line.erase(c);
        std::istringstream ss(line);
        std::string item;
        std::vector<double> &row = rows[(size_t)i];
        while (ss >> item) row.push_back(std::strtod(item.c_str(), nullptr));
        if (row.size() != ncol_a && row.size() != ncol_b)
            throw TabFail{path + ": row " + std::to_string(i + 1) + " has " + std::to_string(row.size()) + " items"};
    }
    return rows;
}

std::string table_base(const ParamSet &P, const std::string &name)
{
    const size_t slash = P.path.find_last_of('/');
    return slash == std::string::npos ? name : P.path.substr(0, slash + 1) + name;
}

Table2 read_2b(const std::string &base)
{
    const auto e = read_rows(base + ".energy", 2, 2);
    const auto f = read_rows(base + ".force", 2, 2);
    if (e.size() != f.size()) throw TabFail{base + ": energy and force tables differ in length"};
    if (e.size() < 3) throw TabFail{base + ": a 2-body table needs at least 3 rows"};
    Table2 t;
    for (size_t i = 0; i < e.size(); i++) {
        t.r.push_back(e[i][0]);
        t.e.push_back(e[i][1]);
        t.f.push_back(f[i][1]);
    }
    return t;
}

Table3 read_3b(const std::string &base)
{
    const auto e = read_rows(base + ".energy", 4, 6);
    const auto f = read_rows(base + ".force", 6, 6);
    if (e.size() != f.size()) throw TabFail{base + ": energy and force tables differ in length"};
    Table3 t;
    t.n = (int)std::cbrt((double)e.size());
    const long n = t.n;
    if (n < 2 || (size_t)(n * n + n + 1) >= e.size())
        throw TabFail{base + ": a 3-body table needs at least 2 points per axis"};
    for (size_t i = 0; i < e.size(); i++) {
        t.rij.push_back(e[i][0]);
        t.rik.push_back(e[i][1]);
        t.rjk.push_back(e[i][2]);
        t.v.insert(t.v.end(), {e[i][3], f[i][3], f[i][4], f[i][5]});
    }
    t.dr = 1 / (t.rjk[(size_t)(n * n + n + 1)] - t.rjk[0]);
    return t;
}

}  // namespace

bool load_tables(const ParamSet &P, TableSet &T, std::string &err)
{
    T = TableSet{};
    try {
        if (P.tab2_mixed || (!P.tab2_files.empty() && (int)P.tab2_files.size() != P.npairs()))
            throw TabFail{"all 2-body parameters must be tabulated, or none"};
        for (const std::string &name : P.tab2_files) T.t2.push_back(read_2b(table_base(P, name)));
        for (const std::string &name : P.tab3_files) T.t3.push_back(read_3b(table_base(P, name)));
        if (!T.t3.empty())
            for (size_t t = 0; t < P.trip.size(); t++)
                if (!P.trip[t].excluded && t >= T.t3.size())
                    throw TabFail{"triplet type " + std::to_string(t) + " has no 3-body table"};
    } catch (const TabFail &f) {
        err = f.msg;
        return false;
    }
    return true;
}

}  // namespace chx
