// ChIMES parameter parser (host).  Follows the multi-pass scan of
// chimesFF::read_parameters (chimesFF/src/chimesFF.cpp:354-1570): every pass
// rewinds to the first line and stops at the first line containing ENDFILE;
// keys are recognised by substring, tokens by whitespace after stripping
// comments that start with '!' or "##" (chimesFF.cpp:148-179).
#include "params.h"

#include <algorithm>
#include <cstdlib>
#include <fstream>
#include <sstream>

namespace chx {

namespace {

struct Lines {
    std::vector<std::string> v;
    size_t pos = 0;
    bool eof = false;
    const std::string &next()
    {
        static const std::string empty;
        if (pos >= v.size()) { eof = true; return empty; }
        return v[pos++];
    }
};

std::vector<std::string> tokens(std::string line)
{
    size_t p = line.find('!');
    if (p != std::string::npos) line.erase(p);
    p = line.find("##");
    if (p != std::string::npos) line.erase(p);
    std::istringstream ss(line);
    std::vector<std::string> t;
    std::string s;
    while (ss >> s) t.push_back(s);
    return t;
}

bool has(const std::string &s, const char *key) { return s.find(key) != std::string::npos; }

struct Fail {
    std::string msg;
};

const std::string &tok(const std::vector<std::string> &t, size_t i)
{
    if (i >= t.size()) throw Fail{"missing token in parameter line"};
    return t[i];
}
int toi(const std::vector<std::string> &t, size_t i) { return std::stoi(tok(t, i)); }
double tod(const std::vector<std::string> &t, size_t i) { return std::stod(tok(t, i)); }

int proper_pair(const ParamSet &P, const std::string &a, const std::string &b)
{
    for (int i = 0; i < P.npairs(); i++) {
        if (a == P.pair_a1[i] && b == P.pair_a2[i]) return i;
        if (b == P.pair_a1[i] && a == P.pair_a2[i]) return i;
    }
    throw Fail{"no proper pair name for atom types " + a + ", " + b};
}

// Permutation search of chimesFF.h:438-539: the first permutation (in
// next_permutation order) of the cluster's atoms whose proper pair names
// equal the parameter block's pair order.
bool first_match(const ParamSet &P, int natoms, const int *typ, const std::vector<std::string> &clu_pairs,
                 int *perm_out, int *mapped)
{
    int perm[4] = {0, 1, 2, 3};
    std::vector<std::pair<int, int>> tp;
    for (int i = 0; i < natoms; i++)
        for (int j = i + 1; j < natoms; j++) tp.push_back({i, j});
    do {
        bool match = true;
        for (size_t q = 0; q < tp.size() && match; q++) {
            const std::string &nm = P.proper[typ[perm[tp[q].first]] * P.natmtyps + typ[perm[tp[q].second]]];
            if (nm != clu_pairs[q]) match = false;
        }
        if (match) {
            for (int a = 0; a < natoms; a++) perm_out[a] = perm[a];
            for (size_t i = 0; i < tp.size(); i++) {
                int ra = perm[tp[i].first], rb = perm[tp[i].second];
                for (size_t j = 0; j < tp.size(); j++)
                    if ((ra == tp[j].first && rb == tp[j].second) || (ra == tp[j].second && rb == tp[j].first))
                        mapped[j] = (int)i;
            }
            return true;
        }
    } while (std::next_permutation(perm, perm + natoms));
    return false;
}

void read_cluster_block(ParamSet &P, Lines &L, int nb)
{
    const int npr = nb == 3 ? 3 : 6;
    std::vector<ClusterType> &types = nb == 3 ? P.trip : P.quad;
    const char *kcount = nb == 3 ? "ATOM PAIR TRIPLETS:" : "ATOM PAIR QUADRUPLETS:";
    const char *kblock = nb == 3 ? "TRIPLETTYPE PARAMS:" : "QUADRUPLETYPE PARAMS:";
    const char *kmaps = nb == 3 ? "TRIPMAPS:" : "QUADMAPS:";
    std::vector<std::string> slow_name;
    std::vector<int> slow_idx;
    L.pos = 0;
    for (;;) {
        const std::string line = L.next();
        if (L.eof) throw Fail{"unexpected end of parameter file"};
        if (has(line, "ENDFILE")) break;
        if (has(line, kcount)) types.resize(toi(tokens(line), 3));
        if (has(line, kblock)) {
            std::vector<std::string> t = tokens(L.next());
            // eight items on a triplet's INDEX line: it names a 3-body table
            // (chimesFF.cpp:882-897), kept in reading order
            if (nb == 3 && t.size() == 8) P.tab3_files.push_back(t[7]);
            int ci = toi(t, 1);
            if (ci < 0 || ci >= (int)types.size()) throw Fail{"cluster type index out of range"};
            ClusterType &T = types[ci];
            for (int a = 0; a < nb; a++) T.atoms.push_back(tok(t, 3 + a));
            t = tokens(L.next());
            for (int q = 0; q < npr; q++) T.pairs.push_back(tok(t, 1 + q));
            if (tok(t, 1 + npr) != "EXCLUDED:") {
                int K = toi(t, npr + 4);
                L.next();
                L.next();
                T.powers.resize((size_t)K * npr);
                T.coef.resize(K);
                for (int k = 0; k < K; k++) {
                    std::vector<std::string> c = tokens(L.next());
                    for (int q = 0; q < npr; q++) T.powers[(size_t)k * npr + q] = toi(c, 1 + q);
                    T.coef[k] = tod(c, npr + 3);
                }
            } else {
                T.excluded = true;
            }
        }
        if (has(line, kmaps)) {
            int n = toi(tokens(line), 1);
            for (int i = 0; i < n; i++) {
                std::vector<std::string> t = tokens(L.next());
                slow_idx.push_back(toi(t, 0));
                slow_name.push_back(tok(t, 1));
            }
        }
    }
    if (L.eof || slow_name.empty()) throw Fail{std::string("missing ") + kmaps};
    const int nt = P.natmtyps;
    const int ntup = nb == 3 ? nt * nt * nt : nt * nt * nt * nt;
    std::vector<int> &fast = nb == 3 ? P.tripmap : P.quadmap;
    fast.assign(ntup, -1);
    for (int t = 0; t < ntup; t++) {
        int ty[4], r = t;
        for (int a = nb - 1; a >= 0; a--) { ty[a] = r % nt; r /= nt; }
        std::string nm;
        for (int a = 0; a < nb; a++)
            for (int b = a + 1; b < nb; b++) nm += P.proper[ty[a] * nt + ty[b]];
        auto it = std::find(slow_name.begin(), slow_name.end(), nm);
        if (it == slow_name.end()) throw Fail{"cluster map entry missing for " + nm};
        fast[t] = slow_idx[it - slow_name.begin()];
        if (fast[t] >= (int)types.size()) throw Fail{"cluster map index out of range"};
    }
    // default inner/outer cutoffs from the pair cutoffs (chimesFF.cpp:1014-1039, 1358-1393)
    for (ClusterType &T : types) {
        int at[4];
        for (int a = 0; a < nb; a++) {
            at[a] = type_index(P, T.atoms.size() > (size_t)a ? T.atoms[a] : "");
            if (at[a] < 0) throw Fail{"unknown element in cluster type"};
        }
        int q = 0;
        for (int a = 0; a < nb; a++)
            for (int b = a + 1; b < nb; b++, q++) {
                int pt = P.pairmap[at[a] * nt + at[b]];
                T.rin[q] = P.pair_rin[pt];
                T.rout[q] = P.pair_rout[pt];
            }
    }
    // SPECIAL nB S_MAXIM / S_MINIM (chimesFF.cpp:1055-1167, 1409-1565)
    const std::string kmax = nb == 3 ? "SPECIAL 3B S_MAXIM:" : "SPECIAL 4B S_MAXIM:";
    const std::string kmin = nb == 3 ? "SPECIAL 3B S_MINIM:" : "SPECIAL 4B S_MINIM:";
    L.pos = 0;
    for (;;) {
        const std::string line = L.next();
        if (L.eof) throw Fail{"unexpected end of parameter file"};
        if (has(line, "ENDFILE")) break;
        for (int which = 1; which >= 0; which--) {
            if (!has(line, which ? kmax.c_str() : kmin.c_str())) continue;
            std::vector<std::string> t = tokens(line);
            if (tok(t, 3) == "ALL") {
                double v = tod(t, 4);
                for (ClusterType &T : types)
                    for (int q = 0; q < npr; q++) (which ? T.rout : T.rin)[q] = v;
                continue;
            }
            int nent = toi(t, 4);
            if (nb == 4 && which == 0)
                for (ClusterType &T : types)
                    for (int q = 0; q < 6; q++) T.rin[q] = -1.0;
            for (int e = 0; e < nent; e++) {
                std::vector<std::string> s = tokens(L.next());
                auto it = std::find(slow_name.begin(), slow_name.end(), tok(s, 0));
                if (it == slow_name.end()) throw Fail{"special cutoff for unknown cluster " + tok(s, 0)};
                int ci = slow_idx[it - slow_name.begin()];
                if (ci < 0 || ci >= (int)types.size()) throw Fail{"special cutoff index out of range"};
                ClusterType &T = types[ci];
                bool used[6] = {false, false, false, false, false, false};
                for (int q = 0; q < npr; q++) {
                    int hit = -1;
                    for (int r = 0; r < npr; r++)
                        if (!used[r] && T.pairs[r] == tok(s, 1 + q)) { hit = r; used[r] = true; break; }
                    if (hit < 0) throw Fail{"special cutoff pair not in cluster"};
                    (which ? T.rout : T.rin)[hit] = tod(s, 1 + npr + q);
                }
            }
        }
    }
    // permutation tables (chimesFF.cpp:2958-3080)
    std::vector<int> &mapped = nb == 3 ? P.trip_mapped : P.quad_mapped;
    std::vector<int> &perm = nb == 3 ? P.trip_perm : P.quad_perm;
    std::vector<int> &ok = nb == 3 ? P.trip_perm_ok : P.quad_perm_ok;
    mapped.assign((size_t)ntup * npr, 0);
    perm.assign((size_t)ntup * nb, 0);
    ok.assign(ntup, 0);
    for (int t = 0; t < ntup; t++) {
        if (fast[t] < 0) continue;
        int ty[4], r = t;
        for (int a = nb - 1; a >= 0; a--) { ty[a] = r % nt; r /= nt; }
        ok[t] = first_match(P, nb, ty, types[fast[t]].pairs, &perm[(size_t)t * nb], &mapped[(size_t)t * npr]);
    }
}

}  // namespace

int type_index(const ParamSet &P, const std::string &name)
{
    for (int i = 0; i < P.natmtyps; i++)
        if (P.atmtyps[i] == name) return i;
    return -1;
}

double ParamSet::max_cutoff(int body) const
{
    if (body == 2) {
        double m = pair_rout.empty() ? 0.0 : pair_rout[0];
        for (double v : pair_rout) if (v > m) m = v;
        return m;
    }
    const std::vector<ClusterType> &T = body == 3 ? trip : quad;
    if (order[body - 2] == 0 || T.empty()) return 0.0;
    int npr = body == 3 ? 3 : 6;
    double m = T[0].rout[0];
    for (const ClusterType &c : T)
        for (int q = 0; q < npr; q++) if (c.rout[q] > m) m = c.rout[q];
    return m;
}

bool load_params(const std::string &path, ParamSet &P, std::string &err)
{
    std::ifstream f(path);
    if (!f.is_open()) { err = "Cannot open parameter file: " + path; return false; }
    Lines L;
    for (std::string s; std::getline(f, s);) L.v.push_back(s);
    P = ParamSet();
    P.path = path;
    try {
        // pass 1: orders (chimesFF.cpp:390-437)
        for (;;) {
            const std::string line = L.next();
            if (L.eof || has(line, "ENDFILE")) throw Fail{"no \"PAIRTYP: CHEBYSHEV\" line"};
            if (has(line, "PAIRTYP: CHEBYSHEV")) {
                std::vector<std::string> t = tokens(line);
                P.order[0] = toi(t, 2);
                if (t.size() >= 4) P.order[1] = toi(t, 3);
                if (t.size() >= 5) P.order[2] = toi(t, 4);
                break;
            }
        }
        // pass 2: types, pairs, cutoff style, penalty, offsets (chimesFF.cpp:445-686)
        L.pos = 0;
        int no_pairs = 0;
        for (;;) {
            const std::string line = L.next();
            if (L.eof) throw Fail{"unexpected end of parameter file"};
            if (has(line, "ENDFILE")) break;
            if (has(line, "ATOM TYPES:")) {
                P.natmtyps = toi(tokens(line), 2);
                P.energy_offsets.assign(P.natmtyps, 0.0);
            }
            if (has(line, "# TYPEIDX #")) {
                P.atmtyps.resize(P.natmtyps);
                P.masses.resize(P.natmtyps);
                for (int i = 0; i < P.natmtyps; i++) {
                    std::vector<std::string> t = tokens(L.next());
                    P.atmtyps[i] = tok(t, 1);
                    P.masses[i] = tod(t, 3);
                }
            }
            if (has(line, "ATOM PAIRS:")) no_pairs = toi(tokens(line), 2);
            if (has(line, "# PAIRIDX #")) {
                if (has(line, "# USEOVRP #")) continue;
                std::string style0;
                for (int i = 0; i < no_pairs; i++) {
                    std::vector<std::string> t = tokens(L.next());
                    int sidx, midx;
                    if (t.size() == 8) { sidx = 6; midx = 7; }
                    else if (t.size() == 7) { sidx = 5; midx = 6; }
                    else throw Fail{"pair line must have 7 or 8 entries"};
                    P.pair_a1.push_back(t[1]);
                    P.pair_a2.push_back(t[2]);
                    P.pair_rin.push_back(tod(t, 3));
                    P.pair_rout.push_back(tod(t, 4));
                    if (i == 0) style0 = t[sidx];
                    else if (t[sidx] != style0) throw Fail{"distance transformation must be the same for all pairs"};
                    P.morse.push_back(style0 == "MORSE" ? tod(t, midx) : 0.0);
                }
                if (style0 != "MORSE") throw Fail{"only the MORSE distance transformation is supported"};
            }
            if (has(line, "FCUT TYPE:")) {
                std::vector<std::string> t = tokens(line);
                if (tok(t, 2) == "CUBIC") P.tersoff = false;
                else if (tok(t, 2) == "TERSOFF") { P.tersoff = true; P.fcut_var = tod(t, 3); }
                else throw Fail{"unknown FCUT TYPE: " + tok(t, 2)};
            }
            if (has(line, "PAIR CHEBYSHEV PENALTY DIST:")) P.penalty_dist = tod(tokens(line), 4);
            if (has(line, "PAIR CHEBYSHEV PENALTY SCALING:")) P.penalty_scale = tod(tokens(line), 4);
            if (has(line, "NO ENERGY OFFSETS:")) {
                std::vector<std::string> t = tokens(line);
                if (std::stoi(t.back()) != P.natmtyps) throw Fail{"number of energy offsets does not match atom types"};
                for (int i = 0; i < P.natmtyps; i++) {
                    std::vector<std::string> s = tokens(L.next());
                    int k = toi(s, 2);
                    if (k < 1 || k > P.natmtyps) throw Fail{"energy offset index out of range"};
                    P.energy_offsets[k - 1] = tod(s, 3);
                }
            }
        }
        if (P.natmtyps <= 0 || P.npairs() == 0) throw Fail{"no atom types or pairs"};
        // pass 3: 2-body coefficients and pair maps (chimesFF.cpp:704-830)
        P.pow2.assign(P.npairs(), {});
        P.c2.assign(P.npairs(), {});
        L.pos = 0;
        for (;;) {
            const std::string line = L.next();
            if (L.eof) throw Fail{"unexpected end of parameter file"};
            if (has(line, "ENDFILE")) break;
            if (has(line, "PAIRTYPE PARAMS:")) {
                // seven items: the entry also names a 2-body table
                // (chimesFF.cpp:719-745), kept in reading order
                const std::vector<std::string> pt = tokens(line);
                if (pt.size() == 7) {
                    if (P.tab2_plain) P.tab2_mixed = true;
                    P.tab2_files.push_back(pt[6]);
                } else {
                    if (!P.tab2_files.empty()) P.tab2_mixed = true;
                    P.tab2_plain = true;
                }
                int p = toi(pt, 2);
                if (p < 0 || p >= P.npairs()) throw Fail{"pair parameter index out of range"};
                L.next();
                P.pow2[p].clear();
                P.c2[p].clear();
                for (int i = 0; i < P.order[0]; i++) {
                    std::vector<std::string> t = tokens(L.next());
                    P.pow2[p].push_back(toi(t, 0));
                    P.c2[p].push_back(tod(t, 1));
                }
            }
            if (has(line, "PAIRMAPS:")) {
                int n = toi(tokens(line), 1);
                std::vector<std::string> nm;
                std::vector<int> idx;
                for (int i = 0; i < n; i++) {
                    std::vector<std::string> t = tokens(L.next());
                    idx.push_back(toi(t, 0));
                    nm.push_back(tok(t, 1));
                }
                const int nt = P.natmtyps;
                P.pairmap_slow = idx;
                P.pairmap.assign(nt * nt, 0);
                P.proper.assign(nt * nt, "");
                for (int i = 0; i < nt; i++)
                    for (int j = 0; j < nt; j++) {
                        auto it = std::find(nm.begin(), nm.end(), P.atmtyps[i] + P.atmtyps[j]);
                        if (it == nm.end()) throw Fail{"pair map entry missing for " + P.atmtyps[i] + P.atmtyps[j]};
                        int pi = idx[it - nm.begin()];
                        if (pi < 0 || pi >= P.npairs()) throw Fail{"pair map index out of range"};
                        P.pairmap[i * nt + j] = pi;
                        int pp = proper_pair(P, P.atmtyps[i], P.atmtyps[j]);
                        P.proper[i * nt + j] = P.pair_a1[pp] + P.pair_a2[pp];
                    }
            }
        }
        if (P.pairmap.empty()) throw Fail{"missing PAIRMAPS"};
        for (int p = 0; p < P.npairs(); p++)
            for (int v : P.pow2[p])
                if (v < 0 || v + 1 > P.order[0]) throw Fail{"2-body power out of range"};
        if (P.order[1] > 0) read_cluster_block(P, L, 3);
        if (P.order[2] > 0) read_cluster_block(P, L, 4);
        for (int nb = 3; nb <= 4; nb++) {
            const std::vector<ClusterType> &T = nb == 3 ? P.trip : P.quad;
            for (const ClusterType &c : T)
                for (int v : c.powers)
                    if (v < 0 || v > P.order[nb - 2]) throw Fail{"many-body power out of range"};
        }
    } catch (const Fail &e) {
        err = e.msg;
        return false;
    } catch (const std::exception &e) {
        err = std::string("malformed parameter file: ") + e.what();
        return false;
    }
    return true;
}

}  // namespace chx
