"""Seeded synthetic ChIMES workloads: parameter files and geometries.

The parameter files follow the ChIMES parameter grammar the reference parses
(chimesFF/src/chimesFF.cpp:354-1570): a 2-body block per pair type, and
permutation-expanded 3-body / 4-body term lists per cluster type, whose
coefficients are equal on every power tuple related by a type-preserving atom
permutation, so the energy of a cluster does not depend on member order.

Choices the workload descriptions leave open, fixed here and written into each
file's header:
  * coefficients are N(0, sigma) * 1/(1 + d), d = total Chebyshev degree of the
    term (2-body: pow + 1; 3/4-body: sum of the pair powers);
  * the 3-body term set is every power triple in [0, n)^3 covering all three
    atoms (>= 2 nonzero powers); the 4-body set is every 6-tuple in [0, n)^6
    covering all four atoms (636 terms for n = 3, 3861 for n = 4);
  * Morse lambda = 1.4 A for every pair, cubic smooth cutoff;
  * inner cutoffs are chosen below the closest approach each generator allows,
    so the penalty and out-of-range branches stay quiet on these inputs.
"""
import itertools
import math

import numpy as np

AMU_PER_A3_TO_GCC = 1.66053906660  # 1 amu / A^3 in g/cc
MASS = {"C": 12.011, "O": 15.999, "H": 1.008, "N": 14.007}

# ---------------------------------------------------------------------------
# Term sets
# ---------------------------------------------------------------------------

_PAIRS3 = [(0, 1), (0, 2), (1, 2)]
_PAIRS4 = [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]


def _covers(powers, pairs, natoms):
    hit = [False] * natoms
    for p, (a, b) in zip(powers, pairs):
        if p > 0:
            hit[a] = hit[b] = True
    return all(hit)


def term_set(nbody, order):
    """Power tuples of the permutation-expanded term list, lexicographic."""
    pairs = _PAIRS3 if nbody == 3 else _PAIRS4
    return [t for t in itertools.product(range(order), repeat=len(pairs))
            if _covers(t, pairs, nbody)]


def _pair_perm(sigma, pairs):
    """Pair-slot permutation induced by atom permutation sigma: new slot q takes
    the power of the old slot holding the pair (sigma[a], sigma[b])."""
    idx = {frozenset(p): k for k, p in enumerate(pairs)}
    return [idx[frozenset((sigma[a], sigma[b]))] for (a, b) in pairs]


def _orbit_rep(t, perms):
    return min(tuple(t[k] for k in pp) for pp in perms)


# ---------------------------------------------------------------------------
# Parameter file writer
# ---------------------------------------------------------------------------

class Model:
    """A synthetic ChIMES model description."""

    def __init__(self, name, elements, orders, rout, rin, lam=1.4, sigma=0.05,
                 seed=0, offsets=None, fcut="CUBIC"):
        self.name = name
        self.elements = list(elements)
        self.orders = tuple(orders)
        self.rout = tuple(rout)          # (2B, 3B, 4B) outer cutoffs
        self.rin = dict(rin)             # frozenset pair -> inner cutoff
        self.lam = lam
        self.sigma = sigma if isinstance(sigma, (tuple, list)) else (sigma,) * 3
        self.seed = seed
        self.offsets = offsets
        self.fcut = fcut

    def proper(self, a, b):
        ia, ib = self.elements.index(a), self.elements.index(b)
        if ia > ib:
            a, b = b, a
        return a + b

    def pair_list(self):
        E = self.elements
        return [(E[i], E[j]) for i in range(len(E)) for j in range(i, len(E))]

    def write(self, path):
        rng = np.random.default_rng(self.seed)
        E = self.elements
        nt = len(E)
        o2, o3, o4 = self.orders
        pairs = self.pair_list()
        trips = list(itertools.combinations_with_replacement(range(nt), 3)) if o3 else []
        quads = list(itertools.combinations_with_replacement(range(nt), 4)) if o4 else []
        L = []
        w = L.append
        w("! Synthetic ChIMES model '%s' (seed %d)" % (self.name, self.seed))
        w("! coefficients N(0, sigma)/(1 + degree), sigma (2B,3B,4B) = %s" % (self.sigma,))
        w("! Morse lambda %.3f A for all pairs; %s cutoff" % (self.lam, self.fcut))
        w("")
        w("PAIRTYP: CHEBYSHEV  %d %d %d -1 1" % (o2, o3, o4))
        w("")
        w("ATOM TYPES: %d" % nt)
        w("")
        w("# TYPEIDX #\t# ATM_TYP #\t# ATMCHRG #\t# ATMMASS #")
        for i, e in enumerate(E):
            w("%d\t\t%s\t\t0\t\t%.4f" % (i, e, MASS[e]))
        w("")
        w("ATOM PAIRS: %d" % len(pairs))
        w("")
        w("# PAIRIDX #\t# ATM_TY1 #\t# ATM_TY1 #\t# S_MINIM #\t# S_MAXIM #\t# S_DELTA #\t# CHBDIST #\t# MORSE_LAMBDA #")
        for i, (a, b) in enumerate(pairs):
            rin = self.rin[frozenset((a, b))]
            w("\t%d\t%s\t%s\t%.6f\t%.6f\t0.01\tMORSE\t%.6f" % (i, a, b, rin, self.rout[0], self.lam))
        w("")
        w("FCUT TYPE: %s" % self.fcut)
        w("")
        if o3:
            w("SPECIAL 3B S_MAXIM: ALL %.6f" % self.rout[1])
            w("")
        if o4:
            w("SPECIAL 4B S_MAXIM: ALL %.6f" % self.rout[2])
            w("")
        if o3:
            w("ATOM PAIR TRIPLETS: %d" % len(trips))
        if o4:
            w("ATOM PAIR QUADRUPLETS: %d" % len(quads))
        w("")
        w("PAIR CHEBYSHEV PARAMS")
        w("")
        for i, (a, b) in enumerate(pairs):
            w("PAIRTYPE PARAMS: %d %s %s" % (i, a, b))
            w("")
            for p in range(o2):
                c = rng.normal(0.0, self.sigma[0]) / (1.0 + (p + 1))
                w("%d %.17g" % (p, c))
            w(" ")
        for nb, clusters, order, label in ((3, trips, o3, "TRIPLET"), (4, quads, o4, "QUADRUPLET")):
            if not clusters:
                continue
            w("%s CHEBYSHEV PARAMS" % label)
            w("")
            apairs = _PAIRS3 if nb == 3 else _PAIRS4
            terms = term_set(nb, order)
            for ci, tys in enumerate(clusters):
                names = [E[t] for t in tys]
                sym = [s for s in itertools.permutations(range(nb))
                       if all(tys[s[a]] == tys[a] for a in range(nb))]
                perms = [_pair_perm(s, apairs) for s in sym]
                reps, coef = {}, {}
                for t in terms:
                    r = _orbit_rep(t, perms)
                    if r not in coef:
                        reps[r] = len(reps)
                        coef[r] = rng.normal(0.0, self.sigma[nb - 2]) / (1.0 + sum(r))
                first = {}
                for k, t in enumerate(terms):
                    first.setdefault(_orbit_rep(t, perms), k)
                pnames = [self.proper(names[a], names[b]) for a, b in apairs]
                w("%sTYPE PARAMS:" % ("TRIPLET" if nb == 3 else "QUADRUPLE"))
                w("   INDEX: %d ATOMS: %s" % (ci, " ".join(names)))
                w("PAIRS: %s UNIQUE: %d TOTAL: %d" % (" ".join(pnames), len(reps), len(terms)))
                w("     index  |  powers  |  equiv index  |  param index  |       parameter       ")
                w("   ----------------------------------------------------------------------------")
                for k, t in enumerate(terms):
                    r = _orbit_rep(t, perms)
                    w("%d %s %d %d %.17g" % (k, " ".join(str(v) for v in t), first[r], reps[r], coef[r]))
                w("")
        # maps
        pm = {}
        for i in range(nt):
            for j in range(nt):
                pm.setdefault(E[i] + E[j], pairs.index(tuple(sorted((E[i], E[j]), key=E.index))))
        w("PAIRMAPS: %d" % len(pm))
        for k, v in pm.items():
            w("%d %s" % (v, k))
        w("")
        for nb, clusters in ((3, trips), (4, quads)):
            if not clusters:
                continue
            apairs = _PAIRS3 if nb == 3 else _PAIRS4
            mp = {}
            for tup in itertools.product(range(nt), repeat=nb):
                nm = "".join(self.proper(E[tup[a]], E[tup[b]]) for a, b in apairs)
                mp.setdefault(nm, clusters.index(tuple(sorted(tup))))
            w("%s: %d" % ("TRIPMAPS" if nb == 3 else "QUADMAPS", len(mp)))
            for k, v in mp.items():
                w("%d %s" % (v, k))
            w("")
        if self.offsets:
            w("NO ENERGY OFFSETS: %d" % nt)
            for i, e in enumerate(E):
                w("ENERGY OFFSET %d %.10f" % (i + 1, self.offsets[e]))
        w("ENDFILE")
        with open(path, "w") as f:
            f.write("\n".join(L) + "\n")
        return path


def _rin_all(elements, v):
    return {frozenset((a, b)): v for a in elements for b in elements}


# Models of the five workloads (0-based: C0 ... C4).
MODELS = {
    # C0: single-element carbon 2B+3B, orders 12/7, rc 4.0 / 3.5
    "carbon_2b3b": Model("carbon_2b3b", ["C"], (12, 7, 0), (4.0, 3.5, 0.0),
                         _rin_all(["C"], 1.0), sigma=0.1, seed=0xC0DE00),
    # C1: O/H water-like, full 2B+3B+4B at orders 12/8/4, rc 5.0 / 4.5 / 4.0
    "water_2b3b4b": Model("water_2b3b4b", ["O", "H"], (12, 8, 4), (5.0, 4.5, 4.0),
                          {frozenset(("O",)): 1.0, frozenset(("O", "H")): 0.7,
                           frozenset(("H",)): 0.7},
                          sigma=0.05, seed=0xC0DE01, offsets={"O": -10.0, "H": -2.0}),
    # C2 (batch) and C4 set A: C/O, orders 12/7/3, rc 4.5 / 4.0 / 3.5
    "co_2b3b4b": Model("co_2b3b4b", ["C", "O"], (12, 7, 3), (4.5, 4.0, 3.5),
                       _rin_all(["C", "O"], 0.6), sigma=0.05, seed=0xC0DE02),
    # C3: carbon at scale, orders 16/10/4, rc 4.5 / 4.0 / 3.5
    "carbon_16_10_4": Model("carbon_16_10_4", ["C"], (16, 10, 4), (4.5, 4.0, 3.5),
                            _rin_all(["C"], 1.0), sigma=0.05, seed=0xC0DE03),
    # test case, not a workload: a 7.5 A 2-body cutoff at 3.5 g/cc gives each
    # owner ~155 half-list images (crowded-owner enumeration path)
    "carbon_wide": Model("carbon_wide", ["C"], (8, 5, 3), (7.5, 4.0, 3.5),
                         _rin_all(["C"], 1.0), sigma=0.05, seed=0xC0DE05),
}


# ---------------------------------------------------------------------------
# Geometries
# ---------------------------------------------------------------------------

def _min_image_min_dist(pos, L):
    d = pos[:, None, :] - pos[None, :, :]
    d -= L * np.round(d / L)
    r = np.sqrt((d * d).sum(-1))
    np.fill_diagonal(r, np.inf)
    return r.min()


def carbon_lattice(nside=4, a=2.1525, jitter=0.25, seed=0xC0DE00, min_dist=1.1):
    """Jittered simple-cubic carbon lattice (C0: nside 4, L = 8.61 A;
    C3: nside 32, L = 68.88 A).  Returns (pos, types, cell)."""
    rng = np.random.default_rng(seed)
    L = nside * a
    g = (np.stack(np.meshgrid(*[np.arange(nside)] * 3, indexing="ij"), -1).reshape(-1, 3) + 0.5) * a
    while True:
        pos = g + rng.uniform(-jitter, jitter, g.shape)
        if nside > 8 or _min_image_min_dist(pos, L) >= min_dist:
            break
    return pos, ["C"] * len(pos), np.diag([L, L, L])


def dense_carbon(aspect=(1.0, 1.0, 1.0), natoms=64, density=3.5, dmin=1.2, seed=0xC0DE05):
    """natoms carbon atoms at seeded random sites (minimum image distance >=
    dmin) in an orthorhombic box of the given density whose edge lengths are
    in the ratio aspect.  Returns (pos, types, cell)."""
    rng = np.random.default_rng(seed)
    vol = natoms * MASS["C"] / density * AMU_PER_A3_TO_GCC
    a = np.asarray(aspect, dtype=float)
    Lv = a * (vol / a.prod()) ** (1.0 / 3.0)
    pos = []
    while len(pos) < natoms:
        p = rng.uniform(0.0, 1.0, 3) * Lv
        ok = True
        for q in pos:
            d = p - q
            d -= Lv * np.round(d / Lv)
            if np.sqrt((d * d).sum()) < dmin:
                ok = False
                break
        if ok:
            pos.append(p)
    return np.array(pos), ["C"] * natoms, np.diag(Lv)


def water_box(nmol=512, L=24.8, roh=0.97, angle=104.5, doo=2.5, dmin=1.0, seed=0xC0DE01):
    """O at seeded random non-overlapping sites (O-O >= doo, any intermolecular
    distance >= dmin), each with two H at roh and the given H-O-H angle in a
    random orientation (C1).  Types are ordered O, H, H per molecule."""
    rng = np.random.default_rng(seed)
    half = math.radians(angle) / 2.0
    local = np.array([[0.0, 0.0, 0.0],
                      [roh * math.sin(half), 0.0, roh * math.cos(half)],
                      [-roh * math.sin(half), 0.0, roh * math.cos(half)]])
    mols = np.zeros((0, 3, 3))
    while len(mols) < nmol:
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        w_, x_, y_, z_ = q
        R = np.array([[1 - 2 * (y_ * y_ + z_ * z_), 2 * (x_ * y_ - z_ * w_), 2 * (x_ * z_ + y_ * w_)],
                      [2 * (x_ * y_ + z_ * w_), 1 - 2 * (x_ * x_ + z_ * z_), 2 * (y_ * z_ - x_ * w_)],
                      [2 * (x_ * z_ - y_ * w_), 2 * (y_ * z_ + x_ * w_), 1 - 2 * (x_ * x_ + y_ * y_)]])
        m = local @ R.T + rng.uniform(0, L, 3)
        if len(mols):
            oo = mols[:, 0, :] - m[0]
            oo -= L * np.round(oo / L)
            if (oo * oo).sum(-1).min() < doo * doo:
                continue
            d = mols.reshape(-1, 1, 3) - m[None, :, :]
            d -= L * np.round(d / L)
            if (d * d).sum(-1).min() < dmin * dmin:
                continue
        mols = np.concatenate([mols, m[None]], 0)
    pos = mols.reshape(-1, 3)
    types = ["O", "H", "H"] * nmol
    return pos, types, np.diag([L, L, L])


def co_config(rng, natoms, density, frac_c, jitter=0.15, elements=("C", "O")):
    """One jittered-lattice C/O configuration (C2 / C4 style)."""
    n_c = int(round(frac_c * natoms))
    types = [elements[0]] * n_c + [elements[1]] * (natoms - n_c)
    rng.shuffle(types)
    mass = sum(MASS[t] for t in types)
    L = (mass / density * AMU_PER_A3_TO_GCC) ** (1.0 / 3.0)
    ns = int(math.ceil(natoms ** (1.0 / 3.0) - 1e-9))
    a = L / ns
    sites = (np.stack(np.meshgrid(*[np.arange(ns)] * 3, indexing="ij"), -1).reshape(-1, 3) + 0.5) * a
    sel = rng.choice(len(sites), natoms, replace=False)
    pos = sites[np.sort(sel)] + rng.uniform(-jitter * a, jitter * a, (natoms, 3))
    return pos, types, np.diag([L, L, L])


def co_batch(nconf=4096, nmin=8, nmax=512, rho=(1.5, 3.5), frac=(0.3, 0.7), seed=0xC0DE02):
    """C2: nconf C/O configurations, sizes log-uniform in [nmin, nmax],
    densities uniform in rho (g/cc), carbon fraction uniform in frac."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(nconf):
        n = int(round(math.exp(rng.uniform(math.log(nmin), math.log(nmax)))))
        out.append(co_config(rng, n, rng.uniform(*rho), rng.uniform(*frac)))
    return out


def serving_stream(nreq=256, rate=200.0, nmin=16, nmax=4096, seed=0xC0DE04):
    """C4: Poisson arrivals; each request is a C/O configuration for model
    'co_2b3b4b' or a carbon configuration for model 'carbon_2b3b', with size
    log-uniform in [nmin, nmax].  Returns [(t_arrival, model, (pos, types, cell))]."""
    rng = np.random.default_rng(seed)
    t = 0.0
    reqs = []
    for _ in range(nreq):
        t += rng.exponential(1.0 / rate)
        n = int(round(math.exp(rng.uniform(math.log(nmin), math.log(nmax)))))
        if rng.uniform() < 0.5:
            cfg = co_config(rng, n, rng.uniform(1.5, 3.5), rng.uniform(0.3, 0.7))
            reqs.append((t, "co_2b3b4b", cfg))
        else:
            cfg = co_config(rng, n, 2.0, 1.0, elements=("C", "C"))
            reqs.append((t, "carbon_2b3b", cfg))
    return reqs


# Test cases, not workloads: two-type C/O models whose 3-/4-body dense
# dimensions (D3, D4) together cover every dimension the evaluation kernels
# are instantiated for (3-body 2..13, 4-body 2..6), each with one small
# jittered-lattice configuration.
DSCAN = [(2, 2), (3, 3), (4, 4), (5, 5), (6, 6), (7, 2), (8, 3), (9, 4), (10, 2), (11, 3), (12, 4), (13, 2)]


def dscan_model(d3, d4):
    return Model("dscan_%d_%d" % (d3, d4), ["C", "O"], (8, d3, d4), (4.5, 4.0, 3.5), _rin_all(["C", "O"], 0.6),
                 sigma=(0.05, 0.5, 2.0), seed=0xC0DE10 + 16 * d3 + d4)


def dscan_config(d3, d4):
    """One C/O configuration for the (d3, d4) model: 96 atoms at 2.4 g/cc
    (48 for the large 4-body dimensions, whose reference evaluation is slow)."""
    rng = np.random.default_rng(0xC0DE20 + 16 * d3 + d4)
    return co_config(rng, 48 if d4 >= 5 else 96, 2.4, 0.5)


WORKLOADS = {
    "C0": "carbon 64 atoms, L=8.61 A, 2B+3B 12/7, rc 4.0/3.5, CPU-runnable",
    "C1": "O/H 1536 atoms (512 H2O-like), L=24.8 A, 2B+3B+4B 12/8/4, rc 5.0/4.5/4.0",
    "C2": "4096 C/O configurations, 8-512 atoms, 1.5-3.5 g/cc, 12/7/3, rc 4.5/4.0/3.5",
    "C3": "carbon 32768 atoms, L=68.88 A (2.0 g/cc), 16/10/4, rc 4.5/4.0/3.5",
    "C4": "Poisson stream of 16-4096-atom C/O and carbon requests, two resident models",
}


def workload(name):
    """(model name, [(pos, types, cell), ...]) for the single-system workloads."""
    if name == "C0":
        return "carbon_2b3b", [carbon_lattice(4, seed=0xC0DE00)]
    if name == "C1":
        return "water_2b3b4b", [water_box()]
    if name == "C2":
        return "co_2b3b4b", co_batch()
    if name == "C3":
        return "carbon_16_10_4", [carbon_lattice(32, a=2.1525, seed=0xC0DE03)]
    raise KeyError(name)
