"""LAMMPS-side inputs for the pair-style tests (test infrastructure only).

Builds what LAMMPS hands PairCHIMES::compute on one process with an
orthogonal periodic box (etc/lmp/src/pair_chimes.cpp:271-288 asks for a full
neighbour list with ghosts): local atoms followed by ghost images within the
ghost cutoff of the box faces, added dimension by dimension as
Comm::borders does, each ghost carrying its atom's tag and type; and for every
local atom the indices of all other atoms (local or ghost) closer than the
type pair's 2-body cutoff plus the skin (init_one, pair_chimes.cpp:290-301).
"""
import gzip

import numpy as np


def read_lammps_data(path):
    """Atoms (atomic style), masses and an orthogonal box of a LAMMPS data file."""
    op = gzip.open if path.endswith(".gz") else open
    lines = op(path, "rt").read().splitlines()
    lo, hi, masses = np.zeros(3), np.zeros(3), {}
    atoms = []
    sec = None
    for raw in lines:
        s = raw.split("#")[0].strip()
        if not s:
            continue
        w = s.split()
        if len(w) >= 4 and w[2].endswith("lo") and w[3].endswith("hi"):
            d = "xyz".index(w[2][0])
            lo[d], hi[d] = float(w[0]), float(w[1])
            continue
        if w[0] in ("Masses", "Atoms", "Velocities"):
            sec = w[0]
            continue
        if sec == "Masses":
            masses[int(w[0])] = float(w[1])
        elif sec == "Atoms":
            atoms.append((int(w[0]), int(w[1]), float(w[2]), float(w[3]), float(w[4])))
    atoms.sort()
    tag = np.array([a[0] for a in atoms], np.int32)
    typ = np.array([a[1] for a in atoms], np.int32)
    x = np.array([a[2:] for a in atoms], np.float64)
    return x, typ, tag, lo, hi, masses


def match_types(masses, chimes_masses):
    """LAMMPS type -> ChIMES type by mass within 1e-3 (set_chimes_type,
    pair_chimes.cpp:833-879); every LAMMPS type must match."""
    out = []
    for t in sorted(masses):
        hit = [j for j, m in enumerate(chimes_masses) if abs(masses[t] - m) < 1e-3]
        assert hit, "no ChIMES type for LAMMPS type %d" % t
        out.append(hit[0])
    return np.array(out, np.int32)


def add_ghosts(x, typ, tag, lo, hi, cutghost):
    """Local atoms then ghost images x + n L (n != 0) inside
    [lo - cutghost, hi + cutghost), dimension by dimension."""
    X, T, G = x.copy(), typ.copy(), tag.copy()
    L = hi - lo
    for d in range(3):
        n_max = int(np.ceil(cutghost / L[d])) + 1
        adds = []
        for n in range(-n_max, n_max + 1):
            if n == 0:
                continue
            y = X[:, d] + n * L[d]
            sel = np.nonzero((y >= lo[d] - cutghost) & (y < hi[d] + cutghost))[0]
            if len(sel):
                Xn = X[sel].copy()
                Xn[:, d] = y[sel]
                adds.append((Xn, T[sel], G[sel]))
        for Xn, Tn, Gn in adds:
            X, T, G = np.vstack([X, Xn]), np.concatenate([T, Tn]), np.concatenate([G, Gn])
    return X, T, G


class LmpSystem:
    """x[nall][3], type (LAMMPS, 1-based), tag, ctype (ChIMES type index),
    ilist, numneigh (per atom index), packed neighbour rows nbr_off/nbr."""

    def __init__(self, x, typ, tag, lo, hi, chimes_type, rc2, skin, shuffle_seed=None):
        self.nlocal = len(x)
        cutneigh = rc2 + skin  # rc2[ct_i][ct_j]
        self.x, self.type, self.tag = add_ghosts(x, typ, tag, lo, hi, float(cutneigh.max()))
        self.x = np.ascontiguousarray(self.x)
        self.nall = len(self.x)
        self.chimes_type = np.asarray(chimes_type, np.int32)
        self.ctype = self.chimes_type[self.type - 1].astype(np.int32)
        self.ilist = np.arange(self.nlocal, dtype=np.int32)
        self.numneigh = np.zeros(self.nall, np.int32)
        rows = []
        rng = np.random.default_rng(shuffle_seed) if shuffle_seed is not None else None
        for i in range(self.nlocal):
            d = self.x - self.x[i]
            rsq = (d * d).sum(1)
            cut = cutneigh[self.ctype[i], self.ctype]
            js = np.nonzero(rsq < cut * cut)[0]
            js = js[js != i].astype(np.int32)
            if rng is not None:
                js = rng.permutation(js).astype(np.int32)
            self.numneigh[i] = len(js)
            rows.append(js)
        self.nbr_off = np.zeros(self.nlocal + 1, np.int32)
        self.nbr_off[1:] = np.cumsum([len(r) for r in rows])
        self.nbr = np.concatenate(rows).astype(np.int32) if rows else np.zeros(0, np.int32)
        self.inum = self.nlocal
        self.skin = skin

    def firstneigh_rows(self):
        return [self.nbr[self.nbr_off[ii]:self.nbr_off[ii + 1]] for ii in range(self.inum)]
