"""The cluster fingerprint oracle (oracle/chimes_fp_oracle.c, a restatement
of chimesFF/src/FP/calc_histograms.cpp) pinned on the CPU.

- read_flat_clusters' Morse transform: the inner cutoff of every cluster
  slot maps to 1 and the outer one to -1, and the map decreases in between,
  for a two-type force field with special 3-/4-body cutoffs (so the slot
  cutoffs, not the pair cutoffs, must be the ones used); edges come out
  sorted;
- gen_flat_hists: against a vectorised brute force (full distance matrix,
  floor, bincount) on random clusters of 1, 3 and 6 edges, one set and two;
  the sample count is n1 n2 or n (n - 1) / 2.
"""
import gzip
import os
import tempfile

import numpy as np
import pytest

from conftest import GOLDEN
from oracle_bindings import Oracle, fp_edges, fp_hist

FF = "published_params.CO2400K.2+3+4b.Tersoff.special.offsets.txt.gz"


@pytest.fixture(scope="module")
def co():
    f = tempfile.NamedTemporaryFile(suffix=".txt", delete=False)
    f.write(gzip.open(os.path.join(GOLDEN, "ref_inputs", FF)).read())
    f.close()
    return Oracle(f.name)


def _cutoffs(O, nb, t, p):
    import ctypes as C
    rin, rout = C.c_double(), C.c_double()
    assert O.L.oc_cluster_pair_cutoffs(O.h, nb, np.asarray(t, np.int32), p, C.byref(rin), C.byref(rout)) == 0
    return rin.value, rout.value


@pytest.mark.parametrize("nb", [2, 3, 4])
def test_fp_transform_endpoints(co, nb):
    import itertools
    P = nb * (nb - 1) // 2
    nt = len(co.type_names())
    special = False
    for t in itertools.product(range(nt), repeat=nb):
        try:
            cuts = [_cutoffs(co, nb, t, p) for p in range(P)]
        except AssertionError:
            continue  # excluded cluster type
        if nb > 2:
            pair = [_cutoffs(co, 2, (t[a], t[b]), 0) for a in range(nb) for b in range(a + 1, nb)]
            special |= any(abs(c[1] - q[1]) > 1e-12 for c, q in zip(cuts, pair))
        for k, want in ((0, 1.0), (1, -1.0)):
            d = np.array([c[k] for c in cuts])
            e = fp_edges(co, nb, d[None, :], np.array(t)[None, :])[0]
            assert np.allclose(e, want, atol=1e-12), (t, e)
        # interior: strictly between, sorted
        d = np.array([c[0] + 0.37 * (c[1] - c[0]) for c in cuts])
        e = fp_edges(co, nb, d[None, :], np.array(t)[None, :])[0]
        assert np.all(e > -1) and np.all(e < 1) and np.all(np.diff(e) >= 0)
    if nb > 2:
        assert special  # the force field's special cutoffs differ from the pair cutoffs


def _brute(A, B, same, nbin, maxd):
    D = np.sqrt(((A[:, None, :] - B[None, :, :]) ** 2).sum(-1))
    if same:
        D = D[np.triu_indices(len(A), 1)]
    D = D.ravel()
    binw = maxd / nbin
    bins = np.floor(D / binw).astype(np.int64)
    bins[D == maxd] -= 1
    return np.bincount(bins, minlength=nbin), D.size


@pytest.mark.parametrize("P", [1, 3, 6])
@pytest.mark.parametrize("same", [True, False])
def test_fp_hist_brute_force(P, same):
    rng = np.random.default_rng(0xC0DE30 + P)
    nb = {1: 2, 3: 3, 6: 4}[P]
    A = np.sort(rng.uniform(-1, 1, (300, P)), 1)
    B = None if same else np.sort(rng.uniform(-1, 1, (250, P)), 1)
    counts, ns, binw = fp_hist(nb, A, B, nbin=100)
    want, n = _brute(A, A if same else B, same, 100, (P * 4.0) ** 0.5 + 1.0)
    assert ns == n == (300 * 299 // 2 if same else 300 * 250)
    assert counts.sum() == ns
    assert np.array_equal(counts, want)


def _co_system(O, skin=0.5):
    from lmp_setup import LmpSystem
    from oracle_bindings import read_xyz
    d = tempfile.mkdtemp()
    xyz = os.path.join(d, "co.xyz")
    with gzip.open(os.path.join(GOLDEN, "ref_inputs", "CO.9GPa_2400K.OUTCAR_#000.xyz.gz")) as a, open(xyz, "wb") as b:
        b.write(a.read())
    pos, names, cell = read_xyz(xyz)
    Lv = np.diag(np.asarray(cell, float).reshape(3, 3))
    pos = np.mod(pos, Lv)
    tn = O.type_names()
    ct = np.array([tn.index(n) for n in names], np.int32)
    return LmpSystem(pos, ct + 1, np.arange(1, len(pos) + 1, dtype=np.int32), np.zeros(3), Lv,
                     np.arange(len(tn), dtype=np.int32), O.rc2(), skin)


def _dist(x, a, b):
    d = x[b] - x[a]
    return np.sqrt(d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2])


def test_fp_rows_oracle():
    """oc_lmp_fp_rows against the row rule restated here: the pair style's
    loops (2-body pairs tag[j] > tag[i]; the 3-/4-mer lists), kept when the
    members are in order (2-body i < j; strictly increasing tags) and every
    pair is inside its slot's outer cutoff (type not excluded); distances
    as get_dist forms them; rows in loop order."""
    from oracle_bindings import LmpOracle
    f = tempfile.NamedTemporaryFile(suffix=".txt", delete=False)
    f.write(gzip.open(os.path.join(GOLDEN, "ref_inputs", FF)).read())
    f.close()
    O = LmpOracle(f.name)
    S = _co_system(O)
    trips, quads = O.build_lists(S)
    pairs = np.array([(i, j) for ii, i in enumerate(S.ilist) for j in S.nbr[S.nbr_off[ii]:S.nbr_off[ii + 1]]
                      if S.tag[j] > S.tag[i]], np.int32)
    tag = S.tag
    for nb, lst in ((2, pairs), (3, trips), (4, quads)):
        if nb == 2:
            order = lst[:, 0] < lst[:, 1]
        elif nb == 3:
            order = (tag[lst[:, 0]] < tag[lst[:, 1]]) & (tag[lst[:, 0]] < tag[lst[:, 2]]) & \
                    (tag[lst[:, 1]] < tag[lst[:, 2]])
        else:
            order = (tag[lst[:, 0]] < tag[lst[:, 1]]) & (tag[lst[:, 1]] < tag[lst[:, 2]]) & \
                    (tag[lst[:, 2]] < tag[lst[:, 3]])
        lst = lst[order]
        idx = [(a, b) for a in range(nb) for b in range(a + 1, nb)]
        dist = np.stack([_dist(S.x, lst[:, a], lst[:, b]) for a, b in idx], 1)
        types = S.ctype[lst]
        keep = np.ones(len(lst), bool)
        for k, t in enumerate(types):
            for p in range(len(idx)):
                import ctypes as C
                rin, rout = C.c_double(), C.c_double()
                if O.L.oc_cluster_pair_cutoffs(O.h, nb, np.asarray(t, np.int32), p, C.byref(rin), C.byref(rout)):
                    keep[k] = False
                    break
                if dist[k, p] >= rout.value:
                    keep[k] = False
                    break
        got_d, got_t = O.fp_rows(S, trips, quads, nb)
        assert len(got_d) > 50, nb
        assert np.array_equal(got_d, dist[keep]) and np.array_equal(got_t, types[keep]), nb
