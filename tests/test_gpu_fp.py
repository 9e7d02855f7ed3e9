"""Cluster fingerprint histograms on the GPU (chimes_gpu_fp_edges /
chimes_gpu_fp_histogram) against the oracle's restatement of
chimesFF/src/FP/calc_histograms.cpp.

Cluster rows are the pairs, triplets and quadruplets of a 128-atom C/O
configuration (two types, special 3-/4-body cutoffs) under the LAMMPS
neighbour rules, as chimesFF's FINGERPRINT mode would dump them.  Edges must
be bit-equal; histogram counts must be equal bin for bin (one set and two
sets).  At sizes the oracle cannot finish, the sample count is checked.
"""
import gzip
import itertools
import os

import numpy as np
import pytest

from conftest import GOLDEN
from lmp_setup import LmpSystem
from oracle_bindings import LmpOracle, fp_edges, fp_hist, read_xyz

pytestmark = pytest.mark.gpu

FF = "published_params.CO2400K.2+3+4b.Tersoff.special.offsets.txt.gz"
XYZ = "CO.9GPa_2400K.OUTCAR_#000.xyz.gz"


def _ungz(path, d):
    out = os.path.join(str(d), os.path.basename(path)[:-3])
    with gzip.open(path) as a, open(out, "wb") as b:
        b.write(a.read())
    return out


@pytest.fixture(scope="module")
def rows(tmp_path_factory):
    import ctypes as C
    d = tmp_path_factory.mktemp("fp")
    params = _ungz(os.path.join(GOLDEN, "ref_inputs", FF), d)
    O = LmpOracle(params)
    pos, names, cell = read_xyz(_ungz(os.path.join(GOLDEN, "ref_inputs", XYZ), d))
    Lv = np.diag(np.asarray(cell, float).reshape(3, 3))
    pos = np.mod(pos, Lv)
    tn = O.type_names()
    ct = np.array([tn.index(n) for n in names], np.int32)
    S = LmpSystem(pos, ct + 1, np.arange(1, len(pos) + 1, dtype=np.int32), np.zeros(3), Lv,
                  np.arange(len(tn), dtype=np.int32), O.rc2(), 0.5)
    trips, quads = O.build_lists(S)
    pairs = np.array([(i, j) for ii, i in enumerate(S.ilist) for j in S.nbr[S.nbr_off[ii]:S.nbr_off[ii + 1]]
                      if S.tag[j] > S.tag[i]], np.int32)
    out = {}
    for nb, lst in ((2, pairs), (3, trips), (4, quads)):
        idx = [(a, b) for a in range(nb) for b in range(a + 1, nb)]
        dist = np.stack([np.linalg.norm(S.x[lst[:, b]] - S.x[lst[:, a]], axis=1) for a, b in idx], 1)
        types = S.ctype[lst]
        ok = {}
        for t in set(map(tuple, types)):
            rin, rout = C.c_double(), C.c_double()
            ok[t] = O.L.oc_cluster_pair_cutoffs(O.h, nb, np.asarray(t, np.int32), 0, C.byref(rin),
                                                C.byref(rout)) == 0
        keep = np.array([ok[tuple(t)] for t in types])
        out[nb] = (dist[keep], types[keep])
    return params, O, out


def test_fp_edges_bit_equal(rows):
    from chimes_calculator_amd import GpuModel
    params, O, out = rows
    m = GpuModel(params)
    for nb, (dist, types) in out.items():
        assert len(dist) > 100
        got = m.fp_edges(nb, dist, types)
        want = fp_edges(O, nb, dist, types)
        assert np.array_equal(got, want), nb
    m.close()


@pytest.mark.parametrize("nb,limit", [(2, None), (3, 6000), (4, 5000)])
def test_fp_histogram_counts(rows, nb, limit):
    from chimes_calculator_amd import GpuModel
    params, O, out = rows
    m = GpuModel(params)
    dist, types = out[nb]
    e = fp_edges(O, nb, dist[:limit], types[:limit])
    for nbin in (100, 2000):
        got, ns, _ = m.fp_histogram(nb, e, nbin=nbin)
        want, wns, _ = fp_hist(nb, e, nbin=nbin)
        assert ns == wns == len(e) * (len(e) - 1) // 2 and got.sum() == ns
        assert np.array_equal(got, want), (nb, nbin)
        h = len(e) // 2
        got, ns, _ = m.fp_histogram(nb, e[:h], e[h:], nbin=nbin)
        want, wns, _ = fp_hist(nb, e[:h], e[h:], nbin=nbin)
        assert ns == wns == h * (len(e) - h) and np.array_equal(got, want), (nb, nbin)
    m.close()


def test_fp_histogram_large_and_errors(rows):
    from chimes_calculator_amd import ChimesError, GpuModel
    params, O, out = rows
    m = GpuModel(params)
    rng = np.random.default_rng(0xC0DE31)
    A = np.sort(rng.uniform(-1, 1, (60000, 3)), 1)
    counts, ns, _ = m.fp_histogram(3, A, nbin=100)
    assert ns == 60000 * 59999 // 2 and counts.sum() == ns
    B = np.sort(rng.uniform(-1, 1, (70001, 3)), 1)
    counts, ns, _ = m.fp_histogram(3, A[:50000], B, nbin=777)
    assert ns == 50000 * 70001 and counts.sum() == ns
    counts, ns, _ = m.fp_histogram(3, A[:0], nbin=100)
    assert ns == 0 and not counts.any()
    counts, ns, _ = m.fp_histogram(3, A[:1], nbin=100)
    assert ns == 0 and not counts.any()
    counts, ns, _ = m.fp_histogram(3, A[:0], B[:10], nbin=100)
    assert ns == 0 and not counts.any()
    with pytest.raises(ChimesError, match="bins"):
        m.fp_histogram(3, A[:100], nbin=9000)
    with pytest.raises(ChimesError, match="outside the histogram"):
        m.fp_histogram(3, A[:1000], nbin=10, maxd=0.5)
    m.close()
