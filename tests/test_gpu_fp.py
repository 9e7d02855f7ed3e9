"""Cluster fingerprints on the GPU against the oracle.

  * Rows (chimes_lmp_fp_rows): the pair style's FINGERPRINT records, made on
    the GPU from its own neighbour rows and many-body lists after a compute,
    of a 128-atom C/O configuration (two types, special 3-/4-body cutoffs)
    under the LAMMPS neighbour rules, for 2, 3 and 4 bodies: equal to the
    oracle's rows (oc_lmp_fp_rows), bit for bit and in order.
  * Edges (chimes_gpu_fp_edges, read_flat_clusters on the GPU) of those rows:
    within 2e-15 of the oracle's (the device exponential may round an ulp
    differently from the host libm's, and an edge near 0 is a difference of
    numbers near the Morse variable's midpoint, so only the absolute error
    is small), sorted, and the histograms of the GPU's edges equal the
    oracle's histograms of its own edges bin for bin.
  * Histograms (chimes_gpu_fp_histogram): counts equal to the oracle's
    restatement of gen_flat_hists bin for bin (one set and two sets); at
    sizes the oracle cannot finish, the sample count is checked.
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
    tag = S.tag
    out = {"system": S}
    # the FINGERPRINT record rule: members in order (2-body: i < j; else
    # strictly increasing tags), type not excluded, every pair inside its
    # slot's outer cutoff (pair_chimes.cpp:625-626, 703-704, 777-778;
    # chimesFF.cpp:1691-1703, 1989-2022, 2575-2613)
    for nb, lst in ((2, pairs), (3, trips), (4, quads)):
        if nb == 2:
            lst = lst[lst[:, 0] < lst[:, 1]]
        else:
            order = np.ones(len(lst), bool)
            for a in range(nb - 1):
                order &= tag[lst[:, a]] < tag[lst[:, a + 1]]
            if nb == 3:
                order &= tag[lst[:, 0]] < tag[lst[:, 2]]
            lst = lst[order]
        idx = [(a, b) for a in range(nb) for b in range(a + 1, nb)]
        dist = np.zeros((len(lst), len(idx)))
        for q, (a, b) in enumerate(idx):  # get_dist's operation order
            d = S.x[lst[:, b]] - S.x[lst[:, a]]
            dist[:, q] = np.sqrt(d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2])
        types = S.ctype[lst]
        keep = np.ones(len(lst), bool)
        for k, t in enumerate(types):
            for q in range(len(idx)):
                rin, rout = C.c_double(), C.c_double()
                if O.L.oc_cluster_pair_cutoffs(O.h, nb, np.asarray(t, np.int32), q, C.byref(rin), C.byref(rout)) \
                        or dist[k, q] >= rout.value:
                    keep[k] = False
                    break
        out[nb] = (dist[keep], types[keep])
    return params, O, out


def test_fp_rows_from_gpu_lists(rows):
    """Rows made on the GPU from the pair style's device lists equal the
    oracle's rows of its own lists, and the oracle's rows equal the rows
    assembled in this module from the definitions (fixture `rows`)."""
    from chimes_calculator_amd import GpuModel, LmpPair
    params, O, out = rows
    S = out["system"]
    m = GpuModel(params)
    masses = {t + 1: mass for t, mass in enumerate(O.masses())}
    P = LmpPair(m, masses)
    P.compute(S.x, S.type, S.tag, S.ilist, S.numneigh, S.nbr_off, S.nbr, ago=0, skin=S.skin)
    trips, quads = O.build_lists(S)
    for nb in (2, 3, 4):
        gd, gt = P.fp_rows(nb)
        od, ot = O.fp_rows(S, trips, quads, nb)
        print("fingerprint rows, %d-body: %d" % (nb, len(gd)), flush=True)
        assert len(gd) > 100
        assert np.array_equal(gd, od) and np.array_equal(gt, ot), nb
        assert np.array_equal(od, out[nb][0]) and np.array_equal(ot, out[nb][1]), nb
    P.close()
    m.close()


def test_fp_edges_on_gpu(rows):
    from chimes_calculator_amd import GpuModel
    params, O, out = rows
    m = GpuModel(params)
    for nb in (2, 3, 4):
        dist, types = out[nb]
        assert len(dist) > 100
        got = m.fp_edges(nb, dist, types)
        want = fp_edges(O, nb, dist, types)
        err = float(np.abs(got - want).max())
        print("edges, %d-body: %d rows, %d of %d values differ, largest difference %.3g"
              % (nb, len(dist), int((got != want).sum()), got.size, err), flush=True)
        assert err <= 2e-15, nb
        assert np.array_equal(np.sort(got, 1), got), nb
        e1 = got[:4000]
        g, gns, _ = m.fp_histogram(nb, e1, nbin=100)
        w, wns, _ = fp_hist(nb, want[:4000], nbin=100)
        assert gns == wns and np.array_equal(g, w), nb
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
