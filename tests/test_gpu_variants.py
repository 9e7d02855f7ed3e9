"""Every 3-/4-body kernel variant the autotuner can choose, pinned to the
reference's goldens.

The variants are the matrix-core kernels with 64- and 32-cluster tiles, the
vector-unit kernel, and the one-cluster-per-lane kernel with the 3-body
contraction on the matrix cores (FP32-mixed, D <= 10; the 4-body matrix-core
kernels cover D <= 4; outside those ranges the request runs on the vector
units).  Each is forced in turn through
chimes_gpu_set_variants and evaluated in FP64 and FP32-mixed on:
  * the dense-dimension scan (tests/golden/dscan_cases.json): two-type C/O
    models covering every instantiated dense dimension, 3-body D = 2..13
    and 4-body D = 2..6;
  * every reference fixture (3-body D = 4, 5, 7, 8; 4-body D = 3, 4);
  * the seeded O/H box (1536 atoms, D = 8 / 4) and the first 24
    configurations of the C/O batch (D = 7 / 3) as one batch.
Bars: FP64 energy, forces and stress within 1e-10 relative (forces plus
1e-9 kcal/mol/A absolute) and cluster lists identical; FP32-mixed within
1e-6 eV/atom, 1e-4 eV/A and 1e-5 of the largest stress component.
"""
import os

import numpy as np
import pytest

from conftest import assert_close, assert_mixed, golden_arrays
from oracle_bindings import read_xyz
from test_gpu_parity import check_lists

pytestmark = pytest.mark.gpu

RTOL = 1e-10
FATOL = 1e-9
VARIANTS = (0, 1, 2, 3)


@pytest.fixture(scope="module")
def cc():
    import chimes_calculator_amd as cc
    v = cc.get_variants()
    yield cc
    cc.set_variants(*v)


def check(cc, tag, m, cfgs, recs, precision, small=False, lists=False):
    E, F, S = cc.eval_batch(m, cfgs, precision=precision, small=small)
    worst = [0.0, 0.0, 0.0]
    for k, rec in enumerate(recs):
        t = "%s[%d]" % (tag, k)
        if precision == cc.FP64:
            e0, f0, s0 = golden_arrays(rec)
            assert abs(E[k] - e0) <= RTOL * abs(e0), "%s: energy %.17g vs %.17g" % (t, E[k], e0)
            assert_close(t + " forces", F[k], f0, RTOL, atol=FATOL)
            assert_close(t + " stress", S[k], s0, RTOL)
            if lists:
                check_lists(m, rec, conf=k)
        else:
            errs = assert_mixed(t, E[k], F[k], S[k], rec, len(cfgs[k][0]))
            worst = [max(a, b) for a, b in zip(worst, errs)]
    return worst


@pytest.mark.parametrize("precision", [0, 1])
def test_dense_dimension_scan(cc, dscan_dir, dscan_cases, precision):
    from chimes_calculator_amd import workloads as W
    for d3, d4 in W.DSCAN:
        m = cc.GpuModel(os.path.join(dscan_dir, "dscan_%d_%d.txt" % (d3, d4)))
        rec = dscan_cases["DSCAN|%d|%d" % (d3, d4)]
        cfg = [W.dscan_config(d3, d4)]
        for v in VARIANTS:
            cc.set_variants(v, v)
            print("dscan D3=%d D4=%d variant %s precision %d" % (d3, d4, cc.VARIANTS[v], precision), flush=True)
            w = check(cc, "dscan %d/%d %s" % (d3, d4, cc.VARIANTS[v]), m, cfg, [rec], precision,
                      lists=precision == cc.FP64)
            if precision:
                print("  mixed errors: %.3g eV/atom, %.3g eV/A, stress %.3g" % tuple(w))
        m.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_reference_fixtures(cc, ref_dir, ref_cases, precision):
    for key, rec in ref_cases.items():
        pf, xyz, small = key.split("|")
        pos, types, cell = read_xyz(os.path.join(ref_dir, xyz))
        m = cc.GpuModel(os.path.join(ref_dir, pf))
        for v in VARIANTS:
            cc.set_variants(v, v)
            print("fixture", key, "variant", cc.VARIANTS[v], "precision", precision, flush=True)
            check(cc, key, m, [(pos, types, cell)], [rec], precision, small=bool(int(small)))
        m.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_workload_goldens(cc, model_dir, synth_cases, precision):
    from chimes_calculator_amd import workloads as W
    for wl, n in (("C1", 1), ("C2", 24)):
        mname, cfgs = W.workload(wl)
        m = cc.GpuModel(os.path.join(model_dir, mname + ".txt"))
        recs = [synth_cases["%s|%d" % (wl, k)] for k in range(n)]
        for v3 in VARIANTS:
            for v4 in VARIANTS:
                cc.set_variants(v3, v4)
                print(wl, "variants", cc.VARIANTS[v3], cc.VARIANTS[v4], "precision", precision, flush=True)
                check(cc, "%s %s/%s" % (wl, cc.VARIANTS[v3], cc.VARIANTS[v4]), m, cfgs[:n], recs, precision)
        m.close()


def test_set_variants_rejects_unknown(cc):
    with pytest.raises(cc.ChimesError):
        cc.set_variants(4, 0)
    with pytest.raises(cc.ChimesError):
        cc.set_variants(0, -2)
