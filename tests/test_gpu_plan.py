"""Batch plans (captured-graph replays) against direct evaluation.

Every replay is compared with an eager evaluation of the same inputs made
between replays; the plan's own kernels (tune off) give identical bits, the
tuned variants agree to 1e-12 relative.  The first 24 configurations of the
batch are also checked against the reference goldens at 1e-10.
"""
import os

import numpy as np
import pytest

from conftest import assert_close, golden_arrays

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cc():
    import chimes_calculator_amd as cc
    return cc


@pytest.fixture(scope="module")
def c2(cc, model_dir):
    from chimes_calculator_amd import workloads as W
    mname, cfgs = W.workload("C2")
    return cc.GpuModel(os.path.join(model_dir, mname + ".txt")), cfgs[:96]


def jitter(cfgs, seed, amp):
    rng = np.random.default_rng(seed)
    return [(np.asarray(p) + rng.uniform(-amp, amp, np.shape(p)), t, c) for p, t, c in cfgs]


def same(a, b, rtol):
    Ea, Fa, Sa = a
    Eb, Fb, Sb = b
    if rtol == 0:
        assert np.array_equal(Ea, Eb)
        assert all(np.array_equal(x, y) for x, y in zip(Fa, Fb))
        assert np.array_equal(Sa, Sb)
    else:
        assert np.allclose(Ea, Eb, rtol=rtol, atol=0)
        for x, y in zip(Fa, Fb):
            assert np.abs(x - y).max() <= rtol * max(np.abs(y).max(), 1.0)
        assert np.abs(Sa - Sb).max() <= rtol * np.abs(Sb).max()


@pytest.mark.parametrize("precision", [0, 1])
def test_plan_replays_match_eager(cc, c2, precision):
    m, cfgs = c2
    plan = cc.GpuPlan(m, cfgs, precision=precision, tune=False)
    same(plan.first, cc.eval_batch(m, cfgs, precision=precision), 0)
    for rep in range(4):
        new = jitter(cfgs, 100 + rep, 0.02)
        got = plan.run(new)
        ref = cc.eval_batch(m, new, precision=precision)   # eager work between replays
        print("replay", rep, "precision", precision, flush=True)
        same(got, ref, 0)
    plan.close()


def test_plan_tuned_and_goldens(cc, c2, synth_cases):
    m, cfgs = c2
    plan = cc.GpuPlan(m, cfgs, tune=True)
    v3, v4, tms, counts = plan.info()
    print("tuned variants: 3-body", v3, "4-body", v4, "ms", tms.tolist(), "counts", counts)
    assert (tms > 0).any()
    E, F, S = plan.run(cfgs)
    same((E, F, S), cc.eval_batch(m, cfgs), 1e-12)
    for k in range(24):
        e0, f0, s0 = golden_arrays(synth_cases["C2|%d" % k])
        assert abs(E[k] - e0) <= 1e-10 * abs(e0)
        assert_close("C2 plan [%d] forces" % k, F[k], f0, 1e-10, atol=1e-9)
        assert_close("C2 plan [%d] stress" % k, S[k], s0, 1e-10)
    plan.close()


def test_plan_recaptures_when_counts_grow(cc, c2):
    """Compressing every cell raises the cluster counts beyond the captured
    buffers; the run falls back to a direct evaluation and captures again."""
    m, cfgs = c2
    plan = cc.GpuPlan(m, cfgs[:16], tune=False)
    dense = [(np.asarray(p) * 0.93, t, np.asarray(c) * 0.93) for p, t, c in cfgs[:16]]
    got = plan.run(dense)
    same(got, cc.eval_batch(m, dense), 0)
    got2 = plan.run(dense)
    same(got2, got, 0)
    plan.close()
