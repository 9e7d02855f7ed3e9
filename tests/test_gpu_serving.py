"""Serving: continuous batching over resident models.

  * Each request's result is compared bit for bit between the request served
    alone and the same request served inside a burst of mixed requests of
    both models and every size of the serving stream (16-4096 atoms; batched
    with others, overlapping with other batches), and with a direct
    single-configuration evaluation.
  * A sample of the served results is compared with the reference's goldens
    (tests/golden/c4_cases.json.gz): FP64 at 1e-10, FP32-mixed at 1e-6
    eV/atom and 1e-4 eV/A.
  * Failure isolation: requests that must fail (an atom with more neighbour
    images than the lists hold; a resident parameter set whose 4-body cutoff
    exceeds its 3-body cutoff) sit inside a burst; they fail alone, with
    their own error, and every other request completes with the bits it has
    alone.
  * HBM budget: under a small budget the estimated bytes of the batches in
    flight stay within it, and no batch asks for more device memory than its
    estimate.
  * result() polled without wait() returns only finished requests, and a
    released request is not returned by a later wait().
"""
import gzip
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN, assert_close, assert_mixed, golden_arrays

pytestmark = pytest.mark.gpu

C4_SAMPLE = (2, 5, 11, 16, 20, 35, 38, 44)


@pytest.fixture(scope="module")
def cc():
    import chimes_calculator_amd as cc
    return cc


@pytest.fixture(scope="module")
def served(cc, model_dir):
    from chimes_calculator_amd import workloads as W
    names = ["co_2b3b4b", "carbon_2b3b"]
    models = [cc.GpuModel(os.path.join(model_dir, n + ".txt")) for n in names]
    reqs = W.serving_stream(nreq=48, nmax=4096, seed=0xC0DE14)
    return models, names, reqs


@pytest.fixture(scope="module")
def c4_cases():
    with gzip.open(os.path.join(GOLDEN, "c4_cases.json.gz"), "rt") as f:
        return json.load(f)


def same(a, b):
    return a[0] == b[0] and np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])


@pytest.mark.parametrize("precision", [0, 1])
def test_alone_equals_in_the_mix(cc, served, c4_cases, precision):
    models, names, reqs = served
    srv = cc.GpuServer(models, precision=precision, nslots=2, max_batch_atoms=16384, hbm_budget_gb=4.0)
    alone = []
    for k, (_, mname, cfg) in enumerate(reqs):
        rid = srv.submit(names.index(mname), cfg)
        assert srv.wait(rid, 60.0) == rid
        alone.append(srv.result(rid))
    print("served alone", len(alone), flush=True)
    single_batches = srv.stats()["batches"]
    ids = [srv.submit(names.index(mname), cfg) for _, mname, cfg in reqs]
    mixed = {}
    for _ in ids:
        rid = srv.wait(0, 60.0)
        assert rid > 0
        mixed[rid] = srv.result(rid)
    st = srv.stats()
    print("precision", precision, "stats", st, "burst batches", st["batches"] - single_batches, flush=True)
    assert st["max_batch"] > 1, "the burst was never batched"
    assert st["failed"] == 0
    for k, rid in enumerate(ids):
        assert same(alone[k], mixed[rid]), "request %d differs" % k
    # the same bits as a direct evaluation of the configuration alone
    for k in (0, 7, 11, 23):
        _, mname, cfg = reqs[k]
        E, F, S = cc.eval_batch(models[names.index(mname)], [cfg], precision=precision)
        assert same((E[0], F[0], S[0]), alone[k])
    # and the reference's values
    for k in C4_SAMPLE:
        rec = c4_cases["C4|%d" % k]
        e, f, s, _ = alone[k]
        tag = "request %d (%d atoms)" % (k, len(f))
        if precision == cc.FP64:
            e0, f0, s0 = golden_arrays(rec)
            assert abs(e - e0) <= 1e-10 * abs(e0), tag
            assert_close(tag + " forces", f, f0, 1e-10, atol=1e-9)
            assert_close(tag + " stress", s, s0, 1e-10)
        else:
            assert_mixed(tag, e, f, s, rec, len(f))
    srv.close()


def crowded_box():
    """A carbon box dense enough that, under the C/O model's 4.5 A search
    radius, an atom has about 800 neighbour images, more than the 512 the
    lists hold (an error the list kernel reports)."""
    rng = np.random.default_rng(0xC0DE15)
    L = 6.0
    pos = rng.uniform(0.0, L, (900, 3))
    return pos, ["C"] * len(pos), np.diag([L, L, L])


def test_failure_isolation(cc, served, model_dir, tmp_path):
    from chimes_calculator_amd import workloads as W
    models, names, reqs = served
    # a parameter set the engine refuses at evaluation: 4-body cutoff above
    # the 3-body cutoff
    bad = W.Model("bad_rc4", ["C"], (8, 4, 3), (4.0, 3.0, 3.5), {frozenset(("C",)): 1.0}, seed=7)
    bad_model = cc.GpuModel(bad.write(str(tmp_path / "bad_rc4.txt")))
    srv = cc.GpuServer(models + [bad_model], precision=0, nslots=2, max_batch_atoms=65536, hbm_budget_gb=4.0)
    sub = reqs[:16]
    alone = []
    for _, mname, cfg in sub:
        rid = srv.submit(names.index(mname), cfg)
        assert srv.wait(rid, 60.0) == rid
        alone.append(srv.result(rid))
    ids = []
    bad_ids = []
    for k, (_, mname, cfg) in enumerate(sub):
        ids.append(srv.submit(names.index(mname), cfg))
        if k == 5:
            bad_ids.append(srv.submit(0, crowded_box()))
        if k == 9:
            bad_ids.append(srv.submit(2, W.carbon_lattice(4)))
    done = {}
    for _ in range(len(ids) + len(bad_ids)):
        rid = srv.wait(0, 60.0)
        assert rid > 0
        try:
            done[rid] = srv.result(rid)
        except cc.ChimesError as e:
            done[rid] = str(e)
    print("failures:", [done[r] for r in bad_ids], "stats", srv.stats(), flush=True)
    assert "neighbour images" in done[bad_ids[0]]
    assert "4-body cutoff" in done[bad_ids[1]]
    for k, rid in enumerate(ids):
        assert not isinstance(done[rid], str), "request %d failed: %s" % (k, done[rid])
        assert same(alone[k], done[rid]), "request %d differs" % k
    st = srv.stats()
    assert st["failed"] == 2
    srv.close()


def test_budget_bounds_bytes_in_flight(cc, served):
    models, names, reqs = served
    budget_gb = 0.25
    srv = cc.GpuServer(models, precision=0, nslots=3, max_batch_atoms=65536, hbm_budget_gb=budget_gb)
    ids = [srv.submit(names.index(mname), cfg) for _, mname, cfg in reqs]
    for _ in ids:
        rid = srv.wait(0, 60.0)
        assert rid > 0
        srv.result(rid)
    mem, st = srv.memory(), srv.stats()
    print("budget", mem, "stats", st, flush=True)
    assert st["batches"] > 3, "the budget should split the burst into several batches"
    assert mem["peak_inflight"] <= mem["budget"]
    assert mem["underestimated"] == 0 and mem["worst_ratio"] <= 1.0
    srv.close()


def test_result_polling(cc, served):
    models, names, reqs = served
    srv = cc.GpuServer(models, precision=0, nslots=2)
    _, mname, cfg = reqs[11]  # the largest request
    rid = srv.submit(names.index(mname), cfg)
    got = None
    for _ in range(200000):
        try:
            got = srv.result(rid)
            break
        except cc.ChimesError as e:
            assert "not complete" in str(e)
    assert got is not None
    # released: a later wait() returns neither it nor anything else
    assert srv.wait(0, 0.2) == 0
    with pytest.raises(cc.ChimesError):
        srv.wait(rid, 0.1)
    srv.close()
