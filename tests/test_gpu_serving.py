"""Serving: continuous batching over two resident models.

Each request's result is compared bit for bit between (a) the request served
alone and (b) the same request served inside a burst of mixed requests of
both models and many sizes (batched with others, overlapping with other
batches), and with a direct single-configuration evaluation.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cc():
    import chimes_calculator_amd as cc
    return cc


@pytest.fixture(scope="module")
def served(cc, model_dir):
    from chimes_calculator_amd import workloads as W
    names = ["co_2b3b4b", "carbon_2b3b"]
    models = [cc.GpuModel(os.path.join(model_dir, n + ".txt")) for n in names]
    reqs = W.serving_stream(nreq=48, nmax=700, seed=0xC0DE14)
    return models, names, reqs


@pytest.mark.parametrize("precision", [0, 1])
def test_alone_equals_in_the_mix(cc, served, precision):
    models, names, reqs = served
    srv = cc.GpuServer(models, precision=precision, nslots=2, max_batch_atoms=8192, hbm_budget_gb=4.0)
    alone = []
    for _, mname, cfg in reqs:
        rid = srv.submit(names.index(mname), cfg)
        assert srv.wait(rid, 60.0) == rid
        alone.append(srv.result(rid))
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
    for k, rid in enumerate(ids):
        e0, f0, s0, _ = alone[k]
        e1, f1, s1, _ = mixed[rid]
        assert e0 == e1 and np.array_equal(f0, f1) and np.array_equal(s0, s1), "request %d differs" % k
    # and the same bits as a direct evaluation of the configuration alone
    for k in (0, 7, 23):
        _, mname, cfg = reqs[k]
        E, F, S = cc.eval_batch(models[names.index(mname)], [cfg], precision=precision)
        assert E[0] == alone[k][0] and np.array_equal(F[0], alone[k][1]) and np.array_equal(S[0], alone[k][2])
    srv.close()
