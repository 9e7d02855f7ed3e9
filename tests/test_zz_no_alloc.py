"""The allocation-free explicit-cluster entry point
(chimes_gpu_compute_clusters_static, include/chimes_gpu.h) against
chimes_gpu_compute_clusters, through the C ABI: two calls in a row, the
second smaller than the first, for 2-, 3- and 4-body clusters of the seeded
O/H model.  Both run the same kernel on the same inputs, so the results must
be identical; a call above the scratch array's capacity is an error."""
import ctypes as C
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def clusters(rng, nb, n, rmin, rmax, ntypes):
    pairs = [(a, b) for a in range(nb) for b in range(a + 1, nb)]
    out = np.zeros((0, nb, 3))
    while len(out) < n:
        x = rng.uniform(-rmax / 2, rmax / 2, (2 * n, nb, 3))
        d = np.stack([np.linalg.norm(x[:, b] - x[:, a], axis=1) for a, b in pairs], 1)
        out = np.concatenate([out, x[((d > rmin) & (d < rmax)).all(1)]])[:n]
    dr = np.stack([out[:, b] - out[:, a] for a, b in pairs], 1)
    return (np.ascontiguousarray(np.linalg.norm(dr, axis=2).reshape(-1)), np.ascontiguousarray(dr.reshape(-1)),
            np.ascontiguousarray(rng.integers(0, ntypes, (n, nb)).astype(np.int32).reshape(-1)))


def call(L, fn, h, nb, n, dist, dr, types):
    E, F, S = np.zeros(n), np.zeros(n * nb * 3), np.zeros(n * 6)
    rc = getattr(L, fn)(h, nb, n, dist, dr, types, E, F, S)
    return rc, E, F, S


def test_static_entry_matches_allocating_entry(model_dir):
    import chimes_calculator_amd as cc
    L = cc.lib()
    m = cc.GpuModel(os.path.join(model_dir, "water_2b3b4b.txt"))
    rng = np.random.default_rng(0xC0DE07)
    for nb, rc_out in ((2, 5.0), (3, 4.5), (4, 4.0)):
        for n in (100000, 3001):  # the second call smaller than the first
            dist, dr, types = clusters(rng, nb, n, 1.0, rc_out, 2)
            print("static entry: %d-body, %d clusters" % (nb, n), flush=True)
            r0, E0, F0, S0 = call(L, "chimes_gpu_compute_clusters", m.h, nb, n, dist, dr, types)
            r1, E1, F1, S1 = call(L, "chimes_gpu_compute_clusters_static", m.h, nb, n, dist, dr, types)
            assert r0 == 0 and r1 == 0, cc._err()
            assert np.array_equal(E0, E1) and np.array_equal(F0, F1) and np.array_equal(S0, S1)
            assert np.abs(E1).max() > 0
    # above the scratch array's capacity: an error, no evaluation
    big = 131072 + 1  # CHIMES_GPU_STATIC_MAX_CLUSTERS + 1
    dist, dr, types = np.zeros(big * 6), np.zeros(big * 18), np.zeros(big * 4, np.int32)
    rc, _, _, _ = call(L, "chimes_gpu_compute_clusters_static", m.h, 4, big, dist, dr, types)
    assert rc != 0 and "out of range" in cc._err()
    m.close()
