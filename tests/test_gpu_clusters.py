"""Per-cluster API (chimescalc_C.h) on the GPU against the reference's own
compute_2B/3B/4B outputs (tests/golden/cluster_vectors.json: random clusters
of fitted force fields, including Tersoff cutoffs, special cutoffs, offsets
and the penalty region).  Energy, forces and packed stress to 1e-10 relative
(forces/stress relative to the largest component of the cluster, plus
1e-12 absolute for values that vanish).
"""
import ctypes as C
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def vectors():
    with open(os.path.join(GOLDEN, "cluster_vectors.json")) as f:
        return json.load(f)


def test_per_cluster_c_abi(ref_dir, vectors):
    import chimes_calculator_amd as cc
    from oracle_bindings import Oracle
    L = cc.lib()
    D3, D6, D9, D12 = C.c_double * 3, C.c_double * 6, C.c_double * 9, C.c_double * 12
    L.chimes_read_params.argtypes = [C.c_char_p]
    L.init_chimes.argtypes = [C.c_int]
    checked = 0
    for key, rows in sorted(vectors.items()):
        pf, nb = key.split("|")
        nb = int(nb)
        path = os.path.join(ref_dir, pf)
        names = Oracle(path).type_names()
        L.set_chimes()
        L.init_chimes(0)
        L.chimes_read_params(os.fsencode(path))
        L.chimes_build_pair_int_trip_map()
        L.chimes_build_pair_int_quad_map()
        for row in rows:
            dx = np.array(row["dx"], dtype=np.float64)
            dr = np.array(row["dr"], dtype=np.float64)
            tn = (C.c_char_p * nb)(*[names[t].encode() for t in row["types"]])
            force = (C.c_double * (3 * nb))()
            stress = D9()
            epot = C.c_double(0.0)
            if nb == 2:
                L.chimes_compute_2b_props(C.c_double(dx[0]), D3(*dr), tn, force, stress, C.byref(epot))
            elif nb == 3:
                L.chimes_compute_3b_props(D3(*dx), D9(*dr), tn, force, stress, C.byref(epot))
            else:
                L.chimes_compute_4b_props(D6(*dx), (C.c_double * 18)(*dr), tn, force, stress, C.byref(epot))
            e0 = float(row["energy"])
            f0 = np.array(row["force"], dtype=np.float64)
            s0 = np.array(row["stress"], dtype=np.float64)
            f1 = np.array(force[:])
            s1 = np.array(stress[:6])
            assert abs(epot.value - e0) <= 1e-10 * abs(e0) + 1e-14, (key, epot.value, e0)
            fs = max(np.abs(f0).max(), 1e-300)
            assert np.abs(f1 - f0).max() <= 1e-10 * fs + 1e-12, (key, f1, f0)
            ss = max(np.abs(s0).max(), 1e-300)
            assert np.abs(s1 - s0).max() <= 1e-10 * ss + 1e-12, (key, s1, s0)
            checked += 1
    print("per-cluster vectors checked:", checked)
    assert checked >= 200



def test_batched_clusters(ref_dir, vectors):
    """chimes_gpu_compute_clusters: every golden cluster of a (force field,
    body count) set in one call, against the reference's compute_nB outputs
    at the same bars as the single-cluster ABI."""
    import chimes_calculator_amd as cc
    from oracle_bindings import Oracle
    checked = 0
    for key, rows in sorted(vectors.items()):
        pf, nb = key.split("|")
        nb = int(nb)
        path = os.path.join(ref_dir, pf)
        names = Oracle(path).type_names()
        m = cc.GpuModel(path)
        dist = np.array([row["dx"] for row in rows], dtype=np.float64)
        dr = np.array([row["dr"] for row in rows], dtype=np.float64)
        types = np.stack([m.types_to_index([names[t] for t in row["types"]]) for row in rows])
        E, F, S = m.compute_clusters(nb, dist, dr, types)
        for k, row in enumerate(rows):
            e0 = float(row["energy"])
            f0 = np.array(row["force"], dtype=np.float64)
            s0 = np.array(row["stress"], dtype=np.float64)
            assert abs(E[k] - e0) <= 1e-10 * abs(e0) + 1e-14, (key, k, E[k], e0)
            fs = max(np.abs(f0).max(), 1e-300)
            assert np.abs(F[k].reshape(-1) - f0).max() <= 1e-10 * fs + 1e-12, (key, k)
            ss = max(np.abs(s0).max(), 1e-300)
            assert np.abs(S[k] - s0).max() <= 1e-10 * ss + 1e-12, (key, k)
            checked += 1
        m.close()
    print("batched cluster vectors checked:", checked)
    assert checked >= 200
