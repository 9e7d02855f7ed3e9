"""GPU parity: the HIP path against the reference's own outputs.

Golden values (tests/golden/*.json) were produced by the compiled reference
(tests/golden/make_golden.py).  FP64 energy, forces and stress must agree to
1e-10 relative (forces and stress relative to the largest component of the
quantity, plus 1e-9 kcal/mol/A absolute for forces that vanish by symmetry);
cluster lists must be identical (sorted canonical keys, compared
in full for C0 and by SHA-256 for the larger systems).
"""
import hashlib
import os

import numpy as np
import pytest

from conftest import assert_close, golden_arrays
from oracle_bindings import read_xyz, sort_keys

pytestmark = pytest.mark.gpu

RTOL = 1e-10
FATOL = 1e-9  # kcal/mol/A: forces that cancel by symmetry (perfect crystals) are summation noise


def sha(keys):
    return hashlib.sha256(np.ascontiguousarray(sort_keys(keys), dtype="<i4").tobytes()).hexdigest()


def check_lists(model, rec, conf=0):
    p2, p3, p4 = model.cluster_lists(conf)
    assert len(p2) == rec["n_pairs"] and len(p3) == rec["n_trips"] and len(p4) == rec["n_quads"]
    assert sha(p2) == rec["sha_pairs"], "pair list differs"
    assert sha(p3) == rec["sha_trips"], "triplet list differs"
    assert sha(p4) == rec["sha_quads"], "quadruplet list differs"


def check_efs(tag, E, F, S, rec):
    e0, f0, s0 = golden_arrays(rec)
    assert abs(E - e0) <= RTOL * abs(e0), "%s: energy %.17g vs %.17g" % (tag, E, e0)
    assert_close(tag + " forces", F, f0, RTOL, atol=FATOL)
    assert_close(tag + " stress", S, s0, RTOL)


@pytest.fixture(scope="module")
def cc():
    import chimes_calculator_amd as cc
    return cc


def test_reference_cases_c_abi(cc, ref_dir, ref_cases):
    """Every reference fixture through the system C ABI (calculate_chimes)."""
    for key, rec in ref_cases.items():
        pf, xyz, small = key.split("|")
        pos, types, cell = read_xyz(os.path.join(ref_dir, xyz))
        print("case", key, flush=True)
        s = cc.ChimesSerial(small=int(small))
        s.init_chimes(os.path.join(ref_dir, pf))
        fx, fy, fz, st, e = s.calculate_chimes(len(pos), pos[:, 0], pos[:, 1], pos[:, 2], types, cell[0], cell[1],
                                               cell[2])
        s.close()
        check_efs(key, e, np.stack([fx, fy, fz], 1), st, rec)


def test_reference_cases_lists(cc, ref_dir, ref_cases):
    """Cluster lists of every reference fixture equal the reference's."""
    for key, rec in ref_cases.items():
        pf, xyz, small = key.split("|")
        pos, types, cell = read_xyz(os.path.join(ref_dir, xyz))
        print("lists", key, flush=True)
        m = cc.GpuModel(os.path.join(ref_dir, pf))
        E, F, S = cc.eval_batch(m, [(pos, types, cell)], small=int(small))
        check_efs(key, E[0], F[0], S[0], rec)
        check_lists(m, rec)


def test_c0_full_lists(cc, model_dir, synth_cases):
    from chimes_calculator_amd import workloads as W
    mname, cfgs = W.workload("C0")
    m = cc.GpuModel(os.path.join(model_dir, mname + ".txt"))
    for small in (0, 1):
        rec = synth_cases["C0|%d" % small]
        E, F, S = cc.eval_batch(m, cfgs, small=small)
        check_efs("C0", E[0], F[0], S[0], rec)
        p2, p3, p4 = m.cluster_lists(0)
        assert np.array_equal(p2, np.array(rec["pairs"], dtype=np.int32).reshape(-1, 5))
        assert np.array_equal(p3, np.array(rec["trips"], dtype=np.int32).reshape(-1, 9))
        assert len(p4) == 0


def test_c1_water(cc, model_dir, synth_cases):
    from chimes_calculator_amd import workloads as W
    mname, cfgs = W.workload("C1")
    m = cc.GpuModel(os.path.join(model_dir, mname + ".txt"))
    rec = synth_cases["C1|0"]
    E, F, S = cc.eval_batch(m, cfgs)
    check_efs("C1", E[0], F[0], S[0], rec)
    check_lists(m, rec)


def test_tersoff_wide_ramp(cc, tmp_path):
    """Carbon 2.0 Small (Tersoff ramp over 75% of the cutoff, even-only 2-body
    powers, special 3- and 4-body cutoffs) on the 512-atom diamond of
    etc/lmp/tests/test_suite-ev_tally_mb, jittered: the serial path against
    the oracle.  The fixtures above all use a ramp fraction of 0.5, where
    the ramp's width and its start coincide."""
    import gzip

    from conftest import GOLDEN
    from lmp_setup import read_lammps_data
    from oracle_bindings import Oracle
    src = os.path.join(GOLDEN, "lmp", "published_params.Carbon-2.0.Small.2+3+4b.Tersoff.txt.gz")
    params = os.path.join(str(tmp_path), "carbon_small.txt")
    with gzip.open(src) as a, open(params, "wb") as b:
        b.write(a.read())
    x, _, _, lo, hi, _ = read_lammps_data(os.path.join(GOLDEN, "lmp", "diamond_3.21gcc_512atoms.data.in.gz"))
    rng = np.random.default_rng(0xC0DE31)
    pos = x - lo + rng.uniform(-0.08, 0.08, x.shape)
    cell = np.diag(hi - lo)
    types = ["C"] * len(pos)
    R = Oracle(params).calculate(pos, types, cell)
    m = cc.GpuModel(params)
    E, F, S = cc.eval_batch(m, [(pos, types, cell)])
    m.close()
    assert abs(E[0] - R.energy) <= RTOL * abs(R.energy), "energy %.17g vs %.17g" % (E[0], R.energy)
    assert_close("forces", F[0], R.forces, RTOL, atol=FATOL)
    assert_close("stress", S[0], R.stress, RTOL)


def test_c2_batch_subset(cc, model_dir, synth_cases):
    """The first 24 configurations of the C2 batch, evaluated as one batch."""
    from chimes_calculator_amd import workloads as W
    mname, cfgs = W.workload("C2")
    m = cc.GpuModel(os.path.join(model_dir, mname + ".txt"))
    sub = cfgs[:24]
    E, F, S = cc.eval_batch(m, sub)
    for k in range(24):
        rec = synth_cases["C2|%d" % k]
        check_efs("C2[%d]" % k, E[k], F[k], S[k], rec)
        check_lists(m, rec, conf=k)


def test_dense_owners(cc, model_dir, synth_cases):
    """Owners with ~155 half-list images, more than the small enumeration
    working set holds: a cube (every searched cell in one flattened range)
    and a thin slab (more searched cells than that range takes), in one
    batch."""
    from chimes_calculator_amd import workloads as W
    m = cc.GpuModel(os.path.join(model_dir, "carbon_wide.txt"))
    tags = ("cube", "slab")
    cfgs = [W.dense_carbon((1.0, 1.0, 1.0)), W.dense_carbon((4.0, 0.5, 0.5))]
    E, F, S = cc.eval_batch(m, cfgs)
    for k, tag in enumerate(tags):
        rec = synth_cases["DENSE|" + tag]
        check_efs("dense " + tag, E[k], F[k], S[k], rec)
        check_lists(m, rec, conf=k)


def test_dense_owners_large_batch(cc, model_dir, synth_cases):
    """The same cube and slab, 129 of each in one batch (16,512 atoms): above
    16,384 atoms the neighbour search and the enumeration run one owner per
    wavefront, and owners with up to ~300 images go through them (and the
    enumeration's queue for crowded owners).  Every copy must give the
    reference's values and lists."""
    from chimes_calculator_amd import workloads as W
    m = cc.GpuModel(os.path.join(model_dir, "carbon_wide.txt"))
    tags = ("cube", "slab")
    pair = [W.dense_carbon((1.0, 1.0, 1.0)), W.dense_carbon((4.0, 0.5, 0.5))]
    cfgs = pair * 129
    assert sum(len(c[0]) for c in cfgs) > 16384
    E, F, S = cc.eval_batch(m, cfgs)
    for k in (0, 1, 128, 255, 256, 257):
        rec = synth_cases["DENSE|" + tags[k % 2]]
        check_efs("dense %s [%d]" % (tags[k % 2], k), E[k], F[k], S[k], rec)
    for k in (0, 257):
        check_lists(m, synth_cases["DENSE|" + tags[k % 2]], conf=k)


def test_lists_after_capacity_change(cc, model_dir, synth_cases):
    """A denser 64-atom box first, then the reference's dense cube (the same
    batch shape, so the slab capacity carries over and is tightened after the
    second run): the cube's exported lists must still be the reference's."""
    from chimes_calculator_amd import workloads as W
    m = cc.GpuModel(os.path.join(model_dir, "carbon_wide.txt"))
    cc.eval_batch(m, [W.dense_carbon((1.0, 1.0, 1.0), density=5.0, dmin=1.0)])
    rec = synth_cases["DENSE|cube"]
    E, F, S = cc.eval_batch(m, [W.dense_carbon((1.0, 1.0, 1.0))])
    check_efs("dense cube after a denser box", E[0], F[0], S[0], rec)
    check_lists(m, rec, conf=0)


@pytest.mark.parametrize("reps", [(4, 4, 4), (8, 8, 4)])
def test_tiled_extensivity(cc, model_dir, synth_cases, reps):
    """C0 tiled 4 x 4 x 4 (4096 atoms, more than one energy-reduction chunk)
    and 8 x 8 x 4 (16384 atoms, the 2-body kernel's large-workgroup shape):
    energy n x the reference's C0 energy, each image's forces the reference's
    C0 forces, stress the C0 stress (per-volume, so equal)."""
    from chimes_calculator_amd import workloads as W
    mname, cfgs = W.workload("C0")
    pos, types, cell = cfgs[0]
    L = np.diag(cell)
    shifts = np.array([(i, j, k) for i in range(reps[0]) for j in range(reps[1]) for k in range(reps[2])],
                      dtype=float)
    n = len(shifts)
    big = (shifts[:, None, :] * L[None, None, :] + pos[None, :, :]).reshape(-1, 3)
    m = cc.GpuModel(os.path.join(model_dir, mname + ".txt"))
    E, F, S = cc.eval_batch(m, [(big, types * n, np.diag(np.array(reps) * L))])
    e0, f0, s0 = golden_arrays(synth_cases["C0|0"])
    assert abs(E[0] - n * e0) <= 1e-10 * abs(n * e0), (E[0], n * e0)
    assert_close("tiled forces", F[0], np.tile(f0, (n, 1)), 1e-10, atol=FATOL)
    assert_close("tiled stress", S[0], s0, 1e-10)


def test_clustered_capacity_retry(cc, model_dir):
    """64 carbon atoms packed in an 8 A corner of a 20 A box: the capacity
    estimates from the mean density fall short (neighbour slabs and list
    segments), the pipeline reruns with the sizes it measured, and the result
    equals the oracle's."""
    from chimes_calculator_amd import workloads as W
    from oracle_bindings import Oracle
    rng = np.random.default_rng(0xC0DE06)
    pos = []
    while len(pos) < 64:
        p = rng.uniform(0.0, 8.0, 3)
        if all(np.linalg.norm(p - q) >= 1.2 for q in pos):
            pos.append(p)
    pos = np.array(pos) + 2.0
    cell = np.diag([20.0, 20.0, 20.0])
    types = ["C"] * 64
    pf = os.path.join(model_dir, "carbon_2b3b.txt")
    m = cc.GpuModel(pf)
    E, F, S = cc.eval_batch(m, [(pos, types, cell)])
    r = Oracle(pf).calculate(pos, types, cell)
    assert abs(E[0] - r.energy) <= RTOL * abs(r.energy), (E[0], r.energy)
    assert_close("clustered forces", F[0], r.forces, RTOL, atol=FATOL)
    assert_close("clustered stress", S[0], r.stress, RTOL)


def test_c3_carbon(cc, model_dir, c3_case):
    """The 32768-atom carbon box at 2B/3B/4B orders 16/10/4 (3-body D = 10,
    4-body D = 4, 2-body order 16): energy and forces at 1e-10 of the
    reference's, lists identical.  This box's stress is small against its
    terms (cancellation), and the reference adds ~1e10 per-term values into
    running double totals, whose rounding reaches ~2e-10 of it; the stress is
    held to 1e-10 of the oracle's compensated sum of the same terms
    (per-cluster sums added in long double), and the reference's own value
    to the same sum within 1e-9."""
    from chimes_calculator_amd import workloads as W
    mname, cfgs = W.workload("C3")
    m = cc.GpuModel(os.path.join(model_dir, mname + ".txt"))
    E, F, S = cc.eval_batch(m, cfgs)
    e0, f0, s0 = golden_arrays(c3_case)
    sx = np.array([float(v) for v in c3_case["stress_exact"]])
    print("C3 stress: GPU vs compensated %.3g, reference vs compensated %.3g (relative)"
          % (np.abs(S[0] - sx).max() / np.abs(sx).max(), np.abs(s0 - sx).max() / np.abs(sx).max()))
    assert abs(E[0] - e0) <= RTOL * abs(e0), "C3: energy %.17g vs %.17g" % (E[0], e0)
    assert_close("C3 forces", F[0], f0, RTOL, atol=FATOL)
    assert_close("C3 stress", S[0], sx, RTOL)
    assert_close("C3 reference stress", s0, sx, 1e-9)
    check_lists(m, c3_case)
