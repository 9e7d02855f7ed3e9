"""LAMMPS pair style on the GPU (chimes_lmp.h) against the oracle's
restatement of pair_chimes.cpp (oracle/chimes_lmp_oracle.c) and the
reference's own LAMMPS test log.

- many-body lists: equal to the oracle's, entry for entry, in the reference's
  loop order;
- forces (local and ghost), global energy and virial, per-atom energy and
  virial: within 1e-10 of the oracle, relative to the largest component;
- the 512-atom diamond of etc/lmp/tests/test_suite-ev_tally_mb: pressure and
  summed per-atom stress of its log;
- small boxes whose neighbour rows hold several images of one atom, rows in
  shuffled order, a two-type 4-body force field with special cutoffs and
  energy offsets, list reuse between rebuilds (ago > 0) and the tally flags.
"""
import gzip
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN
from lmp_setup import LmpSystem, match_types, read_lammps_data
from oracle_bindings import LmpOracle, read_xyz

pytestmark = pytest.mark.gpu

G = os.path.join(GOLDEN, "lmp")
NKTV2P_REAL = 68568.415


def _ungz(path, tmp_path):
    out = os.path.join(str(tmp_path), os.path.basename(path)[:-3])
    with gzip.open(path) as a, open(out, "wb") as b:
        b.write(a.read())
    return out


def _close(name, got, ref, rtol=1e-10):
    got, ref = np.asarray(got, float), np.asarray(ref, float)
    scale = max(np.abs(ref).max(), 1e-300)
    err = np.abs(got - ref).max() / scale
    assert err <= rtol, "%s: relative error %.3e" % (name, err)


def _gpu(params):
    from chimes_calculator_amd import GpuModel
    return GpuModel(params)


def _run_case(params, S, masses, ago_steps=0, seed=0):
    from chimes_calculator_amd import LmpPair
    O = LmpOracle(params)
    model = _gpu(params)
    P = LmpPair(model, masses)
    trips, quads = O.build_lists(S)
    out = P.compute(S.x, S.type, S.tag, S.ilist, S.numneigh, S.nbr_off, S.nbr, ago=0, skin=S.skin)
    g3, g4 = P.mb_lists()
    assert np.array_equal(g3, trips), "3-mer lists differ"
    assert np.array_equal(g4, quads), "4-mer lists differ"
    _compare(O, S, trips, quads, out)
    # later steps reuse the lists (ago > 0) with moved atoms
    rng = np.random.default_rng(seed)
    for step in range(ago_steps):
        S.x = np.ascontiguousarray(S.x + rng.uniform(-0.02, 0.02, S.x.shape))
        out = P.compute(S.x, S.type, S.tag, S.ilist, S.numneigh, S.nbr_off, S.nbr, ago=step + 1, skin=S.skin)
        _compare(O, S, trips, quads, out)
    P.close()
    model.close()
    return out, len(trips), len(quads)


def _compare(O, S, trips, quads, out):
    f, eng, virial, eatom, vatom = out
    R = O.compute(S, trips, quads)
    _close("forces", f, R.f)
    _close("energy", [eng], [R.eng])
    _close("virial", virial, R.virial)
    _close("eatom", eatom, R.eatom)
    _close("vatom", vatom, R.vatom)


def test_lmp_diamond_reference_log(tmp_path):
    rec = json.load(open(os.path.join(G, "ev_tally_mb.json")))
    params = _ungz(os.path.join(G, rec["params"]), tmp_path)
    O = LmpOracle(params)
    x, typ, tag, lo, hi, masses = read_lammps_data(os.path.join(G, rec["data"]))
    S = LmpSystem(x, typ, tag, lo, hi, match_types(masses, O.masses()), O.rc2(), rec["neighbor_skin"])
    (f, eng, virial, eatom, vatom), n3, n4 = _run_case(params, S, masses)
    assert n3 > 0 and n4 > 0
    V = float(np.prod(hi - lo))
    th = rec["thermo"]
    assert abs((virial[0] + virial[1] + virial[2]) / (3 * V) * NKTV2P_REAL - th["Press"]) < 1e-5
    Ssum = -vatom.sum(0) * NKTV2P_REAL
    for k, name in enumerate(("v_Sxx", "v_Syy", "v_Szz")):
        assert abs(Ssum[k] - th[name]) <= 1e-10 * abs(th[name])


@pytest.mark.parametrize("L,natoms,skin", [(7.0, 24, 1.0), (9.5, 60, 1.5)])
def test_lmp_small_box_images(tmp_path, L, natoms, skin):
    rec = json.load(open(os.path.join(G, "ev_tally_mb.json")))
    params = _ungz(os.path.join(G, rec["params"]), tmp_path)
    O = LmpOracle(params)
    rng = np.random.default_rng(0xC0DE20 + natoms)
    pos = []
    while len(pos) < natoms:
        p = rng.uniform(0, L, 3)
        if all(np.linalg.norm((p - q + L / 2) % L - L / 2) >= 1.3 for q in pos):
            pos.append(p)
    masses = {1: O.masses()[0]}
    S = LmpSystem(np.array(pos), np.ones(natoms, np.int32), np.arange(1, natoms + 1, dtype=np.int32),
                  np.zeros(3), np.full(3, L), [0], O.rc2(), skin, shuffle_seed=natoms)
    assert any(len(set(S.tag[r])) < len(r) for r in S.firstneigh_rows())
    _run_case(params, S, masses, ago_steps=2, seed=natoms)


@pytest.mark.parametrize("ff,xyz", [
    # two types, special 3-/4-body cutoffs, energy offsets
    ("published_params.CO2400K.2+3+4b.Tersoff.special.offsets.txt.gz", "CO.9GPa_2400K.OUTCAR_#000.xyz.gz"),
    # two types (N, H), 2+3+4-body, special cutoffs, offsets, 512 atoms
    ("published_params.HN3.2+3+4b.Tersoff.special.offsets.txt.gz", "HN3.2gcc_3000K.OUTCAR_#000.xyz.gz"),
    # one type, 2+3-body only (no quadruplets), cubic smoothing
    ("published_params.liqC.2+3b.cubic.txt.gz", "liqC.2.5gcc_6000K.OUTCAR_#000.xyz.gz"),
])
def test_lmp_force_fields(tmp_path, ff, xyz):
    params = _ungz(os.path.join(GOLDEN, "ref_inputs", ff), tmp_path)
    O = LmpOracle(params)
    pos, names, cell = read_xyz(_ungz(os.path.join(GOLDEN, "ref_inputs", xyz), tmp_path))
    Lv = np.diag(np.asarray(cell, float).reshape(3, 3))
    pos = np.mod(np.asarray(pos, float), Lv)
    tn = O.type_names()
    ct = np.array([tn.index(n) for n in names], np.int32)
    # LAMMPS types 1..nt in the parameter file's order, masses from it
    masses = {t + 1: m for t, m in enumerate(O.masses())}
    S = LmpSystem(pos, ct + 1, np.arange(1, len(pos) + 1, dtype=np.int32), np.zeros(3), Lv,
                  np.arange(len(tn), dtype=np.int32), O.rc2(), 1.0, shuffle_seed=3)
    _, n3, n4 = _run_case(params, S, masses, ago_steps=1, seed=5)
    assert n3 > 0 and (n4 > 0) == (O.order(4) > 0)


def test_lmp_tally_flags(tmp_path):
    """Forces alone, or with global tallies only, are the full call's."""
    from chimes_calculator_amd import LmpPair
    rec = json.load(open(os.path.join(G, "ev_tally_mb.json")))
    params = _ungz(os.path.join(G, rec["params"]), tmp_path)
    O = LmpOracle(params)
    x, typ, tag, lo, hi, masses = read_lammps_data(os.path.join(G, rec["data"]))
    S = LmpSystem(x, typ, tag, lo, hi, match_types(masses, O.masses()), O.rc2(), 1.0)
    model = _gpu(params)
    P = LmpPair(model, masses)
    full = P.compute(S.x, S.type, S.tag, S.ilist, S.numneigh, S.nbr_off, S.nbr, ago=0, skin=S.skin)
    bare = P.compute(S.x, S.type, S.tag, S.ilist, S.numneigh, S.nbr_off, S.nbr, ago=1, skin=S.skin,
                     eflag_global=0, eflag_atom=0, vflag_global=0, vflag_atom=0)
    glob = P.compute(S.x, S.type, S.tag, S.ilist, S.numneigh, S.nbr_off, S.nbr, ago=1, skin=S.skin,
                     eflag_atom=0, vflag_atom=0)
    assert np.array_equal(full[0], bare[0]) and np.array_equal(full[0], glob[0])
    assert bare[1] == 0.0 and not bare[2].any() and not bare[3].any() and not bare[4].any()
    assert glob[1] == full[1] and np.array_equal(glob[2], full[2])
    assert not glob[3].any() and not glob[4].any()
    P.close()
    model.close()


def test_lmp_unmatched_type_is_an_error(tmp_path):
    from chimes_calculator_amd import ChimesError, LmpPair
    rec = json.load(open(os.path.join(G, "ev_tally_mb.json")))
    model = _gpu(_ungz(os.path.join(G, rec["params"]), tmp_path))
    with pytest.raises(ChimesError, match="matches no ChIMES atom type"):
        LmpPair(model, {1: 12.0107, 2: 39.948})
    model.close()


def test_lmp_empty_and_ghost_only(tmp_path):
    """A process without atoms, and one whose atoms are all ghosts (inum = 0),
    add nothing, as the reference's loops do."""
    from chimes_calculator_amd import LmpPair
    rec = json.load(open(os.path.join(G, "ev_tally_mb.json")))
    params = _ungz(os.path.join(G, rec["params"]), tmp_path)
    model = _gpu(params)
    P = LmpPair(model, {1: 12.0107})
    e = np.zeros(0, np.int32)
    f, eng, vir, ea, va = P.compute(np.zeros((0, 3)), e, e, e, e, np.zeros(1, np.int32), e, ago=0, skin=1.0)
    assert f.size == 0 and eng == 0.0 and not vir.any()
    x = np.array([[0.0, 0.0, 0.0], [1.5, 0.0, 0.0]])
    one = np.ones(2, np.int32)
    f, eng, vir, ea, va = P.compute(x, one, np.array([1, 2], np.int32), e, np.zeros(2, np.int32),
                                    np.zeros(1, np.int32), e, ago=0, skin=1.0)
    assert not f.any() and eng == 0.0 and not vir.any() and not ea.any() and not va.any()
    P.close()
    model.close()
