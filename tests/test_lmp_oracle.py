"""The LAMMPS pair-style oracle (oracle/chimes_lmp_oracle.c) pinned against
the reference's own LAMMPS test, the serial path and the definitions of its
lists (CPU only).

- etc/lmp/tests/test_suite-ev_tally_mb (512-atom diamond, Carbon-2.0 2+3+4-body,
  neighbor skin 1.5, real units): the step-0 pressure and the summed per-atom
  stress Sxx, Syy, Szz of its expected log (tests/golden/lmp/ev_tally_mb.json);
- the serial oracle on the same box (wider than twice the padded cutoffs, so
  the tag rule and the parent rule pick the same clusters);
- the 3-/4-mer lists of a small box (images of one atom in one neighbour row)
  against a vectorised enumeration from the definitions of pair_chimes.cpp:
  321-478, in the reference's loop order;
- the per-atom split of ev_tally_mb adds up to the global tallies.
"""
import gzip
import json
import os
import tempfile

import numpy as np
import pytest

from lmp_setup import LmpSystem, match_types, read_lammps_data
from oracle_bindings import LmpOracle

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "lmp")
NKTV2P_REAL = 68568.415  # LAMMPS real units: kcal/mol/A^3 -> atm


def _params(name):
    data = gzip.open(os.path.join(G, name)).read()
    f = tempfile.NamedTemporaryFile(suffix=".txt", delete=False)
    f.write(data)
    f.close()
    return f.name


@pytest.fixture(scope="module")
def diamond():
    rec = json.load(open(os.path.join(G, "ev_tally_mb.json")))
    O = LmpOracle(_params(rec["params"]))
    x, typ, tag, lo, hi, masses = read_lammps_data(os.path.join(G, rec["data"]))
    ct = match_types(masses, O.masses())
    S = LmpSystem(x, typ, tag, lo, hi, ct, O.rc2(), rec["neighbor_skin"])
    trips, quads = O.build_lists(S)
    R = O.compute(S, trips, quads)
    return dict(rec=rec, O=O, x=x, typ=typ, ct=ct, lo=lo, hi=hi, S=S, trips=trips, quads=quads, R=R)


def test_lmp_oracle_reference_log(diamond):
    rec, R, lo, hi = diamond["rec"], diamond["R"], diamond["lo"], diamond["hi"]
    V = float(np.prod(hi - lo))
    th = rec["thermo"]
    # no velocities in the data file: the pressure is the virial part alone
    press = (R.virial[0] + R.virial[1] + R.virial[2]) / (3.0 * V) * NKTV2P_REAL
    assert abs(press - th["Press"]) < 1e-5
    # compute stress/atom sums: -(per-atom virial) * nktv2p, LAMMPS order
    S = -R.vatom.sum(0) * NKTV2P_REAL
    for k, name in enumerate(("v_Sxx", "v_Syy", "v_Szz")):
        assert abs(S[k] - th[name]) <= 1e-10 * abs(th[name]), (name, S[k], th[name])


def test_lmp_oracle_matches_serial(diamond):
    O, R, S = diamond["O"], diamond["R"], diamond["S"]
    names = O.type_names()
    cell = np.diag(diamond["hi"] - diamond["lo"])
    res = O.calculate(diamond["x"] - diamond["lo"], [names[c] for c in diamond["ct"][diamond["typ"] - 1]], cell)
    assert abs(res.energy - R.eng) <= 1e-10 * abs(R.eng)
    V = float(np.prod(diamond["hi"] - diamond["lo"]))
    sv = res.stress.reshape(3, 3) * V
    lmp = np.array([sv[0, 0], sv[1, 1], sv[2, 2], sv[0, 1], sv[0, 2], sv[1, 2]])
    assert np.abs(lmp - R.virial).max() <= 1e-10 * np.abs(R.virial).max()
    F = np.zeros_like(res.forces)
    np.add.at(F, S.tag - 1, R.f)  # ghost forces folded onto their atoms
    assert np.abs(F - res.forces).max() <= 1e-10 * np.abs(res.forces).max()
    # the per-atom split adds up to the global tallies
    assert abs(R.eatom.sum() - R.eng) <= 1e-10 * abs(R.eng)
    assert np.abs(R.vatom.sum(0) - R.virial).max() <= 1e-10 * np.abs(R.virial).max()


def _lists_by_definition(S, pad3, pad4, has4):
    """Vectorised 3-/4-mers of every owner in (jj, kk, ll) order."""
    T3, T4 = [], []
    x, tag = S.x, S.tag
    for ii in range(S.inum):
        i = S.ilist[ii]
        r = S.nbr[S.nbr_off[ii]:S.nbr_off[ii + 1]]
        d_i = np.sqrt(((x[r] - x[i]) ** 2).sum(1))
        J = (r != i) & (tag[r] >= tag[i]) & ((d_i < pad3) | (d_i < pad4))
        D = np.sqrt(((x[r][None, :, :] - x[r][:, None, :]) ** 2).sum(2))  # D[a, b] = |x_b - x_a|
        tg = tag[r]
        ne = r[:, None] != r[None, :]
        ok2 = J[:, None] & J[None, :] & ne & (tg[None, :] >= tg[:, None])
        t3 = ok2 & (d_i[:, None] < pad3) & (d_i[None, :] < pad3) & (D < pad3)
        for a, b in zip(*np.nonzero(t3)):
            T3.append((i, r[a], r[b]))
        if not has4:
            continue
        q2 = ok2 & (d_i[:, None] < pad4) & (d_i[None, :] < pad4) & (D < pad4)
        q3 = (q2[:, :, None] & J[None, None, :] & ne[:, None, :] & ne[None, :, :]
              & (tg[None, None, :] >= tg[None, :, None]) & (d_i[None, None, :] < pad4)
              & (D[:, None, :] < pad4) & (D[None, :, :] < pad4))
        for a, b, c in zip(*np.nonzero(q3)):
            T4.append((i, r[a], r[b], r[c]))
    return np.array(T3, np.int32).reshape(-1, 3), np.array(T4, np.int32).reshape(-1, 4)


@pytest.mark.parametrize("L,natoms", [(7.0, 24), (9.5, 60)])
def test_lmp_oracle_small_box_lists(diamond, L, natoms):
    O = diamond["O"]
    rng = np.random.default_rng(0xC0DE10 + natoms)
    pos = []
    while len(pos) < natoms:
        p = rng.uniform(0, L, 3)
        if all(np.linalg.norm((p - q + L / 2) % L - L / 2) >= 1.3 for q in pos):
            pos.append(p)
    x = np.array(pos)
    typ = np.ones(natoms, np.int32)
    tag = np.arange(1, natoms + 1, dtype=np.int32)
    S = LmpSystem(x, typ, tag, np.zeros(3), np.full(3, L), diamond["ct"], O.rc2(), 1.0, shuffle_seed=natoms)
    # small box: some neighbour row holds two images of one atom
    assert any(len(set(S.tag[r])) < len(r) for r in S.firstneigh_rows())
    trips, quads = O.build_lists(S)
    pad3, pad4 = O.max_cutoff(3) + S.skin, O.max_cutoff(4) + S.skin
    T3, T4 = _lists_by_definition(S, pad3, pad4, O.order(4) > 0)
    assert len(T3) > 0 and len(T4) > 0
    assert np.array_equal(trips, T3)
    assert np.array_equal(quads, T4)
    R = O.compute(S, trips, quads)
    assert abs(R.eatom.sum() - R.eng) <= 1e-10 * abs(R.eng)
    assert np.abs(R.vatom.sum(0) - R.virial).max() <= 1e-10 * np.abs(R.virial).max()
    # forces sum to zero over all atoms and images (every term is internal)
    assert np.abs(R.f.sum(0)).max() <= 1e-9 * np.abs(R.f).max()
