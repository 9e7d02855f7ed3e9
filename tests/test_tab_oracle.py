"""Pins the tabulation restatement (oracle/chimes_tab_oracle.c) on the CPU.

  * the reference's chimesFF compiled with -DTABULATION (oracle/_ref), call by
    call and bit for bit: get_tab_2B and compute_2B_tab on the reference's own
    2-body table (etc/lmp/tests/example-tabulate/chimes_scan_2b.type_0.dat),
    interpolateTricubic and compute_3B_tab on a generated 3-body table;
  * what the mathematics fixes, without the reference: a linear 2-body table
    is reproduced by the spline, including the piece the reference evaluates
    left of its interval; the Catmull-Rom tricubic reproduces polynomials of
    degree two in each distance inside the grid and returns the table's value
    at a grid point, so a transposed axis or stencil index fails;
  * the parameter-file rules: table names in reading order, all 2-body pairs
    or none, rows of the wrong width and missing files are errors.
"""
import os

import numpy as np
import pytest

from oracle_bindings import Oracle, ReferenceTab, TabOracle, ref_tab_available
from tab_setup import TABLE_2B, TABLE_3B, grid, tab_dir, write_3b_table

need_ref = pytest.mark.skipif(not ref_tab_available(),
                              reason="oracle/_ref/libchimes_ref_tab.so is not built (reference sources absent)")


@pytest.fixture(scope="module")
def tabdir(tmp_path_factory):
    return tab_dir(str(tmp_path_factory.mktemp("tab")))


def _load(params):
    O = Oracle(params)
    return O, TabOracle(O, params)


def _knots(params):
    rows = open(os.path.join(os.path.dirname(params), TABLE_2B + ".energy")).read().split("\n")[1:]
    return np.array([float(r.split()[0]) for r in rows if r.strip()])


def _same(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return a.shape == b.shape and np.array_equal(a.view(np.int64), b.view(np.int64))


@need_ref
def test_2b_lookup_matches_reference(tabdir):
    O, T = _load(tabdir)
    R = ReferenceTab(tabdir)
    assert R.flags() == 3 and T.bodies() == 3
    x = _knots(tabdir)
    rng = np.random.default_rng(0xC0DE40)
    rs = np.concatenate([rng.uniform(x[1], x[-1] + 0.05, 3000), x[2:], 0.5 * (x[1:-1] + x[2:]),
                         np.nextafter(x[2:-1], 0.0), np.nextafter(x[2:-1], 9.0)])
    rs = rs[rs > x[1]]
    for r in rs:
        for e in (1, 0):
            v, err = T.value_2b(0, r, e)
            assert err == 0
            assert _same(v, R.value_2b(0, r, e)), (r, e)


@need_ref
def test_2b_compute_matches_reference(tabdir):
    O, T = _load(tabdir)
    R = ReferenceTab(tabdir)
    x = _knots(tabdir)
    rng = np.random.default_rng(0xC0DE41)
    for k in range(2000):
        r = rng.uniform(x[1] + 1e-9, 3.3)
        u = rng.normal(size=3)
        dr = u / np.linalg.norm(u) * r
        dx = float(np.sqrt(dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2]))
        if dx <= x[1]:
            continue
        f, s, e, rc = T.compute_2b(dx, dr, [0, 0])
        f0, s0, e0 = R.compute_2b(dx, dr, [0, 0])
        assert rc == 0
        assert _same(f, f0) and _same(s, s0) and _same(e, e0), (k, dx)


@need_ref
def test_3b_interpolation_matches_reference(tabdir):
    O, T = _load(tabdir)
    R = ReferenceTab(tabdir)
    rng = np.random.default_rng(0xC0DE42)
    g = grid()
    pts = list(rng.uniform(0.85, 3.14, (3000, 3)))
    pts += [(g[i], g[j], g[k]) for i, j, k in rng.integers(0, 22, (200, 3))]
    for rij, rik, rjk in pts:
        v, err = T.interp_3b(0, rij, rik, rjk)
        assert err == 0
        assert _same(v, R.interp_3b(0, rij, rik, rjk)), (rij, rik, rjk)


@need_ref
def test_3b_compute_matches_reference(tabdir):
    O, T = _load(tabdir)
    R = ReferenceTab(tabdir)
    rng = np.random.default_rng(0xC0DE43)
    n = 0
    while n < 2000:
        p = rng.uniform(0.0, 2.8, (3, 3))
        dr = np.concatenate([p[1] - p[0], p[2] - p[0], p[2] - p[1]])
        dx = np.array([np.sqrt(dr[3 * q] ** 2 + dr[3 * q + 1] ** 2 + dr[3 * q + 2] ** 2) for q in range(3)])
        if dx.min() < 0.95:
            continue
        f, s, e, rc = T.compute_3b(dx, dr, [0, 0, 0])
        f0, s0, e0 = R.compute_3b(dx, dr, [0, 0, 0])
        assert rc == 0
        assert _same(f, f0) and _same(s, s0) and _same(e, e0), dx
        n += 1


def test_2b_spline_reproduces_a_linear_table(tmp_path):
    params = tab_dir(str(tmp_path))
    x = np.array([1.0 + 0.013 * i + 0.0004 * i * i for i in range(120)])  # non-uniform knots
    for suffix, y in ((".energy", 3.0 - 2.0 * x), (".force", 0.5 + 0.25 * x)):
        with open(os.path.join(str(tmp_path), TABLE_2B + suffix), "w") as f:
            f.write("%d\n" % len(x))
            for a, b in zip(x, y):
                f.write("%r %r\n" % (float(a), float(b)))
    O, T = _load(params)
    rng = np.random.default_rng(0xC0DE44)
    # lower_bound's interval: a distance inside the last interval gives nothing
    for r in np.concatenate([rng.uniform(x[1], x[-2], 500), x[2:-1]]):
        if r <= x[1]:
            continue
        e, err = T.value_2b(0, r, 1)
        f, _ = T.value_2b(0, r, 0)
        assert err == 0
        assert abs(e - (3.0 - 2.0 * r)) <= 1e-12 and abs(f - (0.5 + 0.25 * r)) <= 1e-12, r
    # past the last knot the table gives nothing; at or below the second knot
    # the reference exits, the restatement reports an error
    assert T.value_2b(0, x[-1] + 0.01, 1) == (0.0, 0)
    assert T.value_2b(0, 0.5 * (x[-2] + x[-1]), 1) == (0.0, 0)
    assert T.value_2b(0, x[1], 1)[1] == 1 and T.value_2b(0, x[0], 0)[1] == 1


def test_tricubic_reproduces_quadratics(tmp_path):
    def quad(a, b, c):
        E = 1.0 + 0.5 * a - 0.3 * b + 0.2 * c + 0.1 * a * b - 0.05 * c * c + 0.02 * a * a * b * c
        return E, 2.0 - a * c, 0.7 * b * b - c, a + b + c * c

    params = tab_dir(str(tmp_path), table3={"values": quad})
    O, T = _load(params)
    rng = np.random.default_rng(0xC0DE45)
    g = grid()
    # every stencil point inside the grid: index 1 .. n - 3 on each axis
    for rij, rik, rjk in rng.uniform(g[1] + 1e-9, g[-3] - 1e-9, (500, 3)):
        v, err = T.interp_3b(0, rij, rik, rjk)
        ref = np.array(quad(rij, rik, rjk))
        assert err == 0
        assert np.abs(v - ref).max() <= 1e-10 * np.abs(ref).max(), (rij, rik, rjk)


def test_tricubic_at_grid_points(tabdir):
    O, T = _load(tabdir)
    base = os.path.join(os.path.dirname(tabdir), TABLE_3B)
    rows = open(base + ".energy").read().split("\n")[1:]
    E = np.array([float(r.split()[3]) for r in rows if r.strip()]).reshape(24, 24, 24)
    g = grid()
    rng = np.random.default_rng(0xC0DE46)
    for i, j, k in rng.integers(1, 23, (300, 3)):
        v, err = T.interp_3b(0, g[i], g[j], g[k])
        assert err == 0
        assert abs(v[0] - E[i, j, k]) <= 1e-12 * np.abs(E).max(), (i, j, k)


def test_3b_grid_index_past_the_table_is_an_error(tabdir):
    O, T = _load(tabdir)
    g = grid()
    assert T.interp_3b(0, g[-1] + 0.2, 1.5, 1.5)[1] == 1


def test_table_entries_and_errors(tmp_path):
    params = tab_dir(str(tmp_path / "both"))
    O, T = _load(params)
    assert T.bodies() == 3
    only2 = tab_dir(str(tmp_path / "two"), tab3=False)
    assert _load(only2)[1].bodies() == 1
    only3 = tab_dir(str(tmp_path / "three"), tab2=False)
    assert _load(only3)[1].bodies() == 2
    # a force table with energy-table rows, and a missing table
    d = str(tmp_path / "bad")
    params = tab_dir(d)
    write_3b_table(os.path.join(d, TABLE_3B), n=5)
    with open(os.path.join(d, TABLE_3B + ".force"), "w") as f:
        f.write("1\n1.0 1.0 1.0 2.0\n")
    with pytest.raises(ValueError, match="items"):
        TabOracle(Oracle(params), params)
    os.remove(os.path.join(d, TABLE_3B + ".force"))
    with pytest.raises(ValueError, match="could not open"):
        TabOracle(Oracle(params), params)


def test_2b_table_is_a_scan_of_the_polynomial(tmp_path):
    """The reference's 2-body table holds its example model's own 2-body
    polynomial: on the example's 200-atom carbon box (pair style loops of the
    oracle, 3-body polynomial) the tabulated 2-body path gives the polynomial
    energy to 1e-8 relative and forces to 1e-5 of the largest component."""
    from lmp_setup import LmpSystem, match_types, read_lammps_data
    from oracle_bindings import LmpOracle
    params = tab_dir(str(tmp_path), tab3=False)
    O = LmpOracle(params)
    T = TabOracle(O, params)
    assert T.bodies() == 1
    x, typ, tag, lo, hi, masses = read_lammps_data(os.path.join(str(tmp_path), "data.in"))
    S = LmpSystem(x, typ, tag, lo, hi, match_types(masses, O.masses()), O.rc2(), 2.0)
    trips, quads = O.build_lists(S)
    tab = O.compute(S, trips, quads, tables=T)
    assert O.tab_errors == 0
    poly = O.compute(S, trips, quads)
    assert abs(tab.eng - poly.eng) <= 1e-8 * abs(poly.eng)
    assert np.abs(tab.f - poly.f).max() <= 1e-5 * np.abs(poly.f).max()
    assert np.abs(tab.virial - poly.virial).max() <= 1e-6 * np.abs(poly.virial).max()
