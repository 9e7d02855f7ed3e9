"""Tabulated force fields for the tests.

The reference's own tabulation example (etc/lmp/tests/example-tabulate:
params.txt, the 2-body table chimes_scan_2b.type_0.dat{.energy,.force} and
the 200-atom carbon data.in) is held in tests/golden/tab.  Its parameter file
also names a 3-body table, chimes_scan_3b.type_0.dat, that the reference does
not ship; write_3b_table generates one in the format read_3B_tab reads
(chimesFF.cpp:259-352): a row count, then rows "rij rik rjk E" (.energy) or
"rij rik rjk Fij Fik Fjk" (.force) on a uniform grid, rij slowest and rjk
fastest.  Its values are a smooth three-body energy and the force scalars
(dE/dr_q) / r_q of its pairs, so forces from the table approach -grad E.
"""
import gzip
import os
import shutil

import numpy as np

from conftest import GOLDEN

TAB = os.path.join(GOLDEN, "tab")
PARAMS = "params.txt"
TABLE_2B = "chimes_scan_2b.type_0.dat"
TABLE_3B = "chimes_scan_3b.type_0.dat"


def e3(a, b, c):
    """A smooth three-body energy (kcal/mol) of the pair distances and its
    partial derivatives."""
    g = np.exp(-(a + b + c) / 2.0)
    p = 1.0 + 0.2 * a * b - 0.1 * c * c
    E = 40.0 * g * p
    da = 40.0 * g * (-0.5 * p + 0.2 * b)
    db = 40.0 * g * (-0.5 * p + 0.2 * a)
    dc = 40.0 * g * (-0.5 * p - 0.2 * c)
    return E, da, db, dc


def grid(r0=0.9, h=0.1, n=24):
    return np.array([r0 + i * h for i in range(n)], dtype=float)


def write_3b_table(base, r0=0.9, h=0.1, n=24, values=None):
    """base.energy and base.force on the grid r0 + i h (i < n) in each pair
    distance.  values(a, b, c) -> (E, Fij, Fik, Fjk) (default: e3 and its
    force scalars)."""
    r = grid(r0, h, n)
    A, B, Cc = np.meshgrid(r, r, r, indexing="ij")
    a, b, c = A.reshape(-1), B.reshape(-1), Cc.reshape(-1)
    if values is None:
        E, da, db, dc = e3(a, b, c)
        fij, fik, fjk = da / a, db / b, dc / c
    else:
        E, fij, fik, fjk = values(a, b, c)
    # plain Python floats, so %r writes the shortest round-trip digits
    a, b, c, E, fij, fik, fjk = (np.asarray(v, float).tolist() for v in (a, b, c, E, fij, fik, fjk))
    with open(base + ".energy", "w") as f:
        f.write("%d\n" % len(a))
        for k in range(len(a)):
            f.write("%r %r %r %r\n" % (a[k], b[k], c[k], E[k]))
    with open(base + ".force", "w") as f:
        f.write("%d\n" % len(a))
        for k in range(len(a)):
            f.write("%r %r %r %r %r %r\n" % (a[k], b[k], c[k], fij[k], fik[k], fjk[k]))


def tab_dir(dest, tab3=True, tab2=True, table3=None):
    """The reference's tabulation example in dest: its parameter file (the
    3-body TABULATED entry dropped when tab3 is false, the 2-body one when tab2
    is false), its 2-body tables, a generated 3-body table (table3: keyword
    arguments of write_3b_table) and data.in.  Returns the parameter file."""
    os.makedirs(dest, exist_ok=True)
    text = open(os.path.join(TAB, PARAMS)).read()
    if not tab3:
        text = text.replace(" TABULATED " + TABLE_3B, "")
    if not tab2:
        text = text.replace(" TABULATED " + TABLE_2B, "")
    with open(os.path.join(dest, PARAMS), "w") as f:
        f.write(text)
    for suffix in (".energy", ".force"):
        with gzip.open(os.path.join(TAB, TABLE_2B + suffix + ".gz")) as a, \
                open(os.path.join(dest, TABLE_2B + suffix), "wb") as b:
            shutil.copyfileobj(a, b)
    write_3b_table(os.path.join(dest, TABLE_3B), **(table3 or {}))
    with gzip.open(os.path.join(TAB, "data.in.gz")) as a, open(os.path.join(dest, "data.in"), "wb") as b:
        shutil.copyfileobj(a, b)
    return os.path.join(dest, PARAMS)
