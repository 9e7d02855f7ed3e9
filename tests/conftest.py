import gzip
import json
import os
import shutil
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TESTS = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(TESTS, "golden")
sys.path.insert(0, ROOT)
sys.path.insert(0, TESTS)
# library errors exit the process (as the reference does); make that visible
os.environ.setdefault("CHIMES_ERROR_EXIT", "86")
os.environ.setdefault("CHIMES_QUIET", "1")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU and the built HIP library")
    config.addinivalue_line("markers", "slow: CPU test taking more than a few seconds")


@pytest.fixture(scope="session")
def ref_dir(tmp_path_factory):
    """Decompressed copies of the reference test inputs held as fixtures."""
    d = tmp_path_factory.mktemp("ref_inputs")
    src = os.path.join(GOLDEN, "ref_inputs")
    for f in os.listdir(src):
        if f.endswith(".gz"):
            with gzip.open(os.path.join(src, f), "rb") as g, open(os.path.join(d, f[:-3]), "wb") as o:
                shutil.copyfileobj(g, o)
    return str(d)


@pytest.fixture(scope="session")
def model_dir(tmp_path_factory):
    """Parameter files of the seeded synthetic workloads."""
    from chimes_calculator_amd import workloads as W
    d = tmp_path_factory.mktemp("models")
    for k, m in W.MODELS.items():
        m.write(os.path.join(d, k + ".txt"))
    return str(d)


@pytest.fixture(scope="session")
def dscan_dir(tmp_path_factory):
    """Parameter files of the dense-dimension scan models (workloads.DSCAN)."""
    from chimes_calculator_amd import workloads as W
    d = tmp_path_factory.mktemp("dscan")
    for d3, d4 in W.DSCAN:
        W.dscan_model(d3, d4).write(os.path.join(d, "dscan_%d_%d.txt" % (d3, d4)))
    return str(d)


@pytest.fixture(scope="session")
def dscan_cases():
    with open(os.path.join(GOLDEN, "dscan_cases.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def c3_case():
    with gzip.open(os.path.join(GOLDEN, "c3_case.json.gz"), "rt") as f:
        return json.load(f)["C3|0"]


@pytest.fixture(scope="session")
def ref_cases():
    with open(os.path.join(GOLDEN, "ref_cases.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def synth_cases():
    with open(os.path.join(GOLDEN, "synth_cases.json")) as f:
        return json.load(f)


def golden_arrays(rec):
    return (float(rec["energy"]), np.array([[float(v) for v in r] for r in rec["forces"]]),
            np.array([float(v) for v in rec["stress"]]))


EV = 0.0433641       # eV per kcal/mol
MIXED_E_TOL = 1e-6   # eV/atom
MIXED_F_TOL = 1e-4   # eV/A
MIXED_S_TOL = 1e-5   # of the largest stress component


def mixed_errors(E, F, S, rec, natoms):
    """Errors of one configuration's result against its FP64 golden record:
    energy in eV/atom, largest force component error in eV/A, stress relative
    to the largest golden component."""
    e0, f0, s0 = golden_arrays(rec)
    de = abs(E - e0) * EV / natoms
    df = float(np.abs(np.asarray(F) - f0).max()) * EV
    ds = float(np.abs(np.asarray(S) - s0).max()) / max(np.abs(s0).max(), 1e-300)
    return de, df, ds


def assert_mixed(tag, E, F, S, rec, natoms):
    """The FP32-mixed bars against the FP64 golden: energy within 1e-6 eV/atom,
    every force component within 1e-4 eV/A, stress within 1e-5 of its
    largest component."""
    de, df, ds = mixed_errors(E, F, S, rec, natoms)
    assert de <= MIXED_E_TOL, "%s: energy error %.3g eV/atom" % (tag, de)
    assert df <= MIXED_F_TOL, "%s: force error %.3g eV/A" % (tag, df)
    assert ds <= MIXED_S_TOL, "%s: stress error %.3g (relative)" % (tag, ds)
    return de, df, ds


def assert_close(name, got, ref, rtol, scale=None, atol=0.0):
    """max |got - ref| <= rtol * scale + atol, scale = max(|ref|) of the
    quantity (a component-wise relative test would be ill-posed for
    components that cancel to ~0).  atol covers quantities that vanish by
    symmetry (forces in a perfect crystal are summation noise, ~1e-11)."""
    got, ref = np.asarray(got, dtype=np.float64), np.asarray(ref, dtype=np.float64)
    s = scale if scale is not None else max(np.abs(ref).max(), 1e-300)
    err = np.abs(got - ref).max()
    assert err <= rtol * s + atol, "%s: max error %.3e (scale %.3e) exceeds %.1e relative + %.1e" % (
        name, err, s, rtol, atol)
    return err / s


DRIVERS = os.path.join(TESTS, "drivers")
LIBDIR = os.path.join(ROOT, "chimes_calculator_amd", "lib")


def build_driver(name, outdir):
    """Compiles tests/drivers/<name>.{c,cpp} against include/ and links it to
    the built libchimescalc.so (rpath set); returns the executable's path."""
    import subprocess
    src = os.path.join(DRIVERS, name + (".cpp" if os.path.exists(os.path.join(DRIVERS, name + ".cpp")) else ".c"))
    exe = os.path.join(str(outdir), name)
    cc = ["g++", "-std=c++11"] if src.endswith(".cpp") else ["gcc", "-std=c99"]
    cmd = cc + ["-O1", "-I", os.path.join(ROOT, "include"), "-o", exe, src, "-L", LIBDIR, "-lchimescalc",
                "-Wl,-rpath," + LIBDIR]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, "compile %s failed:\n%s" % (name, r.stderr)
    return exe


def parse_driver_output(text):
    """(energy, ntypes or None, stress[9], forces[n, 3]) from a driver's output."""
    e, nt, st, F = None, None, None, []
    for line in text.splitlines():
        f = line.split()
        if not f:
            continue
        if f[0] == "energy":
            e = float(f[1])
        elif f[0] == "ntypes":
            nt = int(f[1])
        elif f[0] == "stress":
            st = np.array([float(v) for v in f[1:10]])
        elif f[0] == "force":
            F.append([float(v) for v in f[1:4]])
    return e, nt, st, np.array(F)
