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
