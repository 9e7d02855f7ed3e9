"""CPU checks of the product library: it loads, exports every symbol the
headers under include/ declare, and refuses to run without a GPU (no CPU
fallback)."""
import ctypes as C
import os
import re
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions(path):
    txt = open(path).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w \*]*?\b([a-z_][a-z0-9_]*)\s*\(", txt, flags=re.M)
    return sorted(set(n for n in names if n not in ("if", "while", "for")))


@pytest.fixture(scope="module")
def lib():
    import chimes_calculator_amd as cc
    return cc.lib()


@pytest.mark.parametrize("header", ["chimescalc_serial_C.h", "chimescalc_C.h", "chimes_gpu.h"])
def test_exports(lib, header):
    import chimes_calculator_amd as cc
    names = header_functions(os.path.join(ROOT, "include", header))
    assert names, header
    assert set(names) == set(cc.EXPORTS[header]), set(names) ^ set(cc.EXPORTS[header])
    for n in names:
        assert hasattr(lib, n), "%s does not export %s" % (header, n)


def test_dl_and_shared_names():
    for name in ("libchimescalc.so", "libchimescalc_dl.so"):
        assert os.path.exists(os.path.join(ROOT, "chimes_calculator_amd", "lib", name))


def test_no_cpu_fallback():
    """Without a visible GPU the model cannot be created and says why."""
    code = ("import sys; sys.path.insert(0, %r); import chimes_calculator_amd as cc\n"
            "try:\n    cc.GpuModel(%r)\nexcept cc.ChimesError as e:\n    print('ERR', e)\n" %
            (ROOT, os.path.join(ROOT, "tests", "golden", "make_golden.py")))
    env = dict(os.environ, HIP_VISIBLE_DEVICES="", ROCR_VISIBLE_DEVICES="")
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=env, timeout=120)
    assert "ERR" in out.stdout


@pytest.mark.parametrize("name", ["serial_cpp_main", "serial_c_main"])
def test_caller_programs_build(tmp_path, name):
    """A C++ program of the serial_chimes_interface class and a C program of
    the system C ABI compile against include/ and link libchimescalc.so
    alone (they run on the GPU in test_gpu_boundary.py)."""
    from conftest import build_driver
    exe = build_driver(name, tmp_path)
    assert os.access(exe, os.X_OK)
