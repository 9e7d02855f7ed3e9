"""The drop-in boundary as callers use it, against the reference's goldens.

  * A C++ program of the serial_chimes_interface class (tests/drivers/
    serial_cpp_main.cpp, the call shape of a user program: construct with
    the replication flag, init_chimesFF, calculate), compiled against
    include/ and linked to libchimescalc.so, on reference fixtures: a
    4-body Tersoff force field with special cutoffs and offsets, and
    replicated small cells.
  * The same calls with caller-seeded energy and forces: calculate() adds
    onto them and divides the energy, seed included, by the replica count
    (serial_chimes_interface.cpp:1105); the stress is overwritten; the
    caller's type list grows by the replicas' types (cpp:187, 257).
  * A C program of the static C ABI path (set_chimes_serial ->
    init_chimes_serial -> calculate_chimes) on the same fixtures.
Bars: 1e-10 relative, as the parity tests.
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import assert_close, build_driver, golden_arrays, parse_driver_output

pytestmark = pytest.mark.gpu

CASES = [
    "published_params.HN3.2+3+4b.Tersoff.special.offsets.txt|HN3.2gcc_3000K.OUTCAR_#000.xyz|0",
    "published_params.liqCO.2+3b.cubic.txt|diam.2_#000.xyz|1",
    "published_params.liqCO.2+3b.cubic.txt|diam.8_#000.xyz|1",
    "published_params.liqCO.2+3b.cubic.txt|CO.2.5gcc_6500K.OUTCAR_#000.scramble.xyz|1",
]


@pytest.fixture(scope="module")
def drivers(tmp_path_factory):
    d = tmp_path_factory.mktemp("drivers")
    return {n: build_driver(n, d) for n in ("serial_cpp_main", "serial_c_main")}


def run(exe, *args):
    env = dict(os.environ)
    r = subprocess.run([exe] + [str(a) for a in args], capture_output=True, text=True, timeout=120, env=env)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return parse_driver_output(r.stdout)


def check(tag, E, F, S, e0, f0, s0):
    assert abs(E - e0) <= 1e-10 * abs(e0), "%s: energy %.17g vs %.17g" % (tag, E, e0)
    assert_close(tag + " forces", F, f0, 1e-10, atol=1e-9)
    assert_close(tag + " stress", S, s0, 1e-10)


@pytest.mark.parametrize("key", CASES)
def test_cpp_program(drivers, ref_dir, ref_cases, key):
    pf, xyz, small = key.split("|")
    rec = ref_cases[key]
    e0, f0, s0 = golden_arrays(rec)
    print("C++ program on", key, flush=True)
    E, nt, S, F = run(drivers["serial_cpp_main"], os.path.join(ref_dir, pf), os.path.join(ref_dir, xyz), small)
    check(key, E, F, S, e0, f0, s0)
    assert nt == rec["meta"][0], "type list has %d entries, the replicated system %d atoms" % (nt, rec["meta"][0])


@pytest.mark.parametrize("key", CASES)
def test_cpp_program_seeded(drivers, ref_dir, ref_cases, key):
    pf, xyz, small = key.split("|")
    rec = ref_cases[key]
    e0, f0, s0 = golden_arrays(rec)
    ncopy = (rec["meta"][2] + 1) ** 3
    eseed, fseed = 123.25, -0.75
    print("C++ program, seeded, on", key, flush=True)
    E, nt, S, F = run(drivers["serial_cpp_main"], os.path.join(ref_dir, pf), os.path.join(ref_dir, xyz), small,
                      eseed, fseed)
    n = len(f0)
    seedF = fseed * (1.0 + np.arange(n)[:, None] + 0.25 * np.arange(3)[None, :])
    check(key + " seeded", E, F, S, e0 + eseed / ncopy, f0 + seedF, s0)


@pytest.mark.parametrize("key", CASES)
def test_c_static_path(drivers, ref_dir, ref_cases, key):
    pf, xyz, small = key.split("|")
    e0, f0, s0 = golden_arrays(ref_cases[key])
    print("C program (static instance) on", key, flush=True)
    E, _, S, F = run(drivers["serial_c_main"], os.path.join(ref_dir, pf), os.path.join(ref_dir, xyz), small)
    check(key + " static C", E, F, S, e0, f0, s0)
