"""Writes the LAMMPS pair-style fixtures under tests/golden/lmp/ from the
reference's own LAMMPS test (etc/lmp/tests/test_suite-ev_tally_mb): its data
file and force field, gzipped, and the step-0 thermo line of its expected log
(pressure and the summed per-atom stress Sxx, Syy, Szz, in atm and atm*A^3).

Run in a checkout that has the reference at /root/reference:
    python tests/golden/make_lmp_golden.py
"""
import gzip
import json
import os
import shutil

REF = "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lmp")
SUITE = os.path.join(REF, "etc/lmp/tests/test_suite-ev_tally_mb")
FF = os.path.join(REF, "serial_interface/tests/force_fields/published_params.Carbon-2.0.Small.2+3+4b.Tersoff.txt")
DATA = os.path.join(SUITE, "diamond_3.21gcc_512atoms.data.in")


def gz_copy(src, dst):
    with open(src, "rb") as a, gzip.open(dst, "wb") as b:
        shutil.copyfileobj(a, b)


def main():
    os.makedirs(OUT, exist_ok=True)
    gz_copy(DATA, os.path.join(OUT, os.path.basename(DATA) + ".gz"))
    gz_copy(FF, os.path.join(OUT, os.path.basename(FF) + ".gz"))
    lines = open(os.path.join(SUITE, "expected_output/log.lammps")).read().splitlines()
    head = next(i for i, l in enumerate(lines) if l.split()[:2] == ["Step", "Press"])
    cols = lines[head].split()
    row0 = [float(v) for v in lines[head + 1].split()]
    rec = {
        "source": "etc/lmp/tests/test_suite-ev_tally_mb/expected_output/log.lammps (step 0)",
        "units": "real", "newton": "on", "neighbor_skin": 1.5,
        "data": os.path.basename(DATA) + ".gz", "params": os.path.basename(FF) + ".gz",
        "thermo": dict(zip(cols, row0)),
        "format": "%20.5f",
    }
    json.dump(rec, open(os.path.join(OUT, "ev_tally_mb.json"), "w"), indent=1)
    print(rec)


if __name__ == "__main__":
    main()
