"""Benchmark: atom-evaluations per second of the ChIMES evaluation path.

Default workload (--workload C2): 4096 independent C/O configurations, 8-512
atoms each (489,549 atoms), densities 1.5-3.5 g/cc, 2B/3B/4B orders 12/7/3,
rc 4.5/4.0/3.5 A.  One step = one full evaluation of the whole batch (ghost
images, cluster lists, 1-4-body energies, forces and stress of every
configuration) by replaying the batch plan's captured graph, with positions,
types and outputs resident in GPU memory.  The headline value is the
FP32-mixed plan; the same line carries the FP64 plan, the plain FP64 drop-in
path (one calculate() per configuration), the error of the measured result
against the reference's FP64 goldens, a roofline entry for the batch's
dominant kernel (term-model FLOPs over its HIP-event-timed duration), the CPU
baseline (the oracle's C restatement, single-threaded and as one process per
host core over the same configurations), and nested lines for the 1536-atom
O/H box (C1) and the served request stream (C4, p50/p99 latency).

Other forms (one JSON line each):
  --workload C1  1536 O/H atoms (512 H2O-like, L = 24.8 A), 12/8/4, rc
                 5.0/4.5/4.0 A, FP64: direct evaluations and an autotuned plan.
  --workload C3  32768-atom carbon, orders 16/10/4: FP64 and FP32-mixed, with
                 the error of each form against the FP64 golden.
  --workload C4  Poisson request stream served with continuous batching:
                 p50/p99 latency and throughput, FP64 and FP32-mixed.
  --workload K   batched explicit 3-/4-body clusters (per-cluster API).
  --workload F   cluster fingerprint histogram (calc_histograms.cpp): all
                 117k triplets of a 128-atom C/O configuration against each
                 other (6.8e9 cluster pairs), edges in host memory.
  --workload L   the LAMMPS pair style (chimes_lmp.h): 4096 local carbon atoms
                 (the reference's 512-atom diamond test box repeated 2x2x2),
                 Carbon-2.0 2+3+4-body, skin 1.5 A; steps reuse the many-body
                 lists as LAMMPS does between rebuilds (host arrays in and out).
"""
import argparse
import atexit
import ctypes as C
import gzip
import json
import multiprocessing
import os
import shutil
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64 = 78.6e12   # MI355X dense FP64 (vector and matrix), FLOP/s
PEAK_FP32 = 157.3e12  # MI355X FP32 (vector and matrix), FLOP/s
PEAK_HBM = 8.0e12     # MI355X HBM3E, bytes/s
EV = 0.0433641        # eV per kcal/mol
METRIC = "atom-evals/s (E+F+stress) at stated error vs FP64; p50/p99 request latency"
_TMP = []


def model_file(name):
    """Parameter file of a synthetic workload model, written once into a
    temporary directory removed at exit."""
    from chimes_calculator_amd import workloads as W
    if not _TMP:
        _TMP.append(tempfile.mkdtemp(prefix="chimes_bench_"))
        atexit.register(shutil.rmtree, _TMP[0], True)
    path = os.path.join(_TMP[0], name + ".txt")
    if not os.path.exists(path):
        W.MODELS[name].write(path)
    return path


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def term_flops(nterms, order, npairs):
    """Algorithmic FLOPs of one cluster (term model): one FMA per listed
    coefficient for the energy and each pair gradient, plus per-pair setup."""
    return 2 * nterms * (npairs + 1) + npairs * (6 * order + 20)


def param_terms(path):
    """TOTAL term counts of the 3-body and 4-body blocks of a parameter file."""
    t3, t4 = [], []
    cur = None
    for line in open(path):
        if "TRIPLETTYPE PARAMS" in line:
            cur = t3
        elif "QUADRUPLETYPE PARAMS" in line:
            cur = t4
        elif line.startswith("PAIRS:") and cur is not None:
            cur.append(int(line.split()[-1]))
            cur = None
    return t3, t4


class Hip:
    """Minimal device-memory plumbing (hipMalloc / hipMemcpy) for resident inputs."""

    def __init__(self):
        self.L = C.CDLL("libamdhip64.so")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipDeviceSynchronize.argtypes = []

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), max(a.nbytes, 8)) == 0
        assert self.L.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0
        return p

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), max(nbytes, 8)) == 0
        return p

    def download(self, p, shape, dtype):
        out = np.empty(shape, dtype=dtype)
        assert self.L.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0
        return out

    def sync(self):
        assert self.L.hipDeviceSynchronize() == 0


def profile_figures(summary, kernel):
    """key=value figures of `kernel` from a committed profile summary (lines
    'traffic <kernel> fetch_bytes=.. write_bytes=..' and 'counters <kernel>
    k=v ...'), or {}."""
    path = os.path.join(ROOT, "profiles", summary)
    out = {}
    if not os.path.exists(path):
        return out
    for line in open(path):
        f = line.split()
        if len(f) > 2 and f[0] in ("traffic", "counters") and f[1] == kernel:
            out.update({k: float(v) for k, v in (x.split("=") for x in f[2:])})
    return out


def measured_traffic(kernel, summary="k_eval4m_water1536_fp64_traffic.md"):
    """HBM bytes per launch of `kernel` from the committed counter summary
    (FETCH_SIZE and WRITE_SIZE runs, copy-calibrated), or None."""
    f = profile_figures(summary, kernel)
    if "fetch_bytes" in f and "write_bytes" in f:
        return f["fetch_bytes"] + f["write_bytes"]
    return None


def cpu_baseline(param_file, cfg, stride=3):
    """The oracle (C restatement, single thread) on 1/stride of the clusters
    (owners i with i % stride == 0), scaled to a full evaluation."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_bindings import Oracle
    o = Oracle(param_file)
    t = time.perf_counter()
    o.calculate(*cfg, owner_stride=stride)
    dt = time.perf_counter() - t
    n = len(cfg[0])
    return {"value": n / (dt * stride), "unit": "atom-evals/s", "cores": 1, "kind": "port",
            "sample": "oracle calculate() on the same %d-atom system, clusters of every %d-th owner atom "
                      "(1/%d of the work, %.1f s; lists built in full), scaled by %d" % (n, stride, stride, dt,
                                                                                       stride)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", choices=["C1", "C2", "C3", "C4", "K", "L", "F"], default="C2",
                    help="K: batched explicit 3-/4-body clusters (chimes_gpu_compute_clusters); "
                         "L: the LAMMPS pair style (chimes_lmp.h); F: fingerprint histogram")
    ap.add_argument("--precision", choices=["fp64", "mixed"], default=None,
                    help="C2 headline precision (default mixed); C1 precision (default fp64)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-nested", action="store_true", help="C2: skip the nested C1 and C4 lines")
    ap.add_argument("--direct-only", action="store_true",
                    help="C1: time direct evaluations only (no plan), e.g. for a kernel trace of plain steps")
    ap.add_argument("--plan-only", action="store_true",
                    help="C2: only the batch plan of --precision (no plain per-configuration path), for profiling")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="run one device-to-device copy of this many MiB before timing (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: every configuration of this workload fits one MI355X and the "
              "multi-GPU decomposition is not part of this build (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    if args.precision is None:
        args.precision = "mixed" if args.workload == "C2" else "fp64"
    if args.workload == "C2":
        return bench_c2(args)
    if args.workload == "C3":
        return bench_c3(args)
    if args.workload == "C4":
        return bench_c4(args)
    if args.workload == "K":
        return bench_clusters(args)
    if args.workload == "L":
        return bench_lmp(args)
    if args.workload == "F":
        return bench_fp(args)
    bench_c1(args)


def bench_c1(args):
    print(json.dumps(c1_form(args, cpu=not args.no_cpu)))


def c1_form(args, cpu=False, precision=None):
    """The 1536-atom O/H box, FP64: direct evaluations (HIP-event timed
    kernels) and an autotuned device-buffer plan (the value)."""
    import chimes_calculator_amd as cc
    from chimes_calculator_amd import workloads as W

    mname, cfgs = W.workload("C1")
    pf = model_file(mname)
    model = cc.GpuModel(pf)
    pos, types, cell = cfgs[0]
    n = len(pos)
    hip = Hip()
    d_pos = hip.upload(np.ascontiguousarray(pos, dtype=np.float64).reshape(-1))
    d_typ = hip.upload(model.types_to_index(types))
    d_E, d_F, d_V = hip.alloc(8), hip.alloc(8 * 3 * n), hip.alloc(8 * 6)
    natoms = np.array([n], np.int32)
    model_of = np.zeros(1, np.int32)
    cells = np.ascontiguousarray(cell, dtype=np.float64).reshape(9)
    prec = cc.MIXED if (precision or args.precision) == "mixed" else cc.FP64
    arr = (C.c_void_p * 1)(model.h)
    L = cc.lib()

    def step():
        rc = L.chimes_gpu_eval_batch_device(arr, 1, 1, model_of, natoms, d_pos, d_typ, cells, 0, prec, d_E, d_F,
                                            d_V)
        if rc:
            raise cc.ChimesError(cc._err())

    # (1) direct evaluations: each call runs the pipeline with HIP events
    # around the kernel groups (the 4-body kernel's time for the roofline)
    for _ in range(args.warmup):
        step()
    hip.sync()
    if args.calib_copy_mib > 0:  # a plain copy of known size, for HBM counter calibration
        nb = args.calib_copy_mib << 20
        a, b = hip.alloc(nb), hip.alloc(nb)
        assert hip.L.hipMemcpy(b, a, nb, 3) == 0
        hip.sync()
    t4, tall = [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
        tm = model.timing()
        t4.append(tm[1])
        tall.append(tm[0])
    hip.sync()
    dt_direct = time.perf_counter() - t0
    counts = model.counts()
    # (2) the same evaluation as a batch plan: autotuned kernels, one captured
    # graph replayed per step (the repeated-evaluation form: an MD loop)
    if args.direct_only:
        dt, tp, vv = dt_direct, [1e-3 * x for x in tall], (2, 0)
    else:
        L.chimes_gpu_plan_create_device.restype = C.c_void_p
        plan = L.chimes_gpu_plan_create_device(arr, 1, 1, model_of, natoms, d_pos, d_typ, cells, 0, prec, 1, d_E, d_F,
                                               d_V)
        if not plan:
            raise cc.ChimesError(cc._err())
        L.chimes_gpu_plan_run_device.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        cells_c = cells.ctypes.data_as(C.c_void_p)

        def pstep():
            if L.chimes_gpu_plan_run_device(plan, cells_c, d_E, d_F, d_V):
                raise cc.ChimesError(cc._err())

        for _ in range(args.warmup):
            pstep()
        hip.sync()
        tp = []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            t1 = time.perf_counter()
            pstep()
            tp.append(time.perf_counter() - t1)
        hip.sync()
        dt = time.perf_counter() - t0
        vv, tune_ms, pcounts = (C.c_int * 2)(), (C.c_double * (2 * len(cc.VARIANTS)))(), (C.c_longlong * 3)()
        L.chimes_gpu_plan_info.argtypes = [C.c_void_p] * 4
        L.chimes_gpu_plan_info(plan, vv, tune_ms, pcounts)
        L.chimes_gpu_plan_free.argtypes = [C.c_void_p]
        L.chimes_gpu_plan_free(plan)
    variants = cc.VARIANTS

    # error of the measured result against the reference's own FP64 values
    E = hip.download(d_E, (1,), np.float64)[0]
    F = hip.download(d_F, (n, 3), np.float64)
    gold = json.load(open(os.path.join(ROOT, "tests", "golden", "synth_cases.json")))["C1|0"]
    e0 = float(gold["energy"])
    f0 = np.array([[float(v) for v in r] for r in gold["forces"]])
    ev_per_kcal = 0.0433641
    err = {"energy_eV_per_atom": abs(E - e0) * ev_per_kcal / n,
           "force_max_eV_per_A": float(np.abs(F - f0).max() * ev_per_kcal),
           "energy_rel": abs(E - e0) / abs(e0)}

    t3terms, t4terms = param_terms(pf)
    order4 = model.order(4)
    flops4 = counts[2] * term_flops(t4terms[0], order4, 6)
    t4_s = np.mean(t4) * 1e-3
    achieved = flops4 / t4_s
    peak = PEAK_FP64 if prec == cc.FP64 else PEAK_FP32
    # counters of the chosen 4-body kernel from the committed summaries
    v4 = vv[1]
    k4 = ("k_eval4m<%d,%s,%d,0>" % (order4, "double" if prec == cc.FP64 else "float", 4 if v4 == 0 else 2)
          if v4 in (0, 1) and order4 <= 4 else "k_evalv<4,%d,%s>" % (order4, "double" if prec == cc.FP64 else "float"))
    prof4 = "k_eval4m_water1536_%s.md" % ("fp64_traffic" if prec == cc.FP64 else "mixed_counters")
    figs4 = profile_figures(prof4, k4)
    out = {
        "metric": METRIC,
        "value": n * args.steps / dt,
        "unit": "atom-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == cc.FP64 else "f32-mixed",
        "data": "synthetic: seeded O/H water-like box and seeded random-coefficient ChIMES model",
        "config": {"workload": "O/H 1536 atoms (512 H2O-like, L=24.8 A), 2B+3B+4B orders 12/8/4, "
                               "rc 5.0/4.5/4.0 A, single system, FP64",
                   "natoms": n, "pairs": counts[0], "triplets": counts[1], "quadruplets": counts[2]},
        "latency_ms": {"p50": 1e3 * float(np.percentile(tp, 50)), "p99": 1e3 * float(np.percentile(tp, 99))},
        "forms": {"plan" if not args.direct_only else "direct (timed)": {
                      "value": n * args.steps / dt, "ms_per_step": dt / args.steps * 1e3,
                      "variant3": variants[vv[0]], "variant4": variants[vv[1]]},
                  "direct": {"value": n * args.steps / dt_direct, "ms_per_step": dt_direct / args.steps * 1e3,
                             "pipeline_ms_p50": float(np.percentile(tall, 50)), "variant3": "vector",
                             "variant4": "mfma64"}},
        "error_vs_reference_fp64": err,
        "roofline": {"bound": "mfma", "achieved": achieved / 1e12, "peak": peak / 1e12, "unit": "TFLOP/s",
                     "frac": achieved / peak,
                     "traffic": measured_traffic(k4) if prec == cc.FP64 else None,
                     "traffic_source": "profiles/k_eval4m_water1536_fp64_traffic.md (bytes per launch)",
                     "kernel": "%s (4-body evaluation, %d cluster types)" % (k4, len(t4terms)),
                     "flops_per_step": flops4, "kernel_ms_per_step": t4_s * 1e3,
                     "counters_source": "profiles/" + prof4 if figs4 else None, "counters": figs4 or None},
    }
    if v4 in (0, 1) and order4 <= 4:
        # executed matrix-core work of the 4-body kernel: per 16-cluster tile
        # two products of RB row blocks x D^2 k-steps of 16 x 16 x 4 MFMAs
        # (2048 FLOP each), RB = the 4 D^2 rows padded to 16, over 16
        rb = -(-4 * order4 * order4 // 16)
        ex = counts[2] * 2 * rb * order4 * order4 * 2048 / 16
        out["roofline"]["executed_mfma_tflops"] = ex / t4_s / 1e12
        out["roofline"]["executed_mfma_frac"] = ex / t4_s / peak
        if figs4:
            out["roofline"]["mfma_util_pct"] = figs4.get("MfmaUtil_pct")
            out["roofline"]["valu_per_mfma"] = figs4["valu_per_wave"] / figs4["mfma_per_wave"] \
                if figs4.get("mfma_per_wave") else None
    if cpu:
        out["cpu_baseline"] = cpu_baseline(pf, cfgs[0])
    return out


def atoms_of(cfgs):
    return sum(len(c[0]) for c in cfgs)


_C2 = {}


def _c2_cpu_worker(job):
    """The oracle on configurations idx of the batch until budget_s has
    passed; returns (atoms evaluated, seconds)."""
    pf, idx, budget_s = job
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_bindings import Oracle
    cfgs = _C2["cfgs"]
    o = Oracle(pf)
    atoms, t0 = 0, time.perf_counter()
    for k in idx:
        o.calculate(*cfgs[k])
        atoms += len(cfgs[k][0])
        if time.perf_counter() - t0 >= budget_s:
            break
    return atoms, time.perf_counter() - t0


def c2_cpu_baseline(pf, cfgs, budget_s=10.0):
    """The oracle (C restatement of the reference's serial path, -O2, one
    thread per process) over the batch's configurations in order: once in
    one process, and as one process per host core (at most 16) taking every
    P-th configuration, each for about budget_s seconds.  Runs before the GPU
    is touched (the worker processes are forked)."""
    _C2["cfgs"] = cfgs
    n = len(cfgs)
    a1, t1 = _c2_cpu_worker((pf, range(n), budget_s))
    P = max(1, min(16, os.cpu_count() or 1))
    ctx = multiprocessing.get_context("fork")
    t0 = time.perf_counter()
    with ctx.Pool(P) as pool:
        res = pool.map(_c2_cpu_worker, [(pf, range(w, n, P), budget_s) for w in range(P)])
    wall = time.perf_counter() - t0
    aP = sum(r[0] for r in res)
    return {"value": aP / max(r[1] for r in res), "unit": "atom-evals/s", "cores": P, "kind": "port",
            "cpu_model": cpu_model(),
            "sample": "the oracle's calculate() on the C/O batch's configurations, %d processes (one per "
                      "core) taking every %d-th configuration for %.0f s each: %d atoms (%.1f s wall)"
                      % (P, P, budget_s, aP, wall),
            "single_thread": {"value": a1 / t1, "cores": 1,
                              "sample": "configurations 0.. in order for %.0f s: %d atoms" % (budget_s, a1)}}


def c2_goldens_error(E, F, V, cfgs):
    """Error of a batch result (raw energies, forces, virials of the first
    configurations) against the reference's FP64 goldens (24 configurations)."""
    gold = json.load(open(os.path.join(ROOT, "tests", "golden", "synth_cases.json")))
    de = df = ds = 0.0
    off = 0
    for k in range(24):
        n = len(cfgs[k][0])
        rec = gold["C2|%d" % k]
        e0 = float(rec["energy"])
        f0 = np.array([[float(v) for v in r] for r in rec["forces"]])
        s0 = np.array([float(v) for v in rec["stress"]])
        vol = abs(np.linalg.det(np.asarray(cfgs[k][2], float).reshape(3, 3)))
        v = V[6 * k: 6 * k + 6]
        s9 = np.array([v[0], v[1], v[2], v[1], v[3], v[4], v[2], v[4], v[5]]) / vol
        de = max(de, abs(E[k] - e0) * EV / n)
        df = max(df, float(np.abs(F[off: off + n] - f0).max()) * EV)
        ds = max(ds, float(np.abs(s9 - s0).max() / np.abs(s0).max()))
        off += n
    return {"energy_eV_per_atom_max": de, "force_max_eV_per_A": df, "stress_max_rel": ds,
            "configurations": 24, "reference": "tests/golden/synth_cases.json (compiled reference, FP64)"}


def c2_plan(cc, L, hip, model, cfgs, prec, steps, warmup):
    """A device-buffer batch plan of the whole batch (autotuned): K replays
    bracketed by device synchronisations.  Returns the form's figures and
    the downloaded outputs of the last replay."""
    nconf = len(cfgs)
    model_of = np.zeros(nconf, np.int32)
    natoms, pos, types, cells = cc._flatten(cfgs, model_of, [model])
    N = pos.size // 3
    d_pos, d_typ = hip.upload(pos), hip.upload(types)
    d_E, d_F, d_V = hip.alloc(8 * nconf), hip.alloc(8 * 3 * N), hip.alloc(8 * 6 * nconf)
    arr = (C.c_void_p * 1)(model.h)
    t0 = time.perf_counter()
    plan = L.chimes_gpu_plan_create_device(arr, 1, nconf, model_of, natoms, d_pos, d_typ, cells, 0, prec, 1, d_E,
                                           d_F, d_V)
    if not plan:
        raise cc.ChimesError(cc._err())
    t_create = time.perf_counter() - t0
    cells_c = cells.ctypes.data_as(C.c_void_p)

    def run():
        if L.chimes_gpu_plan_run_device(plan, cells_c, d_E, d_F, d_V):
            raise cc.ChimesError(cc._err())

    for _ in range(warmup):
        run()
    hip.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        run()
    hip.sync()
    dt = time.perf_counter() - t0
    vv, tune_ms, counts = (C.c_int * 2)(), (C.c_double * (2 * len(cc.VARIANTS)))(), (C.c_longlong * 3)()
    L.chimes_gpu_plan_info(plan, vv, tune_ms, counts)
    L.chimes_gpu_plan_free(plan)
    E = hip.download(d_E, (nconf,), np.float64)
    F = hip.download(d_F, (N, 3), np.float64)
    V = hip.download(d_V, (6 * nconf,), np.float64)
    for p in (d_pos, d_typ, d_E, d_F, d_V):
        hip.L.hipFree(p)
    form = {"value": N * steps / dt, "ms_per_step": dt / steps * 1e3, "variant3": cc.VARIANTS[vv[0]],
            "variant4": cc.VARIANTS[vv[1]], "tune_ms": [round(x, 4) for x in tune_ms],
            "plan_create_s": t_create}
    return form, (E, F, V), (vv[0], vv[1]), tuple(counts)


def c2_kernel_timing(cc, L, hip, model, cfgs, prec, variants, reps=5):
    """Per-kernel-group milliseconds of the batch (HIP events on the engine's
    stream around the 3- and 4-body evaluation launches) from direct device
    evaluations with the plan's kernel variants."""
    nconf = len(cfgs)
    model_of = np.zeros(nconf, np.int32)
    natoms, pos, types, cells = cc._flatten(cfgs, model_of, [model])
    N = pos.size // 3
    d_pos, d_typ = hip.upload(pos), hip.upload(types)
    d_E, d_F, d_V = hip.alloc(8 * nconf), hip.alloc(8 * 3 * N), hip.alloc(8 * 6 * nconf)
    arr = (C.c_void_p * 1)(model.h)
    old = cc.get_variants()
    cc.set_variants(*variants)
    ts = []
    for k in range(reps + 2):
        if L.chimes_gpu_eval_batch_device(arr, 1, nconf, model_of, natoms, d_pos, d_typ, cells, 0, prec, d_E, d_F,
                                          d_V):
            raise cc.ChimesError(cc._err())
        if k >= 2:
            ts.append(model.timing())
    cc.set_variants(*old)
    for p in (d_pos, d_typ, d_E, d_F, d_V):
        hip.L.hipFree(p)
    t = np.mean(np.array(ts), 0)
    return {"pipeline": t[0], "4body": t[1], "3body": t[2], "lists": t[3]}


def bench_c2(args):
    """4096 uneven configurations as one batch plan (device buffers): the
    FP32-mixed plan (headline), the FP64 plan, the plain FP64 drop-in path."""
    import chimes_calculator_amd as cc
    from chimes_calculator_amd import workloads as W

    mname, cfgs = W.workload("C2")
    pf = model_file(mname)
    natoms = atoms_of(cfgs)
    cpu = None if args.no_cpu else c2_cpu_baseline(pf, cfgs)  # before the GPU is touched
    model = cc.GpuModel(pf)
    L = cc.lib()
    L.chimes_gpu_plan_info.argtypes = [C.c_void_p] * 4
    L.chimes_gpu_plan_free.argtypes = [C.c_void_p]
    L.chimes_gpu_plan_run_device.argtypes = [C.c_void_p] * 5
    hip = Hip()
    if args.calib_copy_mib > 0:  # a plain copy of known size, for HBM counter calibration
        nb = args.calib_copy_mib << 20
        a, b = hip.alloc(nb), hip.alloc(nb)
        assert hip.L.hipMemcpy(b, a, nb, 3) == 0
        hip.sync()
        hip.L.hipFree(a)
        hip.L.hipFree(b)
    head_prec = cc.MIXED if args.precision == "mixed" else cc.FP64
    forms = {}
    head, outs, variants, counts = c2_plan(cc, L, hip, model, cfgs, head_prec, args.steps, args.warmup)
    err = c2_goldens_error(*outs, cfgs)
    if not args.plan_only:
        other = cc.FP64 if head_prec == cc.MIXED else cc.MIXED
        f2, o2, _, _ = c2_plan(cc, L, hip, model, cfgs, other, args.steps, args.warmup)
        f2["error_vs_reference_fp64"] = c2_goldens_error(*o2, cfgs)
        forms["fp64_plan" if other == cc.FP64 else "mixed_plan"] = f2
        # the same plan with host buffers: positions in and results out cross PCIe in every replay
        plan = cc.GpuPlan(model, cfgs, precision=head_prec, tune=True)
        _, pos, types, cells = cc._flatten(cfgs, plan.model_of, [model])
        E, F, S = np.zeros(len(cfgs)), np.zeros(pos.shape[0]), np.zeros(9 * len(cfgs))
        for _ in range(args.warmup):
            plan.run_flat(pos, types, cells, E, F, S)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            plan.run_flat(pos, types, cells, E, F, S)
        dt = time.perf_counter() - t0
        plan.close()
        forms["host_buffer_plan"] = {"value": natoms * args.steps / dt, "ms_per_step": dt / args.steps * 1e3,
                                     "note": "positions, types and results cross PCIe inside each replay"}
        # plain FP64 GPU path: the drop-in calculate() once per configuration (host buffers)
        s = cc.ChimesSerial(small=False)
        s.init_chimes(pf)
        for p, t, c in cfgs[:8]:
            s.calculate_chimes(len(p), p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy(), t, c[0], c[1], c[2])
        t0 = time.perf_counter()
        for p, t, c in cfgs:
            s.calculate_chimes(len(p), p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy(), t, c[0], c[1], c[2])
        plain = natoms / (time.perf_counter() - t0)
        s.close()
        forms["plain_fp64_per_config"] = {"value": plain, "note": "calculate_chimes_instance once per "
                                          "configuration, host buffers (the drop-in path)"}
        key = "mixed_plan" if head_prec == cc.FP64 else "fp64_plan"
        fp64v = head["value"] if head_prec == cc.FP64 else forms[key]["value"]
        mixv = head["value"] if head_prec == cc.MIXED else forms[key]["value"]
        forms["speedup_mixed_plan_vs_plain_fp64"] = mixv / plain
        forms["speedup_mixed_plan_vs_fp64_plan"] = mixv / fp64v
    # roofline of the batch's dominant kernel (the larger of the 3- and 4-body
    # evaluations), HIP-event timed on the engine's stream
    kt = c2_kernel_timing(cc, L, hip, model, cfgs, head_prec, variants)
    t3, t4 = param_terms(pf)
    d3, d4 = model.order(3), model.order(4)
    body = 3 if kt["3body"] >= kt["4body"] else 4
    var = variants[body - 3]
    f32 = head_prec == cc.MIXED
    if var == 2 or (body == 4 and d4 > 4) or (var == 3 and (body == 4 or not f32 or d3 > 10)):
        kname = "k_evalv<%d, %d, %s>" % (body, d3 if body == 3 else d4, "float" if f32 else "double")
        unit = "vector ALU"
    elif var == 3:
        kname = "k_eval3v<%d>" % d3
        unit = "matrix cores (one cluster per lane)"
    else:
        kname = "%s<%d, %s, %d, 0>" % ("k_eval3m" if body == 3 else "k_eval4m", d3 if body == 3 else d4,
                                       "float" if f32 else "double", 4 if var == 0 else 2)
        unit = "matrix cores"
    nclus = counts[1] if body == 3 else counts[2]
    terms = (t3 if body == 3 else t4)[0]
    flops = nclus * term_flops(terms, d3 if body == 3 else d4, 3 if body == 3 else 6)
    ms = kt["3body"] if body == 3 else kt["4body"]
    peak = PEAK_FP32 if f32 else PEAK_FP64
    prof = "%s_co4096_%s_counters.md" % ("k_evalv3" if body == 3 else "k_evalv4", "mixed" if f32 else "fp64")
    figs = profile_figures(prof, kname.replace(" ", ""))
    traffic = (figs["fetch_bytes"] + figs["write_bytes"]) if "fetch_bytes" in figs else None
    # bound from the committed counters: HBM when the measured traffic would
    # take over half the launch at 8 TB/s, else the arithmetic units
    bound = "hbm" if traffic is not None and traffic / PEAK_HBM > 0.5 * ms * 1e-3 else "mfma"
    roof = {"bound": bound, "achieved": flops / (ms * 1e-3) / 1e12, "peak": peak / 1e12, "unit": "TFLOP/s",
            "frac": flops / (ms * 1e-3) / peak, "traffic": traffic,
            "kernel": "%s (%d-body evaluation on the %s)" % (kname, body, unit),
            "compute_unit": unit, "flops_per_launch": flops, "kernel_ms": ms,
            "clusters_per_launch": nclus, "term_flops_per_cluster": flops / max(nclus, 1),
            "kernel_groups_ms": kt, "counters_source": "profiles/" + prof if figs else None,
            "counters": figs or None}
    if traffic is not None:
        roof["hbm_TBps_at_measured_traffic"] = traffic / (ms * 1e-3) / 1e12
    if unit == "vector ALU":
        # FP32 arithmetic is built without packed-FP32 instructions, so each
        # FP32 operation issues at the FP64 vector rate (78.6 TFLOP/s)
        roof["frac_of_unpacked_vector_peak"] = flops / (ms * 1e-3) / PEAK_FP64
        if "VALUBusy_pct" in figs:
            roof["valu_issue_busy_pct"] = figs["VALUBusy_pct"]
    out = {"metric": METRIC, "value": head["value"], "unit": "atom-evals/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": head["ms_per_step"], "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32-mixed" if f32 else "f64",
           "data": "synthetic: seeded jittered-lattice C/O configurations and seeded random-coefficient ChIMES model",
           "config": {"workload": "C/O batch: 4096 configurations of 8-512 atoms (489,549 atoms), 1.5-3.5 g/cc, "
                                  "2B/3B/4B orders 12/7/3, rc 4.5/4.0/3.5 A, one batch plan (captured graph), "
                                  "inputs and outputs in HBM, %s" % ("FP32-mixed" if f32 else "FP64"),
                      "nconf": len(cfgs), "natoms": natoms, "pairs": counts[0], "triplets": counts[1],
                      "quadruplets": counts[2]},
           "error_vs_reference_fp64": err,
           "variants": {"3body": head["variant3"], "4body": head["variant4"], "tune_ms": head["tune_ms"]},
           "forms": forms, "roofline": roof}
    if not args.plan_only and not args.no_nested:
        out["forms"]["C1"] = c1_form(args, precision="fp64")
        out["forms"]["C4"] = c4_form(cc, steps_scale=1.0)
        out["latency_ms"] = {"p50": out["forms"]["C4"]["mixed_rate200"]["p50_ms"],
                             "p99": out["forms"]["C4"]["mixed_rate200"]["p99_ms"],
                             "stream": "C4, 200 requests/s, FP32-mixed"}
    if cpu:
        out["cpu_baseline"] = cpu
    print(json.dumps(out))


def random_clusters(rng, nb, n, rmin, rmax, ntypes):
    """n clusters of nb atoms with every pair distance in (rmin, rmax): pair
    distances (order 01 02 03 12 13 23), pair vectors x_b - x_a, types."""
    pairs = [(a, b) for a in range(nb) for b in range(a + 1, nb)]
    out = np.zeros((0, nb, 3))
    while len(out) < n:
        x = rng.uniform(-rmax / 2, rmax / 2, (2 * n, nb, 3))
        d = np.stack([np.linalg.norm(x[:, b] - x[:, a], axis=1) for a, b in pairs], 1)
        out = np.concatenate([out, x[((d > rmin) & (d < rmax)).all(1)]])[:n]
    dr = np.stack([out[:, b] - out[:, a] for a, b in pairs], 1)
    dist = np.linalg.norm(dr, axis=2)
    return dist, dr.reshape(n, -1), rng.integers(0, ntypes, (n, nb)).astype(np.int32)


def bench_clusters(args):
    """Batched explicit clusters (the per-cluster API over arrays): random
    water-model 3- and 4-body clusters, host arrays in and out."""
    import chimes_calculator_amd as cc
    from chimes_calculator_amd import workloads as W

    m = W.MODELS["water_2b3b4b"]
    pf = model_file("water_2b3b4b")
    model = cc.GpuModel(pf)
    rng = np.random.default_rng(0xC0DE05)
    n = 1 << 20
    res = {}
    for nb, rc in ((3, m.rout[1]), (4, m.rout[2])):
        dist, dr, types = random_clusters(rng, nb, n, 1.0, rc, 2)
        for _ in range(args.warmup):
            model.compute_clusters(nb, dist, dr, types)
        ts = []
        for _ in range(args.steps):
            t0 = time.perf_counter()
            model.compute_clusters(nb, dist, dr, types)
            ts.append(time.perf_counter() - t0)
        res["%d-body" % nb] = {"clusters_per_s": n / float(np.median(ts)), "ms_per_call": 1e3 * float(np.median(ts)),
                              "clusters_per_call": n}
    out = {"metric": "explicit clusters/s (E+F+stress per cluster, host arrays in and out)",
           "value": res["4-body"]["clusters_per_s"], "unit": "clusters/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": res["4-body"]["ms_per_call"], "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic: seeded random O/H clusters, pair distances in (1.0, rc) A",
           "config": {"workload": "2^20 clusters per call, water model 12/8/4, rc 5.0/4.5/4.0"}, "forms": res}
    if not args.no_cpu:
        # the oracle's per-cluster routine, one C call per cluster, on a sample
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from oracle_bindings import Oracle
        import ctypes as C
        o = Oracle(pf)
        dist, dr, types = random_clusters(rng, 4, 20000, 1.0, m.rout[2], 2)
        f, st, e = np.zeros(12), np.zeros(9), C.c_double(0.0)
        t0 = time.perf_counter()
        for k in range(len(dist)):
            o.L.oc_compute_4b(o.h, dist[k], dr[k], types[k], f, st, C.byref(e))
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": len(dist) / dt, "unit": "clusters/s", "cores": 1, "kind": "port",
                               "sample": "oracle oc_compute_4b on 20,000 of the same kind of clusters, one call each"}
    print(json.dumps(out))


def lmp_inputs(reps, skin):
    """The reference's 512-atom diamond LAMMPS test box (tests/golden/lmp),
    repeated reps^3 times, as one process's local atoms, ghosts and full
    neighbour list."""
    import gzip
    import tempfile
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from lmp_setup import LmpSystem, read_lammps_data
    G = os.path.join(ROOT, "tests", "golden", "lmp")
    rec = json.load(open(os.path.join(G, "ev_tally_mb.json")))
    pf = os.path.join(tempfile.mkdtemp(), "carbon.txt")
    with gzip.open(os.path.join(G, rec["params"])) as a, open(pf, "wb") as b:
        b.write(a.read())
    x, typ, tag, lo, hi, masses = read_lammps_data(os.path.join(G, rec["data"]))
    L = hi - lo
    shifts = [np.array([i, j, k]) * L for i in range(reps) for j in range(reps) for k in range(reps)]
    X = np.vstack([x + s for s in shifts])
    T = np.tile(typ, len(shifts))
    Tg = np.arange(1, len(X) + 1, dtype=np.int32)
    return pf, masses, X, T, Tg, lo, lo + reps * L, LmpSystem


def bench_lmp(args):
    """The LAMMPS pair style: PairCHIMES::compute's work on one process's
    local and ghost atoms, as a LAMMPS step hands it over."""
    import chimes_calculator_amd as cc

    skin = 1.5
    pf, masses, X, T, Tg, lo, hi, LmpSystem = lmp_inputs(2, skin)
    model = cc.GpuModel(pf)
    rc2 = np.array([[model.max_cutoff(2)]])
    S = LmpSystem(X, T, Tg, lo, hi, [0], rc2, skin)
    P = cc.LmpPair(model, masses)
    args_c = (S.x, S.type, S.tag, S.ilist, S.numneigh, S.nbr_off, S.nbr)
    P.compute(*args_c, ago=0, skin=skin)
    g3, g4 = P.mb_lists()
    # prebuilt row pointers, as a C++ caller passes LAMMPS's firstneigh as is
    L = cc.lib()
    nall = S.nall
    fn, keep = cc.LmpPair._rows(nall, S.ilist, S.nbr_off, S.nbr)
    f, vir, ea, va = np.zeros(nall * 3), np.zeros(6), np.zeros(nall), np.zeros(nall * 6)
    eng = C.c_double(0.0)
    xs = np.ascontiguousarray(S.x.reshape(-1))

    def step(ago, fl):
        if L.chimes_lmp_compute(P.h, nall, xs, S.type, S.tag, S.inum, S.ilist, S.numneigh, fn, ago, skin,
                                fl[0], fl[1], fl[2], fl[3], f, C.byref(eng), vir, ea, va):
            raise cc.ChimesError(cc._err())

    # tally flags (eflag_global, eflag_atom, vflag_global, vflag_atom) of
    # three kinds of LAMMPS step: forces only, a thermo step (energy and
    # pressure), and a step whose computes ask for per-atom energy and stress
    kinds = {"forces": (0, 0, 0, 0), "forces_energy_virial": (1, 0, 1, 0), "all_tallies": (1, 1, 1, 1)}
    forms = {}
    for name, fl in kinds.items():
        for _ in range(args.warmup):
            step(1, fl)
        ts, kt = [], []
        for _ in range(args.steps):
            t0 = time.perf_counter()
            step(1, fl)
            ts.append(time.perf_counter() - t0)
            kt.append(P.timing())
        msk = 1e3 * float(np.median(ts))
        kt = np.median(np.array(kt), 0)
        forms[name] = {"value": S.inum / (msk * 1e-3), "ms_per_step": msk,
                       "phases_ms": {"positions_in": float(kt[0]), "pair_1_2_body": float(kt[2]),
                                     "many_body_and_readback": float(kt[3])}}
    # a rebuild step (ago == 0): row packing and checks, list build, evaluation
    tb, lb = [], []
    for _ in range(max(3, args.steps // 4)):
        t0 = time.perf_counter()
        step(0, kinds["forces_energy_virial"])
        tb.append(time.perf_counter() - t0)
        lb.append(P.timing()[1])
    t_build_call, t_build_lists = float(np.median(tb)), float(np.median(lb))
    ms = forms["forces_energy_virial"]["ms_per_step"]
    kt = forms["forces_energy_virial"]["phases_ms"]
    out = {"metric": "atom-evals/s (LAMMPS pair style: E, F, virial; lists reused between rebuilds)",
           "value": S.inum / (ms * 1e-3), "unit": "atom-evals/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64",
           "data": "the reference's 512-atom diamond LAMMPS test box repeated 2x2x2 (4096 local atoms)",
           "config": {"workload": "4096 local carbon atoms + %d ghosts, Carbon-2.0 2+3+4-body, skin %.1f A, "
                                  "full neighbour list with ghosts" % (nall - S.inum, skin),
                      "local_atoms": int(S.inum), "ghost_atoms": int(nall - S.inum),
                      "neighbour_entries": int(len(S.nbr)), "triplets": int(len(g3)), "quadruplets": int(len(g4))},
           "phases_ms": kt, "forms": forms,
           "rebuild_step_ms": {"call": 1e3 * t_build_call, "lists": t_build_lists}}
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from oracle_bindings import LmpOracle
        from lmp_setup import read_lammps_data, match_types
        O = LmpOracle(pf)
        pf1, masses1, x1, t1, g1, lo1, hi1, _ = lmp_inputs(1, skin)
        S1 = LmpSystem(x1, t1, g1, lo1, hi1, match_types(masses1, O.masses()), O.rc2(), skin)
        tr, qu = O.build_lists(S1)
        t0 = time.perf_counter()
        O.compute(S1, tr, qu)
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": S1.inum / dt, "unit": "atom-evals/s", "cores": 1, "kind": "port",
                               "sample": "the oracle's restatement of PairCHIMES::compute on the 512-atom box "
                                         "(one step, lists prebuilt)"}
    del keep
    P.close()
    model.close()
    print(json.dumps(out))


def fp_rows():
    """Triplet rows (distances ij ik jk, types) of a 128-atom C/O configuration
    under the LAMMPS neighbour rules, and the force field's path."""
    import gzip
    import tempfile
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from lmp_setup import LmpSystem
    from oracle_bindings import read_xyz
    import chimes_calculator_amd as cc
    G = os.path.join(ROOT, "tests", "golden", "ref_inputs")
    d = tempfile.mkdtemp()

    def ungz(name):
        out = os.path.join(d, name[:-3])
        with gzip.open(os.path.join(G, name)) as a, open(out, "wb") as b:
            b.write(a.read())
        return out

    pf = ungz("published_params.CO2400K.2+3+4b.Tersoff.special.offsets.txt.gz")
    pos, names, cell = read_xyz(ungz("CO.9GPa_2400K.OUTCAR_#000.xyz.gz"))
    Lv = np.diag(np.asarray(cell, float).reshape(3, 3))
    pos = np.mod(pos, Lv)
    model = cc.GpuModel(pf)
    ct = model.types_to_index(names)
    nt = model.ntypes
    rc2 = np.full((nt, nt), model.max_cutoff(2))
    S = LmpSystem(pos, ct + 1, np.arange(1, len(pos) + 1, dtype=np.int32), np.zeros(3), Lv,
                  np.arange(nt, dtype=np.int32), rc2, 0.5)
    # triplets as build_mb_neighlists lists them (tag order), within the 3-body cutoff plus the skin
    rc3 = model.max_cutoff(3) + 0.5
    T = []
    for ii in range(S.inum):
        i = S.ilist[ii]
        r = S.nbr[S.nbr_off[ii]:S.nbr_off[ii + 1]]
        d_i = np.linalg.norm(S.x[r] - S.x[i], axis=1)
        J = r[(S.tag[r] >= S.tag[i]) & (d_i < rc3)]
        for a in range(len(J)):
            for b in range(len(J)):
                j, k = J[a], J[b]
                if j != k and S.tag[k] >= S.tag[j] and np.linalg.norm(S.x[k] - S.x[j]) < rc3:
                    T.append((i, j, k))
    T = np.array(T, np.int32)
    dist = np.stack([np.linalg.norm(S.x[T[:, b]] - S.x[T[:, a]], axis=1) for a, b in ((0, 1), (0, 2), (1, 2))], 1)
    return model, dist, S.ctype[T]


def bench_fp(args):
    """Cluster fingerprint histogram: every triplet against every later one."""
    model, dist, types = fp_rows()
    # excluded cluster types are not dumped: keep the tuples the model evaluates
    keep = np.ones(len(types), bool)
    for t in set(map(tuple, types)):
        if model.L.chimes_gpu_fp_edges(model.h, 3, 1, np.zeros(3), np.asarray(t, np.int32), np.zeros(3)):
            keep &= ~np.all(types == np.asarray(t), axis=1)
    dist, types = dist[keep], types[keep]
    t0 = time.perf_counter()
    E = model.fp_edges(3, dist, types)
    t_edges = time.perf_counter() - t0
    n = len(E)
    for _ in range(args.warmup):
        model.fp_histogram(3, E, nbin=100)
    ts = []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        counts, ns, binw = model.fp_histogram(3, E, nbin=100)
        ts.append(time.perf_counter() - t0)
    ms = 1e3 * float(np.median(ts))
    out = {"metric": "cluster pairs binned/s (fingerprint histogram, 3-body, one set against itself)",
           "value": ns / (ms * 1e-3), "unit": "cluster-pairs/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64",
           "data": "triplets of the reference's 128-atom C/O test configuration (CO.9GPa_2400K #000), "
                   "CO2400K force field, LAMMPS tag rule, skin 0.5 A",
           "config": {"workload": "%d triplets, all pairs j > i (%d samples), 100 bins" % (n, ns),
                      "clusters": n, "samples": int(ns)},
           "edges_ms": 1e3 * t_edges}
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from oracle_bindings import fp_hist
        m = 12000
        t0 = time.perf_counter()
        _, ns_cpu, _ = fp_hist(3, E[:m], nbin=100)
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": ns_cpu / dt, "unit": "cluster-pairs/s", "cores": 1, "kind": "port",
                               "sample": "the oracle's gen_flat_hists on the first %d triplets (%d pairs)" % (m, ns_cpu)}
    model.close()
    print(json.dumps(out))


def bench_c3(args):
    """32768-atom carbon at orders 16/10/4, FP64 and FP32-mixed."""
    import chimes_calculator_amd as cc
    from chimes_calculator_amd import workloads as W

    mname, cfgs = W.workload("C3")
    pf = model_file(mname)
    model = cc.GpuModel(pf)
    n = atoms_of(cfgs)
    out = {"metric": "atom-evals/s (E+F+stress) at stated error vs FP64; p50/p99 request latency",
           "unit": "atom-evals/s", "n_gpus": 1, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "data": "synthetic: seeded jittered carbon lattice and seeded ChIMES model",
           "config": {"workload": W.WORKLOADS["C3"], "natoms": n}}
    res = {}
    for prec, tag in ((cc.FP64, "fp64"), (cc.MIXED, "mixed"), (cc.BF16, "bf16"), (cc.BF16X3, "bf16x3")):
        plan = cc.GpuPlan(model, cfgs, precision=prec, tune=True)
        _, pos, types, cells = cc._flatten(cfgs, plan.model_of, [model])
        E, F, S = np.zeros(1), np.zeros(pos.shape[0]), np.zeros(9)
        for _ in range(args.warmup):
            plan.run_flat(pos, types, cells, E, F, S)
        ts = []
        for _ in range(max(3, args.steps // 4)):
            t0 = time.perf_counter()
            plan.run_flat(pos, types, cells, E, F, S)
            ts.append(time.perf_counter() - t0)
        v3, v4, tms, counts = plan.info()
        res[tag] = {"value": n / np.median(ts), "ms_per_step": 1e3 * float(np.median(ts)), "variant3": v3,
                    "variant4": v4, "counts": counts, "E": E[0], "F": F.copy(), "S": S.copy()}
        plan.close()
    a = res["fp64"]
    for tag in ("mixed", "bf16", "bf16x3"):
        b = res[tag]
        out["%s_error_vs_fp64" % tag] = {
            "energy_eV_per_atom": abs(b["E"] - a["E"]) * 0.0433641 / n,
            "force_max_eV_per_A": float(np.abs(b["F"] - a["F"]).max() * 0.0433641),
            "force_rms_eV_per_A": float(np.sqrt(np.mean((b["F"] - a["F"]) ** 2)) * 0.0433641),
            "stress_max_rel": float(np.abs(b["S"] - a["S"]).max() / np.abs(a["S"]).max())}
    for r in res.values():
        del r["E"], r["F"], r["S"]
    out["forms"] = res
    out["value"] = res["mixed"]["value"]
    out["ms_per_step"] = res["mixed"]["ms_per_step"]
    out["dtype"] = "f32-mixed"
    print(json.dumps(out))


def serve_stream(cc, srv, names, reqs, rate_scale=1.0):
    """Replays the arrival times (divided by rate_scale) into the server;
    returns per-request latency (ms, submit to done) and atoms/s."""
    for _, mname, cfg in reqs[:8]:  # warm the slots
        srv.wait(srv.submit(names.index(mname), cfg), 60.0)
    ids, atoms = [], 0
    t0 = time.perf_counter()
    for t_arr, mname, cfg in reqs:
        dt = t0 + t_arr / rate_scale - time.perf_counter()
        if dt > 0:
            time.sleep(dt)
        ids.append(srv.submit(names.index(mname), cfg))
        atoms += len(cfg[0])
    lat, first, last = [], 1e300, 0.0
    for rid in ids:
        srv.wait(rid, 120.0)
        _, _, _, (ts, _, td) = srv.result(rid)
        lat.append(td - ts)
        first, last = min(first, ts), max(last, td)
    return np.array(lat), atoms / ((last - first) * 1e-3)


def c4_form(cc, steps_scale=1.0):
    """The served request stream (256 requests at 200/s, two resident models),
    FP64 and FP32-mixed: latency percentiles; and a burst's throughput."""
    from chimes_calculator_amd import workloads as W
    names = ["co_2b3b4b", "carbon_2b3b"]
    models = [cc.GpuModel(model_file(n)) for n in names]
    reqs = W.serving_stream(nreq=256, rate=200.0)
    forms = {"requests": len(reqs), "atoms": sum(len(r[2][0]) for r in reqs)}
    for prec, tag in ((cc.FP64, "fp64"), (cc.MIXED, "mixed")):
        for scale in (1.0, 1e9):
            srv = cc.GpuServer(models, precision=prec, nslots=2, max_batch_atoms=65536, hbm_budget_gb=32.0)
            lat, thr = serve_stream(cc, srv, names, reqs, scale)
            st = srv.stats()
            srv.close()
            key = "%s_rate200" % tag if scale < 1e6 else "%s_burst" % tag
            forms[key] = {"p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                          "atom_evals_per_s": thr, "batches": st["batches"] - 8, "max_batch": st["max_batch"]}
    for m in models:
        m.close()
    return forms


def bench_c4(args):
    """Poisson request stream (two resident models, 16-4096 atoms) served
    with continuous batching: latency percentiles and throughput."""
    import chimes_calculator_amd as cc
    from chimes_calculator_amd import workloads as W

    names = ["co_2b3b4b", "carbon_2b3b"]
    models = [cc.GpuModel(model_file(n)) for n in names]
    reqs = W.serving_stream(nreq=256, rate=200.0)
    out = {"metric": "atom-evals/s (E+F+stress) at stated error vs FP64; p50/p99 request latency",
           "unit": "atom-evals/s", "n_gpus": 1, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "data": "synthetic: seeded Poisson stream of C/O and carbon configurations",
           "config": {"workload": W.WORKLOADS["C4"], "requests": len(reqs), "rate_per_s": 200.0,
                      "atoms": sum(len(r[2][0]) for r in reqs)}}
    forms = {}
    sizes = np.array([len(r[2][0]) for r in reqs])
    for prec, tag in ((cc.FP64, "fp64"), (cc.MIXED, "mixed")):
        for scale in (1.0, 10.0, 1e9):
            srv = cc.GpuServer(models, precision=prec, nslots=2, max_batch_atoms=65536, hbm_budget_gb=32.0)
            lat, thr = serve_stream(cc, srv, names, reqs, scale)
            st = srv.stats()
            srv.close()
            key = "%s_rate%d" % (tag, int(200 * scale)) if scale < 1e6 else "%s_burst" % tag
            forms[key] = {
                "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                "mean_ms": float(lat.mean()), "atom_evals_per_s": thr, "batches": st["batches"] - 8,
                "max_batch": st["max_batch"],
                # what the tail is made of: the slowest requests and their sizes
                "slowest": [{"ms": round(float(lat[k]), 3), "atoms": int(sizes[k]), "request": int(k)}
                            for k in np.argsort(lat)[::-1][:4]],
                "p50_ms_2048_atoms_up": float(np.percentile(lat[sizes >= 2048], 50)) if (sizes >= 2048).any() else None}
    out["forms"] = forms
    best = forms["mixed_burst"]
    out["value"] = best["atom_evals_per_s"]
    out["latency_ms"] = {"p50": forms["mixed_rate200"]["p50_ms"], "p99": forms["mixed_rate200"]["p99_ms"]}
    out["dtype"] = "f32-mixed"
    print(json.dumps(out))


if __name__ == "__main__":
    main()
