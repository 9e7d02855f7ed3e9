"""MI355X-native ChIMES evaluator: Python host interface.

Loads the in-tree HIP library (chimes_calculator_amd/lib/libchimescalc_dl.so)
and exposes
  * ChimesSerial -- the reference's Python wrapper interface
    (serial_interface/api/chimescalc_serial_py.py:24-101): init_chimes(),
    calculate_chimes() returning (fx, fy, fz, stress, energy);
  * GpuModel / eval_batch -- batched configurations, FP64 or FP32-mixed;
  * cluster evaluation helpers for the per-cluster API.
There is no CPU fallback: without the built library or a GPU every entry
point raises.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libchimescalc_dl.so")

FP64, MIXED, BF16, BF16X3 = 0, 1, 2, 3

_lib = None
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")

# Every symbol the C headers under include/ declare.
EXPORTS = {
    "chimescalc_serial_C.h": ["chimes_open_instance", "chimes_close_instance", "set_chimes_serial",
                              "set_chimes_serial_instance", "init_chimes_serial",
                              "init_chimes_serial_instance", "calculate_chimes", "calculate_chimes_instance"],
    "chimescalc_C.h": ["get_chimes_max_2b_cutoff", "get_chimes_max_3b_cutoff", "get_chimes_max_4b_cutoff",
                       "get_chimes_2b_order", "get_chimes_3b_order", "get_chimes_4b_order", "set_chimes",
                       "init_chimes", "chimes_read_params", "chimes_build_pair_int_trip_map",
                       "chimes_build_pair_int_quad_map", "chimes_compute_2b_props", "chimes_compute_3b_props",
                       "chimes_compute_4b_props", "chimes_compute_2b_props_fromf90",
                       "chimes_compute_3b_props_fromf90", "chimes_compute_4b_props_fromf90"],
    "chimes_gpu.h": ["chimes_gpu_model_load", "chimes_gpu_model_free", "chimes_gpu_last_error",
                     "chimes_gpu_model_ntypes", "chimes_gpu_model_type_index", "chimes_gpu_model_order",
                     "chimes_gpu_model_max_cutoff", "chimes_gpu_eval_batch", "chimes_gpu_eval_batch_device",
                     "chimes_gpu_plan_create", "chimes_gpu_plan_run", "chimes_gpu_plan_create_device",
                     "chimes_gpu_plan_run_device", "chimes_gpu_plan_info", "chimes_gpu_plan_free",
                     "chimes_gpu_set_variants", "chimes_gpu_get_variants", "chimes_gpu_plan_set_variants",
                     "chimes_gpu_server_create", "chimes_gpu_server_submit", "chimes_gpu_server_wait",
                     "chimes_gpu_server_result", "chimes_gpu_server_stats", "chimes_gpu_server_memory",
                     "chimes_gpu_server_failures", "chimes_gpu_server_last_error", "chimes_gpu_server_free",
                     "chimes_gpu_last_timing", "chimes_gpu_last_counts", "chimes_gpu_compute_clusters",
                     "chimes_gpu_compute_clusters_static",
                     "chimes_gpu_cluster_lists", "chimes_gpu_free", "chimes_gpu_fp_edges",
                     "chimes_gpu_fp_histogram"],
    "chimes_lmp.h": ["chimes_lmp_create", "chimes_lmp_free", "chimes_lmp_build_mb_neighlists",
                     "chimes_lmp_compute", "chimes_lmp_mb_lists", "chimes_lmp_fp_rows", "chimes_lmp_timing",
                     "chimes_lmp_set_tabulation"],
}


class ChimesError(RuntimeError):
    pass


def lib():
    """The loaded HIP library (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ChimesError("libchimescalc_dl.so is not built; run __graft_entry__.build() "
                              "or make -C chimes_calculator_amd/csrc")
        L = C.CDLL(LIB_PATH)
        L.chimes_open_instance.restype = C.c_void_p
        L.chimes_close_instance.argtypes = [C.c_void_p]
        L.set_chimes_serial.argtypes = [C.c_int]
        L.set_chimes_serial_instance.argtypes = [C.c_void_p, C.c_int]
        L.init_chimes_serial.argtypes = [C.c_char_p, C.POINTER(C.c_int)]
        L.init_chimes_serial_instance.argtypes = [C.c_void_p, C.c_char_p, C.c_int]
        cargs = [C.c_int, _dp, _dp, _dp, C.c_void_p, _dp, _dp, _dp, C.POINTER(C.c_double), _dp, _dp, _dp, _dp]
        L.calculate_chimes.argtypes = cargs
        L.calculate_chimes_instance.argtypes = [C.c_void_p] + cargs
        L.chimes_gpu_model_load.restype = C.c_void_p
        L.chimes_gpu_model_load.argtypes = [C.c_char_p, C.c_int]
        L.chimes_gpu_model_free.argtypes = [C.c_void_p]
        L.chimes_gpu_last_error.restype = C.c_char_p
        L.chimes_gpu_model_ntypes.argtypes = [C.c_void_p]
        L.chimes_gpu_model_type_index.argtypes = [C.c_void_p, C.c_char_p]
        L.chimes_gpu_model_order.argtypes = [C.c_void_p, C.c_int]
        L.chimes_gpu_model_max_cutoff.restype = C.c_double
        L.chimes_gpu_model_max_cutoff.argtypes = [C.c_void_p, C.c_int]
        L.chimes_gpu_eval_batch.argtypes = [C.c_void_p, C.c_int, C.c_int, _ip, _ip, _dp, _ip, _dp, C.c_int,
                                            C.c_int, _dp, _dp, _dp]
        L.chimes_gpu_eval_batch_device.argtypes = [C.c_void_p, C.c_int, C.c_int, _ip, _ip, C.c_void_p,
                                                   C.c_void_p, _dp, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                                   C.c_void_p]
        L.chimes_gpu_plan_create.restype = C.c_void_p
        L.chimes_gpu_plan_create.argtypes = [C.c_void_p, C.c_int, C.c_int, _ip, _ip, _dp, _ip, _dp, C.c_int,
                                             C.c_int, C.c_int, _dp, _dp, _dp]
        L.chimes_gpu_plan_run.argtypes = [C.c_void_p, _dp, _ip, _dp, _dp, _dp, _dp]
        L.chimes_gpu_plan_create_device.restype = C.c_void_p
        L.chimes_gpu_plan_create_device.argtypes = [C.c_void_p, C.c_int, C.c_int, _ip, _ip, C.c_void_p,
                                                    C.c_void_p, _dp, C.c_int, C.c_int, C.c_int, C.c_void_p,
                                                    C.c_void_p, C.c_void_p]
        L.chimes_gpu_plan_run_device.argtypes = [C.c_void_p, _dp, C.c_void_p, C.c_void_p, C.c_void_p]
        L.chimes_gpu_plan_info.argtypes = [C.c_void_p, _ip, _dp,
                                           np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")]
        L.chimes_gpu_plan_free.argtypes = [C.c_void_p]
        L.chimes_gpu_set_variants.argtypes = [C.c_int, C.c_int, C.c_int]
        L.chimes_gpu_get_variants.argtypes = [C.c_int, _ip]
        L.chimes_gpu_plan_set_variants.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.chimes_gpu_server_create.restype = C.c_void_p
        L.chimes_gpu_server_create.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_long, C.c_double]
        L.chimes_gpu_server_submit.restype = C.c_long
        L.chimes_gpu_server_submit.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp, _ip, _dp, C.c_int]
        L.chimes_gpu_server_wait.restype = C.c_long
        L.chimes_gpu_server_wait.argtypes = [C.c_void_p, C.c_long, C.c_double]
        L.chimes_gpu_server_result.argtypes = [C.c_void_p, C.c_long, C.POINTER(C.c_double), _dp, _dp, _dp]
        L.chimes_gpu_server_stats.argtypes = [C.c_void_p, np.ctypeslib.ndpointer(dtype=np.int64,
                                                                                  flags="C_CONTIGUOUS")]
        L.chimes_gpu_server_memory.argtypes = [C.c_void_p, _dp]
        L.chimes_gpu_server_failures.argtypes = [C.c_void_p, np.ctypeslib.ndpointer(dtype=np.int64,
                                                                                     flags="C_CONTIGUOUS")]
        L.chimes_gpu_server_last_error.restype = C.c_char_p
        L.chimes_gpu_server_free.argtypes = [C.c_void_p]
        L.chimes_gpu_last_timing.argtypes = [C.c_void_p, _dp]
        L.chimes_gpu_last_counts.argtypes = [C.c_void_p, np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")]
        L.chimes_gpu_compute_clusters.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp, _dp, _ip, _dp, _dp, _dp]
        L.chimes_gpu_compute_clusters_static.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp, _dp, _ip, _dp, _dp,
                                                         _dp]
        pp = C.POINTER(C.POINTER(C.c_int))
        L.chimes_gpu_cluster_lists.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_long), pp,
                                               C.POINTER(C.c_long), pp, C.POINTER(C.c_long), pp]
        L.chimes_gpu_free.argtypes = [C.c_void_p]
        L.chimes_gpu_fp_edges.argtypes = [C.c_void_p, C.c_int, C.c_long, _dp, _ip, _dp]
        _lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
        L.chimes_gpu_fp_histogram.argtypes = [C.c_void_p, C.c_int, C.c_long, _dp, C.c_long, _dp, C.c_int, C.c_int,
                                              C.c_double, C.c_double, _lp, C.POINTER(C.c_longlong)]
        ipp = C.POINTER(C.POINTER(C.c_int))
        L.chimes_lmp_create.restype = C.c_void_p
        L.chimes_lmp_create.argtypes = [C.c_void_p, C.c_int, _dp]
        L.chimes_lmp_free.argtypes = [C.c_void_p]
        L.chimes_lmp_build_mb_neighlists.argtypes = [C.c_void_p, C.c_int, _dp, _ip, C.c_int, _ip, _ip, ipp,
                                                     C.c_double]
        L.chimes_lmp_compute.argtypes = [C.c_void_p, C.c_int, _dp, _ip, _ip, C.c_int, _ip, _ip, ipp, C.c_int,
                                         C.c_double, C.c_int, C.c_int, C.c_int, C.c_int, _dp,
                                         C.POINTER(C.c_double), _dp, _dp, _dp]
        L.chimes_lmp_mb_lists.argtypes = [C.c_void_p, C.POINTER(C.c_long), pp, C.POINTER(C.c_long), pp]
        L.chimes_lmp_timing.argtypes = [C.c_void_p, _dp]
        L.chimes_lmp_set_tabulation.argtypes = [C.c_void_p, C.c_int]
        L.chimes_lmp_fp_rows.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_long),
                                         C.POINTER(C.POINTER(C.c_double))]
        _lib = L
    return _lib


def _err():
    return lib().chimes_gpu_last_error().decode()


class GpuModel:
    """A ChIMES parameter set resident on one GPU (chimes_gpu.h)."""

    def __init__(self, param_file, device=0):
        self.L = lib()
        self.h = self.L.chimes_gpu_model_load(os.fsencode(param_file), device)
        if not self.h:
            raise ChimesError(_err())
        self.param_file = param_file

    def close(self):
        if getattr(self, "h", None):
            self.L.chimes_gpu_model_free(self.h)
            self.h = None

    def __del__(self):
        self.close()

    @property
    def ntypes(self):
        return self.L.chimes_gpu_model_ntypes(self.h)

    def type_index(self, name):
        return self.L.chimes_gpu_model_type_index(self.h, name.encode())

    def types_to_index(self, types):
        idx = np.array([self.type_index(t) for t in types], dtype=np.int32)
        if (idx < 0).any():
            raise ChimesError("unknown atom type(s): %s" % sorted({t for t, k in zip(types, idx) if k < 0}))
        return idx

    def order(self, body):
        return self.L.chimes_gpu_model_order(self.h, body)

    def max_cutoff(self, body):
        return self.L.chimes_gpu_model_max_cutoff(self.h, body)

    def timing(self):
        t = np.zeros(4)
        self.L.chimes_gpu_last_timing(self.h, t)
        return t

    def counts(self):
        """(pair images, triplets, quadruplets) of the last evaluation."""
        c = np.zeros(3, np.int64)
        self.L.chimes_gpu_last_counts(self.h, c)
        return tuple(int(v) for v in c)

    def compute_clusters(self, nb, dist, dr, types):
        """Batched explicit clusters; returns (energy[n], force[n, nb, 3], stress[n, 6])."""
        dist = np.ascontiguousarray(dist, dtype=np.float64)
        n = dist.shape[0]
        dr = np.ascontiguousarray(dr, dtype=np.float64).reshape(-1)
        types = np.ascontiguousarray(types, dtype=np.int32).reshape(-1)
        E, F, S = np.zeros(n), np.zeros(n * nb * 3), np.zeros(n * 6)
        if self.L.chimes_gpu_compute_clusters(self.h, nb, n, dist.reshape(-1), dr, types, E, F, S):
            raise ChimesError(_err())
        return E, F.reshape(n, nb, 3), S.reshape(n, 6)

    def fp_edges(self, nb, dist, types):
        """Sorted Morse-transformed edges of dumped cluster rows (calc_histograms.cpp:140-249)."""
        dist = np.ascontiguousarray(dist, np.float64).reshape(-1)
        types = np.ascontiguousarray(types, np.int32).reshape(-1)
        n = types.size // nb
        out = np.zeros(max(dist.size, 1))
        if self.L.chimes_gpu_fp_edges(self.h, nb, n, dist if dist.size else np.zeros(1), types
                                      if types.size else np.zeros(1, np.int32), out):
            raise ChimesError(_err())
        return out[: dist.size].reshape(n, -1)

    def fp_histogram(self, nb, edges1, edges2=None, nbin=100, maxd=None):
        """(counts[nbin], nsamples, binw) of the cluster-distance histogram of
        gen_flat_hists (calc_histograms.cpp:351-489); edges2=None: one set,
        pairs j > i.  maxd defaults to the reference's sqrt(P * 2^2) + 1."""
        P = nb * (nb - 1) // 2
        if maxd is None:
            maxd = (P * 4.0) ** 0.5 + 1.0
        binw = maxd / nbin
        e1 = np.ascontiguousarray(edges1, np.float64).reshape(-1)
        same = edges2 is None
        e2 = e1 if same else np.ascontiguousarray(edges2, np.float64).reshape(-1)
        counts = np.zeros(nbin, np.int64)
        ns = C.c_longlong(0)
        if self.L.chimes_gpu_fp_histogram(self.h, nb, e1.size // P, e1 if e1.size else np.zeros(1), e2.size // P,
                                          e2 if e2.size else np.zeros(1), int(same), nbin, binw, maxd, counts,
                                          C.byref(ns)):
            raise ChimesError(_err())
        return counts, ns.value, binw

    def cluster_lists(self, conf=0):
        """Sorted canonical cluster keys of configuration `conf` of the last evaluation."""
        n2, n3, n4 = C.c_long(), C.c_long(), C.c_long()
        p2, p3, p4 = (C.POINTER(C.c_int)() for _ in range(3))
        if self.L.chimes_gpu_cluster_lists(self.h, conf, C.byref(n2), C.byref(p2), C.byref(n3), C.byref(p3),
                                           C.byref(n4), C.byref(p4)):
            raise ChimesError(_err())
        out = []
        for p, n, w in ((p2, n2.value, 5), (p3, n3.value, 9), (p4, n4.value, 13)):
            a = np.ctypeslib.as_array(p, shape=(max(n, 1) * w,))[: n * w].reshape(n, w).copy()
            self.L.chimes_gpu_free(C.cast(p, C.c_void_p))
            out.append(a)
        return tuple(out)


class LmpPair:
    """The LAMMPS pair style chimesFF on the GPU (chimes_lmp.h), with the
    arrays PairCHIMES::compute reads (etc/lmp/src/pair_chimes.cpp:480-831).

    masses: {LAMMPS type: mass}.  Neighbour rows are given packed: the row of
    owner ilist[ii] is nbr[nbr_off[ii]:nbr_off[ii + 1]] (firstneigh[ilist[ii]]
    in LAMMPS); numneigh is indexed by atom like LAMMPS's.
    """

    def __init__(self, model, masses):
        self.L = lib()
        self.model = model
        nt = max(masses)
        m = np.zeros(nt + 1)
        for t, v in masses.items():
            m[t] = v
        self.h = self.L.chimes_lmp_create(model.h, nt, m)
        if not self.h:
            raise ChimesError(_err())

    def set_tabulation(self, on=True):
        """As the pair style built with -DTABULATION: 2-/3-body interactions
        from the tables the parameter file names (chimes_lmp_set_tabulation)."""
        if self.L.chimes_lmp_set_tabulation(self.h, int(bool(on))):
            raise ChimesError(_err())

    @staticmethod
    def _rows(nall, ilist, nbr_off, nbr):
        rows = (C.POINTER(C.c_int) * nall)()
        base = nbr.ctypes.data
        for ii, i in enumerate(ilist):
            rows[int(i)] = C.cast(base + 4 * int(nbr_off[ii]), C.POINTER(C.c_int))
        return C.cast(rows, C.POINTER(C.POINTER(C.c_int))), rows

    def compute(self, x, type, tag, ilist, numneigh, nbr_off, nbr, ago=0, skin=1.0, eflag_global=1,
                eflag_atom=1, vflag_global=1, vflag_atom=1):
        """Returns (f[nall, 3], eng, virial[6], eatom[nall], vatom[nall, 6])."""
        x = np.ascontiguousarray(x, np.float64).reshape(-1)
        nall = x.size // 3
        type, tag = (np.ascontiguousarray(a, np.int32) for a in (type, tag))
        ilist, numneigh = (np.ascontiguousarray(a, np.int32) for a in (ilist, numneigh))
        nbr = np.ascontiguousarray(nbr, np.int32)
        if nbr.size == 0:
            nbr = np.zeros(1, np.int32)
        fn, keep = self._rows(nall, ilist, nbr_off, nbr)
        f, virial, eatom, vatom = np.zeros(nall * 3), np.zeros(6), np.zeros(nall), np.zeros(nall * 6)
        eng = C.c_double(0.0)
        if self.L.chimes_lmp_compute(self.h, nall, x, type, tag, len(ilist), ilist, numneigh, fn, int(ago),
                                     float(skin), eflag_global, eflag_atom, vflag_global, vflag_atom, f,
                                     C.byref(eng), virial, eatom, vatom):
            raise ChimesError(_err())
        del keep
        return f.reshape(-1, 3), eng.value, virial, eatom, vatom.reshape(-1, 6)

    def build_mb_neighlists(self, x, tag, ilist, numneigh, nbr_off, nbr, skin):
        x = np.ascontiguousarray(x, np.float64).reshape(-1)
        nall = x.size // 3
        tag, ilist, numneigh = (np.ascontiguousarray(a, np.int32) for a in (tag, ilist, numneigh))
        nbr = np.ascontiguousarray(nbr, np.int32)
        if nbr.size == 0:
            nbr = np.zeros(1, np.int32)
        fn, keep = self._rows(nall, ilist, nbr_off, nbr)
        if self.L.chimes_lmp_build_mb_neighlists(self.h, nall, x, tag, len(ilist), ilist, numneigh, fn,
                                                 float(skin)):
            raise ChimesError(_err())
        del keep

    def mb_lists(self):
        """(trips[n3, 3], quads[n4, 4]) of the last build, in the reference's order."""
        n3, n4 = C.c_long(), C.c_long()
        p3, p4 = C.POINTER(C.c_int)(), C.POINTER(C.c_int)()
        if self.L.chimes_lmp_mb_lists(self.h, C.byref(n3), C.byref(p3), C.byref(n4), C.byref(p4)):
            raise ChimesError(_err())
        out = []
        for p, n, w in ((p3, n3.value, 3), (p4, n4.value, 4)):
            a = np.ctypeslib.as_array(p, shape=(max(n, 1) * w,))[: n * w].reshape(n, w).copy()
            self.L.chimes_gpu_free(C.cast(p, C.c_void_p))
            out.append(a)
        return tuple(out)

    def timing(self):
        t = np.zeros(4)
        self.L.chimes_lmp_timing(self.h, t)
        return t

    def fp_rows(self, nb):
        """(distances[n, P], types[n, nb]) of the FINGERPRINT rows of the last
        compute, made on the GPU (chimes_lmp_fp_rows)."""
        n = C.c_long()
        p = C.POINTER(C.c_double)()
        if self.L.chimes_lmp_fp_rows(self.h, nb, C.byref(n), C.byref(p)):
            raise ChimesError(_err())
        P = nb * (nb - 1) // 2
        if n.value == 0:
            return np.zeros((0, P)), np.zeros((0, nb), np.int32)
        a = np.ctypeslib.as_array(p, shape=(n.value * (P + nb),)).reshape(n.value, P + nb).copy()
        self.L.chimes_gpu_free(C.cast(p, C.c_void_p))
        return a[:, :P], a[:, P:].astype(np.int32)

    def close(self):
        if getattr(self, "h", None):
            self.L.chimes_lmp_free(self.h)
            self.h = None

    def __del__(self):
        self.close()


def _flatten(configs, model_of, models):
    natoms = np.array([len(c[0]) for c in configs], dtype=np.int32)
    pos = np.ascontiguousarray(np.concatenate([np.asarray(c[0], dtype=np.float64).reshape(-1, 3)
                                               for c in configs]).reshape(-1))
    types = np.concatenate([models[m].types_to_index(c[1]) for c, m in zip(configs, model_of)]).astype(np.int32)
    cells = np.ascontiguousarray(np.concatenate([np.asarray(c[2], dtype=np.float64).reshape(9) for c in configs]))
    return natoms, pos, types, cells


def eval_batch(models, configs, model_of=None, small=False, precision=FP64):
    """Evaluates configurations [(pos[n,3], types[n], cell[3,3]), ...].

    Returns (energy[nconf], [forces[n_c, 3]], stress[nconf, 9])."""
    if isinstance(models, GpuModel):
        models = [models]
    nconf = len(configs)
    model_of = np.zeros(nconf, np.int32) if model_of is None else np.asarray(model_of, dtype=np.int32)
    natoms, pos, types, cells = _flatten(configs, model_of, models)
    arr = (C.c_void_p * len(models))(*[m.h for m in models])
    E, F, S = np.zeros(nconf), np.zeros(pos.shape[0]), np.zeros(9 * nconf)
    rc = lib().chimes_gpu_eval_batch(arr, len(models), nconf, model_of, natoms, pos, types, cells, int(small),
                                     int(precision), E, F, S)
    if rc:
        raise ChimesError(_err())
    F = F.reshape(-1, 3)
    offs = np.concatenate([[0], np.cumsum(natoms)])
    return E, [F[offs[k]:offs[k + 1]] for k in range(nconf)], S.reshape(nconf, 9)


VARIANTS = ("mfma64", "mfma32", "vector", "mfmav")


def set_variants(var3=-1, var4=-1, device=0):
    """Kernel variants of the 3- and 4-body terms (indices into VARIANTS, -1
    keeps) for every later evaluation on `device` outside plans."""
    if lib().chimes_gpu_set_variants(int(device), int(var3), int(var4)):
        raise ChimesError(_err())


def get_variants(device=0):
    v = np.zeros(2, np.int32)
    if lib().chimes_gpu_get_variants(int(device), v):
        raise ChimesError(_err())
    return int(v[0]), int(v[1])


class GpuPlan:
    """A batch plan (chimes_gpu.h): one batch shape evaluated by replaying a
    captured HIP graph.  Construction evaluates `configs` once (see
    `first`); run() takes new configurations of the same shape."""

    def __init__(self, models, configs, model_of=None, small=False, precision=FP64, tune=True):
        self.L = lib()
        if isinstance(models, GpuModel):
            models = [models]
        self.models = models
        self.nconf = len(configs)
        self.model_of = np.zeros(self.nconf, np.int32) if model_of is None else np.asarray(model_of, np.int32)
        self.natoms, pos, types, cells = _flatten(configs, self.model_of, models)
        self._arr = (C.c_void_p * len(models))(*[m.h for m in models])
        E, F, S = np.zeros(self.nconf), np.zeros(pos.shape[0]), np.zeros(9 * self.nconf)
        self.h = self.L.chimes_gpu_plan_create(self._arr, len(models), self.nconf, self.model_of, self.natoms, pos,
                                               types, cells, int(small), int(precision), int(bool(tune)), E, F, S)
        if not self.h:
            raise ChimesError(_err())
        self.first = self._split(E, F, S)

    def _split(self, E, F, S):
        F = F.reshape(-1, 3)
        offs = np.concatenate([[0], np.cumsum(self.natoms)])
        return E, [F[offs[k]:offs[k + 1]] for k in range(self.nconf)], S.reshape(self.nconf, 9)

    def run(self, configs):
        _, pos, types, cells = _flatten(configs, self.model_of, self.models)
        E, F, S = np.zeros(self.nconf), np.zeros(pos.shape[0]), np.zeros(9 * self.nconf)
        if self.L.chimes_gpu_plan_run(self.h, pos, types, cells, E, F, S):
            raise ChimesError(_err())
        return self._split(E, F, S)

    def run_flat(self, pos, types, cells, E, F, S):
        """Replays on pre-flattened arrays (pos[3n], types[n], cells[9c]) into E[c], F[3n], S[9c]."""
        if self.L.chimes_gpu_plan_run(self.h, pos, types, cells, E, F, S):
            raise ChimesError(_err())

    def set_variants(self, var3=-1, var4=-1):
        """Forces the kernel variants of this plan (the next run captures again)."""
        if self.L.chimes_gpu_plan_set_variants(self.h, int(var3), int(var4)):
            raise ChimesError(_err())

    def info(self):
        """(3-body variant, 4-body variant, tuning ms per (body, variant) - negative
        for a variant the tuner rejected - , (pairs, triplets, quads))."""
        v, t, c = np.zeros(2, np.int32), np.zeros(2 * len(VARIANTS)), np.zeros(3, np.int64)
        self.L.chimes_gpu_plan_info(self.h, v, t, c)
        return VARIANTS[v[0]], VARIANTS[v[1]], t.reshape(2, len(VARIANTS)), tuple(int(x) for x in c)

    def close(self):
        if getattr(self, "h", None):
            self.L.chimes_gpu_plan_free(self.h)
            self.h = None

    def __del__(self):
        self.close()


class GpuServer:
    """Continuous-batching request loop over resident models (chimes_gpu.h).
    submit() returns a request id at once; wait()/result() collect it."""

    def __init__(self, models, precision=FP64, nslots=2, max_batch_atoms=65536, hbm_budget_gb=16.0):
        self.L = lib()
        self.models = list(models)
        arr = (C.c_void_p * len(self.models))(*[m.h for m in self.models])
        self.h = self.L.chimes_gpu_server_create(arr, len(self.models), int(precision), int(nslots),
                                                 int(max_batch_atoms), float(hbm_budget_gb))
        if not self.h:
            raise ChimesError(self.L.chimes_gpu_server_last_error().decode())
        self.natoms = {}

    def submit(self, model, config, small=False):
        pos, types, cell = config
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1)
        idx = self.models[model].types_to_index(types)
        cell = np.ascontiguousarray(cell, dtype=np.float64).reshape(9)
        rid = self.L.chimes_gpu_server_submit(self.h, int(model), len(idx), pos, idx, cell, int(small))
        if rid < 0:
            raise ChimesError(self.L.chimes_gpu_server_last_error().decode())
        self.natoms[rid] = len(idx)
        return rid

    def wait(self, rid=0, timeout=60.0):
        """Id of a completed request (rid, or any if rid == 0); 0 on timeout."""
        r = self.L.chimes_gpu_server_wait(self.h, int(rid), float(timeout))
        if r < 0:
            raise ChimesError(self.L.chimes_gpu_server_last_error().decode())
        return r

    def result(self, rid):
        """(energy, forces[n, 3], stress[9], (t_submit, t_start, t_done) in ms)."""
        n = self.natoms[rid]
        E = C.c_double()
        F, S, T = np.zeros(3 * n), np.zeros(9), np.zeros(3)
        if self.L.chimes_gpu_server_result(self.h, int(rid), C.byref(E), F, S, T):
            msg = self.L.chimes_gpu_server_last_error().decode()
            if msg != "request not complete":  # a failed request is released too
                self.natoms.pop(rid, None)
            raise ChimesError(msg)
        self.natoms.pop(rid, None)
        return E.value, F.reshape(n, 3), S, tuple(T)

    def stats(self):
        s = np.zeros(4, np.int64)
        self.L.chimes_gpu_server_stats(self.h, s)
        f = np.zeros(2, np.int64)
        self.L.chimes_gpu_server_failures(self.h, f)
        return dict(requests=int(s[0]), batches=int(s[1]), atoms=int(s[2]), max_batch=int(s[3]),
                    failed=int(f[0]), isolated_batches=int(f[1]))

    def memory(self):
        """Bytes: budget, peak estimated in flight, peak held by the slots; the
        largest needed/estimated ratio of a batch; batches above their estimate."""
        m = np.zeros(5)
        self.L.chimes_gpu_server_memory(self.h, m)
        return dict(budget=m[0], peak_inflight=m[1], peak_held=m[2], worst_ratio=m[3], underestimated=int(m[4]))

    def close(self):
        if getattr(self, "h", None):
            self.L.chimes_gpu_server_free(self.h)
            self.h = None

    def __del__(self):
        self.close()


class ChimesSerial:
    """System-level interface with the reference Python wrapper's calls
    (serial_interface/api/chimescalc_serial_py.py:24-101) over the C ABI of
    chimescalc_serial_C.h, one handle per object."""

    def __init__(self, small=False):
        self.L = lib()
        self.h = self.L.chimes_open_instance()
        self.L.set_chimes_serial_instance(self.h, int(bool(small)))

    def init_chimes(self, param_file, rank=0):
        self.L.init_chimes_serial_instance(self.h, os.fsencode(param_file), int(rank))

    def calculate_chimes(self, natom, xc, yc, zc, atom_types, ca, cb, cc, energy=0.0, fx=None, fy=None, fz=None,
                         stress=None):
        """Returns (fx, fy, fz, stress, energy) like the reference wrapper."""
        f = [np.ascontiguousarray(v if v is not None else np.zeros(natom), dtype=np.float64)
             for v in (fx, fy, fz)]
        st = np.ascontiguousarray(stress if stress is not None else np.zeros(9), dtype=np.float64)
        e = C.c_double(energy)
        ts = (C.c_char_p * natom)(*[t.encode() for t in atom_types])
        arrs = [np.ascontiguousarray(v, dtype=np.float64) for v in (xc, yc, zc, ca, cb, cc)]
        self.L.calculate_chimes_instance(self.h, natom, arrs[0], arrs[1], arrs[2], C.cast(ts, C.c_void_p),
                                         arrs[3], arrs[4], arrs[5], C.byref(e), f[0], f[1], f[2], st)
        return f[0], f[1], f[2], st, e.value

    def close(self):
        if getattr(self, "h", None):
            self.L.chimes_close_instance(self.h)
            self.h = None

    def __del__(self):
        self.close()
