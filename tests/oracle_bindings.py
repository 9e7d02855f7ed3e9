"""ctypes bindings for the oracle (test infrastructure only).

Loads oracle/liboracle.so (the C restatement of the reference path) and, when
it was built in this checkout, oracle/_ref/libchimes_ref.so (the reference's
own classes compiled from its sources).  Only tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg import this module.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libchimes_ref.so")
REF_TAB_SO = os.path.join(ROOT, "oracle", "_ref", "libchimes_ref_tab.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


class _Lists(C.Structure):
    _fields_ = [("n_atoms", C.c_int), ("n_layers", C.c_int), ("n_rep", C.c_int),
                ("n2", C.c_long), ("n3", C.c_long), ("n4", C.c_long),
                ("pairs", C.POINTER(C.c_int)), ("trips", C.POINTER(C.c_int)),
                ("quads", C.POINTER(C.c_int))]


def _cstrs(types):
    arr = (C.c_char_p * len(types))()
    arr[:] = [t.encode() for t in types]
    return arr


def _keys(ptr, n, width):
    if n == 0:
        return np.zeros((0, width), np.int32)
    return np.ctypeslib.as_array(ptr, shape=(n * width,)).reshape(n, width).copy()


_oracle = None


def oracle_lib():
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle/liboracle.so is not built (run make -C oracle)")
        L = C.CDLL(ORACLE_SO)
        L.oc_load.restype = C.c_void_p
        L.oc_load.argtypes = [C.c_char_p]
        L.oc_set_exact_sums.argtypes = [C.c_int]
        L.oc_free.argtypes = [C.c_void_p]
        L.oc_order.argtypes = [C.c_void_p, C.c_int]
        L.oc_max_cutoff.restype = C.c_double
        L.oc_max_cutoff.argtypes = [C.c_void_p, C.c_int]
        L.oc_ntypes.argtypes = [C.c_void_p]
        L.oc_type_name.restype = C.c_char_p
        L.oc_type_name.argtypes = [C.c_void_p, C.c_int]
        L.oc_calculate.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp, C.c_void_p, _dp, _dp, _dp,
                                   C.c_int, C.POINTER(C.c_double), _dp, _dp, _dp, _dp,
                                   C.POINTER(_Lists), C.c_int]
        L.oc_lists_free.argtypes = [C.POINTER(_Lists)]
        L.oc_compute_2b.argtypes = [C.c_void_p, C.c_double, _dp, _ip, _dp, _dp, C.POINTER(C.c_double)]
        L.oc_compute_3b.argtypes = [C.c_void_p, _dp, _dp, _ip, _dp, _dp, C.POINTER(C.c_double)]
        L.oc_compute_4b.argtypes = [C.c_void_p, _dp, _dp, _ip, _dp, _dp, C.POINTER(C.c_double)]
        L.oc_cheby.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double, C.c_int, _dp, _dp]
        L.oc_fcut.argtypes = [C.c_void_p, C.c_double, C.c_double, C.POINTER(C.c_double),
                              C.POINTER(C.c_double)]
        L.oc_last_penalty_count.restype = C.c_long
        L.oc_last_out_of_range_count.restype = C.c_long
        L.oc_mass.restype = C.c_double
        L.oc_mass.argtypes = [C.c_void_p, C.c_int]
        L.oc_cutoff_2b.restype = C.c_double
        L.oc_cutoff_2b.argtypes = [C.c_void_p, C.c_int, C.c_int]
        _lp = C.POINTER(C.c_long)
        _ipp = C.POINTER(C.POINTER(C.c_int))
        L.oc_lmp_build_lists.argtypes = [C.c_void_p, _dp, _ip, C.c_int, _ip, _ip, _ip, _ip, C.c_double,
                                         _lp, _ipp, _lp, _ipp]
        L.oc_lmp_compute.argtypes = [C.c_void_p, _dp, _ip, _ip, C.c_int, _ip, _ip, _ip, _ip, C.c_long, _ip,
                                     C.c_long, _ip, C.c_int, C.c_int, C.c_int, C.c_int, _dp,
                                     C.POINTER(C.c_double), _dp, _dp, _dp]
        L.oc_lmp_free.argtypes = [C.c_void_p]
        L.oc_lmp_fp_rows.restype = C.c_long
        L.oc_lmp_fp_rows.argtypes = [C.c_void_p, C.c_int, _dp, _ip, _ip, C.c_int, _ip, _ip, _ip, _ip, C.c_long, _ip,
                                     C.c_long, _ip, C.POINTER(C.POINTER(C.c_double))]
        L.oc_cluster_pair_cutoffs.argtypes = [C.c_void_p, C.c_int, _ip, C.c_int, C.POINTER(C.c_double),
                                              C.POINTER(C.c_double)]
        L.oc_fp_edges.argtypes = [C.c_void_p, C.c_int, C.c_long, _dp, _ip, _dp]
        _llp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
        L.oc_fp_hist.argtypes = [C.c_int, C.c_long, _dp, C.c_long, _dp, C.c_int, C.c_int, C.c_double, C.c_double,
                                 _llp, C.POINTER(C.c_longlong)]
        L.oc_tab_load.restype = C.c_void_p
        L.oc_tab_load.argtypes = [C.c_void_p, C.c_char_p]
        L.oc_tab_free.argtypes = [C.c_void_p]
        L.oc_tab_error.restype = C.c_char_p
        L.oc_tab_bodies.argtypes = [C.c_void_p]
        L.oc_tab_2b_value.restype = C.c_double
        L.oc_tab_2b_value.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_int, C.POINTER(C.c_int)]
        L.oc_tab_3b_interp.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_double, _dp,
                                       C.POINTER(C.c_int)]
        L.oc_compute_2b_tab.argtypes = [C.c_void_p, C.c_void_p, C.c_double, _dp, _ip, _dp, _dp,
                                        C.POINTER(C.c_double)]
        L.oc_compute_3b_tab.argtypes = [C.c_void_p, C.c_void_p, _dp, _dp, _ip, _dp, _dp, C.POINTER(C.c_double)]
        L.oc_lmp_use_tables.argtypes = [C.c_void_p]
        L.oc_lmp_tab_errors.restype = C.c_long
        _oracle = L
    return _oracle


class Result:
    def __init__(self, energy, forces, stress, pairs=None, trips=None, quads=None, meta=None):
        self.energy = energy
        self.forces = forces
        self.stress = stress
        self.pairs, self.trips, self.quads = pairs, trips, quads
        self.meta = meta


class Oracle:
    """The C restatement of the reference's calculate() path."""

    def __init__(self, param_file):
        self.L = oracle_lib()
        self.h = self.L.oc_load(param_file.encode())
        if not self.h:
            raise ValueError("oracle could not parse %s" % param_file)

    def __del__(self):
        if getattr(self, "h", None):
            self.L.oc_free(self.h)
            self.h = None

    def order(self, body):
        return self.L.oc_order(self.h, body)

    def max_cutoff(self, body):
        return self.L.oc_max_cutoff(self.h, body)

    def type_names(self):
        return [self.L.oc_type_name(self.h, t).decode() for t in range(self.L.oc_ntypes(self.h))]

    def calculate(self, pos, types, cell, small=False, lists=False, owner_stride=1, exact_sums=False):
        """exact_sums: per-cluster energy and stress sums added in long double
        (oc_set_exact_sums) instead of the reference's running double totals."""
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        n = pos.shape[0]
        x, y, z = (np.ascontiguousarray(pos[:, k]) for k in range(3))
        cell = np.asarray(cell, dtype=np.float64).reshape(3, 3)
        ca, cb, cc = (np.ascontiguousarray(cell[k]) for k in range(3))
        fx, fy, fz = np.zeros(n), np.zeros(n), np.zeros(n)
        st = np.zeros(9)
        e = C.c_double(0.0)
        L = _Lists()
        ts = _cstrs(types)
        self.L.oc_set_exact_sums(int(bool(exact_sums)))
        try:
            rc = self.L.oc_calculate(self.h, n, x, y, z, C.cast(ts, C.c_void_p), ca, cb, cc,
                                     int(bool(small)), C.byref(e), fx, fy, fz, st,
                                     C.byref(L) if lists else None, int(owner_stride))
        finally:
            self.L.oc_set_exact_sums(0)
        if rc != 0:
            raise RuntimeError("oracle calculate failed with code %d" % rc)
        res = Result(e.value, np.stack([fx, fy, fz], 1), st)
        if lists:
            res.pairs = _keys(L.pairs, L.n2, 5)
            res.trips = _keys(L.trips, L.n3, 9)
            res.quads = _keys(L.quads, L.n4, 13)
            res.meta = (L.n_atoms, L.n_layers, L.n_rep)
            self.L.oc_lists_free(C.byref(L))
        return res


def fp_edges(oracle, nb, dist, types):
    """The oracle's read_flat_clusters edges of dumped rows."""
    dist = np.ascontiguousarray(dist, np.float64).reshape(-1)
    types = np.ascontiguousarray(types, np.int32).reshape(-1)
    n = types.size // nb
    out = np.zeros(max(dist.size, 1))
    if oracle.L.oc_fp_edges(oracle.h, nb, n, dist if dist.size else np.zeros(1), types, out):
        raise RuntimeError("oracle fp_edges failed")
    return out[: dist.size].reshape(n, -1)


def fp_hist(nb, e1, e2=None, nbin=100, maxd=None):
    """The oracle's gen_flat_hists: (counts, nsamples, binw)."""
    L = oracle_lib()
    P = nb * (nb - 1) // 2
    if maxd is None:
        maxd = (P * 4.0) ** 0.5 + 1.0
    binw = maxd / nbin
    a = np.ascontiguousarray(e1, np.float64).reshape(-1)
    same = e2 is None
    b = a if same else np.ascontiguousarray(e2, np.float64).reshape(-1)
    counts = np.zeros(nbin, np.int64)
    ns = C.c_longlong(0)
    if L.oc_fp_hist(P, a.size // P, a, b.size // P, b, int(same), nbin, binw, maxd, counts, C.byref(ns)):
        raise RuntimeError("oracle histogram: distance outside the bins")
    return counts, ns.value, binw


class TabOracle:
    """The tabulated 2-/3-body restatement (chimes_tab_oracle.c) over the
    tables a parameter file names, with the oracle's parameter set."""

    def __init__(self, oracle, param_file):
        self.O = oracle
        self.L = oracle.L
        self.h = self.L.oc_tab_load(oracle.h, param_file.encode())
        if not self.h:
            raise ValueError(self.L.oc_tab_error().decode())

    def __del__(self):
        if getattr(self, "h", None):
            self.L.oc_tab_free(self.h)
            self.h = None

    def bodies(self):
        return self.L.oc_tab_bodies(self.h)

    def value_2b(self, pair, r, for_energy):
        err = C.c_int(0)
        v = self.L.oc_tab_2b_value(self.h, pair, r, int(for_energy), C.byref(err))
        return v, err.value

    def interp_3b(self, trip, rij, rik, rjk):
        out, err = np.zeros(4), C.c_int(0)
        self.L.oc_tab_3b_interp(self.h, trip, rij, rik, rjk, out, C.byref(err))
        return out, err.value

    def compute_2b(self, dx, dr, t):
        f, s, e = np.zeros(6), np.zeros(6), C.c_double(0.0)
        rc = self.L.oc_compute_2b_tab(self.O.h, self.h, dx, np.ascontiguousarray(dr, np.float64),
                                      np.ascontiguousarray(t, np.int32), f, s, C.byref(e))
        return f, s, e.value, rc

    def compute_3b(self, dx, dr, t):
        f, s, e = np.zeros(9), np.zeros(6), C.c_double(0.0)
        rc = self.L.oc_compute_3b_tab(self.O.h, self.h, np.ascontiguousarray(dx, np.float64),
                                      np.ascontiguousarray(dr, np.float64), np.ascontiguousarray(t, np.int32), f, s,
                                      C.byref(e))
        return f, s, e.value, rc


class LmpResult:
    def __init__(self, f, eng, virial, eatom, vatom):
        self.f, self.eng, self.virial, self.eatom, self.vatom = f, eng, virial, eatom, vatom


class LmpOracle(Oracle):
    """The C restatement of the LAMMPS pair style (pair_chimes.cpp) on one
    process, over an LmpSystem (tests/lmp_setup.py)."""

    def masses(self):
        return [self.L.oc_mass(self.h, t) for t in range(self.L.oc_ntypes(self.h))]

    def rc2(self):
        nt = self.L.oc_ntypes(self.h)
        return np.array([[self.L.oc_cutoff_2b(self.h, a, b) for b in range(nt)] for a in range(nt)])

    def build_lists(self, S):
        n3, n4 = C.c_long(0), C.c_long(0)
        p3, p4 = C.POINTER(C.c_int)(), C.POINTER(C.c_int)()
        rc = self.L.oc_lmp_build_lists(self.h, S.x.reshape(-1), S.tag, S.inum, S.ilist, S.numneigh, S.nbr_off,
                                       S.nbr, S.skin, C.byref(n3), C.byref(p3), C.byref(n4), C.byref(p4))
        if rc != 0:
            raise RuntimeError("oracle list build failed")
        trips = _keys(p3, n3.value, 3)
        quads = _keys(p4, n4.value, 4)
        self.L.oc_lmp_free(C.cast(p3, C.c_void_p))
        self.L.oc_lmp_free(C.cast(p4, C.c_void_p))
        return trips, quads

    def fp_rows(self, S, trips, quads, nb):
        """(distances[n, P], types[n, nb]) of the FINGERPRINT rows
        (oc_lmp_fp_rows) over the system's pairs and the given lists."""
        trips = np.ascontiguousarray(trips, np.int32).reshape(-1)
        quads = np.ascontiguousarray(quads, np.int32).reshape(-1)
        p = C.POINTER(C.c_double)()
        n = self.L.oc_lmp_fp_rows(self.h, nb, S.x.reshape(-1), S.ctype, S.tag, S.inum, S.ilist, S.numneigh,
                                  S.nbr_off, S.nbr, len(trips) // 3, trips if len(trips) else np.zeros(3, np.int32),
                                  len(quads) // 4, quads if len(quads) else np.zeros(4, np.int32), C.byref(p))
        P = nb * (nb - 1) // 2
        if n == 0:
            return np.zeros((0, P)), np.zeros((0, nb), np.int32)
        a = np.ctypeslib.as_array(p, shape=(n * (P + nb),)).reshape(n, P + nb).copy()
        self.L.oc_lmp_free(C.cast(p, C.c_void_p))
        return a[:, :P], a[:, P:].astype(np.int32)

    def compute(self, S, trips, quads, eflag_global=1, eflag_atom=1, vflag_global=1, vflag_atom=1, tables=None):
        """tables: a TabOracle whose tabulated interactions replace the
        polynomial ones (a -DTABULATION build of the pair style)."""
        self.L.oc_lmp_use_tables(tables.h if tables is not None else None)
        try:
            return self._compute(S, trips, quads, eflag_global, eflag_atom, vflag_global, vflag_atom)
        finally:
            self.tab_errors = self.L.oc_lmp_tab_errors()
            self.L.oc_lmp_use_tables(None)

    def _compute(self, S, trips, quads, eflag_global, eflag_atom, vflag_global, vflag_atom):
        trips = np.ascontiguousarray(trips, np.int32).reshape(-1)
        quads = np.ascontiguousarray(quads, np.int32).reshape(-1)
        f = np.zeros(S.nall * 3)
        virial, eatom, vatom = np.zeros(6), np.zeros(S.nall), np.zeros(S.nall * 6)
        eng = C.c_double(0.0)
        self.L.oc_lmp_compute(self.h, S.x.reshape(-1), S.ctype, S.tag, S.inum, S.ilist, S.numneigh, S.nbr_off,
                              S.nbr, len(trips) // 3, trips if len(trips) else np.zeros(3, np.int32),
                              len(quads) // 4, quads if len(quads) else np.zeros(4, np.int32),
                              eflag_global, eflag_atom, vflag_global, vflag_atom, f, C.byref(eng), virial, eatom,
                              vatom)
        return LmpResult(f.reshape(-1, 3), eng.value, virial, eatom, vatom.reshape(-1, 6))


_ref = None


def ref_available():
    return os.path.exists(REF_SO)


def ref_lib():
    global _ref
    if _ref is None:
        L = C.CDLL(REF_SO)
        L.ref_open.restype = C.c_void_p
        L.ref_open.argtypes = [C.c_char_p, C.c_int]
        L.ref_close.argtypes = [C.c_void_p]
        pp = C.POINTER(C.POINTER(C.c_int))
        L.ref_calculate.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp, C.c_void_p, _dp, _dp, _dp,
                                    C.POINTER(C.c_double), _dp, _dp, _dp, _dp,
                                    C.POINTER(C.c_long), pp, C.POINTER(C.c_long), pp,
                                    C.POINTER(C.c_long), pp, _ip]
        L.ref_free.argtypes = [C.c_void_p]
        L.ref_compute_2b.argtypes = [C.c_void_p, C.c_double, _dp, _ip, _dp, _dp, C.POINTER(C.c_double)]
        L.ref_compute_3b.argtypes = [C.c_void_p, _dp, _dp, _ip, _dp, _dp, C.POINTER(C.c_double)]
        L.ref_compute_4b.argtypes = [C.c_void_p, _dp, _dp, _ip, _dp, _dp, C.POINTER(C.c_double)]
        _ref = L
    return _ref


class Reference:
    """The reference's own classes (compiled from its sources) behind a C shim."""

    def __init__(self, param_file, small=False):
        self.L = ref_lib()
        self.h = self.L.ref_open(param_file.encode(), int(bool(small)))

    def __del__(self):
        if getattr(self, "h", None):
            self.L.ref_close(self.h)
            self.h = None

    def calculate(self, pos, types, cell, lists=False):
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        n = pos.shape[0]
        x, y, z = (np.ascontiguousarray(pos[:, k]) for k in range(3))
        cell = np.asarray(cell, dtype=np.float64).reshape(3, 3)
        ca, cb, cc = (np.ascontiguousarray(cell[k]) for k in range(3))
        fx, fy, fz = np.zeros(n), np.zeros(n), np.zeros(n)
        st = np.zeros(9)
        e = C.c_double(0.0)
        ts = _cstrs(types)
        n2, n3, n4 = C.c_long(0), C.c_long(0), C.c_long(0)
        p2, p3, p4 = C.POINTER(C.c_int)(), C.POINTER(C.c_int)(), C.POINTER(C.c_int)()
        meta = np.zeros(3, np.int32)
        if lists:
            self.L.ref_calculate(self.h, n, x, y, z, C.cast(ts, C.c_void_p), ca, cb, cc, C.byref(e),
                                 fx, fy, fz, st, C.byref(n2), C.byref(p2), C.byref(n3),
                                 C.byref(p3), C.byref(n4), C.byref(p4), meta)
        else:
            self.L.ref_calculate(self.h, n, x, y, z, C.cast(ts, C.c_void_p), ca, cb, cc, C.byref(e),
                                 fx, fy, fz, st, None, None, None, None, None, None, meta)
        res = Result(e.value, np.stack([fx, fy, fz], 1), st, meta=tuple(meta))
        if lists:
            res.pairs = _keys(p2, n2.value, 5)
            res.trips = _keys(p3, n3.value, 9)
            res.quads = _keys(p4, n4.value, 13)
            for p in (p2, p3, p4):
                self.L.ref_free(C.cast(p, C.c_void_p))
        return res


_ref_tab = None


def ref_tab_available():
    return os.path.exists(REF_TAB_SO)


def ref_tab_lib():
    global _ref_tab
    if _ref_tab is None:
        L = C.CDLL(REF_TAB_SO)
        L.reft_open.restype = C.c_void_p
        L.reft_open.argtypes = [C.c_char_p, C.c_char_p]
        L.reft_close.argtypes = [C.c_void_p]
        L.reft_flags.argtypes = [C.c_void_p]
        L.reft_get_tab_2b.restype = C.c_double
        L.reft_get_tab_2b.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_int]
        L.reft_interp_3b.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_double, _dp]
        L.reft_compute_2b.argtypes = [C.c_void_p, C.c_double, _dp, _ip, _dp, _dp, C.POINTER(C.c_double)]
        L.reft_compute_3b.argtypes = [C.c_void_p, _dp, _dp, _ip, _dp, _dp, C.POINTER(C.c_double)]
        _ref_tab = L
    return _ref_tab


class ReferenceTab:
    """The reference's chimesFF compiled with -DTABULATION, one call at a time
    (oracle/ref_tab_driver.cpp).  Its table readers exit the process on a
    missing file and its lookups on a distance below a 2-body table, so
    callers keep to inputs inside the tables."""

    def __init__(self, param_file):
        self.L = ref_tab_lib()
        d, f = os.path.split(os.path.abspath(param_file))
        self.h = self.L.reft_open(d.encode(), f.encode())
        if not self.h:
            raise RuntimeError("reference could not open %s" % param_file)

    def __del__(self):
        if getattr(self, "h", None):
            self.L.reft_close(self.h)
            self.h = None

    def flags(self):
        return self.L.reft_flags(self.h)

    def value_2b(self, pair, r, for_energy):
        return self.L.reft_get_tab_2b(self.h, pair, r, int(for_energy))

    def interp_3b(self, trip, rij, rik, rjk):
        out = np.zeros(4)
        self.L.reft_interp_3b(self.h, trip, rij, rik, rjk, out)
        return out

    def compute_2b(self, dx, dr, t):
        f, s, e = np.zeros(6), np.zeros(6), C.c_double(0.0)
        self.L.reft_compute_2b(self.h, dx, np.ascontiguousarray(dr, np.float64), np.ascontiguousarray(t, np.int32),
                               f, s, C.byref(e))
        return f, s, e.value

    def compute_3b(self, dx, dr, t):
        f, s, e = np.zeros(9), np.zeros(6), C.c_double(0.0)
        self.L.reft_compute_3b(self.h, np.ascontiguousarray(dx, np.float64), np.ascontiguousarray(dr, np.float64),
                               np.ascontiguousarray(t, np.int32), f, s, C.byref(e))
        return f, s, e.value


def read_xyz(path):
    """The reference driver's input format: natoms, 9 cell numbers, then
    'type x y z' lines (serial_interface/examples/cpp/main.cpp:84-110)."""
    with open(path) as f:
        lines = f.read().splitlines()
    n = int(lines[0].split()[0])
    cell = np.array([float(v) for v in lines[1].split()[:9]]).reshape(3, 3)
    types, pos = [], []
    for ln in lines[2:2 + n]:
        t = ln.split()
        types.append(t[0])
        pos.append([float(t[1]), float(t[2]), float(t[3])])
    return np.array(pos), types, cell


def sort_keys(keys):
    """Canonical (lexicographic) order of a cluster key array."""
    if len(keys) == 0:
        return keys
    order = np.lexsort(keys.T[::-1])
    return keys[order]
