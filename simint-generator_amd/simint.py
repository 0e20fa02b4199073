"""ctypes mirror of the simint C API exported by libsimint.so.

Python-side helpers for tests, the benchmark and the autotuner.  Struct
layouts match include/simint/shell/shell.h (the reference's
skel/simint/shell/shell.h:18-85 with alpha2/beta2 present).  Loading fails
loudly if the library has not been built: there is no fallback path.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsimint.so")

c_int, c_double, c_size_t, c_void_p = ctypes.c_int, ctypes.c_double, ctypes.c_size_t, ctypes.c_void_p
dptr = ctypes.POINTER(ctypes.c_double)
iptr = ctypes.POINTER(ctypes.c_int)
llptr = ctypes.POINTER(ctypes.c_longlong)

SCREEN_NONE, SCREEN_SCHWARZ, SCREEN_FASTSCHWARZ = 0, 1, 2
ERROR_NO_DEVICE, ERROR_UNSUPPORTED, ERROR_HIP, ERROR_ARGUMENT = -2, -3, -4, -5


class Shell(ctypes.Structure):
    _fields_ = [("am", c_int), ("nprim", c_int), ("x", c_double), ("y", c_double), ("z", c_double),
                ("alpha", dptr), ("coef", dptr), ("memsize", c_size_t), ("ptr", c_void_p)]


class MultiShellpair(ctypes.Structure):
    _fields_ = [("am1", c_int), ("am2", c_int), ("nprim", c_int), ("nshell12", c_int), ("nshell12_clip", c_int),
                ("nprim12", iptr),
                ("AB_x", dptr), ("AB_y", dptr), ("AB_z", dptr),
                ("x", dptr), ("y", dptr), ("z", dptr),
                ("PA_x", dptr), ("PA_y", dptr), ("PA_z", dptr),
                ("PB_x", dptr), ("PB_y", dptr), ("PB_z", dptr),
                ("alpha", dptr), ("alpha2", dptr), ("beta2", dptr),
                ("prefac", dptr), ("screen", dptr), ("screen_max", c_double),
                ("memsize", c_size_t), ("ptr", c_void_p)]


# every exported symbol of include/simint/*.h
EXPORTS = [
    "simint_init", "simint_finalize",
    "simint_compute_eri", "simint_compute_eri_deriv", "simint_eri_worksize", "simint_eri_workmem",
    "simint_compute_ostei", "simint_compute_ostei_deriv",
    "simint_initialize_shell", "simint_initialize_shells", "simint_allocate_shell", "simint_free_shell",
    "simint_free_shells", "simint_copy_shell", "simint_normalize_shells", "simint_create_shell",
    "simint_create_zero_shell", "simint_initialize_multi_shellpair", "simint_initialize_multi_shellpairs",
    "simint_allocate_multi_shellpair", "simint_allocate_multi_shellpair2", "simint_free_multi_shellpair",
    "simint_free_multi_shellpairs", "simint_fill_multi_shellpair", "simint_fill_multi_shellpair2",
    "simint_create_multi_shellpair", "simint_create_multi_shellpair2", "simint_cat_multi_shellpair",
    "simint_cat_shellpairs",
    "simint_shellscreen", "simint_primscreen", "simint_shellscreen_schwarz", "simint_shellscreen_fastschwarz",
    "simint_primscreen_schwarz", "simint_primscreen_fastschwarz",
    "simint_hip_pairs_create", "simint_hip_pairs_free", "simint_hip_pairs_count",
    "simint_hip_compute_eri_prefix", "simint_hip_compute_eri_diag", "simint_hip_synchronize",
    "simint_hip_stream", "simint_hip_last_error", "simint_hip_class_kernel", "simint_hip_set_class_kernel",
    "simint_hip_compute_eri_list", "simint_hip_compute_eri_range", "simint_hip_deriv1_worksize",
    "simint_hip_compute_eri_deriv1_range", "simint_hip_set_stream", "simint_hip_class_list_kernel",
    "simint_hip_set_list_split",
    "simint_hip_set_class_list_kernel", "simint_hip_compute_eri_deriv1_list",
    "simint_hip_capture_begin", "simint_hip_capture_end", "simint_hip_graph_launch", "simint_hip_graph_free",
    "simint_compute_overlap", "simint_compute_ke", "simint_compute_potential",
    "simint_compute_osoei_overlap", "simint_compute_osoei_ke", "simint_compute_osoei_potential",
    "simint_hip_shells_create", "simint_hip_shells_free", "simint_hip_compute_osoei",
    "simint_hip_compute_eri_deriv1_list_noalloc", "simint_hip_pairs_bytes",
    "simint_hip_set_ordering", "simint_hip_ordering",
    "simint_hip_pairs_from_shells", "simint_hip_pair_bounds", "simint_hip_order_pairs", "simint_hip_screen_counts",
    "simint_hip_pairs_export", "simint_hip_compute_jk_range", "simint_hip_jk_symmetrize", "simint_hip_set_jk_rows",
]
ORDERING_PSI4, ORDERING_GAMESS = 0, 1

_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libsimint.so is not built; run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        SP, MP = ctypes.POINTER(Shell), ctypes.POINTER(MultiShellpair)
        L.simint_create_shell.argtypes = [c_int, c_int, c_double, c_double, c_double, dptr, dptr, SP]
        L.simint_normalize_shells.argtypes = [c_int, SP]
        L.simint_free_shell.argtypes = [SP]
        L.simint_initialize_shell.argtypes = [SP]
        L.simint_copy_shell.argtypes = [SP, SP]
        L.simint_initialize_multi_shellpair.argtypes = [MP]
        L.simint_free_multi_shellpair.argtypes = [MP]
        L.simint_create_multi_shellpair.argtypes = [c_int, SP, c_int, SP, MP, c_int]
        L.simint_create_multi_shellpair2.argtypes = [c_int, SP, MP, c_int]
        L.simint_cat_shellpairs.argtypes = [c_int, ctypes.POINTER(MP), MP, c_int]
        L.simint_cat_multi_shellpair.argtypes = [c_int, ctypes.POINTER(MP), MP, c_int]
        L.simint_compute_eri.argtypes = [MP, MP, c_double, dptr, dptr]
        L.simint_compute_eri_deriv.argtypes = [c_int, MP, MP, c_double, dptr, dptr]
        L.simint_eri_worksize.argtypes = [c_int, c_int]
        L.simint_eri_worksize.restype = c_size_t
        L.simint_primscreen.argtypes = [SP, SP, dptr, c_int]
        L.simint_primscreen.restype = c_double
        L.simint_shellscreen.argtypes = [SP, SP, c_int]
        L.simint_shellscreen.restype = c_double
        L.simint_hip_pairs_create.argtypes = [MP, ctypes.POINTER(c_void_p)]
        L.simint_hip_pairs_free.argtypes = [c_void_p]
        L.simint_hip_pairs_count.argtypes = [c_void_p]
        L.simint_hip_pairs_bytes.argtypes = [c_void_p]
        L.simint_hip_pairs_bytes.restype = c_size_t
        L.simint_hip_compute_eri_prefix.argtypes = [c_void_p, c_void_p, c_int, c_int, iptr, llptr, c_double, c_void_p]
        L.simint_hip_compute_eri_prefix.restype = ctypes.c_longlong
        L.simint_hip_compute_eri_diag.argtypes = [c_void_p, c_void_p]
        L.simint_hip_compute_eri_diag.restype = ctypes.c_longlong
        ll = ctypes.c_longlong
        L.simint_hip_compute_eri_list.argtypes = [c_void_p, c_void_p, ll, c_void_p, c_void_p, c_void_p, c_double,
                                                  c_void_p]
        L.simint_hip_compute_eri_list.restype = ll
        L.simint_hip_compute_eri_range.argtypes = [c_void_p, c_void_p, c_int, c_int, iptr, iptr, llptr, c_double,
                                                   c_int, c_void_p]
        L.simint_hip_compute_eri_range.restype = ll
        L.simint_hip_compute_eri_deriv1_range.argtypes = [c_void_p, c_void_p, c_int, c_int, iptr, iptr, c_double,
                                                          c_void_p, c_void_p]
        L.simint_hip_compute_eri_deriv1_range.restype = ll
        L.simint_hip_deriv1_worksize.argtypes = [c_void_p, c_void_p, ll]
        L.simint_hip_deriv1_worksize.restype = ll
        L.simint_hip_compute_eri_deriv1_list.argtypes = [c_void_p, c_void_p, ll, c_void_p, c_void_p, c_void_p,
                                                         c_double, c_void_p, c_void_p]
        L.simint_hip_compute_eri_deriv1_list.restype = ll
        L.simint_hip_capture_end.argtypes = [ctypes.POINTER(c_void_p)]
        L.simint_hip_graph_launch.argtypes = [c_void_p]
        L.simint_hip_graph_free.argtypes = [c_void_p]
        L.simint_hip_stream.restype = c_void_p
        L.simint_hip_set_stream.argtypes = [c_void_p]
        L.simint_hip_set_list_split.argtypes = [c_int]
        L.simint_hip_set_ordering.argtypes = [c_int]
        L.simint_hip_last_error.restype = ctypes.c_char_p
        L.simint_hip_class_kernel.argtypes = [c_int] * 4
        L.simint_hip_set_class_kernel.argtypes = [c_int] * 5
        L.simint_hip_class_list_kernel.argtypes = [c_int] * 4
        L.simint_hip_set_class_list_kernel.argtypes = [c_int] * 5
        for f in ("simint_compute_overlap", "simint_compute_ke", "simint_compute_osoei_overlap",
                  "simint_compute_osoei_ke"):
            getattr(L, f).argtypes = [SP, SP, dptr]
        for f in ("simint_compute_potential", "simint_compute_osoei_potential"):
            getattr(L, f).argtypes = [c_int, dptr, dptr, dptr, dptr, SP, SP, dptr]
        L.simint_hip_shells_create.argtypes = [c_int, SP, ctypes.POINTER(c_void_p)]
        L.simint_hip_pairs_from_shells.argtypes = [c_void_p, c_int, iptr, c_int, ctypes.POINTER(c_void_p)]
        L.simint_hip_pair_bounds.argtypes = [c_void_p, c_void_p]
        L.simint_hip_pair_bounds.restype = ll
        L.simint_hip_order_pairs.argtypes = [c_void_p, c_void_p, c_void_p]
        L.simint_hip_screen_counts.argtypes = [c_int, c_void_p, c_int, c_void_p, c_int, c_double, c_void_p]
        L.simint_hip_pairs_export.argtypes = [c_void_p, dptr, iptr, dptr]
        L.simint_hip_pairs_export.restype = ll
        L.simint_hip_compute_jk_range.argtypes = [c_void_p, c_void_p, c_int, c_int, iptr, iptr, llptr, c_double,
                                                  c_int, c_void_p, ctypes.POINTER(JK)]
        L.simint_hip_compute_jk_range.restype = ll
        L.simint_hip_jk_symmetrize.argtypes = [c_int, c_void_p, c_void_p]
        L.simint_hip_shells_free.argtypes = [c_void_p]
        L.simint_hip_compute_osoei.argtypes = [c_int, c_void_p, ll, c_void_p, c_void_p, c_int, c_void_p, c_void_p]
        L.simint_hip_compute_osoei.restype = ll
        _LIB = L
    return _LIB


def last_error():
    return lib().simint_hip_last_error().decode()


def ncart(l):
    return (l + 1) * (l + 2) // 2


class Basis:
    """A list of simint_shell structures created through the C API.

    shells: iterable of (am, (x, y, z), alphas, coefs).  Shells are
    normalised with simint_normalize_shells when normalize=True."""

    def __init__(self, shells, normalize=True):
        shells = list(shells)
        self.n = len(shells)
        self.arr = (Shell * max(self.n, 1))()
        L = lib()
        for i, (am, xyz, al, co) in enumerate(shells):
            a = np.ascontiguousarray(al, dtype=np.float64)
            c = np.ascontiguousarray(co, dtype=np.float64)
            L.simint_initialize_shell(ctypes.byref(self.arr[i]))
            L.simint_create_shell(len(a), int(am), float(xyz[0]), float(xyz[1]), float(xyz[2]),
                                  a.ctypes.data_as(dptr), c.ctypes.data_as(dptr), ctypes.byref(self.arr[i]))
        if normalize and self.n:
            L.simint_normalize_shells(self.n, self.arr)

    def am(self, i):
        return self.arr[i].am

    def subset(self, idx):
        """ctypes array of copies of the structs (same storage, so the same
        shell object for compare_shell)."""
        out = (Shell * max(len(idx), 1))()
        for k, i in enumerate(idx):
            out[k] = self.arr[i]
        return out

    def free(self):
        L = lib()
        for i in range(self.n):
            L.simint_free_shell(ctypes.byref(self.arr[i]))
        self.n = 0

    def coefs(self, i):
        s = self.arr[i]
        return np.ctypeslib.as_array(s.coef, shape=(s.nprim,)).copy()


class Pairs:
    """A simint_multi_shellpair built with simint_create_multi_shellpair."""

    def __init__(self, basis, a_idx, b_idx, screen=SCREEN_NONE):
        self.mp = MultiShellpair()
        L = lib()
        L.simint_initialize_multi_shellpair(ctypes.byref(self.mp))
        A = basis.subset(a_idx)
        B = basis.subset(b_idx)
        L.simint_create_multi_shellpair(len(a_idx), A, len(b_idx), B, ctypes.byref(self.mp), screen)
        self.n = len(a_idx) * len(b_idx)

    @classmethod
    def from_pairs(cls, basis, pairs, screen=SCREEN_NONE):
        """simint_create_multi_shellpair2 on an explicit list of (i, j)."""
        self = cls.__new__(cls)
        self.mp = MultiShellpair()
        L = lib()
        L.simint_initialize_multi_shellpair(ctypes.byref(self.mp))
        AB = (Shell * max(2 * len(pairs), 1))()
        for k, (i, j) in enumerate(pairs):
            AB[2 * k] = basis.arr[i]
            AB[2 * k + 1] = basis.arr[j]
        L.simint_create_multi_shellpair2(len(pairs), AB, ctypes.byref(self.mp), screen)
        self.n = len(pairs)
        return self

    def field(self, name):
        n = self.mp.nprim if name not in ("AB_x", "AB_y", "AB_z", "nprim12") else self.mp.nshell12
        p = getattr(self.mp, name)
        if not p:
            return None
        return np.ctypeslib.as_array(p, shape=(n,)).copy()

    def free(self):
        lib().simint_free_multi_shellpair(ctypes.byref(self.mp))


def compute_eri(P, Q, screen_tol=0.0, deriv=0):
    """Drop-in call; returns (rc, integrals as a flat numpy array)."""
    la, lb, lc, ld = P.mp.am1, P.mp.am2, Q.mp.am1, Q.mp.am2
    n = P.mp.nshell12_clip * Q.mp.nshell12_clip * ncart(la) * ncart(lb) * ncart(lc) * ncart(ld)
    n *= 12 if deriv else 1
    out = np.zeros(max(n, 1), dtype=np.float64)
    if deriv:
        rc = lib().simint_compute_eri_deriv(deriv, ctypes.byref(P.mp), ctypes.byref(Q.mp), screen_tol, None,
                                            out.ctypes.data_as(dptr))
    else:
        rc = lib().simint_compute_eri(ctypes.byref(P.mp), ctypes.byref(Q.mp), screen_tol, None,
                                      out.ctypes.data_as(dptr))
    return rc, out[:n]


class DevicePairs:
    """simint_hip_pairs: shell pairs resident in HBM."""

    def __init__(self, pairs):
        h = c_void_p()
        rc = lib().simint_hip_pairs_create(ctypes.byref(pairs.mp), ctypes.byref(h))
        if rc != 0:
            raise RuntimeError("simint_hip_pairs_create failed (%d): %s" % (rc, last_error()))
        self.h = h
        self.am = (pairs.mp.am1, pairs.mp.am2)
        self.n = pairs.mp.nshell12_clip
        self.bytes = int(lib().simint_hip_pairs_bytes(h))

    @classmethod
    def from_shells(cls, dshells, pairs, am, screen=SCREEN_NONE):
        """simint_hip_pairs_from_shells: the pairs (i, j) of a DeviceShells
        set, filled on the device (all of class `am`)."""
        pr = np.ascontiguousarray(np.asarray(pairs, dtype=np.int32).reshape(-1, 2))
        h = c_void_p()
        rc = lib().simint_hip_pairs_from_shells(dshells.h, len(pr), pr.ctypes.data_as(iptr), int(screen),
                                                ctypes.byref(h))
        if rc != 0:
            raise RuntimeError("simint_hip_pairs_from_shells failed (%d): %s" % (rc, last_error()))
        self = cls.__new__(cls)
        self.h = h
        self.am = tuple(am)
        self.n = len(pr)
        self.bytes = int(lib().simint_hip_pairs_bytes(h))
        return self

    def export(self):
        """The linear view copied back: dict of the 14 per-primitive fields
        (x .. beta2, screen), nprim12 and AB."""
        L = lib()
        m = L.simint_hip_pairs_export(self.h, None, None, None)
        if m < 0:
            raise RuntimeError("simint_hip_pairs_export failed (%d): %s" % (m, last_error()))
        f = np.zeros(14 * max(m, 1))
        npr = np.zeros(max(self.n, 1), dtype=np.int32)
        ab = np.zeros(3 * max(self.n, 1))
        r = L.simint_hip_pairs_export(self.h, f.ctypes.data_as(dptr), npr.ctypes.data_as(iptr), ab.ctypes.data_as(dptr))
        if r < 0:
            raise RuntimeError("simint_hip_pairs_export failed (%d): %s" % (r, last_error()))
        names = ["x", "y", "z", "PA_x", "PA_y", "PA_z", "PB_x", "PB_y", "PB_z", "alpha", "prefac", "alpha2",
                 "beta2", "screen"]
        d = {k: f[i * m:(i + 1) * m] for i, k in enumerate(names)}
        d["nprim12"] = npr[:self.n]
        d["AB"] = ab[:3 * self.n].reshape(3, self.n)
        return d

    def bounds(self, d_bound):
        r = lib().simint_hip_pair_bounds(self.h, c_void_p(d_bound))
        if r < 0:
            raise RuntimeError("simint_hip_pair_bounds failed (%d): %s" % (r, last_error()))

    def order(self, d_bound, d_order):
        r = lib().simint_hip_order_pairs(self.h, c_void_p(d_bound), c_void_p(d_order))
        if r < 0:
            raise RuntimeError("simint_hip_order_pairs failed (%d): %s" % (r, last_error()))

    def free(self):
        if self.h:
            lib().simint_hip_pairs_free(self.h)
            self.h = None


FAMILIES = {"lpq": 1, "wpq": 2, "lpqs_a": 3, "lpqs_b": 4, "lpqsf_a": 5, "lpqsf_b": 6}


_LIST_SPLIT = None   # class name -> lanes per quartet of its list launches


def list_split_of(cls):
    """Lanes per quartet a batch gives the lane-per-quartet list launches of
    class cls: the tuned value ("list_split" of the package's tuning.json,
    or of the last file passed to apply_tuning), 0 = the library's
    automatic choice (about 65536 lanes per launch)."""
    global _LIST_SPLIT
    if _LIST_SPLIT is None:
        import json
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.json")
        _LIST_SPLIT = dict(json.load(open(path)).get("list_split", {})) if os.path.exists(path) else {}
    return int(_LIST_SPLIT.get("_".join("spdfghik"[x] for x in cls), 0))


def apply_tuning(path):
    """Set each class's kernel family (and the list launches' lanes per
    quartet) from a tuning file (as autotune.py writes it) at run time;
    returns the number of classes changed."""
    import json
    global _LIST_SPLIT
    n = 0
    d = json.load(open(path))
    for name, fam in d.get("family", {}).items():
        am = ["spdfghik".index(c) for c in name.split("_")]
        if lib().simint_hip_set_class_kernel(*am, FAMILIES[fam]) >= 0:
            n += 1
    for name, fam in d.get("list_family", {}).items():
        am = ["spdfghik".index(c) for c in name.split("_")]
        if lib().simint_hip_set_class_list_kernel(*am, FAMILIES[fam]) >= 0:
            n += 1
    if "list_split" in d:
        for name, lanes in d["list_split"].items():
            if lanes not in (0, 1, 2, 4, 8, 16):
                raise ValueError("list_split %s: %r lanes (0, 1, 2, 4, 8 or 16)" % (name, lanes))
        _LIST_SPLIT = dict(d["list_split"])
        n += len(_LIST_SPLIT)
    return n


def compute_list(P, Q, n, d_qbra, d_qket, d_qout, d_out, screen_tol=0.0):
    """simint_hip_compute_eri_list on device arrays (pointers as ints;
    d_qout may be 0 for contiguous output)."""
    r = lib().simint_hip_compute_eri_list(P.h, Q.h, int(n), c_void_p(d_qbra), c_void_p(d_qket),
                                          c_void_p(d_qout or None), float(screen_tol), c_void_p(d_out))
    if r < 0:
        raise RuntimeError("simint_hip_compute_eri_list failed (%d): %s" % (r, last_error()))
    return r


def deriv1_worksize(P, Q, n):
    return lib().simint_hip_deriv1_worksize(P.h, Q.h, int(n))


def compute_deriv1_list(P, Q, n, d_qbra, d_qket, d_qout, d_work, d_out, screen_tol=0.0):
    r = lib().simint_hip_compute_eri_deriv1_list(P.h, Q.h, int(n), c_void_p(d_qbra), c_void_p(d_qket),
                                                 c_void_p(d_qout or None), float(screen_tol), c_void_p(d_work),
                                                 c_void_p(d_out))
    if r < 0:
        raise RuntimeError("simint_hip_compute_eri_deriv1_list failed (%d): %s" % (r, last_error()))
    return r


class Graph:
    """A batch of list calls recorded on the library stream and replayed as
    one graph launch: with Graph.capture() as g: ...calls...; g.launch()."""

    def __init__(self):
        self.h = None

    @classmethod
    def capture(cls):
        return _Capture(cls())

    def launch(self):
        r = lib().simint_hip_graph_launch(self.h)
        if r != 0:
            raise RuntimeError("simint_hip_graph_launch failed (%d): %s" % (r, last_error()))

    def free(self):
        if self.h:
            lib().simint_hip_graph_free(self.h)
            self.h = None


class _Capture:
    def __init__(self, g):
        self.g = g

    def __enter__(self):
        r = lib().simint_hip_capture_begin()
        if r != 0:
            raise RuntimeError("simint_hip_capture_begin failed (%d): %s" % (r, last_error()))
        return self.g

    def __exit__(self, et, ev, tb):
        h = c_void_p()
        r = lib().simint_hip_capture_end(ctypes.byref(h))
        if r != 0 and et is None:
            raise RuntimeError("simint_hip_capture_end failed (%d): %s" % (r, last_error()))
        self.g.h = h
        return False


def compute_range(P, Q, b0, ket_lo, ket_hi, out_off, d_out_ptr, screen_tol=0.0, precision=0):
    """simint_hip_compute_eri_range: kets [ket_lo[i], ket_hi[i]) of bra b0+i
    (ket_lo None: from 0); precision 1 = FP32 VRR/HRR where available."""
    ket_hi = np.ascontiguousarray(ket_hi, dtype=np.int32)
    lo = None if ket_lo is None else np.ascontiguousarray(ket_lo, dtype=np.int32)
    out_off = np.ascontiguousarray(out_off, dtype=np.int64)
    r = lib().simint_hip_compute_eri_range(P.h, Q.h, int(b0), len(ket_hi),
                                           None if lo is None else lo.ctypes.data_as(iptr),
                                           ket_hi.ctypes.data_as(iptr), out_off.ctypes.data_as(llptr),
                                           float(screen_tol), int(precision), c_void_p(d_out_ptr))
    if r < 0:
        raise RuntimeError("simint_hip_compute_eri_range failed (%d): %s" % (r, last_error()))
    return r


class JK(ctypes.Structure):
    """simint_hip_jk: the J/K digestion's device arrays."""
    _fields_ = [("nbf", c_int), ("same", c_int), ("bra_fun", c_void_p), ("ket_fun", c_void_p),
                ("density", c_void_p), ("J", c_void_p), ("K", c_void_p)]


def compute_jk_range(P, Q, b0, ket_lo, ket_hi, out_off, d_out_ptr, jk, screen_tol=0.0, precision=0):
    """simint_hip_compute_jk_range: compute_range, then the J/K digestion of
    the same quartets with the density and accumulators of `jk` (a JK)."""
    ket_hi = np.ascontiguousarray(ket_hi, dtype=np.int32)
    lo = None if ket_lo is None else np.ascontiguousarray(ket_lo, dtype=np.int32)
    out_off = np.ascontiguousarray(out_off, dtype=np.int64)
    r = lib().simint_hip_compute_jk_range(P.h, Q.h, int(b0), len(ket_hi),
                                          None if lo is None else lo.ctypes.data_as(iptr),
                                          ket_hi.ctypes.data_as(iptr), out_off.ctypes.data_as(llptr),
                                          float(screen_tol), int(precision), c_void_p(d_out_ptr), ctypes.byref(jk))
    if r < 0:
        raise RuntimeError("simint_hip_compute_jk_range failed (%d): %s" % (r, last_error()))
    return r


def jk_symmetrize(nbf, d_J, d_K):
    r = lib().simint_hip_jk_symmetrize(int(nbf), c_void_p(d_J or None), c_void_p(d_K or None))
    if r < 0:
        raise RuntimeError("simint_hip_jk_symmetrize failed (%d): %s" % (r, last_error()))


def compute_deriv1_range(P, Q, b0, ket_lo, ket_hi, d_work, d_out_ptr, screen_tol=0.0):
    """simint_hip_compute_eri_deriv1_range: output dense in (bra, ket) order."""
    ket_hi = np.ascontiguousarray(ket_hi, dtype=np.int32)
    lo = None if ket_lo is None else np.ascontiguousarray(ket_lo, dtype=np.int32)
    r = lib().simint_hip_compute_eri_deriv1_range(P.h, Q.h, int(b0), len(ket_hi),
                                                  None if lo is None else lo.ctypes.data_as(iptr),
                                                  ket_hi.ctypes.data_as(iptr), float(screen_tol),
                                                  c_void_p(d_work), c_void_p(d_out_ptr))
    if r < 0:
        raise RuntimeError("simint_hip_compute_eri_deriv1_range failed (%d): %s" % (r, last_error()))
    return r


def compute_prefix(P, Q, b0, nket, out_off, d_out_ptr, screen_tol=0.0):
    nket = np.ascontiguousarray(nket, dtype=np.int32)
    out_off = np.ascontiguousarray(out_off, dtype=np.int64)
    r = lib().simint_hip_compute_eri_prefix(P.h, Q.h, int(b0), len(nket), nket.ctypes.data_as(iptr),
                                            out_off.ctypes.data_as(llptr), float(screen_tol), c_void_p(d_out_ptr))
    if r < 0:
        raise RuntimeError("simint_hip_compute_eri_prefix failed (%d): %s" % (r, last_error()))
    return r


def compute_diag(P, d_out_ptr):
    r = lib().simint_hip_compute_eri_diag(P.h, c_void_p(d_out_ptr))
    if r < 0:
        raise RuntimeError("simint_hip_compute_eri_diag failed (%d): %s" % (r, last_error()))
    return r


OEI_KINDS = {"overlap": 0, "ke": 1, "potential": 2}


def compute_oei(kind, basis, i, j, charges=None):
    """Drop-in one-electron call on shells i and j of `basis`:
    simint_compute_overlap / _ke / _potential; charges = (Z, xyz) arrays for
    the potential.  Returns (rc, ncart(a) x ncart(b) values, flat)."""
    n = ncart(basis.arr[i].am) * ncart(basis.arr[j].am)
    out = np.zeros(max(n, 1), dtype=np.float64)
    A, B = ctypes.byref(basis.arr[i]), ctypes.byref(basis.arr[j])
    if kind == "potential":
        Z = np.ascontiguousarray(charges[0], dtype=np.float64)
        C = np.ascontiguousarray(charges[1], dtype=np.float64).reshape(-1, 3)
        xs, ys, zs = (np.ascontiguousarray(C[:, k]) for k in range(3))
        rc = lib().simint_compute_potential(len(Z), Z.ctypes.data_as(dptr), xs.ctypes.data_as(dptr),
                                            ys.ctypes.data_as(dptr), zs.ctypes.data_as(dptr), A, B,
                                            out.ctypes.data_as(dptr))
    else:
        fn = lib().simint_compute_overlap if kind == "overlap" else lib().simint_compute_ke
        rc = fn(A, B, out.ctypes.data_as(dptr))
    return rc, out[:n]


class DeviceShells:
    """simint_hip_shells: a shell set resident in HBM for batched
    one-electron integrals."""

    def __init__(self, basis):
        h = c_void_p()
        rc = lib().simint_hip_shells_create(basis.n, basis.arr, ctypes.byref(h))
        if rc != 0:
            raise RuntimeError("simint_hip_shells_create failed (%d): %s" % (rc, last_error()))
        self.h = h

    def compute(self, kind, npair, d_pairs, d_off, d_out, ncenter=0, d_charges=0):
        r = lib().simint_hip_compute_osoei(OEI_KINDS[kind], self.h, int(npair), c_void_p(d_pairs),
                                           c_void_p(d_off), int(ncenter), c_void_p(d_charges or None),
                                           c_void_p(d_out))
        if r < 0:
            raise RuntimeError("simint_hip_compute_osoei failed (%d): %s" % (r, last_error()))
        return r

    def free(self):
        if self.h:
            lib().simint_hip_shells_free(self.h)
            self.h = None


def screen_counts(nbra, d_bra_bound, nket, d_ket_bound, same, tol, d_counts):
    r = lib().simint_hip_screen_counts(int(nbra), c_void_p(d_bra_bound), int(nket), c_void_p(d_ket_bound),
                                       1 if same else 0, float(tol), c_void_p(d_counts))
    if r < 0:
        raise RuntimeError("simint_hip_screen_counts failed (%d): %s" % (r, last_error()))


def synchronize():
    rc = lib().simint_hip_synchronize()
    if rc != 0:
        raise RuntimeError("simint_hip_synchronize failed (%d): %s" % (rc, last_error()))
