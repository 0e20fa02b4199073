"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module, and only as the checker or as the timed
CPU baseline.  The product path never imports it.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run build())")
        L = ctypes.CDLL(path)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        llp = ctypes.POINTER(ctypes.c_longlong)
        L.oracle_eri_quartets.argtypes = [ctypes.c_int, ip, ip, ip, dp, dp, dp, ctypes.c_int, ip, llp,
                                          ctypes.c_int, ctypes.c_double, ctypes.c_int, dp]
        L.md_eri_quartets.argtypes = [ctypes.c_int, ip, ip, ip, dp, dp, dp, ctypes.c_int, ip, llp, dp]
        L.md_eri_deriv1_quartets.argtypes = [ctypes.c_int, ip, ip, ip, dp, dp, dp, ctypes.c_int, ip, llp,
                                             ctypes.c_double, dp]
        L.md_eri_deriv1_terms.argtypes = [ctypes.c_int, ip, ip, ip, dp, dp, dp, ctypes.c_int, ip, llp,
                                          ctypes.c_int, dp]
        L.oracle_fill_pair.argtypes = [ctypes.c_int, ip, ip, ip, dp, dp, dp, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_int, dp, dp]
        L.oracle_normalize.argtypes = [ctypes.c_int, ip, ip, ip, dp, dp]
        L.oracle_boys.argtypes = [ctypes.c_double, ctypes.c_int, dp]
        L.oracle_boys_ld.argtypes = [ctypes.c_int, ctypes.c_double]
        L.oracle_boys_ld.restype = ctypes.c_double
        L.oracle_norm_fac.argtypes = [ctypes.c_int]
        L.oracle_norm_fac.restype = ctypes.c_double
        L.oracle_boys_grid.argtypes = [dp]
        L.oracle_boys_longfac.argtypes = [dp]
        L.oracle_class_counts.argtypes = [ctypes.c_int] * 4 + [llp]
        L.md_set_reference_boys.argtypes = [ctypes.c_int]
        L.md_boys_value.argtypes = [ctypes.c_int, ctypes.c_double]
        L.md_boys_value.restype = ctypes.c_double
        L.oracle_fastschwarz.argtypes = [ctypes.c_int, ip, ip, ip, dp, dp, dp, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_int, ctypes.c_int]
        L.oracle_fastschwarz.restype = ctypes.c_double
        L.oracle_osoei.argtypes = [ctypes.c_int, ctypes.c_int, ip, ip, ip, dp, dp, dp, ctypes.c_int, ip, llp,
                                   ctypes.c_int, dp, dp, dp]
        L.md_oei_pairs.argtypes = L.oracle_osoei.argtypes
        _LIB = L
    return _LIB


def _p(a, t):
    return a.ctypes.data_as(ctypes.POINTER(t))


def ncart(l):
    return (l + 1) * (l + 2) // 2


class ShellArrays:
    """Flat shell description: am, nprim, prim offsets, centres, exponents, coefs."""

    def __init__(self, shells):
        # shells: list of (am, (x,y,z), alphas, coefs)
        self.am = np.array([s[0] for s in shells], dtype=np.int32)
        self.nprim = np.array([len(s[2]) for s in shells], dtype=np.int32)
        self.off = np.zeros(len(shells), dtype=np.int32)
        if len(shells) > 1:
            self.off[1:] = np.cumsum(self.nprim)[:-1]
        self.xyz = np.array([s[1] for s in shells], dtype=np.float64).reshape(-1, 3).copy()
        self.alpha = np.concatenate([np.asarray(s[2], dtype=np.float64) for s in shells]) if shells else np.zeros(0)
        self.coef = np.concatenate([np.asarray(s[3], dtype=np.float64) for s in shells]) if shells else np.zeros(0)

    def __len__(self):
        return len(self.am)

    def args(self):
        ip = ctypes.c_int
        dp = ctypes.c_double
        return (len(self.am), _p(self.am, ip), _p(self.nprim, ip), _p(self.off, ip), _p(self.xyz, dp),
                _p(self.alpha, dp), _p(self.coef, dp))


def normalize(sa):
    """simint_normalize_shells restated (in place on sa.coef)."""
    ip = ctypes.c_int
    dp = ctypes.c_double
    lib().oracle_normalize(len(sa.am), _p(sa.am, ip), _p(sa.nprim, ip), _p(sa.off, ip), _p(sa.alpha, dp),
                           _p(sa.coef, dp))


def quartet_sizes(sa, quartets, deriv=0):
    q = np.asarray(quartets, dtype=np.int32).reshape(-1, 4)
    am = sa.am[q]
    n = ((am + 1) * (am + 2) // 2).prod(axis=1) * (12 if deriv else 1)
    off = np.zeros(len(q), dtype=np.int64)
    if len(q) > 1:
        off[1:] = np.cumsum(n)[:-1]
    return q, n, off


def eri(sa, quartets, screen_method=0, screen_tol=0.0, deriv=0):
    """Contracted integrals per quartet, concatenated; returns (flat, offsets, sizes)."""
    q, n, off = quartet_sizes(sa, quartets, deriv)
    out = np.zeros(int(n.sum()), dtype=np.float64)
    lib().oracle_eri_quartets(*sa.args(), len(q), _p(q, ctypes.c_int), _p(off, ctypes.c_longlong),
                              screen_method, screen_tol, deriv, _p(out, ctypes.c_double))
    return out, off, n


def md_eri(sa, quartets, reference_boys=False):
    """Brute-force McMurchie-Davidson values in long double.  With
    reference_boys=True the Boys values are simint's split scheme (so only
    the recurrence algebra is checked)."""
    q, n, off = quartet_sizes(sa, quartets, 0)
    out = np.zeros(int(n.sum()), dtype=np.float64)
    lib().md_set_reference_boys(1 if reference_boys else 0)
    rc = lib().md_eri_quartets(*sa.args(), len(q), _p(q, ctypes.c_int), _p(off, ctypes.c_longlong),
                               _p(out, ctypes.c_double))
    lib().md_set_reference_boys(0)
    if rc != 0:
        raise ValueError("md_eri: angular momentum too high")
    return out, off, n


def md_eri_deriv1_analytic(sa, quartets, scale=False):
    """Long-double first derivatives from 2a (a+1) - a_x (a-1) on every
    centre, McMurchie-Davidson integrals (exact Boys values).  scale=True
    returns instead, per element, the summed magnitudes of the primitive
    terms the derivative is formed from (its conditioning scale)."""
    q, n, off = quartet_sizes(sa, quartets, 1)
    out = np.zeros(int(n.sum()), dtype=np.float64)
    rc = lib().md_eri_deriv1_terms(*sa.args(), len(q), _p(q, ctypes.c_int), _p(off, ctypes.c_longlong),
                                   1 if scale else 0, _p(out, ctypes.c_double))
    if rc != 0:
        raise ValueError("md_eri_deriv1_analytic: angular momentum too high")
    return out, off, n


def md_eri_deriv1(sa, quartets, h=1e-4):
    q, n, off = quartet_sizes(sa, quartets, 1)
    out = np.zeros(int(n.sum()), dtype=np.float64)
    lib().md_eri_deriv1_quartets(*sa.args(), len(q), _p(q, ctypes.c_int), _p(off, ctypes.c_longlong), h,
                                 _p(out, ctypes.c_double))
    return out, off, n


def fill_pair(sa, i, j, screen_method=0):
    n = int(sa.nprim[i] * sa.nprim[j])
    out = np.zeros(14 * max(n, 1), dtype=np.float64)
    ab = np.zeros(3)
    m = lib().oracle_fill_pair(*sa.args(), i, j, screen_method, _p(out, ctypes.c_double),
                               _p(ab, ctypes.c_double))
    names = ["x", "y", "z", "PA_x", "PA_y", "PA_z", "PB_x", "PB_y", "PB_z", "alpha", "prefac", "alpha2",
             "beta2", "screen"]
    d = {k: out[i * m:(i + 1) * m].copy() for i, k in enumerate(names)}
    d["AB"] = ab
    return d


def boys(x, n):
    F = np.zeros(n + 1)
    lib().oracle_boys(float(x), int(n), _p(F, ctypes.c_double))
    return F


def boys_grid():
    g = np.zeros((366, 32))
    lib().oracle_boys_grid(_p(g, ctypes.c_double))
    return g


def boys_longfac():
    g = np.zeros(32)
    lib().oracle_boys_longfac(_p(g, ctypes.c_double))
    return g


def class_counts(la, lb, lc, ld):
    c = (ctypes.c_longlong * 4)()
    lib().oracle_class_counts(la, lb, lc, ld, c)
    return list(c)


def fastschwarz(sa, ia, ib, i, j):
    """The fast-Schwarz value of primitive pair (i of shell ia, j of shell ib)
    (shell_screen.c:226-270): the s-type (ij|ij), not square-rooted."""
    return lib().oracle_fastschwarz(*sa.args(), ia, ib, i, j)


def _oei(fn, kind, sa, pairs, Z=None, cxyz=None):
    pr = np.asarray(pairs, dtype=np.int32).reshape(-1, 2)
    am = sa.am[pr]
    n = ((am + 1) * (am + 2) // 2).prod(axis=1)
    off = np.zeros(len(pr), dtype=np.int64)
    if len(pr) > 1:
        off[1:] = np.cumsum(n)[:-1]
    out = np.zeros(max(int(n.sum()), 1), dtype=np.float64)
    Zc = np.ascontiguousarray(np.zeros(1) if Z is None else Z, dtype=np.float64)
    C = np.ascontiguousarray(np.zeros((1, 3)) if cxyz is None else cxyz, dtype=np.float64).reshape(-1, 3)
    nc = 0 if Z is None else len(Zc)
    rc = fn(kind, *sa.args(), len(pr), _p(pr, ctypes.c_int), _p(off, ctypes.c_longlong), nc,
            _p(Zc, ctypes.c_double), _p(C, ctypes.c_double), _p(out, ctypes.c_double))
    if rc != 0:
        raise ValueError("one-electron integrals: angular momentum too high")
    return out[:int(n.sum())], off, n


def osoei(sa, pairs, kind, Z=None, cxyz=None):
    """One-electron integrals restated from skel/simint/osoei: kind 0
    overlap, 1 kinetic energy, 2 nuclear attraction of point charges Z at
    cxyz; ncart(a) x ncart(b) per shell pair, concatenated."""
    return _oei(lib().oracle_osoei, kind, sa, pairs, Z, cxyz)


def md_oei(sa, pairs, kind, Z=None, cxyz=None, reference_boys=False):
    """The same integrals in long double from the Hermite expansion (exact
    Boys values, or simint's split scheme with reference_boys=True),
    independent of the Obara-Saika restatement."""
    lib().md_set_reference_boys(1 if reference_boys else 0)
    try:
        return _oei(lib().md_oei_pairs, kind, sa, pairs, Z, cxyz)
    finally:
        lib().md_set_reference_boys(0)


def eri_tensor(sa, exact=False):
    """The full four-index tensor (mu nu|lambda sigma) over the Cartesian
    functions of every shell (Psi4 order within a shell, shells in input
    order), from every ordered shell quartet -- no permutational symmetry
    used.  exact=True: the long-double McMurchie-Davidson values."""
    ns = len(sa.am)
    nf = [ncart(int(l)) for l in sa.am]
    f0 = np.zeros(ns + 1, dtype=np.int64)
    f0[1:] = np.cumsum(nf)
    nbf = int(f0[-1])
    quartets = [(a, b, c, d) for a in range(ns) for b in range(ns) for c in range(ns) for d in range(ns)]
    vals, off, n = (md_eri(sa, quartets, reference_boys=True) if exact else eri(sa, quartets))
    g = np.zeros((nbf,) * 4)
    for (a, b, c, d), o, m in zip(quartets, off, n):
        g[f0[a]:f0[a + 1], f0[b]:f0[b + 1], f0[c]:f0[c + 1], f0[d]:f0[d + 1]] = \
            vals[o:o + m].reshape(nf[a], nf[b], nf[c], nf[d])
    return g, f0


def jk(g, D):
    """Coulomb and exchange matrices of a density D from the full tensor:
    J_mn = sum_ls (mn|ls) D_ls, K_mn = sum_ls (ml|ns) D_ls (the digestion a
    Fock build performs; the reference has no such consumer)."""
    J = np.einsum("mnls,ls->mn", g, D)
    K = np.einsum("mlns,ls->mn", g, D)
    return J, K
