"""Pins the CPU oracle (oracle/oracle.c) -- CPU only.

What fixes each part:
  * Boys grid / long-range factors: mpmath (gammainc), closed forms, and the
    reference's own tables compiled from its sources (oracle/_ref) where built;
  * Boys split scheme: mpmath, with the asymptotic-branch error bound the
    reference's scheme carries near its cutoff;
  * (ss|ss): closed form 2 pi^(5/2) / (p q sqrt(p+q)) K_AB K_CD F_0(T);
  * every class up to f: an independent long-double McMurchie-Davidson brute
    force (oracle/md_ref.c), fed the reference's Boys values so that the
    recurrence algebra is compared exactly, and fed exact Boys values with
    the Boys-scheme error allowance;
  * invariants: bra/ket and centre permutations, translation, Schwarz bound;
  * first derivatives: Richardson-extrapolated finite differences of the
    long-double brute force;
  * normalisation: a normalised shell's x^l component has unit self-overlap
    (long-double Hermite overlap);
  * fast-Schwarz values: the exact (ij|ij) of the s-type primitive pair;
  * one-electron integrals (overlap, kinetic energy, nuclear attraction):
    long-double Hermite (McMurchie-Davidson) expansions, closed forms for s.
"""
import ctypes
import itertools
import math
import os

import numpy as np
import pytest

import oracle
import workloads
from helpers import mol_shells, normalized_arrays, per_quartet_errors

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_TABLES = os.path.join(ROOT, "oracle", "_ref", "libsimint_ref_tables.so")
PRODUCT_GRID = os.path.join(ROOT, "simint-generator_amd", "csrc", "gen", "boys_grid.inc")


def _mp_boys(n, x):
    from mpmath import mp
    mp.dps = 40
    x = mp.mpf(x)
    if x == 0:
        return mp.mpf(1) / (2 * n + 1)
    N = n + mp.mpf("0.5")
    return mp.gammainc(N, 0, x) / (2 * mp.power(x, N))


def test_boys_grid_against_mpmath():
    g = oracle.boys_grid()
    assert np.all(g[0] == 1.0 / (2 * np.arange(32) + 1))
    from mpmath import mp
    for i in (1, 7, 50, 123, 250, 365):
        for m in (0, 1, 5, 13, 31):
            v = float(_mp_boys(m, mp.mpf(i) / 10))     # grid point x_i = i/10 exactly
            assert abs(g[i, m] - v) <= 1.2e-16 * abs(v), (i, m)


def test_boys_longfac_closed_form():
    lf = oracle.boys_longfac()
    assert lf[0] == pytest.approx(math.sqrt(math.pi) / 2, rel=1e-16)
    for n in range(1, 32):
        assert lf[n] == pytest.approx(lf[n - 1] * (2 * n - 1) / 2, rel=4e-16)


@pytest.mark.skipif(not os.path.exists(REF_TABLES), reason="reference tables not built here")
def test_tables_match_reference_sources():
    L = ctypes.CDLL(REF_TABLES)
    ref = np.ctypeslib.as_array((ctypes.c_double * (366 * 32)).in_dll(L, "boys_shortgrid")).reshape(366, 32)
    g = oracle.boys_grid()
    ulp = np.spacing(np.abs(ref))
    assert np.all(np.abs(g - ref) <= ulp), np.abs(g - ref).max()
    L.simint_ref_boys_longfac.restype = ctypes.POINTER(ctypes.c_double)
    lf = np.ctypeslib.as_array(L.simint_ref_boys_longfac(), shape=(32,))
    assert np.all(np.abs(oracle.boys_longfac() - lf) <= np.spacing(lf))
    nf = np.ctypeslib.as_array((ctypes.c_double * 21).in_dll(L, "norm_fac"))
    for l in range(21):
        assert abs(oracle.lib().oracle_norm_fac(l) - nf[l]) <= np.spacing(nf[l])


@pytest.mark.skipif(not os.path.exists(PRODUCT_GRID), reason="generated sources not present")
def test_product_grid_matches_oracle_grid():
    import re
    text = re.sub(r"/\*.*?\*/", "", open(PRODUCT_GRID).read(), flags=re.S)
    vals = [float.fromhex(t) for t in text.replace("\n", " ").split(",") if t.strip()]
    g = np.array(vals).reshape(366, 32)
    o = oracle.boys_grid()
    assert np.all(np.abs(g - o) <= np.spacing(np.abs(o)))


@pytest.mark.parametrize("x", [0.0, 1e-9, 0.04999, 0.05, 3.3333, 17.2, 30.0, 36.49, 36.5, 40.0, 120.0, 5e3])
def test_boys_split_scheme(x):
    for L in (0, 2, 3, 4, 8, 12):
        F = oracle.boys(x, L)
        for m in range(L + 1):
            exact = float(_mp_boys(m, x))
            if x < 36.5:
                tol = 2e-14
            else:
                # the asymptotic form drops Gamma(m+1/2, x)/Gamma(m+1/2)
                from mpmath import mp
                mp.dps = 40
                tol = float(mp.gammainc(m + mp.mpf("0.5"), x) / mp.gamma(m + mp.mpf("0.5"))) * 4 + 2e-14
                if L >= 4:
                    tol = max(tol, float(mp.gammainc(L + mp.mpf("0.5"), x) / mp.gamma(L + mp.mpf("0.5"))) * 4)
            assert abs(F[m] - exact) <= tol * exact, (x, L, m, F[m], exact)


def test_ssss_closed_form():
    rng = np.random.default_rng(4)
    shells = []
    for _ in range(6):
        shells.append((0, tuple(rng.uniform(-3, 3, 3)), [float(rng.uniform(0.1, 30))], [1.0]))
    sa = normalized_arrays(shells)
    for q in itertools.combinations_with_replacement(range(6), 4):
        got, _, _ = oracle.eri(sa, [q])
        a, b, c, d = [shells[i] for i in q]
        ca, cb, cc, cd = [sa.coef[sa.off[i]] for i in q]
        A, B, C, D = [np.array(s[1]) for s in (a, b, c, d)]
        ea, eb, ec, ed = a[2][0], b[2][0], c[2][0], d[2][0]
        p, qq = ea + eb, ec + ed
        P, Q = (ea * A + eb * B) / p, (ec * C + ed * D) / qq
        T = p * qq / (p + qq) * np.sum((P - Q) ** 2)
        kab = math.exp(-ea * eb / p * np.sum((A - B) ** 2))
        kcd = math.exp(-ec * ed / qq * np.sum((C - D) ** 2))
        want = 2 * math.pi ** 2.5 / (p * qq * math.sqrt(p + qq)) * kab * kcd * float(_mp_boys(0, T)) * ca * cb * cc * cd
        # exp() turns the rounding of its argument into a relative error of
        # |argument| ulps, so the bound grows with the exponents' arguments
        args = ea * eb / p * np.sum((A - B) ** 2) + ec * ed / qq * np.sum((C - D) ** 2) + T
        assert abs(got[0] - want) <= (2e-15 + 4e-16 * args) * abs(want) + 1e-300


def _cases():
    out = []
    shells = mol_shells("water.6-31gss")
    byam = {}
    for n, s in enumerate(shells):
        byam.setdefault(s[0], []).append(n)
    rng = np.random.default_rng(1)
    qs = []
    for am in itertools.product(sorted(byam), repeat=4):
        for _ in range(2):
            qs.append(tuple(int(rng.choice(byam[a])) for a in am))
    out.append(("water.6-31gss", shells, qs))
    shells = mol_shells("water.aug-cc-pvtz")
    byam = {}
    for n, s in enumerate(shells):
        byam.setdefault(s[0], []).append(n)
    qs = []
    for am in itertools.product(sorted(byam), repeat=4):
        if sum(am) > 9:
            continue
        qs.append(tuple(int(rng.choice(byam[a])) for a in am))
    out.append(("water.aug-cc-pvtz (to f)", shells, qs))
    shells = workloads.grid_molecule(natom=2, seed=7)
    qs = [tuple(int(x) for x in rng.integers(0, len(shells), 4)) for _ in range(300)]
    out.append(("grid_molecule 2 atoms", shells, qs))
    shells, qs1 = workloads.pshell_quartets()
    out.append(("pshell_quartets", shells, qs1[:300]))
    return out


@pytest.mark.parametrize("case", _cases(), ids=lambda c: c[0])
def test_oracle_matches_brute_force(case):
    _, shells, qs = case
    sa = normalized_arrays(shells)
    o, offs, sizes = oracle.eri(sa, qs)
    m, _, _ = oracle.md_eri(sa, qs, reference_boys=True)
    d = np.abs(o - m)
    # same Boys values: only FP64 rounding of the recurrences separates them
    assert np.all(d <= 1e-14 * np.maximum(1.0, np.abs(m))), d.max()
    e = per_quartet_errors(o, m, offs, sizes)
    assert np.median(e) < 1e-14
    # exact Boys values: the reference scheme's own approximation error
    mx, _, _ = oracle.md_eri(sa, qs)
    assert np.all(np.abs(o - mx) <= 1e-12 * np.maximum(1.0, np.abs(mx)))


def _block(out, offs, sizes, k, shape):
    return out[offs[k]:offs[k] + sizes[k]].reshape(shape)


def test_invariants():
    shells = workloads.grid_molecule(natom=2, seed=3)
    sa = normalized_arrays(shells)
    rng = np.random.default_rng(2)
    nc = [(s[0] + 1) * (s[0] + 2) // 2 for s in shells]
    for _ in range(40):
        a, b, c, d = (int(x) for x in rng.integers(0, len(shells), 4))
        qs = [(a, b, c, d), (b, a, c, d), (a, b, d, c), (c, d, a, b)]
        o, offs, sizes = oracle.eri(sa, qs)
        sh = (nc[a], nc[b], nc[c], nc[d])
        x0 = _block(o, offs, sizes, 0, sh)
        x1 = _block(o, offs, sizes, 1, (nc[b], nc[a], nc[c], nc[d])).transpose(1, 0, 2, 3)
        x2 = _block(o, offs, sizes, 2, (nc[a], nc[b], nc[d], nc[c])).transpose(0, 1, 3, 2)
        x3 = _block(o, offs, sizes, 3, (nc[c], nc[d], nc[a], nc[b])).transpose(2, 3, 0, 1)
        s = max(1.0, np.abs(x0).max())
        for x in (x1, x2, x3):
            assert np.abs(x - x0).max() <= 1e-14 * s
        # Schwarz: |(ab|cd)| <= sqrt((ab|ab)) sqrt((cd|cd)) element-wise on the diagonal blocks
        d1, o1, s1 = oracle.eri(sa, [(a, b, a, b), (c, d, c, d)])
        gab = _block(d1, o1, s1, 0, (nc[a] * nc[b], nc[a] * nc[b])).diagonal()
        gcd = _block(d1, o1, s1, 1, (nc[c] * nc[d], nc[c] * nc[d])).diagonal()
        bound = np.sqrt(np.outer(np.abs(gab), np.abs(gcd))).reshape(sh)
        assert np.all(np.abs(x0) <= bound * (1 + 1e-12) + 1e-300)
    # translation of every centre by the same vector
    moved = [(s[0], tuple(np.array(s[1]) + np.array([1.25, -0.5, 3.0])), s[2], s[3]) for s in shells]
    sb = normalized_arrays(moved)
    qs = [tuple(int(x) for x in rng.integers(0, len(shells), 4)) for _ in range(30)]
    o, offs, sizes = oracle.eri(sa, qs)
    t, _, _ = oracle.eri(sb, qs)
    assert np.all(np.abs(o - t) <= 1e-12 * np.maximum(1.0, np.abs(o)))


def test_deriv1_against_finite_differences():
    shells = workloads.grid_molecule(natom=2, seed=11)
    sa = normalized_arrays(shells)
    rng = np.random.default_rng(5)
    qs = [tuple(int(x) for x in rng.integers(0, len(shells), 4)) for _ in range(25)]
    qs += [(0, 0, 3, 3), (5, 5, 5, 5), (2, 4, 2, 4)]
    g, offs, sizes = oracle.eri(sa, qs, deriv=1)
    f, _, _ = oracle.md_eri_deriv1(sa, qs, h=1e-4)
    for k in range(len(qs)):
        a = g[offs[k]:offs[k] + sizes[k]]
        b = f[offs[k]:offs[k] + sizes[k]]
        scale = np.abs(b).max()
        # finite differences carry ~1e-9 relative error; components that
        # vanish by symmetry are compared absolutely
        assert np.abs(a - b).max() <= 1e-8 * scale + 1e-14, (qs[k], np.abs(a - b).max(), scale)


def test_md_analytic_deriv1_against_finite_differences():
    """The long-double analytic first derivatives (the exact values the GPU
    derivative tests use) against Richardson finite differences of the
    long-double integrals, and translational invariance."""
    shells = [
        (1, (0.1, -0.3, 0.2), [1.3, 0.4], [0.6, 0.5]),
        (2, (1.0, 0.7, -0.4), [0.9], [1.0]),
        (0, (-0.6, 0.2, 1.1), [2.1, 0.3], [0.4, 0.7]),
        (1, (0.3, 1.2, 0.5), [0.7], [1.0]),
    ]
    sa = oracle.ShellArrays(shells)
    oracle.normalize(sa)
    qs = [(0, 1, 2, 3), (1, 1, 3, 2), (3, 0, 0, 2), (2, 2, 2, 2), (0, 0, 0, 0)]
    ex, offs, sizes = oracle.md_eri_deriv1_analytic(sa, qs)
    fd, _, _ = oracle.md_eri_deriv1(sa, qs, h=1e-3)
    assert np.abs(ex - fd).max() <= 1e-9 * max(1.0, np.abs(ex).max())
    inv = ex.reshape(-1, 4, 3).sum(axis=1)
    assert np.abs(inv).max() <= 1e-15 * max(1.0, np.abs(ex).max())
    # the oracle (deriv = 1) is the same quantity
    ref, _, _ = oracle.eri(sa, qs, deriv=1)
    assert np.abs(ref - ex).max() <= 1e-13 * max(1.0, np.abs(ex).max())


def _random_shells(seed, lmax=3, per_l=3, depth=(1, 6), lo=0.05, hi=50.0, box=2.0):
    rng = np.random.default_rng(seed)
    shells = []
    for l in range(lmax + 1):
        for _ in range(per_l):
            n = int(rng.integers(depth[0], depth[1] + 1))
            shells.append((l, tuple(rng.uniform(-box, box, 3)), list(np.exp(rng.uniform(np.log(lo), np.log(hi), n))),
                           list(rng.uniform(0.1, 1.0, n))))
    return shells


def test_normalize_unit_self_overlap():
    """oracle_normalize (shell.c:248-278): after normalisation the x^l
    Cartesian of every shell has <x^l|x^l> = 1, computed independently in
    long double.  A wrong exponent m, a missing norm_fac[l] or coefficients
    left unscaled by alpha^(m/2) all move this by far more than 1e-13."""
    shells = _random_shells(3, lmax=5, per_l=4)
    sa = oracle.ShellArrays(shells)
    oracle.normalize(sa)
    d, off, n = oracle.md_oei(sa, [(i, i) for i in range(len(shells))], 0)
    selfov = d[off]
    assert np.all(np.abs(selfov - 1.0) < 1e-13), selfov
    # and the unnormalised shells are not (the check has teeth)
    raw = oracle.ShellArrays(shells)
    d0, off0, _ = oracle.md_oei(raw, [(i, i) for i in range(len(shells))], 0)
    assert np.abs(d0[off0] - 1.0).max() > 1e-2


def test_fastschwarz_is_exact_s_type_diagonal():
    """fastschwarz_value (shell_screen.c:226-270) is, for any shell AM, the
    (ij|ij) integral of the s-type primitives i, j with their contraction
    coefficients -- compared with the long-double brute force of that
    single-primitive s quartet."""
    shells = _random_shells(5, lmax=3, per_l=2, depth=(1, 4))
    sa = oracle.ShellArrays(shells)
    oracle.normalize(sa)
    worst = 0.0
    for ia in range(len(shells)):
        for ib in range(len(shells)):
            for i in range(sa.nprim[ia]):
                for j in range(sa.nprim[ib]):
                    v = oracle.fastschwarz(sa, ia, ib, i, j)
                    s2 = [(0, tuple(sa.xyz[ia]), [sa.alpha[sa.off[ia] + i]], [sa.coef[sa.off[ia] + i]]),
                          (0, tuple(sa.xyz[ib]), [sa.alpha[sa.off[ib] + j]], [sa.coef[sa.off[ib] + j]])]
                    ex, _, _ = oracle.md_eri(oracle.ShellArrays(s2), [(0, 1, 0, 1)])
                    # FP64 exp(-2 rho R^2) carries the rounding of its argument:
                    # relative error ~ eps * (1 + 2 rho R^2)
                    a, b = s2[0][2][0], s2[1][2][0]
                    x = 2.0 * a * b / (a + b) * float(((sa.xyz[ia] - sa.xyz[ib]) ** 2).sum())
                    worst = max(worst, abs(v - ex[0]) / abs(ex[0]) / (1.0 + x))
    assert worst < 1e-14, worst


def test_osoei_against_hermite():
    """One-electron restatement (osoei_overlap/ke/potential.c) against the
    long-double Hermite expansions: overlap and kinetic energy to 1e-14
    normwise per block; nuclear attraction to 1e-14 with simint's Boys values
    (algebra) and within the Boys scheme's own error with exact ones."""
    shells = _random_shells(7, lmax=3, per_l=2, depth=(1, 4))
    sa = oracle.ShellArrays(shells)
    oracle.normalize(sa)
    rng = np.random.default_rng(1)
    Z = np.array([1.0, 6.0, 8.0, 92.0])
    C = rng.uniform(-2.5, 2.5, (4, 3))
    pairs = [(i, j) for i in range(len(shells)) for j in range(len(shells))]

    def worst(a, b, off, n):
        return max(np.abs(a[o:o + m] - b[o:o + m]).max() / np.abs(b[o:o + m]).max() for o, m in zip(off, n))
    for kind in (0, 1):
        a, off, n = oracle.osoei(sa, pairs, kind)
        b, _, _ = oracle.md_oei(sa, pairs, kind)
        assert worst(a, b, off, n) < 1e-14, (kind, worst(a, b, off, n))
    a, off, n = oracle.osoei(sa, pairs, 2, Z, C)
    b, _, _ = oracle.md_oei(sa, pairs, 2, Z, C, reference_boys=True)
    assert worst(a, b, off, n) < 1e-14
    b, _, _ = oracle.md_oei(sa, pairs, 2, Z, C)
    assert worst(a, b, off, n) < 1e-11


def test_osoei_s_closed_forms():
    """s-type closed forms: S = (pi/p)^(3/2) e^{-mu R^2}, T = mu (3 - 2 mu
    R^2) S, V = -Z 2 pi/p e^{-mu R^2} F_0(p |P-C|^2) (unit coefficients)."""
    from math import erf, exp, pi, sqrt
    a, b = 0.7, 1.9
    A, B, Cc = np.array([0.1, -0.4, 0.3]), np.array([-0.5, 0.6, 1.1]), np.array([0.9, 0.2, -0.7])
    sa = oracle.ShellArrays([(0, tuple(A), [a], [1.0]), (0, tuple(B), [b], [1.0])])
    p, mu = a + b, a * b / (a + b)
    R2 = float(((A - B) ** 2).sum())
    S = (pi / p) ** 1.5 * exp(-mu * R2)
    T = mu * (3.0 - 2.0 * mu * R2) * S
    P = (a * A + b * B) / p
    x = p * float(((P - Cc) ** 2).sum())
    V = -3.0 * 2.0 * pi / p * exp(-mu * R2) * 0.5 * sqrt(pi / x) * erf(sqrt(x))
    got = [oracle.osoei(sa, [(0, 1)], k, np.array([3.0]), Cc.reshape(1, 3))[0][0] for k in range(3)]
    for g, e in zip(got, (S, T, V)):
        assert abs(g - e) <= 1e-14 * abs(e), (g, e)


def _rank_one_case(l, seed):
    """Shells of mixed AM on several centres, followed by three shells of
    AM l on one centre; D = c c^T on the x^l components of those three.
    phi = sum_s c_s chi_s is then the x^l component of one merged contracted
    shell (appended last to the returned arrays, unnormalised), so
    J_mn = (mn|phi phi) and K_mn = (m phi|n phi) are single quartets of the
    long-double brute force, independent of any tensor assembly."""
    rng = np.random.default_rng(seed)
    others = _random_shells(seed, lmax=2, per_l=1, depth=(1, 3), box=1.5)
    cen = tuple(rng.uniform(-1, 1, 3))
    phis = [(l, cen, list(np.exp(rng.uniform(np.log(0.1), np.log(20), k))), list(rng.uniform(0.2, 1.0, k)))
            for k in (1, 2, 3)]
    shells = others + phis
    sa = oracle.ShellArrays(shells)
    oracle.normalize(sa)
    c = rng.uniform(-1, 1, len(phis))
    first = len(others)
    al = np.concatenate([sa.alpha[sa.off[s]:sa.off[s] + sa.nprim[s]] for s in range(first, len(shells))])
    co = np.concatenate([c[k] * sa.coef[sa.off[first + k]:sa.off[first + k] + sa.nprim[first + k]]
                         for k in range(len(phis))])
    full = oracle.ShellArrays([(int(sa.am[s]), tuple(sa.xyz[s]), list(sa.alpha[sa.off[s]:sa.off[s] + sa.nprim[s]]),
                                list(sa.coef[sa.off[s]:sa.off[s] + sa.nprim[s]])) for s in range(len(shells))]
                              + [(l, cen, list(al), list(co))])
    nf = [oracle.ncart(int(x)) for x in sa.am]
    f0 = np.r_[0, np.cumsum(nf)]
    vec = np.zeros(f0[-1])
    vec[f0[first:-1]] = c
    return sa, full, np.outer(vec, vec), nf, f0


@pytest.mark.parametrize("l", [0, 1, 2])
def test_jk_rank_one_density(l):
    """oracle.jk on the full tensor against single long-double quartets:
    for D = c c^T with phi = sum c_mu chi_mu one merged shell's component,
    J_mn = (mn|phi phi) and K_mn = (m phi|n phi) for every pair of functions
    on any centre.  A swapped J/K index pattern, a transposed block in the
    tensor or a wrong function offset fails this."""
    sa, full, D, nf, f0 = _rank_one_case(l, 40 + l)
    g, _ = oracle.eri_tensor(sa)
    J, K = oracle.jk(g, D)
    ns, m = len(nf), len(nf)
    qJ = [(a, b, m, m) for a in range(ns) for b in range(ns)]
    qK = [(a, m, b, m) for a in range(ns) for b in range(ns)]
    vJ, oJ, _ = oracle.md_eri(full, qJ, reference_boys=True)
    vK, oK, _ = oracle.md_eri(full, qK, reference_boys=True)
    nphi = oracle.ncart(l)
    Jx = np.zeros_like(J)
    Kx = np.zeros_like(K)
    for (a, b, _, _), o in zip(qJ, oJ):
        Jx[f0[a]:f0[a + 1], f0[b]:f0[b + 1]] = vJ[o:o + nf[a] * nf[b] * nphi * nphi].reshape(
            nf[a], nf[b], nphi, nphi)[:, :, 0, 0]
    for (a, _, b, _), o in zip(qK, oK):
        Kx[f0[a]:f0[a + 1], f0[b]:f0[b + 1]] = vK[o:o + nf[a] * nf[b] * nphi * nphi].reshape(
            nf[a], nphi, nf[b], nphi)[:, 0, :, 0]
    assert np.abs(J - Jx).max() <= 1e-12 * np.abs(Jx).max()
    assert np.abs(K - Kx).max() <= 1e-12 * np.abs(Kx).max()


def test_jk_tensor_symmetry_and_psd():
    """The assembled tensor has the 8-fold permutational symmetry of real
    integrals (it is built from every ordered quartet, none inferred), and
    for a random positive semidefinite density over shells on several
    centres K is positive semidefinite and tr(D J) > 0."""
    shells = _random_shells(8, lmax=2, per_l=2, depth=(1, 3), box=1.5)
    sa = oracle.ShellArrays(shells)
    oracle.normalize(sa)
    g, f0 = oracle.eri_tensor(sa)
    for perm in [(1, 0, 2, 3), (0, 1, 3, 2), (2, 3, 0, 1)]:
        assert np.abs(g - g.transpose(perm)).max() <= 1e-13 * np.abs(g).max()
    rng = np.random.default_rng(2)
    C = rng.uniform(-1, 1, (g.shape[0], 3))
    D = C @ C.T
    J, K = oracle.jk(g, D)
    assert np.trace(D @ J) > 0
    assert np.linalg.eigvalsh(0.5 * (K + K.T)).min() >= -1e-12 * np.abs(K).max()
    assert np.abs(K - K.T).max() <= 1e-12 * np.abs(K).max()
