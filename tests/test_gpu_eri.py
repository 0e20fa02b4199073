"""Parity of the GPU ERI path (through the C ABI) against the CPU oracle.

Bar (helpers.parity_report): elementwise |gpu - oracle| <= 1e-14 max(1,|oracle|)
(the reference's own acceptance bar, skel/test/test_eri.cpp:227), and where
the long-double McMurchie-Davidson value ("exact", fed the reference's Boys
values) is computed, per shell quartet
max|gpu - oracle| <= max(1e-14 max|oracle|, 4 max|oracle - exact|): the GPU
may differ from the reference algorithm by no more than that algorithm's own
FP64 rounding error.  A flat 1e-14 normwise bar per quartet is not attainable
in FP64: the oracle itself is up to 4.6e-13 (normwise) away from the exact
algebra on the 1000 (pp|pp) quartets, from cancellation in quartets whose integrals are tiny
(absolute error <= 2e-16); the normwise statistics are printed.
"""
import itertools

import numpy as np
import pytest

import oracle
import simint
import workloads
from helpers import mol_shells, normalized_arrays, parity_report, parity_report_deriv

pytestmark = pytest.mark.gpu

TOL = 1e-14


def _oracle_for_dropin(sa, A, B, C, D, screen_method=0, screen_tol=0.0, deriv=0):
    """Oracle output laid out as simint_compute_eri writes P=(A x B), Q=(C x D)."""
    quartets = [(a, b, c, d) for a in A for b in B for c in C for d in D]
    out, offs, sizes = oracle.eri(sa, quartets, screen_method, screen_tol, deriv)
    return out, offs, sizes


def _check(sa, basis, A, B, C, D, screen=0, tol=0.0, exact=False, deriv=0):
    """Run simint_compute_eri (or _deriv) on P = A x B, Q = C x D and compare
    with the oracle; returns the parity stats (asserting the bar)."""
    P = simint.Pairs(basis, A, B, screen)
    Q = simint.Pairs(basis, C, D, screen)
    try:
        rc, got = simint.compute_eri(P, Q, tol, deriv)
    finally:
        P.free()
        Q.free()
    ref, offs, sizes = _oracle_for_dropin(sa, A, B, C, D, screen, tol, deriv)
    if rc == -1:
        got = np.zeros_like(ref)
    else:
        assert rc == len(A) * len(B) * len(C) * len(D), (rc, simint.last_error())
    quartets = [(a, b, c, d) for a in A for b in B for c in C for d in D]
    if deriv:
        ex, _, _ = oracle.md_eri_deriv1_analytic(sa, quartets)
        ok, st = parity_report_deriv(got, ref, offs, sizes, ex)
        assert ok, ((A, B, C, D), st)
        return st
    ex = None
    if exact:
        ex, _, _ = oracle.md_eri(sa, quartets, reference_boys=True)
    ok, st = parity_report(got, ref, offs, sizes, ex)
    assert ok, ((A, B, C, D), st)
    return st


def test_pppp_1000_single_quartets(gpu):
    """1000 (pp|pp) quartets of 40 p shells, one simint_compute_eri call each."""
    shells, quartets = workloads.pshell_quartets()
    basis = simint.Basis(shells)
    sa = normalized_arrays(shells)
    worst = 0.0
    for (a, b, c, d) in quartets:
        st = _check(sa, basis, [a], [b], [c], [d], exact=True)
        worst = max(worst, st["max_normwise"])
    print("pshell_quartets worst normwise rel vs oracle", worst)


def test_pppp_1000_batched(gpu):
    shells, _ = workloads.pshell_quartets()
    basis = simint.Basis(shells)
    sa = normalized_arrays(shells)
    st = _check(sa, basis, list(range(0, 6)), list(range(0, 6)), list(range(10, 25)), list(range(20, 40)))
    print(st)


@pytest.mark.parametrize("mol", ["water.6-31gss", "water.sto-3g"])
def test_all_classes_molecule(gpu, mol):
    """Every class up to d for a molecule from simint's test fixtures:
    P = all shells of AM i x all of AM j, Q = all of AM k x all of AM l
    (the shape skel/test/test_eri.cpp:103-168 uses, with P batched)."""
    shells = mol_shells(mol)
    basis = simint.Basis(shells)
    sa = normalized_arrays(shells)
    byam = {}
    for n, s in enumerate(shells):
        byam.setdefault(s[0], []).append(n)
    ams = sorted(a for a in byam if a <= 2)
    worst = {}
    for i, j, k, l in itertools.product(ams, repeat=4):
        st = _check(sa, basis, byam[i], byam[j], byam[k], byam[l], exact=True)
        worst[(i, j, k, l)] = st["max_normwise"]
    print("worst normwise over classes", max(worst.values()))


@pytest.mark.parametrize("mol", ["benzene.sto-3g", "ethane.6-31gss", "water.dzp"])
def test_fixture_molecules_all_classes(gpu, mol):
    """The other molecules of simint's test fixtures (skel/test/dat/): every
    class of their shells, P and Q built from the first and last shell of
    each AM (up to 16 quartets per class), against the oracle and the
    long-double values."""
    shells = mol_shells(mol)
    basis = simint.Basis(shells)
    sa = normalized_arrays(shells)
    byam = {}
    for n, s in enumerate(shells):
        byam.setdefault(s[0], []).append(n)
    byam = {l: sorted({v[0], v[-1]}) for l, v in byam.items()}
    worst = 0.0
    for i, j, k, l in itertools.product(sorted(byam), repeat=4):
        st = _check(sa, basis, byam[i], byam[j], byam[k], byam[l], exact=True)
        worst = max(worst, st["max_normwise"])
    print(mol, "worst normwise over classes", worst)


def _f_shells(seed=11):
    """Two shells of each AM 0..3 on random centres in a 4-bohr cube,
    contraction depth 1-3, exponents log-uniform in [0.1, 50]."""
    rng = np.random.default_rng(seed)
    shells = []
    for l in range(4):
        for _ in range(2):
            k = int(rng.integers(1, 4))
            shells.append((l, tuple(rng.uniform(0, 4, 3)), list(np.exp(rng.uniform(np.log(0.1), np.log(50), k))),
                           list(rng.uniform(0.1, 1.0, k))))
    return shells


def test_f_classes(gpu):
    """Every class with an f shell: four-centre quartets of random shells,
    and water aug-cc-pVTZ (contracted 8-primitive s, f on oxygen, d on
    hydrogen) with the first and last shell of each AM."""
    for shells, pick in [(_f_shells(), None), (mol_shells("water.aug-cc-pvtz"), "ends")]:
        basis = simint.Basis(shells)
        sa = normalized_arrays(shells)
        byam = {}
        for n, s in enumerate(shells):
            byam.setdefault(s[0], []).append(n)
        if pick == "ends":
            byam = {l: sorted({v[0], v[-1]}) for l, v in byam.items()}
        worst = 0.0
        for am in itertools.product(range(4), repeat=4):
            if 3 not in am:
                continue
            st = _check(sa, basis, byam[am[0]], byam[am[1]], byam[am[2]], byam[am[3]], exact=True)
            worst = max(worst, st["max_normwise"])
        print("f classes worst normwise", worst)


@pytest.mark.parametrize("mol", ["water.6-31gss"])
def test_deriv1_all_classes(gpu, mol):
    """First derivatives (simint_compute_eri_deriv, 12 values per Cartesian
    [Ax Ay Az Bx ... Dz], ostei_deriv1_generator.cpp) for every class up to
    d against the oracle's deriv = 1 path."""
    shells = mol_shells(mol)
    basis = simint.Basis(shells)
    sa = normalized_arrays(shells)
    byam = {}
    for n, s in enumerate(shells):
        byam.setdefault(s[0], []).append(n)
    ams = sorted(a for a in byam if a <= 2)
    worst = 0.0
    for i, j, k, l in itertools.product(ams, repeat=4):
        st = _check(sa, basis, byam[i], byam[j], byam[k], byam[l], deriv=1)
        worst = max(worst, st["max_abs_vs_oracle"])
    print("deriv1 worst |gpu - oracle| over classes", worst)


def test_deriv1_tight_exponents(gpu):
    """The 64-atom recipe's shells (core exponents up to 5000) on two atoms,
    all classes up to d: the ill-conditioned derivatives."""
    shells = workloads.grid_molecule(natom=2, seed=5)
    basis = simint.Basis(shells)
    sa = normalized_arrays(shells)
    byam = {}
    for n, s in enumerate(shells):
        byam.setdefault(s[0], []).append(n)
    for am in itertools.product(range(3), repeat=4):
        _check(sa, basis, byam[am[0]], byam[am[1]], byam[am[2]], byam[am[3]], deriv=1)


@pytest.mark.parametrize("screen", [simint.SCREEN_FASTSCHWARZ, simint.SCREEN_SCHWARZ])
def test_deriv1_screened_and_limits(gpu, screen):
    """Primitive screening on the derivative path, the same-shell pair
    weighting (alpha2 = beta2 = a_i + a_j on a triangle pair), and a class
    above the derivative maximum rejected with SIMINT_ERROR_UNSUPPORTED."""
    shells = workloads.grid_molecule(natom=2, seed=5)
    basis = simint.Basis(shells)
    sa = normalized_arrays(shells)
    byam = {}
    for n, s in enumerate(shells):
        byam.setdefault(s[0], []).append(n)
    for am in [(0, 0, 0, 0), (1, 0, 1, 0), (1, 1, 1, 1), (2, 1, 0, 1), (2, 2, 2, 1)]:
        _check(sa, basis, byam[am[0]], byam[am[1]], byam[am[2]], byam[am[3]], screen, 1e-7, deriv=1)
    # same shell on both sides of a pair
    s0 = byam[1][0]
    _check(sa, basis, [s0], [s0], [s0], [s0], deriv=1)
    # a g shell is above the compiled maximum AM (3), for integrals and
    # derivatives alike
    d = byam[2][0]
    b2 = simint.Basis([(4, shells[d][1], shells[d][2], shells[d][3])])
    P = simint.Pairs(b2, [0], [0])
    for der in (0, 1):
        rc, _ = simint.compute_eri(P, P, 0.0, der)
        assert rc == simint.ERROR_UNSUPPORTED, rc
    P.free()


def test_deriv1_f_classes(gpu):
    """First derivatives of every class with an f shell (SIMINT_OSTEI_DERIV1_MAXAM
    equals the maximum AM, as in the reference, ostei_config.h.in:7): their
    (a+1) terms are the classes with a g shell, built for derivatives only.
    Two shells of each AM, every combination per class, against the oracle
    and the exact long-double analytic derivatives."""
    shells = _f_shells(seed=23)
    basis = simint.Basis(shells)
    sa = normalized_arrays(shells)
    byam = {}
    for n, s in enumerate(shells):
        byam.setdefault(s[0], []).append(n)
    assert simint.lib().simint_eri_worksize(1, 3) > 0
    worst = 0.0
    for am in itertools.product(range(4), repeat=4):
        if 3 not in am:
            continue
        st = _check(sa, basis, byam[am[0]], byam[am[1]], byam[am[2]], byam[am[3]], deriv=1)
        worst = max(worst, st["max_abs_vs_oracle"])
    print("f-class deriv1 worst |gpu - oracle|", worst)


def test_lane_and_wave_kernels_agree(gpu):
    """Every kernel family (lane-per-quartet in registers, wave-per-quartet
    in LDS, the split lane-per-quartet kernels) against the oracle for every
    class up to d that has it."""
    shells = workloads.grid_molecule(natom=2)
    basis = simint.Basis(shells)
    sa = normalized_arrays(shells)
    byam = {}
    for n, s in enumerate(shells):
        byam.setdefault(s[0], []).append(n)
    L = simint.lib()
    checked = 0
    for am in itertools.product(range(3), repeat=4):
        prev = L.simint_hip_class_kernel(*am)
        for fam in (1, 2, 3, 4):
            if L.simint_hip_set_class_kernel(*am, fam) < 0:
                continue
            _check(sa, basis, byam[am[0]], byam[am[1]], byam[am[2]], byam[am[3]])
            checked += 1
        L.simint_hip_set_class_kernel(*am, prev)
    assert checked >= 81


@pytest.mark.parametrize("screen", [simint.SCREEN_FASTSCHWARZ, simint.SCREEN_SCHWARZ])
@pytest.mark.parametrize("tol", [1e-4, 1e-7])
def test_primitive_screening(gpu, screen, tol):
    """Primitive screening (OSTEI_Writer.cpp:504-587, one primitive quartet at
    a time) matches the oracle's restatement, including the primitive sort the
    screened pairs get (shell.c:579-581)."""
    shells = workloads.grid_molecule(natom=2, seed=3)
    basis = simint.Basis(shells)
    sa = normalized_arrays(shells)
    byam = {}
    for n, s in enumerate(shells):
        byam.setdefault(s[0], []).append(n)
    for am in [(0, 0, 0, 0), (1, 0, 1, 0), (1, 1, 1, 1), (2, 0, 1, 0), (2, 1, 2, 0)]:
        _check(sa, basis, byam[am[0]], byam[am[1]], byam[am[2]], byam[am[3]], screen, tol)


def test_empty_and_clipped(gpu):
    shells, _ = workloads.pshell_quartets()
    basis = simint.Basis(shells)
    P = simint.Pairs(basis, [0, 1], [2, 3])
    Q = simint.Pairs(basis, [4], [5])
    P.mp.nshell12_clip = 0
    rc, _ = simint.compute_eri(P, Q)
    assert rc == 0
    P.mp.nshell12_clip = 3       # first three of the four pairs
    rc, got = simint.compute_eri(P, Q)
    assert rc == 3
    sa = normalized_arrays(shells)
    ref, offs, sizes = oracle.eri(sa, [(0, 2, 4, 5), (0, 3, 4, 5), (1, 2, 4, 5)])
    assert parity_report(got, ref, offs, sizes)[0]
    P.free()
    Q.free()


def test_extreme_geometry(gpu):
    """Coincident centres (Boys argument 0), far-apart tight shells (long-range
    branch x >= 36.5) and very diffuse exponents in one set."""
    shells = [
        (0, (0.0, 0.0, 0.0), [5000.0, 0.05], [0.3, 0.7]),
        (1, (0.0, 0.0, 0.0), [2.0], [1.0]),
        (2, (0.0, 0.0, 0.0), [0.8], [1.0]),
        (1, (25.0, -3.0, 11.0), [30.0, 4.0], [0.5, 0.5]),
        (2, (25.0, -3.0, 11.0), [12.0], [1.0]),
        (0, (40.0, 40.0, 40.0), [1e4], [1.0]),
    ]
    basis = simint.Basis(shells)
    sa = normalized_arrays(shells)
    for q in itertools.product(range(len(shells)), repeat=4):
        _check(sa, basis, [q[0]], [q[1]], [q[2]], [q[3]], exact=True)


def test_device_prefix_and_diag(gpu):
    """simint_hip_compute_eri_prefix / _diag: device-resident pairs, per-bra
    ket prefixes, integrals left in HBM."""
    import torch
    shells = workloads.grid_molecule(natom=3, seed=1)
    basis = simint.Basis(shells)
    sa = normalized_arrays(shells)
    byam = {}
    for n, s in enumerate(shells):
        byam.setdefault(s[0], []).append(n)
    rng = np.random.default_rng(0)
    for (la, lb, lc, ld) in [(0, 0, 0, 0), (1, 1, 0, 0), (2, 1, 1, 0), (2, 2, 2, 2), (1, 0, 2, 1)]:
        A, B, C, D = byam[la], byam[lb], byam[lc], byam[ld]
        Pp = simint.Pairs(basis, A, B)
        Qp = simint.Pairs(basis, C, D)
        P, Q = simint.DevicePairs(Pp), simint.DevicePairs(Qp)
        nc = simint.ncart(la) * simint.ncart(lb) * simint.ncart(lc) * simint.ncart(ld)
        b0 = 1
        nbra = P.n - 1
        nket = rng.integers(0, Q.n + 1, size=nbra)
        off = np.zeros(nbra, dtype=np.int64)
        off[1:] = np.cumsum(nket[:-1]) * nc
        out = torch.zeros(int(nket.sum()) * nc + 1, dtype=torch.float64, device="cuda")
        simint.compute_prefix(P, Q, b0, nket, off, out.data_ptr())
        simint.synchronize()
        got = out.cpu().numpy()[:int(nket.sum()) * nc]
        pairsP = [(a, b) for a in A for b in B]
        pairsQ = [(c, d) for c in C for d in D]
        quartets = []
        for i in range(nbra):
            for k in range(nket[i]):
                quartets.append(pairsP[b0 + i] + pairsQ[k])
        ref, offs, sizes = oracle.eri(sa, quartets)
        assert parity_report(got, ref, offs, sizes)[0]
        # diagonal (bb|bb)
        if (la, lb) == (lc, ld):
            outd = torch.zeros(P.n * nc, dtype=torch.float64, device="cuda")
            simint.compute_diag(P, outd.data_ptr())
            simint.synchronize()
            refd, offs, sizes = oracle.eri(sa, [p + p for p in pairsP])
            assert parity_report(outd.cpu().numpy(), refd, offs, sizes)[0]
        P.free(); Q.free(); Pp.free(); Qp.free()


def test_split_kernels_at_scale(gpu):
    """Thousands of work items in flight at once (the split kernels' staging
    areas must not overlap): (bb|bb) of every (dp) and (dd) pair of a 16-atom
    system through each family, equal to the oracle on a sample and to each
    other everywhere."""
    import torch
    shells = workloads.grid_molecule(natom=16, seed=0)
    basis = simint.Basis(shells)
    sa = normalized_arrays(shells)
    L = simint.lib()
    rng = np.random.default_rng(5)
    for cls in [(2, 1), (2, 2)]:
        pairs = [(i, j) for i in range(len(shells)) for j in range(len(shells))
                 if shells[i][0] == cls[0] and shells[j][0] == cls[1] and (cls[0] != cls[1] or i >= j)]
        tmp = simint.Pairs.from_pairs(basis, pairs)
        dev = simint.DevicePairs(tmp)
        nab = simint.ncart(cls[0]) * simint.ncart(cls[1])
        am = (cls[0], cls[1], cls[0], cls[1])
        prev = L.simint_hip_class_kernel(*am)
        res = {}
        for fam in (1, 2, 3, 4):
            if L.simint_hip_set_class_kernel(*am, fam) < 0:
                continue
            out = torch.full((len(pairs) * nab * nab,), float("nan"), dtype=torch.float64, device="cuda")
            torch.cuda.synchronize()   # fill done before the library stream writes
            simint.compute_diag(dev, out.data_ptr())
            simint.synchronize()
            res[fam] = out.cpu().numpy()
        L.simint_hip_set_class_kernel(*am, prev)
        sel = rng.choice(len(pairs), size=24, replace=False)
        ref, offs, sizes = oracle.eri(sa, [pairs[k] + pairs[k] for k in sel])
        for fam, v in res.items():
            got = np.concatenate([v[k * nab * nab:(k + 1) * nab * nab] for k in sel])
            ok, st = parity_report(got, ref, offs, sizes)
            assert ok, (cls, fam, st)
            assert np.abs(v - res[2]).max() <= 1e-13 * max(1.0, np.abs(res[2]).max()), (cls, fam)
        dev.free()
        tmp.free()


def test_split_kernels_both_orientations(gpu):
    """The split kernels against the long-double values for classes whose
    bra outweighs the ket (bra-top groups, ket HRR first) and the reverse
    (ket-top groups, bra HRR first), each family that is built."""
    shells = _f_shells(seed=13)
    basis = simint.Basis(shells)
    sa = normalized_arrays(shells)
    byam = {}
    for n, s in enumerate(shells):
        byam.setdefault(s[0], []).append(n)
    L = simint.lib()
    checked = 0
    for am in [(2, 2, 1, 1), (1, 1, 2, 2), (2, 1, 2, 2), (3, 2, 2, 1), (1, 2, 2, 3), (1, 1, 3, 3), (2, 2, 2, 2)]:
        prev = L.simint_hip_class_kernel(*am)
        for fam in (3, 4):
            if L.simint_hip_set_class_kernel(*am, fam) < 0:
                continue
            _check(sa, basis, byam[am[0]], byam[am[1]], byam[am[2]], byam[am[3]], exact=True)
            checked += 1
        L.simint_hip_set_class_kernel(*am, prev)
    assert checked >= 7
