"""GAMESS Cartesian order of the drop-in calls (simint_hip_set_ordering):
ERIs, first derivatives and one-electron integrals come out as the oracle's
Psi4-ordered values with every d and f shell's components permuted into
GAMESS order; switching back restores Psi4 order."""
import itertools

import numpy as np
import pytest

import oracle
import simint
from helpers import normalized_arrays

pytestmark = pytest.mark.gpu


def _exponents(l, gamess):
    """Cartesian exponents of AM l in output order.  Psi4: x from l down,
    then y from l - x down.  GAMESS (its published convention for d and f;
    the same as Psi4 elsewhere): xx yy zz xy xz yz and
    xxx yyy zzz xxy xxz xyy yyz xzz yzz xyz."""
    if gamess and l == 2:
        return [(2, 0, 0), (0, 2, 0), (0, 0, 2), (1, 1, 0), (1, 0, 1), (0, 1, 1)]
    if gamess and l == 3:
        return [(3, 0, 0), (0, 3, 0), (0, 0, 3), (2, 1, 0), (2, 0, 1), (1, 2, 0), (0, 2, 1), (1, 0, 2), (0, 1, 2),
                (1, 1, 1)]
    return [(x, y, l - x - y) for x in range(l, -1, -1) for y in range(l - x, -1, -1)]


def _to_gamess(block, ams, width=1):
    """Reorder one Psi4-ordered block of len prod(ncart) * width."""
    psi = [_exponents(l, False) for l in ams]
    gam = [_exponents(l, True) for l in ams]
    n = [len(p) for p in psi]
    src = np.asarray(block).reshape(n + [width])
    idx = [[p.index(e) for e in g] for p, g in zip(psi, gam)]
    return src[np.ix_(*idx, range(width))].ravel()


def _shells():
    rng = np.random.default_rng(31)
    return [(l, tuple(rng.uniform(-1.5, 1.5, 3)), list(np.exp(rng.uniform(np.log(0.3), np.log(9.0), 2))),
             list(rng.uniform(0.3, 1.0, 2))) for l in range(4)]


def test_gamess_order_dropin(gpu):
    shells = _shells()
    basis = simint.Basis(shells)
    sa = normalized_arrays(shells)
    L = simint.lib()
    prev = L.simint_hip_set_ordering(simint.ORDERING_GAMESS)
    try:
        assert L.simint_hip_ordering() == simint.ORDERING_GAMESS
        for q in [(2, 0, 0, 0), (2, 1, 3, 0), (3, 3, 2, 1), (1, 2, 2, 3), (0, 1, 0, 1)]:
            for deriv in (0, 1):
                P = simint.Pairs(basis, [q[0]], [q[1]])
                Q = simint.Pairs(basis, [q[2]], [q[3]])
                rc, got = simint.compute_eri(P, Q, 0.0, deriv)
                P.free()
                Q.free()
                assert rc == 1, (rc, simint.last_error())
                ref, _, _ = oracle.eri(sa, [q], deriv=deriv)
                want = _to_gamess(ref, [shells[i][0] for i in q], 12 if deriv else 1)
                tol = 1e-10 if deriv else 1e-14 * max(1.0, np.abs(ref).max())
                assert np.abs(got - want).max() <= tol, (q, deriv)
        Z, C = np.array([2.0, 7.0]), np.array([[0.1, 0.2, 0.3], [-0.4, 0.5, -0.6]])
        for i, j in itertools.product(range(4), repeat=2):
            for kind, k in simint.OEI_KINDS.items():
                rc, got = simint.compute_oei(kind, basis, i, j, (Z, C))
                assert rc == 1
                ref, _, _ = oracle.osoei(sa, [(i, j)], k, Z if k == 2 else None, C if k == 2 else None)
                want = _to_gamess(ref, [shells[i][0], shells[j][0]])
                assert np.abs(got - want).max() <= 1e-13 * max(1.0, np.abs(ref).max()), (i, j, kind)
    finally:
        L.simint_hip_set_ordering(prev)
    assert L.simint_hip_ordering() == simint.ORDERING_PSI4
    # Psi4 order again
    P = simint.Pairs(basis, [2], [3])
    rc, got = simint.compute_eri(P, P, 0.0, 0)
    P.free()
    ref, _, _ = oracle.eri(sa, [(2, 3, 2, 3)])
    assert np.abs(got - ref).max() <= 1e-14 * max(1.0, np.abs(ref).max())
    assert L.simint_hip_set_ordering(7) == simint.ERROR_ARGUMENT
    basis.free()
