"""simint_hip_compute_eri_deriv1_list_noalloc: the derivative list with its
scratch in the library's fixed device array agrees bit for bit with
simint_hip_compute_eri_deriv1_list on a caller's work buffer (same kernels,
same order), on two calls in a row (the second smaller, so stale values of
the first are left in the array), and with the oracle; a list too large for
the array is refused before anything is queued."""
import ctypes

import numpy as np
import pytest

import oracle
import simint
import workloads
from helpers import normalized_arrays

pytestmark = pytest.mark.gpu


def _noalloc(P, Q, n, qb, qk, qo, out):
    L = simint.lib()
    f = L.simint_hip_compute_eri_deriv1_list_noalloc
    f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_longlong, ctypes.c_void_p, ctypes.c_void_p,
                  ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p]
    f.restype = ctypes.c_longlong
    return f(P.h, Q.h, int(n), ctypes.c_void_p(qb), ctypes.c_void_p(qk), ctypes.c_void_p(qo or None), 0.0,
             ctypes.c_void_p(out))


def test_noalloc_matches_work_buffer_entry(gpu):
    import torch
    shells = workloads.grid_molecule(natom=2, seed=17)
    basis = simint.Basis(shells)
    sa = normalized_arrays(shells)
    byam = {}
    for i, s in enumerate(shells):
        byam.setdefault(s[0], []).append(i)
    A, B, C, D = byam[2], byam[1], byam[1], byam[0]
    Ph = simint.Pairs(basis, A, B)
    Qh = simint.Pairs(basis, C, D)
    P, Q = simint.DevicePairs(Ph), simint.DevicePairs(Qh)
    nc12 = 6 * 3 * 3 * 1 * 12
    rng = np.random.default_rng(4)
    for n in (300, 57):
        qb = rng.integers(0, P.n, n).astype(np.int32)
        qk = rng.integers(0, Q.n, n).astype(np.int32)
        qo = (np.arange(n, dtype=np.int64)[::-1] * nc12).copy()     # reversed placement
        d_qb = torch.tensor(qb, device="cuda")
        d_qk = torch.tensor(qk, device="cuda")
        d_qo = torch.tensor(qo, device="cuda")
        ref = torch.full((n * nc12,), float("nan"), dtype=torch.float64, device="cuda")
        got = torch.full((n * nc12,), float("nan"), dtype=torch.float64, device="cuda")
        work = torch.empty(max(simint.deriv1_worksize(P, Q, n), 1), dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        simint.compute_deriv1_list(P, Q, n, d_qb.data_ptr(), d_qk.data_ptr(), d_qo.data_ptr(), work.data_ptr(),
                                   ref.data_ptr())
        r = _noalloc(P, Q, n, d_qb.data_ptr(), d_qk.data_ptr(), d_qo.data_ptr(), got.data_ptr())
        assert r == n, (r, simint.last_error())
        simint.synchronize()
        assert torch.equal(got, ref), float((got - ref).abs().max())
        # and the oracle, for the first few quartets
        h = got.cpu().numpy()
        pairs_b = [(a, b) for a in A for b in B]
        pairs_k = [(c, d) for c in C for d in D]
        qs = [pairs_b[qb[i]] + pairs_k[qk[i]] for i in range(12)]
        orc, _, _ = oracle.eri(sa, qs, deriv=1)
        mine = np.concatenate([h[qo[i]:qo[i] + nc12] for i in range(12)])
        assert np.abs(mine - orc).max() <= 1e-10
    # a list whose scratch exceeds the fixed array is refused, nothing queued
    per = simint.deriv1_worksize(P, Q, 1)
    big = (8 << 20) // per + 1
    r = _noalloc(P, Q, big, d_qb.data_ptr(), d_qk.data_ptr(), d_qo.data_ptr(), got.data_ptr())
    assert r == simint.ERROR_ARGUMENT
    assert "fixed area" in simint.last_error()
    simint.synchronize()
    P.free()
    Q.free()
    Ph.free()
    Qh.free()
    basis.free()
