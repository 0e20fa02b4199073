"""Coulomb/exchange digestion on the GPU (simint_hip_compute_jk_range and
simint_hip_jk_symmetrize) against the oracle's J and K from the full
four-index tensor (every ordered shell quartet, no permutational symmetry
used, oracle.jk).

The GPU digests each unique shell quartet of a ScreenedSet once (no
screening: tol 0) with the weight of its permutational images; a wrong
weight for same-shell pairs, for a quartet that is its own bra-ket image,
or a wrong image's index pattern moves J or K by O(1) relative to their
size.  The GPU sums with atomics in varying order, so the bar is normwise:
max|gpu - oracle| <= 1e-12 max|oracle| (the integrals themselves meet the
1e-14 bar; J and K are sums of up to nbf^2 terms each), or, where that
fails, at most 4x the oracle's own normwise distance from J and K of the
long-double tensor (for f shells with tight exponents the oracle's K is
itself ~1e-12 away from the exact one)."""
import numpy as np
import pytest

import oracle
import quartets
import simint
import workloads
from helpers import mol_shells, normalized_arrays

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _f_set(seed=3):
    """Two shells of each AM 0..3 on random centres, contraction depth 1-3."""
    rng = np.random.default_rng(seed)
    shells = []
    for l in range(4):
        for _ in range(2):
            k = int(rng.integers(1, 4))
            shells.append((l, tuple(rng.uniform(0, 3, 3)), list(np.exp(rng.uniform(np.log(0.1), np.log(30), k))),
                           list(rng.uniform(0.1, 1.0, k))))
    return shells


def _gpu_jk(shells, D, mixed=None, chunk=1 << 22, want=("J", "K")):
    import torch
    ss = quartets.ScreenedSet(shells, tol=0.0, out_doubles=chunk, mixed=mixed)
    try:
        nbf = ss.nbf()
        dD = torch.from_numpy(D).cuda()
        J = torch.zeros((nbf, nbf), dtype=torch.float64, device="cuda")
        K = torch.zeros((nbf, nbf), dtype=torch.float64, device="cuda")
        out = torch.empty(chunk, dtype=torch.float64, device="cuda")
        jp = J.data_ptr() if "J" in want else None
        kp = K.data_ptr() if "K" in want else None
        ss.run_jk(out.data_ptr(), dD.data_ptr(), jp, kp)
        simint.jk_symmetrize(nbf, jp, kp)
        simint.synchronize()
        return J.cpu().numpy(), K.cpu().numpy(), ss.total_quartets
    finally:
        ss.free()


def _density(nbf, seed):
    rng = np.random.default_rng(seed)
    C = rng.uniform(-1, 1, (nbf, max(2, nbf // 4)))
    return np.ascontiguousarray(C @ C.T / nbf)


def _compare(shells, seed, **kw):
    sa = normalized_arrays(shells)
    g, _ = oracle.eri_tensor(sa)
    D = _density(g.shape[0], seed)
    Jr, Kr = oracle.jk(g, D)
    J, K, nq = _gpu_jk(shells, D, **kw)
    n = len(shells)
    # unique quartets: pairs i >= j, then pair-of-pairs ij >= kl
    npair = n * (n + 1) // 2
    assert nq == npair * (npair + 1) // 2
    ej = np.abs(J - Jr).max() / np.abs(Jr).max()
    ek = np.abs(K - Kr).max() / np.abs(Kr).max()
    print("nbf %d  J normwise %.2e  K normwise %.2e" % (g.shape[0], ej, ek))
    ok = [ej <= TOL, ek <= TOL]
    if not all(ok):
        ge, _ = oracle.eri_tensor(sa, exact=True)
        for i, (got, ref, ex) in enumerate(zip((J, K), (Jr, Kr), oracle.jk(ge, D))):
            o = np.abs(ref - ex).max() / np.abs(ex).max()
            e = np.abs(got - ex).max() / np.abs(ex).max()
            print("%s vs long double: oracle %.2e gpu %.2e" % ("JK"[i], o, e))
            ok[i] = ok[i] or e <= max(TOL, 4 * o)
    return all(ok)


def test_jk_water(gpu):
    """Water 6-31G** (s, p and d shells; same-shell pairs; every quartet that
    is its own bra-ket image)."""
    assert _compare(mol_shells("water.6-31gss"), 1)


def test_jk_f_shells_small_chunks(gpu):
    """Shells up to f (split and cooperative class kernels among them) with
    an output buffer small enough that launches are cut into ket ranges."""
    assert _compare(_f_set(), 2, chunk=1 << 14)


def test_jk_tight_exponents(gpu):
    """Two atoms of the 64-atom recipe (core exponents up to 5000, depth-6
    contractions)."""
    assert _compare(workloads.grid_molecule(natom=2, seed=7), 3)


def test_jk_only_one_of_them(gpu):
    """J alone and K alone (the other accumulator NULL) give the same
    matrices as both together."""
    shells = mol_shells("water.sto-3g")
    D = _density(sum(simint.ncart(s[0]) for s in shells), 4)
    J, K, _ = _gpu_jk(shells, D)
    J1, _, _ = _gpu_jk(shells, D, want=("J",))
    _, K1, _ = _gpu_jk(shells, D, want=("K",))
    assert np.abs(J - J1).max() <= TOL * np.abs(J).max()
    assert np.abs(K - K1).max() <= TOL * np.abs(K).max()


@pytest.mark.parametrize("rows", [0, 1])
def test_jk_exchange_paths(gpu, rows):
    """The exchange updates summed per bra in the local data share (rows 1,
    the default) and added to K quartet by quartet (rows 0) both give the
    oracle's K, on shells up to f with tight exponents."""
    L = simint.lib()
    prev = L.simint_hip_set_jk_rows(rows)
    try:
        assert _compare(_f_set(seed=9) + workloads.grid_molecule(natom=1, seed=9), 5)
    finally:
        L.simint_hip_set_jk_rows(prev)


def test_jk_mixed_plan(gpu):
    """A mixed plan (quartets with small Schwarz bounds on the FP32-VRR
    kernels) digested the same way, on two atoms of the 64-atom recipe: J
    and K within 1e-9 absolute of the oracle's -- each FP32 integral is
    within the mixed path's 1e-10, and a J or K element sums such errors
    weighted by density elements of at most ~0.1."""
    shells = workloads.grid_molecule(natom=2, seed=11)
    sa = normalized_arrays(shells)
    g, _ = oracle.eri_tensor(sa)
    D = _density(g.shape[0], 6)
    Jr, Kr = oracle.jk(g, D)
    J, K, _ = _gpu_jk(shells, D, mixed=1e-4)
    ej, ek = np.abs(J - Jr).max(), np.abs(K - Kr).max()
    print("mixed plan: max |dJ| %.2e, max |dK| %.2e" % (ej, ek))
    assert ej <= 1e-9 and ek <= 1e-9


def test_group_function_arrays_shared_by_threads(gpu):
    """Threads that start digesting one set side by side (the serving
    workers) get one device function array per pair group, with the basis
    function offsets of the group's pairs: an array built twice would be
    freed under the first thread's digestion."""
    import threading

    import torch
    shells = mol_shells("water.6-31gss")
    ss = quartets.ScreenedSet(shells, tol=0.0)
    try:
        f0 = np.r_[0, np.cumsum([simint.ncart(s[0]) for s in shells])]
        nthr = 8
        go = threading.Barrier(nthr)
        got = [None] * nthr

        def run(t):
            s = torch.cuda.Stream()
            go.wait()
            with torch.cuda.stream(s):
                got[t] = [ss._group_fun(g) for g in (ss.groups if t % 2 else ss.groups[::-1])]

        th = [threading.Thread(target=run, args=(t,)) for t in range(nthr)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        for t in range(nthr):
            arrs = got[t] if t % 2 else got[t][::-1]
            for g, a in zip(ss.groups, arrs):
                assert a is g.dfun
                want = f0[np.asarray(g.pairs, dtype=np.int64).reshape(-1)]
                assert np.array_equal(a.cpu().numpy(), want)
    finally:
        ss.free()


def test_jk_rows_beyond_64k_of_lds(gpu):
    """An 80-atom system (1200 basis functions): a (dd| bra's twelve exchange
    rows take 115 KB of the local data share, more than 64 KB and within
    the device's 160 KB; J and K with the rows match those with per-quartet
    atomics (simint_hip_set_jk_rows(0)) to rounding."""
    import torch
    shells = workloads.grid_molecule(natom=80, seed=11)
    ss = quartets.ScreenedSet(shells, tol=1e-12, out_doubles=1 << 26)
    L = simint.lib()
    prev = L.simint_hip_set_jk_rows(1)
    try:
        nbf = ss.nbf()
        assert 12 * nbf * 8 > 96 * 1024
        D = torch.from_numpy(_density(nbf, 11)).cuda()
        out = torch.empty(ss.out_doubles, dtype=torch.float64, device="cuda")
        res = []
        for rows in (1, 0):
            L.simint_hip_set_jk_rows(rows)
            J = torch.zeros((nbf, nbf), dtype=torch.float64, device="cuda")
            K = torch.zeros_like(J)
            ss.run_jk(out.data_ptr(), D.data_ptr(), J.data_ptr(), K.data_ptr())
            simint.jk_symmetrize(nbf, J.data_ptr(), K.data_ptr())
            simint.synchronize()
            res.append((J.cpu().numpy(), K.cpu().numpy()))
        (J1, K1), (J0, K0) = res
        ej = np.abs(J1 - J0).max() / np.abs(J0).max()
        ek = np.abs(K1 - K0).max() / np.abs(K0).max()
        print("nbf %d: rows vs per-quartet atomics, J %.2e K %.2e normwise" % (nbf, ej, ek))
        assert ej <= 1e-13 and ek <= 1e-13
    finally:
        L.simint_hip_set_jk_rows(prev)
        ss.free()
