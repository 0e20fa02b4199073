"""Mixed precision (FP32 VRR/HRR, FP64 Boys, prefactors and contraction)
for quartets whose Schwarz bound is below the threshold: the absolute error
against the FP64 path and against the oracle stays within 1e-10, class by
class; FP64 quartets of a mixed pass are unchanged bit for bit; ket ranges
(simint_hip_compute_eri_range) place every quartet where the prefix path
does."""
import numpy as np
import pytest

import oracle
import quartets
import simint
import workloads
from helpers import normalized_arrays, parity_report

pytestmark = pytest.mark.gpu

MIXED_BOUND = 1e-4
ABS_BAR = 1e-10


def _chunk_outputs(ss, chunk, out):
    """Run one output chunk of a screened set; returns {launch index: (lo,
    hi) of the chunk region it wrote} and the buffer on the host."""
    import torch
    out.fill_(float("nan"))
    torch.cuda.synchronize()   # fill done before the library stream writes
    ss.run_pass(out.data_ptr(), chunk=chunk)
    simint.synchronize()
    return out.cpu().numpy()


def test_mixed_vs_fp64_and_oracle(gpu):
    import torch
    shells = workloads.grid_molecule(natom=8, seed=2)
    ss64 = quartets.ScreenedSet(shells, tol=1e-12, out_doubles=1 << 24)
    ssm = quartets.ScreenedSet(shells, tol=1e-12, out_doubles=1 << 24, mixed=MIXED_BOUND)
    assert ss64.nchunk == ssm.nchunk and ss64.total_quartets == ssm.total_quartets
    nmixed = sum(ln.nq for ln in ssm.launches if ln.precision == 1)
    assert nmixed > 0
    print("FP32 quartets: %d of %d" % (nmixed, ssm.total_quartets))
    a = torch.empty(ss64.out_doubles, dtype=torch.float64, device="cuda")
    b = torch.empty(ssm.out_doubles, dtype=torch.float64, device="cuda")
    worst = {}
    for c in range(ssm.nchunk):
        x = _chunk_outputs(ss64, c, a)
        y = _chunk_outputs(ssm, c, b)
        for ln in ssm.launches:
            if ln.chunk != c:
                continue
            nc = simint.ncart(ln.cls[0]) * simint.ncart(ln.cls[1]) * simint.ncart(ln.cls[2]) * simint.ncart(ln.cls[3])
            lo = ln.lo if ln.lo is not None else np.zeros_like(ln.nket)
            for i in range(len(ln.nket)):
                n = int(ln.nket[i] - lo[i]) * nc
                if n <= 0:
                    continue
                s0 = int(ln.off[i])
                d = np.abs(y[s0:s0 + n] - x[s0:s0 + n])
                assert np.all(np.isfinite(y[s0:s0 + n]))
                if ln.precision == 0:
                    assert np.array_equal(y[s0:s0 + n], x[s0:s0 + n])
                else:
                    worst[ln.cls] = max(worst.get(ln.cls, 0.0), float(d.max()))
    print("max |mixed - fp64| per class:", {quartets.simint_class_name(k): "%.1e" % v for k, v in worst.items()})
    assert max(worst.values()) <= ABS_BAR
    # the FP32 quartets against the oracle directly (a sample)
    sa = normalized_arrays(shells)
    qs = ssm.sample(400, seed=1, precision=1)
    basis = ssm.basis
    got = []
    for q in qs:
        P = simint.Pairs(basis, [q[0]], [q[1]])
        Q = simint.Pairs(basis, [q[2]], [q[3]])
        Pd, Qd = simint.DevicePairs(P), simint.DevicePairs(Q)
        nc = simint.ncart(shells[q[0]][0]) * simint.ncart(shells[q[1]][0]) * simint.ncart(shells[q[2]][0]) \
            * simint.ncart(shells[q[3]][0])
        o = torch.full((nc,), float("nan"), dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()   # fill done before the library stream writes
        simint.compute_range(Pd, Qd, 0, None, [1], [0], o.data_ptr(), 0.0, 1)
        simint.synchronize()
        got.append(o.cpu().numpy())
        Pd.free(); Qd.free(); P.free(); Q.free()
    got = np.concatenate(got)
    ref, offs, sizes = oracle.eri(sa, qs)
    err = np.abs(got - ref).max()
    print("sampled FP32 quartets: max |mixed - oracle| = %.2e" % err)
    assert err <= ABS_BAR
    ss64.free()
    ssm.free()


def test_range_matches_prefix(gpu):
    """Ket ranges [lo, hi) per bra in FP64: the same values at the same
    places as the prefix call over [0, hi)."""
    import torch
    shells = workloads.grid_molecule(natom=3, seed=4)
    basis = simint.Basis(shells)
    sa = normalized_arrays(shells)
    byam = {}
    for n, s in enumerate(shells):
        byam.setdefault(s[0], []).append(n)
    rng = np.random.default_rng(3)
    for (la, lb, lc, ld) in [(1, 0, 1, 0), (2, 1, 1, 1), (0, 0, 0, 0), (2, 2, 1, 0)]:
        A, B, C, D = byam[la], byam[lb], byam[lc], byam[ld]
        Pp, Qp = simint.Pairs(basis, A, B), simint.Pairs(basis, C, D)
        P, Q = simint.DevicePairs(Pp), simint.DevicePairs(Qp)
        nc = simint.ncart(la) * simint.ncart(lb) * simint.ncart(lc) * simint.ncart(ld)
        hi = rng.integers(0, Q.n + 1, size=P.n)
        lo = np.minimum(rng.integers(0, 70, size=P.n), hi)
        cnt = hi - lo
        off = np.zeros(P.n, dtype=np.int64)
        off[1:] = np.cumsum(cnt[:-1]) * nc
        out = torch.full((int(cnt.sum()) * nc + 1,), float("nan"), dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()   # fill done before the library stream writes
        for prec in (0, 1):
            simint.compute_range(P, Q, 0, lo, hi, off, out.data_ptr(), 0.0, prec)
            simint.synchronize()
            got = out.cpu().numpy()[:int(cnt.sum()) * nc]
            pairsP = [(a, b) for a in A for b in B]
            pairsQ = [(c, d) for c in C for d in D]
            qs = [(*pairsP[i], *pairsQ[k]) for i in range(P.n) for k in range(lo[i], hi[i])]
            if not qs:
                continue
            ref, offs, sizes = oracle.eri(sa, qs)
            if prec == 0:
                ok, st = parity_report(got, ref, offs, sizes)
                assert ok, ((la, lb, lc, ld), st)
            else:
                assert np.abs(got - ref).max() <= 1e-6 * max(1.0, np.abs(ref).max())
        P.free(); Q.free(); Pp.free(); Qp.free()
