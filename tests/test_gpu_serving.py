"""The serving loop: requests of 10-20 atoms of both kinds admitted while
others run, overlapping on the device; each result (an FP64 digest of every
integral of the request) is the same alone and in the mix; the integrals
served in the mix match the oracle on a sample of every request's quartets;
first-derivative passes over a screened set match the oracle."""
import numpy as np
import pytest

import oracle
import simint
import quartets
import serving
import workloads
from helpers import normalized_arrays, parity_report, parity_report_deriv

pytestmark = pytest.mark.gpu


def _reqs():
    spec = [(10, 11, "eri"), (12, 12, "deriv"), (10, 13, "eri"), (20, 14, "eri"), (10, 15, "deriv"), (14, 16, "eri")]
    return [serving.Request(i, n, s, k, 0.0) for i, (n, s, k) in enumerate(spec)]


def test_same_alone_and_in_mix(gpu):
    alone = {}
    for r in _reqs():
        srv = serving.Server(hbm_budget=8 << 30, slots=1, tol=1e-10, chunk_doubles=1 << 22)
        srv.serve([r])
        assert r.error is None, r.error
        alone[r.rid] = (r.digest, r.quartets)
    mix = _reqs()
    for i, r in enumerate(mix):
        r.arrival = 0.05 * i            # arrive while the first ones run
    srv = serving.Server(hbm_budget=8 << 30, slots=4, tol=1e-10, chunk_doubles=1 << 22)
    srv.serve(mix)
    st = serving.summarize(mix, srv.wall)
    print(st, "max concurrent", srv.max_concurrent)
    assert srv.max_concurrent >= 2
    for r in mix:
        assert r.error is None, r.error
        assert (r.digest, r.quartets) == alone[r.rid], r.rid


def test_served_results_vs_oracle(gpu):
    """Two ERI requests and two first-derivative requests of 10-16 atoms,
    served concurrently; 48 quartets of each, read from the output chunks
    the server streams, against the oracle (ERIs: the parity bar with exact
    values; derivatives: 1e-10 absolute and as accurate as the oracle)."""
    spec = [(16, 41, "eri"), (10, 42, "deriv"), (12, 43, "eri"), (11, 44, "deriv")]
    reqs = [serving.Request(i, n, s, k, 0.02 * i) for i, (n, s, k) in enumerate(spec)]
    srv = serving.Server(hbm_budget=8 << 30, slots=4, tol=1e-10, chunk_doubles=1 << 22, probe=48)
    srv.serve(reqs)
    assert srv.max_concurrent >= 2
    for r in reqs:
        assert r.error is None, r.error
        assert len(r.samples) == 48
        sa = normalized_arrays(workloads.config2(natom=r.natom, seed=r.seed))
        qs = [q for q, _ in r.samples]
        got = np.concatenate([v for _, v in r.samples])
        ref, offs, sizes = oracle.eri(sa, qs, deriv=1 if r.kind == "deriv" else 0)
        if r.kind == "deriv":
            ex, _, _ = oracle.md_eri_deriv1_analytic(sa, qs)
            ok, st = parity_report_deriv(got, ref, offs, sizes, ex)
        else:
            ex, _, _ = oracle.md_eri(sa, qs, reference_boys=True)
            ok, st = parity_report(got, ref, offs, sizes, ex, floor=1e-2)
        print(r.rid, r.natom, r.kind, st)
        assert ok, (r.rid, st)


def test_admission_respects_budget(gpu):
    """A budget that holds one request at a time serialises them."""
    reqs = [serving.Request(i, 6, 30 + i, "eri", 0.0) for i in range(3)]
    fp = serving.footprint(6, "eri", 1 << 22)
    srv = serving.Server(hbm_budget=fp * 1.5, slots=4, tol=1e-10, chunk_doubles=1 << 22)
    srv.serve(reqs)
    assert srv.max_concurrent == 1
    assert all(r.error is None for r in reqs)


def test_screened_deriv_pass_vs_oracle(gpu):
    """A first-derivative pass of the screened set (the serving deriv path)
    against the oracle on a sample of its quartets."""
    import torch
    shells = workloads.config2(natom=3, seed=21)
    ss = quartets.ScreenedSet(shells, tol=1e-10, out_doubles=1 << 22, deriv=1)
    work = torch.empty(ss.deriv_worksize(), dtype=torch.float64, device="cuda")
    out = torch.empty(ss.out_doubles, dtype=torch.float64, device="cuda")
    sa = normalized_arrays(shells)
    rng = np.random.default_rng(0)
    got, qs = [], []
    for c in range(ss.nchunk):
        out.fill_(float("nan"))
        torch.cuda.synchronize()
        ss.run_pass(out.data_ptr(), chunk=c, work_ptr=work.data_ptr())
        simint.synchronize()
        host = out.cpu().numpy()
        for ln in ss.launches:
            if ln.chunk != c:
                continue
            nc = ln.nout // ln.nq
            for i in rng.choice(len(ln.nket), size=min(3, len(ln.nket)), replace=False):
                for k in range(min(int(ln.nket[i]), 2)):
                    bp = ln.bra.pairs[ln.b0 + i]
                    kp = ln.ket.pairs[k]
                    qs.append((bp[0], bp[1], kp[0], kp[1]))
                    s0 = int(ln.off[i]) + k * nc
                    got.append(host[s0:s0 + nc])
    assert len(qs) > 20
    ref, offs, sizes = oracle.eri(sa, qs, deriv=1)
    ex, _, _ = oracle.md_eri_deriv1_analytic(sa, qs)
    ok, st = parity_report_deriv(np.concatenate(got), ref, offs, sizes, ex)
    print(st)
    assert ok, st
    ss.free()
