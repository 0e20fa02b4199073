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
        sa = normalized_arrays(workloads.grid_molecule(natom=r.natom, seed=r.seed))
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
    """A budget that holds one request at a time serialises them: a request
    is admitted on its estimate and then charged what its set-up allocated."""
    reqs = [serving.Request(i, 6, 30 + i, "eri", 0.0) for i in range(3)]
    est = serving.footprint(36, "eri", 1 << 22)
    ss = quartets.ScreenedSet(workloads.grid_molecule(natom=6, seed=30), tol=1e-10, out_doubles=1 << 22)
    act = ss.device_bytes()
    ss.free()
    assert act > 0
    srv = serving.Server(hbm_budget=est + 0.5 * act, slots=4, tol=1e-10, chunk_doubles=1 << 22)
    srv.serve(reqs)
    assert srv.max_concurrent == 1
    assert all(r.error is None for r in reqs)


def test_malformed_and_timed_out_requests_leave_others_intact(gpu):
    """A NaN exponent, an f shell, a zero time limit, an unknown kind and a
    density of the wrong shape end those requests with an error; the
    requests beside them give the same result as alone."""
    good = [serving.Request(0, 8, 51, "eri", 0.0), serving.Request(1, 6, 52, "deriv", 0.0)]
    alone = {}
    for r in good:
        srv = serving.Server(hbm_budget=8 << 30, slots=1, tol=1e-10, chunk_doubles=1 << 22)
        srv.serve([r])
        assert r.error is None, r.error
        alone[r.rid] = (r.digest, r.quartets)
    nan_shells = workloads.grid_molecule(natom=2, seed=53)
    am, xyz, al, co = nan_shells[3]
    nan_shells[3] = (am, xyz, [float("nan")] + list(al[1:]), co)
    f_shells = workloads.grid_molecule(natom=2, seed=54) + [(3, (0.0, 0.0, 0.0), [1.0], [1.0])]
    bad_density = np.ones((3, 3))
    mix = [serving.Request(0, 8, 51, "eri", 0.0), serving.Request(1, 6, 52, "deriv", 0.01),
           serving.Request(2, 2, 0, "eri", 0.0, shells=nan_shells),
           serving.Request(3, 2, 0, "eri", 0.02, shells=f_shells),
           serving.Request(4, 10, 55, "eri", 0.0, timeout=0.0),
           serving.Request(5, 2, 56, "hessian", 0.0),
           serving.Request(6, 2, 57, "jk", 0.0, density=bad_density)]
    srv = serving.Server(hbm_budget=8 << 30, slots=4, tol=1e-10, chunk_doubles=1 << 22)
    srv.serve(mix)
    assert mix[2].error.startswith("rejected") and "exponent" in mix[2].error
    assert mix[3].error.startswith("rejected") and "angular momentum" in mix[3].error
    assert mix[4].error.startswith("timeout")
    assert mix[5].error.startswith("rejected") and "kind" in mix[5].error
    assert mix[6].error.startswith("rejected") and "density" in mix[6].error
    for r in mix[:2]:
        assert r.error is None, r.error
        assert (r.digest, r.quartets) == alone[r.rid], r.rid


def test_small_request_overtakes_large_one(gpu):
    """Tiles are taken across requests, least work left first: a small
    request arriving just after a large one finishes first."""
    big = serving.Request(0, 48, 61, "eri", 0.0)
    small = serving.Request(1, 4, 62, "eri", 0.05)
    srv = serving.Server(hbm_budget=8 << 30, slots=2, tol=1e-10, chunk_doubles=1 << 20)
    srv.serve([big, small])
    assert big.error is None and small.error is None
    print("big", big.t_done - big.arrival, "small", small.t_done - small.arrival, "tiles", srv.tiles_run)
    # the small one arrived while the big one was still being served, was
    # admitted at once and finished first
    assert small.arrival < big.t_done
    assert small.t_admit - small.arrival < 1.0, (small.t_admit, small.arrival)
    assert small.t_done < big.t_done


def test_screened_deriv_pass_vs_oracle(gpu):
    """A first-derivative pass of the screened set (the serving deriv path)
    against the oracle on a sample of its quartets."""
    import torch
    shells = workloads.grid_molecule(natom=3, seed=21)
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


def test_jk_requests_vs_oracle_alone_and_in_mix(gpu):
    """"jk" requests (Coulomb and exchange matrices of a density, the
    consumer a Fock build puts after the integrals) served among ERI and
    derivative requests: J and K match the oracle's full-tensor J and K, and
    the same request alone and in the mix gives the same matrices (up to
    the atomic sums' order).  Screening at 1e-10 drops quartets the oracle
    keeps, so the bar against it is 1e-9 absolute."""
    from test_gpu_jk import _density

    def mk():
        shells_a = workloads.grid_molecule(natom=2, seed=21)
        shells_b = workloads.grid_molecule(natom=3, seed=22)
        rs = [serving.Request(0, 2, 21, "jk", 0.0, shells=shells_a),
              serving.Request(1, 12, 23, "eri", 0.0),
              serving.Request(2, 3, 22, "jk", 0.0, shells=shells_b),
              serving.Request(3, 10, 24, "deriv", 0.0)]
        for r in rs[::2]:
            nbf = sum(simint.ncart(s[0]) for s in r.shells)
            r.density = _density(nbf, r.seed)
        return rs

    alone = {}
    for r in mk()[::2]:
        srv = serving.Server(hbm_budget=8 << 30, slots=1, tol=1e-10, chunk_doubles=1 << 20)
        srv.serve([r])
        assert r.error is None, r.error
        alone[r.rid] = (r.J, r.K)
    mix = mk()
    for i, r in enumerate(mix):
        r.arrival = 0.02 * i
    srv = serving.Server(hbm_budget=8 << 30, slots=4, tol=1e-10, chunk_doubles=1 << 20)
    srv.serve(mix)
    for r in mix:
        assert r.error is None, r.error
    for r in mix[::2]:
        sa = normalized_arrays(r.shells)
        g, _ = oracle.eri_tensor(sa)
        Jr, Kr = oracle.jk(g, r.density)
        ej, ek = np.abs(r.J - Jr).max(), np.abs(r.K - Kr).max()
        print("request %d: |dJ| %.2e |dK| %.2e (max |J| %.2e, |K| %.2e)" % (r.rid, ej, ek, np.abs(Jr).max(),
                                                                            np.abs(Kr).max()))
        assert ej <= 1e-9 and ek <= 1e-9
        J0, K0 = alone[r.rid]
        assert np.abs(r.J - J0).max() <= 1e-12 * np.abs(J0).max()
        assert np.abs(r.K - K0).max() <= 1e-12 * np.abs(K0).max()
