"""End-to-end checks of the two host-facing paths the benchmark and callers
use, against the oracle:

  * the device-resident screened pass (quartets.ScreenedSet.run_pass, the
    path the benchmark's value is measured on), FP64, chunk by chunk, on a
    sample of its quartets located in the output chunks;
  * the drop-in simint_compute_eri called from several host threads at once
    on a shared ket pair set (the reference's tests do this with OpenMP,
    skel/test/test_eri.cpp:120-134): every thread's results match the oracle;
  * the same pass with the primitive Schwarz screen (pairs built with
    SIMINT_SCREEN_SCHWARZ, screen_tol in the primitive loop): every
    integral within K * tol of the unscreened pass (K its primitive
    quartets; each skipped one has a bound below tol), and something is
    skipped at a loose tolerance.
"""
import threading

import numpy as np
import pytest

import oracle
import quartets
import simint
import workloads
from helpers import normalized_arrays, parity_report

pytestmark = pytest.mark.gpu


def test_screened_pass_fp64_vs_oracle(gpu):
    import torch
    shells = workloads.grid_molecule(natom=12, seed=5)
    ss = quartets.ScreenedSet(shells, tol=1e-12, out_doubles=1 << 24)
    sa = normalized_arrays(shells)
    sample = ss.located_sample(600, seed=1)
    bychunk = {}
    for c, off, nc, q in sample:
        bychunk.setdefault(c, []).append((off, nc, q))
    out = torch.empty(ss.out_doubles, dtype=torch.float64, device="cuda")
    got, qs = [], []
    for c in sorted(bychunk):
        out.fill_(float("nan"))
        torch.cuda.synchronize()
        ss.run_pass(out.data_ptr(), chunk=c)
        simint.synchronize()
        for off, nc, q in bychunk[c]:
            got.append(out[off:off + nc].cpu().numpy())
            qs.append(q)
    ss.free()
    classes = {tuple(shells[i][0] for i in q) for q in qs}
    assert len(classes) >= 20, classes
    ref, offs, sizes = oracle.eri(sa, qs)
    ex, _, _ = oracle.md_eri(sa, qs, reference_boys=True)
    # the 64-atom recipe: the relative bar holds from 1e-2 up
    ok, st = parity_report(np.concatenate(got), ref, offs, sizes, ex, floor=1e-2)
    print("screened FP64 pass", st)
    assert ok, st


def test_dropin_concurrent_host_threads(gpu):
    """Four threads, each computing (P_t | Q) for its own bra pairs P_t and
    one shared ket set Q, eight times in a row, all at once."""
    shells = workloads.grid_molecule(natom=4, seed=9)
    basis = simint.Basis(shells)
    sa = normalized_arrays(shells)
    byam = {}
    for i, s in enumerate(shells):
        byam.setdefault(s[0], []).append(i)
    # ket set: (d p) pairs shared by all threads; bras: per-thread classes
    C, D = byam[2], byam[1]
    Q = simint.Pairs(basis, C, D)
    bras = [(byam[0], byam[0]), (byam[1], byam[0]), (byam[1], byam[1]), (byam[2], byam[2])]
    refs = []
    for A, B in bras:
        qs = [(a, b, c, d) for a in A for b in B for c in C for d in D]
        refs.append((qs,) + tuple(oracle.eri(sa, qs)))
    errors = []

    def work(t):
        try:
            A, B = bras[t]
            P = simint.Pairs(basis, A, B)
            qs, ref, offs, sizes = refs[t]
            for _ in range(8):
                rc, got = simint.compute_eri(P, Q)
                if rc != len(qs):
                    errors.append((t, "rc", rc, simint.last_error()))
                    return
                ok, st = parity_report(got, ref, offs, sizes)
                if not ok:
                    errors.append((t, st))
                    return
            P.free()
        except Exception as e:   # reported by the main thread
            errors.append((t, repr(e)))

    th = [threading.Thread(target=work, args=(t,)) for t in range(len(bras))]
    for x in th:
        x.start()
    for x in th:
        x.join()
    Q.free()
    basis.free()
    assert not errors, errors


@pytest.mark.parametrize("tol", [1e-12, 1e-9])
def test_primitive_schwarz_screen_within_bound(gpu, tol):
    import torch
    shells = workloads.grid_molecule(natom=6, seed=7)
    plain = quartets.ScreenedSet(shells, tol=1e-12, out_doubles=1 << 24)
    scr = quartets.ScreenedSet(shells, tol=1e-12, out_doubles=1 << 24, prim_tol=tol)
    try:
        sig = [(ln.cls, ln.b0, ln.nq, ln.chunk) for ln in plain.launches]
        assert sig == [(ln.cls, ln.b0, ln.nq, ln.chunk) for ln in scr.launches]
        a = torch.empty(plain.out_doubles, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        worst, moved = 0.0, 0.0
        for c in range(plain.nchunk):
            plain.run_pass(a.data_ptr(), chunk=c)
            scr.run_pass(b.data_ptr(), chunk=c)
            simint.synchronize()
            for lp, ls in zip(plain.chunk_launches(c), scr.chunk_launches(c)):
                o = int(lp.off[0])
                x, y = a[o:o + lp.nout], b[o:o + lp.nout]
                K = lp.bra.nprim * lp.ket.nprim
                d = float((x - y).abs().max())
                bound = K * tol + 1e-14 * max(1.0, float(x.abs().max()))
                worst = max(worst, d / bound)
                moved = max(moved, d)
        print("prim_tol %g: max change %.3e, worst change / bound %.3f" % (tol, moved, worst))
        assert worst <= 1.0
        if tol >= 1e-9:
            assert moved > 1e-12   # the screen skipped primitive quartets
    finally:
        plain.free()
        scr.free()
