"""Device-side front end (simint_hip_pairs_from_shells, _pair_bounds,
_order_pairs, _screen_counts): shell pairs filled on the GPU against the
oracle's restatement of the host fill (skel/simint/shell/shell.c:425-582,
shell_screen.c:226-270), and a screened quartet set set up entirely on the
device against the host-side set up and against the oracle."""
import numpy as np
import pytest

import oracle
import quartets
import simint
import workloads
from helpers import normalized_arrays, parity_report

pytestmark = pytest.mark.gpu

ULP = np.finfo(np.float64).eps


def _ulps(a, b):
    d = np.abs(a - b)
    return np.max(d / np.maximum(np.abs(b), np.finfo(np.float64).tiny) / ULP) if d.size else 0.0


@pytest.mark.parametrize("screen", [simint.SCREEN_NONE, simint.SCREEN_FASTSCHWARZ])
def test_device_pair_fill_matches_oracle(gpu, screen):
    """Every ordered shell pair of one atom of the 64-atom recipe plus an f
    shell, class by class: every field bit-identical to the oracle's fill
    (same arithmetic, no fused multiply-adds), except the two behind
    exp()/pow() -- the prefactor and the screening value -- which may differ
    in the last two bits; the descending sort by screening value gives the
    same primitive order."""
    shells = workloads.grid_molecule(natom=1, seed=5) + [(3, (0.4, -0.3, 0.2), [3.0, 0.6, 0.15], [0.3, 0.6, 0.4])]
    basis = simint.Basis(shells)
    sa = normalized_arrays(shells)
    ds = simint.DeviceShells(basis)
    n = len(shells)
    byclass = {}
    for i in range(n):
        for j in range(n):
            byclass.setdefault((shells[i][0], shells[j][0]), []).append((i, j))
    exact = ["x", "y", "z", "PA_x", "PA_y", "PA_z", "PB_x", "PB_y", "PB_z", "alpha", "alpha2", "beta2"]
    worst = {"prefac": 0.0, "screen": 0.0}
    for cls, pairs in sorted(byclass.items()):
        P = simint.DevicePairs.from_shells(ds, pairs, cls, screen)
        got = P.export()
        P.free()
        start = 0
        for k, (i, j) in enumerate(pairs):
            ref = oracle.fill_pair(sa, i, j, screen)
            m = len(ref["x"])
            assert got["nprim12"][k] == m
            for f in exact:
                np.testing.assert_array_equal(got[f][start:start + m], ref[f], err_msg="%s %s" % (f, (i, j)))
            np.testing.assert_array_equal(got["AB"][:, k], ref["AB"])
            for f in ("prefac", "screen") if screen else ("prefac",):
                u = _ulps(got[f][start:start + m], ref[f])
                worst[f] = max(worst[f], u)
                assert u <= 2.0, (f, (i, j), u)
            start += m
    print("prefactor / screening value worst ulps", worst)
    ds.free()
    basis.free()


def test_device_front_end_screened_set(gpu):
    """A 12-atom set of the 64-atom recipe set up on the device (pairs,
    bounds, sort, counts) holds the same quartets as the host-side set up,
    and its integrals match the oracle on a sample."""
    import torch
    shells = workloads.grid_molecule(natom=12, seed=7)
    host = quartets.ScreenedSet(shells, tol=1e-12, out_doubles=1 << 24)
    dev = quartets.ScreenedSet(shells, tol=1e-12, out_doubles=1 << 24, front_end="device")
    print("set-up s: host %.3f, device %.3f" % (host.setup_s, dev.setup_s))
    # bounds agree to rounding, so the surviving sets agree to within the
    # quartets whose bound product sits on the threshold
    assert abs(dev.total_quartets - host.total_quartets) <= 1e-6 * host.total_quartets
    hb = {p: b for g in host.groups for p, b in zip(g.pairs, g.bounds)}
    db = {p: b for g in dev.groups for p, b in zip(g.pairs, g.bounds)}
    assert hb.keys() == db.keys()
    # the two fills differ in the last bit or two of the prefactor (exp on
    # the device and on the host); the HRR of far-apart pairs with tight
    # exponents amplifies that by up to ~1e4 (measured 4e-12 relative on
    # (dp|dp) diagonals), far below what moves a screening decision
    rel = sorted(((abs(hb[p] - db[p]) / max(hb[p], 1e-300), p, hb[p], db[p]) for p in hb), reverse=True)
    print("worst bound differences", rel[:6])
    assert rel[0][0] <= 1e-10
    for g in dev.groups:
        assert np.all(np.diff(g.bounds) <= 0)
    host.free()
    sa = normalized_arrays(shells)
    sample = dev.located_sample(400, seed=3)
    bychunk = {}
    for c, off, nc, q in sample:
        bychunk.setdefault(c, []).append((off, nc, q))
    out = torch.empty(dev.out_doubles, dtype=torch.float64, device="cuda")
    got, qs = [], []
    for c in sorted(bychunk):
        out.fill_(float("nan"))
        torch.cuda.synchronize()
        dev.run_pass(out.data_ptr(), chunk=c)
        simint.synchronize()
        for off, nc, q in bychunk[c]:
            got.append(out[off:off + nc].cpu().numpy())
            qs.append(q)
    dev.free()
    ref, offs, sizes = oracle.eri(sa, qs)
    ex, _, _ = oracle.md_eri(sa, qs, reference_boys=True)
    ok, st = parity_report(np.concatenate(got), ref, offs, sizes, ex, floor=1e-2)
    print("device front end pass", st)
    assert ok, st


def test_screen_counts_kernel(gpu):
    """The binary-search count kernel against numpy on random sorted bounds,
    both for two groups and within one group (at most b + 1 kets)."""
    import torch
    rng = np.random.default_rng(2)
    for nb, nk in [(1, 1), (37, 1000), (1000, 37), (513, 513)]:
        bb = np.exp(rng.uniform(-30, 2, nb))
        kb = np.sort(np.exp(rng.uniform(-30, 2, nk)))[::-1].copy()
        bb[0] = 0.0
        tol = 1e-10
        for same in (False, True):
            if same and nb != nk:
                continue
            b2 = kb.copy() if same else bb
            out = torch.empty(nb, dtype=torch.int32, device="cuda")
            db = torch.tensor(b2, device="cuda")
            dk = torch.tensor(kb, device="cuda")
            simint.screen_counts(nb, db.data_ptr(), nk, dk.data_ptr(), same, tol, out.data_ptr())
            simint.synchronize()
            with np.errstate(divide="ignore"):
                need = np.where(b2 > 0, tol * tol / b2, np.inf)
            want = np.searchsorted(-kb, -need, side="right")
            if same:
                want = np.minimum(want, np.arange(1, nb + 1))
            np.testing.assert_array_equal(out.cpu().numpy(), want)
