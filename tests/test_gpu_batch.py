"""Quartet-list path (simint_hip_compute_eri_list / _deriv1_list, graph
capture) against the CPU oracle: a mixed stream over every class up to f
with contraction depths 1-6, the shape of the batched configuration."""
import itertools

import numpy as np
import pytest

import batch
import oracle
import simint
import workloads
from helpers import normalized_arrays, parity_report, parity_report_deriv

pytestmark = pytest.mark.gpu


def _pool(seed=3, per_am=12, cube=6.0, maxk=6):
    """Shells of the batched recipe (depth U{1..6}, exponents log-uniform in
    [0.05, 1e4]) in a small cube, so that most integrals are not negligible."""
    r = workloads.SplitMix64(seed)
    shells = []
    for am in range(4):
        for _ in range(per_am):
            k = 1 + r.index(maxk)
            xyz = (r.uniform(0, cube), r.uniform(0, cube), r.uniform(0, cube))
            al = [r.loguniform(0.05, 1e4) for _ in range(k)]
            co = [r.uniform(0.1, 1.0) for _ in range(k)]
            shells.append((am, xyz, al, co))
    return shells


def _stream(shells, per_class=3, seed=0):
    """per_class random quartets of every class (la, lb, lc, ld) <= f, plus
    same-shell pairs and a repeated quartet, in shuffled stream order."""
    rng = np.random.default_rng(seed)
    byam = {}
    for n, s in enumerate(shells):
        byam.setdefault(s[0], []).append(n)
    qs = []
    for am in itertools.product(range(4), repeat=4):
        for _ in range(per_class):
            qs.append([rng.choice(byam[a]) for a in am])
    qs.append([0, 0, 0, 0])
    qs.append([byam[2][0], byam[2][0], byam[1][1], byam[1][1]])
    qs.append(list(qs[5]))
    qs = np.array(qs, dtype=np.int64)
    return qs[rng.permutation(len(qs))]


def test_list_mixed_stream_all_classes(gpu):
    import torch
    shells = _pool()
    basis = simint.Basis(shells)
    sa = normalized_arrays(shells)
    qs = _stream(shells)
    b = batch.QuartetBatch(basis, qs)
    assert len(b.groups) == 256
    out = torch.full((b.out_doubles,), float("nan"), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()   # the fill (torch's stream) ends before the library stream writes
    b.run(out.data_ptr())
    simint.synchronize()
    got = out.cpu().numpy()
    ql = [tuple(x) for x in qs]
    ref, offs, sizes = oracle.eri(sa, ql)
    assert np.array_equal(offs, b.offsets) and np.array_equal(sizes, b.sizes)
    ex, _, _ = oracle.md_eri(sa, ql, reference_boys=True)
    ok, st = parity_report(got, ref, offs, sizes, ex)
    print(st)
    assert ok, st
    b.free()


@pytest.mark.parametrize("family", [1, 2])
def test_list_both_families(gpu, family):
    """Each kernel family's list variant on the classes where it is built."""
    import torch
    shells = _pool(seed=4, per_am=6)
    basis = simint.Basis(shells)
    sa = normalized_arrays(shells)
    qs = _stream(shells, per_class=2, seed=1)
    L = simint.lib()
    prev = {}
    for am in itertools.product(range(4), repeat=4):
        prev[am] = L.simint_hip_class_list_kernel(*am)
        L.simint_hip_set_class_list_kernel(*am, family)
    try:
        b = batch.QuartetBatch(basis, qs)
        out = torch.zeros(b.out_doubles, dtype=torch.float64, device="cuda")
        b.run(out.data_ptr())
        simint.synchronize()
        got = out.cpu().numpy()
    finally:
        for am, f in prev.items():
            L.simint_hip_set_class_list_kernel(*am, f)
    ql = [tuple(x) for x in qs]
    ref, offs, sizes = oracle.eri(sa, ql)
    ex, _, _ = oracle.md_eri(sa, ql, reference_boys=True)
    ok, st = parity_report(got, ref, offs, sizes, ex)
    assert ok, st
    b.free()


@pytest.mark.parametrize("lanes", [1, 2, 4, 8, 16])
def test_list_lanes_per_quartet(gpu, lanes):
    """Lane-per-quartet list kernels with 1, 2, 4, 8 or 16 lanes per quartet
    (each lane a strided share of the primitive quartets, shares added across
    the lanes) against the oracle on the all-classes stream."""
    import torch
    shells = _pool(seed=6, per_am=4, maxk=4)
    basis = simint.Basis(shells)
    sa = normalized_arrays(shells)
    qs = _stream(shells, per_class=1, seed=3)
    L = simint.lib()
    prev = L.simint_hip_set_list_split(lanes)
    try:
        b = batch.QuartetBatch(basis, qs)
        out = torch.zeros(b.out_doubles, dtype=torch.float64, device="cuda")
        b.run(out.data_ptr())
        simint.synchronize()
        got = out.cpu().numpy()
    finally:
        L.simint_hip_set_list_split(prev)
    ql = [tuple(x) for x in qs]
    ref, offs, sizes = oracle.eri(sa, ql)
    ex, _, _ = oracle.md_eri(sa, ql, reference_boys=True)
    ok, st = parity_report(got, ref, offs, sizes, ex)
    print(st)
    assert ok, st
    b.free()


def test_list_graph_replay(gpu):
    """The batch captured into one graph: every replay (with eager work and
    synchronisation between replays) equals the eager result bit for bit."""
    import torch
    shells = _pool(seed=5, per_am=8)
    basis = simint.Basis(shells)
    qs = _stream(shells, per_class=2, seed=2)
    b = batch.QuartetBatch(basis, qs)
    eager = torch.zeros(b.out_doubles, dtype=torch.float64, device="cuda")
    b.run(eager.data_ptr())
    simint.synchronize()
    for streams in (None, 4):
        out = torch.zeros(b.out_doubles, dtype=torch.float64, device="cuda")
        g = b.capture(out.data_ptr(), streams=streams)
        for rep in range(3):
            out.zero_()
            torch.cuda.synchronize()
            g.launch()
            simint.synchronize()
            assert torch.equal(out, eager), (streams, rep)
            b.run(eager.data_ptr())          # eager work between replays
            simint.synchronize()
        g.free()
        # eager, spread over side streams
        out.zero_()
        torch.cuda.synchronize()
        b.run(out.data_ptr(), streams=streams)
        simint.synchronize()
        torch.cuda.synchronize()
        assert torch.equal(out, eager), streams
    b.free()


@pytest.mark.parametrize("maxam", [2, 3])
def test_deriv1_list(gpu, maxam):
    """First derivatives through the list path against the oracle: every
    class up to d (the (a+1) classes of d are f), and 48 classes with an f
    shell (their (a+1) classes have a g shell)."""
    import torch
    shells = [s for s in _pool(seed=6, per_am=6, maxk=3) if s[0] <= maxam]
    basis = simint.Basis(shells)
    sa = normalized_arrays(shells)
    rng = np.random.default_rng(7)
    byam = {}
    for n, s in enumerate(shells):
        byam.setdefault(s[0], []).append(n)
    if maxam == 2:
        classes = [am for am in itertools.product(range(3), repeat=4) for _ in range(2)]
    else:
        fc = [am for am in itertools.product(range(4), repeat=4) if 3 in am]
        classes = [fc[i] for i in rng.choice(len(fc), size=48, replace=False)]
    qs = np.array([[rng.choice(byam[a]) for a in am] for am in classes], dtype=np.int64)
    b = batch.QuartetBatch(basis, qs, deriv=1)
    out = torch.zeros(b.out_doubles, dtype=torch.float64, device="cuda")
    b.run(out.data_ptr())
    simint.synchronize()
    got = out.cpu().numpy()
    ql = [tuple(x) for x in qs]
    ref, offs, sizes = oracle.eri(sa, ql, deriv=1)
    ex, _, _ = oracle.md_eri_deriv1_analytic(sa, ql)
    ok, st = parity_report_deriv(got, ref, offs, sizes, ex)
    print(st)
    assert ok, st
    b.free()
