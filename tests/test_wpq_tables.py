"""The cooperative kernels' recurrence tables (CPU): every thread of every
VRR level reads a descriptor row of that level, each row either a node of
the level (written to its own slot) or padding (zero sources, written to the
sink slot), and every LDS offset a row holds stays inside the work area or
the primitive-constant block; evaluating the rows level by level the way a
thread does reproduces the recurrence graph's own evaluation."""
import itertools
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "simint-generator_amd"))

from emitter import hip_wpq  # noqa: E402
from emitter.rrgraph import I, J, K, L  # noqa: E402

CLASSES = [am for am in itertools.product(range(4), repeat=4) if sum(am) % 3 == 0 or max(am) == 3][:60]


@pytest.mark.parametrize("am", CLASSES + [(3, 3, 3, 3), (2, 2, 2, 2), (1, 0, 0, 0), (0, 0, 0, 0)])
def test_rows_cover_levels_in_bounds(am):
    g, order, starts, desc, tops = hip_wpq.plan(am)
    B, ws, npc = hip_wpq.batch_size(am)
    LPB = hip_wpq.THREADS // B
    rows, first = hip_wpq.byte_desc(desc, starts, B, LPB)
    zero, sink = len(order), len(order) + 2
    assert sink < ws
    work_bytes = ws * B * 8
    for lv in range(1, len(starts) - 1):
        S, E = starts[lv], starts[lv + 1]
        nj = (E - S + LPB - 1) // LPB
        for r in range(LPB):
            for j in range(nj):
                row = first[lv] + r + j * LPB
                assert row < len(rows)
                k = S + r + j * LPB
                d = rows[row]
                if k < E:
                    assert d[5] == k * B * 8
                else:
                    assert d[5] == sink * B * 8 and d[:5] == (zero * B * 8,) * 5 and d[7] == 0
                for off in d[:5]:
                    # pair reads touch off and off + 8B, plus the quartet slot b < B
                    assert off + 8 * B + 8 * (B - 1) < work_bytes
                assert d[5] + 8 * (B - 1) < work_bytes
                assert d[6] + 24 * B + 8 * (B - 1) < npc * B * 8   # constants at +0, 8B, 16B, 24B bytes


@pytest.mark.parametrize("am", [(1, 1, 1, 1), (2, 1, 2, 0), (3, 1, 0, 2), (2, 2, 1, 1)])
def test_rows_evaluate_the_recurrence(am):
    """Random constants and bases; the rows, read as the kernel reads them
    (B = 1, byte offsets / 8), give the same values as evaluating each node
    from the graph's sources and factors directly."""
    g, order, starts, desc, tops = hip_wpq.plan(am)
    rows, first = hip_wpq.byte_desc(desc, starts, 1, hip_wpq.THREADS)
    rng = np.random.default_rng(7)
    c = rng.uniform(-1, 1, hip_wpq.NPC_FIXED)
    nb = starts[1]
    n = len(order) + 3
    work = np.zeros(n)
    work[:nb] = rng.uniform(0.1, 1, nb)
    # direct evaluation from the graph
    slot = {(nd.q, nd.m): i for i, nd in enumerate(order)}
    ref = np.zeros(n)
    ref[:nb] = work[:nb]
    for i, nd in enumerate(order[nb:], start=nb):
        isket = 1 if nd.step in (K, L) else 0
        s = isket * 3 + nd.dir
        kk, aa, ao, oo = c[4 * s], c[4 * s + 1], c[4 * s + 2], c[4 * s + 3]
        src = [ref[slot[x]] if x is not None else 0.0 for x in nd.src]
        f = [x if nd.src[p] is not None else 0 for x, p in zip(nd.fac, (2, 4, 6, 7))]
        t = kk * src[0] + aa * src[1]
        t += f[0] * oo * (src[2] + ao * src[3]) + f[1] * oo * (src[4] + ao * src[5])
        t += f[2] * c[24] * src[6] + f[3] * c[24] * src[7]
        ref[i] = t
    # the rows, level by level
    for lv in range(1, len(starts) - 1):
        S, E = starts[lv], starts[lv + 1]
        new = {}
        for r in range(first[lv], first[lv + 1] if lv + 1 < len(first) else len(rows)):
            d = [x // 8 for x in rows[r][:7]]
            fac = rows[r][7]
            f = [(fac >> sh) & 0xff for sh in (0, 8, 16, 24)]
            kk, aa, ao, oo = c[d[6]], c[d[6] + 1], c[d[6] + 2], c[d[6] + 3]
            t = kk * work[d[0]] + aa * work[d[0] + 1]
            t += f[0] * oo * (work[d[1]] + ao * work[d[1] + 1]) + f[1] * oo * (work[d[2]] + ao * work[d[2] + 1])
            t += f[2] * c[24] * work[d[3]] + f[3] * c[24] * work[d[4]]
            new[d[5]] = t
        for k, t in new.items():
            work[k] = t
        assert work[len(order)] == 0.0 and work[len(order) + 1] == 0.0
    np.testing.assert_allclose(work[:len(order)], ref[:len(order)], rtol=1e-13, atol=0)
