"""The benchmark's bookkeeping, on the CPU: the plan shares a timed step
covers, the measured traffic read from the committed rocprofv3 summary,
the FP64/FP32 flop split of a mixed launch and the whole-job roofline
time (each launch at the slower of its flop time and its byte time)."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


class _G:
    def __init__(self, nprim):
        self.nprim = nprim


class _L:
    def __init__(self, nq, cls=(1, 0, 0, 0), precision=0, nprim=(2, 3), flops=None, nc=3):
        self.nq, self.cls, self.precision = nq, cls, precision
        self.bra, self.ket = _G(nprim[0]), _G(nprim[1])
        self.flops = float(nq) * 100.0 if flops is None else flops
        self.nout = nq * nc


def test_shares_cover_the_plan_once_in_order():
    ls = [_L(n) for n in (5, 1, 40, 7, 7, 7, 30, 2, 1, 100)]
    for k in (1, 2, 3, 4, 7):
        sh = bench.shares(ls, k)
        assert len(sh) == k
        flat = [x for s in sh for x in s]
        assert flat == ls                              # every launch once, plan order kept
        tot = sum(l.nq for l in ls)
        big = max(l.nq for l in ls)
        for s in sh:                                    # about tot/k each, within one launch
            assert sum(l.nq for l in s) <= tot / k + big


def test_measured_traffic_comes_from_the_committed_summary():
    assert bench.measured_traffic("eri_lpqsf8_d_d_d_p") == pytest.approx(6.70793e9)
    assert bench.measured_traffic("eri_no_such_kernel") is None


def test_flop_split_of_mixed_and_fp64_launches():
    info = {"p_s_s_s": {"f_vrr": 6.0}}
    fp64 = _L(10, precision=0, flops=5000.0)
    assert bench.launch_split_flops(info, fp64) == (5000.0, 0.0)
    mixed = _L(10, precision=1, flops=5000.0, nprim=(2, 3))
    f64, f32 = bench.launch_split_flops(info, mixed)
    assert f32 == 10 * 6 * 6.0 and f64 + f32 == 5000.0


def test_whole_job_prices_each_launch_at_its_bound():
    info = {"p_s_s_s": {"f_vrr": 6.0}}
    flop_bound = _L(1000, flops=1e12, nc=3)             # 1e12 flops vs 24 kB
    byte_bound = _L(10 ** 9, flops=1.0, nc=3)           # 1 flop-ish vs 24 GB
    res = bench.whole_job([(flop_bound, 1.0), (byte_bound, 1.0)], info, wall=2.0)
    t = 1e12 / bench.PEAK_FP64 + 8.0 * 3 * 10 ** 9 / bench.PEAK_HBM
    assert res["roofline_time_s"] == pytest.approx(t)
    assert res["frac"] == pytest.approx(t / 2.0)


def test_profile_commands_run_the_committed_bench():
    """Every `bench.py ...` command a profile summary gives parses with the
    committed bench's command line (a summary must say how to make it again
    with committed files)."""
    import re
    import shlex
    pdir = os.path.join(ROOT, "profiles")
    seen = 0
    for f in sorted(os.listdir(pdir)):
        for line in open(os.path.join(pdir, f)):
            for m in re.finditer(r"bench\.py((?:\s+--?[\w-]+(?:\s+[^\s`-][^\s`]*)?)*)", line):
                args = shlex.split(m.group(1))
                bench.make_parser().parse_args(args)      # exits (SystemExit) on an unknown flag
                seen += 1
    assert seen >= 5
