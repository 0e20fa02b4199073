"""The screened-set launch planner (CPU): every screened quartet planned
exactly once, every launch's output inside the chunk buffer -- including a
bra whose own kets exceed the buffer (cut into ket ranges) -- and the
mixed-precision split covering each range without overlap; the options
of the primitive screen and of serving's derivative tiles checked without a
device."""
import json
import os

import numpy as np
import pytest

import quartets

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INFO = os.path.join(ROOT, "simint-generator_amd", "csrc", "gen", "classes.json")


class _Group:
    def __init__(self, cls, nprim, bounds):
        self.cls, self.nprim = cls, nprim
        self.bounds = np.sort(np.asarray(bounds, dtype=np.float64))[::-1].copy()
        self.pairs = [(i, i) for i in range(len(bounds))]


def _planned(groups, out_doubles, mixed=None, deriv=0, tol=1e-12):
    ss = quartets.ScreenedSet.__new__(quartets.ScreenedSet)
    ss.tol, ss.mixed, ss.deriv, ss.out_doubles, ss.groups = tol, mixed, deriv, out_doubles, groups
    ss.front_end = "host"
    ss.info = json.load(open(INFO))
    ss._plan()
    return ss


@pytest.mark.skipif(not os.path.exists(INFO), reason="generated class table not built")
@pytest.mark.parametrize("mixed,deriv", [(None, 0), (1e-4, 0), (None, 1)])
def test_plan_covers_each_quartet_once_within_buffer(mixed, deriv):
    rng = np.random.default_rng(0)
    groups = [_Group((2, 2), 1, 10 ** rng.uniform(-14, 1, 300)),
              _Group((2, 1), 3, 10 ** rng.uniform(-14, 1, 500)),
              _Group((1, 0), 1, 10 ** rng.uniform(-14, 1, 200))]
    cap = 40000 if not deriv else 300000      # (dd|dd): 1296 doubles, x12 for derivatives
    ss = _planned(groups, cap, mixed, deriv)
    assert any(ln.lo is not None for ln in ss.launches)          # ranges were needed
    seen = {}
    for ln in ss.launches:
        nc = ln.nout // ln.nq
        lo = np.zeros_like(ln.nket) if ln.lo is None else ln.lo
        assert (ln.off + (ln.nket - lo) * nc).max() <= cap
        for i in range(len(ln.nket)):
            for k in range(int(lo[i]), int(ln.nket[i])):
                key = (id(ln.bra), id(ln.ket), ln.b0 + i, k)
                assert key not in seen
                seen[key] = (ln.chunk, int(ln.off[i]) + (k - int(lo[i])) * nc)
    # the screened set itself
    want = 0
    for gi, gb in enumerate(groups):
        for gk in groups[:gi + 1]:
            for b, vb in enumerate(gb.bounds):
                n = int((vb * gk.bounds >= 1e-24).sum())
                want += min(n, b + 1) if gk is gb else n
    assert len(seen) == want == ss.total_quartets
    # no two quartets share an output slot within a chunk
    slots = set(seen.values())
    assert len(slots) == len(seen)


@pytest.mark.skipif(not os.path.exists(INFO), reason="generated class table not built")
def test_plan_rejects_a_quartet_larger_than_the_buffer():
    groups = [_Group((2, 2), 1, [1.0, 0.5])]
    with pytest.raises(RuntimeError):
        _planned(groups, 1000, deriv=1)


@pytest.mark.parametrize("kw", [dict(prim_tol=-1e-12), dict(prim_tol=1e-12, front_end="device"),
                                dict(prim_tol=1e-12, deriv=1)])
def test_primitive_screen_options_rejected_before_any_device_work(kw):
    """The primitive Schwarz screen takes a tolerance >= 0, the host front
    end (its pairs carry SIMINT_SCREEN_SCHWARZ bounds) and plain ERIs; other
    combinations are refused before the set touches the library."""
    with pytest.raises(ValueError, match="prim_tol"):
        quartets.ScreenedSet([(0, (0.0, 0.0, 0.0), [1.0], [1.0])], **kw)


def test_serving_derivative_tiles_are_larger():
    """Derivative requests are cut into tiles four times the ERI tile unless
    a size is given (their calls carry per-term launch costs)."""
    import serving
    assert serving.Server(chunk_doubles=1 << 20).deriv_chunk == 4 << 20
    assert serving.Server(chunk_doubles=1 << 20, deriv_chunk_doubles=1 << 21).deriv_chunk == 1 << 21


@pytest.mark.parametrize("cls,div", [((0, 0, 0, 0), 1), ((1, 0, 0, 0), 1), ((2, 0, 2, 0), 1),
                                     ((1, 1, 0, 0), 400), ((2, 0, 1, 1), 400), ((2, 1, 2, 0), 400),
                                     ((1, 1, 1, 1), 1e5), ((2, 2, 2, 1), 1e5)])
def test_mixed_threshold_by_hrr_sides(cls, div):
    """The FP32-VRR threshold: the base for classes without HRR, /400 with
    one HRR side (both shells of a side above s), /1e5 with both."""
    assert quartets.mixed_threshold(cls, 1e-4) == pytest.approx(1e-4 / div, rel=1e-15)


def test_list_split_tuning_applied_per_class(tmp_path):
    """Lanes per quartet of a class's list launches: 0 (the library's
    automatic choice) unless a tuning file names a count, which
    apply_tuning takes (1, 2, 4, 8 or 16 only)."""
    import json

    import simint
    saved = simint._LIST_SPLIT
    try:
        simint._LIST_SPLIT = None
        pkg = json.load(open(os.path.join(ROOT, "simint-generator_amd", "tuning.json")))
        assert simint.list_split_of((1, 0, 0, 0)) == pkg.get("list_split", {}).get("p_s_s_s", 0)
        f = tmp_path / "t.json"
        f.write_text(json.dumps({"list_split": {"p_s_s_s": 4, "d_d_s_s": 16}}))
        assert simint.apply_tuning(str(f)) == 2
        assert simint.list_split_of((1, 0, 0, 0)) == 4 and simint.list_split_of((2, 2, 0, 0)) == 16
        assert simint.list_split_of((1, 1, 1, 1)) == 0
        f.write_text(json.dumps({"list_split": {"p_s_s_s": 3}}))
        with pytest.raises(ValueError, match="lanes"):
            simint.apply_tuning(str(f))
    finally:
        simint._LIST_SPLIT = saved
