"""Generated lane-per-quartet and split kernels, checked on the CPU from
their source text: every contracted input is accumulated exactly once per
primitive quartet, and the ket-top split kernels' local-data-share use (the
per-primitive-quartet setup cache, the output rows and their offsets) stays
within one wavefront's share of a CU, with every cache access inside the
declared array and guarded by the same slot count on both sides; the
refined reciprocal square root of p+q is used only where no HRR transfer
can amplify its few ulp; every group of a split kernel evaluates the Boys
function at the class's total AM; and the FP32-VRR kernels round only the
VRR (setup, Boys, accumulation, HRR and stores stay FP64)."""
import re

import pytest

from emitter import generate, hip_lpq, hip_lpqs
from emitter.rrgraph import graph

LPQ_CLASSES = [(1, 0, 0, 0), (1, 1, 1, 1), (2, 1, 1, 0), (2, 2, 1, 0), (2, 0, 2, 0)]


@pytest.mark.parametrize("am", LPQ_CLASSES)
def test_lpq_accumulates_every_top_once(am):
    src = hip_lpq.emit_lpq(am)
    body = src[src.index("for (int j = 0; j < nq; ++j)"):src.index("if (lane < wi.lane_lo")]
    ntop = len(graph(*am).top_keys)
    for t in range(ntop):
        assert len(re.findall(r"\bacc%d \+= " % t, body)) == 1, (am, t)
    assert not re.search(r"\bacc%d \+= " % ntop, body)


def _split_classes():
    out = []
    for am in generate.classes(2):
        for G in generate.lpqsf_variants(am):
            out.append((am, G))
    return out


@pytest.mark.parametrize("am,G", _split_classes())
def test_ket_top_split_kernel_lds_and_cache(am, G):
    src = hip_lpqs.emit_lpqs(am, G, "f")
    g = graph(*am)
    # declared LDS arrays and their sizes (doubles / long longs, 8 bytes each)
    sizes = {}
    for m in re.finditer(r"__shared__ (double|long long) (\w+)\[([^\]]+)\];", src):
        sizes[m.group(2)] = eval(m.group(3))
    total = 8 * sum(sizes.values())
    assert total <= hip_lpqs.LDS_PER_WAVE, (am, G, sizes)
    KC = hip_lpqs.setup_cache_slots(am) if G > 1 else 0
    NV = g.Ltot + 10
    if KC == 0:
        assert "pc[" not in src
        return
    assert sizes["pc"] == KC * NV * 64
    # every access is pc[(t * NV + v) * 64 + lane] with 0 <= v < NV, inside
    # a branch on t < KC
    acc = re.findall(r"pc\[\(t \* (\d+) \+ (\d+)\) \* 64 \+ lane\]", src)
    assert acc and all(int(nv) == NV and 0 <= int(v) < NV for nv, v in acc)
    assert len(re.findall(r"pc\[", src)) == len(acc) + 1   # and the declaration
    writes = re.findall(r"pc\[\(t \* \d+ \+ (\d+)\) \* 64 \+ lane\] = ", src)
    assert sorted(int(v) for v in writes) == list(range(NV))          # group 0 writes every value once
    guards = re.findall(r"if \(t < (\d+)\)", src)
    assert guards and all(int(k) == KC for k in guards) and len(guards) == G   # group 0 + G-1 readers


@pytest.mark.parametrize("am", [am for am in generate.classes(2)])
def test_fast_reciprocal_square_root_only_without_hrr(am):
    """1/(p+q) and sqrt(1/(p+q)) come from one refined reciprocal square root
    (a few ulp) only in classes without an HRR transfer: an HRR transfer
    (a,b+1| = (a+1,b| + AB (a,b| cancels strongly for far-apart pairs and
    would amplify those ulp; classes with HRR divide and take the square
    root exactly."""
    g = graph(*am)
    src = hip_lpq.emit_lpq(am, mixed=True)
    fast = "rsqrt_nr(PQalpha_sum)" in src
    assert fast == (not g.bra_hrr_needed() and not g.ket_hrr_needed()), am
    if not fast:
        assert "1.0 / PQalpha_sum" in src


def _split_all():
    """Every split kernel the build emits for classes up to d: the default
    kind's group counts and the ket-top variants."""
    out = []
    for am in generate.classes(2):
        for G in generate.lpqs_variants(am):
            out.append((am, G, None))
        for G in generate.lpqsf_variants(am):
            out.append((am, G, "f"))
    return out


@pytest.mark.parametrize("am,G,kind", _split_all())
def test_split_groups_evaluate_boys_at_the_class_total(am, G, kind):
    """Every group of a split kernel evaluates F_m at the class's total AM,
    as one evaluation per primitive quartet does: with per-group orders the
    long-range branch's error pattern differs between groups, and the HRR
    transfers between tops of different groups no longer cancel it."""
    src = hip_lpqs.emit_lpqs(am, G, kind)
    orders = [int(x) for x in re.findall(r"boys_eval(?:_rs)?<(\d+)>\(grid", src)]
    assert orders and set(orders) == {graph(*am).Ltot}, (am, G, kind, sorted(set(orders)))


def _kernel_text(src, name):
    i = src.index("__global__ void __launch_bounds__(64) %s(" % name)
    j = src.find("\n}\n", i)
    return src[i:j + 2]


@pytest.mark.parametrize("am", generate.classes(2))
def test_mixed_kernels_keep_setup_boys_accumulation_and_hrr_fp64(am):
    """The FP32-VRR kernels round only the VRR: the Gaussian-product setup,
    the Boys function and the prefactor are FP64, the VRR results are
    accumulated into FP64 accumulators, and everything after the primitive
    loops (the HRR transfers and the stores) is FP64 -- FP32 there lost
    5e-7 on the 64-atom shells."""
    name = "eri_lpqm_" + "_".join("spdf"[x] for x in am)
    src = hip_lpq.emit_lpq(am, mixed=True)
    if name + "(" not in src:
        pytest.skip("no FP32-VRR variant for this class")
    k = _kernel_text(src, name)
    loop = k.index("    for (int i = ib0; i < ib1; ++i) {")
    end = k.index("\n    }\n", loop)
    setup, body, after = k[:loop], k[loop:end], k[end:]
    assert "float" not in after and "fmaf" not in after
    assert re.search(r"\bdouble acc0 = 0\.0;", setup)
    assert not re.search(r"\bfloat acc\d+", k)
    for v in ("PQalpha_sum", "one_over_PQalpha_sum", "F_x", "prefac"):
        assert not re.search(r"\bfloat %s\b" % v, body), v
    assert "boys_eval" in body                          # FP64 Boys inside the loops
    if graph(*am).Ltot > 0:
        assert "fmaf(" in body                          # and the FP32 VRR
