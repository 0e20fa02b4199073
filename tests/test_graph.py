"""Pins the emitter's recurrence graphs (CPU only).

The emitter's VRR/HRR graphs must be the reference generator's: the sizes
the reference reports for every class ("WORK SIZE: bcont prim cont",
generator/ostei/OSTEI_Writer.cpp:753, parsed by create.py:266-271) are
reproduced from the graph, compared against the reference generator itself
(compiled from its sources into oracle/_ref) where it is built, against the
oracle's independent C construction, and against the worksize table the
reference ships (skel/simint/ostei/ostei_config.h.in:14-17).
"""
import itertools
import os
import re
import subprocess

import pytest

import oracle
from emitter import config_header
from emitter.rrgraph import graph

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_GEN = os.path.join(ROOT, "oracle", "_ref", "ostei_generator")

CLASSES3 = list(itertools.product(range(4), repeat=4))


def test_graph_matches_oracle_construction():
    for am in CLASSES3:
        c = graph(*am).counts()
        oc = oracle.class_counts(*am)
        assert c["bcont"] == oc[0], am
        assert len(graph(*am).node_list) == oc[1], am
        assert c["prim"] == oc[3], am


def test_worksize_maxima_match_reference_table():
    # per-AM maxima of (bcont, cont, prim) in the reference's shipped table
    want = {0: (1, 0, 1), 1: (81, 81, 149), 2: (961, 4332, 2405), 3: (5476, 57512, 17273)}
    for L, (b, c, p) in want.items():
        cls = list(itertools.product(range(L + 1), repeat=4))
        assert max(graph(*a).counts()["bcont"] for a in cls) == b
        assert max(graph(*a).counts()["cont"] for a in cls) == c
        assert max(graph(*a).counts()["prim"] for a in cls) == p


def test_flop_counts_follow_reference_emission():
    # FLOPs the reference generator emits per primitive quartet (setup + VRR,
    # FMA = 2, counted from its inlined output; one MUL of the screening
    # test not counted here) and per contracted quartet (HRR)
    for am, f_prim_ref, f_hrr_ref in [((0, 0, 0, 0), 23, 0), ((1, 1, 1, 1), 706, 324),
                                      ((2, 0, 2, 0), 586, 0), ((2, 1, 2, 1), 3317, 1224),
                                      ((2, 2, 2, 2), 12757, 10586), ((3, 3, 3, 3), 95187, 111012)]:
        fc = graph(*am).flop_counts()
        L = sum(am)
        # reference text counts the (L+1) prefactor MULs once and includes the
        # screening MUL: ours = ref - 1 - 1 + (L + 1)
        assert fc["f_vrr"] + fc["f_setup"] == f_prim_ref - 2 + (L + 1), am
        assert fc["f_hrr"] == f_hrr_ref, am


@pytest.mark.skipif(not os.path.exists(REF_GEN), reason="reference generator not built here")
def test_graph_matches_reference_generator(tmp_path):
    for am in CLASSES3:
        out = subprocess.run([REF_GEN, "-q", *map(str, am), "-o", str(tmp_path / "k.c"), "-oh", str(tmp_path / "k.h"),
                              "-ve", "100", "-vg", "100", "-he", "100", "-hg", "100"],
                             capture_output=True, text=True, check=True).stdout
        m = re.search(r"WORK SIZE:\s+(\d+)\s+(\d+)\s+(\d+)", out)
        ref = tuple(int(x) for x in m.groups())
        c = graph(*am).counts()
        assert (c["bcont"], c["prim"], c["cont"]) == ref, am


def test_committed_config_header_matches_emitter():
    text = open(os.path.join(ROOT, "include", "simint", "ostei", "ostei_config.h")).read()
    maxam = int(re.search(r"#define SIMINT_OSTEI_MAXAM (\d+)", text).group(1))
    assert text == config_header.text(maxam, 1)
