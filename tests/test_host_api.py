"""Host side of the drop-in boundary (no GPU work): the library loads, exports
every symbol the headers declare, and its shell / shell-pair code produces
the same data as the oracle's restatement of skel/simint/shell/shell.c."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle
import simint
import workloads
from helpers import mol_shells

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    names = set()
    inc = os.path.join(ROOT, "include", "simint")
    for dirpath, _, files in os.walk(inc):
        for f in files:
            if not f.endswith(".h"):
                continue
            text = open(os.path.join(dirpath, f)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            for m in re.finditer(r"^\s*(?:const\s+)?[\w\s\*]+?\b(simint_\w+)\s*\(", text, flags=re.M):
                line = text[m.start():m.end()]
                if "static inline" in line or "#define" in line:
                    continue
                names.add(m.group(1))
    # static inline helpers in headers are not exported symbols
    return names - {"simint_ostei_worksize", "simint_ostei_workmem"}


def test_library_exports_every_declared_symbol():
    L = simint.lib()
    declared = _declared_functions()
    assert len(declared) >= 40
    missing = [n for n in sorted(declared) if not hasattr(L, n)]
    assert not missing, missing
    for n in simint.EXPORTS:
        assert hasattr(L, n), n


def test_struct_layout():
    # field order and offsets of the reference structs (shell.h:18-85, MAXDER > 0)
    assert ctypes.sizeof(simint.Shell) == 64
    # five ints (padded to 24 bytes), then 18 pointers
    assert simint.MultiShellpair.screen_max.offset == 24 + 8 * 18
    assert ctypes.sizeof(simint.MultiShellpair) == simint.MultiShellpair.ptr.offset + 8


def test_normalization_matches_oracle():
    shells = mol_shells("water.6-31gss") + workloads.grid_molecule(natom=1)
    basis = simint.Basis(shells)
    sa = oracle.ShellArrays(shells)
    oracle.normalize(sa)
    for i in range(len(shells)):
        np.testing.assert_array_equal(basis.coefs(i), sa.coef[sa.off[i]:sa.off[i] + sa.nprim[i]])


@pytest.mark.parametrize("screen", [simint.SCREEN_NONE, simint.SCREEN_FASTSCHWARZ])
def test_pair_fill_matches_oracle(screen):
    shells = workloads.grid_molecule(natom=1, seed=5)
    basis = simint.Basis(shells)
    sa = oracle.ShellArrays(shells)
    oracle.normalize(sa)
    n = len(shells)
    pairs = [(i, j) for i in range(n) for j in range(n)]
    P = simint.Pairs.from_pairs(basis, pairs, screen)
    start = 0
    names = ["x", "y", "z", "PA_x", "PA_y", "PA_z", "PB_x", "PB_y", "PB_z", "alpha", "prefac", "alpha2", "beta2"]
    got = {k: P.field(k) for k in names}
    got_screen = P.field("screen")
    nprim12 = P.field("nprim12")
    for k, (i, j) in enumerate(pairs):
        ref = oracle.fill_pair(sa, i, j, screen)
        m = len(ref["x"])
        assert nprim12[k] == m
        for f in names:
            np.testing.assert_array_equal(got[f][start:start + m], ref[f], err_msg=f)
        if screen:
            np.testing.assert_array_equal(got_screen[start:start + m], ref["screen"])
        np.testing.assert_array_equal([P.field("AB_x")[k], P.field("AB_y")[k], P.field("AB_z")[k]], ref["AB"])
        start += m
    assert P.mp.screen_max == 1e99          # shell.c:434 quirk kept
    assert P.mp.nprim == start
    P.free()


def test_cat_shellpairs():
    shells = workloads.grid_molecule(natom=1, seed=2)
    basis = simint.Basis(shells)
    singles = [simint.Pairs(basis, [i], [j], simint.SCREEN_FASTSCHWARZ) for i, j in [(0, 1), (3, 3), (2, 0)]]
    arr = (ctypes.POINTER(simint.MultiShellpair) * 3)(*[ctypes.pointer(p.mp) for p in singles])
    out = simint.MultiShellpair()
    L = simint.lib()
    L.simint_initialize_multi_shellpair(ctypes.byref(out))
    for fn in (L.simint_cat_shellpairs, L.simint_cat_multi_shellpair):
        fn(3, arr, ctypes.byref(out), simint.SCREEN_FASTSCHWARZ)
        assert out.nshell12 == 3 and out.nshell12_clip == 3
        tot = sum(p.mp.nprim for p in singles)
        assert out.nprim == tot
        got = np.ctypeslib.as_array(out.prefac, shape=(tot,))
        want = np.concatenate([p.field("prefac") for p in singles])
        np.testing.assert_array_equal(got, want)
        assert out.screen_max == max(p.mp.screen_max for p in singles)
    L.simint_free_multi_shellpair(ctypes.byref(out))
    for p in singles:
        p.free()


def test_worksize_table():
    L = simint.lib()
    # deriv-0 maxima: ROUND(2*bcont) + ROUND(cont) + prim with the reference's
    # per-AM maxima (ostei_config.h.in:14-17): bcont/cont/prim 1/0/1,
    # 81/81/149, 961/4332/2405
    assert L.simint_eri_worksize(0, 0) == 2 * 1 + 0 + 1
    assert L.simint_eri_worksize(0, 1) == 2 * 81 + 81 + 149
    assert L.simint_eri_worksize(0, 2) == 2 * 961 + 4332 + 2405
    assert L.simint_eri_worksize(1, 1) > L.simint_eri_worksize(0, 1)


def test_device_entry_points_reject_bad_arguments_without_a_device():
    """The device-resident entry points check their arguments before they
    touch the device: NULL pair sets, a missing density or function table,
    an out-of-range bra range or precision are SIMINT_ERROR_ARGUMENT, here
    as on a GPU box; with valid-looking arguments and no GPU the answer is
    SIMINT_ERROR_NO_DEVICE, never a CPU result."""
    L = simint.lib()
    jk = simint.JK(4, 0, None, None, None, None, None)
    one = (ctypes.c_int * 1)(1)
    off = (ctypes.c_longlong * 1)(0)
    assert L.simint_hip_compute_jk_range(None, None, 0, 1, None, one, off, 0.0, 0, None,
                                         ctypes.byref(jk)) == simint.ERROR_ARGUMENT
    assert L.simint_hip_compute_jk_range(None, None, 0, 1, None, one, off, 0.0, 0, None, None) \
        == simint.ERROR_ARGUMENT
    assert L.simint_hip_jk_symmetrize(0, None, None) == simint.ERROR_ARGUMENT
    assert L.simint_hip_compute_eri_range(None, None, 0, 1, None, one, off, 0.0, 0, None) == simint.ERROR_ARGUMENT
    assert L.simint_hip_pair_bounds(None, None) == simint.ERROR_ARGUMENT
    prev = L.simint_hip_set_jk_rows(0)
    assert L.simint_hip_set_jk_rows(prev) == 0
    import torch
    if not torch.cuda.is_available():
        assert L.simint_hip_jk_symmetrize(4, None, None) == simint.ERROR_NO_DEVICE
