"""The Fortran binding (include/simint/simint.f90, module SimintFortran, the
reference's skel/simint/simint.f90 interface): tests/fortran/client.f90 uses
only the module, is compiled with the image's Fortran compiler and linked
against libsimint.so, and runs the same jobs as the C client.

CPU: the module and the client compile and link.  GPU: the client's ERIs,
first derivatives and one-electron integrals match the oracle."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle
import simint
from helpers import REL_FLOOR, mol_shells, normalized_arrays, parity_report
from test_c_client import _read_results, _write_jobs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MOD = os.path.join(ROOT, "include", "simint", "simint.f90")
SRC = os.path.join(ROOT, "tests", "fortran", "client.f90")
LIBDIR = os.path.dirname(simint.LIB_PATH)


def _fc():
    for c in ("/opt/rocm/bin/amdflang", shutil.which("amdflang"), shutil.which("flang"), shutil.which("gfortran")):
        if c and os.path.exists(c):
            return c
    return None


def _build(tmpdir):
    fc = _fc()
    if fc is None:
        pytest.skip("no Fortran compiler")
    d = str(tmpdir)
    p = subprocess.run([fc, "-c", MOD, "-o", os.path.join(d, "simint_f.o")], cwd=d, capture_output=True, text=True)
    assert p.returncode == 0, p.stderr
    exe = os.path.join(d, "client_f")
    p = subprocess.run([fc, SRC, os.path.join(d, "simint_f.o"), "-o", exe, "-L" + LIBDIR, "-lsimint",
                        "-Wl,-rpath," + LIBDIR, "-Wl,-rpath-link,/opt/rocm/lib"], cwd=d, capture_output=True,
                       text=True)
    assert p.returncode == 0, p.stderr
    return exe


def test_fortran_module_compiles_and_links(tmp_path):
    if not os.path.exists(simint.LIB_PATH):
        pytest.skip("libsimint.so not built")
    exe = _build(tmp_path)
    und = subprocess.run(["nm", "-u", exe], capture_output=True, text=True).stdout
    used = sorted({ln.split()[-1] for ln in und.splitlines() if ln.split()[-1].startswith("simint_")})
    assert len(used) >= 12, used
    L = simint.lib()
    for name in used:
        assert hasattr(L, name), name


@pytest.mark.gpu
def test_fortran_client_vs_oracle(tmp_path, gpu):
    """Water/6-31G** plus one f shell through the Fortran client: ERIs of a
    sample of quartets, first derivatives (an f class among them) and the
    one-electron integrals of every shell pair."""
    shells = mol_shells("water.6-31gss") + [(3, (0.3, -0.2, 0.9), [2.1, 0.45], [0.6, 0.5])]
    sa = normalized_arrays(shells)
    n = len(shells)
    rng = np.random.default_rng(4)
    eri_q = [tuple(int(x) for x in rng.integers(0, n, 4)) for _ in range(40)] + [(n - 1, 0, n - 1, 2)]
    der_q = eri_q[:10] + [(n - 1, 1, 0, 0)]
    oei = [(i, j) for i in range(n) for j in range(n)]
    jobs = [(0,) + q for q in eri_q] + [(1,) + q for q in der_q]
    jobs += [(k, i, j, 0, 0) for k in (2, 3, 4) for (i, j) in oei]
    jf = os.path.join(str(tmp_path), "jobs.txt")
    of = os.path.join(str(tmp_path), "out.bin")
    _write_jobs(jf, shells, jobs)
    exe = _build(tmp_path)
    p = subprocess.run([exe, jf, of], capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr
    res = _read_results(of, len(jobs))
    k = 0
    ref, offs, sizes = oracle.eri(sa, eri_q)
    got = np.concatenate([r[1] for r in res[k:k + len(eri_q)]])
    assert all(r[0] == 1 for r in res[k:k + len(eri_q)])
    ex, _, _ = oracle.md_eri(sa, eri_q, reference_boys=True)
    ok, st = parity_report(got, ref, offs, sizes, ex)
    print("Fortran client ERIs", st)
    assert ok, st
    k += len(eri_q)
    ref, offs, sizes = oracle.eri(sa, der_q, deriv=1)
    got = np.concatenate([r[1] for r in res[k:k + len(der_q)]])
    assert all(r[0] == 1 for r in res[k:k + len(der_q)])
    assert np.abs(got - ref).max() <= 1e-10
    k += len(der_q)
    cen = np.array([s[1] for s in shells], dtype=np.float64)
    for kind in (0, 1, 2):
        Z, C = (np.ones(n), cen) if kind == 2 else (None, None)
        r, o, m = oracle.osoei(sa, oei, kind, Z, C)
        assert all(res[k + q][0] == 1 for q in range(len(oei)))
        got = np.concatenate([res[k + q][1] for q in range(len(oei))])
        ex, _, _ = oracle.md_oei(sa, oei, kind, Z, C, reference_boys=True)
        ok, st = parity_report(got, r, o, m, ex, floor=None if kind == 1 else REL_FLOOR)
        assert ok, (kind, st)
        k += len(oei)
