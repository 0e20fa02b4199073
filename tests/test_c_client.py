"""The drop-in boundary from C: tests/c_api/client.c includes only the
headers under include/, is compiled with gcc and linked against
libsimint.so as a simint program would be (skel/examples/example1.c), and
its results are compared with the oracle.

The CPU test checks that the headers compile as C99 and that every symbol
the client uses resolves at link time; the GPU test runs the client."""
import os
import struct
import subprocess

import numpy as np
import pytest

import oracle
import simint
from helpers import REL_FLOOR, mol_shells, normalized_arrays, parity_report

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "c_api", "client.c")
LIBDIR = os.path.dirname(simint.LIB_PATH)


def _build(tmpdir):
    exe = os.path.join(str(tmpdir), "client")
    cmd = ["gcc", "-std=c99", "-Wall", "-Wextra", "-Werror", "-O2", "-I" + os.path.join(ROOT, "include"), SRC,
           "-o", exe, "-L" + LIBDIR, "-lsimint", "-Wl,-rpath," + LIBDIR, "-Wl,-rpath-link,/opt/rocm/lib",
           "-Wl,--unresolved-symbols=report-all"]
    p = subprocess.run(cmd, capture_output=True, text=True)
    assert p.returncode == 0, p.stderr
    return exe


def test_client_compiles_and_links(tmp_path):
    if not os.path.exists(simint.LIB_PATH):
        pytest.skip("libsimint.so not built")
    exe = _build(tmp_path)
    assert os.path.exists(exe)
    # every simint_* symbol the client references is defined by the library
    und = subprocess.run(["nm", "-u", exe], capture_output=True, text=True).stdout
    used = sorted({ln.split()[-1] for ln in und.splitlines() if "simint_" in ln})
    assert len(used) >= 15, used
    L = simint.lib()
    for name in used:
        assert hasattr(L, name), name


def _write_jobs(path, shells, jobs):
    with open(path, "w") as f:
        f.write("%d\n" % len(shells))
        for am, xyz, al, co in shells:
            f.write("%d %d %r %r %r %s %s\n" % (am, len(al), float(xyz[0]), float(xyz[1]), float(xyz[2]),
                                                 " ".join(repr(float(a)) for a in al),
                                                 " ".join(repr(float(c)) for c in co)))
        f.write("%d\n" % len(jobs))
        for j in jobs:
            f.write("%d %d %d %d %d\n" % j)


def _read_results(path, njob):
    res = []
    with open(path, "rb") as f:
        for _ in range(njob):
            rc, n = struct.unpack("<iq", f.read(12))
            res.append((rc, np.frombuffer(f.read(8 * n), dtype=np.float64)))
    return res


@pytest.mark.gpu
def test_client_results_vs_oracle(tmp_path, gpu):
    """Water/6-31G** (s, p, d shells) plus one f shell: ERIs of every class
    that occurs, first derivatives of a sample, and the three one-electron
    integrals of every shell pair, all through the C client."""
    shells = mol_shells("water.6-31gss") + [(3, (0.3, -0.2, 0.9), [2.1, 0.45], [0.6, 0.5])]
    sa = normalized_arrays(shells)
    n = len(shells)
    rng = np.random.default_rng(3)
    eri_q = [tuple(int(x) for x in rng.integers(0, n, 4)) for _ in range(60)]
    # make sure the f shell appears in several classes
    eri_q += [(n - 1, 0, 1, 2), (3, n - 1, n - 1, 0), (n - 1, n - 1, n - 1, n - 1)]
    der_q = [q for q in eri_q if max(shells[i][0] for i in q) <= 2][:12]
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
    print("C client ERIs", st)
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
        print("C client one-electron kind", kind, st)
        assert ok, (kind, st)
        k += len(oei)
