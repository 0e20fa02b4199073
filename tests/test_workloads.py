"""The synthetic workloads, on the CPU, against the recipes they state: the
(pp|pp) quartet set, the grid molecule of the 64- and 256-atom systems and
the all-classes pool and stream (shell counts, angular momenta, contraction
depths, exponent / coefficient / centre ranges, determinism), and the
SplitMix64 stream they draw from against its published first outputs."""
import numpy as np

import workloads


def test_splitmix64_known_answers_and_block():
    r = workloads.SplitMix64(0)
    assert [r.next() for _ in range(3)] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    a, b = workloads.SplitMix64(12345), workloads.SplitMix64(12345)
    blk = a.block(1000)
    assert [int(x) for x in blk] == [b.next() for _ in range(1000)]
    assert a.next() == b.next()       # the state advances alike


def test_pshell_quartets_recipe():
    shells, q = workloads.pshell_quartets()
    assert len(shells) == 40 and len(q) == 1000
    for am, xyz, al, co in shells:
        assert am == 1 and len(al) == 3 and len(co) == 3
        assert all(0.0 <= x <= 10.0 for x in xyz)
        assert all(0.1 <= a <= 20.0 for a in al) and all(0.1 <= c <= 1.0 for c in co)
    assert all(0 <= i < 40 for t in q for i in t)
    assert workloads.pshell_quartets() == (shells, q)


def _atoms(shells):
    assert len(shells) % 6 == 0
    return [shells[i:i + 6] for i in range(0, len(shells), 6)]


def test_grid_molecule_recipe():
    for natom, dims in ((64, (4, 4, 4)), (256, (8, 8, 4))):
        shells = workloads.grid_molecule(natom=natom, seed=0)
        atoms = _atoms(shells)
        assert len(atoms) == natom
        sites = set()
        for at in atoms:
            assert [(s[0], len(s[2])) for s in at] == [(0, 6), (0, 3), (1, 3), (0, 1), (1, 1), (2, 1)]
            assert len({s[1] for s in at}) == 1          # one centre per atom
            for k, (am, xyz, al, co) in enumerate(at):
                lo, hi = (0.1, 5000.0) if k < 3 else (0.05, 2.0)
                assert all(lo <= a <= hi for a in al)
                assert all(0.1 <= c <= 1.0 for c in co)
            xyz = np.array(at[0][1])
            site = np.round(xyz / 3.0).astype(int)
            assert np.all(np.abs(xyz - 3.0 * site) <= 0.3)   # jitter U[-0.3, 0.3] about a 3-bohr grid
            sites.add(tuple(site))
        assert sites == {(i, j, k) for i in range(dims[0]) for j in range(dims[1]) for k in range(dims[2])}
    assert workloads.grid_molecule(natom=8, seed=4) == workloads.grid_molecule(natom=8, seed=4)
    assert workloads.grid_molecule(natom=8, seed=4) != workloads.grid_molecule(natom=8, seed=5)


def test_all_class_stream_recipe():
    shells, q = workloads.all_class_stream(nquartet=20000, seed=1)
    assert len(shells) == 4096
    assert [s[0] for s in shells] == [l for l in range(4) for _ in range(1024)]
    depths = set()
    for am, xyz, al, co in shells:
        depths.add(len(al))
        assert len(al) == len(co)
        assert all(0.0 <= x <= 30.0 for x in xyz)
        assert all(0.05 <= a <= 1e4 for a in al) and all(0.1 <= c <= 1.0 for c in co)
    assert depths == set(range(1, 7))
    assert q.shape == (20000, 4) and q.min() >= 0 and q.max() < 4096
    classes = {tuple(shells[i][0] for i in t) for t in q}
    assert len(classes) == 256                             # every class up to f
    s2, q2 = workloads.all_class_stream(nquartet=20000, seed=1)
    assert np.array_equal(q, q2)
