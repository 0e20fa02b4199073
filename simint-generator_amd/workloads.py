"""Synthetic shell sets of the benchmark and test workloads.

Random numbers come from SplitMix64 (Steele, Lea, Flood 2014) so that the
same stream can be reproduced by any host code: u = (next() >> 11) * 2^-53.
Log-uniform means exp(U(ln lo, ln hi)).  Draw order per shell: centre
x, y, z (or per atom), then the exponents, then the coefficients.

  pshell_quartets(seed=0): 40 p shells of contraction depth 3, exponents log-uniform
      in [0.1, 20], coefficients uniform in [0.1, 1], centres uniform in a
      10-bohr cube; 1000 quartets = 1000 draws of four shell indices.
  grid_molecule(natom=64, seed=0): atoms on a jittered cubic grid (spacing 3 bohr,
      jitter U[-0.3, 0.3] per coordinate), each with s(6), s(3), p(3) core
      shells (exponents log-uniform in [0.1, 5000]) and s(1), p(1), d(1)
      valence shells (log-uniform in [0.05, 2]); coefficients U[0.1, 1].
      natom=64 is a 4x4x4 grid, natom=256 an 8x8x4 grid.
  all_class_stream(nquartet=10**6, seed=0): a pool of 4096 shells, 1024 of each AM
      0..3 (in that order); per shell the contraction depth U{1..6}, the
      centre uniform in a 30-bohr cube, exponents log-uniform in
      [0.05, 1e4], coefficients U[0.1, 1]; then nquartet quartets of four
      independent uniform pool indices (every class up to f).
"""
import math

import numpy as np

MASK = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed):
        self.s = seed & MASK

    def next(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & MASK
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
        return z ^ (z >> 31)

    def uniform(self, lo=0.0, hi=1.0):
        return lo + (hi - lo) * ((self.next() >> 11) * (1.0 / 9007199254740992.0))

    def loguniform(self, lo, hi):
        return math.exp(self.uniform(math.log(lo), math.log(hi)))

    def index(self, n):
        return min(int(self.uniform() * n), n - 1)

    def block(self, n):
        """The next n outputs as a uint64 array (same values as n next())."""
        k = np.arange(1, n + 1, dtype=np.uint64)
        with np.errstate(over="ignore"):
            z = np.uint64(self.s) + k * np.uint64(0x9E3779B97F4A7C15)
            self.s = int(z[-1]) if n else self.s
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))

    def indices(self, count, n):
        u = (self.block(count) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
        return np.minimum((u * n).astype(np.int64), n - 1)


def pshell_quartets(seed=0, nshell=40, nquartet=1000):
    r = SplitMix64(seed)
    shells = []
    for _ in range(nshell):
        xyz = (r.uniform(0, 10), r.uniform(0, 10), r.uniform(0, 10))
        al = [r.loguniform(0.1, 20.0) for _ in range(3)]
        co = [r.uniform(0.1, 1.0) for _ in range(3)]
        shells.append((1, xyz, al, co))
    quartets = [tuple(r.index(nshell) for _ in range(4)) for _ in range(nquartet)]
    return shells, quartets


CORE = [(0, 6), (0, 3), (1, 3)]
VALENCE = [(0, 1), (1, 1), (2, 1)]


def grid_dims(natom):
    if natom == 64:
        return (4, 4, 4)
    if natom == 256:
        return (8, 8, 4)
    n = round(natom ** (1.0 / 3.0))
    while n * n * n < natom:
        n += 1
    return (n, n, n)


def grid_molecule(natom=64, seed=0, spacing=3.0, jitter=0.3):
    """Shells of the grid-molecule recipe, returned as a list ordered atom by
    atom: (am, xyz, alphas, coefs)."""
    r = SplitMix64(seed)
    nx, ny, nz = grid_dims(natom)
    shells = []
    k = 0
    for ix in range(nx):
        for iy in range(ny):
            for iz in range(nz):
                if k >= natom:
                    break
                k += 1
                xyz = (ix * spacing + r.uniform(-jitter, jitter),
                       iy * spacing + r.uniform(-jitter, jitter),
                       iz * spacing + r.uniform(-jitter, jitter))
                for am, n in CORE:
                    al = [r.loguniform(0.1, 5000.0) for _ in range(n)]
                    co = [r.uniform(0.1, 1.0) for _ in range(n)]
                    shells.append((am, xyz, al, co))
                for am, n in VALENCE:
                    al = [r.loguniform(0.05, 2.0) for _ in range(n)]
                    co = [r.uniform(0.1, 1.0) for _ in range(n)]
                    shells.append((am, xyz, al, co))
    return shells


def mixed_pool(seed=0, per_am=1024, maxam=3, cube=30.0):
    r = SplitMix64(seed)
    shells = []
    for am in range(maxam + 1):
        for _ in range(per_am):
            k = 1 + r.index(6)
            xyz = (r.uniform(0, cube), r.uniform(0, cube), r.uniform(0, cube))
            al = [r.loguniform(0.05, 1e4) for _ in range(k)]
            co = [r.uniform(0.1, 1.0) for _ in range(k)]
            shells.append((am, xyz, al, co))
    return shells, r


def all_class_stream(nquartet=10 ** 6, seed=0):
    """(shells, quartets[nquartet, 4]) of the mixed-class stream."""
    shells, r = mixed_pool(seed)
    q = r.indices(4 * nquartet, len(shells)).reshape(nquartet, 4)
    return shells, q


def unique_pairs_by_class(shells, maxam=2):
    """All unique shell pairs (i >= j by index), oriented so am1 >= am2,
    grouped by (am1, am2).  Returns {(am1, am2): [(i, j), ...]}."""
    out = {}
    for i in range(len(shells)):
        for j in range(i + 1):
            a, b = (i, j) if shells[i][0] >= shells[j][0] else (j, i)
            key = (shells[a][0], shells[b][0])
            out.setdefault(key, []).append((a, b))
    return out
