"""Shared test helpers: fixture loading and error metrics."""
import os

import numpy as np

import molfile
import oracle

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def mol_shells(name):
    shells, _ = molfile.read_mol(os.path.join(GOLDEN, "mol", name + ".mol"))
    return shells


def normalized_arrays(shells):
    sa = oracle.ShellArrays(shells)
    oracle.normalize(sa)
    return sa


def normwise_rel(a, b, floor=1e-300):
    """max|a-b| / max|b| per block; blocks that are zero by symmetry fall back
    to the absolute difference."""
    d = np.abs(a - b).max() if a.size else 0.0
    m = np.abs(b).max() if b.size else 0.0
    return d / m if m > 1e-12 else d


def per_quartet_errors(x, ref, offs, sizes):
    return np.array([normwise_rel(x[o:o + n], ref[o:o + n]) for o, n in zip(offs, sizes)])
