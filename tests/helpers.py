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


ATOL = 1e-14     # skel/test/test_eri.cpp:227 flags |diff| > 1e-14 (with rel > 1e-8)
NORMWISE = 1e-14


def parity_report(got, ref, offs, sizes, exact=None):
    """Compare a result with the oracle per shell quartet.

    Pass criteria (both must hold):
      * elementwise |got - ref| <= 1e-14 * max(1, |ref|);
      * per quartet, max|got - ref| <= max(1e-14 * max|ref|, 4 * max|ref - exact| + 1e-300)
        when `exact` (long-double McMurchie-Davidson values) is given, i.e. the
        two FP64 results may differ by no more than the oracle's own error;
        without `exact` the first bound alone applies per quartet as well as
        elementwise.
    Returns (ok, stats)."""
    got = np.asarray(got)
    ref = np.asarray(ref)
    d = np.abs(got - ref)
    elem_ok = bool(np.all(d <= ATOL * np.maximum(1.0, np.abs(ref))))
    nw = per_quartet_errors(got, ref, offs, sizes)
    quart_ok = True
    worst_ratio = 0.0
    if exact is not None:
        for o, n in zip(offs, sizes):
            dq = d[o:o + n].max() if n else 0.0
            rq = np.abs(ref[o:o + n]).max() if n else 0.0
            eq = np.abs(ref[o:o + n] - exact[o:o + n]).max() if n else 0.0
            bound = max(NORMWISE * rq, 4.0 * eq + 1e-300)
            worst_ratio = max(worst_ratio, dq / bound)
            if dq > bound:
                quart_ok = False
    stats = {"max_abs": float(d.max()) if d.size else 0.0,
             "max_normwise": float(nw.max()) if nw.size else 0.0,
             "frac_normwise_gt_1e-14": float((nw > NORMWISE).mean()) if nw.size else 0.0,
             "median_normwise": float(np.median(nw)) if nw.size else 0.0,
             "worst_vs_exact_bound": worst_ratio}
    return elem_ok and quart_ok, stats
