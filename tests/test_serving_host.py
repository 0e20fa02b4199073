"""Host-side parts of the serving loop, on the CPU: request validation
(malformed requests are rejected before they touch the device), the
admission footprint, the seeded request stream and the default density of
a "jk" request."""
import math

import numpy as np
import pytest

import serving

GOOD = (1, (0.0, 0.0, 0.0), [1.0, 0.2], [0.5, 0.5])


def test_validate_accepts_well_formed_shells():
    assert serving.validate([GOOD, (0, (1.0, 2.0, 3.0), [3.0], [1.0]), (2, (0.5, 0.0, 0.0), [0.8], [1.0])]) is None


@pytest.mark.parametrize("bad,what", [
    ([], "no shells"),
    ([(3, (0.0, 0.0, 0.0), [1.0], [1.0])], "angular momentum"),       # above the served maximum
    ([(-1, (0.0, 0.0, 0.0), [1.0], [1.0])], "angular momentum"),
    ([(1.0, (0.0, 0.0, 0.0), [1.0], [1.0])], "angular momentum"),     # not an integer
    ([(0, (0.0, 0.0, 0.0), [], [])], "exponents"),
    ([(0, (0.0, 0.0, 0.0), [1.0, 2.0], [1.0])], "exponents"),
    ([(0, (0.0, math.nan, 0.0), [1.0], [1.0])], "centre"),
    ([(0, (0.0, 0.0, 0.0), [math.nan], [1.0])], "exponent"),
    ([(0, (0.0, 0.0, 0.0), [-1.0], [1.0])], "exponent"),
    ([(0, (0.0, 0.0, 0.0), [0.0], [1.0])], "exponent"),
    ([(0, (0.0, 0.0, 0.0), [1.0], [math.inf])], "coefficient"),
])
def test_validate_rejects_malformed_shells(bad, what):
    reason = serving.validate([GOOD] + bad if bad else bad)
    assert reason is not None and what in reason, reason
    if bad:
        assert reason.startswith("shell 1")     # the offending shell is named


def test_footprint_grows_with_the_request_and_counts_jk_matrices():
    f = [serving.footprint(n, "eri", 1 << 27) for n in (60, 120, 240)]
    assert f[0] < f[1] < f[2]
    assert abs(f[2] / f[1] - (240 * 241) / (120 * 121)) < 1e-12     # pairs grow as n(n+1)/2
    n = 120
    assert serving.footprint(n, "jk", 1 << 27) - serving.footprint(n, "eri", 1 << 27) == 3 * 8 * (6 * n) ** 2
    assert serving.footprint(n, "deriv", 1 << 27) == serving.footprint(n, "eri", 1 << 27)


def test_request_stream_is_seeded_and_poisson_ordered():
    a = serving.request_stream(40, 0.5, seed=3, natom_range=(10, 48), deriv_frac=0.3, jk_frac=0.5)
    b = serving.request_stream(40, 0.5, seed=3, natom_range=(10, 48), deriv_frac=0.3, jk_frac=0.5)
    assert [(r.natom, r.seed, r.kind, r.arrival) for r in a] == [(r.natom, r.seed, r.kind, r.arrival) for r in b]
    arr = [r.arrival for r in a]
    assert all(x < y for x, y in zip(arr, arr[1:])) and arr[0] > 0
    assert all(10 <= r.natom <= 48 for r in a)
    kinds = {r.kind for r in a}
    assert kinds <= {"eri", "deriv", "jk"} and {"deriv", "jk"} <= kinds
    assert {r.kind for r in serving.request_stream(40, 0.5, seed=3, jk_frac=0.0)} <= {"eri", "deriv"}


def test_request_density_is_symmetric_positive_semidefinite():
    D = serving.request_density(30, 7)
    assert D.shape == (30, 30)
    assert np.array_equal(D, D.T)
    assert np.linalg.eigvalsh(D).min() > -1e-12
    assert np.array_equal(D, serving.request_density(30, 7))
