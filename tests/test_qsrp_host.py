"""Host-side pins of the QSRP screen (no GPU): the oracle's CDF against the
published error of Abramowitz & Stegun 7.1.26, and the library's minimax line
fit (the per-user regression of rkr_qsrp_build) against a brute-force
enumeration of the vertices of NormalLinearRegressionLP's LP."""
import ctypes as C
import math

import numpy as np


def _fit(lib, x, y):
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    a, b, e = C.c_double(), C.c_double(), C.c_double()
    rc = lib.rkr_qsrp_fit_line(x.ctypes.data, y.ctypes.data, len(x), C.byref(a), C.byref(b), C.byref(e))
    return rc, a.value, b.value, e.value


def test_cdf_phi_within_published_error(orc):
    # A&S 7.1.26: |erf error| <= 1.5e-7, so |Phi error| <= 7.5e-8
    for x in np.linspace(-9, 9, 20001):
        exact = 0.5 * (1.0 + math.erf(x / math.sqrt(2.0)))
        assert abs(orc.cdf_phi(float(x)) - exact) <= 7.5e-8 + 1e-15, x


def test_lp_bruteforce_closed_forms(orc):
    x = np.array([0.0, 1.0, 2.0, 5.0])
    assert orc.lp_min_error_bruteforce(x, 3 * x - 2) < 1e-12           # collinear: e = 0
    # three points: the minimax line is parallel to the outer chord, e = half the middle gap
    assert abs(orc.lp_min_error_bruteforce([0.0, 1.0, 2.0], [0.0, 1.0, 0.0]) - 0.5) < 1e-12
    # a vertical pair at x = 1 with y 0 and 4 forces e >= 2
    assert abs(orc.lp_min_error_bruteforce([0.0, 1.0, 1.0, 2.0], [2.0, 0.0, 4.0, 2.0]) - 2.0) < 1e-12


def test_minimax_fit_equals_lp_vertex_enumeration(orc):
    from rkr import _lib
    lib = _lib.load()
    rng = np.random.default_rng(21)
    for case in range(60):
        if case % 2 == 0:
            # the regression's own points: a decreasing score row and its sample positions
            S = int(rng.integers(3, 8))
            T = np.sort(rng.standard_normal(S).astype(np.float32))[::-1]
            mu = np.float32(T.mean())
            inv = np.float32(1.0 / T.std())
            x, y = orc.regression_points(T, mu, inv)
        else:
            n = int(rng.integers(3, 12))
            x = rng.standard_normal(n)
            y = rng.standard_normal(n) * 3
        rc, a, b, e = _fit(lib, x, y)
        assert rc == 0
        ref = orc.lp_min_error_bruteforce(x, y)
        assert abs(e - ref) <= 1e-9 * max(1.0, ref), (case, e, ref)
        # the returned line attains its error on the points
        assert np.max(np.abs(a * x + b - y)) <= e * (1 + 1e-12) + 1e-12


def test_minimax_fit_rejects_single_x():
    from rkr import _lib
    lib = _lib.load()
    rc, *_ = _fit(lib, [1.0, 1.0, 1.0], [0.0, 1.0, 2.0])
    assert rc == _lib.RKR_ERR_INVALID
