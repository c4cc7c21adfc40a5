"""The V-cycle's Chebyshev smoother (host-side coefficients, rustyfem_chebyshev;
no counterpart in the reference, whose smoother is Gauss-Seidel) against the
closed form of Chebyshev iteration: after `degree` steps from x = 0 the error
of an eigencomponent lambda of D^-1 A is T_deg((theta - lambda) / delta) /
T_deg(theta / delta) (Saad, Iterative Methods for Sparse Linear Systems,
2nd ed., section 12.3), which equioscillates with amplitude 1 / T_deg(theta /
delta) on [hi / ratio, hi] and stays below 1 in magnitude on (0, hi]."""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def rf():
    from rustyfem_amd import ffi
    ffi.lib()
    return ffi


def smoother_error(c0, c1, c2, lam):
    """Error factor of the smoother's recurrence (x_0 = 0, D^-1 A -> lam, D^-1 b -> lam)."""
    e = np.ones_like(lam)            # e = x* - x with x* = 1
    d = c0 * lam                     # d_0 = c0 D^-1 b
    e = e - d
    for a, b in zip(c1, c2):
        d = a * d + b * lam * e      # D^-1 r = lam e
        e = e - d
    return e


@pytest.mark.parametrize("degree", [1, 2, 3, 4, 6])
@pytest.mark.parametrize("ratio", [20.0, 8.0])
def test_chebyshev_error_polynomial(rf, degree, ratio):
    hi = 2.37
    lo = hi / ratio
    c0, c1, c2 = rf.chebyshev(hi, ratio, degree)
    assert len(c1) == degree - 1
    theta, delta = 0.5 * (hi + lo), 0.5 * (hi - lo)
    lam = np.linspace(lo, hi, 401)
    t = np.polynomial.chebyshev.Chebyshev.basis(degree)
    want = t((theta - lam) / delta) / t(theta / delta)
    got = smoother_error(c0, c1, c2, lam)
    assert np.max(np.abs(got - want)) < 1e-12
    amp = 1.0 / t(theta / delta)
    assert abs(np.max(np.abs(got)) - amp) < 1e-12
    # the smoother never amplifies a component with 0 < lambda <= hi
    small = np.linspace(1e-6, lo, 101)
    assert np.all(np.abs(smoother_error(c0, c1, c2, small)) <= 1.0 + 1e-12)


def test_chebyshev_rejects_bad_interval(rf):
    with pytest.raises(ValueError):
        rf.chebyshev(1.0, 1.0, 2)
    with pytest.raises(ValueError):
        rf.chebyshev(-1.0, 20.0, 2)
