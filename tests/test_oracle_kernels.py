"""The oracle's radial kernels B_n of erf(a r)/r against closed-form identities (CPU)."""

import math

import numpy as np

from oracle.kernels import erf_kernels, erf_kernels_closed, erf_kernels_series


def test_series_and_closed_form_agree_near_the_switch():
    """Around (a r)^2 = 1 both branches agree (closed form loses ~1 digit per order in float64 too)."""
    a = 5.0
    r = np.linspace(0.7, 1.3, 13) / a
    s = erf_kernels_series(a, r, 4)
    c = erf_kernels_closed(a, r, 4)
    for n, (u, v) in enumerate(zip(s, c)):
        assert np.allclose(u, v, rtol=1e-10 * 10**n, atol=0), n


def test_derivative_identity_by_finite_differences():
    """B_{n+1} = -(1/r) dB_n/dr on both branches (a wrong order, sign or factor fails)."""
    a = 6.2
    h = 1e-6
    for r in (0.05, 0.12, 0.3, 0.8):
        B = erf_kernels(a, np.array([r]), 4)
        Bp = erf_kernels(a, np.array([r + h]), 4)
        Bm = erf_kernels(a, np.array([r - h]), 4)
        for n in range(3):
            fd = -(Bp[n] - Bm[n]) / (2 * h) / r
            assert abs(fd[0] - B[n + 1][0]) < 1e-6 * abs(B[n + 1][0]), (r, n)


def test_zero_distance_limit():
    """B_n(0) = 2 a^(2n+1) 2^n / ((2n+1) sqrt(pi)) (the m = 0 term of the series)."""
    a = 4.0
    B = erf_kernels(a, np.array([1e-9]), 4)
    for n in range(4):
        ref = 2 * a ** (2 * n + 1) * 2**n / ((2 * n + 1) * math.sqrt(math.pi))
        assert abs(B[n][0] - ref) < 1e-9 * ref


def test_b0_is_erf_over_r():
    """B0 = erf(a r)/r on the closed-form branch (scipy's erf)."""
    from scipy.special import erf

    r = np.array([0.3, 0.5, 0.9])
    assert np.allclose(erf_kernels(4.0, r, 1)[0], erf(4.0 * r) / r, rtol=1e-15)
