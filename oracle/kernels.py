"""Radial kernels B_n of erf(a r)/r in float64 (restates pgm_jax/md/kernels.py:46-118).

For phi_a(r) = erf(a r)/r:  B0 = phi_a,  B_n = -(1/r) dB_{n-1}/dr, so grad_x B_n = -B_{n+1} x.
Closed form (kernels.py:67-73):  B_n = [(2n-1) B_{n-1} - (2a^2)^(n-1) c] / r^2, c = 2a exp(-a^2 r^2)/sqrt(pi).
Series for s = (a r)^2 < 1 (kernels.py:99-118):
    B_n = (2 a/sqrt(pi)) (2 a^2)^n sum_{m=0}^{14} (-s)^m / (m! (2m + 2n + 1)).

Parity: pinned by closed-form identities in tests/test_oracle_kernels.py -- the series and the
closed form agree near the switch, B_n = -(1/r) dB_{n-1}/dr by finite differences (a wrong sign,
order or factor of 2 fails it), and the r -> 0 limit B_n(0) = 2 a^(2n+1) 2^n / ((2n+1) sqrt(pi)).
"""

from __future__ import annotations

import math

import numpy as np
from scipy.special import erf

SQRT_PI = math.sqrt(math.pi)
NSERIES = 15


def erf_kernels_closed(a, r, nmax: int = 3) -> tuple:
    """Return (B0..B_{nmax-1}) of erf(a r)/r by the closed-form recursion (kernels.py:46-73)."""
    a = np.asarray(a, float)
    r = np.asarray(r, float)
    c = (2.0 * a / SQRT_PI) * np.exp(-((a * r) ** 2))
    ir2 = 1.0 / (r * r)
    B = [erf(a * r) / r]
    for n in range(1, nmax):
        B.append(((2 * n - 1) * B[-1] - (2.0 * a * a) ** (n - 1) * c) * ir2)
    return tuple(B)


def erf_kernels_series(a, r, nmax: int = 3) -> tuple:
    """Return (B0..B_{nmax-1}) of erf(a r)/r from the 15-term Taylor series (kernels.py:99-110)."""
    a = np.asarray(a, float)
    s = (a * np.asarray(r, float)) ** 2
    terms = [np.ones_like(s)]
    for m in range(1, NSERIES):
        terms.append(terms[-1] * (-s) / m)
    c0 = 2.0 * a / SQRT_PI
    return tuple(c0 * (2.0 * a * a) ** n * sum(t / (2 * m + 2 * n + 1) for m, t in enumerate(terms)) for n in range(nmax))


def erf_kernels(a, r, nmax: int = 3) -> tuple:
    """Return (B0..B_{nmax-1}): series where (a r)^2 < 1, closed form elsewhere (kernels.py:76-118)."""
    a = np.broadcast_to(np.asarray(a, float), np.shape(r)) if np.ndim(a) < np.ndim(r) else np.asarray(a, float)
    r = np.asarray(r, float)
    small = (a * r) ** 2 < 1.0
    rl = np.where(small, 1.0 / np.where(a > 0, a, 1.0), r)
    Bs = erf_kernels_series(a, np.where(small, r, 0.0), nmax)
    Bl = erf_kernels_closed(a, rl, nmax)
    return tuple(np.where(small, u, v) for u, v in zip(Bs, Bl))


def pair_kernels(a, b0: float, r, nmax: int = 3) -> tuple:
    """Return G_n = B_n[erf(a r)/r] - B_n[erf(b0 r)/r] (forcefield.py:1318-1338, the direct-space kernels)."""
    A = erf_kernels(a, r, nmax)
    B = erf_kernels(np.full(np.shape(r), b0), r, nmax)
    return tuple(u - v for u, v in zip(A, B))
