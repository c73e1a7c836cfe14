"""The FP32 closed-form erfc of the pair kernels (csrc/common.h erfc_ge1) against scipy (CPU).

erfc(x) = exp(-x^2) h(1/x^2) / (x sqrt(pi)) with h = 1 - u R(u), R the (4,3) rational whose
coefficients common.h holds (read from the source here, not retyped).  Evaluated in float32 the
way the kernel does (Horner, reciprocal of Q), h must be within 2e-7 of sqrt(pi) x exp(x^2) erfc(x)
on x in [1, 20] (and the limit h -> 1 as x grows); a wrong coefficient or sign breaks the bound by
orders of magnitude.
"""

import os
import re

import numpy as np
from scipy.special import erfc, erfcx

SRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "pgm-jax_amd", "csrc", "common.h")


def _coeffs(name):
    text = open(SRC).read()
    body = re.search(name + r"\[\d\] = \{([^}]*)\}", text).group(1)
    return [np.float32(float(v.strip().rstrip("f"))) for v in body.split(",")]


def _h32(x):
    P, Q = _coeffs("ERFC_RP"), _coeffs("ERFC_RQ")
    ix = (np.float32(1) / x.astype(np.float32)).astype(np.float32)
    u = (ix * ix).astype(np.float32)
    p = np.zeros_like(u) + P[4]
    for c in P[3::-1]:
        p = (p * u + c).astype(np.float32)
    q = np.zeros_like(u) + Q[3]
    for c in Q[2::-1]:
        q = (q * u + c).astype(np.float32)
    return (np.float32(1) - u * (p / q)).astype(np.float32), ix


def test_scaled_erfc_rational_within_float32_bound():
    x = np.concatenate([np.linspace(1.0, 6.0, 20001), np.geomspace(6.0, 20.0, 2001)])
    h, _ = _h32(x)
    ref = np.sqrt(np.pi) * x * erfcx(x)
    rel = np.abs(h.astype(np.float64) / ref - 1.0)
    assert rel.max() < 2e-7, rel.max()


def test_erfc_from_the_gaussian():
    """The full formula with exp(-x^2) in float64: erfc itself to 2e-7 relative."""
    x = np.linspace(1.0, 9.0, 4001)
    h, ix = _h32(x)
    val = np.exp(-x * x) * h.astype(np.float64) * ix.astype(np.float64) / np.sqrt(np.pi)
    assert np.abs(val / erfc(x) - 1.0).max() < 2e-7
