"""Fit the float32 erfc used by the pair kernels for x >= 1 (writes the coefficients for common.h).

    erfc(x) = exp(-x^2) * t * P(t),   t = 1 / (1 + x/2),   x >= 1  (t in (0, 2/3])

P is a polynomial fitted here by weighted least squares on the relative error of R(x) =
erfc(x) exp(x^2) (scipy.special.erfcx), over t in (0, 2/3] (x from 1 to infinity; at t -> 0,
R -> 1/(x sqrt(pi)), so P(0) = 1/(2 sqrt(pi))).  The form exp(-x^2) t P(t) follows the
rational-Chebyshev idea of W. J. Cody, Math. Comp. 23, 631 (1969) (erfc as exp(-x^2) times a smooth
function of 1/x); the coefficients are this project's own fit.  The pair kernels already need
exp(-x^2) for the Gaussian kernels, so erfc costs one reciprocal and the Horner sum.

    python pgm-jax_amd/csrc/fit_erfc.py      # prints the C array and the max relative error
"""

from __future__ import annotations

import numpy as np
from scipy.special import erfcx

DEG = 10


def fit(deg: int = DEG) -> np.ndarray:
    """Return the coefficients c_0..c_deg (P(t) = sum c_k t^k) minimising the relative error."""
    t = np.linspace(1e-6, 2.0 / 3.0, 20001)
    x = 2.0 * (1.0 - t) / t
    target = erfcx(x) / t
    V = np.vander(t, deg + 1, increasing=True) / target[:, None]  # relative residual
    c, *_ = np.linalg.lstsq(V, np.ones_like(t), rcond=None)
    return c


def max_rel_error(c: np.ndarray, dtype=np.float32) -> float:
    """Max relative error of exp(-x^2) t P(t) vs erfcx(x) exp(-x^2) for x in [1, 12], Horner in `dtype`."""
    x = np.linspace(1.0, 12.0, 200001).astype(dtype)
    t = (dtype(1) / (dtype(1) + dtype(0.5) * x)).astype(dtype)
    p = np.full_like(t, dtype(c[-1]))
    for ck in c[-2::-1]:
        p = (p * t + dtype(ck)).astype(dtype)
    R = (t * p).astype(np.float64)
    ref = erfcx(x.astype(np.float64))
    return float(np.max(np.abs(R / ref - 1.0)))


if __name__ == "__main__":
    c = fit()
    print("// P(t), t = 1/(1 + x/2): c0 .. c%d" % DEG)
    print("{" + ", ".join(f"{v:.9e}f" for v in c) + "}")
    print("max rel error float64 Horner:", max_rel_error(c, np.float64))
    print("max rel error float32 Horner:", max_rel_error(c, np.float32))
