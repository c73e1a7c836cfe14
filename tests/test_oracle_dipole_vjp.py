"""The oracle's backward pass of the induced-dipole solve against finite differences (CPU).

Loss L(theta) = 1e2 sum |mu(theta)|^2 (the dipole part of the reference's differentiable-path
loss, tests/regression/regression_cases.py:282-287).  dL/dtheta from dipole_vjp with mu_bar =
2e2 mu must equal central differences of L over single atoms' charge, Gaussian radius and
polarizability and over single covalent-dipole strengths (mu re-solved to 1e-12 at every
perturbed parameter), to 1e-6 relative.  A wrong sign, a dropped reciprocal term or a wrong
kernel derivative fails by orders of magnitude.
"""

import dataclasses

import numpy as np
import pytest

from oracle import systems
from oracle.forcefield import ForceField, Settings

S = Settings(cutoff=0.6, ewald_beta=6.0, pme_grid=(32, 32, 32), pme_order=6, lj_lrc=True, dipole_tol=1e-13,
             max_iter=1000, predictor="none", peek=0.0)


def _loss(top, pos, H):
    mu = ForceField(top, H, S).compute(pos, H).mu
    return 1e2 * float(np.sum(mu * mu)), mu


def _perturbed(top, field, k, h):
    arr = getattr(top, field).copy()
    arr[k] += h
    return dataclasses.replace(top, **{field: arr})


@pytest.mark.parametrize("field,key,rel", [("q", "q", 1e-4), ("radius", "radius", 1e-4), ("alpha", "alpha", 1e-4),
                                           ("cov_c", "cov", 1e-3)])
def test_vjp_matches_finite_differences(field, key, rel):
    top, pos, H = systems.small_box(0)
    L0, mu = _loss(top, pos, H)
    vjp = ForceField(top, H, S).dipole_vjp(pos, H, mu, 2e2 * mu, tol=1e-13)
    rng = np.random.default_rng(7)
    n = len(getattr(top, field))
    for k in rng.choice(n, 3, replace=False):
        h = rel * max(abs(getattr(top, field)[k]), 1e-3)
        Lp, _ = _loss(_perturbed(top, field, k, h), pos, H)
        Lm, _ = _loss(_perturbed(top, field, k, -h), pos, H)
        fd = (Lp - Lm) / (2 * h)
        assert abs(vjp[key][k] - fd) <= 1e-6 * max(abs(fd), 1e-3 * np.abs(vjp[key]).max()), (field, k, vjp[key][k], fd)
