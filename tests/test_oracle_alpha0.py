"""Atoms without polarizability in the oracle (forcefield.py:767-780 `_div_alpha`).

With alpha = 0 an atom's induced dipole stays 0 and it adds nothing to |mu|^2 / (2 alpha).  Pinned by
continuity: as alpha -> 0 the induced dipole vanishes like alpha E and mu^2/(2 alpha) like alpha E^2 / 2,
so the masked alpha = 0 result equals the limit of small positive alpha: energies to 1e-9 relative
and the other atoms' dipoles to 1e-8 at alpha = 1e-12 nm^3, with that atom's dipole below 1e-10.
"""

import dataclasses

import numpy as np

from oracle import systems
from oracle.forcefield import ForceField, Settings


def _with_alpha(top, idx, value):
    a = top.alpha.copy()
    a[idx] = value
    return dataclasses.replace(top, alpha=a)


def test_zero_alpha_is_the_small_alpha_limit():
    top, pos, H = systems.small_box(0)
    idx = np.arange(0, top.n, 7)  # a scattering of atoms made non-polarizable
    kw = dict(cutoff=0.6, ewald_beta=6.0, pme_grid=(24, 24, 24), pme_order=6, lj_lrc=True, dipole_tol=1e-12, max_iter=500)
    r0 = ForceField(_with_alpha(top, idx, 0.0), H, Settings(**kw)).compute(pos, H)
    rs = ForceField(_with_alpha(top, idx, 1e-12), H, Settings(**kw)).compute(pos, H)
    assert np.all(r0.mu[idx] == 0.0)
    assert np.max(np.abs(rs.mu[idx])) < 1e-10
    keep = np.setdiff1d(np.arange(top.n), idx)
    assert np.max(np.abs(r0.mu[keep] - rs.mu[keep])) < 1e-8
    assert abs(r0.energy["total"] - rs.energy["total"]) <= 1e-9 * abs(rs.energy["total"])
    assert np.max(np.abs(r0.forces - rs.forces)) < 1e-6 * np.max(np.abs(rs.forces))
