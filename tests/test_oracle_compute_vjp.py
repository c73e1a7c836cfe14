"""The oracle's parameter cotangents of (E, F, mu) against the reference's golden gradient (CPU).

The reference's ff_small_box case (tests/regression/regression_cases.py:253-287) stores dloss.*,
jax.grad of L(theta) = 1e-6 sum F^2 + 1e2 sum mu^2 over the tied parameter table
(pgm_jax/system.py:345-447), computed with differentiable=True, float64, dipole_tol 1e-10 and the
default adjoint_tol 1e-6 (forcefield.py:414-425).  ForceField.compute_vjp with f_bar = 2e-6 F,
mu_bar = 2e2 mu and the same adjoint tolerance, summed over the atoms (covalent dipoles) of each
key, must reproduce every quantity to 1e-9 of its largest entry.  The energy part (e_bar), which
the golden loss does not reach, is pinned by central differences of the energy at fixed mu.
"""

import dataclasses
import os

import numpy as np
import pytest

from oracle import systems
from oracle.forcefield import ForceField, Settings

GOLD = os.path.join(os.path.dirname(__file__), "golden", "ff_small_box.npz")
S = Settings(cutoff=0.6, ewald_beta=6.0, pme_grid=(32, 32, 32), pme_order=6, lj_lrc=True, dipole_tol=1e-10,
             max_iter=500, predictor="none")

# Tying keys of the toy box (models/toy.py:26-78, system.py:300-343): by-type quantities use the atom
# type, q "<molecule>:<type>" (one symmetry class per type here), cov "<molecule>:<type_i>><type_j>";
# keys in first-occurrence order over the molecules (4 MeOH, then 30 WAT).
TYPES = {"MeOH": ["c3", "oh", "h1", "h1", "h1", "ho"], "WAT": ["OW", "HW", "HW"]}
COV = {"MeOH": [(0, 1), (1, 0), (1, 5), (5, 1)] + [c for h in (2, 3, 4) for c in ((0, h), (h, 0))],
       "WAT": [(0, 1), (0, 2), (1, 0), (2, 0)]}
MOLS = ["MeOH"] * 4 + ["WAT"] * 30


def _keys():
    atom_type, q_key, cov_key = [], [], []
    for m in MOLS:
        atom_type += TYPES[m]
        q_key += [f"{m}:{t}" for t in TYPES[m]]
        cov_key += [f"{m}:{TYPES[m][i]}>{TYPES[m][j]}" for i, j in COV[m]]
    by = {k: atom_type for k in ("radius", "alpha", "lj_rmin_half", "lj_sqrt_eps")}
    return dict(by, q=q_key, cov=cov_key)


def _reduce(vals, keys):
    order = list(dict.fromkeys(keys))
    out = np.zeros(len(order))
    np.add.at(out, [order.index(k) for k in keys], vals)
    return out


def test_dloss_matches_the_reference_gradient():
    top, pos, H = systems.small_box(0)
    ff = ForceField(top, H, S)
    res = ff.compute(pos, H)
    out = ff.compute_vjp(pos, H, res.mu, 0.0, 2e-6 * res.forces, 2e2 * res.mu, tol=1e-6)
    keys = _keys()
    with np.load(GOLD) as gold:
        for qn in ("q", "radius", "alpha", "lj_rmin_half", "lj_sqrt_eps", "cov"):
            got = _reduce(out[qn], keys[qn])
            ref = gold["dloss." + qn]
            err = np.abs(got - ref).max() / np.abs(ref).max()
            print(f"{qn}: {err:.2e}")
            assert got.shape == ref.shape and err <= 1e-9, qn
        for qn in ("gvdw_b", "gvdw_sqrt_a", "gvdw_sqrt_c6"):  # Gaussian vdW is off: zero gradients
            assert not np.any(gold["dloss." + qn])


def _energy(top, pos, H, mu):
    ff = ForceField(top, H, S)
    g = ff.geometry(pos, H)
    e, _ = ff.energy_forces(pos, H, mu, g, ff.pme.setup(pos, H), ff.pme.influence(H))
    return e["total"]


@pytest.mark.parametrize("field,key,k", [("q", "q", 5), ("radius", "radius", 3), ("alpha", "alpha", 7),
                                         ("lj_rmin_half", "lj_rmin_half", 1), ("lj_sqrt_eps", "lj_sqrt_eps", 4),
                                         ("lj_sqrt_eps", "lj_sqrt_eps", 30), ("cov_c", "cov", 11)])
def test_energy_part_matches_finite_differences(field, key, k):
    """e_bar = 1: dE/dtheta at fixed mu against fourth-order central differences of the total energy
    (h 1e-4 of the value), including the self, background and LJ-tail terms."""
    top, pos, H = systems.small_box(0)
    ff = ForceField(top, H, S)
    mu = ff.compute(pos, H).mu
    got = ff.compute_vjp(pos, H, mu, 1.0)[key][k]
    arr = getattr(top, field)
    h = 1e-4 * max(abs(arr[k]), 1e-2)
    e = {}
    for s in (1, -1, 2, -2):
        a = arr.copy()
        a[k] += s * h
        e[s] = _energy(dataclasses.replace(top, **{field: a}), pos, H, mu)
    fd = (8.0 * (e[1] - e[-1]) - (e[2] - e[-2])) / (12.0 * h)
    print(f"{key}[{k}]: {got:.10e} fd {fd:.10e}")
    assert abs(got - fd) <= 1e-6 * max(abs(fd), 1.0)
