"""The oracle's MD force field against the reference's golden single points (CPU).

ff_small_box: toy water/methanol triclinic box, PME 32^3 order 6, beta 6, cutoff 0.6, tol 1e-10
(reference tests/regression/regression_cases.py:255-289).  ff_pgm3p25_512: the 512-water pGM3P-25
truncated octahedron of examples/water512 with the default settings (regression_cases.py:292-309).
Both were recorded by the reference (jax 0.11.2, float64); the oracle must agree to rounding.
"""

import os

import numpy as np
import pytest

from oracle import systems
from oracle.forcefield import ForceField, Settings

GOLD = os.path.join(os.path.dirname(__file__), "golden")
SMALL = Settings(cutoff=0.6, ewald_beta=6.0, pme_grid=(32, 32, 32), pme_order=6, lj_lrc=True,
                 dipole_tol=1e-10, max_iter=500)


def _check(res, gold, tag, rtol_e=1e-12, rtol_f=1e-10):
    for k in ("elec", "vdw", "total"):
        ref = float(gold[f"{tag}.energy.{k}"])
        assert abs(res.energy[k] - ref) <= rtol_e * max(1.0, abs(ref)), (k, res.energy[k], ref)
    F = gold[f"{tag}.forces"]
    assert np.abs(res.forces - F).max() <= rtol_f * np.abs(F).max()
    mu = gold[f"{tag}.mu"]
    assert np.abs(res.mu - mu).max() <= 1e-10 * np.abs(mu).max()
    assert res.iterations == int(gold[f"{tag}.iterations"])
    assert abs(res.residual - float(gold[f"{tag}.residual"])) <= 1e-6 * float(gold[f"{tag}.residual"])


def test_small_box_matches_golden_double():
    """Energies, forces, induced dipoles, CG iterations and residual of ff_small_box (double)."""
    top, pos, H = systems.small_box(0)
    res = ForceField(top, H, SMALL).compute(pos, H)
    with np.load(os.path.join(GOLD, "ff_small_box.npz")) as gold:
        _check(res, gold, "double")


def test_water512_matches_golden_double():
    """The 512-water box with default settings (grid 30^3 from spacing 0.08 nm) against ff_pgm3p25_512."""
    top, pos, H, _ = systems.water512()
    ff = ForceField(top, H, Settings())
    assert ff.pme.K == (30, 30, 30)
    res = ff.compute(pos, H)
    with np.load(os.path.join(GOLD, "ff_pgm3p25_512.npz")) as gold:
        _check(res, gold, "double")


def test_water512_matches_pmemd_within_its_settings_difference():
    """Induced dipoles agree with pmemd-pgm's (Fortran) to the difference of their settings.

    pmemd ran cut 1.0 nm, ew_coeff 3.8 /nm, 72^3 order 8, tol 1e-9; the oracle is run with those
    settings and must match its dipoles to 1e-5 e A RMS (the reference reports 1.2e-10 for its own
    exact settings; here the grid rule differs slightly).
    """
    top, pos, H, _ = systems.water512()
    s = Settings(cutoff=1.0, ewald_beta=3.8, pme_grid=(72, 72, 72), pme_order=8, dipole_tol=1e-9,
                 max_iter=500, peek=0.0, lj_lrc=False)
    res = ForceField(top, H, s).compute(pos, H)
    with np.load(os.path.join(GOLD, "pmemd_pbc.npz")) as pm:
        mu_ref = pm["mu_eA"] * 0.1
    rms = np.sqrt(np.mean((res.mu - mu_ref) ** 2))
    assert rms < 1e-6, rms


def test_forces_are_minus_the_energy_gradient():
    """Forces at the converged dipoles equal central differences of the fixed-mu energy (variational)."""
    top, pos, H = systems.small_box(1)
    ff = ForceField(top, H, SMALL)
    res = ff.compute(pos, H)
    h = 1e-6
    for a, c in [(0, 0), (5, 1), (40, 2), (77, 0)]:
        d = np.zeros_like(pos)
        d[a, c] = h
        fd = -(ff.energy_fixed_mu(pos + d, H, res.mu) - ff.energy_fixed_mu(pos - d, H, res.mu)) / (2 * h)
        assert abs(fd - res.forces[a, c]) < 1e-6 * max(1.0, abs(fd)), (a, c, fd, res.forces[a, c])


@pytest.mark.parametrize("case", ["small_box", "water512"])
def test_strain_derivative_matches_golden(case):
    """The analytic molecular strain derivative (full 3x3, tail impulse included) vs the golden `strain`."""
    if case == "small_box":
        top, pos, H = systems.small_box(0)
        ff, gold = ForceField(top, H, SMALL), "ff_small_box.npz"
    else:
        top, pos, H, _ = systems.water512()
        ff, gold = ForceField(top, H, Settings()), "ff_pgm3p25_512.npz"
    res = ff.compute(pos, H)
    W = ff.strain_derivative(pos, H, res.mu)
    with np.load(os.path.join(GOLD, gold)) as g:
        ref = g["double.strain"]
    assert np.abs(W - ref).max() < 1e-9 * np.abs(ref).max()


def test_strain_derivative_trace_by_isotropic_scaling():
    """tr(W) + 3 E_lrc equals the finite difference of the fixed-mu energy under isotropic molecular scaling."""
    top, pos, H = systems.small_box(2)
    ff = ForceField(top, H, SMALL)
    res = ff.compute(pos, H)
    W = ff.strain_derivative(pos, H, res.mu)
    m = top.masses
    com = np.array([np.average(pos[top.mol == k], 0, weights=m[top.mol == k]) for k in range(top.nmol)])
    e = lambda s: ff.energy_fixed_mu(pos + (s * com)[top.mol], H * (1 + s), res.mu)  # noqa: E731
    h = 1e-6
    fd = (e(h) - e(-h)) / (2 * h)
    assert abs(fd - (np.trace(W) + 3 * ff.vdw_tail(H))) < 1e-6 * max(1.0, abs(fd))


@pytest.mark.parametrize("molecular", [True, False])
def test_strain_derivative_by_finite_differences(molecular):
    """W_ab (a <= b: the box stays lower triangular) + E_lrc delta_ab equals the central difference of the
    fixed-mu energy under the strain eps = h e_a e_b^T, molecular or atomic scaling."""
    top, pos, H = systems.small_box(2)
    ff = ForceField(top, H, SMALL)
    res = ff.compute(pos, H)
    W = ff.strain_derivative(pos, H, res.mu, molecular=molecular)
    m = top.masses
    com = np.array([np.average(pos[top.mol == k], 0, weights=m[top.mol == k]) for k in range(top.nmol)])

    def e(eps):
        F = np.eye(3) + eps
        x = pos + (com @ eps.T)[top.mol] if molecular else pos @ F.T
        return ff.energy_fixed_mu(x, H @ F.T, res.mu)

    h = 1e-6
    tail = ff.vdw_tail(H)
    for a in range(3):
        for b in range(a, 3):
            E = np.zeros((3, 3))
            E[a, b] = h
            fd = (e(E) - e(-E)) / (2 * h)
            assert abs(fd - (W[a, b] + (tail if a == b else 0.0))) < 2e-6 * np.abs(W).max(), (a, b, fd, W[a, b])
