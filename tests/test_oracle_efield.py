"""The oracle with a uniform static external field against the reference's golden (CPU).

efield_point (tests/regression/regression_cases.py:311-331): the toy water box (small_box(13),
30 waters), PME 32^3 order 6, beta 6, cutoff 0.6 nm, no LJ tail, float64, CG to 1e-10, with the
field E = (0.1, -0.3, 0.8) V/nm (forcefield.py:2048-2097, md/efield.py): energies elec / vdw /
field / total, forces, induced dipoles, the total dipole M, the CG iterations and residual.
"""

import os

import numpy as np

from oracle import systems
from oracle.forcefield import ForceField, Settings

GOLD = os.path.join(os.path.dirname(__file__), "golden", "efield_point.npz")
EFIELD = Settings(cutoff=0.6, ewald_beta=6.0, pme_grid=(32, 32, 32), pme_order=6, lj_lrc=False, dipole_tol=1e-10,
                  max_iter=500, predictor="none")
E_VNM = np.array([0.1, -0.3, 0.8])


def test_efield_point_matches_golden():
    top, pos, H = systems.small_box(13, nm=0)
    r = ForceField(top, H, EFIELD).compute(pos, H, efield=(E_VNM, None))
    with np.load(GOLD) as g:
        for k in ("elec", "vdw", "field", "total"):
            ref = float(g[f"E.energy.{k}"])
            assert abs(r.energy[k] - ref) <= 1e-12 * abs(ref), (k, r.energy[k], ref)
        assert np.abs(r.forces - g["E.forces"]).max() <= 1e-12 * np.abs(g["E.forces"]).max()
        assert np.abs(r.mu - g["E.mu"]).max() <= 1e-12 * np.abs(g["E.mu"]).max()
        assert np.allclose(r.dipole, g["E.dipole"], rtol=1e-12, atol=0)
        assert r.iterations == int(g["E.iterations"])
        assert abs(r.residual - float(g["E.residual"])) <= 1e-6 * float(g["E.residual"])


def test_field_forces_are_minus_the_energy_gradient():
    """With the field the forces at the converged dipoles are still -dE/dx of the energy at fixed mu
    (the q F0 pull and the torque on the covalent dipoles included): central differences, h 1e-6 nm."""
    from oracle.forcefield import VNM_TO_INTERNAL

    top, pos, H = systems.small_box(13, nm=0)
    ff = ForceField(top, H, EFIELD)
    r = ff.compute(pos, H, efield=(E_VNM, None))
    F0 = VNM_TO_INTERNAL * E_VNM
    Gk = ff.pme.influence(H)

    def energy(x):
        return ff.energy_forces(x, H, r.mu, ff.geometry(x, H), ff.pme.setup(x, H), Gk, F0)[0]["total"]

    h = 1e-6
    for a, c in [(0, 2), (1, 0), (44, 1), (89, 2)]:
        d = np.zeros_like(pos)
        d[a, c] = h
        fd = -(energy(pos + d) - energy(pos - d)) / (2 * h)
        assert abs(fd - r.forces[a, c]) < 1e-6 * max(1.0, abs(fd)), (a, c, fd, r.forces[a, c])
