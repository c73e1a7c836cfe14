"""The exact Ewald models on the GPU (PeriodicPGM / PeriodicModel, FP64) against the reference's golden
periodic_model.npz and the oracle (GPU).

small_box(1), cutoff 0.6 nm, beta 3.8 /nm, k_tol 1e-12, dipoles to 1e-12, LJ with the long-range
correction (reference tests/regression/regression_cases.py:231-252).  Bounds: energies 1e-10
relative, forces and dipoles 1e-9 of the largest component, strain derivatives 1e-9 of the largest
component, the pressure 1e-9 relative.
"""

import os

import numpy as np
import pytest

from oracle import systems

GOLD = os.path.join(os.path.dirname(__file__), "golden", "periodic_model.npz")

pytestmark = pytest.mark.gpu


def _system(top):
    import pgm_jax_amd as P

    return P.System(q=top.q, radius=top.radius, alpha=top.alpha, lj_rmin_half=top.lj_rmin_half,
                    lj_sqrt_eps=top.lj_sqrt_eps, masses=top.masses, mol=top.mol, cov_i=top.cov_i,
                    cov_j=top.cov_j, cov_c=top.cov_c)


@pytest.fixture(scope="module")
def model():
    import pgm_jax_amd as P

    top, pos, H = systems.small_box(1)
    print("periodic model small_box(1)", flush=True)
    return P.PeriodicModel(_system(top), H, pos, cutoff=0.6, lj_lrc=True), top, pos, H


def test_energy_forces_dipoles_match_golden(model):
    m, top, pos, H = model
    e = m.energy(pos)
    with np.load(GOLD) as g:
        for k in ("perm", "ind", "elec", "vdw", "total"):
            ref = float(g[f"E.{k}"])
            assert abs(e[k] - ref) < 1e-10 * abs(ref), (k, e[k], ref)
        F = m.forces(pos).cpu().numpy()
        assert np.abs(F - g["forces"]).max() < 1e-9 * np.abs(g["forces"]).max()
        mu = m.elec.induced_dipoles(pos).cpu().numpy()
        assert np.abs(mu - g["mu"]).max() < 1e-9 * np.abs(g["mu"]).max()


def test_strain_and_pressure_match_golden(model):
    m, top, pos, H = model
    with np.load(GOLD) as g:
        for mol, key in ((True, "strain_mol"), (False, "strain_atom")):
            W = m.strain_derivative(pos, molecular=mol)
            assert np.abs(W - g[key]).max() < 1e-9 * np.abs(g[key]).max(), (key, W, g[key])
        p = m.pressure(pos)
        assert abs(p - float(g["pressure"])) < 1e-9 * abs(float(g["pressure"])), (p, float(g["pressure"]))


def test_qp_without_vdw_matches_golden():
    import pgm_jax_amd as P

    top, pos, H = systems.small_box(1)
    e = P.PeriodicModel(_system(top), H, pos, cutoff=0.6, elec="qp", vdw="none").energy(pos)
    with np.load(GOLD) as g:
        for k in ("perm", "ind", "elec", "vdw", "total"):
            ref = float(g[f"E_qp_none.{k}"])
            assert abs(e[k] - ref) < 1e-10 * max(abs(ref), 1.0), (k, e[k], ref)


def test_periodic_pgm_energy_is_elec_only(model):
    """PeriodicPGM.energy's total equals PeriodicModel's elec; its forces exclude LJ (oracle LJ forces added back)."""
    from oracle.ewald import EwaldPGM

    m, top, pos, H = model
    e, aux = m.elec.energy(pos)
    assert abs(e["total"] - m.energy(pos)["elec"]) < 1e-10 * abs(e["total"])
    ew = EwaldPGM(top, H, ewald_beta=3.8, cutoff=0.6, lj_lrc=True)
    F_elec = m.elec.forces(pos).cpu().numpy()
    F_all = ew.forces(pos)
    assert np.abs(F_elec - (F_all + ew.lj(pos)[1])).max() < 1e-9 * np.abs(F_all).max()
    assert np.abs(aux["p"].cpu().numpy() - ew.perm_dipoles(pos)).max() < 1e-12
