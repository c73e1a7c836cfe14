"""Rigid-water MD on the GPU: one step against the oracle's SHAKE/RATTLE step, and NVE conservation (GPU)."""

import numpy as np
import pytest

from oracle import md as omd
from oracle import systems
from oracle.forcefield import ForceField, Settings

pytestmark = pytest.mark.gpu
KW = dict(pme_grid=(32, 32, 32), dipole_tol=1e-10, max_iter=300)


def _setup(precision, nw=512, edge=2.483, seed=5, **kw):
    import pgm_jax_amd as P

    sys_, pos, H = P.models.water_box(nw, edge, seed=seed)
    vel = P.models.maxwell_boltzmann_rigid(pos, sys_.masses, 298.0, seed=6)
    s = dict(KW)
    s.update(kw)
    ff = P.PGMForceField(sys_, H, P.MDSettings(precision=precision, predictor="previous", **s))
    return P, ff, sys_, pos, vel, H


def test_one_step_matches_oracle_step():
    """Positions and velocities after one FP64 step equal the numpy SHAKE/RATTLE step with oracle forces."""
    P, ff, sys_, pos, vel, H = _setup("double")
    print("md one step", flush=True)
    dyn = P.RigidWaterMD(ff, pos, vel)
    dyn.run(1, 1.0)
    top, _, _ = systems.water_box(512, 2.483, seed=5)
    oracle = ForceField(top, H, Settings(pme_grid=(32, 32, 32), dipole_tol=1e-10, max_iter=300, predictor="none", peek=0.65))
    fn = lambda x: oracle.compute(x, H).forces  # noqa: E731
    p1, v1, _ = omd.step(pos, vel, fn(pos), top.masses, omd.constraint_lengths2(pos), 1e-3, fn)
    assert np.abs(dyn.pos.cpu().numpy() - p1).max() < 1e-12
    assert np.abs(dyn.vel.cpu().numpy() - v1).max() < 1e-8 * np.abs(v1).max()


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_nve_energy_conservation(precision):
    """200 NVE steps of 1 fs: total-energy drift well below the kinetic-energy scale; constraints hold."""
    P, ff, sys_, pos, vel, H = _setup(precision, dipole_tol=1e-7)
    print(f"nve {precision}", flush=True)
    dyn = P.RigidWaterMD(ff, pos, vel)
    r0 = dyn.run(20, 1.0)
    r1 = dyn.run(200, 1.0)
    print(r0, r1, flush=True)
    drift = abs(r1["etot"] - r0["etot"])
    assert drift < 2e-3 * r1["ekin"], (drift, r1["ekin"])
    x = dyn.pos.cpu().numpy()
    assert np.abs(omd.constraint_lengths2(x) / omd.constraint_lengths2(pos) - 1).max() < 1e-10
    assert r1["cg_residual_max"] <= 1e-7


def test_langevin_step_matches_oracle_step():
    """One FP64 g-BAOAB step (same Philox noise) equals the numpy step with oracle forces."""
    P, ff, sys_, pos, vel, H = _setup("double")
    print("langevin one step", flush=True)
    dyn = P.RigidWaterMD(ff, pos, vel, thermostat=P.Langevin(2.0), temperature=298.0, seed=12345)
    rep = dyn.run(1, 1.0)
    top, _, _ = systems.water_box(512, 2.483, seed=5)
    oracle = ForceField(top, H, Settings(pme_grid=(32, 32, 32), dipole_tol=1e-10, max_iter=300, predictor="none", peek=0.65))
    fn = lambda x: oracle.compute(x, H).forces  # noqa: E731
    kT = 0.0083144626181532 * 298.0
    p1, v1, _, heat = omd.langevin_step(pos, vel, fn(pos), top.masses, omd.constraint_lengths2(pos), 1e-3, 2.0, kT,
                                        12345, 0, fn)
    assert np.abs(dyn.pos.cpu().numpy() - p1).max() < 1e-12
    assert np.abs(dyn.vel.cpu().numpy() - v1).max() < 1e-8 * np.abs(v1).max()
    assert abs(rep["heat"] - heat) < 1e-6 * abs(heat)  # a difference of kinetic energies ~250x larger


def test_langevin_holds_temperature():
    """Mixed precision, gamma 10/ps at 298 K from a 150 K start: the mean temperature of the last 300 fs is 298 K
    within 4 %, constraints hold."""
    P, ff, sys_, pos, vel, H = _setup("mixed", dipole_tol=1e-7)
    print("langevin temperature", flush=True)
    dyn = P.RigidWaterMD(ff, pos, vel * np.sqrt(0.5), thermostat=P.Langevin(10.0), temperature=298.0, seed=3)
    dyn.run(400, 1.0)
    temps = [dyn.run(10, 1.0)["temperature"] for _ in range(30)]
    print(np.mean(temps), flush=True)
    assert abs(np.mean(temps) - 298.0) < 0.04 * 298.0
    x = dyn.pos.cpu().numpy()
    assert np.abs(omd.constraint_lengths2(x) / omd.constraint_lengths2(pos) - 1).max() < 1e-10


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_water512_ten_steps_known_answer(precision):
    """examples/water512, 10 NVE steps of 1 fs with the reference's MD defaults on the GPU: Epot(10) and T(10)
    against the reference's recorded run (-2114884.256504 kJ/mol, 291.356783 K; tests/test_water512.py:26-27).
    SHAKE/RATTLE vs the reference's rigid-body integrator: 1e-6 relative in Epot, 1e-4 in T."""
    import pgm_jax_amd as P

    top, pos, H, vel = systems.water512()
    print(f"water512 md {precision}", flush=True)
    sys_ = P.System(q=top.q, radius=top.radius, alpha=top.alpha, lj_rmin_half=top.lj_rmin_half, lj_sqrt_eps=top.lj_sqrt_eps,
                    masses=top.masses, mol=top.mol, cov_i=top.cov_i, cov_j=top.cov_j, cov_c=top.cov_c)
    ff = P.PGMForceField(sys_, H, P.MDSettings(precision=precision))
    dyn = P.RigidWaterMD(ff, pos, vel)
    r = dyn.run(10, 1.0)
    print(r, flush=True)
    M = len(pos) // 3
    T = 2 * r["ekin"] / ((6 * M - 3) * 0.0083144626181532)
    assert abs(r["epot"] - (-2114884.256504)) < 1e-6 * 2114884.256504, r["epot"]
    assert abs(T - 291.356783) < 1e-4 * 291.356783, T
