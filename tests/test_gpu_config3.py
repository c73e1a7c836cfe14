"""The 32,768-water system at full size (GPU): mixed against FP64, NVE with the learned guess, and the
learned guess against the previous-step guess.

System: 32,768 rigid pGM3P-25 waters (98,304 atoms) in a 9.93 nm cubic box, seeded jittered lattice
with random orientations, qpi + LJ (long-range correction), cutoff 0.9 nm, Ewald beta 4.0/nm, PME
128^3 order 6, CG tolerance 1e-7.  The lattice start is first relaxed by 300 Langevin steps
(10/ps, 298 K, mixed precision).  Tolerances follow the mixed path's bars (energy 1e-7 relative,
CG residual 1e-7), as pgm_jax's own tests compare the mixed path with float64
(tests/test_water512.py:10-12 gives the convention: energies relative, forces and dipoles as RMS).
"""

import numpy as np
import pytest

from oracle import md as omd

pytestmark = pytest.mark.gpu

NW, EDGE, K = 32768, 9.93, 128
BASE = dict(cutoff=0.9, skin=0.1, ewald_beta=4.0, pme_grid=(K, K, K), pme_order=6, lj_lrc=True, dipole_tol=1e-7,
            max_iter=200)


@pytest.fixture(scope="module")
def relaxed():
    """(P, system, box, positions, velocities) after 300 Langevin steps from the lattice start."""
    import pgm_jax_amd as P

    print("32k waters: building the lattice", flush=True)
    sys_, pos, H = P.models.water_box(NW, EDGE, seed=3)
    vel = P.models.maxwell_boltzmann_rigid(pos, sys_.masses, 298.0, seed=4)
    ff = P.PGMForceField(sys_, H, P.MDSettings(precision="mixed", predictor="mu4", precond="molecule", **BASE))
    md = P.RigidWaterMD(ff, pos, vel, thermostat=P.Langevin(10.0), temperature=298.0, seed=5)
    rep = md.run(300, 1.0)
    print("relaxed:", {k: rep[k] for k in ("epot", "temperature", "cg_iterations", "cg_residual_max")}, flush=True)
    assert rep["cg_residual_max"] <= 1e-7
    out = (P, sys_, H, md.pos.clone(), md.vel.clone())
    del md, ff
    return out


def _rel_rms(a, b):
    return float(np.sqrt(np.mean((a - b) ** 2)) / np.sqrt(np.mean(b**2)))


def test_mixed_against_fp64_at_32k_waters(relaxed):
    """Fresh solves (from alpha b) of mixed and FP64 at the relaxed configuration: energy within 1e-7
    relative, force and dipole RMS within 2e-6, both CG residuals within 1e-7."""
    P, sys_, H, x, _ = relaxed
    fresh = dict(BASE, predictor="none")
    print("32k: mixed single point", flush=True)
    rm = P.PGMForceField(sys_, H, P.MDSettings(precision="mixed", **fresh)).compute(x, H)
    print("32k: fp64 single point", flush=True)
    rd = P.PGMForceField(sys_, H, P.MDSettings(precision="double", **fresh)).compute(x, H)
    e_rel = abs(rm.energy["total"] - rd.energy["total"]) / abs(rd.energy["total"])
    f_rel = _rel_rms(rm.forces.cpu().numpy(), rd.forces.cpu().numpy())
    d_rel = _rel_rms(rm.induction.mu.cpu().numpy(), rd.induction.mu.cpu().numpy())
    print(f"32k mixed vs fp64: energy {e_rel:.2e}, force rms {f_rel:.2e}, dipole rms {d_rel:.2e}, residuals "
          f"{rm.residual:.2e} / {rd.residual:.2e}, CG {rm.iterations} / {rd.iterations} it", flush=True)
    assert e_rel <= 1e-7
    assert f_rel <= 2e-6 and d_rel <= 2e-6
    assert rm.residual <= 1e-7 and rd.residual <= 1e-7


@pytest.fixture(scope="module")
def learned_md(relaxed):
    """Mixed NVE with the learned guess from the relaxed state: 100 steps harvesting every 10 solves,
    then 300 Adam steps of the network on the harvest; returns (P, ff, md)."""
    P, sys_, H, x, v = relaxed
    ff = P.PGMForceField(sys_, H, P.MDSettings(precision="mixed", predictor="learned", precond="molecule", **BASE))
    ff.guess_config(hidden_layers=3, harvest_every=10, seed=11)
    md = P.RigidWaterMD(ff, x, v)
    print("32k learned: 100 harvesting steps", flush=True)
    md.run(100, 1.0)
    loss = ff.guess_train(steps=300, batch=4096, lr=1e-3)
    print("32k learned: training loss", loss, ff.guess_status(), flush=True)
    return P, ff, md


def test_nve_with_learned_guess_at_32k_waters(learned_md, relaxed):
    """200 NVE steps with the learned guess: every solve within the tolerance, constraints to 1e-10,
    total-energy change within 2e-3 of the kinetic energy, dipoles of the last step equal to a fresh
    FP64 solve within 2e-6 RMS."""
    P, ff, md = learned_md
    _, sys_, H, x0, _ = relaxed
    r0 = md.run(1, 1.0)
    r1 = md.run(200, 1.0)
    drift = abs(r1["etot"] - r0["etot"])
    print(f"32k NVE learned: Etot {r0['etot']:.3f} -> {r1['etot']:.3f} (drift {drift:.3f} kJ/mol, Ekin {r1['ekin']:.1f}), "
          f"CG {r1['cg_iterations'] / 200:.2f} it/step, residual max {r1['cg_residual_max']:.2e}, "
          f"{ff.guess_status()['mode']}", flush=True)
    assert r1["cg_residual_max"] <= 1e-7 and not r1["not_converged"]
    assert drift < 2e-3 * r1["ekin"]
    x = md.pos.cpu().numpy()
    assert np.abs(omd.constraint_lengths2(x) / omd.constraint_lengths2(x0.cpu().numpy()) - 1).max() < 1e-10
    rd = P.PGMForceField(sys_, H, P.MDSettings(precision="double", **dict(BASE, predictor="none"))).compute(md.pos, H)
    d_rel = _rel_rms(md.ind.mu.cpu().numpy(), rd.induction.mu.cpu().numpy())
    print(f"32k NVE learned: dipoles vs fp64 {d_rel:.2e}", flush=True)
    assert d_rel <= 2e-6


def test_learned_guess_against_previous_step_guess_at_32k_waters(learned_md):
    """From the same state, 50 NVE steps: the learned guess needs fewer CG iterations per step than the
    previous-step guess and no more than mu4 (same preconditioner, same tolerance)."""
    P, ff, md = learned_md
    x, v = md.pos.clone(), md.vel.clone()
    its = {}
    for pred in ("previous", "mu4"):
        ffp = P.PGMForceField(ff.sys, ff._H, P.MDSettings(precision="mixed", predictor=pred, precond="molecule", **BASE))
        mdp = P.RigidWaterMD(ffp, x, v)
        mdp.run(5, 1.0)
        its[pred] = mdp.run(50, 1.0)["cg_iterations"] / 50
        del mdp, ffp
    md.run(5, 1.0)
    its["learned"] = md.run(50, 1.0)["cg_iterations"] / 50
    print("32k CG it/step:", its, ff.guess_status(), flush=True)
    assert its["learned"] < its["previous"]
    assert its["learned"] <= its["mu4"]
