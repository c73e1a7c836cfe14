"""The GPU force field with a uniform static external field (GPU).

pgm_set_efield / pgm_field_result behind PGMForceField.compute(..., efield=(E, offset))
(forcefield.py:2048-2097, :2489-2560; md/efield.py):
  * efield_point, the reference's golden (tests/regression/regression_cases.py:311-331; the toy
    water box, PME 32^3 order 6, beta 6, cutoff 0.6 nm, no LJ tail, CG to 1e-10,
    E = (0.1, -0.3, 0.8) V/nm): FP64 energies elec / vdw / field / total within 1e-9 relative,
    forces within 1e-7 of the largest, dipoles within 1e-8, the total dipole M within 1e-10
    relative, the same CG iterations (the bars of the field-free golden test); mixed precision
    energies within 1e-6 relative, forces and dipoles within 1e-5 relative RMS;
  * 256 waters on atom rows and on molecule blocks (the record kernels of the CG and the force
    pass), with equal and with split cutoffs, a field and a dipole offset, against the oracle's
    compute(efield=...) on the same inputs: FP64 energies within 1e-9 relative, forces within 1e-8
    of the largest, M within 1e-10;
  * E = 0 gives the field-free results bit for bit, and switching the field off again restores
    them; the trial energy with the field from the converged dipoles equals compute's total;
  * MD in a field (RigidWaterMD(efield=...)): one FP64 step against the numpy SHAKE/RATTLE step
    with the oracle's forces in the same field, and NVE conservation of the energy with -E . M;
  * energy_forces / energy_fixed_mu in the field at the golden's dipoles: the golden energies to
    1e-10 (stationary in the dipoles) and forces to 1e-8.
"""

import os

import numpy as np
import pytest

from oracle import systems
from oracle.forcefield import ForceField, Settings

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden", "efield_point.npz")
EFIELD = dict(cutoff=0.6, ewald_beta=6.0, pme_grid=(32, 32, 32), pme_order=6, lj_lrc=False, dipole_tol=1e-10, max_iter=500)
E_VNM = np.array([0.1, -0.3, 0.8])


def _system(top):
    import pgm_jax_amd as P

    return P.System(q=top.q, radius=top.radius, alpha=top.alpha, lj_rmin_half=top.lj_rmin_half,
                    lj_sqrt_eps=top.lj_sqrt_eps, masses=top.masses, mol=top.mol, cov_i=top.cov_i,
                    cov_j=top.cov_j, cov_c=top.cov_c)


def _rel_rms(a, b):
    return float(np.sqrt(np.mean((a - b) ** 2)) / np.sqrt(np.mean(b**2)))


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_efield_point_against_golden(precision):
    import pgm_jax_amd as P

    top, pos, H = systems.small_box(13, nm=0)
    print(f"efield_point {precision}", flush=True)
    ff = P.PGMForceField(_system(top), H, P.MDSettings(precision=precision, predictor="none", skin=0.05, **EFIELD))
    r = ff.compute(pos, H, efield=(E_VNM, None))
    F = r.forces.cpu().numpy()
    mu = r.induction.mu.cpu().numpy()
    with np.load(GOLD) as g:
        print(f"  field {r.energy['field']:.12f} (golden {float(g['E.energy.field']):.12f}), it {r.iterations}", flush=True)
        rt = 1e-9 if precision == "double" else 1e-6
        for k in ("elec", "vdw", "field", "total"):
            ref = float(g[f"E.energy.{k}"])
            assert abs(r.energy[k] - ref) <= rt * abs(ref), (k, r.energy[k], ref)
        if precision == "double":
            assert np.abs(F - g["E.forces"]).max() <= 1e-7 * np.abs(g["E.forces"]).max()
            assert np.abs(mu - g["E.mu"]).max() <= 1e-8 * np.abs(g["E.mu"]).max()
            assert np.abs(r.dipole - g["E.dipole"]).max() <= 1e-10 * np.abs(g["E.dipole"]).max()
            assert r.iterations == int(g["E.iterations"])
        else:
            assert _rel_rms(F, g["E.forces"]) < 1e-5
            assert _rel_rms(mu, g["E.mu"]) < 1e-5
            assert np.abs(r.dipole - g["E.dipole"]).max() <= 1e-6 * np.abs(g["E.dipole"]).max()
            assert abs(r.iterations - int(g["E.iterations"])) <= 1


CASES = {
    "atom rows": (dict(cutoff=0.8, ewald_beta=4.0), ("atom", 4)),
    "molecule blocks": (dict(cutoff=0.8, ewald_beta=4.0), ("molecule", 8)),
    "molecule blocks, split cutoffs": (dict(cutoff=0.9, elec_cutoff=0.7, ewald_beta=4.6), ("molecule", 8)),
}


@pytest.mark.parametrize("case", list(CASES))
def test_waters_against_oracle(case):
    import pgm_jax_amd as P

    kw, layout = CASES[case]
    kw = dict(kw, pme_grid=(40, 40, 40), pme_order=6, lj_lrc=True, dipole_tol=1e-10, max_iter=500)
    top, pos, H = systems.water_box(256, 1.971, seed=3)
    efield = (np.array([0.3, -0.2, 0.5]), np.array([0.01, 0.02, -0.03]))
    print(f"256 waters, {case}", flush=True)
    ref = ForceField(top, H, Settings(predictor="none", **kw)).compute(pos, H, efield=efield)
    ff = P.PGMForceField(_system(top), H, P.MDSettings(precision="double", predictor="none", skin=0.05, **kw))
    ff.set_row_layout(*layout)
    r = ff.compute(pos, H, efield=efield)
    assert ff.row_layout()[0] == layout[0]
    for k in ("elec", "vdw", "field", "total"):
        assert abs(r.energy[k] - ref.energy[k]) <= 1e-9 * abs(ref.energy[k]), (k, r.energy[k], ref.energy[k])
    F = r.forces.cpu().numpy()
    assert np.abs(F - ref.forces).max() <= 1e-8 * np.abs(ref.forces).max()
    assert np.abs(r.dipole - ref.dipole).max() <= 1e-10 * np.abs(ref.dipole).max()
    assert r.iterations == ref.iterations


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_zero_field_and_switching_off(precision):
    import pgm_jax_amd as P

    top, pos, H = systems.small_box(13, nm=0)
    ff = P.PGMForceField(_system(top), H, P.MDSettings(precision=precision, predictor="none", skin=0.05, **EFIELD))
    r0 = ff.compute(pos, H)
    rz = ff.compute(pos, H, efield=(np.zeros(3), None))
    assert rz.energy["field"] == 0.0
    for k in ("elec", "vdw", "total"):
        assert rz.energy[k] == r0.energy[k], k
    assert bool((rz.forces == r0.forces).all()) and bool((rz.induction.mu == r0.induction.mu).all())
    r1 = ff.compute(pos, H, efield=(E_VNM, None))
    assert abs(r1.energy["total"] - r0.energy["total"]) > 1.0
    # the trial energy from the converged dipoles: the same total, with the field
    e, _, it, _ = ff.energy(pos, H, ind=r1.induction, efield=(E_VNM, None))
    assert abs(e - r1.energy["total"]) <= (1e-10 if precision == "double" else 1e-6) * abs(e), (e, r1.energy["total"])
    r2 = ff.compute(pos, H)
    assert "field" not in r2.energy and r2.dipole is None
    assert r2.energy == r0.energy and bool((r2.forces == r0.forces).all())


def test_refusals():
    import pgm_jax_amd as P

    top, pos, H = systems.small_box(13, nm=0)
    ff = P.PGMForceField(_system(top), H, P.MDSettings(precision="double", predictor="none", skin=0.05, **EFIELD))
    r = ff.compute(pos, H, efield=(E_VNM, None))
    with pytest.raises(NotImplementedError):
        ff.strain_derivative(pos, H, r.induction.mu, efield=(E_VNM, None))
    with pytest.raises(NotImplementedError):
        ff.compute(pos, H, efield=(E_VNM, None, "D"))
    # the derivatives are those of the field-free force field: the field is switched off for them
    out = ff.compute_vjp(pos, H, r.induction.mu, 1.0)
    assert out["iterations"] == 0


def test_md_with_field():
    """RigidWaterMD(efield=...): one FP64 step equals the numpy SHAKE/RATTLE step with the oracle's
    forces in the same field; over 200 NVE steps of 1 fs the total energy with the field's -E . M
    term is conserved as well as without a field (drift below 2e-3 of the kinetic energy)."""
    import pgm_jax_amd as P
    from oracle import md as omd

    field = (np.array([0.5, -0.2, 0.3]), None)
    sys_, pos, H = P.models.water_box(512, 2.483, seed=5)
    vel = P.models.maxwell_boltzmann_rigid(pos, sys_.masses, 298.0, seed=6)
    kw = dict(pme_grid=(32, 32, 32), dipole_tol=1e-10, max_iter=300)
    ff = P.PGMForceField(sys_, H, P.MDSettings(precision="double", predictor="previous", **kw))
    print("md step with a field", flush=True)
    dyn = P.RigidWaterMD(ff, pos, vel, efield=field)
    dyn.run(1, 1.0)
    top, _, _ = systems.water_box(512, 2.483, seed=5)
    oracle = ForceField(top, H, Settings(predictor="none", peek=0.65, **kw))
    fn = lambda x: oracle.compute(x, H, efield=field).forces  # noqa: E731
    p1, v1, _ = omd.step(pos, vel, fn(pos), top.masses, omd.constraint_lengths2(pos), 1e-3, fn)
    assert np.abs(dyn.pos.cpu().numpy() - p1).max() < 1e-12
    assert np.abs(dyn.vel.cpu().numpy() - v1).max() < 1e-8 * np.abs(v1).max()
    # a field of 0.5 V/nm turns the waters: the step differs from the field-free one
    free = P.RigidWaterMD(P.PGMForceField(sys_, H, P.MDSettings(precision="double", predictor="previous", **kw)), pos, vel)
    free.run(1, 1.0)
    assert np.abs(free.vel.cpu().numpy() - v1).max() > 1e-4 * np.abs(v1).max()
    print("nve with a field", flush=True)
    ffm = P.PGMForceField(sys_, H, P.MDSettings(precision="mixed", predictor="previous", **dict(kw, dipole_tol=1e-7)))
    md = P.RigidWaterMD(ffm, pos, vel, efield=field)
    r0 = md.run(20, 1.0)
    r1 = md.run(200, 1.0)
    print(r0, r1, flush=True)
    assert abs(r1["etot"] - r0["etot"]) < 2e-3 * r1["ekin"], (r0["etot"], r1["etot"], r1["ekin"])


def test_energy_at_fixed_dipoles_in_the_field():
    """energy_forces / energy_fixed_mu with efield at the golden's own dipoles (energy_fixed_mu of
    forcefield.py:2249-2292 with an external field): the total is stationary in the dipoles, so it
    equals the golden total to 1e-10 relative; forces within 1e-8 of the largest (FP64)."""
    import pgm_jax_amd as P

    top, pos, H = systems.small_box(13, nm=0)
    ff = P.PGMForceField(_system(top), H, P.MDSettings(precision="double", predictor="none", skin=0.05, **EFIELD))
    with np.load(GOLD) as g:
        e, F = ff.energy_forces(pos, g["E.mu"], H, efield=(E_VNM, None))
        for k in ("elec", "vdw", "field", "total"):
            ref = float(g[f"E.energy.{k}"])
            assert abs(e[k] - ref) <= 1e-10 * abs(ref), (k, e[k], ref)
        assert np.abs(F.cpu().numpy() - g["E.forces"]).max() <= 1e-8 * np.abs(g["E.forces"]).max()
        assert ff.energy_fixed_mu(pos, H, g["E.mu"], efield=(E_VNM, None)) == e["total"]
