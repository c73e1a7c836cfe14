"""Split cutoffs on the GPU (electrostatics at elec_cutoff, LJ at cutoff) against the oracle (GPU).

The reference's split rows (pgm_jax/md/forcefield.py:44-58, :1197-1238): the CG field, the
electrostatic energies, forces and strain use the pairs inside elec_cutoff, LJ those inside cutoff.
The oracle's split is pinned on the CPU (tests/test_oracle_split.py: single cutoff + brute-force LJ
shell + tail).  Double precision must match it to rounding (energies 1e-9, forces 1e-7 of the
largest, dipoles 1e-8, the same CG iteration count), mixed within 1e-6 (energy) / 1e-5 (force RMS),
on the atom rows (the water/methanol box) and on the molecule blocks (water; fewer records than
with one cutoff, by about (elec_cutoff / cutoff)^3).
"""

import numpy as np
import pytest

from oracle import systems
from oracle.forcefield import ForceField, Settings

pytestmark = pytest.mark.gpu


def _system(P, top):
    return P.System(q=top.q, radius=top.radius, alpha=top.alpha, lj_rmin_half=top.lj_rmin_half,
                    lj_sqrt_eps=top.lj_sqrt_eps, masses=top.masses, mol=top.mol, cov_i=top.cov_i,
                    cov_j=top.cov_j, cov_c=top.cov_c)


def _rel_rms(a, b):
    return float(np.sqrt(np.mean((a - b) ** 2)) / np.sqrt(np.mean(b**2)))


def _compare(res, ref, precision):
    e, er = res.energy, ref.energy
    F = res.forces.cpu().numpy()
    mu = res.induction.mu.cpu().numpy()
    print(f"  E {e['total']:.9f} (oracle {er['total']:.9f}), vdw {e['vdw']:.9f} ({er['vdw']:.9f}), "
          f"CG {res.iterations} ({ref.iterations}), force rms {_rel_rms(F, ref.forces):.2e}", flush=True)
    if precision == "double":
        for k in ("elec", "vdw", "total"):
            assert abs(e[k] - er[k]) <= 1e-9 * max(1.0, abs(er[k])), k
        assert np.abs(F - ref.forces).max() <= 1e-7 * np.abs(ref.forces).max()
        assert np.abs(mu - ref.mu).max() <= 1e-8 * np.abs(ref.mu).max()
        assert res.iterations == ref.iterations
    else:
        assert abs(e["total"] - er["total"]) <= 1e-6 * abs(er["total"])
        assert _rel_rms(F, ref.forces) <= 1e-5


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_split_atom_rows_against_oracle(precision):
    import pgm_jax_amd as P

    top, pos, H = systems.small_box(0)
    kw = dict(cutoff=0.6, elec_cutoff=0.45, ewald_beta=7.5, pme_grid=(32, 32, 32), pme_order=6, lj_lrc=True,
              dipole_tol=1e-10, max_iter=500)
    print(f"split atom rows {precision}", flush=True)
    ref = ForceField(top, H, Settings(predictor="none", **kw)).compute(pos, H)
    ff = P.PGMForceField(_system(P, top), H, P.MDSettings(precision=precision, skin=0.05, predictor="none", **kw))
    _compare(ff.compute(pos, H), ref, precision)


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_split_molecule_blocks_against_oracle(precision):
    """512 waters, LJ at 0.9 nm, electrostatics at 0.7 nm with beta 5.19/nm and a 48^3 grid."""
    import pgm_jax_amd as P

    top, pos, H = systems.water_box(512, 2.483, seed=5)
    es = P.elec_cutoff_settings(0.7)
    kw = dict(cutoff=0.9, elec_cutoff=0.7, ewald_beta=es["ewald_beta"], pme_grid=(48, 48, 48), pme_order=6, lj_lrc=True,
              dipole_tol=1e-10, max_iter=500)
    print(f"split molecule blocks {precision}", flush=True)
    ref = ForceField(top, H, Settings(predictor="none", **kw)).compute(pos, H)
    ff = P.PGMForceField(_system(P, top), H, P.MDSettings(precision=precision, predictor="none", **kw))
    assert ff.row_layout()[0] == "molecule"
    res = ff.compute(pos, H)
    _compare(res, ref, precision)
    blocks_split = ff.row_blocks()
    one = P.PGMForceField(_system(P, top), H, P.MDSettings(precision=precision, predictor="none",
                                                          **dict(kw, elec_cutoff=None, ewald_beta=4.0)))
    one.compute(pos, H)
    ratio = blocks_split / one.row_blocks()
    print(f"  records {blocks_split} against {one.row_blocks()} with one cutoff: {ratio:.3f}", flush=True)
    assert 0.4 < ratio < 0.65  # (0.7 / 0.9)^3 = 0.47 for pairs; molecule pairs reach a little further


def test_split_md_conserves_energy():
    """200 NVE steps of 1 fs, mixed, split 0.7 / 0.9 nm: total-energy drift below 2e-3 of the kinetic energy."""
    import pgm_jax_amd as P

    sys_, pos, H = P.models.water_box(512, 2.483, seed=5)
    vel = P.models.maxwell_boltzmann_rigid(pos, sys_.masses, 298.0, seed=6)
    es = P.elec_cutoff_settings(0.7)
    ff = P.PGMForceField(sys_, H, P.MDSettings(precision="mixed", cutoff=0.9, elec_cutoff=0.7, ewald_beta=es["ewald_beta"],
                                               pme_grid=(48, 48, 48), dipole_tol=1e-7, max_iter=300))
    md = P.RigidWaterMD(ff, pos, vel)
    print("split md", flush=True)
    r0 = md.run(20, 1.0)
    r1 = md.run(200, 1.0)
    print(r0["etot"], r1["etot"], r1["ekin"], flush=True)
    assert abs(r1["etot"] - r0["etot"]) < 2e-3 * r1["ekin"]
    assert r1["cg_residual_max"] <= 1e-7
