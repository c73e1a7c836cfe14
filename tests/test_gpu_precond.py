"""The molecule-block CG preconditioner (GPU): same dipoles and energies, fewer iterations.

settings precond="molecule" replaces the reference's Jacobi z = alpha r by the inverse 9x9 block of
each 3-site molecule; the convergence test (max|alpha r| / mean|alpha b| <= tol) is unchanged, so
both solves end within the tolerance of the same linear system: at tol 1e-9 the dipoles agree to
1e-7 relative RMS and the energies to 1e-9 relative (FP64), with fewer CG iterations.  Beads use
the same kernels bead by bead.
"""

import numpy as np
import pytest

from oracle import systems

pytestmark = pytest.mark.gpu

CFG1 = dict(cutoff=0.9, ewald_beta=4.0, pme_grid=(32, 32, 32), pme_order=6, max_iter=200)


def _system(top):
    import pgm_jax_amd as P

    return P.System(q=top.q, radius=top.radius, alpha=top.alpha, lj_rmin_half=top.lj_rmin_half,
                    lj_sqrt_eps=top.lj_sqrt_eps, masses=top.masses, mol=top.mol, cov_i=top.cov_i,
                    cov_j=top.cov_j, cov_c=top.cov_c)


def _rel_rms(a, b):
    return float(np.sqrt(np.mean((a - b) ** 2)) / np.sqrt(np.mean(b**2)))


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_molecule_blocks_same_solution_fewer_iterations(precision):
    import pgm_jax_amd as P

    top, pos, H = systems.water_box(512, 2.483, seed=8)
    tol = 1e-9 if precision == "double" else 1e-7
    res = {}
    for pc in ("jacobi", "molecule"):
        ff = P.PGMForceField(_system(top), H, P.MDSettings(precision=precision, precond=pc, dipole_tol=tol, **CFG1))
        res[pc] = ff.compute(pos, H)
        print(f"precond {pc} {precision}: {res[pc].iterations} it, residual {res[pc].residual:.2e}, "
              f"E {res[pc].energy['total']:.9f}", flush=True)
    j, m = res["jacobi"], res["molecule"]
    assert m.residual <= tol and j.residual <= tol
    assert m.iterations < j.iterations
    assert _rel_rms(m.induction.mu.cpu().numpy(), j.induction.mu.cpu().numpy()) < (1e-7 if precision == "double" else 1e-5)
    et = 1e-9 if precision == "double" else 1e-6
    assert abs(m.energy["total"] - j.energy["total"]) <= et * abs(j.energy["total"])


def test_molecule_blocks_on_beads():
    import pgm_jax_amd as P

    P_BEADS = 3
    qs = []
    for b in range(P_BEADS):
        top, pos, H = systems.water_box(512, 2.483, seed=20 + b)
        qs.append(pos)
    q = np.stack(qs)
    s = P.MDSettings(precision="mixed", precond="molecule", dipole_tol=1e-7, **CFG1)
    beads = P.PGMBeads(_system(top), H, s, beads=P_BEADS)
    forces, U, e = beads.compute(q, H)
    per, its, res = beads.bead_results()
    for b in range(P_BEADS):
        r = P.PGMForceField(_system(top), H, s).compute(q[b], H)
        assert its[b] == r.iterations, (b, its[b], r.iterations)
        assert abs(per[b]["total"] - r.energy["total"]) <= 1e-7 * abs(r.energy["total"])
        assert _rel_rms(forces[b].cpu().numpy(), r.forces.cpu().numpy()) < 1e-6


def test_molecule_blocks_need_three_site_molecules():
    import pgm_jax_amd as P

    top, pos, H = systems.small_box(0)
    with pytest.raises(P.PGMError):
        P.PGMForceField(_system(top), H, P.MDSettings(cutoff=0.6, ewald_beta=6.0, pme_grid=(32,) * 3, skin=0.05,
                                                      precond="molecule"))
