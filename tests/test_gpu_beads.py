"""Beads in one context against one context per bead (GPU).

PGMBeads runs P copies of a system in the same launches with bead-wise FP64 CG reductions; every
bead's energies, forces, induced dipoles and CG iteration count must match a single-system
PGMForceField on the same positions, whose path is pinned to the reference's goldens and the oracle
(test_gpu_forcefield.py).  The beads differ (different seeds), so they converge after different
iteration counts.  Tolerances: double 1e-10 relative in energy and forces (the batched rocFFT plan
may order its sums differently from the single plan), mixed 1e-7 relative in energy and 1e-6
relative RMS in forces and dipoles; iterations equal.
"""

import numpy as np
import pytest

from oracle import systems

pytestmark = pytest.mark.gpu

CFG1 = dict(cutoff=0.9, ewald_beta=4.0, pme_grid=(32, 32, 32), pme_order=6, dipole_tol=1e-7, skin=0.1)


def _system(top):
    import pgm_jax_amd as P

    return P.System(q=top.q, radius=top.radius, alpha=top.alpha, lj_rmin_half=top.lj_rmin_half,
                    lj_sqrt_eps=top.lj_sqrt_eps, masses=top.masses, mol=top.mol, cov_i=top.cov_i,
                    cov_j=top.cov_j, cov_c=top.cov_c)


def _rel_rms(a, b):
    return float(np.sqrt(np.mean((a - b) ** 2)) / np.sqrt(np.mean(b**2)))


def _beads(n_waters, edge, P, seed0=0):
    tops, qs = None, []
    for b in range(P):
        top, pos, H = systems.water_box(n_waters, edge, seed=seed0 + b)
        tops = top
        qs.append(pos)
    return tops, np.stack(qs), H


def _compare(beads, singles, q, H, e, inds, precision, step):
    forces, U, e = beads.compute(q, H, e)
    per, its, res = beads.bead_results()
    F = forces.cpu().numpy()
    mu = e.induction.mu.cpu().numpy()
    et = 1e-10 if precision == "double" else 1e-7
    ft = 1e-10 if precision == "double" else 1e-6
    total = 0.0
    for b, ff in enumerate(singles):
        r = ff.compute(q[b], H, None, inds[b])
        inds[b] = r.induction
        print(f"step {step} bead {b}: it {its[b]} vs {r.iterations}, E {per[b]['total']:.9f} vs {r.energy['total']:.9f}",
              flush=True)
        for k in ("elec", "vdw", "total"):
            assert abs(per[b][k] - r.energy[k]) <= et * abs(r.energy[k]), (b, k, per[b][k], r.energy[k])
        assert _rel_rms(F[b], r.forces.cpu().numpy()) < ft, b
        assert _rel_rms(mu[b], r.induction.mu.cpu().numpy()) < ft, b
        assert its[b] == r.iterations, (b, its[b], r.iterations)
        assert res[b] <= 1e-7
        total += r.energy["total"]
    assert abs(U - total) <= et * abs(total)
    assert e.iters == max(its)
    return e


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_beads_match_single_contexts(precision):
    import pgm_jax_amd as P

    P_BEADS = 3
    top, q, H = _beads(512, 2.483, P_BEADS)
    s = P.MDSettings(precision=precision, predictor="mu4", **CFG1)
    print(f"beads P={P_BEADS} n={top.n} precision={precision}", flush=True)
    beads = P.PGMBeads(_system(top), H, s, beads=P_BEADS)
    singles = [P.PGMForceField(_system(top), H, s) for _ in range(P_BEADS)]
    e = beads.init_state()
    inds = [ff.init_induction() for ff in singles]
    rng = np.random.default_rng(7)
    for step in range(3):  # the later calls start from the predictor and the fused residual
        e = _compare(beads, singles, q, H, e, inds, precision, step)
        q = q + rng.normal(scale=2e-4, size=q.shape)


def test_beads_of_the_batched_pimd_size():
    """32 beads of 1024 waters in a 3.94 nm box with a 48^3 PME grid, two beads checked singly."""
    import pgm_jax_amd as P

    P_BEADS = 32
    top, q, H = _beads(1024, 3.94, P_BEADS, seed0=100)
    s = P.MDSettings(precision="mixed", predictor="mu4", cutoff=0.9, ewald_beta=4.0, pme_grid=(48, 48, 48),
                     pme_order=6, dipole_tol=1e-7)
    print(f"beads P={P_BEADS} n={top.n}", flush=True)
    beads = P.PGMBeads(_system(top), H, s, beads=P_BEADS)
    forces, U, e = beads.compute(q, H)
    per, its, res = beads.bead_results()
    assert np.all(res <= 1e-7) and np.all(its > 0)
    assert np.isfinite(forces.cpu().numpy()).all()
    for b in (0, 31):
        r = P.PGMForceField(_system(top), H, s).compute(q[b], H)
        assert abs(per[b]["total"] - r.energy["total"]) <= 1e-7 * abs(r.energy["total"]), b
        assert _rel_rms(forces[b].cpu().numpy(), r.forces.cpu().numpy()) < 1e-6, b
        assert its[b] == r.iterations, (b, its[b], r.iterations)
    assert abs(U - sum(p["total"] for p in per)) <= 1e-12 * abs(U)
