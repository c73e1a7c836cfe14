"""The GPU force field at the snapshots of the reference's rigid-water trajectories (GPU).

The ten single points of tests/test_oracle_md_points.py (md_rigid_nve, md_rigid_npt: 64 pGM3P-25
waters, tests/regression/regression_cases.py:356-411; the NPT snapshots in changed boxes), computed
by one context that moves from box to box: FP64 energies within 1e-12 of the recorded ones and
dipoles within 1e-9 of the largest (the recording solved from the MD's predictor to 1e-10;
measured 3e-15 and 2e-11); mixed precision (FP32 pair and PME kernels) energies within 2e-7
relative and dipoles within 5e-6 (measured 1.0e-7 and 5.9e-7).
"""

import numpy as np
import pytest

from test_oracle_md_points import POINTS, TIGHT, load, water64

pytestmark = pytest.mark.gpu


def _system(top):
    import pgm_jax_amd as P

    return P.System(q=top.q, radius=top.radius, alpha=top.alpha, lj_rmin_half=top.lj_rmin_half,
                    lj_sqrt_eps=top.lj_sqrt_eps, masses=top.masses, mol=top.mol, cov_i=top.cov_i,
                    cov_j=top.cov_j, cov_c=top.cov_c)


@pytest.mark.parametrize("precision,tol_e,tol_mu", [("double", 1e-12, 1e-9), ("mixed", 2e-7, 5e-6)])
def test_gpu_at_trajectory_snapshots(precision, tol_e, tol_mu):
    import pgm_jax_amd as P

    first = load(POINTS[0])
    kw = dict(TIGHT, skin=0.05)
    ff = P.PGMForceField(_system(water64()), np.asarray(first["box"]), P.MDSettings(precision=precision, **kw))
    worst_e, worst_mu = 0.0, 0.0
    for tag in POINTS:
        g = load(tag)
        print(f"snapshot {tag} {precision}", flush=True)
        r = ff.compute(g["pos"], np.asarray(g["box"]))
        for k, ref in (("elec", g["obs.elec"]), ("vdw", g["obs.vdw"]), ("total", g["obs.epot"])):
            err = abs(r.energy[k] - float(ref)) / abs(float(ref))
            worst_e = max(worst_e, err)
            assert err <= tol_e, (tag, k, r.energy[k], float(ref))
        mu = r.induction.mu.cpu().numpy()
        err = np.abs(mu - g["mu"]).max() / np.abs(g["mu"]).max()
        worst_mu = max(worst_mu, err)
        assert err <= tol_mu, (tag, err)
    print(f"  {precision}: energies {worst_e:.2e}, dipoles {worst_mu:.2e}", flush=True)
