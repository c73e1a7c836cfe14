"""Neighbour rows of a sheared (triclinic) box against the oracle (GPU).

With at least three cells per axis the cell-list row build wraps positions into the box and adds
one lattice shift per neighbour cell instead of a minimum image per pair; with fewer cells it keeps
the per-pair minimum image.  A 512-water box sheared into a reduced triclinic cell (rc 0.6 nm +
skin 0.05 nm: three cells per axis) takes the shifted path: a pair missed or counted twice would
move the FP64 energy far beyond 1e-9 relative, and the CG iteration count is the oracle's.  The
same box with rc 0.9 nm (two cells per axis) takes the minimum-image path.
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


def _sheared():
    top, pos, H = systems.water_box(512, 2.483, seed=12)
    L = H[0, 0]
    Ht = np.array([[L, 0.0, 0.0], [0.3 * L, L, 0.0], [-0.2 * L, 0.25 * L, L]])
    u = pos @ np.linalg.inv(H)
    return top, u @ Ht, Ht


@pytest.mark.parametrize("cutoff,skin", [(0.6, 0.05), (0.9, 0.1)])
def test_triclinic_rows_match_oracle(cutoff, skin):
    import pgm_jax_amd as P

    top, pos, H = _sheared()
    kw = dict(cutoff=cutoff, ewald_beta=5.0, pme_grid=(32, 32, 32), pme_order=6, lj_lrc=True, dipole_tol=1e-9,
              max_iter=500)
    ref = ForceField(top, H, Settings(**kw)).compute(pos, H)
    ff = P.PGMForceField(_system(P, top), H, P.MDSettings(skin=skin, precision="double", **kw))
    res = ff.compute(pos, H)
    e, eref = res.energy["total"], ref.energy["total"]
    F = res.forces.cpu().numpy()
    ferr = float(np.sqrt(np.mean((F - ref.forces) ** 2)) / np.sqrt(np.mean(ref.forces**2)))
    print(f"triclinic rc {cutoff}: E {e:.9f} (oracle {eref:.9f}), force rms rel {ferr:.2e}, "
          f"CG {res.iterations} (oracle {ref.iterations})", flush=True)
    assert abs(e - eref) <= 1e-9 * abs(eref)
    assert ferr < 1e-8
    assert res.iterations == ref.iterations
