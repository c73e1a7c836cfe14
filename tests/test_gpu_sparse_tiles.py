"""A water cluster in a box mostly of vacuum against the oracle (GPU).

216 waters on a 1.87 nm lattice inside a 3.6 nm cubic box: with a 32^3 grid most spread tiles
(8^3 points, 4 per axis) and gather tiles hold no atom, and the sorted tile keys jump over runs of
empty tiles.  Every tile's atom range comes from those sorted keys, so an empty tile that received
a stale or overlapping range would add atoms twice or drop them: the FP64 energy would move far
beyond 1e-9 relative of the oracle's, and the CG iteration count would differ.
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


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_cluster_in_vacuum_matches_oracle(precision):
    import pgm_jax_amd as P

    top, pos, _ = systems.water_box(216, 1.87, seed=21)
    pos = pos + 0.3  # the cluster away from the box faces, inside [0, 3.6) nm
    H = np.diag([3.6, 3.6, 3.6])
    kw = dict(cutoff=0.9, ewald_beta=4.0, pme_grid=(32, 32, 32), pme_order=6, lj_lrc=True, dipole_tol=1e-9,
              max_iter=500)
    ref = ForceField(top, H, Settings(**kw)).compute(pos, H)
    if precision == "mixed":
        kw["dipole_tol"] = 1e-7
    ff = P.PGMForceField(_system(P, top), H, P.MDSettings(precision=precision, **kw))
    res = ff.compute(pos, H)
    e, eref = res.energy["total"], ref.energy["total"]
    F = res.forces.cpu().numpy()
    ferr = float(np.sqrt(np.mean((F - ref.forces) ** 2)) / np.sqrt(np.mean(ref.forces**2)))
    print(f"cluster {precision}: E {e:.9f} (oracle {eref:.9f}), force rms rel {ferr:.2e}, "
          f"CG {res.iterations} (oracle {ref.iterations})", flush=True)
    if precision == "double":
        assert abs(e - eref) <= 1e-9 * abs(eref)
        assert ferr < 1e-8
        assert res.iterations == ref.iterations
    else:
        assert abs(e - eref) <= 1e-6 * abs(eref)
        assert ferr < 1e-5
