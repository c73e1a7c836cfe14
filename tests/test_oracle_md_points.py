"""The oracle at the snapshots of the reference's rigid-water trajectories (CPU).

md_rigid_nve and md_rigid_npt (tests/regression/regression_cases.py:356-411): 64 pGM3P-25 waters in
a 1.24 nm box (the NPT run's box changes), float64 with the regression harness's tight settings
(cutoff 0.55 nm, CG to 1e-10, max 300 iterations, tests/regression/regression_cases.py:92-96) and
the MDSettings defaults otherwise (beta 4.0 /nm, PME spacing 0.08 nm -> 16^3, order 6, LJ tail
correction, peek 0.65).  Every 50 steps the reference recorded the positions, box, induced dipoles
and energies; each snapshot is a single point at a geometry the other fixtures do not reach.  The
dipoles there came from the MD's predictor-started CG, so a fresh solve agrees to the CG tolerance:
dipoles within 1e-9 of the largest, energies (variational in the dipoles) within 1e-12 relative
(measured: 2e-11 and 3e-15).
"""

import os

import numpy as np
import pytest

from oracle import systems
from oracle.forcefield import ForceField, Settings

GOLD = os.path.join(os.path.dirname(__file__), "golden", "md_rigid_points.npz")
TIGHT = dict(cutoff=0.55, dipole_tol=1e-10, max_iter=300, predictor="none")
POINTS = [f"{run}.t{s:04d}" for run in ("nve", "npt") for s in range(0, 201, 50)]


def load(tag):
    with np.load(GOLD) as z:
        return {k: z[f"{tag}.{k}"] for k in ("pos", "box", "mu", "obs.epot", "obs.elec", "obs.vdw")}


def water64():
    return systems.Topology.from_molecules([systems.pgm3p25_water()] * 64)


@pytest.mark.parametrize("tag", POINTS)
def test_oracle_at_trajectory_snapshots(tag):
    g = load(tag)
    H = np.asarray(g["box"], dtype=np.float64)
    ff = ForceField(water64(), H, Settings(**TIGHT))
    assert ff.pme.K == (16, 16, 16)
    res = ff.compute(g["pos"], H)
    for k, ref in (("elec", g["obs.elec"]), ("vdw", g["obs.vdw"]), ("total", g["obs.epot"])):
        assert abs(res.energy[k] - float(ref)) <= 1e-12 * abs(float(ref)), (tag, k, res.energy[k], float(ref))
    assert np.abs(res.mu - g["mu"]).max() <= 1e-9 * np.abs(g["mu"]).max(), tag
