"""The C++/OpenMP float64 restatement (oracle/cpu_ref.cpp, the CPU baseline) against the numpy oracle (CPU).

It must reproduce the oracle (itself pinned by the reference's goldens) to rounding: energies
1e-11 relative, forces 1e-8 of the largest, dipoles 1e-9, the same CG iteration counts; the
golden ff_pgm3p25_512 double single point directly; split cutoffs; and one SHAKE/RATTLE step.
"""

import os

import numpy as np
import pytest

from oracle import md as omd
from oracle import systems
from oracle.forcefield import ForceField, Settings

GOLD = os.path.join(os.path.dirname(__file__), "golden")
LIB = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "libcpu_ref.so")
pytestmark = pytest.mark.skipif(not os.path.exists(LIB), reason="oracle/libcpu_ref.so not built")


def _cmp(c, ref, rtol_e=1e-11):
    for k in ("elec", "vdw", "total"):
        assert abs(c["energy"][k] - ref.energy[k]) <= rtol_e * max(1.0, abs(ref.energy[k])), (k, c["energy"][k], ref.energy[k])
    assert np.abs(c["forces"] - ref.forces).max() <= 1e-8 * np.abs(ref.forces).max()
    assert np.abs(c["mu"] - ref.mu).max() <= 1e-9 * np.abs(ref.mu).max()
    assert c["iterations"] == ref.iterations


def test_small_box_matches_oracle():
    """The triclinic water / methanol box of the golden ff_small_box (beta 6: the reciprocal sum's
    Nyquist planes, whose frequency labels of a triclinic box the two restatements alias
    differently, hold exp(-31) of the weight)."""
    from oracle.cpu_ref import CpuRef

    top, pos, H = systems.small_box(0)
    s = Settings(cutoff=0.6, ewald_beta=6.0, pme_grid=(32, 32, 32), pme_order=6, lj_lrc=True, dipole_tol=1e-10,
                 max_iter=500, predictor="none")
    _cmp(CpuRef(top, H, s).compute(pos), ForceField(top, H, s).compute(pos, H))


def test_split_cutoffs_match_oracle():
    """64 waters in a cubic box, LJ at 0.6 nm, electrostatics at 0.45 nm (beta 7.5)."""
    from oracle.cpu_ref import CpuRef

    top, pos, H = systems.water_box(64, 1.2416, seed=3)
    s = Settings(cutoff=0.6, elec_cutoff=0.45, ewald_beta=7.5, pme_grid=(20, 20, 20), pme_order=6, lj_lrc=True,
                 dipole_tol=1e-10, max_iter=500, predictor="none")
    _cmp(CpuRef(top, H, s).compute(pos), ForceField(top, H, s).compute(pos, H))


def test_water512_matches_golden_double():
    """The truncated octahedron of ff_pgm3p25_512: energies to 1e-10 relative (2e-11 measured: the
    triclinic Nyquist-plane labels, see above, at beta 4 and a 30^3 grid)."""
    from oracle.cpu_ref import CpuRef

    top, pos, H, _ = systems.water512()
    c = CpuRef(top, H, Settings(predictor="none")).compute(pos)
    with np.load(os.path.join(GOLD, "ff_pgm3p25_512.npz")) as g:
        for k in ("elec", "vdw", "total"):
            assert abs(c["energy"][k] - float(g[f"double.energy.{k}"])) <= 1e-10 * abs(float(g[f"double.energy.{k}"]))
        F = g["double.forces"]
        assert np.abs(c["forces"] - F).max() <= 1e-8 * np.abs(F).max()
        assert c["iterations"] == int(g["double.iterations"])


def test_md_step_matches_oracle_step():
    """Two NVE steps of 64 waters (mu4 predictor from the first solve on): positions to 1e-12 nm."""
    from oracle.cpu_ref import CpuRef

    top, pos, H = systems.water_box(64, 1.2416, seed=3)
    import importlib

    P = importlib.import_module("pgm_jax_amd")
    vel = P.models.maxwell_boltzmann_rigid(pos, top.masses, 298.0, seed=4)
    s = Settings(cutoff=0.55, ewald_beta=5.0, pme_grid=(16, 16, 16), dipole_tol=1e-10, max_iter=300)
    ff = ForceField(top, H, s)
    st = {"ind": ff.init_induction()}

    def fn(x):
        r = ff.compute(x, H, st["ind"])
        st["ind"] = r.induction
        return r.forces

    d2 = omd.constraint_lengths2(pos)
    p1, v1, F1 = pos.copy(), vel.copy(), fn(pos)
    for _ in range(2):
        p1, v1, F1 = omd.step(p1, v1, F1, top.masses, d2, 1e-3, fn)
    c = CpuRef(top, H, s)
    p2, v2 = np.ascontiguousarray(pos.copy()), np.ascontiguousarray(vel.copy())
    F2 = np.ascontiguousarray(c.compute(p2)["forces"])
    rep = c.md(p2, v2, F2, top.masses, np.ascontiguousarray(d2), 2, 1e-3)
    assert np.abs(p2 - p1).max() < 1e-12
    assert np.abs(v2 - v1).max() < 1e-8 * np.abs(v1).max()
    assert np.abs(F2 - F1).max() < 1e-8 * np.abs(F1).max()
    assert rep["cg_iterations"] > 0
