"""MD state across calls: neighbour rows, captured graphs and the solver normaliser (GPU).

* Many short pgm_md_run calls must follow the same trajectory as one long call: the displacement
  test that triggers a row rebuild also runs at the first step of every call.
* Captured CG iterations of the per-atom PME path (grids too small for tiles) and of the Ewald
  path hold the positions' address: compute() with positions at a new address must not replay them.
* A Monte Carlo trial energy (pgm_trial_energy) tests convergence against its own geometry's
  normaliser and leaves the state's in place (pgm_jax/md/forcefield.py:2612-2624): an MD run after
  a trial equals one without it, normaliser included (the trial's row build at the scaled
  positions must not leave a displacement reading that triggers an extra rebuild either).
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _water(P, nw=512, edge=2.483, seed=5):
    sys_, pos, H = P.models.water_box(nw, edge, seed=seed)
    vel = P.models.maxwell_boltzmann_rigid(pos, sys_.masses, 298.0, seed=6)
    return sys_, pos, vel, H


def test_many_short_runs_equal_one_long_run():
    """40 calls of one step against one call of 40 steps, skin 0.04 nm (rebuild at 0.01 nm): the same
    rebuilds and the same positions and energies, bit for bit."""
    import pgm_jax_amd as P

    sys_, pos, vel, H = _water(P)
    kw = dict(precision="mixed", predictor="mu4", pme_grid=(32, 32, 32), dipole_tol=1e-7, max_iter=300, skin=0.04)
    print("md calls: 40 x run(1) against run(40), skin 0.04", flush=True)
    ffa = P.PGMForceField(sys_, H, P.MDSettings(**kw))
    mda = P.RigidWaterMD(ffa, pos, vel)
    ra = mda.run(40, 1.0)
    ffb = P.PGMForceField(sys_, H, P.MDSettings(**kw))
    mdb = P.RigidWaterMD(ffb, pos, vel)
    rebuilds = 0
    for _ in range(40):
        rb = mdb.run(1, 1.0)
        rebuilds += rb["row_rebuilds"]
    print("one call:", ra["row_rebuilds"], "rebuilds, Etot", ra["etot"], "; 40 calls:", rebuilds, "rebuilds, Etot", rb["etot"],
          flush=True)
    assert ra["row_rebuilds"] > 0
    assert rebuilds == ra["row_rebuilds"]
    assert rb["etot"] == ra["etot"]
    assert np.array_equal(mdb.pos.cpu().numpy(), mda.pos.cpu().numpy())


@pytest.mark.parametrize("recip,grid", [("pme", 12), ("ewald", None)])
def test_graphs_with_new_position_tensors(recip, grid):
    """compute() with graphs on, at two position tensors of the same box (a fresh device tensor per
    call), against the eager context: FP64 energies to 1e-12 relative (the per-atom spread adds
    atomically, in no fixed order), the same CG iteration counts."""
    import pgm_jax_amd as P
    import torch

    sys_, pos, vel, H = _water(P, nw=64, edge=1.2416, seed=9)
    kw = dict(precision="double", predictor="mu4", dipole_tol=1e-9, max_iter=300, cutoff=0.55, skin=0.05, ewald_beta=5.0,
              recip=recip)
    if grid:
        kw["pme_grid"] = (grid,) * 3
    pos2 = pos + np.random.default_rng(3).normal(scale=2e-3, size=pos.shape)
    print(f"graphs with new position tensors, {recip}", flush=True)
    out = []
    for graphs in (False, True):
        ff = P.PGMForceField(sys_, H, P.MDSettings(**kw))
        if graphs:
            ff.set_graphs(True)
        r1 = ff.compute(torch.as_tensor(pos, device="cuda"), H)
        r2 = ff.compute(torch.as_tensor(pos2, device="cuda"), H, None, r1.induction)
        r3 = ff.compute(torch.as_tensor(pos, device="cuda"), H, None, r2.induction)
        out.append([(r.energy["total"], r.iterations) for r in (r1, r2, r3)])
    print(out, flush=True)
    for (ee, ie), (eg, ig) in zip(*out):
        assert ie == ig
        assert abs(ee - eg) <= 1e-12 * abs(ee)


def test_trial_energy_leaves_the_normaliser():
    """MD after a trial energy at a 1 % larger box (then the box restored) against MD after the same
    box changes without the trial: identical positions, energies, iteration counts and normaliser."""
    import pgm_jax_amd as P

    sys_, pos, vel, H = _water(P)
    kw = dict(precision="mixed", predictor="mu4", pme_grid=(32, 32, 32), dipole_tol=1e-7, max_iter=300)
    H2 = H * 1.01
    print("trial energy then md", flush=True)
    res = []
    for trial in (False, True):
        ff = P.PGMForceField(sys_, H, P.MDSettings(**kw))
        md = P.RigidWaterMD(ff, pos, vel)
        md.run(10, 1.0)
        x2 = md.pos * 1.01
        if trial:
            e, _, it, _ = ff.energy(x2, H2, None, md.ind)
            print("trial energy", e, it, flush=True)
        else:
            ff._box(H2)
        ff._box(H)
        r = md.run(5, 1.0)
        res.append((r["etot"], r["cg_iterations"], md.ind.norm, md.pos.cpu().numpy()))
    print([x[:3] for x in res], flush=True)
    assert res[0][:3] == res[1][:3]
    assert np.array_equal(res[0][3], res[1][3])
