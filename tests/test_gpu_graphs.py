"""CG iterations replayed as hipGraphs reproduce the eager run bit for bit (GPU).

With set_graphs(True) every CG iteration index is captured once and replayed with hipGraphLaunch,
between the eager kernels of the rest of each step and the host's convergence reads.  The graphs
hold the same kernels with the same arguments in the same order, so an MD run with graphs must give
the eager run's positions, energies and iteration counts exactly, after every block of steps (each
block replays the graphs many times with eager work and synchronisations in between).  A change of
box drops the graphs; the next solve captures them again.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _md(P, sys_, pos, vel, H, precision, graphs, precond="jacobi"):
    ff = P.PGMForceField(sys_, H, P.MDSettings(precision=precision, predictor="mu4", precond=precond, pme_grid=(32, 32, 32),
                                               dipole_tol=1e-7, max_iter=300))
    if graphs:
        ff.set_graphs(True)
    return ff, P.RigidWaterMD(ff, pos, vel)


@pytest.mark.parametrize("precision,precond", [("double", "jacobi"), ("mixed", "jacobi"), ("mixed", "molecule")])
def test_graph_md_equals_eager_md(precision, precond):
    import pgm_jax_amd as P

    sys_, pos, H = P.models.water_box(512, 2.483, seed=5)
    vel = P.models.maxwell_boltzmann_rigid(pos, sys_.masses, 298.0, seed=6)
    print(f"graphs precision={precision} precond={precond}", flush=True)
    ffe, eager = _md(P, sys_, pos, vel, H, precision, False, precond)
    ffg, graph = _md(P, sys_, pos, vel, H, precision, True, precond)
    for block in range(4):
        re = eager.run(5, 1.0)
        rg = graph.run(5, 1.0)
        print(f"block {block}: eager {re['cg_iterations']} it E {re['etot']:.9f}, graphs {rg['cg_iterations']} it "
              f"E {rg['etot']:.9f}", flush=True)
        assert rg["cg_iterations"] == re["cg_iterations"]
        assert rg["etot"] == re["etot"]
        assert np.array_equal(graph.pos.cpu().numpy(), eager.pos.cpu().numpy())


def test_graphs_follow_a_box_change():
    """After a new box the graphs are captured again: the solve still equals the eager one."""
    import pgm_jax_amd as P

    sys_, pos, H = P.models.water_box(512, 2.483, seed=7)
    H2 = H * 1.01
    res = []
    for graphs in (False, True):
        ff = P.PGMForceField(sys_, H, P.MDSettings(precision="mixed", pme_grid=(32, 32, 32), dipole_tol=1e-7))
        if graphs:
            ff.set_graphs(True)
        r1 = ff.compute(pos, H)
        r2 = ff.compute(pos * 1.01, H2, None, r1.induction)
        res.append((r1.energy["total"], r1.iterations, r2.energy["total"], r2.iterations))
    assert res[0] == res[1], res
