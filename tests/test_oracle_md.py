"""The oracle's SHAKE/RATTLE step against invariants (CPU)."""

import numpy as np

from oracle import md, systems


def _harmonic(centers, k=500.0):
    return lambda x: -k * (x - centers)


def test_constraints_and_tangent_velocities_hold():
    top, pos, H = systems.water_box(27, 0.93, seed=1)
    rng = np.random.default_rng(0)
    vel = rng.normal(scale=0.3, size=pos.shape)
    M = len(pos) // 3
    vel = md.rattle(pos.reshape(M, 3, 3), vel.reshape(M, 3, 3), (1 / top.masses).reshape(M, 3)).reshape(-1, 3)
    d2 = md.constraint_lengths2(pos)
    f = _harmonic(pos + 0.01)
    F = f(pos)
    for _ in range(20):
        pos, vel, F = md.step(pos, vel, F, top.masses, d2, 1e-3, f)
        assert np.abs(md.constraint_lengths2(pos) / d2 - 1).max() < 1e-12
        r = pos.reshape(M, 3, 3)
        v = vel.reshape(M, 3, 3)
        for c in range(3):
            s = r[:, md.CA[c]] - r[:, md.CB[c]]
            assert np.abs(np.sum((v[:, md.CA[c]] - v[:, md.CB[c]]) * s, -1)).max() < 1e-12


def test_energy_error_is_second_order_in_dt():
    """NVE with a smooth potential: halving dt quarters the energy fluctuation (velocity Verlet)."""
    top, pos0, H = systems.water_box(8, 0.62, seed=2)
    k = 2000.0
    c = pos0 + 0.02
    f = _harmonic(c, k)
    M = len(pos0) // 3
    rng = np.random.default_rng(3)
    v0 = md.rattle(pos0.reshape(M, 3, 3), rng.normal(scale=0.5, size=(M, 3, 3)), (1 / top.masses).reshape(M, 3)).reshape(-1, 3)
    d2 = md.constraint_lengths2(pos0)

    def spread(dt, T=0.2):
        pos, vel, F = pos0.copy(), v0.copy(), f(pos0)
        es = []
        for _ in range(int(round(T / dt))):
            pos, vel, F = md.step(pos, vel, F, top.masses, d2, dt, f)
            es.append(0.5 * k * np.sum((pos - c) ** 2) + 0.5 * np.sum(top.masses[:, None] * vel**2))
        return np.ptp(es)

    e1, e2 = spread(2e-3), spread(1e-3)
    assert 3.0 < e1 / e2 < 5.0, (e1, e2)
