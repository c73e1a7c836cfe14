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


# Known-answer vectors of Philox4x32-10 from the Random123 distribution (kat_vectors):
# (counter, key) -> output.
PHILOX_KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF, 0xFFFFFFFF), (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0), (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


def test_philox_known_answers():
    for ctr, key, want in PHILOX_KAT:
        assert tuple(int(x) for x in md.philox4x32_10(np.array(ctr), key)) == want


def test_langevin_normals_are_standard():
    g = np.concatenate([md.langevin_normals(20000, s, seed=7).ravel() for s in range(3)])
    assert abs(g.mean()) < 5e-3 and abs(g.std() - 1.0) < 5e-3
    assert abs(np.mean(g**4) - 3.0) < 0.05  # Gaussian kurtosis
    a, b = md.langevin_normals(1000, 0, 7), md.langevin_normals(1000, 1, 7)
    assert abs(np.corrcoef(a.ravel(), b.ravel())[0, 1]) < 0.03  # steps independent


def test_langevin_thermalises_free_rigid_water():
    """Force-free rigid water under g-BAOAB: constraints hold, temperature -> target over 6 dof per molecule,
    and the heat booked by the O steps equals the kinetic-energy change."""
    top, pos, H = systems.water_box(216, 1.86, seed=4)
    M = len(pos) // 3
    vel = np.zeros_like(pos)
    d2 = md.constraint_lengths2(pos)
    kT = 0.0083144626181532 * 300.0
    zero = lambda x: np.zeros_like(x)  # noqa: E731
    F = zero(pos)
    temps, heat = [], 0.0
    ke0 = 0.0
    for s in range(600):
        pos, vel, F, q = md.langevin_step(pos, vel, F, top.masses, d2, 2e-3, 10.0, kT, 11, s, zero)
        heat += q
        assert np.abs(md.constraint_lengths2(pos) / d2 - 1).max() < 1e-12
        ke = 0.5 * np.sum(top.masses[:, None] * vel**2)
        if s >= 200:
            temps.append(2 * ke / (6 * M * 0.0083144626181532))
    assert abs(np.mean(temps) - 300.0) < 12.0, np.mean(temps)
    assert abs(heat - (ke - ke0)) < 1e-9 * abs(heat)  # no forces: only the O steps change the kinetic energy


def test_water512_ten_steps_known_answer():
    """examples/water512, 10 NVE steps of 1 fs with the reference's MD defaults: Epot(10) and T(10) against the
    reference's recorded run (tests/test_water512.py:26-27: -2114884.256504 kJ/mol, 291.356783 K).

    The reference integrates rigid bodies (NO_SQUISH); this is SHAKE/RATTLE, so agreement is to the
    integrator difference: 5e-7 relative in Epot, 1e-4 in T (measured 1.5e-7 and 4e-5)."""
    from oracle.forcefield import ForceField, Settings

    top, pos, H, vel = systems.water512()
    ff = ForceField(top, H, Settings())
    state = {"ind": ff.init_induction()}

    def fn(x):
        r = ff.compute(x, H, state["ind"])
        state["ind"], state["e"] = r.induction, r.energy["total"]
        return r.forces

    M = len(pos) // 3
    minv = (1 / top.masses).reshape(M, 3)
    v = md.rattle(pos.reshape(M, 3, 3), vel.reshape(M, 3, 3), minv).reshape(-1, 3)
    d2 = md.constraint_lengths2(pos)
    F = fn(pos)
    for _ in range(10):
        pos, v, F = md.step(pos, v, F, top.masses, d2, 1e-3, fn)
    T = 2 * 0.5 * np.sum(top.masses[:, None] * v**2) / ((6 * M - 3) * 0.0083144626181532)
    assert abs(state["e"] - (-2114884.256504)) < 5e-7 * 2114884.256504, state["e"]
    assert abs(T - 291.356783) < 1e-4 * 291.356783, T
