"""Rigid 3-site molecules: velocity Verlet with SHAKE/RATTLE in float64 (numpy, vectorised over molecules).

Restates the constrained step the MD engine takes for rigid water (pgm_jax/md/flexible.py:776-799
kick-drift with pgm_jax/md/constraints.py:254-299: three distance constraints O-H1, O-H2, H1-H2 per
molecule, Newton iterations on the Lagrange multipliers with 3x3 solves; RATTLE projects the
velocities tangent to the constraints).

Parity: pinned by invariants in tests/test_oracle_md.py -- constraints hold to 1e-12 after every
step, velocities are tangent to them, and NVE total energy with a smooth test potential drifts by
O(dt^2) (halving dt quarters the energy error); the reference's own rigid-body (NO_SQUISH) engine
is not restated, so trajectories are compared by energy conservation only ("parity unpinned" for
trajectory-level agreement with the reference's JAX-MD integrator).
"""

from __future__ import annotations

import numpy as np

CA = (0, 0, 1)
CB = (1, 2, 2)


def _g(a: int, c: int) -> float:
    return 1.0 if CA[c] == a else (-1.0 if CB[c] == a else 0.0)


def shake(r0: np.ndarray, r1: np.ndarray, minv: np.ndarray, d2: np.ndarray, tol: float = 1e-14, max_iter: int = 50):
    """Return (constrained r1 (M,3,3), displacement (M,3,3), iterations) for old positions r0 (M,3,3).

    Displacement of atom a: sum_c lam_c G(a, c) s_c / m_a with s_c = r0_A - r0_B; Newton on
    sigma_c = |x_c|^2 - d_c^2 in molecule-local coordinates.
    """
    org = r0[:, :1]
    r0 = r0 - org
    r1 = r1 - org
    s = np.stack([r0[:, CA[c]] - r0[:, CB[c]] for c in range(3)], 1)  # (M, 3 constraints, 3)
    lam = np.zeros(d2.shape)
    G = np.array([[_g(a, c) for c in range(3)] for a in range(3)])  # (atom, constraint)
    for it in range(max_iter):
        disp = np.einsum("ac,mc,mck->mak", G, lam, s) * minv[:, :, None]
        rr = r1 + disp
        x = np.stack([rr[:, CA[c]] - rr[:, CB[c]] for c in range(3)], 1)
        sig = np.sum(x * x, -1) - d2
        if np.max(np.abs(sig) / d2) < tol:
            return rr + org, disp, it
        w = np.stack([[G[CA[c], cp] * minv[:, CA[c]] - G[CB[c], cp] * minv[:, CB[c]] for cp in range(3)] for c in range(3)],
                     0).transpose(2, 0, 1)  # (M, c, c')
        J = 2.0 * w * np.einsum("mck,mpk->mcp", x, s)
        lam = lam + np.linalg.solve(J, -sig[..., None])[..., 0]
    raise RuntimeError("SHAKE did not converge")


def rattle(r: np.ndarray, v: np.ndarray, minv: np.ndarray) -> np.ndarray:
    """Return v (M,3,3) projected tangent to the constraints of positions r (M,3,3)."""
    s = np.stack([r[:, CA[c]] - r[:, CB[c]] for c in range(3)], 1)
    G = np.array([[_g(a, c) for c in range(3)] for a in range(3)])
    w = np.stack([[G[CA[c], cp] * minv[:, CA[c]] - G[CB[c], cp] * minv[:, CB[c]] for cp in range(3)] for c in range(3)],
                 0).transpose(2, 0, 1)
    K = w * np.einsum("mck,mpk->mcp", s, s)
    rhs = -np.stack([np.sum((v[:, CA[c]] - v[:, CB[c]]) * s[:, c], -1) for c in range(3)], 1)
    mu = np.linalg.solve(K, rhs[..., None])[..., 0]
    return v + np.einsum("ac,mc,mck->mak", G, mu, s) * minv[:, :, None]


def step(pos: np.ndarray, vel: np.ndarray, forces: np.ndarray, masses: np.ndarray, d2: np.ndarray, dt: float, force_fn):
    """One velocity-Verlet step with SHAKE/RATTLE; returns (pos, vel, forces) (N,3) each."""
    M = len(pos) // 3
    minv = (1.0 / masses).reshape(M, 3)
    r0 = pos.reshape(M, 3, 3)
    v = vel.reshape(M, 3, 3) + 0.5 * dt * forces.reshape(M, 3, 3) * minv[:, :, None]
    r1 = r0 + dt * v
    r1, disp, _ = shake(r0, r1, minv, d2)
    v = v + disp / dt
    f = force_fn(r1.reshape(-1, 3))
    v = v + 0.5 * dt * f.reshape(M, 3, 3) * minv[:, :, None]
    v = rattle(r1, v, minv)
    return r1.reshape(-1, 3), v.reshape(-1, 3), f


def constraint_lengths2(pos: np.ndarray) -> np.ndarray:
    """Return the squared O-H1, O-H2, H1-H2 distances (M, 3) of consecutive 3-site molecules."""
    r = pos.reshape(-1, 3, 3)
    return np.stack([np.sum((r[:, CA[c]] - r[:, CB[c]]) ** 2, -1) for c in range(3)], 1)


# ---------------------------------------------------------------------------- Langevin (g-BAOAB)
PHILOX_M = (0xD2511F53, 0xCD9E8D57)
PHILOX_W = (0x9E3779B9, 0xBB67AE85)


def philox4x32_10(ctr: np.ndarray, key: tuple[int, int]) -> np.ndarray:
    """Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC 2011) of counters (..., 4) uint32 and key (k0, k1).

    Pinned by the known-answer vectors of the Random123 distribution (tests/test_oracle_md.py).
    """
    c = np.array(ctr, dtype=np.uint64) & 0xFFFFFFFF
    k0, k1 = int(key[0]) & 0xFFFFFFFF, int(key[1]) & 0xFFFFFFFF
    mask = np.uint64(0xFFFFFFFF)
    for _ in range(10):
        p0 = np.uint64(PHILOX_M[0]) * c[..., 0]
        p1 = np.uint64(PHILOX_M[1]) * c[..., 2]
        hi0, lo0 = p0 >> np.uint64(32), p0 & mask
        hi1, lo1 = p1 >> np.uint64(32), p1 & mask
        c = np.stack([hi1 ^ c[..., 1] ^ np.uint64(k0), lo1, hi0 ^ c[..., 3] ^ np.uint64(k1), lo0], -1)
        k0 = (k0 + PHILOX_W[0]) & 0xFFFFFFFF
        k1 = (k1 + PHILOX_W[1]) & 0xFFFFFFFF
    return c.astype(np.uint32)


def langevin_normals(nmol: int, step: int, seed: int) -> np.ndarray:
    """Return the (nmol, 3 atoms, 3) standard normals of the O step: Philox counter (m, step, k, 0),
    k = 0..2, keyed by the 64-bit seed; Box-Muller on pairs of uniforms (x + 1/2) / 2^32."""
    m = np.arange(nmol, dtype=np.uint64)
    g = []
    for k in range(3):
        ctr = np.stack([m, np.full(nmol, step & 0xFFFFFFFF, np.uint64), np.full(nmol, k, np.uint64),
                        np.zeros(nmol, np.uint64)], -1)
        u = (philox4x32_10(ctr, (seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF)).astype(np.float64) + 0.5) * 2.0**-32
        r0, t0 = np.sqrt(-2.0 * np.log(u[:, 0])), 2.0 * np.pi * u[:, 1]
        r1, t1 = np.sqrt(-2.0 * np.log(u[:, 2])), 2.0 * np.pi * u[:, 3]
        g.append(np.stack([r0 * np.cos(t0), r0 * np.sin(t0), r1 * np.cos(t1), r1 * np.sin(t1)], -1))
    return np.concatenate(g, -1)[:, :9].reshape(nmol, 3, 3)


def langevin_step(pos, vel, forces, masses, d2, dt, friction, kT, seed, step, force_fn):
    """One constrained g-BAOAB step (pgm_jax/md/flexible.py:776-799 with the O step of
    md/thermostats.py:195-242): kick dt/2; drift dt/2 + SHAKE, v = dq/(dt/2), RATTLE; O step
    v <- c v + sqrt(kT (1 - c^2)/m) xi, c = exp(-friction dt), RATTLE; drift dt/2 + SHAKE,
    v = dq/(dt/2); forces; kick dt/2 + RATTLE.  Returns (pos, vel, forces, heat)."""
    M = len(pos) // 3
    minv = (1.0 / masses).reshape(M, 3)
    m3 = masses.reshape(M, 3)
    h = 0.5 * dt
    r0 = pos.reshape(M, 3, 3)
    v = vel.reshape(M, 3, 3) + 0.5 * dt * forces.reshape(M, 3, 3) * minv[:, :, None]
    r1, disp, _ = shake(r0, r0 + h * v, minv, d2)
    v = rattle(r1, v + disp / h, minv)
    c = np.exp(-friction * dt)
    e0 = 0.5 * np.sum(m3[:, :, None] * v * v)
    v = c * v + np.sqrt(kT * (1.0 - c * c) * minv)[:, :, None] * langevin_normals(M, step, seed)
    v = rattle(r1, v, minv)
    e1 = 0.5 * np.sum(m3[:, :, None] * v * v)
    r2, disp, _ = shake(r1, r1 + h * v, minv, d2)
    v = v + disp / h
    f = force_fn(r2.reshape(-1, 3))
    v = rattle(r2, v + 0.5 * dt * f.reshape(M, 3, 3) * minv[:, :, None], minv)
    return r2.reshape(-1, 3), v.reshape(-1, 3), f, e1 - e0
