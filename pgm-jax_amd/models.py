"""Water models and synthetic boxes (restates pgm_jax/models/toy.py:81-140 for pGM3P-25 water).

`pgm3p25_water()`: the rigid 3-site pGM3P-25 water of examples/water512/pgm3p25_512.prmtop
(charges, Gaussian radii, polarizabilities, covalent dipoles, LJ on O from the A/B coefficients,
masses 16.0 / 1.008).  `water_box(n_waters, edge, seed)`: waters on a jittered cubic lattice with
random proper orientations (QR of Gaussian matrices, numpy default_rng(seed)), molecules whole.
`maxwell_boltzmann_rigid`: rigid-body velocities (centre-of-mass + angular) at temperature T.
"""

from __future__ import annotations

import numpy as np

from .system import Molecule, System

KB = 0.0083144626181532  # kJ/mol/K (pgm_jax/units.py:15)
OH, HOH_DEG = 0.09572, 104.52
PGM3P25 = {
    "q": (-2.0405622, 1.0202811, 1.0202811),
    "radius": (0.0605150752, 0.0536231183, 0.0536231183),
    "alpha": (1.11765163e-3, 3.2963077e-4, 3.2963077e-4),
    "cov": ((0, 1, -0.0191201350), (0, 2, -0.0191201350), (1, 0, 0.00858770326), (2, 0, 0.00858770326)),
    "lj_A_kcal_A12": 5.81935564e5,
    "lj_B_kcal_A6": 5.94825035e2,
    "masses": (16.0, 1.008, 1.008),
}


def pgm3p25_water() -> Molecule:
    """Return the rigid pGM3P-25 water template (nm, e, e nm, nm^3; LJ R*/2 [nm], sqrt(eps) [sqrt(kJ/mol)])."""
    A, B = PGM3P25["lj_A_kcal_A12"], PGM3P25["lj_B_kcal_A6"]
    rmin_A = (2.0 * A / B) ** (1.0 / 6.0)
    eps = B * B / (4.0 * A) * 4.184
    return Molecule("WAT", ["O", "H", "H"], PGM3P25["q"], PGM3P25["radius"], PGM3P25["alpha"],
                    [tuple(c) for c in PGM3P25["cov"]], [rmin_A / 2 * 0.1, 0.0, 0.0], [np.sqrt(eps), 0.0, 0.0],
                    masses=np.array(PGM3P25["masses"]))


def water_geometry(r: float = OH, theta_deg: float = HOH_DEG) -> np.ndarray:
    """Return O, H, H (3, 3) [nm], O at the origin, bisector along +y (models/toy.py:81-87)."""
    t = np.radians(theta_deg / 2)
    return np.array([[0, 0, 0], [r * np.sin(t), r * np.cos(t), 0], [-r * np.sin(t), r * np.cos(t), 0]])


def water_box(n_waters: int, edge: float, seed: int = 0, jitter: float = 0.01) -> tuple[System, np.ndarray, np.ndarray]:
    """Return (System, positions (3 n_waters, 3) [nm], cubic box [nm]) of rigid pGM3P-25 waters."""
    rng = np.random.default_rng(seed)
    s = int(np.ceil(n_waters ** (1.0 / 3.0) - 1e-9))
    spacing = edge / s
    w = water_geometry()
    w = w - (w * np.array([16.0, 1.008, 1.008])[:, None]).sum(0) / 18.016
    sites = np.array([[i, j, k] for i in range(s) for j in range(s) for k in range(s)], float)[:n_waters]
    pos = np.empty((n_waters, 3, 3))
    for n in range(n_waters):
        Q = np.linalg.qr(rng.normal(size=(3, 3)))[0]
        if np.linalg.det(Q) < 0:
            Q[:, 0] = -Q[:, 0]
        pos[n] = w @ Q.T + (sites[n] + 0.5) * spacing + rng.normal(scale=jitter, size=3)
    return System([pgm3p25_water()] * n_waters), pos.reshape(-1, 3), np.eye(3) * edge


def maxwell_boltzmann_rigid(pos: np.ndarray, masses: np.ndarray, temperature: float, seed: int = 0) -> np.ndarray:
    """Return velocities [nm/ps] of rigid 3-site molecules: COM and angular velocities at T [K].

    Centre-of-mass velocities ~ N(0, kT/M) per component (total momentum removed); angular velocity
    omega ~ N(0, kT I^-1) from each molecule's inertia tensor; v_a = V + omega x (r_a - R).  The
    velocities satisfy the rigid constraints exactly.
    """
    rng = np.random.default_rng(seed)
    x = pos.reshape(-1, 3, 3)
    m = masses.reshape(-1, 3)
    M = m.sum(1)
    R = (x * m[:, :, None]).sum(1) / M[:, None]
    V = rng.normal(size=R.shape) * np.sqrt(KB * temperature / M)[:, None]
    V -= (V * M[:, None]).sum(0) / M.sum()
    d = x - R[:, None]
    I = np.zeros((len(m), 3, 3))
    for a in range(3):
        da = d[:, a]
        I += m[:, a, None, None] * (np.einsum("mk,mk->m", da, da)[:, None, None] * np.eye(3) - da[:, :, None] * da[:, None, :])
    evals, evecs = np.linalg.eigh(I)
    xi = rng.normal(size=R.shape) * np.sqrt(KB * temperature / np.maximum(evals, 1e-300))
    w = np.einsum("mab,mb->ma", evecs, xi)
    v = V[:, None] + np.cross(w[:, None, :], d)
    return v.reshape(-1, 3)
