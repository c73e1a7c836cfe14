"""Test systems of the oracle: molecules flattened to per-atom arrays, boxes and coordinates.

Restates the builders the reference's tests use (numpy only, seeded):

* the toy water / methanol templates and `small_box` (pgm_jax/models/toy.py:26-78, :81-92, :143-180)
  with `reduce_box` (pgm_jax/md/box.py:32-59) -- the inputs of the golden files ff_small_box.npz and
  periodic_model.npz;
* `box_from_cell` (pgm_jax/md/box.py:291-314) and the 512-water pGM3P-25 box of examples/water512
  (read from the committed fixture tests/golden/water512.npz);
* `water_box`: rigid pGM3P-25 waters on a jittered cubic lattice with random orientations, the
  synthetic inputs of the benchmark configurations (the pattern of models/toy.py:107-140).

Parity: `small_box` is pinned by the golden energies of ff_small_box / periodic_model (a wrong atom
order, seed use or parameter fails them); the pGM3P-25 constants are pinned against the values the
fixture script read from the prmtop (tests/test_oracle_systems.py).

Units: nm, e, e nm, nm^3, amu, LJ R*/2 [nm] and sqrt(eps) [sqrt(kJ/mol)].
"""

from __future__ import annotations

import dataclasses
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tests", "golden")

MASSES = {"H": 1.008, "C": 12.011, "O": 15.999}


@dataclasses.dataclass
class Molecule:
    """One molecule template (the fields of pgm_jax/system.py Molecule that the hot path reads)."""

    name: str
    elements: list
    q: np.ndarray
    radius: np.ndarray
    alpha: np.ndarray
    cov: list  # (i, j, c): covalent dipole c [e nm] on atom i along unit(r_j - r_i)
    lj_rmin_half: np.ndarray
    lj_sqrt_eps: np.ndarray
    masses: np.ndarray | None = None

    def __post_init__(self) -> None:
        """Convert to float arrays and fill default element masses."""
        for f in ("q", "radius", "alpha", "lj_rmin_half", "lj_sqrt_eps"):
            setattr(self, f, np.asarray(getattr(self, f), float))
        if self.masses is None:
            self.masses = np.array([MASSES[e] for e in self.elements])
        self.masses = np.asarray(self.masses, float)

    @property
    def n(self) -> int:
        """Number of atoms."""
        return len(self.elements)


@dataclasses.dataclass
class Topology:
    """A system flattened to per-atom arrays (System.expand of pgm_jax/system.py:561-577).

    Attributes: q, radius, alpha, lj_rmin_half, lj_sqrt_eps, masses (N,); mol (N,) molecule index;
    cov_i, cov_j (C,) int and cov_c (C,) covalent dipoles; first (M,) first atom of each molecule.
    """

    q: np.ndarray
    radius: np.ndarray
    alpha: np.ndarray
    lj_rmin_half: np.ndarray
    lj_sqrt_eps: np.ndarray
    masses: np.ndarray
    mol: np.ndarray
    cov_i: np.ndarray
    cov_j: np.ndarray
    cov_c: np.ndarray
    first: np.ndarray

    @property
    def n(self) -> int:
        """Number of atoms N."""
        return len(self.q)

    @property
    def nmol(self) -> int:
        """Number of molecules."""
        return len(self.first)

    @classmethod
    def from_molecules(cls, mols: list) -> Topology:
        """Flatten a list of Molecule templates (System.__init__, pgm_jax/system.py:483-523)."""
        offs = np.cumsum([0] + [m.n for m in mols])
        cat = lambda f: np.concatenate([getattr(m, f) for m in mols])  # noqa: E731
        ci, cj, cc = [], [], []
        for k, m in enumerate(mols):
            for i, j, c in m.cov:
                ci.append(offs[k] + i)
                cj.append(offs[k] + j)
                cc.append(c)
        return cls(
            q=cat("q"),
            radius=cat("radius"),
            alpha=cat("alpha"),
            lj_rmin_half=cat("lj_rmin_half"),
            lj_sqrt_eps=cat("lj_sqrt_eps"),
            masses=cat("masses"),
            mol=np.concatenate([np.full(m.n, k, np.int32) for k, m in enumerate(mols)]),
            cov_i=np.array(ci, np.int32),
            cov_j=np.array(cj, np.int32),
            cov_c=np.array(cc, float),
            first=np.asarray(offs[:-1], np.int32),
        )


# ----------------------------------------------------------------------------- boxes
def reduce_box(H: np.ndarray) -> np.ndarray:
    """Return the reduced form of a lower-triangular box (pgm_jax/md/box.py:32-59)."""
    H = np.array(H, float)
    a, b, c = H
    c = c - b * np.round(c[1] / b[1])
    c = c - a * np.round(c[0] / a[0])
    b = b - a * np.round(b[0] / a[0])
    return np.array([a, b, c])


def box_from_cell(lengths, angles) -> np.ndarray:
    """Return the lower-triangular box of cell lengths and angles [deg] (pgm_jax/md/box.py:291-314)."""
    a, b, c = lengths
    al, be, ga = np.radians(angles)
    cx = c * np.cos(be)
    cy = c * (np.cos(al) - np.cos(be) * np.cos(ga)) / np.sin(ga)
    return np.array([[a, 0, 0], [b * np.cos(ga), b * np.sin(ga), 0], [cx, cy, np.sqrt(c * c - cx * cx - cy * cy)]])


# ----------------------------------------------------------------------------- toy molecules
def toy_water() -> Molecule:
    """Return the toy pGM water template (pgm_jax/models/toy.py:26-43)."""
    return Molecule(
        "WAT",
        ["O", "H", "H"],
        [-0.8, 0.4, 0.4],
        [0.06, 0.05, 0.05],
        [1.0e-3, 0.3e-3, 0.3e-3],
        [(0, 1, -0.02), (0, 2, -0.02), (1, 0, 0.008), (2, 0, 0.008)],
        [0.178, 0.0, 0.0],
        [0.80, 0.0, 0.0],
    )


def toy_methanol() -> tuple[Molecule, np.ndarray]:
    """Return the toy methanol template and its geometry [nm] (pgm_jax/models/toy.py:46-78)."""
    x = (
        np.array(
            [
                [-0.0467, 0.6590, 0.0],
                [-0.0467, -0.7598, 0.0],
                [-1.0830, 0.9930, 0.0],
                [0.4406, 1.0735, 0.8902],
                [0.4406, 1.0735, -0.8902],
                [0.8785, -1.0591, 0.0],
            ]
        )
        * 0.1
    )
    cov = [(0, 1, 0.01), (1, 0, -0.01), (1, 5, -0.02), (5, 1, 0.005)]
    cov += [c for h in (2, 3, 4) for c in ((0, h, 0.002), (h, 0, -0.002))]
    m = Molecule(
        "MeOH",
        ["C", "O", "H", "H", "H", "H"],
        [0.12, -0.62, 0.02, 0.02, 0.02, 0.44],
        [0.07, 0.06, 0.05, 0.05, 0.05, 0.05],
        [1.2e-3, 0.8e-3, 0.4e-3, 0.4e-3, 0.4e-3, 0.3e-3],
        cov,
        [0.19, 0.172, 0.139, 0.139, 0.139, 0.02],
        [0.33, 0.85, 0.20, 0.20, 0.20, 0.1],
    )
    return m, x


def water_geometry(r: float = 0.09572, theta_deg: float = 104.52) -> np.ndarray:
    """Return O, H, H (3, 3) [nm], O at the origin, bisector along +y (pgm_jax/models/toy.py:81-87)."""
    t = np.radians(theta_deg / 2)
    return np.array([[0, 0, 0], [r * np.sin(t), r * np.cos(t), 0], [-r * np.sin(t), r * np.cos(t), 0]])


def small_box(seed: int = 0, nw: int = 30, nm: int = 4) -> tuple[Topology, np.ndarray, np.ndarray]:
    """Return (topology, positions [nm], box [nm]) of the skewed water / methanol box.

    Restates pgm_jax/models/toy.py:143-180: methanols first, then waters, on a jittered 4 x 4 x 3
    lattice of a reduced triclinic box, orientations from QR of Gaussian matrices (numpy
    default_rng(seed), draws in the reference's order).
    """
    rng = np.random.default_rng(seed)
    H = reduce_box(np.array([[1.75, 0.0, 0.0], [0.45, 1.70, 0.0], [-0.40, 0.50, 1.65]]))
    w = water_geometry()
    m, xm = toy_methanol()
    xm = xm - xm.mean(0)
    W = toy_water()
    mols, pos = [], []
    grid = np.array([[i, j, k] for i in range(4) for j in range(4) for k in range(3)], float)
    grid = (grid + 0.5) / np.array([4, 4, 3])
    for n, f in enumerate(grid[rng.permutation(len(grid))][: nw + nm]):
        R = np.linalg.qr(rng.normal(size=(3, 3)))[0]
        c = f @ H + rng.normal(scale=0.01, size=3)
        if n < nm:
            mols.append(m)
            pos.append(xm @ R.T + c)
        else:
            mols.append(W)
            pos.append((w - w.mean(0)) @ R.T + c)
    return Topology.from_molecules(mols), np.concatenate(pos), H


# ----------------------------------------------------------------------------- pGM3P-25 water
# pGM3P-25 water, read from examples/water512/pgm3p25_512.prmtop (POL_GAUSS_* sections, LJ A/B
# of the O type, MASS) and converted as pgm_jax/param.py:113-233 does.  tests/test_oracle_systems.py
# checks these constants against tests/golden/water512.npz (written by make_fixtures.py).
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
    """Return the rigid 3-site pGM3P-25 water template (LJ on O only)."""
    A, B = PGM3P25["lj_A_kcal_A12"], PGM3P25["lj_B_kcal_A6"]
    rmin_A = (2.0 * A / B) ** (1.0 / 6.0)
    eps_kcal = B * B / (4.0 * A)
    return Molecule(
        "WAT",
        ["O", "H", "H"],
        PGM3P25["q"],
        PGM3P25["radius"],
        PGM3P25["alpha"],
        [tuple(c) for c in PGM3P25["cov"]],
        [rmin_A / 2 * 0.1, 0.0, 0.0],
        [np.sqrt(eps_kcal * 4.184), 0.0, 0.0],
        masses=np.array(PGM3P25["masses"]),
    )


def water512() -> tuple[Topology, np.ndarray, np.ndarray, np.ndarray]:
    """Return (topology, positions [nm], box [nm], velocities [nm/ps]) of examples/water512.

    The truncated-octahedron cell of the rst7 as a lower-triangular box (box_from_cell), as the
    reference's golden ff_pgm3p25_512 builds it (tests/regression/regression_cases.py:292-309).
    """
    with np.load(os.path.join(GOLDEN, "water512.npz"), allow_pickle=False) as z:
        pos, vel = z["pos"], z["vel"]
        H = box_from_cell(z["cell_lengths"], z["cell_angles"])
    top = Topology.from_molecules([pgm3p25_water()] * (len(pos) // 3))
    return top, pos, H, vel


def water_box(n_waters: int, edge: float, seed: int = 0, jitter: float = 0.01) -> tuple[Topology, np.ndarray, np.ndarray]:
    """Return (topology, positions [nm], cubic box [nm]) of rigid pGM3P-25 waters.

    Oxygens sit on the sites of a cubic lattice of the smallest side s with s^3 >= n_waters (the
    first n_waters sites in lattice order), displaced by Gaussian noise of `jitter` nm; each water
    gets a random orientation (QR of a Gaussian matrix, made proper).  O-H 0.09572 nm, H-O-H 104.52 deg
    (pgm_jax/models/toy.py:81-87, :107-140 pattern).  Molecules are whole (not wrapped).
    """
    rng = np.random.default_rng(seed)
    s = int(np.ceil(n_waters ** (1.0 / 3.0) - 1e-9))
    spacing = edge / s
    w = water_geometry()
    w = w - (w * np.array([16.0, 1.008, 1.008])[:, None]).sum(0) / 18.016  # centre of mass at 0
    sites = np.array([[i, j, k] for i in range(s) for j in range(s) for k in range(s)], float)[:n_waters]
    pos = np.empty((n_waters, 3, 3))
    for n in range(n_waters):
        Q = np.linalg.qr(rng.normal(size=(3, 3)))[0]
        if np.linalg.det(Q) < 0:
            Q[:, 0] = -Q[:, 0]
        c = (sites[n] + 0.5) * spacing + rng.normal(scale=jitter, size=3)
        pos[n] = w @ Q.T + c
    top = Topology.from_molecules([pgm3p25_water()] * n_waters)
    return top, pos.reshape(-1, 3), np.eye(3) * edge
