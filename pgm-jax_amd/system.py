"""Molecules flattened to the per-atom arrays the hot path reads (System.expand of the reference).

Mirrors pgm_jax/system.py: a `Molecule` template (elements, charges, Gaussian radii,
polarizabilities, covalent dipoles (i, j, c), LJ R*/2 and sqrt(eps), masses) and a `System` of
molecules in order, flattened to (N,) arrays, the molecule index of every atom and the covalent
dipole index arrays cov_i, cov_j, cov_c.
"""

from __future__ import annotations

import dataclasses

import numpy as np

MASSES = {"H": 1.008, "C": 12.011, "N": 14.007, "O": 15.999, "S": 32.06}


@dataclasses.dataclass
class Molecule:
    """A molecule template (units nm, e, e nm, nm^3, amu)."""

    name: str
    elements: list
    q: np.ndarray
    radius: np.ndarray
    alpha: np.ndarray
    cov: list
    lj_rmin_half: np.ndarray
    lj_sqrt_eps: np.ndarray
    masses: np.ndarray | None = None

    def __post_init__(self) -> None:
        """Arrays as float; element masses by default."""
        for f in ("q", "radius", "alpha", "lj_rmin_half", "lj_sqrt_eps"):
            setattr(self, f, np.asarray(getattr(self, f), float))
        self.masses = np.array([MASSES[e] for e in self.elements]) if self.masses is None else np.asarray(self.masses, float)

    @property
    def n(self) -> int:
        """Number of atoms."""
        return len(self.elements)


class System:
    """Flattened system: per-atom arrays q, radius, alpha, lj_rmin_half, lj_sqrt_eps, masses, mol."""

    def __init__(self, molecules: list | None = None, **arrays) -> None:
        """Flatten `molecules`, or take the arrays directly (q, radius, alpha, lj_rmin_half,
        lj_sqrt_eps, masses, mol, cov_i, cov_j, cov_c)."""
        if molecules is not None:
            offs = np.cumsum([0] + [m.n for m in molecules])
            cat = lambda f: np.concatenate([getattr(m, f) for m in molecules])  # noqa: E731
            ci, cj, cc = [], [], []
            for k, m in enumerate(molecules):
                for i, j, c in m.cov:
                    ci.append(offs[k] + i)
                    cj.append(offs[k] + j)
                    cc.append(c)
            arrays = dict(
                q=cat("q"), radius=cat("radius"), alpha=cat("alpha"), lj_rmin_half=cat("lj_rmin_half"),
                lj_sqrt_eps=cat("lj_sqrt_eps"), masses=cat("masses"),
                mol=np.concatenate([np.full(m.n, k) for k, m in enumerate(molecules)]),
                cov_i=np.array(ci), cov_j=np.array(cj), cov_c=np.array(cc, float),
            )
        self.q = np.ascontiguousarray(arrays["q"], float)
        self.radius = np.ascontiguousarray(arrays["radius"], float)
        self.alpha = np.ascontiguousarray(arrays["alpha"], float)
        self.lj_rmin_half = np.ascontiguousarray(arrays["lj_rmin_half"], float)
        self.lj_sqrt_eps = np.ascontiguousarray(arrays["lj_sqrt_eps"], float)
        self.masses = np.ascontiguousarray(arrays.get("masses", np.ones(len(self.q))), float)
        self.mol = np.ascontiguousarray(arrays["mol"], np.int32)
        self.cov_i = np.ascontiguousarray(arrays["cov_i"], np.int32)
        self.cov_j = np.ascontiguousarray(arrays["cov_j"], np.int32)
        self.cov_c = np.ascontiguousarray(arrays["cov_c"], float)
        n = len(self.q)
        for k in ("radius", "alpha", "lj_rmin_half", "lj_sqrt_eps", "masses", "mol"):
            if len(getattr(self, k)) != n:
                raise ValueError(f"System: {k} has {len(getattr(self, k))} entries, expected {n}")

    @property
    def n(self) -> int:
        """Number of atoms N."""
        return len(self.q)

    @property
    def nmol(self) -> int:
        """Number of molecules."""
        return int(self.mol.max()) + 1 if self.n else 0
