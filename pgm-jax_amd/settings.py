"""MDSettings of the hot path: the reference's settings under their flat names.

Mirrors pgm_jax/md/forcefield.py:142-454 (MDSettings with its groups Terms, Cutoffs, NeighborList,
PMESettings, Induction) restricted to what the electrostatics hot path reads, with the same
defaults and the same flat names that MDSettings.replace accepts (FLAT_SETTINGS, forcefield.py:331-367).
"""

from __future__ import annotations

import dataclasses

PREDICTORS = {"none": 0, "mu3": 1, "mu4": 2, "previous": 3, "learned": 4}


@dataclasses.dataclass(frozen=True)
class MDSettings:
    """Frozen settings; change them with `replace` (flat names).

    cutoff [nm], skin [nm], ewald_beta [1/nm], pme_grid (K1, K2, K3) or None (from pme_spacing [nm]),
    pme_order, lj_lrc, dipole_tol (max|alpha r| / mean|alpha b|), max_iter, predictor
    ("mu4", "mu3", "previous", "none"; "learned": the MLP guess of PGMForceField.guess_config, not in
    the reference), fused, norm_refresh, peek, precision ("mixed" | "double").
    Beyond the reference's MD settings: recip ("pme" | "ewald", the exact Ewald sum of the
    PeriodicPGM reference model), k_tol (its k-vector truncation) and precond ("jacobi", the
    reference's z = alpha r, or "molecule": the inverse 9x9 block of each 3-site molecule).
    elec_cutoff [nm] (None: cutoff) is the real-space electrostatics cutoff of the reference's split
    rows (forcefield.py:44-58): electrostatic pair terms stop there, LJ at cutoff; choose
    ewald_beta and the PME spacing for it (elec_cutoff_settings).
    adjoint_tol: the CG tolerance of the backward pass of the solve (PGMForceField.dipole_vjp;
    MDSettings.adjoint_tol, forcefield.py:414-425).
    """

    elec: str = "qpi"
    vdw: str = "lj"
    cutoff: float = 0.9
    skin: float = 0.1
    ewald_beta: float = 4.0
    pme_grid: tuple | None = None
    pme_spacing: float = 0.08
    pme_order: int = 6
    lj_lrc: bool = True
    dipole_tol: float = 1e-5
    max_iter: int = 50
    predictor: str = "mu4"
    fused: bool = True
    norm_refresh: int = 1000
    peek: float = 0.65
    precision: str = "mixed"
    recip: str = "pme"
    k_tol: float = 1e-12
    precond: str = "jacobi"
    elec_cutoff: float | None = None
    adjoint_tol: float = 1e-6

    def __post_init__(self) -> None:
        """Validate like PGMForceField.__init__ (forcefield.py:917-1000)."""
        if self.elec != "qpi" or self.vdw != "lj":
            raise ValueError("the HIP path implements elec='qpi' with vdw='lj'")
        if self.predictor not in PREDICTORS:
            raise ValueError(f"unknown predictor {self.predictor!r}")
        if self.precision not in ("mixed", "double"):
            raise ValueError(f"precision must be 'mixed' or 'double', got {self.precision!r}")
        if not (self.cutoff > 0 and self.ewald_beta > 0):
            raise ValueError("cutoff and ewald_beta must be positive")
        if not 4 <= int(self.pme_order) <= 8:
            raise ValueError("pme_order must be in 4..8")
        if self.recip not in ("pme", "ewald"):
            raise ValueError(f"recip must be 'pme' or 'ewald', got {self.recip!r}")
        if not 0.0 < self.k_tol < 1.0:
            raise ValueError("k_tol must be in (0, 1)")
        if self.precond not in ("jacobi", "molecule"):
            raise ValueError(f"precond must be 'jacobi' or 'molecule', got {self.precond!r}")
        if not self.adjoint_tol > 0:
            raise ValueError(f"adjoint_tol must be positive, got {self.adjoint_tol}")
        if self.elec_cutoff is not None and not self.elec_cutoff > 0:
            raise ValueError(f"elec_cutoff must be positive, got {self.elec_cutoff}")

    @property
    def elec_rc(self) -> float:
        """The electrostatics cutoff [nm] (cutoff unless elec_cutoff is set; forcefield.py:475)."""
        return float(self.cutoff if self.elec_cutoff is None else self.elec_cutoff)

    def replace(self, **changes) -> MDSettings:
        """Return a copy with some settings changed (TypeError for an unknown name, as the reference)."""
        names = {f.name for f in dataclasses.fields(self)}
        for k in changes:
            if k not in names:
                raise TypeError(f"MDSettings.replace: unknown setting {k!r}")
        return dataclasses.replace(self, **changes)


def grid_size(H, spacing: float = 0.05, factors=(2, 3, 5)) -> tuple[int, int, int]:
    """Return the smallest even (2,3,5)-smooth grid with plane spacing <= spacing (pgm_jax/md/pme.py:134-172)."""
    import numpy as np

    H = np.asarray(H, float)
    V = abs(np.linalg.det(H))
    heights = [V / np.linalg.norm(np.cross(H[(a + 1) % 3], H[(a + 2) % 3])) for a in range(3)]

    def smooth(n: int) -> bool:
        for f in factors:
            while n % f == 0:
                n //= f
        return n == 1

    out = []
    for h in heights:
        n = max(8, int(np.ceil(h / spacing)))
        while not smooth(n) or n % 2:
            n += 1
        out.append(n)
    return tuple(out)


# Direct-sum tolerance (Amber's convention, per Angstrom) of the default pair 0.9 nm / 4.0 nm^-1.
DSUM_TOL = 3.95e-8


def ewald_beta_for(elec_cutoff: float, dsum_tol: float = DSUM_TOL) -> float:
    """Return beta [1/nm] with erfc(beta rc) / rc_A = dsum_tol (rc_A = 10 rc), by bisection as Amber
    does (pgm_jax/md/forcefield.py:523-567): 4.0 at 0.9 nm, 5.19 at 0.7 nm for the default tolerance.
    Raises ValueError for a non-positive cutoff or dsum_tol outside (0, 1/rc_A)."""
    import math

    rc = float(elec_cutoff)
    if not rc > 0.0:
        raise ValueError(f"elec_cutoff must be positive, got {elec_cutoff}")

    def f(b: float) -> float:
        return math.erfc(b * rc) / (10.0 * rc) - dsum_tol

    if not (0.0 < dsum_tol and f(0.0) > 0.0):
        raise ValueError(f"dsum_tol must be in (0, {1.0 / (10.0 * rc):.3g}), got {dsum_tol}")
    lo, hi = 0.0, 1.0
    while f(hi) > 0.0:
        hi *= 2.0
    for _ in range(100):
        mid = 0.5 * (lo + hi)
        if f(mid) > 0.0:
            lo = mid
        else:
            hi = mid
    return 0.5 * (lo + hi)


def elec_cutoff_settings(elec_cutoff: float, dsum_tol: float = DSUM_TOL, exponent: float = 1.6) -> dict:
    """MDSettings arguments for an electrostatics cutoff: beta from ewald_beta_for and the PME spacing
    0.08 nm (4.0 / beta)^exponent (pgm_jax/md/forcefield.py:570-601)."""
    b = ewald_beta_for(elec_cutoff, dsum_tol)
    return {"elec_cutoff": float(elec_cutoff), "ewald_beta": b, "pme_spacing": 0.08 * (4.0 / b) ** exponent}
