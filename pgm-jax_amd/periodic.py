"""The exact Ewald reference models on the MI355X: PeriodicPGM and PeriodicModel over the HIP C ABI.

Mirrors pgm_jax/ewald.py `PeriodicPGM` (:186-470) and pgm_jax/periodic.py `PeriodicModel`
(:101-281), `strain_derivative` / `pressure_bar` (:43-106): Gaussian charges, covalent dipoles and
induced Gaussian dipoles with the plain Ewald reciprocal sum (k-vectors fixed at the reference box,
exp(-k^2/(4 beta^2)) >= k_tol), a hard real-space cutoff, LJ between molecules with the optional
long-range correction.  Everything runs in FP64 in the HIP library (recip = Ewald): the
real-space rows and CG of the MD path with the structure-factor kernels in place of PME.

    model = PeriodicModel(system, box, positions, cutoff=0.9, lj_lrc=True)
    e = model.energy(positions)            # {"perm", "ind", "elec", "vdw", "total"} kJ/mol
    F = model.forces(positions)            # kJ/mol/nm
    W = model.strain_derivative(positions) # dE/d eps (3, 3), molecular scaling
    p = model.pressure(positions)          # bar

Differences from the reference, by design: the induced dipoles come from the library's Jacobi-
preconditioned CG (converged when max|alpha r| / mean|alpha b| <= dipole_tol) instead of
jax.scipy's unpreconditioned CG; at the default 1e-12 both reach the solution to rounding.  The
neighbour rows are rebuilt from the current positions (cutoff + skin) instead of being frozen at
positions_ref, which gives the same result whenever the reference's frozen list is complete.
Parameter pytrees are taken as the reference takes them (the System's tied table, or per-atom
arrays for a System given as arrays; None: the System's own values) and set on the device context
(pgm_set_params); the GVDW / DE van der Waals forms are not provided (NotImplementedError).
"""

from __future__ import annotations

from math import comb

import numpy as np

from .forcefield import PGMForceField, _dev, _torch
from .settings import MDSettings
from .system import System

BAR_PER_KJMOL_NM3 = 16.605390671738466  # 1 kJ/mol/nm^3 in bar (pgm_jax/units.py:17)
KE = 138.935458
ELEC_LEVELS = {"q": (False, False), "qp": (True, False), "qi": (False, True), "qpi": (True, True)}


def _variant(sys: System, perm: bool, lj: bool) -> System:
    """Return sys with the covalent dipoles and/or the LJ parameters switched off."""
    arrays = dict(q=sys.q, radius=sys.radius, alpha=sys.alpha, lj_rmin_half=sys.lj_rmin_half,
                  lj_sqrt_eps=sys.lj_sqrt_eps if lj else np.zeros_like(sys.lj_sqrt_eps), masses=sys.masses, mol=sys.mol,
                  cov_i=sys.cov_i if perm else np.zeros(0, np.int32), cov_j=sys.cov_j if perm else np.zeros(0, np.int32),
                  cov_c=sys.cov_c if perm else np.zeros(0))
    return System(**arrays)


class _EwaldField:
    """One FP64 Ewald context (elec level, with or without LJ) and its calls."""

    def __init__(self, sys: System, H, cutoff, ewald_beta, skin, k_tol, dipole_tol, elec, lj, lj_lrc) -> None:
        if elec not in ELEC_LEVELS:
            raise ValueError(f"unknown electrostatics level {elec!r} (expected q, qp, qi or qpi)")
        self.pd, self.ind = ELEC_LEVELS[elec]
        self.sys, self.lj = sys, bool(lj)
        self.n = sys.n
        self.custom = False  # parameters other than the System's own are set on the context
        s = MDSettings(cutoff=cutoff, skin=skin, ewald_beta=ewald_beta, pme_grid=(8, 8, 8), pme_order=4,
                       lj_lrc=bool(lj and lj_lrc), dipole_tol=dipole_tol, max_iter=2000, predictor="none", fused=False,
                       peek=0.0, precision="double", recip="ewald", k_tol=k_tol)
        self.ff = PGMForceField(_variant(sys, self.pd, lj), H, s)
        self.H = np.array(H, float)

    def box(self, H):
        return self.H if H is None else np.asarray(H, float)

    def set_params(self, params) -> dict:
        """Put the parameters of a pytree (None: the System's own) on the context (with this level's
        covalent dipoles and LJ switched off where the context has none); return them per atom."""
        ex = self.sys.expand(params)
        if params is None and not self.custom:
            return ex
        arrays = dict(ex)
        if not self.lj:
            arrays["lj_sqrt_eps"] = np.zeros_like(ex["lj_sqrt_eps"])
        if not self.pd:
            arrays["cov"] = np.zeros(0)
        self.ff.set_params(arrays)
        self.custom = params is not None
        return ex

    def mu(self, pos, H):
        """Induced dipoles at pos (zeros without induction)."""
        torch = _torch()
        if not self.ind:
            return torch.zeros((self.n, 3), dtype=torch.float64, device="cuda")
        return self.ff.compute(pos, self.box(H)).induction.mu

    def energy_forces(self, pos, H, mu):
        return self.ff.energy_forces(pos, mu, self.box(H))


class PeriodicPGM:
    """pGM electrostatic energy, forces and induced dipoles of one System in a periodic box (Ewald, FP64)."""

    def __init__(self, system: System, box, positions_ref, ewald_beta: float = 3.8, cutoff: float = 1.0,
                 skin: float = 0.0, k_tol: float = 1e-12, dipole_tol: float = 1e-12, nlist=None, elec: str = "qpi",
                 _lj: bool = False, _lj_lrc: bool = False) -> None:
        """Build the model (ewald.py:213-266); ValueError for an unknown `elec`.  `nlist` is accepted and ignored."""
        self.sys, self.H = system, np.asarray(box, float)
        self.b0, self.rc, self.cg_tol = float(ewald_beta), float(cutoff), float(dipole_tol)
        self._f = _EwaldField(system, self.H, cutoff, ewald_beta, skin, k_tol, dipole_tol, elec, _lj, _lj_lrc)
        self.pd, self.ind = self._f.pd, self._f.ind

    def _p(self, pos):
        torch = _torch()
        if not self.pd:
            return torch.zeros((self.sys.n, 3), dtype=torch.float64, device="cuda")
        return self._f.ff.perm_dipoles(pos)

    def induced_dipoles(self, pos, params=None, H=None):
        """Return the induced dipoles mu* (N, 3) [e nm] (ewald.py:421-430), computed even without induction."""
        self._f.set_params(params)
        pos = _dev(pos, (self.sys.n, 3))
        return self._f.ff.compute(pos, self._f.box(H)).induction.mu

    def U(self, pos, q, d, R=None, H=None, tens=None) -> float:
        """Return U(q, d) [kJ/mol] of the system's charges q and total dipoles d (ewald.py:344-379)."""
        if not np.allclose(np.asarray(q, float), self.sys.q, rtol=0, atol=0):
            raise NotImplementedError("U with charges other than the system's")
        if R is not None and not np.array_equal(np.asarray(R, float), self.sys.radius):
            raise NotImplementedError("U with radii other than the system's")
        pos = _dev(pos, (self.sys.n, 3))
        d = _dev(d, (self.sys.n, 3))
        mu = d - self._p(pos)
        e, _ = self._f.energy_forces(pos, H, mu)
        upol = KE * float(((mu * mu).sum(1) / (2.0 * _dev(self.sys.alpha))).sum().item())
        return e["elec"] - upol

    def energy(self, pos, params=None, H=None) -> tuple[dict, dict]:
        """Return ({"perm", "ind", "total"} [kJ/mol], {"p": (N, 3)}) (ewald.py:432-464)."""
        self._f.set_params(params)
        torch = _torch()
        pos = _dev(pos, (self.sys.n, 3))
        zero = torch.zeros((self.sys.n, 3), dtype=torch.float64, device="cuda")
        e_perm = self._f.energy_forces(pos, H, zero)[0]["elec"]
        e_tot = self._f.energy_forces(pos, H, self._f.mu(pos, H))[0]["elec"] if self.ind else e_perm
        return {"perm": e_perm, "ind": e_tot - e_perm, "total": e_tot}, {"p": self._p(pos)}

    def forces(self, pos, params=None, H=None):
        """Return -dE_total/dpos (N, 3) [kJ/mol/nm] at mu* (ewald.py:466-470)."""
        self._f.set_params(params)
        pos = _dev(pos, (self.sys.n, 3))
        return self._f.energy_forces(pos, H, self._f.mu(pos, H))[1]


class PeriodicModel:
    """pGM (Ewald) + LJ in a periodic box (periodic.py:101-281), FP64 on the GPU."""

    def __init__(self, system: System, box, positions_ref, cutoff: float = 1.0, ewald_beta: float = 3.8,
                 skin: float = 0.0, vdw_cutoff: float | None = None, lj_lrc: bool = False, k_tol: float = 1e-12,
                 dipole_tol: float = 1e-12, elec: str = "qpi", vdw: str = "lj", gvdw_rep: str = "gauss",
                 de_alpha: float | None = None, de_beta: float | None = None) -> None:
        """Build the model; ValueError for unknown levels / forms, NotImplementedError for GVDW, DE or a
        van der Waals cutoff other than the electrostatics cutoff."""
        if vdw not in ("lj", "gvdw", "de", "none"):
            raise ValueError(f"unknown van der Waals form {vdw!r}")
        if gvdw_rep not in ("gauss", "slater"):
            raise ValueError(f"unknown GVDW repulsion {gvdw_rep!r}")
        if vdw in ("gvdw", "de"):
            raise NotImplementedError(f"vdw={vdw!r} is not part of the HIP path (lj or none)")
        if vdw_cutoff is not None and vdw_cutoff != cutoff:
            raise NotImplementedError("a van der Waals cutoff other than the electrostatics cutoff")
        self.sys, self.H = system, np.asarray(box, float)
        self._args = (positions_ref, cutoff, ewald_beta, skin, k_tol, dipole_tol, elec)
        self.vdw_form = vdw
        self.lj_lrc = bool(lj_lrc) and vdw == "lj"
        self._full = _EwaldField(system, self.H, cutoff, ewald_beta, skin, k_tol, dipole_tol, elec, vdw == "lj", self.lj_lrc)
        self._elec = None

    @property
    def elec(self) -> PeriodicPGM:
        """The electrostatics alone (ewald.PeriodicPGM), built on first use."""
        if self._elec is None:
            ref, cutoff, beta, skin, k_tol, tol, elec = self._args
            self._elec = PeriodicPGM(self.sys, self.H, ref, ewald_beta=beta, cutoff=cutoff, skin=skin, k_tol=k_tol,
                                     dipole_tol=tol, elec=elec)
        return self._elec

    def _state(self, pos, params, H):
        self._lj_now = self._full.set_params(params)
        pos = _dev(pos, (self.sys.n, 3))
        return pos, self._full.mu(pos, H)

    def energy(self, pos, params=None, H=None) -> dict:
        """Return {"perm", "ind", "elec", "vdw", "total"} [kJ/mol] (periodic.py:205-228)."""
        torch = _torch()
        pos, mu = self._state(pos, params, H)
        zero = torch.zeros((self.sys.n, 3), dtype=torch.float64, device="cuda")
        e0 = self._full.energy_forces(pos, H, zero)[0]
        e = self._full.energy_forces(pos, H, mu)[0] if self._full.ind else e0
        return {"perm": e0["elec"], "ind": e["elec"] - e0["elec"], "elec": e["elec"], "vdw": e["vdw"], "total": e["total"]}

    def forces(self, pos, params=None, H=None):
        """Return -dE_total/dpos (N, 3) [kJ/mol/nm] (periodic.py:230-234)."""
        pos, mu = self._state(pos, params, H)
        return self._full.energy_forces(pos, H, mu)[1]

    def _tail(self, H) -> float:
        if not self.lj_lrc:
            return 0.0
        ex = getattr(self, "_lj_now", None) or self.sys.expand(None)
        V = abs(np.linalg.det(self._full.box(H)))
        rc = self._args[1]
        # E_lrc = -2 pi / (3 V rc^3) sum_{i,j} 2 eps_ij rmin_ij^6 (pgm_jax/lj.py:87-114), with
        # eps_ij rmin_ij^6 = sqrt(e_i) sqrt(e_j) (r_i + r_j)^6 summed by the binomial moments of r
        e, r = ex["lj_sqrt_eps"], ex["lj_rmin_half"]
        mom = [np.sum(e * r**k) for k in range(7)]
        total = 2.0 * sum(comb(6, k) * mom[k] * mom[6 - k] for k in range(7))
        return -2.0 * np.pi / (3.0 * V * rc**3) * total

    def strain_derivative(self, pos, params=None, H=None, molecular: bool = True) -> np.ndarray:
        """Return dE_total/d eps (3, 3) [kJ/mol] (periodic.py:236-251; molecular or atomic scaling)."""
        pos, mu = self._state(pos, params, H)
        W = self._full.ff.strain_derivative(pos, self._full.box(H), mu, molecular=molecular)
        return W + self._tail(H) * np.eye(3)  # the library includes the tail impulse -E_lrc I; this method does not

    def virial_derivative(self, pos, params=None, H=None, molecular: bool = True) -> np.ndarray:
        """Return the strain derivative plus the LJ tail impulse -E_lrc I (periodic.py:253-268)."""
        return self.strain_derivative(pos, params, H, molecular) - self._tail(H) * np.eye(3)

    def pressure(self, pos, params=None, H=None, molecular: bool = True) -> float:
        """Return the static pressure -tr(virial_derivative) / (3 V) [bar] (periodic.py:270-281)."""
        Hb = self._full.box(H)
        return pressure_bar(self.virial_derivative(pos, params, H, molecular), Hb)


def pressure_bar(dE_deps, H) -> float:
    """Return the static pressure P = -tr(dE/d eps) / (3 V) [bar] (periodic.py:84-106)."""
    V = abs(np.linalg.det(np.asarray(H, float)))
    return float(-np.trace(np.asarray(dE_deps, float)) / (3.0 * V) * BAR_PER_KJMOL_NM3)
