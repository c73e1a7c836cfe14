"""The pGM + LJ MD force field in float64 (restates pgm_jax/md/forcefield.py, the hot path).

E = KE [U_dir + U_rec + U_self + U_bg + sum |mu|^2/(2 alpha)] + E_LJ,   d = p(R) + mu   (forcefield.py:8-16)
U_dir = sum_{pairs < rc} q_i q_j G0 + (q_i d_j.x - q_j d_i.x) G1 - G2 (d_i.x)(d_j.x) + G1 d_i.d_j,
G_n = B_n[erf(a_ij r)/r] - B_n[erf(b0 r)/r], a_ij = 1/sqrt(2 (R_i^2 + R_j^2)).

Restated functions (reference file:line):
  perm_dipoles        forcefield.py:1032-1042        p_i = sum c_m unit(min_image(r_j - r_i))
  perm_dipoles_vjp    the jax.vjp of forcefield.py:2412 written out
  row_field           forcefield.py:1395-1402        sum_k de_ik/dd_i (minus the direct field)
  field               forcefield.py:1428-1433        -(row + PME - c_self d)
  operator            forcefield.py:1694-1712        A v = v/alpha - field(0, v)
  cg                  forcefield.py:1714-1821        Jacobi PCG, Polak-Ribiere beta, max|alpha r|/norm, peek
  solve               forcefield.py:1917-2008        mu4 / mu3 predictor, fused residual, history push
  pair_terms          forcefield.py:2294-2360        pair energies and analytic row gradients, LJ rows
  nonpair             forcefield.py:2011-2045        PME + self + background + polarisation energies
  energy_forces       forcefield.py:2362-2430        forces = -(KE g_el + g_lj + grad_np + VJP_p(dE/dd))
  lj_long_range       lj.py:87-114                   continuum tail -2 pi/(3 V rc^3) sum_ij 2 eps rmin^6
  compute             forcefield.py:2489-2560        one force evaluation -> Result
  strain_derivative   forcefield.py:2628-2695        molecular dE/d eps at fixed mu (+ tail impulse)

With elec_cutoff (< or > cutoff) the electrostatic pair terms stop at elec_cutoff and LJ at cutoff, the
pairs spanning both (the reference's split rows, forcefield.py:44-58, :1197-1238, :2355); the LJ tail
keeps cutoff.

Every pair is kept once (i < j); the reference keeps it in both rows and halves the sums, which is
the same arithmetic in float64.  Only the "qpi" electrostatics with "lj" van der Waals, the PME
long-range treatment, the Jacobi preconditioner and nonzero polarizabilities are restated (the
settings of every benchmark and golden case here).

Parity: pinned by the golden ff_small_box (double: energies, forces, induced dipoles, CG iterations,
residual, strain) and ff_pgm3p25_512 (double) single points of the reference, by the exact Ewald
oracle (dense induction solve) as the PME grid grows, and by finite differences of the energy for
the forces and the strain derivative.
"""

from __future__ import annotations

import dataclasses
import math

import numpy as np

from .kernels import pair_kernels
from .pairs import min_image, pairs
from .pme import PME, grid_size

KE = 138.935458  # Coulomb constant, kJ/mol nm / e^2 (pgm_jax/units.py:31)
FARADAY_KJ = 96.48533212331002  # kJ/mol per (e V) (pgm_jax/units.py:45)
VNM_TO_INTERNAL = FARADAY_KJ / KE  # a field of 1 V/nm in e/nm^2 (pgm_jax/md/efield.py:104)
SQRT_PI = math.sqrt(math.pi)
PRED = {"mu3": (3.0, -3.0, 1.0), "mu4": (4.0, -6.0, 4.0, -1.0)}  # forcefield.py:605


def div_alpha(x, alpha):
    """x / alpha, 0 where alpha = 0 (forcefield.py:767-780 `_div_alpha`): atoms without polarizability
    keep a zero induced dipole (z = alpha r never moves it) and add nothing to |mu|^2 / (2 alpha)."""
    alpha = np.asarray(alpha, float)
    pol = alpha != 0
    return np.where(pol, np.asarray(x, float) / np.where(pol, alpha, 1.0), 0.0)


@dataclasses.dataclass(frozen=True)
class Settings:
    """The MDSettings fields of the hot path (forcefield.py:142-454 defaults)."""

    cutoff: float = 0.9
    ewald_beta: float = 4.0
    pme_grid: tuple | None = None
    pme_spacing: float = 0.08
    pme_order: int = 6
    lj_lrc: bool = True
    elec_cutoff: float | None = None
    dipole_tol: float = 1e-5
    max_iter: int = 50
    predictor: str = "mu4"
    fused: bool = True
    norm_refresh: int = 1000
    peek: float = 0.65


@dataclasses.dataclass
class InductionState:
    """mu (N, 3), hist (4, N, 3) newest first, count, norm (forcefield.py:679-713)."""

    mu: np.ndarray
    hist: np.ndarray
    count: int
    norm: float


@dataclasses.dataclass
class Result:
    """energy {"elec", "vdw", "total"} [kJ/mol], forces (N, 3), mu, iterations, residual (forcefield.py:716-746)."""

    energy: dict
    forces: np.ndarray
    mu: np.ndarray
    iterations: int
    residual: float
    induction: InductionState
    dipole: np.ndarray | None = None  # M = sum q r + sum (p + mu) + offset with an external field


def lj_long_range(rh: np.ndarray, se: np.ndarray, volume: float, rc: float) -> float:
    """Return E_lrc = -2 pi/(3 V rc^3) sum_{i,j} 2 eps_ij rmin_ij^6 in O(N) by moments (lj.py:87-114)."""
    M = [np.sum(se * rh**k) for k in range(7)]
    SB = 2.0 * sum(math.comb(6, k) * M[k] * M[6 - k] for k in range(7))
    return -2.0 * math.pi * SB / (3.0 * volume * rc**3)


class ForceField:
    """The force field of one system (PGMForceField, forcefield.py:810-1000), float64."""

    def __init__(self, top, H: np.ndarray, settings: Settings = Settings()) -> None:
        """Set up PME (grid from settings or grid_size(H, spacing)), b0 and c_self (forcefield.py:940-960)."""
        self.top, self.s = top, settings
        self.b0 = float(settings.ewald_beta)
        self.c_self = 4.0 * self.b0**3 / (3.0 * SQRT_PI)
        grid = settings.pme_grid or grid_size(H, settings.pme_spacing)
        self.pme = PME(grid, settings.pme_order, self.b0)
        self.rc = float(settings.cutoff)  # van der Waals cutoff (and the LJ tail's)
        # electrostatics cutoff (forcefield.py:44-58, :475): the real-space pair terms stop there,
        # LJ at rc; the pairs span both
        self.rc_e = self.rc if settings.elec_cutoff is None else float(settings.elec_cutoff)
        self.rc_pair = max(self.rc, self.rc_e)

    # ------------------------------------------------------------------ geometry
    def geometry(self, pos: np.ndarray, H: np.ndarray) -> dict:
        """Return the pairs within rc with their kernels G0..G3 and LJ parameters (forcefield.py:1340-1380)."""
        t = self.top
        g = pairs(pos, H, self.rc_pair, t.mol, t.first)
        i, j = g["i"], g["j"]
        a = 1.0 / np.sqrt(2.0 * (t.radius[i] ** 2 + t.radius[j] ** 2))
        inside = g["r"] < self.rc_e  # split rows (forcefield.py:1197-1238): electrostatics within rc_e only
        g["inside"] = inside
        g["G"] = tuple(Gn * inside for Gn in pair_kernels(a, self.b0, g["r"], 4))
        # rigid topology: intramolecular vdW weight 0; van der Waals only within rc
        g["wv"] = np.where(g["intra"] | (g["r"] >= self.rc), 0.0, 1.0)
        g["rmin"] = t.lj_rmin_half[i] + t.lj_rmin_half[j]
        g["eps"] = t.lj_sqrt_eps[i] * t.lj_sqrt_eps[j]
        return g

    def perm_dipoles(self, pos: np.ndarray, H: np.ndarray) -> np.ndarray:
        """Return p_i = sum_m c_m unit(min_image(r_j - r_i)) (forcefield.py:1032-1042)."""
        t = self.top
        v = min_image(pos[t.cov_j] - pos[t.cov_i], H)
        u = v / np.linalg.norm(v, axis=-1, keepdims=True)
        p = np.zeros((t.n, 3))
        np.add.at(p, t.cov_i, t.cov_c[:, None] * u)
        return p

    def perm_dipoles_vjp(self, pos: np.ndarray, H: np.ndarray, gp: np.ndarray) -> np.ndarray:
        """Return sum_i gp_i . dp_i/dr (N, 3): the pull-back through the covalent frames."""
        t = self.top
        v = min_image(pos[t.cov_j] - pos[t.cov_i], H)
        L = np.linalg.norm(v, axis=-1, keepdims=True)
        u = v / L
        gi = gp[t.cov_i]
        tv = t.cov_c[:, None] * (gi - np.sum(gi * u, -1, keepdims=True) * u) / L
        out = np.zeros((t.n, 3))
        np.add.at(out, t.cov_j, tv)
        np.add.at(out, t.cov_i, -tv)
        return out

    # ------------------------------------------------------------------ fields and operator
    def row_field(self, g: dict, q: np.ndarray | None, d: np.ndarray) -> np.ndarray:
        """Return sum_k de_ik/dd_i (N, 3), minus the direct-space field (forcefield.py:1395-1402)."""
        i, j, x = g["i"], g["j"], g["x"]
        G1, G2 = g["G"][1], g["G"][2]
        di, dj = d[i], d[j]
        dix = np.sum(di * x, -1)
        djx = np.sum(dj * x, -1)
        ci = -G2 * djx
        cj = -G2 * dix
        if q is not None:
            ci = ci - q[j] * G1
            cj = cj + q[i] * G1
        out = np.zeros_like(d, dtype=float)
        np.add.at(out, i, ci[:, None] * x + G1[:, None] * dj)
        np.add.at(out, j, cj[:, None] * x + G1[:, None] * di)
        return out

    def field(self, g: dict, S: dict, Gk: np.ndarray, q: np.ndarray | None, d: np.ndarray) -> np.ndarray:
        """Return the total field -dU/dd = -(row + PME - c_self d) [e/nm^2] (forcefield.py:1428-1433)."""
        qq = np.zeros(self.top.n) if q is None else q
        rec, _ = self.pme.gradients(S, Gk, qq, d)
        return -(self.row_field(g, q, d) + rec - self.c_self * d)

    def operator(self, g: dict, S: dict, Gk: np.ndarray):
        """Return A: v -> v/alpha - field(0, v) (forcefield.py:1694-1712), v/alpha = 0 where alpha = 0."""
        inv_a = div_alpha(1.0, self.top.alpha[:, None])
        return lambda v: v * inv_a - self.field(g, S, Gk, None, v)

    def cg(self, A, x: np.ndarray, r: np.ndarray, norm: float, tol=None, peek=None, max_iter=None) -> tuple:
        """Jacobi PCG from (x0, r0 = b - A x0) (forcefield.py:1714-1821); returns (mu, iterations, residual)."""
        s = self.s
        tol = s.dipole_tol if tol is None else tol
        peek = s.peek if peek is None else peek
        max_iter = s.max_iter if max_iter is None else max_iter
        a_c = self.top.alpha[:, None]
        z = r * a_c
        p = z.copy()
        rz = float(np.sum(r * z))
        it = 0
        err = float(np.max(np.abs(r * a_c))) / norm
        while err > tol and it < max_iter:
            Ap = A(p)
            al = rz / max(float(np.sum(p * Ap)), 1e-300)
            x = x + al * p
            r_new = r - al * Ap
            z_new = r_new * a_c
            beta = float(np.sum(z_new * (r_new - r))) / max(rz, 1e-300)
            p = z_new + beta * p
            rz = float(np.sum(r_new * z_new))
            r = r_new
            it += 1
            err = float(np.max(np.abs(r * a_c))) / norm
        if peek:
            x = x + peek * r * a_c
        return x, it, err

    def init_induction(self) -> InductionState:
        """Return a fresh solver state: zero dipoles and history, normaliser 1 (forcefield.py:1436-1450)."""
        n = self.top.n
        return InductionState(np.zeros((n, 3)), np.zeros((4, n, 3)), 0, 1.0)

    def solve(self, g: dict, S: dict, Gk: np.ndarray, p: np.ndarray, ind: InductionState, F0=None) -> tuple:
        """Predictor + (fused) initial residual + CG; returns (mu, it, err, new state) (forcefield.py:1917-2008).

        F0: a uniform external field (3,) [e/nm^2] added to the permanent field of every atom (the
        right-hand side of the induction equations, forcefield.py:1966-1997 with _ext_at :1823-1832)."""
        t, s = self.top, self.s
        a64 = t.alpha[:, None]
        A = self.operator(g, S, Gk)
        q = t.q
        ext = np.zeros(3) if F0 is None else np.asarray(F0, float)

        def plain(x0h, have):
            b = self.field(g, S, Gk, q, p) + ext
            ab = a64 * b
            x0 = x0h if have else ab
            return x0, b - A(x0), float(np.mean(np.abs(ab))) + 1e-300

        if s.predictor in PRED:
            c = PRED[s.predictor]
            have = ind.count >= len(c)
            x0h = sum(ci * ind.hist[k] for k, ci in enumerate(c))
            if s.fused and have and ind.count % s.norm_refresh != 0:
                x0 = x0h
                r0 = self.field(g, S, Gk, q, p + x0h) + ext - div_alpha(x0h, a64)
                norm = ind.norm
            else:
                x0, r0, norm = plain(x0h, have)
        else:
            x0, r0, norm = plain(None, False)
        mu, it, err = self.cg(A, x0, r0, norm)
        hist = np.concatenate([mu[None], ind.hist[:-1]], 0)
        return mu, it, err, InductionState(mu, hist, ind.count + 1, norm)

    # ------------------------------------------------------------------ energies and forces
    def pair_terms(self, g: dict, q: np.ndarray, d: np.ndarray) -> tuple:
        """Return (sum e_elec, sum e_lj, g_el (N,3), g_lj (N,3)) over unique pairs (forcefield.py:2294-2360, :2116-2146)."""
        i, j, x, r = g["i"], g["j"], g["x"], g["r"]
        G0, G1, G2, G3 = g["G"]
        qi, qk = q[i], q[j]
        di, dk = d[i], d[j]
        dix = np.sum(di * x, -1)
        dkx = np.sum(dk * x, -1)
        didk = np.sum(di * dk, -1)
        tt = qi * dkx - qk * dix
        e = qi * qk * G0 + tt * G1 - G2 * dix * dkx + G1 * didk
        radial = -qi * qk * G1 - tt * G2 + G3 * dix * dkx - G2 * didk
        cross = di * dkx[:, None] + dk * dix[:, None]
        gx = radial[:, None] * x + (qi * G1)[:, None] * dk - (qk * G1)[:, None] * di - G2[:, None] * cross
        s6 = (g["rmin"] / r) ** 6
        on = (g["wv"] != 0) & (g["eps"] != 0)
        elj = np.where(on, g["eps"] * (s6 * s6 - 2.0 * s6), 0.0)
        dlj = np.where(on, g["eps"] * 12.0 * (s6 - s6 * s6) / (r * r), 0.0)
        glj = dlj[:, None] * x
        n = len(q)
        g_el = np.zeros((n, 3))
        g_lj = np.zeros((n, 3))
        np.add.at(g_el, i, gx)
        np.add.at(g_el, j, -gx)
        np.add.at(g_lj, i, glj)
        np.add.at(g_lj, j, -glj)
        return float(np.sum(e)), float(np.sum(elj)), g_el, g_lj

    def vdw_tail(self, H: np.ndarray) -> float:
        """Return the LJ long-range correction [kJ/mol] (forcefield.py:2148-2168 -> lj.py:87-114)."""
        if not self.s.lj_lrc:
            return 0.0
        t = self.top
        return lj_long_range(t.lj_rmin_half, t.lj_sqrt_eps, abs(np.linalg.det(H)), self.rc)

    def nonpair(self, S: dict, Gk: np.ndarray, d: np.ndarray, mu: np.ndarray, H: np.ndarray) -> float:
        """Return KE [U_rec + U_self + U_bg + U_pol] [kJ/mol] (forcefield.py:2011-2045)."""
        t = self.top
        u_rec = self.pme.energy(S, Gk, t.q, d)
        u_self = -(self.b0 / SQRT_PI) * np.sum(t.q**2) - 0.5 * self.c_self * np.sum(d * d)
        u_bg = -math.pi * np.sum(t.q) ** 2 / (2.0 * abs(np.linalg.det(H)) * self.b0**2)
        u_pol = np.sum(div_alpha(mu * mu, 2.0 * t.alpha[:, None]))
        return KE * (u_rec + u_self + u_bg + u_pol)

    def energy_forces(self, pos: np.ndarray, H: np.ndarray, mu: np.ndarray, g: dict, S: dict, Gk: np.ndarray,
                      F0=None, offset=None) -> tuple:
        """Return ({"elec", "vdw", "total"}, forces (N, 3)) at fixed mu (forcefield.py:2362-2430).

        With a uniform external field F0 (3,) [e/nm^2] (forcefield.py:2417-2430, _ext_terms :2075-2088,
        field_dipole :2090-2097): the energy gains "field" = -KE F0 . M, M = sum q r + sum (p + mu)
        + offset (molecules whole), the total includes it, dE/dd of every atom gains -KE F0 (the
        torque on the covalent dipoles) and every atom the force KE q F0; also returns M."""
        t = self.top
        p = self.perm_dipoles(pos, H)
        d = p + mu
        se, sl, g_el, g_lj = self.pair_terms(g, t.q, d)
        dEdd = KE * self.row_field(g, t.q, d)
        if F0 is not None:
            dEdd = dEdd - KE * np.asarray(F0, float)[None, :]
        rec_dd, rec_dr = self.pme.gradients(S, Gk, t.q, d)
        gd_np = KE * (rec_dd - self.c_self * d)
        gpos_np = KE * rec_dr
        forces = -(KE * g_el + g_lj + gpos_np + self.perm_dipoles_vjp(pos, H, dEdd + gd_np))
        e_elec = KE * se + self.nonpair(S, Gk, d, mu, H)
        e_lj = sl + self.vdw_tail(H)
        if F0 is None:
            return {"elec": e_elec, "vdw": e_lj, "total": e_elec + e_lj}, forces
        F0 = np.asarray(F0, float)
        M = np.sum(t.q[:, None] * pos + d, axis=0) + (np.zeros(3) if offset is None else np.asarray(offset, float))
        e_f = -KE * float(np.dot(F0, M))
        forces = forces + KE * t.q[:, None] * F0[None, :]
        return {"elec": e_elec, "vdw": e_lj, "field": e_f, "total": e_elec + e_lj + e_f}, forces, M

    def compute(self, pos: np.ndarray, H: np.ndarray, ind: InductionState | None = None, efield=None) -> Result:
        """Solve the induced dipoles from the predictor, then return energy and forces (forcefield.py:2489-2560).

        efield = (E (3,) [V/nm], offset (3,) [e nm] or None): a uniform static external field (the
        constant-field kind of forcefield.py:2048-2073, md/efield.py); Result.dipole = M."""
        pos = np.asarray(pos, float)
        H = np.asarray(H, float)
        ind = self.init_induction() if ind is None else ind
        g = self.geometry(pos, H)
        p = self.perm_dipoles(pos, H)
        S = self.pme.setup(pos, H)
        Gk = self.pme.influence(H)
        F0 = None if efield is None else VNM_TO_INTERNAL * np.asarray(efield[0], float)
        mu, it, err, ind = self.solve(g, S, Gk, p, ind, F0)
        if F0 is None:
            energy, forces = self.energy_forces(pos, H, mu, g, S, Gk)
            return Result(energy, forces, mu, it, err, ind)
        energy, forces, M = self.energy_forces(pos, H, mu, g, S, Gk, F0, efield[1])
        return Result(energy, forces, mu, it, err, ind, M)

    def strain_derivative(self, pos: np.ndarray, H: np.ndarray, mu: np.ndarray, molecular: bool = True) -> np.ndarray:
        """Return dE/d eps (3, 3) [kJ/mol] at fixed mu, minus the tail impulse E_lrc I.

        molecular=False: atomic scaling (every atom r -> (1 + eps) r): o = 0 below, and the covalent
        dipole frames strain, adding sum_cov t (x) v with t = c (g - (g.u) u)/|v|, g = dE/dd of the
        frame's owner atom (the pull-back of perm_dipoles_vjp, applied to eps v instead of dr).

        forcefield.py:2628-2695 takes it by autodiff of the whole energy (full_strain_derivative,
        :608-662); here it is written out.  Under x -> x + (com eps^T)[mol], H -> H (1 + eps)^T:
          pairs: the displacement x_ij moves by eps X_ij, X_ij = x_ij - o_i + o_j (o = r - com):
                 W += sum de/dx (x) X;
          PME:   fractional coordinates move by -o eps^T H^-1 and e = d H^-1 K by -d eps^T H^-1 K:
                 W -= sum dE/dr (x) o + sum dE/dd_rec (x) d; the influence function G(m*, V) gives
                 W_ab += sum |F Q|^2 G [2 (pi^2/beta^2 + 1/|m*|^2) m*_a m*_b - delta_ab];
          background and LJ tail scale as 1/V: -E delta_ab each, and the tail impulse -E_lrc I.
        """
        t = self.top
        pos, H = np.asarray(pos, float), np.asarray(H, float)
        g = self.geometry(pos, H)
        S = self.pme.setup(pos, H)
        Gk = self.pme.influence(H)
        p = self.perm_dipoles(pos, H)
        d = p + mu
        m = t.masses
        com = np.zeros((t.nmol, 3))
        np.add.at(com, t.mol, m[:, None] * pos)
        msum = np.bincount(t.mol, weights=m, minlength=t.nmol)
        com /= msum[:, None]
        o = pos - com[t.mol] if molecular else np.zeros_like(pos)
        i, j, x, r = g["i"], g["j"], g["x"], g["r"]
        X = x - o[i] + o[j]
        # pair gradients de/dx (electrostatic in KE units, LJ)
        G0, G1, G2, G3 = g["G"]
        qi, qk, di, dk = t.q[i], t.q[j], d[i], d[j]
        dix, dkx, didk = np.sum(di * x, -1), np.sum(dk * x, -1), np.sum(di * dk, -1)
        tt = qi * dkx - qk * dix
        radial = -qi * qk * G1 - tt * G2 + G3 * dix * dkx - G2 * didk
        gx = radial[:, None] * x + (qi * G1)[:, None] * dk - (qk * G1)[:, None] * di - G2[:, None] * (
            di * dkx[:, None] + dk * dix[:, None])
        s6 = (g["rmin"] / r) ** 6
        on = (g["wv"] != 0) & (g["eps"] != 0)
        dlj = np.where(on, g["eps"] * 12.0 * (s6 - s6 * s6) / (r * r), 0.0)
        W = KE * gx.T @ X + (dlj[:, None] * x).T @ X
        # reciprocal space
        gd, gr = self.pme.gradients(S, Gk, t.q, d)
        W -= KE * (gr.T @ o + gd.T @ d)
        FQ = np.fft.rfftn(self.pme.spread(S, t.q, d))
        F2 = FQ.real**2 + FQ.imag**2
        R = np.linalg.inv(H).T
        m1, m2, m3 = self.pme.m
        mv = m1[:, None, None, None] * R[0] + m2[None, :, None, None] * R[1] + m3[None, None, :, None] * R[2]
        msq = np.sum(mv * mv, -1)
        safe = np.where(msq > 0, msq, 1.0)
        fac = np.where(msq > 0, 2.0 * (np.pi**2 / self.b0**2 + 1.0 / safe), 0.0) * Gk * F2
        W += KE * (np.einsum("abc,abci,abcj->ij", fac, mv, mv) - np.sum(Gk * F2) * np.eye(3))
        V = abs(np.linalg.det(H))
        e_bg = -math.pi * np.sum(t.q) ** 2 / (2.0 * V * self.b0**2) * KE
        W -= e_bg * np.eye(3)
        tail = self.vdw_tail(H)
        W -= 2.0 * tail * np.eye(3)
        if not molecular:
            gdd = KE * (self.row_field(g, t.q, d) + gd - self.c_self * d)
            v = min_image(pos[t.cov_j] - pos[t.cov_i], H)
            L = np.linalg.norm(v, axis=-1, keepdims=True)
            u = v / L
            gi = gdd[t.cov_i]
            tv = t.cov_c[:, None] * (gi - np.sum(gi * u, -1, keepdims=True) * u) / L
            W += tv.T @ v
        return W

    def energy(self, pos: np.ndarray, H: np.ndarray, ind: InductionState) -> tuple:
        """Energy only, for Monte Carlo barostat trials (forcefield.py:2562-2626): CG from ind.mu (no
        predictor; the field b and the norm mean|alpha b| evaluated afresh); returns (E_total, mu, it).
        The state's history is not touched."""
        pos = np.asarray(pos, float)
        H = np.asarray(H, float)
        g = self.geometry(pos, H)
        p = self.perm_dipoles(pos, H)
        S = self.pme.setup(pos, H)
        Gk = self.pme.influence(H)
        A = self.operator(g, S, Gk)
        b = self.field(g, S, Gk, self.top.q, p)
        norm = float(np.mean(np.abs(self.top.alpha[:, None] * b))) + 1e-300
        x0 = np.asarray(ind.mu, float)
        mu, it, _ = self.cg(A, x0, b - A(x0), norm)
        e = self.energy_forces(pos, H, mu, g, S, Gk)[0]["total"]
        return e, mu, it

    def energy_fixed_mu(self, pos: np.ndarray, H: np.ndarray, mu: np.ndarray) -> float:
        """Return the total energy [kJ/mol] at fixed induced dipoles (forcefield.py:2249-2292)."""
        g = self.geometry(pos, H)
        S = self.pme.setup(pos, H)
        Gk = self.pme.influence(H)
        return self.energy_forces(pos, H, mu, g, S, Gk)[0]["total"]

    # ------------------------------------------------------------------ the solve's backward pass
    def potential_from_dipoles(self, g: dict, S: dict, Gk: np.ndarray, lam: np.ndarray) -> np.ndarray:
        """Return phi_j = dU/dq_j of dipoles lam alone (N,) [e/nm]: sum_k G1 lam_k . (r_j - r_k) over the
        pairs plus the reciprocal potential sum_stencil theta1 theta2 theta3 dE_rec/dQ."""
        i, j, x = g["i"], g["j"], g["x"]
        G1 = g["G"][1]
        phi = np.zeros(self.top.n)
        np.add.at(phi, i, G1 * np.sum(lam[j] * x, -1))
        np.add.at(phi, j, -G1 * np.sum(lam[i] * x, -1))
        grid = self.pme.potential(S, Gk, np.zeros(self.top.n), lam)
        P = grid.reshape(-1)[self.pme._flat(S)].reshape(len(phi), *(self.pme.order,) * 3)
        t = [S["th"][:, a] for a in range(3)]
        return phi + np.einsum("nabc,na,nb,nc->n", P, t[0], t[1], t[2], optimize=True)

    def dipole_vjp(self, pos: np.ndarray, H: np.ndarray, mu: np.ndarray, mu_bar: np.ndarray, tol: float = 1e-12) -> dict:
        """The backward pass of the differentiable solve (forcefield.py:1853-1915): the adjoint CG
        A lam = mu_bar (same operator, x0 = 0, no peek), then lam pulled back through the residual
        r = b(theta) - A(theta) mu at fixed mu.  With E(mu) = U(q, p + mu) + sum |mu|^2 / (2 alpha) and
        r = -dE/dmu, theta_bar = lam . dr/dtheta = -d/dtheta (lam . dE/dmu):

          q_bar_j     = -phi_j(lam)                      (potential of the dipoles lam at j)
          alpha_bar_i = mu_i . lam_i / alpha_i^2
          p_bar_i     = field_i(0, lam)                  (the permanent dipoles; cov_bar by the frames)
          R_bar_i     = 2 R_i sum_j a_ij^3 d/da [dlam e_ij],  dG_n/da = (2/sqrt(pi)) (2a^2)^n exp(-a^2 r^2)

        where dlam e_ij is the derivative of the pair energy along d -> d + eps lam.  Returns per-atom
        arrays and cov_bar (per covalent dipole) with lam and the CG iteration count."""
        t = self.top
        pos, H = np.asarray(pos, float), np.asarray(H, float)
        g = self.geometry(pos, H)
        S = self.pme.setup(pos, H)
        Gk = self.pme.influence(H)
        A = self.operator(g, S, Gk)
        a64 = t.alpha[:, None]
        norm = float(np.mean(np.abs(a64 * mu_bar))) + 1e-300
        lam, it, _ = self.cg(A, np.zeros_like(mu_bar), np.array(mu_bar, float), norm, tol=tol, peek=0.0)
        q_bar = -self.potential_from_dipoles(g, S, Gk, lam)
        alpha_bar = np.sum(mu * lam, -1) / t.alpha**2
        p_bar = self.field(g, S, Gk, None, lam)
        # radii: pair sums of the a-derivative of the lam-directional pair energy
        d = self.perm_dipoles(pos, H) + mu
        i, j, x, r = g["i"], g["j"], g["x"], g["r"]
        a = 1.0 / np.sqrt(2.0 * (t.radius[i] ** 2 + t.radius[j] ** 2))
        inside = r < self.rc_e
        c = (2.0 / SQRT_PI) * np.exp(-(a * r) ** 2) * inside
        H1, H2 = c * 2.0 * a * a, c * (2.0 * a * a) ** 2
        qi, qk = t.q[i], t.q[j]
        li, lk, di, dk = lam[i], lam[j], d[i], d[j]
        lix, lkx, dix, dkx = (np.sum(v * x, -1) for v in (li, lk, di, dk))
        de = (qi * lkx - qk * lix) * H1 - H2 * (lix * dkx + dix * lkx) + H1 * (np.sum(li * dk, -1) + np.sum(di * lk, -1))
        R_bar = np.zeros(t.n)
        np.add.at(R_bar, i, 2.0 * t.radius[i] * a**3 * de)
        np.add.at(R_bar, j, 2.0 * t.radius[j] * a**3 * de)
        v = min_image(pos[t.cov_j] - pos[t.cov_i], H)
        u = v / np.linalg.norm(v, axis=-1, keepdims=True)
        cov_bar = np.sum(p_bar[t.cov_i] * u, -1)
        return {"lam": lam, "iterations": it, "q": q_bar, "alpha": alpha_bar, "p": p_bar, "radius": R_bar, "cov": cov_bar}

    # ------------------------------------------------------------------ parameter derivatives
    def geometry_frozen(self, pos: np.ndarray, H: np.ndarray, g0: dict, pos0: np.ndarray) -> dict:
        """The pairs of g0 (built at pos0) at positions pos: the same pairs, lattice shifts and cutoff
        masks, the kernels evaluated at the new distances.  Autodiff of the reference's masked pair
        sums (forcefield.py:2232-2292) differentiates exactly this: the masks carry no derivative."""
        t = self.top
        i, j = g0["i"], g0["j"]
        x = pos[i] - pos[j] + (g0["x"] - (pos0[i] - pos0[j]))
        r = np.sqrt(np.sum(x * x, -1))
        a = 1.0 / np.sqrt(2.0 * (t.radius[i] ** 2 + t.radius[j] ** 2))
        G = tuple(Gn * g0["inside"] for Gn in pair_kernels(a, self.b0, r, 4))
        return dict(g0, x=x, r=r, G=G)

    def param_grad(self, pos: np.ndarray, H: np.ndarray, mu: np.ndarray, g: dict) -> dict:
        """Return dE/dtheta at fixed mu [kJ/mol per unit] for every per-atom parameter and covalent
        dipole strength (the Hellmann-Feynman derivative: E is stationary in mu), and the total field
        -dU/dd (N, 3) of q and d = p + mu.  E = KE [pairs + U_rec + U_self + U_bg + U_pol] + LJ + tail
        (forcefield.py:2011-2045, 2294-2430; lj.py:87-114):
          q_i:      KE [sum_k (q_k G0 + d_k.x G1) + phi_rec_i - 2 beta q_i / sqrt(pi) - pi Q / (V beta^2)]
          R_i:      KE sum_k de_ik/da da/dR_i,  da/dR_i = -2 R_i a^3,  dG_n/da = (2/sqrt(pi)) (2a^2)^n exp(-a^2 r^2)
          alpha_i:  -KE |mu_i|^2 / (2 alpha_i^2)
          c_m:      dE/dp_i . u_m = -KE field_i . u_m
          R*_i/2:   sum_k 12 eps (s^12 - s^6) / rmin (+ tail),  sqrt(eps_i): sum_k sqrt(eps_k) (s^12 - 2 s^6) (+ tail)
        with x = r_i - r_k, s = rmin / r and the pairs of g."""
        t = self.top
        n = t.n
        q = t.q
        p = self.perm_dipoles(pos, H)
        d = p + mu
        S = self.pme.setup(pos, H)
        Gk = self.pme.influence(H)
        i, j, x, r = g["i"], g["j"], g["x"], g["r"]
        G0, G1, G2, _ = g["G"]
        dix, djx = np.sum(d[i] * x, -1), np.sum(d[j] * x, -1)
        didj = np.sum(d[i] * d[j], -1)
        gq = np.zeros(n)
        np.add.at(gq, i, q[j] * G0 + djx * G1)
        np.add.at(gq, j, q[i] * G0 - dix * G1)
        phi = self.pme.potential(S, Gk, q, d)
        P = phi.reshape(-1)[self.pme._flat(S)].reshape(n, *(self.pme.order,) * 3)
        th = [S["th"][:, a] for a in range(3)]
        rec = np.einsum("nabc,na,nb,nc->n", P, th[0], th[1], th[2], optimize=True)
        V = abs(np.linalg.det(H))
        gq = KE * (gq + rec - 2.0 * self.b0 * q / SQRT_PI - math.pi * np.sum(q) / (V * self.b0**2))
        a = 1.0 / np.sqrt(2.0 * (t.radius[i] ** 2 + t.radius[j] ** 2))
        H0 = (2.0 / SQRT_PI) * np.exp(-((a * r) ** 2)) * g["inside"]
        H1 = H0 * 2.0 * a * a
        H2 = H1 * 2.0 * a * a
        dea = q[i] * q[j] * H0 + (q[i] * djx - q[j] * dix) * H1 - H2 * dix * djx + H1 * didj
        gR = np.zeros(n)
        np.add.at(gR, i, -2.0 * t.radius[i] * a**3 * dea)
        np.add.at(gR, j, -2.0 * t.radius[j] * a**3 * dea)
        galpha = -KE * div_alpha(np.sum(mu * mu, -1), 2.0 * t.alpha**2)
        field = self.field(g, S, Gk, q, d)
        v = min_image(pos[t.cov_j] - pos[t.cov_i], H)
        u = v / np.sqrt(np.sum(v * v, -1, keepdims=True))
        gcov = -KE * np.sum(field[t.cov_i] * u, -1)
        on = g["wv"] != 0
        rmin, eps = g["rmin"], g["eps"]
        s6 = np.where(on, (rmin / r) ** 6, 0.0)
        drmin = np.where(on & (eps != 0), 12.0 * eps * (s6 * s6 - s6) / np.where(rmin > 0, rmin, 1.0), 0.0)
        grh, gse = np.zeros(n), np.zeros(n)
        np.add.at(grh, i, drmin)
        np.add.at(grh, j, drmin)
        ese = s6 * s6 - 2.0 * s6
        np.add.at(gse, i, t.lj_sqrt_eps[j] * ese)
        np.add.at(gse, j, t.lj_sqrt_eps[i] * ese)
        if self.s.lj_lrc:  # E_lrc = -2 pi SB / (3 V rc^3), SB = 2 sum_k C(6,k) M_k M_(6-k), M_k = sum se rh^k
            rh, se = t.lj_rmin_half, t.lj_sqrt_eps
            M = [np.sum(se * rh**k) for k in range(7)]
            c0 = -2.0 * math.pi / (3.0 * V * self.rc**3)
            for k in range(7):
                dSB = 4.0 * math.comb(6, k) * M[6 - k]
                gse += c0 * dSB * rh**k
                if k:
                    grh += c0 * dSB * k * se * rh ** (k - 1)
        return {"q": gq, "radius": KE * gR, "alpha": galpha, "lj_rmin_half": grh, "lj_sqrt_eps": gse, "cov": gcov,
                "field": field}

    def compute_vjp(self, pos: np.ndarray, H: np.ndarray, mu: np.ndarray, e_bar: float = 0.0, f_bar=None, mu_bar=None,
                    tol: float = 1e-12, step: float = 1e-4) -> dict:
        """Return the parameter cotangents of (E, F, mu) of the differentiable force field
        (forcefield.py:1853-1915, 2489-2560 with differentiable=True): for a loss with dL/dE = e_bar,
        dL/dF = f_bar (N, 3), dL/dmu = mu_bar (N, 3) at the solved dipoles mu,

          theta_bar = e_bar dE/dtheta + f_bar . dF/dtheta|_mu + mu_bar_eff . dmu/dtheta,
          f_bar . dF/dtheta|_mu = -D_v (dE/dtheta),   mu_bar_eff = mu_bar + KE D_v field,   v = f_bar,

        with D_v the derivative along r -> r + eps v at fixed mu and fixed pair set (F = -dE/dr at
        fixed mu, dF/dmu = KE dfield/dr), and the last term the adjoint solve (dipole_vjp).  D_v is
        taken here by fourth-order central differences (largest displacement `step` nm, the pairs and
        masks of the unperturbed geometry).  Returns per-atom arrays and "cov" (per covalent dipole)."""
        pos, H, mu = np.asarray(pos, float), np.asarray(H, float), np.asarray(mu, float)
        keys = ("q", "radius", "alpha", "lj_rmin_half", "lj_sqrt_eps", "cov")
        g0 = self.geometry(pos, H)
        base = self.param_grad(pos, H, mu, g0)
        out = {k: e_bar * base[k] for k in keys}
        mub = np.zeros_like(mu) if mu_bar is None else np.array(mu_bar, float)
        if f_bar is not None and np.any(f_bar):
            v = np.asarray(f_bar, float)
            h = step / np.abs(v).max()
            f = {}
            for sgn in (1, -1, 2, -2):
                x = pos + sgn * h * v
                f[sgn] = self.param_grad(x, H, mu, self.geometry_frozen(x, H, g0, pos))
            D = {k: (8.0 * (f[1][k] - f[-1][k]) - (f[2][k] - f[-2][k])) / (12.0 * h) for k in keys + ("field",)}
            for k in keys:
                out[k] = out[k] - D[k]
            mub = mub + KE * D["field"]
        if np.any(mub):
            adj = self.dipole_vjp(pos, H, mu, mub, tol=tol)
            for k in ("q", "radius", "alpha", "cov"):
                out[k] = out[k] + adj[k]
        return out
