"""Exact Ewald pGM reference in float64 (restates pgm_jax/ewald.py, periodic.py, solver.py, lj.py).

U = KE (U_dir + U_rec + U_self + U_bg)                                     (ewald.py:344-379)
U_dir = sum_{pairs, images < rc} q_i q_j g - q_i d_j.grad g + q_j d_i.grad g - d_i.(grad grad g).d_j
        g = (erf(b_ij r) - erf(b0 r))/r                                    (ewald.py:161-183, :301-315)
U_rec = (4 pi/V) sum_{half k} exp(-k^2/(4 b0^2))/k^2 |S(k)|^2,  S = sum (q + i k.d) exp(i k.r)  (ewald.py:317-333)
U_self, U_bg                                                                 (ewald.py:335-342)
Induction: mu* minimises G(mu) = U(q, p + mu) + KE sum |mu|^2/(2 alpha)   (ewald.py:386-413)
solved here by a dense 3N x 3N solve (solver.py:29-62 `solve_linear_induction`, the exact form of
the reference's CG), energies {"perm", "ind", "total"} (ewald.py:432-464), forces -dE/dr at
fixed mu* (variational, solver.py:126-167), and PeriodicModel's LJ with its tail (periodic.py:205-228,
lj.py:117-201).

Parity: pinned by the golden periodic_model.npz of the reference (small_box(1), cutoff 0.6,
beta 3.8, k_tol 1e-12, dipole_tol 1e-12, LJ with lrc): energies, forces and induced dipoles.
"""

from __future__ import annotations

import math

import numpy as np

from .forcefield import KE, lj_long_range
from .kernels import pair_kernels
from .pairs import min_image

BAR_PER_KJMOL_NM3 = 16.605390671738466  # 1 kJ/mol/nm^3 in bar (units.py:17)


def kvector_indices(H: np.ndarray, kcut: float) -> np.ndarray:
    """Return half-space integer triples m with 0 < |2 pi m H^-T| < kcut (ewald.py:129-158)."""
    H = np.asarray(H, float)
    B = 2 * np.pi * np.linalg.inv(H).T
    Mx = int(np.ceil(kcut * np.max(np.linalg.norm(H, axis=1)) / (2 * np.pi))) + 1
    rng = np.arange(-Mx, Mx + 1)
    m = np.stack(np.meshgrid(rng, rng, rng, indexing="ij"), -1).reshape(-1, 3).astype(float)
    half = (m[:, 0] > 0) | ((m[:, 0] == 0) & (m[:, 1] > 0)) | ((m[:, 0] == 0) & (m[:, 1] == 0) & (m[:, 2] > 0))
    m = m[half]
    kk = np.linalg.norm(m @ B, axis=1)
    return m[(kk > 0) & (kk < kcut)]


class EwaldPGM:
    """PeriodicPGM + LJ (PeriodicModel) of a small system, float64 (ewald.py:186-470, periodic.py:101-281)."""

    def __init__(self, top, H: np.ndarray, ewald_beta: float = 3.8, cutoff: float = 1.0, k_tol: float = 1e-12,
                 lj_lrc: bool = False) -> None:
        """Set the k-vectors (kcut = 2 beta sqrt(-ln k_tol), ewald.py:264-265) and cutoffs."""
        self.top, self.H = top, np.asarray(H, float)
        self.b0, self.rc, self.lj_lrc = float(ewald_beta), float(cutoff), lj_lrc
        self.m = kvector_indices(self.H, 2 * self.b0 * math.sqrt(-math.log(k_tol)))

    def _pairs(self, pos: np.ndarray) -> dict:
        """Return all pairs i < j with minimum-image displacement below rc (rc < half the box height)."""
        n = self.top.n
        i, j = np.triu_indices(n, k=1)
        x = min_image(pos[i] - pos[j], self.H)
        r = np.linalg.norm(x, axis=-1)
        keep = r < self.rc
        i, j, x, r = i[keep], j[keep], x[keep], r[keep]
        R = self.top.radius
        a = 1.0 / np.sqrt(2.0 * (R[i] ** 2 + R[j] ** 2))
        return {"i": i, "j": j, "x": x, "r": r, "G": pair_kernels(a, self.b0, r, 4)}

    def _k(self) -> tuple[np.ndarray, np.ndarray, float]:
        """Return k-vectors (K, 3), the factors (4 pi/V) exp(-k^2/4b0^2)/k^2, and V."""
        B = 2 * np.pi * np.linalg.inv(self.H).T
        k = self.m @ B
        V = abs(np.linalg.det(self.H))
        k2 = np.sum(k * k, 1)
        return k, (4 * np.pi / V) * np.exp(-k2 / (4 * self.b0**2)) / k2, V

    def perm_dipoles(self, pos: np.ndarray) -> np.ndarray:
        """Return p_i = sum c unit(r_j - r_i), plain differences (channels.py:70-96)."""
        t = self.top
        v = pos[t.cov_j] - pos[t.cov_i]
        u = v / np.linalg.norm(v, axis=-1, keepdims=True)
        p = np.zeros((t.n, 3))
        np.add.at(p, t.cov_i, t.cov_c[:, None] * u)
        return p

    def _rec(self, pos: np.ndarray, q: np.ndarray, d: np.ndarray) -> tuple[float, np.ndarray, np.ndarray]:
        """Return U_rec and its gradients with respect to d and r (structure-factor sums)."""
        k, f, _ = self._k()
        ph = pos @ k.T
        kd = d @ k.T
        c, s = np.cos(ph), np.sin(ph)
        A = np.sum(q[:, None] * c - kd * s, 0)
        Bs = np.sum(q[:, None] * s + kd * c, 0)
        U = float(np.sum(f * (A * A + Bs * Bs)))
        wA, wB = 2 * f * A, 2 * f * Bs
        gd = (-s * wA[None] + c * wB[None]) @ k  # sum_k (2 f A dA/dd + 2 f B dB/dd)
        coef = wA[None] * (-q[:, None] * s - kd * c) + wB[None] * (q[:, None] * c - kd * s)
        gr = coef @ k
        return U, gd, gr

    def _dir(self, g: dict, q: np.ndarray, d: np.ndarray) -> tuple[float, np.ndarray, np.ndarray]:
        """Return U_dir, dU_dir/dd (N, 3) and dU_dir/dr (N, 3)."""
        i, j, x = g["i"], g["j"], g["x"]
        G0, G1, G2, G3 = g["G"]
        qi, qk, di, dk = q[i], q[j], d[i], d[j]
        dix, dkx, didk = np.sum(di * x, -1), np.sum(dk * x, -1), np.sum(di * dk, -1)
        tt = qi * dkx - qk * dix
        e = qi * qk * G0 + tt * G1 - G2 * dix * dkx + G1 * didk
        radial = -qi * qk * G1 - tt * G2 + G3 * dix * dkx - G2 * didk
        gx = radial[:, None] * x + (qi * G1)[:, None] * dk - (qk * G1)[:, None] * di - G2[:, None] * (
            di * dkx[:, None] + dk * dix[:, None]
        )
        n = len(q)
        gd = np.zeros((n, 3))
        np.add.at(gd, i, (-qk * G1 - G2 * dkx)[:, None] * x + G1[:, None] * dk)
        np.add.at(gd, j, (qi * G1 - G2 * dix)[:, None] * x + G1[:, None] * di)
        gr = np.zeros((n, 3))
        np.add.at(gr, i, gx)
        np.add.at(gr, j, -gx)
        return float(np.sum(e)), gd, gr

    def U(self, pos: np.ndarray, q: np.ndarray, d: np.ndarray, g: dict | None = None) -> tuple:
        """Return (U [kJ/mol], dU/dd, dU/dr) of charges q and dipoles d (ewald.py:344-379)."""
        g = self._pairs(pos) if g is None else g
        ud, gdd, gdr = self._dir(g, q, d)
        ur, grd, grr = self._rec(pos, q, d)
        b0 = self.b0
        cs = 4 * b0**3 / (3 * math.sqrt(math.pi))
        us = -(b0 / math.sqrt(math.pi)) * np.sum(q**2) - 0.5 * cs * np.sum(d * d)
        ub = -np.pi * np.sum(q) ** 2 / (2 * abs(np.linalg.det(self.H)) * b0**2)
        return KE * (ud + ur + us + ub), KE * (gdd + grd - cs * d), KE * (gdr + grr)

    def induced_dipoles(self, pos: np.ndarray) -> np.ndarray:
        """Return mu* (N, 3) by the dense solve of grad G = 0 (ewald.py:396-430, solver.py:29-62)."""
        t = self.top
        n = t.n
        g = self._pairs(pos)
        p = self.perm_dipoles(pos)
        _, g0, _ = self.U(pos, t.q, p, g)  # dU/dd at mu = 0
        cols = []
        zq = np.zeros(n)
        for c in range(3 * n):  # Hessian of U in d (U is quadratic in d): columns of dU/dd(e_c)
            e = np.zeros(3 * n)
            e[c] = 1.0
            cols.append(self.U(pos, zq, e.reshape(n, 3), g)[1].reshape(-1))
        T = np.stack(cols, 1)
        A = T + np.diag(np.repeat(KE / t.alpha, 3))
        return np.linalg.solve(A, -g0.reshape(-1)).reshape(n, 3)

    def energy(self, pos: np.ndarray) -> dict:
        """Return {"perm", "ind", "elec", "vdw", "total"} [kJ/mol] (ewald.py:432-464, periodic.py:205-228)."""
        t = self.top
        pos = np.asarray(pos, float)
        g = self._pairs(pos)
        p = self.perm_dipoles(pos)
        e_perm = self.U(pos, t.q, p, g)[0]
        mu = self.induced_dipoles(pos)
        e_tot = self.U(pos, t.q, p + mu, g)[0] + KE * np.sum(mu * mu / (2 * t.alpha[:, None]))
        vdw = self.lj(pos)[0]
        return {"perm": e_perm, "ind": e_tot - e_perm, "elec": e_tot, "vdw": vdw, "total": e_tot + vdw}

    def lj(self, pos: np.ndarray) -> tuple[float, np.ndarray]:
        """Return (E_LJ, dE/dr) of the intermolecular pairs within rc, with the tail (lj.py:117-201)."""
        t = self.top
        n = t.n
        i, j = np.triu_indices(n, k=1)
        inter = t.mol[i] != t.mol[j]
        i, j = i[inter], j[inter]
        x = min_image(pos[i] - pos[j], self.H)
        r = np.linalg.norm(x, axis=-1)
        keep = r < self.rc
        i, j, x, r = i[keep], j[keep], x[keep], r[keep]
        rmin = t.lj_rmin_half[i] + t.lj_rmin_half[j]
        eps = t.lj_sqrt_eps[i] * t.lj_sqrt_eps[j]
        s6 = (rmin / r) ** 6
        e = float(np.sum(eps * (s6 * s6 - 2 * s6)))
        gx = (eps * 12.0 * (s6 - s6 * s6) / (r * r))[:, None] * x
        gr = np.zeros((n, 3))
        np.add.at(gr, i, gx)
        np.add.at(gr, j, -gx)
        if self.lj_lrc:
            e += lj_long_range(t.lj_rmin_half, t.lj_sqrt_eps, abs(np.linalg.det(self.H)), self.rc)
        return e, gr

    def forces(self, pos: np.ndarray) -> np.ndarray:
        """Return -dE_total/dr (N, 3) [kJ/mol/nm] at the converged mu* (ewald.py:466-470, periodic.py:230-234)."""
        t = self.top
        pos = np.asarray(pos, float)
        g = self._pairs(pos)
        p = self.perm_dipoles(pos)
        mu = self.induced_dipoles(pos)
        _, gd, gr = self.U(pos, t.q, p + mu, g)
        v = pos[t.cov_j] - pos[t.cov_i]
        L = np.linalg.norm(v, axis=-1, keepdims=True)
        u = v / L
        gi = gd[t.cov_i]
        tv = t.cov_c[:, None] * (gi - np.sum(gi * u, -1, keepdims=True) * u) / L
        np.add.at(gr, t.cov_j, tv)
        np.add.at(gr, t.cov_i, -tv)
        return -(gr + self.lj(pos)[1])


    def strain_derivative(self, pos: np.ndarray, molecular: bool = True) -> np.ndarray:
        """Return dE_total/d eps (3, 3) [kJ/mol] at mu* (periodic.py:43-81, :236-251), without the tail impulse.

        Written out (the reference takes it by autodiff).  Molecular scaling moves atoms with their
        molecule's centre of mass (offsets o = r - com fixed); atomic scaling moves every atom
        (o = 0), so the covalent dipole frames strain too:
          pairs        sum (KE de/dx + de_LJ/dx) (x) (x - o_i + o_j)
          reciprocal   -KE sum (dU/dr (x) o + dU/dd (x) d)
                       + KE sum_k f |S|^2 [2 k k^T (1/(4 b0^2) + 1/k^2) - I]
          background   -E_bg I;  LJ tail (1/V) -E_lrc I
          atomic only  sum_cov t (x) v,  t = c (g - (g.u) u)/|v|, g = dE/dd of the frame's owner.
        """
        t = self.top
        pos = np.asarray(pos, float)
        n = t.n
        g = self._pairs(pos)
        p = self.perm_dipoles(pos)
        mu = self.induced_dipoles(pos)
        d = p + mu
        if molecular:
            m = t.masses
            com = np.zeros((t.nmol, 3))
            np.add.at(com, t.mol, m[:, None] * pos)
            com /= np.bincount(t.mol, weights=m, minlength=t.nmol)[:, None]
            o = pos - com[t.mol]
        else:
            o = np.zeros_like(pos)
        i, j, x = g["i"], g["j"], g["x"]
        G0, G1, G2, G3 = g["G"]
        q = t.q
        qi, qk, di, dk = q[i], q[j], d[i], d[j]
        dix, dkx, didk = np.sum(di * x, -1), np.sum(dk * x, -1), np.sum(di * dk, -1)
        tt = qi * dkx - qk * dix
        radial = -qi * qk * G1 - tt * G2 + G3 * dix * dkx - G2 * didk
        gx = radial[:, None] * x + (qi * G1)[:, None] * dk - (qk * G1)[:, None] * di - G2[:, None] * (
            di * dkx[:, None] + dk * dix[:, None])
        W = KE * gx.T @ (x - o[i] + o[j])
        # LJ pairs (intermolecular, within rc)
        li, lj = np.triu_indices(n, k=1)
        inter = t.mol[li] != t.mol[lj]
        li, lj = li[inter], lj[inter]
        lx = min_image(pos[li] - pos[lj], self.H)
        lr = np.linalg.norm(lx, axis=-1)
        keep = lr < self.rc
        li, lj, lx, lr = li[keep], lj[keep], lx[keep], lr[keep]
        s6 = ((t.lj_rmin_half[li] + t.lj_rmin_half[lj]) / lr) ** 6
        eps = t.lj_sqrt_eps[li] * t.lj_sqrt_eps[lj]
        glj = (eps * 12.0 * (s6 - s6 * s6) / (lr * lr))[:, None] * lx
        W += glj.T @ (lx - o[li] + o[lj])
        # reciprocal space
        k, f, V = self._k()
        ph = pos @ k.T
        kd = d @ k.T
        c, s = np.cos(ph), np.sin(ph)
        A = np.sum(q[:, None] * c - kd * s, 0)
        Bs = np.sum(q[:, None] * s + kd * c, 0)
        _, gd, gr = self._rec(pos, q, d)
        W -= KE * (gr.T @ o + gd.T @ d)
        S2 = A * A + Bs * Bs
        k2 = np.sum(k * k, 1)
        fac = f * S2 * 2.0 * (1.0 / (4.0 * self.b0**2) + 1.0 / k2)
        W += KE * (np.einsum("k,ka,kb->ab", fac, k, k) - np.sum(f * S2) * np.eye(3))
        e_bg = -KE * np.pi * np.sum(q) ** 2 / (2.0 * V * self.b0**2)
        W -= e_bg * np.eye(3)
        if self.lj_lrc:
            W -= lj_long_range(t.lj_rmin_half, t.lj_sqrt_eps, V, self.rc) * np.eye(3)
        if not molecular:  # covalent frames strain with the atoms
            _, gdd, _ = self.U(pos, q, d, g)
            v = pos[t.cov_j] - pos[t.cov_i]
            L = np.linalg.norm(v, axis=-1, keepdims=True)
            u = v / L
            gi = gdd[t.cov_i]
            tv = t.cov_c[:, None] * (gi - np.sum(gi * u, -1, keepdims=True) * u) / L
            W += tv.T @ v
        return W

    def virial_derivative(self, pos: np.ndarray, molecular: bool = True) -> np.ndarray:
        """Return the strain derivative plus the LJ tail impulse -E_lrc I (periodic.py:253-268, lj.py:190-201)."""
        W = self.strain_derivative(pos, molecular)
        if self.lj_lrc:
            W = W - lj_long_range(self.top.lj_rmin_half, self.top.lj_sqrt_eps, abs(np.linalg.det(self.H)), self.rc) * np.eye(3)
        return W

    def pressure(self, pos: np.ndarray, molecular: bool = True) -> float:
        """Return the static pressure -tr(virial_derivative)/(3V) [bar] (periodic.py:84-106, :270-281)."""
        V = abs(np.linalg.det(self.H))
        return float(-np.trace(self.virial_derivative(pos, molecular)) / (3.0 * V) * BAR_PER_KJMOL_NM3)
