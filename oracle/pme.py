"""Smooth PME for Gaussian charges and point dipoles, float64 (restates pgm_jax/md/pme.py).

E_rec = sum_{m != 0} exp(-pi^2 |m*|^2 / beta^2) / (2 pi V |m*|^2) B(m) |F[Q](m)|^2       (pme.py:9-25)
Q(k)  = sum_i q_i M1 M2 M3 + e_i . grad_w(M1 M2 M3),  w_i = K frac(r_i H^-1),  e_i = K (d_i H^-1)

Functions: `bspline` (weights, first and second derivatives; OpenMM recursion, pme.py:57-103),
`bspline_moduli` (pme.py:106-131), `grid_size` (pme.py:134-172), `influence` (pme.py:243-267),
`setup` (pme.py:269-304), `spread` (pme.py:306-337), `potential` (the grid dU/dQ of pme.py:373-376),
`grad_dipoles` (pme.py:339-388), `energy` (pme.py:390-411) and `gradients` (dU/dd and dU/dr at
fixed q, d: the autodiff of `energy` that pgm_jax/md/forcefield.py:2423 takes, written out).

Parity: the splines are pinned by partition of unity, the derivative identities (finite
differences of the weights) and the moduli by a direct DFT of M_p; the reciprocal energy and its
gradients by the exact Ewald structure-factor sum (oracle/ewald.py) as the grid grows (the
reference's own test_md.py:39-48 criterion) and by finite differences of `energy`; the whole
reciprocal part by the golden ff_small_box energies (PME 32^3, order 6).
"""

from __future__ import annotations

import numpy as np


def bspline(f: np.ndarray, order: int) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Return theta, dtheta, d2theta (..., p) at offsets f in [0, 1) (pme.py:57-103 and its derivative).

    theta[..., k] = M_p(f + p - 1 - k), the weight of grid point floor(w) + k; dtheta = d theta/df
    (from the order p-1 values: M_p' (x) = M_{p-1}(x) - M_{p-1}(x-1)); d2theta from the order p-2
    values (M_p'' (x) = M_{p-2}(x) - 2 M_{p-2}(x-1) + M_{p-2}(x-2)).
    """
    if order < 4:
        raise ValueError("PME order must be >= 4")
    f = np.asarray(f, float)
    data = [np.zeros_like(f) for _ in range(order)]
    data[0] = 1.0 - f
    data[1] = f.copy()
    saved = {2: list(data)}
    for j in range(3, order + 1):  # raise to order j: data[0..j-1] = M_j(f + j - 1 - k)
        div = 1.0 / (j - 1.0)
        data[j - 1] = div * f * data[j - 2]
        for k in range(1, j - 1):
            data[j - k - 1] = div * ((f + k) * data[j - k - 2] + (j - k - f) * data[j - k - 1])
        data[0] = div * (1.0 - f) * data[0]
        saved[j] = list(data)
    m1, m2 = saved[order - 1], saved[order - 2]
    zero = np.zeros_like(f)

    def at(d: list, k: int, n: int) -> np.ndarray:  # d[k] of an order-n array, 0 outside 0..n-1
        return d[k] if 0 <= k < n else zero

    th = np.stack(saved[order][:order], -1)
    dth = np.stack([at(m1, k - 1, order - 1) - at(m1, k, order - 1) for k in range(order)], -1)
    d2th = np.stack(
        [at(m2, k - 2, order - 2) - 2.0 * at(m2, k - 1, order - 2) + at(m2, k, order - 2) for k in range(order)], -1
    )
    return th, dth, d2th


def bspline_moduli(K: int, order: int) -> np.ndarray:
    """Return |sum_j M_p(j + 1) exp(2 pi i m j / K)|^2, m = 0..K-1, zeros patched (pme.py:106-131)."""
    data = bspline(np.zeros(()), order)[0]
    j = np.arange(order)
    m = np.arange(K)[:, None]
    arg = 2 * np.pi * m * j[None, :] / K
    mod = (data * np.cos(arg)).sum(1) ** 2 + (data * np.sin(arg)).sum(1) ** 2
    for i in range(K):
        if mod[i] < 1e-7:
            mod[i] = 0.5 * (mod[(i - 1) % K] + mod[(i + 1) % K])
    return mod


def grid_size(H: np.ndarray, spacing: float = 0.05, factors=(2, 3, 5)) -> tuple[int, int, int]:
    """Return the smallest even (2,3,5)-smooth grid with plane spacing <= spacing (pme.py:134-172)."""
    H = np.asarray(H, float)
    V = abs(np.linalg.det(H))
    heights = [
        V / np.linalg.norm(np.cross(H[1], H[2])),
        V / np.linalg.norm(np.cross(H[2], H[0])),
        V / np.linalg.norm(np.cross(H[0], H[1])),
    ]

    def ok(n: int) -> bool:
        for f in factors:
            while n % f == 0:
                n //= f
        return n == 1

    out = []
    for h in heights:
        n = max(8, int(np.ceil(h / spacing)))
        while not ok(n) or n % 2:
            n += 1
        out.append(n)
    return tuple(out)


class PME:
    """Reciprocal-space energy and gradients of charges and point dipoles (pme.py:175-411), float64."""

    def __init__(self, grid, order: int, beta: float) -> None:
        """Grid (K1, K2, K3), B-spline order p, Ewald coefficient beta [1/nm] (pme.py:209-241)."""
        self.K = tuple(int(k) for k in grid)
        self.order, self.beta = int(order), float(beta)
        K1, K2, K3 = self.K
        mods = [bspline_moduli(k, self.order) for k in self.K]
        Binv = 1.0 / (mods[0][:, None, None] * mods[1][None, :, None] * mods[2][None, None, : K3 // 2 + 1])
        w3 = np.full(K3 // 2 + 1, 2.0)
        w3[0] = 1.0
        if K3 % 2 == 0:
            w3[-1] = 1.0
        self.Bw = Binv * w3[None, None, :]
        self.w3 = w3
        self.m = [np.fft.fftfreq(K1, 1.0 / K1), np.fft.fftfreq(K2, 1.0 / K2), np.arange(K3 // 2 + 1, dtype=float)]

    def influence(self, H: np.ndarray) -> np.ndarray:
        """Return G(m) on the half spectrum, with moduli and half-space weights (pme.py:243-267)."""
        H = np.asarray(H, float)
        R = np.linalg.inv(H).T
        m1, m2, m3 = self.m
        mv = m1[:, None, None, None] * R[0] + m2[None, :, None, None] * R[1] + m3[None, None, :, None] * R[2]
        msq = np.sum(mv * mv, -1)
        V = abs(np.linalg.det(H))
        safe = np.where(msq > 0, msq, 1.0)
        G = np.where(msq > 0, np.exp(-(np.pi**2) * safe / self.beta**2) / (2 * np.pi * V * safe), 0.0)
        return G * self.Bw

    def setup(self, pos: np.ndarray, H: np.ndarray) -> dict:
        """Return the per-atom stencil base, weights (3 derivatives) and e_scale (pme.py:269-304)."""
        H = np.asarray(H, float)
        Hinv = np.linalg.inv(H)
        u = np.asarray(pos, float) @ Hinv
        Kf = np.array(self.K, float)
        w = (u - np.floor(u)) * Kf
        base = np.floor(w)
        f = w - base
        th, dth, d2th = bspline(f, self.order)  # (N, 3, p)
        return {
            "base": base.astype(np.int64),
            "th": th,
            "dth": dth,
            "d2th": d2th,
            "e_scale": Hinv * Kf[None, :],
        }

    def _flat(self, S: dict) -> np.ndarray:
        """Return the flat grid indices (N, p, p, p) of every atom's stencil (pme.py:293-298)."""
        K1, K2, K3 = self.K
        ar = np.arange(self.order)
        idx = (S["base"][:, :, None] + ar[None, None, :]) % np.array(self.K)[None, :, None]
        return (idx[:, 0, :, None, None] * K2 + idx[:, 1, None, :, None]) * K3 + idx[:, 2, None, None, :]

    def spread(self, S: dict, q: np.ndarray, d: np.ndarray) -> np.ndarray:
        """Return the grid Q (K1, K2, K3) of charges q and dipoles d (pme.py:306-337)."""
        e = np.asarray(d, float) @ S["e_scale"]
        t1, t2, t3 = (S["th"][:, a] for a in range(3))
        d1, d2, d3 = (S["dth"][:, a] for a in range(3))
        a1 = np.asarray(q, float)[:, None] * t1 + e[:, 0:1] * d1
        val = (
            a1[:, :, None, None] * t2[:, None, :, None] * t3[:, None, None, :]
            + (e[:, 1:2] * t1)[:, :, None, None] * d2[:, None, :, None] * t3[:, None, None, :]
            + (e[:, 2:3] * t1)[:, :, None, None] * t2[:, None, :, None] * d3[:, None, None, :]
        )
        Q = np.zeros(int(np.prod(self.K)))
        np.add.at(Q, self._flat(S).reshape(-1), val.reshape(-1))
        return Q.reshape(self.K)

    def potential(self, S: dict, G: np.ndarray, q: np.ndarray, d: np.ndarray) -> np.ndarray:
        """Return phi = dE_rec/dQ on the grid: 2 K1 K2 K3 irfft(G/w3 rfft(Q)) (pme.py:373-376)."""
        Q = self.spread(S, q, d)
        K1, K2, K3 = self.K
        return np.fft.irfftn(np.fft.rfftn(Q) * (G / self.w3[None, None, :]), Q.shape, axes=(0, 1, 2)) * (2.0 * K1 * K2 * K3)

    def energy(self, S: dict, G: np.ndarray, q: np.ndarray, d: np.ndarray) -> float:
        """Return E_rec = sum G |F[Q]|^2 [e^2/nm] (pme.py:390-411)."""
        FQ = np.fft.rfftn(self.spread(S, q, d))
        return float(np.sum(G * (FQ.real**2 + FQ.imag**2)))

    def gradients(self, S: dict, G: np.ndarray, q: np.ndarray, d: np.ndarray, phi: np.ndarray | None = None) -> tuple:
        """Return (dE_rec/dd (N, 3), dE_rec/dr (N, 3)) at fixed q and d [e/nm^2, e^2/nm^2].

        dE/dd is pme.py:339-388 (`grad_dipoles`); dE/dr is the position derivative through the
        spline weights (what forcefield.py:2423 gets by autodiff of `energy`): with phi = dE/dQ,
        dE/dw_a = sum_stencil phi dQ/dw_a and dE/dr = e_scale dE/dw.
        """
        if phi is None:
            phi = self.potential(S, G, q, d)
        n = len(q)
        p = self.order
        P = phi.reshape(-1)[self._flat(S)].reshape(n, p, p, p)
        e = np.asarray(d, float) @ S["e_scale"]
        t = [S["th"][:, a] for a in range(3)]
        dt = [S["dth"][:, a] for a in range(3)]
        d2t = [S["d2th"][:, a] for a in range(3)]

        def contract(w1, w2, w3):  # sum_abc P[n,a,b,c] w1[n,a] w2[n,b] w3[n,c]
            return np.einsum("nabc,na,nb,nc->n", P, w1, w2, w3, optimize=True)

        g = np.stack([contract(dt[0], t[1], t[2]), contract(t[0], dt[1], t[2]), contract(t[0], t[1], dt[2])], -1)
        qv = np.asarray(q, float)
        dw = []
        for a in range(3):
            w = [t[0], t[1], t[2]]
            w[a] = dt[a]
            s = qv * contract(*w)
            for b in range(3):  # e_b times the derivative along a of the dipole spline along b
                ww = [t[0], t[1], t[2]]
                if a == b:
                    ww[a] = d2t[a]
                else:
                    ww[a] = dt[a]
                    ww[b] = dt[b]
                s = s + e[:, b] * contract(*ww)
            dw.append(s)
        dw = np.stack(dw, -1)
        return g @ S["e_scale"].T, dw @ S["e_scale"].T
