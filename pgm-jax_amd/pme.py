"""The PME object of the reference on the GPU (pgm_jax/md/pme.py:209-411).

`PME(grid, order, beta, dtype)` with `influence(H)`, `setup(pos, H)`, `spread(S, q, d)`,
`grad_dipoles(S, G, q, d)` and `energy(S, G, q, d)`, the reference's methods with their argument
meaning and units, computed by the library's PME kernels (pgm_pme_influence, pgm_pme_spread,
pgm_pme_energy) on a context that holds N point sites with the given charges and nothing else
(no pair terms are evaluated).  dtype float32 runs the FP32 kernels with FP64 sums ("mixed"),
float64 the FP64 ones.  S holds the positions (device float64) and the box; G is the reference's
influence array, accepted for the signature: the kernels use the influence of S's box, which is
what every caller of the reference passes.
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .forcefield import PGMForceField, _dev, _torch
from .settings import MDSettings
from .system import System


def _heights(H: np.ndarray) -> np.ndarray:
    """Distances between opposite faces of the box with lattice vectors as rows [nm]."""
    V = abs(np.linalg.det(H))
    return np.array([V / np.linalg.norm(np.cross(H[(a + 1) % 3], H[(a + 2) % 3])) for a in range(3)])


class PME:
    """Smooth PME of point charges and dipoles (pgm_jax/md/pme.py:209-241)."""

    def __init__(self, grid, order: int, beta: float, dtype=np.float32) -> None:
        self.K = tuple(int(k) for k in grid)
        self.order, self.beta = int(order), float(beta)
        self.precision = "double" if np.dtype(dtype) == np.float64 else "mixed"
        self._ff: PGMForceField | None = None
        self._n = -1
        self._q = None

    def _context(self, n: int, H) -> PGMForceField:
        H = np.ascontiguousarray(H, float)
        if self._ff is None or self._n != n:
            sys_ = System(q=np.zeros(n), radius=np.ones(n), alpha=np.zeros(n), lj_rmin_half=np.zeros(n),
                          lj_sqrt_eps=np.zeros(n), masses=np.ones(n), mol=np.arange(n, dtype=np.int32),
                          cov_i=np.zeros(0, np.int32), cov_j=np.zeros(0, np.int32), cov_c=np.zeros(0))
            rc = float(min(0.3, 0.2 * _heights(H).min()))
            st = MDSettings(precision=self.precision, cutoff=rc, skin=0.0, ewald_beta=self.beta, pme_grid=self.K,
                            pme_order=self.order, lj_lrc=False)
            self._ff, self._n, self._q = PGMForceField(sys_, H, st), n, None
        else:
            self._ff._box(H)
        return self._ff

    def _charges(self, ff: PGMForceField, q) -> int:
        """Put the charges q on the context; returns with_charges (0 when every charge is zero)."""
        q = np.ascontiguousarray(q.cpu().numpy() if hasattr(q, "cpu") else np.asarray(q, float), float).reshape(-1)
        if q.shape != (self._n,):
            raise ValueError(f"charges of shape {q.shape}, expected ({self._n},)")
        if not q.any():
            return 0
        if self._q is None or not np.array_equal(q, self._q):
            ff.set_params({"q": q})
            self._q = q.copy()
        return 1

    def influence(self, H):
        """G (K1, K2, K3//2 + 1) float64 [1/nm]: exp(-pi^2 |m*|^2 / beta^2) / (2 pi V |m*|^2) times the
        spline moduli and the half-space weights, 0 at m = 0 (pgm_jax/md/pme.py:243-267)."""
        ff = self._context(max(self._n, 1), H)
        K1, K2, K3 = self.K
        G = np.empty((K1, K2, K3 // 2 + 1))
        _lib.check(_lib.lib().pgm_pme_influence(ff._ctx, G.ctypes.data_as(C.POINTER(C.c_double))), "pgm_pme_influence")
        return G

    def setup(self, pos, H) -> dict:
        """S for the positions pos (N, 3) [nm] in the box H (pgm_jax/md/pme.py:269-304): the splines
        are computed by each call from these positions."""
        pos = _dev(pos)
        if pos.ndim != 2 or pos.shape[1] != 3:
            raise ValueError(f"positions of shape {tuple(pos.shape)}, expected (N, 3)")
        H = np.ascontiguousarray(H, float)
        self._context(pos.shape[0], H)
        return {"pos": pos, "H": H}

    def spread(self, S: dict, q, d):
        """Q (K1, K2, K3) float64 device tensor of the charges q (N,) [e] and dipoles d (N, 3) [e nm]
        (pgm_jax/md/pme.py:306-337)."""
        torch = _torch()
        ff = self._context(S["pos"].shape[0], S["H"])
        wq = self._charges(ff, q)
        d = _dev(d, (self._n, 3))
        Q = torch.empty(self.K, dtype=torch.float64, device="cuda")
        _lib.check(_lib.lib().pgm_pme_spread(ff._ctx, C.c_void_p(S["pos"].data_ptr()), wq, C.c_void_p(d.data_ptr()),
                                             C.c_void_p(Q.data_ptr())), "pgm_pme_spread")
        return Q

    def _energy_grad(self, S: dict, G, q, d) -> tuple:
        K1, K2, K3 = self.K
        if G is not None and tuple(np.shape(G)) != (K1, K2, K3 // 2 + 1):
            raise ValueError(f"influence of shape {np.shape(G)}, expected {(K1, K2, K3 // 2 + 1)}")
        ff = self._context(S["pos"].shape[0], S["H"])
        wq = self._charges(ff, q)
        return ff.pme_energy(S["pos"], d, bool(wq), S["H"])

    def grad_dipoles(self, S: dict, G, q, d):
        """dU_rec/dd (N, 3) float64 device tensor [e/nm^2] (pgm_jax/md/pme.py:339-388)."""
        return self._energy_grad(S, G, q, d)[1]

    def energy(self, S: dict, G, q, d) -> float:
        """E_rec [e^2/nm] accumulated in float64 (pgm_jax/md/pme.py:390-411)."""
        return float(self._energy_grad(S, G, q, d)[0])
