"""PGMBeads on the MI355X: the pGM force field of every ring-polymer bead in the same launches.

Mirrors pgm_jax/md/pimd.py `PGMBeads` (:816-1172) and `PGMBeadState` (:777-813) for the force beads
of the qpi + LJ + PME hot path:

    beads = PGMBeads(system, H, MDSettings(precision="mixed"), beads=32)
    e = beads.init_state()
    forces, U, e = beads.compute(q, H, e)        # q, forces: (P, N, 3); U: sum over the beads

The reference evaluates the beads with jax.vmap (in lax.map chunks of 8, pimd.py:1068-1102).  Here
the P beads are one device context whose atom axis holds the beads one after the other: every
stage (rows, pair tensors, row field, PME spread / FFT / gather, CG updates, forces) is one launch
for all beads, while each bead keeps its own rows, PME grid, induced dipoles, predictor history and
CG scalars (bead-wise FP64 reductions, bead-wise convergence).  Every bead has its own neighbour
rows from the library's cell list (the reference shares the centroid's list with a margin), so the
result per bead is the force field's own on that bead.  Path-integral propagation (normal modes,
thermostats) stays with the caller; no CPU fallback.
"""

from __future__ import annotations

import ctypes as C
import dataclasses

import numpy as np

from . import _lib
from .forcefield import InductionState, _autotune, _dev, _make_context, _ptr, _set_transform, _torch, _transform
from .settings import MDSettings, grid_size
from .system import System


@dataclasses.dataclass
class PGMBeadState:
    """Engine state of PGMBeads (pimd.py:777-813).

    induction: InductionState batched over the beads (mu (P, N, 3), hist (4, P, N, 3), count shared);
    iters: CG iterations of the last solve (largest over the beads); max_iters / resid: largest
    since the block started; overflow: row capacity grown; cg_total: iters summed over the calls;
    elec, vdw: bead averages of the force field's parts [kJ/mol].
    """

    induction: InductionState
    iters: int = 0
    max_iters: int = 0
    resid: float = 0.0
    overflow: bool = False
    cg_total: int = 0
    elec: float = 0.0
    vdw: float = 0.0

    def reset_block(self) -> "PGMBeadState":
        """Return the state with the block maxima reset (pimd.py:1058-1060)."""
        return dataclasses.replace(self, max_iters=0, resid=0.0, overflow=False)


class PGMBeads:
    """The pGM + LJ force field on P beads of one system (pimd.py:816-1172, force beads only)."""

    def __init__(self, sys: System, H, settings: MDSettings = MDSettings(), beads: int = 32) -> None:
        """Build the bead context: P copies of the topology, P PME grids, one batched FFT plan per direction.

        Raises ValueError for fewer than one bead, the learned guess or the Ewald reciprocal sum
        (beads run with smooth PME and a polynomial or previous-step predictor).
        """
        if int(beads) < 1:
            raise ValueError("beads must be at least 1")
        if settings.predictor == "learned" or settings.recip != "pme":
            raise ValueError("beads run with smooth PME and the mu4 / mu3 / previous / none predictors")
        self.sys, self.s = sys, settings
        self.P = self.nf = int(beads)
        self.n = sys.n
        H = np.ascontiguousarray(H, float)
        self.K = tuple(int(k) for k in (settings.pme_grid or grid_size(H, settings.pme_spacing)))
        self._ctx, self._stream = _make_context(sys, H, settings, self.K, self.P)
        self._H = H.copy()
        self._built_at = None

    def __del__(self) -> None:
        ctx = getattr(self, "_ctx", None)
        if ctx:
            try:
                _lib.lib().pgm_ctx_destroy(ctx)
            except Exception:  # noqa: BLE001 (interpreter shutdown)
                pass
            self._ctx = None

    def init_state(self) -> PGMBeadState:
        """Return a fresh engine state: zero dipoles and histories on every bead."""
        torch = _torch()
        mu = torch.zeros((self.P, self.n, 3), dtype=torch.float64, device="cuda")
        hist = torch.zeros((4, self.P, self.n, 3), dtype=torch.float64, device="cuda")
        return PGMBeadState(InductionState(mu, hist, 0, 1.0))

    def _box(self, H) -> None:
        if H is None:
            return
        H = np.ascontiguousarray(H, float)
        if not np.array_equal(H, self._H):
            _lib.check(_lib.lib().pgm_set_box(self._ctx, _ptr(H, C.c_double)), "pgm_set_box")
            self._H = H.copy()
            self._built_at = None

    def _ensure_rows(self, q) -> None:
        torch = _torch()
        if self._built_at is not None:
            disp = torch.max(torch.linalg.norm((q - self._built_at).reshape(-1, 3), dim=1)).item()
            if disp <= 0.5 * self.s.skin:
                return
        _lib.check(_lib.lib().pgm_build_rows(self._ctx, C.c_void_p(q.data_ptr())), "pgm_build_rows")
        self._built_at = q.clone()

    def autotune(self, q, reps: int = 20) -> dict:
        """Time every row-field layout, PME spread-tile width and reciprocal transform on the beads q (P, N, 3); keep the fastest."""
        q = _dev(q, (self.P, self.n, 3))
        self._built_at = None
        return _autotune(self._ctx, q, reps)

    def set_pme_transform(self, transform: str = "fused") -> None:
        """Reciprocal transform of the CG field evaluations: "fused" kernels or the "library" FFT plans."""
        _set_transform(self._ctx, transform)

    def pme_transform(self) -> tuple[str, bool]:
        """Return the transform in use and whether the fused one is available for this grid."""
        return _transform(self._ctx)

    def time_kernels(self, q, reps: int = 20) -> dict:
        """Average duration [ms] of the hot kernels of one launch for all beads at q (P, N, 3) (pgm_time_kernels):
        "spread", "row_field", "gather", the transform kernels and "matvec" (one field evaluation)."""
        q = _dev(q, (self.P, self.n, 3))
        self._ensure_rows(q)
        ms = (C.c_double * 8)()
        _lib.check(_lib.lib().pgm_time_kernels(self._ctx, C.c_void_p(q.data_ptr()), int(reps), ms), "pgm_time_kernels")
        if self.pme_transform()[0] == "fused":
            names = ("merge", "spread", "fwd_planes", "x_conv", "bwd_planes", "row_field", "gather", "matvec")
            return {k: v for k, v in zip(names, list(ms)) if k != "merge"}
        names = ("merge", "spread", "fft_r2c", "convolve", "fft_c2r", "row_field", "gather", "matvec")
        return dict(zip(names, list(ms)))

    def bead_results(self) -> tuple[list[dict], np.ndarray, np.ndarray]:
        """Per-bead energies {"elec","vdw","total"}, CG iterations and final residuals of the last compute."""
        P = self.P
        e = (_lib.Energy * P)()
        it = np.zeros(P, np.int32)
        res = np.zeros(P, np.float64)
        _lib.check(_lib.lib().pgm_bead_results(self._ctx, C.cast(e, C.c_void_p), it.ctypes.data_as(C.c_void_p),
                                               res.ctypes.data_as(C.c_void_p)), "pgm_bead_results")
        return [{"elec": x.elec, "vdw": x.vdw, "total": x.total} for x in e], it, res

    def compute(self, q, box=None, e: PGMBeadState | None = None):
        """Return the forces (P, N, 3) [kJ/mol/nm], U [kJ/mol] and the engine state at beads q (P, N, 3).

        U is the sum of the beads' total energies (pimd.py:1104-1172 without contraction or bonded
        terms); the state's elec / vdw are bead averages, iters the largest CG count over the beads.
        """
        torch = _torch()
        q = _dev(q, (self.P, self.n, 3))
        self._box(box)
        self._ensure_rows(q)
        e = self.init_state() if e is None else e
        ind = e.induction
        forces = torch.empty((self.P, self.n, 3), dtype=torch.float64, device="cuda")
        cind = _lib.Induction(ind.mu.data_ptr(), ind.hist.data_ptr(), ind.count, ind.norm)
        en = _lib.Energy()
        it = C.c_int()
        res = C.c_double()
        st = _lib.check(_lib.lib().pgm_compute(self._ctx, C.c_void_p(q.data_ptr()), C.byref(cind), C.byref(en),
                                               C.c_void_p(forces.data_ptr()), C.byref(it), C.byref(res)), "pgm_compute")
        new = dataclasses.replace(
            e,
            induction=InductionState(ind.mu, ind.hist, cind.count, cind.norm),
            iters=it.value,
            max_iters=max(e.max_iters, it.value),
            resid=max(e.resid, res.value),
            overflow=e.overflow or st == _lib.PGM_ROW_OVERFLOW,
            cg_total=e.cg_total + it.value,
            elec=en.elec / self.P,
            vdw=en.vdw / self.P,
        )
        return forces, en.total, new
