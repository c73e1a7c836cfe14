"""PGMForceField on the MI355X: the reference's MD force-field entry points over the HIP C ABI.

Mirrors pgm_jax/md/forcefield.py `PGMForceField` (:810-2695) for the qpi + LJ + PME hot path:

    ff = PGMForceField(system, H, MDSettings(precision="mixed"))
    res = ff.compute(pos, H, None, ff.init_induction())
    res.energy["total"], res.forces, res.induction.mu, res.iterations, res.residual

Arrays are torch float64 tensors on the GPU (positions, forces, dipoles); numpy inputs are copied to
the device.  Rows come from the library's own cell list (`idx` of the reference is accepted and
ignored); they are rebuilt when an atom has moved by more than half the skin since the last build.
Errors raise PGMError; there is no CPU fallback.
"""

from __future__ import annotations

import ctypes as C
import dataclasses

import numpy as np

from . import _lib
from .settings import PREDICTORS, MDSettings, grid_size
from .system import System


def _torch():
    import torch

    if not torch.cuda.is_available():
        raise _lib.PGMError("no GPU visible: the HIP path needs an AMD GPU")
    return torch


def _dev(x, shape=None):
    """Return x as a contiguous float64 tensor on the GPU."""
    torch = _torch()
    t = x if isinstance(x, torch.Tensor) else torch.as_tensor(np.asarray(x, float))
    t = t.to(device="cuda", dtype=torch.float64).contiguous()
    if shape is not None and tuple(t.shape) != tuple(shape):
        raise ValueError(f"expected shape {shape}, got {tuple(t.shape)}")
    return t


def _ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


@dataclasses.dataclass
class InductionState:
    """Induced-dipole solver state (forcefield.py:679-713): mu (N,3), hist (4,N,3) newest first, count, norm."""

    mu: object
    hist: object
    count: int
    norm: float


@dataclasses.dataclass
class Result:
    """Result of compute (forcefield.py:716-746): energy {"elec","vdw","total"} (and "field" with an
    external field) [kJ/mol], forces (N,3) [kJ/mol/nm]."""

    energy: dict
    forces: object
    induction: InductionState
    iterations: int
    residual: float
    overflow: bool = False
    dipole: object = None  # with an external field: M = sum q r + sum (p + mu) + offset (3,) [e nm]


def _make_context(sys: System, H: np.ndarray, settings: MDSettings, K: tuple, nbead: int = 1):
    """Create the device context of sys (nbead copies) on torch's current stream; return (ctx, stream)."""
    torch = _torch()
    L = _lib.lib()
    top = _lib.Topology(
        sys.n, _ptr(sys.q, C.c_double), _ptr(sys.radius, C.c_double), _ptr(sys.alpha, C.c_double),
        _ptr(sys.lj_rmin_half, C.c_double), _ptr(sys.lj_sqrt_eps, C.c_double), _ptr(sys.mol, C.c_int),
        len(sys.cov_i), _ptr(sys.cov_i, C.c_int), _ptr(sys.cov_j, C.c_int), _ptr(sys.cov_c, C.c_double),
    )
    cs = _lib.Settings()
    cs.cutoff, cs.skin, cs.ewald_beta = settings.cutoff, settings.skin, settings.ewald_beta
    cs.grid[:] = list(K)
    cs.order = settings.pme_order
    cs.lj_lrc = int(settings.lj_lrc)
    cs.dipole_tol, cs.max_iter = settings.dipole_tol, settings.max_iter
    cs.predictor = PREDICTORS[settings.predictor]
    cs.fused, cs.norm_refresh, cs.peek = int(settings.fused), settings.norm_refresh, settings.peek
    cs.precision = 1 if settings.precision == "mixed" else 0
    cs.recip = 1 if settings.recip == "ewald" else 0
    cs.k_tol = settings.k_tol
    cs.nbead = int(nbead)
    cs.precond = 1 if settings.precond == "molecule" else 0
    cs.elec_cutoff = 0.0 if settings.elec_cutoff is None else float(settings.elec_cutoff)
    stream = torch.cuda.current_stream()
    ctx = C.c_void_p()
    _lib.check(L.pgm_ctx_create(C.byref(ctx), C.byref(top), C.byref(cs), _ptr(H, C.c_double),
                                C.c_void_p(stream.cuda_stream)), "pgm_ctx_create")
    return ctx, stream


FORKS = ("start", "after_spread", "after_forward_planes", "start_join_before_planes", "none")


def _autotune(ctx, pos, reps: int) -> dict:
    """pgm_autotune on a context; return the chosen shapes and their timings."""
    rep = (C.c_double * 12)()
    _lib.check(_lib.lib().pgm_autotune(ctx, C.c_void_p(pos.data_ptr()), int(reps), rep), "pgm_autotune")
    return {"row_layout": ("atom", "molecule")[int(rep[0])], "row_lanes": int(rep[1]),
            "spread_tile": tuple(int(x) for x in rep[2:5]), "row_field_ms": rep[5], "pme_ms": rep[6],
            "layouts_tried": int(rep[7]), "pme_transform": ("library", "fused")[int(rep[8])],
            "other_transform_ms": rep[9], "row_fork": FORKS[int(rep[10])],
            "field_eval_ms": rep[11]}


def _set_transform(ctx, transform: str) -> None:
    code = {"library": 0, "fused": 1}[transform]
    _lib.check(_lib.lib().pgm_set_pme_transform(ctx, code), "pgm_set_pme_transform")


def _transform(ctx) -> tuple[str, bool]:
    fused, avail = C.c_int(), C.c_int()
    _lib.check(_lib.lib().pgm_pme_transform(ctx, C.byref(fused), C.byref(avail)), "pgm_pme_transform")
    return ("library", "fused")[fused.value], bool(avail.value)


class PGMForceField:
    """The pGM + LJ force field of one system on the GPU (forcefield.py:810-1000)."""

    def __init__(self, sys: System, H, settings: MDSettings = MDSettings()) -> None:
        """Build the device context: topology, PME grid (settings or grid_size(H, spacing)), plans."""
        self.sys, self.s = sys, settings
        self.n = sys.n
        H = np.ascontiguousarray(H, float)
        self.K = tuple(int(k) for k in (settings.pme_grid or grid_size(H, settings.pme_spacing)))
        self._ctx, self._stream = _make_context(sys, H, settings, self.K)
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

    # ------------------------------------------------------------------ state
    def init_induction(self) -> InductionState:
        """Return a fresh solver state: zero dipoles and history, normaliser 1 (forcefield.py:1436-1450)."""
        torch = _torch()
        z = torch.zeros((self.n, 3), dtype=torch.float64, device="cuda")
        return InductionState(z, torch.zeros((4, self.n, 3), dtype=torch.float64, device="cuda"), 0, 1.0)

    def _box(self, H) -> None:
        if H is None:
            return
        H = np.ascontiguousarray(H, float)
        if not np.array_equal(H, self._H):
            _lib.check(_lib.lib().pgm_set_box(self._ctx, _ptr(H, C.c_double)), "pgm_set_box")
            self._H = H.copy()
            self._built_at = None

    def build_rows(self, pos) -> None:
        """Rebuild the neighbour rows (cell list, cutoff + skin) at pos."""
        _lib.check(_lib.lib().pgm_build_rows(self._ctx, C.c_void_p(pos.data_ptr())), "pgm_build_rows")
        self._built_at = pos.clone()

    def _ensure_rows(self, pos) -> None:
        torch = _torch()
        if self._built_at is None:
            self.build_rows(pos)
            return
        disp = torch.max(torch.linalg.norm(pos - self._built_at, dim=1)).item()
        if disp > 0.5 * self.s.skin:
            self.build_rows(pos)

    def _efield(self, efield) -> bool:
        """Put the external field of an efield argument on the context (pgm_set_efield): None, or
        (E [V/nm] (3,), offset [e nm] (3,) or None[, "E"]), the constant-field kind of
        PGMForceField._ext (forcefield.py:2048-2073).  Returns whether a field is on."""
        if efield is None:
            key = None
        else:
            kind = efield[2] if len(efield) > 2 else "E"
            if kind == "D":
                raise NotImplementedError("the constant-displacement field (efield kind 'D') is not built")
            if kind != "E":
                raise ValueError(f"field kind {kind!r}: 'E' (constant field) or 'D' (constant displacement)")
            E = np.ascontiguousarray(np.asarray(efield[0], float).reshape(3))
            off = None if efield[1] is None else np.ascontiguousarray(np.asarray(efield[1], float).reshape(3))
            key = (tuple(E), None if off is None else tuple(off))
        if key != getattr(self, "_efield_now", None):
            if key is None:
                _lib.check(_lib.lib().pgm_set_efield(self._ctx, None, None), "pgm_set_efield")
            else:
                _lib.check(_lib.lib().pgm_set_efield(self._ctx, _ptr(E, C.c_double),
                                                     None if off is None else _ptr(off, C.c_double)), "pgm_set_efield")
            self._efield_now = key
        return key is not None

    def _field_result(self) -> tuple[float, np.ndarray]:
        """(energy["field"] [kJ/mol], M (3,) [e nm]) of the last evaluation (pgm_field_result)."""
        e, M = (C.c_double * 1)(), (C.c_double * 3)()
        _lib.check(_lib.lib().pgm_field_result(self._ctx, e, M), "pgm_field_result")
        return e[0], np.array(list(M), float)

    # ------------------------------------------------------------------ entry points
    def compute(self, pos, H=None, idx=None, ind: InductionState | None = None, params=None, efield=None) -> Result:
        """Solve the induced dipoles from the predictor, then return energy and forces (forcefield.py:2489-2560).

        params: a parameter pytree {quantity: (K,) array} of the system's tied table (System.expand;
        None: the parameters in use), set on the device before the evaluation (set_params).  Its
        gradient for a loss of the outputs is grad_params.  efield: None or (E [V/nm], offset [e nm]
        or None), a uniform static field (md/efield.py): energy["field"] = -E . M and Result.dipole =
        M; molecules whole in pos."""
        if params is not None:
            self.set_params(params)
        on = self._efield(efield)
        pos = _dev(pos, (self.n, 3))
        self._box(H)
        self._ensure_rows(pos)
        ind = self.init_induction() if ind is None else ind
        torch = _torch()
        forces = torch.empty((self.n, 3), dtype=torch.float64, device="cuda")
        cind = _lib.Induction(ind.mu.data_ptr(), ind.hist.data_ptr(), ind.count, ind.norm)
        e = _lib.Energy()
        it = C.c_int()
        res = C.c_double()
        st = _lib.check(_lib.lib().pgm_compute(self._ctx, C.c_void_p(pos.data_ptr()), C.byref(cind), C.byref(e),
                                               C.c_void_p(forces.data_ptr()), C.byref(it), C.byref(res)), "pgm_compute")
        new = InductionState(ind.mu, ind.hist, cind.count, cind.norm)
        energy = {"elec": e.elec, "vdw": e.vdw, "total": e.total}
        M = None
        if on:
            energy["field"], M = self._field_result()
        return Result(energy, forces, new, it.value, res.value, st == _lib.PGM_ROW_OVERFLOW, M)

    def energy(self, pos, H=None, idx=None, ind: InductionState | None = None, params=None, efield=None) -> tuple:
        """Return (total energy [kJ/mol], state, CG iterations, overflow) for Monte Carlo barostat trials
        (forcefield.py:2562-2626): the dipoles are solved from ind.mu (no predictor) into a new mu;
        the state's history, count and norm are those passed in.  efield as in compute (the total
        includes the field's energy)."""
        if params is not None:
            self.set_params(params)
        self._efield(efield)
        pos = _dev(pos, (self.n, 3))
        self._box(H)
        self._ensure_rows(pos)
        ind = self.init_induction() if ind is None else ind
        mu = ind.mu.clone()  # the caller's state stays as it was (a rejected trial keeps it)
        cind = _lib.Induction(mu.data_ptr(), ind.hist.data_ptr(), ind.count, ind.norm)
        e = _lib.Energy()
        it = C.c_int()
        res = C.c_double()
        st = _lib.check(_lib.lib().pgm_trial_energy(self._ctx, C.c_void_p(pos.data_ptr()), C.byref(cind), C.byref(e), C.byref(it),
                                              C.byref(res)), "pgm_trial_energy")
        return e.total, InductionState(mu, ind.hist, ind.count, ind.norm), it.value, st == _lib.PGM_ROW_OVERFLOW

    def energy_forces(self, pos, mu, H=None, efield=None) -> tuple[dict, object]:
        """Return (energy parts, forces) at fixed induced dipoles mu (forcefield.py:2249-2292, :2362-2430);
        efield as in compute (energy["field"])."""
        torch = _torch()
        on = self._efield(efield)
        pos = _dev(pos, (self.n, 3))
        mu = _dev(mu, (self.n, 3))
        self._box(H)
        self._ensure_rows(pos)
        forces = torch.empty((self.n, 3), dtype=torch.float64, device="cuda")
        e = _lib.Energy()
        _lib.check(_lib.lib().pgm_energy_forces(self._ctx, C.c_void_p(pos.data_ptr()), C.c_void_p(mu.data_ptr()), C.byref(e),
                                                C.c_void_p(forces.data_ptr())), "pgm_energy_forces")
        energy = {"elec": e.elec, "vdw": e.vdw, "total": e.total}
        if on:
            energy["field"] = self._field_result()[0]
        return energy, forces

    def energy_fixed_mu(self, pos, H, mu, efield=None) -> float:
        """Return the total energy [kJ/mol] with the induced dipoles held at mu (forcefield.py:2249-2292)."""
        return self.energy_forces(pos, mu, H, efield)[0]["total"]

    def strain_derivative(self, pos, H, mu, molecular: bool = True, efield=None):
        """Return dE/d eps (3, 3) [kJ/mol] at fixed mu, plus the tail impulse -E_lrc I with lj_lrc
        (forcefield.py:2628-2695), as a float64 numpy array.

        molecular=True: molecules are translated with their centres of mass (the system's masses;
        molecules whole in `pos`).  molecular=False: atomic scaling, every atom r -> (1 + eps) r (the
        covalent dipole frames strain with them).  An external field (efield) is not built here.
        """
        if efield is not None:
            raise NotImplementedError("strain_derivative with an external field is not built")
        self._efield(None)
        torch = _torch()
        pos = _dev(pos, (self.n, 3))
        mu = _dev(mu, (self.n, 3))
        self._box(H)
        self._ensure_rows(pos)
        if getattr(self, "_masses_dev", None) is None:
            self._masses_dev = torch.as_tensor(np.asarray(self.sys.masses, float), device="cuda")
        W = (C.c_double * 9)()
        e = _lib.Energy()
        _lib.check(_lib.lib().pgm_strain_derivative(self._ctx, C.c_void_p(pos.data_ptr()), C.c_void_p(mu.data_ptr()),
                                                    C.c_void_p(self._masses_dev.data_ptr()), int(bool(molecular)), C.byref(e),
                                                    None, W),
                   "pgm_strain_derivative")
        return np.array(list(W), float).reshape(3, 3)

    def perm_dipoles(self, pos, H=None):
        """Return the permanent dipoles p (N,3) [e nm] (forcefield.py:1032-1042)."""
        torch = _torch()
        pos = _dev(pos, (self.n, 3))
        self._box(H)
        p = torch.empty((self.n, 3), dtype=torch.float64, device="cuda")
        _lib.check(_lib.lib().pgm_perm_dipoles(self._ctx, C.c_void_p(pos.data_ptr()), C.c_void_p(p.data_ptr())),
                   "pgm_perm_dipoles")
        torch.cuda.current_stream().synchronize()
        return p

    def field(self, pos, d, with_charges: bool = True, H=None):
        """Return the total field -dU/dd (N,3) [e/nm^2] of the charges and dipoles d (forcefield.py:1428-1433)."""
        torch = _torch()
        pos = _dev(pos, (self.n, 3))
        d = _dev(d, (self.n, 3))
        self._box(H)
        self._ensure_rows(pos)
        out = torch.empty((self.n, 3), dtype=torch.float64, device="cuda")
        _lib.check(_lib.lib().pgm_field(self._ctx, C.c_void_p(pos.data_ptr()), int(with_charges), C.c_void_p(d.data_ptr()),
                                        C.c_void_p(out.data_ptr())), "pgm_field")
        return out

    def dipole_vjp(self, pos, H, mu, mu_bar, tol: float | None = None) -> dict:
        """The backward pass of the differentiable solve (the custom VJP of PGMForceField._solve,
        forcefield.py:1853-1915), through pgm_dipole_vjp: the adjoint CG A lam = mu_bar (the
        forward operator, x0 = 0, no peek, to `tol`, default settings.adjoint_tol), then lam pulled
        back through the residual b - A mu at the solved dipoles mu.

        Returns {"lam" (N,3), "q" (N,), "radius" (N,), "alpha" (N,), "p" (N,3), "cov" (n_cov,)}: the
        cotangents of the charges, Gaussian radii, polarizabilities, permanent dipoles and covalent
        dipole magnitudes (device float64 tensors; field and potential without the Coulomb constant,
        as `field`), plus "iterations" and "residual" of the adjoint CG.  Raises on a status < 0;
        a CG that stops at max_iter returns its results (status 3, as Result.residual reports).
        """
        torch = _torch()
        pos = _dev(pos, (self.n, 3))
        mu = _dev(mu, (self.n, 3))
        mu_bar = _dev(mu_bar, (self.n, 3))
        self._box(H)
        self._ensure_rows(pos)
        tol = float(self.s.adjoint_tol if tol is None else tol)
        self._efield(None)  # the derivatives are those of the field-free force field
        f64 = dict(dtype=torch.float64, device="cuda")
        out = {"lam": torch.empty((self.n, 3), **f64), "q": torch.empty(self.n, **f64), "radius": torch.empty(self.n, **f64),
               "alpha": torch.empty(self.n, **f64), "p": torch.empty((self.n, 3), **f64),
               "cov": torch.empty(max(len(self.sys.cov_i), 1), **f64)}
        it, res = C.c_int(), C.c_double()
        _lib.check(_lib.lib().pgm_dipole_vjp(self._ctx, C.c_void_p(pos.data_ptr()), C.c_void_p(mu.data_ptr()),
                                             C.c_void_p(mu_bar.data_ptr()), tol,
                                             *(C.c_void_p(out[k].data_ptr()) for k in ("lam", "q", "radius", "alpha", "p", "cov")),
                                             C.byref(it), C.byref(res)),
                   "pgm_dipole_vjp")
        out["cov"] = out["cov"][:len(self.sys.cov_i)]
        out["iterations"], out["residual"] = it.value, res.value
        return out

    def set_params(self, params: dict) -> None:
        """Use the parameter pytree `params` ({quantity: (K,) array} of the system's table, or per-atom
        arrays for a system given as arrays; missing quantities keep their values) from now on:
        pgm_set_params with the expanded per-atom arrays, skipped when they are the ones in use."""
        arrays = self.sys.expand(params)
        cur = getattr(self, "_params_now", None)
        if cur is not None and all(np.array_equal(arrays[k], cur[k]) for k in arrays):
            return
        keep = [np.ascontiguousarray(arrays[k], float) for k in ("q", "radius", "alpha", "lj_rmin_half", "lj_sqrt_eps", "cov")]
        if keep[5].size == 0:
            keep[5] = np.zeros(1)
        _lib.check(_lib.lib().pgm_set_params(self._ctx, *(_ptr(a, C.c_double) for a in keep)), "pgm_set_params")
        self._params_now = {k: np.array(v, float) for k, v in arrays.items()}

    def grad_params(self, pos, H, mu, e_bar: float = 0.0, f_bar=None, mu_bar=None, tol: float | None = None) -> dict:
        """The gradient over the system's parameter table of a loss with dL/dE = e_bar, dL/dF = f_bar,
        dL/dmu = mu_bar at the dipoles mu solved at pos with the parameters in use (what jax.grad of
        such a loss gives with differentiable=True; regression_cases.py:282-287): compute_vjp
        summed per tying key (System.reduce_params).  {quantity: (K,) numpy array}."""
        return self.sys.reduce_params(self.compute_vjp(pos, H, mu, e_bar, f_bar, mu_bar, tol))

    def compute_vjp(self, pos, H, mu, e_bar: float = 0.0, f_bar=None, mu_bar=None, tol: float | None = None) -> dict:
        """Parameter cotangents of compute's outputs (differentiable=True with a parameter pytree,
        forcefield.py:1853-1915, 2489-2560), through pgm_compute_vjp: for a loss with dL/dE = e_bar,
        dL/dF = f_bar (N,3), dL/dmu = mu_bar (N,3) at the dipoles mu solved at pos, returns
        {"q", "radius", "alpha", "lj_rmin_half", "lj_sqrt_eps"} (N,) and "cov" (n_cov,) device float64
        tensors, plus "iterations" and "residual" of the adjoint CG (tol: settings.adjoint_tol by
        default).  System.reduce_params sums them into the tied parameter table's gradient.
        """
        torch = _torch()
        pos = _dev(pos, (self.n, 3))
        mu = _dev(mu, (self.n, 3))
        fb = None if f_bar is None else _dev(f_bar, (self.n, 3))
        mb = None if mu_bar is None else _dev(mu_bar, (self.n, 3))
        self._box(H)
        self._ensure_rows(pos)
        tol = float(self.s.adjoint_tol if tol is None else tol)
        self._efield(None)  # the derivatives are those of the field-free force field
        f64 = dict(dtype=torch.float64, device="cuda")
        keys = ("q", "radius", "alpha", "lj_rmin_half", "lj_sqrt_eps")
        out = {k: torch.empty(self.n, **f64) for k in keys}
        out["cov"] = torch.empty(max(len(self.sys.cov_i), 1), **f64)
        ptr = lambda t: None if t is None else C.c_void_p(t.data_ptr())  # noqa: E731
        it, res = C.c_int(), C.c_double()
        _lib.check(_lib.lib().pgm_compute_vjp(self._ctx, ptr(pos), ptr(mu), float(e_bar), ptr(fb), ptr(mb), tol,
                                              *(ptr(out[k]) for k in keys + ("cov",)), C.byref(it), C.byref(res)),
                   "pgm_compute_vjp")
        out["cov"] = out["cov"][:len(self.sys.cov_i)]
        out["iterations"], out["residual"] = it.value, res.value
        return out

    def pme_energy(self, pos, d, with_charges: bool = True, H=None) -> tuple[float, object]:
        """Return (E_rec [e^2/nm], dE_rec/dd (N,3)) of PME (pgm_jax/md/pme.py:339-411)."""
        torch = _torch()
        pos = _dev(pos, (self.n, 3))
        d = _dev(d, (self.n, 3))
        self._box(H)
        g = torch.empty((self.n, 3), dtype=torch.float64, device="cuda")
        e = C.c_double()
        _lib.check(_lib.lib().pgm_pme_energy(self._ctx, C.c_void_p(pos.data_ptr()), int(with_charges), C.c_void_p(d.data_ptr()),
                                             C.byref(e), C.c_void_p(g.data_ptr())), "pgm_pme_energy")
        return e.value, g

    # ------------------------------------------------------------------ learned CG guess
    def guess_config(self, hidden_layers: int = 3, dipole_scale: float | None = None, harvest_every: int = 10,
                     capacity: int | None = None, seed: int = 0) -> None:
        """Set up the learned initial guess (used with predictor "learned"): pgm_guess_config.

        dipole_scale [e nm] defaults to alpha_O x 100 e/nm^2 (a typical induced dipole); capacity
        defaults to the larger of 200,000 rows and two harvests.
        """
        scale = float(dipole_scale) if dipole_scale else float(np.max(self.sys.alpha) * 100.0)
        cap = int(capacity) if capacity else max(200_000, 2 * self.n)
        _lib.check(_lib.lib().pgm_guess_config(self._ctx, int(hidden_layers), scale, int(harvest_every), cap, int(seed)),
                   "pgm_guess_config")
        self.guess_nparams = int(_lib.lib().pgm_guess_nparams(self._ctx))

    def guess_params(self, params: np.ndarray | None = None) -> np.ndarray:
        """Return (params None) or set the flat FP32 network parameters (per layer W (in x out), b)."""
        buf = np.zeros(self.guess_nparams, np.float32) if params is None else np.ascontiguousarray(params, np.float32)
        _lib.check(_lib.lib().pgm_guess_params(self._ctx, buf.ctypes.data_as(C.c_void_p), buf.size, int(params is not None)),
                   "pgm_guess_params")
        return buf

    def guess_forward(self, x, bf16: bool = True):
        """Network outputs (n, 3) of feature rows x (n, 16): matrix-core BF16 inference or the FP32 forward."""
        torch = _torch()
        x = torch.as_tensor(np.asarray(x, np.float32) if not isinstance(x, torch.Tensor) else x).to("cuda", torch.float32).contiguous()
        y = torch.empty((x.shape[0], 3), dtype=torch.float32, device="cuda")
        _lib.check(_lib.lib().pgm_guess_forward(self._ctx, C.c_void_p(x.data_ptr()), x.shape[0], C.c_void_p(y.data_ptr()),
                                                int(bool(bf16))), "pgm_guess_forward")
        return y

    def guess_grad(self, x, t) -> tuple[np.ndarray, float]:
        """Return (flat gradients, loss) of the mean squared error of rows x (n, 16) against t (n, 3)."""
        torch = _torch()
        x = torch.as_tensor(np.asarray(x, np.float32)).to("cuda").contiguous()
        t = torch.as_tensor(np.asarray(t, np.float32)).to("cuda").contiguous()
        g = np.zeros(self.guess_nparams, np.float32)
        loss = C.c_double()
        _lib.check(_lib.lib().pgm_guess_grad(self._ctx, C.c_void_p(x.data_ptr()), C.c_void_p(t.data_ptr()), x.shape[0],
                                             g.ctypes.data_as(C.c_void_p), C.byref(loss)), "pgm_guess_grad")
        return g, loss.value

    def guess_train(self, steps: int = 200, batch: int = 4096, lr: float = 1e-3) -> float:
        """Run Adam steps on the harvested rows; return the mean loss."""
        loss = C.c_double()
        _lib.check(_lib.lib().pgm_guess_train(self._ctx, int(steps), int(batch), float(lr), C.byref(loss)), "pgm_guess_train")
        return loss.value

    def guess_status(self) -> dict:
        """Rows harvested, sigma, squared errors of mu4 / the linear predictor / linear + network at the last
        harvest, the selector's mode, trained, number of fits, the linear coefficients."""
        out = (C.c_double * 16)()
        _lib.check(_lib.lib().pgm_guess_status(self._ctx, out), "pgm_guess_status")
        modes = ("mu4", "linear", "linear+network")
        return {"rows": int(out[0]), "sigma": out[1], "err_mu4": out[2], "err_linear": out[3], "err_net": out[4],
                "mode": modes[int(out[5])], "trained": bool(out[6]), "fits": int(out[7]),
                "coefficients": [round(v, 6) for v in out[8:16]]}

    def set_timing(self, on: bool = True) -> None:
        """Record HIP events around the CG matvec kernels from now on (pgm_set_timing); clears the sums."""
        _lib.check(_lib.lib().pgm_set_timing(self._ctx, int(bool(on))), "pgm_set_timing")

    def kernel_timing(self) -> dict:
        """Average duration [ms] and launch count of the row field, spread and gather since set_timing."""
        ms = (C.c_double * 3)()
        cnt = (C.c_longlong * 3)()
        _lib.check(_lib.lib().pgm_kernel_timing(self._ctx, ms, cnt), "pgm_kernel_timing")
        return {k: {"ms": ms[i], "launches": cnt[i]} for i, k in enumerate(("row_field", "spread", "gather"))}

    def row_entries(self) -> int:
        """Number of stored row entries within the cutoff at the last evaluation (pairs counted twice)."""
        return int(_lib.lib().pgm_row_entries(self._ctx))

    def row_blocks(self) -> int:
        """Number of molecule-block records (molecule pairs within the cutoff, both directions) at the last evaluation."""
        return int(_lib.lib().pgm_row_blocks(self._ctx))

    def set_row_layout(self, layout: str = "molecule", lanes: int = 4) -> None:
        """Choose the CG row-field layout: "atom" rows or "molecule" blocks (3-site molecules; 4, 8 or 16 lanes)."""
        code = {"atom": 0, "molecule": 1}[layout]
        _lib.check(_lib.lib().pgm_set_row_layout(self._ctx, code, int(lanes)), "pgm_set_row_layout")
        self._built_at = None

    def autotune(self, pos, reps: int = 20) -> dict:
        """Time every row-field layout, PME spread-tile width and reciprocal transform at pos; keep the fastest."""
        pos = _dev(pos, (self.n, 3))
        self._built_at = None
        return _autotune(self._ctx, pos, reps)

    def set_graphs(self, on: bool = True) -> None:
        """Replay every CG iteration as a captured hipGraph (same kernels, same results)."""
        _lib.check(_lib.lib().pgm_set_graphs(self._ctx, int(bool(on))), "pgm_set_graphs")

    def set_pme_transform(self, transform: str = "fused") -> None:
        """Reciprocal transform of the CG field, energy and force evaluations: "fused" kernels or the "library" FFT plans."""
        _set_transform(self._ctx, transform)

    def pme_transform(self) -> tuple[str, bool]:
        """Return the transform in use and whether the fused one is available for this grid."""
        return _transform(self._ctx)

    def set_row_fork(self, stage: str = "start") -> None:
        """Start the CG row field beside the PME chain at "start", "after_spread", "after_forward_planes" or
        "start_join_before_planes" (the forward planes wait for it), or run it after the spread on the
        context's own stream ("none": no side stream, no fork and join events)."""
        code = FORKS.index(stage)
        _lib.check(_lib.lib().pgm_set_row_fork(self._ctx, code), "pgm_set_row_fork")

    def row_layout(self) -> tuple[str, int]:
        """Return the CG row-field layout and its lanes per row."""
        lay, lanes = C.c_int(), C.c_int()
        _lib.check(_lib.lib().pgm_row_layout(self._ctx, C.byref(lay), C.byref(lanes)), "pgm_row_layout")
        return ("atom", "molecule")[lay.value], lanes.value
