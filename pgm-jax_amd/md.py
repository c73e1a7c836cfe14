"""Rigid-water MD on the GPU: NVE velocity Verlet or Langevin g-BAOAB with SHAKE/RATTLE in FP64.

Mirrors the MD engine around pgm_jax/md/forcefield.py's compute (pgm_jax/md/integrate.py:743-761,
the constrained step of md/flexible.py:776-799 with md/constraints.py:254-299, the Langevin O step
of md/thermostats.py:195-242): the whole step runs in the HIP library (pgm_md_run), the host only
launches blocks of steps and reads the report.
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .forcefield import PGMForceField, _dev, _torch


class Langevin:
    """Langevin thermostat (pgm_jax/md/thermostats.py:165-242): friction gamma [1/ps].

    Raises ValueError for a negative friction, as the reference does.
    """

    name = "langevin"

    def __init__(self, friction: float = 1.0) -> None:
        if float(friction) < 0.0:
            raise ValueError(f"Langevin: friction must be >= 0 ({friction!r} 1/ps)")
        self.friction = float(friction)

    def describe(self) -> str:
        return f"Langevin {self.friction:g}/ps"


class RigidWaterMD:
    """MD of consecutive 3-site rigid molecules (constraint lengths from the initial geometry).

    NVE without a thermostat; with `thermostat=Langevin(gamma)` and `temperature` [K], the g-BAOAB
    step (noise from Philox4x32-10 keyed by `seed`).  efield: None or (E [V/nm], offset [e nm] or
    None), a uniform static field on every step (the constant-field kind of md/efield.py).
    """

    def __init__(self, ff: PGMForceField, pos, vel, masses=None, thermostat: Langevin | None = None,
                 temperature: float = 298.0, seed: int = 0, efield=None) -> None:
        torch = _torch()
        self.ff = ff
        self.efield = efield
        n = ff.n
        self.pos = _dev(pos, (n, 3)).clone()
        self.vel = _dev(vel, (n, 3)).clone()
        self.forces = torch.zeros((n, 3), dtype=torch.float64, device="cuda")
        self.ind = ff.init_induction()
        m = np.ascontiguousarray(ff.sys.masses if masses is None else masses, float)
        L = _lib.lib()
        _lib.check(L.pgm_md_setup(ff._ctx, m.ctypes.data_as(C.POINTER(C.c_double)), C.c_void_p(self.pos.data_ptr())),
                   "pgm_md_setup")
        ff.build_rows(self.pos)
        self._state = _lib.MDState(self.pos.data_ptr(), self.vel.data_ptr(), self.forces.data_ptr(),
                                   _lib.Induction(self.ind.mu.data_ptr(), self.ind.hist.data_ptr(), 0, 1.0), 0, 0)
        self.cg_total = 0
        self.steps = 0
        self.heat = 0.0
        self.set_thermostat(thermostat, temperature, seed)

    def set_thermostat(self, thermostat: Langevin | None, temperature: float = 298.0, seed: int = 0) -> None:
        """Use `thermostat` (None: NVE) at `temperature` [K] for the following steps."""
        gamma = 0.0 if thermostat is None else thermostat.friction
        _lib.check(_lib.lib().pgm_md_set_thermostat(self.ff._ctx, gamma, float(temperature), int(seed) & (2**64 - 1)),
                   "pgm_md_set_thermostat")
        self.thermostat, self.temperature = thermostat, float(temperature)

    def run(self, nsteps: int, dt_fs: float = 1.0) -> dict:
        """Advance nsteps of dt_fs femtoseconds; return the report of the last step (energies in kJ/mol)."""
        rep = _lib.MDReport()
        self.ff._efield(self.efield)
        st = _lib.check(_lib.lib().pgm_md_run(self.ff._ctx, C.byref(self._state), int(nsteps), dt_fs * 1e-3, C.byref(rep)),
                        "pgm_md_run")
        self.cg_total += rep.cg_iterations_total
        self.steps += nsteps
        self.ind.count = self._state.ind.count
        self.ind.norm = self._state.ind.norm
        self.heat += rep.heat
        return {
            "epot": rep.epot, "elec": rep.elec, "vdw": rep.vdw, "ekin": rep.ekin, "etot": rep.etot,
            "temperature": rep.temperature, "cg_iterations": rep.cg_iterations_total,
            "cg_iterations_max": rep.cg_iterations_max, "cg_residual_max": rep.cg_residual_max,
            "row_rebuilds": rep.row_rebuilds, "not_converged": st == _lib.PGM_NOT_CONVERGED, "heat": rep.heat,
            "cg_iter_last": rep.cg_iterations_last,
        }

    def invalidate_forces(self) -> None:
        """Mark the stored forces stale (positions or box changed outside run): the next run recomputes them first."""
        self._state.have_forces = 0

    def equilibrate(self, nsteps: int, temperature: float, dt_fs: float = 1.0, every: int = 10) -> dict:
        """Run nsteps with the velocities rescaled to `temperature` [K] every `every` steps; return the last report."""
        rep = None
        done = 0
        while done < nsteps:
            k = min(every, nsteps - done)
            rep = self.run(k, dt_fs)
            done += k
            if rep["temperature"] > 0:
                _lib.check(_lib.lib().pgm_md_rescale(self.ff._ctx, C.byref(self._state),
                                                     (temperature / rep["temperature"]) ** 0.5), "pgm_md_rescale")
        return rep

    def time_kernels(self, reps: int = 20) -> dict:
        """Average duration [ms] of the hot kernels (pgm_time_kernels) at the current positions.

        With the fused reciprocal transform the block merge is part of the forward planes
        ("fwd_planes": merge, z and y transforms; "x_conv": x transform, influence function and
        inverse; "bwd_planes": inverse y and z transforms)."""
        ms = (C.c_double * 8)()
        _lib.check(_lib.lib().pgm_time_kernels(self.ff._ctx, C.c_void_p(self.pos.data_ptr()), int(reps), ms), "pgm_time_kernels")
        if self.ff.pme_transform()[0] == "fused":
            names = ("merge", "spread", "fwd_planes", "x_conv", "bwd_planes", "row_field", "gather", "matvec")
            return {k: v for k, v in zip(names, list(ms)) if k != "merge"}
        names = ("merge", "spread", "fft_r2c", "convolve", "fft_c2r", "row_field", "gather", "matvec")
        return dict(zip(names, list(ms)))
