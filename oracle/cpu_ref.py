"""ctypes wrapper of oracle/libcpu_ref.so (cpu_ref.cpp): TEST INFRASTRUCTURE AND CPU BASELINE ONLY.

The C++/OpenMP float64 restatement of the force field, CG solve and rigid-water step of this
oracle, on all host cores (or `threads`).  Used by the CPU tests (against the numpy oracle) and by
bench.py's cpu_baseline leg; never by the product.
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .forcefield import Settings
from .pme import grid_size

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libcpu_ref.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise OSError(f"{LIB} not built (python oracle/build_cpu_ref.py)")
        L = C.CDLL(LIB)
        P, I, D = C.c_void_p, C.c_int, C.c_double
        L.cref_create.restype = P
        L.cref_create.argtypes = [I, P, P, P, P, P, P, I, P, P, P, P, D, D, D, P, I, I, D, I, D, I]
        L.cref_destroy.argtypes = [P]
        L.cref_threads.argtypes = [P]
        L.cref_compute.argtypes = [P, P, P, P, P, C.POINTER(I), C.POINTER(D)]
        L.cref_md.argtypes = [P, P, P, P, P, P, I, D, P, P]
        L.cref_timings.argtypes = [P, I]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class CpuRef:
    """Force field of `top` at box H with oracle Settings (mu4 predictor state inside)."""

    def __init__(self, top, H, s: Settings = Settings(), threads: int = 0) -> None:
        self._keep = [np.ascontiguousarray(v, float) for v in (top.q, top.radius, top.alpha, top.lj_rmin_half, top.lj_sqrt_eps)]
        self._keep += [np.ascontiguousarray(top.mol, np.int32), np.ascontiguousarray(top.cov_i, np.int32),
                       np.ascontiguousarray(top.cov_j, np.int32), np.ascontiguousarray(top.cov_c, float)]
        self.H = np.ascontiguousarray(H, float)
        self.K = np.array(s.pme_grid or grid_size(self.H, s.pme_spacing), np.int32)
        q, R, a, rh, se, mol, ci, cj, cc = self._keep
        self.n = len(q)
        elec = 0.0 if getattr(s, "elec_cutoff", None) is None else float(s.elec_cutoff)
        self.h = lib().cref_create(self.n, _p(q), _p(R), _p(a), _p(rh), _p(se), _p(mol), len(ci), _p(ci), _p(cj), _p(cc),
                                   _p(self.H), float(s.cutoff), elec, float(s.ewald_beta), _p(self.K), int(s.pme_order),
                                   int(bool(s.lj_lrc)), float(s.dipole_tol), int(s.max_iter), float(s.peek), int(threads))
        self.mu = np.zeros((self.n, 3))

    def __del__(self) -> None:
        if getattr(self, "h", None):
            lib().cref_destroy(self.h)
            self.h = None

    @staticmethod
    def timings(reset: bool = False) -> dict:
        """Seconds spent so far per section (all contexts): rows, spread+fft, gathers, row_field."""
        t = np.zeros(6)
        lib().cref_timings(_p(t), int(reset))
        return {"rows": t[0], "spread_fft": t[1], "gathers": t[2], "row_field": t[3]}

    @property
    def threads(self) -> int:
        return int(lib().cref_threads(self.h))

    def compute(self, pos) -> dict:
        """One force evaluation (mu4 predictor from this object's history): energies, forces, mu, CG."""
        pos = np.ascontiguousarray(pos, float)
        en = np.zeros(2)
        F = np.zeros((self.n, 3))
        it, res = C.c_int(), C.c_double()
        lib().cref_compute(self.h, _p(pos), _p(self.mu), _p(en), _p(F), C.byref(it), C.byref(res))
        return {"energy": {"elec": en[0], "vdw": en[1], "total": en[0] + en[1]}, "forces": F, "mu": self.mu.copy(),
                "iterations": it.value, "residual": res.value}

    def md(self, pos, vel, forces, masses, d2, nsteps: int, dt: float) -> dict:
        """nsteps NVE steps (in place on pos, vel, forces: float64 (N,3) C-contiguous arrays)."""
        rep = np.zeros(4)
        lib().cref_md(self.h, _p(pos), _p(vel), _p(forces), _p(np.ascontiguousarray(masses, float)),
                      _p(np.ascontiguousarray(d2, float)), int(nsteps), float(dt), _p(self.mu), _p(rep))
        return {"epot": rep[0], "ekin": rep[1], "cg_iterations": int(rep[2]), "cg_residual_max": rep[3]}
