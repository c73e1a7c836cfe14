"""Induction solves of pgm_jax/solver.py on the MI355X.

`solve_linear_induction` (solver.py:29-62) solves (T + diag(1/alpha)) mu = F for dense
dipole-dipole tensors T (n, n, 3, 3) with the library's FP64 CG (pgm_solve_linear_induction; the
reference uses a dense LU solve, and the CG stops at |r| <= tol |F| with tol 1e-13 by default, so
both agree to rounding).  `variational` (solver.py:126-167) wraps an energy functional as
E(theta) = G(solve(theta), theta); the reference's custom JVP (first derivatives at fixed x*) is a
JAX transformation, which the HIP path replaces by the analytic Hellmann-Feynman forces and strain
derivatives of PGMForceField / PeriodicModel.
"""

from __future__ import annotations

import ctypes as C

from . import _lib
from .forcefield import _dev, _torch


def solve_linear_induction(T, alpha, F, tol: float = 1e-13, max_iter: int = 10000):
    """Return the induced dipoles (n, 3) [e nm] minimising sum |mu|^2/(2 alpha) - mu.F + 1/2 mu T mu.

    T (n, n, 3, 3) [1/nm^3] with zero diagonal blocks, alpha (n,) [nm^3] nonzero, F (n, 3) [e/nm^2];
    arrays (numpy or torch) are moved to the GPU as float64.  Raises PGMError when the CG does not
    reach tol within max_iter iterations.
    """
    torch = _torch()
    a = _dev(alpha)
    n = a.shape[0]
    Tt = _dev(T, (n, n, 3, 3))
    Ft = _dev(F, (n, 3))
    if not bool(torch.all(a != 0).item()):
        raise ValueError("solve_linear_induction: every polarizability must be nonzero")
    mu = torch.empty((n, 3), dtype=torch.float64, device="cuda")
    it, res = C.c_int(), C.c_double()
    st = _lib.check(_lib.lib().pgm_solve_linear_induction(n, C.c_void_p(Tt.data_ptr()), C.c_void_p(a.data_ptr()),
                                                          C.c_void_p(Ft.data_ptr()), C.c_void_p(mu.data_ptr()), float(tol),
                                                          int(max_iter), C.c_void_p(torch.cuda.current_stream().cuda_stream),
                                                          C.byref(it), C.byref(res)), "pgm_solve_linear_induction")
    if st == _lib.PGM_NOT_CONVERGED:
        raise _lib.PGMError(f"solve_linear_induction: CG stopped at {it.value} iterations, |r|/|F| = {res.value:.3e}")
    return mu


def variational(G, solve):
    """Return E(theta) = G(solve(theta), theta) (solver.py:126-167), without the JAX derivative rule."""

    def E(theta):
        return G(solve(theta), theta)

    return E


__all__ = ["solve_linear_induction", "variational"]
