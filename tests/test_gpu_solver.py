"""Dense induction solve on the GPU (solver.solve_linear_induction) against numpy's LU solve (GPU)."""

import numpy as np
import pytest

from oracle import systems
from oracle.kernels import erf_kernels

pytestmark = pytest.mark.gpu


def _cluster_tensors(nw=40, seed=2):
    """Gaussian-screened dipole tensors T_ij = B1 I - B2 x x^T (a_ij = 1/sqrt(2(Ri^2 + Rj^2))) of a water
    cluster (zero diagonal blocks), the polarizabilities, and a random field."""
    top, pos, H = systems.water_box(nw, 1.2, seed=seed)
    n = len(pos)
    i, j = np.triu_indices(n, k=1)
    x = pos[i] - pos[j]
    r = np.linalg.norm(x, axis=-1)
    a = 1.0 / np.sqrt(2.0 * (top.radius[i] ** 2 + top.radius[j] ** 2))
    B = erf_kernels(a, r, 3)
    blk = B[1][:, None, None] * np.eye(3)[None] - B[2][:, None, None] * x[:, :, None] * x[:, None, :]
    T = np.zeros((n, n, 3, 3))
    T[i, j] = blk
    T[j, i] = blk
    F = np.random.default_rng(seed).normal(scale=50.0, size=(n, 3))
    return T, top.alpha, F


def test_dense_solve_matches_lu():
    import pgm_jax_amd as P

    T, alpha, F = _cluster_tensors()
    n = len(alpha)
    print("dense solve n =", n, flush=True)
    A = T.transpose(0, 2, 1, 3).reshape(3 * n, 3 * n) + np.diag(np.repeat(1.0 / alpha, 3))
    ref = np.linalg.solve(A, F.reshape(-1)).reshape(n, 3)
    mu = P.solver.solve_linear_induction(T, alpha, F).cpu().numpy()
    assert np.abs(mu - ref).max() < 1e-11 * np.abs(ref).max()
    # the functional's gradient vanishes at the solution
    g = (A @ mu.reshape(-1) - F.reshape(-1))
    assert np.linalg.norm(g) < 1e-12 * np.linalg.norm(F)
