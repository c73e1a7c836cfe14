"""The dense induction solve without allocations against the allocating one, through the C ABI (GPU).

pgm_solve_linear_induction_noalloc runs the kernels of pgm_solve_linear_induction on a fixed
device array instead of per-call buffers, so on the same inputs it must return the same dipoles,
iteration count and residual bit for bit.  Two calls in a row, the second with fewer atoms (the
array keeps the first call's contents), and a size beyond the array's limit is refused.
"""

import ctypes as C

import numpy as np
import pytest

from oracle import systems
from oracle.kernels import erf_kernels

pytestmark = pytest.mark.gpu


def _cluster(nw, seed):
    """Screened dipole tensors T (n, n, 3, 3) of a water cluster, polarizabilities and a random field."""
    top, pos, H = systems.water_box(nw, 1.2 * (nw / 40) ** (1 / 3), seed=seed)
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
    return T, np.asarray(top.alpha, float), F


def _solve(L, name, T, alpha, F, torch):
    n = len(alpha)
    Tt, at, Ft = (torch.as_tensor(v, device="cuda").contiguous() for v in (T, alpha, F))
    mu = torch.full((n, 3), np.nan, dtype=torch.float64, device="cuda")
    it, res = C.c_int(), C.c_double()
    st = getattr(L, name)(n, C.c_void_p(Tt.data_ptr()), C.c_void_p(at.data_ptr()), C.c_void_p(Ft.data_ptr()),
                          C.c_void_p(mu.data_ptr()), 1e-13, 10000, C.c_void_p(torch.cuda.current_stream().cuda_stream),
                          C.byref(it), C.byref(res))
    return st, mu.cpu().numpy(), it.value, res.value


def test_noalloc_solve_equals_allocating_solve():
    import pgm_jax_amd as P
    import torch

    L = P._lib.lib()
    for nw, seed in ((60, 4), (25, 5)):  # the second call is smaller than the first
        T, alpha, F = _cluster(nw, seed)
        print(f"no-alloc dense solve: n = {len(alpha)} atoms", flush=True)
        sa, mua, ita, ra = _solve(L, "pgm_solve_linear_induction", T, alpha, F, torch)
        sb, mub, itb, rb = _solve(L, "pgm_solve_linear_induction_noalloc", T, alpha, F, torch)
        print(f"  allocating: status {sa}, {ita} it, residual {ra:.3e}; no-alloc: status {sb}, {itb} it, residual {rb:.3e}",
              flush=True)
        assert sa == 0 and sb == 0
        assert ita == itb and ra == rb
        assert np.array_equal(mua, mub)
        # and the solution solves the system (the allocating entry point's own check)
        n = len(alpha)
        A = T.transpose(0, 2, 1, 3).reshape(3 * n, 3 * n) + np.diag(np.repeat(1.0 / alpha, 3))
        assert np.linalg.norm(A @ mub.reshape(-1) - F.reshape(-1)) < 1e-12 * np.linalg.norm(F)


def test_noalloc_solve_refuses_sizes_beyond_its_array():
    import pgm_jax_amd as P
    import torch

    L = P._lib.lib()
    n = 16385  # PGM_DENSE_NOALLOC_MAX_ATOMS + 1; rejected before any pointer is read
    one = torch.zeros(3, dtype=torch.float64, device="cuda")
    p = C.c_void_p(one.data_ptr())
    st = L.pgm_solve_linear_induction_noalloc(n, p, p, p, p, 1e-10, 10, C.c_void_p(torch.cuda.current_stream().cuda_stream),
                                              None, None)
    assert st < 0
    assert b"scratch limit" in L.pgm_last_error()
