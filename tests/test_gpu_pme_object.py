"""The PME object mirror (pgm_jax/md/pme.py:209-411) on the GPU against the oracle's PME (GPU).

PME(grid, order, beta, dtype).influence / setup / spread / grad_dipoles / energy through
pgm_pme_influence, pgm_pme_spread and pgm_pme_energy, on the skewed toy box (138 atoms) with its
charges and seeded dipoles of 0.02 e nm: a 32^3 grid of order 6, a 12^3 grid of order 4 and a
24 x 20 x 18 grid of odd order 5 (axes that are not powers of two: the library transform);
charges switched between calls and a call without charges.  FP64: influence to 1e-13 of
its largest value, Q to 1e-12, E_rec to 1e-11 relative, dU/dd to 1e-10 of its largest entry;
float32 (FP32 kernels, FP64 sums): Q to 1e-6, E_rec to 1e-6, dU/dd to 1e-5 (influence is FP64
on the host in both).
"""

import numpy as np
import pytest

from oracle import pme as opme
from oracle import systems

pytestmark = pytest.mark.gpu

CASES = [((32, 32, 32), 6, 6.0), ((12, 12, 12), 4, 3.5), ((24, 20, 18), 5, 4.0)]


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.parametrize("dtype,tol_q,tol_e,tol_g", [(np.float64, 1e-12, 1e-11, 1e-10), (np.float32, 1e-6, 1e-6, 1e-5)])
@pytest.mark.parametrize("grid,order,beta", CASES)
def test_pme_object_against_oracle(grid, order, beta, dtype, tol_q, tol_e, tol_g):
    import pgm_jax_amd as P

    top, pos, H = systems.small_box(0)
    rng = np.random.default_rng(7)
    d = 0.02 * rng.standard_normal((len(top.q), 3))
    print(f"PME {grid} order {order} {np.dtype(dtype).name}", flush=True)
    ref = opme.PME(grid, order, beta)
    Gr = ref.influence(H)
    Sr = ref.setup(pos, H)
    pme = P.PME(grid, order, beta, dtype)
    G = pme.influence(H)
    assert G.shape == Gr.shape and _rel(G, Gr) <= 1e-13
    S = pme.setup(pos, H)
    for q in (top.q, 0.5 * top.q, np.zeros_like(top.q)):
        Q = pme.spread(S, q, d).cpu().numpy()
        assert Q.shape == tuple(grid) and _rel(Q, ref.spread(Sr, q, d)) <= tol_q
        e, er = pme.energy(S, G, q, d), ref.energy(Sr, Gr, q, d)
        assert abs(e - er) <= tol_e * abs(er), (e, er)
        g = pme.grad_dipoles(S, G, q, d).cpu().numpy()
        assert _rel(g, ref.gradients(Sr, Gr, q, d)[0]) <= tol_g
