"""The oracle's test systems: pGM3P-25 constants against the prmtop values in the fixture (CPU)."""

import os

import numpy as np

from oracle import systems

GOLD = os.path.join(os.path.dirname(__file__), "golden", "water512.npz")


def test_pgm3p25_constants_match_the_prmtop():
    w = systems.pgm3p25_water()
    with np.load(GOLD) as z:
        for f in ("q", "radius", "alpha", "lj_rmin_half", "lj_sqrt_eps", "masses"):
            assert np.allclose(getattr(w, f), z[f], rtol=1e-9, atol=0), f
        cov = [(int(i), int(j), float(c)) for i, j, c in zip(z["cov_i"], z["cov_j"], z["cov_c"])]
        assert np.allclose(np.array(w.cov), np.array(cov), rtol=1e-9)
        assert np.allclose(z["all_q"], np.tile(z["q"], len(z["all_q"]) // 3))


def test_water_box_geometry():
    """Rigid waters: O-H 0.09572 nm, H-O-H 104.52 deg, whole molecules, density ~1 g/cm^3 at 9.93 nm."""
    top, pos, H = systems.water_box(64, 1.241, seed=1)
    x = pos.reshape(-1, 3, 3)
    oh = np.linalg.norm(x[:, 1:] - x[:, :1], axis=-1)
    assert np.allclose(oh, 0.09572, atol=1e-12)
    c = np.einsum("na,na->n", x[:, 1] - x[:, 0], x[:, 2] - x[:, 0]) / oh[:, 0] / oh[:, 1]
    assert np.allclose(np.degrees(np.arccos(c)), 104.52, atol=1e-9)
    assert top.n == 192 and np.allclose(H, np.eye(3) * 1.241)


def test_product_builder_matches_oracle_builder():
    """The product's synthetic water box equals the oracle's (same seed use, same parameters)."""
    import pgm_jax_amd as P

    s1, p1, H1 = P.models.water_box(64, 1.241, seed=1)
    t2, p2, H2 = systems.water_box(64, 1.241, seed=1)
    assert np.allclose(p1, p2, rtol=0, atol=1e-14) and np.array_equal(H1, H2)
    for f in ("q", "radius", "alpha", "lj_rmin_half", "lj_sqrt_eps", "masses"):
        assert np.allclose(getattr(s1, f), getattr(t2, f), rtol=1e-15)
    assert np.array_equal(s1.cov_i, t2.cov_i) and np.array_equal(s1.cov_j, t2.cov_j) and np.allclose(s1.cov_c, t2.cov_c)
