"""The oracle's exact Ewald model against the reference's golden periodic_model.npz (CPU).

small_box(1), cutoff 0.6 nm, beta 3.8 /nm, k_tol 1e-12, induced dipoles to 1e-12, LJ with the
long-range correction (reference tests/regression/regression_cases.py:231-252).
"""

import os

import numpy as np

from oracle import systems
from oracle.ewald import EwaldPGM
from oracle.pairs import min_image

GOLD = os.path.join(os.path.dirname(__file__), "golden", "periodic_model.npz")


def test_periodic_model_matches_golden():
    top, pos, H = systems.small_box(1)
    ew = EwaldPGM(top, H, ewald_beta=3.8, cutoff=0.6, lj_lrc=True)
    e = ew.energy(pos)
    with np.load(GOLD) as g:
        for k in ("perm", "ind", "elec", "vdw", "total"):
            assert abs(e[k] - float(g[f"E.{k}"])) < 1e-11 * abs(float(g[f"E.{k}"])), k
        mu = ew.induced_dipoles(pos)
        assert np.abs(mu - g["mu"]).max() < 1e-10 * np.abs(g["mu"]).max()
        F = ew.forces(pos)
        assert np.abs(F - g["forces"]).max() < 1e-10 * np.abs(g["forces"]).max()


def test_min_image_is_the_nearest_of_27_images():
    """Sequential reduction = minimum over the 27 images in a truncated octahedron (test_md.py:137-157)."""
    rng = np.random.default_rng(3)
    H = systems.reduce_box(np.array([[2.72, 0, 0], [-0.9067, 2.5645, 0], [-0.9067, -1.2823, 2.2210]]))
    pos = rng.uniform(size=(200, 3)) @ H
    d = min_image(pos[:, None] - pos[None], H)
    r = np.linalg.norm(d, axis=-1)
    imgs = np.array([[a, b, c] for a in (-1, 0, 1) for b in (-1, 0, 1) for c in (-1, 0, 1)]) @ H
    r27 = np.min(np.linalg.norm((pos[:, None, None] - pos[None, :, None]) + imgs[None, None], axis=-1), axis=-1)
    assert np.allclose(np.minimum(r, 1.0), np.minimum(r27, 1.0))


def test_periodic_model_strain_and_pressure_match_golden():
    """Molecular and atomic strain derivatives (full 3x3) and the pressure vs periodic_model.npz."""
    top, pos, H = systems.small_box(1)
    ew = EwaldPGM(top, H, ewald_beta=3.8, cutoff=0.6, lj_lrc=True)
    with np.load(GOLD) as g:
        for mol, key in ((True, "strain_mol"), (False, "strain_atom")):
            W = ew.strain_derivative(pos, molecular=mol)
            assert np.abs(W - g[key]).max() < 1e-10 * np.abs(g[key]).max(), key
        assert abs(ew.pressure(pos) - float(g["pressure"])) < 1e-10 * abs(float(g["pressure"]))
