"""The oracle's smooth PME against invariants, finite differences and the exact Ewald sum (CPU)."""

import numpy as np
import pytest

from oracle import systems
from oracle.ewald import EwaldPGM
from oracle.forcefield import ForceField, Settings
from oracle.pme import PME, bspline, bspline_moduli


@pytest.mark.parametrize("order", [4, 6, 8])
def test_bspline_partition_of_unity_and_derivatives(order):
    """sum theta = 1, sum dtheta = 0, sum d2theta = 0; dtheta, d2theta match finite differences."""
    f = np.linspace(0.01, 0.99, 17)
    th, dth, d2th = bspline(f, order)
    assert np.allclose(th.sum(-1), 1.0, atol=1e-14)
    assert np.allclose(dth.sum(-1), 0.0, atol=1e-13)
    assert np.allclose(d2th.sum(-1), 0.0, atol=1e-12)
    h = 1e-6
    tp, dp, _ = bspline(f + h, order)
    tm, dm, _ = bspline(f - h, order)
    assert np.allclose((tp - tm) / (2 * h), dth, atol=1e-8)
    assert np.allclose((dp - dm) / (2 * h), d2th, atol=1e-7)


def test_moduli_are_the_dft_of_the_spline_at_integers():
    """|b(m)|^-2 equals |FFT of M_p(1..p)|^2 (numpy FFT, even order: no zeros)."""
    K, p = 24, 6
    w = np.zeros(K)
    w[:p] = bspline(np.zeros(()), p)[0]
    ref = np.abs(np.fft.fft(w)) ** 2
    assert np.allclose(bspline_moduli(K, p), ref, rtol=1e-12)


def test_pme_energy_and_dipoles_match_exact_ewald():
    """Fine-grid PME (48^3, order 8) reproduces the exact Ewald energy to 2e-6 (reference test_md.py:39-48)."""
    top, pos, H = systems.small_box(0)
    s = Settings(cutoff=0.6, ewald_beta=6.0, pme_grid=(48, 48, 48), pme_order=8, lj_lrc=False,
                 dipole_tol=1e-12, max_iter=500, peek=0.0)
    res = ForceField(top, H, s).compute(pos, H)
    ew = EwaldPGM(top, H, ewald_beta=6.0, cutoff=0.6)
    e = ew.energy(pos)
    assert abs(res.energy["elec"] - e["elec"]) < 2e-6 * abs(e["elec"])
    mu = ew.induced_dipoles(pos)
    assert np.allclose(res.mu, mu, atol=1e-6 * np.abs(mu).max())


def test_reciprocal_gradients_match_finite_differences():
    """dE_rec/dd and dE_rec/dr of `gradients` are the derivatives of `energy`."""
    top, pos, H = systems.small_box(2)
    pme = PME((20, 20, 20), 6, 4.0)
    rng = np.random.default_rng(0)
    d = rng.normal(scale=0.01, size=pos.shape)
    G = pme.influence(H)
    S = pme.setup(pos, H)
    gd, gr = pme.gradients(S, G, top.q, d)
    h = 1e-6
    for a, c in [(0, 0), (7, 1), (50, 2), (113, 0)]:
        dp, dm = d.copy(), d.copy()
        dp[a, c] += h
        dm[a, c] -= h
        fd = (pme.energy(S, G, top.q, dp) - pme.energy(S, G, top.q, dm)) / (2 * h)
        assert abs(fd - gd[a, c]) < 1e-6 * max(1.0, abs(fd))
        xp, xm = pos.copy(), pos.copy()
        xp[a, c] += h
        xm[a, c] -= h
        fd = (pme.energy(pme.setup(xp, H), G, top.q, d) - pme.energy(pme.setup(xm, H), G, top.q, d)) / (2 * h)
        assert abs(fd - gr[a, c]) < 1e-5 * max(1.0, abs(fd))
