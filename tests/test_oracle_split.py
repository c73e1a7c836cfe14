"""The oracle's split cutoffs (electrostatics at elec_cutoff, LJ at cutoff) (CPU).

Pins (pgm_jax/md/forcefield.py:44-58, :1197-1238, :2355):
* elec_cutoff = cutoff is the single-cutoff force field, exactly;
* with elec_cutoff < cutoff the electrostatics are those of a single cutoff at elec_cutoff (same
  beta and grid: identical dipoles, CG iterations and electrostatic energy), and the van der Waals
  energy and forces are that run's plus the LJ pairs between the two radii, summed here by brute
  force over the minimum images, plus the tail difference -2 pi/(3V) sum 2 eps rmin^6 (1/rc^3 - 1/rc_e^3);
* forces are minus the gradient of the split energy (central differences).
"""

import numpy as np

from oracle import systems
from oracle.forcefield import ForceField, Settings, lj_long_range
from oracle.pairs import min_image

BASE = dict(ewald_beta=6.0, pme_grid=(32, 32, 32), pme_order=6, lj_lrc=True, dipole_tol=1e-11, max_iter=500,
            predictor="none")


def _box():
    return systems.small_box(0)


def test_equal_cutoffs_are_the_single_cutoff():
    top, pos, H = _box()
    a = ForceField(top, H, Settings(cutoff=0.6, **BASE)).compute(pos, H)
    b = ForceField(top, H, Settings(cutoff=0.6, elec_cutoff=0.6, **BASE)).compute(pos, H)
    assert a.energy == b.energy and np.array_equal(a.forces, b.forces) and np.array_equal(a.mu, b.mu)


def _lj_shell(top, pos, H, r0, r1):
    """Brute-force LJ energy and forces of intermolecular pairs with r0 <= r < r1 (minimum image)."""
    n = len(pos)
    E, F = 0.0, np.zeros_like(pos)
    for i in range(n - 1):
        j = np.arange(i + 1, n)
        x = min_image(pos[i] - pos[j], H)
        r2 = np.sum(x * x, -1)
        m = (r2 >= r0 * r0) & (r2 < r1 * r1) & (top.mol[j] != top.mol[i])
        eps = top.lj_sqrt_eps[i] * top.lj_sqrt_eps[j]
        rmin = top.lj_rmin_half[i] + top.lj_rmin_half[j]
        s6 = (rmin * rmin / r2) ** 3
        e = np.where(m, eps * (s6 * s6 - 2 * s6), 0.0)
        dl = np.where(m, eps * 12.0 * (s6 - s6 * s6) / r2, 0.0)  # de/dr / r
        E += e.sum()
        f = -dl[:, None] * x  # force on i from each j
        F[i] += f.sum(0)
        np.subtract.at(F, j, f)
    return E, F


def test_split_is_single_cutoff_plus_lj_shell_and_tail():
    top, pos, H = _box()
    rc_e, rc_v = 0.5, 0.6
    split = ForceField(top, H, Settings(cutoff=rc_v, elec_cutoff=rc_e, **BASE)).compute(pos, H)
    single = ForceField(top, H, Settings(cutoff=rc_e, **BASE)).compute(pos, H)
    assert split.iterations == single.iterations
    assert np.abs(split.mu - single.mu).max() < 1e-13 * np.abs(single.mu).max()
    assert abs(split.energy["elec"] - single.energy["elec"]) < 1e-12 * abs(single.energy["elec"])
    e_sh, f_sh = _lj_shell(top, pos, H, rc_e, rc_v)
    V = abs(np.linalg.det(H))
    tail = lj_long_range(top.lj_rmin_half, top.lj_sqrt_eps, V, rc_v) - lj_long_range(top.lj_rmin_half, top.lj_sqrt_eps, V, rc_e)
    assert abs(e_sh) > 1e-3  # the shell holds LJ pairs
    assert abs(split.energy["vdw"] - (single.energy["vdw"] + e_sh + tail)) < 1e-10 * abs(single.energy["vdw"])
    assert np.abs(split.forces - (single.forces + f_sh)).max() < 1e-9 * np.abs(single.forces).max()


def test_split_forces_are_minus_the_energy_gradient():
    top, pos, H = _box()
    ff = ForceField(top, H, Settings(cutoff=0.6, elec_cutoff=0.45, **dict(BASE, ewald_beta=7.5)))
    F = ff.compute(pos, H).forces
    rng = np.random.default_rng(1)
    h = 1e-5
    for i in rng.choice(len(pos), 4, replace=False):
        for a in range(3):
            p1, p2 = pos.copy(), pos.copy()
            p1[i, a] += h
            p2[i, a] -= h
            fd = -(ff.compute(p1, H).energy["total"] - ff.compute(p2, H).energy["total"]) / (2 * h)
            assert abs(fd - F[i, a]) < 2e-5 * np.abs(F).max() + 1e-6, (i, a, fd, F[i, a])
