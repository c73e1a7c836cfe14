"""The fused reciprocal transform of the CG field evaluations against the library FFT plans (GPU).

Both compute phi = c2r(Gc r2c(Q)) (pme.py:373-376) from the same merged spread blocks; the fused
kernels transform in LDS (radix-16/8/4/2 Stockham stages, FP64-tabulated twiddles), the library plans
with rocFFT.  Fields agree to rounding: 1e-12 relative RMS in FP64, 2e-6 in FP32; full evaluations
take the same CG iterations (FP64) and energies agree to 1e-11 / 1e-7 relative (the fused
leg sums the reciprocal energy in the x-transform kernel, the library leg in the convolution kernel).  Grids of unequal
axes check the axis bookkeeping; grids with an axis that is not a power of two keep the library
plans and refuse the fused transform.
"""

import numpy as np
import pytest

from oracle import systems

pytestmark = pytest.mark.gpu


def _system(top):
    import pgm_jax_amd as P

    return P.System(q=top.q, radius=top.radius, alpha=top.alpha, lj_rmin_half=top.lj_rmin_half,
                    lj_sqrt_eps=top.lj_sqrt_eps, masses=top.masses, mol=top.mol, cov_i=top.cov_i,
                    cov_j=top.cov_j, cov_c=top.cov_c)


def _rel_rms(a, b):
    return float(np.sqrt(np.mean((a - b) ** 2)) / np.sqrt(np.mean(b**2)))


@pytest.mark.parametrize("precision", ["double", "mixed"])
@pytest.mark.parametrize("grid", [(32, 32, 32), (32, 64, 128), (128, 32, 64), (32, 32, 256), (256, 32, 32)])
def test_fused_transform_matches_library(precision, grid):
    import pgm_jax_amd as P

    top, pos, H = systems.water_box(512, 2.483, seed=3)
    print(f"fused transform precision={precision} grid={grid}", flush=True)
    ff = P.PGMForceField(_system(top), H, P.MDSettings(precision=precision, cutoff=0.9, ewald_beta=4.0, pme_grid=grid,
                                                       pme_order=6, dipole_tol=1e-8 if precision == "double" else 1e-7))
    assert ff.pme_transform() == ("fused", True)
    rng = np.random.default_rng(4)
    d = rng.normal(scale=0.01, size=pos.shape)
    res = {}
    for tr in ("library", "fused"):
        ff.set_pme_transform(tr)
        assert ff.pme_transform()[0] == tr
        res[tr] = (ff.field(pos, d).cpu().numpy(), ff.compute(pos, H))
    (f_lib, r_lib), (f_fu, r_fu) = res["library"], res["fused"]
    tol_f = 1e-12 if precision == "double" else 2e-6
    assert _rel_rms(f_fu, f_lib) < tol_f
    if precision == "double":
        assert r_fu.iterations == r_lib.iterations
    else:
        assert abs(r_fu.iterations - r_lib.iterations) <= 1
    tol_e = 1e-11 if precision == "double" else 1e-7
    for k in ("elec", "total"):
        assert abs(r_fu.energy[k] - r_lib.energy[k]) <= tol_e * abs(r_lib.energy[k]), k
    assert _rel_rms(r_fu.forces.cpu().numpy(), r_lib.forces.cpu().numpy()) < (1e-10 if precision == "double" else 1e-5)


def test_fused_transform_needs_power_of_two_axes():
    import pgm_jax_amd as P

    top, pos, H = systems.water_box(512, 2.483, seed=3)
    ff = P.PGMForceField(_system(top), H, P.MDSettings(precision="mixed", pme_grid=(36, 32, 32), dipole_tol=1e-7))
    assert ff.pme_transform() == ("library", False)
    with pytest.raises(P.PGMError):
        ff.set_pme_transform("fused")
    r = ff.compute(pos, H)
    assert r.residual <= 1e-7


def test_fused_transform_on_beads():
    """Bead contexts run the fused kernels bead by bead: every bead equals its own single-system context."""
    import pgm_jax_amd as P

    nb = 3
    qs = []
    for b in range(nb):
        top, pos, H = systems.water_box(512, 2.483, seed=30 + b)
        qs.append(pos)
    q = np.stack(qs)
    s = P.MDSettings(precision="mixed", cutoff=0.9, ewald_beta=4.0, pme_grid=(32, 32, 32), pme_order=6, dipole_tol=1e-7)
    beads = P.PGMBeads(_system(top), H, s, beads=nb)
    assert beads.pme_transform() == ("fused", True)
    forces, U, e = beads.compute(q, H)
    per, its, _ = beads.bead_results()
    for b in range(nb):
        ff = P.PGMForceField(_system(top), H, s)
        ff.set_pme_transform("library")
        r = ff.compute(q[b], H)
        assert abs(int(its[b]) - r.iterations) <= 1, (b, its[b], r.iterations)
        assert abs(per[b]["total"] - r.energy["total"]) <= 1e-7 * abs(r.energy["total"])
        assert _rel_rms(forces[b].cpu().numpy(), r.forces.cpu().numpy()) < 1e-5


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_row_fork_point_changes_nothing(precision):
    """The row field's start on the side stream (before / after the spread, after the forward planes,
    joined before them, or no side stream) only moves the overlap: fields, iterations and energies are
    bitwise the same."""
    import pgm_jax_amd as P

    top, pos, H = systems.water_box(512, 2.483, seed=9)
    ff = P.PGMForceField(_system(top), H, P.MDSettings(precision=precision, cutoff=0.9, ewald_beta=4.0,
                                                       pme_grid=(32, 32, 32), pme_order=6, dipole_tol=1e-7))
    d = np.random.default_rng(5).normal(scale=0.01, size=pos.shape)
    out = []
    for stage in ("start", "after_spread", "after_forward_planes", "start_join_before_planes", "none"):
        ff.set_row_fork(stage)
        r = ff.compute(pos, H)
        out.append((ff.field(pos, d).cpu().numpy(), r.iterations, r.energy["total"]))
    for f, it, e in out[1:]:
        assert np.array_equal(f, out[0][0]) and it == out[0][1] and e == out[0][2]
    ff.set_pme_transform("library")
    with pytest.raises(P.PGMError):
        ff.set_row_fork("after_forward_planes")
    ff.set_row_fork("none")  # the library transform without a side stream: the same field again
    assert np.allclose(ff.field(pos, d).cpu().numpy(), out[0][0], rtol=1e-5 if precision == "mixed" else 1e-12,
                       atol=1e-6 * np.abs(out[0][0]).max())
