"""CG row-field layouts agree (GPU): atom rows against molecule blocks (4, 8 and 16 lanes).

The molecule blocks hold the same pair kernels G1, G2 (pair_G of the same FP64 minimum-image
displacements rounded to the compute precision) and sum the same pair terms in molecule-pair order,
so double precision agrees to 1e-12 relative RMS in the field and mixed precision to 1e-6; a full
evaluation (CG from alpha b) takes the same iteration count with energies equal to 1e-12 (double) /
1e-7 (mixed) relative.  Atoms wrapped one by one into the box split molecules across the boundary:
their pairs then disagree on the lattice shift and take their own minimum images.
"""

import numpy as np
import pytest

from oracle import systems

pytestmark = pytest.mark.gpu

CFG1 = dict(cutoff=0.9, ewald_beta=4.0, pme_grid=(32, 32, 32), pme_order=6, dipole_tol=1e-7)


def _system(top):
    import pgm_jax_amd as P

    return P.System(q=top.q, radius=top.radius, alpha=top.alpha, lj_rmin_half=top.lj_rmin_half,
                    lj_sqrt_eps=top.lj_sqrt_eps, masses=top.masses, mol=top.mol, cov_i=top.cov_i,
                    cov_j=top.cov_j, cov_c=top.cov_c)


def _rel_rms(a, b):
    return float(np.sqrt(np.mean((a - b) ** 2)) / np.sqrt(np.mean(b**2)))


@pytest.mark.parametrize("precision", ["double", "mixed"])
@pytest.mark.parametrize("wrapped", [False, True])
def test_molecule_blocks_match_atom_rows(precision, wrapped):
    import pgm_jax_amd as P

    top, pos, H = systems.water_box(512, 2.483, seed=5)
    if wrapped:  # every atom into [0, L): molecules on the faces are split
        pos = pos - np.floor(pos / np.diag(H)) * np.diag(H)
    print(f"layouts precision={precision} wrapped={wrapped}", flush=True)
    ff = P.PGMForceField(_system(top), H, P.MDSettings(precision=precision, **CFG1))
    assert ff.row_layout() == ("molecule", 4)
    rng = np.random.default_rng(2)
    d = rng.normal(scale=0.01, size=pos.shape)
    ff.set_row_layout("atom")
    f_atom = ff.field(pos, d).cpu().numpy()
    r_atom = ff.compute(pos, H)
    tol_f = 1e-12 if precision == "double" else 1e-6
    tol_e = 1e-12 if precision == "double" else 1e-7
    for lanes in (4, 8, 16):
        ff.set_row_layout("molecule", lanes)
        assert ff.row_layout() == ("molecule", lanes)
        f_mol = ff.field(pos, d).cpu().numpy()
        assert _rel_rms(f_mol, f_atom) < tol_f, lanes
        r = ff.compute(pos, H)
        assert r.iterations == r_atom.iterations, (lanes, r.iterations, r_atom.iterations)
        for k in ("elec", "vdw", "total"):
            assert abs(r.energy[k] - r_atom.energy[k]) <= tol_e * abs(r_atom.energy[k]), (lanes, k)
        assert _rel_rms(r.forces.cpu().numpy(), r_atom.forces.cpu().numpy()) < 10 * tol_f


def test_atom_layout_only_for_other_molecules():
    """The toy water/methanol box is not all 3-site molecules: atom rows, and molecule blocks refused."""
    import pgm_jax_amd as P

    top, pos, H = systems.small_box(0)
    ff = P.PGMForceField(_system(top), H, P.MDSettings(cutoff=0.6, ewald_beta=6.0, pme_grid=(32,) * 3, skin=0.05))
    assert ff.row_layout()[0] == "atom"
    with pytest.raises(P.PGMError):
        ff.set_row_layout("molecule", 4)


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_autotune_keeps_the_results(precision):
    """The autotuner only picks among shapes of the same sums: energies and iterations stay (to rounding)."""
    import pgm_jax_amd as P

    top, pos, H = systems.water_box(512, 2.483, seed=6)
    print(f"autotune precision={precision}", flush=True)
    ff = P.PGMForceField(_system(top), H, P.MDSettings(precision=precision, **CFG1))
    r0 = ff.compute(pos, H)
    rep = ff.autotune(pos, reps=5)
    print("autotune:", rep, flush=True)
    assert rep["layouts_tried"] == 4 and rep["row_layout"] in ("atom", "molecule")
    assert ff.row_layout() == (rep["row_layout"], rep["row_lanes"])
    assert all(5 <= t <= 8 for t in rep["spread_tile"]) and rep["row_field_ms"] > 0 and rep["pme_ms"] > 0
    r1 = ff.compute(pos, H)
    tol = 1e-11 if precision == "double" else 1e-6
    assert abs(r1.energy["total"] - r0.energy["total"]) <= tol * abs(r0.energy["total"])
    assert abs(r1.iterations - r0.iterations) <= (0 if precision == "double" else 1)
