"""The backward pass of the induced-dipole solve on the GPU against the oracle (GPU).

pgm_dipole_vjp (the custom VJP of PGMForceField._solve, pgm_jax/md/forcefield.py:1853-1915): the
adjoint CG A lam = mu_bar, then lam pulled back to the cotangents of the charges, Gaussian radii,
polarizabilities, permanent dipoles and covalent-dipole strengths.  The oracle's dipole_vjp is
pinned against finite differences of a dipole loss on the CPU (tests/test_oracle_dipole_vjp.py).
Both sides get the same mu (the oracle's converged dipoles) and mu_bar.  Double precision, CG to
1e-11: every output within 1e-8 of the oracle's largest entry of that output.  Mixed (FP32 CG and
fields, FP64 pull-back sums), CG to 1e-6: within 2e-5.  On the atom rows (water/methanol box,
Jacobi) and on the molecule blocks (512 waters, molecule preconditioner).
"""

import numpy as np
import pytest

from oracle import systems
from oracle.forcefield import ForceField, Settings

pytestmark = pytest.mark.gpu

KEYS = ("lam", "q", "radius", "alpha", "p", "cov")


def _system(P, top):
    return P.System(q=top.q, radius=top.radius, alpha=top.alpha, lj_rmin_half=top.lj_rmin_half,
                    lj_sqrt_eps=top.lj_sqrt_eps, masses=top.masses, mol=top.mol, cov_i=top.cov_i,
                    cov_j=top.cov_j, cov_c=top.cov_c)


def _check(top, pos, H, kw, precision, precond, mu_bar_of):
    import pgm_jax_amd as P

    ref_ff = ForceField(top, H, Settings(predictor="none", **kw))
    mu = ref_ff.compute(pos, H).mu
    mu_bar = mu_bar_of(mu)
    tol = 1e-11 if precision == "double" else 1e-6
    ref = ref_ff.dipole_vjp(pos, H, mu, mu_bar, tol=1e-11)
    ff = P.PGMForceField(_system(P, top), H, P.MDSettings(precision=precision, predictor="none", precond=precond,
                                                         skin=0.05, **kw))
    got = ff.dipole_vjp(pos, H, mu, mu_bar, tol=tol)
    bar = 1e-8 if precision == "double" else 2e-5
    print(f"  CG {got['iterations']} (oracle {ref['iterations']}), residual {got['residual']:.2e}", flush=True)
    assert got["residual"] <= tol
    for k in KEYS:
        g = got[k].cpu().numpy()
        r = np.asarray(ref[k])
        assert g.shape == r.shape, k
        err = np.abs(g - r).max() / np.abs(r).max()
        print(f"    {k}: max |diff| / max |oracle| = {err:.2e}", flush=True)
        assert err <= bar, k
    return ff, got


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_vjp_atom_rows_against_oracle(precision):
    top, pos, H = systems.small_box(0)
    kw = dict(cutoff=0.6, ewald_beta=6.0, pme_grid=(32, 32, 32), pme_order=6, lj_lrc=True, dipole_tol=1e-10, max_iter=500)
    print(f"vjp atom rows {precision}", flush=True)
    _check(top, pos, H, kw, precision, "jacobi", lambda mu: 2e2 * mu)


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_vjp_molecule_blocks_against_oracle(precision):
    """512 waters, split cutoffs (electrostatics to 0.7 nm, LJ to 0.9 nm), a random mu_bar."""
    import pgm_jax_amd as P

    top, pos, H = systems.water_box(512, 2.483, seed=5)
    es = P.elec_cutoff_settings(0.7)
    kw = dict(cutoff=0.9, elec_cutoff=0.7, ewald_beta=es["ewald_beta"], pme_grid=(48, 48, 48), pme_order=6, lj_lrc=True,
              dipole_tol=1e-10, max_iter=500)
    rng = np.random.default_rng(11)
    print(f"vjp molecule blocks {precision}", flush=True)
    ff, _ = _check(top, pos, H, kw, precision, "molecule", lambda mu: rng.standard_normal(mu.shape))
    assert ff.row_layout()[0] == "molecule"


def test_vjp_leaves_the_forward_solve_unchanged():
    """A compute after a backward pass gives the same result as one before it (the adjoint shares the
    CG buffers and the CG state's normaliser)."""
    import pgm_jax_amd as P

    top, pos, H = systems.small_box(0)
    kw = dict(cutoff=0.6, ewald_beta=6.0, pme_grid=(32, 32, 32), pme_order=6, lj_lrc=True, dipole_tol=1e-10, max_iter=500)
    ff = P.PGMForceField(_system(P, top), H, P.MDSettings(precision="double", predictor="mu4", **kw))
    ind = ff.init_induction()
    out = []
    for step in range(6):
        res = ff.compute(pos, H, ind=ind)
        ind = res.induction
        out.append((res.energy["total"], res.iterations))
        if step == 3:
            ff.dipole_vjp(pos, H, res.induction.mu, 2e2 * res.induction.mu, tol=1e-9)
    twin = P.PGMForceField(_system(P, top), H, P.MDSettings(precision="double", predictor="mu4", **kw))
    ind2 = twin.init_induction()
    for step in range(6):
        res = twin.compute(pos, H, ind=ind2)
        ind2 = res.induction
        assert (res.energy["total"], res.iterations) == out[step], step


def test_vjp_rejects_bad_arguments():
    import pgm_jax_amd as P

    top, pos, H = systems.small_box(0)
    kw = dict(cutoff=0.6, ewald_beta=6.0, pme_grid=(32, 32, 32), pme_order=6, dipole_tol=1e-10, max_iter=500)
    ff = P.PGMForceField(_system(P, top), H, P.MDSettings(precision="double", predictor="none", **kw))
    mu = np.zeros((top.n, 3))
    with pytest.raises(P.PGMError):
        ff.dipole_vjp(pos, H, mu, mu, tol=0.0)
