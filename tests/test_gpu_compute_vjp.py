"""Parameter cotangents of (E, F, mu) on the GPU against the oracle and the reference's gradient (GPU).

pgm_compute_vjp: theta_bar = e_bar dE/dtheta + f_bar . dF/dtheta + mu_bar . dmu/dtheta for the
charges, radii, polarizabilities, LJ rmin/2 and sqrt(eps) and covalent-dipole strengths, the force
part by forward-mode derivatives along f_bar (pair set and masks fixed), the dipole part by the
adjoint solve.  Checks:
  * the golden dloss.* of ff_small_box (jax.grad of 1e-6 sum F^2 + 1e2 sum mu^2 over the tied
    table, tests/regression/regression_cases.py:253-287; the reference's adjoint_tol 1e-6): the
    GPU's own FP64 compute and VJP, summed per key, within 1e-8 of the largest entry;
  * the oracle's compute_vjp (pinned to the same golden on the CPU, tests/test_oracle_compute_vjp.py)
    with e_bar, f_bar and mu_bar all non-zero, the same mu on both sides: FP64 within 1e-8 of each
    output's largest entry (atom rows; 512 waters on molecule blocks with split cutoffs), mixed
    (FP32 CG and grids) within 1e-4.
"""

import os

import numpy as np
import pytest

from oracle import systems
from oracle.forcefield import ForceField, Settings

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden", "ff_small_box.npz")
KEYS = ("q", "radius", "alpha", "lj_rmin_half", "lj_sqrt_eps", "cov")
SMALL = dict(cutoff=0.6, ewald_beta=6.0, pme_grid=(32, 32, 32), pme_order=6, lj_lrc=True, dipole_tol=1e-10, max_iter=500)
TYPES = {"MeOH": ["c3", "oh", "h1", "h1", "h1", "ho"], "WAT": ["OW", "HW", "HW"]}
COV = {"MeOH": [(0, 1), (1, 0), (1, 5), (5, 1)] + [c for h in (2, 3, 4) for c in ((0, h), (h, 0))],
       "WAT": [(0, 1), (0, 2), (1, 0), (2, 0)]}
MOLS = ["MeOH"] * 4 + ["WAT"] * 30


def _system(P, top):
    return P.System(q=top.q, radius=top.radius, alpha=top.alpha, lj_rmin_half=top.lj_rmin_half,
                    lj_sqrt_eps=top.lj_sqrt_eps, masses=top.masses, mol=top.mol, cov_i=top.cov_i,
                    cov_j=top.cov_j, cov_c=top.cov_c)


def _tied_keys():
    """Tying keys of the toy box (system.py:300-343): types, "<molecule>:<type>", "<molecule>:<ti>><tj>"."""
    atom_type, q_key, cov_key = [], [], []
    for m in MOLS:
        atom_type += TYPES[m]
        q_key += [f"{m}:{t}" for t in TYPES[m]]
        cov_key += [f"{m}:{TYPES[m][i]}>{TYPES[m][j]}" for i, j in COV[m]]
    by = {k: atom_type for k in ("radius", "alpha", "lj_rmin_half", "lj_sqrt_eps")}
    return dict(by, q=q_key, cov=cov_key)


def _reduce(vals, keys):
    order = list(dict.fromkeys(keys))
    out = np.zeros(len(order))
    np.add.at(out, [order.index(k) for k in keys], vals)
    return out


def test_dloss_of_the_reference():
    import pgm_jax_amd as P

    top, pos, H = systems.small_box(0)
    ff = P.PGMForceField(_system(P, top), H, P.MDSettings(precision="double", predictor="none", skin=0.05, **SMALL))
    res = ff.compute(pos, H)
    F, mu = res.forces, res.induction.mu
    out = ff.compute_vjp(pos, H, mu, 0.0, 2e-6 * F, 2e2 * mu, tol=1e-6)
    print(f"dloss: adjoint CG {out['iterations']}, residual {out['residual']:.2e}", flush=True)
    keys = _tied_keys()
    with np.load(GOLD) as gold:
        for qn in KEYS:
            got = _reduce(out[qn].cpu().numpy(), keys[qn])
            ref = gold["dloss." + qn]
            err = np.abs(got - ref).max() / np.abs(ref).max()
            print(f"  {qn}: {err:.2e}", flush=True)
            assert got.shape == ref.shape and err <= 1e-8, qn


def _against_oracle(top, pos, H, kw, precision, bar, precond="jacobi"):
    import pgm_jax_amd as P

    ref_ff = ForceField(top, H, Settings(predictor="none", **kw))
    r = ref_ff.compute(pos, H)
    e_bar, f_bar, mu_bar = 0.7, 2e-6 * r.forces, 2e2 * r.mu
    ref = ref_ff.compute_vjp(pos, H, r.mu, e_bar, f_bar, mu_bar, tol=1e-11)
    ff = P.PGMForceField(_system(P, top), H, P.MDSettings(precision=precision, predictor="none", precond=precond,
                                                         skin=0.05, **kw))
    tol = 1e-11 if precision == "double" else 1e-6
    got = ff.compute_vjp(pos, H, r.mu, e_bar, f_bar, mu_bar, tol=tol)
    print(f"  adjoint CG {got['iterations']}, residual {got['residual']:.2e}", flush=True)
    for k in KEYS:
        g = got[k].cpu().numpy()
        rk = np.asarray(ref[k])
        err = np.abs(g - rk).max() / np.abs(rk).max()
        print(f"    {k}: {err:.2e}", flush=True)
        assert g.shape == rk.shape and err <= bar, k
    return ff


@pytest.mark.parametrize("precision,bar", [("double", 1e-8), ("mixed", 1e-4)])
def test_atom_rows_against_oracle(precision, bar):
    top, pos, H = systems.small_box(0)
    print(f"compute_vjp atom rows {precision}", flush=True)
    _against_oracle(top, pos, H, SMALL, precision, bar)


def test_molecule_blocks_split_cutoffs_against_oracle():
    import pgm_jax_amd as P

    top, pos, H = systems.water_box(512, 2.483, seed=5)
    es = P.elec_cutoff_settings(0.7)
    kw = dict(cutoff=0.9, elec_cutoff=0.7, ewald_beta=es["ewald_beta"], pme_grid=(48, 48, 48), pme_order=6, lj_lrc=True,
              dipole_tol=1e-10, max_iter=500)
    print("compute_vjp molecule blocks double", flush=True)
    ff = _against_oracle(top, pos, H, kw, "double", 1e-8, precond="molecule")
    assert ff.row_layout()[0] == "molecule"


def test_energy_cotangent_alone():
    """f_bar = mu_bar = None: dE/dtheta at fixed mu only (no adjoint solve)."""
    import pgm_jax_amd as P

    top, pos, H = systems.small_box(0)
    ref_ff = ForceField(top, H, Settings(predictor="none", **SMALL))
    mu = ref_ff.compute(pos, H).mu
    ref = ref_ff.compute_vjp(pos, H, mu, 1.0)
    ff = P.PGMForceField(_system(P, top), H, P.MDSettings(precision="double", predictor="none", skin=0.05, **SMALL))
    got = ff.compute_vjp(pos, H, mu, 1.0)
    assert got["iterations"] == 0
    for k in KEYS:
        g, rk = got[k].cpu().numpy(), np.asarray(ref[k])
        assert np.abs(g - rk).max() <= 1e-10 * np.abs(rk).max(), k


def _toy_system(P):
    """The toy box's System from its molecule templates (types and bonds of pgm_jax/models/toy.py:26-78),
    in the order of systems.small_box: 4 methanols, then 30 waters."""
    from test_param_table import toy_molecules

    M, W = toy_molecules(P)
    return P.System([M] * 4 + [W] * 30)


def test_grad_params_through_the_tied_table():
    """The product's own tying (System from templates) and grad_params give the golden dloss.*."""
    import pgm_jax_amd as P

    _, pos, H = systems.small_box(0)
    sys_ = _toy_system(P)
    ff = P.PGMForceField(sys_, H, P.MDSettings(precision="double", predictor="none", skin=0.05, **SMALL))
    res = ff.compute(pos, H, params=sys_.initial_params())
    grad = ff.grad_params(pos, H, res.induction.mu, 0.0, 2e-6 * res.forces, 2e2 * res.induction.mu, tol=1e-6)
    with np.load(GOLD) as gold:
        for qn in KEYS:
            ref = gold["dloss." + qn]
            err = np.abs(grad[qn] - ref).max() / np.abs(ref).max()
            print(f"  grad_params {qn}: {err:.2e}", flush=True)
            assert grad[qn].shape == ref.shape and err <= 1e-8, qn


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_params_pytree_in_compute(precision):
    """compute(params=p) equals a context built with p's per-atom values; going back to the initial
    table restores the first result exactly."""
    import pgm_jax_amd as P

    _, pos, H = systems.small_box(0)
    sys_ = _toy_system(P)
    st = P.MDSettings(precision=precision, predictor="none", skin=0.05, **SMALL)
    ff = P.PGMForceField(sys_, H, st)
    p0 = sys_.initial_params()
    r0 = ff.compute(pos, H, params=p0)
    p = {"q": p0["q"] * 1.01, "radius": p0["radius"] * 1.02, "alpha": p0["alpha"] * 0.97,
         "lj_rmin_half": p0["lj_rmin_half"] * 1.01, "lj_sqrt_eps": p0["lj_sqrt_eps"] * 0.99, "cov": p0["cov"] * 1.05}
    r1 = ff.compute(pos, H, params=p)
    ex = sys_.expand(p)
    fresh = P.PGMForceField(P.System(q=ex["q"], radius=ex["radius"], alpha=ex["alpha"], lj_rmin_half=ex["lj_rmin_half"],
                                     lj_sqrt_eps=ex["lj_sqrt_eps"], masses=sys_.masses, mol=sys_.mol, cov_i=sys_.cov_i,
                                     cov_j=sys_.cov_j, cov_c=ex["cov"]), H, st)
    rf = fresh.compute(pos, H)
    print(f"  {precision}: E {r0.energy['total']:.10f} -> {r1.energy['total']:.10f} (fresh {rf.energy['total']:.10f})", flush=True)
    assert abs(r1.energy["total"] - r0.energy["total"]) > 1e-3 * abs(r0.energy["total"])
    for k in ("elec", "vdw", "total"):
        assert abs(r1.energy[k] - rf.energy[k]) <= 1e-10 * abs(rf.energy[k]), k
    assert (r1.forces - rf.forces).abs().max().item() <= 1e-9 * rf.forces.abs().max().item()
    assert r1.iterations == rf.iterations
    r2 = ff.compute(pos, H, params=p0)
    assert r2.energy == r0.energy and bool((r2.forces == r0.forces).all())
