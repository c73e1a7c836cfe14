"""Parameter pytrees in the exact Ewald models (periodic.PeriodicModel, ewald.PeriodicPGM) (GPU).

energy / forces / strain_derivative / pressure with params = p (the System's tied table,
pgm_jax/system.py:345-447, as periodic.py:205-281 and ewald.py:421-470 take it) must equal a model
built from a System holding p's per-atom values (FP64: energies to 1e-12 relative, forces and
strain to 1e-11 of their largest entry), and a call with params=None must give the System's own
values again, bit for bit.
"""

import numpy as np
import pytest

from oracle import systems

pytestmark = pytest.mark.gpu


def _toy_system(P):
    from test_param_table import toy_molecules

    M, W = toy_molecules(P)
    return P.System([M] * 4 + [W] * 30)


def _perturbed(p0):
    return {"q": p0["q"] * 1.01, "radius": p0["radius"] * 1.02, "alpha": p0["alpha"] * 0.97,
            "lj_rmin_half": p0["lj_rmin_half"] * 1.01, "lj_sqrt_eps": p0["lj_sqrt_eps"] * 0.99, "cov": p0["cov"] * 1.05}


def _arrays_system(P, sys_, ex):
    return P.System(q=ex["q"], radius=ex["radius"], alpha=ex["alpha"], lj_rmin_half=ex["lj_rmin_half"],
                    lj_sqrt_eps=ex["lj_sqrt_eps"], masses=sys_.masses, mol=sys_.mol, cov_i=sys_.cov_i, cov_j=sys_.cov_j,
                    cov_c=ex["cov"])


def _close(a, b, rel):
    a, b = np.asarray(a), np.asarray(b)
    return np.abs(a - b).max() <= rel * max(np.abs(b).max(), 1e-300)


def test_periodic_model_params():
    import pgm_jax_amd as P

    _, pos, H = systems.small_box(0)
    sys_ = _toy_system(P)
    kw = dict(cutoff=0.6, ewald_beta=4.5, lj_lrc=True)
    model = P.PeriodicModel(sys_, H, pos, **kw)
    e0 = model.energy(pos)
    p = _perturbed(sys_.initial_params())
    e1 = model.energy(pos, params=p)
    f1 = model.forces(pos, params=p).cpu().numpy()
    w1 = model.strain_derivative(pos, params=p)
    pr1 = model.pressure(pos, params=p)
    fresh = P.PeriodicModel(_arrays_system(P, sys_, sys_.expand(p)), H, pos, **kw)
    ef = fresh.energy(pos)
    print(f"E {e0['total']:.8f} -> {e1['total']:.8f} (fresh {ef['total']:.8f})", flush=True)
    assert abs(e1["total"] - e0["total"]) > 1e-3 * abs(e0["total"])
    for k in ("perm", "ind", "elec", "vdw", "total"):
        assert abs(e1[k] - ef[k]) <= 1e-12 * max(abs(ef[k]), 1.0), k
    assert _close(f1, fresh.forces(pos).cpu().numpy(), 1e-11)
    assert _close(w1, fresh.strain_derivative(pos), 1e-11)
    assert abs(pr1 - fresh.pressure(pos)) <= 1e-9 * max(abs(pr1), 1.0)
    assert model.energy(pos) == e0


def test_periodic_pgm_params():
    import pgm_jax_amd as P

    _, pos, H = systems.small_box(0)
    sys_ = _toy_system(P)
    kw = dict(cutoff=0.6, ewald_beta=4.5)
    m = P.PeriodicPGM(sys_, H, pos, **kw)
    mu0 = m.induced_dipoles(pos).cpu().numpy()
    p = _perturbed(sys_.initial_params())
    e1, _ = m.energy(pos, params=p)
    mu1 = m.induced_dipoles(pos, params=p).cpu().numpy()
    fresh = P.PeriodicPGM(_arrays_system(P, sys_, sys_.expand(p)), H, pos, **kw)
    ef, _ = fresh.energy(pos)
    for k in ("perm", "ind", "total"):
        assert abs(e1[k] - ef[k]) <= 1e-12 * max(abs(ef[k]), 1.0), k
    assert _close(mu1, fresh.induced_dipoles(pos).cpu().numpy(), 1e-11)
    assert _close(m.forces(pos, params=p).cpu().numpy(), fresh.forces(pos).cpu().numpy(), 1e-11)
    assert np.array_equal(m.induced_dipoles(pos).cpu().numpy(), mu0)
