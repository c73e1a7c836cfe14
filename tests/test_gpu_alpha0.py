"""Atoms without polarizability on the GPU against the oracle (forcefield.py:767-780 `_div_alpha`).

alpha = 0 atoms keep a zero induced dipole and add nothing to |mu|^2 / (2 alpha); the reciprocal
1/alpha of the operator is 0 for them.  FP64: energies to 1e-9 relative, dipoles and forces to
1e-8 relative RMS, the same CG iteration count; mixed: 1e-6 / 1e-5.  The molecule-block
preconditioner (which inverts diag(1/alpha) blocks) is refused for such systems.
"""

import dataclasses

import numpy as np
import pytest

from oracle import systems
from oracle.forcefield import ForceField, Settings

pytestmark = pytest.mark.gpu


def _system(P, top):
    return P.System(q=top.q, radius=top.radius, alpha=top.alpha, lj_rmin_half=top.lj_rmin_half,
                    lj_sqrt_eps=top.lj_sqrt_eps, masses=top.masses, mol=top.mol, cov_i=top.cov_i,
                    cov_j=top.cov_j, cov_c=top.cov_c)


def _rel_rms(a, b):
    return float(np.sqrt(np.mean((a - b) ** 2)) / np.sqrt(np.mean(b**2)))


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_zero_alpha_atoms_match_oracle(precision):
    import pgm_jax_amd as P

    top, pos, H = systems.water_box(512, 2.483, seed=14)
    a = top.alpha.copy()
    idx = np.arange(1, top.n, 9)  # one hydrogen in every third water
    a[idx] = 0.0
    top = dataclasses.replace(top, alpha=a)
    kw = dict(cutoff=0.9, ewald_beta=4.0, pme_grid=(32, 32, 32), pme_order=6, lj_lrc=True, dipole_tol=1e-10, max_iter=300)
    ref = ForceField(top, H, Settings(**kw)).compute(pos, H)
    ff = P.PGMForceField(_system(P, top), H, P.MDSettings(precision=precision, **kw))
    r = ff.compute(pos, H)
    mu = r.induction.mu.cpu().numpy()
    print(f"alpha0 {precision}: E {r.energy['total']:.9f} (oracle {ref.energy['total']:.9f}), CG {r.iterations} "
          f"(oracle {ref.iterations})", flush=True)
    assert np.all(mu[idx] == 0.0)
    tol_e, tol_f = (1e-9, 1e-8) if precision == "double" else (1e-6, 1e-5)
    assert abs(r.energy["total"] - ref.energy["total"]) <= tol_e * abs(ref.energy["total"])
    assert _rel_rms(mu, ref.mu) < tol_f
    assert _rel_rms(r.forces.cpu().numpy(), ref.forces) < tol_f
    if precision == "double":
        assert r.iterations == ref.iterations
    with pytest.raises(P.PGMError):
        P.PGMForceField(_system(P, top), H, P.MDSettings(precision=precision, precond="molecule", **kw))
