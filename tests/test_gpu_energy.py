"""PGMForceField.energy, the Monte Carlo barostat's trial energy (GPU).

forcefield.py:2562-2626: the induced dipoles are solved by CG from the state's mu (no predictor;
field and norm afresh) and only the energy is returned, with the state's history untouched.  FP64
against the oracle's restatement of the same call after an isotropic volume move (energy to
1e-9 relative, the same CG iteration count); the caller's state is not modified; the energy equals
a full compute from scratch at the trial geometry to the CG tolerance.
"""

import numpy as np
import pytest

from oracle import systems
from oracle.forcefield import ForceField, InductionState, Settings

pytestmark = pytest.mark.gpu

KW = dict(cutoff=0.9, ewald_beta=4.0, pme_grid=(32, 32, 32), pme_order=6, lj_lrc=True, dipole_tol=1e-10, max_iter=300)


def _system(P, top):
    return P.System(q=top.q, radius=top.radius, alpha=top.alpha, lj_rmin_half=top.lj_rmin_half,
                    lj_sqrt_eps=top.lj_sqrt_eps, masses=top.masses, mol=top.mol, cov_i=top.cov_i,
                    cov_j=top.cov_j, cov_c=top.cov_c)


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_trial_energy_after_volume_move(precision):
    import pgm_jax_amd as P

    top, pos, H = systems.water_box(512, 2.483, seed=13)
    ff = P.PGMForceField(_system(P, top), H, P.MDSettings(precision=precision, skin=0.1, **KW))
    r1 = ff.compute(pos, H)
    mu_before = r1.induction.mu.clone()
    s = 1.003  # molecular centres scaled, as a barostat trial moves them
    com = np.stack([pos[top.mol == m].mean(0) for m in range(top.mol.max() + 1)])
    pos2 = pos + (s - 1.0) * com[top.mol]
    H2 = H * s
    e, st, it, ovf = ff.energy(pos2, H2, None, r1.induction)
    assert not ovf
    assert np.array_equal(r1.induction.mu.cpu().numpy(), mu_before.cpu().numpy())  # caller's state kept
    assert st.count == r1.induction.count and st.hist is r1.induction.hist
    ref = ForceField(top, H2, Settings(**KW))
    start = InductionState(r1.induction.mu.cpu().numpy(), np.zeros((4, top.n, 3)), 0, 1.0)
    e_ref, mu_ref, it_ref = ref.energy(pos2, H2, start)
    print(f"trial {precision}: E {e:.9f} (oracle {e_ref:.9f}), CG {it} (oracle {it_ref})", flush=True)
    tol = 1e-9 if precision == "double" else 1e-6
    assert abs(e - e_ref) <= tol * abs(e_ref)
    if precision == "double":
        assert it == it_ref
        assert float(np.max(np.abs(st.mu.cpu().numpy() - mu_ref))) < 1e-8
    # the same energy as a full evaluation from scratch at the trial geometry
    fresh = P.PGMForceField(_system(P, top), H2, P.MDSettings(precision=precision, skin=0.1, **KW)).compute(pos2, H2)
    assert abs(e - fresh.energy["total"]) <= tol * abs(e)
    assert it < fresh.iterations
