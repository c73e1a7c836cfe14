"""The tied parameter table of the Python mirror (System / Molecule; pgm_jax/system.py:260-447) (CPU).

Keys of the toy water / methanol box in the reference's first-occurrence order (the golden dloss.*
vectors of ff_small_box have exactly these lengths: 6 per atom quantity, 8 covalent dipoles); a type
split into two symmetry classes gets "<type>.<k>" labels; the keys do not depend on the atom order;
expand and reduce_params are each other's transposes.
"""

import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden", "ff_small_box.npz")


@pytest.fixture
def pgm():
    import pgm_jax_amd

    return pgm_jax_amd


def toy_molecules(P):
    """The toy templates of pgm_jax/models/toy.py:26-78 (types and bonds as there)."""
    W = P.Molecule("WAT", ["O", "H", "H"], [-0.8, 0.4, 0.4], [0.06, 0.05, 0.05], [1.0e-3, 0.3e-3, 0.3e-3],
                   [(0, 1, -0.02), (0, 2, -0.02), (1, 0, 0.008), (2, 0, 0.008)], [0.178, 0.0, 0.0], [0.80, 0.0, 0.0],
                   types=["OW", "HW", "HW"], bonds=[(0, 1), (0, 2)])
    cov = [(0, 1, 0.01), (1, 0, -0.01), (1, 5, -0.02), (5, 1, 0.005)]
    cov += [c for h in (2, 3, 4) for c in ((0, h, 0.002), (h, 0, -0.002))]
    M = P.Molecule("MeOH", ["C", "O", "H", "H", "H", "H"], [0.12, -0.62, 0.02, 0.02, 0.02, 0.44],
                   [0.07, 0.06, 0.05, 0.05, 0.05, 0.05], [1.2e-3, 0.8e-3, 0.4e-3, 0.4e-3, 0.4e-3, 0.3e-3], cov,
                   [0.19, 0.172, 0.139, 0.139, 0.139, 0.02], [0.33, 0.85, 0.20, 0.20, 0.20, 0.1],
                   types=["c3", "oh", "h1", "h1", "h1", "ho"], bonds=[(0, 1), (0, 2), (0, 3), (0, 4), (1, 5)])
    return M, W


def test_toy_box_keys(pgm):
    M, W = toy_molecules(pgm)
    S = pgm.System([M] * 4 + [W] * 30)
    by_type = ["c3", "oh", "h1", "ho", "OW", "HW"]
    assert S.param_keys["radius"] == by_type and S.param_keys["lj_sqrt_eps"] == by_type
    assert S.param_keys["q"] == ["MeOH:c3", "MeOH:oh", "MeOH:h1", "MeOH:ho", "WAT:OW", "WAT:HW"]
    assert S.param_keys["cov"] == ["MeOH:c3>oh", "MeOH:oh>c3", "MeOH:oh>ho", "MeOH:ho>oh", "MeOH:c3>h1", "MeOH:h1>c3",
                                   "WAT:OW>HW", "WAT:HW>OW"]
    with np.load(GOLD) as gold:
        for qn in ("q", "radius", "alpha", "lj_rmin_half", "lj_sqrt_eps", "cov"):
            assert len(S.param_keys[qn]) == gold["dloss." + qn].shape[0]
    p0 = S.initial_params()
    assert np.allclose(p0["q"], [0.12, -0.62, 0.02, 0.44, -0.8, 0.4])
    ex = S.expand(p0)
    for qn, arr in (("q", S.q), ("radius", S.radius), ("alpha", S.alpha), ("cov", S.cov_c)):
        assert np.allclose(ex[qn], arr, rtol=1e-14, atol=0), qn


def test_two_classes_of_one_type_are_split(pgm):
    """A chain X-C-C-Y with both carbons of type "ct": different neighbours, so two classes."""
    m = pgm.Molecule("XY", ["F", "C", "C", "Cl"], [0.1, 0.0, 0.0, -0.1], [0.05] * 4, [1e-3] * 4, [(1, 2, 0.01)],
                     [0.1] * 4, [0.1] * 4, types=["f", "ct", "ct", "cl"], bonds=[(0, 1), (1, 2), (2, 3)])
    k = m.tying_keys()
    assert sorted(k["q"][1:3]) == ["XY:ct.0", "XY:ct.1"]
    assert k["radius"][1] == k["radius"][2] == "ct"
    sym = pgm.Molecule("CC", ["H", "C", "C", "H"], [0.1, -0.1, -0.1, 0.1], [0.05] * 4, [1e-3] * 4, [], [0.1] * 4, [0.1] * 4,
                       types=["h", "ct", "ct", "h"], bonds=[(0, 1), (1, 2), (2, 3)])
    assert sym.tying_keys()["q"] == ["CC:h", "CC:ct", "CC:ct", "CC:h"]


def test_keys_do_not_depend_on_the_atom_order(pgm):
    M, _ = toy_molecules(pgm)
    perm = [5, 3, 0, 4, 1, 2]
    inv = np.argsort(perm)
    cov = [(int(inv[i]), int(inv[j]), c) for i, j, c in M.cov]
    Mp = pgm.Molecule("MeOH", [M.elements[k] for k in perm], M.q[perm], M.radius[perm], M.alpha[perm], cov,
                      M.lj_rmin_half[perm], M.lj_sqrt_eps[perm], types=[M.types[k] for k in perm],
                      bonds=[(int(inv[i]), int(inv[j])) for i, j in M.bonds])
    a, b = M.tying_keys(), Mp.tying_keys()
    assert [a["q"][k] for k in perm] == b["q"]
    assert a["cov"] == b["cov"]


def test_expand_and_reduce_are_transposes(pgm):
    M, W = toy_molecules(pgm)
    S = pgm.System([M] * 4 + [W] * 30)
    rng = np.random.default_rng(3)
    p = {qn: rng.standard_normal(len(k)) for qn, k in S.param_keys.items()}
    ex = S.expand(p)
    c = {qn: rng.standard_normal(len(v)) for qn, v in ex.items()}
    red = S.reduce_params(c)
    for qn in p:
        assert np.isclose(np.dot(red[qn], p[qn]), np.dot(c[qn], ex[qn])), qn
