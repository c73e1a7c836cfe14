"""Amber readers and the log table of the MD runner (CPU, no GPU calls).

The product's readers (pgm-jax_amd/amber.py) on byte copies of examples/water512's prmtop and
NetCDF restart: 512 three-site residues with the pGM3P-25 parameters (the values of
pgm_jax/param.py's reading recorded in tests/golden/water512.npz, and the prmtop constants of
the synthetic-box template), coordinates / velocities / cell equal to the fixture's, the
truncated-octahedron box from the cell (a along x, b in xy; volume = a b c sqrt(1 - cos^2 al -
cos^2 be - cos^2 ga + 2 cos al cos be cos ga)).  The RATTLE velocity projection keeps each
molecule's momentum and angular momentum and leaves no bond-stretching velocity.
"""

import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TOP = os.path.join(GOLDEN, "pgm3p25_512.prmtop")
CRD = os.path.join(GOLDEN, "pgm3p25_512.rst7")


def test_prmtop_reads_the_pgm3p25_box():
    import pgm_jax_amd as P

    mols = P.amber.read_prmtop_pgm(TOP, first_residue_only=False)
    assert len(mols) == 512 and all(m.n == 3 for m in mols)
    w = mols[0]
    with np.load(os.path.join(GOLDEN, "water512.npz"), allow_pickle=False) as z:
        for f in ("q", "radius", "alpha", "lj_rmin_half", "lj_sqrt_eps", "masses"):
            assert np.allclose(getattr(w, f), z[f], rtol=1e-12, atol=0), f
        cov = sorted(zip(z["cov_i"].tolist(), z["cov_j"].tolist(), z["cov_c"].tolist()))
    assert sorted(w.cov) == cov
    ref = P.models.pgm3p25_water()
    assert np.allclose(w.q, ref.q) and np.allclose(w.radius, ref.radius) and np.allclose(w.alpha, ref.alpha)
    assert np.allclose(w.lj_rmin_half, ref.lj_rmin_half, rtol=1e-9) and np.allclose(w.lj_sqrt_eps, ref.lj_sqrt_eps, rtol=1e-9)
    assert w.elements == ["O", "H", "H"]
    sys_ = P.System(mols)
    assert sys_.n == 1536 and sys_.nmol == 512 and abs(sys_.q.sum()) < 1e-9


def test_restart_reads_coordinates_velocities_and_box():
    import pgm_jax_amd as P

    pos, vel, H = P.amber.read_coordinates_nm(CRD)
    with np.load(os.path.join(GOLDEN, "water512.npz"), allow_pickle=False) as z:
        assert np.array_equal(pos, z["pos"]) and np.allclose(vel, z["vel"], rtol=1e-14)
        L, ang = z["cell_lengths"], z["cell_angles"]
    assert np.allclose(ang, 109.4712206, atol=1e-6)  # truncated octahedron
    assert H[0, 1] == H[0, 2] == H[1, 2] == 0.0
    assert np.allclose(np.linalg.norm(H, axis=1), L, rtol=1e-13)
    c = np.cos(np.radians(ang))
    vol = np.prod(L) * np.sqrt(1 - np.sum(c**2) + 2 * np.prod(c))
    assert abs(abs(np.linalg.det(H)) - vol) < 1e-12 * vol
    cosg = np.dot(H[0], H[1]) / (L[0] * L[1])
    assert abs(cosg - c[2]) < 1e-13


def test_rigid_velocity_projection():
    import pgm_jax_amd as P

    pos, vel, H = P.amber.read_coordinates_nm(CRD)
    m = np.tile([16.0, 1.008, 1.008], 512)
    v = P.driver.rigid_velocities(pos, vel, m)
    x = pos.reshape(-1, 3, 3)
    vv, v0, mm = v.reshape(-1, 3, 3), vel.reshape(-1, 3, 3), m.reshape(-1, 3)
    for a, b in ((0, 1), (0, 2), (1, 2)):
        g = np.einsum("mi,mi->m", x[:, a] - x[:, b], vv[:, a] - vv[:, b])
        assert np.abs(g).max() < 1e-12
    p1, p0 = (mm[:, :, None] * vv).sum(1), (mm[:, :, None] * v0).sum(1)
    assert np.abs(p1 - p0).max() < 1e-12
    l1 = (mm[:, :, None] * np.cross(x, vv)).sum(1)
    l0 = (mm[:, :, None] * np.cross(x, v0)).sum(1)
    assert np.abs(l1 - l0).max() < 1e-11 * np.abs(l0).max()


def test_log_table_format(tmp_path):
    import pgm_jax_amd as P

    path = tmp_path / "md.log"
    with P.driver.LogTable(str(path), title=["run"]) as log:
        log.write({"step": 5, "epot": -2114884.256504, "ok": True, "name": "x"})
        log.write({"step": 10, "epot": 1.5, "ok": False, "name": "yy"})
    lines = path.read_text().splitlines()
    assert lines[0] == "# run"
    assert lines[1] == "# " + " ".join(f"{c:>14s}" for c in ("step", "epot", "ok", "name"))
    assert lines[2] == "  " + " ".join([f"{5:14d}", f"{-2114884.256504:14.6f}", f"{1:14d}", f"{'x':>14s}"])
    names = lines[1].lstrip("#").split()
    row = dict(zip(names, lines[3].split()))
    assert row["step"] == "10" and float(row["epot"]) == 1.5
