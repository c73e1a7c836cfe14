"""Write the committed test fixtures from data files that the reference repository holds.

Run in the build container only (the reference is not present on GPU machines):

    python tests/golden/make_fixtures.py [/root/reference]

Every output is data (inputs and expected outputs), never reference source:

* ``water512.npz``   -- the 512-water pGM3P-25 box of ``examples/water512`` (prmtop + NetCDF rst7):
  coordinates and velocities (nm, nm/ps), cell lengths/angles, and the per-water pGM parameters
  read from the POL_GAUSS_* / LENNARD_JONES_* / MASS sections of the prmtop (converted to nm, e,
  e nm, nm^3, R*/2 nm, sqrt(kJ/mol) the way the reference's ``param.read_prmtop_pgm``
  (pgm_jax/param.py:113-233, :281-307) defines them).
* ``ff_small_box.npz``, ``ff_pgm3p25_512.npz``, ``periodic_model.npz``, ``efield_point.npz`` -- copies
  of the golden outputs of the reference's regression harness (tests/regression/golden/), recorded
  with jax 0.11.2.
* ``pgm3p25_512.prmtop``, ``pgm3p25_512.rst7`` -- byte copies of the two input files of
  ``examples/water512`` (the reference's install check, tests/test_water512.py), read by the
  product's own Amber readers in the tests of the MD runner.
* ``md_rigid_points.npz`` -- the snapshots of two of the reference's rigid-water trajectory goldens
  (md_rigid_nve, md_rigid_npt: 64 pGM3P-25 waters, float64, CG to 1e-10, tests/regression/
  regression_cases.py:356-411): per snapshot the positions, box, induced dipoles and the recorded
  elec / vdw / potential energies, as single points at new geometries (the NPT ones in changed boxes).
* ``pmemd_pbc.npz`` -- the independent Fortran (pmemd-pgm) single point of
  data/validation/amber_ref/pmemd_pbc: EELEC / VDWAALS (kcal/mol), forces (kcal/mol/A) and induced
  dipoles (e A).

Only numpy / scipy text and NetCDF readers are used (no pickles, nothing executed from the files).
"""

from __future__ import annotations

import os
import re
import shutil
import sys

import numpy as np
from scipy.io import netcdf_file

HERE = os.path.dirname(os.path.abspath(__file__))
KCAL = 4.184
ANG_NM = 0.1


def read_prmtop_sections(path: str) -> dict:
    """Return {flag: list of str tokens} of an Amber prmtop (fixed-width FORMAT fields)."""
    out, flag, fmt = {}, None, None
    with open(path) as fh:
        for line in fh:
            if line.startswith("%FLAG"):
                flag = line.split()[1]
                out[flag] = []
                fmt = None
            elif line.startswith("%FORMAT"):
                fmt = re.search(r"\((\d+)([aIE])(\d+)", line)
            elif line.startswith("%COMMENT") or line.startswith("%VERSION"):
                continue
            elif flag is not None and fmt is not None:
                width = int(fmt.group(3))
                s = line.rstrip("\n")
                out[flag] += [s[k : k + width].strip() for k in range(0, len(s), width) if s[k : k + width].strip()]
    return out


def water512(ref: str) -> dict:
    """Return the water512 box: coordinates, velocities, cell and one water's pGM parameters."""
    top = read_prmtop_sections(os.path.join(ref, "examples/water512/pgm3p25_512.prmtop"))
    q = np.array(top["POL_GAUSS_MONOPOLES_LIST"], float)
    rad = np.array(top["POL_GAUSS_RADII_LIST"], float) * ANG_NM
    alp = np.array(top["POL_GAUSS_POLARIZABILITY_LIST"], float) * ANG_NM**3
    nptr = np.array(top["POL_GAUSS_COVALENT_POINTERS_LIST"], int)
    catm = np.array(top["POL_GAUSS_COVALENT_ATOMS_LIST"], int) - 1
    cdip = np.array(top["POL_GAUSS_COVALENT_DIPOLES_LIST"], float) * ANG_NM
    mass = np.array(top["MASS"], float)
    ntypes = int(top["POINTERS"][1])
    ti = np.array(top["ATOM_TYPE_INDEX"], int) - 1
    nbi = np.array(top["NONBONDED_PARM_INDEX"], int).reshape(ntypes, ntypes) - 1
    A = np.array(top["LENNARD_JONES_ACOEF"], float)[nbi]
    B = np.array(top["LENNARD_JONES_BCOEF"], float)[nbi]
    a, b = np.diag(A), np.diag(B)
    ok = (a > 0) & (b > 0)
    rmin = np.where(ok, (2 * np.where(ok, a, 1) / np.where(ok, b, 1)) ** (1 / 6), 0.0)
    eps = np.where(ok, b**2 / (4 * np.where(ok, a, 1)), 0.0)
    rh = (rmin / 2 * ANG_NM)[ti]
    se = np.sqrt(eps * KCAL)[ti]
    start = np.concatenate([[0], np.cumsum(nptr)])
    cov = [(i, catm[k], cdip[k]) for i in range(3) for k in range(start[i], start[i + 1])]
    nc = netcdf_file(os.path.join(ref, "examples/water512/pgm3p25_512.rst7"), "r", mmap=False)
    xyz = np.array(nc.variables["coordinates"].data, float)
    # Amber stores velocities in A / (1/20.455 ps): the file's scale_factor (default 20.455) gives A/ps
    vel = np.array(nc.variables["velocities"].data, float) * float(getattr(nc.variables["velocities"], "scale_factor", 20.455))
    lengths = np.array(nc.variables["cell_lengths"].data, float)
    angles = np.array(nc.variables["cell_angles"].data, float)
    nc.close()
    return {
        "pos": xyz * ANG_NM,
        "vel": vel * ANG_NM,  # A/ps -> nm/ps
        "cell_lengths": lengths * ANG_NM,
        "cell_angles": angles,
        "q": q[:3],
        "radius": rad[:3],
        "alpha": alp[:3],
        "lj_rmin_half": rh[:3],
        "lj_sqrt_eps": se[:3],
        "masses": mass[:3],
        "cov_i": np.array([c[0] for c in cov], np.int32),
        "cov_j": np.array([c[1] for c in cov], np.int32),
        "cov_c": np.array([c[2] for c in cov], float),
        "all_q": q,
        "all_radius": rad,
        "all_alpha": alp,
    }


def pmemd_pbc(ref: str) -> dict:
    """Return pmemd-pgm's single point of the 512-water box (data/validation/amber_ref/pmemd_pbc)."""
    d = os.path.join(ref, "data/validation/amber_ref/pmemd_pbc")
    text = open(os.path.join(d, "mdout")).read()
    eelec = float(re.findall(r"EELEC\s*=\s*(-?\d+\.\d+)", text)[0])
    vdw = float(re.findall(r"VDWAALS\s*=\s*(-?\d+\.\d+)", text)[0])
    nc = netcdf_file(os.path.join(d, "mdfrc"), "r", mmap=False)
    frc = np.array(nc.variables["forces"].data, float)[0]
    nc.close()
    mu = []
    for line in open(os.path.join(d, "cpu_cg_iter_041.dat")):
        parts = line.split()
        try:
            vals = [float(v) for v in parts]
        except ValueError:
            continue
        if len(vals) == 5:  # atom, x, y, z, |x|
            mu.append(vals[1:4])
    return {"eelec_kcal": np.array(eelec), "vdw_kcal": np.array(vdw), "forces_kcal_A": frc, "mu_eA": np.array(mu)}


def md_rigid_points(gold: str) -> dict:
    """Snapshots (positions, box, induced dipoles, energies) of the md_rigid_nve / md_rigid_npt goldens."""
    out = {}
    for case, short in (("md_rigid_nve", "nve"), ("md_rigid_npt", "npt")):
        with np.load(os.path.join(gold, case + ".npz"), allow_pickle=False) as z:
            tags = sorted({k.split(".")[0] for k in z.files if k.startswith("t")})
            for t in tags:
                for k in ("pos", "box", "mu", "obs.epot", "obs.elec", "obs.vdw", "obs.cg_resid_max"):
                    out[f"{short}.{t}.{k}"] = z[f"{t}.{k}"]
    return out


def copy_water512_inputs(ref: str) -> None:
    """Copy the prmtop and NetCDF restart of examples/water512 into tests/golden/."""
    for name in ("pgm3p25_512.prmtop", "pgm3p25_512.rst7"):
        shutil.copyfile(os.path.join(ref, "examples/water512", name), os.path.join(HERE, name))


def main(ref: str) -> None:
    """Write every fixture into tests/golden/."""
    copy_water512_inputs(ref)
    np.savez_compressed(os.path.join(HERE, "water512.npz"), **water512(ref))
    gold = os.path.join(ref, "tests/regression/golden")
    for name in ("ff_small_box", "ff_pgm3p25_512", "periodic_model", "efield_point"):
        with np.load(os.path.join(gold, name + ".npz"), allow_pickle=False) as z:
            np.savez_compressed(os.path.join(HERE, name + ".npz"), **{k: z[k] for k in z.files})
    np.savez_compressed(os.path.join(HERE, "md_rigid_points.npz"), **md_rigid_points(gold))
    try:
        np.savez_compressed(os.path.join(HERE, "pmemd_pbc.npz"), **pmemd_pbc(ref))
    except (OSError, IndexError, KeyError) as exc:  # optional independent check
        print("pmemd_pbc fixture skipped:", exc)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
