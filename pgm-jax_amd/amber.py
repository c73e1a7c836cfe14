"""Amber inputs of the MD runner: pGM prmtop and coordinate / restart files.

Mirrors the readers the reference's `pgm-jax md` uses for its inputs:

* `read_prmtop` / `read_prmtop_pgm` follow pgm_jax/prmtop.py:125-200 (sections by %FLAG, values by
  the %FORMAT field width) and pgm_jax/param.py:113-233 (one Molecule per residue: POL_GAUSS_*
  charges, Gaussian radii and polarizabilities, covalent dipoles, LJ R*/2 and sqrt(eps) from the
  type-pair A/B tables, masses);
* `read_coordinates` / `read_coordinates_nm` follow pgm_jax/md/io.py:30-106 (NetCDF restart by its
  magic number, else ASCII inpcrd / restart in 12-character fields; velocities in Amber's unit
  A / (1/20.455 ps)) and `box_from_cell` pgm_jax/md/box.py:291-314.

Units: prmtop lengths are Angstrom, dipoles e A, polarizabilities A^3, LJ kcal/mol; Molecule values
are nm, e, e nm, nm^3, R*/2 nm and sqrt(kJ/mol).  Only text and NetCDF-3 readers are used: nothing
in the files is executed.
"""

from __future__ import annotations

import re

import numpy as np

from .system import Molecule

ANG_NM = 0.1
KCAL = 4.184  # kJ per kcal
AMBER_VEL = 20.455  # Amber velocity unit A / (1/20.455 ps): v [A/ps] = v_file * 20.455
_FORMAT = re.compile(r"\((\d*)([aAiIeEfF])(\d+)")


def read_prmtop(path: str) -> dict:
    """Return {flag: list of str fields} of an Amber prmtop.

    Every %FLAG section is split by the field width of its %FORMAT line (e.g. 5E16.8, 10I8, 20a4);
    blank fields are dropped.  Raises ValueError for a section without a usable %FORMAT.
    """
    out, flag, width = {}, None, None
    with open(path) as fh:
        for line in fh:
            line = line.rstrip("\n")
            if line.startswith("%FLAG"):
                flag, width = line.split()[1], None
                out[flag] = []
            elif line.startswith("%FORMAT"):
                m = _FORMAT.search(line)
                if m is None:
                    raise ValueError(f"{path}: unreadable %FORMAT for {flag}: {line!r}")
                width = int(m.group(3))
            elif line.startswith("%"):
                continue
            elif flag is not None:
                if width is None:
                    raise ValueError(f"{path}: section {flag} has no %FORMAT")
                out[flag] += [line[k:k + width].strip() for k in range(0, len(line), width) if line[k:k + width].strip()]
    return out


def _lj_per_type(s: dict) -> tuple[np.ndarray, np.ndarray]:
    """Return per-atom (R*/2 [nm], sqrt(eps) [sqrt(kJ/mol)]) from the LJ A/B tables (param.py:281-307).

    Types with A = B = 0 have no LJ.  Pairs of different types must follow Lorentz-Berthelot
    (R*_ij = R*_i/2 + R*_j/2, eps_ij = sqrt(eps_i eps_j)); otherwise ValueError.
    """
    ntypes = int(s["POINTERS"][1])
    ti = np.array(s["ATOM_TYPE_INDEX"], int) - 1
    nbi = np.array(s["NONBONDED_PARM_INDEX"], int).reshape(ntypes, ntypes) - 1
    A = np.array(s["LENNARD_JONES_ACOEF"], float)[nbi]
    B = np.array(s["LENNARD_JONES_BCOEF"], float)[nbi]
    ok = (A > 0) & (B > 0)
    rmin = np.where(ok, (2.0 * np.where(ok, A, 1.0) / np.where(ok, B, 1.0)) ** (1.0 / 6.0), 0.0)  # A
    eps = np.where(ok, B**2 / (4.0 * np.where(ok, A, 1.0)), 0.0)  # kcal/mol
    rh, ee = np.diag(rmin) / 2.0, np.diag(eps)
    lb_r = rh[:, None] + rh[None, :]
    lb_e = np.sqrt(ee[:, None] * ee[None, :])
    if not (np.allclose(np.where(ok, rmin, lb_r), lb_r, rtol=1e-5, atol=1e-8)
            and np.allclose(np.where(ok, eps, lb_e), lb_e, rtol=1e-5, atol=1e-12)):
        raise ValueError("prmtop LJ tables are not Lorentz-Berthelot (NBFIX pairs are not supported)")
    return (rh * ANG_NM)[ti], np.sqrt(ee * KCAL)[ti]


def read_prmtop_pgm(path: str, first_residue_only: bool = True) -> list:
    """Return the molecules (one per residue, in prmtop order) of an Amber pGM prmtop (param.py:113-233).

    Raises ValueError if the prmtop has no POL_GAUSS_* sections or a covalent dipole crosses
    residues; NotImplementedError for extra points (atom type EP), which the HIP path does not model.
    """
    s = read_prmtop(path)
    if "POL_GAUSS_MONOPOLES_LIST" not in s:
        raise ValueError(f"{path} has no pGM sections (POL_GAUSS_*)")
    names = s["ATOM_NAME"]
    if any(t.upper() == "EP" for t in s.get("AMBER_ATOM_TYPE", [])):
        raise NotImplementedError(f"{path}: extra points (type EP) are not supported by the HIP path")
    res_ptr = [int(x) - 1 for x in s["RESIDUE_POINTER"]] + [len(names)]
    res_lab = s["RESIDUE_LABEL"]
    q = np.array(s["POL_GAUSS_MONOPOLES_LIST"], float)
    rad = np.array(s["POL_GAUSS_RADII_LIST"], float) * ANG_NM
    alp = np.array(s["POL_GAUSS_POLARIZABILITY_LIST"], float) * ANG_NM**3
    nptr = np.array(s["POL_GAUSS_COVALENT_POINTERS_LIST"], int)
    catm = np.array(s["POL_GAUSS_COVALENT_ATOMS_LIST"], int) - 1
    cdip = np.array(s["POL_GAUSS_COVALENT_DIPOLES_LIST"], float) * ANG_NM
    start = np.concatenate([[0], np.cumsum(nptr)])
    mass = np.array(s["MASS"], float)
    rh, se = _lj_per_type(s)
    z = s.get("ATOMIC_NUMBER")
    zel = {1: "H", 6: "C", 7: "N", 8: "O", 9: "F", 15: "P", 16: "S", 17: "Cl"}
    # bonds (the symmetry classes of the parameter tying): atom pointers are 3 (i - 1) in the prmtop
    bl = [int(v) for k in ("BONDS_INC_HYDROGEN", "BONDS_WITHOUT_HYDROGEN") for v in s.get(k, [])]
    bond_pairs = [(bl[t] // 3, bl[t + 1] // 3) for t in range(0, len(bl) - 2, 3)]
    mols = []
    for r in range(1 if first_residue_only else len(res_lab)):
        a0, a1 = res_ptr[r], res_ptr[r + 1]
        cov = []
        for i in range(a0, a1):
            for k in range(start[i], start[i + 1]):
                if not a0 <= catm[k] < a1:
                    raise ValueError(f"atom {i + 1} has a covalent dipole to atom {catm[k] + 1} of another residue")
                cov.append((i - a0, int(catm[k]) - a0, float(cdip[k])))
        if z is not None:
            el = [zel.get(int(v), "X") for v in z[a0:a1]]
        else:
            el = [re.sub(r"\d+", "", n)[:1].upper() for n in names[a0:a1]]
        types = [str(t) for t in s["AMBER_ATOM_TYPE"][a0:a1]] if "AMBER_ATOM_TYPE" in s else None
        bonds = [(i - a0, j - a0) for i, j in bond_pairs if a0 <= i < a1 and a0 <= j < a1]
        mols.append(Molecule(res_lab[r], el, q[a0:a1].copy(), rad[a0:a1].copy(), alp[a0:a1].copy(), cov,
                             rh[a0:a1].copy(), se[a0:a1].copy(), masses=mass[a0:a1].copy(), types=types, bonds=bonds))
    return mols


def box_from_cell(lengths, angles) -> np.ndarray:
    """Return the lower-triangular box (lattice vectors as rows) of cell lengths and angles [deg]
    (a along x, b in the xy plane; pgm_jax/md/box.py:291-314)."""
    a, b, c = (float(v) for v in lengths)
    al, be, ga = np.radians(np.asarray(angles, float))
    cx = c * np.cos(be)
    cy = c * (np.cos(al) - np.cos(be) * np.cos(ga)) / np.sin(ga)
    return np.array([[a, 0.0, 0.0], [b * np.cos(ga), b * np.sin(ga), 0.0], [cx, cy, np.sqrt(c * c - cx * cx - cy * cy)]])


def read_coordinates(path: str) -> tuple:
    """Return (xyz [A] (N,3), velocities [A/ps] (N,3) or None, (lengths [A], angles [deg]) or None)
    of an Amber NetCDF restart or ASCII inpcrd / restart (md/io.py:30-80)."""
    with open(path, "rb") as fh:
        magic = fh.read(4)
    if magic[:3] == b"CDF":
        from scipy.io import netcdf_file

        f = netcdf_file(path, "r", mmap=False)
        try:
            v = f.variables
            xyz = np.array(v["coordinates"][:], float).reshape(-1, 3)
            vel = None
            if "velocities" in v:
                vel = np.array(v["velocities"][:], float).reshape(-1, 3) * float(getattr(v["velocities"], "scale_factor", AMBER_VEL))
            box = None
            if "cell_lengths" in v:
                box = (np.array(v["cell_lengths"][:], float).reshape(-1)[:3], np.array(v["cell_angles"][:], float).reshape(-1)[:3])
        finally:
            f.close()
        return xyz, vel, box
    lines = open(path).read().splitlines()
    n = int(lines[1].split()[0])
    vals = [float(ln[k:k + 12]) for ln in lines[2:] for k in range(0, len(ln.rstrip()), 12) if ln[k:k + 12].strip()]
    xyz = np.array(vals[:3 * n]).reshape(n, 3)
    rest = vals[3 * n:]
    vel = None
    if len(rest) >= 3 * n:
        vel = np.array(rest[:3 * n]).reshape(n, 3) * AMBER_VEL
        rest = rest[3 * n:]
    box = (np.array(rest[:3]), np.array(rest[3:6])) if len(rest) >= 6 else None
    return xyz, vel, box


def read_coordinates_nm(path: str) -> tuple:
    """Return (positions [nm], velocities [nm/ps] or None, box (3,3) [nm] or None) (md/io.py:83-106)."""
    xyz, vel, box = read_coordinates(path)
    H = None if box is None else box_from_cell(*box) * ANG_NM
    return xyz * ANG_NM, None if vel is None else vel * ANG_NM, H


__all__ = ["read_prmtop", "read_prmtop_pgm", "read_coordinates", "read_coordinates_nm", "box_from_cell"]
