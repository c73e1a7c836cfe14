"""The `pgm-jax md` run on the MI355X: Amber inputs, rigid-water MD on the GPU, a log table.

Mirrors the reference's command (pgm_jax/cli/main.py:35 -> scripts/md/run_md.py) for what the HIP
path covers: a pGM prmtop of rigid 3-site molecules and an Amber restart with a box, Amber-style
options in Angstrom (--cut, --skin, --ew-coeff [1/A], --pme-spacing / --nfft, --order, --vdwmeth),
NVE (--thermostat none) or Langevin (--thermostat langevin, --gamma, --temp), optionally the Monte
Carlo barostat (--barostat mc: isotropic volume moves every --barostat-interval steps, accepted by
the Metropolis rule at --press bar, molecules moved with their centres of mass; the trial energy
is pgm_trial_energy, the reference's PGMForceField.energy), and a log file <out>.log with one row
every --report-every steps in the reference's format (pgm_jax/md/driver.py:190-305 `format_row`,
`LogTable`; columns of pgm_jax/md/engine.py:173-231 `observables`, plus the speed in ns/day).

    python pgm-jax_amd/run_md.py -p water.prmtop -c water.rst7 -o runs/md --thermostat none \
        --barostat none --nsteps 10 --report-every 5 --dt-fs 1.0 --precision double

Defaults are the reference command's (cut 9 A, skin 1 A, ew-coeff 0.4/A, PME spacing 0.5 A, order 8,
LJ long-range correction, dipole tolerance 1e-5, 50 CG iterations, mu4 predictor, peek 0.65).  The
input velocities are projected onto the rigid-body motions (the RATTLE velocity projection, which
keeps each molecule's linear and angular momentum).
"""

from __future__ import annotations

import argparse
import math
import sys
import time

import numpy as np

from . import amber
from .md import Langevin, RigidWaterMD
from .settings import MDSettings, grid_size
from .system import System

KB = 0.0083144626181532  # kJ/mol/K
AMU_NM3_TO_G_CM3 = 1.66053906660e-3  # 1 amu/nm^3 in g/cm^3
BAR_NM3 = 0.0602214076  # 1 bar nm^3 in kJ/mol


def format_header(columns) -> str:
    """``#`` and the column names right-aligned in 14 characters (pgm_jax/md/driver.py:216-229)."""
    return "# " + " ".join(f"{c:>14s}" for c in columns)


def format_row(values) -> str:
    """One log row: floats %14.6f, integers and booleans %14d, else right-aligned strings, indented by
    two spaces under the header's ``# `` (pgm_jax/md/driver.py:190-214)."""
    out = []
    for v in values:
        if isinstance(v, (bool, np.bool_)):
            out.append(f"{int(v):14d}")
        elif isinstance(v, (float, np.floating)):
            out.append(f"{float(v):14.6f}")
        elif isinstance(v, (int, np.integer)):
            out.append(f"{int(v):14d}")
        else:
            out.append(f"{v!s:>14s}")
    return "  " + " ".join(out)


class LogTable:
    """A text table: optional title lines and a header once, then one fixed-width row per report
    (pgm_jax/md/driver.py:233-305).  The first row's keys are the columns; rows may also be echoed."""

    def __init__(self, path: str | None, append: bool = False, title=(), echo=None) -> None:
        self.path, self.echo = path, echo
        self.fh = open(path, "a" if append else "w") if path is not None else None
        self._needs_header = self.fh is not None and self.fh.tell() == 0
        self.title = [str(t) for t in title]
        self.columns = None

    def write(self, row: dict) -> str:
        """Append one row (the first also writes the title and header to a new file); return its text."""
        if self.columns is None:
            self.columns = list(row)
            header = format_header(self.columns)
            if self._needs_header:
                self.fh.write("".join(f"# {t}\n" for t in self.title) + header + "\n")
            if self.echo is not None:
                print(header, file=self.echo, flush=True)
        line = format_row(row[c] for c in self.columns)
        if self.fh is not None:
            self.fh.write(line + "\n")
            self.fh.flush()
        if self.echo is not None:
            print(line, file=self.echo, flush=True)
        return line

    def close(self) -> None:
        if self.fh is not None:
            self.fh.close()

    def __enter__(self) -> LogTable:
        return self

    def __exit__(self, *exc) -> None:
        self.close()


def rigid_velocities(pos: np.ndarray, vel: np.ndarray, masses: np.ndarray) -> np.ndarray:
    """Project velocities of consecutive 3-site rigid molecules onto rigid-body motions: per molecule
    the three bond constraints d_k . (v_a - v_b) = 0 are imposed by central impulses (RATTLE's
    velocity step), which keep the molecule's linear and angular momentum."""
    x = pos.reshape(-1, 3, 3)
    v = vel.reshape(-1, 3, 3).copy()
    w = 1.0 / masses.reshape(-1, 3)
    pairs = ((0, 1), (0, 2), (1, 2))
    d = np.stack([x[:, a] - x[:, b] for a, b in pairs], 1)  # (M, 3 constraints, 3)
    # impulse lam_k on a along +d_k, on b along -d_k: g_k = d_k . (v_a - v_b) changes by sum_l K_kl lam_l
    K = np.zeros((len(x), 3, 3))
    for k, (a, b) in enumerate(pairs):
        for m, (c, e) in enumerate(pairs):
            coef = ((a == c) - (a == e)) * w[:, a] - ((b == c) - (b == e)) * w[:, b]
            K[:, k, m] = coef * np.einsum("mi,mi->m", d[:, k], d[:, m])
    g = np.stack([np.einsum("mi,mi->m", d[:, k], v[:, a] - v[:, b]) for k, (a, b) in enumerate(pairs)], 1)
    lam = np.linalg.solve(K, -g[:, :, None])[:, :, 0]
    for k, (a, b) in enumerate(pairs):
        v[:, a] += (lam[:, k] * w[:, a])[:, None] * d[:, k]
        v[:, b] -= (lam[:, k] * w[:, b])[:, None] * d[:, k]
    return v.reshape(-1, 3)


def build_parser() -> argparse.ArgumentParser:
    """The options of `pgm-jax md` that the HIP path implements (Amber units, see the module docstring)."""
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("-p", "--prmtop", required=True, help="Amber pGM prmtop of rigid 3-site molecules")
    ap.add_argument("-c", "--coords", required=True, help="inpcrd / rst7 (ASCII or NetCDF) with a box")
    ap.add_argument("-o", "--out", default="md", help="output prefix: <out>.log")
    ap.add_argument("--nsteps", type=int, default=10000)
    ap.add_argument("--dt-fs", type=float, default=1.0)
    ap.add_argument("--report-every", type=int, default=1000)
    ap.add_argument("--thermostat", default="langevin", choices=["none", "langevin"])
    ap.add_argument("--barostat", default="none", choices=["none", "mc"])
    ap.add_argument("--temp", type=float, default=298.0, help="temperature [K] (temp0)")
    ap.add_argument("--gamma", type=float, default=2.0, help="Langevin friction [1/ps] (gamma_ln)")
    ap.add_argument("--press", type=float, default=1.0, help="pressure [bar] (pres0)")
    ap.add_argument("--barostat-interval", type=int, default=100, help="steps between MC volume moves (mcbarint)")
    ap.add_argument("--max-volume-change", type=float, default=0.01, help="largest relative MC volume move")
    ap.add_argument("--cut", type=float, default=9.0, help="cutoff [A]")
    ap.add_argument("--skin", type=float, default=1.0, help="neighbour-list skin [A]")
    ap.add_argument("--ew-coeff", type=float, default=0.4, help="Ewald coefficient [1/A]")
    ap.add_argument("--nfft", type=int, nargs=3, help="PME grid; default from --pme-spacing")
    ap.add_argument("--pme-spacing", type=float, default=0.5, help="PME grid spacing [A]")
    ap.add_argument("--order", type=int, default=8, help="PME B-spline order")
    ap.add_argument("--vdwmeth", type=int, default=1, choices=[0, 1], help="1: LJ long-range correction")
    ap.add_argument("--dipole-tol", type=float, default=1e-5)
    ap.add_argument("--max-iter", type=int, default=50)
    ap.add_argument("--peek", type=float, default=0.65)
    ap.add_argument("--predictor", default="mu4", choices=["mu4", "mu3", "previous", "none", "learned"])
    ap.add_argument("--precond", default="jacobi", choices=["jacobi", "molecule"])
    ap.add_argument("--precision", default="mixed", choices=["mixed", "double"])
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-velocities", action="store_true", help="ignore the file's velocities (Maxwell-Boltzmann at --temp)")
    return ap


def settings_from_args(a: argparse.Namespace) -> MDSettings:
    """MDSettings of the Amber-style options (Angstrom -> nm; scripts/md/run_md.py:200-240)."""
    return MDSettings(cutoff=a.cut / 10, skin=a.skin / 10, ewald_beta=a.ew_coeff * 10,
                      pme_grid=tuple(a.nfft) if a.nfft else None, pme_spacing=a.pme_spacing / 10, pme_order=a.order,
                      lj_lrc=bool(a.vdwmeth), dipole_tol=a.dipole_tol, max_iter=a.max_iter, peek=a.peek,
                      predictor=a.predictor, precision=a.precision, precond=a.precond)


def _com_scale(pos: np.ndarray, masses: np.ndarray, s: float) -> np.ndarray:
    """Positions with every 3-site molecule's centre of mass scaled by s (shapes kept)."""
    x = pos.reshape(-1, 3, 3)
    m = masses.reshape(-1, 3)
    com = (x * m[:, :, None]).sum(1) / m.sum(1)[:, None]
    return (x + (s - 1.0) * com[:, None, :]).reshape(-1, 3)


def run(argv=None, echo=sys.stdout) -> list:
    """Run `pgm-jax md` with argv (see build_parser); return the log rows as dicts."""
    import torch

    a = build_parser().parse_args(argv)
    mols = amber.read_prmtop_pgm(a.prmtop, first_residue_only=False)
    sys_ = System(mols)
    pos, vel, H = amber.read_coordinates_nm(a.coords)
    if H is None:
        raise SystemExit(f"{a.coords} has no box: the MD run is periodic")
    if len(pos) != sys_.n:
        raise SystemExit(f"{a.coords} has {len(pos)} atoms, {a.prmtop} {sys_.n}")
    st = settings_from_args(a)
    if st.pme_grid is None:
        st = st.replace(pme_grid=grid_size(H, st.pme_spacing))
    if vel is None or a.no_velocities:
        from .models import maxwell_boltzmann_rigid

        vel = maxwell_boltzmann_rigid(pos, sys_.masses, a.temp, seed=a.seed)
    else:
        vel = rigid_velocities(pos, vel, sys_.masses)
    from .forcefield import PGMForceField

    ff = PGMForceField(sys_, H, st)
    if a.predictor == "learned":
        ff.guess_config(hidden_layers=3, harvest_every=10, seed=a.seed)
    thermo = Langevin(a.gamma) if a.thermostat == "langevin" else None
    md = RigidWaterMD(ff, pos, vel, thermostat=thermo, temperature=a.temp, seed=a.seed)
    mass_total = float(np.sum(sys_.masses))
    rng = np.random.default_rng(a.seed + 7)
    title = [f"pgm-jax md on the MI355X: {a.prmtop} {a.coords}", f"{sys_.n} atoms, PME {st.pme_grid} order {st.pme_order}, "
             f"{a.precision}, thermostat {a.thermostat}, barostat {a.barostat}, dt {a.dt_fs} fs"]
    rows, done, cg_total, heat, tries, accepted = [], 0, 0, 0.0, 0, 0
    Hcur = np.array(H, float)
    with LogTable(a.out + ".log", title=title, echo=echo) as log:
        t_last = time.perf_counter()
        steps_since = 0
        while done < a.nsteps:
            k = min(a.report_every - done % a.report_every, a.nsteps - done)
            if a.barostat == "mc":
                k = min(k, a.barostat_interval - done % a.barostat_interval)
            rep = md.run(k, a.dt_fs)
            done += k
            steps_since += k
            cg_total += rep["cg_iterations"]
            heat += rep["heat"]
            if a.barostat == "mc" and done % a.barostat_interval == 0:
                tries += 1
                accepted += _mc_volume_move(ff, md, Hcur, rep["epot"], a, rng)
                Hcur = ff._H.copy()
            if done % a.report_every == 0 or done == a.nsteps:
                torch.cuda.synchronize()
                now = time.perf_counter()
                V = abs(float(np.linalg.det(Hcur)))
                row = {
                    "step": done, "time_ps": done * a.dt_fs * 1e-3, "temp_K": rep["temperature"], "etot": rep["etot"],
                    "ekin": rep["ekin"], "epot": rep["epot"], "elec": rep["elec"], "vdw": rep["vdw"],
                    "econs": rep["etot"] - heat, "volume_nm3": V, "density_g_cm3": mass_total / V * AMU_NM3_TO_G_CM3,
                    "cg_iter": int(rep["cg_iter_last"]), "cg_iter_max": int(rep["cg_iterations_max"]),
                    "cg_resid_max": rep["cg_residual_max"], "cg_mean": cg_total / max(done, 1),
                    "ns_day": 86400.0 / ((now - t_last) / steps_since) * a.dt_fs * 1e-6,
                }
                if a.barostat == "mc":
                    row["mc_accept"] = accepted / max(tries, 1)
                rows.append(row)
                log.write(row)
                t_last, steps_since = now, 0
    return rows


def _mc_volume_move(ff, md, H: np.ndarray, epot: float, a: argparse.Namespace, rng) -> int:
    """One isotropic Monte Carlo volume move (molecular scaling); return 1 if accepted.

    Acceptance min(1, exp(-(dE + P dV - N_mol kT ln(V'/V)) / kT)) with the trial energy of the
    scaled positions from the current dipoles (pgm_trial_energy); a rejected move leaves the
    state, box and neighbour rows as they were.
    """
    import torch

    V = abs(float(np.linalg.det(H)))
    dV = (2.0 * rng.random() - 1.0) * a.max_volume_change * V
    s = ((V + dV) / V) ** (1.0 / 3.0)
    H2 = H * s
    x0 = md.pos.cpu().numpy()
    x2 = _com_scale(x0, ff.sys.masses, s)
    e2, ind2, _, _ = ff.energy(x2, H2, None, md.ind)
    kT = KB * a.temp
    nmol = ff.sys.nmol
    w = (e2 - epot) + a.press * BAR_NM3 * dV - nmol * kT * math.log((V + dV) / V)
    if w <= 0.0 or rng.random() < math.exp(-w / kT):
        md.pos.copy_(torch.as_tensor(x2, device="cuda"))
        md.ind.mu.copy_(ind2.mu)
        md.invalidate_forces()
        return 1
    ff._box(H)  # back to the current box (the trial switched the context to H2)
    return 0


def main(argv=None) -> int:
    run(argv)
    return 0


__all__ = ["LogTable", "format_row", "format_header", "rigid_velocities", "build_parser", "settings_from_args", "run"]
