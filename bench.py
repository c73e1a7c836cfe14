#!/usr/bin/env python
"""Benchmark: ns/day and CG iterations per step of 32,768-water pGM (qpi) MD on one MI355X.

Workload (one GPU): 32,768 rigid pGM3P-25 waters (98,304 atoms) in a 9.93 nm cubic box (1.0
g/cm^3), seeded jittered lattice with random orientations, Maxwell-Boltzmann rigid-body velocities
at 298 K; qpi + LJ (long-range correction), cutoff 0.9 nm, Ewald beta 4.0/nm, PME 128^3 order 6,
CG tolerance 1e-7, NVE velocity Verlet + SHAKE/RATTLE (FP64), dt 1 fs.  A step is one MD step
(force field with the induced-dipole solve, constraints, integrator); inputs are resident in GPU
memory.  The lattice start is first equilibrated (untimed, --equil steps with the velocities
rescaled to 298 K every 10 steps); with the learned guess (--guess learned, the default) the
equilibration also harvests converged dipoles and trains the network.

`value` is the mixed-precision path (FP32 pair/PME kernels, FP64 accumulation, FP64 CG reductions)
with the learned CG initial guess and the molecule-block preconditioner (--precond; "jacobi" is the
reference's); beside it: the same mixed path from the previous-step guess, the plain FP64 path
(Jacobi, previous-step guess: the reference's algorithm; "fp64", the baseline of the metric), the energy /
force / dipole errors of mixed vs FP64 at the final configuration, the HBM roofline of the
dominant kernel (the CG matvec's row field) with the PME spread / gather beside it, from HIP events
attached to every launch of these kernels (hipExtLaunchKernel) in a second block of timed steps right
after the measured one (--timing-steps; the events cost about a tenth of the step time, so the block
that gives `value` runs without them), and the float64 C++/OpenMP restatement of the same algorithm on
every host core (cpu_baseline, oracle/cpu_ref.cpp) from the same equilibrated state.

Beads (--bead-evals > 0): 32 beads of 1,024 waters (3.94 nm box, PME 48^3, mixed, previous-step
guess), force evaluations of all beads in one context (every stage one launch for all beads,
bead-wise CG) against 32 single-bead contexts called one after the other, on the same sequence of
positions (every evaluation moves every atom by 2e-4 nm Gaussian noise, on the GPU).

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--guess learned|mu4|previous]
"""

from __future__ import annotations

import argparse
import importlib.util
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
METRIC = "ns/day + CG iters/step, 32k-water pGM qpi; learned-guess vs plain FP64 path"
HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, 8 TB/s (AMD Instinct MI355X data sheet)
FP32_PEAK_TFLOPS = 157.3  # MI355X vector FP32 (packed), the peak the pair kernels are held against


def _package():
    if "pgm_jax_amd" in sys.modules:
        return sys.modules["pgm_jax_amd"]
    pkg = os.path.join(ROOT, "pgm-jax_amd")
    spec = importlib.util.spec_from_file_location("pgm_jax_amd", os.path.join(pkg, "__init__.py"),
                                                  submodule_search_locations=[pkg])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["pgm_jax_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def log(*a) -> None:
    print("#", *a, file=sys.stderr, flush=True)


def ns_per_day(ms_per_step: float, dt_fs: float) -> float:
    return 86400.0 / (ms_per_step * 1e-3) * dt_fs * 1e-6


def _host_cores() -> int:
    """Cores this process may use: the affinity set, capped by OMP_NUM_THREADS when set."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    try:
        n = min(n, int(os.environ.get("OMP_NUM_THREADS", n)))
    except ValueError:
        pass
    return max(1, n)


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(pos, vel, masses, H, grid: int, steps: int = 6, timed: int = 2) -> dict:
    """Time the float64 C++/OpenMP restatement (oracle/cpu_ref.cpp) on every host core: `steps` NVE
    steps of the bench's own system from its equilibrated state; the first steps fill the mu4
    predictor's history (the first four solve from alpha b), the last `timed` are the sample."""
    sys.path.insert(0, ROOT)
    import numpy as np

    from oracle import systems
    from oracle import md as omd
    from oracle.cpu_ref import CpuRef
    from oracle.forcefield import Settings

    n = len(pos)
    top, _, _ = systems.water_box(n // 3, float(H[0, 0]), seed=3)  # the topology of n // 3 waters
    cores = _host_cores()
    c = CpuRef(top, H, Settings(cutoff=0.9, ewald_beta=4.0, pme_grid=(grid,) * 3, pme_order=6, lj_lrc=True, dipole_tol=1e-7,
                                max_iter=200), threads=cores)
    x = np.ascontiguousarray(pos, float).copy()
    v = np.ascontiguousarray(vel, float).copy()
    d2 = np.ascontiguousarray(omd.constraint_lengths2(x))
    t0 = time.perf_counter()
    F = np.ascontiguousarray(c.compute(x)["forces"])
    first = time.perf_counter() - t0
    per, its = [], []
    for _ in range(steps):
        t0 = time.perf_counter()
        rep = c.md(x, v, F, masses, d2, 1, 1e-3)
        per.append(time.perf_counter() - t0)
        its.append(rep["cg_iterations"])
    ms = 1e3 * float(np.mean(per[-timed:]))
    return {
        "value": round(ns_per_day(ms, 1.0), 6),
        "unit": "ns/day",
        "cores": c.threads,
        "kind": "port",
        "sample": (f"float64 C++/OpenMP restatement of the same algorithm (oracle/cpu_ref.cpp: cell-list rows, PME "
                   f"{grid}^3 order 6, Jacobi CG to 1e-7 with the mu4 predictor, SHAKE/RATTLE NVE 1 fs), {n // 3} waters from the "
                   f"bench's equilibrated state, {c.threads} threads on '{_cpu_model()}': {steps} steps after a first force "
                   f"evaluation ({first:.1f} s), CG iterations per step {its}; value from the mean of the last {timed} steps "
                   f"({ms / 1e3:.2f} s/step).  Reference's published CPU numbers (JAX CPU path): 3.14 ns/day on 32 cores for "
                   f"512 waters (docs/efield.md:261), 40.3 ms/step on 16 cores for 1,536 atoms (docs/interfaces.md:235)"),
        "seconds_per_step": [round(t, 3) for t in per],
        "cpu_model": _cpu_model(),
    }


def pmc_traffic(kernel: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC summary (profiles/), or None.

    The summary holds FETCH_SIZE and WRITE_SIZE per dispatch [KiB], each already divided by the
    ratio the same counter reads for a copy kernel of known size.
    """
    path = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    try:
        with open(path) as fh:
            d = json.load(fh)
        k = d["kernels"][kernel]
        return round((k["fetch_kib"] + k["write_kib"]) * 1024.0)
    except (OSError, KeyError, ValueError):
        return None


def roofline_of(live: dict, kern: dict, entries: int, blocks: int, layout: str, n: int, K3: int, fused: bool = False,
                lanes: int = 16, floor_us: float | None = None) -> dict:
    """HBM roofline of the dominant kernel (the CG matvec's row field) with the PME spread / gather beside it.

    Launch durations are the averages of the HIP event pairs attached to every launch of the row
    field, the spread kernel and the gather in the instrumented block of timed MD steps.

    Algorithmic bytes per launch (each logical read or write once):
      row field, molecule blocks (3-site molecules, the default for water): per record (molecule
                 pair with a pair inside the cutoff) the partner word (4 B) + 9 G1 + 9 G2 (FP32,
                 72 B); per atom its position (32 B) and dipole (16 B) read and the field written
                 (16 B);
      row field, atom rows: per row entry j (4 B) + G1, G2 (FP32, 8 B); per atom the dipole read
                 (16 B) and the field written (16 B);
      spread:    per atom positions 12 + q 4 + d 12 B, the FP32 grid written once (K^3 x 4 B), for the
                 tiled spread kernel (its blocks are merged by the forward planes of the fused
                 transform or by k_merge_table before the library transform);
      gather:    the FP32 grid read once (K^3 x 4 B), per atom d in 12 B and the field out 12 B;
      transform: (fused kernels, isolated timings) forward planes: the grid read and the half
                 spectrum (K^3/2 complex x 8 B) written; x transform: the spectrum read and written
                 and Gc read (K^3/2 x 4 B); inverse planes: the spectrum read and the grid written.

    Also the FP32 view of the row field ("fp32"): 66 flops per unique pair per matvec (the kernel
    evaluation, ~30, plus the matvec's 36; FMA = 2 flops), unique pairs = row entries / 2, against the
    vector FP32 peak.  The kernel evaluation runs once per geometry in the record pass, so the row
    field itself does the matvec's 36 ("matvec_part").
    """
    if layout == "molecule":
        row_bytes = blocks * 76 + n * 64
        kname, kdesc = "k_row_field_mol", "k_row_field_mol (CG matvec real-space rows, FP32, molecule blocks of 9 pairs)"
    else:
        row_bytes = entries * 12 + n * 32
        kname, kdesc = "k_row_field", "k_row_field (CG matvec real-space rows, FP32, sliced-ELL j/G1/G2 stream)"
    spread_bytes = n * 28 + K3 * 4
    gather_bytes = K3 * 4 + n * 24
    row_ms, spread_ms, gather_ms = live["row_field"]["ms"], live["spread"]["ms"], live["gather"]["ms"]

    def gbs(b, ms):
        return b / (ms * 1e-3) / 1e9

    out = {
        "kernel": kdesc,
        "bound": "hbm",
        "achieved": round(gbs(row_bytes, row_ms), 2),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(gbs(row_bytes, row_ms) / HBM_PEAK_GBS, 5),
        "traffic": pmc_traffic(kname if layout != "molecule" or lanes == 16 else f"{kname}_{lanes}_lanes"),
        "algorithmic_bytes": row_bytes,
        "ms_per_launch": round(row_ms, 5),
        "launches": live["row_field"]["launches"],
        "row_entries": entries,
        "row_blocks": blocks,
        "spread": {"kernels": "k_spread_tiles",
                   "achieved": round(gbs(spread_bytes, spread_ms), 2),
                   "frac": round(gbs(spread_bytes, spread_ms) / HBM_PEAK_GBS, 5), "algorithmic_bytes": spread_bytes,
                   "ms_per_launch": round(spread_ms, 5)},
        "gather": {"kernels": "k_gather_tiles", "achieved": round(gbs(gather_bytes, gather_ms), 2),
                   "frac": round(gbs(gather_bytes, gather_ms) / HBM_PEAK_GBS, 5),
                   "algorithmic_bytes": gather_bytes, "ms_per_launch": round(gather_ms, 5)},
        "isolated_ms": {k: round(v, 5) for k, v in kern.items()},
    }
    pairs = entries // 2
    tf = 66 * pairs / (row_ms * 1e-3) / 1e12
    tm = 36 * pairs / (row_ms * 1e-3) / 1e12
    out["fp32"] = {"flop_per_unique_pair": 66, "unique_pairs": pairs, "achieved": round(tf, 3), "peak": FP32_PEAK_TFLOPS,
                   "unit": "TFLOP/s", "frac": round(tf / FP32_PEAK_TFLOPS, 5),
                   "matvec_part": {"flop_per_unique_pair": 36, "achieved": round(tm, 3),
                                   "frac": round(tm / FP32_PEAK_TFLOPS, 5)}}
    if floor_us is not None:  # the spread / gather bytes at the launch floor: their HBM fraction's ceiling
        out["launch_floor_us"] = round(floor_us, 3)
        out["spread"]["ceiling_frac_at_launch_floor"] = round(gbs(spread_bytes, floor_us * 1e-3) / HBM_PEAK_GBS, 4)
        out["gather"]["ceiling_frac_at_launch_floor"] = round(gbs(gather_bytes, floor_us * 1e-3) / HBM_PEAK_GBS, 4)
    if fused:
        half = K3 // 2  # complex points of the half spectrum (to K^2 (K/2 + 1) ~ K^3 / 2)
        tb = {"fwd_planes": K3 * 4 + half * 8, "x_conv": 2 * half * 8 + half * 4, "bwd_planes": half * 8 + K3 * 4}
        t_ms = sum(kern[k] for k in tb)
        out["transform"] = {"kernels": "k_pme_fwd_plane + k_pme_x_conv + k_pme_bwd_plane (isolated)",
                            "achieved": round(gbs(sum(tb.values()), t_ms), 2),
                            "frac": round(gbs(sum(tb.values()), t_ms) / HBM_PEAK_GBS, 5),
                            "algorithmic_bytes": sum(tb.values()), "ms_per_launch": round(t_ms, 5)}
    return out


def ctypes_floor(P, reps: int = 2000) -> float:
    """Microseconds per back-to-back launch of an empty kernel on torch's stream (pgm_launch_floor)."""
    import ctypes

    import torch

    us = ctypes.c_double()
    P._lib.check(P._lib.lib().pgm_launch_floor(ctypes.c_void_p(torch.cuda.current_stream().cuda_stream), reps,
                                               ctypes.byref(us)), "pgm_launch_floor")
    return us.value


def bead_leg(P, evals: int, warmup: int, beads: int = 32) -> dict:
    """Force evaluations of `beads` beads of 1,024 waters: one bead context vs one context per bead."""
    import numpy as np
    import torch

    nw, edge, K = 1024, 3.94, 48
    qs, sys_, H = [], None, None
    for b in range(beads):
        sys_, pos, H = P.models.water_box(nw, edge, seed=100 + b)
        qs.append(pos)
    q0 = torch.as_tensor(np.stack(qs), device="cuda")
    s = P.MDSettings(precision="mixed", predictor="previous", cutoff=0.9, skin=0.1, ewald_beta=4.0, pme_grid=(K,) * 3,
                     pme_order=6, dipole_tol=1e-7, max_iter=200)
    gen = torch.Generator(device="cuda").manual_seed(5)
    moves = [torch.randn(q0.shape, generator=gen, device="cuda", dtype=torch.float64) * 2e-4 for _ in range(warmup + evals)]

    kern = {}

    def batched():
        pb = P.PGMBeads(sys_, H, s, beads=beads)
        log("beads autotune", json.dumps(pb.autotune(q0, reps=5)))
        e, q, its = pb.init_state(), q0.clone(), 0
        e = pb.compute(q, H, e)[2]  # first solve from alpha * b (untimed)
        for k in range(warmup + evals):
            if k == warmup:
                torch.cuda.synchronize()
                t0, its = time.perf_counter(), 0
            q += moves[k]
            _, U, e = pb.compute(q, H, e)
            its += e.iters
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / evals
        kern.update(pb.time_kernels(q, reps=20))  # one launch of each kernel for all beads
        return ms, its / evals

    def sequential():
        ffs = [P.PGMForceField(sys_, H, s) for _ in range(beads)]
        q = q0.clone()
        inds = [ff.compute(q[b], H).induction for b, ff in enumerate(ffs)]
        its = 0
        for k in range(warmup + evals):
            if k == warmup:
                torch.cuda.synchronize()
                t0, its = time.perf_counter(), 0
            q += moves[k]
            m = 0
            for b, ff in enumerate(ffs):
                r = ff.compute(q[b], H, None, inds[b])
                inds[b] = r.induction
                m = max(m, r.iterations)
            its += m
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3 / evals, its / evals

    ms_b, it_b = batched()
    log(f"beads: {beads} x {nw} waters, one context: {ms_b:.3f} ms per evaluation of all beads, {it_b:.2f} CG it (max over beads)")
    ms_s, it_s = sequential()
    log(f"beads: {beads} single-bead contexts in turn: {ms_s:.3f} ms per evaluation of all beads, {it_s:.2f} CG it (max)")
    n3 = 3 * nw
    spread_bytes = beads * (n3 * 28 + K**3 * 4)
    gather_bytes = beads * (K**3 * 4 + n3 * 24)
    sp, ga = kern.get("spread", 0.0), kern.get("gather", 0.0)
    return {"workload": f"{beads} beads x {nw} rigid pGM3P-25 waters ({3 * nw} atoms per bead), {edge} nm box, PME {K}^3 "
                        f"order 6, mixed, previous-step guess, CG tol 1e-7; force evaluation of every bead",
            "kernels_ms": {k: round(v, 5) for k, v in kern.items()},
            "spread_all_beads": {"algorithmic_bytes": spread_bytes, "ms_per_launch": round(sp, 5),
                                 "achieved": round(spread_bytes / (sp * 1e-3) / 1e9, 2) if sp else None,
                                 "frac": round(spread_bytes / (sp * 1e-3) / 1e9 / HBM_PEAK_GBS, 5) if sp else None},
            "gather_all_beads": {"algorithmic_bytes": gather_bytes, "ms_per_launch": round(ga, 5),
                                 "achieved": round(gather_bytes / (ga * 1e-3) / 1e9, 2) if ga else None,
                                 "frac": round(gather_bytes / (ga * 1e-3) / 1e9 / HBM_PEAK_GBS, 5) if ga else None},
            "evals": evals, "ms_per_eval_batched": round(ms_b, 4), "ms_per_eval_sequential": round(ms_s, 4),
            "speedup": round(ms_s / ms_b, 3), "cg_iters_batched": round(it_b, 3), "cg_iters_sequential": round(it_s, 3)}


def main() -> int:  # noqa: C901
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--fp64-steps", type=int, default=40)
    ap.add_argument("--waters", type=int, default=32768)
    ap.add_argument("--edge", type=float, default=9.93)
    ap.add_argument("--grid", type=int, default=128)
    ap.add_argument("--dt-fs", type=float, default=1.0)
    ap.add_argument("--equil", type=int, default=300, help="equilibration steps (velocity rescaling to 298 K), untimed")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--guess", default="learned", choices=["learned", "mu4", "previous"],
                    help="CG initial guess of the measured mixed path")
    ap.add_argument("--train-steps", type=int, default=300, help="Adam steps per 100 equilibration steps (learned)")
    ap.add_argument("--timing-steps", type=int, default=100,
                    help="MD steps after the timed ones with HIP events on the row field, spread and gather (roofline)")
    ap.add_argument("--prev-steps", type=int, default=30, help="timed steps of the mixed path from the previous-step guess")
    ap.add_argument("--bead-evals", type=int, default=10, help="timed force evaluations of the 32-bead leg (0: off)")
    ap.add_argument("--graph-steps", type=int, default=50,
                    help="timed steps of the measured context again with its CG iterations replayed as hipGraphs (0: off)")
    ap.add_argument("--precond", default="molecule", choices=["jacobi", "molecule"],
                    help="CG preconditioner of the mixed paths (the FP64 baseline keeps the reference's Jacobi)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run the library's copy kernel 3 times over MIB MiB (HBM counter calibration) and exit")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {a.gpus}); this build has no multi-GPU path", file=sys.stderr)
        return 2
    import numpy as np
    import torch

    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        return 1
    P = _package()
    if a.copy_probe > 0:
        import ctypes

        nbytes = a.copy_probe << 20
        src = torch.zeros(nbytes // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        L = P._lib.lib()
        for _ in range(3):
            P._lib.check(L.pgm_copy_probe(ctypes.c_void_p(dst.data_ptr()), ctypes.c_void_p(src.data_ptr()), nbytes,
                                          ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)), "pgm_copy_probe")
        torch.cuda.synchronize()
        log(f"copy probe: 3 x {a.copy_probe} MiB read + written")
        return 0
    log("building system", a.waters, "waters")
    sys_, pos, H = P.models.water_box(a.waters, a.edge, seed=3)
    vel = P.models.maxwell_boltzmann_rigid(pos, sys_.masses, 298.0, seed=4)
    base = dict(cutoff=0.9, skin=0.1, ewald_beta=4.0, pme_grid=(a.grid,) * 3, pme_order=6, lj_lrc=True,
                dipole_tol=1e-7, max_iter=200)

    # Equilibrate once (mixed precision, velocity rescaling to 298 K) from the lattice start.  With the
    # learned guess, this run harvests (features, converged dipoles) every 10 solves and trains the
    # network between blocks of steps (untimed), so the timed run starts with a trained guess.
    ffe = P.PGMForceField(sys_, H, P.MDSettings(precision="mixed", predictor=a.guess, precond=a.precond, **base))
    if a.guess == "learned":
        ffe.guess_config(hidden_layers=3, harvest_every=10, seed=11)
    mde = P.RigidWaterMD(ffe, pos, vel)
    if a.equil > 0:
        log("equilibrating", a.equil, "steps at 298 K")
        done = 0
        while done < a.equil:
            k = min(100, a.equil - done)
            r = mde.equilibrate(k, 298.0, a.dt_fs)
            done += k
            if a.guess == "learned" and ffe.guess_status()["rows"] > 0:
                loss = ffe.guess_train(steps=a.train_steps, batch=4096, lr=1e-3)
                log(f"  step {done}: guess training loss {loss:.4f}, status {json.dumps(ffe.guess_status())}")
        log(f"equilibrated: T {r['temperature']:.1f} K, Epot {r['epot']:.1f}")
    pos = mde.pos.cpu().numpy()
    vel = mde.vel.cpu().numpy()

    def run(precision: str, steps: int, warmup: int, timing_steps: int = 0, predictor: str = "previous",
            ff=None, md=None, precond: str = "jacobi"):
        if ff is None:
            ff = P.PGMForceField(sys_, H, P.MDSettings(precision=precision, predictor=predictor, precond=precond, **base))
            md = P.RigidWaterMD(ff, pos, vel)
        tune = ff.autotune(md.pos, reps=10)  # row-field layout and spread tiles for this context (untimed)
        log(precision, predictor, "autotune", json.dumps(tune))
        rep0 = md.run(warmup, a.dt_fs)
        torch.cuda.synchronize()
        log(precision, predictor, "timed", steps)
        t0 = time.perf_counter()
        rep = md.run(steps, a.dt_fs)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / steps
        log(precision, f"{ms:.3f} ms/step, {rep['cg_iterations'] / steps:.2f} CG it/step, T {rep['temperature']:.1f} K, "
            f"Etot {rep['etot']:.3f}, rebuilds {rep['row_rebuilds']}")
        if timing_steps > 0:  # the same trajectory on, with the kernels' event pairs
            ff.set_timing(True)
            t0 = time.perf_counter()
            rt = md.run(timing_steps, a.dt_fs)
            torch.cuda.synchronize()
            ms_t = (time.perf_counter() - t0) * 1e3 / timing_steps
            rep["kernel_timing"] = ff.kernel_timing()
            ff.set_timing(False)
            rep["instrumented"] = {"steps": timing_steps, "ms_per_step": round(ms_t, 4),
                                   "cg_iters_per_step": round(rt["cg_iterations"] / timing_steps, 3)}
            log(precision, f"instrumented: {ms_t:.3f} ms/step, {rt['cg_iterations'] / timing_steps:.2f} CG it/step")
        rep["autotune"] = tune
        rep["etot_start"] = rep0["etot"]
        return ff, md, rep, ms

    # the measured path: mixed precision with the guess (the equilibrated context keeps the trained network)
    ffm, mdm, repm, ms_m = run("mixed", a.steps, a.warmup, max(1, a.timing_steps), predictor=a.guess, ff=ffe, md=mde)
    guess_status = ffm.guess_status() if a.guess == "learned" else None
    guess_label = a.guess
    if guess_status:
        log("learned guess after the timed steps:", json.dumps(guess_status))
        guess_label = f"learned ({guess_status['mode']} at the end of the timed steps)"
    graphs = None
    if a.graph_steps > 0:  # the same context and trajectory, CG iterations as hipGraphs (no per-kernel events)
        ffm.set_graphs(True)
        mdm.run(3, a.dt_fs)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        repg = mdm.run(a.graph_steps, a.dt_fs)
        torch.cuda.synchronize()
        ms_g = (time.perf_counter() - t0) * 1e3 / a.graph_steps
        ffm.set_graphs(False)
        graphs = {"ms_per_step": round(ms_g, 4), "value": round(ns_per_day(ms_g, a.dt_fs), 3), "steps": a.graph_steps,
                  "cg_iters_per_step": round(repg["cg_iterations"] / a.graph_steps, 3)}
        log(f"mixed with CG hipGraphs: {ms_g:.3f} ms/step, {repg['cg_iterations'] / a.graph_steps:.2f} CG it/step")
    live = repm["kernel_timing"]
    log("live kernel timing over the instrumented steps:", json.dumps(live))
    kern = mdm.time_kernels(reps=50)
    entries = ffm.row_entries()
    layout, lanes = ffm.row_layout()
    blocks = ffm.row_blocks()
    log("kernels (ms):", json.dumps({k: round(v, 5) for k, v in kern.items()}), "row entries", entries)
    prev = None
    if a.guess != "previous" and a.prev_steps > 0:  # the same mixed path from the previous-step guess
        _, _, repp, ms_p = run("mixed", a.prev_steps, max(2, a.warmup // 4), predictor="previous", precond=a.precond)
        prev = {"ms_per_step": round(ms_p, 4), "value": round(ns_per_day(ms_p, a.dt_fs), 3), "steps": a.prev_steps,
                "cg_iters_per_step": round(repp["cg_iterations"] / a.prev_steps, 3)}
    errors, fp64 = None, None
    if a.fp64_steps > 0:
        ffd, mdd, repd, ms_d = run("double", a.fp64_steps, max(2, a.warmup // 4))
        # mixed vs FP64 at the final mixed configuration (same positions, fresh solves)
        x = mdm.pos
        rm = ffm.compute(x)
        rd = ffd.compute(x)
        Fm, Fd = rm.forces.cpu().numpy(), rd.forces.cpu().numpy()
        mum, mud = rm.induction.mu.cpu().numpy(), rd.induction.mu.cpu().numpy()
        errors = {
            "energy_rel": abs(rm.energy["total"] - rd.energy["total"]) / abs(rd.energy["total"]),
            "force_rms_rel": float(np.sqrt(np.mean((Fm - Fd) ** 2)) / np.sqrt(np.mean(Fd**2))),
            "dipole_rms_rel": float(np.sqrt(np.mean((mum - mud) ** 2)) / np.sqrt(np.mean(mud**2))),
            "cg_residual_mixed": rm.residual,
            "cg_residual_fp64": rd.residual,
        }
        log("errors vs fp64:", json.dumps(errors))
        fp64 = {"value": round(ns_per_day(ms_d, a.dt_fs), 3), "ms_per_step": round(ms_d, 4), "steps": a.fp64_steps,
                "cg_iters_per_step": round(repd["cg_iterations"] / a.fp64_steps, 3),
                "speedup_mixed_vs_fp64": round(ms_d / ms_m, 3)}

    beads = bead_leg(P, a.bead_evals, 3) if a.bead_evals > 0 else None
    floor_us = ctypes_floor(P)
    log(f"launch floor: {floor_us:.2f} us per empty-kernel launch")
    n = 3 * a.waters
    K3 = a.grid**3
    roofline = roofline_of(live, kern, entries, blocks, layout, n, K3, ffm.pme_transform()[0] == "fused", lanes, floor_us)
    roofline["events_block"] = repm["instrumented"]  # the MD steps the launch durations come from
    value = ns_per_day(ms_m, a.dt_fs)
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "ns/day",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_m, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": (f"{a.waters} rigid pGM3P-25 waters ({n} atoms), {a.edge} nm cubic box, qpi+LJ(lrc), rc 0.9 nm, "
                         f"beta 4.0/nm, PME {a.grid}^3 order 6, CG tol 1e-7 from the {guess_label} guess ({a.precond} "
                         f"preconditioner), NVE "
                         f"dt {a.dt_fs} fs, FP32 kernels with FP64 accumulation; baseline: the FP64 path from the "
                         f"previous-step guess"),
            "waters": a.waters, "atoms": n, "pme_grid": a.grid, "dt_fs": a.dt_fs, "guess": a.guess, "precond": a.precond,

        },
        "cg_iters_per_step": round(repm["cg_iterations"] / a.steps, 3),
        "autotune": repm["autotune"],
        "learned_guess": guess_status,
        "mixed_previous_guess": prev,
        "mixed_cg_graphs": graphs,
        "fp64": fp64,
        "errors_vs_fp64": None if errors is None else {k: float(f"{v:.3e}") for k, v in errors.items()},
        "energy_drift_check": {"etot_start_kJmol": repm["etot_start"], "etot_end_kJmol": repm["etot"],
                               "drift_kJmol": repm["etot"] - repm["etot_start"],
                               "drift_per_ekin": (repm["etot"] - repm["etot_start"]) / repm["ekin"],
                               "steps": a.steps, "ekin_kJmol": repm["ekin"], "temperature_K": round(repm["temperature"], 2)},
        "kernels_ms": {k: round(v, 5) for k, v in kern.items()},
        "beads": beads,
        "roofline": roofline,
    }
    if not a.no_cpu_baseline:
        log("cpu baseline (C++/OpenMP float64 restatement, all host cores)")
        out["cpu_baseline"] = cpu_baseline(pos, vel, sys_.masses, H, a.grid)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
