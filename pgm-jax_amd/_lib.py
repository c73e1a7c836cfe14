"""ctypes binding of libpgm_hip.so (include/pgm_hip.h).

The library is loaded from this package directory.  There is no fallback: if it is missing or a
GPU is absent, calls raise.  Device pointers come from torch CUDA (HIP) tensors; torch is only the
device-memory and stream plumbing.
"""

from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpgm_hip.so")

PGM_OK = 0
PGM_ROW_OVERFLOW = 1
PGM_NOT_CONVERGED = 3

EXPORTS = (
    "pgm_last_error",
    "pgm_version",
    "pgm_ctx_create",
    "pgm_ctx_destroy",
    "pgm_set_stream",
    "pgm_set_box",
    "pgm_build_rows",
    "pgm_row_entries",
    "pgm_row_blocks",
    "pgm_perm_dipoles",
    "pgm_field",
    "pgm_pme_energy",
    "pgm_compute",
    "pgm_trial_energy",
    "pgm_bead_count",
    "pgm_bead_results",
    "pgm_set_graphs",
    "pgm_set_row_layout",
    "pgm_row_layout",
    "pgm_autotune",
    "pgm_set_pme_transform",
    "pgm_pme_transform",
    "pgm_set_row_fork",
    "pgm_energy_forces",
    "pgm_strain_derivative",
    "pgm_dipole_vjp",
    "pgm_compute_vjp",
    "pgm_set_params",
    "pgm_set_efield",
    "pgm_pme_spread",
    "pgm_pme_influence",
    "pgm_field_result",
    "pgm_solve_linear_induction",
    "pgm_solve_linear_induction_noalloc",
    "pgm_md_setup",
    "pgm_md_run",
    "pgm_md_set_thermostat",
    "pgm_md_rescale",
    "pgm_time_kernels",
    "pgm_set_timing",
    "pgm_kernel_timing",
    "pgm_guess_config",
    "pgm_guess_nparams",
    "pgm_guess_params",
    "pgm_guess_forward",
    "pgm_guess_grad",
    "pgm_guess_train",
    "pgm_guess_status",
    "pgm_copy_probe",
    "pgm_launch_floor",
)


class Topology(C.Structure):
    """pgm_topology."""

    _fields_ = [
        ("n_atoms", C.c_int),
        ("q", C.POINTER(C.c_double)),
        ("radius", C.POINTER(C.c_double)),
        ("alpha", C.POINTER(C.c_double)),
        ("lj_rmin_half", C.POINTER(C.c_double)),
        ("lj_sqrt_eps", C.POINTER(C.c_double)),
        ("mol", C.POINTER(C.c_int)),
        ("n_cov", C.c_int),
        ("cov_i", C.POINTER(C.c_int)),
        ("cov_j", C.POINTER(C.c_int)),
        ("cov_c", C.POINTER(C.c_double)),
    ]


class Settings(C.Structure):
    """pgm_settings."""

    _fields_ = [
        ("cutoff", C.c_double),
        ("skin", C.c_double),
        ("ewald_beta", C.c_double),
        ("grid", C.c_int * 3),
        ("order", C.c_int),
        ("lj_lrc", C.c_int),
        ("dipole_tol", C.c_double),
        ("max_iter", C.c_int),
        ("predictor", C.c_int),
        ("fused", C.c_int),
        ("norm_refresh", C.c_int),
        ("peek", C.c_double),
        ("precision", C.c_int),
        ("recip", C.c_int),
        ("k_tol", C.c_double),
        ("nbead", C.c_int),
        ("precond", C.c_int),
        ("elec_cutoff", C.c_double),
    ]


class Induction(C.Structure):
    """pgm_induction."""

    _fields_ = [("mu", C.c_void_p), ("hist", C.c_void_p), ("count", C.c_int), ("norm", C.c_double)]


class Energy(C.Structure):
    """pgm_energy."""

    _fields_ = [("elec", C.c_double), ("vdw", C.c_double), ("total", C.c_double)]


class MDState(C.Structure):
    """pgm_md_state."""

    _fields_ = [
        ("pos", C.c_void_p),
        ("vel", C.c_void_p),
        ("forces", C.c_void_p),
        ("ind", Induction),
        ("have_forces", C.c_int),
        ("step", C.c_longlong),
    ]


class MDReport(C.Structure):
    """pgm_md_report."""

    _fields_ = [
        ("epot", C.c_double),
        ("elec", C.c_double),
        ("vdw", C.c_double),
        ("ekin", C.c_double),
        ("etot", C.c_double),
        ("temperature", C.c_double),
        ("cg_iterations_total", C.c_longlong),
        ("cg_iterations_max", C.c_int),
        ("cg_iterations_last", C.c_int),
        ("cg_residual_max", C.c_double),
        ("row_rebuilds", C.c_int),
        ("heat", C.c_double),
    ]


class PGMError(RuntimeError):
    """An error status of the HIP library."""


_lib = None


def lib() -> C.CDLL:
    """Return the loaded library (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise PGMError(f"{LIB_PATH} not found: build it first (__graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    P, I, D = C.c_void_p, C.c_int, C.c_double
    L.pgm_last_error.restype = C.c_char_p
    L.pgm_version.restype = I
    L.pgm_ctx_create.argtypes = [C.POINTER(P), C.POINTER(Topology), C.POINTER(Settings), C.POINTER(D), P]
    L.pgm_ctx_destroy.argtypes = [P]
    L.pgm_set_stream.argtypes = [P, P]
    L.pgm_set_box.argtypes = [P, C.POINTER(D)]
    L.pgm_build_rows.argtypes = [P, P]
    L.pgm_row_entries.argtypes = [P]
    L.pgm_row_entries.restype = C.c_longlong
    L.pgm_row_blocks.argtypes = [P]
    L.pgm_row_blocks.restype = C.c_longlong
    L.pgm_copy_probe.argtypes = [P, P, C.c_size_t, P]
    L.pgm_launch_floor.argtypes = [P, I, C.POINTER(D)]
    L.pgm_perm_dipoles.argtypes = [P, P, P]
    L.pgm_field.argtypes = [P, P, I, P, P]
    L.pgm_pme_energy.argtypes = [P, P, I, P, C.POINTER(D), P]
    L.pgm_compute.argtypes = [P, P, C.POINTER(Induction), C.POINTER(Energy), P, C.POINTER(I), C.POINTER(D)]
    L.pgm_trial_energy.argtypes = [P, P, C.POINTER(Induction), C.POINTER(Energy), C.POINTER(I), C.POINTER(D)]
    L.pgm_bead_count.argtypes = [P]
    L.pgm_bead_results.argtypes = [P, P, P, P]
    L.pgm_set_graphs.argtypes = [P, I]
    L.pgm_set_row_layout.argtypes = [P, I, I]
    L.pgm_row_layout.argtypes = [P, C.POINTER(I), C.POINTER(I)]
    L.pgm_autotune.argtypes = [P, P, I, C.POINTER(D)]
    L.pgm_set_pme_transform.argtypes = [P, I]
    L.pgm_pme_transform.argtypes = [P, C.POINTER(I), C.POINTER(I)]
    L.pgm_set_row_fork.argtypes = [P, I]
    L.pgm_energy_forces.argtypes = [P, P, P, C.POINTER(Energy), P]
    L.pgm_strain_derivative.argtypes = [P, P, P, P, I, C.POINTER(Energy), P, C.POINTER(D)]
    L.pgm_dipole_vjp.argtypes = [P, P, P, P, D, P, P, P, P, P, P, C.POINTER(I), C.POINTER(D)]
    L.pgm_compute_vjp.argtypes = [P, P, P, D, P, P, D, P, P, P, P, P, P, C.POINTER(I), C.POINTER(D)]
    L.pgm_set_params.argtypes = [P, P, P, P, P, P, P]
    L.pgm_set_efield.argtypes = [P, P, P]
    L.pgm_pme_spread.argtypes = [P, P, I, P, P]
    L.pgm_pme_influence.argtypes = [P, P]
    L.pgm_field_result.argtypes = [P, P, P]
    L.pgm_solve_linear_induction.argtypes = [I, P, P, P, P, D, I, P, C.POINTER(I), C.POINTER(D)]
    L.pgm_solve_linear_induction_noalloc.argtypes = [I, P, P, P, P, D, I, P, C.POINTER(I), C.POINTER(D)]
    L.pgm_md_setup.argtypes = [P, C.POINTER(D), P]
    L.pgm_md_run.argtypes = [P, C.POINTER(MDState), I, D, C.POINTER(MDReport)]
    L.pgm_md_set_thermostat.argtypes = [P, D, D, C.c_ulonglong]
    L.pgm_time_kernels.argtypes = [P, P, I, C.POINTER(D)]
    L.pgm_md_rescale.argtypes = [P, C.POINTER(MDState), D]
    L.pgm_set_timing.argtypes = [P, I]
    L.pgm_guess_config.argtypes = [P, I, D, I, I, C.c_ulonglong]
    L.pgm_guess_nparams.argtypes = [P]
    L.pgm_guess_nparams.restype = C.c_longlong
    L.pgm_guess_params.argtypes = [P, P, C.c_longlong, I]
    L.pgm_guess_forward.argtypes = [P, P, I, P, I]
    L.pgm_guess_grad.argtypes = [P, P, P, I, P, C.POINTER(D)]
    L.pgm_guess_train.argtypes = [P, I, I, D, C.POINTER(D)]
    L.pgm_guess_status.argtypes = [P, C.POINTER(D)]
    L.pgm_kernel_timing.argtypes = [P, C.POINTER(D), C.POINTER(C.c_longlong)]
    for name in EXPORTS:
        if name not in ("pgm_last_error", "pgm_version", "pgm_row_entries", "pgm_row_blocks", "pgm_guess_nparams"):
            getattr(L, name).restype = I
    _lib = L
    return L


def check(status: int, what: str) -> int:
    """Raise PGMError for a negative status; return the status otherwise."""
    if status < 0:
        raise PGMError(f"{what} failed ({status}): {lib().pgm_last_error().decode(errors='replace')}")
    return status
