"""pgm-jax_amd: the pGM electrostatics hot path of pgm-jax, MI355X-native (HIP for gfx950).

Drop-in for the force-field entry points of the reference (pgm_jax/md/forcefield.py
PGMForceField.compute / energy_fixed_mu / _field, pgm_jax/md/pme.py) on the MI355X: real-space
pair rows of Gaussian charges and dipoles, smooth PME (order 4..8) around rocFFT, and the
induced-dipole conjugate gradients with FP64 reductions, in FP64 ("double") or FP32 kernels with
FP64 accumulation ("mixed").  The kernels live in libpgm_hip.so (csrc/, C ABI in include/pgm_hip.h),
called through ctypes; there is no CPU fallback.
"""

from .settings import MDSettings, elec_cutoff_settings, ewald_beta_for, grid_size
from .system import Molecule, System
from .forcefield import InductionState, PGMForceField, Result
from ._lib import PGMError

__all__ = ["MDSettings", "elec_cutoff_settings", "ewald_beta_for", "grid_size", "Molecule", "System", "PGMForceField", "InductionState", "Result", "PGMError"]
from .md import Langevin, RigidWaterMD  # noqa: E402
from .periodic import PeriodicModel, PeriodicPGM, pressure_bar  # noqa: E402
from .beads import PGMBeads, PGMBeadState  # noqa: E402
from .pme import PME  # noqa: E402
from . import amber, driver, models, solver  # noqa: E402

__all__ += ["Langevin", "RigidWaterMD", "PeriodicModel", "PeriodicPGM", "pressure_bar", "PGMBeads", "PGMBeadState", "PME",
            "amber", "driver", "models", "solver"]
