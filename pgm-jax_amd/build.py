"""Build libpgm_hip.so (the HIP kernels + C ABI) in-tree for gfx950 with hipcc.

    python -c "import runpy; runpy.run_path('pgm-jax_amd/build.py', run_name='__main__')"

The library lands next to this file (pgm-jax_amd/libpgm_hip.so), where `_lib.py` loads it.
"""

from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libpgm_hip.so")
SOURCES = [os.path.join(HERE, "csrc", "context.hip")]
HEADERS = ["common.h", "pairs.h", "pme.h", "fft.h", "cg.h", "md.h", "virial.h", "ewald.h", "dense.h", "mlp.h", "guess.h", "adjoint.h"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # no packed FP32 (v_pk_*_f32) code generation
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def command(out: str = LIB) -> list:
    """Return the hipcc command line that builds the library."""
    inc = ["-I", os.path.join(ROOT, "include"), "-I", os.path.join(HERE, "csrc")]
    return [HIPCC, *FLAGS, *inc, *SOURCES, "-o", out, "-lrocfft"]


def build(verbose: bool = False) -> str:
    """Compile the library (always) and return its path; raises CalledProcessError on failure."""
    tmp = LIB + ".tmp"
    cmd = command(tmp)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise subprocess.CalledProcessError(res.returncode, cmd, res.stdout, res.stderr)
    if verbose:
        sys.stderr.write(res.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
