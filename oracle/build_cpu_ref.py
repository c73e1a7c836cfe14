"""Build oracle/libcpu_ref.so, the C++/OpenMP float64 restatement (test infrastructure and CPU baseline).

    python oracle/build_cpu_ref.py

g++ -O3 for x86-64-v3 (AVX2 + FMA, present on every current server CPU) with OpenMP; the library is
loaded only by the tests and by bench.py's cpu_baseline leg, never by the product.
"""

from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "cpu_ref.cpp")
LIB = os.path.join(HERE, "libcpu_ref.so")


def build() -> str:
    """Compile the library (always); raises CalledProcessError on failure."""
    tmp = LIB + ".tmp"
    cmd = ["g++", "-O3", "-march=x86-64-v3", "-fopenmp", "-std=c++17", "-fPIC", "-shared", SRC, "-o", tmp]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise subprocess.CalledProcessError(res.returncode, cmd, res.stdout, res.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build())
