"""The C ABI library loads and exports every function include/pgm_hip.h declares (CPU, no GPU calls)."""

import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pgm_hip.h")
LIB = os.path.join(ROOT, "pgm-jax_amd", "libpgm_hip.so")


def declared() -> list:
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+char\s*\*|int|long long|void)\s*\**\s*(pgm_\w+)\s*\(", text, re.M)))


def test_header_declares_the_entry_points():
    names = declared()
    for n in ("pgm_ctx_create", "pgm_ctx_destroy", "pgm_compute", "pgm_field", "pgm_build_rows", "pgm_energy_forces"):
        assert n in names


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    lib.pgm_version.restype = ctypes.c_int
    assert lib.pgm_version() >= 1


def test_python_binding_lists_every_symbol():
    import pgm_jax_amd._lib as L

    assert sorted(L.EXPORTS) == declared()
