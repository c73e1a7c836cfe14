"""Shared test setup: the `gpu` marker, import paths and the product package.

The product package lives in the directory ``pgm-jax_amd/`` and is imported as ``pgm_jax_amd``.
Tests marked ``gpu`` need an AMD GPU (MI355X, gfx950) and the built HIP library; without a GPU they
are skipped.
"""

from __future__ import annotations

import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def _load_package():
    """Import ``pgm-jax_amd/`` as the module ``pgm_jax_amd`` (the directory name has a hyphen)."""
    if "pgm_jax_amd" in sys.modules:
        return sys.modules["pgm_jax_amd"]
    pkg = os.path.join(ROOT, "pgm-jax_amd")
    spec = importlib.util.spec_from_file_location("pgm_jax_amd", os.path.join(pkg, "__init__.py"),
                                                  submodule_search_locations=[pkg])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["pgm_jax_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


if os.path.exists(os.path.join(ROOT, "pgm-jax_amd", "__init__.py")):
    _load_package()


def pytest_configure(config):
    """Register the markers."""
    config.addinivalue_line("markers", "gpu: needs an AMD GPU and the built HIP library")
    config.addinivalue_line("markers", "slow: longer CPU test")


def _has_gpu() -> bool:
    try:
        import torch

        return torch.cuda.is_available()
    except Exception:  # noqa: BLE001
        return False


def pytest_collection_modifyitems(config, items):
    """Skip `gpu` tests when no GPU is visible."""
    if _has_gpu():
        return
    skip = pytest.mark.skip(reason="no GPU visible")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)
