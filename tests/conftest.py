"""Shared test setup: load the host package as `kitbase_amd` and the oracle binding."""
import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT / "kitbase.jl_amd"))
sys.path.insert(0, str(ROOT / "oracle"))

from loader import load_package  # noqa: E402

kb = load_package()
import oracle  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


@pytest.fixture(scope="session")
def kbpkg():
    return kb


@pytest.fixture(scope="session")
def orc():
    oracle.load()
    return oracle


def gpu_available() -> bool:
    try:
        import ctypes
        lib = kb._lib.load()
        h = ctypes.c_void_p()
        rc = lib.kbg_ctx_create(0, ctypes.byref(h))
        if rc == 0:
            lib.kbg_ctx_destroy(h)
            return True
    except Exception:
        return False
    return False


@pytest.fixture(scope="session")
def gpu_ctx():
    if os.environ.get("KBG_REQUIRE_GPU", "1") == "1":
        # a GPU test run without a usable device fails instead of skipping (no silent fallback)
        return kb.default_context()
    if not gpu_available():
        pytest.skip("no HIP device")
    return kb.default_context()
