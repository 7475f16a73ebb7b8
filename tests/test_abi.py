"""The C ABI (include/kbgpu.h) on the CPU: libkbgpu.so loads and exports every declared entry point,
the Python binding declares each of them, and the descriptor mirror matches the header's field
order.  No compute calls here (no GPU in this environment)."""
import ctypes as C
import re
from pathlib import Path

import pytest

from conftest import kb

ROOT = Path(__file__).resolve().parents[1]
HEADER = (ROOT / "include" / "kbgpu.h").read_text()


def declared_functions():
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_]+\s*\**\s*(kbg_[a-z0-9_]+)\s*\(", HEADER, re.M)))


def test_header_declares_the_entry_points():
    names = declared_functions()
    for expect in ("kbg_ctx_create", "kbg_solve_create", "kbg_upload", "kbg_download", "kbg_reconstruct", "kbg_evolve",
                   "kbg_update", "kbg_step", "kbg_batch_step", "kbg_last_error"):
        assert expect in names


def test_library_exports_every_declared_symbol():
    lib = kb._lib.load()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_the_header():
    assert set(declared_functions()) == set(kb._lib.EXPORTED)


@pytest.mark.parametrize("name,mirror", [("kbg_desc", "KbgDesc"), ("kbg_udesc", "KbgUDesc")])
def test_desc_mirror_field_order(name, mirror):
    body = HEADER[HEADER.index(f"typedef struct {name} "):]
    body = body[: body.index(f"}} {name};")]
    fields = re.findall(r"([a-zA-Z_][a-zA-Z0-9_]*)\s*(?:\[\d+\])?\s*[,;]", re.sub(r"/\*.*?\*/", "", body, flags=re.S))
    assert fields == [f[0] for f in getattr(kb._lib, mirror)._fields_]


def test_no_gpu_means_an_error_not_a_fallback():
    """Without a device kbg_ctx_create fails with KBG_ERR_HIP; nothing falls back to the CPU."""
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    lib = kb._lib.load()
    h = C.c_void_p()
    assert lib.kbg_ctx_create(0, C.byref(h)) != 0
    assert b"device" in lib.kbg_last_error()


def test_version_and_footprint_without_gpu():
    lib = kb._lib.load()
    assert lib.kbg_version() >= 10000
    ks = kb.SolverSet(kb.configs.sod_1d1f1v())
    n32, n64 = kb.solve_bytes(ks, "fp32"), kb.solve_bytes(ks, "fp64")
    assert n64 > n32 > 2 * 100 * 64 * 4
