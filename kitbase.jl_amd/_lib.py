"""ctypes binding of libkbgpu.so (include/kbgpu.h).

The library is loaded from kitbase.jl_amd/lib/libkbgpu.so (built in-tree by build.py).  A missing
library or a missing GPU is an error: there is no CPU fallback behind these entry points.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libkbgpu.so"

KBG_OK = 0
KBG_FP32, KBG_FP64 = 0, 1
KBG_KFVS, KBG_KCU, KBG_UGKS = 0, 1, 2
KBG_BGK, KBG_SHAKHOV, KBG_FSM = 0, 1, 2
KBG_BC_FIX, KBG_BC_MAXWELL = 0, 1
KBG_VANLEER, KBG_MINMOD, KBG_SUPERBEE, KBG_VANALBABA, KBG_LINEAR = 0, 1, 2, 3, 4
KBG_EV_STEPS_BEGIN, KBG_EV_STEPS_END = 14, 15  # event slots around kbg_step / kbg_batch_step loops
KBG_HEALTH_NONFINITE, KBG_HEALTH_RHO, KBG_HEALTH_LAMBDA = 1, 2, 4  # kbg_solve_health flag bits

FLUX = {"kfvs": KBG_KFVS, "kcu": KBG_KCU, "ugks": KBG_UGKS}
COLL = {"bgk": KBG_BGK, "shakhov": KBG_SHAKHOV, "fsm": KBG_FSM}
BC = {"fix": KBG_BC_FIX, "maxwell": KBG_BC_MAXWELL}
LIMITER = {"vanleer": KBG_VANLEER, "minmod": KBG_MINMOD, "superbee": KBG_SUPERBEE,
           "vanalbaba": KBG_VANALBABA, "linear": KBG_LINEAR}

_dp = C.POINTER(C.c_double)


class KbgDesc(C.Structure):
    """Mirror of kbg_desc (include/kbgpu.h)."""

    _fields_ = [
        ("dim", C.c_int), ("nf", C.c_int), ("vdim", C.c_int), ("nv", C.c_int),
        ("nx", C.c_int), ("ny", C.c_int), ("ng", C.c_int), ("precision", C.c_int),
        ("flux", C.c_int), ("collision", C.c_int), ("limiter", C.c_int), ("interp_order", C.c_int),
        ("bc", C.c_int * 4),
        ("K", C.c_double), ("gamma", C.c_double), ("mu_ref", C.c_double), ("omega", C.c_double),
        ("Pr", C.c_double),
        ("u", _dp), ("v", _dp), ("w", _dp), ("weights", _dp),
        ("dx", _dp), ("dy", _dp), ("bc_prim", _dp),
        ("xgeo", _dp), ("ygeo", _dp), ("xwall", _dp), ("ywall", _dp),
        ("fsm_nm", C.c_int), ("fsm_shape", C.c_int * 3), ("fsm_kn", C.c_double),
        ("fsm_phi", _dp), ("fsm_psi", _dp), ("fsm_phipsi", _dp),
    ]


_ip = C.POINTER(C.c_int)


class KbgUDesc(C.Structure):
    """Mirror of kbg_udesc (include/kbgpu.h)."""

    _fields_ = [
        ("nf", C.c_int), ("nv", C.c_int), ("ncell", C.c_int), ("nface", C.c_int),
        ("precision", C.c_int), ("collision", C.c_int), ("interp_order", C.c_int),
        ("K", C.c_double), ("gamma", C.c_double), ("mu_ref", C.c_double), ("omega", C.c_double),
        ("Pr", C.c_double),
        ("u", _dp), ("v", _dp), ("weights", _dp),
        ("face_cells", _ip), ("face_kind", _ip), ("face_geo", _dp), ("face_wall", _dp),
        ("cell_faces", _ip), ("cell_dirc", _dp), ("cell_area", _dp), ("cell_type", _ip),
        ("cell_nb", _ip), ("cell_dmin", _dp),
    ]


# name -> (restype, argtypes)
_PROTOS = {
    "kbg_version": (C.c_int, []),
    "kbg_last_error": (C.c_char_p, []),
    "kbg_ctx_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "kbg_ctx_destroy": (C.c_int, [C.c_void_p]),
    "kbg_ctx_sync": (C.c_int, [C.c_void_p]),
    "kbg_ctx_set_schedule": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "kbg_ctx_get_schedule": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "kbg_ctx_set_graphs": (C.c_int, [C.c_void_p, C.c_int]),
    "kbg_ctx_set_sweep_steps": (C.c_int, [C.c_void_p, C.c_int]),
    "kbg_ctx_set_sweep_vw": (C.c_int, [C.c_void_p, C.c_int]),
    "kbg_ctx_event_record": (C.c_int, [C.c_void_p, C.c_int]),
    "kbg_ctx_event_elapsed": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_float)]),
    "kbg_solve_create": (C.c_int, [C.c_void_p, C.POINTER(KbgDesc), C.POINTER(C.c_void_p)]),
    "kbg_solve_destroy": (C.c_int, [C.c_void_p]),
    "kbg_solve_bytes": (C.c_size_t, [C.POINTER(KbgDesc)]),
    "kbg_usolve_create": (C.c_int, [C.c_void_p, C.POINTER(KbgUDesc), C.POINTER(C.c_void_p)]),
    "kbg_usolve_bytes": (C.c_size_t, [C.POINTER(KbgUDesc)]),
    "kbg_upload": (C.c_int, [C.c_void_p, _dp, _dp, _dp]),
    "kbg_download": (C.c_int, [C.c_void_p, _dp, _dp, _dp]),
    "kbg_upload_slopes": (C.c_int, [C.c_void_p, _dp, _dp]),
    "kbg_download_slopes": (C.c_int, [C.c_void_p, _dp, _dp]),
    "kbg_upload_faces": (C.c_int, [C.c_void_p, _dp, _dp]),
    "kbg_download_faces": (C.c_int, [C.c_void_p, _dp, _dp]),
    "kbg_reconstruct": (C.c_int, [C.c_void_p]),
    "kbg_evolve": (C.c_int, [C.c_void_p, C.c_double]),
    "kbg_update": (C.c_int, [C.c_void_p, C.c_double, _dp]),
    "kbg_timestep": (C.c_int, [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, _dp]),
    "kbg_step": (C.c_int, [C.c_void_p, C.c_double, C.c_int, _dp]),
    "kbg_batch_step": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.c_int, _dp, C.c_int, _dp]),
    "kbg_solve_health": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.c_int]),
    "kbg_batch_step_noalloc": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.c_int, _dp, C.c_int, _dp]),
    "kbg_batch_timestep": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.c_int, _dp, _dp, _dp, _dp, _dp, _dp]),
    "kbg_batch_step_cfl": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.c_int, _dp, _dp, _dp, _dp, _dp, C.c_int, _dp, _dp]),
}

EXPORTED = tuple(_PROTOS)

_lib = None


class KbgError(RuntimeError):
    pass


def load(path: str | os.PathLike | None = None) -> C.CDLL:
    """Load libkbgpu.so once; raise if it is missing (no fallback)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else Path(os.environ.get("KBG_LIBRARY", LIB_PATH))
    if not p.exists():
        raise KbgError(f"{p} not found: build it with `python kitbase.jl_amd/build.py`")
    lib = C.CDLL(str(p))
    for name, (res, args) in _PROTOS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def check(rc: int) -> None:
    if rc != KBG_OK:
        msg = load().kbg_last_error().decode(errors="replace")
        raise KbgError(f"kbgpu error {rc}: {msg}")


def dptr(a: np.ndarray | None):
    """Pointer to a C-contiguous float64 array (or NULL)."""
    if a is None:
        return None
    if a.dtype != np.float64 or not a.flags["C_CONTIGUOUS"]:
        raise TypeError("expected a C-contiguous float64 array")
    return a.ctypes.data_as(_dp)


def iptr(a: np.ndarray | None):
    """Pointer to a C-contiguous int32 array (or NULL)."""
    if a is None:
        return None
    if a.dtype != np.int32 or not a.flags["C_CONTIGUOUS"]:
        raise TypeError("expected a C-contiguous int32 array")
    return a.ctypes.data_as(_ip)
