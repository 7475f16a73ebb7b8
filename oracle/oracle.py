"""ctypes binding of the FP64 CPU oracle (oracle/liboracle.so, built from kitbase_oracle.c).

TEST INFRASTRUCTURE ONLY: the HIP path is tested against it.  Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg may import this module.  The oracle restates KitBase.jl v0.9.31 (Julia,
which is not installed here and cannot be built or run), function by function; the citations are
in kitbase_oracle.h / kitbase_oracle.c.
"""
from __future__ import annotations

import ctypes as C
import subprocess
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"

_dp = C.POINTER(C.c_double)
_lib = None


def build() -> Path:
    subprocess.run(["make", "-C", str(HERE), "-s"], check=True)
    return LIB


def load() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        _lib = C.CDLL(str(LIB))
        for name in ("orc_heaviside", "orc_fortsign", "orc_newton_cotes", "orc_ref_vhs_vis",
                     "orc_vhs_collision_time", "orc_heat_capacity_ratio", "orc_sound_speed"):
            getattr(_lib, name).restype = C.c_double
        _lib.orc_limiter.restype = C.c_double
        _lib.orc_limiter.argtypes = [C.c_int, C.c_double, C.c_double]
        _lib.orc_heaviside.argtypes = [C.c_double]
        _lib.orc_fortsign.argtypes = [C.c_double, C.c_double]
        _lib.orc_newton_cotes.argtypes = [C.c_int, C.c_int]
        _lib.orc_ref_vhs_vis.argtypes = [C.c_double] * 3
        _lib.orc_vhs_collision_time.argtypes = [C.c_double] * 4
    return _lib


def p(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"], "oracle arrays must be C-contiguous float64"
    return a.ctypes.data_as(_dp)


def vec(x, n=None):
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float64))
    return a


class Solver1DStruct(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ng", C.c_int), ("nf", C.c_int), ("vdim", C.c_int), ("nv", C.c_int),
        ("flux", C.c_int), ("coll", C.c_int), ("limiter", C.c_int), ("interp_order", C.c_int),
        ("bc", C.c_int * 2),
        ("K", C.c_double), ("gam", C.c_double), ("mu", C.c_double), ("om", C.c_double), ("Pr", C.c_double),
        ("u", _dp), ("v", _dp), ("wv", _dp), ("weights", _dp), ("dx", _dp), ("bc_prim", _dp * 2),
        ("w", _dp), ("prim", _dp), ("f", _dp), ("sw", _dp), ("sf", _dp), ("fw", _dp), ("ff", _dp),
        ("fsm_nm", C.c_int), ("fsm_nu", C.c_int), ("fsm_nv", C.c_int), ("fsm_nw", C.c_int),
        ("fsm_kn", C.c_double), ("fsm_phi", _dp), ("fsm_psi", _dp), ("fsm_phipsi", _dp),
    ]


class Solver2DStruct(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nv", C.c_int),
        ("flux", C.c_int), ("coll", C.c_int), ("limiter", C.c_int), ("interp_order", C.c_int),
        ("bc", C.c_int * 4),
        ("K", C.c_double), ("gam", C.c_double), ("mu", C.c_double), ("om", C.c_double), ("Pr", C.c_double),
        ("u", _dp), ("v", _dp), ("weights", _dp), ("dx", _dp), ("dy", _dp),
        ("xn", _dp), ("xlen", _dp), ("xwall", _dp), ("yn", _dp), ("ylen", _dp), ("ywall", _dp),
        ("w", _dp), ("prim", _dp), ("f", _dp), ("sw", _dp), ("sf", _dp),
        ("xfw", _dp), ("xff", _dp), ("yfw", _dp), ("yff", _dp),
        ("nf", C.c_int),
    ]


class SolverUSStruct(C.Structure):
    _fields_ = [
        ("ncell", C.c_int), ("nface", C.c_int), ("nv", C.c_int), ("nf", C.c_int), ("coll", C.c_int),
        ("K", C.c_double), ("gam", C.c_double), ("mu", C.c_double), ("om", C.c_double), ("Pr", C.c_double),
        ("u", _dp), ("v", _dp), ("weights", _dp),
        ("fc", C.POINTER(C.c_int)), ("fkind", C.POINTER(C.c_int)), ("cf", C.POINTER(C.c_int)),
        ("ctype", C.POINTER(C.c_int)), ("nb", C.POINTER(C.c_int)),
        ("fn", _dp), ("flen", _dp), ("fwall", _dp), ("dirc", _dp), ("area", _dp),
        ("w", _dp), ("prim", _dp), ("f", _dp), ("fw", _dp), ("ff", _dp),
    ]


FLUX = {"kfvs": 0, "kcu": 1, "ugks": 2}
COLL = {"bgk": 0, "shakhov": 1, "fsm": 2}
BC = {"fix": 0, "maxwell": 1}
LIM = {"vanleer": 0, "minmod": 1, "superbee": 2, "vanalbaba": 3, "linear": 4}


class OracleSolve:
    """The oracle's view of a SolverSet plus host arrays (modified in place)."""

    def __init__(self, ks, ctr, face):
        self.lib = load()
        self.ks, self.ctr, self.face = ks, ctr, face
        self.keep = []
        g = ks.gas
        bcs = ks.boundaries()

        def k(a):
            a = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
            self.keep.append(a)
            return p(a)

        if ks.dim == 1:
            S = Solver1DStruct()
            ps, vs = ks.ps, ks.vs
            S.nx, S.ng, S.nf, S.vdim, S.nv = ps.nx, ps.ng, ks.nf, ks.vdim, vs.nv
            S.flux, S.coll, S.limiter, S.interp_order = FLUX[ks.set.flux], COLL[ks.set.collision], LIM[ks.set.limiter], ks.set.interpOrder
            S.bc[0], S.bc[1] = BC[bcs[0]], BC[bcs[1]]
            S.K, S.gam, S.mu, S.om, S.Pr = g.K, g.gamma, g.mu_ref, g.omega, g.Pr
            S.u, S.weights = k(vs.u), k(vs.weights)
            if ks.vdim == 3:
                S.v, S.wv = k(vs.v), k(vs.w)
            S.dx = k(ps.dx)
            S.bc_prim[0], S.bc_prim[1] = k(ks.ib.bc(ps.x0)), k(ks.ib.bc(ps.x1))
            S.w, S.prim, S.f, S.sw, S.sf = p(ctr.w), p(ctr.prim), p(ctr.f), p(ctr.sw), p(ctr.sf)
            S.fw, S.ff = p(face.fw), p(face.ff)
            if ks.set.collision == "fsm":
                fk = ks.gas.fsm
                S.fsm_nm, (S.fsm_nu, S.fsm_nv, S.fsm_nw) = fk.nm, vs.shape
                S.fsm_kn, S.fsm_phi, S.fsm_psi, S.fsm_phipsi = fk.Kn, k(fk.phi), k(fk.psi), k(fk.phipsi)
        else:
            kb = sys.modules["kitbase_amd"]  # the host mirror supplies the face geometry
            S = Solver2DStruct()
            ps, vs = ks.ps, ks.vs
            S.nx, S.ny, S.nv, S.nf = ps.nx, ps.ny, vs.nv, ks.nf
            S.flux, S.coll, S.limiter, S.interp_order = FLUX[ks.set.flux], COLL[ks.set.collision], LIM[ks.set.limiter], ks.set.interpOrder
            for i in range(4):
                S.bc[i] = BC[bcs[i]]
            S.K, S.gam, S.mu, S.om, S.Pr = g.K, g.gamma, g.mu_ref, g.omega, g.Pr
            S.u, S.v, S.weights = k(vs.u), k(vs.v), k(vs.weights)
            S.dx, S.dy = k(ps.dx), k(ps.dy)
            xg, yg = kb.geometry.face_geometry_2d(ps, ks.nf)
            xw, yw = kb.solver.wall_prims_2d(ks)
            S.xn, S.xlen, S.xwall = k(xg[..., 0:2]), k(xg[..., 2]), k(xw)
            S.yn, S.ylen, S.ywall = k(yg[..., 0:2]), k(yg[..., 2]), k(yw)
            S.w, S.prim, S.f, S.sw, S.sf = p(ctr.w), p(ctr.prim), p(ctr.f), p(ctr.sw), p(ctr.sf)
            S.xfw, S.xff, S.yfw, S.yff = p(face.fw), p(face.ff), p(face.yfw), p(face.yff)
        self.S = S
        self.dim = ks.dim

    def _call(self, name, *args):
        rc = getattr(self.lib, f"orc_{name}_{self.dim}d")(C.byref(self.S), *args)
        if rc != 0:
            raise RuntimeError(f"oracle {name} returned {rc}")

    def reconstruct(self):
        self._call("reconstruct")

    def evolve(self, dt):
        self._call("evolve", C.c_double(dt))

    def update(self, dt, res=None):
        r = np.zeros(5) if res is None else res
        self._call("update", C.c_double(dt), p(r))
        return r

    def run(self, dt, nsteps, res=None):
        r = np.zeros(5) if res is None else res
        self._call("run", C.c_double(dt), C.c_int(nsteps), p(r))
        return r


def set_threads(n: int):
    load().orc_set_num_threads(int(n))


def num_threads() -> int:
    return int(load().orc_num_threads())


class OracleSolveUS:
    """The oracle's unstructured triangle solve (orc_run_us) over a mesh from kitbase_amd.mesh; the
    state (an UnstructState) is modified in place.  It restates KitBase's interpOrder-1 scheme (the
    cells' slopes zero), so SolverSets of another order are refused."""

    def __init__(self, ks, ps, st, fwall=None, fkind=None):
        if ks.set.interpOrder != 1:
            raise ValueError("the oracle's unstructured step restates interpOrder 1 only")
        kb = sys.modules["kitbase_amd"]
        self.lib = load()
        self.keep = []
        g = ks.gas

        def k(a, t=np.float64):
            a = np.ascontiguousarray(np.asarray(a, dtype=t))
            self.keep.append(a)
            return a.ctypes.data_as(C.POINTER(C.c_int)) if t == np.int32 else p(a)

        fn, fl = kb.mesh.face_geometry(ps)
        S = SolverUSStruct()
        S.ncell, S.nface, S.nv, S.nf, S.coll = ps.ncell, ps.nface, ks.vs.nv, ks.nf, COLL[ks.set.collision]
        S.K, S.gam, S.mu, S.om, S.Pr = g.K, g.gamma, g.mu_ref, g.omega, g.Pr
        S.u, S.v, S.weights = k(ks.vs.u.ravel()), k(ks.vs.v.ravel()), k(ks.vs.weights.ravel())
        S.fc, S.fkind = k(ps.faceCells, np.int32), k(kb.mesh.face_kinds(ps) if fkind is None else fkind, np.int32)
        S.cf, S.ctype, S.nb = k(ps.cellFaces, np.int32), k(ps.cellType, np.int32), k(ps.cellNeighbors, np.int32)
        S.fn, S.flen, S.fwall = k(fn), k(fl), k(kb.mesh.wall_prims(ks, ps) if fwall is None else fwall)
        S.dirc, S.area = k(kb.mesh.cell_dirc(ps, fn)), k(ps.cellArea)
        S.w, S.prim, S.f, S.fw, S.ff = p(st.w), p(st.prim), p(st.f), p(st.fw), p(st.ff)
        self.S, self.st = S, st

    def run(self, dt, nsteps, res=None):
        r = np.zeros(4) if res is None else res
        rc = self.lib.orc_run_us(C.byref(self.S), C.c_double(dt), C.c_int(nsteps), p(r))
        if rc != 0:
            raise RuntimeError(f"oracle orc_run_us returned {rc}")
        return r

    def evolve(self, dt):
        rc = self.lib.orc_evolve_us(C.byref(self.S), C.c_double(dt))
        if rc != 0:
            raise RuntimeError(f"oracle orc_evolve_us returned {rc}")

    def update(self, dt, res=None):
        r = np.zeros(4) if res is None else res
        rc = self.lib.orc_update_us(C.byref(self.S), C.c_double(dt), p(r))
        if rc != 0:
            raise RuntimeError(f"oracle orc_update_us returned {rc}")
        return r
