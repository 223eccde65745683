"""Host-side state arrays and ctypes mirrors of the C structs in include/pgdg.h."""
from __future__ import annotations

import ctypes as C

import numpy as np

PGDG_MAX_PHASE = 3
SEM_T2, SEM_S3, SEM_UX = 2, 3, 4
BC_INTERIOR, BC_WALL, BC_INLET, BC_OUTLET, BC_SYMMETRY = 0, 1, -1, -2, -3


class Case(C.Structure):
    _fields_ = [
        ("dim", C.c_int), ("nphase", C.c_int), ("order", C.c_int), ("rk_stages", C.c_int),
        ("semantics", C.c_int), ("n", C.c_int * 3), ("bc", C.c_int * 6),
        ("box_l", C.c_double * 3), ("box_r", C.c_double * 3),
        ("p_sz", C.c_double), ("p_g", C.c_double),
        ("gamma", C.c_double * PGDG_MAX_PHASE), ("pinf", C.c_double * PGDG_MAX_PHASE),
        ("ini_rho", C.c_double * PGDG_MAX_PHASE),
        ("inlet_rho", C.c_double), ("inlet_u", C.c_double * 3), ("inlet_p", C.c_double),
        ("p2g_cut", C.c_double), ("output_dt", C.c_double), ("max_v0", C.c_double),
        ("precision", C.c_int),
        ("cfl", C.c_double), ("ppc_min", C.c_int), ("ppc_max", C.c_int),
    ]

    @property
    def U(self):
        return self.dim + 2

    @property
    def M(self):
        return self.dim + 1 if self.order == 1 else (6 if self.dim == 2 else 10)

    @property
    def nc(self):
        return self.n[0] * self.n[1] * self.n[2]


_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class Particles(C.Structure):
    _fields_ = [("n", C.c_int64), ("cap", C.c_int64),
                ("x", _dp), ("v", _dp), ("tv", _dp), ("rhov", _dp),
                ("vol", _dp), ("mass", _dp), ("h", _dp), ("P", _dp), ("rho", _dp),
                ("E", _dp), ("c", _dp), ("coef", _dp), ("phase", _ip), ("mark_in", _ip)]


class Walls(C.Structure):
    _fields_ = [("n", C.c_int64), ("x", _dp), ("norm", _dp), ("vol", _dp)]


class Cells(C.Structure):
    _fields_ = [("coef", _dp), ("coef0", _dp), ("dcoef", _dp), ("lim", _dp),
                ("alph", _dp), ("rho", _dp), ("E", _dp), ("c", _dp), ("ac", _dp),
                ("rhov", _dp), ("ph_state", _ip), ("ph_state_b", _ip),
                ("v", _dp), ("a", _dp), ("P", _dp), ("G_v", _dp)]


class Scalars(C.Structure):
    _fields_ = [("dt", C.c_double), ("max_v", C.c_double), ("integral_time", C.c_double),
                ("run_time", C.c_double), ("iterate_num", C.c_int64),
                ("n_split", C.c_int64), ("n_inlet", C.c_int64), ("n_outlet", C.c_int64),
                ("n_merge", C.c_int64)]


def _p(a: np.ndarray, t=_dp):
    assert a.flags.c_contiguous
    return a.ctypes.data_as(t)


PARTICLE_F64 = ["x", "v", "tv", "rhov", "vol", "mass", "h", "P", "rho", "E", "c", "coef"]
PARTICLE_I32 = ["phase", "mark_in"]
CELL_F64 = ["coef", "coef0", "dcoef", "lim", "alph", "rho", "E", "c", "ac", "rhov", "v", "a",
            "P", "G_v"]
CELL_I32 = ["ph_state", "ph_state_b"]


class HostState:
    """numpy-backed particle/wall/cell/scalar arrays for one case."""

    def __init__(self, case: Case, cap: int, nwall: int):
        self.case = case
        D, F, U, M, nc = case.dim, case.nphase, case.U, case.M, case.nc
        self.cap = cap
        self.n = 0
        z = lambda *s: np.zeros(s, dtype=np.float64)
        zi = lambda *s: np.zeros(s, dtype=np.int32)
        self.p = {
            "x": z(3, cap), "v": z(3, cap), "tv": z(3, cap), "rhov": z(3, cap),
            "vol": z(cap), "mass": z(cap), "h": z(cap), "P": z(cap), "rho": z(cap),
            "E": z(cap), "c": z(cap), "coef": z(U * M, cap),
            "phase": zi(cap), "mark_in": zi(cap),
        }
        self.w = {"x": z(3, nwall), "norm": z(3, nwall), "vol": z(nwall)}
        self.nwall = nwall
        self.c = {
            "coef": z(F, U, M, nc), "coef0": z(F, U, M, nc), "dcoef": z(F, U, M, nc),
            "lim": z(F, U, nc), "alph": z(F, nc), "rho": z(F, nc), "E": z(F, nc),
            "c": z(F, nc), "ac": z(F, nc), "rhov": z(F, 3, nc),
            "ph_state": zi(F, nc), "ph_state_b": zi(F, nc),
            "v": z(3, nc), "a": z(3, nc), "P": z(nc), "G_v": z(9, nc),
        }
        self.s = Scalars()

    # ctypes views ------------------------------------------------------------
    def particles(self) -> Particles:
        P = Particles()
        P.n, P.cap = self.n, self.cap
        for k in PARTICLE_F64:
            setattr(P, k, _p(self.p[k]))
        for k in PARTICLE_I32:
            setattr(P, k, _p(self.p[k], _ip))
        self._pk = P
        return P

    def walls(self) -> Walls:
        W = Walls()
        W.n = self.nwall
        for k in ("x", "norm", "vol"):
            setattr(W, k, _p(self.w[k]))
        self._wk = W
        return W

    def cells(self) -> Cells:
        Cc = Cells()
        for k in CELL_F64:
            setattr(Cc, k, _p(self.c[k]))
        for k in CELL_I32:
            setattr(Cc, k, _p(self.c[k], _ip))
        self._ck = Cc
        return Cc

    def copy(self) -> "HostState":
        o = HostState.__new__(HostState)
        o.case, o.cap, o.n, o.nwall = self.case, self.cap, self.n, self.nwall
        o.p = {k: v.copy() for k, v in self.p.items()}
        o.w = {k: v.copy() for k, v in self.w.items()}
        o.c = {k: v.copy() for k, v in self.c.items()}
        o.s = Scalars()
        C.pointer(o.s)[0] = self.s
        return o

    def live(self, key: str) -> np.ndarray:
        """Particle array restricted to the live count n."""
        a = self.p[key]
        return a[..., : self.n]
