"""ctypes binding of the oracle's step API (TEST INFRASTRUCTURE ONLY)."""
from __future__ import annotations

import ctypes as C

from pgdg_p2_amd.state import Case, Cells, HostState, Particles, Scalars, Walls

FN = {"set_timestep": 0, "used_state_update": 1, "bin": 2, "split": 3, "merge": 4,
      "membership": 5, "neighbors": 6, "p2g": 7, "limiter": 8, "rhs": 9, "update_grid": 10,
      "update_particle": 11, "run_simulation": 12}


def declare(L: C.CDLL) -> None:
    L.oracle_create.restype = C.c_void_p
    L.oracle_create.argtypes = [C.POINTER(Case)]
    L.oracle_destroy.argtypes = [C.c_void_p]
    L.oracle_load.argtypes = [C.c_void_p, C.POINTER(Particles), C.POINTER(Walls),
                              C.POINTER(Cells), C.POINTER(Scalars)]
    L.oracle_save.argtypes = [C.c_void_p, C.POINTER(Particles), C.POINTER(Cells),
                              C.POINTER(Scalars)]
    L.oracle_step.argtypes = [C.c_void_p, C.c_int]
    L.oracle_call.argtypes = [C.c_void_p, C.c_int, C.c_int]
    L.oracle_neighbor_counts.argtypes = [C.c_void_p, C.POINTER(C.c_int32)]
    L.oracle_used_cells.argtypes = [C.c_void_p, C.POINTER(C.c_int32)]
    L.oracle_bins.restype = C.c_int64
    L.oracle_bins.argtypes = [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.c_int64]
    L.oracle_list.argtypes = [C.c_void_p, C.c_int64, C.POINTER(C.c_int32), C.c_int]
    L.oracle_kernel.argtypes = [C.c_int, C.c_double, C.c_double, C.POINTER(C.c_double),
                                C.POINTER(C.c_double)]
    L.oracle_limiter_fn.restype = C.c_double
    L.oracle_limiter_fn.argtypes = [C.c_double, C.c_int]
    L.oracle_face_flux.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                   C.POINTER(C.c_double), C.POINTER(C.c_double)]


def _dptr(a):
    import numpy as np
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(C.POINTER(C.c_double))


def kernel(dim: int, dist: float, h: float):
    """The restatement's Wendland kernel: (W, (dW/dr)/r)."""
    from oracle.pyoracle import lib
    w, dw = C.c_double(), C.c_double()
    assert lib().oracle_kernel(dim, dist, h, C.byref(w), C.byref(dw)) == 0
    return w.value, dw.value


def limiter_fn(ds: float, bd_fg: int) -> float:
    """The restatement's Limiter(ds, bd_fg)."""
    from oracle.pyoracle import lib
    return lib().oracle_limiter_fn(ds, bd_fg)


def face_flux(dim: int, kind: int, n, left, right):
    """Interior HLLC (kind 0) or boundary HLL (kind 1) flux of the restatement.
    left/right: (rho, rho v[3], E, P, c); returns the U x D flux tensor."""
    import numpy as np
    from oracle.pyoracle import lib
    n, pn = _dptr(n)
    l, pl = _dptr(left)
    r, pr = _dptr(right)
    F = np.zeros((dim + 2, dim))
    assert lib().oracle_face_flux(dim, kind, pn, pl, pr,
                                  F.ctypes.data_as(C.POINTER(C.c_double))) == 0
    return F


class OracleSim:
    """The oracle's restatement of the reference step, driven from Python."""

    def __init__(self, case: Case, st: HostState):
        from oracle.pyoracle import lib
        self.L = lib()
        self.case = case
        self.h = self.L.oracle_create(C.byref(case))
        if not self.h:
            raise RuntimeError("oracle_create failed")
        self.load(st)

    def load(self, st: HostState):
        self._cap = st.cap
        P, W, Cc = st.particles(), st.walls(), st.cells()
        rc = self.L.oracle_load(self.h, C.byref(P), C.byref(W), C.byref(Cc), C.byref(st.s))
        assert rc == 0, rc

    def save(self, st: HostState):
        P, Cc = st.particles(), st.cells()
        rc = self.L.oracle_save(self.h, C.byref(P), C.byref(Cc), C.byref(st.s))
        assert rc == 0, rc
        st.n = P.n
        return st

    def scalars(self):
        """The time-loop scalars (the state is saved into scratch arrays)."""
        from pgdg_p2_amd.state import HostState
        st = HostState(self.case, self._cap, 0)
        self.save(st)
        return st.s

    def step(self, n=1):
        rc = self.L.oracle_step(self.h, n)
        assert rc == 0, rc

    def call(self, fn: str, arg: int = 0):
        rc = self.L.oracle_call(self.h, FN[fn], arg)
        assert rc == 0, rc

    def neighbor_counts(self, n: int):
        """particle->neighbor.size() of the n fluid particles, particle order."""
        import numpy as np
        out = np.zeros(n, dtype=np.int32)
        rc = self.L.oracle_neighbor_counts(self.h, out.ctypes.data_as(C.POINTER(C.c_int32)))
        assert rc == 0, rc
        return out

    def bins(self, nc: int, cap: int):
        """Cell lists as (off[nc+1], entries): fluid positions, walls -(w+2)."""
        import numpy as np
        off = np.zeros(nc + 1, dtype=np.int32)
        ent = np.zeros(max(cap, 1), dtype=np.int32)
        k = self.L.oracle_bins(self.h, off.ctypes.data_as(C.POINTER(C.c_int32)),
                               ent.ctypes.data_as(C.POINTER(C.c_int32)), cap)
        assert k >= 0, "bins: capacity"
        return off, ent[:k]

    def neighbor_list(self, i: int, cap: int = 4096):
        """particle->neighbor of the i-th fluid particle (fluid positions, walls
        -(w+2), -1 for a reference to a removed particle)."""
        import numpy as np
        out = np.zeros(cap, dtype=np.int32)
        k = self.L.oracle_list(self.h, i, out.ctypes.data_as(C.POINTER(C.c_int32)), cap)
        assert k >= 0, "neighbor_list: capacity or index"
        return out[:k]

    def used_cells(self, nc: int):
        """1 for each cell in the used-cell list, else 0 (cell order)."""
        import numpy as np
        out = np.zeros(nc, dtype=np.int32)
        rc = self.L.oracle_used_cells(self.h, out.ctypes.data_as(C.POINTER(C.c_int32)))
        assert rc == 0, rc
        return out

    def __del__(self):
        try:
            if getattr(self, "h", None):
                self.L.oracle_destroy(self.h)
                self.h = None
        except Exception:
            pass
