"""MLS particle-to-grid fit with a per-cell least-squares solve on the GPU
(C ABI ``pgdg_mls_ls_fit``).

The fit generalises MPM::MLS_P2G (S3:mpm.cpp:274-382): the reference forms a
Wendland-weighted average of the particles within the P2G cut-off; here the
weighted least-squares problem is solved in the zero-mean P2 basis of the
cell (10 modes), with the same weights W(r, h) * vol.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import DeviceArray, check, device_sync, lib

PREC_FP64, PREC_MIXED = 0, 1


class MlsProblem(C.Structure):
    _fields_ = [("n", C.c_int * 3), ("dx", C.c_double), ("cut", C.c_double), ("h", C.c_double),
                ("nfields", C.c_int), ("order", C.c_int), ("np", C.c_int64),
                ("off", C.c_void_p), ("x", C.c_void_p), ("vol", C.c_void_p), ("val", C.c_void_p),
                ("u", C.c_void_p), ("volf", C.c_void_p), ("valf", C.c_void_p)]


def _declare(L):
    L.pgdg_mls_ls_fit.argtypes = [C.POINTER(MlsProblem), C.c_int, C.c_void_p, C.c_void_p,
                                  C.c_void_p]


class MlsFit:
    """Device-resident inputs of one fit problem (cell-sorted particles).

    data: dict with n, dx, off[nc+1], x[3][np], u[3][np] (cell-local in
    [0,1)), vol[np], val[nf][np] (as made by workloads.mls_particles).
    precisions: which input forms to upload (PREC_FP64 / PREC_MIXED)."""

    def __init__(self, data, order=2, cut=None, h=None, precisions=(PREC_FP64, PREC_MIXED)):
        L = lib()
        _declare(L)
        n = data["n"]
        self.n = (n, n, n) if np.isscalar(n) else tuple(int(v) for v in n)
        self.nc = self.n[0] * self.n[1] * self.n[2]
        dx = float(data["dx"])
        self.nf = int(data["val"].shape[0])
        self.np = int(data["x"].shape[1] if "x" in data else data["u"].shape[1])
        self._keep = []
        off = DeviceArray.from_numpy(np.asarray(data["off"], dtype=np.int64))
        self._keep.append(off)
        p = MlsProblem()
        p.n[:] = self.n
        p.dx = dx
        p.cut = 1.5 * dx if cut is None else cut
        p.h = 1.6 * dx if h is None else h
        p.nfields = self.nf
        p.order = order
        p.np = self.np
        p.off = off.ptr.value

        def up(a, dt):
            d = DeviceArray.from_numpy(np.ascontiguousarray(a, dtype=dt))
            self._keep.append(d)
            return d.ptr.value

        if PREC_FP64 in precisions:
            p.x, p.vol, p.val = up(data["x"], np.float64), up(data["vol"], np.float64), \
                up(data["val"], np.float64)
        if PREC_MIXED in precisions:
            p.u, p.volf, p.valf = up(data["u"], np.float32), up(data["vol"], np.float32), \
                up(data["val"], np.float32)
        self.problem = p
        self.coef = DeviceArray((self.nf, 10, self.nc), np.float64)
        self.order_used = DeviceArray((self.nc,), np.int32)

    def launch(self, precision, stream=None):
        """Enqueue one fit (asynchronous)."""
        check(lib().pgdg_mls_ls_fit(C.byref(self.problem), precision, self.coef.ptr,
                                    self.order_used.ptr, stream), "pgdg_mls_ls_fit")

    def run(self, precision):
        """Fit and return (coef [nf][10][nc], order used [nc]) on the host."""
        self.launch(precision)
        device_sync()
        return self.coef.numpy(), self.order_used.numpy()

    def close(self):
        for d in self._keep:
            d.free()
        self.coef.free()
        self.order_used.free()
        self._keep = []


def center_values(coef):
    """Value of the fitted P2 polynomial at the cell centre (xi = 0):
    c0 - (c4 + c5 + c6) / 12.  coef [nf][10][nc] -> [nf][nc]."""
    return coef[:, 0] - (coef[:, 4] + coef[:, 5] + coef[:, 6]) / 12.0
