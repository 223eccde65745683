"""ctypes binding of oracle/liboracle.so (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module, as the checker of the HIP library.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "liboracle.so"
_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} not built (make -C oracle)")
        L = C.CDLL(str(LIB_PATH))
        L.oracle_mgfm_batch.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64,
                                        C.c_double, C.c_double, C.c_double, C.c_double,
                                        C.c_int, C.c_int, C.c_double, C.c_int, C.c_void_p]
        L.oracle_root_dim.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_int]
        try:
            from . import pyoracle_step
            pyoracle_step.declare(L)
        except ImportError:
            pass
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    return C.c_void_p(a.ctypes.data)


def mgfm_batch(inp: np.ndarray, gl, pinfl, gr, pinfr, variant=0, mode=0, rtol=1e-12,
               max_iter=50, guess: np.ndarray | None = None):
    inp = np.ascontiguousarray(inp, dtype=np.float64)
    assert inp.ndim == 2 and inp.shape[0] == 6
    n = inp.shape[1]
    out = np.empty((3, n), dtype=np.float64)
    it = np.empty((n,), dtype=np.int32)
    g = None
    if guess is not None:
        guess = np.ascontiguousarray(guess, dtype=np.float64)
        g = _ptr(guess)
    lib().oracle_mgfm_batch(_ptr(inp), _ptr(out), _ptr(it), n, gl, pinfl, gr, pinfr, variant,
                            mode, rtol, max_iter, g)
    return out, it


def root_dim(x: np.ndarray, dim: int) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty_like(x)
    lib().oracle_root_dim(_ptr(x), _ptr(y), x.size, dim)
    return y


def mls_ls(n, dx, off, x, vol, val, order=2, cut=None, h=None):
    """FP64 MLS least-squares fit per cell: coef [nf][10][nc], order used [nc]."""
    n3 = (C.c_int * 3)(*([n] * 3 if np.isscalar(n) else n))
    nc = n3[0] * n3[1] * n3[2]
    off = np.ascontiguousarray(off, dtype=np.int64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    vol = np.ascontiguousarray(vol, dtype=np.float64)
    val = np.ascontiguousarray(val, dtype=np.float64)
    nf = val.shape[0]
    npart = x.shape[1]
    cut = 1.5 * dx if cut is None else cut
    h = 1.6 * dx if h is None else h
    coef = np.zeros((nf, 10, nc))
    ordv = np.zeros(nc, dtype=np.int32)
    L = lib()
    L.oracle_mls_ls.argtypes = [C.c_void_p, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p,
                                C.c_void_p, C.c_int64, C.c_int, C.c_int, C.c_double, C.c_double,
                                C.c_void_p, C.c_void_p]
    L.oracle_mls_ls(n3, dx, _ptr(off), _ptr(x), _ptr(vol), _ptr(val), npart, nf, order, cut, h,
                    _ptr(coef), _ptr(ordv))
    return coef, ordv
