"""pgdg_p2_amd - MI355X-native PG-DG time step (particle-grid discontinuous
Galerkin for compressible multiphase flow).

The compute path is the HIP library ``libpgdg.so`` built from ``csrc/``; this
module only binds its C ABI (``include/pgdg.h``) with ctypes.  There is no
Python or CPU fallback: if the library is missing, ``lib()`` raises.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

_PKG = Path(__file__).resolve().parent
LIB_PATH = _PKG / "libpgdg.so"

_lib = None

PGDG_OK = 0
STATUS = {
    0: "ok",
    -1: "invalid argument",
    -2: "HIP error",
    -3: "capacity exceeded",
    -4: "non-finite value",
    -5: "call out of order",
    -6: "no HIP device",
}

MGFM_S3, MGFM_T2 = 0, 1
TOL_REFERENCE, TOL_RELATIVE = 0, 1


class PgdgError(RuntimeError):
    pass


def check(rc: int, what: str = "pgdg call") -> None:
    if rc != PGDG_OK:
        raise PgdgError(f"{what} failed: {rc} ({STATUS.get(rc, '?')})")


class EosPair(C.Structure):
    _fields_ = [("gamma_l", C.c_double), ("pinf_l", C.c_double),
                ("gamma_r", C.c_double), ("pinf_r", C.c_double)]


class MgfmOpts(C.Structure):
    _fields_ = [("variant", C.c_int), ("tol_mode", C.c_int),
                ("rtol", C.c_double), ("max_iter", C.c_int)]


def lib() -> C.CDLL:
    """Load libpgdg.so (built in-tree by __graft_entry__.build())."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise PgdgError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = C.CDLL(str(LIB_PATH))
        L.pgdg_version.restype = C.c_char_p
        L.pgdg_device_count.argtypes = [C.POINTER(C.c_int)]
        L.pgdg_mgfm_batch.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64,
                                      C.POINTER(EosPair), C.POINTER(MgfmOpts), C.c_void_p,
                                      C.c_void_p]
        from . import _bind  # noqa: F401  (declares the step API signatures)
        _bind.declare(L)
        _lib = L
    return _lib


def _stream_ptr(torch_stream=None):
    import torch
    s = torch_stream if torch_stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


def mgfm_batch(inp, gamma_l, pinf_l, gamma_r, pinf_r, variant=MGFM_S3, tol_mode=TOL_REFERENCE,
               rtol=1e-12, max_iter=50, pstar0=None, out=None, iters=None, stream=None):
    """Batched MGFM (replaces MPM::MGFM, S3:mpm.cpp:882).

    inp: CUDA float64 tensor [6, n] = (ul, ur, pl, pr, rhol, rhor).
    Returns (out [3, n] = (rho*_L, p*, u*_L), iters [n] int32).
    """
    import torch
    if inp.dtype != torch.float64 or inp.dim() != 2 or inp.shape[0] != 6 or not inp.is_cuda:
        raise ValueError("inp must be a CUDA float64 tensor of shape [6, n]")
    inp = inp.contiguous()
    n = inp.shape[1]
    if out is None:
        out = torch.empty((3, n), dtype=torch.float64, device=inp.device)
    if iters is None:
        iters = torch.empty((n,), dtype=torch.int32, device=inp.device)
    eos = EosPair(gamma_l, pinf_l, gamma_r, pinf_r)
    opts = MgfmOpts(variant, tol_mode, rtol, max_iter)
    p0 = C.c_void_p(pstar0.data_ptr()) if pstar0 is not None else None
    check(lib().pgdg_mgfm_batch(C.c_void_p(inp.data_ptr()), C.c_void_p(out.data_ptr()),
                                C.c_void_p(iters.data_ptr()), n, C.byref(eos), C.byref(opts),
                                p0, _stream_ptr(stream)), "pgdg_mgfm_batch")
    return out, iters
