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
        L.pgdg_malloc.argtypes = [C.POINTER(C.c_void_p), C.c_int64]
        L.pgdg_free.argtypes = [C.c_void_p]
        L.pgdg_memcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_int]
        L.pgdg_root_dim.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_int, C.c_void_p]
        L.pgdg_copy_kernel.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p]
        L.pgdg_event_create.argtypes = [C.POINTER(C.c_void_p)]
        L.pgdg_event_record.argtypes = [C.c_void_p, C.c_void_p]
        L.pgdg_event_elapsed_ms.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(C.c_float)]
        L.pgdg_event_destroy.argtypes = [C.c_void_p]
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


class DeviceArray:
    """A device buffer owned through the library's C ABI (no GPU framework
    needed).  Holds `shape`/`dtype` for round trips with numpy."""

    def __init__(self, shape, dtype):
        import numpy as np
        self.shape = tuple(shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.ptr = C.c_void_p()
        check(lib().pgdg_malloc(C.byref(self.ptr), self.nbytes), "pgdg_malloc")

    @classmethod
    def from_numpy(cls, a):
        import numpy as np
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        check(lib().pgdg_memcpy(d.ptr, C.c_void_p(a.ctypes.data), d.nbytes, 0), "pgdg_memcpy")
        return d

    def numpy(self):
        import numpy as np
        a = np.empty(self.shape, dtype=self.dtype)
        check(lib().pgdg_memcpy(C.c_void_p(a.ctypes.data), self.ptr, self.nbytes, 1), "pgdg_memcpy")
        return a

    def free(self):
        if self.ptr and self.ptr.value:
            lib().pgdg_free(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def device_sync():
    check(lib().pgdg_device_sync(), "pgdg_device_sync")


def mgfm_batch_np(inp, gamma_l, pinf_l, gamma_r, pinf_r, variant=MGFM_S3,
                  tol_mode=TOL_REFERENCE, rtol=1e-12, max_iter=50, pstar0=None):
    """Batched MGFM on host numpy input [6, n]; returns (out [3, n], iters [n])."""
    import numpy as np
    inp = np.ascontiguousarray(inp, dtype=np.float64)
    if inp.ndim != 2 or inp.shape[0] != 6:
        raise ValueError("inp must have shape [6, n]")
    n = inp.shape[1]
    d_in = DeviceArray.from_numpy(inp)
    d_out = DeviceArray((3, n), np.float64)
    d_it = DeviceArray((n,), np.int32)
    d_p0 = DeviceArray.from_numpy(np.asarray(pstar0, dtype=np.float64)) if pstar0 is not None else None
    mgfm_batch_dev(d_in, d_out, d_it, n, gamma_l, pinf_l, gamma_r, pinf_r, variant, tol_mode,
                   rtol, max_iter, d_p0)
    device_sync()
    return d_out.numpy(), d_it.numpy()


def mgfm_batch_dev(d_in, d_out, d_it, n, gamma_l, pinf_l, gamma_r, pinf_r, variant=MGFM_S3,
                   tol_mode=TOL_REFERENCE, rtol=1e-12, max_iter=50, d_p0=None):
    """Batched MGFM on DeviceArray buffers (null stream, asynchronous)."""
    eos = EosPair(gamma_l, pinf_l, gamma_r, pinf_r)
    opts = MgfmOpts(variant, tol_mode, rtol, max_iter)
    check(lib().pgdg_mgfm_batch(d_in.ptr, d_out.ptr, d_it.ptr if d_it is not None else None, n,
                                C.byref(eos), C.byref(opts), d_p0.ptr if d_p0 is not None else None,
                                None), "pgdg_mgfm_batch")


def root_dim_np(x, dim):
    """x^(1./dim) on the device for a host numpy array."""
    import numpy as np
    x = np.ascontiguousarray(x, dtype=np.float64)
    dx = DeviceArray.from_numpy(x)
    dy = DeviceArray(x.shape, np.float64)
    check(lib().pgdg_root_dim(dx.ptr, dy.ptr, x.size, dim, None), "pgdg_root_dim")
    device_sync()
    return dy.numpy()


class EventTimer:
    """Times work enqueued on `stream` (None: the null stream) between
    start() and stop() with HIP events; ms() waits for the stop event."""

    def __init__(self, stream=None):
        self.stream = stream
        self.a, self.b = C.c_void_p(), C.c_void_p()
        check(lib().pgdg_event_create(C.byref(self.a)), "pgdg_event_create")
        check(lib().pgdg_event_create(C.byref(self.b)), "pgdg_event_create")

    def start(self):
        check(lib().pgdg_event_record(self.a, self.stream), "pgdg_event_record")

    def stop(self):
        check(lib().pgdg_event_record(self.b, self.stream), "pgdg_event_record")

    def ms(self) -> float:
        v = C.c_float()
        check(lib().pgdg_event_elapsed_ms(self.a, self.b, C.byref(v)), "pgdg_event_elapsed_ms")
        return float(v.value)

    def close(self):
        for e in (self.a, self.b):
            if e.value:
                lib().pgdg_event_destroy(e)
        self.a, self.b = C.c_void_p(), C.c_void_p()


def device_count() -> int:
    n = C.c_int(0)
    check(lib().pgdg_device_count(C.byref(n)), "pgdg_device_count")
    return n.value
