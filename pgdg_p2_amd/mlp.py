"""Learned initial guess for the MGFM Newton solve (C ABI pgdg_mlp_*).

A 12 -> 32 -> 32 -> 32 -> 2 ReLU MLP trained on the GPU (BF16 MFMA forward
and backward kernels, FP32 Adam) on faces labelled by the exact MGFM solve;
its p* prediction seeds the Newton iteration, which still runs to the
requested tolerance.  The reference has no counterpart: its Newton solve
starts from max(pL, pR) (S3:mpm.cpp:891).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import DeviceArray, EosPair, MgfmOpts, check, device_sync, lib, MGFM_S3, TOL_RELATIVE

NIN, NPAD, NOUT, NPARAM = 12, 16, 2, 3184


class MlpNorm(C.Structure):
    _fields_ = [("feat_mu", C.c_float * NIN), ("feat_sd", C.c_float * NIN),
                ("out_mu", C.c_float * NOUT), ("out_sd", C.c_float * NOUT), ("keep", C.c_float)]


class MlpTrainOpts(C.Structure):
    _fields_ = [("epochs", C.c_int), ("batch", C.c_int), ("lr", C.c_float),
                ("seed", C.c_uint64), ("variant", C.c_int)]


def _declare(L):
    vp, i64 = C.c_void_p, C.c_int64
    L.pgdg_mlp_create.argtypes = [C.POINTER(C.c_void_p)]
    L.pgdg_mlp_destroy.argtypes = [vp]
    L.pgdg_mlp_train.argtypes = [vp, vp, i64, C.POINTER(EosPair), C.POINTER(MlpTrainOpts), vp,
                                 C.c_int, C.POINTER(C.c_int), vp]
    L.pgdg_mlp_train_pairs.argtypes = [vp, vp, vp, i64, C.POINTER(EosPair), C.c_int,
                                       C.POINTER(MlpTrainOpts), vp, C.c_int, C.POINTER(C.c_int), vp]
    L.pgdg_mgfm_batch_pairs.argtypes = [vp, vp, vp, vp, vp, i64, C.POINTER(EosPair), C.c_int,
                                        C.POINTER(MgfmOpts), vp]
    L.pgdg_mlp_guess.argtypes = [vp, vp, i64, C.POINTER(EosPair), vp, vp]
    L.pgdg_mgfm_batch_learned.argtypes = [vp, vp, vp, vp, i64, C.POINTER(EosPair),
                                          C.POINTER(MgfmOpts), vp]
    L.pgdg_mlp_get_params.argtypes = [vp, vp, C.POINTER(MlpNorm)]
    L.pgdg_mlp_set_params.argtypes = [vp, vp, C.POINTER(MlpNorm)]
    L.pgdg_mlp_forward.argtypes = [vp, vp, vp, i64, vp]
    L.pgdg_mlp_loss_grad.argtypes = [vp, vp, vp, i64, vp, vp, vp]


class Mlp:
    def __init__(self):
        self.L = lib()
        _declare(self.L)
        self.h = C.c_void_p()
        check(self.L.pgdg_mlp_create(C.byref(self.h)), "pgdg_mlp_create")

    def close(self):
        if self.h:
            self.L.pgdg_mlp_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def train(self, d_in: DeviceArray, n: int, eos, epochs=8, batch=4096, lr=2e-3, seed=0x5EED0004,
              variant=MGFM_S3, hist_cap=1 << 16):
        """Label the device faces d_in [6][n] with the exact solve and train.
        Returns the per-step mean loss history."""
        hist = np.zeros(hist_cap, dtype=np.float32)
        steps = C.c_int(0)
        opts = MlpTrainOpts(epochs, batch, lr, seed, variant)
        check(self.L.pgdg_mlp_train(self.h, d_in.ptr, n, C.byref(EosPair(*eos)), C.byref(opts),
                                    C.c_void_p(hist.ctypes.data), hist_cap, C.byref(steps), None),
              "pgdg_mlp_train")
        return hist[:min(steps.value, hist_cap)].copy()

    def train_pairs(self, d_in: DeviceArray, d_pair: DeviceArray | None, n: int, eos_list, epochs=8,
                    batch=4096, lr=2e-3, seed=0x5EED0004, variant=MGFM_S3, hist_cap=1 << 16):
        """As train, over faces whose EOS pair d_pair[i] indexes eos_list."""
        hist = np.zeros(hist_cap, dtype=np.float32)
        steps = C.c_int(0)
        opts = MlpTrainOpts(epochs, batch, lr, seed, variant)
        tab = (EosPair * len(eos_list))(*[EosPair(*e) for e in eos_list])
        check(self.L.pgdg_mlp_train_pairs(self.h, d_in.ptr, d_pair.ptr if d_pair is not None else None, n, tab,
                                          len(eos_list), C.byref(opts), C.c_void_p(hist.ctypes.data), hist_cap,
                                          C.byref(steps), None),
              "pgdg_mlp_train_pairs")
        return hist[:min(steps.value, hist_cap)].copy()

    def guess_dev(self, d_in, n, eos, d_p0):
        check(self.L.pgdg_mlp_guess(self.h, d_in.ptr, n, C.byref(EosPair(*eos)), d_p0.ptr, None),
              "pgdg_mlp_guess")

    def mgfm_dev(self, d_in, d_out, d_it, n, eos, variant=MGFM_S3, tol_mode=TOL_RELATIVE,
                 rtol=1e-12, max_iter=100):
        """pgdg_mgfm_batch with the learned initial p* (asynchronous)."""
        opts = MgfmOpts(variant, tol_mode, rtol, max_iter)
        check(self.L.pgdg_mgfm_batch_learned(self.h, d_in.ptr, d_out.ptr,
                                             d_it.ptr if d_it is not None else None, n,
                                             C.byref(EosPair(*eos)), C.byref(opts), None),
              "pgdg_mgfm_batch_learned")

    def get_params(self):
        P = np.zeros(NPARAM, dtype=np.float32)
        nm = MlpNorm()
        check(self.L.pgdg_mlp_get_params(self.h, C.c_void_p(P.ctypes.data), C.byref(nm)),
              "pgdg_mlp_get_params")
        return P, nm

    def set_params(self, P, nm: MlpNorm | None = None):
        P = np.ascontiguousarray(P, dtype=np.float32)
        check(self.L.pgdg_mlp_set_params(self.h, C.c_void_p(P.ctypes.data),
                                         C.byref(nm) if nm is not None else None),
              "pgdg_mlp_set_params")

    def forward(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        n = X.shape[0]
        dx = DeviceArray.from_numpy(X)
        dy = DeviceArray((n, NOUT), np.float32)
        check(self.L.pgdg_mlp_forward(self.h, dx.ptr, dy.ptr, n, None), "pgdg_mlp_forward")
        device_sync()
        return dy.numpy()

    def loss_grad(self, X, T):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        n = X.shape[0]
        dx, dt = DeviceArray.from_numpy(X), DeviceArray.from_numpy(T)
        dg = DeviceArray((NPARAM,), np.float32)
        dl = DeviceArray((1,), np.float32)
        check(self.L.pgdg_mlp_loss_grad(self.h, dx.ptr, dt.ptr, n, dg.ptr, dl.ptr, None),
              "pgdg_mlp_loss_grad")
        device_sync()
        return float(dl.numpy()[0]), dg.numpy()


def mgfm_pairs_dev(mlp: Mlp | None, d_in, d_pair, d_out, d_it, n, eos_list, variant=MGFM_S3,
                   tol_mode=TOL_RELATIVE, rtol=1e-12, max_iter=100):
    """pgdg_mgfm_batch_pairs: faces with per-face EOS pairs, from the
    network's p*0 when mlp is given (asynchronous)."""
    L = lib()
    _declare(L)
    opts = MgfmOpts(variant, tol_mode, rtol, max_iter)
    tab = (EosPair * len(eos_list))(*[EosPair(*e) for e in eos_list])
    check(L.pgdg_mgfm_batch_pairs(mlp.h if mlp is not None else None, d_in.ptr,
                                  d_pair.ptr if d_pair is not None else None, d_out.ptr,
                                  d_it.ptr if d_it is not None else None, n, tab, len(eos_list),
                                  C.byref(opts), None),
          "pgdg_mgfm_batch_pairs")
