"""Host-side argument checks of pgdg_mls_ls_fit (no device work happens)."""
import ctypes as C

import numpy as np

import pgdg_p2_amd as P
from pgdg_p2_amd import mls


def _problem(**kw):
    p = mls.MlsProblem()
    p.n[:] = (4, 4, 4)
    p.dx, p.cut, p.h = 0.25, 0.375, 0.4
    p.nfields, p.order, p.np = 6, 2, 10
    p.off = p.x = p.vol = p.val = 8  # never dereferenced: rejected first
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def test_mls_rejects_bad_arguments():
    L = P.lib()
    mls._declare(L)
    bad = [dict(nfields=0), dict(nfields=7), dict(order=3), dict(order=-1), dict(dx=0.0),
           dict(h=-1.0), dict(cut=0.0), dict(off=None), dict(np=-1)]
    for kw in bad:
        assert L.pgdg_mls_ls_fit(C.byref(_problem(**kw)), 0, 8, None, None) == -1, kw
    # missing buffers of the requested form
    assert L.pgdg_mls_ls_fit(C.byref(_problem()), 1, 8, None, None) == -1
    assert L.pgdg_mls_ls_fit(C.byref(_problem(x=None)), 0, 8, None, None) == -1
    assert L.pgdg_mls_ls_fit(C.byref(_problem()), 2, 8, None, None) == -1
    assert L.pgdg_mls_ls_fit(None, 0, 8, None, None) == -1
    assert L.pgdg_mls_ls_fit(C.byref(_problem()), 0, None, None, None) == -1
    assert np.dtype(np.float64).itemsize == 8
