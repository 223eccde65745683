"""Exact comparison of two host states (particle order, counts and every
field; -0.0 == 0.0 and NaN == NaN)."""
from __future__ import annotations

import numpy as np

from .state import HostState

P_FIELDS_3 = ["x", "v", "tv", "rhov"]
P_FIELDS_1 = ["vol", "mass", "h", "P", "rho", "E", "c", "phase", "mark_in"]
C_FIELDS = ["coef", "coef0", "dcoef", "lim", "alph", "rho", "E", "c", "ac", "rhov", "v", "a", "P",
            "G_v", "ph_state", "ph_state_b"]


def _eq(a, b):
    return np.array_equal(a, b, equal_nan=True) if a.dtype.kind == "f" else np.array_equal(a, b)


def compare_states(a: HostState, b: HostState, cells: bool = True, rtol: float = 0.0):
    """List of (field, first bad index, a value, b value); empty when equal.

    rtol > 0 compares floating fields with that relative tolerance instead of
    exactly."""
    bad = []
    if a.n != b.n:
        return [("n", -1, a.n, b.n)]
    n = a.n
    D = a.case.dim

    def chk(name, x, y):
        if rtol > 0 and x.dtype.kind == "f":
            ok = np.isclose(x, y, rtol=rtol, atol=0.0, equal_nan=True)
            if not ok.all():
                idx = np.argwhere(~ok)[0]
                bad.append((name, tuple(int(t) for t in idx), x[tuple(idx)], y[tuple(idx)]))
        elif not _eq(x, y):
            idx = np.argwhere(~((x == y) | (np.isnan(x) & np.isnan(y)) if x.dtype.kind == "f" else (x == y)))[0]
            bad.append((name, tuple(int(t) for t in idx), x[tuple(idx)], y[tuple(idx)]))

    for k in P_FIELDS_3:
        chk("p." + k, a.p[k][:D, :n], b.p[k][:D, :n])
    for k in P_FIELDS_1:
        chk("p." + k, a.p[k][:n], b.p[k][:n])
    chk("p.coef", a.p["coef"][:, :n], b.p["coef"][:, :n])
    if cells:
        for k in C_FIELDS:
            chk("c." + k, a.c[k], b.c[k])
    for k in ("dt", "max_v", "integral_time", "iterate_num", "n_split", "n_inlet", "n_outlet",
              "n_merge"):
        if getattr(a.s, k) != getattr(b.s, k):
            bad.append(("s." + k, -1, getattr(a.s, k), getattr(b.s, k)))
    return bad
