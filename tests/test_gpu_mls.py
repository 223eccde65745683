"""GPU MLS fit (pgdg_mls_ls_fit) against the FP64 oracle.

FP64 form: identical solved orders and every coefficient within 1e-10 of
the oracle, relative to the field's largest coefficient of that mode (the
kernel performs the oracle's operations in the oracle's order with correctly
rounded division and square root, so it is in fact bit-identical, which is
asserted too).
Mixed form (FP32 assembly, FP64 solve): per-field L1 and Linf deviation
from the FP64 fit stated and bounded, and the L1 error of the fitted cell
centre values against the exact field equal to the FP64 fit's within 1%.
"""
import numpy as np
import pytest

from oracle import pyoracle
from pgdg_p2_amd import workloads

FP64_RTOL = 1e-10


def _rel_err(a, b):
    # per (field, mode): max |a-b| / max |b|
    scale = np.maximum(np.abs(b).max(axis=2, keepdims=True), 1e-300)
    return (np.abs(a - b) / scale).max()


def _sparse(d, keep_every):
    """Keep every k-th particle (cell-sorted order preserved)."""
    idx = np.arange(0, d["x"].shape[1], keep_every)
    cell = d["cell"][idx]
    nc = d["n"] ** 3
    off = np.zeros(nc + 1, dtype=np.int64)
    np.cumsum(np.bincount(cell, minlength=nc), out=off[1:])
    return {"n": d["n"], "dx": d["dx"], "off": off, "cell": cell, "u": d["u"][:, idx],
            "x": d["x"][:, idx], "vol": d["vol"][idx], "val": d["val"][:, idx]}


@pytest.mark.gpu
@pytest.mark.parametrize("order", [2, 1, 0])
def test_mls_fp64_matches_oracle(cuda, order):
    from pgdg_p2_amd import mls
    d = workloads.mls_particles(n=14, seed=11)
    print(f"mls fp64 n=14 order={order} np={d['x'].shape[1]}", flush=True)
    ref, ordr = pyoracle.mls_ls(d["n"], d["dx"], d["off"], d["x"], d["vol"], d["val"], order=order)
    f = mls.MlsFit(d, order=order)
    coef, ordg = f.run(mls.PREC_FP64)
    f.close()
    np.testing.assert_array_equal(ordg, ordr)
    assert (ordr == order).all()
    assert _rel_err(coef, ref) <= FP64_RTOL
    np.testing.assert_array_equal(coef, ref)


@pytest.mark.gpu
def test_mls_fallback_orders(cuda):
    # ~0.7 particles per cell: many neighbourhoods cannot carry P2 (or P1),
    # and some are empty; the GPU must take the same fallbacks.
    from pgdg_p2_amd import mls
    d = _sparse(workloads.mls_particles(n=12, seed=12), 11)
    print(f"mls fallback np={d['x'].shape[1]}", flush=True)
    ref, ordr = pyoracle.mls_ls(d["n"], d["dx"], d["off"], d["x"], d["vol"], d["val"])
    f = mls.MlsFit(d)
    coef, ordg = f.run(mls.PREC_FP64)
    f.close()
    np.testing.assert_array_equal(ordg, ordr)
    assert len(set(ordr.tolist())) >= 3, set(ordr.tolist())
    assert _rel_err(coef, ref) <= FP64_RTOL
    np.testing.assert_array_equal(coef, ref)


@pytest.mark.gpu
def test_mls_mixed_error_budget(cuda):
    from pgdg_p2_amd import mls
    d = workloads.mls_particles(n=24, seed=13)
    print(f"mls mixed n=24 np={d['x'].shape[1]}", flush=True)
    f = mls.MlsFit(d)
    c64, o64 = f.run(mls.PREC_FP64)
    c32, o32 = f.run(mls.PREC_MIXED)
    f.close()
    np.testing.assert_array_equal(o32, o64)
    v64, v32 = mls.center_values(c64), mls.center_values(c32)
    for fld in range(v64.shape[0]):
        l1 = np.abs(v32[fld] - v64[fld]).mean() / np.abs(v64[fld]).mean()
        linf = np.abs(v32[fld] - v64[fld]).max() / np.abs(v64[fld]).max()
        print(f"field {fld}: mixed vs fp64 centre values L1 {l1:.2e} Linf {linf:.2e}")
        assert l1 <= 1e-6 and linf <= 1e-5
    n = d["n"]
    g = (np.arange(n) + 0.5) * d["dx"]
    X = np.stack(np.meshgrid(g, g, g, indexing="ij")).reshape(3, -1)
    exact = workloads.mls_fields(X, v64.shape[0], 13)
    e64 = np.abs(v64 - exact).mean(axis=1)
    e32 = np.abs(v32 - exact).mean(axis=1)
    print("L1 err vs exact fp64", e64, "mixed", e32)
    np.testing.assert_allclose(e32, e64, rtol=1e-2)


@pytest.mark.gpu
def test_mls_empty_and_single_cell(cuda):
    from pgdg_p2_amd import mls
    # one cell, one particle: P0 = the particle's value
    d = {"n": 1, "dx": 1.0, "off": np.array([0, 1], np.int64), "x": np.full((3, 1), 0.5),
         "u": np.full((3, 1), 0.5), "vol": np.ones(1), "val": np.arange(1.0, 7.0)[:, None]}
    f = mls.MlsFit(d)
    coef, o = f.run(mls.PREC_FP64)
    f.close()
    assert o.tolist() == [0]
    np.testing.assert_allclose(coef[:, 0, 0], np.arange(1.0, 7.0), rtol=1e-15)
    assert not coef[:, 1:].any()


@pytest.mark.gpu
def test_mls_empty_neighbourhood(cuda):
    # 4x1x1 cells, one particle near the left face of cell 0: cells 0 and 1
    # fall back to P0, cells 2 and 3 see no particle (order -1, zero fit).
    from pgdg_p2_amd import mls
    x = np.array([[0.1], [0.5], [0.5]])
    d = {"n": (4, 1, 1), "dx": 1.0, "off": np.array([0, 1, 1, 1, 1], np.int64), "x": x,
         "u": x.copy(), "vol": np.ones(1), "val": np.full((2, 1), 3.0)}
    ref, ordr = pyoracle.mls_ls((4, 1, 1), 1.0, d["off"], x, d["vol"], d["val"])
    f = mls.MlsFit(d)
    coef, o = f.run(mls.PREC_FP64)
    cm, om = f.run(mls.PREC_MIXED)
    f.close()
    assert ordr.tolist() == [0, 0, -1, -1]
    np.testing.assert_array_equal(o, ordr)
    np.testing.assert_array_equal(om, ordr)
    np.testing.assert_array_equal(coef, ref)
    np.testing.assert_allclose(cm, ref, rtol=1e-6, atol=0)



@pytest.mark.gpu
def test_mls_mixed_error_budget_128(cuda):
    # the isolated-MLS workload at its full size, 128^3 cells: the mixed
    # form's centre values against the FP64 form's, per field
    from pgdg_p2_amd import mls
    d = workloads.mls_particles(n=128)
    print(f"mls mixed n=128 np={d['x'].shape[1]}", flush=True)
    f = mls.MlsFit(d)
    del d
    c64, o64 = f.run(mls.PREC_FP64)
    c32, o32 = f.run(mls.PREC_MIXED)
    f.close()
    assert (o32 != o64).sum() == 0
    v64, v32 = mls.center_values(c64), mls.center_values(c32)
    for fld in range(v64.shape[0]):
        l1 = np.abs(v32[fld] - v64[fld]).mean() / np.abs(v64[fld]).mean()
        linf = np.abs(v32[fld] - v64[fld]).max() / np.abs(v64[fld]).max()
        print(f"field {fld}: mixed vs fp64 centre values L1 {l1:.2e} Linf {linf:.2e}", flush=True)
        assert l1 <= 1e-6 and linf <= 1e-5
