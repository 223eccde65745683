"""Pins the oracle's MGFM restatement (S3:mpm.cpp:882-926, T2:mpm.cpp:1373-1416)
against closed forms that do not share its code:

* the root p* of the two-shock function in Toro's form
  f_K(p) = (p - p_K) sqrt(A_K / (p + B_K)) (stiffened gas: p -> p + P_inf),
  solved independently with scipy's brentq;
* the Rankine-Hugoniot shock density ratio at p*;
* equal states (p* = p, u* = u after one Newton step);
* the p* >= 1e-8 clamp for a strong rarefaction and the iteration cap.
"""
import numpy as np
import pytest
from scipy.optimize import brentq

from oracle import pyoracle
from pgdg_p2_amd import workloads


def toro_f(p, pk, rk, g, pinf):
    A = 2.0 / ((g + 1.0) * rk)
    B = (g - 1.0) / (g + 1.0) * (pk + pinf) + pinf  # shift by P_inf
    return (p - pk) * np.sqrt(A / (p + B))


def two_shock_root(ul, ur, pl, pr, rl, rr, gl, pil, gr, pir):
    F = lambda p: toro_f(p, pl, rl, gl, pil) + toro_f(p, pr, rr, gr, pir) + (ur - ul)
    lo, hi = 1e-8, max(pl, pr) * 10 + 1.0
    while F(hi) < 0:
        hi *= 10
    if F(lo) > 0:
        return lo
    return brentq(F, lo, hi, xtol=1e-300, rtol=1e-15, maxiter=500)


def test_relative_mode_matches_independent_root():
    inp = workloads.mgfm_faces(400, seed=11)
    gl, pil, gr, pir = workloads.MGFM_EOS
    out, it = pyoracle.mgfm_batch(inp, gl, pil, gr, pir, variant=0, mode=1, rtol=1e-14,
                                  max_iter=200)
    for i in range(inp.shape[1]):
        ref = two_shock_root(*inp[:, i], gl, pil, gr, pir)
        if ref <= 1e-8:
            assert out[1, i] == pytest.approx(1e-8, rel=1e-6)
            continue
        assert out[1, i] == pytest.approx(ref, rel=1e-10), i


def test_shock_density_ratio_closed_form():
    inp = workloads.mgfm_faces(200, seed=12)
    gl, pil, gr, pir = workloads.MGFM_EOS
    out, _ = pyoracle.mgfm_batch(inp, gl, pil, gr, pir, variant=0, mode=1, rtol=1e-14,
                                 max_iter=200)
    ps, pl, rl = out[1], inp[2], inp[4]
    ratio = ((gl + 1) * ps + (gl - 1) * pl + 2 * gl * pil) / ((gl - 1) * ps + (gl + 1) * pl + 2 * gl * pil)
    np.testing.assert_allclose(out[0], rl * ratio, rtol=1e-12)
    # u*_L = u_L - (p* - p_L)/W_L
    W = np.sqrt(0.5 * rl * ((gl + 1) * ps + (gl - 1) * pl + 2 * gl * pil))
    np.testing.assert_allclose(out[2], inp[0] - (ps - pl) / W, rtol=1e-9, atol=1e-9)


@pytest.mark.parametrize("variant", [0, 1])
def test_equal_states(variant):
    n = 50
    rng = np.random.default_rng(3)
    u = rng.uniform(-10, 10, n)
    p = np.exp(rng.uniform(0, 10, n))
    r = np.exp(rng.uniform(-2, 2, n))
    inp = np.stack([u, u, p, p, r, r])
    out, it = pyoracle.mgfm_batch(inp, 1.4, 0.0, 1.4, 0.0, variant=variant, mode=0)
    assert np.all(it == 1)
    np.testing.assert_array_equal(out[1], p)      # dp = 0 exactly
    np.testing.assert_allclose(out[0], r, rtol=4e-16)  # rho*_L(p*=p) = rho_L up to rounding
    np.testing.assert_array_equal(out[2], u)


def test_reference_tolerance_residual():
    # Reference rule: the loop stops once |f(p*)| <= max(1e-8, 1e-3|uL-uR|)
    # at the pre-update p*; check the two-shock residual at the returned
    # p* against the relative-mode root.
    inp = workloads.mgfm_faces(300, seed=13)
    gl, pil, gr, pir = workloads.MGFM_EOS
    ref, it_ref = pyoracle.mgfm_batch(inp, gl, pil, gr, pir, variant=0, mode=0)
    tight, _ = pyoracle.mgfm_batch(inp, gl, pil, gr, pir, variant=0, mode=1, rtol=1e-14,
                                   max_iter=200)
    conv = it_ref < 50
    assert conv.mean() > 0.95
    # Newton is quadratically convergent near the root: the step after a
    # residual below tol lands well inside it.
    rel = np.abs(ref[1] - tight[1]) / tight[1]
    assert np.median(rel[conv]) < 1e-6


def test_strong_rarefaction_clamps_pstar():
    inp = np.array([[-500.0], [500.0], [1e3], [1e3], [1.0], [1.0]])
    out, it = pyoracle.mgfm_batch(inp, 1.4, 0.0, 1.4, 0.0, variant=0, mode=0)
    assert out[1, 0] == 1e-8
    out, it = pyoracle.mgfm_batch(inp, 1.4, 0.0, 1.4, 0.0, variant=0, mode=0, max_iter=3)
    assert it[0] == 3


def test_t2_variant_initial_guess_and_usl():
    # T2 starts from (pL+pR)/2 and evaluates u*_L with the updated p* and the
    # last W_L; with max_iter=1 both are closed forms.
    ul, ur, pl, pr, rl, rr = 1.0, -1.0, 2.0, 1.0, 1.0, 0.5
    g = 1.4
    inp = np.array([[ul], [ur], [pl], [pr], [rl], [rr]])
    out, it = pyoracle.mgfm_batch(inp, g, 0.0, g, 0.0, variant=1, mode=0, max_iter=1)
    p0 = (pl + pr) * 0.5
    W = lambda r, p, ps: np.sqrt(0.5 * r * ((g + 1) * ps + (g - 1) * p))
    Wl, Wr = W(rl, pl, p0), W(rr, pr, p0)
    f = ul - ur - (p0 - pl) / Wl - (p0 - pr) / Wr
    dK = lambda w, p, r, ps: 1 / w - (ps - p) / w ** 2 * r * (g + 1) / 4 / w
    J = -dK(Wl, pl, rl, p0) - dK(Wr, pr, rr, p0)
    dp = -f / J
    if abs(dp) > 0.6 * p0:          # step clamp of the reference loop
        dp = np.copysign(0.6 * p0, dp)
    p1 = max(p0 + dp, 1e-8)
    assert out[1, 0] == pytest.approx(p1, rel=1e-15)
    assert out[2, 0] == pytest.approx(ul - (p1 - pl) / Wl, rel=1e-14)
    # Jacobian against a central finite difference of the residual
    fr = lambda ps: ul - ur - (ps - pl) / W(rl, pl, ps) - (ps - pr) / W(rr, pr, ps)
    h = 1e-6
    assert J == pytest.approx((fr(p0 + h) - fr(p0 - h)) / (2 * h), rel=1e-7)
