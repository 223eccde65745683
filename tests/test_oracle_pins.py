"""Closed-form, brute-force and invariant checks of the oracle's step routines.

The GPU step is compared bit for bit with the oracle (tests/test_gpu_*.py);
these tests pin the oracle itself, without a reference build, by checks that
a plausible mistake in the restatement (a wrong index, sign, weight, order or
threshold) fails:

* binning (S3:mpm.cpp:65-96): brute-force floor/clamp cells, fluid entries in
  particle order then walls, particles outside the box clamped;
* neighbour lists, status 2 and 3 (S3:level_infor.cpp:153-195): every entry
  within the cut-off found by an O(N^2) search, ordered by cell (i, j, k
  ascending) and then bin order;
* Wendland kernel (S3:particle.cpp:131-149, 297-315): integral pi/PI over the
  support (PI = 3.1415926, S3:glbparam.h:36), derivative against a central
  difference, zero beyond 2h;
* Limiter and Calculate_limiter (S3:glbfunc.h:283-299, S3:mpm.cpp:392-738):
  the map's special values, Limiter(0) = 1 through the whole routine on a
  uniform state, no limiting of a linear field, the 2 ds branch on a steep
  slope, the positivity probe;
* HLLC / HLL (S3:particle_flow.cpp:232-272, 367-381): consistency
  F(U, U) = F(U), supersonic upwinding, exact resolution of an isolated
  contact, conservation between the two sides of a face;
* Set_timestep (S3:mpm.cpp:19-57): the closed form on uniform states;
* Split_particle_fast (S3:mpm.cpp:1036-1417, S3:particle.cpp:575-609):
  children reproduce a global linear (P1) or quadratic (P2) field, conserve
  mass, volume, momentum and energy, sit at the reference offsets in the
  reference order; count identity;
* Merge_particle_fast (S3:mpm.cpp:1536-1706, S3:particle.cpp:614-661): the
  mutual-nearest pairs of a brute-force search, stable removal, conservation;
* the whole step: n' = n + (2^D - 1) splits + inlet - outlet - merges.
"""
import math

import numpy as np
import pytest
from scipy import integrate

from oracle import pyoracle
from oracle.pyoracle_step import OracleSim, face_flux, kernel, limiter_fn
from pgdg_p2_amd import cases

PI_REF = 3.1415926


# ---------------------------------------------------------------------------
# binning and neighbour lists
# ---------------------------------------------------------------------------
def _cells_of(case, x):
    """Brute-force get_cell_id: floor((x - box_l) / dcell), clamped."""
    D = case.dim
    N = [case.n[a] for a in range(3)]
    idx = np.zeros(x.shape[1], dtype=np.int64)
    for a in range(3):
        if a < D:
            dcell = (case.box_r[a] - case.box_l[a]) / N[a]
            p = np.floor((x[a] - case.box_l[a]) / dcell).astype(np.int64)
            p = np.clip(p, 0, N[a] - 1)
        else:
            p = np.zeros_like(idx)
        idx = idx * N[a] + p
    return idx


def _brute_bins(case, st):
    """Entries per cell: fluid particles in particle order, then walls."""
    n, nw = st.n, st.nwall
    cf = _cells_of(case, st.p["x"][:, :n])
    cw = _cells_of(case, st.w["x"])
    keys = np.concatenate([cf, cw])
    refs = np.concatenate([np.arange(n), -(np.arange(nw) + 2)])
    rank = np.arange(n + nw)  # fluid first, then walls, each in index order
    order = np.lexsort((rank, keys))
    off = np.zeros(case.nc + 1, dtype=np.int64)
    np.add.at(off, keys + 1, 1)
    return np.cumsum(off), refs[order], keys[order]


def _small_2d(jitter=0.25, nx=12, ny=6, ppc=3, seed=3):
    case, st = cases.triple_point_2d(nx=nx, ny=ny, ppc_axis=ppc, jitter=jitter, seed=seed)
    return case, st


def _small_3d(nx=8, ny=4, nz=4, ppc=2):
    return cases.shock_bubble_3d(nx=nx, ny=ny, nz=nz, ppc_axis=ppc, jitter=0.25)


@pytest.mark.parametrize("dim", [2, 3])
def test_binning_matches_brute_force(dim):
    case, st = _small_2d() if dim == 2 else _small_3d()
    # a few particles outside the box on either side are clamped into the
    # boundary cells
    rng = np.random.default_rng(11)
    k = rng.choice(st.n, 6, replace=False)
    st.p["x"][0, k[:3]] = case.box_l[0] - 0.3 * case.p_g
    st.p["x"][1, k[3:]] = case.box_r[1] + 0.7 * case.p_g
    o = OracleSim(case, st)
    o.call("bin")
    off, ent = o.bins(case.nc, st.n + st.nwall)
    boff, bent, _ = _brute_bins(case, st)
    assert np.array_equal(off, boff)
    assert np.array_equal(ent, bent)


@pytest.mark.parametrize("dim,status", [(2, 2), (2, 3), (3, 2), (3, 3)])
def test_neighbor_lists_match_brute_force(dim, status):
    case, st = _small_2d() if dim == 2 else _small_3d()
    n = st.n
    # smoothing lengths vary per particle (status 2 cut-off is 2 h_i)
    rng = np.random.default_rng(5)
    st.p["h"][:n] *= rng.uniform(0.8, 1.0, n)
    o = OracleSim(case, st)
    o.call("bin")
    o.call("neighbors", status)
    D = case.dim
    X = np.concatenate([st.p["x"][:D, :n], st.w["x"][:D]], axis=1)
    _, bent, bkeys = _brute_bins(case, st)
    # position of every entry in (cell, bin order)
    pos_of = {}
    for q, r in enumerate(bent):
        pos_of[int(r)] = q
    refs = np.concatenate([np.arange(n), -(np.arange(st.nwall) + 2)])
    order_key = np.array([pos_of[int(r)] for r in refs])
    for i in range(0, n, max(1, n // 60)):
        d = np.sqrt(((X - st.p["x"][:D, i:i + 1]) ** 2).sum(axis=0))
        lim = st.p["h"][i] * 2.0 if status == 2 else case.p_sz * 2.0
        sel = np.nonzero(d <= lim)[0]
        want = refs[sel[np.argsort(order_key[sel])]]
        got = o.neighbor_list(i)
        assert np.array_equal(got, want), (i, got[:8], want[:8])
        assert i in got  # a particle is its own neighbour at distance 0


# ---------------------------------------------------------------------------
# Wendland kernel
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("dim", [2, 3])
def test_kernel_integral_derivative_support(dim):
    h = 0.37
    W = lambda r: kernel(dim, r, h)[0]
    if dim == 2:
        val, _ = integrate.quad(lambda r: 2 * math.pi * r * W(r), 0, 2 * h, epsabs=1e-14,
                                epsrel=1e-13)
    else:
        val, _ = integrate.quad(lambda r: 4 * math.pi * r * r * W(r), 0, 2 * h, epsabs=1e-14,
                                epsrel=1e-13)
    # normalised up to the reference's PI constant
    assert abs(val - math.pi / PI_REF) < 1e-11
    for r in (0.05, 0.2, 0.4, 0.6, 0.73):
        e = 1e-6
        fd = (W(r + e) - W(r - e)) / (2 * e)
        assert abs(kernel(dim, r, h)[1] * r - fd) < 1e-6 * max(1.0, abs(fd))
    assert kernel(dim, 2 * h, h)[0] == 0.0 and kernel(dim, 2.5 * h, h)[0] == 0.0
    assert kernel(dim, 2 * h * (1 - 1e-9), h)[0] > 0.0


# ---------------------------------------------------------------------------
# Limiter
# ---------------------------------------------------------------------------
def test_limiter_map_special_values():
    assert limiter_fn(0.0, 0) == 1.0            # Limiter(0) = 1 (S3:glbfunc.h:289-291)
    assert limiter_fn(0.25, 0) == 0.5           # ds in (0, 0.5): 2 ds
    assert limiter_fn(0.5, 0) == 1.0 and limiter_fn(3.0, 0) == 1.0
    assert limiter_fn(1.0, 1) == 1.0            # boundary or mixed cell: 2 ds / (ds + 1)
    assert abs(limiter_fn(0.5, 1) - 2.0 / 3.0) < 1e-16
    assert limiter_fn(0.0, 1) == 0.0
    assert abs(limiter_fn(3.0, 1) - 1.5) < 1e-16


def _uniform_cells(case, st, rho=1.0, p=1.0):
    """Phase 0 everywhere, uniform at rest."""
    F, nc = case.nphase, case.nc
    c = st.c
    c["alph"][:] = 0.0
    c["alph"][0] = 1.0
    c["ph_state"][:] = 0
    c["ph_state"][0] = 1
    g = case.gamma[0]
    E = p / (g - 1)
    c["coef"][:] = 0.0
    c["coef"][0, 0, 0] = rho
    c["coef"][0, case.U - 1, 0] = E
    c["rho"][0] = rho
    c["E"][0] = E
    c["rhov"][:] = 0.0
    c["lim"][:] = 1.0
    return c


def _limiter_run(case, st):
    o = OracleSim(case, st)
    o.call("used_state_update")
    o.call("limiter")
    out = st.copy()
    o.save(out)
    return out.c["lim"][0], out


def test_limiter_uniform_state_is_unlimited():
    # zero differences everywhere: every ratio is 0; Limiter(0) = 1 in the
    # interior, while boundary cells use 2 ds / (ds + 1) = 0
    case, st = _small_2d(nx=8, ny=6)
    _uniform_cells(case, st)
    lim, _ = _limiter_run(case, st)
    _, i, j = _cell_x(case)
    inner = (i > 0) & (i < case.n[0] - 1) & (j > 0) & (j < case.n[1] - 1)
    assert np.all(lim[:, inner] == 1.0)
    assert np.all(lim[:, ~inner] == 0.0)


def _cell_x(case):
    nx, ny = case.n[0], case.n[1]
    dx = (case.box_r[0] - case.box_l[0]) / nx
    i = np.arange(case.nc) // ny
    j = np.arange(case.nc) % ny
    return (i + 0.5) * dx, i, j


@pytest.mark.parametrize("slope_factor,expect", [(1.0, 1.0), (4.0, 0.5)])
def test_limiter_linear_density(slope_factor, expect):
    # cell means of rho = 1 + g x; slope g (consistent: ratio 1 -> 1) or 4 g
    # (ratio 1/4 -> 2 * 1/4 = 1/2) in the interior cells
    case, st = _small_2d(nx=8, ny=6)
    c = _uniform_cells(case, st)
    xc, i, j = _cell_x(case)
    g = 0.05
    c["rho"][0] = 1.0 + g * xc
    c["coef"][0, 0, 0] = c["rho"][0]
    c["coef"][0, 0, 1] = slope_factor * g
    lim, _ = _limiter_run(case, st)
    inner = (i > 0) & (i < case.n[0] - 1) & (j > 0) & (j < case.n[1] - 1)
    assert np.allclose(lim[0][inner], expect, rtol=0, atol=1e-12)
    assert np.all(lim[1:][:, inner] == 1.0)  # momentum and energy are uniform


def test_limiter_positivity_probe():
    # uniform means, one cell with a slope that makes rho negative at the
    # probe points (+-p_g, +-p_g): every variable of the phase is limited to
    # 0; a cell with half that slope keeps 1 (its ratios are 0 -> Limiter 1)
    case, st = _small_2d(nx=8, ny=6)
    c = _uniform_cells(case, st)
    a = 3 * case.n[1] + 2
    b = 5 * case.n[1] + 3
    c["coef"][0, 0, 1, a] = 1.5 / case.p_g
    c["coef"][0, 0, 1, b] = 0.5 / case.p_g
    lim, _ = _limiter_run(case, st)
    assert np.all(lim[:, a] == 0.0)
    assert np.all(lim[:, b] == 1.0)
    _, i, j = _cell_x(case)
    others = (i > 0) & (i < case.n[0] - 1) & (j > 0) & (j < case.n[1] - 1)
    others[a] = False
    assert np.all(lim[:, others] == 1.0)


# ---------------------------------------------------------------------------
# HLLC / HLL
# ---------------------------------------------------------------------------
def _state(dim, rho, v, p, gamma=1.4):
    v = np.asarray(v, dtype=float)
    rv = np.zeros(3)
    rv[:dim] = rho * v[:dim]
    E = p / (gamma - 1) + 0.5 * rho * float(v[:dim] @ v[:dim])
    c = math.sqrt(gamma * p / rho)
    return np.array([rho, rv[0], rv[1], rv[2], E, p, c]), v


def _euler_flux(dim, rho, v, p, gamma=1.4):
    """F(U) as a U x D tensor: U v^T + p I (+ p v in the energy row)."""
    U = dim + 2
    F = np.zeros((U, dim))
    E = p / (gamma - 1) + 0.5 * rho * float(v[:dim] @ v[:dim])
    F[0] = rho * v[:dim]
    for a in range(dim):
        F[1 + a] = rho * v[a] * v[:dim]
        F[1 + a, a] += p
    F[U - 1] = (E + p) * v[:dim]
    return F


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("kind", [0, 1])
def test_face_flux_consistency_and_upwinding(dim, kind):
    rng = np.random.default_rng(dim * 10 + kind)
    for trial in range(20):
        n = np.zeros(3)
        ax = trial % dim
        n[ax] = 1.0 if trial % 2 else -1.0
        rho, p = rng.uniform(0.2, 3.0), rng.uniform(0.1, 5.0)
        v = np.zeros(3)
        v[:dim] = rng.uniform(-0.5, 0.5, dim)
        st, v = _state(dim, rho, v, p)
        F = face_flux(dim, kind, n, st, st)
        # F(U, U) = F(U) contracted with n
        Fe = _euler_flux(dim, rho, v, p)
        assert np.allclose(F @ n[:dim], Fe @ n[:dim], rtol=1e-13, atol=1e-13)
        # supersonic towards +n: upwind left state; towards -n: right state
        c = st[6]
        vl = np.zeros(3)
        vl[:dim] = rng.uniform(-0.1, 0.1, dim)
        vl[ax] = n[ax] * 3.0 * c
        stl, vl = _state(dim, rho, vl, p)
        str_, vr = _state(dim, 0.5 * rho, vl, 0.6 * p)
        F = face_flux(dim, kind, n, stl, str_)
        assert np.allclose(F @ n[:dim], _euler_flux(dim, rho, vl, p) @ n[:dim], rtol=1e-13,
                           atol=1e-13)
        vr2 = -vl
        stl2, _ = _state(dim, rho, vr2, p)
        str2, vr2 = _state(dim, 0.5 * rho, vr2, 0.6 * p)
        F = face_flux(dim, kind, n, stl2, str2)
        assert np.allclose(F @ n[:dim], _euler_flux(dim, 0.5 * rho, vr2, 0.6 * p) @ n[:dim],
                           rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("dim", [2, 3])
def test_hllc_resolves_isolated_contact(dim):
    # equal pressure and normal velocity, different densities and tangential
    # velocities: the exact flux is F of the upwind side (S* = u.n); HLLC
    # (unlike HLL) reproduces it, which fixes the star-state formulas
    rng = np.random.default_rng(7 + dim)
    for trial in range(20):
        n = np.zeros(3)
        n[trial % dim] = 1.0
        un = rng.uniform(-0.3, 0.3)
        p = rng.uniform(0.2, 2.0)
        vl = rng.uniform(-0.4, 0.4, 3)
        vr = rng.uniform(-0.4, 0.4, 3)
        vl[trial % dim] = un
        vr[trial % dim] = un
        vl[dim:] = vr[dim:] = 0
        stl, vl = _state(dim, rng.uniform(0.5, 2.0), vl, p)
        str_, vr = _state(dim, rng.uniform(0.05, 0.4), vr, p)
        F = face_flux(dim, 0, n, stl, str_) @ n[:dim]
        up = (stl, vl) if un >= 0 else (str_, vr)
        want = _euler_flux(dim, up[0][0], up[1], p) @ n[:dim]
        assert np.allclose(F, want, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("kind", [0, 1])
def test_face_flux_is_conservative(dim, kind):
    # the flux leaving L through n equals the flux entering R through -n:
    # F(L, R, n).n == F(R, L, -n).n
    rng = np.random.default_rng(99 + dim + kind)
    for trial in range(40):
        n = np.zeros(3)
        n[trial % dim] = 1.0
        stl, _ = _state(dim, rng.uniform(0.1, 3), rng.uniform(-1, 1, 3), rng.uniform(0.05, 5))
        str_, _ = _state(dim, rng.uniform(0.1, 3), rng.uniform(-1, 1, 3), rng.uniform(0.05, 5))
        a = face_flux(dim, kind, n, stl, str_) @ n[:dim]
        b = face_flux(dim, kind, -n, str_, stl) @ n[:dim]
        assert np.allclose(a, b, rtol=1e-12, atol=1e-12)


# ---------------------------------------------------------------------------
# Set_timestep
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("dim", [2, 3])
def test_timestep_closed_form(dim):
    case, st = _small_2d() if dim == 2 else _small_3d()
    n = st.n
    u = np.zeros(3)
    u[:dim] = np.array([0.3, -0.2, 0.1])[:dim]
    rho, p, g = 1.0, 1.0, case.gamma[0]
    st.p["phase"][:n] = 1
    st.p["rho"][:n] = rho
    st.p["v"][:dim, :n] = u[:dim, None]
    c = math.sqrt(g * p / rho)
    st.p["c"][:n] = c
    st.p["h"][:n] = 1.6 * case.p_sz
    o = OracleSim(case, st)
    o.call("set_timestep")  # iteration 1: particles, no acceleration limit
    out = st.copy()
    o.save(out)
    speed = float(np.sqrt(u @ u))
    assert abs(out.s.dt - 0.3 * case.p_sz / (c + speed)) < 1e-15 * out.s.dt * 4
    assert abs(out.s.max_v - speed) < 1e-15
    # later iterations: cells in use, with their acceleration limit
    st2 = out.copy()
    _uniform_cells(case, st2)
    st2.c["v"][:dim] = u[:dim, None]
    st2.c["c"][0] = c
    st2.c["c"][1] = 0.5 * c
    acc = 2.0
    st2.c["ac"][0] = acc
    st2.s.iterate_num = 5
    o2 = OracleSim(case, st2)
    o2.call("used_state_update")
    o2.call("set_timestep")
    out2 = st2.copy()
    o2.save(out2)
    want = min(0.3 * case.p_sz / (c + speed), 0.3 * math.sqrt(case.p_sz / acc))
    assert abs(out2.s.dt - want) < 4e-15 * want


# ---------------------------------------------------------------------------
# Split
# ---------------------------------------------------------------------------
def _poly(case, order, rng):
    """Global polynomials f_u(x) (u = rho, rho v, E): a + b.x (+ x^T Q x)."""
    D, U = case.dim, case.U
    a = np.array([1.0] + [0.2] * D + [3.0])[:U]
    b = rng.uniform(-0.3, 0.3, (U, D))
    Q = np.zeros((U, D, D))
    if order == 2:
        for u in range(U):
            A = rng.uniform(-0.2, 0.2, (D, D))
            Q[u] = 0.5 * (A + A.T)
    return a, b, Q


def _poly_eval(a, b, Q, x):
    """values [U][k] and gradients [U][D][k] at positions x [D][k]."""
    val = a[:, None] + b @ x + np.einsum("ak,uab,bk->uk", x, Q, x)
    grad = b[:, :, None] + 2.0 * np.einsum("uab,bk->uak", Q, x)
    return val, grad


def _set_particle_poly(case, st, order, rng):
    """Particle state and coefficient copies of a global polynomial."""
    D, U, M, n = case.dim, case.U, case.M, st.n
    a, b, Q = _poly(case, order, rng)
    x = st.p["x"][:D, :n]
    val, grad = _poly_eval(a, b, Q, x)
    st.p["rho"][:n] = val[0]
    st.p["rhov"][:D, :n] = val[1:1 + D]
    st.p["E"][:n] = val[U - 1]
    coef = st.p["coef"].reshape(U, M, -1)
    coef[:, :, :n] = 0.0
    coef[:, 1:1 + D, :n] = grad
    if order == 2:
        # tail modes after the gradient: cross terms (xy[, xz, yz]) with
        # coefficient 2 Q_ab, squares with Q_aa (psi = r_a r_b, r_a^2)
        pairs = [(0, 1)] if D == 2 else [(0, 1), (0, 2), (1, 2)]
        for q, (i, j) in enumerate(pairs):
            coef[:, 1 + D + q, :n] = 2.0 * Q[:, i, j][:, None]
        for i in range(D):
            coef[:, M - D + i, :n] = Q[:, i, i][:, None]
    return a, b, Q


@pytest.mark.parametrize("dim,order", [(2, 1), (2, 2), (3, 1), (3, 2)])
def test_split_children_reproduce_field_and_conserve(dim, order):
    sch = dict(order=order, rk_stages=2 if order == 1 else 3)
    if dim == 2:
        case, st = cases.triple_point_2d(nx=10, ny=6, ppc_axis=2, jitter=0.25, scheme=sch)
    else:
        case, st = cases.shock_bubble_3d(nx=8, ny=4, nz=4, ppc_axis=2, jitter=0.25, scheme=sch)
    D, U, M, n = case.dim, case.U, case.M, st.n
    rng = np.random.default_rng(17 * dim + order)
    _set_particle_poly(case, st, order, rng)
    ps = case.p_sz ** D if D == 2 else (case.p_sz * case.p_sz) * case.p_sz
    # interior particles only (no inlet copy or outlet removal in 3-D)
    inner = np.nonzero((st.p["x"][0, :n] > 2 * case.p_g)
                       & (st.p["x"][0, :n] < case.box_r[0] - 2 * case.p_g))[0]
    parents = np.sort(rng.choice(inner, 9, replace=False))
    st.p["vol"][:n] = ps
    st.p["vol"][parents] = ps * rng.uniform(1.7, 3.0, parents.size)
    st.p["mass"][:n] = st.p["vol"][:n] * st.p["rho"][:n]
    st.p["mark_in"][:n] = 0
    before = st.copy()
    o = OracleSim(case, st)
    o.call("bin")
    o.call("split")
    out = st.copy()
    o.save(out)
    nch = 2 ** D
    assert out.s.n_split == parents.size
    assert out.n == n + (nch - 1) * parents.size
    a, b, Q = _poly(case, order, np.random.default_rng(17 * dim + order))
    # untouched particles are unchanged
    others = np.setdiff1d(np.arange(n), parents)
    for k in ("x", "rho", "E", "vol", "mass"):
        assert np.array_equal(out.p[k][..., others], before.p[k][..., others])
    for q, p in enumerate(parents):
        dp = pyoracle.root_dim(np.array([before.p["vol"][p]]), D)[0]
        kids = [n + q * (nch - 1) + k for k in range(nch - 1)] + [p]
        # children at (x - dp/4) + dp/2 (a, b[, c]) in i, j[, k] order; the
        # parent takes the last child's place
        for k, idx in enumerate(kids):
            bits = [(k >> (D - 1 - t)) & 1 for t in range(D)]
            want = before.p["x"][:D, p] - dp / 4.0 + dp / 2.0 * np.array(bits)
            assert np.allclose(out.p["x"][:D, idx], want, rtol=0, atol=1e-15)
        xk = out.p["x"][:D, kids]
        val, grad = _poly_eval(a, b, Q, xk)
        assert np.allclose(out.p["rho"][kids], val[0], rtol=1e-13, atol=1e-14)
        assert np.allclose(out.p["rhov"][:D][:, kids], val[1:1 + D], rtol=1e-13, atol=1e-14)
        assert np.allclose(out.p["E"][kids], val[U - 1], rtol=1e-13, atol=1e-14)
        if order == 2:  # the child's copy carries the gradient at the child
            cc = out.p["coef"].reshape(U, M, -1)[:, 1:1 + D][:, :, kids]
            assert np.allclose(cc, grad, rtol=1e-12, atol=1e-13)
        # volume, mass and the volume-weighted state are conserved
        v = out.p["vol"][kids]
        assert abs(v.sum() - before.p["vol"][p]) <= 1e-15 * before.p["vol"][p]
        assert abs(out.p["mass"][kids].sum() - before.p["mass"][p]) <= 1e-15 * abs(before.p["mass"][p])
        assert np.allclose((out.p["x"][:D, kids] * v).sum(axis=1), before.p["x"][:D, p] * v.sum(),
                           rtol=1e-14, atol=1e-15)
        if order == 1:  # symmetric placement: slope terms cancel
            for key, rows in (("rho", None), ("E", None), ("rhov", slice(0, D))):
                arr = out.p[key] if rows is None else out.p[key][rows]
                ref = before.p[key][p] if rows is None else before.p[key][rows, p]
                tot = (arr[..., kids] * v).sum(axis=-1)
                assert np.allclose(tot, ref * before.p["vol"][p], rtol=1e-13, atol=1e-15)


# ---------------------------------------------------------------------------
# Merge
# ---------------------------------------------------------------------------
def test_merge_pairs_match_brute_force_and_conserve():
    case, st = cases.triple_point_2d(nx=12, ny=6, ppc_axis=3, jitter=0.3, seed=21)
    D, n = case.dim, st.n
    ps = case.p_sz * case.p_sz
    vol_min = ps * 2.0 / 3.0
    rng = np.random.default_rng(4)
    st.p["vol"][:n] = ps * rng.uniform(0.4, 0.9, n)  # about half below vol_min
    st.p["mass"][:n] = st.p["vol"][:n] * st.p["rho"][:n]
    o = OracleSim(case, st)
    o.call("bin")
    o.call("neighbors", 2)
    before = st.copy()
    o.save(before)  # same state; lists built from these positions
    o.call("merge")
    out = st.copy()
    o.save(out)
    # brute force: nearest small same-phase particle within the status-2
    # cut-off and closer than vol_min^(1/D) (strict), mutual pairs, the
    # lower index survives
    x = before.p["x"][:D, :n]
    vol, ph = before.p["vol"][:n], before.p["phase"][:n]
    small = vol < vol_min
    dss0 = math.sqrt(vol_min)
    cand = -np.ones(n, dtype=int)
    for i in np.nonzero(small)[0]:
        d = np.sqrt(((x - x[:, i:i + 1]) ** 2).sum(axis=0))
        ok = small & (ph == ph[i]) & (d < dss0) & (d <= 2 * before.p["h"][i])
        ok[i] = False
        if ok.any():
            j = np.nonzero(ok)[0]
            cand[i] = j[np.argmin(d[j])]
    pairs = [(i, cand[i]) for i in range(n) if cand[i] > i and cand[cand[i]] == i]
    assert len(pairs) > 5
    assert out.s.n_merge == len(pairs)
    removed = {p for _, p in pairs}
    keep = np.array([i for i in range(n) if i not in removed])
    assert out.n == n - len(pairs)
    # stable removal: survivors and untouched particles keep their order
    surv = {s: p for s, p in pairs}
    for new, old in enumerate(keep):
        if old in surv:
            p = surv[old]
            v0, v1 = vol[old], vol[p]
            assert abs(out.p["vol"][new] - (v0 + v1)) <= 1e-16 * (v0 + v1)
            assert out.p["mass"][new] == before.p["mass"][old] + before.p["mass"][p]
            want = (x[:, old] * v0 + x[:, p] * v1) / (v0 + v1)
            assert np.allclose(out.p["x"][:D, new], want, rtol=1e-15, atol=1e-15)
        else:
            assert np.array_equal(out.p["x"][:, new], before.p["x"][:, old])
    # totals of the volume-weighted state and mass are conserved
    for key in ("rho", "E"):
        a = (before.p[key][:n] * vol).sum()
        b = (out.p[key][:out.n] * out.p["vol"][:out.n]).sum()
        assert abs(a - b) <= 1e-13 * abs(a)
    a = (before.p["rhov"][:D, :n] * vol).sum(axis=1)
    b = (out.p["rhov"][:D, :out.n] * out.p["vol"][:out.n]).sum(axis=1)
    assert np.allclose(a, b, rtol=1e-13, atol=1e-13)
    assert abs(before.p["mass"][:n].sum() - out.p["mass"][:out.n].sum()) <= 1e-13 * before.p["mass"][:n].sum()


# ---------------------------------------------------------------------------
# whole step
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("scheme", [dict(order=1, rk_stages=2), dict(order=2, rk_stages=3)])
def test_step_count_identity_with_forced_events(scheme):
    from forced_events import shock_bubble_forced
    case, st = shock_bubble_forced(scheme=scheme)
    o = OracleSim(case, st)
    a = st.copy()
    tot = dict(split=0, merge=0, outlet=0)
    for k in range(3):
        n0 = a.n
        o.step(1)
        o.save(a)
        s = a.s
        assert a.n == n0 + 7 * s.n_split + s.n_inlet - s.n_outlet - s.n_merge
        tot["split"] += s.n_split
        tot["merge"] += s.n_merge
        tot["outlet"] += s.n_outlet
    assert tot["split"] > 0 and tot["merge"] > 0 and tot["outlet"] > 0, tot


# ---------------------------------------------------------------------------
# per-cell split/merge bounds (no reference routine)
# ---------------------------------------------------------------------------
def _fluid_counts(o, nc, cap):
    off, ent = o.bins(nc, cap)
    fl = (ent >= 0).astype(np.int64)
    c = np.zeros(nc, dtype=np.int64)
    cell = np.repeat(np.arange(nc), np.diff(off))
    np.add.at(c, cell, fl)
    return c


@pytest.mark.parametrize("lo,hi", [(4, 16), (8, 12)])
def test_per_cell_bounds_hold_after_every_pass(lo, hi):
    from forced_events import shock_bubble_forced
    res = {}
    for bounded in (False, True):
        sch = dict(order=2, rk_stages=3, ppc_min=lo if bounded else 0, ppc_max=hi if bounded else 0)
        case, st = shock_bubble_forced(scheme=sch)
        o = OracleSim(case, st)
        cap = st.cap + st.nwall
        o.call("bin")
        before = _fluid_counts(o, case.nc, cap)
        o.call("split")
        out = st.copy()
        o.save(out)
        after = _fluid_counts(o, case.nc, cap)
        assert out.s.n_inlet == 0
        if bounded:  # no split takes a cell above ppc_max
            assert np.all(after <= np.maximum(hi, before))
        merges = 0
        for p in range(4):
            before = after
            o.call("merge")
            after = _fluid_counts(o, case.nc, cap)
            o.save(out)
            if bounded:  # no merge takes a cell below ppc_min
                assert np.all(after >= np.minimum(lo, before))
        res[bounded] = (out.s.n_split, out.s.n_merge)
    # the bounds refuse some splits and merges of this state
    assert 0 < res[True][0] < res[False][0], res
    assert 0 < res[True][1] < res[False][1], res


# ---------------------------------------------------------------------------
# P2 appearance positivity (no reference routine)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("cxx,fallback", [(-4000.0, True), (-1.0, False)])
def test_p2_appearance_falls_back_to_particle_average(cxx, fallback):
    # every phase appears in the first MLS_P2G (ph_state_b = 0); with a large
    # negative curvature of the particles' density copies the quadratic
    # extrapolation to the cell centre is negative, and the cell takes the
    # plain weighted average of the particles' own values with zero slopes
    case, st = cases.triple_point_2d(nx=8, ny=4, ppc_axis=3, jitter=0.2,
                                     scheme=dict(order=2, rk_stages=3))
    n, U, M = st.n, case.U, case.M
    coef = st.p["coef"].reshape(U, M, -1)
    coef[:, :, :n] = 0.0
    coef[0, 4, :n] = cxx  # x^2 mode of the density row
    o = OracleSim(case, st)
    o.call("used_state_update")
    o.call("bin")
    o.call("membership")
    o.call("p2g")
    out = st.copy()
    o.save(out)
    c = out.c
    ph = c["ph_state"][0] == 1
    assert ph.sum() > 10
    tails = c["coef"][0, :, 1:, :][:, :, ph]
    rho = c["coef"][0, 0, 0, ph]
    # phase 1 carries rho = 1 in the left block and bottom right
    if fallback:
        assert np.all(tails == 0.0)
        assert np.all(rho > 0)
        # the plain average of particle densities 1 (and 1 where mixed with
        # nothing else of phase 1)
        assert np.allclose(rho, 1.0, rtol=0, atol=1e-12)
    else:
        assert np.any(tails != 0.0)
        assert np.all(rho > 0) and not np.allclose(rho, 1.0, rtol=0, atol=1e-12)
