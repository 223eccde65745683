"""Closed-form checks of the oracle's P2 / SSP-RK3 / per-cell-bound
extensions (no reference routine exists for them).

* DG right-hand side: a state that is a global polynomial of degree <= 2 in
  one coordinate (or the xy cross term), advected by a uniform velocity at
  uniform pressure, is continuous across faces, so the HLLC face states
  agree and the 3-point rules are exact: the rates of every mode follow from
  d(rho)/dt = -u . grad(rho) in closed form (interior cells).
* G2P: the cell polynomial evaluated at a particle reproduces the global
  quadratic; the particle's copy holds the gradient at the particle.
* P2G: particles carrying an exact quadratic give the exact cell mean (the
  centre value plus the square-mode constants) and the square coefficient.
* SSP-RK3: the three stages on a state whose rate is constant advance the
  mean by dt * rate (the Shu-Osher weights; a wrong weight changes it).
* Per-cell bounds: the bounded run merges fewer particles than the
  unbounded one on the same lattice.
"""
import numpy as np
import pytest

from oracle.pyoracle_step import OracleSim
from pgdg_p2_amd import cases


def _centres(case):
    D = case.dim
    n = [case.n[a] for a in range(3)]
    ps = [(case.box_r[a] - case.box_l[a]) / n[a] if a < D else 1.0 for a in range(3)]
    i, j, k = np.meshgrid(np.arange(n[0]), np.arange(n[1]), np.arange(n[2]), indexing="ij")
    xc = [case.box_l[a] + ps[a] * ((i, j, k)[a].reshape(-1) + 0.5) if a < D else np.zeros(i.size)
          for a in range(3)]
    return np.array(xc), ps, (i.reshape(-1), j.reshape(-1), k.reshape(-1))


def _interior(case, idx):
    D = case.dim
    m = np.ones(idx[0].size, dtype=bool)
    for a in range(D):
        m &= (idx[a] >= 1) & (idx[a] <= case.n[a] - 2)
    return m


def _one_phase_cells(case, st, rho_modes, vel, p0):
    """Phase 0 present everywhere with density modes rho_modes[m][nc],
    momentum vel * rho, E = p0/(g-1) + |vel|^2/2 rho; other phases absent."""
    D, U, M = case.dim, case.U, case.M
    g = case.gamma[0]
    c = st.c
    c["coef"][:] = 0
    for m in range(M):
        c["coef"][0, 0, m] = rho_modes[m]
        for a in range(D):
            c["coef"][0, 1 + a, m] = vel[a] * rho_modes[m]
        c["coef"][0, U - 1, m] = 0.5 * float(np.dot(vel[:D], vel[:D])) * rho_modes[m]
    c["coef"][0, U - 1, 0] += p0 / (g - 1)
    c["coef0"][:] = c["coef"]
    c["alph"][:] = 0
    c["alph"][0] = 1
    c["ph_state"][:] = 0
    c["ph_state"][0] = 1
    c["ph_state_b"][:] = c["ph_state"]
    c["lim"][:] = 1.0
    c["rho"][0] = rho_modes[0]
    for a in range(D):
        c["rhov"][0, a] = vel[a] * rho_modes[0]
    c["E"][0] = c["coef"][0, U - 1, 0]


def _rates(case, st):
    o = OracleSim(case, st)
    o.call("used_state_update")
    o.call("rhs")
    out = st.copy()
    o.save(out)
    return out.c["dcoef"][0]


P2 = dict(order=2, rk_stages=3)


@pytest.mark.parametrize("axis", [0, 1])
def test_p2_rhs_advected_quadratic_2d(axis):
    case, st = cases.triple_point_2d(nx=10, ny=8, ppc_axis=1, jitter=0.0, scheme=P2)
    xc, ps, idx = _centres(case)
    a1, b2, u0, p0 = 0.05, 0.02, 0.7, 1.0
    M = case.M
    modes = [np.zeros(case.nc) for _ in range(M)]
    q12 = ps[axis] ** 2 / 12
    t = xc[axis]
    modes[0] = 1 + a1 * t + b2 * t * t + b2 * q12
    modes[1 + axis] = a1 + 2 * b2 * t
    modes[4 + axis] = b2 * np.ones(case.nc)          # x^2 or y^2 mode
    vel = np.zeros(3)
    vel[axis] = u0
    _one_phase_cells(case, st, modes, vel, p0)
    dc = _rates(case, st)
    m = _interior(case, idx)
    rate0 = -u0 * (a1 + 2 * b2 * t)                  # rho rate at the centre
    fac = [1.0, vel[0], vel[1], 0.5 * u0 * u0]
    for u in range(case.U):
        if fac[u] == 0:
            continue
        exp = np.zeros((M, case.nc))
        exp[0] = fac[u] * rate0
        exp[1 + axis] = fac[u] * (-2 * u0 * b2)
        got = dc[u][:, m]
        np.testing.assert_allclose(got, exp[:, m], rtol=0, atol=2e-12)


def test_p2_rhs_cross_term_2d():
    case, st = cases.triple_point_2d(nx=10, ny=8, ppc_axis=1, jitter=0.0, scheme=P2)
    xc, ps, idx = _centres(case)
    cxy, u0 = 0.03, 0.6
    M = case.M
    modes = [np.zeros(case.nc) for _ in range(M)]
    modes[0] = 1 + cxy * xc[0] * xc[1]
    modes[1] = cxy * xc[1]
    modes[2] = cxy * xc[0]
    modes[3] = cxy * np.ones(case.nc)
    _one_phase_cells(case, st, modes, np.array([u0, 0, 0]), 1.0)
    dc = _rates(case, st)[0]
    m = _interior(case, idx)
    np.testing.assert_allclose(dc[0][m], (-u0 * cxy * xc[1])[m], atol=2e-12)
    np.testing.assert_allclose(dc[2][m], -u0 * cxy, atol=2e-12)
    for mode in (1, 3, 4, 5):
        np.testing.assert_allclose(dc[mode][m], 0, atol=2e-12)


@pytest.mark.parametrize("axis", [0, 2])
def test_p2_rhs_advected_quadratic_3d(axis):
    case, st = cases.shock_bubble_3d(nx=8, ny=4, nz=4, ppc_axis=1, jitter=0.0, scheme=P2)
    xc, ps, idx = _centres(case)
    a1, b2, u0, p0 = 0.05, 0.02, 0.5, 1.0
    M = case.M
    modes = [np.zeros(case.nc) for _ in range(M)]
    q12 = ps[axis] ** 2 / 12
    t = xc[axis]
    modes[0] = 1 + a1 * t + b2 * t * t + b2 * q12
    modes[1 + axis] = a1 + 2 * b2 * t
    modes[7 + axis] = b2 * np.ones(case.nc)
    vel = np.zeros(3)
    vel[axis] = u0
    _one_phase_cells(case, st, modes, vel, p0)
    dc = _rates(case, st)[0]
    m = _interior(case, idx)
    exp = np.zeros((M, case.nc))
    exp[0] = -u0 * (a1 + 2 * b2 * t)
    exp[1 + axis] = -2 * u0 * b2
    np.testing.assert_allclose(dc[:, m], exp[:, m], atol=2e-12)


def _quadratic_particles(case, st, a1, b2, axis):
    """Every particle in phase 1 carrying rho = 1 + a1 t + b2 t^2 (t = x_axis)
    with its copy holding the gradient at the particle."""
    D, U, M, n = case.dim, case.U, case.M, st.n
    P = st.p
    P["phase"][:n] = 1
    t = P["x"][axis, :n]
    rho = 1 + a1 * t + b2 * t * t
    g = case.gamma[0]
    P["rho"][:n] = rho
    P["rhov"][:, :n] = 0
    P["v"][:, :n] = 0
    P["tv"][:, :n] = 0
    P["E"][:n] = 1.0 / (g - 1)
    P["coef"][:, :n] = 0
    P["coef"][0 * M + 0, :n] = rho
    P["coef"][0 * M + 1 + axis, :n] = a1 + 2 * b2 * t
    P["coef"][0 * M + (4 if D == 2 else 7) + axis, :n] = b2
    P["coef"][(U - 1) * M + 0, :n] = P["E"][:n]
    return rho


def test_p2_p2g_exact_quadratic_mean_2d():
    case, st = cases.triple_point_2d(nx=10, ny=8, ppc_axis=2, jitter=0.2, scheme=P2)
    xc, ps, idx = _centres(case)
    a1, b2 = 0.04, 0.03
    _quadratic_particles(case, st, a1, b2, 0)
    o = OracleSim(case, st)
    for fn in ("used_state_update", "bin", "membership", "p2g"):
        o.call(fn)
    out = st.copy()
    o.save(out)
    cf = out.c["coef"][0, 0]
    t = xc[0]
    np.testing.assert_allclose(cf[0], 1 + a1 * t + b2 * t * t + b2 * ps[0] ** 2 / 12, atol=1e-12)
    np.testing.assert_allclose(cf[4], b2, atol=1e-13)
    np.testing.assert_allclose(cf[3], 0, atol=1e-13)


def test_p2_g2p_reproduces_cell_quadratic_2d():
    case, st = cases.triple_point_2d(nx=10, ny=8, ppc_axis=2, jitter=0.2, scheme=P2)
    xc, ps, idx = _centres(case)
    a1, b2 = 0.05, 0.02
    M = case.M
    modes = [np.zeros(case.nc) for _ in range(M)]
    t = xc[0]
    modes[0] = 1 + a1 * t + b2 * t * t + b2 * ps[0] ** 2 / 12
    modes[1] = a1 + 2 * b2 * t
    modes[4] = b2 * np.ones(case.nc)
    _one_phase_cells(case, st, modes, np.zeros(3), 1.0)
    st.p["phase"][: st.n] = 1
    st.p["tv"][:, : st.n] = 0
    o = OracleSim(case, st)
    for fn in ("used_state_update", "bin", "membership", "neighbors"):
        o.call(fn)
    o.call("update_particle", 0)
    out = st.copy()
    o.save(out)
    x = out.live("x")[0]
    np.testing.assert_allclose(out.live("rho"), 1 + a1 * x + b2 * x * x, atol=1e-12)
    cpy = out.live("coef")
    np.testing.assert_allclose(cpy[1], a1 + 2 * b2 * x, atol=1e-13)   # gradient at the particle
    np.testing.assert_allclose(cpy[4], b2, atol=1e-15)


def test_rk3_constant_rate_advances_by_dt_times_rate():
    # a linear density advected at uniform velocity and pressure has a
    # constant rate -u a; the three Shu-Osher stages must advance the mean
    # by exactly dt * rate (a wrong stage weight gives 1/2 or 3/4 of it)
    case, st = cases.triple_point_2d(nx=10, ny=8, ppc_axis=1, jitter=0.0, scheme=P2)
    xc, ps, idx = _centres(case)
    a1, u0, dt = 0.05, 0.7, 0.01
    M = case.M
    modes = [np.zeros(case.nc) for _ in range(M)]
    modes[0] = 1 + a1 * xc[0]
    modes[1] = np.full(case.nc, a1)
    _one_phase_cells(case, st, modes, np.array([u0, 0, 0]), 1.0)
    st.s.dt = dt
    before = st.c["coef"][0, 0].copy()
    o = OracleSim(case, st)
    o.call("used_state_update")
    for md in (0, 2, 3):
        o.call("rhs")
        o.call("update_grid", md)
    out = st.copy()
    o.save(out)
    m = _interior(case, idx)
    m &= (idx[0] >= 3) & (idx[0] <= case.n[0] - 4)   # away from the walls' influence
    after = out.c["coef"][0, 0]
    np.testing.assert_allclose(after[0][m], (before[0] - dt * u0 * a1)[m], rtol=0, atol=1e-13)
    np.testing.assert_allclose(after[1][m], a1, atol=1e-13)


def test_per_cell_bounds_reduce_merges():
    runs = []
    for sch in (None, dict(ppc_min=1, ppc_max=4)):
        case, st = cases.triple_point_2d(nx=112, ny=48, ppc_axis=1, jitter=0.0, scheme=sch)
        o = OracleSim(case, st)
        a = st.copy()
        merges = 0
        for _ in range(30):
            o.step(1)
            o.save(a)
            merges += a.s.n_merge
        runs.append(merges)
    assert runs[0] > 0 and runs[1] < runs[0], runs
