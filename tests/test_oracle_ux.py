"""Checks of the oracle's underwater-explosion routines (PGDG_SEM_UX) by
hand-derived values and an independent finite-volume restatement:

* Calculate_limiter (UX:mpm.cpp:352-587): the map 2 ds / (ds + 1) in every
  cell (UX:glbfunc.h:291-306), so a uniform state is limited to 0 and a
  slope 4 times the neighbour differences to 2 (1/4) / (5/4) = 0.4; a
  boundary neighbour equal to the cell's means with the momentum normal to
  the domain edge set to 0 (UX:mpm.cpp:478-487); G_v's x columns against
  a zero wall velocity (UX:mpm.cpp:388-399);
* MLS_P2G (UX:mpm.cpp:299-334): an absent phase keeps its density and
  momentum (T2 resets them);
* Grid_acceleration (UX:particle_flow.cpp:16-385): with piecewise-constant
  states (P1, zero slopes) the rate of a cell's means is the finite-volume
  sum of HLL fluxes: interior faces HLL between the two cells, boundary faces
  HLL against a ghost copy of the cell with v_y reflected on the bottom side
  (tag 2) and kept elsewhere (tag -1); the HLL function itself is pinned by
  tests/test_oracle_pins.py and by its left-wave jump condition here;
* pgdg_create's argument checks for the UX routines (they run before any
  device call).
"""
import ctypes as C
import math

import numpy as np
import pytest

from oracle.pyoracle_step import OracleSim, face_flux
from pgdg_p2_amd import cases
from pgdg_p2_amd.state import BC_INLET, BC_OUTLET, BC_WALL

G_W, PINF_W = 4.4, 6e8


def _ux(nx=8, **kw):
    return cases.underwater_explosion_2d(nx=nx, **kw)


def _uniform(case, st, rho=1000.0, p=1e5):
    """Phase 0 (water) everywhere, uniform at rest, zero slopes."""
    c = st.c
    c["alph"][:] = 0.0
    c["alph"][0] = 1.0
    c["ph_state"][:] = 0
    c["ph_state"][0] = 1
    c["ph_state_b"][:] = c["ph_state"]
    E = (p + case.gamma[0] * case.pinf[0]) / (case.gamma[0] - 1)
    c["coef"][:] = 0.0
    c["coef"][0, 0, 0] = rho
    c["coef"][0, case.U - 1, 0] = E
    c["rho"][0] = rho
    c["E"][0] = E
    c["rhov"][:] = 0.0
    c["lim"][:] = 1.0
    return c


def _ij(case):
    ny = case.n[1]
    k = np.arange(case.nc)
    return k // ny, k % ny


def _limiter(case, st):
    o = OracleSim(case, st)
    o.call("used_state_update")
    o.call("limiter")
    out = st.copy()
    o.save(out)
    return out


def test_ux_limiter_map_everywhere():
    # zero differences: every ratio is 0 and 2*0/(0+1) = 0 in every cell
    # (T2 keeps 1 in the interior, tests/test_oracle_pins.py)
    case, st = _ux()
    _uniform(case, st)
    lim = _limiter(case, st).c["lim"][0]
    assert np.all(lim == 0.0)


@pytest.mark.parametrize("slope_factor,expect", [(1.0, 1.0), (4.0, 0.4)])
def test_ux_limiter_linear_density(slope_factor, expect):
    # means rho = rho0 + g x; slope g: ratio 1 -> 2/2 = 1; slope 4 g: ratio
    # 1/4 -> (1/2)/(5/4) = 0.4 (T2's map gives 0.5 there)
    case, st = _ux()
    c = _uniform(case, st)
    i, j = _ij(case)
    dx = (case.box_r[0] - case.box_l[0]) / case.n[0]
    xc = (i + 0.5) * dx
    g = 2.0
    c["rho"][0] = 1000.0 + g * xc
    c["coef"][0, 0, 0] = c["rho"][0]
    c["coef"][0, 0, 1] = slope_factor * g
    lim = _limiter(case, st).c["lim"][0]
    inner = (i > 0) & (i < case.n[0] - 1) & (j > 0) & (j < case.n[1] - 1)
    np.testing.assert_allclose(lim[0][inner], expect, rtol=0, atol=1e-12)


def test_ux_limiter_edge_neighbour_and_wall_gradient():
    # x-momentum means m + g x with slope g.  In cell (0, j) the neighbour
    # differences are g dx (x+), 0 (y+-) and, at the domain edge, the cell's
    # own means with rho v_x = 0: -(m + g x0).  Ratios: x+ face 1 -> 1;
    # y faces dU = 0 -> 2; x- face (dU = -g dx/2) 2 (m + g x0) / (g dx)
    # = 1/2 when m + g x0 = g dx / 4 -> (2 * 1/2) / (3/2) = 2/3.
    case, st = _ux()
    c = _uniform(case, st)
    i, j = _ij(case)
    dx = (case.box_r[0] - case.box_l[0]) / case.n[0]
    xc = (i + 0.5) * dx
    g = 0.2
    m = g * dx / 4 - g * (0.5 * dx)
    c["rhov"][0, 0] = m + g * xc
    c["coef"][0, 1, 0] = c["rhov"][0, 0]
    c["coef"][0, 1, 1] = g
    vx, vy = 0.3, -0.2
    c["v"][0] = vx
    c["v"][1] = vy
    out = _limiter(case, st)
    lim = out.c["lim"][0]
    cell = 0 * case.n[1] + 2
    assert abs(lim[1, cell] - 2.0 / 3.0) < 1e-12
    # G_v x column against a zero wall velocity: (v - 0)/(dx/2) at i = 0,
    # (0 - v)/(dx/2) at i = N-1, central differences (0) inside
    G = out.c["G_v"].reshape(3, 3, -1)
    first, last = i == 0, i == case.n[0] - 1
    np.testing.assert_allclose(G[0, 0, first], vx / (0.5 * dx), rtol=1e-15)
    np.testing.assert_allclose(G[1, 0, first], vy / (0.5 * dx), rtol=1e-15)
    np.testing.assert_allclose(G[0, 0, last], -vx / (0.5 * dx), rtol=1e-15)
    np.testing.assert_allclose(G[1, 0, last], -vy / (0.5 * dx), rtol=1e-15)
    assert np.all(G[:2, 0, ~(first | last)] == 0.0)


def test_ux_p2g_keeps_absent_phase_state():
    # a layer of water only (no phase-2 particles reach the bottom rows):
    # there UX keeps phase 2's density and momentum as they were
    case, st = _ux(nx=16)
    c = st.c
    c["rho"][1] = 7.0
    c["rhov"][1, 0] = 0.25
    c["rhov"][1, 1] = -0.5
    o = OracleSim(case, st)
    for fn in ("used_state_update", "bin", "membership", "p2g"):
        o.call(fn)
    out = st.copy()
    o.save(out)
    absent = out.c["ph_state"][1] == 0
    assert absent.sum() > 10
    assert np.all(out.c["rho"][1][absent] == 7.0)
    assert np.all(out.c["rhov"][1, 0][absent] == 0.25)
    assert np.all(out.c["rhov"][1, 1][absent] == -0.5)
    assert np.all(out.c["coef"][1][:, :, absent] == 0.0)


def _water_state(rho, rv, E):
    """(rho, rho v[3], E, P, c) in the water EOS."""
    v = rv / rho
    P = (E - 0.5 * rho * float(v @ v)) * (G_W - 1) - G_W * PINF_W
    c = math.sqrt(G_W * (PINF_W + P) / rho)
    return np.array([rho, rv[0], rv[1], 0.0, E, P, c])


def test_hll_left_wave_jump_condition():
    # HLL (Toro, Riemann Solvers, 10.21 and 10.13): with the intermediate
    # state U* = (SR Ur - SL Ul - (Fr - Fl)) / (SR - SL) the flux is
    # Fl + SL (U* - Ul) (and Fr + SR (U* - Ur)) when SL < 0 < SR
    rng = np.random.default_rng(11)
    for trial in range(30):
        ax = trial % 2
        n = np.zeros(3)
        n[ax] = 1.0 if trial % 4 < 2 else -1.0
        sts = []
        for _ in range(2):
            rho = rng.uniform(900, 1100)
            v = rng.uniform(-30, 30, 2)
            p = rng.uniform(1e5, 5e7)
            E = (p + G_W * PINF_W) / (G_W - 1) + 0.5 * rho * float(v @ v)
            sts.append(_water_state(rho, rho * v, E))
        L, R = sts
        F = face_flux(2, 1, n, L, R) @ n[:2]

        def phys(s):
            rho, rv, E, P = s[0], s[1:3], s[4], s[5]
            v = rv / rho
            vn = float(v @ n[:2])
            return np.array([rho * vn, rv[0] * vn + P * n[0], rv[1] * vn + P * n[1], (E + P) * vn])

        def cons(s):
            return np.array([s[0], s[1], s[2], s[4]])

        vnl, vnr = (L[1:3] / L[0]) @ n[:2], (R[1:3] / R[0]) @ n[:2]
        SL = min(vnl - L[6], vnr - R[6])
        SR = max(vnl + L[6], vnr + R[6])
        assert SL < 0 < SR
        Us = (SR * cons(R) - SL * cons(L) - (phys(R) - phys(L))) / (SR - SL)
        scale = np.abs(phys(L)).max() + np.abs(phys(R)).max()
        np.testing.assert_allclose(F, phys(L) + SL * (Us - cons(L)), rtol=0, atol=1e-12 * scale)
        np.testing.assert_allclose(F, phys(R) + SR * (Us - cons(R)), rtol=0, atol=1e-12 * scale)


def test_ux_rhs_is_finite_volume_hll_with_ghosts():
    case, st = _ux()
    c = _uniform(case, st)
    nx, ny, nc = case.n[0], case.n[1], case.nc
    rng = np.random.default_rng(5)
    rho = rng.uniform(950, 1050, nc)
    v = rng.uniform(-20, 20, (2, nc))
    p = rng.uniform(1e5, 2e7, nc)
    E = (p + G_W * PINF_W) / (G_W - 1) + 0.5 * rho * (v * v).sum(axis=0)
    c["coef"][0, 0, 0] = rho
    c["coef"][0, 1, 0] = rho * v[0]
    c["coef"][0, 2, 0] = rho * v[1]
    c["coef"][0, 3, 0] = E
    c["rho"][0] = rho
    c["rhov"][0, 0] = rho * v[0]
    c["rhov"][0, 1] = rho * v[1]
    c["E"][0] = E
    o = OracleSim(case, st)
    o.call("used_state_update")
    o.call("rhs")
    out = st.copy()
    o.save(out)
    rate = out.c["dcoef"][0, :, 0]
    dx = (case.box_r[0] - case.box_l[0]) / nx
    dy = (case.box_r[1] - case.box_l[1]) / ny
    vol = dx * dy
    sides = [((0, -1), (0.0, -1.0), dx, BC_WALL), ((1, 0), (1.0, 0.0), dy, BC_OUTLET),
             ((0, 1), (0.0, 1.0), dx, BC_OUTLET), ((-1, 0), (-1.0, 0.0), dy, BC_OUTLET)]
    worst = 0.0
    for i in range(nx):
        for j in range(ny):
            ci = i * ny + j
            me = _water_state(rho[ci], rho[ci] * v[:, ci], E[ci])
            want = np.zeros(4)
            for (di, dj), nrm, S, tag in sides:
                n = np.array([nrm[0], nrm[1], 0.0])
                ni, nj = i + di, j + dj
                if 0 <= ni < nx and 0 <= nj < ny:
                    cj = ni * ny + nj
                    other = _water_state(rho[cj], rho[cj] * v[:, cj], E[cj])
                else:
                    vg = v[:, ci].copy()
                    if tag == BC_WALL:
                        vg[1] = -vg[1]
                    other = me.copy()
                    other[1:3] = rho[ci] * vg
                F = face_flux(2, 1, n, me, other) @ n[:2]
                want -= F * S
            want /= vol
            scale = np.abs(want).max() + 1e-30
            worst = max(worst, np.abs(rate[:, ci] - want).max() / scale)
    assert worst < 1e-9, worst


def _create(case):
    from pgdg_p2_amd import lib
    h = C.c_void_p()
    return lib().pgdg_create(C.byref(case), 1000, 10, C.byref(h))


@pytest.mark.parametrize("change", ["mixed", "3phase", "3d", "inlet", "x_wall"])
def test_create_rejects_unsupported_ux_cases(change):
    case, _ = _ux()
    if change == "mixed":
        case.precision = 1
    elif change == "3phase":
        case.nphase = 3
    elif change == "3d":
        case.dim = 3
    elif change == "inlet":
        case.bc[3] = BC_INLET
    else:
        case.bc[1] = BC_WALL
    assert _create(case) == -1


def test_create_accepts_the_ux_case_arguments():
    # arguments pass; without a device the call stops at the device check
    case, _ = _ux()
    case.bc[2] = BC_WALL
    rc = _create(case)
    assert rc in (0, -6)
