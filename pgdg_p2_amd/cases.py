"""Case set-up: geometry, boundary tags, EOS tables and initial particle/cell
states (the reference's MPM::Initialize_case, S3:mpm_ini.cpp:33-673 and
T2:mpm_ini.cpp:33-482, generalised to several particles per cell with a
seeded jitter).

Particles are placed on a lattice of spacing p_sz = p_g / ppc_axis at
(i + 1/2) p_sz (S3:mpm_ini.cpp:177-184), optionally jittered by
U(-jitter, jitter) * p_sz per axis (numpy PCG64, fixed seed).  Wall particles
sit on the boundary lattice with the reference's outward normals
(S3:mpm_ini.cpp:555-620, T2:mpm_ini.cpp:385-430).
"""
from __future__ import annotations

import math

import numpy as np

from .state import (BC_INLET, BC_OUTLET, BC_SYMMETRY, BC_WALL, Case, HostState, SEM_S3,
                    SEM_T2, SEM_UX)

CELL_RATIO, CUT_OFF = 3.2, 2.0


def _base_case(dim, nphase, n, box_l, box_r, semantics, bc, p_g, ppc_axis, scheme=None):
    """`scheme` = dict(order=1|2, rk_stages=2|3, cfl=0 (reference 0.3),
    ppc_min/ppc_max=0 (reference volume criteria only))."""
    sch = dict(order=1, rk_stages=2, cfl=0.0, ppc_min=0, ppc_max=0)
    sch.update(scheme or {})
    c = Case()
    c.dim, c.nphase, c.order, c.rk_stages, c.semantics = (dim, nphase, sch["order"],
                                                          sch["rk_stages"], semantics)
    c.cfl, c.ppc_min, c.ppc_max = sch["cfl"], sch["ppc_min"], sch["ppc_max"]
    for a in range(3):
        c.n[a] = n[a] if a < dim else 1
        c.box_l[a] = box_l[a] if a < dim else 0.0
        c.box_r[a] = box_r[a] if a < dim else 1.0
    for k in range(6):
        c.bc[k] = bc[k]
    # reference: p_sz = (px py [pz])^(1/D) / round(p_g / p_sz)
    ps = [(c.box_r[a] - c.box_l[a]) / c.n[a] for a in range(dim)]
    prod = ps[0] * ps[1] if dim == 2 else ps[0] * ps[1] * ps[2]
    c.p_g = p_g
    c.p_sz = (math.sqrt(prod) if dim == 2 else prod ** (1.0 / 3.0)) / round(ppc_axis)
    # P2G cut-off: the reference's 1.5*p_sz at one particle per cell; with
    # several particles per cell 1.5*p_g (equal to it at one per cell) so
    # that every particle of a cell reaches the cell centre.
    c.p2g_cut = 1.5 * (c.p_sz if round(ppc_axis) == 1 else c.p_g)
    c.output_dt = 1e30
    c.max_v0 = 1e-7
    return c


def _lattice(case: Case, jitter: float, seed: int):
    D = case.dim
    counts = [round((case.box_r[a] - case.box_l[a]) / case.p_sz) for a in range(D)]
    axes = [(np.arange(counts[a]) + 0.5) * case.p_sz + case.box_l[a] for a in range(D)]
    grids = np.meshgrid(*axes, indexing="ij")  # i, j, k order (S3:mpm_ini.cpp:177-179)
    pos = np.stack([g.reshape(-1) for g in grids])
    if jitter > 0:
        rng = np.random.Generator(np.random.PCG64(seed))
        pos = pos + rng.uniform(-jitter, jitter, pos.shape) * case.p_sz
    return pos


def _fill_particles(st: HostState, case: Case, pos, phase, rho, p, vel):
    D, U, M = case.dim, case.U, case.M
    n = pos.shape[1]
    st.n = n
    P = st.p
    P["x"][:D, :n] = pos
    g = np.array(case.gamma[: case.nphase])[phase - 1]
    pinf = np.array(case.pinf[: case.nphase])[phase - 1]
    vol = case.p_sz ** D if D == 2 else (case.p_sz * case.p_sz) * case.p_sz
    P["vol"][:n] = vol
    P["mass"][:n] = vol * rho
    P["h"][:n] = CELL_RATIO / CUT_OFF * case.p_sz
    P["phase"][:n] = phase
    P["rho"][:n] = rho
    P["P"][:n] = p
    P["v"][:D, :n] = vel
    P["rhov"][:D, :n] = vel * rho
    sq = (vel * vel).sum(axis=0) if vel.size else 0.0
    P["E"][:n] = (p + g * pinf) / (g - 1) + 0.5 * rho * sq
    P["c"][:n] = np.sqrt(g * (p + pinf) / rho)
    P["tv"][:, :n] = 0.0
    P["coef"][:, :n] = 0.0
    P["mark_in"][:n] = 0


def _init_cells(st: HostState, case: Case):
    F = case.nphase
    c = st.c
    for s in range(F):
        c["rho"][s] = case.ini_rho[s]
        c["lim"][s] = 1.0
        c["c"][s] = math.sqrt(case.gamma[s] * (case.pinf[s] + 0) / case.ini_rho[s])
        c["ph_state"][s] = 1
        c["ph_state_b"][s] = 0
    st.s.max_v = case.max_v0


def _walls_nodes(case: Case, spacing, vol):
    """Boundary nodes of a lattice with `spacing` per axis, in sx, sy[, sz]
    order, with the first matching outward normal of the reference
    (sx==0, sx==N, sy==0, sy==N, sz==0, sz==N)."""
    D = case.dim
    cnt = [round((case.box_r[a] - case.box_l[a]) / spacing[a]) for a in range(D)]
    idx = np.meshgrid(*[np.arange(cnt[a] + 1) for a in range(D)], indexing="ij")
    idx = [g.reshape(-1) for g in idx]
    lo = [idx[a] == 0 for a in range(D)]
    hi = [idx[a] == cnt[a] for a in range(D)]
    on = np.zeros_like(lo[0])
    for a in range(D):
        on |= lo[a] | hi[a]
    sel = np.nonzero(on)[0]
    nw = sel.size
    x = np.zeros((3, nw))
    for a in range(D):
        x[a] = case.box_l[a] + spacing[a] * (idx[a][sel] + 0.0)
    nrm = np.zeros((3, nw))
    done = np.zeros(nw, dtype=bool)
    for a in range(D):
        m = lo[a][sel] & ~done
        nrm[a, m] = -1.0
        done |= m
        m = hi[a][sel] & ~done
        nrm[a, m] = 1.0
        done |= m
    return x, nrm, np.full(nw, vol)


def triple_point_2d(nx=112, ny=48, ppc_axis=3, jitter=0.25, seed=0x5EED0000, nphase=2,
                    cap_factor=3.0, gammas=None, scheme=None):
    """2-D triple point on [0,7]x[0,3] with T2 semantics (all sides walls).

    Two phases (default): phase 1 gamma=1.5 holds the left block x<=1
    (rho=1, p=1) and the bottom-right block (rho=1, p=0.1); phase 2
    gamma=1.4 the top-right block (rho=0.125, p=0.1).  Region tests follow
    T2:mpm_ini.cpp:159,200 (x > 1, y < 1.5).
    Three phases, as T3 sets them up (EOS constants T3:mpm_ini.cpp:79-95,
    regions and states T3:mpm_ini.cpp:157-236): phase 1 the bottom-right
    block x > 1, y < 1.5 (gamma 1.4, Pinf 0, rho 1, p 0.1), phase 2 the
    top-right block (gamma 1.5, Pinf 0, rho 0.125, p 0.1), phase 3 the left
    block (gamma 1.6, Pinf 0.1, rho 1, p 1); ini_rho (1, 0.125, 0.125).
    """
    box_l, box_r = (0.0, 0.0, 0.0), (7.0, 3.0, 1.0)
    p_g = 3.0 / ny
    case = _base_case(2, nphase, (nx, ny, 1), box_l, box_r, SEM_T2, [BC_WALL] * 6, p_g, ppc_axis,
                      scheme)
    if nphase == 2:
        g = gammas or (1.5, 1.4)
        pinf = (0.0, 0.0)
        rho0 = (1.0, 0.125)
    else:
        g = gammas or (1.4, 1.5, 1.6)
        pinf = (0.0, 0.0, 0.1)
        rho0 = (1.0, 0.125, 0.125)
    for s in range(nphase):
        case.gamma[s] = g[s]
        case.pinf[s] = pinf[s]
        case.ini_rho[s] = rho0[s]
    pos = _lattice(case, jitter, seed)
    x, y = pos[0], pos[1]
    right = x > 1
    bottom = right & (y < 1.5)
    top = right & ~bottom
    n = pos.shape[1]
    phase = np.ones(n, dtype=np.int32)
    rho = np.ones(n)
    p = np.ones(n)
    rho[top] = 0.125
    p[right] = 0.1
    if nphase == 2:
        phase[top] = 2
    else:
        phase[bottom] = 1
        phase[top] = 2
        phase[~right] = 3
    vel = np.zeros((2, n))
    # wall particles on the particle lattice, vol = p_sz (T2:mpm_ini.cpp:385-409)
    pg = [round((box_r[a] - box_l[a]) / case.p_sz) for a in range(2)]
    spacing = [(box_r[a] - box_l[a]) / pg[a] for a in range(2)]
    wx, wn, wv = _walls_nodes(case, spacing, case.p_sz)
    st = HostState(case, int(n * cap_factor) + 64, wx.shape[1])
    _fill_particles(st, case, pos, phase, rho, p, vel)
    st.w["x"][:] = wx
    st.w["norm"][:] = wn
    st.w["vol"][:] = wv
    _init_cells(st, case)
    return case, st


def underwater_explosion_2d(nx=400, ppc_axis=1, jitter=0.0, seed=0x5EED0006, cap_factor=3.0,
                            scheme=None):
    """The reference's 2-D underwater explosion (UX:mpm_ini.cpp:33-300) in
    its own units, with UX semantics: domain [0,4]x[0,3] with
    p_g = 4 / nx (the reference: nx = 400, giving 400 x 300 cells and one
    particle per cell); water (phase 1: gamma 4.4, Pinf 6e8, rho 1000,
    p 1.01325e5) below y = 1.5, air (phase 2: gamma 1.4, Pinf 0,
    rho 1.225, p 1.01325e5) above, and a gas bubble of phase 2 (rho 1250,
    p 1e9; its mass, as the reference sets it, p_sz^2 * 1.225) of radius
    0.12 around (2, 1.2) (UX:mpm_ini.cpp:152-236).  Side tags: the bottom
    reflects the y velocity (the reference's tag 2, PGDG_BC_WALL), the other
    three sides are transmissive (tag -1, PGDG_BC_OUTLET)
    (UX:mpm_ini.cpp:300-380).  Wall particles on the particle lattice's
    boundary nodes with vol = p_sz (UX:mpm_ini.cpp:386-443); cells start
    with ini_rho and the sound speeds sqrt(gamma (Pinf + p) / ini_rho) at
    p = 1.01325e5 (water) and 1e9 (gas) (UX:mpm_ini.cpp:84-90, 262-282).
    nx must be a multiple of 4."""
    if nx % 4:
        raise ValueError("nx must be a multiple of 4 (3 nx / 4 rows)")
    box_l, box_r = (0.0, 0.0, 0.0), (4.0, 3.0, 1.0)
    p_g = 4.0 / nx
    ny = int(3.0 / p_g)
    bc = [BC_WALL, BC_OUTLET, BC_OUTLET, BC_OUTLET, BC_OUTLET, BC_OUTLET]
    case = _base_case(2, 2, (nx, ny, 1), box_l, box_r, SEM_UX, bc, p_g, ppc_axis, scheme)
    case.gamma[0], case.pinf[0], case.ini_rho[0] = 4.4, 6e8, 1000.0
    case.gamma[1], case.pinf[1], case.ini_rho[1] = 1.4, 0.0, 1.225
    pos = _lattice(case, jitter, seed)
    n = pos.shape[1]
    x, y = pos[0], pos[1]
    water = y < 1.5
    bubble = water & ((x - 2) ** 2 + (y - 1.2) ** 2 < 0.12 * 0.12)
    phase = np.where(water & ~bubble, 1, 2).astype(np.int32)
    rho = np.where(water, 1000.0, 1.225)
    p = np.full(n, 1.01325e5)
    rho[bubble], p[bubble] = 1250.0, 1e9
    vel = np.zeros((2, n))
    pg = [round((box_r[a] - box_l[a]) / case.p_sz) for a in range(2)]
    spacing = [(box_r[a] - box_l[a]) / pg[a] for a in range(2)]
    wx, wn, wv = _walls_nodes(case, spacing, case.p_sz)
    st = HostState(case, int(n * cap_factor) + 64, wx.shape[1])
    _fill_particles(st, case, pos, phase, rho, p, vel)
    vol = st.p["vol"][:n]
    st.p["mass"][:n][bubble] = vol[bubble] * case.ini_rho[1]
    st.w["x"][:] = wx
    st.w["norm"][:] = wn
    st.w["vol"][:] = wv
    _init_cells(st, case)
    for s_, p0 in ((0, 1.01325e5), (1, 1e9)):
        st.c["c"][s_] = math.sqrt(case.gamma[s_] * (case.pinf[s_] + p0) / case.ini_rho[s_])
    return case, st


def rankine_hugoniot(mach, gamma, rho1, p1):
    """Post-shock (rho2, u2, p2) behind a shock of Mach `mach` entering gas at rest."""
    g = gamma
    p2 = p1 * (1 + 2 * g / (g + 1) * (mach * mach - 1))
    rho2 = rho1 * ((g + 1) * mach * mach) / ((g - 1) * mach * mach + 2)
    c1 = math.sqrt(g * p1 / rho1)
    us = mach * c1
    u2 = us * (1 - rho1 / rho2)
    return rho2, u2, p2


def shock_bubble_3d(nx=256, ny=128, nz=128, ppc_axis=2, jitter=0.25, seed=0x5EED0002,
                    x_shock=0.2, cap_factor=2.5, length=2.0, scheme=None):
    """3-D shock-bubble channel [0,L]x[0,1]x[0,1] with S3 semantics.

    Air (gamma 1.4, rho 1, p 1) holds a helium sphere (gamma 1.648,
    rho 0.138, p 1) of radius 0.25 centred at (0.5, 0.5, 0.5); a Mach 1.22
    shock starts at x = x_shock and the inlet (x = 0) feeds the post-shock
    state (Rankine-Hugoniot).  Side tags follow S3:mpm_ini.cpp:447-549:
    yL/zL symmetry, yR/zR wall, xL inlet, xR outlet.
    """
    box_l, box_r = (0.0, 0.0, 0.0), (length, 1.0, 1.0)
    p_g = 1.0 / ny
    bc = [BC_SYMMETRY, BC_OUTLET, BC_WALL, BC_INLET, BC_SYMMETRY, BC_WALL]
    case = _base_case(3, 2, (nx, ny, nz), box_l, box_r, SEM_S3, bc, p_g, ppc_axis, scheme)
    case.gamma[0], case.pinf[0], case.ini_rho[0] = 1.4, 0.0, 1.0
    case.gamma[1], case.pinf[1], case.ini_rho[1] = 1.648, 0.0, 0.138
    rho2, u2, p2 = rankine_hugoniot(1.22, 1.4, 1.0, 1.0)
    case.inlet_rho, case.inlet_p = rho2, p2
    case.inlet_u[0], case.inlet_u[1], case.inlet_u[2] = u2, 0.0, 0.0
    pos = _lattice(case, jitter, seed)
    n = pos.shape[1]
    phase = np.ones(n, dtype=np.int32)
    rho = np.ones(n)
    p = np.ones(n)
    vel = np.zeros((3, n))
    post = pos[0] < x_shock
    rho[post], p[post], vel[0, post] = rho2, p2, u2
    d2 = (pos[0] - 0.5) ** 2 + (pos[1] - 0.5) ** 2 + (pos[2] - 0.5) ** 2
    he = (~post) & (d2 <= 0.25 ** 2)
    phase[he], rho[he] = 2, 0.138
    # walls on the grid nodes, vol = p_sz^(D-1) (S3:mpm_ini.cpp:555-620)
    spacing = [(box_r[a] - box_l[a]) / case.n[a] for a in range(3)]
    wx, wn, wv = _walls_nodes(case, spacing, case.p_sz * case.p_sz)
    st = HostState(case, int(n * cap_factor) + 64, wx.shape[1])
    _fill_particles(st, case, pos, phase, rho, p, vel)
    st.w["x"][:] = wx
    st.w["norm"][:] = wn
    st.w["vol"][:] = wv
    _init_cells(st, case)
    return case, st


def shock_cylinder_3d(ny=30, ppc_axis=1, jitter=0.0, seed=0x5EED0005, cap_factor=2.5, scheme=None):
    """The reference's 3-D cylinder case (C3:mpm_ini.cpp:33-300) in its own
    units: channel 0.35 x 0.0445 x 0.01 with p_g = 0.0445 / ny (the
    reference: ny = 30, giving 235 x 30 x 6 cells); post-shock air
    (rho 1.7756, u 110.62, p 156980) for x < 0.1 and at the inlet, air
    (rho 1.29, p 1e5) elsewhere, and a helium cylinder (gamma 1.648,
    rho 0.2347, p 1e5) of radius 0.025 along z around (0.15, 0) on the
    yL symmetry plane.  Side tags: yL, zL and zR symmetry, yR wall, xL
    inlet, xR outlet (C3:mpm_ini.cpp:390-545; the cylinder's zR face is
    the symmetry tag where the spheric case has a wall)."""
    dom = (0.350, 0.0445, 0.01)
    p_g = dom[1] / ny
    n = tuple(max(int(dom[a] / p_g), 1) for a in range(3))
    bc = [BC_SYMMETRY, BC_OUTLET, BC_WALL, BC_INLET, BC_SYMMETRY, BC_SYMMETRY]
    case = _base_case(3, 2, n, (0.0, 0.0, 0.0), dom, SEM_S3, bc, p_g, ppc_axis, scheme)
    case.gamma[0], case.pinf[0], case.ini_rho[0] = 1.4, 0.0, 1.29
    case.gamma[1], case.pinf[1], case.ini_rho[1] = 1.648, 0.0, 0.2347
    case.inlet_rho, case.inlet_p = 1.7756, 156980.0
    case.inlet_u[0], case.inlet_u[1], case.inlet_u[2] = 110.62, 0.0, 0.0
    pos = _lattice(case, jitter, seed)
    npart = pos.shape[1]
    phase = np.ones(npart, dtype=np.int32)
    rho = np.full(npart, 1.29)
    p = np.full(npart, 100000.0)
    vel = np.zeros((3, npart))
    post = pos[0] < 0.1
    rho[post], p[post], vel[0, post] = 1.7756, 156980.0, 110.62
    he = (~post) & ~((pos[0] - 0.15) ** 2 + pos[1] ** 2 > 0.025 ** 2)
    phase[he], rho[he] = 2, 0.2347
    spacing = [(case.box_r[a] - case.box_l[a]) / case.n[a] for a in range(3)]
    wx, wn, wv = _walls_nodes(case, spacing, case.p_sz * case.p_sz)
    st = HostState(case, int(npart * cap_factor) + 64, wx.shape[1])
    _fill_particles(st, case, pos, phase, rho, p, vel)
    st.w["x"][:] = wx
    st.w["norm"][:] = wn
    st.w["vol"][:] = wv
    _init_cells(st, case)
    return case, st
