"""The reference's output files and Output_point sums, written by the
library's host writers (no GPU).

The expected bytes of the particle file are built here field by field from
the sequence MPM::Output_plt_file writes (S3:mpm_ini.cpp:832-946,
T2:mpm_ini.cpp:662-805): magic, three int32 words, the variable count and
names (one int32 per character, a zero word after each), the zone marker
299.0f, the zone name, the zone words, IMAX/JMAX/KMAX, 357.0f, 299.0f, the
value types, three sharing words, a (-1.0, 1.0) double pair per variable,
then every variable for every particle in turn (double, or int32 for phase
and n_neighbor).  The cell file's expected text uses printf's %g, which is
what C++ streams print for a double at their default precision 6."""
import math
import os
import struct

import numpy as np
import pytest

from pgdg_p2_amd import cases, output


def _state(dim, seed=1, ux=False):
    if ux:
        case, st = cases.underwater_explosion_2d(nx=4)
    elif dim == 2:
        case, st = cases.triple_point_2d(nx=4, ny=2, ppc_axis=1, jitter=0.0)
    else:
        case, st = cases.shock_bubble_3d(nx=2, ny=1, nz=1, ppc_axis=2, jitter=0.25)
    rng = np.random.default_rng(seed)
    n, cap = st.n, st.cap
    for f in ("v", "tv"):
        st.p[f][:, :n] = rng.normal(size=(3, n))
    for f in ("P", "rho", "vol", "E"):
        st.p[f][:n] = rng.uniform(0.5, 2.0, n)
    st.p["phase"][:n] = rng.integers(1, 3, n)
    return case, st


def _name(s):
    return b"".join(struct.pack("<i", ord(ch)) for ch in s) + struct.pack("<i", 0)


def _expected_particle_bytes(case, st, nb, run_time):
    D = case.dim
    names = (["x", "y", "z"] if D == 3 else ["x", "y"]) + ["phase", "vx", "vy", "vol", "rho", "P", "tv", "v",
                                                          "n_neighbor"]
    n = st.n
    b = b"#!TDV112" + struct.pack("<iii", 1, 0, 0) + struct.pack("<i", len(names))
    for s in names:
        b += _name(s)
    b += struct.pack("<f", 299.0) + _name("grid")
    b += struct.pack("<ii", -1, 0) + struct.pack("<d", run_time) + struct.pack("<iiiii", -1, 0, 0, 0, 0)
    b += struct.pack("<iiii", n, 1, 1, 0) + struct.pack("<ff", 357.0, 299.0)
    b += b"".join(struct.pack("<i", 3 if s in ("phase", "n_neighbor") else 2) for s in names)
    b += struct.pack("<iii", 0, 0, -1)
    b += struct.pack("<dd", -1.0, 1.0) * len(names)
    p = st.p
    for s in names:
        for i in range(n):
            if s in ("x", "y", "z"):
                b += struct.pack("<d", p["x"]["xyz".index(s), i])
            elif s == "phase":
                b += struct.pack("<i", int(p["phase"][i]))
            elif s in ("vx", "vy"):
                b += struct.pack("<d", p["v"][0 if s == "vx" else 1, i])
            elif s in ("vol", "rho", "P"):
                b += struct.pack("<d", p[s][i])
            elif s in ("tv", "v"):
                w = p[s][:, i]
                acc = 0.0
                for a in range(D):
                    acc += w[a] * w[a]
                b += struct.pack("<d", math.sqrt(acc))
            else:
                b += struct.pack("<i", int(nb[i]))
    return b


@pytest.mark.parametrize("dim", [2, 3])
def test_particle_file_bytes(tmp_path, dim):
    case, st = _state(dim)
    nb = np.arange(3, 3 + st.n, dtype=np.int32)
    path = os.path.join(tmp_path, "sp.plt")
    output.write_particles_plt(path, case, st, nb, 0.0625)
    got = open(path, "rb").read()
    want = _expected_particle_bytes(case, st, nb, 0.0625)
    assert got == want


def _g(x):
    return "%g" % x


@pytest.mark.parametrize("dim", [2, 3, "ux"])
def test_cell_file_text(tmp_path, dim):
    ux = dim == "ux"
    case, st = _state(2 if ux else dim, ux=ux)
    rng = np.random.default_rng(7)
    c = st.c
    for f in ("v", "rho", "alph", "lim", "coef", "rhov"):
        c[f][...] = rng.uniform(0.1, 1.0, c[f].shape)
    c["P"][...] = rng.uniform(0.5, 2.0, c["P"].shape)
    path = os.path.join(tmp_path, "ele.plt")
    output.write_cells_plt(path, case, st, 0.25)
    lines = open(path).read().split("\n")
    # UX adds the cells' ca column, never assigned (UX:mpm_ini.cpp:771-789)
    assert lines[0] == ('VARIABLES = "x","y","z", "v1","v2", "v","rho0","rho1",        '
                        '"E0","E1","Lim0","Lim1","st", "st1", ' + ('"ca", "P"' if ux else '"P"'))
    assert lines[1] == 'ZONE T="GRID 1",SOLUTIONTIME=0.25'
    D = case.dim
    n0, n1 = case.n[0], case.n[1]
    n2 = case.n[2] if D == 3 else 1
    ps = [(case.box_r[a] - case.box_l[a]) / case.n[a] for a in range(D)] + [0.0] * (3 - D)
    k = 2
    for i in range(n0):
        for j in range(n1):
            for q in range(n2):
                ci = (i * n1 + j) * n2 + q
                x = [ps[0] * (i + 0.5), ps[1] * (j + 0.5), ps[2] * (q + 0.5) if D == 3 else 0.0]
                v = c["v"].reshape(3, -1)[:, ci]
                vn = math.sqrt(sum(v[a] * v[a] for a in range(D)))
                rho = c["rho"].reshape(case.nphase, -1)[:, ci]
                vals = [x[0], x[1], x[2], v[0], v[1], vn, rho[0], rho[1]]
                lim = c["lim"].reshape(case.nphase, D + 2, -1)
                coef = c["coef"].reshape(case.nphase, D + 2, -1, c["coef"].shape[-1])
                if D == 3:
                    vals += [coef[0, 1, 1, ci], coef[0, 1, 2, ci], lim[0, 0, ci], lim[1, 0, ci]]
                elif ux:  # phase 2's y-momentum slopes (UX:mpm_ini.cpp:784)
                    vals += [coef[1, 2, 1, ci], coef[1, 2, 2, ci], lim[0, 3, ci], lim[1, 3, ci]]
                else:
                    rv = c["rhov"].reshape(case.nphase, 3, -1)
                    sp = [math.sqrt(sum((rv[f, a, ci] / rho[f]) ** 2 for a in range(D))) for f in range(2)]
                    vals += [sp[0], sp[1], lim[0, 3, ci], lim[1, 3, ci]]
                al = c["alph"].reshape(case.nphase, -1)[:, ci]
                vals += [al[0], al[1]] + ([0.0] if ux else []) + [c["P"].reshape(-1)[ci]]
                assert lines[k] == " ".join(_g(t) for t in vals) + " ", (k, lines[k])
                k += 1
    assert lines[k] == "" and k == len(lines) - 1


@pytest.mark.parametrize("dim", [2, 3])
def test_point_sums_and_lines(tmp_path, dim):
    case, st = _state(dim)
    p, c = st.p, st.c
    n = st.n
    # move a few phase-2 particles near the axis for the S3 front positions
    p["x"][1, :n:3] = 0.1 * case.p_sz
    p["x"][2, :n:3] = 0.1 * case.p_sz
    rng = np.random.default_rng(3)
    for f in ("rho", "alph", "E"):
        c[f][...] = rng.uniform(0.1, 1.0, c[f].shape)
    nc = c["P"].size
    used = (rng.uniform(size=nc) < 0.7).astype(np.int32)
    d = output.point_diagnostics(case, st, used, 0.125)
    # the reference's loops, in its order
    dw, jet, ups, m2 = 0.0, case.box_r[0], case.box_r[0], 0.0
    pm1 = pm2 = pe = 0.0
    for i in range(n):
        x, y, z = p["x"][:, i]
        if p["phase"][i] == 2:
            if y < 2.5 * case.p_sz and z < 2.5 * case.p_sz:
                dw = x if x > dw else dw
                jet = x if x < jet else jet
            elif y > 4 * case.p_sz or z > 4 * case.p_sz:
                ups = x if x < ups else ups
            m2 += p["rho"][i] * p["vol"][i]
        if p["phase"][i] == 1:
            pm1 += p["vol"][i] * p["rho"][i]
        elif p["phase"][i] == 2:
            pm2 += p["vol"][i] * p["rho"][i]
        pe += p["vol"][i] * p["E"][i]
    cvol = 1.0
    for a in range(case.dim):
        cvol *= (case.box_r[a] - case.box_l[a]) / case.n[a]
    rho = c["rho"].reshape(case.nphase, -1)
    al = c["alph"].reshape(case.nphase, -1)
    E = c["E"].reshape(case.nphase, -1)
    gm1 = gm2 = ge = 0.0
    for ci in range(nc):
        if used[ci]:
            gm1 += cvol * rho[0, ci] * al[0, ci]
            gm2 += cvol * rho[1, ci] * al[1, ci]
            for f in range(case.nphase):
                ge += cvol * E[f, ci] * al[f, ci]
    assert (d.dw, d.jet, d.ups, d.mass2, d.n_fluid) == (dw, jet, ups, m2, n)
    assert (d.p_mass[0], d.p_mass[1], d.p_energy) == (pm1, pm2, pe)
    assert (d.g_mass[0], d.g_mass[1], d.g_energy) == (gm1, gm2, ge)
    assert d.p2g == n / used.sum()
    output.point_append(str(tmp_path), case, d, True)
    output.point_append(str(tmp_path), case, d, False)
    if dim == 3:
        a = open(os.path.join(tmp_path, "pointa.txt")).read().split("\n")
        b = open(os.path.join(tmp_path, "mass.txt")).read().split("\n")
        assert a[0] == a[1] == f"{_g(0.125)} {_g(dw)} {_g(jet)} {_g(ups)}"
        assert b[0] == b[1] == f"{_g(0.125)} {_g(m2)} {n}"
    else:
        a = open(os.path.join(tmp_path, "point_p.txt")).read().split("\n")
        b = open(os.path.join(tmp_path, "point_g.txt")).read().split("\n")
        assert a[0] == a[1] == f"{_g(0.125)} {_g(pm1)} {_g(pm2)} {_g(pe)} {_g(n / used.sum())} "
        assert b[0] == b[1] == f"{_g(0.125)} {_g(gm1)} {_g(gm2)} {_g(ge)} "


def test_run_case_intervals_on_the_oracle(tmp_path):
    # the outer loop on the oracle: files 0..3, one Output_point line per
    # interval, each written at the interval's end time
    from oracle.pyoracle_step import OracleSim
    from pgdg_p2_amd.run import run_case
    case, st = cases.triple_point_2d(nx=28, ny=12, ppc_axis=3, jitter=0.25)
    case.output_dt = 0.05
    ref = OracleSim(case, st)
    t = run_case(ref, st.copy(), 3, str(tmp_path))
    assert t == (0.05 + 0.05) + 0.05
    names = sorted(os.listdir(tmp_path))
    assert names == sorted([f"simulation_sp.{n}.0.plt" for n in range(4)] +
                           [f"ele_file.{n}.0.plt" for n in range(4)] + ["point_p.txt", "point_g.txt"])
    times = [float(l.split()[0]) for l in open(os.path.join(tmp_path, "point_p.txt")).read().split("\n") if l]
    assert times == [float(_g(0.05)), float(_g(0.1)), float(_g(0.15))]
    for n, want in ((0, 0.0), (3, None)):
        head = open(os.path.join(tmp_path, f"ele_file.{n}.0.plt")).read().split("\n")[1]
        if want is not None:
            assert head == f'ZONE T="GRID 1",SOLUTIONTIME={_g(want)}'
        else:
            assert head.startswith('ZONE T="GRID 1",SOLUTIONTIME=0.15')


def test_writers_reject_bad_arguments(tmp_path):
    import ctypes as C
    from pgdg_p2_amd import lib
    case, st = _state(3)
    nb = np.zeros(st.n, dtype=np.int32)
    missing = os.path.join(tmp_path, "no_such_dir", "x.plt")
    with pytest.raises(Exception):
        output.write_particles_plt(missing, case, st, nb, 0.0)
    with pytest.raises(Exception):
        output.write_cells_plt(missing, case, st, 0.0)
    P = st.particles()
    P.n = P.cap + 1  # more particles than the arrays hold
    rc = lib().pgdg_plt_write_particles(os.fsencode(os.path.join(tmp_path, "a.plt")), C.byref(case), C.byref(P),
                                        nb.ctypes.data_as(C.POINTER(C.c_int32)), 0.0)
    assert rc == -1
    d = output.PointDiag()
    used = np.ones(st.c["P"].size, dtype=np.int32)
    rc = lib().pgdg_point_diagnostics(C.byref(case), C.byref(P), C.byref(st.cells()),
                                      used.ctypes.data_as(C.POINTER(C.c_int32)), 0.0, C.byref(d))
    assert rc == -1
    with pytest.raises(Exception):
        output.point_append(os.path.join(tmp_path, "no_such_dir"), case, d, True)


def test_run_case_needs_an_output_interval(tmp_path):
    from oracle.pyoracle_step import OracleSim
    from pgdg_p2_amd.run import run_case
    case, st = cases.triple_point_2d(nx=8, ny=4, ppc_axis=1, jitter=0.0)
    ref = OracleSim(case, st)  # output_dt left at its "no output" default
    with pytest.raises(ValueError):
        run_case(ref, st.copy(), 1, str(tmp_path))
