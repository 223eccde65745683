"""The mixed-precision step (FP32 particle-shifting pair terms in cell-local
coordinates with FP64 sums; FP32 DG quadrature with the MGFM Newton solve,
the mass division and the state in FP64) against the FP64 step on the GPU
(which is bit-identical to the oracle, see test_gpu_step.py).

Stated tolerance (3-D shock-bubble, the workload of the mixed form): cell
mixture density within 1e-6 relative in L1 and 1e-3 relative in Linf at
every step; 2-D triple point (coarse grid, T2 semantics, strong shocks
meeting a contact): 1e-5 in L1 and 1e-2 in Linf over 12 steps.  Particle
counts (split/merge outcomes) identical at every step; particle positions
within 1e-5 particle spacings on average (1e-4 for the triple point) and
0.05 at most (a rounding-level change can flip a limiter or state-fallback
branch at a shock and move a few particles by more)."""
import numpy as np
import pytest

from pgdg_p2_amd import cases

DX_MAX_TOL = 0.05


def _mix_rho(st):
    return (st.c["alph"] * st.c["rho"]).sum(axis=0)


def _compare(make, steps, label, L1_TOL=1e-6, LINF_TOL=1e-3, DX_MEAN_TOL=1e-5):
    from pgdg_p2_amd.step import GpuSim
    case, st = make()
    casem, _ = make()
    casem.precision = 1
    gp, gm = GpuSim(case, st), GpuSim(casem, st)
    a, b = st.copy(), st.copy()
    worst = 0.0
    try:
        for k in range(steps):
            print(f"{label}: step {k}", flush=True)
            gp.step(1)
            gm.step(1)
            gp.save(a)
            gm.save(b)
            assert a.n == b.n, f"{label} step {k}: particle counts {a.n} vs {b.n}"
            rp, rm = _mix_rho(a), _mix_rho(b)
            l1 = np.abs(rm - rp).sum() / np.abs(rp).sum()
            linf = np.abs(rm - rp).max() / np.abs(rp).max()
            d = np.abs(a.live("x") - b.live("x")).max(axis=0) / case.p_sz
            dmean, dmax = d.mean(), d.max()
            assert l1 <= L1_TOL and linf <= LINF_TOL, (label, k, l1, linf)
            assert dmean <= DX_MEAN_TOL and dmax <= DX_MAX_TOL, (label, k, dmean, dmax)
            worst = max(worst, linf)
    finally:
        gp.close()
        gm.close()
    print(f"{label}: worst density Linf {worst:.3e}, last L1 {l1:.3e}")
    assert worst > 0  # the mixed path did run


@pytest.mark.gpu
def test_mixed_step_shock_bubble_3d(cuda):
    _compare(lambda: cases.shock_bubble_3d(nx=32, ny=16, nz=16), 20, "sb3d-mixed")


@pytest.mark.gpu
def test_mixed_step_triple_point_2d(cuda):
    _compare(lambda: cases.triple_point_2d(nx=56, ny=24, ppc_axis=3), 12, "tp2d-mixed",
             L1_TOL=1e-5, LINF_TOL=1e-2, DX_MEAN_TOL=1e-4)


@pytest.mark.gpu
def test_mixed_step_p2_rk3_shock_bubble_3d(cuda):
    sch = dict(order=2, rk_stages=3, ppc_min=4, ppc_max=16)
    _compare(lambda: cases.shock_bubble_3d(nx=32, ny=16, nz=16, scheme=sch), 12, "sb3d-p2-mixed")


def _staged_vs_walk(case, st, steps, tag):
    # the staged windows and the per-lane walk visit the same candidates in
    # the same order, with the records in the same fixed frames (blocks of
    # cells along each line): every particle and cell field identical
    from pgdg_p2_amd.compare import compare_states
    from pgdg_p2_amd.step import GpuSim
    case.precision = 1
    ga, gb = GpuSim(case, st), GpuSim(case, st)
    gb.pst_staging(False)
    a, b = st.copy(), st.copy()
    try:
        for k in range(steps):
            print(f"{tag}: step {k}", flush=True)
            ga.step(1)
            gb.step(1)
            ga.save(a)
            gb.save(b)
            bad = compare_states(a, b)
            assert not bad, f"step {k}: {bad[:6]}"
    finally:
        ga.close()
        gb.close()


@pytest.mark.gpu
def test_mixed_pst_staged_equals_per_lane_walk(cuda):
    # 64x32x32 P2 cells, 8 particles per cell: a workgroup's targets lie on
    # one or two cell lines (candidates staged in LDS) or, where cells hold
    # fewer particles, on more lines (every lane walks its candidates in
    # global memory); both must give the per-lane sums bit for bit
    sch = dict(order=2, rk_stages=3, ppc_min=4, ppc_max=16)
    case, st = cases.shock_bubble_3d(nx=64, ny=32, nz=32, scheme=sch)
    _staged_vs_walk(case, st, 6, "staged-vs-walk")


@pytest.mark.gpu
def test_mixed_pst_staged_equals_walk_with_splits_and_merges(cuda):
    # splits raise the particle count of some cells (candidate ranges wider
    # than the LDS tables are walked in global memory), merges and outlet
    # removal leave removed entries in the bins
    from forced_events import shock_bubble_forced
    sch = dict(order=2, rk_stages=3, ppc_min=4, ppc_max=16)
    case, st = shock_bubble_forced(nx=32, ny=16, nz=16, scheme=sch, big_frac=0.2)
    _staged_vs_walk(case, st, 3, "staged-vs-walk-forced")


@pytest.mark.gpu
def test_mixed_pst_staged_equals_walk_27_per_cell(cuda):
    # 27 particles per cell: a workgroup's targets span several cells of a
    # line and the candidate ranges are wide
    sch = dict(order=2, rk_stages=3, ppc_min=8, ppc_max=40)
    case, st = cases.shock_bubble_3d(nx=16, ny=8, nz=8, ppc_axis=3, scheme=sch)
    _staged_vs_walk(case, st, 2, "staged-vs-walk-27ppc")


@pytest.mark.gpu
def test_mixed_rhs_matrix_cores_match_lane_per_cell(cuda):
    # the right-hand side as 16x16 matrix products on the matrix cores
    # against the one-lane-per-cell kernel: the same point terms summed in
    # another order; the cell densities stay within the mixed form's stated
    # tolerance of each other (the higher modes at the shock and the bubble
    # interface follow the limiter's switching and are compared through the
    # densities, as between the mixed and FP64 forms); the shock-bubble
    # channel has wall, inlet and outlet faces and two phases
    from pgdg_p2_amd.step import GpuSim
    sch = dict(order=2, rk_stages=3, ppc_min=4, ppc_max=16)
    case, st = cases.shock_bubble_3d(nx=32, ny=16, nz=16, scheme=sch)
    case.precision = 1
    ga, gb = GpuSim(case, st), GpuSim(case, st)
    ga.rhs_mfma(True)
    gb.rhs_mfma(False)
    a, b = st.copy(), st.copy()
    try:
        for k in range(6):
            print(f"rhs-mfma-vs-lanes: step {k}", flush=True)
            ga.step(1)
            gb.step(1)
            ga.save(a)
            gb.save(b)
            assert a.n == b.n, (k, a.n, b.n)
            ra, rb = _mix_rho(a), _mix_rho(b)
            l1 = np.abs(ra - rb).sum() / np.abs(rb).sum()
            linf = np.abs(ra - rb).max() / np.abs(rb).max()
            assert l1 <= 1e-6 and linf <= 1e-3, (k, l1, linf)
        assert l1 > 0  # the two kernels did differ
    finally:
        ga.close()
        gb.close()


@pytest.mark.gpu
def test_mixed_step_p2_rk3_shock_cylinder_3d(cuda):
    # the reference's cylinder case in its own units (157x20x4 cells, 8
    # particles per cell, inlet copies every step)
    sch = dict(order=2, rk_stages=3, ppc_min=4, ppc_max=16)
    _compare(lambda: cases.shock_cylinder_3d(ny=20, ppc_axis=2, jitter=0.25, scheme=sch), 12, "c3d-p2-mixed")


@pytest.mark.gpu
def test_mixed_pst_staged_equals_walk_long_lines(cuda):
    # lines of 160 cells: the records sit in the frames of blocks of 64
    # cells and a window that crosses a block edge switches frames there
    sch = dict(order=2, rk_stages=3, ppc_min=4, ppc_max=16)
    case, st = cases.shock_bubble_3d(nx=16, ny=8, nz=160, scheme=sch)
    _staged_vs_walk(case, st, 3, "staged-vs-walk-long-lines")


@pytest.mark.gpu
@pytest.mark.parametrize("grid", [(64, 32, 32), (16, 8, 160)], ids=["line-frames", "block-frames"])
def test_mixed_step_independent_of_workgroup_sizes(cuda, grid):
    # the autotuner picks the workgroup sizes by timing, so the mixed step
    # must give the same state at every size: 64-lane particle-shifting
    # workgroups (windows staged for about eight target cells) against
    # 256-lane ones (32 cells, other line spans and staging chunks), RHS
    # and limiter at 64 and 256; lines of 32 cells (one record frame per
    # line) and of 160 cells (frames of 64-cell blocks)
    from pgdg_p2_amd.compare import compare_states
    from pgdg_p2_amd.step import GpuSim
    sch = dict(order=2, rk_stages=3, ppc_min=4, ppc_max=16)
    nx, ny, nz = grid
    case, st = cases.shock_bubble_3d(nx=nx, ny=ny, nz=nz, scheme=sch)
    case.precision = 1
    ga, gb = GpuSim(case, st), GpuSim(case, st)
    ga.set_tuning(64, 64, 64)
    gb.set_tuning(256, 256, 256)
    a, b = st.copy(), st.copy()
    try:
        for k in range(4):
            print(f"workgroup sizes {grid}: step {k}", flush=True)
            ga.step(1)
            gb.step(1)
            ga.save(a)
            gb.save(b)
            bad = compare_states(a, b)
            assert not bad, f"step {k}: {bad[:6]}"
        assert ga.tuning() == {"pst": 64, "rhs": 64, "limiter": 64}
    finally:
        ga.close()
        gb.close()
