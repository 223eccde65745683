"""GPU parity of the whole PG-DG step against the oracle.

Both sides start from the same seeded initial state and run the reference
step (binning, split/merge, P2G, limiter, DG RHS, RK update, G2P/PST, dt);
after every step the particle count and order, every particle field and every
cell field must be identical (FP64, no contraction; -0.0 == 0.0)."""
import numpy as np
import pytest

from oracle.pyoracle_step import OracleSim
from pgdg_p2_amd import cases
from pgdg_p2_amd.compare import compare_states


def run_pair(case, st, steps, label, graphs=True):
    from pgdg_p2_amd.step import GpuSim
    ref = OracleSim(case, st)
    gpu = GpuSim(case, st)
    gpu.graphs(graphs)
    a, b = st.copy(), st.copy()
    events = 0
    for k in range(steps):
        print(f"{label}: step {k}", flush=True)
        ref.step(1)
        gpu.step(1)
        ref.save(a)
        gpu.save(b)
        bad = compare_states(a, b)
        assert not bad, f"{label} step {k}: {bad[:6]}"
        if k == 0:
            tuned = gpu.tuning()  # the autotuner ran during the first step
            assert tuned is not None and all(v in (64, 128, 256) for v in tuned.values()), tuned
        events += a.s.n_split + a.s.n_merge + a.s.n_inlet + a.s.n_outlet
    gpu.close()
    return a, events


@pytest.mark.gpu
def test_step_parity_triple_point_2d(cuda):
    case, st = cases.triple_point_2d(nx=28, ny=12, ppc_axis=3, jitter=0.25)
    run_pair(case, st, 12, "tp2d-9ppc")


@pytest.mark.gpu
def test_step_parity_triple_point_lattice_1ppc(cuda):
    # the reference's own setting: one particle per cell on an exact lattice;
    # merges start after ~20 steps at this resolution
    case, st = cases.triple_point_2d(nx=112, ny=48, ppc_axis=1, jitter=0.0)
    a, events = run_pair(case, st, 40, "tp2d-1ppc")
    assert events > 0


@pytest.mark.gpu
def test_step_parity_shock_bubble_3d(cuda):
    case, st = cases.shock_bubble_3d(nx=16, ny=8, nz=8, ppc_axis=2, jitter=0.25)
    run_pair(case, st, 6, "sb3d")


@pytest.mark.gpu
def test_step_parity_triple_point_three_phase(cuda):
    # FldNum = 3 (the reference's Triple_point_3phase: the 2-phase routines
    # with a third material, T3:mpm_ini.cpp)
    case, st = cases.triple_point_2d(nx=28, ny=12, ppc_axis=3, jitter=0.25, nphase=3)
    run_pair(case, st, 12, "tp2d-3phase")


@pytest.mark.gpu
def test_step_parity_shock_bubble_3d_without_graphs(cuda):
    # the RK stages launched one kernel at a time (the other tests replay
    # the captured stage graphs, with split/merge work between replays)
    case, st = cases.shock_bubble_3d(nx=16, ny=8, nz=8, ppc_axis=2, jitter=0.25)
    run_pair(case, st, 6, "sb3d-eager", graphs=False)


# P2 basis (3-point rules) and SSP-RK3: no reference routine; the oracle's
# restatement of the same definitions is the checker (bit-exact).
P2_RK3 = dict(order=2, rk_stages=3, cfl=0.2)


@pytest.mark.gpu
def test_step_parity_p2_rk3_triple_point_2d(cuda):
    case, st = cases.triple_point_2d(nx=28, ny=12, ppc_axis=3, jitter=0.25, scheme=P2_RK3)
    run_pair(case, st, 10, "tp2d-p2rk3")


@pytest.mark.gpu
def test_step_parity_p2_rk3_shock_bubble_3d(cuda):
    case, st = cases.shock_bubble_3d(nx=16, ny=8, nz=8, ppc_axis=2, jitter=0.25,
                                     scheme=dict(order=2, rk_stages=3))
    run_pair(case, st, 4, "sb3d-p2rk3")


@pytest.mark.gpu
def test_step_parity_p1_rk3_shock_bubble_3d_without_graphs(cuda):
    case, st = cases.shock_bubble_3d(nx=16, ny=8, nz=8, ppc_axis=2, jitter=0.25,
                                     scheme=dict(order=1, rk_stages=3))
    run_pair(case, st, 4, "sb3d-p1rk3-eager", graphs=False)


@pytest.mark.gpu
def test_step_parity_per_cell_bounds_lattice_1ppc(cuda):
    # per-cell split/merge bounds (no reference routine): merges only in cells
    # above ppc_min, no split past ppc_max; counts and order bit-exact
    case, st = cases.triple_point_2d(nx=112, ny=48, ppc_axis=1, jitter=0.0,
                                     scheme=dict(ppc_min=1, ppc_max=4))
    a, events = run_pair(case, st, 40, "tp2d-1ppc-bounds")
    assert events > 0


@pytest.mark.gpu
def test_step_parity_p2_rk3_triple_point_112x48_100_steps(cuda):
    # the plain-method workload at its full size: 112x48 P2 cells, 3x3
    # jittered particles per cell, SSP-RK3, CFL 0.2, 100 steps
    case, st = cases.triple_point_2d(nx=112, ny=48, ppc_axis=3, jitter=0.25,
                                     scheme=dict(order=2, rk_stages=3, cfl=0.2))
    run_pair(case, st, 100, "tp2d-112x48-p2rk3")


@pytest.mark.gpu
def test_step_parity_shock_cylinder_3d(cuda):
    # the reference's cylinder case (symmetry on both z sides) in its own
    # units, coarsened to 78x10x2 cells at one particle per cell; inlet
    # copies enter every step
    case, st = cases.shock_cylinder_3d(ny=10, ppc_axis=1)
    a, events = run_pair(case, st, 30, "c3d")
    assert events > 0


@pytest.mark.gpu
def test_step_parity_p2_rk3_bounds_shock_cylinder_3d(cuda):
    case, st = cases.shock_cylinder_3d(ny=10, ppc_axis=2, jitter=0.25,
                                       scheme=dict(order=2, rk_stages=3, ppc_min=4, ppc_max=16))
    a, events = run_pair(case, st, 10, "c3d-p2rk3")
    assert events > 0


def _restart_pair(case, st, before, after, label, tmp_path):
    # one context runs before+after steps; a second context restores the
    # first one's checkpoint after `before` steps and runs `after` more
    from pgdg_p2_amd.step import GpuSim
    path = str(tmp_path / f"{label}.ckpt")
    gpu = GpuSim(case, st)
    print(f"{label}: {before} steps, checkpoint", flush=True)
    gpu.step(before)
    gpu.checkpoint(path)
    gpu.step(after)
    a = gpu.save(st.copy())
    gpu.close()
    print(f"{label}: restore, {after} steps", flush=True)
    gpu2 = GpuSim(case, st)
    gpu2.restore(path)
    gpu2.step(after)
    b = gpu2.save(st.copy())
    gpu2.close()
    bad = compare_states(a, b)
    assert not bad, f"{label}: {bad[:6]}"
    assert a.s.iterate_num == b.s.iterate_num == before + after
    return a


@pytest.mark.gpu
def test_restart_continues_bit_for_bit_triple_point_lattice(cuda, tmp_path):
    # merges start after ~20 steps: the restart falls inside the merge phase
    case, st = cases.triple_point_2d(nx=112, ny=48, ppc_axis=1, jitter=0.0)
    _restart_pair(case, st, 25, 15, "restart-tp2d", tmp_path)


@pytest.mark.gpu
def test_restart_continues_bit_for_bit_mixed_p2_shock_cylinder(cuda, tmp_path):
    case, st = cases.shock_cylinder_3d(ny=10, ppc_axis=2, jitter=0.25,
                                       scheme=dict(order=2, rk_stages=3, ppc_min=4, ppc_max=16))
    case.precision = 1
    _restart_pair(case, st, 6, 6, "restart-c3d-mixed", tmp_path)


@pytest.mark.gpu
def test_restart_rejects_checkpoint_of_another_case(cuda, tmp_path):
    from pgdg_p2_amd.step import GpuSim
    case, st = cases.triple_point_2d(nx=28, ny=12, ppc_axis=3, jitter=0.25)
    gpu = GpuSim(case, st)
    gpu.step(1)
    path = str(tmp_path / "tp.ckpt")
    gpu.checkpoint(path)
    gpu.close()
    case2, st2 = cases.triple_point_2d(nx=32, ny=12, ppc_axis=3, jitter=0.25)
    gpu2 = GpuSim(case2, st2)
    with pytest.raises(Exception):
        gpu2.restore(path)
    with pytest.raises(Exception):
        gpu2.restore(str(tmp_path / "missing.ckpt"))
    gpu2.close()


@pytest.mark.gpu
def test_restart_rejects_corrupted_checkpoint(cuda, tmp_path):
    # a phase outside 1..F or a truncated file is refused before upload
    import struct
    from pgdg_p2_amd.step import GpuSim
    case, st = cases.triple_point_2d(nx=28, ny=12, ppc_axis=3, jitter=0.25)
    gpu = GpuSim(case, st)
    gpu.step(2)
    path = tmp_path / "tp.ckpt"
    gpu.checkpoint(str(path))
    raw = bytearray(path.read_bytes())
    # header: magic, 12 int32 (dim ... M, pad), then np, nwall, nc, pool_used
    np_ = struct.unpack_from("<q", raw, 8 + 48)[0]
    phase_at = 8 + 48 + 32 + 8 * (19 + case.U * case.M) * np_
    assert struct.unpack_from("<i", raw, phase_at)[0] in (1, 2)
    bad = bytearray(raw)
    struct.pack_into("<i", bad, phase_at, 7)
    (tmp_path / "bad.ckpt").write_bytes(bytes(bad))
    (tmp_path / "short.ckpt").write_bytes(bytes(raw[:-4]))
    for f in ("bad.ckpt", "short.ckpt"):
        with pytest.raises(Exception):
            gpu.restore(str(tmp_path / f))
    gpu.restore(str(path))  # the intact file still loads
    gpu.step(1)
    gpu.close()


def _forced_pair(scheme, steps, label, precision=0):
    # 3-D state with particles above the split volume (three in one cell),
    # close pairs below the merge volume and particles past the outlet: the
    # first steps split, merge and remove particles (S3:mpm.cpp:1036-1417,
    # 1536-1706); every field bit-exact after every step
    from forced_events import shock_bubble_forced
    from pgdg_p2_amd.step import GpuSim
    case, st = shock_bubble_forced(scheme=scheme)
    ref = OracleSim(case, st)
    gpu = GpuSim(case, st)
    a, b = st.copy(), st.copy()
    tot = dict(split=0, merge=0, outlet=0)
    for k in range(steps):
        print(f"{label}: step {k}", flush=True)
        n0 = a.n
        ref.step(1)
        gpu.step(1)
        ref.save(a)
        gpu.save(b)
        bad = compare_states(a, b)
        assert not bad, f"{label} step {k}: {bad[:6]}"
        s = a.s
        assert a.n == n0 + 7 * s.n_split + s.n_inlet - s.n_outlet - s.n_merge
        tot["split"] += s.n_split
        tot["merge"] += s.n_merge
        tot["outlet"] += s.n_outlet
    gpu.close()
    assert tot["split"] > 0 and tot["merge"] > 0 and tot["outlet"] > 0, tot
    return tot


@pytest.mark.gpu
def test_step_parity_forced_split_merge_outlet_3d(cuda):
    _forced_pair(dict(order=1, rk_stages=2), 4, "forced-p1")


@pytest.mark.gpu
def test_step_parity_forced_split_merge_outlet_3d_p2_rk3_bounds(cuda):
    _forced_pair(dict(order=2, rk_stages=3, ppc_min=4, ppc_max=16), 4, "forced-p2-bounds")


def _cell_counts(case, st):
    from test_oracle_pins import _cells_of
    c = np.zeros(case.nc, dtype=np.int64)
    np.add.at(c, _cells_of(case, st.p["x"][:, :st.n]), 1)
    return c


@pytest.mark.gpu
@pytest.mark.parametrize("lo,hi", [(4, 16), (8, 12)])
def test_per_cell_bounds_hold_after_every_pass_gpu(cuda, lo, hi):
    # stage by stage (bin, split, merge x4) against the oracle; after the
    # split no cell is above max(ppc_max, its count before), after each merge
    # pass none below min(ppc_min, its count before)
    from forced_events import shock_bubble_forced
    from pgdg_p2_amd.step import GpuSim
    case, st = shock_bubble_forced(scheme=dict(order=2, rk_stages=3, ppc_min=lo, ppc_max=hi))
    ref = OracleSim(case, st)
    gpu = GpuSim(case, st)
    a, b = st.copy(), st.copy()
    before = _cell_counts(case, st)
    for name in ["bin", "split", "merge", "merge", "merge", "merge"]:
        print(f"bounds {lo},{hi}: {name}", flush=True)
        ref.call(name)
        gpu.call(name)
        ref.save(a)
        gpu.save(b)
        bad = compare_states(a, b, cells=False)
        assert not bad, f"{name}: {bad[:6]}"
        after = _cell_counts(case, b)
        if name == "split":
            assert b.s.n_split > 0
            assert np.all(after <= np.maximum(hi, before))
        elif name == "merge":
            assert np.all(after >= np.minimum(lo, before))
        before = after
    assert b.s.n_merge > 0
    gpu.close()


@pytest.mark.gpu
def test_stages_reading_bins_refuse_before_binning(cuda):
    # after an upload the cell bins do not exist yet: split, merge, P2G and
    # the particle update return PGDG_ERR_STATE instead of reading them
    from pgdg_p2_amd.step import FN, GpuSim
    case, st = cases.triple_point_2d(nx=28, ny=12, ppc_axis=3, jitter=0.25)
    gpu = GpuSim(case, st)
    for name in ("split", "merge", "p2g", "update_particle"):
        assert gpu.L.pgdg_stage(gpu.h, FN[name], 0) == -5, name
    gpu.call("bin")
    gpu.call("split")
    gpu.load(st)  # a new upload invalidates the bins again
    assert gpu.L.pgdg_stage(gpu.h, FN["merge"], 0) == -5
    gpu.step(1)
    gpu.close()


@pytest.mark.gpu
def test_step_parity_three_phase_p2_rk3_112x48_with_split_merge(cuda):
    # the three-material triple point as T3 sets it up (phase 1 bottom-right
    # gamma 1.4, phase 2 top-right gamma 1.5, phase 3 left gamma 1.6 with
    # Pinf 0.1) with P2, SSP-RK3, CFL 0.2, 3x3 jittered particles per cell and
    # per-cell bounds [1, 27]: merges start after ~45 steps, splits after ~65;
    # every step bit-exact
    case, st = cases.triple_point_2d(nx=112, ny=48, ppc_axis=3, jitter=0.25, nphase=3,
                                     scheme=dict(order=2, rk_stages=3, cfl=0.2, ppc_min=1,
                                                 ppc_max=27))
    from pgdg_p2_amd.step import GpuSim
    ref = OracleSim(case, st)
    gpu = GpuSim(case, st)
    a, b = st.copy(), st.copy()
    splits = merges = 0
    for k in range(80):
        if k % 10 == 0:
            print(f"tp3-p2rk3: step {k}", flush=True)
        ref.step(1)
        gpu.step(1)
        ref.save(a)
        gpu.save(b)
        bad = compare_states(a, b)
        assert not bad, f"step {k}: {bad[:6]}"
        splits += a.s.n_split
        merges += a.s.n_merge
    gpu.close()
    assert splits > 0 and merges > 100, (splits, merges)


@pytest.mark.gpu
def test_step_parity_three_phase_p2_448x192_lattice_positivity(cuda):
    # the three-material junction at 448x192 P2 cells, one particle per cell:
    # phases appearing in cells next to the junction get means extrapolated
    # from quadratic particle copies; the positivity fallback of MLS_P2G
    # keeps them physical, and the state stays finite and bit-exact
    case, st = cases.triple_point_2d(nx=448, ny=192, ppc_axis=1, jitter=0.0, nphase=3,
                                     scheme=dict(order=2, rk_stages=3, cfl=0.2, ppc_min=1,
                                                 ppc_max=27))
    a, events = run_pair(case, st, 20, "tp3-448x192-p2")
    assert np.isfinite(a.p["x"][:2, :a.n]).all() and np.isfinite(a.c["coef"]).all()


@pytest.mark.gpu
def test_step_parity_underwater_explosion_2d(cuda):
    # the reference's UX routines (HLL faces, HLL ghost boundaries, its
    # limiter and P2G) on its case at 80 x 60 cells, P1 and two RK stages as
    # the reference ships them; splits and merges start within 30 steps
    case, st = cases.underwater_explosion_2d(nx=80)
    a, events = run_pair(case, st, 30, "ux-80x60")
    assert events > 0


@pytest.mark.gpu
def test_step_parity_p2_rk3_underwater_explosion_2d(cuda):
    case, st = cases.underwater_explosion_2d(nx=80, scheme=dict(order=2, rk_stages=3, ppc_min=1,
                                                                  ppc_max=9))
    run_pair(case, st, 20, "ux-80x60-p2")


@pytest.mark.gpu
def test_step_parity_underwater_explosion_reference_grid(cuda):
    # the reference's own resolution: 400 x 300 cells, one particle per cell
    case, st = cases.underwater_explosion_2d(nx=400)
    a, _ = run_pair(case, st, 8, "ux-400x300")
    assert np.isfinite(a.c["coef"]).all()
