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
