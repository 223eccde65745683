"""The reference's outer loop with its outputs, for a context created with
`case.output_dt` set to the output interval.

main() (S3:main.cpp:38-52) writes output 0 and then runs num_output
intervals; Solver::MPM_Solver (S3:solver.cpp:29-62) steps until the
interval's time reaches output_dt (the last step is truncated onto the
output instant, which pgdg_step does by itself), MPM_run_simulation calls
Output_point after that last step (S3:mpm_ini.cpp:759-760,
T2:mpm_ini.cpp:584-585), and the interval ends with Output_plt_file(n).
The time written is the reference's mpm.run_time: the interval's start
time (accumulated by adding output_dt) plus the interval's integral time.
"""
from __future__ import annotations

import os

import numpy as np

from . import output
from .state import HostState


def _counts(sim, st: HostState):
    """Neighbour-list sizes and used-cell flags (any driver with the
    neighbor_counts / used_cells accessors of pgdg_p2_amd.step.GpuSim)."""
    nc = int(np.prod([sim.case.n[a] if a < sim.case.dim else 1 for a in range(3)]))
    return sim.neighbor_counts(st.n), sim.used_cells(nc)


def _scalars(sim, st: HostState):
    return sim.scalars()


def _plt(sim, st: HostState, directory: str, n: int, rank: int, run_time: float):
    sim.save(st)
    nb, used = _counts(sim, st)
    case = sim.case
    output.write_particles_plt(os.path.join(directory, f"simulation_sp.{n}.{rank}.plt"), case, st, nb, run_time)
    output.write_cells_plt(os.path.join(directory, f"ele_file.{n}.{rank}.plt"), case, st, run_time)
    return used


def run_case(sim, st: HostState, num_output: int, directory: str, point_dir: str | None = None,
             rank: int = 0) -> float:
    """Runs num_output output intervals of `sim` (a GpuSim, or the oracle
    for parity runs), writing the Tecplot files into `directory`
    (outdata/ in the reference) and the Output_point lines into point_dir
    (the working directory in the reference; default: `directory`).
    `st` is scratch host state of the case.  Returns the final time."""
    case = sim.case
    out_dt = case.output_dt
    if not (out_dt > 0 and out_dt < 1e29):
        raise ValueError("case.output_dt must be set to the output interval")
    point_dir = point_dir or directory
    os.makedirs(directory, exist_ok=True)
    os.makedirs(point_dir, exist_ok=True)
    time = _scalars(sim, st).run_time
    _plt(sim, st, directory, 0, rank, time)
    for n in range(1, num_output + 1):
        while True:
            sim.step(1)
            s = _scalars(sim, st)
            if not (s.integral_time < out_dt - 1.0e-10):
                break
        run_time = time + s.integral_time
        used = _plt(sim, st, directory, n, rank, run_time)
        d = output.point_diagnostics(case, st, used, run_time)
        output.point_append(point_dir, case, d, n == 1)
        time += out_dt
    return time
