"""The reference's output files and diagnostics (MPM::Output_plt_file,
S3:mpm_ini.cpp:832-980 / T2:mpm_ini.cpp:662-840, and MPM::Output_point,
S3:mpm_ini.cpp:774-827 / T2:mpm_ini.cpp:599-654), through the library's
writers; the neighbour-list sizes and the cells in use come from the GPU
context."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import check, lib
from .state import Case, HostState


class PointDiag(C.Structure):
    _fields_ = [("time", C.c_double), ("dw", C.c_double), ("jet", C.c_double), ("ups", C.c_double),
                ("mass2", C.c_double), ("n_fluid", C.c_int64), ("p_mass", C.c_double * 2),
                ("p_energy", C.c_double), ("g_mass", C.c_double * 2), ("g_energy", C.c_double),
                ("p2g", C.c_double)]


def _i32p(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def neighbor_counts(gpu) -> np.ndarray:
    """particle->neighbor.size() of every fluid particle (particle order)."""
    out = np.zeros(max(gpu.count(), 1), dtype=np.int32)
    check(gpu.L.pgdg_neighbor_counts(gpu.h, _i32p(out)), "pgdg_neighbor_counts")
    return out[:gpu.count()]


def used_cells(gpu) -> np.ndarray:
    nc = int(np.prod([gpu.case.n[a] if a < gpu.case.dim else 1 for a in range(3)]))
    out = np.zeros(nc, dtype=np.int32)
    check(gpu.L.pgdg_used_cells(gpu.h, _i32p(out)), "pgdg_used_cells")
    return out


def write_particles_plt(path: str, case: Case, st: HostState, n_neighbor: np.ndarray, run_time: float):
    nb = np.ascontiguousarray(n_neighbor, dtype=np.int32)
    P = st.particles()
    check(lib().pgdg_plt_write_particles(os.fsencode(path), C.byref(case), C.byref(P), _i32p(nb),
                                         run_time), "pgdg_plt_write_particles")


def write_cells_plt(path: str, case: Case, st: HostState, run_time: float):
    Cc = st.cells()
    check(lib().pgdg_plt_write_cells(os.fsencode(path), C.byref(case), C.byref(Cc), run_time),
          "pgdg_plt_write_cells")


def point_diagnostics(case: Case, st: HostState, used: np.ndarray, time: float) -> PointDiag:
    d = PointDiag()
    u = np.ascontiguousarray(used, dtype=np.int32)
    check(lib().pgdg_point_diagnostics(C.byref(case), C.byref(st.particles()), C.byref(st.cells()),
                                       _i32p(u), time, C.byref(d)), "pgdg_point_diagnostics")
    return d


def point_append(directory: str, case: Case, d: PointDiag, first: bool):
    check(lib().pgdg_point_append(os.fsencode(directory), C.byref(case), C.byref(d), 1 if first else 0),
          "pgdg_point_append")


def output_plt(gpu, st: HostState, directory: str, n: int, rank: int = 0):
    """MPM::Output_plt_file for the context's current state: downloads it
    into st and writes simulation_sp.<n>.<rank>.plt and ele_file.<n>.<rank>.plt
    into directory (the reference's outdata/ names)."""
    gpu.save(st)
    nb = neighbor_counts(gpu)
    rt = st.s.run_time
    write_particles_plt(os.path.join(directory, f"simulation_sp.{n}.{rank}.plt"), gpu.case, st, nb, rt)
    write_cells_plt(os.path.join(directory, f"ele_file.{n}.{rank}.plt"), gpu.case, st, rt)
