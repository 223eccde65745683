"""Python driver of the HIP PG-DG step (thin ctypes wrapper over libpgdg.so)."""
from __future__ import annotations

import ctypes as C

from . import check, lib
from .state import Case, HostState, Scalars

FN = {"set_timestep": 0, "used_state_update": 1, "bin": 2, "split": 3, "merge": 4,
      "membership": 5, "neighbors": 6, "p2g": 7, "limiter": 8, "rhs": 9, "update_grid": 10,
      "update_particle": 11, "run_simulation": 12}


PROFILE_SLOTS = ["set_timestep", "used_state_update", "bin", "split", "merge", "membership",
                 "stage_save", "p2g", "limiter", "rhs", "update_grid", "g2p", "pst", "lists",
                 "pst_prep", "compact"]


class GpuSim:
    """One case resident on the GPU (a pgdg_ctx)."""

    def __init__(self, case: Case, st: HostState, cap: int | None = None):
        self.L = lib()
        self.case = case
        self.h = C.c_void_p()
        check(self.L.pgdg_create(C.byref(case), cap or st.cap, st.nwall, C.byref(self.h)),
              "pgdg_create")
        self.load(st)

    def load(self, st: HostState):
        P, W, Cc = st.particles(), st.walls(), st.cells()
        check(self.L.pgdg_upload_state(self.h, C.byref(P), C.byref(W), C.byref(Cc),
                                       C.byref(st.s)), "pgdg_upload_state")

    def save(self, st: HostState) -> HostState:
        P, Cc = st.particles(), st.cells()
        check(self.L.pgdg_download_state(self.h, C.byref(P), C.byref(Cc), C.byref(st.s)),
              "pgdg_download_state")
        st.n = P.n
        return st

    def save_cells(self, st: HostState) -> HostState:
        """Download only the cell arrays and scalars."""
        Cc = st.cells()
        check(self.L.pgdg_download_state(self.h, None, C.byref(Cc), C.byref(st.s)),
              "pgdg_download_state")
        return st

    def scalars(self) -> Scalars:
        """The time-loop scalars and the last step's split/inlet/outlet/merge
        counts, without downloading particle or cell arrays."""
        sc = Scalars()
        check(self.L.pgdg_download_state(self.h, None, None, C.byref(sc)), "pgdg_download_state")
        return sc

    def graphs(self, on: bool):
        """Run the RK stages as captured HIP graphs (default) or launch by launch."""
        check(self.L.pgdg_graphs_enable(self.h, 1 if on else 0), "pgdg_graphs_enable")

    def checkpoint(self, path: str):
        """Write the context's state (arrays, scalars, used cells, carried
        neighbour lists) to `path` (pgdg_checkpoint_write)."""
        import os
        check(self.L.pgdg_checkpoint_write(self.h, os.fsencode(path)), "pgdg_checkpoint_write")

    def restore(self, path: str):
        """Replace the context's state by a checkpoint of the same case
        (pgdg_checkpoint_read); the next step continues the written run."""
        import os
        check(self.L.pgdg_checkpoint_read(self.h, os.fsencode(path)), "pgdg_checkpoint_read")

    def pst_staging(self, on: bool):
        """Mixed form, 3-D: particle-shifting candidates staged in LDS (default)
        or read by every lane from global memory (same sums bit for bit)."""
        check(self.L.pgdg_pst_staging_enable(self.h, 1 if on else 0), "pgdg_pst_staging_enable")

    def rhs_mfma(self, on: bool):
        """Mixed form, 3-D P2: DG right-hand side as matrix products on the
        matrix cores or one lane per cell (default; pgdg_rhs_mfma_enable)."""
        check(self.L.pgdg_rhs_mfma_enable(self.h, 1 if on else 0), "pgdg_rhs_mfma_enable")

    IFACE_INLINE, IFACE_BATCH, IFACE_LEARNED = 0, 1, 2

    def iface_mode(self, mode: int, mlp=None):
        """How the mixed step solves MGFM at interface face points
        (pgdg_iface_mgfm_mode): in the face kernel, as one batch, or as one
        batch from the trained network mlp's initial p*."""
        check(self.L.pgdg_iface_mgfm_mode(self.h, mode, mlp.h if mlp is not None else None),
              "pgdg_iface_mgfm_mode")

    def iface_pairs(self):
        """EOS pair (gamma_l, pinf_l, gamma_r, pinf_r) of each record pair code."""
        from . import EosPair
        from .state import PGDG_MAX_PHASE
        tab = (EosPair * (PGDG_MAX_PHASE * PGDG_MAX_PHASE))()
        check(self.L.pgdg_iface_pairs(self.h, tab), "pgdg_iface_pairs")
        return [(e.gamma_l, e.pinf_l, e.gamma_r, e.pinf_r) for e in tab]

    def iface_records(self, cap: int | None = None):
        """The last batched stage's interface-point records as host arrays:
        (inputs [6][n], solutions [3][n], pair codes [n], iterations [n])."""
        import numpy as np
        n = C.c_int64(0)
        check(self.L.pgdg_iface_records(self.h, None, None, None, None, 0, C.byref(n)), "pgdg_iface_records")
        k = n.value if cap is None else min(cap, n.value)
        a_in = np.zeros((6, k), np.float64)
        a_out = np.zeros((3, k), np.float64)
        a_pair = np.zeros(k, np.int32)
        a_it = np.zeros(k, np.int32)
        if k:
            check(self.L.pgdg_iface_records(self.h, C.c_void_p(a_in.ctypes.data), C.c_void_p(a_out.ctypes.data),
                                            C.c_void_p(a_pair.ctypes.data), C.c_void_p(a_it.ctypes.data), k,
                                            C.byref(n)), "pgdg_iface_records")
        return a_in, a_out, a_pair, a_it

    def autotune(self, on: bool):
        check(self.L.pgdg_autotune_enable(self.h, 1 if on else 0), "pgdg_autotune_enable")

    def tuning(self):
        """Chosen workgroup sizes {pst, rhs, limiter}, or None before tuning."""
        b = (C.c_int * 3)()
        rc = self.L.pgdg_tuning_read(self.h, b)
        return {"pst": b[0], "rhs": b[1], "limiter": b[2]} if rc == 0 else None

    def set_tuning(self, pst: int, rhs: int, limiter: int):
        """Fix the workgroup sizes (64, 128 or 256 each); the autotuner is off."""
        b = (C.c_int * 3)(pst, rhs, limiter)
        check(self.L.pgdg_tuning_set(self.h, b), "pgdg_tuning_set")

    def step(self, n: int = 1):
        check(self.L.pgdg_step(self.h, n), "pgdg_step")

    def call(self, fn: str, arg: int = 0):
        check(self.L.pgdg_stage(self.h, FN[fn], arg), f"pgdg_stage({fn})")

    def sync(self):
        check(self.L.pgdg_sync(self.h), "pgdg_sync")

    def stream(self) -> int:
        s = C.c_void_p()
        check(self.L.pgdg_get_stream(self.h, C.byref(s)))
        return s.value or 0

    def profile(self, on: bool = True):
        check(self.L.pgdg_profile_enable(self.h, 1 if on else 0), "pgdg_profile_enable")

    def profile_read(self) -> dict:
        ms = (C.c_double * len(PROFILE_SLOTS))()
        cnt = (C.c_int64 * len(PROFILE_SLOTS))()
        check(self.L.pgdg_profile_read(self.h, ms, cnt, len(PROFILE_SLOTS)), "pgdg_profile_read")
        return {k: (ms[i], cnt[i]) for i, k in enumerate(PROFILE_SLOTS)}

    def memory_bytes(self) -> int:
        b = C.c_int64()
        check(self.L.pgdg_memory_bytes(self.h, C.byref(b)))
        return b.value

    def neighbor_counts(self, n: int | None = None):
        """particle->neighbor.size() of every fluid particle (particle order)."""
        import numpy as np
        m = self.count()
        out = np.zeros(max(m, 1), dtype=np.int32)
        check(self.L.pgdg_neighbor_counts(self.h, out.ctypes.data_as(C.POINTER(C.c_int32))),
              "pgdg_neighbor_counts")
        return out[:m]

    def used_cells(self, nc: int | None = None):
        """1 for each cell in use (Used_Ele_cen), else 0 (cell order)."""
        import numpy as np
        out = np.zeros(self.case.nc, dtype=np.int32)
        check(self.L.pgdg_used_cells(self.h, out.ctypes.data_as(C.POINTER(C.c_int32))),
              "pgdg_used_cells")
        return out

    def cell_particle_counts(self):
        """Fluid particles in each cell's list (cell order)."""
        import numpy as np
        out = np.zeros(self.case.nc, dtype=np.int32)
        check(self.L.pgdg_cell_particle_counts(self.h, out.ctypes.data_as(C.POINTER(C.c_int32))),
              "pgdg_cell_particle_counts")
        return out

    def count(self) -> int:
        n = C.c_int64()
        check(self.L.pgdg_particle_count(self.h, C.byref(n)))
        return n.value

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            self.L.pgdg_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
