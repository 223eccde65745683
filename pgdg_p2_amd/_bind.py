"""ctypes signatures of the step-level C ABI (include/pgdg.h)."""
import ctypes as C

from .state import Case, Cells, Particles, Scalars, Walls

_CTX = C.c_void_p
_STAGES = ["pgdg_set_timestep", "pgdg_used_state_update", "pgdg_map_particles_to_cells",
           "pgdg_split_particle_fast", "pgdg_merge_particle_fast", "pgdg_refresh_cell_membership",
           "pgdg_mls_p2g", "pgdg_calculate_limiter", "pgdg_grid_acceleration"]


def declare(L: C.CDLL) -> None:
    L.pgdg_create.argtypes = [C.POINTER(Case), C.c_int64, C.c_int64, C.POINTER(_CTX)]
    L.pgdg_destroy.argtypes = [_CTX]
    L.pgdg_upload_state.argtypes = [_CTX, C.POINTER(Particles), C.POINTER(Walls),
                                    C.POINTER(Cells), C.POINTER(Scalars)]
    L.pgdg_download_state.argtypes = [_CTX, C.POINTER(Particles), C.POINTER(Cells),
                                      C.POINTER(Scalars)]
    L.pgdg_step.argtypes = [_CTX, C.c_int]
    L.pgdg_stage.argtypes = [_CTX, C.c_int, C.c_int]
    L.pgdg_sync.argtypes = [_CTX]
    L.pgdg_get_stream.argtypes = [_CTX, C.POINTER(C.c_void_p)]
    L.pgdg_particle_count.argtypes = [_CTX, C.POINTER(C.c_int64)]
    L.pgdg_memory_bytes.argtypes = [_CTX, C.POINTER(C.c_int64)]
    L.pgdg_profile_enable.argtypes = [_CTX, C.c_int]
    L.pgdg_graphs_enable.argtypes = [_CTX, C.c_int]
    L.pgdg_autotune_enable.argtypes = [_CTX, C.c_int]
    L.pgdg_pst_staging_enable.argtypes = [_CTX, C.c_int]
    L.pgdg_rhs_mfma_enable.argtypes = [_CTX, C.c_int]
    L.pgdg_iface_mgfm_mode.argtypes = [_CTX, C.c_int, C.c_void_p]
    from . import EosPair
    L.pgdg_iface_pairs.argtypes = [_CTX, C.POINTER(EosPair)]
    L.pgdg_iface_records.argtypes = [_CTX, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64,
                                     C.POINTER(C.c_int64)]
    L.pgdg_neighbor_counts.argtypes = [_CTX, C.POINTER(C.c_int32)]
    L.pgdg_neighbor_counts_noalloc.argtypes = [_CTX, C.POINTER(C.c_int32)]
    L.pgdg_used_cells.argtypes = [_CTX, C.POINTER(C.c_int32)]
    L.pgdg_cell_particle_counts.argtypes = [_CTX, C.POINTER(C.c_int32)]
    L.pgdg_checkpoint_write.argtypes = [_CTX, C.c_char_p]
    L.pgdg_checkpoint_read.argtypes = [_CTX, C.c_char_p]
    from .output import PointDiag
    L.pgdg_plt_write_particles.argtypes = [C.c_char_p, C.POINTER(Case), C.POINTER(Particles),
                                           C.POINTER(C.c_int32), C.c_double]
    L.pgdg_plt_write_cells.argtypes = [C.c_char_p, C.POINTER(Case), C.POINTER(Cells), C.c_double]
    L.pgdg_point_diagnostics.argtypes = [C.POINTER(Case), C.POINTER(Particles), C.POINTER(Cells),
                                         C.POINTER(C.c_int32), C.c_double, C.POINTER(PointDiag)]
    L.pgdg_point_append.argtypes = [C.c_char_p, C.POINTER(Case), C.POINTER(PointDiag), C.c_int]
    L.pgdg_tuning_read.argtypes = [_CTX, C.POINTER(C.c_int)]
    L.pgdg_tuning_set.argtypes = [_CTX, C.POINTER(C.c_int)]
    L.pgdg_profile_read.argtypes = [_CTX, C.POINTER(C.c_double), C.POINTER(C.c_int64), C.c_int]
    for name in _STAGES:
        getattr(L, name).argtypes = [_CTX]
    L.pgdg_refresh_neighbor_info.argtypes = [_CTX, C.c_int]
    L.pgdg_update_grid.argtypes = [_CTX, C.c_int]
    L.pgdg_update_particle.argtypes = [_CTX, C.c_int]
    L.pgdg_run_simulation.argtypes = [_CTX, C.c_int64]
