"""pgdg_neighbor_counts_noalloc: the neighbour-list sizes of
pgdg_neighbor_counts staged through a fixed device array instead of a
per-call allocation (for callers that cannot allocate device memory).
Checked against pgdg_neighbor_counts through the C ABI on two calls in a
row, the second on a smaller input, and its size limit."""
import ctypes as C

import numpy as np
import pytest

from pgdg_p2_amd import cases

NOALLOC_MAX = 1 << 21


def _counts(gpu, fn):
    n = gpu.count()
    out = np.full(max(n, 1), -7, dtype=np.int32)
    f = getattr(gpu.L, fn)
    f.argtypes = [C.c_void_p, C.POINTER(C.c_int32)]
    rc = f(gpu.h, out.ctypes.data_as(C.POINTER(C.c_int32)))
    return rc, out[:n]


@pytest.mark.gpu
def test_noalloc_neighbor_counts_equal_allocating_entry(cuda):
    from pgdg_p2_amd.step import GpuSim
    for nx, ny in ((56, 24), (28, 12)):  # the second input is smaller
        case, st = cases.triple_point_2d(nx=nx, ny=ny, ppc_axis=3, jitter=0.25)
        gpu = GpuSim(case, st)
        gpu.step(2)
        rc_a, want = _counts(gpu, "pgdg_neighbor_counts")
        rc_b, got = _counts(gpu, "pgdg_neighbor_counts_noalloc")
        assert rc_a == 0 and rc_b == 0
        assert want.size == st.n and want.min() > 0
        assert np.array_equal(got, want), (nx, ny)
        gpu.close()


@pytest.mark.gpu
def test_noalloc_neighbor_counts_refuse_beyond_limit(cuda):
    from pgdg_p2_amd.step import GpuSim
    case, st = cases.triple_point_2d(nx=560, ny=240, ppc_axis=4, jitter=0.0, cap_factor=1.02)
    assert st.n > NOALLOC_MAX
    gpu = GpuSim(case, st)
    rc, _ = _counts(gpu, "pgdg_neighbor_counts_noalloc")
    assert rc == -3  # PGDG_ERR_CAPACITY
    gpu.close()
