"""The mixed-precision step (FP32 particle-shifting pair terms in cell-local
coordinates, FP64 sums and state) against the FP64 step on the GPU (which
is bit-identical to the oracle, see test_gpu_step.py).

Stated tolerance: cell mixture density within 1e-8 relative in L1 and 1e-6
relative in Linf after 20 steps, particle counts (split/merge outcomes)
identical at every step, positions within 1e-5 particle spacings."""
import numpy as np
import pytest

from pgdg_p2_amd import cases

L1_TOL, LINF_TOL, DX_TOL = 1e-8, 1e-6, 1e-5


def _mix_rho(st):
    return (st.c["alph"] * st.c["rho"]).sum(axis=0)


def _compare(make, steps, label):
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
            dx = np.abs(a.live("x") - b.live("x")).max() / case.p_sz
            assert l1 <= L1_TOL and linf <= LINF_TOL and dx <= DX_TOL, (label, k, l1, linf, dx)
            worst = max(worst, linf)
    finally:
        gp.close()
        gm.close()
    print(f"{label}: worst density Linf {worst:.3e}")
    assert worst > 0  # the mixed path did run


@pytest.mark.gpu
def test_mixed_step_shock_bubble_3d(cuda):
    _compare(lambda: cases.shock_bubble_3d(nx=32, ny=16, nz=16), 20, "sb3d-mixed")


@pytest.mark.gpu
def test_mixed_step_triple_point_2d(cuda):
    _compare(lambda: cases.triple_point_2d(nx=56, ny=24, ppc_axis=3), 20, "tp2d-mixed")
