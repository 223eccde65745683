"""The mixed step's interface points (the MGFM ghost state at every face
point whose two cells hold different phases, S3:particle_flow.cpp:192-204)
solved as one batch per stage, from the reference's initial p* and from the
trained network's, against the solve inside the face kernel.

Batch from the reference's initial p*: every point runs the same Newton
iterations from the same inputs, so the step is the in-kernel one bit for
bit.  Network's p*0: the Newton iteration still runs to the reference's
tolerance, so the solutions, and the step, differ only within it: cell
mixture density within the mixed form's stated 1e-6 relative in L1 and
1e-3 in Linf at every step, particle counts identical.  The network is
trained on the GPU on records the step's own exact solves produced.  What
it gains is measured on the records of the steps that follow: solved to
1e-12 relative, its starting points take fewer Newton iterations than the
reference's max(pL, pR).  At the reference's tolerance (|f| <= 1e-3 |uL -
uR|, about 5e-4 |uL - uR| in p*) these interface problems converge in ~2.3
iterations from either start: a BF16 network's p*0 (~1e-3 relative) is not
close enough to stop after the first evaluation, so there the iteration
counts are reported, not asserted."""
import numpy as np
import pytest

import pgdg_p2_amd as P
from pgdg_p2_amd import cases

# the shock starts inside the helium sphere's span, so the interface points
# meet a strong pressure and velocity jump from the first step
SCH = dict(order=2, rk_stages=3, ppc_min=4, ppc_max=16)


def _mix_rho(st):
    return (st.c["alph"] * st.c["rho"]).sum(axis=0)


@pytest.mark.gpu
def test_iface_batch_equals_inline(cuda):
    from pgdg_p2_amd.compare import compare_states
    from pgdg_p2_amd.step import GpuSim
    case, st = cases.shock_bubble_3d(nx=32, ny=16, nz=16, x_shock=0.3, scheme=SCH)
    case.precision = 1
    ga, gb = GpuSim(case, st), GpuSim(case, st)
    gb.iface_mode(GpuSim.IFACE_BATCH)
    a, b = st.copy(), st.copy()
    total = 0
    try:
        for k in range(6):
            print(f"iface batch vs inline: step {k}", flush=True)
            ga.step(1)
            gb.step(1)
            ga.save(a)
            gb.save(b)
            bad = compare_states(a, b)
            assert not bad, f"step {k}: {bad[:6]}"
            rin, rout, rpair, rit = gb.iface_records()
            total += rin.shape[1]
            if rin.shape[1]:
                assert rit.min() >= 1 and np.all(np.isfinite(rout))
                # codes name (bubble, air) and (air, bubble) pairs only
                assert set(np.unique(rpair)) <= {0 * 3 + 1, 1 * 3 + 0}
        assert total > 0, "no interface point was solved"
    finally:
        ga.close()
        gb.close()


@pytest.mark.gpu
def test_iface_batch_equals_inline_three_phase_2d(cuda):
    # the T2/T3 MGFM variant (initial p* (pL + pR) / 2), three media (six
    # ordered pairs), P2 / SSP-RK3 with splits and merges, and a 2-D grid
    from pgdg_p2_amd.compare import compare_states
    from pgdg_p2_amd.step import GpuSim
    case, st = cases.triple_point_2d(nx=112, ny=48, ppc_axis=3, jitter=0.25, nphase=3,
                                     scheme=dict(order=2, rk_stages=3, ppc_min=1, ppc_max=27))
    case.precision = 1
    ga, gb = GpuSim(case, st), GpuSim(case, st)
    gb.iface_mode(GpuSim.IFACE_BATCH)
    a, b = st.copy(), st.copy()
    codes = set()
    try:
        for k in range(20):
            print(f"iface batch vs inline, three phases: step {k}", flush=True)
            ga.step(1)
            gb.step(1)
            ga.save(a)
            gb.save(b)
            bad = compare_states(a, b)
            assert not bad, f"step {k}: {bad[:6]}"
            _, _, rpair, _ = gb.iface_records()
            codes |= set(np.unique(rpair).tolist())
        # records of at least two of the three interfaces, each pair code
        # naming two different media
        assert len(codes) >= 2 and all(c // 3 != c % 3 for c in codes), codes
    finally:
        ga.close()
        gb.close()


@pytest.mark.gpu
def test_iface_learned_guess_in_step(cuda):
    from pgdg_p2_amd.mlp import Mlp, mgfm_pairs_dev
    from pgdg_p2_amd.step import GpuSim
    case, st = cases.shock_bubble_3d(nx=64, ny=32, nz=32, x_shock=0.3, scheme=SCH)
    case.precision = 1
    ga, gb = GpuSim(case, st), GpuSim(case, st)
    gb.iface_mode(GpuSim.IFACE_BATCH)
    a, b = st.copy(), st.copy()
    m = Mlp()
    try:
        # a first step sizes the record buffer (faces past its capacity solve
        # their points in the face kernel), then the records of the next
        # steps' exact solves are the training set
        ga.step(1)
        gb.step(1)
        rec_in, rec_pair = [], []
        for k in range(3):
            print(f"learned guess: collect step {k}", flush=True)
            ga.step(1)
            gb.step(1)
            rin, _, rpair, _ = gb.iface_records()
            rec_in.append(rin)
            rec_pair.append(rpair)
        X = np.ascontiguousarray(np.concatenate(rec_in, axis=1))
        codes = np.ascontiguousarray(np.concatenate(rec_pair)).astype(np.int32)
        n = X.shape[1] // 1024 * 1024
        print(f"learned guess: training on {n} step records", flush=True)
        assert n >= 4096
        d_in = P.DeviceArray.from_numpy(np.ascontiguousarray(X[:, :n]))
        d_pair = P.DeviceArray.from_numpy(codes[:n])
        hist = m.train_pairs(d_in, d_pair, n, gb.iface_pairs(), epochs=30, batch=1024, lr=3e-3)
        assert hist[-1] < 0.5 * hist[0], (hist[0], hist[-1])
        gb.iface_mode(GpuSim.IFACE_LEARNED, m)
        its_l, its_p = [], []
        for k in range(6):
            print(f"learned guess: step {k}", flush=True)
            ga.step(1)
            gb.step(1)
            ga.save(a)
            gb.save(b)
            assert a.n == b.n, (k, a.n, b.n)
            ra, rb = _mix_rho(a), _mix_rho(b)
            l1 = np.abs(rb - ra).sum() / np.abs(ra).sum()
            linf = np.abs(rb - ra).max() / np.abs(ra).max()
            assert l1 <= 1e-6 and linf <= 1e-3, (k, l1, linf)
            # the stage's records: the network's starting points against the
            # reference's on the same inputs (the reference's tolerance)
            rin, rout, rpair, rit = gb.iface_records()
            nr = rin.shape[1]
            assert nr > 0
            d_x = P.DeviceArray.from_numpy(np.ascontiguousarray(rin))
            d_c = P.DeviceArray.from_numpy(np.ascontiguousarray(rpair))
            d_o = P.DeviceArray((3, nr), np.float64)
            d_it = P.DeviceArray((nr,), np.int32)
            its = {}
            for who in ("ref", "net"):
                for tol, (mode, rtol, mx) in (("tol_ref", (P.TOL_REFERENCE, 0.0, 50)),
                                              ("1e-12", (P.TOL_RELATIVE, 1e-12, 100))):
                    mgfm_pairs_dev(m if who == "net" else None, d_x, d_c, d_o, d_it, nr, gb.iface_pairs(),
                                   tol_mode=mode, rtol=rtol, max_iter=mx)
                    P.device_sync()
                    its[who, tol] = d_it.numpy().astype(np.float64).mean()
            # the step's own learned solves took the network's iterations
            assert abs(rit.astype(np.float64).mean() - its["net", "tol_ref"]) < 1e-12
            its_l.append(its["net", "1e-12"])
            its_p.append(its["ref", "1e-12"])
            print(f"learned guess: step {k}: {nr} interface points, Newton iterations per point at the "
                  f"reference's tolerance {its['net', 'tol_ref']:.3f} (reference's p*: {its['ref', 'tol_ref']:.3f}), "
                  f"to 1e-12 {its['net', '1e-12']:.3f} (reference's p*: {its['ref', '1e-12']:.3f})", flush=True)
        assert np.mean(its_l) < np.mean(its_p)
    finally:
        ga.close()
        gb.close()
        m.close()
