"""Benchmark of the PG-DG time step on one MI355X.

Workload (default): the 3-D shock-bubble step, channel 256x128x128 P2 DG cells
(10 modes, 3-point Gauss rules) with 8 particles per cell (2x2x2 sub-lattice,
seeded jitter), air with a helium sphere hit by a Mach 1.22 shock; SSP-RK3
stages, split/merge with per-cell bounds [4, 16].  Plain FP64 and mixed forms
are both run; `value` is the mixed form's whole-job cell-steps/s with the state
resident in HBM.  `--scheme p1` runs the reference's P1 basis and two-stage
step instead.  Secondary measurements: the two forms side by side for 200
steps on a reduced grid (density error horizon), the 2-D triple point at
112x48 (plain-method workload), the 2-D three-phase triple point at
1792x768 after 500 steps of split/merge, the batched MGFM solve on 1e7
interface face pairs (plain and learned initial guess), the MLS fit, and the
CPU oracle on a bounded sample (all host worker threads and one thread).

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--size full|medium|small]
                       [--scheme p2rk3|p1]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
SIZES = {"full": (256, 128, 128), "medium": (128, 64, 64), "small": (32, 16, 16)}
SCHEMES = {
    # the named workload: P2 basis, SSP-RK3, per-cell particle bounds [4, 16]
    "p2rk3": dict(order=2, rk_stages=3, ppc_min=4, ppc_max=16),
    # the reference's own scheme: P1 basis, two-stage Heun, volume criteria
    "p1": dict(order=1, rk_stages=2),
}


def stage_bytes(stage: str, case, n: int, nwall: int) -> float:
    """Algorithmic HBM bytes of one launch of a stage: every input element
    read once, every output element written once (FP64 = 8 B, int32 = 4 B)."""
    D, F = case.dim, case.nphase
    U, M = D + 2, case.M
    nc = case.n[0] * case.n[1] * case.n[2]
    bins = (n + nwall) * 4 + (nc + 1) * 4
    if stage == "pst":
        # x, xb (2D), h, hb, rho, c, vol, v (D) read; phase, mark_in, in_ele;
        # tv (D) written
        return n * ((3 * D + 5) * 8 + 12 + D * 8) + bins
    if stage == "g2p":
        # x, tv, x0 (3D) + vol, vol0 read; in_ele, phase; cell coef/G/rho.. ;
        # x, vol, rho, rhov(3), v(3), E, P, c, mass, h, coef written
        return n * ((3 * D + 2) * 8 + 8 + (D + 11 + 6 + U * M) * 8) + nc * (F * U * M + 9 + 4 * F) * 8
    if stage == "rhs":
        # coef, lim, alph, ph_state read; dcoef written
        return nc * (F * U * M * 16 + F * U * 8 + F * 12)
    if stage == "p2g":
        # particle x, vol, rho, rhov, E, coef, phase read; cell alph, ph_state,
        # ph_state_b, v, P written
        return n * ((2 * D + 3 + U * M) * 8 + 4) + bins + nc * (F * 16 + 32)
    if stage == "limiter":
        return nc * (F * (U * M + 2 + D + 1) * 8 + 24 + F * U * 8 + 72)
    if stage == "update_grid":
        return nc * (F * U * M * 8 * 5 + F * U * 8 + F * 8 * 6 + 48)
    if stage == "lists":
        return n * ((D + 1) * 8 + 8) + bins
    if stage == "bin":
        return (n + nwall) * (D * 8 + 12) + bins
    return 0.0


def sec8d_cell_step_bytes(D: int, U: int, M: int, F: int, S: int, ppc: float) -> float:
    """Algorithmic HBM bytes per cell-step of the whole step (plain FP64
    state, particle coefficient copy by reference, one present phase):
    B = S*B_cs + ppc*(S*B_ps + B_pstep)."""
    b_cs = 8 * (U * M + F + U) + 8 * (3 * U * M + U + F + D * D)
    b_ps = (8 * (2 * D + 1) + 8) + 8 * (D + 1 + U) + (8 * (2 * D + 3) + 4) + 8 * D
    b_pstep = (8 * (D + 1) + 4) + (8 * D + 8) + (8 * (D + 1) + 4)
    return S * b_cs + ppc * (S * b_ps + b_pstep)


def sec8d_pst_bytes(D: int) -> float:
    """Particle shifting per particle-stage: x, v, vol, rho, c, phase read;
    tv written."""
    return 8 * (2 * D + 3) + 4 + 8 * D


def sec8d_mfma_flops(D: int, U: int, M: int, S: int) -> float:
    """Quadrature-by-mode contraction flops per cell-step (P2, 3-point rules)."""
    qv, qf = 3 ** D, 3 ** (D - 1)
    return S * 2 * U * M * (D * qv + 2 * D * qf)


def worker_threads() -> int:
    """Host worker threads: the process's share of the host's cores (the
    OMP_NUM_THREADS of the environment, else the core count)."""
    try:
        return max(1, int(os.environ.get("OMP_NUM_THREADS", "")))
    except ValueError:
        return os.cpu_count() or 1


TRAFFIC_PROFILES = {
    "p2rk3": ROOT / "profiles" / "step_shock_bubble_256x128x128_p2_rk3_kernels_and_traffic.txt",
}


def profiled_traffic(kernel: str, size: str, scheme: str):
    """HBM bytes per dispatch of `kernel` from the committed rocprofv3 --pmc
    summary (FETCH_SIZE and WRITE_SIZE scaled by the copy-kernel
    calibration), for the full-size workload only; None if absent."""
    prof = TRAFFIC_PROFILES.get(scheme)
    if size != "full" or prof is None or not prof.exists():
        return None
    for line in prof.read_text().splitlines():
        cols = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cols) == 4 and cols[0] == f"`{kernel}`":
            try:
                return (float(cols[2]) + float(cols[3])) * 1e9
            except ValueError:
                return None
    return None


def cpu_baseline(scheme, sample=(48, 24, 24), steps=2):
    """The oracle (the C++ FP64 restatement of the reference step; its
    per-cell and per-particle loops run on worker threads, split and merge
    sequentially, as in the reference) on a reduced grid of the same
    workload, timed on this host with all worker threads and with one."""
    import ctypes as C
    from oracle.pyoracle import lib as olib
    from oracle.pyoracle_step import OracleSim
    from pgdg_p2_amd import cases
    L = olib()
    L.oracle_set_threads.argtypes = [C.c_int]
    nc = sample[0] * sample[1] * sample[2]
    res = {}
    threads = worker_threads()
    for t in (threads, 1):
        L.oracle_set_threads(t)
        case, st = cases.shock_bubble_3d(nx=sample[0], ny=sample[1], nz=sample[2], ppc_axis=2,
                                         jitter=0.25, scheme=scheme)
        sim = OracleSim(case, st)
        sim.step(1)  # first step (particle dt path) excluded
        t0 = time.perf_counter()
        sim.step(steps)
        res[t] = nc * steps / (time.perf_counter() - t0)
        npart = st.n
        print(f"[cpu] {t} threads: {res[t]:.4g} cell-steps/s", file=sys.stderr, flush=True)
    L.oracle_set_threads(1)
    return {"value": res[threads], "unit": "cell-steps/s", "cores": threads, "kind": "port",
            "value_1_thread": res[1],
            "sample": f"{sample[0]}x{sample[1]}x{sample[2]} cells x 8 ppc ({npart} particles), "
                      f"steps 2..{steps + 1} of the same workload, oracle FP64 on {threads} worker "
                      f"threads (and 1)"}


def three_phase_bench(steps=20, warmup=500, nx=1792, ny=768):
    """The 2-D three-phase triple point at 1792x768 P2 cells (gamma 1.5 /
    1.4 / 1.67), 3x3 jittered particles per cell, SSP-RK3, CFL 0.2,
    per-cell bounds [1, 27], mixed form: `warmup` steps of split/merge
    spread the particle counts per cell; then cell-steps/s, particles/s and
    split/merge particles/s over `steps` timed steps, and the distribution
    of particles per cell."""
    import numpy as np
    from pgdg_p2_amd import cases
    from pgdg_p2_amd.step import GpuSim
    t0 = time.perf_counter()
    case, st = cases.triple_point_2d(nx=nx, ny=ny, ppc_axis=3, jitter=0.25, nphase=3, cap_factor=2.0,
                                     scheme=dict(order=2, rk_stages=3, cfl=0.2, ppc_min=1, ppc_max=27))
    case.precision = 1
    gpu = GpuSim(case, st)
    n_init = st.n
    st.p = {}
    print(f"[3-phase] setup {time.perf_counter() - t0:.1f}s: {nx}x{ny} cells, {n_init} particles",
          file=sys.stderr, flush=True)
    t1 = time.perf_counter()
    ev_w = [0, 0]
    for k in range(warmup):
        gpu.step(1)
        if k % 100 == 99:
            sc = gpu.scalars()
            print(f"[3-phase] warm-up step {k + 1}: {gpu.count()} particles", file=sys.stderr, flush=True)
    gpu.sync()
    t_w = time.perf_counter() - t1
    n0 = gpu.count()
    split = merge = 0
    t = time.perf_counter()
    for _ in range(steps):
        gpu.step(1)
        sc = gpu.scalars()
        split += sc.n_split
        merge += sc.n_merge
    gpu.sync()
    el = time.perf_counter() - t
    n1 = gpu.count()
    cnt = gpu.cell_particle_counts()
    gpu.close()
    hist = np.bincount(cnt, minlength=28)
    nz = cnt[cnt > 0]
    ppc_avg = (n0 + n1) / 2 / (nx * ny)
    b_cs = sec8d_cell_step_bytes(2, 4, 6, 3, 3, ppc_avg)
    return {"workload": f"2-D three-phase triple point, {nx}x{ny} P2 cells, 3x3 jittered particles "
                        "per cell, SSP-RK3, CFL 0.2, per-cell bounds [1,27], mixed form",
            "cells": nx * ny, "particles_initial": n_init, "particles_start": n0, "particles_end": n1,
            "warmup_steps": warmup, "warmup_s": t_w, "steps": steps,
            "ms_per_step": el / steps * 1e3, "cell_steps_per_s": nx * ny * steps / el,
            "particles_per_s": (n0 + n1) / 2 * steps / el,
            "splits": split, "merges": merge,
            "split_merge_particles_per_s": (split * 4 + merge * 2) / el,
            "ppc": {"min": int(cnt.min()), "max": int(cnt.max()), "mean": float(cnt.mean()),
                    "empty_cells": int((cnt == 0).sum()),
                    "p01_p50_p99": [float(np.percentile(nz, q)) for q in (1, 50, 99)],
                    "histogram_0_to_27_and_more": [int(v) for v in hist[:28]] + [int(hist[28:].sum())]},
            "roofline_step": {"bound": "hbm", "bytes_per_cell_step": b_cs,
                              "achieved_GBs": b_cs * nx * ny * steps / el / 1e9,
                              "frac": b_cs * nx * ny * steps / el / 1e9 / HBM_PEAK_GBS},
            "counts_vs_oracle": "bit-exact counts and order at 112x48 over 60 steps (GPU test); the "
                                "oracle is not run at this size"}


def triple_point_bench(steps=100, warmup=2, cpu_steps=10):
    """The plain-method workload: the 2-D triple point on [0,7]x[0,3] with
    112x48 P2 cells, 3x3 jittered particles per cell, SSP-RK3, CFL 0.2,
    100 steps, FP64 and mixed forms on the GPU; the oracle on the host
    worker threads for the first cpu_steps steps."""
    import ctypes as C
    from oracle.pyoracle import lib as olib
    from oracle.pyoracle_step import OracleSim
    from pgdg_p2_amd import cases
    from pgdg_p2_amd.step import GpuSim
    nx, ny = 112, 48
    res = {"workload": "2-D triple point, 112x48 P2 cells, 3x3 jittered particles per cell, "
                       "SSP-RK3, CFL 0.2"}
    for prec, name in ((0, "fp64"), (1, "mixed")):
        case, st = cases.triple_point_2d(nx=nx, ny=ny, ppc_axis=3, jitter=0.25,
                                         scheme=dict(order=2, rk_stages=3, cfl=0.2))
        case.precision = prec
        gpu = GpuSim(case, st)
        gpu.step(warmup)
        gpu.sync()
        t0 = time.perf_counter()
        gpu.step(steps)
        gpu.sync()
        el = time.perf_counter() - t0
        res[name] = {"steps": steps, "ms_per_step": el / steps * 1e3,
                     "cell_steps_per_s": nx * ny * steps / el, "particles_end": gpu.count()}
        gpu.close()
    L = olib()
    L.oracle_set_threads.argtypes = [C.c_int]
    threads = worker_threads()
    L.oracle_set_threads(threads)
    case, st = cases.triple_point_2d(nx=nx, ny=ny, ppc_axis=3, jitter=0.25,
                                     scheme=dict(order=2, rk_stages=3, cfl=0.2))
    sim = OracleSim(case, st)
    sim.step(1)
    t0 = time.perf_counter()
    sim.step(cpu_steps)
    el = time.perf_counter() - t0
    L.oracle_set_threads(1)
    res["cpu_oracle"] = {"cell_steps_per_s": nx * ny * cpu_steps / el, "threads": threads,
                         "steps": f"2..{cpu_steps + 1}"}
    res["mixed_speedup"] = res["fp64"]["ms_per_step"] / res["mixed"]["ms_per_step"]
    return res


def error_horizon(size=(64, 32, 32), steps=200, every=50, scheme=None, perturb=1e-10):
    """FP64 and mixed forms of the shock-bubble step side by side from the
    same initial state for `steps` steps on a reduced grid: relative L1 and
    Linf of the cell density (alpha-weighted over phases) every `every`
    steps, and the particle counts.  A third run, the FP64 form from the
    same state with every particle's density, energy and mass scaled by
    1 + perturb * U(-1, 1) (seeded), shows how far the method itself
    amplifies a perturbation of that size over the same steps."""
    import numpy as np
    from pgdg_p2_amd import cases
    from pgdg_p2_amd.step import GpuSim
    nx, ny, nz = size
    sims = []
    for prec, pert in ((0, 0.0), (1, 0.0), (0, perturb)):
        case, st = cases.shock_bubble_3d(nx=nx, ny=ny, nz=nz, ppc_axis=2, jitter=0.25,
                                         scheme=scheme)
        case.precision = prec
        if pert > 0:
            n = st.n
            f = 1.0 + pert * np.random.Generator(np.random.PCG64(0x5EED00A0)).uniform(-1, 1, n)
            for k in ("rho", "E", "mass"):
                st.p[k][:n] *= f
        sims.append((GpuSim(case, st), st.copy()))
    rows = []
    t0 = time.perf_counter()
    for k in range(every, steps + 1, every):
        rho = []
        for g, st in sims:
            g.step(every)
            g.save_cells(st)
            rho.append((st.c["alph"] * st.c["rho"]).sum(axis=0))
        d = np.abs(rho[1] - rho[0])
        dp = np.abs(rho[2] - rho[0])
        rows.append({"step": k, "density_L1_rel": float(d.sum() / np.abs(rho[0]).sum()),
                     "density_Linf_rel": float(d.max() / np.abs(rho[0]).max()),
                     "fp64_perturbed_L1_rel": float(dp.sum() / np.abs(rho[0]).sum()),
                     "fp64_perturbed_Linf_rel": float(dp.max() / np.abs(rho[0]).max()),
                     "particles": [g.count() for g, _ in sims]})
        print(f"[horizon] step {k}: L1 {rows[-1]['density_L1_rel']:.3g}", file=sys.stderr, flush=True)
    for g, _ in sims:
        g.close()
    last_ok = max([r["step"] for r in rows if r["density_L1_rel"] <= 1e-6], default=0)
    return {"grid": f"{nx}x{ny}x{nz}", "steps": steps, "rows": rows, "perturbation": perturb,
            "last_row_with_L1_le_1e-6": last_ok,
            "wall_s": time.perf_counter() - t0,
            "L1_le_1e-6_at_every_row": all(r["density_L1_rel"] <= 1e-6 for r in rows)}


# MFMA flops of one network pass on one face (padded operands: 16 inputs,
# three hidden layers of 32 units, 16 outputs)
MLP_FLOPS_PER_FACE = 2.0 * (16 * 32 + 32 * 32 + 32 * 32 + 32 * 16)


def mgfm_bench(n=10_000_000, n_train=2_000_000, epochs=12, mlp_save=None, mlp_load=None):
    """Batched MGFM on n stiffened-gas interface face pairs resident on the
    device: faces/s and Newton iterations per face for the reference
    tolerance and for 1e-12 relative on p*, from the reference's initial
    guess and from the learned guess (MLP trained on n_train solver-labelled
    faces, seed 0x5EED0004; training time reported): network and Newton
    solve in one kernel ("learned") and as two kernels with the guesses
    passing through memory ("learned_two_pass").  mlp_save: write the
    trained network (parameters and normalisation, .npz); mlp_load: use such
    a file instead of training (for counter runs that leave training out)."""
    import numpy as np
    import pgdg_p2_amd as P
    from pgdg_p2_amd import workloads
    from pgdg_p2_amd.mlp import Mlp
    eos = workloads.MGFM_EOS
    gl, pil, gr, pir = eos
    res = {"faces": n, "train_faces": n_train, "epochs": epochs}
    mlp = Mlp()
    if mlp_load:
        from pgdg_p2_amd.mlp import MlpNorm
        z = np.load(mlp_load)
        nm = MlpNorm()
        for f in ("feat_mu", "feat_sd", "out_mu", "out_sd"):
            getattr(nm, f)[:] = z[f].tolist()
        nm.keep = float(z["keep"]) if "keep" in z else 0.0
        mlp.set_params(z["params"], nm)
        res["trained_network"] = str(mlp_load)
    else:
        d_tr = P.DeviceArray.from_numpy(workloads.mgfm_faces(n_train, seed=0x5EED0004))
        P.device_sync()
        t0 = time.perf_counter()
        hist = mlp.train(d_tr, n_train, eos, epochs=epochs, batch=4096, lr=3e-3)
        res["train_s"] = time.perf_counter() - t0
        res["train_loss_first_last"] = [float(hist[:32].mean()), float(hist[-32:].mean())]
        d_tr.free()
        if mlp_save:
            Pm, nm = mlp.get_params()
            np.savez(mlp_save, params=Pm, feat_mu=np.array(nm.feat_mu), feat_sd=np.array(nm.feat_sd),
                     out_mu=np.array(nm.out_mu), out_sd=np.array(nm.out_sd), keep=np.float32(nm.keep))
    inp = P.DeviceArray.from_numpy(workloads.mgfm_faces(n))
    out = P.DeviceArray((3, n), np.float64)
    its = P.DeviceArray((n,), np.int32)
    tm = P.EventTimer()
    reps = 5
    p_ref = None
    g0 = P.DeviceArray((n,), np.float64)
    for name, mode, it in (("reference_tol", P.TOL_REFERENCE, 50), ("rel_1e-12", P.TOL_RELATIVE, 100)):
        for guess in ("plain", "learned", "learned_two_pass"):
            def run():
                if guess == "plain":
                    P.mgfm_batch_dev(inp, out, its, n, gl, pil, gr, pir, tol_mode=mode, rtol=1e-12,
                                     max_iter=it)
                elif guess == "learned":
                    mlp.mgfm_dev(inp, out, its, n, eos, tol_mode=mode, rtol=1e-12, max_iter=it)
                else:
                    mlp.guess_dev(inp, n, eos, g0)
                    P.mgfm_batch_dev(inp, out, its, n, gl, pil, gr, pir, tol_mode=mode, rtol=1e-12,
                                     max_iter=it, d_p0=g0)
            run()
            tm.start()
            for _ in range(reps):
                run()
            tm.stop()
            ms = tm.ms() / reps
            o, itv = out.numpy(), its.numpy()
            r = {"faces_per_s": n / (ms * 1e-3), "newton_its_per_face": float(itv.mean()), "ms": ms}
            if mode == P.TOL_RELATIVE:
                if guess == "plain":
                    p_ref = o[1].copy()
                else:
                    r["pstar_max_rel_diff_vs_plain"] = float(np.max(np.abs(o[1] - p_ref) /
                                                                    np.abs(p_ref)))
            res[f"{name}_{guess}"] = r
    # the network alone: one inference pass over the n faces
    tm.start()
    for _ in range(reps):
        mlp.guess_dev(inp, n, eos, g0)
    tm.stop()
    ms = tm.ms() / reps
    res["mlp_inference"] = {"ms": ms, "faces_per_s": n / (ms * 1e-3),
                            "bf16_mfma_TFLOPs": MLP_FLOPS_PER_FACE * n / (ms * 1e-3) / 1e12}
    for key in ("reference_tol", "rel_1e-12"):
        for v in ("learned", "learned_two_pass"):
            res[f"{key}_speedup_{v}"] = res[f"{key}_{v}"]["faces_per_s"] / res[f"{key}_plain"]["faces_per_s"]
    tm.close()
    for d in (inp, out, its, g0):
        d.free()
    mlp.close()
    return res


def mls_bench(n=128, reps=5):
    """MLS fit of 6 fields on n^3 cells (Poisson(8) particles per cell,
    inputs resident on the device): FP64 and mixed forms, with the mixed
    form's per-field deviation from the FP64 fit at the cell centres."""
    import numpy as np
    import pgdg_p2_amd as P
    from pgdg_p2_amd import mls, workloads
    t0 = time.perf_counter()
    d = workloads.mls_particles(n)
    npart = int(d["x"].shape[1])
    f = mls.MlsFit(d)
    print(f"mls setup {time.perf_counter() - t0:.1f}s: {n}^3 cells, {npart} particles",
          file=sys.stderr, flush=True)
    del d
    nc = n ** 3
    res = {"cells": nc, "particles": npart, "fields": f.nf, "modes": 10}
    tm = P.EventTimer()
    coefs = {}
    for name, prec, bpp in (("fp64", mls.PREC_FP64, 80), ("mixed", mls.PREC_MIXED, 40)):
        f.launch(prec)
        tm.start()
        for _ in range(reps):
            f.launch(prec)
        tm.stop()
        ms = tm.ms() / reps
        coefs[name] = f.coef.numpy()
        alg = npart * bpp + nc * (8 + 4 + f.nf * 10 * 8)
        res[name] = {"ms": ms, "cells_per_s": nc / (ms * 1e-3),
                     "particles_per_s": npart / (ms * 1e-3),
                     "alg_GBs": alg / (ms * 1e-3) / 1e9}
    tm.close()
    v64 = mls.center_values(coefs["fp64"])
    v32 = mls.center_values(coefs["mixed"])
    res["mixed_vs_fp64"] = {
        "centre_L1_rel": [float(np.abs(v32[i] - v64[i]).mean() / np.abs(v64[i]).mean())
                          for i in range(f.nf)],
        "centre_Linf_rel": [float(np.abs(v32[i] - v64[i]).max() / np.abs(v64[i]).max())
                            for i in range(f.nf)]}
    res["speedup_mixed"] = res["fp64"]["ms"] / res["mixed"]["ms"]
    f.close()
    return res


def _save_mlp(mlp, path):
    import numpy as np
    Pm, nm = mlp.get_params()
    np.savez(path, params=Pm, feat_mu=np.array(nm.feat_mu), feat_sd=np.array(nm.feat_sd),
             out_mu=np.array(nm.out_mu), out_sd=np.array(nm.out_sd), keep=np.float32(nm.keep))


def _load_mlp(path):
    import numpy as np
    from pgdg_p2_amd.mlp import Mlp, MlpNorm
    z = np.load(path)
    nm = MlpNorm()
    for f in ("feat_mu", "feat_sd", "out_mu", "out_sd"):
        getattr(nm, f)[:] = z[f].tolist()
    nm.keep = float(z["keep"]) if "keep" in z else 0.0
    mlp = Mlp()
    mlp.set_params(z["params"], nm)
    return mlp


def load_iface_guess(gpu, path):
    """Switch the step to the learned initial guess of a saved network."""
    from pgdg_p2_amd.step import GpuSim
    mlp = _load_mlp(path)
    gpu.iface_mode(GpuSim.IFACE_LEARNED, mlp)
    return {"trained_network": str(path), "mlp": mlp}


def train_iface_guess(gpu, save=None):
    """Train the network on the interface-point records of the stage just
    solved (the step's own exact MGFM solves, on the GPU) and switch the
    step to the learned initial guess."""
    import numpy as np
    import pgdg_p2_amd as P
    from pgdg_p2_amd.mlp import Mlp
    from pgdg_p2_amd.step import GpuSim
    rin, _, rpair, _ = gpu.iface_records()
    extra = 0
    if rin.shape[1] < 4096:  # the record buffer has just grown: one more step's records
        gpu.step(1)
        extra = 1
        rin, _, rpair, _ = gpu.iface_records()
    n = rin.shape[1] // 4096 * 4096
    res = {"train_records": n, "steps_taken": extra}
    if n < 4096:
        res["note"] = "too few interface records to train; the batch keeps the reference's initial p*"
        return res
    mlp = Mlp()
    d_in = P.DeviceArray.from_numpy(np.ascontiguousarray(rin[:, :n]))
    d_pair = P.DeviceArray.from_numpy(np.ascontiguousarray(rpair[:n]))
    P.device_sync()
    t0 = time.perf_counter()
    hist = mlp.train_pairs(d_in, d_pair, n, gpu.iface_pairs(), epochs=8, batch=4096, lr=3e-3)
    res["train_s"] = time.perf_counter() - t0
    res["train_loss_first_last"] = [float(hist[:8].mean()), float(hist[-8:].mean())]
    _, nm = mlp.get_params()
    res["keep_below_abs_ln_correction"] = float(nm.keep)
    if save:
        _save_mlp(mlp, save)
    d_in.free()
    d_pair.free()
    gpu.iface_mode(GpuSim.IFACE_LEARNED, mlp)
    res["mlp"] = mlp
    return res


def iface_iterations(gpu, mlp):
    """Newton iterations per interface point on the last stage's records:
    the step's own solves against the same inputs solved from the
    reference's initial p* (and, to 1e-12 relative, from both starts)."""
    import numpy as np
    import pgdg_p2_amd as P
    from pgdg_p2_amd.mlp import mgfm_pairs_dev
    rin, _, rpair, rit = gpu.iface_records()
    n = rin.shape[1]
    res = {"points_last_stage": n}
    if n == 0:
        return res
    d_x = P.DeviceArray.from_numpy(np.ascontiguousarray(rin))
    d_c = P.DeviceArray.from_numpy(np.ascontiguousarray(rpair))
    d_o = P.DeviceArray((3, n), np.float64)
    d_it = P.DeviceArray((n,), np.int32)
    eos = gpu.iface_pairs()
    its = {"step": float(rit.mean())}
    for who, m in (("reference_guess", None), ("learned_guess", mlp)):
        if who == "learned_guess" and m is None:
            continue
        for tol, (mode, rtol, mx) in (("reference_tol", (P.TOL_REFERENCE, 0.0, 50)),
                                      ("rel_1e-12", (P.TOL_RELATIVE, 1e-12, 100))):
            mgfm_pairs_dev(m, d_x, d_c, d_o, d_it, n, eos, tol_mode=mode, rtol=rtol, max_iter=mx)
            P.device_sync()
            its[f"{who}_{tol}"] = float(d_it.numpy().astype(np.float64).mean())
    for d in (d_x, d_c, d_o, d_it):
        d.free()
    res["newton_its_per_point"] = its
    return res


def late_window_bench(size=(256, 128, 128), before=190, steps=10, scheme=None, iface="learned"):
    """The headline step (mixed form, the same interface mode) late in the
    200-step problem: `before` steps run untimed from the initial state
    (the network of the learned interface guess trained on the first
    step's records, as in the headline run), then `steps` steps timed one
    by one, each with its split, merge, inlet and outlet counts (the
    headline's window, right after the warm-up steps, has none of the
    splits and merges of the later run)."""
    import numpy as np
    from pgdg_p2_amd import cases
    from pgdg_p2_amd.step import GpuSim
    nx, ny, nz = size
    case, st = cases.shock_bubble_3d(nx=nx, ny=ny, nz=nz, ppc_axis=2, jitter=0.25, cap_factor=1.0,
                                     scheme=scheme)
    case.precision = 1
    n0 = st.n
    gpu = GpuSim(case, st, cap=int(n0 * 2.2) + 4096)
    t_start = time.perf_counter()
    taken, mlp = 0, None
    if iface != "inline":
        gpu.iface_mode(GpuSim.IFACE_BATCH)
        gpu.step(1)
        taken = 1
        if iface == "learned":
            tr = train_iface_guess(gpu)
            taken += tr.get("steps_taken", 0)
            mlp = tr.get("mlp")
    names = ("split", "merge", "inlet", "outlet")
    ev_all = dict.fromkeys(names, 0)

    def add_events(acc):
        sc = gpu.scalars()
        for k, v in zip(names, (sc.n_split, sc.n_merge, sc.n_inlet, sc.n_outlet)):
            acc[k] += int(v)

    done = taken
    while done < before:  # one step at a time, for the event counts
        gpu.step(1)
        add_events(ev_all)
        done += 1
        if done % 50 == 0:
            print(f"[late] step {done}: {gpu.count()} particles", file=sys.stderr, flush=True)
    n_before = gpu.count()
    ms, ev = [], dict.fromkeys(names, 0)
    for _ in range(steps):
        gpu.sync()
        t0 = time.perf_counter()
        gpu.step(1)
        gpu.sync()
        ms.append((time.perf_counter() - t0) * 1e3)
        add_events(ev)
    counts = gpu.cell_particle_counts()
    n_after = gpu.count()
    gpu.close()
    if mlp is not None:
        mlp.close()
    nc = nx * ny * nz
    ms_step = float(np.mean(ms))
    return {"grid": f"{nx}x{ny}x{nz}", "interface_mgfm": iface, "timed_steps": [before + 1, before + steps],
            "ms_per_step": ms_step, "ms_per_step_min_max": [float(min(ms)), float(max(ms))],
            "cell_steps_per_s": nc / (ms_step * 1e-3),
            "particles_start_end": [n0, n_before, n_after],
            "events_in_window": ev, "events_before_window": ev_all,
            "particles_per_cell_min_max": [int(counts.min()), int(counts.max())],
            "wall_s": time.perf_counter() - t_start}


def learned_step_bench(size=(128, 64, 64), x_shock=0.3, steps=10, scheme=None):
    """The learned-guess step where its interface points meet strong jumps:
    the shock-bubble channel with the shock starting inside the helium
    sphere's span, P2 / SSP-RK3 / bounds [4,16].  Three runs from the same
    state: FP64, mixed with MGFM solved in the face kernel, and mixed with
    the stage's interface points solved as one batch from the network's
    initial p* (trained on the GPU on the first warm-up step's own solves).
    Reports ms/step, the cell density deviations and the Newton iterations
    per interface point."""
    import numpy as np
    from pgdg_p2_amd import cases
    from pgdg_p2_amd.step import GpuSim
    nx, ny, nz = size
    res = {"workload": f"3-D shock-bubble, {nx}x{ny}x{nz} P2 cells, 8 particles/cell, SSP-RK3, bounds [4,16], "
                       f"shock starting at x = {x_shock} (inside the bubble's span)", "steps": steps}
    rho = {}
    for name in ("fp64", "mixed_inline", "mixed_learned"):
        case, st = cases.shock_bubble_3d(nx=nx, ny=ny, nz=nz, x_shock=x_shock, cap_factor=1.0, scheme=scheme)
        case.precision = 0 if name == "fp64" else 1
        gpu = GpuSim(case, st, cap=int(st.n * 2.2) + 4096)
        mlp, info = None, {}
        if name == "mixed_learned":
            # the first batched step sizes the record buffer (its records are
            # dropped when it grows); the second's are the training set
            gpu.iface_mode(GpuSim.IFACE_BATCH)
            gpu.step(1)
            info = train_iface_guess(gpu)
            mlp = info.pop("mlp", None)
            gpu.step(1 - info.pop("steps_taken"))  # two warm-up steps in every run
        else:
            gpu.step(2)
        gpu.sync()
        t0 = time.perf_counter()
        gpu.step(steps)
        gpu.sync()
        el = time.perf_counter() - t0
        r = {"ms_per_step": el / steps * 1e3, "cell_steps_per_s": nx * ny * nz * steps / el,
             "particles_end": gpu.count()}
        if name == "mixed_learned":
            info.update(iface_iterations(gpu, mlp))
            r.update(info)
        gpu.save_cells(st)
        rho[name] = (st.c["alph"] * st.c["rho"]).sum(axis=0)
        gpu.close()
        if mlp is not None:
            mlp.close()
        res[name] = r
        print(f"[learned-step] {name}: {r['ms_per_step']:.2f} ms/step", file=sys.stderr, flush=True)
    for a in ("mixed_inline", "mixed_learned"):
        d = np.abs(rho[a] - rho["fp64"])
        res[a]["density_vs_fp64"] = {"L1_rel": float(d.sum() / np.abs(rho["fp64"]).sum()),
                                     "Linf_rel": float(d.max() / np.abs(rho["fp64"]).max())}
        res[a]["speedup_vs_fp64"] = res["fp64"]["ms_per_step"] / res[a]["ms_per_step"]
    d = np.abs(rho["mixed_learned"] - rho["mixed_inline"])
    res["mixed_learned"]["density_vs_mixed_inline"] = {
        "L1_rel": float(d.sum() / np.abs(rho["mixed_inline"]).sum()),
        "Linf_rel": float(d.max() / np.abs(rho["mixed_inline"]).max())}
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--size", default="full", choices=list(SIZES))
    ap.add_argument("--scheme", default="p2rk3", choices=list(SCHEMES))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mgfm", action="store_true")
    ap.add_argument("--no-mls", action="store_true")
    ap.add_argument("--no-3phase", action="store_true")
    ap.add_argument("--no-tp", action="store_true", help="skip the 112x48 triple-point leg")
    ap.add_argument("--no-horizon", action="store_true", help="skip the 200-step error horizon")
    ap.add_argument("--no-late", action="store_true", help="skip the timed window late in the 200-step run")
    ap.add_argument("--only-late", action="store_true")
    ap.add_argument("--3phase-warmup", dest="tp3_warmup", type=int, default=500)
    ap.add_argument("--only-3phase", action="store_true")
    ap.add_argument("--only-horizon", action="store_true")
    ap.add_argument("--mls-n", type=int, default=128)
    ap.add_argument("--only-mls", action="store_true")
    ap.add_argument("--only-mgfm", action="store_true")
    ap.add_argument("--no-learned-step", action="store_true",
                    help="skip the learned-guess step leg (128x64x64, shock inside the bubble's span)")
    ap.add_argument("--only-learned-step", action="store_true")
    ap.add_argument("--mlp-save", default=None, help="MGFM leg: write the trained network (.npz)")
    ap.add_argument("--mlp-load", default=None, help="MGFM leg: use a saved network instead of training")
    ap.add_argument("--rhs-mfma", action="store_true",
                    help="mixed form: the right-hand side on the matrix cores (k_rhs_mfma)")
    ap.add_argument("--iface", default="learned", choices=["inline", "batch", "learned"],
                    help="mixed form: MGFM at the interface face points solved in the face kernel, "
                         "as one batch per stage, or as one batch from the network's initial p* "
                         "(trained on the first warm-up step's own solves)")
    ap.add_argument("--iface-mlp-save", default=None,
                    help="learned interface guess: write the network trained on the step's records (.npz)")
    ap.add_argument("--iface-mlp-load", default=None,
                    help="learned interface guess: use a saved network instead of training in this run "
                         "(counter runs keep the training's thousands of launches out of the profiler)")
    ap.add_argument("--profile-steps", type=int, default=2)
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="also run one plain copy kernel over this many MiB (byte-counter "
                         "calibration for rocprofv3 --pmc runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU "
              "slab decomposition is a design, not a build", file=sys.stderr)
        sys.exit(2)

    if args.calib_copy_mib > 0:
        import ctypes as C
        import numpy as np
        import pgdg_p2_amd as P
        nbytes = args.calib_copy_mib << 20
        a = P.DeviceArray((nbytes,), np.uint8)
        b = P.DeviceArray((nbytes,), np.uint8)
        P.check(P.lib().pgdg_copy_kernel(b.ptr, a.ptr, nbytes, None), "pgdg_copy_kernel")
        P.device_sync()
        a.free()
        b.free()
        print(f"calibration copy: {nbytes} bytes read, {nbytes} bytes written", file=sys.stderr,
              flush=True)
    if args.profile_steps < 1:
        ap.error("--profile-steps must be at least 1 (the per-stage times and the roofline come from them)")
    if args.only_mls:
        print(json.dumps({"mls": mls_bench(args.mls_n)}))
        return
    if args.only_mgfm:
        print(json.dumps({"mgfm": mgfm_bench(mlp_save=args.mlp_save, mlp_load=args.mlp_load)}))
        return
    if args.only_late:
        print(json.dumps({"late_window": late_window_bench(size=SIZES[args.size], scheme=SCHEMES[args.scheme],
                                                           iface=args.iface)}))
        return
    if args.only_horizon:
        print(json.dumps({"error_horizon": [error_horizon(size=sz, scheme=SCHEMES[args.scheme])
                                            for sz in ((64, 32, 32), (128, 64, 64))]}))
        return
    if args.only_learned_step:
        print(json.dumps({"learned_guess_step": learned_step_bench(scheme=SCHEMES[args.scheme])}))
        return
    if args.only_3phase:
        print(json.dumps({"three_phase_2d": three_phase_bench(warmup=args.tp3_warmup)}))
        return

    from pgdg_p2_amd import cases
    from pgdg_p2_amd.step import GpuSim

    nx, ny, nz = SIZES[args.size]
    nc = nx * ny * nz
    walls = (nx + 1) * (ny + 1) * (nz + 1) - (nx - 1) * (ny - 1) * (nz - 1)

    def run_form(precision: int, iface_mode: str = "inline", profile: bool = True):
        t_setup = time.perf_counter()
        case, st = cases.shock_bubble_3d(nx=nx, ny=ny, nz=nz, ppc_axis=2, jitter=0.25,
                                         cap_factor=1.0, scheme=SCHEMES[args.scheme])
        case.precision = precision
        n0 = st.n
        gpu = GpuSim(case, st, cap=int(n0 * 2.2) + 4096)
        if precision and args.rhs_mfma:
            gpu.rhs_mfma(True)
        st.p = {}  # keep only the cell arrays, for the density comparison
        name = "mixed" if precision else "fp64"
        print(f"[{name}] setup {time.perf_counter() - t_setup:.1f}s: {nx}x{ny}x{nz} cells, "
              f"{n0} particles", file=sys.stderr, flush=True)
        iface = None
        if precision and iface_mode != "inline":
            gpu.iface_mode(GpuSim.IFACE_BATCH)
            gpu.step(1)
            iface = {"mode": iface_mode}
            taken = 1
            if iface_mode == "learned" and args.iface_mlp_load:
                iface.update(load_iface_guess(gpu, args.iface_mlp_load))
            elif iface_mode == "learned":
                iface.update(train_iface_guess(gpu, save=args.iface_mlp_save))
                taken += iface.pop("steps_taken")
            gpu.step(max(args.warmup - taken, 0))
        else:
            gpu.step(args.warmup)
        gpu.sync()
        n_start = gpu.count()
        t0 = time.perf_counter()
        gpu.step(args.steps)
        gpu.sync()
        el = time.perf_counter() - t0
        n_end = gpu.count()
        print(f"[{name}] timed {args.steps} steps: {el / args.steps * 1e3:.2f} ms/step",
              file=sys.stderr, flush=True)
        if not profile:
            gpu.close()
            return {"el": el, "n_end": n_end}
        # per-stage breakdown on extra steps (HIP events on the context stream)
        gpu.profile(True)
        gpu.step(args.profile_steps)
        gpu.sync()
        prof = gpu.profile_read()
        gpu.profile(False)
        mlp = None
        if iface is not None:
            mlp = iface.pop("mlp", None)  # the context reads its parameters until closed
            iface.update(iface_iterations(gpu, mlp))
        stages = {k: {"ms_total": v[0], "launches": v[1]} for k, v in prof.items() if v[1] > 0}
        n_mid = gpu.count()
        tuning = gpu.tuning()
        gpu.save_cells(st)
        rho = (st.c["alph"] * st.c["rho"]).sum(axis=0)
        gpu.close()
        if mlp is not None:
            mlp.close()
        return {"case": case, "el": el, "n_start": n_start, "n_end": n_end, "n_mid": n_mid,
                "stages": stages, "rho": rho, "tuning": tuning, "iface": iface}

    plain = run_form(0)
    mixed = run_form(1, args.iface)
    inline = run_form(1, "inline", profile=False) if args.iface != "inline" else None
    import numpy as np
    drho = np.abs(mixed["rho"] - plain["rho"])
    l1 = float(drho.sum() / np.abs(plain["rho"]).sum())
    linf = float(drho.max() / np.abs(plain["rho"]).max())

    R = mixed
    case, el, stages = R["case"], R["el"], R["stages"]
    ms_step = el / args.steps * 1e3
    value = nc * args.steps / el
    dom = max(stages, key=lambda k: stages[k]["ms_total"])
    per_launch_ms = stages[dom]["ms_total"] / stages[dom]["launches"]
    alg_full = stage_bytes(dom, case, R["n_mid"], walls)
    # algorithmic bytes by the per-unit figure of the particle-shifting
    # stage: 100 B per particle-stage in 3-D (x, v, vol, rho, c, phase read;
    # tv written), times the particles of the launch
    alg = sec8d_pst_bytes(case.dim) * R["n_mid"] if dom == "pst" else alg_full
    achieved = alg / (per_launch_ms * 1e-3) / 1e9 if alg > 0 else None
    order = case.order
    tb = R["tuning"] or {}
    zl = "true" if nz <= 128 else "false"  # one record frame per cell line (lines of up to 128 cells)
    kname = {"pst": f"pgdg::k_pst_mixed<3, {tb.get('pst', 256)}, {zl}>",
             "rhs": f"pgdg::k_rhs<3, {order}, {tb.get('rhs', 64)}, float>",
             "p2g": f"pgdg::k_p2g<3, {order}>", "g2p": f"pgdg::k_g2p<3, {order}>"}.get(dom, "")
    traffic = profiled_traffic(kname, args.size, args.scheme)
    S = 3 if case.rk_stages == 3 else 2
    ppc_avg = (R["n_start"] + R["n_end"]) / 2 / nc
    b_step = sec8d_cell_step_bytes(case.dim, case.U, case.M, case.nphase, S, ppc_avg)
    f_step = sec8d_mfma_flops(case.dim, case.U, case.M, S)
    roof = {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
            "traffic": traffic, "alg_bytes_per_launch": alg,
            "alg_bytes_per_unit": "100 B per particle-stage (particle shifting)" if dom == "pst" else None,
            "alg_bytes_per_launch_with_stage_inputs": alg_full,
            "avg_launch_ms": per_launch_ms,
            "step": {"bytes_per_cell_step": b_step, "mfma_flops_per_cell_step": f_step,
                     "achieved_GBs": b_step * value / 1e9,
                     "frac": b_step * value / 1e9 / HBM_PEAK_GBS,
                     "fp64_form_frac": b_step * nc * args.steps / plain["el"] / 1e9 / HBM_PEAK_GBS}}
    if traffic is not None:
        roof["traffic_source"] = str(TRAFFIC_PROFILES[args.scheme].relative_to(ROOT))
        roof["traffic_kernel"] = kname

    line = {
        "metric": "P2 DG cell-steps/s + MGFM Newton its/face, mixed vs FP64 plain at equal L1 err",
        "metric_note": ("value: cell-steps/s of the mixed form with its interface points' MGFM "
                        + ("solved in the face kernel; " if args.iface == "inline" else
                           "solved as one batch per stage from the "
                           + ("learned" if args.iface == "learned" else "reference's") + " initial p* (")
                        + "interface_mgfm: "
                        "the network, trained on the first warm-up step's own solves, and the Newton "
                        "iterations per point; mixed_in_kernel_mgfm: the same form with MGFM solved "
                        "in the face kernel); its density deviation from the FP64 form is in "
                        "mixed_vs_plain (this run) and error_horizon (200 steps); Newton iterations "
                        "per face of the stiffened-gas batch are in mgfm"),
        "value": value,
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("f64 state, MGFM and mass division; f32 DG quadrature and fluxes, "
                  "f32 particle-shifting pair terms (f64 sums)"),
        "data": "synthetic",
        "config": {"workload": f"3-D shock-bubble step, {nx}x{ny}x{nz} cells, 8 particles/cell, " + (
                               "P2 DG (10 modes, 3-point Gauss), SSP-RK3, split/merge with per-cell "
                               "bounds [4,16]" if args.scheme == "p2rk3" else
                               "P1 DG (reference basis), reference two-stage step with split/merge")
                               + ", mixed form",
                   "scheme": args.scheme,
                   "cells": nc, "particles_start": R["n_start"], "particles_end": R["n_end"],
                   "parallelism": "single GPU"},
        "particles_per_s": (R["n_start"] + R["n_end"]) / 2 * args.steps / el,
        "stages_ms": {k: round(v["ms_total"] / args.profile_steps, 4) for k, v in stages.items()},
        "plain_fp64": {
            "value": nc * args.steps / plain["el"], "ms_per_step": plain["el"] / args.steps * 1e3,
            "particles_end": plain["n_end"],
            "stages_ms": {k: round(v["ms_total"] / args.profile_steps, 4)
                          for k, v in plain["stages"].items()}},
        "mixed_vs_plain": {
            "speedup": plain["el"] / el,
            "density_L1_rel": l1, "density_Linf_rel": linf,
            "after_steps": args.warmup + args.steps + args.profile_steps,
            "same_particle_count": plain["n_end"] == R["n_end"]},
        "roofline": roof,
        "autotuned_workgroups": {"mixed": R["tuning"], "fp64": plain["tuning"]},
        "interface_mgfm": R["iface"] or {"mode": "inline"},
        "mixed_in_kernel_mgfm": ({"ms_per_step": inline["el"] / args.steps * 1e3,
                                  "value": nc * args.steps / inline["el"], "particles_end": inline["n_end"]}
                                 if inline else None),
        "stage_graphs": "RK stages replayed as HIP graphs in the timed steps",
    }
    if not args.no_late:
        line["late_window"] = late_window_bench(size=SIZES[args.size], scheme=SCHEMES[args.scheme],
                                                iface=args.iface)
    if not args.no_horizon:
        line["error_horizon"] = [error_horizon(size=sz, scheme=SCHEMES[args.scheme])
                                 for sz in ((64, 32, 32), (128, 64, 64))]
    if not args.no_learned_step:
        line["learned_guess_step"] = learned_step_bench(scheme=SCHEMES[args.scheme])
    if not args.no_tp:
        line["triple_point_112x48"] = triple_point_bench()
    if not args.no_mgfm:
        line["mgfm"] = mgfm_bench()
    if not args.no_3phase:
        line["three_phase_2d"] = three_phase_bench(warmup=args.tp3_warmup)
    if not args.no_mls:
        line["mls"] = mls_bench(args.mls_n)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(SCHEMES[args.scheme])
        line["cpu_baseline"]["host_cpus"] = os.cpu_count()
        line["cpu_baseline"]["gpu_vs_cpu"] = value / line["cpu_baseline"]["value"]
    print(json.dumps(line))


if __name__ == "__main__":
    main()
