"""Benchmark of the PG-DG time step on one MI355X.

Workload (default): the 3-D shock-bubble step, channel 256x128x128 P2 DG cells
(10 modes, 3-point Gauss rules) with 8 particles per cell (2x2x2 sub-lattice,
seeded jitter), air with a helium sphere hit by a Mach 1.22 shock; SSP-RK3
stages, split/merge with per-cell bounds [4, 16].  Plain FP64 and mixed forms
are both run; `value` is the mixed form's whole-job cell-steps/s with the state
resident in HBM.  `--scheme p1` runs the reference's P1 basis and two-stage
step instead.  Secondary measurements: the batched MGFM solve on 1e7
interface face pairs (plain and learned initial guess) and the MLS fit.

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
    """The oracle (scalar C++ restatement of the reference step) on a reduced
    grid of the same workload, timed on this host."""
    from oracle.pyoracle_step import OracleSim
    from pgdg_p2_amd import cases
    case, st = cases.shock_bubble_3d(nx=sample[0], ny=sample[1], nz=sample[2], ppc_axis=2,
                                     jitter=0.25, scheme=scheme)
    sim = OracleSim(case, st)
    sim.step(1)  # first step (particle dt path) excluded
    t0 = time.perf_counter()
    sim.step(steps)
    el = time.perf_counter() - t0
    nc = sample[0] * sample[1] * sample[2]
    return {"value": nc * steps / el, "unit": "cell-steps/s", "cores": 1, "kind": "port",
            "sample": f"{sample[0]}x{sample[1]}x{sample[2]} cells x 8 ppc ({st.n} particles), "
                      f"steps 2..{steps + 1}, same initial condition, oracle FP64 single thread"}


def three_phase_bench(steps=20, warmup=2, nx=1792, ny=768):
    """The 2-D three-phase triple point at 1792x768 P2 cells (gamma 1.5 / 1.4 /
    1.67), the reference's one-particle-per-cell lattice, SSP-RK3, FP64 form:
    cell-steps/s and split + merge events per second over `steps` steps."""
    from pgdg_p2_amd import cases
    from pgdg_p2_amd.step import GpuSim
    t0 = time.perf_counter()
    case, st = cases.triple_point_2d(nx=nx, ny=ny, ppc_axis=1, jitter=0.0, nphase=3,
                                     scheme=dict(order=2, rk_stages=3))
    gpu = GpuSim(case, st)
    print(f"[3-phase] setup {time.perf_counter() - t0:.1f}s: {nx}x{ny} cells, {st.n} particles",
          file=sys.stderr, flush=True)
    gpu.step(warmup)
    gpu.sync()
    n0 = gpu.count()
    ev = 0
    t = time.perf_counter()
    for _ in range(steps):
        gpu.step(1)
        sc = gpu.scalars()
        ev += sc.n_split + sc.n_merge
    gpu.sync()
    el = time.perf_counter() - t
    n1 = gpu.count()
    gpu.close()
    return {"cells": nx * ny, "particles_start": n0, "particles_end": n1, "steps": steps,
            "ms_per_step": el / steps * 1e3, "cell_steps_per_s": nx * ny * steps / el,
            "split_merge_events": ev, "split_merge_events_per_s": ev / el,
            "scheme": "P2, SSP-RK3, FP64 form, 1 particle per cell lattice, 3 phases"}


def mgfm_bench(n=10_000_000, n_train=2_000_000, epochs=12):
    """Batched MGFM on n stiffened-gas interface face pairs resident on the
    device: faces/s and Newton iterations per face for the reference
    tolerance and for 1e-12 relative on p*, from the reference's initial
    guess and from the learned guess (MLP trained on n_train solver-labelled
    faces, seed 0x5EED0004; training time reported)."""
    import numpy as np
    import pgdg_p2_amd as P
    from pgdg_p2_amd import workloads
    from pgdg_p2_amd.mlp import Mlp
    eos = workloads.MGFM_EOS
    gl, pil, gr, pir = eos
    res = {"faces": n, "train_faces": n_train, "epochs": epochs}
    d_tr = P.DeviceArray.from_numpy(workloads.mgfm_faces(n_train, seed=0x5EED0004))
    mlp = Mlp()
    P.device_sync()
    t0 = time.perf_counter()
    hist = mlp.train(d_tr, n_train, eos, epochs=epochs, batch=4096, lr=3e-3)
    res["train_s"] = time.perf_counter() - t0
    res["train_loss_first_last"] = [float(hist[:32].mean()), float(hist[-32:].mean())]
    d_tr.free()
    inp = P.DeviceArray.from_numpy(workloads.mgfm_faces(n))
    out = P.DeviceArray((3, n), np.float64)
    its = P.DeviceArray((n,), np.int32)
    tm = P.EventTimer()
    reps = 5
    p_ref = None
    for name, mode, it in (("reference_tol", P.TOL_REFERENCE, 50), ("rel_1e-12", P.TOL_RELATIVE, 100)):
        for guess in ("plain", "learned"):
            def run():
                if guess == "plain":
                    P.mgfm_batch_dev(inp, out, its, n, gl, pil, gr, pir, tol_mode=mode, rtol=1e-12,
                                     max_iter=it)
                else:
                    mlp.mgfm_dev(inp, out, its, n, eos, tol_mode=mode, rtol=1e-12, max_iter=it)
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
    g0 = P.DeviceArray((n,), np.float64)
    tm.start()
    for _ in range(reps):
        mlp.guess_dev(inp, n, eos, g0)
    tm.stop()
    ms = tm.ms() / reps
    res["mlp_inference"] = {"ms": ms, "faces_per_s": n / (ms * 1e-3),
                            "bf16_mfma_TFLOPs": 17920.0 * n / (ms * 1e-3) / 1e12}
    for key in ("reference_tol", "rel_1e-12"):
        res[f"{key}_speedup_learned"] = res[f"{key}_learned"]["faces_per_s"] / \
            res[f"{key}_plain"]["faces_per_s"]
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
    ap.add_argument("--mls-n", type=int, default=128)
    ap.add_argument("--only-mls", action="store_true")
    ap.add_argument("--only-mgfm", action="store_true")
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
    if args.only_mls:
        print(json.dumps({"mls": mls_bench(args.mls_n)}))
        return
    if args.only_mgfm:
        print(json.dumps({"mgfm": mgfm_bench()}))
        return

    from pgdg_p2_amd import cases
    from pgdg_p2_amd.step import GpuSim

    nx, ny, nz = SIZES[args.size]
    nc = nx * ny * nz
    walls = (nx + 1) * (ny + 1) * (nz + 1) - (nx - 1) * (ny - 1) * (nz - 1)

    def run_form(precision: int):
        t_setup = time.perf_counter()
        case, st = cases.shock_bubble_3d(nx=nx, ny=ny, nz=nz, ppc_axis=2, jitter=0.25,
                                         cap_factor=1.0, scheme=SCHEMES[args.scheme])
        case.precision = precision
        n0 = st.n
        gpu = GpuSim(case, st, cap=int(n0 * 2.2) + 4096)
        st.p = {}  # keep only the cell arrays, for the density comparison
        name = "mixed" if precision else "fp64"
        print(f"[{name}] setup {time.perf_counter() - t_setup:.1f}s: {nx}x{ny}x{nz} cells, "
              f"{n0} particles", file=sys.stderr, flush=True)
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
        # per-stage breakdown on extra steps (HIP events on the context stream)
        gpu.profile(True)
        gpu.step(args.profile_steps)
        gpu.sync()
        prof = gpu.profile_read()
        gpu.profile(False)
        stages = {k: {"ms_total": v[0], "launches": v[1]} for k, v in prof.items() if v[1] > 0}
        n_mid = gpu.count()
        tuning = gpu.tuning()
        gpu.save_cells(st)
        rho = (st.c["alph"] * st.c["rho"]).sum(axis=0)
        gpu.close()
        return {"case": case, "el": el, "n_start": n_start, "n_end": n_end, "n_mid": n_mid,
                "stages": stages, "rho": rho, "tuning": tuning}

    plain = run_form(0)
    mixed = run_form(1)
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
    alg = stage_bytes(dom, case, R["n_mid"], walls)
    achieved = alg / (per_launch_ms * 1e-3) / 1e9 if alg > 0 else None
    order = case.order
    tb = R["tuning"] or {}
    kname = {"pst": f"pgdg::k_pst_mixed<3, {tb.get('pst', 256)}>",
             "rhs": f"pgdg::k_rhs<3, {order}, {tb.get('rhs', 64)}, float>",
             "p2g": f"pgdg::k_p2g<3, {order}>", "g2p": f"pgdg::k_g2p<3, {order}>"}.get(dom, "")
    traffic = profiled_traffic(kname, args.size, args.scheme)
    roof = {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
            "traffic": traffic, "alg_bytes_per_launch": alg, "avg_launch_ms": per_launch_ms}
    if traffic is not None:
        roof["traffic_source"] = str(TRAFFIC_PROFILES[args.scheme].relative_to(ROOT))
        roof["traffic_kernel"] = kname

    line = {
        "metric": "P2 DG cell-steps/s + MGFM Newton its/face, mixed vs FP64 plain at equal L1 err",
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
        "stage_graphs": "RK stages replayed as HIP graphs in the timed steps",
    }
    if not args.no_mgfm:
        line["mgfm"] = mgfm_bench()
    if not args.no_3phase:
        line["three_phase_2d"] = three_phase_bench()
    if not args.no_mls:
        line["mls"] = mls_bench(args.mls_n)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(SCHEMES[args.scheme])
        line["cpu_baseline"]["host_cpus"] = os.cpu_count()
    print(json.dumps(line))


if __name__ == "__main__":
    main()
