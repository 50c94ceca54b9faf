#!/usr/bin/env python3
"""Benchmark of the SAC perturbation-MHD hot path on one MI355X.

Workload: the stratified 3-D perturbation on a static background at 256^3
(rho_b=exp(-z/H), isothermal, uniform vertical B with beta=1 at the base,
hydrostatic gravity on axis 1, Gaussian velocity pulse plus seeded noise,
fixed/periodic/periodic boundaries, 4th-order central differences + RK4 +
hyperdiffusion), FP64.  A step is one full RK4 time step (four stages, each
with its getboundary) including the per-step dt computation; inputs are
resident in HBM when the timed region starts.

Metric: Mcell-updates/s per RK stage = N_mesh * steps * 4 / wall time.
Roofline: the unit is one RK stage, which the library issues as a fixed
kernel sequence (k_plane for the x and y work, k_line for z, k_final,
k_ghost); algorithmic bytes =
296 B per mesh cell per stage (4*n_evolve*8 + n_background*8 for 8 evolving
and 5 background FP64 fields) x N_mesh, divided by the summed average
durations of that sequence, measured with HIP events on the library's
stream.
cpu_baseline: the C oracle (a scalar CPU restatement of the reference) on a
32^3 sample of the same workload, one copy per host core (at most 16) at
once; the one-core figure beside it.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# algorithmic bytes per mesh cell per RK stage: 4*n_evolve*s_e + n_background*8
BYTES_PER_CELL_STAGE = {"fp64": 4 * 8 * 8 + 5 * 8, "mixed": 4 * 8 * 4 + 5 * 8}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=256, help="mesh cells per axis")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=20.0,
                    help="CPU baseline time: half alone on one core, half on all cores")
    ap.add_argument("--cpu-cores", type=int, default=0,
                    help="processes of the multi-core CPU baseline (0: the host cores "
                         "available to this process, at most 16)")
    ap.add_argument("--no-mixed", action="store_true",
                    help="skip the mixed-precision leg (FP32 evolving fields)")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the tube-ensemble leg (4096 uneven 1-D tubes to t=0.1)")
    ap.add_argument("--tubes", type=int, default=4096)
    ap.add_argument("--no-patches", action="store_true",
                    help="skip the 2-D patch-ensemble leg (1024 uneven patches to t=0.1)")
    ap.add_argument("--patches", type=int, default=1024)
    ap.add_argument("--no-learned", action="store_true",
                    help="skip the learned-closure form (train + run at amplitude 0.1)")
    ap.add_argument("--learned-steps", type=int, default=12)
    ap.add_argument("--no-accuracy", action="store_true",
                    help="skip the 128^3 leg against the oracle fixture")
    ap.add_argument("--calib-copy", action="store_true",
                    help="also run the library's plain copy kernel on one register "
                         "(known bytes) for HBM-counter calibration")
    return ap.parse_args()


def _cpu_worker(n: int, budget_s: float):
    """RK4 steps of the C oracle on a stratified n^3 sample for about
    budget_s seconds (the first step untimed) -> (steps, seconds)"""
    from oracle import oracle as om
    from sac_amd import problems
    prob = problems.stratified(n=n)
    o = om.from_problem(prob)
    o.step(prob.courantpar)  # first step (no dt_visc history) untimed
    t0 = time.perf_counter()
    steps = 0
    while time.perf_counter() - t0 < budget_s:
        o.step(prob.courantpar)
        steps += 1
    return steps, time.perf_counter() - t0


def cpu_baseline(budget_s: float, cores: int):
    """Time the C oracle on a bounded sample of the same workload: one
    process alone, then `cores` processes at once, each on its own copy of
    the sample (how a serial code uses a multi-core host); the value is the
    latter's total throughput."""
    import multiprocessing as mp
    from concurrent.futures import ProcessPoolExecutor
    from oracle import oracle as om
    om.build()
    n = 32
    steps1, el1 = _cpu_worker(n, budget_s / 2)
    one = n ** 3 * steps1 * 4 / el1 / 1e6
    # fresh interpreters: this process has initialised the GPU
    with ProcessPoolExecutor(cores, mp_context=mp.get_context("spawn")) as ex:
        res = list(ex.map(_cpu_worker, [n] * cores, [budget_s / 2] * cores))
    allc = sum(n ** 3 * st * 4 / el / 1e6 for st, el in res)
    return {"value": round(allc, 4), "unit": "Mcell-updates/s", "cores": cores, "kind": "port",
            "sample": f"stratified {n}^3 RK4 steps after one warm-up step, C oracle (gcc -O2): "
                      f"{cores} processes at once, one sample each, {budget_s / 2:.0f} s "
                      f"({sum(r[0] for r in res)} steps in all); alone on one core "
                      f"{one:.4f} Mcell-updates/s ({steps1} steps in {el1:.1f} s)",
            "single_core_value": round(one, 4)}


def timed_run(device, problems, n, steps, warmup, precision, calib_copy=False):
    """Warm-up and timed RK4 steps of the stratified workload on one GPU.
    Returns (wall seconds, kernel timings, final state, copy calibration)."""
    import numpy as np
    prob = problems.stratified(n=n)
    dev = device.Device(prob, precision=precision)
    dev.upload(prob.w)
    dev.getboundary()
    del prob.w
    calib = None
    if calib_copy:
        dev.timing(True)
        for _ in range(3):
            dev.copy_slot(device.SLOT_S2, device.SLOT_W)
        dev.synchronize()
        c_ms, c_n = dev.timing_get(device.K_COPY)
        dev.timing(False)
        G = n + 4
        esz = 8 if precision == "fp64" else 4
        fs = (G + 15) // 16 * 16 * G * G  # padded field size (elements)
        nbytes = 2 * 8 * esz * fs          # 8 evolving fields read + written
        c_avg = c_ms / max(c_n, 1)
        calib = {"copy_avg_ms": round(c_avg, 4), "launches": c_n, "copy_bytes": nbytes,
                 "copy_GBs": round(nbytes / (c_avg / 1e3) / 1e9, 1)}
    for _ in range(warmup):
        dev.step(0.3)
    dev.synchronize()
    dev.timing(True)
    t0 = time.perf_counter()
    for _ in range(steps):
        dev.step(0.3)
    dev.synchronize()
    el = time.perf_counter() - t0
    tim = {k: dev.timing_get(getattr(device, k)) for k in
           ("K_FINAL", "K_AXIS", "K_GHOST", "K_AXIS1", "K_AXIS2", "K_AXIS3", "K_PLANE")}
    dev.timing(False)
    w = dev.download()
    dev.close()
    return el, tim, w, calib, bool(np.isfinite(w).all())


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: runs on one GPU only; --gpus {args.gpus} is not supported "
              "(the domain decomposition over several GPUs is a design, not built)",
              file=sys.stderr)
        return 2
    from sac_amd import device, problems
    from sac_amd.errors import per_var_errors

    n = args.n
    el, tim, w64, calib, finite = timed_run(device, problems, n, args.steps, args.warmup,
                                           "fp64", args.calib_copy)
    f_ms, f_n = tim["K_FINAL"]
    a_ms, a_n = tim["K_AXIS"]
    g_ms, g_n = tim["K_GHOST"]
    p_ms, p_n = tim["K_PLANE"]
    ncell = n ** 3
    stages = args.steps * 4
    value = ncell * stages / el / 1e6
    # kernel time of one stage: k_plane (x and y) + k_line (z) + k_final + k_ghost
    per_stage_ms = (p_ms + a_ms + f_ms + g_ms) / max(f_n, 1)
    alg_bytes = BYTES_PER_CELL_STAGE["fp64"] * ncell
    achieved = alg_bytes / (per_stage_ms / 1e3) / 1e9
    stage_wall_s = el / stages
    traffic = measured_traffic("fp64", n)
    out = {
        "metric": "Mcell-updates/s per RK stage at 256^3 (stratified perturbation MHD)",
        "value": round(value, 3),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(el / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"stratified 3-D perturbation MHD, {n}^3 mesh, FP64, RK4 + "
                               "hyperdiffusion + gravity, fixed/periodic/periodic",
                   "mesh": [n, n, n], "precision": "fp64", "typeadvance": "fourstep"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic,
                     "kernel": "RK stage = k_plane (x and y) + k_line<3,3,2> (z) + k_final<3,3> "
                               "+ k_ghost",
                     "avg_ms": round(per_stage_ms, 4),
                     "launches": f_n,
                     # the measured HBM traffic (rocprof, copy-calibrated) over the
                     # same time: how close the kernels run to the memory system
                     "traffic_rate_GBs": (round(traffic / 1e9 / (per_stage_ms / 1e3), 1)
                                          if traffic else None)},
        "stage_breakdown_ms": {"k_plane": round(p_ms / max(p_n, 1), 4),
                               "k_line_z": round(tim["K_AXIS3"][0] / max(tim["K_AXIS3"][1], 1), 4),
                               "k_final": round(f_ms / max(f_n, 1), 4),
                               "k_ghost": round(g_ms / max(g_n, 1), 4),
                               "stage_wall": round(stage_wall_s * 1e3, 4)},
        "stage_roofline_frac": round(alg_bytes / stage_wall_s / 1e9 / HBM_PEAK_GBS, 4),
        "finite": finite,
    }
    if calib is not None:
        out["copy_calibration"] = calib
    if not args.no_mixed:
        elm, timm, wm, _, finm = timed_run(device, problems, n, args.steps, args.warmup, "mixed")
        prob = problems.stratified(n=n)
        errs = per_var_errors(prob, wm, w64, 8)
        fm, fmn = timm["K_FINAL"]
        am, _ = timm["K_AXIS"]
        gm, _ = timm["K_GHOST"]
        pm, _ = timm["K_PLANE"]
        st_ms = (pm + am + fm + gm) / max(fmn, 1)
        ach_m = BYTES_PER_CELL_STAGE["mixed"] * ncell / (st_ms / 1e3) / 1e9
        out["forms"] = {"mixed": {
            "value": round(ncell * stages / elm / 1e6, 3), "unit": "Mcell-updates/s",
            "speedup_vs_fp64": round(el / elm, 3),
            "rel_l2_vs_fp64_gpu": [float("%.3e" % e) for e in errs],
            "after_steps": args.steps + args.warmup, "finite": finm,
            "roofline": {"achieved": round(ach_m, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(ach_m / HBM_PEAK_GBS, 4),
                         "traffic": measured_traffic("mixed", n),
                         "bytes_per_cell_stage": BYTES_PER_CELL_STAGE["mixed"]}}}
    if not args.no_batch:
        out.setdefault("forms", {})["batched_tubes"] = batched_leg(args.tubes)
    if not args.no_patches:
        out.setdefault("forms", {})["batched_patches_2d"] = patches_leg(args.patches)
    if not args.no_learned:
        out.setdefault("forms", {})["learned_closure"] = learned_leg(n, args.learned_steps)
    acc = None if args.no_accuracy else accuracy_leg()
    if acc is not None:
        out["accuracy_vs_oracle"] = acc
    if not args.no_cpu_baseline:
        cores = args.cpu_cores or min(16, len(os.sched_getaffinity(0)))
        out["cpu_baseline"] = cpu_baseline(args.cpu_seconds, cores)
    print(json.dumps(out))
    return 0


def batched_leg(count):
    """The tube ensemble: `count` 1-D MHD Riemann tubes of 200-2000 cells
    (seeded), each run to t=0.1 with its own dt; autotuned launch shape, then
    hipGraph replays of 32 steps until every tube is done.  Value =
    sum over tubes of cells x steps x 4 stages / wall time of the run.  The
    two shortest tubes are compared bit for bit with single-tube runs."""
    import numpy as np
    from sac_amd import batch, device
    tubes = batch.riemann_ensemble(count=count)
    ens = batch.Ensemble(tubes, courantpar=0.3, tmax=0.1)
    ck, tb, ms = ens.tune(trial_steps=4)
    t0 = time.perf_counter()
    issued = ens.run(steps_per_graph=32)
    el = time.perf_counter() - t0
    t, it, st = ens.status()
    ws = ens.download()
    ens.close()
    cells = np.array([p.nx[0] for p in tubes], dtype=np.int64)
    cu = float(np.sum(cells * it)) * 4
    checks = []
    for i in np.argsort(cells)[:2]:
        w1, dts = device.run(tubes[i], tmax=0.1)
        checks.append(bool(len(dts) == it[i] and np.array_equal(ws[i][:8], w1[:8])))
    alg = BYTES_PER_CELL_STAGE["fp64"] * cu  # 296 B per cell-stage update
    return {"value": round(cu / el / 1e6, 1), "unit": "Mcell-updates/s",
            "workload": f"{count} 1-D MHD Riemann tubes, 200-2000 cells (seed 4096), "
                        "gamma 5/3, cont boundaries, RK4 + hyperdiffusion, to t=0.1, FP64",
            "mesh_cells": int(cells.sum()), "wall_s": round(el, 3), "steps_issued": int(issued),
            "steps_per_tube": [int(it.min()), int(it.max())],
            "tuned_shape": {"cells_per_thread": ck, "threads_per_block": tb,
                            "trial_ms_per_step": round(ms, 4)},
            "reached_tmax": int((st == 1).sum()), "stopped_dt_below_dtmin": int((st == 3).sum()),
            "failed": int((st == 2).sum()),
            "wall_roofline_frac": round(alg / el / 1e9 / HBM_PEAK_GBS, 4),
            "bitwise_equal_to_single_tube_runs": checks}


def patches_leg(count):
    """The 2-D patch ensemble: `count` 2-D MHD Riemann patches of 32-160
    cells per axis (seeded), each run to t=0.1 with its own dt; autotuned
    tile shape, then hipGraph replays of 32 steps until every patch is done.
    Value = sum over patches of cells x steps x 4 stages / wall time of the
    run.  The two smallest patches are run again alone (an ensemble of one)
    and compared bit for bit."""
    import numpy as np
    from sac_amd import batch
    patches = batch.patch_ensemble(count=count)
    ens = batch.Ensemble(patches, courantpar=0.3, tmax=0.1)
    ens.tune(trial_steps=4)
    ck, tx, ty = ens.shape()
    t0 = time.perf_counter()
    issued = ens.run(steps_per_graph=32)
    el = time.perf_counter() - t0
    t, it, st = ens.status()
    ws = ens.download()
    ens.close()
    cells = np.array([p.nx[0] * p.nx[1] for p in patches], dtype=np.int64)
    cu = float(np.sum(cells * it)) * 4
    checks = []
    for i in np.argsort(cells)[:2]:
        one = batch.Ensemble([patches[i]], courantpar=0.3, tmax=0.1)
        one.run(steps_per_graph=32)
        _, it1, _ = one.status()
        w1 = one.download()[0]
        one.close()
        checks.append(bool(it1[0] == it[i] and np.array_equal(ws[i][:8], w1[:8])))
    alg = BYTES_PER_CELL_STAGE["fp64"] * cu
    return {"value": round(cu / el / 1e6, 1), "unit": "Mcell-updates/s",
            "workload": f"{count} 2-D MHD Riemann patches, 32-160 x 32-160 cells (seed 2048), "
                        "gamma 5/3, cont boundaries, RK4 + hyperdiffusion, to t=0.1, FP64",
            "mesh_cells": int(cells.sum()), "wall_s": round(el, 3), "steps_issued": int(issued),
            "steps_per_patch": [int(it.min()), int(it.max())],
            "tuned_tile": {"x": tx, "y": ty},
            "reached_tmax": int((st == 1).sum()), "stopped_dt_below_dtmin": int((st == 3).sum()),
            "failed": int((st == 2).sum()),
            "wall_roofline_frac": round(alg / el / 1e9 / HBM_PEAK_GBS, 4),
            "bitwise_equal_to_single_patch_runs": checks}


def learned_leg(n, steps):
    """The learned hyperdiffusion closure on the stratified setup at
    amplitude 0.1 (n^3).  The network is trained on the GPU (4000 Adam steps
    of 16384 samples) on ~1e8 (stencil, exact setnu ratio) pairs collected
    from runs of the plain FP64 solver on the other configurations: the
    Brio-Wu tube (800 cells, every 10th state of its run to t=0.1), the
    stratified atmosphere at 128^3 (every 25th of 200 steps) and at 256^3
    after 20 steps, both at amplitude 1e-3.  Then `steps` RK4 steps of the
    amplitude-0.1 run are taken with the plain method and with the closure
    on (shock-sensor threshold 0.1) under the plain run's dt sequence, both
    timed.  Reported: per-variable rel-L2 of the learned run against the
    plain FP64 run and the speed ratio of the two."""
    import numpy as np
    from sac_amd import closure, device, problems
    from sac_amd.errors import per_var_errors
    prob = problems.stratified(n=n, amp=0.1)
    ncell = n ** 3

    def fresh(p):
        d = device.Device(p)
        d.upload(p.w)
        d.getboundary()
        return d

    train_dev = device.Device(prob)
    cap = 800 * 200 + 8 * 3 * 128 ** 3 + 3 * 256 ** 3
    cl = closure.Closure(train_dev, capacity=cap, max_blocks=256, seed=1)
    sources = []
    d1 = fresh(problems.brio_wu(n=800))
    t, k = 0.0, 0
    while t < 0.1:
        t += d1.step(0.3)
        k += 1
        if k % 10 == 0:
            cl.collect_from(d1)
    d1.close()
    sources.append(f"Brio-Wu 800 cells, {k // 10} states")
    d2 = fresh(problems.stratified(n=128))
    for k in range(1, 201):
        d2.step(0.3)
        if k % 25 == 0:
            cl.collect_from(d2)
    d2.close()
    sources.append("stratified 128^3 amplitude 1e-3, 8 states over 200 steps")
    d3 = fresh(problems.stratified(n=256))
    for _ in range(20):
        d3.step(0.3)
    cl.collect_from(d3)
    d3.close()
    sources.append("stratified 256^3 amplitude 1e-3, the state after 20 steps")
    t0 = time.perf_counter()
    losses = [cl.train(1, lr=1e-3)]
    for k in range(8):
        losses.append(cl.train(500, lr=1e-3 if k < 5 else 3e-4, base=(k + 1) * 500 * 256 * 64))
    train_s = time.perf_counter() - t0
    samples = cl.n
    params = cl.params
    train_dev.close()

    def timed(learned, dts=None):
        d = fresh(prob)
        if learned:
            c = closure.Closure(d, capacity=0, max_blocks=8, seed=1)
            c.params = params
            c.enable(True, 0.1)
        d.advance(dts[0] if dts else 1e-6)  # warm-up of the launch path (not timed)
        d.close()
        d = fresh(prob)
        if learned:
            c = closure.Closure(d, capacity=0, max_blocks=8, seed=1)
            c.params = params
            c.enable(True, 0.1)
        d.synchronize()
        out = []
        t0 = time.perf_counter()
        if dts is None:
            for _ in range(steps):
                out.append(d.step(0.3))
        else:
            for dt in dts:
                d.advance(dt)
        d.synchronize()
        el = time.perf_counter() - t0
        w = d.download()
        d.close()
        return el, w, out

    _, _, dts = timed(False)  # the plain run's dt sequence
    el_exact, w_exact, _ = timed(False, dts)
    el_learned, w_learned, _ = timed(True, dts)
    errs = per_var_errors(prob, w_learned, w_exact, 8)
    return {"value": round(ncell * steps * 4 / el_learned / 1e6, 3), "unit": "Mcell-updates/s",
            "workload": f"stratified {n}^3, perturbation amplitude 0.1 c_s, {steps} RK4 steps with "
                        "the plain run's dt sequence, sensor threshold 0.1",
            "speedup_vs_fp64": round(el_exact / el_learned, 3),
            "rel_l2_vs_fp64_gpu": [float("%.3e" % e) for e in errs],
            "finite": bool(np.isfinite(w_learned).all()),
            "training": {"samples": int(samples), "sources": sources, "adam_steps": 4000,
                         "batch": 16384, "seconds": round(train_s, 3),
                         "mse_first_last": [float("%.3e" % losses[0]), float("%.3e" % losses[-1])]}}


def accuracy_leg():
    """The FP64 path against the oracle at the stratified 128^3
    configuration: 200 RK4 steps with the oracle's dt sequence, compared on
    the cells the committed oracle fixture holds (every 8th cell along each
    axis, and the plane at axis-3 index 64).  Per-variable rel-L2 and
    max-abs of each, beside the spread of the oracle itself: its FMA build
    against its plain build, and its run from the initial state moved by one
    ulp against its run from the exact one (how far apart two FP64 runs of
    this configuration end when they differ by rounding alone); None without
    the fixture."""
    import numpy as np
    from sac_amd import device, problems
    from sac_amd.errors import mesh
    path = os.path.join(ROOT, "tests", "golden", "stratified_128_200steps.npz")
    if not os.path.exists(path):
        return None
    fx = np.load(path)
    prob = problems.stratified(n=128)
    dev = device.Device(prob)
    dev.upload(prob.w)
    dev.getboundary()
    for dt in fx["dts"]:
        dev.advance(float(dt))
    w = dev.download()
    dev.close()
    m = mesh(prob, w[:8])
    fmt = lambda v: [float("%.3e" % e) for e in v]  # noqa: E731
    out = {"config": "stratified 128^3, 200 RK4 steps, FP64, the oracle's dt sequence",
           "oracle_fma_spread": fmt(fx["floor"])}
    for key, got in (("sub", m[:, ::8, ::8, ::8]), ("slice_k", m[:, 64])):
        ref = fx[key]
        out[key] = {"cells": ("every 8th cell along each axis" if key == "sub"
                              else "the plane at axis-3 index 64"),
                    "rel_l2": fmt(np.linalg.norm((got - ref).reshape(8, -1), axis=1) /
                                  np.linalg.norm(ref.reshape(8, -1), axis=1)),
                    "max_abs": fmt(np.abs(got - ref).reshape(8, -1).max(axis=1))}
        if "ulp_" + key in fx:
            out[key]["oracle_one_ulp_spread"] = fmt(fx["ulp_" + key])
    return out


def measured_traffic(precision, n):
    """HBM bytes per stage measured by rocprofv3 (FETCH_SIZE + WRITE_SIZE of
    the stage kernels, scaled by the copy-kernel calibration), from the
    committed profile summary of this workload; None if absent."""
    path = os.path.join(ROOT, "profiles", f"traffic_stratified{n}_{precision}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("bytes_per_stage")


if __name__ == "__main__":
    sys.exit(main())
