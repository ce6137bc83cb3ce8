#!/usr/bin/env python3
"""Benchmark of the plain gfx950 back-projection on the realistic X-band
workload: 4096 pulses x 8192 range samples (complex64, FP64 positions, nonlinear
track) back-projected onto a 4096 x 4096 ground grid at 0.25 m with 8-tap
Kaiser-sinc interpolation, FP64 tile reference + FP32 differential range,
FP32 accumulation.

One step = one full image.  Inputs are resident in HBM before timing; the
timed region is bracketed by device synchronisation and measured with HIP
events recorded on the stream the kernels are launched on.

Prints one JSON line (see the README of the measurement contract):
  value          pixel-pulse back-projections per second (W/s), one GPU
  roofline       FP32 VALU op roofline of the BP kernel (ops per W: 64 for
                 sinc8, 33 for linear; peak 78.6 Top/s = 256 CU x 64 lanes x
                 2 x 2.4 GHz)
  cpu_baseline   the FP64 CPU oracle on a bounded row sample of the same image
  accuracy       rel-L2 of the GPU rows against the oracle rows of that sample
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

F_W = {"nearest": 25, "linear": 33, "sinc8": 64}   # algorithmic VALU ops per W
PEAK_OPS = 78.6e12                                   # MI355X FP32 VALU op/s
PEAK_TFLOPS = 157.3                                  # published FP32 vector TFLOP/s
FLOP_W = {"nearest": 25, "linear": 33, "sinc8": 86}  # conventional FLOPs per W


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="config1")
    ap.add_argument("--interp", default=None, help="override the config interpolator")
    ap.add_argument("--precision", default="ref")
    ap.add_argument("--dtype", default="c64")
    ap.add_argument("--cpu-rows", type=int, default=8, help="oracle row sample (0: skip)")
    ap.add_argument("--no-plan", action="store_true", help="use sarbp_backproject (no host plan)")
    ap.add_argument("--ppl", type=int, default=0)
    ap.add_argument("--pulses-per-stage", type=int, default=0)
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the BP imager is a single-GPU build "
              "(multi-GPU sharding is a design note only); run with --gpus 1", file=sys.stderr)
        return 2

    import torch
    from sar_cambrianaccelerationcard_amd import device as D
    from sar_cambrianaccelerationcard_amd import metrics as M
    from sar_cambrianaccelerationcard_amd import native
    from sar_cambrianaccelerationcard_amd import scene as S

    if not torch.cuda.is_available():
        print("no HIP device", file=sys.stderr)
        return 3
    t0 = time.time()
    cfg = S.load_config(args.config)
    sc = S.make_scene(cfg)
    interp = args.interp or sc.interp
    print(f"[bench] scene {cfg['name']} simulated in {time.time() - t0:.1f}s: "
          f"{sc.n_pulses} pulses x {sc.n_samples} samples -> {sc.grid['nx']}x{sc.grid['ny']} "
          f"({interp}, {args.precision}, {args.dtype})", file=sys.stderr, flush=True)

    ctx = native.Context(0)
    if args.ppl or args.pulses_per_stage:
        ppl0, ps0 = ctx.get_tuning(native.INTERP_NAMES[interp], native.PREC_NAMES[args.precision],
                                   native.DTYPE_NAMES[args.dtype])
        ctx.set_tuning(native.INTERP_NAMES[interp], native.PREC_NAMES[args.precision],
                       native.DTYPE_NAMES[args.dtype], args.ppl or ppl0,
                       args.pulses_per_stage or ps0)
    ds = D.DeviceScene(sc)
    img = ds.new_image()
    params = sc.native_params(interp=interp, precision=args.precision, dtype=args.dtype)
    stream = torch.cuda.current_stream()
    sp = stream.cuda_stream
    prof_ptr = ds.profiles_ptr(args.dtype)
    plan = None if args.no_plan else ctx.plan(sc.pos, sc.native_grid(), params)
    info = plan.info() if plan else {}

    def step():
        if plan is not None:
            plan.execute(prof_ptr, ds.pos.data_ptr(), img.data_ptr(), sp)
        else:
            ctx.backproject(prof_ptr, ds.pos.data_ptr(), sc.native_grid(), params,
                            img.data_ptr(), sp)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    tw0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        step()
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - tw0
    kern_s = ev0.elapsed_time(ev1) / 1e3 / args.steps
    work = sc.work
    wps = work / kern_s
    print(f"[bench] {args.steps} steps: {kern_s * 1e3:.2f} ms/step (events), wall "
          f"{wall / args.steps * 1e3:.2f} ms/step, {wps:.4e} W/s", file=sys.stderr, flush=True)

    # ---- accuracy spot check + CPU baseline on a bounded row sample ----
    cpu = None
    acc = None
    if args.cpu_rows > 0:
        import oracle.oracle as O
        gpu_img = img.cpu().numpy()
        # rows through the first targets (so the sample holds focused energy)
        rows = sorted({int(r) for r in sc.target_pixels[:, 1]})[: max(1, args.cpu_rows // 2)]
        r_lo = max(0, rows[0] - args.cpu_rows // 2)
        r_hi = min(sc.grid["ny"], r_lo + args.cpu_rows)
        p = O.params(sc.c, sc.fc, sc.r0, sc.dr, interp, sc.kaiser_beta)
        nthr = O.max_threads()
        tc = time.perf_counter()
        ref = O.backproject(sc.profiles, sc.pos, sc.grid, p, rows=(r_lo, r_hi), n_threads=nthr)
        tcpu = time.perf_counter() - tc
        w_cpu = (r_hi - r_lo) * sc.grid["nx"] * sc.n_pulses
        cpu = {"value": w_cpu / tcpu, "unit": "pixel-pulses/s", "cores": nthr, "kind": "port",
               "sample": f"FP64 oracle, image rows [{r_lo},{r_hi}) x {sc.grid['nx']} px x "
                         f"{sc.n_pulses} pulses ({w_cpu:.3e} W, {tcpu:.1f} s)"}
        acc = {"rel_l2_rows": M.rel_l2(gpu_img[r_lo:r_hi], ref), "rows": [r_lo, r_hi]}
        print(f"[bench] oracle rows {r_lo}-{r_hi}: {tcpu:.1f}s, rel-L2 {acc['rel_l2_rows']:.3e}",
              file=sys.stderr, flush=True)

    fw = F_W[interp]
    achieved = wps * fw / 1e12
    line = {
        "metric": "pixel·pulse back-projections/s at rel-L2≤1e-4 vs FP64; p50/p99 job latency",
        "value": wps,
        "unit": "pixel-pulses/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": kern_s * 1e3,
        "wall_ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (FP64 tile reference range, FP32 differential, FP32 accumulation)"
        if args.precision == "ref" else args.precision,
        "data": "synthetic (seeded simulated scene)",
        "config": {"workload": f"{cfg['name']}: {sc.n_pulses} pulses x {sc.n_samples} samples "
                               f"{args.dtype} -> {sc.grid['nx']}x{sc.grid['ny']} image, {interp}, "
                               f"precision {args.precision}",
                   "n_pulses": sc.n_pulses, "n_samples": sc.n_samples, "nx": sc.grid["nx"],
                   "ny": sc.grid["ny"], "interp": interp, "precision": args.precision,
                   "profile_dtype": args.dtype, "plan": info},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_OPS / 1e12,
                     "unit": "Top/s", "frac": achieved / (PEAK_OPS / 1e12),
                     "ops_per_W": fw, "traffic": None,
                     "flop_convention": {"achieved_tflops": wps * FLOP_W[interp] / 1e12,
                                         "peak_tflops": PEAK_TFLOPS,
                                         "frac": wps * FLOP_W[interp] / 1e12 / PEAK_TFLOPS}},
        "cpu_baseline": cpu,
        "accuracy": acc,
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
