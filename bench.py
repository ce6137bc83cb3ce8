#!/usr/bin/env python3
"""Benchmark of the gfx950 back-projection imager on one MI355X.

Headline (value): the plain single-image BP on the realistic X-band workload,
4096 pulses x 8192 range samples (complex64, FP64 positions on a nonlinear
track) back-projected onto a 4096 x 4096 ground grid at 0.25 m with 8-tap
Kaiser-sinc interpolation, FP64 tile reference + FP32 differential range and
FP32 accumulation, after the library autotuner has picked the launch
configuration.  One step = one full image; inputs are resident in HBM before
timing; the timed region is bracketed by device synchronisation and measured
with HIP events recorded on the stream the kernels are launched on.

Extra legs (reported in the same JSON line):
  batched   64 mixed-size jobs: the persistent graph-replayed batched launch
            against the same jobs run one plain launch after another
  served    a Poisson stream of mixed jobs (10% large) through the
            continuous-batching server: throughput and p50 / p99 latency; the
            same stream as raw echoes (range-compressed in HBM, 3.3x fewer
            host-to-device bytes); and the reference's host flow per job
            (copy in, BP, copy out) as the end-to-end baseline
  roofline  FP32 VALU op roofline of the plain kernel (ops per pixel-pulse:
            64 sinc8, 33 linear; peak 78.6 Top/s = 256 CU x 64 lanes x 2 x
            2.4 GHz) and the conventional FLOP view (157.3 TFLOP/s)
  cpu_baseline  the FP64 CPU oracle on a bounded row sample of the image
  accuracy  rel-L2 of GPU rows against those oracle rows, and on 64 x 64
            chips around four point targets (peak pixel, phase, PSLR/ISLR)
  range_compression  the step before BP on the device: raw echoes of the
            same scene size (4096 pulses x 2448 samples) -> 8192 4x-upsampled
            samples, against the HBM roofline (8 TB/s), with the FP64 oracle
            timed on a pulse sample
  omega_k   the reference's second imaging path at its scene size: 3072 x
            2048 raw echoes (synthetic point targets, RADARSAT-1 radar) ->
            4096 x 3414 padded image (pad + down-conversion, 2-D FFT,
            reference function + 6-tap Stolt, 2-D inverse FFT); the Stolt
            kernel alone against the HBM roofline (16 B per sample); the FP64
            oracle timed on the whole image
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


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def timed(torch, stream, fn, reps: int) -> float:
    """Seconds per call of fn() measured with HIP events on `stream`."""
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        fn()
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / 1e3 / reps


def leg_range(torch, ctx, native, D, args) -> dict:
    """Device range compression of raw echoes of the headline scene's size."""
    M_, n_raw, n_out = args.range_pulses, 2448, 8192
    rd = dict(c=299792458.0, bandwidth=150e6, pulse_width=2e-6, fs=200e6, t0=3.2e-5,
              n_raw=n_raw, upsample=4, crop_offset=0, n_out=n_out)
    g = torch.Generator(device="cuda").manual_seed(5)
    raw = torch.randn((M_, n_raw), dtype=torch.complex64, device="cuda", generator=g)
    out = torch.empty((M_, n_out), dtype=torch.complex64, device="cuda")
    out16 = torch.empty((M_, n_out), dtype=torch.int32, device="cuda")
    stream = torch.cuda.current_stream()
    sp = stream.cuda_stream
    rp, rp16 = D.range_params(rd, "c64"), D.range_params(rd, "bf16x2")
    nf = native.range_info(rp)[0]

    def run(p, o):
        return lambda: ctx.range_compress(raw.data_ptr(), M_, p, o.data_ptr(), sp)

    res = {"workload": f"{M_} pulses x {n_raw} raw complex64 samples -> {n_out} samples "
                       f"(FFT {nf}, upsampling 4)", "n_pulses": M_, "fft_size": nf}
    for name, p, o, ob in [("c64", rp, out, 8), ("bf16x2", rp16, out16, 4)]:
        fn = run(p, o)
        for _ in range(2):
            fn()
        t = timed(torch, stream, fn, 10)
        byts = M_ * (n_raw * 8 + n_out * ob)
        res[name] = {"ms": t * 1e3, "GB_per_s": byts / t / 1e9, "frac_hbm": byts / t / 8e12,
                     "bytes": byts, "pulses_per_s": M_ / t}
        log(f"range compression ({name}): {t * 1e3:.3f} ms, {byts / t / 1e9:.0f} GB/s "
            f"({byts / t / 8e12:.1%} of 8 TB/s), {M_ / t:.3e} pulses/s")
    if args.cpu_rows > 0:
        import oracle.range_oracle as R
        n_s = 256
        h_raw = raw[:n_s].cpu().numpy().astype(np.complex128)
        tc = time.perf_counter()
        ref = R.range_compress(h_raw, chunk=64, **rd)
        tcpu = time.perf_counter() - tc
        got = out[:n_s].cpu().numpy()
        res["rel_l2_sample"] = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
        res["cpu_baseline"] = {"value": n_s / tcpu, "unit": "pulses/s", "cores": 1, "kind": "port",
                               "sample": f"FP64 numpy oracle (oracle/range_oracle.py), {n_s} pulses, "
                                         f"{tcpu:.2f} s"}
        log(f"range oracle {n_s} pulses: {tcpu:.2f}s ({n_s / tcpu:.1f} pulses/s), rel-L2 "
            f"{res['rel_l2_sample']:.2e}")
    del raw, out, out16
    torch.cuda.empty_cache()
    return res


def leg_omega_k(torch, ctx, native, args) -> dict:
    """omega-K image formation at the reference's scene size."""
    from sar_cambrianaccelerationcard_amd import device as D
    from sar_cambrianaccelerationcard_amd import omegak as O
    scene = O.radarsat_scene()
    targets = [O.aperture_target(scene, a, s) for a, s in
               ((900, 1400), (1300, 1500), (1700, 1600), (2100, 1450), (2500, 1650), (2900, 1550))]
    echo_h = O.simulate(scene, targets)
    na, nr = scene.padded
    echo = torch.from_numpy(echo_h).to("cuda")
    img = torch.empty((na, nr), dtype=torch.complex64, device="cuda")
    stream = torch.cuda.current_stream()
    sp = stream.cuda_stream
    wp = scene.params()
    fn = lambda: ctx.omega_k(echo.data_ptr(), scene.na_in, scene.nr_in, wp, img.data_ptr(), sp)
    tp = time.perf_counter()
    fn()
    torch.cuda.synchronize()
    t_plan = time.perf_counter() - tp
    fn()
    t = timed(torch, stream, fn, 10)
    n_pad = na * nr
    res = {"workload": f"{scene.na_in} x {scene.nr_in} raw complex64 echoes -> {na} x {nr} padded image, "
                       f"6-tap floor-window Stolt, synthetic point targets (RADARSAT-1 radar, squint "
                       f"f_nc = {scene.f_nc:.0f} Hz)",
           "ms": t * 1e3, "images_per_s": 1.0 / t, "echo_samples_per_s": scene.na_in * scene.nr_in / t,
           "first_call_ms": t_plan * 1e3}
    log(f"omega-K: {t * 1e3:.3f} ms per image ({scene.na_in * scene.nr_in / t:.3e} echo samples/s), "
        f"first call {t_plan * 1e3:.0f} ms")
    # The Stolt kernel alone on the padded spectrum (no reference function):
    # 8 B read + 8 B written per sample.
    fr = np.fft.fftshift(-nr / 2.0 + np.arange(nr)) * scene.fs / nr
    fa = scene.f_nc + np.fft.fftshift(-na / 2.0 + np.arange(na)) * scene.prf / na
    out = torch.empty_like(img)
    fs_ = lambda: ctx.stolt_sinc(img.data_ptr(), out.data_ptr(), fr, fa, f0=scene.f0, c=scene.c, vr=scene.vr,
                                 na=na, nr=nr, fs=scene.fs, taps=6, window=0, stream=sp)
    fs_()
    ts = timed(torch, stream, fs_, 10)
    byts = 16 * n_pad
    res["stolt"] = {"ms": ts * 1e3, "GB_per_s": byts / ts / 1e9, "frac_hbm": byts / ts / 8e12, "bytes": byts}
    log(f"Stolt kernel: {ts * 1e3:.3f} ms, {byts / ts / 1e9:.0f} GB/s ({byts / ts / 8e12:.1%} of 8 TB/s)")
    g = img.cpu().numpy()
    res["focus"] = [O.peak_metrics(g, *tg) for tg in targets]
    if args.cpu_rows > 0:
        import oracle.wk_oracle as WK
        tc = time.perf_counter()
        ref = WK.omega_k(echo_h, c=wp.c, f0=wp.f0, fs=wp.fs, prf=wp.prf, kr=wp.kr, vr=wp.vr, r0=wp.r0,
                         f_nc=wp.f_nc, taps=wp.taps, window=wp.window)
        tcpu = time.perf_counter() - tc
        res["rel_l2"] = float(np.linalg.norm(g - ref) / np.linalg.norm(ref))
        res["cpu_baseline"] = {"value": scene.na_in * scene.nr_in / tcpu, "unit": "echo samples/s", "cores": 1,
                               "kind": "port", "sample": f"FP64 numpy oracle (oracle/wk_oracle.py), the whole "
                                                         f"image, {tcpu:.2f} s"}
        log(f"omega-K oracle: {tcpu:.1f}s, rel-L2 {res['rel_l2']:.2e}")
    del echo, img, out
    torch.cuda.empty_cache()
    return res


def leg_mixed(torch, ctx, native, W, args, tuned_classes=()) -> dict:
    jobs = W.mixed_jobs(args.mixed_jobs)
    dev = torch.device("cuda", 0)
    max_prof = max(j.n_pulses * j.n_samples for j in jobs)
    g = torch.Generator(device=dev).manual_seed(11)
    prof = torch.randn(2 * max_prof, device=dev, generator=g, dtype=torch.float32)
    poss = [j.positions() for j in jobs]
    params = [j.params() for j in jobs]
    grids = [j.grid() for j in jobs]
    pptr = [prof.data_ptr()] * len(jobs)
    # a few jobs carry their own simulated scene (point targets over clutter,
    # as the realistic workload) instead of the shared seeded noise, for the
    # oracle spot checks of the batched images: the smallest job of each
    # quarter of the pixel-spacing range
    sims = {}
    if args.mixed_sims > 0:
        qs = np.quantile([j.spacing for j in jobs], np.linspace(0, 1, args.mixed_sims + 1))
        for a, b in zip(qs[:-1], qs[1:]):
            cand = [i for i, j in enumerate(jobs) if a <= j.spacing <= b and i not in sims]
            if not cand:
                continue
            i = min(cand, key=lambda k: jobs[k].work)
            sc = jobs[i].simulate(n_targets=6, n_clutter=512)
            d = torch.from_numpy(np.ascontiguousarray(sc.profiles)).to(dev)
            sims[i] = (sc, d)
            poss[i], params[i], grids[i], pptr[i] = sc.pos, sc.native_params(interp="sinc8"), sc.native_grid(), \
                d.data_ptr()
    dpos = [torch.from_numpy(p).to(dev) for p in poss]
    imgs = [torch.empty((j.side, j.side), dtype=torch.complex64, device=dev) for j in jobs]
    stream = torch.cuda.current_stream()
    sp = stream.cuda_stream
    # autotune every size class of the mix on its largest job (the classes'
    # tunings are kept by the context and used by plans, batches and servers)
    classes = {}
    for i in range(len(jobs)):
        classes.setdefault(native.size_class(grids[i], params[i], poss[i]), []).append(i)
    tuned = {}
    t_tune = time.time()
    shape = None
    if args.class_shapes == "common":
        # one tile shape for every class of the mix, the headline's class
        # included (a batch then needs one persistent launch per sinc8
        # variant): the candidate with the least work-weighted time over the
        # classes' largest jobs
        reps = {c: max(idx, key=lambda i: jobs[i].work) for c, idx in classes.items()}
        best_total = None
        for cand in ((1, 4), (1, 8)):
            total = 0.0
            for c, k in reps.items():
                r = tune_stages(torch, ctx, native, stream, cand, c, prof, dpos[k], poss[k], grids[k], params[k],
                                imgs[k], keep=False)
                total += r["best_ms"] * sum(jobs[i].work for i in classes[c]) / jobs[k].work
            log(f"common tile shape {cand[0]}x{cand[1]}: {total:.1f} ms work-weighted")
            if best_total is None or total < best_total[0]:
                best_total = (total, cand)
        shape = best_total[1]
    for c, idx in sorted(classes.items()):
        kept = ctx.get_class_tuning(native.INTERP_SINC8, native.PREC_REF, native.DTYPE_C64, c)
        if c in tuned_classes and (shape is None or (kept[1], kept[0]) == shape):
            # tuned by the headline leg on a whole image, at the mix's shape
            tuned[c] = {"jobs": len(idx), "work": int(sum(jobs[i].work for i in idx)), "tuned_on": "headline"}
            continue
        k = max(idx, key=lambda i: jobs[i].work)
        if args.class_shapes == "free":
            r = ctx.autotune(prof.data_ptr(), dpos[k].data_ptr(), poss[k], grids[k], params[k], imgs[k].data_ptr(),
                             jobs[k].side ** 2, args.class_crop_px, sp)
        else:
            r = tune_stages(torch, ctx, native, stream, shape, c, prof, dpos[k], poss[k], grids[k], params[k],
                            imgs[k])
        tuned[c] = {"jobs": len(idx), "work": int(sum(jobs[i].work for i in idx)), "tuned_on": k,
                    "spacing_m": jobs[k].spacing, "tile": [r["cols_per_lane"], r["rows_per_lane"]],
                    "pulses_per_stage": r["pulses_per_stage"], "best_ms": r["best_ms"],
                    "default_ms": r["default_ms"]}
        log(f"class {c} ({len(idx)} jobs): {r['cols_per_lane']}x{r['rows_per_lane']} P{r['pulses_per_stage']}, "
            f"{r['best_ms']:.2f} ms vs built-in {r['default_ms']:.2f} ms on job {k}")
    log(f"per-class autotuning: {time.time() - t_tune:.1f}s")
    plans = [ctx.plan(p, gr, pa) for p, gr, pa in zip(poss, grids, params)]
    for c, idx in sorted(classes.items()):
        infos = [plans[i].info() for i in idx]
        lds = max(i["lds_bytes"] for i in infos)
        tuned[c].update({"windows": [min(i["window"] for i in infos), max(i["window"] for i in infos)],
                         "max_lds_bytes": lds, "workgroups_per_cu_by_lds": 163840 // (lds + 1024),
                         "sinc8_halves": sorted({i["sinc8_halves"] for i in infos})})
        log(f"class {c}: windows {tuned[c]['windows']}, LDS <= {lds} B "
            f"({tuned[c]['workgroups_per_cu_by_lds']} workgroups/CU), sinc8 halves {tuned[c]['sinc8_halves']}")
    work = sum(j.work for j in jobs)

    def plain_all():
        for pl, pp, dp, im in zip(plans, pptr, dpos, imgs):
            pl.execute(pp, dp.data_ptr(), im.data_ptr(), sp)

    def job(i, im):
        jb = native.Job()
        jb.d_profiles = pptr[i]
        jb.d_pos = dpos[i].data_ptr()
        jb.h_pos = poss[i].ctypes.data
        jb.grid = grids[i]
        jb.params = params[i]
        jb.d_image = im.data_ptr()
        return jb

    bimgs = [torch.empty_like(im) for im in imgs]
    batch = [job(i, bimgs[i]) for i in range(len(jobs))]

    def batched():
        ctx.backproject_batch(batch, sp)

    plain_all()
    batched()
    t_plain = timed(torch, stream, plain_all, args.mixed_reps)
    t_batch = timed(torch, stream, batched, args.mixed_reps)
    same = all(torch.equal(a.view(torch.float32), b.view(torch.float32)) for a, b in zip(imgs, bimgs))
    log(f"mixed {len(jobs)} jobs ({work:.3e} W): plain sequential {t_plain * 1e3:.1f} ms "
        f"({work / t_plain:.3e} W/s), batched {t_batch * 1e3:.1f} ms ({work / t_batch:.3e} W/s), "
        f"x{t_plain / t_batch:.2f}, bitwise equal {same}")
    checks = []
    if sims:
        import oracle.oracle as O
        import oracle.spot as SP
        for i, (sc, _) in sorted(sims.items()):
            p = O.params(sc.c, sc.fc, sc.r0, sc.dr, "sinc8", sc.kaiser_beta)
            tp = sc.target_pixels
            r = SP.spot_check(sc.profiles, sc.pos, sc.grid, p, bimgs[i].cpu().numpy(), [int(tp[0][1])],
                              [tuple(int(v) for v in tp[0])])
            r.update({"job": i, "side": jobs[i].side, "n_pulses": jobs[i].n_pulses,
                      "spacing_m": jobs[i].spacing,
                      "size_class": native.size_class(grids[i], params[i], poss[i])})
            checks.append(r)
            log(f"batched spot check job {i} ({jobs[i].side}^2 at {jobs[i].spacing:.2f} m, {jobs[i].n_pulses} "
                f"pulses, class {r['size_class']}): rel-L2 {r['rel_l2']:.2e}, "
                f"peak {'ok' if all(r['chip_peak_match']) else 'MISMATCH'}")
    del prof, imgs, bimgs, dpos, sims
    torch.cuda.empty_cache()
    return {"workload": f"{len(jobs)} mixed jobs: side logU[256,2048] px at U[0.25,1] m, "
                        f"pulses U[512,4096], samples U[2048,8192], sinc8; {len(checks)} of them simulated "
                        f"scenes (point targets over clutter, oracle spot checks), the rest seeded noise",
            "n_jobs": len(jobs), "work": work,
            "plain_sequential_W_per_s": work / t_plain, "batched_W_per_s": work / t_batch,
            "speedup_vs_plain": t_plain / t_batch, "plain_ms": t_plain * 1e3,
            "batched_ms": t_batch * 1e3, "batched_equals_plain_bitwise": bool(same),
            "size_classes": tuned,
            "spot_checks": checks,
            "spot_check_max_rel_l2": max((c["rel_l2"] for c in checks), default=None)}


def tune_stages(torch, ctx, native, stream, shape, cls, prof, dpos, pos, grid, params, img, keep=True) -> dict:
    """Pulses per stage for one size class at a fixed tile shape (kept as the
    class's tuning unless keep is False): jobs of all classes then share a
    shape, so a batch or a server epoch needs one persistent launch per sinc8
    variant rather than one per shape.  As in the library autotuner, a
    built-in default of the same shape stays unless a candidate is more than
    2% faster on the tuning job: near-ties on one job do not carry to the
    class's other jobs (a deeper stage can cost them occupancy)."""
    I, Pm, Dt = native.INTERP_SINC8, native.PREC_REF, native.DTYPE_C64
    default = ctx.get_class_tuning(I, Pm, Dt, cls)
    best = None
    for P in (1, 2, 3, 4, 6, 8):
        ctx.set_class_tuning(I, Pm, Dt, cls, shape[1], shape[0], P)
        pl = ctx.plan(pos, grid, params)
        if pl.info()["pulses_per_stage"] != P:
            continue  # the LDS budget cut it: the same as a smaller P
        fn = lambda: pl.execute(prof.data_ptr(), dpos.data_ptr(), img.data_ptr(), stream.cuda_stream)
        fn()
        t = timed(torch, stream, fn, 2)
        if best is None or t < best[0]:
            best = (t, P)
    ctx.set_class_tuning(I, Pm, Dt, cls, *default)
    pl = ctx.plan(pos, grid, params)
    fn = lambda: pl.execute(prof.data_ptr(), dpos.data_ptr(), img.data_ptr(), stream.cuda_stream)
    fn()
    t_def = timed(torch, stream, fn, 2)
    if (default[1], default[0]) == tuple(shape) and best[0] > 0.98 * t_def:
        best = (t_def, default[2])
    if keep:
        ctx.set_class_tuning(I, Pm, Dt, cls, shape[1], shape[0], best[1])
    return {"cols_per_lane": shape[0], "rows_per_lane": shape[1], "pulses_per_stage": best[1],
            "best_ms": best[0] * 1e3, "default_ms": t_def * 1e3}


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def median_time(fn, runs: int = 3) -> float:
    """Median wall time of `runs` calls of fn() (the caller runs a warm-up first)."""
    ts = []
    for _ in range(runs):
        t = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t)
    return float(np.median(ts))


def headline_chips(sc, gpu_img, p, nthr, n_chips: int = 4, half: int = 32) -> dict:
    """The GPU image against the FP64 oracle on 2half x 2half chips centred on
    the first n_chips point targets at least `half` pixels inside the image:
    combined rel-L2, and per chip its rel-L2, whether the brightest pixel is
    the oracle's, the phase error at that peak and the largest PSLR / ISLR
    difference along the peak's cuts."""
    import oracle.oracle as O
    import oracle.spot as SP
    from sar_cambrianaccelerationcard_amd import metrics as M
    nx, ny = sc.grid["nx"], sc.grid["ny"]
    tgt = [(int(tx), int(ty)) for tx, ty in sc.target_pixels
           if half <= tx < nx - half and half <= ty < ny - half][:n_chips]
    num = den = 0.0
    chips = []
    for tx, ty in tgt:
        x0, y0 = tx - half, ty - half
        ref = O.backproject(sc.profiles, sc.pos, SP.chip_grid(sc.grid, x0, y0, 2 * half), p, n_threads=nthr)
        got = gpu_img[y0:y0 + 2 * half, x0:x0 + 2 * half]
        num += float(np.sum(np.abs(got.astype(np.complex128) - ref) ** 2))
        den += float(np.sum(np.abs(ref) ** 2))
        rec = M.compare_targets(got, ref, np.array([[half, half]]))[0]
        chips.append({"target": [tx, ty], "rel_l2": M.rel_l2(got, ref), "peak_match": bool(rec["peak_match"]),
                      "phase_err_deg": rec["phase_err_deg"], "max_pslr_islr_diff_db": rec["max_quality_diff_db"]})
    out = {"rel_l2_chips": float(np.sqrt(num / den)) if den > 0 else None, "chips": chips}
    if chips:
        log(f"GPU target chips ({len(chips)} x {2 * half}^2 px) vs FP64 oracle: rel-L2 {out['rel_l2_chips']:.3e}, "
            f"peaks {'exact' if all(c['peak_match'] for c in chips) else 'MISMATCH'}, max |phase err| "
            f"{max(abs(c['phase_err_deg']) for c in chips):.2e} deg, max PSLR/ISLR diff "
            f"{max(c['max_pslr_islr_diff_db'] for c in chips):.2e} dB")
    return out


def leg_cpu(sc, gpu_img, interp, args):
    """The reference's CPU BP semantics (BPA_SAR.m:178-197) restated in C
    (oracle/sarbp_oracle.c): FP64 and FP32, all host threads and one thread,
    median of 3 timed runs after a warm-up, on a fixed row sample of the
    headline image and on the whole small plain image (config 0).  The FP64
    warm-up rows are also the accuracy reference of the GPU image."""
    import oracle.oracle as O
    from sar_cambrianaccelerationcard_amd import metrics as M
    from sar_cambrianaccelerationcard_amd import scene as S
    nthr = O.max_threads()
    rows = sorted({int(r) for r in sc.target_pixels[:, 1]})
    r_lo = max(0, rows[0] - args.cpu_rows // 2)
    r_hi = min(sc.grid["ny"], r_lo + args.cpu_rows)
    p = O.params(sc.c, sc.fc, sc.r0, sc.dr, interp, sc.kaiser_beta)
    nx = sc.grid["nx"]
    w_rows = (r_hi - r_lo) * nx * sc.n_pulses
    tc = time.perf_counter()
    ref = O.backproject(sc.profiles, sc.pos, sc.grid, p, rows=(r_lo, r_hi), n_threads=nthr)
    t_warm = time.perf_counter() - tc
    acc = {"rel_l2_rows": M.rel_l2(gpu_img[r_lo:r_hi], ref), "rows": [r_lo, r_hi]}
    acc.update(headline_chips(sc, gpu_img, p, nthr))
    t64 = median_time(lambda: O.backproject(sc.profiles, sc.pos, sc.grid, p, rows=(r_lo, r_hi),
                                            n_threads=nthr))
    f32 = O.backproject_f32(sc.profiles, sc.pos, sc.grid, p, rows=(r_lo, r_hi), n_threads=nthr)
    t32 = median_time(lambda: O.backproject_f32(sc.profiles, sc.pos, sc.grid, p, rows=(r_lo, r_hi),
                                                n_threads=nthr))
    # one thread: a 512-pixel segment of the first sample row
    seg = {"origin": list(sc.grid["origin"]), "ux": list(sc.grid["ux"]), "uy": list(sc.grid["uy"]),
           "nx": min(512, nx), "ny": sc.grid["ny"]}
    w_seg = seg["nx"] * sc.n_pulses
    O.backproject(sc.profiles, sc.pos, seg, p, rows=(r_lo, r_lo + 1), n_threads=1)
    t64_1 = median_time(lambda: O.backproject(sc.profiles, sc.pos, seg, p, rows=(r_lo, r_lo + 1), n_threads=1))
    O.backproject_f32(sc.profiles, sc.pos, seg, p, rows=(r_lo, r_lo + 1), n_threads=1)
    t32_1 = median_time(lambda: O.backproject_f32(sc.profiles, sc.pos, seg, p, rows=(r_lo, r_lo + 1),
                                                  n_threads=1))
    log(f"CPU oracle rows {r_lo}-{r_hi} ({w_rows:.3e} W): FP64 {w_rows / t64:.3e} W/s on {nthr} threads "
        f"(warm-up {t_warm:.1f}s), FP32 {w_rows / t32:.3e} W/s; one thread FP64 {w_seg / t64_1:.3e}, "
        f"FP32 {w_seg / t32_1:.3e} W/s; GPU rows rel-L2 {acc['rel_l2_rows']:.3e}")
    # config 0 (512 pulses x 2048 samples -> 512 x 512 at 0.5 m, linear) in full
    s0 = S.make_scene(S.load_config("config0"))
    p0 = O.params(s0.c, s0.fc, s0.r0, s0.dr, s0.interp, s0.kaiser_beta)
    w0 = s0.work
    O.backproject(s0.profiles, s0.pos, s0.grid, p0, n_threads=nthr)
    c0_64 = median_time(lambda: O.backproject(s0.profiles, s0.pos, s0.grid, p0, n_threads=nthr))
    O.backproject_f32(s0.profiles, s0.pos, s0.grid, p0, n_threads=nthr)
    c0_32 = median_time(lambda: O.backproject_f32(s0.profiles, s0.pos, s0.grid, p0, n_threads=nthr))
    rows0 = (0, 32)  # one thread: the first 32 rows
    w0_1 = 32 * s0.grid["nx"] * s0.n_pulses
    O.backproject(s0.profiles, s0.pos, s0.grid, p0, rows=rows0, n_threads=1)
    c0_64_1 = median_time(lambda: O.backproject(s0.profiles, s0.pos, s0.grid, p0, rows=rows0, n_threads=1))
    acc["_ref_rows"] = ref
    log(f"CPU oracle config 0 in full ({w0:.3e} W, {s0.interp}): FP64 {w0 / c0_64:.3e} W/s, FP32 "
        f"{w0 / c0_32:.3e} W/s on {nthr} threads; one thread FP64 {w0_1 / c0_64_1:.3e} W/s")
    cpu = {"value": w_rows / t64, "unit": "pixel-pulses/s", "cores": nthr, "kind": "port",
           "sample": f"FP64 oracle (oracle/sarbp_oracle.c), {interp}, image rows [{r_lo},{r_hi}) x "
                     f"{nx} px x {sc.n_pulses} pulses ({w_rows:.3e} W), median of 3 after a warm-up",
           "cpu_model": cpu_model(), "nproc": os.cpu_count(),
           "affinity_cpus": len(os.sched_getaffinity(0)),
           "runs": {
               "headline_rows_fp64": {"W_per_s": w_rows / t64, "threads": nthr, "s": t64},
               "headline_rows_fp32": {"W_per_s": w_rows / t32, "threads": nthr, "s": t32,
                                      "rel_l2_vs_fp64": M.rel_l2(f32, ref)},
               "headline_segment_fp64_1thread": {"W_per_s": w_seg / t64_1, "threads": 1, "s": t64_1,
                                                 "sample": f"row {r_lo}, {seg['nx']} px"},
               "headline_segment_fp32_1thread": {"W_per_s": w_seg / t32_1, "threads": 1, "s": t32_1,
                                                 "sample": f"row {r_lo}, {seg['nx']} px"},
               "config0_full_fp64": {"W_per_s": w0 / c0_64, "threads": nthr, "s": c0_64,
                                     "interp": s0.interp, "work": w0},
               "config0_full_fp32": {"W_per_s": w0 / c0_32, "threads": nthr, "s": c0_32,
                                     "interp": s0.interp, "work": w0},
               "config0_rows_fp64_1thread": {"W_per_s": w0_1 / c0_64_1, "threads": 1, "s": c0_64_1,
                                             "sample": "rows [0,32)"}}}
    return cpu, acc


def leg_modes(torch, ctx, native, sc, ds, interp, args, acc) -> dict:
    """The other accurate paths on the headline scene: the unplanned drop-in
    entry point (sarbp_backproject: no host positions, so windows come from
    the triangle inequality), and the precision modes F64 (FP64 range and
    phase per pixel-pulse), DF (FP64 tile reference + double-float per-pixel
    differential) and F32 (FP32 range; inaccurate), each planned.  HIP-event
    time per launch, and rel-L2 on the oracle rows of the accuracy check."""
    from sar_cambrianaccelerationcard_amd import metrics as M
    stream = torch.cuda.current_stream()
    sp = stream.cuda_stream
    img = ds.new_image()
    out = {}
    ref = acc.get("_ref_rows") if acc else None
    for name, prec, planned in (("ref_unplanned", "ref", False), ("df", "df", True),
                                ("f64", "f64", True), ("f32", "f32", True)):
        params = sc.native_params(interp=interp, precision=prec)
        if planned:
            plan = ctx.plan(sc.pos, sc.native_grid(), params)
            info = plan.info()

            def fn():
                plan.execute(ds.prof.data_ptr(), ds.pos.data_ptr(), img.data_ptr(), sp)
        else:
            info = None

            def fn():
                ctx.backproject(ds.prof.data_ptr(), ds.pos.data_ptr(), sc.native_grid(), params,
                                img.data_ptr(), sp)
        fn()
        t = timed(torch, stream, fn, max(1, args.mode_reps))
        r = {"W_per_s": sc.work / t, "ms": t * 1e3, "precision": prec, "planned": planned}
        if info:
            r["tile"] = f"{info['cols_per_lane']}x{info['rows_per_lane']} px/lane, " \
                        f"{info['pulses_per_stage']} pulses/stage, window {info['window']}"
        if ref is not None:
            lo, hi = acc["rows"]
            r["rel_l2_rows"] = M.rel_l2(img[lo:hi].cpu().numpy(), ref)
        log(f"mode {name}: {t * 1e3:.2f} ms/launch, {r['W_per_s']:.4e} W/s"
            + (f", rows rel-L2 {r['rel_l2_rows']:.2e}" if "rel_l2_rows" in r else ""))
        out[name] = r
    del img
    return out


class ServedInputs:
    """A served stream: the jobs, shared seeded noise profiles for the bulk,
    simulated scenes for the spot-checked jobs, FP64 positions and host image
    buffers (one shared buffer per size class for the bulk, whose images are
    not inspected; an own buffer for each spot-checked job)."""

    def __init__(self, W, args):
        self.jobs = W.mixed_jobs(args.served_jobs, seed=20260041, large_fraction=0.1)
        jobs = self.jobs
        small_max = max((j.n_pulses * j.n_samples for j in jobs if not j.large), default=1)
        large_max = max((j.n_pulses * j.n_samples for j in jobs if j.large), default=0)
        t0 = time.time()
        self.small_prof = W.random_profiles(small_max, 41)
        self.large_prof = W.random_profiles(large_max, 42) if large_max else None
        self.pos = [j.positions() for j in jobs]
        self.small_img = np.empty(max((j.side ** 2 for j in jobs if not j.large), default=1), np.complex64)
        self.large_img = np.empty(max((j.side ** 2 for j in jobs if j.large), default=1), np.complex64)
        # spot-checked jobs: the first large job and the first two small ones
        self.sims = {}
        if args.cpu_rows > 0:
            larges = [k for k, j in enumerate(jobs) if j.large][:1]
            smalls = [k for k, j in enumerate(jobs) if not j.large][:2]
            for k in larges + smalls:
                j = jobs[k]
                self.sims[k] = j.simulate(n_targets=12 if j.large else 6, n_clutter=4096 if j.large else 512)
        self.images = {}
        log(f"served: {len(jobs)} jobs ({sum(j.large for j in jobs)} large, "
            f"{sum(j.work for j in jobs):.3e} W), host inputs and {len(self.sims)} simulated scenes in "
            f"{time.time() - t0:.1f}s")

    def data(self, k):
        """(profiles, positions, grid, params, host image) of job k."""
        j = self.jobs[k]
        if k in self.sims:
            sc = self.sims[k]
            img = self.images.setdefault(k, np.empty((j.side, j.side), np.complex64))
            return sc.profiles, sc.pos, sc.native_grid(), sc.native_params(interp="sinc8"), img
        src = self.large_prof if j.large else self.small_prof
        buf = self.large_img if j.large else self.small_img
        return (src[: j.n_pulses * j.n_samples], self.pos[k], j.grid(), j.params(),
                buf[: j.side * j.side].reshape(j.side, j.side))


def plain_e2e(torch, ctx, native, inp, n_sample: int, class_tunings=None) -> dict:
    """The reference host flow per image (SAR_Processor, BPA_CUDA.cpp:51-89):
    copy the job's profiles and positions to the device, back-project, copy
    the image back, synchronise -- one job after another, on a sample of the
    stream that keeps its share of large jobs (the first ones of each class in
    stream order).  The back-projection is the strongest one-job call: a plan
    made from the host positions (the reference's host holds them before it
    copies them), planning time included, as the server plans each job.
    Launch configurations: the context's per-class tunings with class_tunings
    over them, as the server of the served leg runs them."""
    I, Pm, Dt = native.INTERP_SINC8, native.PREC_REF, native.DTYPE_C64
    saved = {c: ctx.get_class_tuning(I, Pm, Dt, c) for c in (class_tunings or {})}
    for c, t in (class_tunings or {}).items():
        ctx.set_class_tuning(I, Pm, Dt, c, *t)
    try:
        return _plain_e2e(torch, ctx, inp, n_sample)
    finally:
        for c, t in saved.items():
            ctx.set_class_tuning(I, Pm, Dt, c, *t)


def _plain_e2e(torch, ctx, inp, n_sample: int) -> dict:
    jobs = inp.jobs
    n_large = max(1, round(0.1 * n_sample)) if any(j.large for j in jobs) else 0
    ks = [k for k, j in enumerate(jobs) if j.large][:n_large] + \
         [k for k, j in enumerate(jobs) if not j.large][:n_sample - n_large]
    ks.sort()
    dev = torch.device("cuda", 0)
    max_prof = max(jobs[k].n_pulses * jobs[k].n_samples for k in ks)
    max_pix = max(jobs[k].side ** 2 for k in ks)
    d_prof = torch.empty(max_prof, dtype=torch.complex64, device=dev)
    d_img = torch.empty(max_pix, dtype=torch.complex64, device=dev)
    stream = torch.cuda.current_stream().cuda_stream
    t0 = time.perf_counter()
    for k in ks:
        prof, pos, grid, params, _ = inp.data(k)
        j = jobs[k]
        d_prof[: j.n_pulses * j.n_samples].copy_(torch.from_numpy(np.ascontiguousarray(prof).reshape(-1)))
        d_pos = torch.from_numpy(pos).to(dev)
        pl = ctx.plan(pos, grid, params)
        pl.execute(d_prof.data_ptr(), d_pos.data_ptr(), d_img.data_ptr(), stream)
        pl.close()
        out = d_img[: j.side * j.side].cpu()
        del out
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    work = sum(jobs[k].work for k in ks)
    del d_prof, d_img
    torch.cuda.empty_cache()
    log(f"plain end-to-end, one job after another ({len(ks)} jobs, {n_large} large): {wall:.2f}s, "
        f"{work / wall:.3e} W/s")
    return {"W_per_s": work / wall, "wall_s": wall, "n_jobs": len(ks), "n_large": n_large, "work": work}


def served_raw_inputs(W, jobs):
    """Seeded raw echoes (one buffer per size class) for raw-echo submissions."""
    small_max = max((j.n_pulses * j.n_raw for j in jobs if not j.large), default=1)
    large_max = max((j.n_pulses * j.n_raw for j in jobs if j.large), default=0)
    return W.random_profiles(small_max, 43), (W.random_profiles(large_max, 44) if large_max else None)


def leg_served(native, W, args, rate: float, inp, n_jobs: int, raw_inputs=None, ctx=None,
               class_tunings=None) -> dict:
    """The first n_jobs of the stream as Poisson arrivals (rate jobs/s; 0 =
    all at once) through the server.  With raw_inputs the jobs are submitted
    as raw echoes and range-compressed on the device at admission.  The
    server takes ctx's per-class tunings, then class_tunings ({class:
    (rows, cols, pulses per stage)}) over them."""
    from sar_cambrianaccelerationcard_amd import device as D
    jobs = inp.jobs[:n_jobs]
    srv = native.Server(interp="sinc8", hbm_budget_bytes=args.hbm_budget,
                        pulses_per_epoch=args.pulses_per_epoch)
    if ctx is not None:
        srv.tune_like(ctx)  # the context's per-class (autotuned) launch configurations
    for c, t in (class_tunings or {}).items():
        srv.set_class_tuning(c, *t)
    rng = np.random.default_rng(7)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, len(jobs))) if rate > 0 else np.zeros(len(jobs))
    tag = f"served{' raw' if raw_inputs is not None else ''} @ {rate if rate > 0 else 'burst'}/s"
    ids = []
    step = max(1, len(jobs) // 5)
    start = time.perf_counter()
    for k, (j, at) in enumerate(zip(jobs, arrivals)):
        now = time.perf_counter() - start
        if at > now:
            time.sleep(at - now)
        prof, pos, grid, params, img = inp.data(k)
        if raw_inputs is None:
            ids.append(srv.submit(prof, pos, grid, params, img))
        else:
            src = raw_inputs[1] if j.large else raw_inputs[0]
            ids.append(srv.submit_raw(src[: j.n_pulses * j.n_raw], D.range_params(j.radar()), pos,
                                      grid, params, img))
        if (k + 1) % step == 0 and k + 1 < len(jobs):
            log(f"{tag}: {k + 1}/{len(jobs)} submitted at {time.perf_counter() - start:.1f}s")
    infos = []
    for k, i in enumerate(ids):
        infos.append(srv.wait(i, timeout_s=900))
        if (k + 1) % step == 0:
            log(f"{tag}: {k + 1}/{len(jobs)} done at {time.perf_counter() - start:.1f}s")
    st = srv.stats()
    srv.close()
    lat = np.array([i["done_s"] - i["submit_s"] for i in infos])
    lat_small = np.array([i["done_s"] - i["submit_s"] for i, j in zip(infos, jobs) if not j.large])
    lat_large = np.array([i["done_s"] - i["submit_s"] for i, j in zip(infos, jobs) if j.large])
    span = max(i["done_s"] for i in infos) - min(i["submit_s"] for i in infos)
    work = sum(j.work for j in jobs)
    pct = lambda a, q: float(np.percentile(a, q)) if a.size else None
    out = {"rate_jobs_per_s": rate, "n_jobs": len(jobs), "n_large": int(sum(j.large for j in jobs)),
           "work": work, "span_s": span, "W_per_s": work / span,
           "W_per_s_gpu_busy": work / max(st["gpu_busy_s"], 1e-9),
           "p50_latency_s": pct(lat, 50), "p99_latency_s": pct(lat, 99),
           "p50_latency_small_s": pct(lat_small, 50), "p99_latency_small_s": pct(lat_small, 99),
           "p50_latency_large_s": pct(lat_large, 50), "p99_latency_large_s": pct(lat_large, 99),
           "epochs": st["epochs"], "hbm_peak_bytes": st["hbm_peak"], "failed": st["failed"],
           "pulses_per_epoch": args.pulses_per_epoch, "hbm_budget_bytes": args.hbm_budget,
           "input": "raw echoes" if raw_inputs is not None else "profiles",
           "h2d_input_bytes": int(sum(j.n_pulses * (j.n_raw if raw_inputs is not None else j.n_samples) * 8
                                      for j in jobs))}
    log(f"{tag}: {out['W_per_s']:.3e} W/s over {span:.1f}s (GPU-busy {out['W_per_s_gpu_busy']:.3e}), "
        f"p50 {out['p50_latency_s']:.3f}s p99 {out['p99_latency_s']:.3f}s, {st['epochs']} epochs, "
        f"failed {st['failed']}")
    return out


def served_spot_checks(inp, images) -> list:
    """Oracle spot checks of the served jobs that carry simulated scenes: whole
    rows plus 64 x 64 chips centred on point targets (2 rows + 2 chips for a
    large job, 1 + 1 for a small one)."""
    import oracle.oracle as O
    import oracle.spot as SP
    out = []
    for k, sc in sorted(inp.sims.items()):
        j = inp.jobs[k]
        img = images[k]
        p = O.params(sc.c, sc.fc, sc.r0, sc.dr, "sinc8", sc.kaiser_beta)
        tp = sc.target_pixels
        n = 2 if j.large else 1
        rows = [int(tp[0][1]), sc.grid["ny"] // 2][:n]
        chips = [tuple(int(v) for v in t) for t in tp[:n]]
        t0 = time.perf_counter()
        r = SP.spot_check(sc.profiles, sc.pos, sc.grid, p, img, rows, chips)
        r.update({"job": k, "large": j.large, "side": j.side, "n_pulses": j.n_pulses,
                  "n_samples": j.n_samples, "spacing_m": j.spacing, "oracle_s": time.perf_counter() - t0})
        out.append(r)
        log(f"served spot check job {k} ({'large' if j.large else 'small'} {j.side}^2 x {j.n_pulses} "
            f"pulses): rel-L2 {r['rel_l2']:.2e}, peaks {'ok' if all(r['chip_peak_match']) else 'MISMATCH'}")
    return out


def valu_probe(args) -> int:
    """Issue rate of independent FP32 FMA chains, hardware-sine chains and their
    mixes (sarbp_valu_probe): wave64 VALU instructions per CU per cycle at the
    2.4 GHz peak clock, from HIP events."""
    import torch
    from sar_cambrianaccelerationcard_amd import native
    props = torch.cuda.get_device_properties(0)
    n_cu = props.multi_processor_count
    n_blocks, iters = n_cu * 16, 2048
    out = torch.empty(n_blocks * 256, dtype=torch.float32, device="cuda")
    stream = torch.cuda.current_stream()
    res = {"n_cu": n_cu, "n_blocks": n_blocks, "threads_per_block": 256, "iters": iters, "kinds": {}}
    names = {0: "128 fma (VOP3)", 1: "128 sin", 2: "120 fma + 8 sin", 3: "112 fma + 16 sin",
             4: "128 fmac (VOP2)", 5: "128 mul (VOP2)", 6: "128 add (VOP2)", 7: "128 fract (VOP1)",
             8: "128 med3 (VOP3)", 9: "64 fma + 64 fmac", 10: "64 fma + 64 add_u32",
             11: "64 fma + 64 mad_u32_u24", 12: "64 fma + 64 cvt_f32_i32", 13: "64 fma + 64 fract",
             14: "64 fma + 64 med3", 15: "64 fma + 64 mul", 16: "64 fma + 64 mov",
             17: "64 fma + 64 add_f32", 18: "64 fma + 64 lshrrev_b32", 19: "64 fma + 64 and_b32",
             20: "64 fma + 64 or_b32", 21: "64 fma + 64 min_u32", 22: "64 fma + 64 cndmask_b32",
             23: "64 fma + 64 lshl_add_u32", 24: "64 fma + 64 add3_u32", 25: "64 fma + 64 cvt_i32_f32",
             26: "64 fma + 64 max_f32", 27: "64 fma + 64 floor_f32", 28: "64 fma + 64 sub_u32",
             29: "64 fma + 64 bfe_u32", 30: "64 fma + 64 ldexp_f32", 31: "64 fma + 64 ashrrev_i32"}
    for kind in range(32):
        fn = lambda: native.valu_probe(kind, n_blocks, iters, out.data_ptr(), stream.cuda_stream)
        fn()
        t = timed(torch, stream, fn, 5)
        wave_instr = n_blocks * 4 * iters * 128
        per_cu_cycle = wave_instr / (n_cu * 2.4e9 * t)
        res["kinds"][names[kind]] = {"ms": t * 1e3, "wave_instr": wave_instr,
                                     "wave_instr_per_cu_cycle_at_2p4GHz": per_cu_cycle}
        log(f"VALU probe {names[kind]}: {t * 1e3:.3f} ms, {per_cu_cycle:.3f} wave-instructions per "
            f"CU-cycle at 2.4 GHz")
    print(json.dumps({"valu_probe": res}))
    return 0


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
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--rows", type=int, default=0, help="rows per lane (skips the autotuner)")
    ap.add_argument("--cols", type=int, default=0, help="columns per lane (skips the autotuner)")
    ap.add_argument("--pulses-per-stage", type=int, default=0)
    ap.add_argument("--range-pulses", type=int, default=4096, help="range-compression leg (0: skip)")
    ap.add_argument("--omega-k", type=int, default=1, help="omega-K leg (0: skip)")
    ap.add_argument("--omega-k-only", action="store_true",
                    help="only run the omega-K leg and print its JSON (profiling)")
    ap.add_argument("--mode-reps", type=int, default=2,
                    help="timed launches of the unplanned entry point and the DF / F64 / F32 modes (0: skip)")
    ap.add_argument("--mixed-jobs", type=int, default=64, help="batched leg jobs (0: skip)")
    ap.add_argument("--mixed-reps", type=int, default=2)
    ap.add_argument("--mixed-sims", type=int, default=4,
                    help="batched jobs given simulated scenes and oracle spot checks (0: none)")
    ap.add_argument("--class-shapes", default="common", choices=("common", "free"),
                    help="per-class tuning of the batched and served legs: the headline's tile shape "
                         "with pulses per stage tuned per class (common), or the autotuner's free choice")
    ap.add_argument("--class-crop-px", type=int, default=1536,
                    help="autotuner crop (pixels per side budget) of the per-class tuning")
    ap.add_argument("--served-jobs", type=int, default=2000, help="served leg jobs (0: skip)")
    ap.add_argument("--served-rates", default="20",
                    help="Poisson arrival rates (jobs/s) of the served stream, comma separated; "
                         "0 = every job submitted at once")
    ap.add_argument("--served-sweep", default="8:400",
                    help="further points of the rate sweep as rate:jobs (a prefix of the stream), comma "
                         "separated")
    ap.add_argument("--e2e-jobs", type=int, default=100,
                    help="jobs of the stream run through the reference host flow (baseline)")
    ap.add_argument("--hbm-budget", type=int, default=200 * 10**9)
    ap.add_argument("--served-raw-jobs", type=int, default=96,
                    help="also serve the first jobs of the stream as raw echoes (burst), "
                         "range-compressed on the device (0: skip)")
    ap.add_argument("--pulses-per-epoch", type=int, default=1024)
    ap.add_argument("--valu-probe", action="store_true",
                    help="only run the vector-ALU issue-rate probes (roofline calibration)")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="only run a device-to-device copy of this many bytes (counter calibration)")
    args = ap.parse_args()

    if args.copy_bytes:
        import torch
        src = torch.ones(args.copy_bytes // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copied {args.copy_bytes} bytes 3 times")
        return 0

    if args.valu_probe:
        return valu_probe(args)

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the BP imager is a single-GPU build "
              "(multi-GPU sharding is a design note only); run with --gpus 1", file=sys.stderr)
        return 2

    import torch
    from sar_cambrianaccelerationcard_amd import device as D
    from sar_cambrianaccelerationcard_amd import metrics as M
    from sar_cambrianaccelerationcard_amd import native
    from sar_cambrianaccelerationcard_amd import scene as S
    from sar_cambrianaccelerationcard_amd import workloads as W

    if not torch.cuda.is_available():
        print("no HIP device", file=sys.stderr)
        return 3
    if args.omega_k_only:
        print(json.dumps({"omega_k": leg_omega_k(torch, native.Context(0), native, args)}))
        return 0
    t0 = time.time()
    cfg = S.load_config(args.config)
    sc = S.make_scene(cfg)
    interp = args.interp or sc.interp
    I, P, Dt = native.INTERP_NAMES[interp], native.PREC_NAMES[args.precision], native.DTYPE_NAMES[args.dtype]
    log(f"scene {cfg['name']} simulated in {time.time() - t0:.1f}s: {sc.n_pulses} pulses x "
        f"{sc.n_samples} samples -> {sc.grid['nx']}x{sc.grid['ny']} ({interp}, {args.precision}, {args.dtype})")

    ctx = native.Context(0)
    ds = D.DeviceScene(sc)
    img = ds.new_image()
    params = sc.native_params(interp=interp, precision=args.precision, dtype=args.dtype)
    stream = torch.cuda.current_stream()
    sp = stream.cuda_stream
    prof_ptr = ds.profiles_ptr(args.dtype)
    tune = None
    if args.rows or args.cols or args.pulses_per_stage:
        d_r, d_c, d_ps = ctx.get_tuning(I, P, Dt)
        ctx.set_tuning(I, P, Dt, args.rows or d_r, args.cols or d_c, args.pulses_per_stage or d_ps)
    elif not args.no_autotune:
        t1 = time.time()
        tune = ctx.autotune(prof_ptr, ds.pos.data_ptr(), sc.pos, sc.native_grid(), params,
                            img.data_ptr(), sc.grid["nx"] * sc.grid["ny"], 4096, sp)
        log(f"autotune ({tune['n_candidates']} candidates on a {tune['crop']} crop, "
            f"{time.time() - t1:.1f}s): {tune['cols_per_lane']}x{tune['rows_per_lane']} px/lane, pulses/stage "
            f"{tune['pulses_per_stage']}: {tune['best_ms']:.2f} ms vs default {tune['default_ms']:.2f} ms")
    plan = ctx.plan(sc.pos, sc.native_grid(), params)
    info = plan.info()

    def step():
        plan.execute(prof_ptr, ds.pos.data_ptr(), img.data_ptr(), sp)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    tw0 = time.perf_counter()
    kern_s = timed(torch, stream, step, args.steps)
    wall = (time.perf_counter() - tw0) / args.steps
    work = sc.work
    wps = work / kern_s
    log(f"plain: {args.steps} steps, {kern_s * 1e3:.2f} ms/step (events), {wall * 1e3:.2f} ms/step "
        f"(wall), {wps:.4e} W/s")

    cpu = acc = None
    if args.cpu_rows > 0:
        cpu, acc = leg_cpu(sc, img.cpu().numpy(), interp, args)
    modes = leg_modes(torch, ctx, native, sc, ds, interp, args, acc) if args.mode_reps > 0 else None
    if acc:
        acc.pop("_ref_rows", None)
    del ds, img
    torch.cuda.empty_cache()

    rc = leg_range(torch, ctx, native, D, args) if args.range_pulses > 0 else None
    wk = leg_omega_k(torch, ctx, native, args) if args.omega_k else None
    # The legs below start from the built-in per-class tunings: the served
    # stream's large jobs (8192^2 from 8192 pulses, 96% of its work) share the
    # headline's size class but not its best tile (the autotuner's pick on the
    # 4096-pulse image can be 1 x 8 px per lane, whose 64 x 32 tiles need
    # 108-sample windows on them); the class default runs them with the stage
    # depth the planner fits to their 84-sample windows, and the mixed leg
    # tunes every class of its own jobs.
    ctx.set_tuning(I, P, Dt, 0, 0, 0)
    head_tuning = {}
    batched = leg_mixed(torch, ctx, native, W, args) if args.mixed_jobs > 0 else None
    served = None
    if args.served_jobs > 0:
        inp = ServedInputs(W, args)
        runs = [leg_served(native, W, args, float(r), inp, len(inp.jobs), ctx=ctx, class_tunings=head_tuning)
                for r in args.served_rates.split(",") if r]
        # the simulated jobs' images of the profile runs, before the raw-echo
        # run (which reuses the first jobs and their image buffers) overwrites them
        sim_images = {k: img.copy() for k, img in inp.images.items()}
        best = max(runs, key=lambda r: r["W_per_s"])
        served = {"workload": f"{len(inp.jobs)} Poisson jobs (10% large: 8192^2 image, 8192 pulses x 16384 "
                              f"samples; others sized as the batched mix), 200 GB HBM budget, sinc8, REF",
                  "runs": runs, "peak_W_per_s": best["W_per_s"],
                  "p50_latency_s": runs[0]["p50_latency_s"], "p99_latency_s": runs[0]["p99_latency_s"]}
        for spec in (r for r in args.served_sweep.split(",") if r):
            # a lower-load point of the rate sweep on a prefix of the same stream
            rate, n = spec.split(":")
            runs.append(leg_served(native, W, args, float(rate), inp, min(int(n), len(inp.jobs)), ctx=ctx,
                                   class_tunings=head_tuning))
        if args.served_raw_jobs > 0:
            n_raw = min(args.served_raw_jobs, len(inp.jobs))
            raw_inputs = served_raw_inputs(W, inp.jobs[:n_raw])
            served["raw_echo_burst"] = leg_served(native, W, args, 0.0, inp, n_raw, raw_inputs, ctx=ctx,
                                                  class_tunings=head_tuning)
            del raw_inputs
        e2e = plain_e2e(torch, ctx, native, inp, min(args.e2e_jobs, len(inp.jobs)), class_tunings=head_tuning)
        served["plain_end_to_end"] = e2e
        served["peak_speedup_vs_plain_end_to_end"] = served["peak_W_per_s"] / e2e["W_per_s"]
        if batched:
            served["peak_speedup_vs_plain_sequential"] = served["peak_W_per_s"] / batched["plain_sequential_W_per_s"]
        if inp.sims:
            checks = served_spot_checks(inp, sim_images)
            served["spot_checks"] = checks
            served["spot_check_rel_l2"] = [c["rel_l2"] for c in checks]
            served["spot_check_max_rel_l2"] = max(c["rel_l2"] for c in checks)
        del inp

    fw = F_W[interp]
    achieved = wps * fw / 1e12
    traffic = traffic_tuning = None
    tpath = os.path.join(ROOT, "profiles", "bp_sinc8_4096x4096_4096pulses_traffic.json")
    if os.path.exists(tpath):
        t = json.load(open(tpath))
        if (t["workload"], t["interp"], t["precision"], t["dtype"]) == \
                (cfg["name"], interp, args.precision, args.dtype):
            # the counters of the launch the headline ran (its tile shape and
            # stage depth set the re-fetch), if they were measured
            planned = {k: info[k] for k in ("cols_per_lane", "rows_per_lane", "pulses_per_stage")}
            for launch in t["launches"]:
                if launch["tuning"] == planned:
                    traffic, traffic_tuning = launch["traffic_bytes"], launch["tuning"]
    line = {
        "metric": "pixel·pulse back-projections/s at rel-L2≤1e-4 vs FP64; p50/p99 job latency",
        "value": wps,
        "unit": "pixel-pulses/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": kern_s * 1e3,
        "wall_ms_per_step": wall * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (FP64 tile reference range, FP32 differential, FP32 accumulation)"
        if args.precision == "ref" else args.precision,
        "data": "synthetic (seeded simulated scene)",
        "config": {"workload": f"{cfg['name']}: {sc.n_pulses} pulses x {sc.n_samples} samples "
                               f"{args.dtype} -> {sc.grid['nx']}x{sc.grid['ny']} image at "
                               f"{sc.grid['ux'][0]} m, {interp}, precision {args.precision}",
                   "n_pulses": sc.n_pulses, "n_samples": sc.n_samples, "nx": sc.grid["nx"],
                   "ny": sc.grid["ny"], "interp": interp, "precision": args.precision,
                   "profile_dtype": args.dtype, "plan": info, "autotune": tune},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_OPS / 1e12,
                     "unit": "Top/s", "frac": achieved / (PEAK_OPS / 1e12),
                     "ops_per_W": fw, "traffic": traffic, "traffic_tuning": traffic_tuning,
                     "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE x 2.0 + WRITE_SIZE, "
                                     "copy-calibrated; profiles/)",
                     "flop_convention": {"achieved_tflops": wps * FLOP_W[interp] / 1e12,
                                         "peak_tflops": PEAK_TFLOPS,
                                         "frac": wps * FLOP_W[interp] / 1e12 / PEAK_TFLOPS}},
        "cpu_baseline": cpu,
        "accuracy": acc,
        "other_paths": modes,
        "range_compression": rc,
        "omega_k": wk,
        "batched": batched,
        "served": served,
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
