"""Image quality report for a complex image file: peak, PSLR / ISLR of point
targets, optional comparison with a reference image, and a dB PNG.

    python -m sar_cambrianaccelerationcard_amd.quality IMAGE.bin \\
        [--targets X,Y ...] [--auto N] [--ref REF.bin] [--png OUT.png] [--floor-db -40]

IMAGE.bin is raw interleaved complex64, row-major [ny][nx] (what the
reference's SAR_Processor::saveResult writes, code/SAR_BP/CUDA/BPA_CUDA.cpp:92-100),
with the shape from its JSON sidecar or --shape NY,NX.  The reference only
draws a -3 dB contour (code/SAR_BP/matlab/BPA_SAR.m:168-175); this prints the
numbers.  Prints one JSON object.
"""
from __future__ import annotations

import argparse
import json
import sys

import numpy as np

from . import imageio as io
from . import metrics as M


def brightest(img: np.ndarray, n: int, guard: int = 16) -> list:
    """The n brightest pixels at least `guard` pixels apart (and from the edges)."""
    a = np.abs(img).astype(np.float64).copy()
    ny, nx = a.shape
    a[:guard, :] = a[-guard:, :] = 0
    a[:, :guard] = a[:, -guard:] = 0
    out = []
    for _ in range(n):
        k = int(np.argmax(a))
        y, x = divmod(k, nx)
        if a[y, x] <= 0:
            break
        out.append((x, y))
        a[max(0, y - guard):y + guard + 1, max(0, x - guard):x + guard + 1] = 0
    return out


def report(img: np.ndarray, targets: list, ref: np.ndarray | None = None) -> dict:
    a = np.abs(img)
    pk = int(np.argmax(a))
    out = {"shape": list(img.shape), "peak": {"x": pk % img.shape[1], "y": pk // img.shape[1],
                                              "abs": float(a.ravel()[pk])},
           "mean_power": float(np.mean(a.astype(np.float64) ** 2)), "targets": []}
    for x, y in targets:
        gx, gy = M.local_peak(img, int(x), int(y), 3)
        rec = {"at": [int(x), int(y)], "peak": [gx, gy], "abs": float(a[gy, gx])}
        rec.update(M.point_target_quality(img, gx, gy))
        out["targets"].append(rec)
    if ref is not None:
        out["rel_l2_vs_ref"] = M.rel_l2(img, ref)
        if targets:
            cmp = M.compare_targets(img, ref, np.array(targets))
            out["vs_ref"] = [{"peak_match": c["peak_match"], "phase_err_deg": c["phase_err_deg"],
                              "max_quality_diff_db": c["max_quality_diff_db"]} for c in cmp]
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("image")
    ap.add_argument("--shape", help="NY,NX when the image has no sidecar")
    ap.add_argument("--targets", nargs="*", default=[], help="X,Y pixel positions")
    ap.add_argument("--auto", type=int, default=0, help="also report the N brightest points")
    ap.add_argument("--ref", help="reference image (same format) to compare against")
    ap.add_argument("--png", help="write |image| in dB as PNG")
    ap.add_argument("--floor-db", type=float, default=-40.0)
    args = ap.parse_args(argv)
    shape = tuple(int(v) for v in args.shape.split(",")) if args.shape else None
    img = io.load_image(args.image, shape)
    ref = io.load_image(args.ref, img.shape) if args.ref else None
    targets = [tuple(int(v) for v in t.split(",")) for t in args.targets]
    if args.auto:
        targets += brightest(img, args.auto)
    out = report(img, targets, ref)
    if args.png:
        io.save_png(args.png, img, args.floor_db)
        out["png"] = args.png
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
