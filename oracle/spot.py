"""Oracle spot checks of images too large for the FP64 oracle in full.

TEST INFRASTRUCTURE ONLY (tests/ and the accuracy legs of bench.py): compares
an image against the FP64 CPU oracle (sarbp_oracle.c, the BP loops of
code/SAR_BP/matlab/BPA_SAR.m:133-153 / :178-197 of the reference repository)
on whole image rows and on 64 x 64 chips centred on point targets -- the
config-4 spot-check rule: rel-L2 over target windows plus a fixed row sample.
"""
from __future__ import annotations

import numpy as np

try:  # imported as oracle.spot (bench.py)
    from . import oracle as O
except ImportError:  # oracle/ on sys.path (tests)
    import oracle as O


def rel_l2(img, ref) -> float:
    ref = np.asarray(ref, dtype=np.complex128)
    d = np.linalg.norm(ref.ravel())
    return float(np.linalg.norm((np.asarray(img, dtype=np.complex128) - ref).ravel()) / d)


def chip_grid(grid: dict, ix0: int, iy0: int, n: int) -> dict:
    """Grid of the n x n sub-image whose pixel (0, 0) is pixel (ix0, iy0) of grid."""
    o = [grid["origin"][k] + ix0 * grid["ux"][k] + iy0 * grid["uy"][k] for k in range(3)]
    return {"origin": o, "ux": list(grid["ux"]), "uy": list(grid["uy"]), "nx": n, "ny": n}


def spot_check(profiles, pos, grid: dict, p, image: np.ndarray, rows, chips, half: int = 32,
               n_threads: int = 0) -> dict:
    """rel-L2 of `image` against the oracle on each row in `rows` and on each
    2half x 2half chip centred on a pixel in `chips` ((ix, iy) pairs, clipped
    to the image), the combined rel-L2 of all of them, and whether each chip's
    brightest pixel is the oracle's."""
    nx, ny = int(grid["nx"]), int(grid["ny"])
    num = den = 0.0
    out = {"rows": [], "row_rel_l2": [], "chips": [], "chip_rel_l2": [], "chip_peak_match": []}
    for r in rows:
        ref = O.backproject(profiles, pos, grid, p, rows=(int(r), int(r) + 1), n_threads=n_threads)
        got = image[int(r):int(r) + 1]
        out["rows"].append(int(r))
        out["row_rel_l2"].append(rel_l2(got, ref))
        num += float(np.sum(np.abs(got.astype(np.complex128) - ref) ** 2))
        den += float(np.sum(np.abs(ref) ** 2))
    for ix, iy in chips:
        n = 2 * half
        x0 = int(min(max(int(ix) - half, 0), nx - n))
        y0 = int(min(max(int(iy) - half, 0), ny - n))
        ref = O.backproject(profiles, pos, chip_grid(grid, x0, y0, n), p, n_threads=n_threads)
        got = image[y0:y0 + n, x0:x0 + n]
        out["chips"].append([x0, y0])
        out["chip_rel_l2"].append(rel_l2(got, ref))
        out["chip_peak_match"].append(bool(np.argmax(np.abs(got)) == np.argmax(np.abs(ref))))
        num += float(np.sum(np.abs(got.astype(np.complex128) - ref) ** 2))
        den += float(np.sum(np.abs(ref) ** 2))
    out["rel_l2"] = float(np.sqrt(num / den)) if den > 0 else float("nan")
    return out
