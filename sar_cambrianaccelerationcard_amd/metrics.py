"""Image-quality metrics: relative L2 error, peak location, phase error and
PSLR / ISLR of point targets.

The reference only draws a -3 dB contour of the image
(code/SAR_BP/matlab/BPA_SAR.m:168-175, :213-220); these are the quantitative
counterparts.  PSLR / ISLR follow the usual definitions on range and azimuth
cuts through the peak of an 8x FFT-upsampled chip: the main lobe spans the
first nulls either side of the peak; PSLR = highest sidelobe / peak power,
ISLR = sidelobe energy / main-lobe energy, both in dB.
"""
from __future__ import annotations

import numpy as np


def rel_l2(img: np.ndarray, ref: np.ndarray) -> float:
    ref = np.asarray(ref, dtype=np.complex128)
    err = np.asarray(img, dtype=np.complex128) - ref
    den = np.linalg.norm(ref.ravel())
    return float(np.linalg.norm(err.ravel()) / den) if den > 0 else float(np.linalg.norm(err.ravel()))


def local_peak(img: np.ndarray, ix: int, iy: int, radius: int = 3) -> tuple:
    """(ix, iy) of the largest |img| within +-radius of (ix, iy)."""
    ny, nx = img.shape
    y0, y1 = max(0, iy - radius), min(ny, iy + radius + 1)
    x0, x1 = max(0, ix - radius), min(nx, ix + radius + 1)
    sub = np.abs(img[y0:y1, x0:x1])
    j, i = np.unravel_index(int(np.argmax(sub)), sub.shape)
    return x0 + int(i), y0 + int(j)


def upsample_chip(img: np.ndarray, ix: int, iy: int, half: int = 32, factor: int = 8) -> np.ndarray:
    ny, nx = img.shape
    chip = np.zeros((2 * half, 2 * half), dtype=np.complex128)
    y0, x0 = iy - half, ix - half
    ys, xs = max(0, y0), max(0, x0)
    ye, xe = min(ny, y0 + 2 * half), min(nx, x0 + 2 * half)
    chip[ys - y0:ye - y0, xs - x0:xe - x0] = img[ys:ye, xs:xe]
    n = 2 * half
    F = np.fft.fftshift(np.fft.fft2(chip))
    m = n * factor
    G = np.zeros((m, m), dtype=np.complex128)
    o = (m - n) // 2
    G[o:o + n, o:o + n] = F
    return np.fft.ifft2(np.fft.ifftshift(G)) * factor * factor


def _cut_metrics(p: np.ndarray) -> tuple:
    """PSLR, ISLR (dB) of a 1-D power cut."""
    k = int(np.argmax(p))
    l = k
    while l > 0 and p[l - 1] < p[l]:
        l -= 1
    r = k
    while r < len(p) - 1 and p[r + 1] < p[r]:
        r += 1
    main = p[l:r + 1].sum()
    side = np.concatenate([p[:l], p[r + 1:]])
    if side.size == 0 or main <= 0:
        return float("-inf"), float("-inf")
    pslr = 10 * np.log10(side.max() / p[k])
    islr = 10 * np.log10(side.sum() / main)
    return float(pslr), float(islr)


def point_target_quality(img: np.ndarray, ix: int, iy: int, half: int = 32, factor: int = 8) -> dict:
    """PSLR / ISLR along x and y cuts through the upsampled peak near (ix, iy)."""
    up = upsample_chip(img, ix, iy, half, factor)
    pw = np.abs(up) ** 2
    j, i = np.unravel_index(int(np.argmax(pw)), pw.shape)
    px, ix_ = _cut_metrics(pw[j, :])
    py, iy_ = _cut_metrics(pw[:, i])
    return {"pslr_x_db": px, "islr_x_db": ix_, "pslr_y_db": py, "islr_y_db": iy_}


def compare_targets(img: np.ndarray, ref: np.ndarray, target_pixels: np.ndarray,
                    radius: int = 3, quality: bool = True) -> list:
    """Per-target comparison of an image against the reference image."""
    out = []
    for tx, ty in np.asarray(target_pixels):
        rx, ry = local_peak(ref, int(tx), int(ty), radius)
        gx, gy = local_peak(img, int(tx), int(ty), radius)
        ph = np.angle(img[ry, rx] * np.conj(ref[ry, rx]), deg=True)
        rec = {"truth": (int(tx), int(ty)), "ref_peak": (rx, ry), "peak": (gx, gy),
               "peak_match": (rx, ry) == (gx, gy), "phase_err_deg": float(ph)}
        if quality:
            qr = point_target_quality(ref, rx, ry)
            qi = point_target_quality(img, rx, ry)
            rec["quality_ref"] = qr
            rec["quality"] = qi
            rec["max_quality_diff_db"] = float(max(abs(qi[k] - qr[k]) for k in qr))
        out.append(rec)
    return out
