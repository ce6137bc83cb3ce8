#!/usr/bin/env python3
"""Writes tests/golden/bp_tiny.npz: brute-force back-projections of a tiny
seeded scene evaluated with mpmath at 50 significant digits.

The formulation is deliberately independent of the oracle's tap logic:
  nearest : s[round(u)]                 (MATLAB round, halves away from zero,
                                          code/SAR_BP/matlab/BPA_SAR.m:141)
  linear  : sum_k max(0, 1 - |u - k|) s[k]   (triangle kernel, every k)
  sinc8   : sum_k K(u - k) s[k] over every k, K(x) = sinc(x) I0(b sqrt(1-(x/4)^2))/I0(b)
            for |x| < 4, else 0
and the pixel value is sum_m interp_m(u_m) exp(+j 4 pi fc R_m / c)
(BPA_SAR.m:148-150), R_m = |antenna_m - pixel|, u_m = (R_m - r0) / dr.

Run:  python tests/golden/make_golden.py     (about a minute)
"""
import os

import mpmath as mp
import numpy as np

mp.mp.dps = 50
HERE = os.path.dirname(os.path.abspath(__file__))


def scene():
    rng = np.random.default_rng(20260101)
    n_p, n_s = 4, 64
    prof = (rng.standard_normal((n_p, n_s)) + 1j * rng.standard_normal((n_p, n_s))).astype(np.complex64)
    pos = np.stack([np.linspace(-3.0, 3.0, n_p) + rng.uniform(-0.1, 0.1, n_p),
                    rng.uniform(-0.2, 0.2, n_p),
                    600.0 + rng.uniform(-0.3, 0.3, n_p)], 1)
    grid = {"origin": [-0.8, 800.0, 0.0], "ux": [0.31, 0.02, 0.0], "uy": [-0.015, 0.27, 0.01],
            "nx": 6, "ny": 5}
    radar = {"c": 299792458.0, "fc": 9.6e9, "dr": 0.1874, "r0": 994.5}
    return prof, pos, grid, radar


def kaiser_sinc(x, beta):
    if abs(x) >= 4:
        return mp.mpf(0)
    s = mp.mpf(1) if x == 0 else mp.sin(mp.pi * x) / (mp.pi * x)
    return s * mp.besseli(0, beta * mp.sqrt(1 - (x / 4) ** 2)) / mp.besseli(0, beta)


def interp(row, u, kind, beta):
    n = len(row)
    if kind == "nearest":
        k = int(mp.floor(u + mp.mpf("0.5"))) if u >= 0 else -int(mp.floor(-u + mp.mpf("0.5")))
        return row[k] if 0 <= k < n else mp.mpc(0)
    acc = mp.mpc(0)
    for k in range(n):
        x = u - k
        w = max(mp.mpf(0), 1 - abs(x)) if kind == "linear" else kaiser_sinc(x, beta)
        if w != 0:
            acc += w * row[k]
    return acc


def backproject(prof, pos, grid, radar, kind, beta):
    c, fc, dr, r0 = (mp.mpf(repr(radar[k])) for k in ("c", "fc", "dr", "r0"))
    rows = [[mp.mpc(float(v.real), float(v.imag)) for v in r] for r in prof]
    P = [[mp.mpf(repr(float(v))) for v in p] for p in pos]
    o, ux, uy = ([mp.mpf(repr(float(v))) for v in grid[k]] for k in ("origin", "ux", "uy"))
    img = np.zeros((grid["ny"], grid["nx"]), dtype=np.complex128)
    for iy in range(grid["ny"]):
        for ix in range(grid["nx"]):
            px = [o[i] + ix * ux[i] + iy * uy[i] for i in range(3)]
            acc = mp.mpc(0)
            for m, a in enumerate(P):
                R = mp.sqrt(sum((a[i] - px[i]) ** 2 for i in range(3)))
                u = (R - r0) / dr
                acc += interp(rows[m], u, kind, beta) * mp.expj(4 * mp.pi * fc * R / c)
            img[iy, ix] = complex(acc)
    return img


def main():
    prof, pos, grid, radar = scene()
    out = {"profiles": prof, "pos": pos, "origin": np.array(grid["origin"]),
           "ux": np.array(grid["ux"]), "uy": np.array(grid["uy"]),
           "nx": np.int64(grid["nx"]), "ny": np.int64(grid["ny"]),
           "c": np.float64(radar["c"]), "fc": np.float64(radar["fc"]),
           "dr": np.float64(radar["dr"]), "r0": np.float64(radar["r0"])}
    for kind, beta in [("nearest", 0.0), ("linear", 0.0), ("sinc8", 6.0), ("sinc8", 2.5)]:
        key = f"img_{kind}" + (f"_b{beta:g}" if kind == "sinc8" else "")
        out[key] = backproject(prof, pos, grid, radar, kind, mp.mpf(repr(beta)))
        print(key, np.abs(out[key]).max())
    # Kaiser-sinc kernel and I0 samples
    xs = np.linspace(-4.0, 4.0, 161)
    out["ks_x"] = xs
    out["ks_b6"] = np.array([float(kaiser_sinc(mp.mpf(repr(float(x))), mp.mpf(6))) for x in xs])
    bx = np.array([0.0, 0.5, 1.0, 2.5, 6.0, 9.5, 20.0])
    out["i0_x"] = bx
    out["i0"] = np.array([float(mp.besseli(0, mp.mpf(repr(float(x))))) for x in bx])
    np.savez(os.path.join(HERE, "bp_tiny.npz"), **out)
    print("wrote", os.path.join(HERE, "bp_tiny.npz"))


if __name__ == "__main__":
    main()
