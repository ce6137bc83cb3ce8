"""FP64 restatement of the reference's point-target echo simulator -- the
oracle of sarbp_launch_generate_echo.

TEST INFRASTRUCTURE ONLY: tests/ may use it as the checker; the product
package never imports it.

Follows generateEchoKernel (code/SAR_BP/CUDA/BPA_CUDA_kernels.cu:3-45) and the
inputs its host driver builds (code/SAR_BP/CUDA/BPA_CUDA.cpp:25-48):
  sample (i, j) = sum over targets k of sigma_k exp(j (pi Kr t^2 - 4 pi R / lambda))
  for the k with 0 < t < Tr and |Az| < La/2, where
  Az = ta[i] vr - x_k, R = sqrt(Az^2 + y_k^2 + H^2), t = tr[j] - 2 R / c.
The float32 inputs (Ptarget, ta, tr and the SARParams fields) are taken as
their exact FP64 values and every step is FP64 (the reference computes in
FP32, whose phase error at 4 pi R / lambda ~ 1e6 rad is ~0.1 rad).  The range
and gate are evaluated in the order written above, one rounding per
operation.
"""
from __future__ import annotations

import numpy as np


def reference_scene(p) -> tuple:
    """(Ptarget [5][3], ta [Mslow], tr [Nfast]) as float32, the values
    SAR_Processor::initializeArrays builds (BPA_CUDA.cpp:25-48) for SARParams
    p (fields c, Kr, Tr, Fr, vr, H, La, Rg0, Az0, Nfast, Mslow)."""
    f = np.float32
    az0, rg0 = f(p.Az0), f(p.Rg0)
    pt = np.array([[az0 - f(10), rg0 - f(20), f(1.0)],
                   [az0 + f(20), rg0 + f(30), f(0.8)],
                   [az0 - f(30), rg0 + f(10), f(1.2)],
                   [az0 + f(40), rg0 - f(40), f(0.9)],
                   [az0, rg0, f(1.5)]], dtype=np.float32)
    la, vr = f(p.La), f(p.vr)
    ta_start = -la / (f(2) * vr)
    ta_end = la / (f(2) * vr)
    ta_step = (ta_end - ta_start) / f(p.Mslow - 1)
    ta = np.array([ta_start + f(i) * ta_step for i in range(p.Mslow)], dtype=np.float32)
    tr_step = f(1.0) / f(p.Fr)
    tr = np.array([f(i) * tr_step for i in range(p.Nfast)], dtype=np.float32)
    return pt, ta, tr


def generate_echo(ptarget: np.ndarray, ta: np.ndarray, tr: np.ndarray, p) -> np.ndarray:
    """Srnm [len(ta)][len(tr)] complex128 (generateEchoKernel, BPA_CUDA_kernels.cu:3-45)."""
    pt = np.asarray(ptarget, dtype=np.float32).reshape(-1, 3)[: p.Ntarget].astype(np.float64)
    ta = np.asarray(ta, dtype=np.float32).astype(np.float64)
    tr = np.asarray(tr, dtype=np.float32).astype(np.float64)
    c, lam, kr, tw = (float(np.float32(v)) for v in (p.c, p.lambda_, p.Kr, p.Tr))
    vr, h, la = (float(np.float32(v)) for v in (p.vr, p.H, p.La))
    out = np.zeros((ta.size, tr.size), dtype=np.complex128)
    for x, y, sigma in pt:
        az = ta * vr - x                                     # [Mslow]
        r = np.sqrt(az * az + y * y + h * h)                 # [Mslow]
        t = tr[None, :] - (2.0 * r / c)[:, None]             # [Mslow][Nfast]
        gate = (t > 0.0) & (t < tw) & (np.abs(az) < la / 2.0)[:, None]
        # phase in revolutions, reduced before the trigonometric functions
        rev = 0.5 * kr * t * t - (2.0 * r / lam)[:, None]
        rev -= np.rint(rev)
        out += np.where(gate, sigma * np.exp(2j * np.pi * rev), 0.0)
    return out
