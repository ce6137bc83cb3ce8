"""FP64 range compression: FFT matched filter and upsampling by spectral zero
insertion -- the checker of the device range compression.

TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may use it; the product package never imports it.

Restates code/SAR_BP/matlab/BPA_SAR.m (reference repository):
  :93-94   reference chirp starting at the window start t0
  :95-100  Srnm = IFFT(FFT(echo) . conj(FFT(h)))        (matched filter)
  :106-111 zero insertion after bin floor((N+1)/2) of the spectrum, IFFT over
           L*N (upsampling)
with the same three documented deviations as the simulator: the chirp is
centred at baseband, exp(j pi Kr (t - Tp/2)^2) for 0 < t < Tp (the script's
chirp sweeps [0, B] and only fits because it samples at Fr = 2 B,
BPA_SAR.m:30,87); the FFT length nf is the smallest power of two >= n_raw +
ceil(Tp fs), so the correlation is linear; the upsampled IFFT is scaled by 1/nf
(not 1/(L nf), BPA_SAR.m:111), keeping the compressed amplitude.

FFTs are numpy's (pocketfft, FP64).  The tests pin this module against a
direct O(N^2) correlation sum, a direct evaluation of the band-limited
interpolating trigonometric polynomial, the decimation identity
out_L[L q] = out_1[q], the closed-form compressed peak of a point target, and
the simulator's independent radix-2 implementation.
"""
from __future__ import annotations

import math

import numpy as np


def fft_size(n_raw: int, pulse_width: float, fs: float) -> int:
    need = int(n_raw) + int(math.ceil(pulse_width * fs))
    n = 1
    while n < need:
        n <<= 1
    return n


def reference_chirp(nf: int, n_raw: int, bandwidth: float, pulse_width: float, fs: float) -> np.ndarray:
    """h[n] = exp(j 2 pi frac(Kr/2 (n/fs - Tp/2)^2)) for 0 < n/fs < Tp and n < n_raw."""
    kr = bandwidth / pulse_width
    n = np.arange(nf)
    t = n / fs
    tc = t - 0.5 * pulse_width
    ph = 0.5 * kr * tc * tc
    ph = ph - np.floor(ph)
    h = np.exp(2j * np.pi * ph)
    h[~((t > 0.0) & (t < pulse_width) & (n < n_raw))] = 0.0
    return h


def range_axis(c: float, fs: float, t0: float, upsample: int, crop_offset: int) -> tuple:
    """(r0, dr) of the output samples."""
    L = max(1, int(upsample))
    return 0.5 * c * (t0 + crop_offset / (L * fs)), 0.5 * c / (L * fs)


def range_compress(raw: np.ndarray, *, bandwidth: float, pulse_width: float, fs: float,
                   upsample: int, crop_offset: int, n_out: int, chunk: int = 128,
                   **_unused) -> np.ndarray:
    """raw [n_pulses][n_raw] -> [n_pulses][n_out] complex128."""
    raw = np.asarray(raw)
    M, n_raw = raw.shape
    nf = fft_size(n_raw, pulse_width, fs)
    L = int(upsample)
    nu = L * nf
    if crop_offset < 0 or crop_offset + n_out > nu:
        raise ValueError("crop window outside the upsampled transform")
    hc = np.conj(np.fft.fft(reference_chirp(nf, n_raw, bandwidth, pulse_width, fs)))
    half = (nf + 1) // 2
    out = np.empty((M, n_out), dtype=np.complex128)
    for m0 in range(0, M, chunk):
        blk = raw[m0:m0 + chunk].astype(np.complex128)
        Y = np.fft.fft(blk, n=nf, axis=1) * hc
        Z = np.zeros((blk.shape[0], nu), dtype=np.complex128)
        Z[:, :half] = Y[:, :half]
        Z[:, half + nu - nf:] = Y[:, half:]
        x = np.fft.ifft(Z, axis=1) * L        # unscaled IFFT / nf
        out[m0:m0 + chunk] = x[:, crop_offset:crop_offset + n_out]
    return out
