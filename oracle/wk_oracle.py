"""FP64 omega-K (wavenumber-domain) imaging with truncated-sinc Stolt
interpolation -- the oracle of the device omega-K path.

TEST INFRASTRUCTURE ONLY: tests/ and bench.py's cpu_baseline leg may use it;
the product package never imports it.

Follows the reference's omega-K pipeline, code/SAR_WK/wk_v5.2.3 (process_wk,
wk.cc:43-343) and its MATLAB prototype code/SAR_WK/matlab/mian_wka_sinc.m:
  padded_size   preprocess.cc:21-24 / mian_wka_sinc.m:45 (round(Na/6), round(Nr/3)
                zeros before and after)
  axes          wk.cc:197-221 / mian_wka_sinc.m:54-62: fr = fftshift(-Nr/2:Nr/2-1) Fr/Nr,
                fa = f_nc + fftshift(-Na/2:Na/2-1) PRF/Na; ta = (-Na/2:Na/2-1)/PRF
  downconvert   preprocess.cc:37-129 / mian_wka_sinc.m:70-71: line i times
                exp(-j 2 pi f_nc ta_i)
  reference     compression.cc:17-58 / mian_wka_sinc.m:80-84: S times exp(j theta),
                theta = 4 pi R0/c sqrt(max(0, (f0+fr)^2 - c^2 fa^2/(4 Vr^2))) + pi fr^2/Kr
                with Kr the signed chirp rate (-7.2135e11 Hz/s in the sample
                paras.txt; the MATLAB script negates the positive .mat value,
                mian_wka_sinc.m:29, and the C++ negates twice, wk.cc:229 and
                compression.cc:30, so the two agree)
  stolt         fr_new = sqrt((f0+fr)^2 - c^2 fa^2/(4 Vr^2)) - f0 (stolt.cc:19-35),
                centre c_n = n - (fr_new_n - fr_n) Nr/Fr (stolt.cc:447-505); output
                n = sum over taps idx of in[idx] sinc(c_n - idx), with the taps of
                  window "floor"    (stolt.cc:522-583): idx = floor(c_n + p),
                                    p = -P/2 .. P/2-1, each clamped to [0, Nr-1],
                                    the distance taken to the clamped index
                  window "centered" (mian_wka_sinc.m:98-111): every integer idx in
                                    [0, Nr-1] with |c_n - idx| <= P/2
                sinc(x) = sin(pi x)/(pi x), sinc(0) = 1 (MATLAB's sinc; the MLU
                code's DivNoNan gives 0 there, a defect not reproduced).  Centres
                are exact integers only in constructed cases; then, near the
                row ends, each clamped copy of an end sample at distance 0
                counts with weight 1.
  ifft2         fft.cc:200-359 (1/Nr, 1/Na scales) / mian_wka_sinc.m:119
Rows are azimuth (Na), columns range (Nr), C order.  The display shifts and
flip of mian_wka_sinc.m:124-126 are left to the caller.
"""
from __future__ import annotations

import numpy as np

FLOOR, CENTERED = 0, 1


def padded_size(na_in: int, nr_in: int) -> tuple:
    """(Na, Nr, pad_az, pad_rg): zeros before and after = round(n/6), round(n/3), half away from zero."""
    pa = int(np.floor(na_in / 6.0 + 0.5))
    pr = int(np.floor(nr_in / 3.0 + 0.5))
    return na_in + 2 * pa, nr_in + 2 * pr, pa, pr


def fftshift_axis(n: int) -> np.ndarray:
    """MATLAB fftshift(-n/2 : n/2-1)."""
    return np.fft.fftshift(-n / 2.0 + np.arange(n))


def axes(na: int, nr: int, fs: float, prf: float, f_nc: float) -> tuple:
    fr = fftshift_axis(nr) * (fs / nr)
    fa = f_nc + fftshift_axis(na) * (prf / na)
    ta = (-na / 2.0 + np.arange(na)) / prf
    return fr, fa, ta


def centres(fr_axis: np.ndarray, fa_axis: np.ndarray, f0: float, c: float, vr: float, fs: float) -> np.ndarray:
    """c[m, n] = n - (fr_new[m, n] - fr[n]) Nr / Fr, evaluated as n - (s - f) Nr/Fr with
    f = f0 + fr and s = sqrt(max(f^2 - c^2 fa^2 / (4 Vr^2), 0)) (fr_new - fr = s - f): at fa = 0
    the centres are then exactly the integers n, where the floor window's taps jump."""
    nr = fr_axis.size
    k = c * c / (4.0 * vr * vr)
    f = f0 + fr_axis[None, :]
    s = np.sqrt(np.maximum(f * f - k * fa_axis[:, None] ** 2, 0.0))
    return np.arange(nr)[None, :] - (s - f) * (nr / fs)


def stolt_rows(S: np.ndarray, cen: np.ndarray, taps: int = 6, window: int = FLOOR) -> np.ndarray:
    """Sinc interpolation of each row of S at the positions cen (same shape)."""
    na, nr = S.shape
    half = taps // 2
    fl = np.floor(cen)
    rows = np.arange(na)[:, None]
    acc = np.zeros(S.shape, dtype=np.complex128)
    if window == FLOOR:
        for p in range(-half, half):
            idx = np.clip(fl + p, 0, nr - 1)
            acc += S[rows, idx.astype(np.int64)] * np.sinc(cen - idx)
    elif window == CENTERED:
        for q in range(-half, half + 2):
            idx = fl + q
            x = cen - idx
            ok = (np.abs(x) <= half) & (idx >= 0) & (idx <= nr - 1)
            ii = np.clip(idx, 0, nr - 1).astype(np.int64)
            acc += np.where(ok, S[rows, ii] * np.sinc(x), 0.0)
    else:
        raise ValueError("window")
    return acc


def stolt_sinc(S: np.ndarray, fr_axis, fa_axis, *, f0: float, c: float, vr: float, fs: float,
               taps: int = 6, window: int = FLOOR, chunk: int = 256) -> np.ndarray:
    """The reference's StoltInterp_sinc (stolt.cc:324) on [Na][Nr] S with explicit axes."""
    S = np.asarray(S, dtype=np.complex128)
    fr_axis = np.asarray(fr_axis, dtype=np.float64)
    fa_axis = np.asarray(fa_axis, dtype=np.float64)
    out = np.empty_like(S)
    for m0 in range(0, S.shape[0], chunk):
        cen = centres(fr_axis, fa_axis[m0:m0 + chunk], f0, c, vr, fs)
        out[m0:m0 + chunk] = stolt_rows(S[m0:m0 + chunk], cen, taps, window)
    return out


def stolt_gridsample(S: np.ndarray, fr_axis, fa_axis, *, f0: float, c: float, vr: float,
                     fr: float) -> np.ndarray:
    """The reference's GridSample Stolt step, performStoltInterpolation
    (task/stolt.cc:38-318, declared task/task.h:27-34; its call in process_wk,
    wk.cc:272, is commented out), on [Na][Nr] S:
      fr_new[m, n] = sqrt((f0 + fr_axis[n])^2 - c^2 fa[m]^2 / (4 Vr^2)) - f0,
                     FP64, stored as float32 (calculateNewFrequencyMapping,
                     stolt.cc:19-35);
      row m is rotated by mid = Nr/2: S'[m, w] = S[m, (w + Nr - mid) mod Nr]
                     (stolt.cc:121-130);
      x[m, n] = float32(fr_new[m, n] / (fr Nr / 2))  (the grid, :199-206);
      bilinear grid sampling with align_corners and zero padding on S'
                     (:73-78, :246-272), the row coordinate being row m itself:
      ix = ((x + 1) / 2) (Nr - 1) in float32, i0 = floor(ix), t = ix - i0,
      out[m, n] = (1 - t) S'[m, i0] + t S'[m, i0 + 1], samples outside the
      row are zero.  The output is not rotated back.  fr is the bin spacing
      the caller passes (fr_gap = Fr / Nr in wk.cc:272)."""
    S = np.asarray(S, dtype=np.complex128)
    na, nr = S.shape
    fr_axis = np.asarray(fr_axis, dtype=np.float64)
    fa_axis = np.asarray(fa_axis, dtype=np.float64)
    k = c * c / (4.0 * vr * vr)
    f = f0 + fr_axis[None, :]
    fr_new = (np.sqrt(f * f - k * fa_axis[:, None] ** 2) - f0).astype(np.float32)
    x = (fr_new.astype(np.float64) / (fr * nr * 0.5)).astype(np.float32)
    one, two = np.float32(1.0), np.float32(2.0)
    ix = (((x + one) / two) * np.float32(nr - 1)).astype(np.float32)
    i0 = np.floor(ix)
    t = (ix - i0).astype(np.float64)
    i0 = i0.astype(np.int64)
    shifted = np.roll(S, nr // 2, axis=1)
    rows = np.arange(na)[:, None]

    def tap(i):
        ok = (i >= 0) & (i < nr)
        return np.where(ok, shifted[rows, np.clip(i, 0, nr - 1)], 0.0)

    return (1.0 - t) * tap(i0) + t * tap(i0 + 1)


def reference_phase(fr_axis, fa_axis, *, f0: float, c: float, vr: float, kr: float, r0: float) -> np.ndarray:
    k = c * c / (4.0 * vr * vr)
    s = np.sqrt(np.maximum((f0 + fr_axis[None, :]) ** 2 - k * fa_axis[:, None] ** 2, 0.0))
    return 4.0 * np.pi * r0 / c * s + np.pi / kr * fr_axis[None, :] ** 2


def pad_downconvert(echo: np.ndarray, *, prf: float, f_nc: float) -> np.ndarray:
    na_in, nr_in = echo.shape
    na, nr, pa, pr = padded_size(na_in, nr_in)
    x = np.zeros((na, nr), dtype=np.complex128)
    x[pa:pa + na_in, pr:pr + nr_in] = echo
    ta = (-na / 2.0 + np.arange(na)) / prf
    return x * np.exp(-2j * np.pi * f_nc * ta)[:, None]


REF_R0, REF_CENTRE = 0, 1


def time_shift(fr: np.ndarray, t: float) -> np.ndarray:
    """exp(-j 2 pi fr t): moves a range spectrum's time origin t later."""
    return np.exp(-2j * np.pi * fr * t)


def omega_k(echo: np.ndarray, *, c: float, f0: float, fs: float, prf: float, kr: float, vr: float,
            r0: float, f_nc: float, taps: int = 6, window: int = FLOOR, range_ref: int = REF_R0) -> np.ndarray:
    """echo [na_in][nr_in] -> complex128 image [Na][Nr] (padded size, ifft2 output).

    range_ref REF_R0 is the reference's pipeline: the range spectrum keeps the
    DFT's time origin at padded sample 0 and the reference function uses r0.
    It focuses a target only if the fast time of padded sample 0 is a whole
    number of range windows (Nr/Fr) and the target lies in the first half of
    the window.  REF_CENTRE (this library's option, not in the reference)
    takes r0 as the slant range of raw sample 0 (padded sample 0 at
    ts = 2 r0/c - pr/Fr): the spectrum is moved to absolute fast time (times
    exp(-j 2 pi fr ts)), the reference range is that of padded sample
    h = floor(Nr/2), and after the Stolt step the origin is moved to padded
    sample 0 (times exp(-j 2 pi fr h/Fr)), so every range in the window is
    within Nr/2 samples of the reference and the image columns are padded
    range samples."""
    x = pad_downconvert(np.asarray(echo, dtype=np.complex128), prf=prf, f_nc=f_nc)
    na, nr = x.shape
    fr, fa, _ = axes(na, nr, fs, prf, f_nc)
    S = np.fft.fft2(x)
    r_ref = r0
    if range_ref == REF_CENTRE:
        h = nr // 2
        ts = 2.0 * r0 / c - padded_size(*echo.shape)[3] / fs
        r_ref = 0.5 * c * (ts + h / fs)
        S *= time_shift(fr, ts)[None, :]
    S *= np.exp(1j * reference_phase(fr, fa, f0=f0, c=c, vr=vr, kr=kr, r0=r_ref))
    S = stolt_sinc(S, fr, fa, f0=f0, c=c, vr=vr, fs=fs, taps=taps, window=window)
    if range_ref == REF_CENTRE:
        S *= time_shift(fr, (nr // 2) / fs)[None, :]
    return np.fft.ifft2(S)
