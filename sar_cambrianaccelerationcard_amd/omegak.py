"""omega-K image formation on the device: parameter sets, a point-target echo
simulator matched to the transform's conventions, and the display steps.

The device path is sarbp_omega_k (include/sarbp.h), the reference's process_wk
(code/SAR_WK/wk_v5.2.3/wk.cc:43-343).  Conventions of that path, which the
simulator follows:
  * rows are azimuth lines at slow time eta = m / prf (m the padded line
    index), columns range samples at fast time tau = tau0 + n / fs (n the
    padded sample index);
  * the echo of a point target is exp(j pi kr (tau - 2R/c)^2) for
    |tau - 2R/c| <= Tr/2, times exp(-j 4 pi f0 R / c), R = R(eta) its slant
    range along a straight track at speed vr;
  * the reference function exp(j 4 pi r0/c sqrt(...)) assumes absolute fast
    time; the discrete transform sees the same phases when tau0 fs / Nr is an
    integer, and tau0 = 2 r0 / c then puts a target of closest range R at
    image sample (2R/c - tau0) fs, line eta_t prf (both modulo the padded
    size).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, replace

import numpy as np

from . import native


@dataclass
class WkScene:
    c: float = 299792458.0
    f0: float = 1.0e9       # carrier, Hz
    fs: float = 50.0e6      # range sampling rate, Hz
    prf: float = 200.0      # Hz
    kr: float = -2.0e13     # signed chirp rate, Hz/s
    tr: float = 1.5e-6      # chirp duration, s
    vr: float = 274.0       # m/s
    r0: float = 0.0         # reference range; 0: the integer-period range
    f_nc: float = 0.0       # Doppler centroid, Hz
    doppler_band: float = 0.6  # processed Doppler band, fraction of the prf
    na_in: int = 384
    nr_in: int = 384

    @property
    def padded(self) -> tuple:
        return native.wk_padded_size(self.na_in, self.nr_in)

    @property
    def pads(self) -> tuple:
        na, nr = self.padded
        return (na - self.na_in) // 2, (nr - self.nr_in) // 2

    def resolved(self) -> "WkScene":
        """This scene with r0 = the smallest range whose two-way time is a whole
        number of padded range windows (Nr / fs), if r0 was 0."""
        if self.r0 > 0.0:
            return self
        _, nr = self.padded
        return replace(self, r0=0.5 * self.c * nr / self.fs)

    @property
    def first_sample_range(self) -> float:
        """Slant range of raw echo sample 0 (padded sample pad_rg)."""
        s = self.resolved()
        return s.r0 + 0.5 * s.c * s.pads[1] / s.fs

    def params(self, taps: int = 6, window="floor", range_ref="r0") -> native.WkParams:
        """sarbp_omega_k parameters: r0 is the scene's reference range for
        range_ref "r0", the range of raw sample 0 for "centre"."""
        s = self.resolved()
        r0 = s.first_sample_range if range_ref == "centre" else s.r0
        return native.make_wk_params(c=s.c, f0=s.f0, fs=s.fs, prf=s.prf, kr=s.kr, vr=s.vr, r0=r0,
                                     f_nc=s.f_nc, taps=taps, window=window, range_ref=range_ref)


def target_at(scene: WkScene, line: float, sample: float) -> tuple:
    """(eta_t, R_t) of a target that images at padded line `line` and padded
    range sample `sample`."""
    s = scene.resolved()
    return line / s.prf, s.r0 + 0.5 * s.c * sample / s.fs


def aperture_target(scene: WkScene, aperture_line: float, sample: float) -> tuple:
    """Image position (padded line, sample) of a target whose illumination is
    centred (Doppler = f_nc) on padded line `aperture_line`: with squint the
    target's zero-Doppler line lies away from its illuminated lines."""
    s = scene.resolved()
    lam = s.c / s.f0
    r = s.r0 + 0.5 * s.c * sample / s.fs
    sq = s.f_nc * lam / (2.0 * s.vr)
    d_eta = -s.f_nc * lam * r / (2.0 * s.vr * s.vr * math.sqrt(1.0 - sq * sq))
    return aperture_line - d_eta * s.prf, sample


def radarsat_scene(na_in: int = 3072, nr_in: int = 2048, r0: float = 988647.462) -> WkScene:
    """The RADARSAT-1 scene's radar (code/SAR_WK/wk_v5.2.3/data/CD_run_params.mat,
    wk.cc:62-64) at the integer-window reference range nearest r0."""
    sc = WkScene(c=299790000.0, f0=5.3e9, fs=32.317e6, prf=1256.98, kr=-7.2135e11, tr=41.75e-6, vr=7062.0,
                 f_nc=-6900.0, doppler_band=0.6, na_in=na_in, nr_in=nr_in)
    _, nr = sc.padded
    window = 0.5 * sc.c * nr / sc.fs
    return replace(sc, r0=max(1, round(r0 / window)) * window)


def simulate(scene: WkScene, targets, amplitudes=None) -> np.ndarray:
    """Raw echo [na_in][nr_in] complex64 of point targets given as
    (padded line, padded range sample) image positions."""
    s = scene.resolved()
    pa, pr = s.pads
    lam = s.c / s.f0
    tau0 = 2.0 * s.r0 / s.c
    echo = np.zeros((s.na_in, s.nr_in), dtype=np.complex128)
    amps = np.ones(len(targets)) if amplitudes is None else np.asarray(amplitudes, dtype=np.float64)
    m_in = np.arange(s.na_in)
    eta = (m_in + pa) / s.prf
    half_band = 0.5 * s.doppler_band * s.prf
    n_chirp = int(math.ceil(s.tr * s.fs)) + 2
    for (line, sample), a in zip(targets, amps):
        eta_t, r_t = target_at(s, line, sample)
        rng = np.sqrt(r_t * r_t + (s.vr * (eta - eta_t)) ** 2)
        fd = -2.0 * s.vr * s.vr * (eta - eta_t) / (lam * rng)
        lit = np.abs(fd - s.f_nc) <= half_band
        for mi in np.nonzero(lit)[0]:
            d = 2.0 * rng[mi] / s.c - tau0  # delay from padded sample 0
            n0 = int(math.floor(d * s.fs - 0.5 * n_chirp))
            n = np.arange(n0, n0 + n_chirp + 1)
            t = n / s.fs - d
            n_in = n - pr
            ok = (np.abs(t) <= 0.5 * s.tr) & (n_in >= 0) & (n_in < s.nr_in)
            ph = math.pi * s.kr * t[ok] ** 2 - 4.0 * math.pi * s.f0 * rng[mi] / s.c
            echo[mi, n_in[ok]] += a * np.exp(1j * ph)
    return echo.astype(np.complex64)


def radarsat_params(rp, *, f_nc: float | None = None, taps: int = 6, window="floor",
                    range_ref="r0") -> native.WkParams:
    """omega-K parameters of the RADARSAT-1 scene (radarsat.RunParams): the
    down-chirp (kr = -Kr), r0 = R0, wk.cc's vr and Doppler centroid
    (wk.cc:62-64 / mian_wka_sinc.m:32-34) unless f_nc is given.  R0 is the
    slant range of the first sample, which is what range_ref "centre" takes."""
    return native.make_wk_params(c=rp.c, f0=rp.f0, fs=rp.Fr, prf=rp.PRF, kr=-rp.Kr, vr=rp.Vr, r0=rp.R0,
                                 f_nc=rp.fdc_hint if f_nc is None else f_nc, taps=taps, window=window,
                                 range_ref=range_ref)


def centre_display_shift(rp, nr: int, shift_range: int = -1800) -> int:
    """Range display shift that puts a range_ref "centre" image where the
    reference's display (shift_range on a range_ref "r0" image) puts the
    same scene: the "r0" image's columns sit (2 R0/c fs) mod Nr samples
    before the padded samples of the "centre" image."""
    lag = math.fmod(2.0 * rp.R0 / rp.c * rp.Fr, nr)
    return int(round(shift_range - lag))


def display(image: np.ndarray, shift_range: int = -1800, shift_az: int = -3365, flip: bool = True) -> np.ndarray:
    """The prototype's display steps (mian_wka_sinc.m:124-126): circular shifts
    along range then azimuth, then an up-down flip."""
    out = np.roll(np.roll(image, shift_range, axis=1), shift_az, axis=0)
    return out[::-1] if flip else out


def peak_metrics(image: np.ndarray, line: float, sample: float, search: int = 3, box: int = 8) -> dict:
    """Peak of |image| within `search` of (line, sample) (circularly) and the
    fraction of the energy of the (2 box + 1)^2 neighbourhood it holds."""
    na, nr = image.shape
    li = int(round(line)) % na
    si = int(round(sample)) % nr
    rows = (li + np.arange(-box, box + 1)) % na
    cols = (si + np.arange(-box, box + 1)) % nr
    win = np.abs(image[np.ix_(rows, cols)]) ** 2
    c0 = box - search
    sub = win[c0:c0 + 2 * search + 1, c0:c0 + 2 * search + 1]
    k = int(np.argmax(sub))
    dl, ds = divmod(k, sub.shape[1])
    return {"d_line": dl - search, "d_sample": ds - search, "peak": float(np.sqrt(sub.max())),
            "concentration": float(sub.max() / win.sum())}


def bp_image(ctx, scene: WkScene, echo, *, interp: str = "sinc8", upsample: int = 4,
             lines: tuple | None = None, samples: tuple | None = None) -> np.ndarray:
    """Back-projection of the same raw echo onto the omega-K image grid, for
    comparing the two algorithms: pixel (line, sample) lies at along-track
    vr * line / prf and slant range r0 + c/2 * sample / fs in the slant plane
    of a straight track (the simulator's geometry).  Range compression and BP
    run on the device.  Returns |lines| x |samples| complex64, rows = lines."""
    import torch

    from . import device as D
    s = scene.resolved()
    pa, pr = s.pads
    na, nr = s.padded
    l0, l1 = lines or (0, na)
    s0, s1 = samples or (0, nr)
    # The simulated chirp is centred on its delay; the matched filter puts the
    # peak at the chirp's start, half a pulse earlier.
    t_raw0 = 2.0 * s.r0 / s.c + pr / s.fs + 0.5 * s.tr
    radar = dict(c=s.c, bandwidth=s.kr * s.tr, pulse_width=s.tr, fs=s.fs, t0=t_raw0, n_raw=s.nr_in,
                 upsample=upsample, crop_offset=0, n_out=upsample * s.nr_in)
    prof = D.range_compress(ctx, np.ascontiguousarray(echo, dtype=np.complex64), radar)
    _, r0, dr = native.range_info(D.range_params(radar))
    pos = np.zeros((s.na_in, 3))
    pos[:, 0] = s.vr * (np.arange(s.na_in) + pa) / s.prf
    dx, dy = s.vr / s.prf, 0.5 * s.c / s.fs
    grid = native.make_grid((l0 * dx, s.r0 + s0 * dy, 0.0), (dx, 0.0, 0.0), (0.0, dy, 0.0), l1 - l0, s1 - s0)
    params = native.make_params(c=s.c, fc=s.f0, r0=r0, dr=dr, n_pulses=s.na_in, n_samples=upsample * s.nr_in,
                                interp=interp)
    dpos = torch.from_numpy(pos).to(prof.device)
    img = torch.zeros((s1 - s0, l1 - l0), dtype=torch.complex64, device=prof.device)
    stream = torch.cuda.current_stream(prof.device).cuda_stream
    ctx.backproject(prof.data_ptr(), dpos.data_ptr(), grid, params, img.data_ptr(), stream)
    torch.cuda.synchronize(prof.device)
    return img.cpu().numpy().T.copy()
