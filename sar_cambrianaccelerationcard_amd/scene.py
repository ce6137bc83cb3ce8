"""Seeded synthetic stripmap scenes: FP64 antenna tracks, point targets,
complex AWGN and complex-Gaussian clutter, range-compressed and upsampled
into the pulses that back-projection consumes.

Geometry.  The platform flies along +x at y = 0, height z = altitude; the
scene centre is on the ground (z = 0) at (0, ground_range, 0).  The image grid
is a flat ground grid centred on the scene centre with x (azimuth) fastest:
pixel (ix, iy) = centre + (ix - nx/2) * dx * e_x + (iy - ny/2) * dy * e_y.

Processing chain (code/SAR_BP/matlab/BPA_SAR.m:79-111 of the reference):
raw LFM echoes of the point targets plus raw-rate AWGN, FFT matched filter,
FFT zero-pad upsampling, crop to the configured number of samples.  Clutter is
added after compression as a sum of point scatterers with complex-Gaussian
amplitudes whose compressed response is the band-limited sinc of the chirp
bandwidth (csrc/sarsim.h, sarsim_compressed_points), scaled to the configured
clutter-to-noise ratio.
"""
from __future__ import annotations

import ctypes as C
import json
import os
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import native

SIM_PATH = os.path.join(native.LIB_DIR, "libsarsim.so")
CONFIG_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "configs")


class Radar(C.Structure):
    _fields_ = [("c", C.c_double), ("fc", C.c_double), ("bandwidth", C.c_double),
                ("pulse_width", C.c_double), ("fs", C.c_double), ("t0", C.c_double),
                ("n_raw", C.c_int32), ("upsample", C.c_int32), ("crop_offset", C.c_int32),
                ("n_out", C.c_int32), ("beam_half_width", C.c_double)]


_sim: Optional[C.CDLL] = None


def sim() -> C.CDLL:
    global _sim
    if _sim is None:
        if not os.path.exists(SIM_PATH):
            raise native.NativeLibraryMissing(
                f"{SIM_PATH} is missing: build it with `python -m sar_cambrianaccelerationcard_amd.build`")
        L = C.CDLL(SIM_PATH)
        dp, fp, rp = C.POINTER(C.c_double), C.POINTER(C.c_float), C.POINTER(Radar)
        L.sarsim_fft_size.restype = C.c_int32
        L.sarsim_fft_size.argtypes = [rp]
        L.sarsim_fft.restype = C.c_int
        L.sarsim_fft.argtypes = [dp, C.c_int32, C.c_int]
        L.sarsim_track_straight.argtypes = [dp, C.c_int32, C.c_double, C.c_double, C.c_double,
                                            C.c_double, C.c_double]
        L.sarsim_track_perturb.argtypes = [dp, C.c_int32, C.c_double, C.c_int32, C.c_uint64]
        L.sarsim_echo_raw.argtypes = [dp, C.c_int32, dp, C.c_int32, rp, dp, C.c_int]
        L.sarsim_add_noise.argtypes = [dp, C.c_int64, C.c_double, C.c_uint64]
        L.sarsim_compress.restype = C.c_int
        L.sarsim_compress.argtypes = [dp, C.c_int32, rp, dp, fp, C.c_int]
        L.sarsim_out_r0.restype = C.c_double
        L.sarsim_out_r0.argtypes = [rp]
        L.sarsim_out_dr.restype = C.c_double
        L.sarsim_out_dr.argtypes = [rp]
        L.sarsim_compressed_points.argtypes = [dp, C.c_int32, dp, C.c_int32, rp, C.c_int32, dp,
                                               C.c_int]
        _sim = L
    return _sim


def _dp(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


def fft(x: np.ndarray, inverse: bool = False) -> np.ndarray:
    """The simulator's own radix-2 FP64 FFT (unscaled inverse)."""
    buf = np.ascontiguousarray(x, dtype=np.complex128).copy()
    n = buf.shape[-1]
    if sim().sarsim_fft(_dp(buf.view(np.float64)), n, int(inverse)) != 0:
        raise ValueError("FFT length must be a power of two")
    return buf


@dataclass
class Scene:
    name: str
    profiles: np.ndarray            # [n_pulses][n_samples] complex64
    pos: np.ndarray                 # [n_pulses][3] float64
    grid: dict                      # origin, ux, uy, nx, ny
    c: float
    fc: float
    r0: float
    dr: float
    interp: str
    kaiser_beta: float
    targets: np.ndarray             # [K][5] x, y, z, re, im (point targets only)
    target_pixels: np.ndarray       # [K][2] (ix, iy) nearest pixel of each target
    config: dict = field(default_factory=dict)
    profiles64: Optional[np.ndarray] = None  # complex128 before the cast
    raw: Optional[np.ndarray] = None         # [n_pulses][n_raw] complex128 raw echoes (+ noise)
    radar: Optional[dict] = None             # range-compression parameters of `raw`

    @property
    def n_pulses(self) -> int:
        return self.profiles.shape[0]

    @property
    def n_samples(self) -> int:
        return self.profiles.shape[1]

    @property
    def work(self) -> int:
        return self.n_pulses * self.grid["nx"] * self.grid["ny"]

    def native_grid(self) -> native.Grid:
        g = self.grid
        return native.make_grid(g["origin"], g["ux"], g["uy"], g["nx"], g["ny"])

    def native_params(self, interp: Optional[str] = None, precision="ref", dtype="c64",
                      accumulate=False) -> native.Params:
        return native.make_params(c=self.c, fc=self.fc, r0=self.r0, dr=self.dr,
                                  n_pulses=self.n_pulses, n_samples=self.n_samples,
                                  interp=interp or self.interp, precision=precision, dtype=dtype,
                                  kaiser_beta=self.kaiser_beta, accumulate=accumulate)


def load_config(name_or_path: str) -> dict:
    path = name_or_path
    if not os.path.exists(path):
        path = os.path.join(CONFIG_DIR, name_or_path if name_or_path.endswith(".json")
                            else name_or_path + ".json")
    with open(path) as f:
        return json.load(f)


def image_grid(cfg: dict) -> dict:
    im, sc = cfg["image"], cfg["scene"]
    nx, ny, dx = int(im["nx"]), int(im["ny"]), float(im["spacing"])
    dy = float(im.get("spacing_y", dx))
    yc = float(sc["center_ground_range"])
    origin = [-(nx // 2) * dx, yc - (ny // 2) * dy, 0.0]
    return {"origin": origin, "ux": [dx, 0.0, 0.0], "uy": [0.0, dy, 0.0], "nx": nx, "ny": ny}


def _targets(cfg: dict, grid: dict, rng: np.random.Generator) -> np.ndarray:
    sc = cfg["scene"]
    yc = float(sc["center_ground_range"])
    out = []
    if sc.get("target_grid"):
        n = int(sc["target_grid"])
        sp = float(sc["target_spacing"])
        offs = (np.arange(n) - (n - 1) / 2) * sp
        for oy in offs:
            for ox in offs:
                out.append([ox, yc + oy, 0.0, 1.0, 0.0])
    nr = int(sc.get("n_random_targets", 0))
    if nr:
        margin = float(sc.get("target_margin", 16.0))
        dx, dy = grid["ux"][0], grid["uy"][1]
        x0, y0 = grid["origin"][0], grid["origin"][1]
        xs = rng.uniform(x0 + margin, x0 + (grid["nx"] - 1) * dx - margin, nr)
        ys = rng.uniform(y0 + margin, y0 + (grid["ny"] - 1) * dy - margin, nr)
        for x, y in zip(xs, ys):
            out.append([x, y, 0.0, 1.0, 0.0])
    return np.array(out, dtype=np.float64).reshape(-1, 5)


def make_scene(cfg: dict | str, *, keep_f64: bool = False, keep_raw: bool = False,
               n_threads: int = 0) -> Scene:
    """Simulates one scene from a config dict (or config name under configs/).
    keep_raw keeps the raw echoes (before range compression; clutter is added
    to the compressed profiles only) and their compression parameters."""
    if isinstance(cfg, str):
        cfg = load_config(cfg)
    L = sim()
    seed = int(cfg["seed"])
    rng = np.random.default_rng(seed)
    rd, tr, sc, pr, im = cfg["radar"], cfg["track"], cfg["scene"], cfg["processing"], cfg["image"]
    c, fc, B, Tp, fs, prf = (float(rd[k]) for k in ("c", "fc", "bandwidth", "pulse_width", "fs", "prf"))
    n_p = int(tr["n_pulses"])
    alt = float(tr["altitude"])
    yc = float(sc["center_ground_range"])
    grid = image_grid(cfg)

    pos = np.zeros((n_p, 3), dtype=np.float64)
    L.sarsim_track_straight(_dp(pos), n_p, float(tr["velocity"]), prf, 0.0, 0.0, alt)
    if float(tr.get("deviation_sigma", 0.0)) > 0.0:
        L.sarsim_track_perturb(_dp(pos), n_p, float(tr["deviation_sigma"]),
                               int(tr.get("deviation_lowpass", 200)), seed + 1)

    # Range window: centred on the scene-centre slant range from the track centre.
    up = int(pr["upsample"])
    n_out = int(pr["n_samples"])
    r_c = float(np.hypot(yc, alt))
    dr_raw = 0.5 * c / fs
    n_raw_span = int(np.ceil(n_out / up))                 # raw samples covering the window
    chirp = int(np.ceil(Tp * fs))
    n_raw = n_raw_span + chirp
    # raw sample 0 at range r_c - (n_raw_span/2) * dr_raw (echo starts there)
    t0 = 2.0 * (r_c - 0.5 * n_raw_span * dr_raw) / c
    radar = Radar(c, fc, B, Tp, fs, t0, n_raw, up, 0, n_out, float(tr.get("beam_half_width", 0.0)))
    nf = L.sarsim_fft_size(C.byref(radar))
    if n_out > nf * up:
        raise ValueError("n_samples exceeds the upsampled FFT length")
    radar.crop_offset = 0

    targets = _targets(cfg, grid, rng)
    raw = np.zeros((n_p, n_raw), dtype=np.complex128)
    if len(targets):
        L.sarsim_echo_raw(_dp(pos), n_p, _dp(np.ascontiguousarray(targets)), len(targets),
                          C.byref(radar), _dp(raw.view(np.float64)), n_threads)
    snr_db = sc.get("snr_db")
    noise_var_raw = 0.0
    if snr_db is not None:
        noise_var_raw = 10.0 ** (-float(snr_db) / 10.0)   # unit-amplitude chirp sample power
        L.sarsim_add_noise(_dp(raw.view(np.float64)), raw.size, noise_var_raw, seed + 2)
    prof64 = np.zeros((n_p, n_out), dtype=np.complex128)
    if L.sarsim_compress(_dp(raw.view(np.float64)), n_p, C.byref(radar),
                         _dp(prof64.view(np.float64)), None, n_threads) != 0:
        raise ValueError("bad crop window")
    raw_kept = raw if keep_raw else None
    del raw
    r0 = float(L.sarsim_out_r0(C.byref(radar)))
    dr = float(L.sarsim_out_dr(C.byref(radar)))

    cnr_db = sc.get("clutter_cnr_db")
    n_clutter = int(sc.get("n_clutter", 0))
    if cnr_db is not None and n_clutter > 0:
        xs = rng.uniform(grid["origin"][0], grid["origin"][0] + (grid["nx"] - 1) * grid["ux"][0], n_clutter)
        ys = rng.uniform(grid["origin"][1], grid["origin"][1] + (grid["ny"] - 1) * grid["uy"][1], n_clutter)
        amp = (rng.standard_normal(n_clutter) + 1j * rng.standard_normal(n_clutter)) / np.sqrt(2.0)
        scat = np.stack([xs, ys, np.zeros(n_clutter), amp.real, amp.imag], 1)
        clut = np.zeros_like(prof64)
        L.sarsim_compressed_points(_dp(pos), n_p, _dp(np.ascontiguousarray(scat)), n_clutter,
                                   C.byref(radar), int(pr.get("clutter_half_width", 24)),
                                   _dp(clut.view(np.float64)), n_threads)
        # compressed noise power per sample: raw variance x chirp samples
        p_noise = noise_var_raw * chirp if noise_var_raw > 0 else 1.0
        p_clut = float(np.mean(np.abs(clut) ** 2))
        if p_clut > 0:
            clut *= np.sqrt(10.0 ** (float(cnr_db) / 10.0) * p_noise / p_clut)
        prof64 += clut
        del clut

    dx, dy = grid["ux"][0], grid["uy"][1]
    tpix = np.stack([np.rint((targets[:, 0] - grid["origin"][0]) / dx),
                     np.rint((targets[:, 1] - grid["origin"][1]) / dy)], 1).astype(np.int64) \
        if len(targets) else np.zeros((0, 2), np.int64)
    return Scene(name=cfg.get("name", "scene"), profiles=prof64.astype(np.complex64), pos=pos,
                 grid=grid, c=c, fc=fc, r0=r0, dr=dr, interp=im["interp"],
                 kaiser_beta=float(im.get("kaiser_beta", 6.0)), targets=targets,
                 target_pixels=tpix, config=cfg, profiles64=prof64 if keep_f64 else None,
                 raw=raw_kept,
                 radar={"c": c, "bandwidth": B, "pulse_width": Tp, "fs": fs, "t0": t0,
                        "n_raw": n_raw, "upsample": up, "crop_offset": 0, "n_out": n_out})


def scaled_config(cfg: dict, *, n_pulses=None, nx=None, ny=None, n_samples=None,
                  n_clutter=None, n_random_targets=None) -> dict:
    """A copy of cfg with some sizes replaced (for small parity cases)."""
    out = json.loads(json.dumps(cfg))
    if n_pulses is not None:
        out["track"]["n_pulses"] = int(n_pulses)
    if nx is not None:
        out["image"]["nx"] = int(nx)
    if ny is not None:
        out["image"]["ny"] = int(ny)
    if n_samples is not None:
        out["processing"]["n_samples"] = int(n_samples)
    if n_clutter is not None:
        out["scene"]["n_clutter"] = int(n_clutter)
    if n_random_targets is not None:
        out["scene"]["n_random_targets"] = int(n_random_targets)
    return out
