"""RADARSAT-1 raw data (the Vancouver fine-beam scene shipped with the
reference's omega-K processor) as BP input.

The reference keeps the raw signal as MATLAB files
(code/SAR_WK/wk_v5.2.3/data/CDdata{1,2}.mat: [1536 lines][2048 range cells]
complex, the 4-bit I/Q codes of the CEOS file as odd integers -15..15; read by
code/data_process/CEOS/read_CEOS_raw.m and extract_data.m) and the run
parameters in CD_run_params.mat (Fr, Kr, Tr, PRF, R0 = slant range of the
first sample, f0, c).  Its omega-K driver adds the effective radar velocity
Vr = 7062 m/s and the Doppler centroid -6900 Hz (wk_v5.2.3/wk.cc:62-64) and
uses the chirp as a down-chirp, Kr -> -Kr (wk_v5.2.3/wk.cc:229).

BP needs a geometry instead of those spectral parameters.  With the
hyperbolic range equation R(eta) = sqrt(r^2 + Vr^2 (eta - eta0)^2) the
equivalent straight track is exact: pulse m at (Vr (m - (M-1)/2) / PRF, 0, 0),
pixels in the slant plane at (x, r, 0), r the zero-Doppler slant range.  The
beam is squinted: a target at azimuth x is illuminated around antenna
position x - r sin(theta), sin(theta) = f_dc lambda / (2 Vr).  The pulses are
undersampled in Doppler (|f_dc| >> PRF), so back-projecting every pulse onto
every pixel would also focus each target's azimuth ambiguities (every
r lambda PRF / (2 Vr) ~ 5 km); `aperture` limits a region's pulses to those
whose beam footprint covers it.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from . import native

VR = 7062.0            # effective radar velocity, m/s (reference wk_v5.2.3/wk.cc:62)
FDC_HINT = -6900.0     # Doppler centroid of the scene, Hz (reference wk_v5.2.3/wk.cc:64)
ANTENNA_LENGTH = 15.0  # RADARSAT-1 antenna length, m (azimuth beam width lambda / La)


def pack4(raw: np.ndarray) -> np.ndarray:
    """Complex samples with odd-integer parts in [-15, 15] -> one byte per
    sample (I code in the low nibble, Q in the high), code = (v + 15) / 2."""
    re = np.asarray(raw.real)
    im = np.asarray(raw.imag)
    if not (np.all(np.abs(re) <= 15) and np.all(np.abs(im) <= 15)):
        raise ValueError("samples outside the 4-bit range")
    ci = np.rint((re + 15) / 2).astype(np.int64)
    cq = np.rint((im + 15) / 2).astype(np.int64)
    if not (np.array_equal(2 * ci - 15, re) and np.array_equal(2 * cq - 15, im)):
        raise ValueError("samples are not odd integers")
    return (ci | (cq << 4)).astype(np.uint8)


def unpack4(codes: np.ndarray) -> np.ndarray:
    """Inverse of pack4, as complex64."""
    c = np.asarray(codes, dtype=np.uint8)
    re = 2.0 * (c & 15).astype(np.float32) - 15.0
    im = 2.0 * (c >> 4).astype(np.float32) - 15.0
    return (re + 1j * im).astype(np.complex64)


@dataclass
class RunParams:
    """The scene's radar parameters (CD_run_params.mat) plus the omega-K
    velocity and Doppler-centroid constants of wk.cc:62-64."""
    c: float
    f0: float
    Fr: float
    Kr: float            # magnitude as stored; the chirp is a down-chirp
    Tr: float
    PRF: float
    R0: float            # slant range of raw sample 0, m
    Vr: float = VR
    fdc_hint: float = FDC_HINT

    @property
    def wavelength(self) -> float:
        return self.c / self.f0

    @classmethod
    def from_mat(cls, d: dict) -> "RunParams":
        g = lambda k: float(np.asarray(d[k]).ravel()[0])
        return cls(c=g("c"), f0=g("f0"), Fr=g("Fr"), Kr=g("Kr"), Tr=g("Tr"), PRF=g("PRF"), R0=g("R0"))

    def as_dict(self) -> dict:
        return dict(c=self.c, f0=self.f0, Fr=self.Fr, Kr=self.Kr, Tr=self.Tr, PRF=self.PRF, R0=self.R0,
                    Vr=self.Vr, fdc_hint=self.fdc_hint)


def doppler_centroid(raw: np.ndarray, rp: RunParams) -> float:
    """Doppler centroid from the pulse-to-pulse phase of the raw data (its
    value modulo the PRF), resolved to the ambiguity nearest the hint."""
    acc = np.sum(raw[1:].astype(np.complex128) * np.conj(raw[:-1].astype(np.complex128)))
    fb = math.atan2(acc.imag, acc.real) / (2 * math.pi) * rp.PRF
    return fb + round((rp.fdc_hint - fb) / rp.PRF) * rp.PRF


def range_params(rp: RunParams, n_raw: int, upsample: int, n_out: int, dtype: str = "c64"):
    """Range compression of the raw lines: the down-chirp (bandwidth -Kr Tr),
    raw sample 0 at t0 = 2 R0 / c."""
    return native.make_range_params(c=rp.c, bandwidth=-rp.Kr * rp.Tr, pulse_width=rp.Tr, fs=rp.Fr,
                                    t0=2.0 * rp.R0 / rp.c, n_raw=n_raw, upsample=upsample, crop_offset=0,
                                    n_out=n_out, dtype=dtype)


def positions(rp: RunParams, m0: int, m1: int, n_total: int) -> np.ndarray:
    """Antenna positions of lines [m0, m1) of an n_total-line acquisition."""
    pos = np.zeros((m1 - m0, 3))
    pos[:, 0] = rp.Vr * (np.arange(m0, m1) - (n_total - 1) / 2.0) / rp.PRF
    return pos


def squint_sine(rp: RunParams, fdc: float) -> float:
    return fdc * rp.wavelength / (2.0 * rp.Vr)


def region_grid(x_center: float, r_center: float, nx: int, ny: int, dx: float, dy: float) -> dict:
    """Slant-plane image grid: azimuth along x, zero-Doppler slant range along y."""
    return {"origin": [x_center - nx / 2 * dx, r_center - ny / 2 * dy, 0.0],
            "ux": [dx, 0.0, 0.0], "uy": [0.0, dy, 0.0], "nx": int(nx), "ny": int(ny)}


def aperture(rp: RunParams, fdc: float, grid: dict, n_total: int) -> tuple:
    """Lines [m0, m1) whose beam footprint centre lies within half a
    footprint of the region (so the region's pixels see their targets and not
    the azimuth ambiguities of targets outside it)."""
    r = grid["origin"][1] + 0.5 * (grid["ny"] - 1) * grid["uy"][1]
    x0 = grid["origin"][0]
    x1 = x0 + (grid["nx"] - 1) * grid["ux"][0]
    foot = rp.wavelength / ANTENNA_LENGTH * r
    xm = rp.Vr * (np.arange(n_total) - (n_total - 1) / 2.0) / rp.PRF
    xb = xm + squint_sine(rp, fdc) * r
    sel = np.flatnonzero((xb >= x0 - foot / 2) & (xb <= x1 + foot / 2))
    if sel.size == 0:
        raise ValueError("no line illuminates the region")
    return int(sel[0]), int(sel[-1]) + 1


def beam_centre_x(rp: RunParams, fdc: float, m0: int, m1: int, n_total: int, r: float) -> float:
    """Azimuth of the beam footprint centre at the middle of lines [m0, m1)."""
    xm = rp.Vr * ((m0 + m1 - 1) / 2.0 - (n_total - 1) / 2.0) / rp.PRF
    return xm + squint_sine(rp, fdc) * r


def load_crop(path: str) -> tuple:
    """A packed crop (tests/golden/make_radarsat_fixture.py): (raw complex64
    [lines][cells], RunParams, metadata dict)."""
    import json
    with np.load(path, allow_pickle=False) as z:
        codes = z["codes"]
        meta = json.loads(str(z["meta"]))
    p = meta["run_params"]
    rp = RunParams(c=p["c"], f0=p["f0"], Fr=p["Fr"], Kr=p["Kr"], Tr=p["Tr"], PRF=p["PRF"], R0=p["R0"],
                   Vr=p["Vr"], fdc_hint=p["fdc_hint"])
    return unpack4(codes), rp, meta


def main(argv=None) -> int:
    """Images a region of the reference's RADARSAT-1 scene on the device.

        python -m sar_cambrianaccelerationcard_amd.radarsat (--mat-dir DIR | --crop NPZ) --out PREFIX
            [--lines M0 M1] [--nx 1024 --ny 512 --dx 4 --dy 4.64] [--interp sinc8]
            [--omega-k [--f-nc HZ | --f-nc-estimate] [--range-ref r0|centre]]

    DIR holds CD_run_params.mat and CDdata1.mat / CDdata2.mat (read with
    scipy.io.loadmat); NPZ is a packed crop (load_crop).  The region is centred on the beam footprint of lines
    [M0, M1) at mid-swath; its aperture-limited lines are range-compressed
    and back-projected; PREFIX.bin (+ .json sidecar) and PREFIX.png are
    written.  With --omega-k the lines [M0, M1) (default: all) are imaged by
    sarbp_omega_k instead, as the reference's omega-K driver does with this
    scene (Doppler centroid -6900 Hz as in wk.cc:64 unless given or
    estimated); for the full 3072-line scene the image gets the MATLAB
    prototype's display shifts and flip (mian_wka_sinc.m:124-126), and the
    PNG shows the magnitude saturated at 50 (:129-131).  --range-ref centre
    uses the library's window-centred range reference (R0 is the range of the
    first sample) instead of the reference's, with the range display shift
    that lines its image up with the reference's display."""
    import argparse
    import os
    import time

    import scipy.io as sio
    import torch

    from . import device as D
    from . import imageio as io

    ap = argparse.ArgumentParser()
    src = ap.add_mutually_exclusive_group(required=True)
    src.add_argument("--mat-dir")
    src.add_argument("--crop")
    ap.add_argument("--out", required=True)
    ap.add_argument("--lines", type=int, nargs=2, default=None)
    ap.add_argument("--nx", type=int, default=1024)
    ap.add_argument("--ny", type=int, default=512)
    ap.add_argument("--dx", type=float, default=4.0)
    ap.add_argument("--dy", type=float, default=4.64)
    ap.add_argument("--interp", default="sinc8")
    ap.add_argument("--upsample", type=int, default=2)
    ap.add_argument("--omega-k", action="store_true", help="image with sarbp_omega_k instead of BP")
    fnc = ap.add_mutually_exclusive_group()
    fnc.add_argument("--f-nc", type=float, default=None, help="omega-K Doppler centroid, Hz")
    fnc.add_argument("--f-nc-estimate", action="store_true", help="omega-K with the estimated centroid")
    ap.add_argument("--range-ref", choices=("r0", "centre"), default="r0", help="omega-K range reference")
    args = ap.parse_args(argv)
    if args.crop:
        raw, rp, meta = load_crop(args.crop)
        base, n_tot, fdc = meta["line0"], meta["n_lines_total"], meta["fdc"]
    else:
        rp = RunParams.from_mat(sio.loadmat(os.path.join(args.mat_dir, "CD_run_params.mat")))
        blocks = [sio.loadmat(os.path.join(args.mat_dir, "CDdata%d.mat" % k))["data"] for k in (1, 2)
                  if os.path.exists(os.path.join(args.mat_dir, "CDdata%d.mat" % k))]
        raw = np.concatenate(blocks).astype(np.complex64)
        base, n_tot = 0, raw.shape[0]
        fdc = doppler_centroid(raw, rp)
    n_cells = raw.shape[1]
    if args.omega_k:
        return _omega_k_main(args, raw, rp, base, fdc)
    m0, m1 = args.lines or ((base, base + raw.shape[0]) if args.crop else (n_tot // 4, 3 * n_tot // 4))
    valid = n_cells - int(math.ceil(rp.Tr * rp.Fr))
    r = rp.R0 + 0.5 * valid * rp.c / (2 * rp.Fr)
    grid = region_grid(beam_centre_x(rp, fdc, m0, m1, n_tot, r), r, args.nx, args.ny, args.dx, args.dy)
    a0, a1 = aperture(rp, fdc, grid, n_tot)
    a0, a1 = max(a0, base), min(a1, base + raw.shape[0])  # lines present in the input
    L = args.upsample
    n_out = L * n_cells
    ctx = native.Context(0)
    t0 = time.perf_counter()
    rpar = range_params(rp, n_cells, L, n_out)
    radar = dict(c=rp.c, bandwidth=-rp.Kr * rp.Tr, pulse_width=rp.Tr, fs=rp.Fr, t0=2.0 * rp.R0 / rp.c,
                 n_raw=n_cells, upsample=L, crop_offset=0, n_out=n_out)
    prof = D.range_compress(ctx, np.ascontiguousarray(raw[a0 - base:a1 - base]), radar)
    pos = positions(rp, a0, a1, n_tot)
    d_pos = torch.from_numpy(pos).cuda()
    _, r0, dr = native.range_info(rpar)
    params = native.make_params(c=rp.c, fc=rp.f0, r0=r0, dr=dr, n_pulses=a1 - a0, n_samples=n_out,
                                interp=args.interp, precision="ref")
    ng = native.make_grid(grid["origin"], grid["ux"], grid["uy"], grid["nx"], grid["ny"])
    img = torch.zeros((grid["ny"], grid["nx"]), dtype=torch.complex64, device="cuda")
    plan = ctx.plan(pos, ng, params)
    plan.execute(prof.data_ptr(), d_pos.data_ptr(), img.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    h = img.cpu().numpy()
    meta = {"scene": "RADARSAT-1 CD data (reference code/SAR_WK/wk_v5.2.3/data)", "lines": [a0, a1],
            "fdc_hz": fdc, "grid": grid, "interp": args.interp, "seconds": t}
    io.save_image(args.out + ".bin", h, meta)
    io.save_png(args.out + ".png", np.minimum(np.abs(h), np.percentile(np.abs(h), 99.7)), -25.0)
    print(f"lines [{a0}, {a1}) fdc {fdc:.0f} Hz -> {args.out}.bin / .png ({t:.2f} s)")
    return 0


def _omega_k_main(args, raw: np.ndarray, rp: RunParams, base: int, fdc: float) -> int:
    import time

    from . import device as D
    from . import imageio as io
    from . import omegak as OK
    m0, m1 = args.lines or (base, base + raw.shape[0])
    if not (base <= m0 < m1 <= base + raw.shape[0]):
        raise SystemExit(f"--lines must lie in [{base}, {base + raw.shape[0]})")
    f_nc = fdc if args.f_nc_estimate else (rp.fdc_hint if args.f_nc is None else args.f_nc)
    wp = OK.radarsat_params(rp, f_nc=f_nc, range_ref=args.range_ref)
    ctx = native.Context(0)
    echo = np.ascontiguousarray(raw[m0 - base:m1 - base])
    D.omega_k(ctx, echo, wp)  # builds the plans and tables of this size
    t0 = time.perf_counter()
    img = D.omega_k(ctx, echo, wp).cpu().numpy()
    t = time.perf_counter() - t0
    full = img.shape == (4096, 3414)
    if full:
        shift = -1800 if args.range_ref == "r0" else OK.centre_display_shift(rp, img.shape[1])
        img = OK.display(img, shift_range=shift)
    meta = {"scene": "RADARSAT-1 CD data (reference code/SAR_WK/wk_v5.2.3/data)", "algorithm": "omega-K",
            "lines": [m0, m1], "f_nc_hz": f_nc, "taps": wp.taps, "window": "floor", "range_ref": args.range_ref,
            "display_shifts": full, "seconds": t}
    io.save_image(args.out + ".bin", img, meta)
    io.save_png(args.out + ".png", np.minimum(np.abs(img), 50.0), -25.0)
    print(f"omega-K lines [{m0}, {m1}) f_nc {f_nc:.0f} Hz -> {img.shape[0]} x {img.shape[1]} "
          f"{args.out}.bin / .png ({t * 1e3:.2f} ms with the copy back)")
    ctx.close()
    return 0


if __name__ == "__main__":
    import sys
    sys.exit(main())
