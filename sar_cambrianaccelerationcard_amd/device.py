"""Device-side helpers: moving scenes into HBM (torch tensors as plain device
allocations) and running the imager through the C ABI.

torch is used only for device memory and the current HIP stream; all BP
arithmetic runs in libsarbp.so.
"""
from __future__ import annotations

from typing import Optional

import numpy as np

from . import native
from .scene import Scene


def to_bf16x2(profiles: np.ndarray) -> np.ndarray:
    """complex64 -> uint32 words (bf16 real in the low half, bf16 imag in the
    high half), rounded to nearest even."""
    f = np.ascontiguousarray(profiles, dtype=np.complex64).view(np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF
    r = r.reshape(*profiles.shape, 2)
    return (r[..., 0] | (r[..., 1] << 16)).astype(np.uint32)


def from_bf16x2(words: np.ndarray) -> np.ndarray:
    w = np.asarray(words, dtype=np.uint32)
    re = (w << 16).view(np.float32)
    im = (w & 0xFFFF0000).view(np.float32)
    return re.astype(np.complex64) + 1j * im.astype(np.complex64)


def require_gpu():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("no HIP device visible: the imager has no CPU path")
    return torch


class DeviceScene:
    """A scene resident in HBM: profiles (c64 and optionally bf16x2), FP64
    positions."""

    def __init__(self, sc: Scene, device: int = 0, bf16: bool = False):
        torch = require_gpu()
        self.scene = sc
        dev = torch.device("cuda", device)
        self.prof = torch.from_numpy(np.ascontiguousarray(sc.profiles)).to(dev)
        self.pos = torch.from_numpy(np.ascontiguousarray(sc.pos)).to(dev)
        self.prof_bf16 = None
        if bf16:
            self.prof_bf16 = torch.from_numpy(to_bf16x2(sc.profiles).view(np.int32)).to(dev)
        self.device = dev

    def profiles_ptr(self, dtype: str = "c64") -> int:
        if dtype == "c64":
            return self.prof.data_ptr()
        if self.prof_bf16 is None:
            torch = require_gpu()
            self.prof_bf16 = torch.from_numpy(
                to_bf16x2(self.scene.profiles).view(np.int32)).to(self.device)
        return self.prof_bf16.data_ptr()

    def new_image(self):
        torch = require_gpu()
        g = self.scene.grid
        return torch.zeros((g["ny"], g["nx"]), dtype=torch.complex64, device=self.device)


def backproject(ctx: native.Context, ds: DeviceScene, *, interp: Optional[str] = None,
                precision: str = "ref", dtype: str = "c64", image=None,
                accumulate: bool = False, plan: bool = False, sync: bool = True,
                grid: Optional[dict] = None):
    """Runs one BP through the C ABI on torch's current stream; returns the
    device image tensor.  `grid` (a scene-style grid dict) replaces the
    scene's own image grid."""
    torch = require_gpu()
    sc = ds.scene
    if grid is None:
        ng = sc.native_grid()
        img = ds.new_image() if image is None else image
    else:
        ng = native.make_grid(grid["origin"], grid["ux"], grid["uy"], grid["nx"], grid["ny"])
        img = (torch.zeros((grid["ny"], grid["nx"]), dtype=torch.complex64, device=ds.device)
               if image is None else image)
    params = sc.native_params(interp=interp, precision=precision, dtype=dtype,
                              accumulate=accumulate)
    stream = torch.cuda.current_stream(ds.device).cuda_stream
    if plan:
        pl = ctx.plan(sc.pos, ng, params)
        pl.execute(ds.profiles_ptr(dtype), ds.pos.data_ptr(), img.data_ptr(), stream)
    else:
        ctx.backproject(ds.profiles_ptr(dtype), ds.pos.data_ptr(), ng, params,
                        img.data_ptr(), stream)
    if sync:
        torch.cuda.synchronize(ds.device)
    return img


def range_params(radar: dict, dtype: str = "c64", **override) -> native.RangeParams:
    """RangeParams from a scene's `radar` dict (keys of make_range_params)."""
    kw = dict(radar)
    kw.update(override)
    kw["dtype"] = dtype
    return native.make_range_params(**kw)


def range_compress(ctx: native.Context, raw, radar: dict, *, dtype: str = "c64", device: int = 0,
                   out=None, sync: bool = True, **override):
    """Range-compresses raw echoes [n_pulses][n_raw] (numpy complex or a
    complex64 device tensor) on the device; returns the profiles tensor
    ([n_pulses][n_out] complex64, or int32 bf16x2 words)."""
    torch = require_gpu()
    dev = torch.device("cuda", device)
    if not isinstance(raw, torch.Tensor):
        raw = torch.from_numpy(np.ascontiguousarray(raw, dtype=np.complex64)).to(dev)
    rp = range_params(radar, dtype, **override)
    if raw.dim() != 2 or raw.shape[1] != rp.n_raw or raw.dtype != torch.complex64 or not raw.is_contiguous():
        raise ValueError("raw must be a contiguous [n_pulses][n_raw] complex64 tensor")
    n_pulses = raw.shape[0]
    if out is None:
        out = torch.empty((n_pulses, rp.n_out), dtype=torch.complex64 if dtype == "c64" else torch.int32,
                          device=dev)
    stream = torch.cuda.current_stream(dev).cuda_stream
    ctx.range_compress(raw.data_ptr(), n_pulses, rp, out.data_ptr(), stream)
    if sync:
        torch.cuda.synchronize(dev)
    return out


def omega_k(ctx: native.Context, echo, wp: native.WkParams, *, device: int = 0, out=None,
            sync: bool = True):
    """omega-K image of raw echoes [na_in][nr_in] (numpy complex or a complex64
    device tensor); returns the padded complex64 image tensor [na][nr]."""
    torch = require_gpu()
    dev = torch.device("cuda", device)
    if not isinstance(echo, torch.Tensor):
        echo = torch.from_numpy(np.ascontiguousarray(echo, dtype=np.complex64)).to(dev)
    if echo.dim() != 2 or echo.dtype != torch.complex64 or not echo.is_contiguous():
        raise ValueError("echo must be a contiguous [na][nr] complex64 tensor")
    na, nr = native.wk_padded_size(*echo.shape)
    if out is None:
        out = torch.empty((na, nr), dtype=torch.complex64, device=dev)
    stream = torch.cuda.current_stream(dev).cuda_stream
    ctx.omega_k(echo.data_ptr(), echo.shape[0], echo.shape[1], wp, out.data_ptr(), stream)
    if sync:
        torch.cuda.synchronize(dev)
    return out


def stolt_sinc(ctx: native.Context, data, fr_axis, fa_axis, *, f0: float, c: float, vr: float, fs: float,
               taps: int = 6, window="floor", device: int = 0, out=None, sync: bool = True):
    """Stolt interpolation of [na][nr] complex data (numpy or a complex64 device
    tensor; out may be the input tensor) with explicit host axes."""
    torch = require_gpu()
    dev = torch.device("cuda", device)
    if not isinstance(data, torch.Tensor):
        data = torch.from_numpy(np.ascontiguousarray(data, dtype=np.complex64)).to(dev)
    if data.dim() != 2 or data.dtype != torch.complex64 or not data.is_contiguous():
        raise ValueError("data must be a contiguous [na][nr] complex64 tensor")
    na, nr = data.shape
    if out is None:
        out = torch.empty_like(data)
    stream = torch.cuda.current_stream(dev).cuda_stream
    ctx.stolt_sinc(data.data_ptr(), out.data_ptr(), fr_axis, fa_axis, f0=f0, c=c, vr=vr, na=na, nr=nr,
                   fs=fs, taps=taps, window=window, stream=stream)
    if sync:
        torch.cuda.synchronize(dev)
    return out


def generate_echo(ptarget, ta, tr, params: native.SarParams, *, device: int = 0):
    """Point-target raw echoes [len(ta)][len(tr)] on the device
    (sarbp_launch_generate_echo, the reference's launchGenerateEcho): targets
    [Ntarget][3] (azimuth, ground range, sigma), slow times ta and fast times
    tr, all float32; returns the complex64 tensor (synchronous)."""
    torch = require_gpu()
    dev = torch.device("cuda", device)
    pt = torch.from_numpy(np.ascontiguousarray(ptarget, dtype=np.float32).reshape(-1)).to(dev)
    d_ta = torch.from_numpy(np.ascontiguousarray(ta, dtype=np.float32)).to(dev)
    d_tr = torch.from_numpy(np.ascontiguousarray(tr, dtype=np.float32)).to(dev)
    if pt.numel() < 3 * params.Ntarget:
        raise ValueError("ptarget holds fewer than Ntarget targets")
    out = torch.empty((d_ta.numel(), d_tr.numel()), dtype=torch.complex64, device=dev)
    torch.cuda.synchronize(dev)  # the inputs are on the device before the null-stream launch
    lib = native.load()
    native.check(lib.sarbp_launch_generate_echo(native.C.c_void_p(out.data_ptr()), native.C.c_void_p(pt.data_ptr()),
                                                native.C.c_void_p(d_ta.data_ptr()),
                                                native.C.c_void_p(d_tr.data_ptr()), params,
                                                d_ta.numel(), d_tr.numel()),
                 "sarbp_launch_generate_echo")
    return out
