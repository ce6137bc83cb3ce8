"""ctypes loader of the FP64 CPU back-projection oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and the cpu_baseline
leg of bench.py use this module as the checker; the product package never
imports it.  The C source (sarbp_oracle.c) restates the BP loops of
code/SAR_BP/matlab/BPA_SAR.m:133-153 / :178-197 (reference repository).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

INTERP = {"nearest": 0, "linear": 1, "sinc8": 2}


class OrcParams(C.Structure):
    _fields_ = [("c", C.c_double), ("fc", C.c_double), ("r0", C.c_double), ("dr", C.c_double),
                ("kaiser_beta", C.c_double), ("interp", C.c_int)]


_lib: Optional[C.CDLL] = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)
    return LIB_PATH


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        dp, fp, vp = C.POINTER(C.c_double), C.POINTER(C.c_float), C.c_void_p
        L.orc_bessel_i0.restype = C.c_double
        L.orc_bessel_i0.argtypes = [C.c_double]
        L.orc_kaiser_sinc8.restype = C.c_double
        L.orc_kaiser_sinc8.argtypes = [C.c_double, C.c_double]
        L.orc_interp.argtypes = [vp, vp, C.c_int, C.c_double, C.POINTER(OrcParams), dp]
        L.orc_backproject.argtypes = [vp, vp, C.c_int, C.c_int, dp, dp, dp, dp, C.c_int, C.c_int,
                                      C.c_int, C.POINTER(OrcParams), dp, C.c_int, C.c_int]
        L.orc_backproject_f32.argtypes = [fp, C.c_int, C.c_int, dp, dp, dp, dp, C.c_int, C.c_int,
                                          C.c_int, C.POINTER(OrcParams), fp, C.c_int]
        L.orc_max_threads.restype = C.c_int
        _lib = L
    return _lib


def params(c, fc, r0, dr, interp="linear", kaiser_beta=6.0) -> OrcParams:
    return OrcParams(float(c), float(fc), float(r0), float(dr), float(kaiser_beta),
                     INTERP[interp] if isinstance(interp, str) else int(interp))


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def bessel_i0(x: float) -> float:
    return lib().orc_bessel_i0(float(x))


def kaiser_sinc8(x: float, beta: float) -> float:
    return lib().orc_kaiser_sinc8(float(x), float(beta))


def interp(profile: np.ndarray, u: float, p: OrcParams) -> complex:
    prof = np.ascontiguousarray(profile)
    out = (C.c_double * 2)()
    if prof.dtype == np.complex64:
        lib().orc_interp(prof.ctypes.data, None, prof.shape[-1], float(u), C.byref(p), out)
    else:
        prof = prof.astype(np.complex128)
        lib().orc_interp(None, prof.ctypes.data, prof.shape[-1], float(u), C.byref(p), out)
    return complex(out[0], out[1])


def backproject(profiles: np.ndarray, pos: np.ndarray, grid: dict, p: OrcParams,
                rows: Optional[tuple] = None, image: Optional[np.ndarray] = None,
                n_threads: int = 0) -> np.ndarray:
    """FP64 BP of image rows [row0, row1) (all rows by default).

    profiles: [n_pulses][n_samples] complex64 or complex128.  Returns complex128
    [(row1-row0)][nx]; if `image` is given the sums are added into it."""
    prof = np.ascontiguousarray(profiles)
    n_p, n_s = prof.shape
    nx, ny = int(grid["nx"]), int(grid["ny"])
    r0, r1 = rows if rows is not None else (0, ny)
    acc = image is not None
    out = image if acc else np.zeros((r1 - r0, nx), dtype=np.complex128)
    assert out.dtype == np.complex128 and out.flags.c_contiguous and out.shape == (r1 - r0, nx)
    posd = _d(pos)
    o, ux, uy = _d(grid["origin"]), _d(grid["ux"]), _d(grid["uy"])
    dp = C.POINTER(C.c_double)
    if prof.dtype == np.complex64:
        a32, a64 = prof.ctypes.data, None
    else:
        prof = prof.astype(np.complex128)
        a32, a64 = None, prof.ctypes.data
    lib().orc_backproject(a32, a64, n_p, n_s, posd.ctypes.data_as(dp), o.ctypes.data_as(dp),
                          ux.ctypes.data_as(dp), uy.ctypes.data_as(dp), nx, r0, r1, C.byref(p),
                          out.view(np.float64).ctypes.data_as(dp), int(acc), int(n_threads))
    return out


def backproject_f32(profiles: np.ndarray, pos: np.ndarray, grid: dict, p: OrcParams,
                    rows: Optional[tuple] = None, n_threads: int = 0) -> np.ndarray:
    """Single-precision CPU BP (the FP32 CPU baseline), complex64 rows."""
    prof = np.ascontiguousarray(profiles, dtype=np.complex64)
    n_p, n_s = prof.shape
    nx, ny = int(grid["nx"]), int(grid["ny"])
    r0, r1 = rows if rows is not None else (0, ny)
    out = np.zeros((r1 - r0, nx), dtype=np.complex64)
    posd = _d(pos)
    o, ux, uy = _d(grid["origin"]), _d(grid["ux"]), _d(grid["uy"])
    dp, fp = C.POINTER(C.c_double), C.POINTER(C.c_float)
    lib().orc_backproject_f32(prof.view(np.float32).ctypes.data_as(fp), n_p, n_s,
                              posd.ctypes.data_as(dp), o.ctypes.data_as(dp),
                              ux.ctypes.data_as(dp), uy.ctypes.data_as(dp), nx, r0, r1,
                              C.byref(p), out.view(np.float32).ctypes.data_as(fp), int(n_threads))
    return out


def max_threads() -> int:
    return lib().orc_max_threads()
