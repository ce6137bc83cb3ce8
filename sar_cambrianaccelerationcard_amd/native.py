"""ctypes binding of the gfx950 back-projection library (include/sarbp.h).

The library is loaded from this package's _lib/ directory.  There is no
fallback: if libsarbp.so is missing, or no HIP device is present, every entry
point raises.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional, Sequence

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(_HERE, "_lib")
SARBP_PATH = os.path.join(LIB_DIR, "libsarbp.so")

# ---- enums (mirror include/sarbp.h) ----------------------------------------
OK = 0
ERR_INVALID_ARGUMENT = 1
ERR_OUT_OF_MEMORY = 2
ERR_OVER_BUDGET = 3
ERR_DEVICE = 4
ERR_UNSUPPORTED = 5
ERR_NOT_READY = 6
ERR_NOT_FOUND = 7
ERR_CANCELLED = 8
ERR_TIMEOUT = 9
ERR_SHUTDOWN = 10

INTERP_NEAREST, INTERP_LINEAR, INTERP_SINC8 = 0, 1, 2
PREC_REF, PREC_F64, PREC_F32, PREC_DF = 0, 1, 2, 3
DTYPE_C64, DTYPE_BF16X2 = 0, 1

INTERP_NAMES = {"nearest": INTERP_NEAREST, "linear": INTERP_LINEAR, "sinc8": INTERP_SINC8}
PREC_NAMES = {"ref": PREC_REF, "f64": PREC_F64, "f32": PREC_F32, "df": PREC_DF}
DTYPE_NAMES = {"c64": DTYPE_C64, "bf16x2": DTYPE_BF16X2}


class Grid(C.Structure):
    _fields_ = [("origin", C.c_double * 3), ("ux", C.c_double * 3),
                ("uy", C.c_double * 3), ("nx", C.c_int32), ("ny", C.c_int32)]


class Params(C.Structure):
    _fields_ = [("c", C.c_double), ("fc", C.c_double), ("r0", C.c_double),
                ("dr", C.c_double), ("kaiser_beta", C.c_double),
                ("n_pulses", C.c_int32), ("n_samples", C.c_int32),
                ("interp", C.c_int32), ("precision", C.c_int32),
                ("dtype", C.c_int32), ("accumulate", C.c_int32)]


class PlanInfo(C.Structure):
    _fields_ = [("tiles_x", C.c_int32), ("tiles_y", C.c_int32),
                ("tile_nx", C.c_int32), ("tile_ny", C.c_int32),
                ("window", C.c_int32), ("pulses_per_stage", C.c_int32),
                ("rows_per_lane", C.c_int32), ("cols_per_lane", C.c_int32),
                ("range_order", C.c_int32),
                ("lds_bytes", C.c_int32), ("max_phase_err_rad", C.c_double),
                ("sinc8_halves", C.c_int32), ("precision", C.c_int32)]


class Tuning(C.Structure):
    _fields_ = [("rows_per_lane", C.c_int32), ("cols_per_lane", C.c_int32),
                ("pulses_per_stage", C.c_int32)]


class AutotuneResult(C.Structure):
    _fields_ = [("best", Tuning), ("best_ms", C.c_double), ("default_ms", C.c_double),
                ("n_candidates", C.c_int32), ("crop_nx", C.c_int32), ("crop_ny", C.c_int32)]


class SarParams(C.Structure):
    """The reference's SARParams (code/SAR_BP/CUDA/BPA_CUDA.h:28-50)."""
    _fields_ = [(n, C.c_float) for n in ("c", "fc", "lambda_", "Kr", "Tr", "Br", "Fr", "vr", "H",
                                          "La", "Rg0", "Az0")] + \
               [("Nfast", C.c_int32), ("Mslow", C.c_int32), ("Ntarget", C.c_int32)]

    @classmethod
    def reference_defaults(cls) -> "SarParams":
        """The defaults of BPA_CUDA.h:29-48."""
        p = cls()
        p.c, p.fc, p.Kr, p.Tr, p.Br, p.Fr = 3e8, 5.3e9, 4e12, 2.5e-6, 1e8, 1.2e8
        p.vr, p.H, p.La, p.Rg0, p.Az0 = 150.0, 5000.0, 4.0, 10000.0, 0.0
        p.lambda_ = p.c / p.fc
        # static_cast<int>(Tr * Fr) of two floats: 2.5e-6f * 1.2e8f rounds to 300.0f
        import numpy as np
        p.Nfast = int(np.float32(p.Tr) * np.float32(p.Fr))
        p.Mslow = 1024
        p.Ntarget = 5
        return p


class Job(C.Structure):
    _fields_ = [("d_profiles", C.c_void_p), ("d_pos", C.c_void_p), ("h_pos", C.c_void_p),
                ("grid", Grid), ("params", Params), ("d_image", C.c_void_p)]


class ServerConfig(C.Structure):
    _fields_ = [("device", C.c_int32), ("interp", C.c_int32), ("precision", C.c_int32),
                ("dtype", C.c_int32), ("hbm_budget_bytes", C.c_size_t),
                ("max_active_jobs", C.c_int32), ("pulses_per_epoch", C.c_int32),
                ("rows_per_lane", C.c_int32), ("cols_per_lane", C.c_int32),
                ("pulses_per_stage", C.c_int32),
                ("epoch_watchdog_s", C.c_double)]


class Request(C.Structure):
    _fields_ = [("h_profiles", C.c_void_p), ("h_pos", C.c_void_p), ("grid", Grid),
                ("params", Params), ("h_image", C.c_void_p)]


class JobInfo(C.Structure):
    _fields_ = [("state", C.c_int32), ("status", C.c_int32), ("epochs", C.c_int32),
                ("reserved", C.c_int32), ("submit_s", C.c_double), ("admit_s", C.c_double),
                ("first_epoch_s", C.c_double), ("done_s", C.c_double), ("work", C.c_double)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("completed", C.c_int64), ("failed", C.c_int64),
                ("cancelled", C.c_int64), ("epochs", C.c_int64), ("queued", C.c_int32),
                ("active", C.c_int32), ("work_done", C.c_double), ("gpu_busy_s", C.c_double),
                ("uptime_s", C.c_double), ("hbm_in_use", C.c_size_t), ("hbm_peak", C.c_size_t)]


class RangeParams(C.Structure):
    _fields_ = [("c", C.c_double), ("bandwidth", C.c_double), ("pulse_width", C.c_double),
                ("fs", C.c_double), ("t0", C.c_double), ("n_raw", C.c_int32),
                ("upsample", C.c_int32), ("crop_offset", C.c_int32), ("n_out", C.c_int32),
                ("dtype", C.c_int32), ("reserved", C.c_int32)]


class RawRequest(C.Structure):
    _fields_ = [("h_raw", C.c_void_p), ("range", RangeParams), ("h_pos", C.c_void_p),
                ("grid", Grid), ("params", Params), ("h_image", C.c_void_p)]


class WkParams(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("c", "f0", "fs", "prf", "kr", "vr", "r0", "f_nc")] + \
        [("taps", C.c_int32), ("window", C.c_int32), ("range_ref", C.c_int32), ("reserved", C.c_int32)]


STOLT_FLOOR, STOLT_CENTERED = 0, 1
STOLT_WINDOWS = {"floor": STOLT_FLOOR, "centered": STOLT_CENTERED}
WK_RANGE_REF_R0, WK_RANGE_REF_CENTRE = 0, 1
WK_RANGE_REFS = {"r0": WK_RANGE_REF_R0, "centre": WK_RANGE_REF_CENTRE}

JOB_QUEUED, JOB_RESIDENT, JOB_RUNNING, JOB_DONE, JOB_FAILED, JOB_CANCELLED = range(6)


class SarbpError(RuntimeError):
    def __init__(self, status: int, where: str, detail: str):
        self.status = status
        super().__init__(f"{where}: status {status} ({detail})")


class NativeLibraryMissing(RuntimeError):
    pass


_lib: Optional[C.CDLL] = None

# name -> (restype, argtypes)
_SIGS = {
    "sarbp_strerror": (C.c_char_p, [C.c_int]),
    "sarbp_abi_version": (C.c_int32, []),
    "sarbp_last_error_detail": (C.c_char_p, []),
    "sarbp_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int32]),
    "sarbp_destroy": (None, [C.c_void_p]),
    "sarbp_backproject": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(Grid),
                                    C.POINTER(Params), C.c_void_p, C.c_void_p]),
    "sarbp_backproject_sync": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(Grid),
                                         C.POINTER(Params), C.c_void_p, C.c_void_p]),
    "sarbp_plan_create": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(Grid), C.POINTER(Params),
                                    C.POINTER(C.c_void_p)]),
    "sarbp_plan_execute": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "sarbp_plan_get_info": (C.c_int, [C.c_void_p, C.POINTER(PlanInfo)]),
    "sarbp_plan_destroy": (None, [C.c_void_p]),
    "sarbp_set_tuning": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.POINTER(Tuning)]),
    "sarbp_get_tuning": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.POINTER(Tuning)]),
    "sarbp_size_class": (C.c_int32, [C.POINTER(Grid), C.POINTER(Params), C.c_void_p]),
    "sarbp_set_class_tuning": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                         C.POINTER(Tuning)]),
    "sarbp_get_class_tuning": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                         C.POINTER(Tuning)]),
    "sarbp_server_set_class_tuning": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(Tuning)]),
    "sarbp_backproject_batch": (C.c_int, [C.c_void_p, C.POINTER(Job), C.c_int32, C.c_void_p]),
    "sarbp_backproject_batch_noalloc": (C.c_int, [C.c_void_p, C.POINTER(Job), C.c_int32, C.c_void_p]),
    "sarbp_sinc8_coefficients": (C.c_int, [C.c_double, C.POINTER(C.c_float)]),
    "sarbp_sinc8_coefficients_whole": (C.c_int, [C.c_double, C.POINTER(C.c_float)]),
    "sarbp_launch_backprojection": (C.c_int, [C.c_void_p, C.c_void_p, SarParams, C.c_int32,
                                              C.c_int32]),
    "sarbp_launch_generate_echo": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, SarParams,
                                             C.c_int32, C.c_int32]),
    "sarbp_server_create": (C.c_int, [C.POINTER(ServerConfig), C.POINTER(C.c_void_p)]),
    "sarbp_server_submit": (C.c_int, [C.c_void_p, C.POINTER(Request), C.POINTER(C.c_uint64)]),
    "sarbp_server_submit_raw": (C.c_int, [C.c_void_p, C.POINTER(RawRequest), C.POINTER(C.c_uint64)]),
    "sarbp_server_poll": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(JobInfo)]),
    "sarbp_server_wait": (C.c_int, [C.c_void_p, C.c_uint64, C.c_double, C.POINTER(JobInfo)]),
    "sarbp_server_cancel": (C.c_int, [C.c_void_p, C.c_uint64]),
    "sarbp_server_release": (C.c_int, [C.c_void_p, C.c_uint64]),
    "sarbp_server_get_stats": (C.c_int, [C.c_void_p, C.POINTER(ServerStats)]),
    "sarbp_server_drain": (C.c_int, [C.c_void_p, C.c_double]),
    "sarbp_server_destroy": (None, [C.c_void_p]),
    "sarbp_range_info": (C.c_int, [C.POINTER(RangeParams), C.POINTER(C.c_int32),
                                   C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "sarbp_range_compress": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.POINTER(RangeParams),
                                       C.c_void_p, C.c_void_p]),
    "sarbp_stolt_sinc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                   C.c_double, C.c_double, C.c_double, C.c_int32, C.c_int32,
                                   C.c_double, C.c_int32, C.c_int32, C.c_void_p]),
    "sarbp_stolt_gridsample": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                         C.c_double, C.c_double, C.c_double, C.c_int32, C.c_int32,
                                         C.c_double, C.c_void_p]),
    "sarbp_wk_preprocess": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_double, C.c_double,
                                      C.c_void_p, C.c_void_p]),
    "sarbp_wk_fft2": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_void_p]),
    "sarbp_wk_reference_function": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double,
                                              C.c_double, C.c_double, C.c_double, C.c_double, C.c_int32,
                                              C.c_int32, C.c_void_p]),
    "sarbp_wk_padded_size": (C.c_int, [C.c_int32, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "sarbp_omega_k": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.POINTER(WkParams),
                                C.c_void_p, C.c_void_p]),
    "sarbp_valu_probe": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p]),
    "sarbp_rotation_probe": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p]),
    "sarbp_autotune": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(Grid),
                                 C.POINTER(Params), C.c_void_p, C.c_size_t, C.c_int32, C.c_void_p,
                                 C.POINTER(AutotuneResult)]),
}


def load() -> C.CDLL:
    """Loads libsarbp.so (raises NativeLibraryMissing if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    # One HIP runtime per process: torch ships its own libamdhip64 (same
    # SONAME, libamdhip64.so.7).  Importing torch first makes the dynamic
    # linker bind this library to that copy; loading ours first would leave
    # torch to load a second runtime, which then finds no device.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not os.path.exists(SARBP_PATH):
        raise NativeLibraryMissing(
            f"{SARBP_PATH} is missing: build it with "
            "`python -m sar_cambrianaccelerationcard_amd.build`")
    lib = C.CDLL(SARBP_PATH)
    for name, (res, args) in _SIGS.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def exported_symbols() -> Sequence[str]:
    return list(_SIGS)


def check(status: int, where: str) -> None:
    if status != OK:
        lib = load()
        detail = lib.sarbp_last_error_detail().decode()
        name = lib.sarbp_strerror(status).decode()
        raise SarbpError(status, where, f"{name}: {detail}")


def make_grid(origin, ux, uy, nx: int, ny: int) -> Grid:
    g = Grid()
    for k in range(3):
        g.origin[k] = float(origin[k])
        g.ux[k] = float(ux[k])
        g.uy[k] = float(uy[k])
    g.nx, g.ny = int(nx), int(ny)
    return g


def make_range_params(*, c: float, bandwidth: float, pulse_width: float, fs: float,
                      t0: float, n_raw: int, upsample: int, crop_offset: int, n_out: int,
                      dtype="c64") -> RangeParams:
    rp = RangeParams()
    rp.c, rp.bandwidth, rp.pulse_width, rp.fs, rp.t0 = (float(c), float(bandwidth),
                                                        float(pulse_width), float(fs), float(t0))
    rp.n_raw, rp.upsample, rp.crop_offset, rp.n_out = int(n_raw), int(upsample), int(crop_offset), int(n_out)
    rp.dtype = DTYPE_NAMES[dtype] if isinstance(dtype, str) else int(dtype)
    rp.reserved = 0
    return rp


def range_info(rp: RangeParams) -> tuple:
    """(nf, r0, dr) of a range-compression parameter set (host only)."""
    lib = load()
    nf, r0, dr = C.c_int32(), C.c_double(), C.c_double()
    check(lib.sarbp_range_info(C.byref(rp), C.byref(nf), C.byref(r0), C.byref(dr)), "sarbp_range_info")
    return nf.value, r0.value, dr.value


def make_wk_params(*, c: float, f0: float, fs: float, prf: float, kr: float, vr: float, r0: float,
                   f_nc: float, taps: int = 6, window="floor", range_ref="r0") -> WkParams:
    """omega-K parameters; kr is the signed chirp rate (negative for a down-chirp).
    range_ref "r0" is the reference's range origin (r0 the reference range),
    "centre" the window-centred one (r0 the slant range of raw sample 0)."""
    wp = WkParams()
    for k, v in dict(c=c, f0=f0, fs=fs, prf=prf, kr=kr, vr=vr, r0=r0, f_nc=f_nc).items():
        setattr(wp, k, float(v))
    wp.taps = int(taps)
    wp.window = STOLT_WINDOWS[window] if isinstance(window, str) else int(window)
    wp.range_ref = WK_RANGE_REFS[range_ref] if isinstance(range_ref, str) else int(range_ref)
    return wp


def wk_padded_size(na_in: int, nr_in: int) -> tuple:
    """Padded (na, nr) of an na_in x nr_in echo (host only)."""
    lib = load()
    na, nr = C.c_int32(), C.c_int32()
    check(lib.sarbp_wk_padded_size(int(na_in), int(nr_in), C.byref(na), C.byref(nr)), "sarbp_wk_padded_size")
    return na.value, nr.value


def make_params(*, c: float, fc: float, r0: float, dr: float, n_pulses: int,
                n_samples: int, interp="linear", precision="ref", dtype="c64",
                kaiser_beta: float = 6.0, accumulate: bool = False) -> Params:
    p = Params()
    p.c, p.fc, p.r0, p.dr, p.kaiser_beta = float(c), float(fc), float(r0), float(dr), float(kaiser_beta)
    p.n_pulses, p.n_samples = int(n_pulses), int(n_samples)
    p.interp = INTERP_NAMES[interp] if isinstance(interp, str) else int(interp)
    p.precision = PREC_NAMES[precision] if isinstance(precision, str) else int(precision)
    p.dtype = DTYPE_NAMES[dtype] if isinstance(dtype, str) else int(dtype)
    p.accumulate = int(bool(accumulate))
    return p


class Context:
    """Owns a sarbp_context on one HIP device."""

    def __init__(self, device: int = 0):
        self.lib = load()
        h = C.c_void_p()
        check(self.lib.sarbp_create(C.byref(h), int(device)), "sarbp_create")
        self.handle = h
        self.device = device

    def close(self) -> None:
        if self.handle:
            self.lib.sarbp_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def backproject(self, d_profiles: int, d_pos: int, grid: Grid, params: Params,
                    d_image: int, stream: int = 0, sync: bool = False) -> None:
        f = self.lib.sarbp_backproject_sync if sync else self.lib.sarbp_backproject
        check(f(self.handle, C.c_void_p(d_profiles), C.c_void_p(d_pos), C.byref(grid),
                C.byref(params), C.c_void_p(d_image), C.c_void_p(stream)), "sarbp_backproject")

    def range_compress(self, d_raw: int, n_pulses: int, rp: RangeParams, d_profiles: int,
                       stream: int = 0) -> None:
        check(self.lib.sarbp_range_compress(self.handle, C.c_void_p(d_raw), int(n_pulses),
                                            C.byref(rp), C.c_void_p(d_profiles), C.c_void_p(stream)),
              "sarbp_range_compress")

    def stolt_sinc(self, d_in: int, d_out: int, fr_axis, fa_axis, *, f0: float, c: float, vr: float,
                   na: int, nr: int, fs: float, taps: int = 6, window="floor", stream: int = 0) -> None:
        """sarbp_stolt_sinc on device buffers; the axes are host float64 arrays."""
        import numpy as np
        fr = np.ascontiguousarray(fr_axis, dtype=np.float64)
        fa = np.ascontiguousarray(fa_axis, dtype=np.float64)
        if fr.size != nr or fa.size != na:
            raise ValueError("axis lengths must be nr and na")
        w = STOLT_WINDOWS[window] if isinstance(window, str) else int(window)
        check(self.lib.sarbp_stolt_sinc(self.handle, C.c_void_p(d_in), C.c_void_p(d_out),
                                        fr.ctypes.data_as(C.c_void_p), fa.ctypes.data_as(C.c_void_p),
                                        float(f0), float(c), float(vr), int(na), int(nr), float(fs),
                                        int(taps), w, C.c_void_p(stream)), "sarbp_stolt_sinc")

    def stolt_gridsample(self, d_in: int, d_out: int, fr_axis, fa_axis, *, f0: float, c: float, vr: float,
                         na: int, nr: int, fr: float, stream: int = 0) -> None:
        """sarbp_stolt_gridsample (the reference's GridSample Stolt step) on
        device buffers; the axes are host float64 arrays, fr the bin step."""
        import numpy as np
        fra = np.ascontiguousarray(fr_axis, dtype=np.float64)
        faa = np.ascontiguousarray(fa_axis, dtype=np.float64)
        if fra.size != nr or faa.size != na:
            raise ValueError("axis lengths must be nr and na")
        check(self.lib.sarbp_stolt_gridsample(self.handle, C.c_void_p(d_in), C.c_void_p(d_out),
                                              fra.ctypes.data_as(C.c_void_p), faa.ctypes.data_as(C.c_void_p),
                                              float(f0), float(c), float(vr), int(na), int(nr), float(fr),
                                              C.c_void_p(stream)), "sarbp_stolt_gridsample")

    def wk_preprocess(self, d_echo: int, na_in: int, nr_in: int, prf: float, f_nc: float, d_out: int,
                      stream: int = 0) -> None:
        check(self.lib.sarbp_wk_preprocess(self.handle, C.c_void_p(d_echo), int(na_in), int(nr_in), float(prf),
                                           float(f_nc), C.c_void_p(d_out), C.c_void_p(stream)),
              "sarbp_wk_preprocess")

    def wk_fft2(self, d_data: int, na: int, nr: int, inverse: bool = False, stream: int = 0) -> None:
        check(self.lib.sarbp_wk_fft2(self.handle, C.c_void_p(d_data), int(na), int(nr), int(bool(inverse)),
                                     C.c_void_p(stream)), "sarbp_wk_fft2")

    def wk_reference_function(self, d_theta: int, fr_axis, fa_axis, *, f0: float, c: float, vr: float,
                              kr: float, r0: float, na: int, nr: int, stream: int = 0) -> None:
        import numpy as np
        fra = np.ascontiguousarray(fr_axis, dtype=np.float64)
        faa = np.ascontiguousarray(fa_axis, dtype=np.float64)
        if fra.size != nr or faa.size != na:
            raise ValueError("axis lengths must be nr and na")
        check(self.lib.sarbp_wk_reference_function(self.handle, C.c_void_p(d_theta),
                                                   fra.ctypes.data_as(C.c_void_p), faa.ctypes.data_as(C.c_void_p),
                                                   float(f0), float(c), float(vr), float(kr), float(r0), int(na),
                                                   int(nr), C.c_void_p(stream)), "sarbp_wk_reference_function")

    def omega_k(self, d_echo: int, na_in: int, nr_in: int, wp: WkParams, d_image: int,
                stream: int = 0) -> None:
        check(self.lib.sarbp_omega_k(self.handle, C.c_void_p(d_echo), int(na_in), int(nr_in), C.byref(wp),
                                     C.c_void_p(d_image), C.c_void_p(stream)), "sarbp_omega_k")

    def set_tuning(self, interp: int, precision: int, dtype: int, rows: int, cols: int,
                   pulses: int) -> None:
        t = Tuning(rows, cols, pulses)
        check(self.lib.sarbp_set_tuning(self.handle, interp, precision, dtype, C.byref(t)),
              "sarbp_set_tuning")

    def get_tuning(self, interp: int, precision: int, dtype: int) -> tuple:
        t = Tuning()
        check(self.lib.sarbp_get_tuning(self.handle, interp, precision, dtype, C.byref(t)),
              "sarbp_get_tuning")
        return t.rows_per_lane, t.cols_per_lane, t.pulses_per_stage

    def set_class_tuning(self, interp: int, precision: int, dtype: int, size_class: int, rows: int,
                         cols: int, pulses: int) -> None:
        t = Tuning(rows, cols, pulses)
        check(self.lib.sarbp_set_class_tuning(self.handle, interp, precision, dtype, size_class, C.byref(t)),
              "sarbp_set_class_tuning")

    def get_class_tuning(self, interp: int, precision: int, dtype: int, size_class: int) -> tuple:
        t = Tuning()
        check(self.lib.sarbp_get_class_tuning(self.handle, interp, precision, dtype, size_class, C.byref(t)),
              "sarbp_get_class_tuning")
        return t.rows_per_lane, t.cols_per_lane, t.pulses_per_stage

    def autotune(self, d_profiles: int, d_pos: int, h_pos, grid: Grid, params: Params,
                 d_scratch: int, scratch_pixels: int, max_crop_px: int = 1024,
                 stream: int = 0) -> dict:
        """Runs the library autotuner; the winner becomes this context's tuning."""
        import numpy as np
        hp = None
        if h_pos is not None:
            self._tune_pos = np.ascontiguousarray(h_pos, dtype=np.float64)
            hp = self._tune_pos.ctypes.data_as(C.c_void_p)
        r = AutotuneResult()
        check(self.lib.sarbp_autotune(self.handle, C.c_void_p(d_profiles), C.c_void_p(d_pos), hp,
                                      C.byref(grid), C.byref(params), C.c_void_p(d_scratch),
                                      int(scratch_pixels), int(max_crop_px), C.c_void_p(stream),
                                      C.byref(r)), "sarbp_autotune")
        return {"rows_per_lane": r.best.rows_per_lane, "cols_per_lane": r.best.cols_per_lane,
                "pulses_per_stage": r.best.pulses_per_stage, "best_ms": r.best_ms,
                "default_ms": r.default_ms, "n_candidates": r.n_candidates,
                "crop": [r.crop_nx, r.crop_ny]}

    def plan(self, h_pos, grid: Grid, params: Params) -> "Plan":
        return Plan(self, h_pos, grid, params)

    def backproject_batch(self, jobs: Sequence[Job], stream: int = 0) -> None:
        arr = (Job * len(jobs))(*jobs)
        check(self.lib.sarbp_backproject_batch(self.handle, arr, len(jobs), C.c_void_p(stream)),
              "sarbp_backproject_batch")

    def backproject_batch_noalloc(self, jobs: Sequence[Job], stream: int = 0) -> None:
        """sarbp_backproject_batch_noalloc: the same batch without allocating."""
        arr = (Job * max(len(jobs), 1))(*jobs)
        check(self.lib.sarbp_backproject_batch_noalloc(self.handle, arr, len(jobs), C.c_void_p(stream)),
              "sarbp_backproject_batch_noalloc")


def sinc8_coefficients(beta: float):
    """The library's sinc8 polynomial coefficients as a numpy [2 halves][6 powers][8 taps]."""
    import numpy as np
    out = (C.c_float * 96)()
    check(load().sarbp_sinc8_coefficients(float(beta), out), "sarbp_sinc8_coefficients")
    return np.frombuffer(out, dtype=np.float32).reshape(2, 6, 8).copy()


def sinc8_coefficients_whole(beta: float):
    """The whole-bin sinc8 kernel's coefficients as a numpy [8 powers][8 taps]."""
    import numpy as np
    out = (C.c_float * 64)()
    check(load().sarbp_sinc8_coefficients_whole(float(beta), out), "sarbp_sinc8_coefficients_whole")
    return np.frombuffer(out, dtype=np.float32).reshape(8, 8).copy()


SIZE_CLASSES = 5


def size_class(grid: Grid, params: Params, h_pos=None) -> int:
    """The tuning size class of a job (sarbp_size_class); h_pos: optional host
    float64 [n_pulses][3] positions (the planned bound)."""
    import numpy as np
    hp = None
    if h_pos is not None:
        h_pos = np.ascontiguousarray(h_pos, dtype=np.float64)
        hp = h_pos.ctypes.data_as(C.c_void_p)
    return int(load().sarbp_size_class(C.byref(grid), C.byref(params), hp))


def rotation_probe(d_phase: int, d_out: int, n: int, stream: int = 0) -> None:
    """Enqueues the BP kernels' phase rotation on n FP32 phases (revolutions):
    d_out holds n (cos, sin) float pairs (sarbp_rotation_probe)."""
    check(load().sarbp_rotation_probe(C.c_void_p(d_phase), C.c_void_p(d_out), C.c_int64(n),
                                      C.c_void_p(stream)), "sarbp_rotation_probe")


def valu_probe(kind: int, n_blocks: int, iters: int, d_out: int, stream: int = 0) -> None:
    """Enqueues the vector-ALU issue-rate probe (sarbp_valu_probe)."""
    check(load().sarbp_valu_probe(int(kind), int(n_blocks), int(iters), C.c_void_p(d_out),
                                  C.c_void_p(stream)), "sarbp_valu_probe")


class Plan:
    def __init__(self, ctx: Context, h_pos, grid: Grid, params: Params):
        import numpy as np
        self.ctx = ctx
        self.lib = ctx.lib
        ptr = None
        if h_pos is not None:
            self._pos = np.ascontiguousarray(h_pos, dtype=np.float64)
            ptr = self._pos.ctypes.data_as(C.c_void_p)
        h = C.c_void_p()
        check(self.lib.sarbp_plan_create(ctx.handle, ptr, C.byref(grid), C.byref(params),
                                         C.byref(h)), "sarbp_plan_create")
        self.handle = h

    def info(self) -> dict:
        i = PlanInfo()
        check(self.lib.sarbp_plan_get_info(self.handle, C.byref(i)), "sarbp_plan_get_info")
        return {f: getattr(i, f) for f, _ in PlanInfo._fields_}

    def execute(self, d_profiles: int, d_pos: int, d_image: int, stream: int = 0) -> None:
        check(self.lib.sarbp_plan_execute(self.handle, C.c_void_p(d_profiles), C.c_void_p(d_pos),
                                          C.c_void_p(d_image), C.c_void_p(stream)),
              "sarbp_plan_execute")

    def close(self) -> None:
        if self.handle:
            self.lib.sarbp_plan_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _info_dict(i: JobInfo) -> dict:
    return {f: getattr(i, f) for f, _ in JobInfo._fields_}


class Server:
    """Continuous-batching BP server (sarbp_server_*).  Host arrays passed to
    submit() are kept referenced until the job is released."""

    def __init__(self, *, device=0, interp="sinc8", precision="ref", dtype="c64",
                 hbm_budget_bytes=8 << 30, max_active_jobs=0, pulses_per_epoch=1024,
                 rows_per_lane=0, cols_per_lane=0, pulses_per_stage=0, epoch_watchdog_s=0.0):
        self.lib = load()
        c = ServerConfig()
        c.device = device
        c.interp = INTERP_NAMES[interp] if isinstance(interp, str) else int(interp)
        c.precision = PREC_NAMES[precision] if isinstance(precision, str) else int(precision)
        c.dtype = DTYPE_NAMES[dtype] if isinstance(dtype, str) else int(dtype)
        c.hbm_budget_bytes = int(hbm_budget_bytes)
        c.max_active_jobs = int(max_active_jobs)
        c.pulses_per_epoch = int(pulses_per_epoch)
        c.rows_per_lane = int(rows_per_lane)
        c.cols_per_lane = int(cols_per_lane)
        c.pulses_per_stage = int(pulses_per_stage)
        c.epoch_watchdog_s = float(epoch_watchdog_s)
        h = C.c_void_p()
        check(self.lib.sarbp_server_create(C.byref(c), C.byref(h)), "sarbp_server_create")
        self.handle = h
        self.config = c
        self._keep = {}

    def set_class_tuning(self, size_class: int, rows: int, cols: int, pulses: int) -> None:
        t = Tuning(rows, cols, pulses)
        check(self.lib.sarbp_server_set_class_tuning(self.handle, int(size_class), C.byref(t)),
              "sarbp_server_set_class_tuning")

    def tune_like(self, ctx: "Context") -> None:
        """Takes every size class's tuning in effect on ctx (e.g. autotuned)."""
        c = self.config
        for k in range(SIZE_CLASSES):
            r, co, p = ctx.get_class_tuning(c.interp, c.precision, c.dtype, k)
            self.set_class_tuning(k, r, co, p)

    def submit(self, profiles, pos, grid: Grid, params: Params, image) -> int:
        """profiles: host array (complex64, or uint32 bf16x2 words); pos: float64
        [n][3]; image: complex64 [ny][nx] host array written when done."""
        r = Request()
        r.h_profiles = profiles.ctypes.data if profiles.size else None
        r.h_pos = pos.ctypes.data if pos.size else None
        r.grid = grid
        r.params = params
        r.h_image = image.ctypes.data if image.size else None
        jid = C.c_uint64()
        check(self.lib.sarbp_server_submit(self.handle, C.byref(r), C.byref(jid)),
              "sarbp_server_submit")
        self._keep[jid.value] = (profiles, pos, image)
        return jid.value

    def submit_raw(self, raw, range_params: RangeParams, pos, grid: Grid, params: Params, image) -> int:
        """raw: host complex64 [n_pulses][n_raw] echoes, range-compressed on the
        device after admission (params.n_samples / r0 / dr come from
        range_params)."""
        r = RawRequest()
        r.h_raw = raw.ctypes.data if raw.size else None
        r.range = range_params
        r.h_pos = pos.ctypes.data if pos.size else None
        r.grid = grid
        r.params = params
        r.h_image = image.ctypes.data if image.size else None
        jid = C.c_uint64()
        check(self.lib.sarbp_server_submit_raw(self.handle, C.byref(r), C.byref(jid)),
              "sarbp_server_submit_raw")
        self._keep[jid.value] = (raw, pos, image)
        return jid.value

    def poll(self, jid: int) -> dict:
        i = JobInfo()
        st = self.lib.sarbp_server_poll(self.handle, jid, C.byref(i))
        if st not in (OK, ERR_NOT_READY):
            check(st, "sarbp_server_poll")
        return _info_dict(i)

    def wait(self, jid: int, timeout_s: float = -1.0, raise_on_error: bool = True) -> dict:
        i = JobInfo()
        st = self.lib.sarbp_server_wait(self.handle, jid, float(timeout_s), C.byref(i))
        if raise_on_error:
            check(st, "sarbp_server_wait")
        d = _info_dict(i)
        d["wait_status"] = st
        return d

    def cancel(self, jid: int) -> None:
        check(self.lib.sarbp_server_cancel(self.handle, jid), "sarbp_server_cancel")

    def release(self, jid: int) -> None:
        check(self.lib.sarbp_server_release(self.handle, jid), "sarbp_server_release")
        self._keep.pop(jid, None)

    def stats(self) -> dict:
        s = ServerStats()
        check(self.lib.sarbp_server_get_stats(self.handle, C.byref(s)), "sarbp_server_get_stats")
        return {f: getattr(s, f) for f, _ in ServerStats._fields_}

    def drain(self, timeout_s: float = -1.0) -> None:
        check(self.lib.sarbp_server_drain(self.handle, float(timeout_s)), "sarbp_server_drain")

    def close(self) -> None:
        if self.handle:
            self.lib.sarbp_server_destroy(self.handle)
            self.handle = C.c_void_p()
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
