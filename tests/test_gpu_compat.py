"""The reference-compatible entry point sarbp_launch_backprojection(Srnm,
Image, SARParams, Na, Nr) (drop-in for launchBackProjection,
code/SAR_BP/CUDA/BPA_CUDA_kernels.cu:91-99): same arguments, the reference
host driver's geometry (BPA_CUDA.cpp:35-48), BPA_SAR.m's nearest-bin sum."""
import ctypes as C

import numpy as np
import pytest

from sar_cambrianaccelerationcard_amd import metrics as M
from sar_cambrianaccelerationcard_amd import native

import oracle as O

pytestmark = pytest.mark.gpu


def _reference_geometry(p, na, nr):
    m = np.arange(p.Mslow)
    ta = -p.La / (2 * p.vr) + (p.La / p.vr) * m / (p.Mslow - 1)
    pos = np.stack([np.float64(p.vr) * ta, np.zeros(p.Mslow), np.full(p.Mslow, np.float64(p.H))], 1)
    grid = {"origin": [np.float64(p.Az0) - na // 2, np.float64(p.Rg0) - nr // 2, 0.0],
            "ux": [0.0, 1.0, 0.0], "uy": [1.0, 0.0, 0.0], "nx": nr, "ny": na}
    return pos, grid


def test_compat_entry_point_matches_explicit_call_and_oracle():
    import torch
    lib = native.load()
    p = native.SarParams.reference_defaults()
    # The reference's fast-time axis starts at 0 (BPA_CUDA.cpp:44-48), so put
    # the scene in range of it: 300 samples at 1.25 m -> 0..375 m.
    p.H, p.Rg0, p.Mslow, p.Nfast, p.La = 200.0, 250.0, 256, 300, 60.0
    na, nr = 96, 80
    rng = np.random.default_rng(4)
    srnm = (rng.standard_normal((p.Mslow, p.Nfast)) + 1j * rng.standard_normal((p.Mslow, p.Nfast))
            ).astype(np.complex64)
    d_s = torch.from_numpy(srnm).cuda()
    d_img = torch.zeros((na, nr), dtype=torch.complex64, device="cuda")
    native.check(lib.sarbp_launch_backprojection(C.c_void_p(d_s.data_ptr()),
                                                 C.c_void_p(d_img.data_ptr()), p, na, nr),
                 "sarbp_launch_backprojection")
    got = d_img.cpu().numpy()
    pos, grid = _reference_geometry(p, na, nr)
    ctx = native.Context(0)
    params = native.make_params(c=p.c, fc=p.fc, r0=0.0, dr=np.float64(p.c) / (2 * np.float64(p.Fr)),
                                n_pulses=p.Mslow, n_samples=p.Nfast, interp="nearest")
    d_pos = torch.from_numpy(pos).cuda()
    d2 = torch.zeros_like(d_img)
    ctx.backproject(d_s.data_ptr(), d_pos.data_ptr(),
                    native.make_grid(grid["origin"], grid["ux"], grid["uy"], nr, na), params,
                    d2.data_ptr(), sync=True)
    assert np.array_equal(got.view(np.uint64), d2.cpu().numpy().view(np.uint64))
    ref = O.backproject(srnm, pos, grid, O.params(np.float64(p.c), np.float64(p.fc), 0.0,
                                                  np.float64(p.c) / (2 * np.float64(p.Fr)), "nearest"))
    assert M.rel_l2(got, ref) < 1e-3  # nearest bin: a few half-bin ties round differently


def test_compat_rejects_bad_arguments():
    lib = native.load()
    p = native.SarParams.reference_defaults()
    assert lib.sarbp_launch_backprojection(None, None, p, 4, 4) == native.ERR_INVALID_ARGUMENT
