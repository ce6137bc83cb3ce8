"""Accuracy of the BP kernels' phase rotation on the device.

The pixel loop rotates each interpolated sample by exp(+j 2 pi ph) with the
hardware v_sin_f32 / v_cos_f32 taking ph in revolutions (sarbp_rotation.h,
replacing the sincosf of code/SAR_BP/CUDA/BPA_CUDA_kernels.cu:67).
sarbp_rotation_probe runs that same inline code on given FP32 phases; here
it is held against FP64 cos / sin of 2 pi x for the exact FP32 inputs, over
+-128 revolutions (the kernels keep |ph| below ~90: REF's reference phase,
linear term of up to 128 pixels and the quadratic term; DF and F64 reduce
to [0, 1) or a few revolutions).

Bounds: |error| of cos and sin <= 1e-6, phase error <= 1e-6 rad and
amplitude error <= 1e-6 everywhere; an image rel-L2 of 1e-4 corresponds to an
RMS phase error of ~1e-4 rad, so the rotation holds that budget with two
orders of magnitude to spare."""
import numpy as np
import pytest

from sar_cambrianaccelerationcard_amd import native

pytestmark = pytest.mark.gpu


def _rotate(x32):
    import torch
    ph = torch.from_numpy(np.ascontiguousarray(x32, dtype=np.float32)).cuda()
    out = torch.empty(2 * ph.numel(), dtype=torch.float32, device="cuda")
    native.rotation_probe(ph.data_ptr(), out.data_ptr(), ph.numel(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    o = out.cpu().numpy().astype(np.float64)
    return o[0::2], o[1::2]


def _errors(x32):
    c, s = _rotate(x32)
    x = x32.astype(np.float64)
    frac = x - np.floor(x)  # exact: reduce before the FP64 trigonometry
    c64, s64 = np.cos(2 * np.pi * frac), np.sin(2 * np.pi * frac)
    dphi = np.angle((c + 1j * s) * (c64 - 1j * s64))
    return (np.abs(c - c64).max(), np.abs(s - s64).max(), np.abs(dphi).max(),
            np.abs(np.hypot(c, s) - 1.0).max())


def test_rotation_matches_fp64_over_the_kernel_phase_range():
    rng = np.random.default_rng(7)
    cases = {
        "uniform +-128 rev": rng.uniform(-128.0, 128.0, 1 << 22).astype(np.float32),
        "uniform [0, 1)": rng.uniform(0.0, 1.0, 1 << 20).astype(np.float32),
        # near the quadrant points and integer revolutions, where a reduction
        # error shows most
        "quadrants": (np.repeat(np.arange(-128, 128, 0.25), 64)
                      + rng.uniform(-1e-3, 1e-3, 1024 * 64)).astype(np.float32),
        "small |x|": (rng.uniform(-1, 1, 1 << 16) * np.logspace(-30, -1, 1 << 16)).astype(np.float32),
    }
    for name, x in cases.items():
        ec, es, ep, ea = _errors(x)
        print(f"  {name}: |dcos| {ec:.2e} |dsin| {es:.2e} phase {ep:.2e} rad amplitude {ea:.2e}", flush=True)
        assert ec <= 1e-6 and es <= 1e-6 and ep <= 1e-6 and ea <= 1e-6, name


def test_rotation_exact_points():
    x = np.array([0.0, 0.25, 0.5, 0.75, 1.0, -0.25, -0.5, 64.0, -64.0, 127.75], dtype=np.float32)
    c, s = _rotate(x)
    c64, s64 = np.cos(2 * np.pi * (x % 1.0)), np.sin(2 * np.pi * (x % 1.0))
    assert np.abs(c - c64).max() <= 1e-6 and np.abs(s - s64).max() <= 1e-6
    assert c[0] == 1.0 and s[0] == 0.0


def test_rotation_probe_rejects_bad_arguments():
    with pytest.raises(native.SarbpError):
        native.rotation_probe(0, 0, 4)
    native.rotation_probe(0, 0, 0)  # empty: no launch
