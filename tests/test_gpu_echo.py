"""sarbp_launch_generate_echo (drop-in for launchGenerateEcho,
code/SAR_BP/CUDA/BPA_CUDA_kernels.cu:79-89) against the FP64 echo oracle:
the reference driver's own scene (all zero), delayed fast-time windows with
the five targets of BPA_CUDA.cpp:25-48 and with many random targets, argument checks."""
import ctypes as C
import math
import os
import sys

import numpy as np
import pytest

from sar_cambrianaccelerationcard_amd import native

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import echo_oracle as E  # noqa: E402

pytestmark = pytest.mark.gpu


def _gen(pt, ta, tr, p):
    from sar_cambrianaccelerationcard_amd import device
    print(f"[echo] {len(ta)} x {len(tr)}, {p.Ntarget} targets", flush=True)
    return device.generate_echo(pt, ta, tr, p).cpu().numpy()


def _check(got, ref):
    err = np.max(np.abs(got - ref))
    assert err <= 2e-6 * max(1.0, np.max(np.abs(ref))), err


def test_reference_driver_scene():
    p = native.SarParams.reference_defaults()
    pt, ta, tr = E.reference_scene(p)
    got = _gen(pt, ta, tr, p)
    assert got.shape == (1024, 300) and not np.any(got)


def test_reference_targets_in_a_delayed_window():
    # The five targets and slow times of BPA_CUDA.cpp, fast times starting before
    # the nearest echo and long enough for the farthest.
    p = native.SarParams.reference_defaults()
    p.La = 60.0
    pt, ta, _ = E.reference_scene(p)
    r_lo = math.sqrt(np.float64(pt[:, 1].min()) ** 2 + p.H ** 2)
    t0 = np.float32(2 * r_lo / p.c - 20 / p.Fr)
    tr = (t0 + np.arange(1200, dtype=np.float32) / np.float32(p.Fr)).astype(np.float32)
    got = _gen(pt, ta, tr, p)
    ref = E.generate_echo(pt, ta, tr, p)
    assert np.count_nonzero(ref) > 0.2 * ref.size
    _check(got, ref)


def test_random_targets():
    rng = np.random.default_rng(5)
    p = native.SarParams.reference_defaults()
    p.La, p.H, p.Ntarget = 80.0, 3000.0, 24
    pt = np.stack([rng.uniform(-40, 40, 24), rng.uniform(7000, 7300, 24), rng.uniform(0.2, 2.0, 24)],
                  1).astype(np.float32)
    ta = np.linspace(-0.3, 0.3, 333).astype(np.float32)
    t0 = np.float32(2 * math.sqrt(7000.0 ** 2 + p.H ** 2) / p.c - 30 / p.Fr)
    tr = (t0 + np.arange(777, dtype=np.float32) / np.float32(p.Fr)).astype(np.float32)
    got = _gen(pt, ta, tr, p)
    ref = E.generate_echo(pt, ta, tr, p)
    _check(got, ref)


def test_rejects_bad_arguments():
    import torch
    lib = native.load()
    p = native.SarParams.reference_defaults()
    buf = torch.zeros(64, dtype=torch.complex64, device="cuda")
    f = torch.zeros(64, dtype=torch.float32, device="cuda")
    ptr = [C.c_void_p(t.data_ptr()) for t in (buf, f, f, f)]
    assert lib.sarbp_launch_generate_echo(*ptr, p, -1, 4) == native.ERR_INVALID_ARGUMENT
    assert lib.sarbp_launch_generate_echo(None, ptr[1], ptr[2], ptr[3], p, 4, 4) == native.ERR_INVALID_ARGUMENT
    bad = native.SarParams.reference_defaults()
    bad.c = 0.0
    assert lib.sarbp_launch_generate_echo(*ptr, bad, 4, 4) == native.ERR_INVALID_ARGUMENT
    assert lib.sarbp_launch_generate_echo(None, None, None, None, p, 0, 4) == native.OK
