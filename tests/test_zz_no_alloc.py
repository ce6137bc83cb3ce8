"""Allocation-free batched launch (sarbp_backproject_batch_noalloc).

Its images must be bitwise those of sarbp_backproject_batch for the same jobs,
on two calls in a row with the second smaller (fewer, smaller jobs, so stale
work-list entries of the first call would show), with grids that need two tile
shapes and both sinc8 kernels; and it must refuse more jobs than its fixed
scratch holds."""
import numpy as np
import pytest

from sar_cambrianaccelerationcard_amd import device as D
from sar_cambrianaccelerationcard_amd import native
from sar_cambrianaccelerationcard_amd import scene as S

pytestmark = pytest.mark.gpu


def _scenes(sizes, seed0):
    base = S.load_config("config0")
    out = []
    for i, (npl, nx, ny, spacing) in enumerate(sizes):
        cfg = S.scaled_config(base, n_pulses=npl, nx=nx, ny=ny)
        cfg["seed"] = seed0 + i
        cfg["image"]["interp"] = "sinc8"
        cfg["image"]["spacing"] = spacing
        out.append(D.DeviceScene(S.make_scene(cfg)))
    return out


def _job(ds, img):
    sc = ds.scene
    j = native.Job()
    j.d_profiles = ds.profiles_ptr("c64")
    j.d_pos = ds.pos.data_ptr()
    j.h_pos = sc.pos.ctypes.data
    j.grid = sc.native_grid()
    j.params = sc.native_params(interp="sinc8")
    j.d_image = img.data_ptr()
    return j


def _run_both(ctx, dss):
    import torch
    stream = torch.cuda.current_stream().cuda_stream
    ref = [ds.new_image() for ds in dss]
    got = [ds.new_image().fill_(float("nan")) for ds in dss]  # must be overwritten
    ctx.backproject_batch([_job(ds, im) for ds, im in zip(dss, ref)], stream)
    ctx.backproject_batch_noalloc([_job(ds, im) for ds, im in zip(dss, got)], stream)
    torch.cuda.synchronize()
    return [r.cpu().numpy() for r in ref], [g.cpu().numpy() for g in got]


def test_noalloc_equals_batch_on_two_calls(ctx):
    # (pulses, nx, ny, pixel spacing m): 1.4 m needs a smaller tile shape than
    # the default, 1.0 m runs the whole-bin sinc8 kernel
    first = _scenes([(96, 64, 80, 0.5), (160, 128, 40, 0.5), (64, 200, 96, 0.25), (130, 70, 64, 1.4),
                     (100, 150, 150, 1.0)], 3000)
    second = _scenes([(64, 50, 40, 0.5), (80, 90, 30, 1.4)], 3100)
    for call, dss in (("first", first), ("second", second)):
        print(f"[noalloc] {call} call, {len(dss)} jobs", flush=True)
        ref, got = _run_both(ctx, dss)
        for k, (r, g) in enumerate(zip(ref, got)):
            assert np.isfinite(g).all(), (call, k)
            assert np.array_equal(g.view(np.uint64), r.view(np.uint64)), (call, k)
            assert np.abs(r).max() > 0, (call, k)


def test_noalloc_limits(ctx):
    with pytest.raises(native.SarbpError) as e:
        ctx.backproject_batch_noalloc([native.Job() for _ in range(257)], 0)
    assert e.value.status == native.ERR_INVALID_ARGUMENT
    ctx.backproject_batch_noalloc([], 0)  # nothing to do
