"""Real data on the device: the RADARSAT-1 raw crop the reference ships
(tests/golden/radarsat_cd_crop.npz, 256 lines x 2048 cells of 4-bit I/Q)
range-compressed by sarbp_range_compress with the reference's down-chirp and
back-projected in the slant plane (989 km range, 6 m x 4.64 m pixels) with the
REF precision mode, against the FP64 oracle chain on the same codes.
Bounds: profiles rel-L2 <= 1e-5, images rel-L2 <= 1e-4."""
import os

import numpy as np
import pytest

from sar_cambrianaccelerationcard_amd import device as D
from sar_cambrianaccelerationcard_amd import metrics as M
from sar_cambrianaccelerationcard_amd import native
from sar_cambrianaccelerationcard_amd import radarsat as RS

import oracle as O
import range_oracle as R

pytestmark = pytest.mark.gpu

CROP = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "radarsat_cd_crop.npz")
L, N_OUT = 2, 2800


@pytest.fixture(scope="module")
def crop():
    raw, rp, meta = RS.load_crop(CROP)
    radar = dict(c=rp.c, bandwidth=-rp.Kr * rp.Tr, pulse_width=rp.Tr, fs=rp.Fr, t0=2.0 * rp.R0 / rp.c,
                 n_raw=raw.shape[1], upsample=L, crop_offset=0, n_out=N_OUT)
    prof64 = R.range_compress(raw.astype(np.complex128), **radar)
    return raw, rp, meta, radar, prof64


def test_range_compression_of_real_echoes(ctx, crop):
    raw, rp, meta, radar, prof64 = crop
    print(f"[radarsat] {raw.shape} -> {N_OUT} samples", flush=True)
    got = D.range_compress(ctx, raw, radar).cpu().numpy()
    err = np.linalg.norm(got - prof64) / np.linalg.norm(prof64)
    print(f"  rel_l2={err:.3e}", flush=True)
    assert err <= 1e-5


@pytest.mark.parametrize("interp", ["linear", "sinc8"])
def test_slant_plane_bp_of_real_echoes(ctx, crop, interp):
    import torch
    raw, rp, meta, radar, prof64 = crop
    n_tot, l0 = meta["n_lines_total"], meta["line0"]
    pos = RS.positions(rp, l0, l0 + raw.shape[0], n_tot)
    r = rp.R0 + 350 * rp.c / (2 * rp.Fr)
    xc = RS.beam_centre_x(rp, meta["fdc"], l0, l0 + raw.shape[0], n_tot, r)
    g = RS.region_grid(xc, r, 128, 96, 6.0, 4.64)
    dr = rp.c / (2 * L * rp.Fr)
    print(f"[radarsat] BP {interp} {g['nx']}x{g['ny']} at 6 m x 4.64 m, {raw.shape[0]} pulses", flush=True)
    prof = D.range_compress(ctx, raw, radar)
    d_pos = torch.from_numpy(pos).cuda()
    params = native.make_params(c=rp.c, fc=rp.f0, r0=rp.R0, dr=dr, n_pulses=raw.shape[0], n_samples=N_OUT,
                                interp=interp, precision="ref")
    ng = native.make_grid(g["origin"], g["ux"], g["uy"], g["nx"], g["ny"])
    plan = ctx.plan(pos, ng, params)
    info = plan.info()
    img = torch.zeros((g["ny"], g["nx"]), dtype=torch.complex64, device="cuda")
    plan.execute(prof.data_ptr(), d_pos.data_ptr(), img.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    ref = O.backproject(prof64, pos, g, O.params(rp.c, rp.f0, rp.R0, dr, interp))
    err = M.rel_l2(img.cpu().numpy(), ref)
    print(f"  tile {info['tile_nx']}x{info['tile_ny']} window {info['window']} rel_l2={err:.3e}", flush=True)
    assert err <= 1e-4


def test_imaging_command_on_the_crop(tmp_path):
    """python -m sar_cambrianaccelerationcard_amd.radarsat --crop ...: writes the
    image (+ sidecar) and its dB PNG."""
    import subprocess
    import sys
    from sar_cambrianaccelerationcard_amd import imageio as io
    out = str(tmp_path / "cd")
    r = subprocess.run([sys.executable, "-m", "sar_cambrianaccelerationcard_amd.radarsat", "--crop", CROP,
                        "--out", out, "--nx", "128", "--ny", "96", "--dx", "6"],
                       capture_output=True, text=True, timeout=300,
                       cwd=os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    print(r.stdout, r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    img = io.load_image(out + ".bin")
    assert img.shape == (96, 128) and np.all(np.isfinite(img)) and np.abs(img).max() > 0
    assert os.path.getsize(out + ".png") > 0


@pytest.mark.parametrize("range_ref", ["r0", "centre"])
def test_omega_k_imaging_command_on_the_crop(tmp_path, range_ref):
    """The same command with --omega-k: the crop's 256 lines imaged by
    sarbp_omega_k (the reference's omega-K driver runs on this scene), equal
    to the library call on the same lines."""
    import subprocess
    import sys
    from sar_cambrianaccelerationcard_amd import device as D
    from sar_cambrianaccelerationcard_amd import imageio as io
    from sar_cambrianaccelerationcard_amd import native
    from sar_cambrianaccelerationcard_amd import omegak as OK
    out = str(tmp_path / "cdwk")
    r = subprocess.run([sys.executable, "-m", "sar_cambrianaccelerationcard_amd.radarsat", "--crop", CROP,
                        "--out", out, "--omega-k", "--range-ref", range_ref], capture_output=True, text=True,
                       timeout=300,
                       cwd=os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    print(r.stdout, r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    img = io.load_image(out + ".bin")
    raw, rp, _ = RS.load_crop(CROP)
    assert img.shape == native.wk_padded_size(*raw.shape) and os.path.getsize(out + ".png") > 0
    c = native.Context(0)
    ref = D.omega_k(c, raw, OK.radarsat_params(rp, range_ref=range_ref)).cpu().numpy()
    c.close()
    assert np.array_equal(img, ref)
