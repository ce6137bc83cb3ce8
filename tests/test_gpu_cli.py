"""The sarbp_run command-line imager (the reference's SAR_Processor flow on
this library): it runs, writes the raw complex64 image the reference's
saveResult writes plus a sidecar, focuses each unit target at its pixel with
amplitude pulses x Tp fs, and its image equals the FP64 oracle chain (same
echoes, range-compression oracle, BP oracle) to rel-L2 <= 1e-4."""
import ctypes as C
import json
import math
import os
import subprocess

import numpy as np
import pytest

from sar_cambrianaccelerationcard_amd import imageio as io
from sar_cambrianaccelerationcard_amd import metrics as M
from sar_cambrianaccelerationcard_amd import scene as S

import oracle as O
import range_oracle as R

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "sar_cambrianaccelerationcard_amd", "_lib", "sarbp_run")


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def test_sarbp_run_matches_oracle_chain(tmp_path):
    n_p, n_s, n, sp = 512, 2048, 160, 0.5
    out = str(tmp_path / "img.bin")
    print(f"[cli] sarbp_run {n_p} pulses {n}x{n}", flush=True)
    r = subprocess.run([EXE, "--pulses", str(n_p), "--samples", str(n_s), "--nx", str(n), "--ny", str(n),
                        "--spacing", str(sp), "--interp", "linear", "--no-noise", "--out", out],
                       capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    img = io.load_image(out, verify=False)
    meta = json.load(open(out + ".json"))["meta"]
    assert img.shape == (n, n)
    # the same scene on the host: echoes -> FP64 range compression -> FP64 BP
    L = S.sim()
    c, fs, up = 299792458.0, 200e6, 4
    span = -(-n_s // up)
    radar = S.Radar(c, 9.6e9, 150e6, 2e-6, fs, 2.0 * (math.hypot(4000.0, 3000.0) - 0.5 * span * (0.5 * c / fs)) / c,
                    span + 400, up, 0, n_s, 0.0)
    pos = np.zeros((n_p, 3))
    L.sarsim_track_straight(_dp(pos), n_p, 150.0, 1000.0, 0.0, 0.0, 3000.0)
    tg = np.array([[(b - 1) * 20.0, 4000.0 + (a - 1) * 20.0, 0.0, 1.0, 0.0] for a in range(3) for b in range(3)])
    raw = np.zeros((n_p, radar.n_raw), dtype=np.complex128)
    L.sarsim_echo_raw(_dp(pos), n_p, _dp(np.ascontiguousarray(tg)), len(tg), C.byref(radar),
                      _dp(raw.view(np.float64)), 1)
    prof = R.range_compress(raw.astype(np.complex64).astype(np.complex128), bandwidth=150e6, pulse_width=2e-6,
                            fs=fs, upsample=up, crop_offset=0, n_out=n_s)
    assert abs(meta["r0_m"] - float(L.sarsim_out_r0(C.byref(radar)))) < 1e-6
    grid = {"origin": meta["origin"], "ux": [sp, 0, 0], "uy": [0, sp, 0], "nx": n, "ny": n}
    ref = O.backproject(prof, pos, grid, O.params(c, 9.6e9, meta["r0_m"], meta["dr_m"], "linear"))
    err = M.rel_l2(img, ref)
    print(f"  rel_l2 vs oracle chain {err:.3e}", flush=True)
    assert err <= 1e-4
    for t in tg:
        ix = int(round((t[0] - grid["origin"][0]) / sp))
        iy = int(round((t[1] - grid["origin"][1]) / sp))
        gx, gy = M.local_peak(img, ix, iy, 3)
        assert (gx, gy) == (ix, iy)
        # neighbouring targets 20 m away add a few % through their sidelobes
        assert 0.95 < abs(img[iy, ix]) / (n_p * 400.0) < 1.05


def test_sarbp_run_rejects_bad_options(tmp_path):
    r = subprocess.run([EXE, "--interp", "cubic", "--out", str(tmp_path / "x.bin")], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 1 and "bad interp" in r.stderr


WK_EXE = os.path.join(ROOT, "sar_cambrianaccelerationcard_amd", "_lib", "sarwk_run")


def _write_paras(path, s):
    # the key: value format of the reference's data/paras.txt (read.cc:54-82)
    with open(path, "w") as f:
        f.write(f"Fr: {s.fs:.6f}\nPRF: {s.prf:.6f}\nf0: {s.f0:.6f}\nTr: {s.tr:.9f}\nR0: {s.r0:.9f}\n"
                f"Kr: {s.kr:.6f}\nc: {s.c:.6f}\n")


@pytest.mark.parametrize("fmt", ["txt", "bin"])
def test_sarwk_run_matches_oracle(tmp_path, fmt):
    import wk_oracle as W
    from sar_cambrianaccelerationcard_amd import omegak as OK
    scene = OK.WkScene(f_nc=60.0).resolved()
    targets = [(194, 170), (294, 235), (394, 300)]
    echo = OK.simulate(scene, targets)
    paras, ech = str(tmp_path / "paras.txt"), str(tmp_path / f"echo.{fmt}")
    _write_paras(paras, scene)
    if fmt == "txt":
        np.savetxt(ech, echo.view(np.float32), fmt="%.9e")  # "re im re im ..." per line
    else:
        echo.tofile(ech)
    out = str(tmp_path / f"result.{fmt}")
    print(f"[cli] sarwk_run {echo.shape} {fmt}", flush=True)
    r = subprocess.run([WK_EXE, "--params", paras, "--echo", ech, "--na", str(scene.na_in), "--nr",
                        str(scene.nr_in), "--vr", repr(scene.vr), "--f-nc", repr(scene.f_nc), "-i", "-o", out],
                       capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0 and "first 5 x 5 values" in r.stdout
    na, nr = scene.padded
    if fmt == "txt":
        img = np.loadtxt(out, dtype=np.float32).view(np.complex64)
    else:
        img = io.load_image(out, verify=False)
    assert img.shape == (na, nr)
    # the parameter file carries 6 decimals, as the reference's paras.txt
    p = {k: float(v) for k, v in (l.split(":") for l in open(paras))}
    ref = W.omega_k(echo.astype(np.complex128), c=p["c"], f0=p["f0"], fs=p["Fr"], prf=p["PRF"], kr=p["Kr"],
                    vr=scene.vr, r0=p["R0"], f_nc=scene.f_nc)
    err = M.rel_l2(img, ref)
    print(f"  rel_l2 vs omega-K oracle {err:.3e}", flush=True)
    assert err <= 2e-5


def test_sarwk_run_centred_range_reference(tmp_path):
    # --range-ref centre: R0 is the range of the first echo sample, targets
    # in both halves of the range window focus where the geometry puts them.
    import wk_oracle as W
    from dataclasses import replace
    from sar_cambrianaccelerationcard_amd import omegak as OK
    scene = OK.WkScene(f_nc=60.0, r0=2300.0)
    _, nr = scene.padded
    targets = [(194, 170), (294, 235), (330, nr // 2 + 150)]
    echo = OK.simulate(scene, targets)
    paras, ech, out = str(tmp_path / "paras.txt"), str(tmp_path / "echo.bin"), str(tmp_path / "result.bin")
    _write_paras(paras, replace(scene, r0=scene.first_sample_range))
    echo.tofile(ech)
    print(f"[cli] sarwk_run --range-ref centre {echo.shape}", flush=True)
    r = subprocess.run([WK_EXE, "--params", paras, "--echo", ech, "--na", str(scene.na_in), "--nr",
                        str(scene.nr_in), "--vr", repr(scene.vr), "--f-nc", repr(scene.f_nc), "--range-ref", "centre",
                        "-i", "-o", out], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0 and "range reference centre" in r.stdout
    img = io.load_image(out, verify=False)
    p = {k: float(v) for k, v in (l.split(":") for l in open(paras))}
    ref = W.omega_k(echo.astype(np.complex128), c=p["c"], f0=p["f0"], fs=p["Fr"], prf=p["PRF"], kr=p["Kr"],
                    vr=scene.vr, r0=p["R0"], f_nc=scene.f_nc, range_ref=W.REF_CENTRE)
    assert M.rel_l2(img, ref) <= 2e-5
    for t in targets:
        m = OK.peak_metrics(img, *t)
        assert (m["d_line"], m["d_sample"]) == (0, 0) and m["concentration"] > 0.25, m


def test_sarwk_run_rejects_bad_input(tmp_path):
    paras = str(tmp_path / "paras.txt")
    from sar_cambrianaccelerationcard_amd import omegak as OK
    _write_paras(paras, OK.WkScene().resolved())
    ech = str(tmp_path / "echo.bin")
    np.zeros((4, 5), dtype=np.complex64).tofile(ech)
    r = subprocess.run([WK_EXE, "--params", paras, "--echo", ech, "--na", "4", "--nr", "6"], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 1 and "expected 24" in r.stderr
    r = subprocess.run([WK_EXE, "--window", "hann"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "--window" in r.stderr
    r = subprocess.run([WK_EXE, "--range-ref", "middle"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "--range-ref" in r.stderr


def test_sarbp_run_reference_scene(tmp_path):
    # The reference's own chain and inputs: with its defaults the echo and the
    # image are zero (fast time ends 2.5 us after transmission, the echoes
    # arrive after ~67 us); a scene within its 375 m window gives the image of
    # the same two library calls made here, bit for bit.
    import ctypes as C
    import torch
    import echo_oracle as E
    from sar_cambrianaccelerationcard_amd import device, native
    out = str(tmp_path / "ref.bin")
    print("[cli] sarbp_run --reference-scene", flush=True)
    r = subprocess.run([EXE, "--reference-scene", "--out", out], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0 and "(0 non-zero samples)" in r.stdout
    img = io.load_image(out, verify=False)
    assert img.shape == (1024, 300) and not np.any(img)

    out = str(tmp_path / "near.bin")
    print("[cli] sarbp_run --reference-scene --rg0 250 --h 200 --la 60 --mslow 256", flush=True)
    r = subprocess.run([EXE, "--reference-scene", "--rg0", "250", "--h", "200", "--la", "60", "--mslow", "256",
                        "--out", out], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    img = io.load_image(out, verify=False)
    p = native.SarParams.reference_defaults()
    p.Rg0, p.H, p.La, p.Mslow = 250.0, 200.0, 60.0, 256
    pt, ta, tr = E.reference_scene(p)
    echo = device.generate_echo(pt, ta, tr, p)
    assert torch.count_nonzero(echo).item() > 0
    d_img = torch.zeros_like(echo)
    native.check(native.load().sarbp_launch_backprojection(C.c_void_p(echo.data_ptr()), C.c_void_p(d_img.data_ptr()),
                                                          p, p.Mslow, p.Nfast), "sarbp_launch_backprojection")
    ref = d_img.cpu().numpy()
    assert img.shape == ref.shape == (256, 300) and np.any(ref)
    assert np.array_equal(img.view(np.uint64), ref.view(np.uint64))
