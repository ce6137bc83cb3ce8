"""omega-K on the device against the FP64 oracle (oracle/wk_oracle.py):
sarbp_stolt_sinc on random lines (both windows, several tap counts, the
reference's own 2 x 3 test input, the RADARSAT axes), sarbp_omega_k on
simulated point scenes, the RADARSAT-1 crop and the full 3072 x 2048 scene
size.  Tolerances: Stolt alone 3e-6 of the largest output magnitude (FP32 taps
and weights, FP64 centres); whole pipeline 2e-5 relative L2 (FP32 transforms
of up to 4096 x 3414 points)."""
import json
import os

import numpy as np
import pytest

import wk_oracle as W
from sar_cambrianaccelerationcard_amd import native
from sar_cambrianaccelerationcard_amd import omegak as O

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def _torch():
    import torch
    return torch


def _dev(x):
    torch = _torch()
    return torch.from_numpy(np.ascontiguousarray(x, dtype=np.complex64)).to("cuda:0")


def _stolt(ctx, S, fr, fa, out=None, **kw):
    from sar_cambrianaccelerationcard_amd import device
    d = _dev(S)
    r = device.stolt_sinc(ctx, d, fr, fa, out=(d if out == "inplace" else None), **kw)
    return r.cpu().numpy()


def _rand(shape, seed):
    r = np.random.default_rng(seed)
    return (r.standard_normal(shape) + 1j * r.standard_normal(shape)).astype(np.complex64)


def _check_stolt(got, ref):
    err = np.max(np.abs(got - ref))
    assert err <= 3e-6 * np.max(np.abs(ref)), (err, np.max(np.abs(ref)))


@pytest.mark.parametrize("window", ["floor", "centered"])
@pytest.mark.parametrize("taps", [2, 6, 16])
def test_stolt_matches_oracle(ctx, window, taps):
    na, nr = 37, 300
    S = _rand((na, nr), 10 + taps)
    fs, f0, c, vr = 1e6, 1e9, 3e8, 150.0
    fr = W.fftshift_axis(nr) * fs / nr
    # Centres move by 0.01 to 21 bins and stay off integers (the floor window's
    # taps jump where a centre crosses one).
    fa = np.linspace(-11.3, 11.9, na)
    kw = dict(f0=f0, c=c, vr=vr, fs=fs, taps=taps)
    got = _stolt(ctx, S, fr, fa, window=window, **kw)
    ref = W.stolt_sinc(S, fr, fa, window=W.FLOOR if window == "floor" else W.CENTERED, **kw)
    _check_stolt(got, ref)


def test_stolt_in_place_equals_out_of_place(ctx):
    S = _rand((16, 1000), 5)
    fr = W.fftshift_axis(1000) * 2e5
    fa = np.linspace(-30.0, 40.0, 16)
    kw = dict(f0=1e9, c=3e8, vr=150.0, fs=2e8, taps=6)
    a = _stolt(ctx, S, fr, fa, **kw)
    b = _stolt(ctx, S, fr, fa, out="inplace", **kw)
    assert np.array_equal(a, b)


def test_stolt_reference_test_input(ctx):
    # wk_v5.1/wk.cc:798-838 calls StoltInterp_sinc on a 2 x 3 all-ones input
    # with fr_axis {1, 2, 3}, fa_axis {0.1, 0.2}, f0 = 10 GHz, c = 3e8,
    # Vr = 100, fr = 1 MHz, P = 6.
    S = np.ones((2, 3), dtype=np.complex64)
    fr, fa = np.array([1.0, 2.0, 3.0]), np.array([0.1, 0.2])
    kw = dict(f0=10e9, c=3e8, vr=100.0, fs=1e6, taps=6)
    got = _stolt(ctx, S, fr, fa, **kw)
    ref = W.stolt_sinc(S, fr, fa, window=W.FLOOR, **kw)
    _check_stolt(got, ref)


def test_stolt_radarsat_axes(ctx):
    # The RADARSAT scene's padded size and axes: centres move by up to ~300 bins.
    na, nr = 4096, 3414
    rows = np.arange(0, na, 64)
    fr, fa, _ = W.axes(na, nr, 32.317e6, 1256.98, -6900.0)
    S = _rand((rows.size, nr), 7)
    kw = dict(f0=5.3e9, c=299790000.0, vr=7062.0, fs=32.317e6, taps=6)
    for window in ("floor", "centered"):
        got = _stolt(ctx, S, fr, fa[rows], window=window, **kw)
        ref = W.stolt_sinc(S, fr, fa[rows], window=W.FLOOR if window == "floor" else W.CENTERED, **kw)
        _check_stolt(got, ref)


def test_stolt_centres_just_below_an_integer(ctx):
    # A zero range-frequency axis and one azimuth frequency chosen so every
    # centre of row 1 is n + 1 - 1e-8: the fraction rounds to 1 in FP32.  The
    # taps stay those of floor(centre) and the weights finite.
    na, nr = 3, 64
    f0, c, vr, fs = 1e9, 3e8, 150.0, 1e6
    k = c * c / (4.0 * vr * vr)
    s = f0 - (1.0 - 1e-8) * fs / nr
    fa = np.array([0.0, np.sqrt((f0 - s) * (f0 + s) / k), 0.0371])
    fr = np.zeros(nr)
    cen = W.centres(fr, fa, f0, c, vr, fs)
    frac = cen[1] - np.floor(cen[1])
    assert np.all(frac < 1.0) and np.all(frac.astype(np.float32) == 1.0)
    S = _rand((na, nr), 21)
    for window in ("floor", "centered"):
        got = _stolt(ctx, S, fr, fa, f0=f0, c=c, vr=vr, fs=fs, taps=6, window=window)
        ref = W.stolt_sinc(S, fr, fa, f0=f0, c=c, vr=vr, fs=fs, taps=6,
                           window=W.FLOOR if window == "floor" else W.CENTERED)
        assert np.all(np.isfinite(got)), window
        _check_stolt(got, ref)
        # the centre ~ n + 1: the output is S[n + 1] to 1e-6 (short of the
        # line's end, where the floor window counts clamped copies of the last
        # sample)
        assert np.max(np.abs(ref[1, :-2] - S[1, 1:-1])) < 1e-6 * np.max(np.abs(S[1]))


def _gridsample(ctx, S, fr_axis, fa_axis, out=None, **kw):
    d = _dev(S)
    o = d if out == "inplace" else _torch().empty_like(d)
    ctx.stolt_gridsample(d.data_ptr(), o.data_ptr(), fr_axis, fa_axis, na=S.shape[0], nr=S.shape[1], **kw)
    _torch().cuda.synchronize()
    return o.cpu().numpy()


@pytest.mark.parametrize("nr", [300, 301])
def test_stolt_gridsample_matches_oracle(ctx, nr):
    # The reference's GridSample Stolt step (task/stolt.cc:38-318): rotated
    # rows, FP32 grid, linear interpolation with zeros outside.
    na = 37
    S = _rand((na, nr), 31 + nr)
    fs = 1e6
    fr = W.fftshift_axis(nr) * fs / nr
    fa = np.linspace(-11.3, 11.9, na)
    kw = dict(f0=1e9, c=3e8, vr=150.0, fr=fs / nr)
    got = _gridsample(ctx, S, fr, fa, **kw)
    ref = W.stolt_gridsample(S, fr, fa, **kw)
    err = np.max(np.abs(got - ref))
    assert err <= 1e-6 * np.max(np.abs(S)), err
    assert np.array_equal(_gridsample(ctx, S, fr, fa, out="inplace", **kw), got)


def test_stolt_gridsample_radarsat_axes(ctx):
    na, nr = 4096, 3414
    rows = np.arange(0, na, 128)
    fr, fa, _ = W.axes(na, nr, 32.317e6, 1256.98, -6900.0)
    S = _rand((rows.size, nr), 8)
    kw = dict(f0=5.3e9, c=299790000.0, vr=7062.0, fr=32.317e6 / nr)
    got = _gridsample(ctx, S, fr, fa[rows], **kw)
    ref = W.stolt_gridsample(S, fr, fa[rows], **kw)
    assert np.max(np.abs(got - ref)) <= 1e-6 * np.max(np.abs(S))


def test_stolt_gridsample_rejects_bad_arguments(ctx):
    torch = _torch()
    d = torch.zeros((2, 8), dtype=torch.complex64, device="cuda:0")
    fr, fa = np.zeros(8), np.zeros(2)
    for kw in (dict(fr=0.0), dict(c=0.0), dict(vr=0.0)):
        args = dict(f0=1e9, c=3e8, vr=100.0, na=2, nr=8, fr=1e3)
        args.update(kw)
        with pytest.raises(native.SarbpError) as e:
            ctx.stolt_gridsample(d.data_ptr(), d.data_ptr(), fr, fa, **args)
        assert e.value.status == native.ERR_INVALID_ARGUMENT, kw
    with pytest.raises(native.SarbpError) as e:
        ctx.stolt_gridsample(d.data_ptr(), d.data_ptr(), np.zeros(9000), fa, f0=1e9, c=3e8, vr=100.0, na=2,
                             nr=9000, fr=1e3)
    assert e.value.status == native.ERR_UNSUPPORTED


def test_stolt_rejects_bad_arguments(ctx):
    torch = _torch()
    d = torch.zeros((2, 8), dtype=torch.complex64, device="cuda:0")
    fr, fa = np.zeros(8), np.zeros(2)
    for kw, st in ((dict(taps=5), native.ERR_INVALID_ARGUMENT), (dict(taps=18), native.ERR_INVALID_ARGUMENT),
                   (dict(window=3), native.ERR_INVALID_ARGUMENT), (dict(fs=0.0), native.ERR_INVALID_ARGUMENT)):
        args = dict(f0=1e9, c=3e8, vr=100.0, na=2, nr=8, fs=1e6, taps=6, window=0)
        args.update(kw)
        with pytest.raises(native.SarbpError) as e:
            ctx.stolt_sinc(d.data_ptr(), d.data_ptr(), fr, fa, **args)
        assert e.value.status == st
    big = np.zeros(9000)
    with pytest.raises(native.SarbpError) as e:
        ctx.stolt_sinc(d.data_ptr(), d.data_ptr(), big, fa, f0=1e9, c=3e8, vr=100.0, na=2, nr=9000, fs=1e6)
    assert e.value.status == native.ERR_UNSUPPORTED


def _omega_k(ctx, echo, wp):
    from sar_cambrianaccelerationcard_amd import device
    return device.omega_k(ctx, echo, wp).cpu().numpy()


def _oracle(echo, wp):
    return W.omega_k(echo, c=wp.c, f0=wp.f0, fs=wp.fs, prf=wp.prf, kr=wp.kr, vr=wp.vr, r0=wp.r0, f_nc=wp.f_nc,
                     taps=wp.taps, window=wp.window, range_ref=wp.range_ref)


def _rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("f_nc", [0.0, 60.0])
@pytest.mark.parametrize("window", ["floor", "centered"])
def test_omega_k_point_scene(ctx, f_nc, window):
    scene = O.WkScene(f_nc=f_nc).resolved()
    targets = [(194, 170), (294, 235), (394, 300)]
    echo = O.simulate(scene, targets)
    wp = scene.params(window=window)
    got = _omega_k(ctx, echo, wp)
    ref = _oracle(echo, wp)
    assert got.shape == ref.shape == scene.padded
    assert _rel_l2(got, ref) < 2e-5
    for t in targets:
        mg, mr = O.peak_metrics(got, *t), O.peak_metrics(ref, *t)
        assert (mg["d_line"], mg["d_sample"]) == (mr["d_line"], mr["d_sample"]) == (0, 0)
        assert abs(mg["concentration"] - mr["concentration"]) < 1e-3 and mg["concentration"] > 0.25


def test_omega_k_random_echo_odd_sizes(ctx):
    # Unequal pads, odd input sizes (padded sizes odd too), a 2-tap kernel.
    echo = _rand((45, 101), 3)
    scene = O.WkScene(na_in=45, nr_in=101)
    wp = scene.params(taps=2)
    got = _omega_k(ctx, echo, wp)
    ref = _oracle(echo, wp)
    assert got.shape == W.padded_size(45, 101)[:2]
    assert _rel_l2(got, ref) < 2e-5


@pytest.mark.parametrize("range_ref", ["r0", "centre"])
def test_omega_k_radarsat_crop(ctx, range_ref):
    from sar_cambrianaccelerationcard_amd import radarsat as RS
    z = np.load(os.path.join(HERE, "golden", "radarsat_cd_crop.npz"))
    meta = json.loads(str(z["meta"]))
    rp = RS.RunParams(**meta["run_params"])
    echo = RS.unpack4(z["codes"])
    wp = O.radarsat_params(rp, range_ref=range_ref)
    got = _omega_k(ctx, echo, wp)
    ref = _oracle(echo, wp)
    assert got.shape == W.padded_size(*echo.shape)[:2]
    assert _rel_l2(got, ref) < 2e-5


def test_omega_k_full_scene_size(ctx):
    # The reference's scene size (3072 x 2048 -> 4096 x 3414) with its radar
    # and squint; three point targets, oracle at full size.
    scene = O.radarsat_scene()
    targets = [O.aperture_target(scene, a, s) for a, s in ((1500, 1400), (2000, 1550), (2600, 1690))]
    echo = O.simulate(scene, targets)
    wp = scene.params()
    got = _omega_k(ctx, echo, wp)
    ref = _oracle(echo, wp)
    assert got.shape == (4096, 3414)
    assert _rel_l2(got, ref) < 2e-5
    for t in targets:
        mg = O.peak_metrics(got, *t)
        assert (mg["d_line"], mg["d_sample"]) == (0, 0) and mg["concentration"] > 0.4, mg


def test_omega_k_full_scene_size_centred_reference(ctx):
    # The reference's scene size and radar, targets in both halves of the
    # padded range window, padded sample 0 at no whole number of range
    # windows: the window-centred reference focuses all of them.
    from dataclasses import replace
    scene = O.radarsat_scene()
    scene = replace(scene, r0=scene.r0 + 1234.5)
    _, nr = scene.padded
    # whole chirps (1349 samples) inside the echo's range samples [683, 2731),
    # two targets on each side of nr/2 = 1707
    targets = [O.aperture_target(scene, a, s) for a, s in ((1500, 1400), (2000, 1550), (2600, 1900),
                                                           (1800, 2030))]
    echo = O.simulate(scene, targets)
    wp = scene.params(range_ref="centre")
    got = _omega_k(ctx, echo, wp)
    ref = _oracle(echo, wp)
    assert _rel_l2(got, ref) < 2e-5
    for t in targets:
        mg = O.peak_metrics(got, *t)
        assert (mg["d_line"], mg["d_sample"]) == (0, 0) and mg["concentration"] > 0.4, (t, mg)


def test_omega_k_rejects_bad_arguments(ctx):
    torch = _torch()
    d = torch.zeros((8, 8), dtype=torch.complex64, device="cuda:0")
    img = torch.zeros((12, 14), dtype=torch.complex64, device="cuda:0")
    good = O.WkScene(na_in=8, nr_in=8).params()
    for field, val in (("taps", 7), ("window", 2), ("prf", 0.0), ("kr", 0.0), ("range_ref", 2), ("reserved", 1)):
        wp = O.WkScene(na_in=8, nr_in=8).params()
        setattr(wp, field, val)
        with pytest.raises(native.SarbpError) as e:
            ctx.omega_k(d.data_ptr(), 8, 8, wp, img.data_ptr())
        assert e.value.status == native.ERR_INVALID_ARGUMENT, field
    with pytest.raises(native.SarbpError) as e:
        ctx.omega_k(d.data_ptr(), 8, 6000, good, img.data_ptr())
    assert e.value.status == native.ERR_UNSUPPORTED
    with pytest.raises(native.SarbpError):
        ctx.omega_k(0, 8, 8, good, img.data_ptr())


@pytest.mark.parametrize("f_nc", [0.0, 60.0])
def test_bp_and_omega_k_image_the_same_scene(ctx, f_nc):
    # The two imaging algorithms on the same echoes: BP (exact geometry, device
    # range compression) and omega-K put every target at the same line and
    # sample; BP also focuses a target in the far half of the range window,
    # where the reference's omega-K does not.
    scene = O.WkScene(f_nc=f_nc).resolved()
    _, nr = scene.padded
    near = [(194, 170), (294, 235), (394, 300)]
    far = (244, nr // 2 + 60)
    echo = O.simulate(scene, near + [far])
    wk = _omega_k(ctx, echo, scene.params())
    bp = O.bp_image(ctx, scene, echo)
    assert bp.shape == wk.shape
    for t in near:
        mw, mb = O.peak_metrics(wk, *t), O.peak_metrics(bp, *t)
        assert (mw["d_line"], mw["d_sample"]) == (mb["d_line"], mb["d_sample"]) == (0, 0), (t, mw, mb)
        assert mw["concentration"] > 0.25 and mb["concentration"] > 0.25, (t, mw, mb)
    mw, mb = O.peak_metrics(wk, *far), O.peak_metrics(bp, *far)
    assert (mb["d_line"], mb["d_sample"]) == (0, 0) and mb["concentration"] > 0.25, mb
    assert mw["concentration"] < 0.15, mw


def test_omega_k_long_lines(ctx):
    # Padded lines longer than 4096 samples run the line kernel on
    # 16384-point transforms (128 KiB of LDS); same oracle, same tolerance.
    echo = _rand((64, 3100), 11)
    scene = O.WkScene(na_in=64, nr_in=3100)
    wp = scene.params()
    got = _omega_k(ctx, echo, wp)
    ref = _oracle(echo, wp)
    assert got.shape == (86, 5166)
    assert _rel_l2(got, ref) < 2e-5


def test_omega_k_image_at_an_odd_offset(ctx):
    # An image that starts 8 bytes into its buffer (not 16-byte aligned) on the
    # power-of-two azimuth path: the inverse takes one column per workgroup.
    torch = _torch()
    echo = _rand((768, 20), 77)
    wp = O.WkScene(na_in=768, nr_in=20).params()
    na, nr = W.padded_size(768, 20)[:2]
    buf = torch.zeros(na * nr + 1, dtype=torch.complex64, device="cuda:0")
    d_echo = _dev(echo)
    ctx.omega_k(d_echo.data_ptr(), 768, 20, wp, buf.data_ptr() + 8)
    torch.cuda.synchronize()
    got = buf[1:].reshape(na, nr).cpu().numpy()
    assert _rel_l2(got, _oracle(echo, wp)) < 2e-5
    assert buf[0].item() == 0


def test_omega_k_echo_at_an_8_byte_offset(ctx):
    # An echo that starts 8 bytes into its buffer on the power-of-two azimuth
    # path: the forward azimuth transform then takes its one-column kernel
    # (the two-column kernel reads 16-byte pairs).  Against the oracle, and
    # the element before the echo is never written.
    torch = _torch()
    echo = _rand((768, 20), 78)
    wp = O.WkScene(na_in=768, nr_in=20).params()
    na, nr = W.padded_size(768, 20)[:2]
    ebuf = torch.zeros(768 * 20 + 1, dtype=torch.complex64, device="cuda:0")
    ebuf[1:] = torch.from_numpy(np.ascontiguousarray(echo).reshape(-1)).to("cuda:0")
    before = ebuf[0].item()
    out = torch.zeros((na, nr), dtype=torch.complex64, device="cuda:0")
    ctx.omega_k(ebuf.data_ptr() + 8, 768, 20, wp, out.data_ptr())
    torch.cuda.synchronize()
    assert _rel_l2(out.cpu().numpy(), _oracle(echo, wp)) < 2e-5
    assert ebuf[0].item() == before


def test_omega_k_calls_on_two_streams(ctx):
    # Calls on one context share its workspace; issued back to back on two
    # streams they are ordered on the device and both images are right.
    torch = _torch()
    from sar_cambrianaccelerationcard_amd import device
    scene = O.WkScene(f_nc=60.0).resolved()
    e1 = O.simulate(scene, [(194, 170)])
    e2 = O.simulate(scene, [(294, 235)])
    wp = scene.params()
    ref1 = device.omega_k(ctx, e1, wp).cpu().numpy()
    ref2 = device.omega_k(ctx, e2, wp).cpu().numpy()
    d1, d2 = _dev(e1), _dev(e2)
    na, nr = scene.padded
    o1 = torch.empty((na, nr), dtype=torch.complex64, device="cuda:0")
    o2 = torch.empty_like(o1)
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    torch.cuda.synchronize()
    for _ in range(3):
        ctx.omega_k(d1.data_ptr(), scene.na_in, scene.nr_in, wp, o1.data_ptr(), s1.cuda_stream)
        ctx.omega_k(d2.data_ptr(), scene.na_in, scene.nr_in, wp, o2.data_ptr(), s2.cuda_stream)
    torch.cuda.synchronize()
    assert np.array_equal(o1.cpu().numpy(), ref1)
    assert np.array_equal(o2.cpu().numpy(), ref2)


@pytest.mark.parametrize("na_in,nr_in", [(1, 1), (2, 3), (3, 2), (8, 2458), (8, 2459)])
def test_omega_k_edge_sizes(ctx, na_in, nr_in):
    # Single samples, tiny lines, and padded lines of exactly 4096 samples (the
    # longest on 8192-point transforms) and 4099 (16384-point transforms).
    echo = _rand((na_in, nr_in), na_in * 7 + nr_in)
    wp = O.WkScene(na_in=na_in, nr_in=nr_in).params()
    got = _omega_k(ctx, echo, wp)
    ref = _oracle(echo, wp)
    assert got.shape == W.padded_size(na_in, nr_in)[:2]
    assert _rel_l2(got, ref) < 2e-5


@pytest.mark.parametrize("na_in,nr_in", [(768, 20), (1536, 9), (6144, 6), (3100, 6)])
def test_omega_k_power_of_two_azimuth(ctx, na_in, nr_in):
    # Padded azimuth lengths 1024, 2048 and 8192 take the fused pad /
    # down-conversion / LDS azimuth FFT kernel (4096: the full-size test);
    # 4134 (not a power of two, above the chirp-z kernels' 4096) the hipFFT
    # path; the other tests' lengths the chirp-z azimuth kernels.
    echo = _rand((na_in, nr_in), na_in + nr_in)
    wp = O.WkScene(na_in=na_in, nr_in=nr_in, f_nc=60.0).params()
    got = _omega_k(ctx, echo, wp)
    ref = _oracle(echo, wp)
    assert got.shape == W.padded_size(na_in, nr_in)[:2]
    assert _rel_l2(got, ref) < 2e-5


@pytest.mark.parametrize("f_nc", [0.0, 60.0])
@pytest.mark.parametrize("window", ["floor", "centered"])
def test_omega_k_centred_range_reference(ctx, f_nc, window):
    # Targets in both halves of the padded range window and padded sample 0 at
    # no whole number of range windows: the window-centred range reference
    # focuses every target, on the device as in the oracle.
    scene = O.WkScene(f_nc=f_nc, r0=2300.0)
    _, nr = scene.padded
    targets = [(194, 170), (294, 235), (394, 300), (244, nr // 2 + 60), (330, nr // 2 + 150)]
    echo = O.simulate(scene, targets)
    wp = scene.params(window=window, range_ref="centre")
    assert wp.range_ref == native.WK_RANGE_REF_CENTRE
    got = _omega_k(ctx, echo, wp)
    ref = _oracle(echo, wp)
    assert _rel_l2(got, ref) < 2e-5
    for t in targets:
        mg, mr = O.peak_metrics(got, *t), O.peak_metrics(ref, *t)
        assert (mg["d_line"], mg["d_sample"]) == (0, 0) and mg["concentration"] > 0.25, mg
        assert abs(mg["concentration"] - mr["concentration"]) < 1e-3


@pytest.mark.parametrize("na_in,nr_in", [(2, 3), (45, 101), (8, 2458), (8, 2459), (64, 3100), (768, 20)])
def test_omega_k_centred_random_echo(ctx, na_in, nr_in):
    # Odd and even padded line lengths (the output shift's constant phase
    # differs) on 8192- and 16384-point transforms; 768 lines pad to 1024
    # (the fused azimuth kernel).
    echo = _rand((na_in, nr_in), 5 * na_in + nr_in)
    wp = O.WkScene(na_in=na_in, nr_in=nr_in, r0=2300.0).params(range_ref="centre")
    got = _omega_k(ctx, echo, wp)
    ref = _oracle(echo, wp)
    assert got.shape == W.padded_size(na_in, nr_in)[:2]
    assert _rel_l2(got, ref) < 2e-5


# ---- the steps of process_wk one by one (task/task.h:46-73) --------------------------

def _steps_case():
    scene = O.WkScene(na_in=45, nr_in=101, f_nc=60.0)
    echo = _rand((45, 101), 41)
    return scene.resolved(), echo


def test_wk_steps_match_oracle(ctx):
    torch = _torch()
    sc, echo = _steps_case()
    na, nr = sc.padded
    d_echo = _dev(echo)
    pad = torch.empty((na, nr), dtype=torch.complex64, device="cuda:0")
    ctx.wk_preprocess(d_echo.data_ptr(), 45, 101, sc.prf, sc.f_nc, pad.data_ptr())
    torch.cuda.synchronize()
    ref_pad = W.pad_downconvert(echo.astype(np.complex128), prf=sc.prf, f_nc=sc.f_nc)
    assert _rel_l2(pad.cpu().numpy(), ref_pad) < 1e-6
    ctx.wk_fft2(pad.data_ptr(), na, nr)
    torch.cuda.synchronize()
    spec = np.fft.fft2(ref_pad)
    assert _rel_l2(pad.cpu().numpy(), spec) < 2e-6
    ctx.wk_fft2(pad.data_ptr(), na, nr, inverse=True)
    torch.cuda.synchronize()
    assert _rel_l2(pad.cpu().numpy(), ref_pad) < 2e-6
    fr, fa, _ = W.axes(na, nr, sc.fs, sc.prf, sc.f_nc)
    theta = torch.empty((na, nr), dtype=torch.complex64, device="cuda:0")
    kr_arg = -sc.kr  # process_wk passes -Kr (wk.cc:229)
    ctx.wk_reference_function(theta.data_ptr(), fr, fa, f0=sc.f0, c=sc.c, vr=sc.vr, kr=kr_arg, r0=sc.r0,
                              na=na, nr=nr)
    torch.cuda.synchronize()
    want = np.exp(1j * W.reference_phase(fr, fa, f0=sc.f0, c=sc.c, vr=sc.vr, kr=sc.kr, r0=sc.r0))
    assert np.max(np.abs(theta.cpu().numpy() - want)) < 1e-6


def test_wk_steps_compose_to_omega_k(ctx):
    # preprocess -> 2-D FFT -> times the reference function -> sinc Stolt ->
    # 2-D inverse FFT, as process_wk calls them (wk.cc:87-285), gives the
    # image of sarbp_omega_k.
    torch = _torch()
    from sar_cambrianaccelerationcard_amd import device
    sc, echo = _steps_case()
    na, nr = sc.padded
    wp = sc.params()
    d_echo = _dev(echo)
    x = torch.empty((na, nr), dtype=torch.complex64, device="cuda:0")
    ctx.wk_preprocess(d_echo.data_ptr(), 45, 101, sc.prf, sc.f_nc, x.data_ptr())
    ctx.wk_fft2(x.data_ptr(), na, nr)
    fr, fa, _ = W.axes(na, nr, sc.fs, sc.prf, sc.f_nc)
    theta = torch.empty_like(x)
    ctx.wk_reference_function(theta.data_ptr(), fr, fa, f0=sc.f0, c=sc.c, vr=sc.vr, kr=-sc.kr, r0=sc.r0,
                              na=na, nr=nr)
    torch.cuda.synchronize()
    x *= theta
    torch.cuda.synchronize()
    ctx.stolt_sinc(x.data_ptr(), x.data_ptr(), fr, fa, f0=sc.f0, c=sc.c, vr=sc.vr, na=na, nr=nr, fs=sc.fs)
    ctx.wk_fft2(x.data_ptr(), na, nr, inverse=True)
    torch.cuda.synchronize()
    img = device.omega_k(ctx, echo, wp).cpu().numpy()
    assert _rel_l2(x.cpu().numpy(), img) < 2e-5
    assert _rel_l2(x.cpu().numpy(), _oracle(echo, wp)) < 2e-5


def test_wk_steps_reject_bad_arguments(ctx):
    torch = _torch()
    d = torch.zeros((8, 8), dtype=torch.complex64, device="cuda:0")
    for call in (lambda: ctx.wk_preprocess(d.data_ptr(), 4, 4, 0.0, 0.0, d.data_ptr()),
                 lambda: ctx.wk_fft2(0, 8, 8),
                 lambda: ctx.wk_fft2(d.data_ptr(), -1, 8),
                 lambda: ctx.wk_reference_function(d.data_ptr(), np.zeros(8), np.zeros(8), f0=1e9, c=3e8,
                                                   vr=100.0, kr=0.0, r0=1e3, na=8, nr=8)):
        with pytest.raises(native.SarbpError) as e:
            call()
        assert e.value.status == native.ERR_INVALID_ARGUMENT
