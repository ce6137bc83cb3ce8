"""GPU range compression (sarbp_range_compress) against the FP64 oracle
(oracle/range_oracle.py) on the same complex64 raw echoes.

Tolerance: rel-L2 <= 1e-5 over each output (FP32 Stockham FFTs with FP64-built
tables; the FP32 rounding of the complex64 input is shared by both sides).
bf16x2 output must equal the host's round-to-nearest-even conversion of the
complex64 output bit for bit.  Range compression followed by BP must hold the
BP bound (rel-L2 <= 1e-4) against the oracle BP of FP64-compressed profiles.
"""
import numpy as np
import pytest

from sar_cambrianaccelerationcard_amd import device as D
from sar_cambrianaccelerationcard_amd import metrics as M
from sar_cambrianaccelerationcard_amd import native
from sar_cambrianaccelerationcard_amd import scene as S

import oracle as O
import range_oracle as R

pytestmark = pytest.mark.gpu

TOL = 1e-5
C_LIGHT = 299792458.0

# (n_raw, pulse_width, fs, upsample, crop_offset, n_out, n_pulses): FFT sizes 2 .. 8192
CASES = [
    (1, 1e-9, 1e9, 1, 0, 2, 8),            # nf 2
    (2, 1e-9, 1e9, 4, 0, 16, 8),           # nf 4
    (3, 2e-9, 1e9, 2, 1, 6, 8),            # nf 8
    (10, 5e-9, 1e9, 4, 0, 64, 16),         # nf 16
    (20, 8e-9, 1e9, 2, 0, 64, 16),         # nf 32
    (40, 10e-9, 1e9, 3, 7, 150, 32),       # nf 64
    (100, 20e-9, 1e9, 8, 0, 1024, 32),     # nf 128
    (200, 50e-9, 1e9, 16, 100, 3000, 32),  # nf 256
    (400, 100e-9, 1e9, 4, 3, 2000, 64),    # nf 512
    (624, 2e-6, 200e6, 4, 0, 2048, 64),    # nf 1024 (512-pulse X-band scene)
    (300, 2e-6, 200e6, 1, 0, 1024, 64),    # nf 1024, no upsampling
    (1200, 2e-6, 200e6, 4, 5, 4000, 64),   # nf 2048
    (2448, 2e-6, 200e6, 4, 0, 8192, 64),   # nf 4096 (4096-pulse X-band scene)
    (4496, 2e-6, 200e6, 4, 0, 16384, 32),  # nf 8192 (large served job)
    # the register-FFT path (nf >= 2048) without upsampling and with L = 2
    (1200, 2e-6, 200e6, 1, 0, 2048, 32),   # nf 2048, L 1
    (1200, 2e-6, 200e6, 2, 3, 4000, 32),   # nf 2048, L 2, cropped
    (2448, 2e-6, 200e6, 1, 0, 4096, 16),   # nf 4096, L 1
    (4496, 2e-6, 200e6, 2, 0, 16384, 16),  # nf 8192, L 2
]


def _radar(n_raw, Tp, fs, L, crop, n_out):
    return dict(c=C_LIGHT, bandwidth=0.75 * fs, pulse_width=Tp, fs=fs, t0=3.2e-5, n_raw=n_raw,
                upsample=L, crop_offset=crop, n_out=n_out)


def _raw(M, n, seed):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal((M, n)) + 1j * rng.standard_normal((M, n))).astype(np.complex64)


def _rel(got, ref):
    return float(np.linalg.norm(got - ref) / max(np.linalg.norm(ref), 1e-300))


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"nraw{c[0]}_L{c[3]}_out{c[5]}")
def test_random_raw_matches_oracle(ctx, case):
    n_raw, Tp, fs, L, crop, n_out, M_ = case
    rd = _radar(n_raw, Tp, fs, L, crop, n_out)
    nf = native.range_info(D.range_params(rd))[0]
    print(f"[range] n_raw={n_raw} nf={nf} L={L} crop={crop} n_out={n_out} pulses={M_}", flush=True)
    raw = _raw(M_, n_raw, n_raw + L)
    got = D.range_compress(ctx, raw, rd).cpu().numpy()
    ref = R.range_compress(raw.astype(np.complex128), **rd)
    if np.linalg.norm(ref) == 0.0:            # chirp shorter than one sample
        assert np.abs(got).max() == 0.0
        return
    err = _rel(got, ref)
    print(f"  rel_l2={err:.3e}", flush=True)
    assert err <= TOL


def test_full_scene_size(ctx):
    """4096 pulses x 2448 raw samples -> 8192 upsampled samples."""
    rd = _radar(2448, 2e-6, 200e6, 4, 0, 8192)
    raw = _raw(4096, 2448, 11)
    print("[range] 4096 x 2448 -> 8192", flush=True)
    got = D.range_compress(ctx, raw, rd).cpu().numpy()
    ref = R.range_compress(raw.astype(np.complex128), chunk=256, **rd)
    err = _rel(got, ref)
    worst = max(_rel(got[m], ref[m]) for m in range(0, 4096, 97))
    print(f"  rel_l2={err:.3e} worst sampled pulse {worst:.3e}", flush=True)
    assert err <= TOL and worst <= TOL


def test_bf16_output_is_rounded_c64(ctx):
    rd = _radar(624, 2e-6, 200e6, 4, 0, 2048)
    raw = _raw(48, 624, 5)
    c64 = D.range_compress(ctx, raw, rd).cpu().numpy()
    bf = D.range_compress(ctx, raw, rd, dtype="bf16x2").cpu().numpy().view(np.uint32)
    assert np.array_equal(bf, D.to_bf16x2(c64))


def test_writes_only_the_window(ctx):
    """Every output sample is written and nothing past [n_pulses][n_out]."""
    torch = D.require_gpu()
    rd = _radar(1200, 2e-6, 200e6, 4, 5, 4000)
    M_ = 24
    buf = torch.full((M_ * 4000 + 4096,), float("nan"), dtype=torch.complex64, device="cuda")
    raw = torch.from_numpy(_raw(M_, 1200, 8)).cuda()
    ctx.range_compress(raw.data_ptr(), M_, D.range_params(rd), buf.data_ptr(),
                       torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    h = buf.cpu().numpy()
    assert np.all(np.isfinite(h[: M_ * 4000]))
    assert np.all(np.isnan(h[M_ * 4000:].real))


def test_then_backproject_matches_oracle(ctx):
    """Raw echoes -> device range compression -> device BP (sinc8) against the
    oracle BP of the FP64-compressed profiles."""
    torch = D.require_gpu()
    cfg = S.scaled_config(S.load_config("config0"), n_pulses=128, nx=192, ny=160)
    sc = S.make_scene(cfg, keep_f64=True, keep_raw=True)
    print(f"[range+bp] raw {sc.raw.shape} -> {sc.profiles.shape}", flush=True)
    prof = D.range_compress(ctx, sc.raw, sc.radar)
    # the scene's profiles are this compression plus nothing else (no clutter in the config)
    assert _rel(prof.cpu().numpy(), sc.profiles64) <= TOL
    ds = D.DeviceScene(sc)
    ds.prof.copy_(prof)
    img = D.backproject(ctx, ds, interp="sinc8").cpu().numpy()
    p = O.params(sc.c, sc.fc, sc.r0, sc.dr, "sinc8", sc.kaiser_beta)
    ref = O.backproject(sc.profiles64, sc.pos, sc.grid, p)
    err = M.rel_l2(img, ref)
    print(f"  rel_l2={err:.3e}", flush=True)
    assert err <= 1e-4
    torch.cuda.synchronize()


def test_validation(ctx):
    torch = D.require_gpu()
    rd = _radar(624, 2e-6, 200e6, 4, 0, 2048)
    raw = torch.zeros((4, 624), dtype=torch.complex64, device="cuda")
    out = torch.zeros((4, 2048), dtype=torch.complex64, device="cuda")
    s = torch.cuda.current_stream().cuda_stream
    ctx.range_compress(raw.data_ptr(), 0, D.range_params(rd), out.data_ptr(), s)   # no-op
    for bad, status in [(dict(upsample=0), native.ERR_INVALID_ARGUMENT),
                        (dict(upsample=17), native.ERR_INVALID_ARGUMENT),
                        (dict(n_out=8193 - 0, crop_offset=1), native.ERR_INVALID_ARGUMENT),
                        (dict(n_raw=0), native.ERR_INVALID_ARGUMENT),
                        (dict(fs=0.0), native.ERR_INVALID_ARGUMENT),
                        (dict(n_raw=8000), native.ERR_UNSUPPORTED)]:
        with pytest.raises(native.SarbpError) as e:
            ctx.range_compress(raw.data_ptr(), 4, D.range_params(rd, **bad), out.data_ptr(), s)
        assert e.value.status == status, bad
    with pytest.raises(native.SarbpError):
        ctx.range_compress(0, 4, D.range_params(rd), out.data_ptr(), s)
    with pytest.raises(native.SarbpError):
        ctx.range_compress(raw.data_ptr(), -1, D.range_params(rd), out.data_ptr(), s)
    rp = D.range_params(rd)
    rp.dtype = 7
    with pytest.raises(native.SarbpError):
        ctx.range_compress(raw.data_ptr(), 4, rp, out.data_ptr(), s)
    torch.cuda.synchronize()
