"""The FP64 echo oracle (oracle/echo_oracle.py, generateEchoKernel of
code/SAR_BP/CUDA/BPA_CUDA_kernels.cu:3-45) pinned without retyping its
formula: the reference driver's own scene (all zero, its fast-time axis
starts at 0), gate counts in closed form, superposition, and a matched filter
that must put each pulse's peak at the target's delay with phase -4 pi R /
lambda."""
import math
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import echo_oracle as E  # noqa: E402

from sar_cambrianaccelerationcard_amd import native  # noqa: E402


def _params(**over):
    p = native.SarParams.reference_defaults()
    for k, v in over.items():
        setattr(p, k, v)
    return p


def test_reference_driver_scene_is_all_zero():
    # BPA_CUDA.cpp:25-48: tr = j / Fr from 0 while the targets sit ~10 km
    # away (2 R / c ~ 67 us >> Nfast / Fr = 2.5 us): no sample is inside a pulse.
    p = _params()
    pt, ta, tr = E.reference_scene(p)
    assert pt.shape == (5, 3) and ta.shape == (1024,) and tr.shape == (300,)
    assert ta[0] == np.float32(-p.La / (2 * p.vr)) and abs(ta[-1] - p.La / (2 * p.vr)) < 1e-6
    echo = E.generate_echo(pt, ta, tr, p)
    assert echo.shape == (1024, 300) and not np.any(echo)


def _window(p, r_near, n):
    """Fast times starting 40 samples before the echo of range r_near."""
    t0 = np.float32(2 * r_near / p.c - 40 / p.Fr)
    return (t0 + np.arange(n, dtype=np.float32) / np.float32(p.Fr)).astype(np.float32)


def test_gate_counts_match_closed_form():
    p = _params(Ntarget=1, La=40.0)
    pt = np.array([[3.0, 9000.0, 1.0]], dtype=np.float32)
    ta = np.linspace(-0.2, 0.2, 41).astype(np.float32)
    r_mid = math.sqrt(9000.0 ** 2 + p.H ** 2)
    tr = _window(p, r_mid, 500)
    echo = E.generate_echo(pt, ta, tr, p)
    for i, t_a in enumerate(ta.astype(np.float64)):
        az = t_a * np.float64(np.float32(p.vr)) - 3.0
        r = math.sqrt(az * az + 9000.0 ** 2 + np.float64(np.float32(p.H)) ** 2)
        tau = 2 * r / np.float64(np.float32(p.c))
        inside = abs(az) < np.float64(np.float32(p.La)) / 2
        want = int(np.sum((tr.astype(np.float64) > tau) & (tr.astype(np.float64) < tau + np.float64(np.float32(p.Tr)))))
        got = int(np.count_nonzero(echo[i]))
        assert got == (want if inside else 0), (i, got, want, inside)
        if inside:  # one pulse length of samples, +-1 for the float32 grid's jitter
            assert abs(want - np.float64(np.float32(p.Tr)) * np.float64(np.float32(p.Fr))) <= 1.01


def test_superposition_and_linearity():
    p = _params(Ntarget=2, La=60.0)
    a = np.array([[-4.0, 9000.0, 1.0]], dtype=np.float32)
    b = np.array([[7.0, 9012.5, 0.5]], dtype=np.float32)
    ta = np.linspace(-0.15, 0.15, 31).astype(np.float32)
    tr = _window(p, math.sqrt(9000.0 ** 2 + p.H ** 2), 400)
    both = E.generate_echo(np.concatenate([a, b]), ta, tr, p)
    pa = _params(Ntarget=1, La=60.0)
    ea, eb = E.generate_echo(a, ta, tr, pa), E.generate_echo(b, ta, tr, pa)
    assert np.max(np.abs(both - ea - eb)) < 1e-12
    a2 = a.copy()
    a2[0, 2] = 3.0
    assert np.max(np.abs(E.generate_echo(a2, ta, tr, pa) - 3.0 * ea)) < 1e-12


def _matched_filter(p):
    """Broadside pulse (ta = 0) of one target correlated with the transmitted
    chirp exp(j pi Kr t^2), 0 < t < Tr, of the default radar: (peak sample -
    delay in samples, |peak| / chirp length, peak phase - (-4 pi R / lambda))."""
    q = _params(Ntarget=1, La=40.0)
    fs = np.float64(np.float32(q.Fr))
    h = np.float64(np.float32(q.H))
    y = 9000.0
    r = math.sqrt(y * y + h * h)
    tau = 2 * r / np.float64(np.float32(q.c))
    t0 = np.float32(tau - 40 / fs)
    tr = (t0 + np.arange(600, dtype=np.float32) / np.float32(q.Fr)).astype(np.float32)
    echo = E.generate_echo(np.array([[0.0, y, 1.0]], dtype=np.float32), np.zeros(1, dtype=np.float32), tr, p)[0]
    kr, tw = np.float64(np.float32(q.Kr)), np.float64(np.float32(q.Tr))
    n_ch = int(tw * fs)
    tc = np.arange(n_ch) / fs
    chirp = np.exp(1j * np.pi * kr * tc * tc)
    corr = np.array([np.vdot(chirp, echo[k:k + n_ch]) for k in range(echo.size - n_ch)])
    k = int(np.argmax(np.abs(corr)))
    want = -4 * np.pi * r / np.float64(np.float32(q.lambda_))
    return k - (tau - np.float64(t0)) * fs, np.abs(corr[k]) / n_ch, np.angle(corr[k] * np.exp(-1j * want))


def test_matched_filter_peak_at_delay_with_range_phase():
    # The peak is at the sample of the round-trip delay and its phase is
    # -4 pi R / lambda; the residual chirp phase at the sub-sample offset is
    # below pi Kr / fs^2 per sample of offset.
    p = _params(Ntarget=1, La=40.0)
    dk, gain, dph = _matched_filter(p)
    kr, fs = np.float64(np.float32(p.Kr)), np.float64(np.float32(p.Fr))
    assert abs(dk) <= 1.0 and gain > 0.9, (dk, gain)
    assert abs(dph) < 4 * np.pi * kr / fs ** 2 + 0.3, dph


def test_matched_filter_catches_planted_mistakes():
    # A reversed chirp smears the peak; a reversed range phase (lambda -> -lambda)
    # moves its phase by 8 pi R / lambda.
    p = _params(Ntarget=1, La=40.0)
    p.Kr = -p.Kr
    assert _matched_filter(p)[1] < 0.5
    p = _params(Ntarget=1, La=40.0)
    p.lambda_ = -p.lambda_
    dk, gain, dph = _matched_filter(p)
    kr, fs = np.float64(np.float32(p.Kr)), np.float64(np.float32(p.Fr))
    assert gain > 0.9 and abs(dph) >= 4 * np.pi * kr / fs ** 2 + 0.3, dph
