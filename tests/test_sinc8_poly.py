"""The sinc8 interpolator evaluates each Kaiser-sinc tap weight, over each
half of the fractional bin, as a degree-5 polynomial in the half's centred
variable e = 2 f - h - 1/2 (host-fitted coefficients, exported by
sarbp_sinc8_coefficients); the whole-bin kernel uses degree 7 in e = f - 1/2.
These CPU tests pin both fits against the oracle's direct evaluation
K(x) = sinc(x) I0(beta sqrt(1-(x/4)^2)) / I0(beta), and the mirror identity
the kernels' coefficient precompute relies on."""
import numpy as np
import pytest

from sar_cambrianaccelerationcard_amd import native

import oracle as O


def _weights(C, f):
    """Tap weights [t][n] at fractional bins f (the kernel's half-bin rule)."""
    h = np.minimum((2 * f).astype(int), 1)
    e = 2 * f - h - 0.5
    powers = np.stack([e ** k for k in range(6)])             # [k][n]
    return np.stack([(C[h[n]].T @ powers[:, n]) for n in range(f.size)], 1)


@pytest.mark.parametrize("beta", [0.0, 2.5, 6.0, 8.0])
def test_poly_matches_kaiser_sinc(beta):
    C = native.sinc8_coefficients(beta).astype(np.float64)   # [h][k][t]
    f = np.linspace(0.0, 1.0, 1001, endpoint=False)
    w = _weights(C, f)
    ref = np.array([[O.kaiser_sinc8(fi + 3 - t, beta) for fi in f] for t in range(8)])
    err = np.abs(w - ref).max()
    assert err < (5e-6 if beta <= 6 else 2e-5), err


def test_halves_join_continuously():
    """Half 0 at e = +1/2 and half 1 at e = -1/2 are the same bin f = 1/2."""
    C = native.sinc8_coefficients(6.0).astype(np.float64)
    end0 = (C[0] * (0.5 ** np.arange(6))[:, None]).sum(axis=0)
    start1 = (C[1] * ((-0.5) ** np.arange(6))[:, None]).sum(axis=0)
    assert np.abs(end0 - start1).max() < 5e-6


def test_poly_partition_of_sample_at_integer_bins():
    """At f = 0 the interpolator returns the centre sample: weights (0,0,0,1,0,0,0,0)."""
    C = native.sinc8_coefficients(6.0).astype(np.float64)
    w = _weights(C, np.array([0.0]))[:, 0]
    expect = np.zeros(8)
    expect[3] = 1.0
    assert np.abs(w - expect).max() < 5e-6


@pytest.mark.parametrize("beta", [0.0, 6.0, 11.5])
def test_mirror_identity_is_exact(beta):
    """K(x) = K(-x): half 1 is half 0 mirrored (tap 7 - t, e -> -e), and the
    whole-bin taps mirror each other, bit for bit; the kernels' precompute
    from mirror sums and differences of the taps assumes it."""
    C = native.sinc8_coefficients(beta)
    sign = np.where(np.arange(6) % 2 == 0, 1.0, -1.0).astype(np.float32)[:, None]
    assert np.array_equal(C[1], sign * C[0][:, ::-1])
    W = native.sinc8_coefficients_whole(beta)
    signw = np.where(np.arange(8) % 2 == 0, 1.0, -1.0).astype(np.float32)[:, None]
    assert np.array_equal(W, signw * W[:, ::-1])


def test_mirror_sum_form_equals_the_tap_sum():
    """The precompute's form -- D0_k = P_k.sigma + Q_k.delta, D1_k =
    (-1)^k (P_k.sigma - Q_k.delta) with sigma_t = s_t + s_(7-t), delta_t =
    s_t - s_(7-t), P = (C_t + C_(7-t))/2, Q = (C_t - C_(7-t))/2; whole bins
    D_k = C_k[t<4].sigma (k even), .delta (k odd) -- is the plain 8-tap sum,
    in FP64 on random complex samples."""
    rng = np.random.default_rng(5)
    s = rng.standard_normal((100, 8)) + 1j * rng.standard_normal((100, 8))
    sig, dl = s[:, :4] + s[:, 7:3:-1], s[:, :4] - s[:, 7:3:-1]
    C = native.sinc8_coefficients(6.0).astype(np.float64)
    P = 0.5 * (C[0][:, :4] + C[0][:, 7:3:-1])
    Q = 0.5 * (C[0][:, :4] - C[0][:, 7:3:-1])
    sign = np.where(np.arange(6) % 2 == 0, 1.0, -1.0)
    assert np.allclose(sig @ P.T + dl @ Q.T, s @ C[0].T, rtol=0, atol=1e-12)
    assert np.allclose(sign * (sig @ P.T - dl @ Q.T), s @ C[1].T, rtol=0, atol=1e-12)
    W = native.sinc8_coefficients_whole(6.0).astype(np.float64)
    form = np.where(np.arange(8) % 2 == 0, sig @ W[:, :4].T, dl @ W[:, :4].T)
    assert np.allclose(form, s @ W.T, rtol=0, atol=1e-12)


def test_bad_beta_rejected():
    with pytest.raises(native.SarbpError):
        native.sinc8_coefficients(float("nan"))


@pytest.mark.parametrize("beta", [0.0, 6.0, 8.0])
def test_whole_bin_poly_matches_kaiser_sinc(beta):
    """The whole-bin kernel's degree-7 tap polynomials (used for tiles coarse
    against the range bins)."""
    C = native.sinc8_coefficients_whole(beta).astype(np.float64)   # [k][t]
    f = np.linspace(0.0, 1.0, 1001)
    w = C.T @ np.stack([(f - 0.5) ** k for k in range(8)])
    ref = np.array([[O.kaiser_sinc8(fi + 3 - t, beta) for fi in f] for t in range(8)])
    err = np.abs(w - ref).max()
    assert err < (5e-6 if beta <= 6 else 2e-5), err
