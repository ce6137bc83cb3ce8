"""The sinc8 interpolator evaluates each Kaiser-sinc tap weight as a degree-7
polynomial in the fractional bin f (host-fitted coefficients, exported by
sarbp_sinc8_coefficients).  These CPU tests pin that fit against the oracle's
direct evaluation K(x) = sinc(x) I0(beta sqrt(1-(x/4)^2)) / I0(beta)."""
import numpy as np
import pytest

from sar_cambrianaccelerationcard_amd import native

import oracle as O


@pytest.mark.parametrize("beta", [0.0, 2.5, 6.0, 8.0])
def test_poly_matches_kaiser_sinc(beta):
    C = native.sinc8_coefficients(beta).astype(np.float64)   # [k][t]
    f = np.linspace(0.0, 1.0, 1001)
    powers = np.stack([f ** k for k in range(8)])              # [k][n]
    w = C.T @ powers                                           # [t][n]
    ref = np.array([[O.kaiser_sinc8(fi + 3 - t, beta) for fi in f] for t in range(8)])
    err = np.abs(w - ref).max()
    assert err < (5e-6 if beta <= 6 else 2e-5), err


def test_poly_partition_of_sample_at_integer_bins():
    """At f = 0 the interpolator returns the centre sample: weights (0,0,0,1,0,0,0,0)."""
    C = native.sinc8_coefficients(6.0).astype(np.float64)
    w = C.T @ np.array([0.0 ** k for k in range(8)])
    expect = np.zeros(8)
    expect[3] = 1.0
    assert np.abs(w - expect).max() < 5e-6


def test_bad_beta_rejected():
    with pytest.raises(native.SarbpError):
        native.sinc8_coefficients(float("nan"))
