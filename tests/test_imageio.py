"""Image files in the reference's saveResult format (raw interleaved complex64,
row-major; code/SAR_BP/CUDA/BPA_CUDA.cpp:92-100) with a JSON sidecar."""
import numpy as np
import pytest

from sar_cambrianaccelerationcard_amd import imageio


def test_roundtrip_and_raw_bytes(tmp_path):
    rng = np.random.default_rng(0)
    img = (rng.standard_normal((5, 7)) + 1j * rng.standard_normal((5, 7))).astype(np.complex64)
    p = str(tmp_path / "sar_image.bin")
    side = imageio.save_image(p, img, {"interp": "sinc8"})
    assert side["ny"] == 5 and side["nx"] == 7
    raw = np.fromfile(p, dtype=np.float32)           # interleaved re, im
    assert np.array_equal(raw[0::2].reshape(5, 7), img.real)
    assert np.array_equal(raw[1::2].reshape(5, 7), img.imag)
    assert np.array_equal(imageio.load_image(p), img)


def test_corruption_detected(tmp_path):
    p = str(tmp_path / "x.bin")
    imageio.save_image(p, np.ones((2, 2), np.complex64))
    with open(p, "r+b") as f:
        f.write(b"\x01")
    with pytest.raises(ValueError):
        imageio.load_image(p)


def test_magnitude_db():
    img = np.array([[1.0, 0.1], [0.0, 0.01]], dtype=np.complex64)
    db = imageio.magnitude_db(img)
    assert db[0, 0] == 0.0 and abs(db[0, 1] + 20) < 1e-5 and db[1, 0] == -60.0
