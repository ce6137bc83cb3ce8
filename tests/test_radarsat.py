"""RADARSAT-1 ingest on the CPU: the packed crop decodes to the reference's
data exactly, the run parameters and geometry are consistent, and the FP64
chain (range-compression oracle + BP oracle) behaves as the reference's own
processing choices predict."""
import os

import numpy as np
import pytest

import oracle as O
import range_oracle as R
from sar_cambrianaccelerationcard_amd import radarsat as RS

HERE = os.path.dirname(os.path.abspath(__file__))
CROP = os.path.join(HERE, "golden", "radarsat_cd_crop.npz")
SRC = "/root/reference/code/SAR_WK/wk_v5.2.3/data"


def test_pack_round_trip():
    rng = np.random.default_rng(3)
    v = (2 * rng.integers(0, 16, (7, 9)) - 15) + 1j * (2 * rng.integers(0, 16, (7, 9)) - 15)
    assert np.array_equal(RS.unpack4(RS.pack4(v)), v.astype(np.complex64))
    with pytest.raises(ValueError):
        RS.pack4(np.array([2.0 + 1j]))


@pytest.mark.skipif(not os.path.exists(SRC), reason="reference repository not present")
def test_crop_is_the_reference_data():
    import scipy.io as sio
    raw, rp, meta = RS.load_crop(CROP)
    d = np.concatenate([sio.loadmat(os.path.join(SRC, "CDdata%d.mat" % k))["data"] for k in (1, 2)])
    l0, n = meta["line0"], meta["n_lines"]
    assert np.array_equal(raw, d[l0:l0 + n].astype(np.complex64))


def test_run_params():
    raw, rp, meta = RS.load_crop(CROP)
    assert raw.shape == (256, 2048)
    assert abs(rp.Tr * rp.Fr - meta["nrepl"]) < 1        # chirp samples = replica length
    assert abs(meta["fdc"] - rp.fdc_hint) < rp.PRF / 2   # ambiguity resolved near the hint
    assert abs(RS.doppler_centroid(raw, rp) - meta["fdc"]) < 200


def _region_image(raw, rp, meta, kr_sign, interp="linear", nx=64, ny=64):
    L = 2
    prof = R.range_compress(raw.astype(np.complex128), bandwidth=kr_sign * rp.Kr * rp.Tr, pulse_width=rp.Tr,
                            fs=rp.Fr, upsample=L, crop_offset=0, n_out=L * 1400)
    dr = rp.c / (2 * L * rp.Fr)
    n_tot, l0 = meta["n_lines_total"], meta["line0"]
    pos = RS.positions(rp, l0, l0 + raw.shape[0], n_tot)
    r = rp.R0 + 350 * rp.c / (2 * rp.Fr)
    xc = RS.beam_centre_x(rp, meta["fdc"], l0, l0 + raw.shape[0], n_tot, r)
    g = RS.region_grid(xc, r, nx, ny, 6.0, 4.64)
    return O.backproject(prof, pos, g, O.params(rp.c, rp.f0, rp.R0, dr, interp)), prof, pos, g, dr


def _contrast(img):
    a = np.abs(img) ** 2
    return (a ** 2).mean() / a.mean() ** 2


def test_down_chirp_focuses():
    """The reference uses the chirp as a down-chirp (wk_v5.2.3/wk.cc:229).
    Through range compression and BP of the crop's beam footprint, the down-chirp
    focuses point-like returns (intensity contrast ~10, against 2 for fully
    developed speckle) and the up-chirp does not."""
    raw, rp, meta = RS.load_crop(CROP)
    down = _contrast(_region_image(raw, rp, meta, -1)[0])
    up = _contrast(_region_image(raw, rp, meta, +1)[0])
    assert down > 6.0 and up < 3.0, (down, up)


def test_geometry():
    raw, rp, meta = RS.load_crop(CROP)
    fdc = meta["fdc"]
    n_tot = meta["n_lines_total"]
    pos = RS.positions(rp, 0, n_tot, n_tot)
    assert abs(pos[:, 0].mean()) < 1e-6 and abs(pos[1, 0] - pos[0, 0] - rp.Vr / rp.PRF) < 1e-9
    r = rp.R0 + 700.0
    xc = RS.beam_centre_x(rp, fdc, 1000, 2000, n_tot, r)
    g = RS.region_grid(xc, r, 64, 64, 4.0, 4.64)
    m0, m1 = RS.aperture(rp, fdc, g, n_tot)
    assert m0 < 1500 < m1                                 # the region is lit mid-interval
    foot = rp.wavelength / RS.ANTENNA_LENGTH * r
    assert abs((m1 - m0) * rp.Vr / rp.PRF - (foot + 63 * 4.0)) < 3 * rp.Vr / rp.PRF


def test_fp32_range_loses_the_real_scene():
    """The same sum with FP32 positions and range (the oracle's FP32
    restatement) decorrelates at 989 km slant range: the FP64 reference is
    what the device's REF mode is checked against."""
    raw, rp, meta = RS.load_crop(CROP)
    img, prof, pos, g, dr = _region_image(raw, rp, meta, -1, nx=48, ny=40)
    img32 = O.backproject_f32(prof.astype(np.complex64), pos, g, O.params(rp.c, rp.f0, rp.R0, dr, "linear"))
    assert np.linalg.norm(img32 - img) / np.linalg.norm(img) > 0.5
