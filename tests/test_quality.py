"""The quality CLI on a synthetic point response: it finds the target,
reports the PSLR / ISLR of a sinc response, compares with a reference and
writes a PNG."""
import json
import os

import numpy as np

from sar_cambrianaccelerationcard_amd import imageio as io
from sar_cambrianaccelerationcard_amd import quality as Q


def _sinc_image(n=96, x0=40.3, y0=51.7):
    x = np.arange(n) - x0
    y = np.arange(n) - y0
    return (np.sinc(y)[:, None] * np.sinc(x)[None, :]).astype(np.complex64)


def test_quality_report_of_a_sinc(tmp_path, capsys):
    img = _sinc_image()
    p = str(tmp_path / "img.bin")
    io.save_image(p, img)
    assert Q.main([p, "--auto", "1", "--ref", p, "--png", str(tmp_path / "img.png")]) == 0
    out = json.loads(capsys.readouterr().out)
    t = out["targets"][0]
    assert t["peak"] == [40, 52]
    for k in ("pslr_x_db", "pslr_y_db"):
        assert abs(t[k] - (-13.26)) < 0.3       # uniform-aperture sinc: -13.26 dB
    assert out["rel_l2_vs_ref"] == 0.0 and out["vs_ref"][0]["peak_match"]
    with open(tmp_path / "img.png", "rb") as f:
        assert f.read(8) == b"\x89PNG\r\n\x1a\n"


def test_brightest_keeps_distance():
    img = np.zeros((64, 64), np.complex64)
    img[20, 20], img[22, 22], img[40, 45] = 5, 4, 3
    assert Q.brightest(img, 2, guard=8) == [(20, 20), (45, 40)]
