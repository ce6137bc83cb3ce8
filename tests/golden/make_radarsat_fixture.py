"""Writes tests/golden/radarsat_cd_crop.npz: a crop of the RADARSAT-1 raw data
the reference ships (code/SAR_WK/wk_v5.2.3/data/CDdata{1,2}.mat, 2 x 1536
lines of 2048 range cells; lines LINE0..LINE0+N_LINES of the two blocks in
order) as 4-bit I/Q codes (one byte per sample), with the run parameters of CD_run_params.mat and the
Doppler centroid estimated over both data blocks.  The files are read with
scipy.io.loadmat (MAT parsing only, nothing executed).  Run in a container
that has the reference repository; the GPU tests read only the fixture.
"""
import json
import os
import sys

import numpy as np
import scipy.io as sio

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from sar_cambrianaccelerationcard_amd import radarsat as RS  # noqa: E402

SRC = "/root/reference/code/SAR_WK/wk_v5.2.3/data"
LINE0, N_LINES = 1600, 256


def main():
    params = sio.loadmat(os.path.join(SRC, "CD_run_params.mat"))
    rp = RS.RunParams.from_mat(params)
    blocks = [sio.loadmat(os.path.join(SRC, "CDdata%d.mat" % k))["data"] for k in (1, 2)]
    full = np.concatenate(blocks)
    fdc = RS.doppler_centroid(full, rp)
    crop = full[LINE0:LINE0 + N_LINES]
    codes = RS.pack4(crop)
    meta = {"source": "code/SAR_WK/wk_v5.2.3/data/CDdata1.mat + CDdata2.mat", "line0": LINE0, "n_lines": N_LINES,
            "n_lines_total": int(full.shape[0]), "run_params": rp.as_dict(), "fdc": fdc,
            "nrepl": int(np.asarray(params["Nrepl"]).ravel()[0])}
    out = os.path.join(HERE, "radarsat_cd_crop.npz")
    np.savez_compressed(out, codes=codes, meta=np.array(json.dumps(meta)))
    print(out, codes.shape, os.path.getsize(out), "bytes; fdc %.1f Hz" % fdc)


if __name__ == "__main__":
    main()
