"""Config-4 served jobs at their real sizes: a large job (8192 x 8192 image at
0.25 m, 8192 pulses x 16384 samples, its own simulated scene of point targets
over clutter) served alone and in a mix with smaller jobs through the
continuous-batching server.

Each job's image must be bitwise the same alone and in the mix, and match the
FP64 oracle to rel-L2 <= 1e-4 on whole rows plus 64 x 64 chips centred on
point targets (the oracle needs ~30 min for the whole large image), with
every chip's brightest pixel the oracle's."""
import numpy as np
import pytest

from sar_cambrianaccelerationcard_amd import native
from sar_cambrianaccelerationcard_amd import workloads as W

import oracle as O
import spot

pytestmark = pytest.mark.gpu

LARGE = W.JobSpec(8192, 0.25, 8192, 16384, True, 20260401)
SMALL = [W.JobSpec(512, 0.5, 1024, 4096, False, 20260402),
         W.JobSpec(700, 0.75, 2048, 3000, False, 20260403),
         W.JobSpec(300, 1.0, 600, 6000, False, 20260404)]
TOL = 1e-4


@pytest.fixture(scope="module")
def scenes():
    print("[config4] simulating the large scene (8192 pulses x 16384 samples)", flush=True)
    large = LARGE.simulate(n_targets=12, n_clutter=4096)
    small = [j.simulate(n_targets=6, n_clutter=512) for j in SMALL]
    return large, small


def _serve(scs, pulses_per_epoch=1024):
    """Serves scenes in submission order on a fresh server; returns the images."""
    srv = native.Server(interp="sinc8", hbm_budget_bytes=16 << 30, pulses_per_epoch=pulses_per_epoch)
    subs = []
    for sc in scs:
        img = np.zeros((sc.grid["ny"], sc.grid["nx"]), dtype=np.complex64)
        subs.append((srv.submit(sc.profiles, sc.pos, sc.native_grid(), sc.native_params(interp="sinc8"), img),
                     img))
    infos = [srv.wait(j, timeout_s=300) for j, _ in subs]
    st = srv.stats()
    srv.close()
    assert st["failed"] == 0 and st["completed"] == len(scs)
    assert all(i["state"] == native.JOB_DONE for i in infos)
    return [img for _, img in subs], infos


def _check(sc, img, n_rows, n_chips):
    p = O.params(sc.c, sc.fc, sc.r0, sc.dr, "sinc8", sc.kaiser_beta)
    tp = sc.target_pixels
    rows = [int(tp[0][1]), sc.grid["ny"] // 2][:n_rows]
    chips = [tuple(int(v) for v in t) for t in tp[:n_chips]]
    r = spot.spot_check(sc.profiles, sc.pos, sc.grid, p, img, rows, chips)
    print(f"  {sc.name}: {sc.n_pulses}x{sc.n_samples} -> {sc.grid['nx']}^2 rows {r['rows']} "
          f"rel_l2 {['%.2e' % e for e in r['row_rel_l2']]}, chips {r['chips']} "
          f"rel_l2 {['%.2e' % e for e in r['chip_rel_l2']]}, combined {r['rel_l2']:.2e}", flush=True)
    assert r["rel_l2"] <= TOL
    assert max(r["row_rel_l2"] + r["chip_rel_l2"]) <= TOL
    assert all(r["chip_peak_match"])


def test_large_job_alone_and_in_mix(scenes):
    large, small = scenes
    print("[config4] large job alone", flush=True)
    (alone,), (ia,) = _serve([large])
    print(f"  latency {ia['done_s'] - ia['submit_s']:.2f}s, {ia['epochs']} epochs", flush=True)
    print("[config4] large job in a mix of 4", flush=True)
    mix, infos = _serve([small[0], large, small[1], small[2]])
    print(f"  latencies {[round(i['done_s'] - i['submit_s'], 2) for i in infos]}", flush=True)
    assert np.array_equal(alone.view(np.uint64), mix[1].view(np.uint64))
    assert infos[1]["epochs"] == 8  # 8192 pulses in 1024-pulse chunks
    # the small jobs shared epochs with the large one
    assert all(i["first_epoch_s"] < infos[1]["done_s"] for i in infos[2:])
    print("[config4] oracle spot checks", flush=True)
    _check(large, alone, n_rows=2, n_chips=2)
    for k, sc in enumerate(small):
        (a,), _ = _serve([sc])
        assert np.array_equal(a.view(np.uint64), mix[(0, 2, 3)[k]].view(np.uint64))
        _check(sc, a, n_rows=1, n_chips=1)
