"""The continuous-batching server (sarbp_server_*): every served image matches
the FP64 oracle (rel-L2 <= 1e-4) and is bitwise the same alone and in a mix;
requests arriving while others run join mid-flight; admission respects the
HBM budget; cancellation and the epoch watchdog (forced from the host with a
tiny limit -- the kernels themselves always finish) report through status."""
import time

import numpy as np
import pytest

from sar_cambrianaccelerationcard_amd import metrics as M
from sar_cambrianaccelerationcard_amd import native
from sar_cambrianaccelerationcard_amd import scene as S

import oracle as O

pytestmark = pytest.mark.gpu


def _scene(seed, n_pulses, nx, ny):
    cfg = S.scaled_config(S.load_config("config0"), n_pulses=n_pulses, nx=nx, ny=ny)
    cfg["seed"] = seed
    cfg["image"]["interp"] = "sinc8"
    return S.make_scene(cfg)


@pytest.fixture(scope="module")
def scenes():
    sizes = [(300, 160, 96), (64, 64, 64), (200, 96, 200), (128, 130, 70), (96, 64, 32)]
    return [_scene(2000 + i, *s) for i, s in enumerate(sizes)]


def _submit(srv, sc):
    img = np.zeros((sc.grid["ny"], sc.grid["nx"]), dtype=np.complex64)
    jid = srv.submit(sc.profiles, sc.pos, sc.native_grid(), sc.native_params(interp="sinc8"), img)
    return jid, img


def test_served_images_match_oracle_and_alone(scenes):
    print(f"[server] {len(scenes)} jobs, 100-pulse epochs", flush=True)
    srv = native.Server(interp="sinc8", hbm_budget_bytes=1 << 30, pulses_per_epoch=100)
    subs = [_submit(srv, sc) for sc in scenes]
    infos = [srv.wait(j, timeout_s=120) for j, _ in subs]
    st = srv.stats()
    srv.close()
    assert st["completed"] == len(scenes) and st["failed"] == 0
    for sc, (jid, img), info in zip(scenes, subs, infos):
        assert info["state"] == native.JOB_DONE
        assert info["epochs"] == max(1, -(-sc.n_pulses // 100))
        p = O.params(sc.c, sc.fc, sc.r0, sc.dr, "sinc8", sc.kaiser_beta)
        ref = O.backproject(sc.profiles, sc.pos, sc.grid, p)
        err = M.rel_l2(img, ref)
        print(f"  job {jid}: {sc.n_pulses}x{sc.grid['nx']}x{sc.grid['ny']} rel_l2={err:.2e} "
              f"latency={info['done_s'] - info['submit_s']:.3f}s", flush=True)
        assert err <= 1e-4
    # each job alone on a fresh server: bitwise identical
    for sc, (_, img) in zip(scenes[:3], subs[:3]):
        alone = native.Server(interp="sinc8", hbm_budget_bytes=1 << 30, pulses_per_epoch=100)
        jid, a = _submit(alone, sc)
        alone.wait(jid, timeout_s=120)
        alone.close()
        assert np.array_equal(a.view(np.uint64), img.view(np.uint64))


def test_arrivals_join_while_running(scenes):
    big = _scene(3001, 2048, 512, 512)
    srv = native.Server(interp="sinc8", hbm_budget_bytes=1 << 30, pulses_per_epoch=64)
    jb, _ = _submit(srv, big)
    while srv.poll(jb)["state"] != native.JOB_RUNNING:
        time.sleep(0.001)
    late = [_submit(srv, sc) for sc in scenes[1:3]]
    li = [srv.wait(j, timeout_s=120) for j, _ in late]
    bi = srv.wait(jb, timeout_s=120)
    srv.close()
    for i in li:
        assert i["admit_s"] < bi["done_s"]          # admitted while the big job ran
        assert i["first_epoch_s"] < bi["done_s"]    # and shared its epochs
        assert i["done_s"] < bi["done_s"]           # small jobs overtook it


def test_budget_admission_and_rejection(scenes):
    sc = scenes[0]
    need = sc.profiles.nbytes + sc.pos.nbytes + sc.grid["nx"] * sc.grid["ny"] * 8
    with pytest.raises(native.SarbpError) as ei:
        native.Server(interp="sinc8", hbm_budget_bytes=need // 2).submit(
            sc.profiles, sc.pos, sc.native_grid(), sc.native_params(interp="sinc8"),
            np.zeros((sc.grid["ny"], sc.grid["nx"]), np.complex64))
    assert ei.value.status == native.ERR_OVER_BUDGET
    # room for one job at a time: the second waits for the first to finish
    srv = native.Server(interp="sinc8", hbm_budget_bytes=int(need * 1.5) + 4096,
                        pulses_per_epoch=50)
    a, _ = _submit(srv, sc)
    b, _ = _submit(srv, sc)
    ia, ib = srv.wait(a, 120), srv.wait(b, 120)
    st = srv.stats()
    srv.close()
    assert ib["admit_s"] >= ia["done_s"] - 1e-3
    assert st["hbm_peak"] <= int(need * 1.5) + 4096


def test_cancel_and_release(scenes):
    srv = native.Server(interp="sinc8", hbm_budget_bytes=1 << 30, max_active_jobs=1,
                        pulses_per_epoch=10)
    a, _ = _submit(srv, scenes[0])
    b, _ = _submit(srv, scenes[1])   # waits for the single slot
    srv.cancel(b)
    ib = srv.wait(b, 60, raise_on_error=False)
    assert ib["state"] == native.JOB_CANCELLED and ib["wait_status"] == native.ERR_CANCELLED
    srv.wait(a, 120)
    srv.release(a)
    srv.release(b)
    with pytest.raises(native.SarbpError) as ei:
        srv.poll(a)
    assert ei.value.status == native.ERR_NOT_FOUND
    srv.close()


def test_watchdog_reports_timeout(scenes):
    srv = native.Server(interp="sinc8", hbm_budget_bytes=1 << 30, pulses_per_epoch=1000,
                        epoch_watchdog_s=1e-9)
    jid, _ = _submit(srv, scenes[0])
    info = srv.wait(jid, 60, raise_on_error=False)
    srv.drain(60)
    st = srv.stats()
    srv.close()
    assert info["state"] == native.JOB_FAILED and info["wait_status"] == native.ERR_TIMEOUT
    assert st["hbm_in_use"] == 0


def test_request_validation(scenes):
    sc = scenes[1]
    srv = native.Server(interp="sinc8", hbm_budget_bytes=1 << 28)
    img = np.zeros((sc.grid["ny"], sc.grid["nx"]), np.complex64)
    with pytest.raises(native.SarbpError):
        srv.submit(sc.profiles, sc.pos, sc.native_grid(), sc.native_params(interp="linear"), img)
    with pytest.raises(native.SarbpError):
        srv.submit(sc.profiles, sc.pos, sc.native_grid(),
                   sc.native_params(interp="sinc8", accumulate=True), img)
    srv.close()


def test_raw_echo_jobs_match_profile_jobs(ctx, scenes):
    """Jobs submitted as raw echoes (range-compressed in HBM at admission) in a
    mix with a profile job: each image holds the oracle bound against the BP of
    the FP64-compressed profiles and is bitwise the image of the same job
    submitted as its device-compressed profiles on a fresh server."""
    from sar_cambrianaccelerationcard_amd import device as D
    raws = []
    for i, size in enumerate([(300, 160, 96), (128, 130, 70)]):
        cfg = S.scaled_config(S.load_config("config0"), n_pulses=size[0], nx=size[1], ny=size[2])
        cfg["seed"] = 3000 + i
        cfg["image"]["interp"] = "sinc8"
        raws.append(S.make_scene(cfg, keep_f64=True, keep_raw=True))
    print(f"[server raw] {len(raws)} raw-echo jobs + 1 profile job", flush=True)
    srv = native.Server(interp="sinc8", hbm_budget_bytes=1 << 30, pulses_per_epoch=100)
    subs = []
    for sc in raws:
        img = np.zeros((sc.grid["ny"], sc.grid["nx"]), dtype=np.complex64)
        raw = np.ascontiguousarray(sc.raw, dtype=np.complex64)
        jid = srv.submit_raw(raw, D.range_params(sc.radar), sc.pos, sc.native_grid(),
                             sc.native_params(interp="sinc8"), img)
        subs.append((jid, img, raw))
    pj, pimg = _submit(srv, scenes[1])
    for jid, _, _ in subs:
        assert srv.wait(jid, timeout_s=120)["state"] == native.JOB_DONE
    assert srv.wait(pj, timeout_s=120)["state"] == native.JOB_DONE
    srv.close()
    for sc, (jid, img, raw) in zip(raws, subs):
        p = O.params(sc.c, sc.fc, sc.r0, sc.dr, "sinc8", sc.kaiser_beta)
        ref = O.backproject(sc.profiles64, sc.pos, sc.grid, p)
        err = M.rel_l2(img, ref)
        print(f"  raw job {jid}: rel_l2={err:.2e}", flush=True)
        assert err <= 1e-4
        prof = D.range_compress(ctx, raw, sc.radar).cpu().numpy()
        alone = native.Server(interp="sinc8", hbm_budget_bytes=1 << 30, pulses_per_epoch=100)
        a = np.zeros_like(img)
        j2 = alone.submit(prof, sc.pos, sc.native_grid(), sc.native_params(interp="sinc8"), a)
        alone.wait(j2, timeout_s=120)
        alone.close()
        assert np.array_equal(a, img)


def test_raw_request_validation():
    srv = native.Server(interp="sinc8", hbm_budget_bytes=1 << 28, pulses_per_epoch=100)
    sc = _scene(4000, 16, 32, 32)
    rp = native.make_range_params(c=299792458.0, bandwidth=150e6, pulse_width=2e-6, fs=200e6,
                                  t0=3.2e-5, n_raw=624, upsample=0, crop_offset=0, n_out=2048)
    raw = np.zeros((16, 624), dtype=np.complex64)
    img = np.zeros((32, 32), dtype=np.complex64)
    with pytest.raises(native.SarbpError) as e:
        srv.submit_raw(raw, rp, sc.pos, sc.native_grid(), sc.native_params(interp="sinc8"), img)
    assert e.value.status == native.ERR_INVALID_ARGUMENT
    rp.upsample = 4
    rp.n_raw = 9000
    with pytest.raises(native.SarbpError) as e:
        srv.submit_raw(np.zeros((16, 9000), dtype=np.complex64), rp, sc.pos, sc.native_grid(),
                       sc.native_params(interp="sinc8"), img)
    assert e.value.status == native.ERR_UNSUPPORTED
    srv.close()


def test_server_takes_class_tuning(ctx, scenes):
    """sarbp_server_set_class_tuning (here through tune_like, from a context's
    per-class tuning): a job served in one epoch is bitwise the planned plain
    launch with the same configuration."""
    import torch
    from sar_cambrianaccelerationcard_amd import device as D
    sc = scenes[2]
    key = (native.INTERP_SINC8, native.PREC_REF, native.DTYPE_C64)
    cls = native.size_class(sc.native_grid(), sc.native_params(interp="sinc8"), sc.pos)
    try:
        ctx.set_class_tuning(*key, cls, 2, 2, 5)
        srv = native.Server(interp="sinc8", hbm_budget_bytes=1 << 30, pulses_per_epoch=0)
        srv.tune_like(ctx)
        jid, img = _submit(srv, sc)
        srv.wait(jid, timeout_s=120)
        srv.close()
        plain = D.backproject(ctx, D.DeviceScene(sc), interp="sinc8", plan=True).cpu().numpy()
        torch.cuda.synchronize()
        assert np.array_equal(img.view(np.uint64), plain.view(np.uint64))
    finally:
        ctx.set_tuning(*key, 0, 0, 0)


def _coarse(sc, factor):
    """The scene's grid with its pixel spacing scaled (beyond REF's tile-radius
    limit at factor 3.2: 1.6 m pixels)."""
    g = dict(sc.grid)
    g["ux"] = [v * factor for v in sc.grid["ux"]]
    g["uy"] = [v * factor for v in sc.grid["uy"]]
    return g, native.make_grid(g["origin"], g["ux"], g["uy"], g["nx"], g["ny"])


def test_coarse_grid_served_as_double_float(scenes):
    """A job whose grid no REF tile fits runs the DF arithmetic in the server
    as in the plain call: served alone and in a mix with fine-grid (REF)
    jobs, its image is bitwise the same both ways and matches the oracle."""
    sc = scenes[0]
    g, ng = _coarse(sc, 3.2)

    def run(mix):
        srv = native.Server(interp="sinc8", hbm_budget_bytes=1 << 30, pulses_per_epoch=100)
        others = [_submit(srv, s) for s in mix]
        img = np.zeros((g["ny"], g["nx"]), dtype=np.complex64)
        jid = srv.submit(sc.profiles, sc.pos, ng, sc.native_params(interp="sinc8"), img)
        info = srv.wait(jid, timeout_s=120)
        for j, _ in others:
            srv.wait(j, timeout_s=120)
        st = srv.stats()
        srv.close()
        assert info["state"] == native.JOB_DONE and st["failed"] == 0
        return img

    alone = run([])
    mixed = run(scenes[1:4])
    assert np.array_equal(alone.view(np.uint64), mixed.view(np.uint64))
    p = O.params(sc.c, sc.fc, sc.r0, sc.dr, "sinc8", sc.kaiser_beta)
    err = M.rel_l2(alone, O.backproject(sc.profiles, sc.pos, g, p))
    print(f"  1.6 m grid served (DF): rel_l2={err:.2e}", flush=True)
    assert err <= 1e-4
