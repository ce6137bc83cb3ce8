"""Edge cases of the plain entry point through the C ABI, each against the
FP64 oracle where there is an image to compare: zero pulses, empty images,
ragged sizes (not multiples of the tile), accumulation and pulse chunking,
bf16 storage, tilted and non-orthogonal 3-D grids, every tile shape, samples
off both ends of the profile, explicit streams, and argument validation."""
import numpy as np
import pytest

from sar_cambrianaccelerationcard_amd import device as D
from sar_cambrianaccelerationcard_amd import metrics as M
from sar_cambrianaccelerationcard_amd import native
from sar_cambrianaccelerationcard_amd import scene as S

import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def base():
    cfg = S.scaled_config(S.load_config("config0"), n_pulses=96, nx=100, ny=70)
    sc = S.make_scene(cfg)
    return sc, D.DeviceScene(sc, bf16=True)


def _oracle(sc, interp="sinc8", grid=None, prof=None, pos=None):
    p = O.params(sc.c, sc.fc, sc.r0, sc.dr, interp, sc.kaiser_beta)
    return O.backproject(sc.profiles if prof is None else prof, sc.pos if pos is None else pos,
                         grid or sc.grid, p)


def _run(ctx, ds, grid, params, img=None, stream=0):
    import torch
    if img is None:
        img = torch.zeros((grid.ny, grid.nx), dtype=torch.complex64, device="cuda")
    ctx.backproject(ds.prof.data_ptr(), ds.pos.data_ptr(), grid, params, img.data_ptr(), stream, sync=True)
    return img


def test_zero_pulses_writes_zero_and_accumulate_keeps(ctx, base):
    import torch
    sc, ds = base
    params = sc.native_params(interp="sinc8")
    params.n_pulses = 0
    img = torch.full((sc.grid["ny"], sc.grid["nx"]), 3 + 1j, dtype=torch.complex64, device="cuda")
    _run(ctx, ds, sc.native_grid(), params, img)
    assert torch.count_nonzero(img).item() == 0
    img.fill_(2 - 1j)
    params.accumulate = 1
    _run(ctx, ds, sc.native_grid(), params, img)
    assert torch.all(img == (2 - 1j)).item()


def test_empty_image_is_a_no_op(ctx, base):
    sc, ds = base
    g = sc.native_grid()
    g.nx = 0
    ctx.backproject(ds.prof.data_ptr(), ds.pos.data_ptr(), g, sc.native_params(), 0, sync=True)


@pytest.mark.parametrize("nx,ny", [(1, 1), (63, 5), (65, 33), (129, 3)])
def test_ragged_sizes(ctx, base, nx, ny):
    sc, ds = base
    g = dict(sc.grid, nx=nx, ny=ny)
    img = _run(ctx, ds, native.make_grid(g["origin"], g["ux"], g["uy"], nx, ny),
               sc.native_params(interp="linear")).cpu().numpy()
    assert M.rel_l2(img, _oracle(sc, "linear", g)) <= 1e-4


def test_pulse_chunks_accumulate_to_the_whole(ctx, base):
    """Pulse additivity on the device: three accumulated sub-apertures equal
    the full aperture to FP32 rounding and the oracle to 1e-4."""
    import torch
    sc, ds = base
    full = _run(ctx, ds, sc.native_grid(), sc.native_params(interp="sinc8")).cpu().numpy()
    img = torch.zeros((sc.grid["ny"], sc.grid["nx"]), dtype=torch.complex64, device="cuda")
    cuts = [0, 30, 61, sc.n_pulses]
    for k, (a, b) in enumerate(zip(cuts[:-1], cuts[1:])):
        p = sc.native_params(interp="sinc8", accumulate=k > 0)
        p.n_pulses = b - a
        ctx.backproject(ds.prof[a:].data_ptr(), ds.pos[a:].data_ptr(), sc.native_grid(), p, img.data_ptr(),
                        sync=True)
    got = img.cpu().numpy()
    assert M.rel_l2(got, full) <= 1e-6
    assert M.rel_l2(got, _oracle(sc)) <= 1e-4


def test_bf16_storage_matches_oracle_on_bf16_values(ctx, base):
    """bf16x2 storage is exact widening: against the oracle fed the same
    bf16-rounded samples it holds 1e-4."""
    sc, ds = base
    img = D.backproject(ctx, ds, interp="sinc8", dtype="bf16x2").cpu().numpy()
    rounded = D.from_bf16x2(D.to_bf16x2(sc.profiles))
    assert M.rel_l2(img, _oracle(sc, prof=rounded)) <= 1e-4


def test_tilted_non_orthogonal_grid(ctx, base):
    """A slant-plane grid with a vertical component and non-orthogonal steps."""
    sc, ds = base
    o = sc.grid["origin"]
    g = {"origin": [o[0], o[1], 5.0], "ux": [0.45, 0.05, 0.02], "uy": [0.08, 0.4, -0.15], "nx": 70, "ny": 50}
    img = _run(ctx, ds, native.make_grid(g["origin"], g["ux"], g["uy"], g["nx"], g["ny"]),
               sc.native_params(interp="sinc8")).cpu().numpy()
    assert M.rel_l2(img, _oracle(sc, grid=g)) <= 1e-4


@pytest.mark.parametrize("cols,rows", [(1, 2), (1, 4), (1, 8), (2, 2), (2, 4), (2, 8), (4, 2), (4, 4)])
def test_every_tile_shape(base, cols, rows):
    """Every compiled tile shape on a 0.25 m grid (the widest tile, 256 px,
    spans 32 m from its centre)."""
    sc, ds = base
    o = sc.grid["origin"]
    g = {"origin": o, "ux": [0.25, 0, 0], "uy": [0, 0.25, 0], "nx": 300, "ny": 90}
    c = native.Context(0)
    c.set_tuning(native.INTERP_SINC8, native.PREC_REF, native.DTYPE_C64, rows, cols, 5)
    img = _run(c, ds, native.make_grid(g["origin"], g["ux"], g["uy"], g["nx"], g["ny"]),
               sc.native_params(interp="sinc8")).cpu().numpy()
    c.close()
    assert M.rel_l2(img, _oracle(sc, grid=g)) <= 1e-4


def test_tile_shape_falls_back_within_radius_limit(base):
    """A tuned tile shape whose tiles would reach more than 48 m from their
    centre (FP32 pp) runs with the largest compiled shape within it that fits."""
    sc, ds = base
    c = native.Context(0)
    c.set_tuning(native.INTERP_SINC8, native.PREC_REF, native.DTYPE_C64, 2, 4, 5)  # 256 px at 0.5 m
    info = c.plan(sc.pos, sc.native_grid(), sc.native_params(interp="sinc8")).info()
    assert (info["cols_per_lane"], info["rows_per_lane"]) == (2, 2)
    img = _run(c, ds, sc.native_grid(), sc.native_params(interp="sinc8")).cpu().numpy()
    c.close()
    o = O.params(sc.c, sc.fc, sc.r0, sc.dr, "sinc8", sc.kaiser_beta)
    assert M.rel_l2(img, O.backproject(sc.profiles, sc.pos, sc.grid, o)) <= 1e-4


def _scaled_grid(sc, f):
    g = dict(sc.grid)
    g["ux"] = [v * f for v in sc.grid["ux"]]
    g["uy"] = [v * f for v in sc.grid["uy"]]
    return g, native.make_grid(g["origin"], g["ux"], g["uy"], g["nx"], g["ny"])


@pytest.mark.parametrize("interp,factor", [("sinc8", 3.2), ("linear", 8.0)])
def test_grid_too_coarse_for_ref_runs_double_float(ctx, base, interp, factor):
    """At 1.6 m pixels even the smallest tile (64 x 8) spans 51 m, beyond the
    REF differential's radius limit; at 4 m (linear) 129 m.  The job runs the
    double-float differential instead -- planned, unplanned and batched --
    and matches the FP64 oracle."""
    import torch
    sc, ds = base
    g, ng = _scaled_grid(sc, factor)
    params = sc.native_params(interp=interp)
    info = ctx.plan(sc.pos, ng, params).info()
    assert info["precision"] == native.PREC_DF and info["cols_per_lane"] == 1, info
    ref = _oracle(sc, interp, g)
    img = _run(ctx, ds, ng, params).cpu().numpy()  # unplanned: no host positions
    err = M.rel_l2(img, ref)
    print(f"  {interp} at {0.5 * factor} m: DF fallback rel_l2={err:.2e} ({info['tile_nx']}x{info['tile_ny']} tiles)",
          flush=True)
    assert err <= 1e-4
    planned = D.backproject(ctx, ds, interp=interp, plan=True, grid=g).cpu().numpy()
    assert M.rel_l2(planned, ref) <= 1e-4
    # a batch of a fine (REF) job and the coarse (DF) job: each equals its plain planned launch
    jobs = []
    for gg in (sc.grid, g):
        out = torch.zeros((gg["ny"], gg["nx"]), dtype=torch.complex64, device="cuda")
        jb = native.Job()
        jb.d_profiles, jb.d_pos, jb.h_pos = ds.prof.data_ptr(), ds.pos.data_ptr(), sc.pos.ctypes.data
        jb.grid = native.make_grid(gg["origin"], gg["ux"], gg["uy"], gg["nx"], gg["ny"])
        jb.params, jb.d_image = sc.native_params(interp=interp), out.data_ptr()
        jobs.append((jb, out, gg))
    ctx.backproject_batch([j for j, _, _ in jobs], torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    for jb, out, gg in jobs:
        plain = D.backproject(ctx, ds, interp=interp, plan=True, grid=gg).cpu().numpy()
        assert np.array_equal(out.cpu().numpy().view(np.uint64), plain.view(np.uint64))


def test_allocation_free_batch_with_a_double_float_job(ctx, base):
    """The allocation-free batch groups a DF-fallback job apart from REF jobs
    and reproduces the allocating batch bit for bit."""
    import torch
    sc, ds = base
    g, _ = _scaled_grid(sc, 3.2)
    out = {}
    for name in ("batch", "noalloc"):
        jobs = []
        for gg in (sc.grid, g):
            img = torch.zeros((gg["ny"], gg["nx"]), dtype=torch.complex64, device="cuda")
            jb = native.Job()
            jb.d_profiles, jb.d_pos, jb.h_pos = ds.prof.data_ptr(), ds.pos.data_ptr(), sc.pos.ctypes.data
            jb.grid = native.make_grid(gg["origin"], gg["ux"], gg["uy"], gg["nx"], gg["ny"])
            jb.params, jb.d_image = sc.native_params(interp="sinc8"), img.data_ptr()
            jobs.append((jb, img))
        call = ctx.backproject_batch if name == "batch" else ctx.backproject_batch_noalloc
        call([j for j, _ in jobs], torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        out[name] = [im.cpu().numpy() for _, im in jobs]
    for a, b in zip(out["batch"], out["noalloc"]):
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64))


@pytest.mark.parametrize("rows", [2, 4, 8])
def test_double_float_mode_every_shape(ctx, base, rows):
    """SARBP_PREC_DF requested directly, every compiled tile shape, on the
    0.5 m grid and at 2 m: the oracle bound holds."""
    sc, ds = base
    c = native.Context(0)
    try:
        c.set_tuning(native.INTERP_SINC8, native.PREC_DF, native.DTYPE_C64, rows, 1, 8)
        for factor in (1.0, 4.0):
            g, ng = _scaled_grid(sc, factor)
            interp = "sinc8" if factor == 1.0 else "linear"
            img = _run(c, ds, ng, sc.native_params(interp=interp, precision="df")).cpu().numpy()
            err = M.rel_l2(img, _oracle(sc, interp, g))
            print(f"  DF 1x{rows} at {0.5 * factor} m ({interp}): rel_l2={err:.2e}", flush=True)
            assert err <= 1e-4
    finally:
        c.close()


def test_sinc8_kernel_variants(ctx, base):
    """The planner runs the half-bin kernel on a fine grid and the whole-bin
    kernel where tiles are coarse against the range bins; both hold the
    oracle bound, and one batch of both jobs reproduces their planned (host
    positions known, like the batch) launches bit for bit."""
    import torch
    sc, ds = base
    o = sc.grid["origin"]
    grids = [{"origin": o, "ux": [0.5, 0, 0], "uy": [0, 0.5, 0], "nx": 90, "ny": 60},
             {"origin": o, "ux": [1.25, 0, 0], "uy": [0, 1.25, 0], "nx": 70, "ny": 50}]
    halves = []
    imgs = []
    jobs = []
    for g in grids:
        ng = native.make_grid(g["origin"], g["ux"], g["uy"], g["nx"], g["ny"])
        plan = ctx.plan(sc.pos, ng, sc.native_params(interp="sinc8"))
        halves.append(plan.info()["sinc8_halves"])
        out0 = torch.zeros((g["ny"], g["nx"]), dtype=torch.complex64, device="cuda")
        plan.execute(ds.prof.data_ptr(), ds.pos.data_ptr(), out0.data_ptr(), torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        img = out0.cpu().numpy()
        imgs.append(img)
        assert M.rel_l2(img, _oracle(sc, grid=g)) <= 1e-4
        jb = native.Job()
        out = torch.zeros((g["ny"], g["nx"]), dtype=torch.complex64, device="cuda")
        jb.d_profiles, jb.d_pos, jb.h_pos = ds.prof.data_ptr(), ds.pos.data_ptr(), sc.pos.ctypes.data
        jb.grid, jb.params, jb.d_image = ng, sc.native_params(interp="sinc8"), out.data_ptr()
        jobs.append((jb, out))
    assert sorted(halves) == [1, 2], halves
    ctx.backproject_batch([j for j, _ in jobs], torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    for (_, out), img in zip(jobs, imgs):
        assert np.array_equal(out.cpu().numpy(), img)


def test_samples_beyond_the_profile_read_zero(ctx, base):
    """Shift the range axis so pixels fall off both ends of the profile."""
    sc, ds = base
    for shift in (-sc.n_samples * sc.dr * 0.6, sc.n_samples * sc.dr * 0.6):
        p = sc.native_params(interp="sinc8")
        p.r0 = sc.r0 + shift
        img = _run(ctx, ds, sc.native_grid(), p).cpu().numpy()
        o = O.params(sc.c, sc.fc, sc.r0 + shift, sc.dr, "sinc8", sc.kaiser_beta)
        ref = O.backproject(sc.profiles, sc.pos, sc.grid, o)
        assert M.rel_l2(img, ref) <= 1e-4


def test_explicit_stream(ctx, base):
    import torch
    sc, ds = base
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        img = torch.zeros((sc.grid["ny"], sc.grid["nx"]), dtype=torch.complex64, device="cuda")
        ctx.backproject(ds.prof.data_ptr(), ds.pos.data_ptr(), sc.native_grid(),
                        sc.native_params(interp="linear"), img.data_ptr(), s.cuda_stream)
    s.synchronize()
    assert M.rel_l2(img.cpu().numpy(), _oracle(sc, "linear")) <= 1e-4


def test_invalid_arguments(ctx, base):
    sc, ds = base
    bad = [("n_samples", 0), ("interp", 7), ("precision", -1), ("dtype", 3), ("dr", 0.0),
           ("c", -1.0), ("accumulate", 2)]
    for field, value in bad:
        p = sc.native_params()
        setattr(p, field, value)
        with pytest.raises(native.SarbpError) as ei:
            ctx.backproject(ds.prof.data_ptr(), ds.pos.data_ptr(), sc.native_grid(), p, 0)
        assert ei.value.status == native.ERR_INVALID_ARGUMENT, field
    g = sc.native_grid()
    g.ux[0] = g.ux[1] = g.ux[2] = 0.0
    with pytest.raises(native.SarbpError):
        ctx.backproject(ds.prof.data_ptr(), ds.pos.data_ptr(), g, sc.native_params(), 0)
    with pytest.raises(native.SarbpError):
        ctx.set_tuning(native.INTERP_SINC8, native.PREC_F64, native.DTYPE_C64, 4, 2, 8)  # F64 has cols 1 only
