"""GPU parity of TV-ADMM (src/inversion/tv.rs:63-97) against the f64 oracle.

FP32 FFTs/state with FP64 operator construction and norms; the bar is
NRMSE <= 1e-3 vs f64 (stated per test, tighter where the problem allows)."""
import numpy as np
import pytest

from conftest import nrmse_dc

pytestmark = pytest.mark.gpu


def _case(rng, dims, frac=0.8, scale=0.01):
    n = int(np.prod(dims))
    f = rng.standard_normal(n) * scale
    m = (rng.random(n) < frac).astype(np.uint8)
    return f, m


def test_backend_info(qsm):
    info = qsm.backend_info()
    print(info)
    assert "gfx950" in info


@pytest.mark.parametrize("dims,vs,bdir", [
    ((16, 16, 16), (1.0, 1.0, 1.0), (0.0, 0.0, 1.0)),
    ((16, 14, 12), (0.9, 1.1, 1.3), (0.3, -0.2, 1.0)),
    ((20, 18, 22), (1.0, 1.0, 1.0), (0.2, 0.1, 0.97)),   # radices 5, 3, 11
    ((26, 14, 10), (0.8, 0.8, 1.2), (0.0, 0.0, 1.0)),    # radices 13, 7, 5
    ((8, 8, 8), (1.0, 1.0, 1.0), (0.0, 0.0, 1.0)),
    ((7, 9, 5), (1.0, 1.0, 1.0), (0.1, 0.0, 1.0)),       # odd nx (no Nyquist bin)
    ((64, 48, 40), (0.9, 1.0, 1.1), (0.1, 0.2, 0.97)),   # compiled nx = 64, quarter-row x passes
    ((200, 12, 10), (1.0, 1.0, 1.0), (0.0, 0.1, 1.0)),   # compiled nx = 200 (radices 8, 5, 5)
    ((12, 180, 220), (1.0, 0.9, 1.1), (0.1, 0.0, 1.0)),  # compiled y = 180, z = 220 (staged)
    ((512, 10, 8), (0.5, 1.0, 1.0), (0.0, 0.1, 1.0)),    # compiled nx = 512
    ((8, 512, 6), (1.0, 0.5, 1.0), (0.1, 0.0, 1.0)),     # compiled y = 512 (staged)
    ((6, 10, 512), (1.0, 1.0, 0.5), (0.0, 0.0, 1.0)),    # compiled z = 512 (staged)
    ((288, 10, 8), (1.0, 1.0, 1.0), (0.1, 0.0, 1.0)),    # compiled nx = 288
    ((480, 8, 6), (0.5, 1.0, 1.0), (0.0, 0.1, 1.0)),     # compiled nx = 480
    ((8, 416, 6), (1.0, 0.6, 1.0), (0.0, 0.0, 1.0)),     # compiled y = 416 (staged)
    ((6, 10, 448), (1.0, 1.0, 0.6), (0.1, 0.1, 1.0)),    # compiled z = 448 (staged)
])
@pytest.mark.parametrize("iters", [1, 2, 20])
def test_tv_matches_oracle(qsm, oracle, dims, vs, bdir, iters):
    rng = np.random.default_rng(hash((dims, iters)) % 2**32)
    f, m = _case(rng, dims)
    grid = qsm.Grid(dims, vs)
    p = qsm.TvParams(lambda_=1e-3, rho=0.1, tol=0.0, max_iter=iters)
    x, it, _ = qsm.tv_admm(f, m, grid, bdir, p, return_info=True)
    xo = oracle.tv_admm(f, m, dims, vs, bdir, 1e-3, 0.1, 0.0, iters)
    assert it == iters
    e = nrmse_dc(x, xo, m)
    print(dims, iters, "nrmse(dc removed)", e, "raw", oracle.nrmse(x, xo, m))
    assert e < 2e-5, e          # FP32 vs f64 on small volumes
    assert np.all(x[m == 0] == 0.0)


def _compiled_lengths():
    import importlib.util
    import os
    from conftest import PKG_DIR
    spec = importlib.util.spec_from_file_location("qsm_build", os.path.join(PKG_DIR, "build.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod.FIXED_LENGTHS


@pytest.mark.parametrize("sched", [1, 2], ids=["two-stage", "staged"])
@pytest.mark.parametrize("n", _compiled_lengths())
def test_tv_every_compiled_length(qsm, oracle, n, sched):
    # the compiled-length kernels of n on each axis in turn (x: row passes;
    # y, z: column passes under the forced schedule; lengths without the
    # two-stage form run the staged one either way).  Anisotropic voxels:
    # NRMSE with the reference's DC term removed (raw printed); isotropic
    # voxels (no DC term): raw NRMSE.
    try:
        qsm.set_fft_schedule(sched)
        for k, dims in enumerate([(n, 6, 4), (4, n, 6), (6, 4, n)]):
            for vs in ((1.0, 0.9, 1.1), (1.0, 1.0, 1.0)):
                rng = np.random.default_rng(n * 3 + k)
                f, m = _case(rng, dims)
                bdir = (0.1, 0.0, 1.0)
                p = qsm.TvParams(lambda_=1e-3, rho=0.1, tol=0.0, max_iter=3)
                x = qsm.tv_admm(f, m, qsm.Grid(dims, vs), bdir, p)
                xo = oracle.tv_admm(f, m, dims, vs, bdir, 1e-3, 0.1, 0.0, 3)
                e, raw = nrmse_dc(x, xo, m), oracle.nrmse(x, xo, m)
                print(dims, vs, "nrmse(dc removed)", e, "raw", raw)
                assert e < 2e-5, (dims, vs, e)
                if vs == (1.0, 1.0, 1.0):
                    assert raw < 2e-5, (dims, raw)
                assert np.all(x[m == 0] == 0.0)
    finally:
        qsm.set_fft_schedule(0)


def test_tv_zero_field(qsm):
    # src/inversion/tv.rs:113-126
    n = 8
    grid = qsm.Grid((n, n, n), (1.0, 1.0, 1.0))
    p = qsm.TvParams(lambda_=1e-3, rho=0.1, tol=1e-2, max_iter=10)
    x = qsm.tv_admm(np.zeros(n ** 3), np.ones(n ** 3, np.uint8), grid, (0, 0, 1), p)
    assert np.max(np.abs(x)) < 1e-8


def test_tv_finite_ramp(qsm):
    # src/inversion/tv.rs:129-142
    n = 8
    grid = qsm.Grid((n, n, n), (1.0, 1.0, 1.0))
    p = qsm.TvParams(lambda_=1e-3, rho=0.1, tol=1e-2, max_iter=10)
    f = np.arange(n ** 3) * 0.001
    x = qsm.tv_admm(f, np.ones(n ** 3, np.uint8), grid, (0, 0, 1), p)
    assert np.all(np.isfinite(x))


def test_tv_progress_and_early_exit(qsm, oracle):
    # progress(iter+1, max) per iteration, (k, k) on convergence (tv.rs:81, 89)
    rng = np.random.default_rng(5)
    dims = (16, 16, 16)
    f = np.sin(np.arange(16 ** 3) * 0.7) * 0.01
    m = np.ones(16 ** 3, np.uint8)
    grid = qsm.Grid(dims, (1.0, 1.0, 1.0))
    p = qsm.TvParams(lambda_=1e-3, rho=0.1, tol=1e-3, max_iter=50)
    calls, ocalls = [], []
    x, it, rel = qsm.tv_admm(f, m, grid, (0, 0, 1), p, progress=lambda a, b: calls.append((a, b)),
                             return_info=True)
    xo, oit, orel = oracle.tv_admm(f, m, dims, (1, 1, 1), (0, 0, 1), 1e-3, 0.1, 1e-3, 50,
                                   progress=lambda a, b: ocalls.append((a, b)), return_info=True)
    print("gpu iters", it, rel, "oracle iters", oit, orel)
    assert it == oit
    assert calls == ocalls
    assert nrmse_dc(x, xo, m) < 2e-5


def test_tv_phantom_cfg0_default_params(qsm, oracle, phantom):
    # 128x128x64 @ 1 mm, TvParams::default() (early exit at tol 1e-3)
    ph = phantom.make_phantom((128, 128, 64), n_ellipsoids=12, seed=1000)
    grid = qsm.Grid(ph.dims, ph.voxel_size)
    p = qsm.TvParams.default()
    x, it, rel = qsm.tv_admm(ph.local_field, ph.mask, grid, ph.bdir, p, return_info=True)
    xo, oit, orel = oracle.tv_admm(ph.local_field, ph.mask, ph.dims, ph.voxel_size, ph.bdir,
                                   p.lambda_, p.rho, p.tol, p.max_iter, return_info=True)
    e = oracle.nrmse(x, xo, ph.mask)
    print("cfg0: gpu", it, rel, "oracle", oit, orel, "nrmse", e,
          "maxabs", np.max(np.abs(x - xo)))
    assert abs(it - oit) <= 1
    assert e <= 1e-3


def test_tv_batch_equals_single(qsm):
    rng = np.random.default_rng(9)
    shapes = [((16, 16, 16), (1.0, 1.0, 1.0), (0, 0, 1)),
              ((20, 18, 22), (0.9, 1.0, 1.1), (0.1, 0.2, 0.95)),
              ((12, 10, 8), (1.2, 0.7, 1.0), (0.0, 0.3, 1.0))]
    p = qsm.TvParams(lambda_=1e-3, rho=0.1, tol=0.0, max_iter=15)
    ins, outs, vols, singles = [], [], [], []
    for dims, vs, b in shapes:
        f, m = _case(rng, dims)
        singles.append(qsm.tv_admm(f, m, qsm.Grid(dims, vs), b, p))
        tf = qsm.DeviceArray.from_numpy(f)
        tm = qsm.DeviceArray.from_numpy(m)
        tc = qsm.DeviceArray(f.size, np.float64)
        ins.append((tf, tm))
        outs.append(tc)
        vols.append((qsm.Grid(dims, vs), b, tf.ptr, tm.ptr, tc.ptr))
    batch = qsm.TvBatch(vols, p)
    batch.run()
    for v in range(len(shapes)):
        got = outs[v].numpy()
        assert np.array_equal(got, singles[v]), v
        assert batch.status(v)[0] == 15
    batch.close()


def test_tv_batch_graph_replays_match_eager(qsm):
    # Iterations 2.. replay HIP graphs captured at the first run; an eager
    # batch over the same inputs must give bit-identical results on every
    # replay, with other eager GPU work and synchronisations in between, on
    # the no-callback path (tol 0) and on the progress/early-exit path.
    rng = np.random.default_rng(21)
    shapes = [((64, 64, 32), (1.0, 1.0, 1.0), (0, 0, 1)),
              ((20, 18, 22), (0.9, 1.0, 1.1), (0.1, 0.2, 0.95))]
    inputs, outs, vg, ve = [], [], [], []
    for dims, vs, b in shapes:
        f, m = _case(rng, dims)
        tf, tm = qsm.DeviceArray.from_numpy(f), qsm.DeviceArray.from_numpy(m)
        cg, ce = qsm.DeviceArray(f.size, np.float64), qsm.DeviceArray(f.size, np.float64)
        inputs.append((tf, tm))
        outs.append((cg, ce))
        vg.append((qsm.Grid(dims, vs), b, tf.ptr, tm.ptr, cg.ptr))
        ve.append((qsm.Grid(dims, vs), b, tf.ptr, tm.ptr, ce.ptr))
    other = rng.standard_normal(16 * 16 * 16)
    for p in (qsm.TvParams(lambda_=1e-3, rho=0.1, tol=0.0, max_iter=11),
              qsm.TvParams(lambda_=2e-4, rho=2e-2, tol=5e-3, max_iter=40)):
        bg, be = qsm.TvBatch(vg, p), qsm.TvBatch(ve, p)
        be.set_graphs(False)
        for rep in range(3):
            pe, pg = [], []
            be.run((lambda c, t: pe.append((c, t))) if p.tol > 0 else None)
            qsm.dipole_forward(other, qsm.Grid((16, 16, 16)), (0, 0, 1))  # eager work in between
            qsm.synchronize()
            bg.run((lambda c, t: pg.append((c, t))) if p.tol > 0 else None)
            assert pe == pg, (rep, pe[-3:], pg[-3:])
            for v, (cg, ce) in enumerate(outs):
                assert np.array_equal(cg.numpy(), ce.numpy()), (p.tol, rep, v)
                assert bg.status(v) == be.status(v)
        bg.close()
        be.close()


def test_fft_schedules_agree_with_oracle(qsm, oracle):
    # The y and z lengths 144 and 160 have both column-FFT schedules (the
    # two-stage transform, radix 9 / 10 then 16, and the LDS-staged one):
    # forced either way, TV-ADMM matches the f64 oracle; the default mode is
    # the fixed table, and the tuned mode (3) records a choice for both
    # lengths once a batch used them.
    rng = np.random.default_rng(33)
    dims, vs, b = (24, 144, 160), (1.0, 0.9, 1.1), (0.1, 0.1, 1.0)
    f, m = _case(rng, dims)
    grid = qsm.Grid(dims, vs)
    p = qsm.TvParams(lambda_=1e-3, rho=0.1, tol=0.0, max_iter=8)
    xo = oracle.tv_admm(f, m, dims, vs, b, 1e-3, 0.1, 0.0, 8)
    got = {}
    try:
        for mode in (1, 2):
            qsm.set_fft_schedule(mode)
            assert qsm.fft_schedule("y", 144) == ("two-stage" if mode == 1 else "staged")
            got[mode] = qsm.tv_admm(f, m, grid, b, p)
            e = nrmse_dc(got[mode], xo, m)
            print("schedule", mode, "nrmse(dc removed)", e)
            assert e < 2e-5, (mode, e)
        assert nrmse_dc(got[1], got[2], m) < 2e-6
        qsm.set_fft_schedule(0)
        x0 = qsm.tv_admm(f, m, grid, b, p)
        # fixed table: staged y = 144, two-stage z = 160
        assert qsm.fft_schedule("y", 144) == "staged" and qsm.fft_schedule("z", 160) == "two-stage"
        got[0] = x0
        assert np.array_equal(x0, qsm.tv_admm(f, m, grid, b, p))
        qsm.set_fft_schedule(3)
        x3 = qsm.tv_admm(f, m, grid, b, p)
        assert qsm.fft_schedule("y", 144) in ("two-stage", "staged")
        assert qsm.fft_schedule("z", 160) in ("two-stage", "staged")
        assert qsm.fft_schedule("y", 384) == "staged"  # no two-stage form
        for x in (x0, x3):
            assert nrmse_dc(x, xo, m) < 2e-5
    finally:
        qsm.set_fft_schedule(0)
    with pytest.raises(qsm.QsmError):
        qsm.set_fft_schedule(4)


def test_tv_rejects_bad_input(qsm):
    # 521 is a prime above the chirp-z limit (512): unsupported
    grid = qsm.Grid((521, 4, 4), (1.0, 1.0, 1.0))
    p = qsm.TvParams()
    with pytest.raises(qsm.QsmError) as ei:
        qsm.tv_admm(np.zeros(521 * 16), np.ones(521 * 16, np.uint8), grid, (0, 0, 1), p)
    assert ei.value.code == 6
    with pytest.raises(ValueError):
        qsm.tv_admm(np.zeros(10), np.ones(10, np.uint8), qsm.Grid((4, 4, 4)), (0, 0, 1), p)


def test_tv_full_size_256x256x160_vs_oracle(qsm, oracle, phantom):
    # The benchmark workload at full size (0.9x0.9x1.0 mm, 40 ellipsoids,
    # lambda 1e-3, rho 0.1, tol 0), cut to 10 iterations so the f64 oracle
    # finishes in seconds.  Bar: NRMSE <= 1e-3 vs f64 (range-normalised,
    # tests/common/mod.rs:34-51), after removing the mean difference that the
    # reference's DC term leaves for anisotropic voxels; raw NRMSE and
    # max-abs are printed.
    ph = phantom.make_phantom((256, 256, 160), voxel_size=(0.9, 0.9, 1.0), n_ellipsoids=40,
                              seed=1001)
    grid = qsm.Grid(ph.dims, ph.voxel_size)
    p = qsm.TvParams(lambda_=1e-3, rho=0.1, tol=0.0, max_iter=10)
    x, it, rel = qsm.tv_admm(ph.local_field, ph.mask, grid, ph.bdir, p, return_info=True)
    xo, oit, orel = oracle.tv_admm(ph.local_field, ph.mask, ph.dims, ph.voxel_size, ph.bdir,
                                   1e-3, 0.1, 0.0, 10, return_info=True)
    e = nrmse_dc(x, xo, ph.mask)
    print("256x256x160 x10: nrmse(dc removed)", e, "raw", oracle.nrmse(x, xo, ph.mask),
          "maxabs", float(np.max(np.abs(x - xo))), "rel", rel, orel)
    assert it == oit == 10
    assert e <= 1e-3


def test_tv_uneven_batch_fixed_lengths(qsm, oracle, phantom):
    # Volumes of different specialised axis lengths, voxel sizes and tilted
    # B0 in one batch (grouped launches per length): each equals its own
    # single-volume solve bit for bit, and one is checked against the oracle.
    shapes = [((144, 176, 160), (0.7, 1.1, 0.9)), ((176, 144, 192), (1.0, 0.8, 0.6)),
              ((160, 160, 144), (1.2, 1.2, 0.9))]
    p = qsm.TvParams(lambda_=1e-3, rho=0.1, tol=0.0, max_iter=8)
    keep, outs, vols, singles, phs = [], [], [], [], []
    for v, (dims, vs) in enumerate(shapes):
        b = phantom.tilted_bdir(2000 + v)
        ph = phantom.make_phantom(dims, voxel_size=vs, bdir=b, n_ellipsoids=12, seed=2000 + v)
        phs.append(ph)
        singles.append(qsm.tv_admm(ph.local_field, ph.mask, qsm.Grid(dims, vs), b, p))
        tf = qsm.DeviceArray.from_numpy(ph.local_field)
        tm = qsm.DeviceArray.from_numpy(ph.mask)
        tc = qsm.DeviceArray(ph.local_field.size, np.float64)
        keep.append((tf, tm))
        outs.append(tc)
        vols.append((qsm.Grid(dims, vs), b, tf.ptr, tm.ptr, tc.ptr))
    batch = qsm.TvBatch(vols, p)
    batch.run()
    for v in range(len(shapes)):
        assert np.array_equal(outs[v].numpy(), singles[v]), v
        assert batch.status(v)[0] == 8
    batch.close()
    ph = phs[1]
    xo = oracle.tv_admm(ph.local_field, ph.mask, ph.dims, ph.voxel_size, ph.bdir, 1e-3, 0.1, 0.0, 8)
    e = nrmse_dc(singles[1], xo, ph.mask)
    print("uneven batch volume 1 nrmse(dc removed)", e, "raw", oracle.nrmse(singles[1], xo, ph.mask))
    assert e <= 1e-3


@pytest.mark.parametrize("dims", [(256, 1, 8), (128, 6, 1), (160, 1, 1), (144, 2, 3), (240, 3, 2)])
def test_tv_degenerate_axes(qsm, oracle, dims):
    # compiled x lengths (the wave-row divergence fill: 256 and 128 in whole
    # 64-column groups, 160 / 144 / 240 with a partial last group) with y or
    # z of length 1 or 2, where every y or z neighbour is the voxel itself
    # or wraps at once; 5 iterations against the oracle
    rng = np.random.default_rng(sum(dims))
    f, m = _case(rng, dims)
    vs, b = (1.0, 1.0, 1.0), (0.1, 0.2, 1.0)
    p = qsm.TvParams(lambda_=1e-3, rho=0.1, tol=0.0, max_iter=5)
    x = qsm.tv_admm(f, m, qsm.Grid(dims, vs), b, p)
    xo = oracle.tv_admm(f, m, dims, vs, b, 1e-3, 0.1, 0.0, 5)
    e = float(oracle.nrmse(x, xo, m))
    print(dims, f"nrmse {e:.3e}", flush=True)
    assert e < 2e-5, e
    assert np.all(x[m == 0] == 0.0)
