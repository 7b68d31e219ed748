"""GPU closed-form inversions (tkd / tsvd / tikhonov), RTS, SHARP and the
pipeline arms that dispatch to them, against the f64 oracle.  FP32
transforms: relative L2 error <= 1e-5 for the closed forms (one transform
pair); RTS NRMSE <= 1e-3 (tests/common/mod.rs:34-51), tighter on small
volumes."""
import numpy as np
import pytest

from conftest import nrmse_dc

pytestmark = pytest.mark.gpu


def _rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _case(seed, dims, frac=0.8):
    rng = np.random.default_rng(seed)
    n = int(np.prod(dims))
    return rng.standard_normal(n) * 0.02, (rng.random(n) < frac).astype(np.uint8)


SHAPES = [((16, 16, 16), (1.0, 1.0, 1.0), (0.0, 0.0, 1.0)),
          ((20, 18, 22), (0.9, 1.1, 1.3), (0.2, -0.1, 0.97)),
          ((64, 64, 64), (1.0, 1.0, 1.0), (0.1, 0.2, 1.0)),
          ((128, 64, 64), (0.6, 0.6, 0.6), (0.0, 0.0, 1.0))]


@pytest.mark.parametrize("dims,vs,b", SHAPES)
def test_closed_forms_match_oracle(qsm, oracle, dims, vs, b):
    f, m = _case(3, dims)
    g = qsm.Grid(dims, vs)
    got = {
        "tkd": qsm.tkd(f, m, g, b, qsm.TkdParams(0.15)),
        "tsvd": qsm.tsvd(f, m, g, b, qsm.TkdParams(0.15)),
        "tik_id": qsm.tikhonov(f, m, g, b, qsm.TikhonovParams(0.01, qsm.Regularization.Identity)),
        "tik_grad": qsm.tikhonov(f, m, g, b, qsm.TikhonovParams(0.01, qsm.Regularization.Gradient)),
        "tik_lap": qsm.tikhonov(f, m, g, b, qsm.TikhonovParams(0.01, qsm.Regularization.Laplacian)),
    }
    ref = {
        "tkd": oracle.tkd(f, m, dims, vs, b, 0.15),
        "tsvd": oracle.tkd(f, m, dims, vs, b, 0.15, tsvd=True),
        "tik_id": oracle.tikhonov(f, m, dims, vs, b, 0.01, 0),
        "tik_grad": oracle.tikhonov(f, m, dims, vs, b, 0.01, 1),
        "tik_lap": oracle.tikhonov(f, m, dims, vs, b, 0.01, 2),
    }
    for k in got:
        e = _rel_l2(got[k], ref[k])
        print(dims, k, "rel L2", e)
        assert e <= 1e-5, k
        assert np.all(got[k][m == 0] == 0.0)


@pytest.mark.parametrize("dims", [(64, 1, 8), (20, 2, 1), (1, 16, 12), (2, 24, 2)])
def test_degenerate_axes_match_oracle(qsm, oracle, dims):
    # axes of length 1 or 2, where the reference's Laplacian stencil is not
    # the periodic second difference (its DFT keeps 2/vs^2 on a length-1
    # axis and 1/vs^2 or 3/vs^2 on a length-2 one): Tikhonov with the
    # gradient and Laplacian terms, RTS, NLTV, and the Poisson unwrap, whose
    # eigenvalues stay the periodic ones
    vs, b = (0.9, 1.1, 1.2), (0.1, 0.2, 1.0)
    f, m = _case(5, dims)
    g = qsm.Grid(dims, vs)
    for reg, ri in [(qsm.Regularization.Gradient, 1), (qsm.Regularization.Laplacian, 2)]:
        e = _rel_l2(qsm.tikhonov(f, m, g, b, qsm.TikhonovParams(0.01, reg)),
                    oracle.tikhonov(f, m, dims, vs, b, 0.01, ri))
        print(dims, reg, "rel L2", e, flush=True)
        assert e <= 1e-5
    x = qsm.rts(f, m, g, b, qsm.RtsParams(tol=0.0, max_iter=6))
    xo = oracle.rts(f, m, dims, vs, b, 0.15, 1e5, 10.0, 0.0, 6, 4)
    e = float(oracle.nrmse(x, xo, m))
    print(dims, "rts nrmse", e, flush=True)
    assert e <= 2e-5
    ph = np.mod(f * 400.0 + np.pi, 2 * np.pi) - np.pi
    e = _rel_l2(qsm.laplacian_unwrap(ph, m, g), oracle.laplacian_unwrap(ph, m, dims, vs))
    print(dims, "unwrap rel L2", e, flush=True)
    assert e <= 1e-5
    p = qsm.NltvParams(lambda_=2e-3, mu=1.0, tol=0.0, max_iter=4, newton_iter=2)
    e = float(oracle.nrmse(qsm.nltv(f, m, g, b, p), oracle.nltv(f, m, dims, vs, b, 2e-3, 1.0, 0.0, 4, 2), m))
    print(dims, "nltv nrmse", e, flush=True)
    assert e <= 2e-5


@pytest.mark.parametrize("dims,vs,b", SHAPES[:3])
def test_rts_matches_oracle(qsm, oracle, dims, vs, b):
    f, m = _case(4, dims)
    g = qsm.Grid(dims, vs)
    for tol, iters in [(0.0, 10), (1e-2, 20)]:
        p = qsm.RtsParams(tol=tol, max_iter=iters)
        calls, ocalls = [], []
        x = qsm.rts(f, m, g, b, p, progress=lambda a, c: calls.append((a, c)))
        xo, oit = oracle.rts(f, m, dims, vs, b, 0.15, 1e5, 10.0, tol, iters, 4,
                             progress=lambda a, c: ocalls.append((a, c)), return_info=True)
        # anisotropic voxels: the reference's DC term (FFT of the stencil leaves
        # ~1e-16 at k = 0) is compared apart, as for TV
        e = nrmse_dc(x, xo, m)
        print(dims, "tol", tol, "gpu calls", len(calls), "oracle iters", oit, "nrmse(dc)", e,
              "raw", oracle.nrmse(x, xo, m))
        assert e <= (2e-5 if tol == 0.0 else 1e-3)
        if tol == 0.0:
            assert calls == ocalls
        assert np.all(x[m == 0] == 0.0)


def test_rts_default_pipeline_arm(qsm, oracle, phantom):
    # InversionConfig::default() is Rts (src/pipeline/config.rs:251)
    ph = phantom.make_phantom((64, 64, 64), n_ellipsoids=8, seed=1006)
    meta = qsm.ScanMetadata(dims=ph.dims, voxel_size=ph.voxel_size, b0_direction=ph.bdir)
    x = qsm.run_dipole_inversion(ph.local_field, ph.mask, meta, qsm.InversionConfig())
    xo = oracle.rts(ph.local_field, ph.mask, ph.dims, ph.voxel_size, ph.bdir)
    e = oracle.nrmse(x, xo, ph.mask)
    print("pipeline rts 64^3 nrmse", e)
    assert e <= 1e-3
    cfg = qsm.InversionConfig(algorithm=qsm.InversionAlgorithm.Tkd)
    t = qsm.run_dipole_inversion(ph.local_field, ph.mask, meta, cfg)
    assert np.array_equal(t, qsm.tkd(ph.local_field, ph.mask, qsm.Grid(ph.dims, ph.voxel_size),
                                     ph.bdir, qsm.TkdParams()))


def test_rts_batch_equals_single(qsm):
    shapes = [((16, 16, 16), (1.0, 1.0, 1.0), (0, 0, 1)), ((20, 18, 22), (1.0, 1.0, 1.0), (0.1, 0.2, 0.95))]
    p = qsm.RtsParams(tol=0.0, max_iter=8)
    keep, vols, singles, outs = [], [], [], []
    for s, (dims, vs, b) in enumerate(shapes):
        f, m = _case(10 + s, dims)
        singles.append(qsm.rts(f, m, qsm.Grid(dims, vs), b, p))
        tf, tm = qsm.DeviceArray.from_numpy(f), qsm.DeviceArray.from_numpy(m)
        tc = qsm.DeviceArray(f.size, np.float64)
        keep.append((tf, tm))
        outs.append(tc)
        vols.append((qsm.Grid(dims, vs), b, tf.ptr, tm.ptr, tc.ptr))
    batch = qsm.TvBatch(vols, p)
    batch.run()
    for s in range(len(shapes)):
        assert np.array_equal(outs[s].numpy(), singles[s])
    batch.close()


def test_sharp_matches_oracle_and_pipeline(qsm, oracle, phantom):
    ph = phantom.add_background(phantom.make_phantom((64, 64, 64), n_ellipsoids=8, seed=1007),
                                seed=1007)
    g = qsm.Grid(ph.dims, ph.voxel_size)
    p = qsm.SharpParams(0.05, 4.0)
    local, em = qsm.sharp(ph.total_field, ph.mask, g, p)
    lo, mo = oracle.sharp_core(ph.total_field, ph.mask, ph.dims, ph.voxel_size, 0.05, 4.0)
    assert np.array_equal(em, mo)
    e = _rel_l2(local, lo)
    print("sharp 64^3 rel L2", e)
    assert e <= 1e-4
    meta = qsm.ScanMetadata(dims=ph.dims, voxel_size=ph.voxel_size)
    r = qsm.run_bg_removal(ph.total_field, ph.mask, meta,
                           qsm.BgRemovalConfig(algorithm=qsm.BgRemovalAlgorithm.Sharp, sharp=p))
    assert np.array_equal(r.local_field_ppm, local) and np.array_equal(r.eroded_mask, em)


@pytest.mark.parametrize("dims,vs", [((16, 12, 10), (1.0, 0.8, 1.2)), ((64, 64, 64), (1.0, 1.0, 1.0)),
                                     ((128, 128, 64), (0.9, 0.9, 0.9))])
def test_laplacian_unwrap_matches_oracle(qsm, oracle, dims, vs):
    # wrapped smooth phase with mask: GPU vs oracle; with an all-ones mask the
    # unwrapped phase is the true phase minus its mean
    i, j, k = [np.arange(q) for q in dims]
    true = (6.0 * np.sin(2 * np.pi * i[:, None, None] / dims[0])
            + 4.0 * np.cos(2 * np.pi * j[None, :, None] / dims[1])
            + 3.0 * np.sin(2 * np.pi * k[None, None, :] / dims[2])).ravel(order="F")
    wrapped = np.mod(true + np.pi, 2 * np.pi) - np.pi
    rng = np.random.default_rng(1)
    m = (rng.random(true.size) < 0.9).astype(np.uint8)
    g = qsm.Grid(dims, vs)
    got = qsm.laplacian_unwrap(wrapped, m, g)
    ref = oracle.laplacian_unwrap(wrapped, m, dims, vs)
    e = _rel_l2(got, ref)
    ones = np.ones(true.size, np.uint8)
    full = qsm.laplacian_unwrap(wrapped, ones, g)
    ef = float(np.max(np.abs(full - (true - true.mean()))))
    print(dims, "unwrap rel L2 vs oracle", e, "max err vs truth", ef)
    assert e <= 1e-5
    assert ef <= 1e-4 * np.max(np.abs(true))
    assert np.all(got[m == 0] == 0.0)


@pytest.mark.parametrize("dims,vs,b", [((16, 16, 16), (1.0, 1.0, 1.0), (0.0, 0.0, 1.0)),
                                       ((20, 18, 22), (0.9, 1.1, 1.3), (0.2, -0.1, 0.97)),
                                       ((64, 64, 64), (1.0, 1.0, 1.0), (0.1, 0.0, 1.0))])
def test_nltv_matches_oracle(qsm, oracle, dims, vs, b):
    f, m = _case(5, dims)
    g = qsm.Grid(dims, vs)
    p = qsm.NltvParams(lambda_=2e-3, mu=1.0, tol=0.0, max_iter=5, newton_iter=3)
    calls, ocalls = [], []
    x = qsm.nltv(f, m, g, b, p, progress=lambda a, c: calls.append((a, c)))
    xo = oracle.nltv(f, m, dims, vs, b, 2e-3, 1.0, 0.0, 5, 3,
                     progress=lambda a, c: ocalls.append((a, c)))
    e = nrmse_dc(x, xo, m)
    print(dims, "nltv nrmse(dc)", e, "raw", oracle.nrmse(x, xo, m))
    assert calls == ocalls == [(i, 15) for i in range(1, 16)]
    assert e <= 2e-5
    if len(set(vs)) > 1:
        # anisotropic voxels: the reference's DC term enters the relative
        # change, so where rounds stop early depends on its rounding noise
        return
    # with a tolerance the rounds stop early; the result stays within the bar
    p2 = qsm.NltvParams(lambda_=2e-3, mu=1.0, tol=2e-2, max_iter=30, newton_iter=3)
    x2 = qsm.nltv(f, m, g, b, p2)
    xo2 = oracle.nltv(f, m, dims, vs, b, 2e-3, 1.0, 2e-2, 30, 3)
    e2 = nrmse_dc(x2, xo2, m)
    print(dims, "nltv tol 2e-2 nrmse(dc)", e2)
    assert e2 <= 1e-3


def test_nltv_pipeline_arm(qsm):
    dims = (16, 16, 16)
    f, m = _case(6, dims)
    meta = qsm.ScanMetadata(dims=dims, voxel_size=(1.0, 1.0, 1.0))
    p = qsm.NltvParams(lambda_=2e-3, tol=0.0, max_iter=4, newton_iter=2)
    cfg = qsm.InversionConfig(algorithm=qsm.InversionAlgorithm.Nltv, nltv=p)
    x = qsm.run_dipole_inversion(f, m, meta, cfg)
    assert np.array_equal(x, qsm.nltv(f, m, qsm.Grid(dims), (0, 0, 1), p))
