"""Axis lengths with a prime factor above 13 (the chirp-z transform), on the
GPU against the f64 oracle, whose FFT takes any length as rustfft does
(src/fft.rs:68-74).  Each length runs on every axis in turn: x through the
row passes (R2C / C2R), y and z through the column passes.

Lengths: 17 and 97 (primes), 68 = 4 x 17, 164 = 4 x 41, 232 = 8 x 29 (the
matrix sizes a drop-in must accept, where the reference computes a result),
251, 254 = 2 x 127; above 256 the chirp-z primes 263 and 509 (convolution
length 1024), and the direct lengths above 512 (576, 640, 1000, 1024), whose
column passes run in tiles of 8 columns."""
import numpy as np
import pytest

from conftest import nrmse_dc

pytestmark = pytest.mark.gpu

LENGTHS = [17, 68, 97, 164, 232, 251, 254]


def _case(rng, dims, frac=0.8, scale=0.01):
    n = int(np.prod(dims))
    return rng.standard_normal(n) * scale, (rng.random(n) < frac).astype(np.uint8)


@pytest.mark.parametrize("n", LENGTHS)
def test_chirp_lengths_supported(qsm, n):
    assert qsm.fft_length_supported(n)


@pytest.mark.parametrize("n", LENGTHS)
def test_tv_chirp_length_each_axis(qsm, oracle, n):
    # TV-ADMM (src/inversion/tv.rs:63-97), 5 iterations, tol 0, tilted B0;
    # anisotropic voxels: NRMSE with the reference's DC term removed, raw
    # printed; isotropic: raw NRMSE
    rng = np.random.default_rng(n)
    for dims in [(n, 6, 4), (4, n, 6), (6, 4, n)]:
        for vs in ((1.0, 0.9, 1.1), (1.0, 1.0, 1.0)):
            f, m = _case(rng, dims)
            bdir = (0.1, -0.2, 1.0)
            p = qsm.TvParams(lambda_=1e-3, rho=0.1, tol=0.0, max_iter=5)
            x = qsm.tv_admm(f, m, qsm.Grid(dims, vs), bdir, p)
            xo = oracle.tv_admm(f, m, dims, vs, bdir, 1e-3, 0.1, 0.0, 5)
            raw = float(oracle.nrmse(x, xo, m))
            dc = nrmse_dc(x, xo, m)
            print(dims, vs, f"nrmse raw {raw:.3e} dc-removed {dc:.3e}", flush=True)
            assert dc < 2e-5, (dims, vs, dc)
            if vs == (1.0, 1.0, 1.0):
                assert raw < 2e-5, (dims, raw)
            assert np.all(x[m == 0] == 0.0)


@pytest.mark.parametrize("dims", [(68, 164, 24), (97, 20, 232), (36, 251, 17)])
def test_tv_chirp_lengths_together(qsm, oracle, dims):
    # several chirp-z axes in one volume, 20 iterations
    rng = np.random.default_rng(sum(dims))
    f, m = _case(rng, dims)
    vs, bdir = (1.0, 1.0, 1.0), (0.0, 0.1, 1.0)
    p = qsm.TvParams(lambda_=1e-3, rho=0.1, tol=0.0, max_iter=20)
    x = qsm.tv_admm(f, m, qsm.Grid(dims, vs), bdir, p)
    xo = oracle.tv_admm(f, m, dims, vs, bdir, 1e-3, 0.1, 0.0, 20)
    raw = float(oracle.nrmse(x, xo, m))
    print(dims, f"nrmse raw {raw:.3e}", flush=True)
    assert raw < 2e-5, raw


@pytest.mark.parametrize("dims", [(68, 48, 40), (40, 164, 36), (48, 36, 97)])
def test_vsharp_chirp_lengths(qsm, oracle, dims):
    # V-SHARP (src/bgremove/vsharp.rs:88-199): eroded mask identical, local
    # field within 1e-4 relative
    g = [np.arange(n) - n // 2 for n in dims]
    X, Y, Z = np.meshgrid(*g, indexing="ij")
    r = [0.35 * n for n in dims]
    mask = (((X / r[0]) ** 2 + (Y / r[1]) ** 2 + (Z / r[2]) ** 2) <= 1.0).astype(np.uint8).ravel(order="F")
    rng = np.random.default_rng(dims[0])
    field = (0.01 * X + 0.02 * Y - 0.01 * Z).ravel(order="F") + rng.standard_normal(mask.size) * 0.01
    vs = (1.0, 1.0, 1.0)
    radii = [4.0, 3.0, 2.0]
    local, em = qsm.vsharp_with_radii(field, mask, qsm.Grid(dims, vs), radii, 0.05)
    lo, emo = oracle.vsharp_with_radii(field, mask, dims, vs, radii, 0.05)
    assert np.array_equal(em, emo), int(np.sum(em != emo))
    sel = emo > 0
    err = float(np.max(np.abs(local[sel] - lo[sel])) / np.max(np.abs(lo[sel])))
    print(dims, f"max rel err {err:.3e}", flush=True)
    assert err < 1e-4, err


@pytest.mark.parametrize("dims,bdir", [((68, 17, 24), (0.2, 0.1, 1.0)), ((30, 232, 20), (0.0, 0.0, 1.0))])
def test_dipole_chirp_lengths(qsm, oracle, dims, bdir):
    # the dipole forward model Re IFFT(D FFT(chi)) (src/kernels/dipole.rs:24-72)
    rng = np.random.default_rng(7)
    chi = rng.standard_normal(int(np.prod(dims))) * 0.05
    vs = (0.9, 1.0, 1.1)
    out = qsm.dipole_forward(chi, qsm.Grid(dims, vs), bdir)
    ref = oracle.dipole_forward(chi, dims, vs, bdir)
    err = float(np.max(np.abs(out - ref)) / np.max(np.abs(ref)))
    print(dims, f"max rel err {err:.3e}", flush=True)
    assert err < 2e-5, err


LONG = [576, 640, 1000, 1024]
CHIRP_LONG = [263, 509]


@pytest.mark.parametrize("n", LONG + CHIRP_LONG)
def test_long_lengths_supported(qsm, n):
    assert qsm.fft_length_supported(n)


@pytest.mark.parametrize("n", [576, 1024, 509])
def test_tv_long_length_each_axis(qsm, oracle, n):
    # TV-ADMM, 5 iterations, tol 0, tilted B0, isotropic voxels: raw NRMSE
    rng = np.random.default_rng(n + 1)
    for dims in [(n, 4, 3), (3, n, 4), (4, 3, n)]:
        f, m = _case(rng, dims)
        bdir = (0.1, -0.2, 1.0)
        p = qsm.TvParams(lambda_=1e-3, rho=0.1, tol=0.0, max_iter=5)
        x = qsm.tv_admm(f, m, qsm.Grid(dims, (1.0, 1.0, 1.0)), bdir, p)
        xo = oracle.tv_admm(f, m, dims, (1.0, 1.0, 1.0), bdir, 1e-3, 0.1, 0.0, 5)
        raw = float(oracle.nrmse(x, xo, m))
        print(dims, f"nrmse raw {raw:.3e}", flush=True)
        assert raw < 2e-5, (dims, raw)
        assert np.all(x[m == 0] == 0.0)


@pytest.mark.parametrize("dims", [(640, 12, 10), (10, 1000, 8), (12, 8, 576), (263, 20, 16)])
def test_tv_long_lengths_mixed(qsm, oracle, dims):
    # 20 iterations, anisotropic voxels: DC-removed NRMSE asserted, raw printed
    rng = np.random.default_rng(sum(dims))
    f, m = _case(rng, dims)
    vs, bdir = (0.9, 1.0, 1.1), (0.0, 0.1, 1.0)
    p = qsm.TvParams(lambda_=1e-3, rho=0.1, tol=0.0, max_iter=20)
    x = qsm.tv_admm(f, m, qsm.Grid(dims, vs), bdir, p)
    xo = oracle.tv_admm(f, m, dims, vs, bdir, 1e-3, 0.1, 0.0, 20)
    raw = float(oracle.nrmse(x, xo, m))
    dc = nrmse_dc(x, xo, m)
    print(dims, f"nrmse raw {raw:.3e} dc-removed {dc:.3e}", flush=True)
    assert dc < 2e-5, dc


@pytest.mark.parametrize("dims", [(576, 32, 24), (24, 640, 20), (20, 24, 509)])
def test_vsharp_long_lengths(qsm, oracle, dims):
    # V-SHARP: eroded mask identical, local field within 1e-4 of its largest
    # value.  The GPU works in FP32, so its error follows the range of the
    # background it removes: the ramp spans +-0.32..0.64 rad along every axis
    # (as over the 64-voxel axes of the other V-SHARP tests), not 0.02 rad
    # per voxel over 640 voxels.
    g = [np.arange(n) - n // 2 for n in dims]
    X, Y, Z = np.meshgrid(*g, indexing="ij")
    r = [0.35 * n for n in dims]
    mask = (((X / r[0]) ** 2 + (Y / r[1]) ** 2 + (Z / r[2]) ** 2) <= 1.0).astype(np.uint8).ravel(order="F")
    rng = np.random.default_rng(dims[1])
    sx, sy, sz = (64.0 / n for n in dims)
    field = (0.01 * sx * X + 0.02 * sy * Y - 0.01 * sz * Z).ravel(order="F") + rng.standard_normal(mask.size) * 0.01
    vs = (1.0, 1.0, 1.0)
    radii = [3.0, 2.0]
    local, em = qsm.vsharp_with_radii(field, mask, qsm.Grid(dims, vs), radii, 0.05)
    lo, emo = oracle.vsharp_with_radii(field, mask, dims, vs, radii, 0.05)
    assert np.array_equal(em, emo), int(np.sum(em != emo))
    sel = emo > 0
    err = float(np.max(np.abs(local[sel] - lo[sel])) / np.max(np.abs(lo[sel])))
    print(dims, f"max rel err {err:.3e}", flush=True)
    assert err < 1e-4, err


@pytest.mark.parametrize("dims,bdir", [((1024, 6, 4), (0.2, 0.1, 1.0)), ((6, 4, 768), (0.0, 0.3, 1.0))])
def test_dipole_long_lengths(qsm, oracle, dims, bdir):
    rng = np.random.default_rng(11)
    chi = rng.standard_normal(int(np.prod(dims))) * 0.05
    vs = (0.9, 1.0, 1.1)
    out = qsm.dipole_forward(chi, qsm.Grid(dims, vs), bdir)
    ref = oracle.dipole_forward(chi, dims, vs, bdir)
    err = float(np.max(np.abs(out - ref)) / np.max(np.abs(ref)))
    print(dims, f"max rel err {err:.3e}", flush=True)
    assert err < 2e-5, err


def test_batch_mixing_runtime_and_compiled_lengths(qsm, oracle):
    # one TV batch with a chirp-z volume (232: convolution length 512 on the
    # compile-time stages), a direct 1024-point axis (8-column tiles) and a
    # compiled volume: every volume equals its single-volume solve bit for
    # bit (the launch groups do not change the arithmetic), and the chirp-z
    # one matches the oracle at 10 iterations
    rng = np.random.default_rng(31)
    shapes = [((232, 24, 16), (1.0, 0.9, 1.1), (0.1, 0.2, 1.0)),
              ((8, 1024, 6), (1.0, 1.0, 1.0), (0.0, 0.1, 1.0)),
              ((64, 48, 40), (0.9, 0.9, 1.0), (0.0, 0.0, 1.0))]
    p = qsm.TvParams(lambda_=1e-3, rho=0.1, tol=0.0, max_iter=10)
    keep, vols, singles, outs, cases = [], [], [], [], []
    for dims, vs, b in shapes:
        f, m = _case(rng, dims)
        cases.append((f, m))
        singles.append(qsm.tv_admm(f, m, qsm.Grid(dims, vs), b, p))
        tf, tm = qsm.DeviceArray.from_numpy(f), qsm.DeviceArray.from_numpy(m)
        tc = qsm.DeviceArray(f.size, np.float64)
        keep.append((tf, tm))
        outs.append(tc)
        vols.append((qsm.Grid(dims, vs), b, tf.ptr, tm.ptr, tc.ptr))
    batch = qsm.TvBatch(vols, p)
    try:
        batch.run()
        for v in range(len(shapes)):
            assert np.array_equal(outs[v].numpy(), singles[v]), shapes[v][0]
    finally:
        batch.close()
    (dims, vs, b), (f, m) = shapes[0], cases[0]
    xo = oracle.tv_admm(f, m, dims, vs, b, 1e-3, 0.1, 0.0, 10)
    dc = nrmse_dc(singles[0], xo, m)
    print(dims, f"nrmse dc-removed {dc:.3e} raw {float(oracle.nrmse(singles[0], xo, m)):.3e}", flush=True)
    assert dc < 2e-5, dc


@pytest.mark.parametrize("dims", [(68, 40, 24), (24, 164, 20)])
def test_rts_nltv_chirp_lengths(qsm, oracle, dims):
    # RTS (stored q operand, src/inversion/rts.rs:71-212) and NLTV
    # (src/inversion/nltv.rs:63-133) through the chirp-z transforms (68 and
    # 164 run convolutions of length 256 and 512 on the compile-time stages)
    rng = np.random.default_rng(dims[1])
    f, m = _case(rng, dims)
    vs, b = (1.0, 1.0, 1.0), (0.1, 0.0, 1.0)
    g = qsm.Grid(dims, vs)
    x = qsm.rts(f, m, g, b, qsm.RtsParams(tol=0.0, max_iter=10))
    xo = oracle.rts(f, m, dims, vs, b, 0.15, 1e5, 10.0, 0.0, 10, 4)
    e = float(oracle.nrmse(x, xo, m))
    print(dims, f"rts nrmse {e:.3e}", flush=True)
    assert e < 2e-5, e
    p = qsm.NltvParams(lambda_=2e-3, mu=1.0, tol=0.0, max_iter=5, newton_iter=3)
    x = qsm.nltv(f, m, g, b, p)
    xo = oracle.nltv(f, m, dims, vs, b, 2e-3, 1.0, 0.0, 5, 3)
    e = float(oracle.nrmse(x, xo, m))
    print(dims, f"nltv nrmse {e:.3e}", flush=True)
    assert e < 2e-5, e
