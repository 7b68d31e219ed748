"""GPU parity of V-SHARP / SHARP (src/bgremove/vsharp.rs:88-199,
src/bgremove/sharp.rs:64-131) against the f64 oracle.  Eroded masks must be
identical; the local field is compared at FP32 tolerance."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def sphere_mask(n, radius):
    g = np.arange(n) - n // 2
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    return ((X * X + Y * Y + Z * Z) <= radius * radius).astype(np.uint8).ravel(order="F")


def _rel(a, b, m=None):
    if m is None:
        m = np.ones(a.size, bool)
    m = m.astype(bool)
    d = np.max(np.abs(a[m] - b[m])) if m.any() else 0.0
    s = np.max(np.abs(b[m])) if m.any() else 1.0
    return d / (s if s > 0 else 1.0)


@pytest.mark.parametrize("n,vs,radii", [
    (16, (1.0, 1.0, 1.0), [6.0, 4.0, 3.0, 2.0]),
    (16, (0.8, 1.0, 1.2), [4.0, 3.0, 2.0]),
    (24, (1.0, 1.0, 1.0), [5.0, 4.0, 3.0, 2.0, 1.0]),
    (20, (0.6, 0.6, 0.6), [3.0, 2.4, 1.8]),
    # radii on exact voxel multiples where rounding decides the boundary:
    # (4,3,0) * 0.9 mm is at exactly 4.5 mm only without fused multiply-add
    (32, (0.9, 0.9, 0.9), [4.5, 3.6, 2.7, 1.8]),
    (32, (0.9, 0.9, 0.9), [4.5]),
])
def test_vsharp_matches_oracle(qsm, oracle, n, vs, radii):
    rng = np.random.default_rng(n)
    dims = (n, n, n)
    g = np.arange(n)
    Z = np.broadcast_to(g[None, None, :], dims).ravel(order="F")
    field = Z * 0.1 + rng.standard_normal(n ** 3) * 0.01
    mask = sphere_mask(n, n / 3)
    grid = qsm.Grid(dims, vs)
    local, m = qsm.vsharp_with_radii(field, mask, grid, radii, 0.05)
    lo, mo = oracle.vsharp_with_radii(field, mask, dims, vs, radii, 0.05)
    assert np.array_equal(m, mo)
    err = _rel(local, lo, mo)
    print(n, vs, radii, "mask voxels", int(mo.sum()), "max rel err", err)
    assert err < 1e-4
    assert np.all(local[m == 0] == 0.0)


@pytest.mark.parametrize("dims", [(512, 16, 12), (12, 512, 10), (10, 12, 512),
                                  (288, 16, 12), (12, 448, 10), (10, 12, 480)])
def test_vsharp_long_axis_matches_oracle(qsm, oracle, dims):
    # one long axis of a compiled length (x row passes; y and z staged columns)
    rng = np.random.default_rng(sum(dims))
    axes = [np.arange(n) - (n - 1) / 2 for n in dims]
    X, Y, Z = np.meshgrid(*axes, indexing="ij")
    a = [n / 2 - 2 for n in dims]
    mask = ((X / a[0]) ** 2 + (Y / a[1]) ** 2 + (Z / a[2]) ** 2 <= 1.0).astype(np.uint8).ravel(order="F")
    field = (0.01 * X + 0.02 * Y + 0.03 * Z).ravel(order="F") + rng.standard_normal(mask.size) * 0.01
    radii = [3.0, 2.0, 1.0]
    local, m = qsm.vsharp_with_radii(field, mask, qsm.Grid(dims), radii, 0.05)
    lo, mo = oracle.vsharp_with_radii(field, mask, dims, (1.0, 1.0, 1.0), radii, 0.05)
    assert np.array_equal(m, mo)
    err = _rel(local, lo, mo)
    print(dims, "mask voxels", int(mo.sum()), "max rel err", err)
    assert mo.sum() > 0 and err < 1e-4
    assert np.all(local[m == 0] == 0.0)


@pytest.mark.parametrize("dims", [(32, 1, 24), (24, 2, 20), (1, 20, 18), (30, 28, 1), (2, 2, 16)])
def test_vsharp_degenerate_axes_match_oracle(qsm, oracle, dims):
    # axes of length 1 or 2: the sphere wraps onto itself along them (every
    # offset along a length-1 axis is the voxel itself), so the erosion and
    # the deconvolution see the periodic kernel at its smallest
    rng = np.random.default_rng(sum(dims))
    axes = [np.arange(n) - (n - 1) / 2 for n in dims]
    X, Y, Z = np.meshgrid(*axes, indexing="ij")
    a = [max(n / 2 - 1, 1.0) for n in dims]
    mask = ((X / a[0]) ** 2 + (Y / a[1]) ** 2 + (Z / a[2]) ** 2 <= 1.0).astype(np.uint8).ravel(order="F")
    field = (0.1 * X + 0.05 * Z).ravel(order="F") + rng.standard_normal(mask.size) * 0.01
    vs, radii = (1.0, 0.9, 1.1), [3.0, 2.0, 1.0]
    local, m = qsm.vsharp_with_radii(field, mask, qsm.Grid(dims, vs), radii, 0.05)
    lo, mo = oracle.vsharp_with_radii(field, mask, dims, vs, radii, 0.05)
    assert np.array_equal(m, mo)
    err = _rel(local, lo, mo)
    print(dims, "mask voxels", int(mo.sum()), "max rel err", err, flush=True)
    assert err < 1e-4
    assert np.all(local[m == 0] == 0.0)


def test_single_radius_is_sharp(qsm, oracle):
    # src/bgremove/vsharp.rs:341-371
    n = 8
    dims = (n, n, n)
    field = np.arange(n ** 3) * 0.001
    mask = np.ones(n ** 3, np.uint8)
    local, m = qsm.vsharp_with_radii(field, mask, qsm.Grid(dims), [3.0], 0.05)
    lo, mo = oracle.sharp_core(field, mask, dims, (1, 1, 1), 0.05, 3.0)
    assert np.array_equal(m, mo)
    assert np.max(np.abs(local - lo)) < 1e-5 * max(1.0, np.max(np.abs(lo)))


def test_empty_radii(qsm):
    n = 8
    field = np.arange(n ** 3) * 0.001
    mask = np.ones(n ** 3, np.uint8)
    calls = []
    local, m = qsm.vsharp_with_radii(field, mask, qsm.Grid((n, n, n)), [], 0.05,
                                     progress=lambda a, b: calls.append((a, b)))
    assert np.all(local == 0.0) and np.array_equal(m, mask) and calls == []


def test_unsorted_radii(qsm):
    n = 8
    field = np.arange(n ** 3) * 0.001
    mask = np.ones(n ** 3, np.uint8)
    g = qsm.Grid((n, n, n))
    a = qsm.vsharp_with_radii(field, mask, g, [4.0, 3.0, 2.0], 0.05)
    b = qsm.vsharp_with_radii(field, mask, g, [2.0, 4.0, 3.0], 0.05)
    assert np.array_equal(a[1], b[1]) and np.array_equal(a[0], b[0])


def test_progress_sequence(qsm, oracle):
    n = 8
    field = np.arange(n ** 3) * 0.001
    mask = np.ones(n ** 3, np.uint8)
    calls, ocalls = [], []
    qsm.vsharp_with_radii(field, mask, qsm.Grid((n, n, n)), [4.0, 3.0, 2.0], 0.05,
                          progress=lambda a, b: calls.append((a, b)))
    oracle.vsharp_with_radii(field, mask, (n, n, n), (1, 1, 1), [4.0, 3.0, 2.0], 0.05,
                             progress=lambda a, b: ocalls.append((a, b)))
    assert calls == ocalls == [(1, 3), (2, 3), (3, 3)]


def test_vsharp_defaults_phantom(qsm, oracle, phantom):
    # default radii 12..1 mm on a 64^3 phantom with harmonic background
    ph = phantom.make_phantom((64, 64, 64), n_ellipsoids=8, seed=1004)
    ph = phantom.add_background(ph, seed=1004)
    grid = qsm.Grid(ph.dims, ph.voxel_size)
    local, m = qsm.vsharp(ph.total_field, ph.mask, grid, qsm.VsharpParams.default())
    lo, mo = oracle.vsharp(ph.total_field, ph.mask, ph.dims, ph.voxel_size)
    assert np.array_equal(m, mo)
    e = oracle.nrmse(local, lo, mo)
    print("vsharp 64^3 nrmse", e, "maxabs", np.max(np.abs(local - lo)))
    assert e < 1e-4


def test_chain_vsharp_then_tv_device_resident(qsm, oracle, phantom):
    # run_bg_removal -> run_dipole_inversion on device: a V-SHARP batch
    # feeds its local field and eroded mask straight into a TV batch.  Checked
    # against the host entry points (bit for bit) and against the oracle's
    # vsharp -> tv_admm chain (masks identical, NRMSE <= 1e-3).
    dims, vs, b = (128, 128, 64), (0.9, 0.9, 0.9), (0.0, 0.0, 1.0)
    ph = phantom.add_background(phantom.make_phantom(dims, voxel_size=vs, n_ellipsoids=12,
                                                     seed=4100), seed=4100)
    radii = [4.5, 3.6, 2.7, 1.8]
    p = qsm.TvParams(lambda_=1e-3, rho=0.1, tol=0.0, max_iter=15)
    grid = qsm.Grid(dims, vs)
    n = ph.mask.size
    tf = qsm.DeviceArray.from_numpy(ph.total_field)
    tm = qsm.DeviceArray.from_numpy(ph.mask)
    tl = qsm.DeviceArray(n, np.float64)
    te = qsm.DeviceArray(n, np.uint8)
    tc = qsm.DeviceArray(n, np.float64)
    vsb = qsm.VsharpBatch([(grid, tf.ptr, tm.ptr, tl.ptr, te.ptr)], radii, 0.05)
    tvb = qsm.TvBatch([(grid, b, tl.ptr, te.ptr, tc.ptr)], p)
    vsb.run()
    tvb.run()
    local, em, chi = tl.numpy(), te.numpy(), tc.numpy()
    vsb.close()
    tvb.close()
    hl, hm = qsm.vsharp_with_radii(ph.total_field, ph.mask, grid, radii, 0.05)
    hx = qsm.tv_admm(hl, hm, grid, b, p)
    assert np.array_equal(em, hm) and np.array_equal(local, hl) and np.array_equal(chi, hx)
    lo, mo = oracle.vsharp_with_radii(ph.total_field, ph.mask, dims, vs, radii, 0.05)
    assert np.array_equal(em, mo)
    err = _rel(local, lo, mo)
    xo = oracle.tv_admm(lo, mo, dims, vs, b, 1e-3, 0.1, 0.0, 15)
    e = oracle.nrmse(chi, xo, mo)
    print("chain: eroded voxels", int(mo.sum()), "local max rel err", err, "chi nrmse", e)
    assert err < 1e-4
    assert e <= 1e-3
