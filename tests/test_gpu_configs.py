"""GPU parity at the configured workload shapes, against the f64 oracle.

  * TV-ADMM at the benchmark's iteration count (200, tol 0) on 128x128x64
    and on the benchmark volume itself (256x256x160 at 0.9x0.9x1.0 mm);
  * the chained V-SHARP -> TV workload at its own shape (384x384x256 at
    0.6 mm, V-SHARP radii 12..1 mm from VsharpParams::default()), and the
    8-volume device chain against per-volume host calls;
  * the 32-volume uneven batch (axes from {144..256}, voxel sizes
    U(0.6, 1.2) mm, tilted B0) against single-volume solves and the oracle.

The bar is the metric's: range-normalised NRMSE <= 1e-3 against f64
(tests/common/mod.rs:34-51), asserted on the raw difference; the
DC-removed figure (the reference's anisotropic-voxel DC term, see
conftest.nrmse_dc) and the max-abs error are asserted or printed beside it.
"""
import numpy as np
import pytest

from conftest import nrmse_dc

pytestmark = pytest.mark.gpu


def _progress(tag, every):
    """Progress callback for a long oracle run: prints every `every` steps."""
    def cb(cur, total):
        if cur % every == 0 or cur == total:
            print(f"{tag}: oracle {cur}/{total}", flush=True)
    return cb


def _report(tag, x, xo, mask, oracle):
    m = np.asarray(mask).ravel() > 0
    raw = float(oracle.nrmse(x, xo, mask))
    dc = nrmse_dc(x, xo, mask)
    maxabs = float(np.max(np.abs(np.asarray(x).ravel()[m] - np.asarray(xo).ravel()[m])))
    rng = float(np.ptp(np.asarray(xo).ravel()[m]))
    print(f"{tag}: nrmse raw {raw:.3e}, dc-removed {dc:.3e}, max-abs {maxabs:.3e} "
          f"(oracle range {rng:.3e})", flush=True)
    return raw, dc, maxabs


@pytest.mark.timeout(300)
def test_tv_200_iterations_128x128x64(qsm, oracle, phantom):
    # 200 iterations, tol 0, lambda 1e-3, rho 0.1 at 0.9x0.9x1.0 mm: FP32
    # state through the benchmark's full iteration count
    ph = phantom.make_phantom((128, 128, 64), voxel_size=(0.9, 0.9, 1.0), n_ellipsoids=12,
                              seed=1000)
    grid = qsm.Grid(ph.dims, ph.voxel_size)
    p = qsm.TvParams(lambda_=1e-3, rho=0.1, tol=0.0, max_iter=200)
    x, it, _ = qsm.tv_admm(ph.local_field, ph.mask, grid, ph.bdir, p, return_info=True)
    xo, oit, _ = oracle.tv_admm(ph.local_field, ph.mask, ph.dims, ph.voxel_size, ph.bdir, 1e-3, 0.1,
                                0.0, 200, return_info=True)
    assert it == oit == 200
    raw, dc, _ = _report("128x128x64 x200", x, xo, ph.mask, oracle)
    assert raw <= 1e-3
    assert dc <= 1e-4
    assert np.all(x[ph.mask == 0] == 0.0)


@pytest.mark.timeout(900)
def test_tv_200_iterations_256x256x160(qsm, oracle, phantom):
    # The benchmark volume (seed 1001, 40 ellipsoids) and settings: the solve
    # the headline times, against one full f64 solve (about a minute on 16
    # host threads)
    ph = phantom.make_phantom((256, 256, 160), voxel_size=(0.9, 0.9, 1.0), n_ellipsoids=40,
                              seed=1001)
    grid = qsm.Grid(ph.dims, ph.voxel_size)
    p = qsm.TvParams(lambda_=1e-3, rho=0.1, tol=0.0, max_iter=200)
    x, it, _ = qsm.tv_admm(ph.local_field, ph.mask, grid, ph.bdir, p, return_info=True)
    print("gpu solve done; oracle running", flush=True)
    xo, oit, _ = oracle.tv_admm(ph.local_field, ph.mask, ph.dims, ph.voxel_size, ph.bdir, 1e-3, 0.1,
                                0.0, 200, progress=_progress("256x256x160", 25), return_info=True)
    assert it == oit == 200
    raw, dc, _ = _report("256x256x160 x200", x, xo, ph.mask, oracle)
    assert raw <= 1e-3
    assert dc <= 1e-4
    assert np.all(x[ph.mask == 0] == 0.0)


# ---------------------------------------------------------------- chain --
CHAIN_DIMS, CHAIN_VS = (384, 384, 256), (0.6, 0.6, 0.6)


@pytest.fixture(scope="module")
def chain_phantom(phantom):
    # the chained workload's first volume: 12-ellipsoid phantom plus the
    # harmonic field of 3 external dipoles plus noise
    seed = 4004 * 1000
    return phantom.add_background(
        phantom.make_phantom(CHAIN_DIMS, voxel_size=CHAIN_VS, n_ellipsoids=12, seed=seed), seed=seed)


@pytest.mark.timeout(900)
def test_vsharp_then_tv_384x384x256(qsm, oracle, chain_phantom):
    # V-SHARP with VsharpParams::default() (radii 12, 11, ..., 1 mm at
    # 0.6 mm voxels: the largest sphere holds ~33.5 k voxels, where the
    # erosion threshold 1 - sqrt(1e-7) is tightest, src/bgremove/vsharp.rs:
    # 129) through the device batch; then 20 TV-ADMM iterations from the
    # oracle's V-SHARP output.
    ph = chain_phantom
    n = ph.mask.size
    grid = qsm.Grid(CHAIN_DIMS, CHAIN_VS)
    vp = qsm.VsharpParams()
    radii = qsm.vsharp_radii(vp)
    assert radii == [float(r) for r in range(12, 0, -1)]
    tf, tm = qsm.DeviceArray.from_numpy(ph.total_field), qsm.DeviceArray.from_numpy(ph.mask)
    tl, te = qsm.DeviceArray(n, np.float64), qsm.DeviceArray(n, np.uint8)
    vsb = qsm.VsharpBatch([(grid, tf.ptr, tm.ptr, tl.ptr, te.ptr)], radii, vp.threshold)
    vsb.run()
    vsb.close()
    local, emask = tl.numpy(), te.numpy()
    print("gpu V-SHARP done; oracle running", flush=True)
    lo, emo = oracle.vsharp(ph.total_field, ph.mask, CHAIN_DIMS, CHAIN_VS, vp.threshold,
                            vp.max_radius, vp.min_radius, progress=_progress("V-SHARP radius", 1))
    assert np.array_equal(emask, emo), int(np.sum(emask != emo))
    assert int(emo.sum()) > 0
    raw, _, _ = _report("V-SHARP 384x384x256", local, lo, emo, oracle)
    assert raw <= 1e-4
    assert np.all(local[emo == 0] == 0.0)

    p = qsm.TvParams(lambda_=1e-3, rho=0.1, tol=0.0, max_iter=20)
    x = qsm.tv_admm(lo, emo, grid, ph.bdir, p)
    print("gpu TV done; oracle running", flush=True)
    xo = oracle.tv_admm(lo, emo, CHAIN_DIMS, CHAIN_VS, ph.bdir, 1e-3, 0.1, 0.0, 20,
                        progress=_progress("TV 384x384x256", 5))
    raw, dc, _ = _report("TV x20 on V-SHARP output 384x384x256", x, xo, emo, oracle)
    assert raw <= 1e-3
    assert dc <= 1e-4


@pytest.mark.timeout(600)
def test_chain_batch_equals_host_calls_384x384x256(qsm, chain_phantom):
    # 8 volumes (circular shifts of the chained workload's first volume, so
    # each is a different input) through the device chain VsharpBatch ->
    # TvBatch (200 iterations): each equals the host-pointer vsharp and
    # tv_admm calls on that volume alone, bit for bit.
    ph = chain_phantom
    grid = qsm.Grid(CHAIN_DIMS, CHAIN_VS)
    n = ph.mask.size
    vp = qsm.VsharpParams()
    radii = qsm.vsharp_radii(vp)
    p = qsm.TvParams(lambda_=1e-3, rho=0.1, tol=0.0, max_iter=200)
    f3 = ph.total_field.reshape(CHAIN_DIMS, order="F")
    m3 = ph.mask.reshape(CHAIN_DIMS, order="F")
    shifts = [(0, 0, 0), (7, 0, 0), (0, -9, 0), (0, 0, 5), (3, 4, 0), (-5, 0, 6), (0, 11, -3), (2, 2, 2)]
    fields, masks, keep, vvols, tvols, outs = [], [], [], [], [], []
    for s in shifts:
        f = np.ascontiguousarray(np.roll(f3, s, axis=(0, 1, 2)).ravel(order="F"))
        m = np.ascontiguousarray(np.roll(m3, s, axis=(0, 1, 2)).ravel(order="F"))
        fields.append(f)
        masks.append(m)
        tf, tm = qsm.DeviceArray.from_numpy(f), qsm.DeviceArray.from_numpy(m)
        tl, te = qsm.DeviceArray(n, np.float64), qsm.DeviceArray(n, np.uint8)
        tc = qsm.DeviceArray(n, np.float64)
        keep.append((tf, tm))
        outs.append((tl, te, tc))
        vvols.append((grid, tf.ptr, tm.ptr, tl.ptr, te.ptr))
        tvols.append((grid, ph.bdir, tl.ptr, te.ptr, tc.ptr))
    vsb = qsm.VsharpBatch(vvols, radii, vp.threshold)
    tvb = qsm.TvBatch(tvols, p)
    vsb.run()
    tvb.run()
    vsb.close()
    tvb.close()
    for v in range(len(shifts)):
        tl, te, tc = outs[v]
        lo, eo = qsm.vsharp(fields[v], masks[v], grid, vp)
        assert np.array_equal(te.numpy(), eo), v
        assert np.array_equal(tl.numpy(), lo), v
        xo = qsm.tv_admm(lo, eo, grid, ph.bdir, p)
        assert np.array_equal(tc.numpy(), xo), v
        print(f"volume {v} shift {shifts[v]}: chain == host calls", flush=True)


# ---------------------------------------------------------- uneven batch --
@pytest.mark.timeout(900)
def test_uneven_32_volume_batch(qsm, oracle, phantom):
    # The uneven workload: 32 volumes, each axis from {144, 160, ..., 256},
    # voxel sizes U(0.6, 1.2) mm, B0 tilted U(0, 20) deg (seed 3002), one
    # batch with one launch group per axis length.  Every volume equals its
    # single-volume solve bit for bit (200 iterations); three volumes - the
    # first two and the first whose y length runs the staged column FFT - are
    # checked against the oracle at 20 iterations.
    seed = 3002
    shapes = phantom.uneven_batch_shapes(seed, 32)
    assert len({n for d, _ in shapes for n in d}) >= 6
    p = qsm.TvParams(lambda_=1e-3, rho=0.1, tol=0.0, max_iter=200)
    keep, vols, outs, phs = [], [], [], []
    for v, (dims, vs) in enumerate(shapes):
        b = phantom.tilted_bdir(seed * 1000 + v)
        ph = phantom.make_phantom(dims, voxel_size=vs, bdir=b, n_ellipsoids=12, seed=seed * 1000 + v)
        phs.append(ph)
        tf, tm = qsm.DeviceArray.from_numpy(ph.local_field), qsm.DeviceArray.from_numpy(ph.mask)
        tc = qsm.DeviceArray(ph.local_field.size, np.float64)
        keep.append((tf, tm))
        outs.append(tc)
        vols.append((qsm.Grid(dims, vs), b, tf.ptr, tm.ptr, tc.ptr))
    print("32 phantoms ready", flush=True)
    batch = qsm.TvBatch(vols, p)
    batch.run()
    for v in range(32):
        assert batch.status(v)[0] == 200
    batch.close()
    for v, ph in enumerate(phs):
        single = qsm.tv_admm(ph.local_field, ph.mask, qsm.Grid(ph.dims, ph.voxel_size), ph.bdir, p)
        assert np.array_equal(outs[v].numpy(), single), (v, ph.dims)
    print("batch == single for all 32 volumes", flush=True)
    staged = next(v for v, ph in enumerate(phs) if ph.dims[1] in (144, 176))
    assert qsm.fft_schedule("y", phs[staged].dims[1]) == "staged"
    p20 = qsm.TvParams(lambda_=1e-3, rho=0.1, tol=0.0, max_iter=20)
    for v in (0, 1, staged):
        ph = phs[v]
        x = qsm.tv_admm(ph.local_field, ph.mask, qsm.Grid(ph.dims, ph.voxel_size), ph.bdir, p20)
        xo = oracle.tv_admm(ph.local_field, ph.mask, ph.dims, ph.voxel_size, ph.bdir, 1e-3, 0.1, 0.0, 20)
        raw, dc, _ = _report(f"uneven volume {v} {ph.dims} {tuple(round(s, 3) for s in ph.voxel_size)}",
                             x, xo, ph.mask, oracle)
        assert raw <= 1e-3
        assert dc <= 1e-4
