"""Pin the oracle's kernels and element-wise helpers with the reference's own
unit tests (restated) plus closed forms and brute force:
dipole (src/kernels/dipole.rs:145-185), Laplacian (src/kernels/laplacian.rs:
78-98), SMV (src/kernels/smv.rs:175-192), gradient/divergence
(src/utils/gradient.rs:577-621, 895-951, 1226-1300), shrink
(src/inversion/tv.rs:105-110), apply_mask_zero (src/utils/mask.rs:168-174)."""
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_dipole_dc_range_symmetry(oracle):
    d = oracle.dipole_kernel((4, 4, 4), (1, 1, 1), (0, 0, 1))
    assert abs(d[0]) < 1e-10
    d = oracle.dipole_kernel((8, 8, 8), (1, 1, 1), (0, 0, 1))
    assert np.all(d[1:] >= -2 / 3 - 1e-10) and np.all(d[1:] <= 1 / 3 + 1e-10)
    n = 8
    for i in range(1, n // 2):
        for j in range(1, n // 2):
            for k in range(1, n // 2):
                assert abs(d[i * n * n + j * n + k] - d[(n - i) * n * n + (n - j) * n + (n - k)]) < 1e-10


def test_dipole_golden_tilted(oracle):
    k = np.load(os.path.join(GOLDEN, "kernels.npz"))
    d = oracle.dipole_kernel(tuple(k["dipole_dims"]), tuple(k["dipole_vs"]), tuple(k["dipole_b"]))
    assert np.max(np.abs(d - k["dipole"])) < 1e-15


def test_dipole_rotation_closed_form(oracle):
    # b along x: D(k) = 1/3 - kx^2/|k|^2, the z-case with axes permuted
    n = 6
    dx = oracle.dipole_kernel((n, n, n), (1, 1, 1), (1, 0, 0)).reshape((n,) * 3, order="F")
    dz = oracle.dipole_kernel((n, n, n), (1, 1, 1), (0, 0, 1)).reshape((n,) * 3, order="F")
    assert np.allclose(dx, np.transpose(dz, (2, 1, 0)), atol=1e-15)


def test_laplacian_reference_tests(oracle):
    l = oracle.laplacian_kernel((8, 8, 8), (1, 1, 1), False)
    assert abs(l.sum()) < 1e-10
    assert abs(l[0] + 6.0) < 1e-10
    ln = oracle.laplacian_kernel((8, 8, 8), (1, 1, 1), True)
    assert np.allclose(l + ln, 0.0, atol=1e-10)


def test_laplacian_spectrum_closed_form(oracle):
    dims, vs = (10, 7, 6), (0.9, 1.1, 1.3)
    l = oracle.laplacian_kernel(dims, vs, True)
    spec = oracle.fft3d(l, dims).real
    k = np.load(os.path.join(GOLDEN, "kernels.npz"))
    assert np.max(np.abs(spec - k["lap_spec"])) < 1e-12
    g = np.meshgrid(*[np.arange(n) for n in dims], indexing="ij")
    closed = sum((2 - 2 * np.cos(2 * np.pi * g[a] / dims[a])) / vs[a] ** 2 for a in range(3))
    assert np.max(np.abs(spec - closed.ravel(order="F"))) < 1e-12


def test_laplacian_degenerate_axes(oracle):
    # src/kernels/laplacian.rs:20-62 writes the centre for every axis but the
    # neighbours only where n > 1, by assignment: on a length-2 axis both
    # neighbours are voxel 1, written once; on a length-1 axis there is none.
    # So the stencil no longer sums to zero, and its DFT per axis is 2h
    # (n = 1) and 2h - h cos(pi k) (n = 2), h = 1/vs^2, against the periodic
    # 2h - 2h cos(2 pi k / n) of longer axes.
    dims, vs = (1, 2, 5), (0.9, 1.1, 1.3)
    h = [1.0 / v ** 2 for v in vs]
    l = oracle.laplacian_kernel(dims, vs, True).reshape(dims, order="F")
    assert abs(l[0, 0, 0] - 2 * sum(h)) < 1e-12
    assert abs(l[0, 1, 0] + h[1]) < 1e-12
    assert abs(l[0, 0, 1] + h[2]) < 1e-12 and abs(l[0, 0, 4] + h[2]) < 1e-12
    assert np.count_nonzero(l) == 4
    assert abs(l.sum() - (2 * h[0] + h[1])) < 1e-12
    spec = oracle.fft3d(l.ravel(order="F"), dims).real
    _, j, k = np.meshgrid(*[np.arange(n) for n in dims], indexing="ij")
    want = 2 * h[0] + (2 - np.cos(np.pi * j)) * h[1] + (2 - 2 * np.cos(2 * np.pi * k / 5)) * h[2]
    assert np.max(np.abs(spec - want.ravel(order="F"))) < 1e-12


def test_smv_reference_tests_and_counts(oracle):
    s, _ = oracle.smv_kernel((16, 16, 16), (1, 1, 1), 3.0)
    assert abs(s.sum() - 1.0) < 1e-10
    s, _ = oracle.smv_kernel((16, 16, 16), (1, 1, 1), 2.0)
    assert s[0] > 0
    s, cnt = oracle.smv_kernel((8, 8, 8), (1, 1, 1), 0.5)
    assert np.count_nonzero(s) == 1 and cnt == 1
    k = np.load(os.path.join(GOLDEN, "kernels.npz"))
    for row in k["smv_counts"]:
        dims = tuple(int(v) for v in row[:3])
        _, c = oracle.smv_kernel(dims, tuple(row[3:6]), float(row[6]))
        assert c == row[7]


def test_erosion_brute_force(oracle):
    # erode_mask_smv equals counting mask voxels inside the wrapped sphere
    rng = np.random.default_rng(1)
    dims, vs, r = (12, 10, 8), (1.0, 1.2, 0.9), 2.3
    mask = (rng.random(np.prod(dims)) < 0.9).astype(np.uint8)
    s, cnt = oracle.smv_kernel(dims, vs, r)
    sf = oracle.fft3d(s, dims).real
    delta = 1.0 - np.sqrt(1e-7)
    er = oracle.erode_mask_smv(mask, sf, dims, delta)
    m3 = mask.reshape(dims, order="F")
    offs = np.argwhere(s.reshape(dims, order="F") > 0)
    conv = np.zeros(dims)
    for o in offs:
        conv += np.roll(m3, shift=tuple(int(v) for v in o), axis=(0, 1, 2))
    expect = ((conv / cnt) > delta).astype(np.uint8).ravel(order="F")
    assert np.array_equal(er, expect)


def test_gradient_adjoint_anisotropic(oracle):
    # src/utils/gradient.rs:895-951
    dims, vs = (6, 7, 5), (1.5, 0.8, 2.0)
    g = np.meshgrid(*[np.arange(n, dtype=float) for n in dims], indexing="ij")
    x = (np.sin(g[0] * 0.3) + np.cos(g[1] * 0.7) + g[2] * 0.2).ravel(order="F")
    hx = np.sqrt((g[0] + 1) * (g[1] + 1)).ravel(order="F")
    hy = np.sin(g[2] * 0.5 + 0.1).ravel(order="F")
    hz = np.cos(g[0] * 0.4 - g[2] * 0.2).ravel(order="F")
    gx, gy, gz = oracle.fgrad(x, dims, vs)
    div = oracle.bdiv(hx, hy, hz, dims, vs)
    lhs = gx @ hx + gy @ hy + gz @ hz
    rhs = -(x @ div)
    assert abs(lhs - rhs) / max(abs(lhs), abs(rhs), 1e-15) < 1e-10


def test_gradient_constant_ramp_and_wrap(oracle):
    dims = (8, 8, 8)
    gx, gy, gz = oracle.fgrad(np.full(512, 3.0), dims, (1, 1, 1))
    assert np.all(gx == 0) and np.all(gy == 0) and np.all(gz == 0)
    ramp = np.tile(np.arange(8, dtype=float), 64)  # x[i,j,k] = i
    gx, _, _ = oracle.fgrad(ramp, dims, (1, 1, 1))
    gx3 = gx.reshape(dims, order="F")
    assert np.allclose(gx3[:-1], 1.0) and np.allclose(gx3[-1], -7.0)  # periodic wrap
    div = oracle.bdiv(np.ones(512), np.ones(512), np.ones(512), dims, (1, 1, 1))
    assert np.allclose(div, 0.0, atol=1e-10)


def test_shrink_relchange_mask(oracle):
    assert abs(oracle.shrink(1.0, 0.5) - 0.5) < 1e-10
    assert abs(oracle.shrink(-1.0, 0.5) + 0.5) < 1e-10
    assert oracle.shrink(0.3, 0.5) == 0.0 and oracle.shrink(-0.3, 0.5) == 0.0
    x, xp = np.array([3.0, 4.0]), np.array([0.0, 0.0])
    assert abs(oracle.relative_change(x, xp) - 1.0) < 1e-15
    d = oracle.apply_mask_zero([1.0, 2.0, 3.0, 4.0], [1, 0, 1, 0])
    assert list(d) == [1.0, 0.0, 3.0, 0.0]
