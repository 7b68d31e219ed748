"""Pin the oracle's tv_admm (src/inversion/tv.rs:63-97) and vsharp
(src/bgremove/vsharp.rs:55-199) with the reference's unit tests, closed
forms, invariants and an independent numpy restatement (tests/golden/)."""
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


# ------------------------------------------------------------ TV-ADMM ----
def test_tv_zero_field(oracle):
    # src/inversion/tv.rs:113-126
    n = 8
    x = oracle.tv_admm(np.zeros(n ** 3), np.ones(n ** 3, np.uint8), (n, n, n), (1, 1, 1),
                       (0, 0, 1), 1e-3, 0.1, 1e-2, 10)
    assert np.max(np.abs(x)) < 1e-8


def test_tv_numpy_trajectory(oracle):
    t = np.load(os.path.join(GOLDEN, "tv_traj.npz"))
    dims = tuple(int(v) for v in t["dims"])
    for it in (1, 2, 10, 50):
        x = oracle.tv_admm(t["f"], t["mask"], dims, tuple(t["vs"]), tuple(t["b"]), float(t["lam"]),
                           float(t["rho"]), 0.0, it)
        ref = t[f"x{it}"]
        assert np.max(np.abs(x - ref)) < 1e-12 * np.max(np.abs(ref)), it


@pytest.mark.parametrize("vs", [(1.0, 1.0, 1.0), (0.9, 1.1, 1.3)])
def test_tv_one_iteration_is_tikhonov_gradient(oracle, vs):
    # max_iter = 1: x = Re IFFT(D F^/(D^2 + rho L^)) (src/inversion/admm.rs:165-176)
    # == tikhonov(Regularization::Gradient, lambda = rho) (src/inversion/tikhonov.rs:77-130)
    rng = np.random.default_rng(2)
    dims = (12, 10, 14)
    f = rng.standard_normal(np.prod(dims)) * 0.01
    m = (rng.random(np.prod(dims)) < 0.8).astype(np.uint8)
    x = oracle.tv_admm(f, m, dims, vs, (0.2, 0.1, 1.0), 1e-3, 0.1, 0.0, 1)
    t = oracle.tikhonov_gradient(f, m, dims, vs, (0.2, 0.1, 1.0), 0.1)
    assert np.max(np.abs(x - t)) < 1e-14


def test_tv_scale_covariance(oracle):
    # tv(a f, a lambda, rho) = a tv(f, lambda, rho): every step is 1-homogeneous
    rng = np.random.default_rng(4)
    dims = (12, 12, 10)
    f = rng.standard_normal(np.prod(dims)) * 0.01
    m = np.ones(np.prod(dims), np.uint8)
    a = 4.0  # a power of two keeps every operation exactly scaled
    x1 = oracle.tv_admm(f, m, dims, (1, 1, 1), (0, 0, 1), 1e-3, 0.1, 0.0, 15)
    xa = oracle.tv_admm(a * f, m, dims, (1, 1, 1), (0, 0, 1), a * 1e-3, 0.1, 0.0, 15)
    assert np.max(np.abs(xa - a * x1)) < 1e-13 * np.max(np.abs(xa))


def test_tv_shift_equivariance(oracle):
    # all operators are periodic: shifting f circularly shifts chi
    rng = np.random.default_rng(5)
    dims = (10, 8, 12)
    f3 = rng.standard_normal(dims) * 0.01
    m = np.ones(np.prod(dims), np.uint8)
    x = oracle.tv_admm(f3.ravel(order="F"), m, dims, (1, 1, 1), (0.1, 0, 1), 1e-3, 0.1, 0.0, 10)
    fs = np.roll(f3, (3, -2, 5), axis=(0, 1, 2))
    xs = oracle.tv_admm(fs.ravel(order="F"), m, dims, (1, 1, 1), (0.1, 0, 1), 1e-3, 0.1, 0.0, 10)
    x3 = np.roll(x.reshape(dims, order="F"), (3, -2, 5), axis=(0, 1, 2))
    assert np.max(np.abs(xs.reshape(dims, order="F") - x3)) < 1e-12


def test_tv_thread_invariance_and_progress(oracle):
    # src/inversion/tv.rs:171-197 (parallel == 1 thread) and the progress contract
    n = 16
    f = np.sin(np.arange(n ** 3) * 0.7) * 0.01
    m = np.ones(n ** 3, np.uint8)
    nt = oracle.num_threads()
    calls = []
    try:
        oracle.set_num_threads(1)
        a, it, _ = oracle.tv_admm(f, m, (n,) * 3, (1, 1, 1), (0, 0, 1), 1e-3, 0.1, 1e-3, 50,
                                  progress=lambda c, t: calls.append((c, t)), return_info=True)
        oracle.set_num_threads(max(2, nt))
        b = oracle.tv_admm(f, m, (n,) * 3, (1, 1, 1), (0, 0, 1), 1e-3, 0.1, 1e-3, 50)
    finally:
        oracle.set_num_threads(nt)
    assert np.array_equal(a, b)
    assert calls == [(i, 50) for i in range(1, it + 1)] + ([(it, it)] if it < 50 else [])


# ------------------------------------------------------------- V-SHARP ---
def _sphere_mask(n, r):
    g = np.arange(n) - n // 2
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    return ((X * X + Y * Y + Z * Z) <= r * r).astype(np.uint8).ravel(order="F")


def test_vsharp_numpy_fixture(oracle):
    v = np.load(os.path.join(GOLDEN, "vsharp16.npz"))
    l, m = oracle.vsharp_with_radii(v["field"], v["mask"], (16, 16, 16), (1, 1, 1), v["radii"],
                                    float(v["threshold"]))
    assert np.array_equal(m, v["eroded"])
    assert np.max(np.abs(l - v["local"])) < 1e-12


def test_vsharp_reference_unit_tests(oracle):
    n = 8
    dims = (n, n, n)
    ones = np.ones(n ** 3, np.uint8)
    # zero field -> 0 (src/bgremove/vsharp.rs:206-219)
    l, _ = oracle.vsharp_with_radii(np.zeros(n ** 3), ones, dims, (1, 1, 1), [4.0, 3.0, 2.0], 0.05)
    assert np.max(np.abs(l)) < 1e-10
    ramp = np.arange(n ** 3) * 0.001
    # single radius == sharp_core (vsharp.rs:268-299)
    l1, m1 = oracle.vsharp_with_radii(ramp, ones, dims, (1, 1, 1), [3.0], 0.05)
    l2, m2 = oracle.sharp_core(ramp, ones, dims, (1, 1, 1), 0.05, 3.0)
    assert np.max(np.abs(l1 - l2)) < 1e-10 and np.array_equal(m1, m2)
    # empty radii -> zeros + original mask (vsharp.rs:302-318)
    l, m = oracle.vsharp_with_radii(ramp, ones, dims, (1, 1, 1), [], 0.05)
    assert np.all(l == 0) and np.array_equal(m, ones)
    # unsorted == sorted (vsharp.rs:431-459)
    a = oracle.vsharp_with_radii(ramp, ones, dims, (1, 1, 1), [4.0, 3.0, 2.0], 0.05)
    b = oracle.vsharp_with_radii(ramp, ones, dims, (1, 1, 1), [2.0, 4.0, 3.0], 0.05)
    assert np.array_equal(a[1], b[1]) and np.max(np.abs(a[0] - b[0])) < 1e-10
    # V-SHARP keeps at least as many voxels as SHARP (vsharp.rs:222-244)
    n2 = 16
    _, mv = oracle.vsharp_with_radii(np.zeros(n2 ** 3), np.ones(n2 ** 3, np.uint8), (n2,) * 3,
                                     (1, 1, 1), [5.0, 4.0, 3.0, 2.0], 0.05)
    _, ms = oracle.sharp_core(np.zeros(n2 ** 3), np.ones(n2 ** 3, np.uint8), (n2,) * 3, (1, 1, 1),
                              0.05, 5.0)
    assert mv.sum() >= ms.sum()


def test_vsharp_all_ones_mask_closed_form(oracle):
    # with an all-ones mask erosion keeps everything at the largest radius,
    # and the deconvolution cancels (1 - S): out = Re IFFT(F^ 1{|1-S^max| >= thr})
    rng = np.random.default_rng(6)
    dims, vs = (12, 10, 14), (1.0, 0.9, 1.1)
    f = rng.standard_normal(np.prod(dims))
    ones = np.ones(np.prod(dims), np.uint8)
    l, m = oracle.vsharp_with_radii(f, ones, dims, vs, [3.0, 2.0], 0.05)
    assert np.all(m == 1)
    s, _ = oracle.smv_kernel(dims, vs, 3.0)
    S = oracle.fft3d(s, dims).real
    keep = (np.abs(1.0 - S) >= 0.05).astype(float)
    expect = oracle.fft3d(oracle.fft3d(f, dims) * keep, dims, inverse=True).real
    assert np.max(np.abs(l - expect)) < 1e-10


def test_vsharp_constant_field_is_removed(oracle):
    n = 16
    dims = (n, n, n)
    mask = _sphere_mask(n, 6)
    l, m = oracle.vsharp_with_radii(np.full(n ** 3, 0.7), mask, dims, (1, 1, 1), [3.0, 2.0], 0.05)
    assert np.max(np.abs(l)) < 1e-12


def test_vsharp_radii_list(oracle):
    # src/bgremove/vsharp.rs:70-82 (defaults 12 .. 1 mm)
    assert list(oracle.vsharp_radii(12.0, 1.0)) == [float(r) for r in range(12, 0, -1)]
    assert list(oracle.vsharp_radii(0.5, 1.0)) == [0.5]
    assert list(oracle.vsharp_radii(3.0, 1.5)) == [3.0, 1.5]
