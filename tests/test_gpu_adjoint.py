"""GPU vector-Jacobian product of the unrolled TV-ADMM loop and the dipole
operator, against the f64 oracle (whose adjoint is pinned by central
differences in tests/test_oracle_adjoint.py) and against f64 central
differences of the oracle's forward directly.

Loss L = ||chi_K - chi_true||^2 with chi_K = mask * x_K after K iterations
(tol = 0).  Bars (FP32 GPU vs f64): chi NRMSE <= 2e-5; field-bar relative L2
error <= 2e-3; lambda-bar relative error <= 5e-3 (both are sums over the
soft-threshold pattern, which FP32 and f64 may place differently at voxels
within rounding of the threshold)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _problem(seed, dims, scale=0.02):
    rng = np.random.default_rng(seed)
    n = int(np.prod(dims))
    f = rng.standard_normal(n) * scale
    m = (rng.random(n) < 0.85).astype(np.uint8)
    chi_true = rng.standard_normal(n) * 0.01
    return f, m, chi_true


@pytest.mark.parametrize("dims,b", [((16, 16, 16), (0.0, 0.0, 1.0)),
                                    ((20, 18, 22), (0.2, -0.1, 0.97)),
                                    ((64, 64, 64), (0.1, 0.2, 1.0))])
def test_dipole_forward_and_adjoint(qsm, oracle, dims, b):
    rng = np.random.default_rng(1)
    n = int(np.prod(dims))
    a1, a2 = rng.standard_normal(n), rng.standard_normal(n)
    grid = qsm.Grid(dims, (0.9, 1.0, 1.1))
    y = qsm.dipole_forward(a1, grid, b)
    yo = oracle.dipole_forward(a1, dims, (0.9, 1.0, 1.1), b)
    err = _rel_l2(y, yo)
    at = qsm.dipole_forward(a2, grid, b, adjoint=True)
    lhs, rhs = float(y @ a2), float(a1 @ at)
    print(dims, "forward rel err", err, "<Pa,b>", lhs, "<a,P*b>", rhs)
    assert err < 1e-5
    assert abs(lhs - rhs) <= 1e-5 * np.linalg.norm(y) * np.linalg.norm(a2)


@pytest.mark.parametrize("dims,b,K", [((16, 16, 16), (0.0, 0.0, 1.0), 1),
                                      ((16, 16, 16), (0.0, 0.0, 1.0), 5),
                                      ((20, 18, 22), (0.2, -0.1, 0.97), 8),
                                      ((64, 64, 64), (0.1, 0.0, 1.0), 6)])
def test_unrolled_vjp_matches_oracle(qsm, oracle, dims, b, K):
    vs, lam, rho = (1.0, 1.0, 1.0), 2e-3, 0.1
    f, m, chi_true = _problem(7, dims)
    grid = qsm.Grid(dims, vs)
    p = qsm.TvParams(lambda_=lam, rho=rho, tol=0.0, max_iter=K)
    chi0 = oracle.tv_admm(f, m, dims, vs, b, lam, rho, 0.0, K)
    cbar = 2.0 * (chi0 - chi_true)
    chi, fb, lb = qsm.tv_unrolled_vjp(f, m, grid, b, p, cbar)
    chio, fbo, lbo = oracle.tv_unrolled_vjp(f, m, dims, vs, b, lam, rho, K, cbar)
    e_chi = oracle.nrmse(chi, chio, m)
    e_fb = _rel_l2(fb, fbo)
    print(dims, K, "chi nrmse", e_chi, "field-bar rel", e_fb, "lambda-bar", lb, lbo)
    assert e_chi <= 2e-5
    assert e_fb <= 2e-3
    if K > 1:
        assert abs(lb - lbo) <= 5e-3 * abs(lbo)
    else:
        assert lb == 0.0


def test_unrolled_vjp_central_differences(qsm, oracle):
    # 16 random directions, f64 central differences of the oracle's forward
    dims, vs, b, K = (24, 20, 16), (1.0, 1.0, 1.0), (0.1, 0.1, 1.0), 10
    lam, rho = 2e-3, 0.1
    f, m, chi_true = _problem(21, dims)
    grid = qsm.Grid(dims, vs)
    p = qsm.TvParams(lambda_=lam, rho=rho, tol=0.0, max_iter=K)
    chi0 = oracle.tv_admm(f, m, dims, vs, b, lam, rho, 0.0, K)
    _, fb, lb = qsm.tv_unrolled_vjp(f, m, grid, b, p, 2.0 * (chi0 - chi_true))

    def loss(ff, ll=lam):
        return float(np.sum((oracle.tv_admm(ff, m, dims, vs, b, ll, rho, 0.0, K) - chi_true) ** 2))

    rng = np.random.default_rng(8)
    h = 2e-8
    errs = []
    for _ in range(16):
        v = rng.standard_normal(f.size)
        fd = (loss(f + h * v) - loss(f - h * v)) / (2 * h)
        errs.append(abs(fd - float(fb @ v)) / abs(fd))
    hl = 1e-6 * lam
    fdl = (loss(f, lam + hl) - loss(f, lam - hl)) / (2 * hl)
    print("directions: max rel err", max(errs), "lambda fd", fdl, "gpu", lb)
    assert max(errs) <= 2e-3
    assert abs(fdl - lb) <= 5e-3 * abs(fdl)


def test_unrolled_vjp_full_size_192x192x128(qsm, oracle, phantom):
    # The differentiable workload at full size: 192x192x128 at 1 mm, K = 20,
    # loss ||chi_K - chi_true||^2 against the phantom's susceptibility.
    ph = phantom.make_phantom((192, 192, 128), n_ellipsoids=12, seed=3003)
    K, lam, rho = 20, 1e-3, 0.1
    grid = qsm.Grid(ph.dims, ph.voxel_size)
    p = qsm.TvParams(lambda_=lam, rho=rho, tol=0.0, max_iter=K)
    chi_true = ph.chi * ph.mask
    chio0 = oracle.tv_admm(ph.local_field, ph.mask, ph.dims, ph.voxel_size, ph.bdir, lam, rho,
                           0.0, K)
    cbar = 2.0 * (chio0 - chi_true)
    chi, fb, lb = qsm.tv_unrolled_vjp(ph.local_field, ph.mask, grid, ph.bdir, p, cbar)
    chio, fbo, lbo = oracle.tv_unrolled_vjp(ph.local_field, ph.mask, ph.dims, ph.voxel_size,
                                            ph.bdir, lam, rho, K, cbar)
    e_chi = oracle.nrmse(chi, chio, ph.mask)
    e_fb = _rel_l2(fb, fbo)
    print("192x192x128 K=20: chi nrmse", e_chi, "field-bar rel", e_fb, "lambda-bar", lb, lbo)
    assert e_chi <= 1e-3
    assert e_fb <= 2e-3
    assert abs(lb - lbo) <= 5e-3 * abs(lbo)


def test_unrolled_batch_equals_single(qsm):
    shapes = [((16, 16, 16), (1.0, 1.0, 1.0), (0, 0, 1)), ((20, 18, 22), (1.0, 1.0, 1.0), (0.1, 0.2, 0.95))]
    p = qsm.TvParams(lambda_=2e-3, rho=0.1, tol=0.0, max_iter=6)
    keep, vols, singles, cbars, fbars, chis = [], [], [], [], [], []
    for s, (dims, vs, b) in enumerate(shapes):
        f, m, ct = _problem(30 + s, dims)
        cb = np.random.default_rng(40 + s).standard_normal(f.size)
        singles.append(qsm.tv_unrolled_vjp(f, m, qsm.Grid(dims, vs), b, p, cb))
        tf, tm = qsm.DeviceArray.from_numpy(f), qsm.DeviceArray.from_numpy(m)
        tc, tcb = qsm.DeviceArray(f.size, np.float64), qsm.DeviceArray.from_numpy(cb)
        tfb = qsm.DeviceArray(f.size, np.float64)
        keep.append((tf, tm, tc, tcb, tfb))
        vols.append((qsm.Grid(dims, vs), b, tf.ptr, tm.ptr, tc.ptr))
        cbars.append(tcb.ptr)
        fbars.append(tfb.ptr)
        chis.append(tc)
    u = qsm.TvUnrolled(vols, p)
    u.forward()
    lbs = u.backward(cbars, fbars)
    u.close()
    for s in range(len(shapes)):
        assert np.array_equal(chis[s].numpy(), singles[s][0])
        assert np.array_equal(keep[s][4].numpy(), singles[s][1])
        assert lbs[s] == singles[s][2]
