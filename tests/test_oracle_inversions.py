"""Pins the oracle's closed-form inversions (tkd / tsvd / tikhonov,
src/inversion/tkd.rs, src/inversion/tikhonov.rs) and RTS
(src/inversion/rts.rs:71-212) against what the mathematics fixes: on a
single Fourier mode every closed form is a per-frequency multiplier, and RTS's
first stage is the scalar refinement recurrence applied per frequency."""
import numpy as np
import pytest

DIMS = (8, 6, 10)


def _mode(dims, a, b, c):
    i, j, k = [np.arange(n) for n in dims]
    ph = 2 * np.pi * (a * i[:, None, None] / dims[0] + b * j[None, :, None] / dims[1]
                      + c * k[None, None, :] / dims[2])
    return np.cos(ph).ravel(order="F")


def _dk(dims, vs, bdir, a, b, c):
    f = np.array([a / (dims[0] * vs[0]), b / (dims[1] * vs[1]), c / (dims[2] * vs[2])])
    bb = np.asarray(bdir, float) / np.linalg.norm(bdir)
    D = 1.0 / 3.0 - (f @ bb) ** 2 / (f @ f)
    L = sum((2 - 2 * np.cos(2 * np.pi * q / n)) / v ** 2 for q, n, v in zip((a, b, c), dims, vs))
    return D, L


@pytest.mark.parametrize("mode", [(1, 1, 1), (2, 0, 3), (0, 1, 1), (3, 2, 4)])
@pytest.mark.parametrize("bdir", [(0.0, 0.0, 1.0), (0.3, -0.2, 1.0)])
def test_closed_forms_on_one_mode(oracle, mode, bdir):
    vs = (1.0, 0.9, 1.2)
    f = _mode(DIMS, *mode)
    m = np.ones(f.size, np.uint8)
    D, L = _dk(DIMS, vs, bdir, *mode)
    thr = 0.15
    tkd_m = (1 / thr if D >= 0 else -1 / thr) if abs(D) <= thr else 1 / D
    tsvd_m = 0.0 if abs(D) <= thr else 1 / D
    cases = [(oracle.tkd(f, m, DIMS, vs, bdir, thr), tkd_m),
             (oracle.tkd(f, m, DIMS, vs, bdir, thr, tsvd=True), tsvd_m),
             (oracle.tikhonov(f, m, DIMS, vs, bdir, 0.01, 0), D / (D * D + 0.01)),
             (oracle.tikhonov(f, m, DIMS, vs, bdir, 0.01, 1), D / (D * D + 0.01 * L)),
             (oracle.tikhonov(f, m, DIMS, vs, bdir, 0.01, 2), D / (D * D + 0.01 * L * L))]
    for got, mult in cases:
        assert np.max(np.abs(got - mult * f)) <= 1e-12 * max(1.0, abs(mult))


def test_closed_forms_reference_unit_tests(oracle):
    # src/inversion/tkd.rs:160-231, src/inversion/tikhonov.rs:145-199
    n = 8
    d3 = (n, n, n)
    ones = np.ones(n ** 3, np.uint8)
    assert np.max(np.abs(oracle.tkd(np.zeros(n ** 3), ones, d3, (1, 1, 1), (0, 0, 1), 0.15))) < 1e-10
    m = ones.copy()
    m[0] = m[1] = 0
    chi = oracle.tkd(np.ones(n ** 3), m, d3, (1, 1, 1), (0, 0, 1), 0.15)
    assert chi[0] == 0.0 and chi[1] == 0.0
    f16 = np.sin(np.arange(16 ** 3) * 0.1)
    a = oracle.tkd(f16, np.ones(16 ** 3, np.uint8), (16,) * 3, (1, 1, 1), (0, 0, 1), 0.15)
    b = oracle.tkd(f16, np.ones(16 ** 3, np.uint8), (16,) * 3, (1, 1, 1), (0, 0, 1), 0.15, tsvd=True)
    assert np.sum(np.abs(a - b)) > 1e-10
    f8 = np.sin(np.arange(n ** 3) * 0.1)
    r = [oracle.tikhonov(f8, ones, d3, (1, 1, 1), (0, 0, 1), 0.01, g) for g in range(3)]
    assert np.sum(np.abs(r[0] - r[1])) > 1e-10 and np.sum(np.abs(r[1] - r[2])) > 1e-10
    # Gradient regularisation is tikhonov_gradient (already pinned by tv_admm's
    # one-iteration identity)
    g = oracle.tikhonov_gradient(f8, ones, d3, (1, 1, 1), (0, 0, 1), 0.01)
    assert np.max(np.abs(g - r[1])) <= 1e-14


def _rts_stage1_numpy(f, m, dims, vs, bdir, delta, lsmr):
    F = np.fft.fftn(f.reshape(dims, order="F"))
    kx, ky, kz = [np.fft.fftfreq(n, v) for n, v in zip(dims, vs)]
    bb = np.asarray(bdir, float) / np.linalg.norm(bdir)
    KX, KY, KZ = np.meshgrid(kx, ky, kz, indexing="ij")
    k2 = KX ** 2 + KY ** 2 + KZ ** 2
    with np.errstate(divide="ignore", invalid="ignore"):
        D = np.where(k2 > 1e-20, 1 / 3 - (KX * bb[0] + KY * bb[1] + KZ * bb[2]) ** 2 / k2, 0.0)
    cone = np.abs(D) > delta
    g = D / (D * D + 1e-6)
    w = np.where(cone, F * g, 0)
    for _ in range(lsmr):
        w = np.where(cone, w + (F - w * D) * g, w)
    x0 = np.real(np.fft.ifftn(w)).ravel(order="F")
    return np.where(m > 0, x0, 0.0), D, F


def test_rts_first_stage_and_first_step_closed_form(oracle):
    dims, vs, bdir = (12, 10, 8), (1.0, 1.0, 1.0), (0.1, 0.2, 1.0)
    rng = np.random.default_rng(2)
    n = int(np.prod(dims))
    f = rng.standard_normal(n) * 0.05
    m = (rng.random(n) < 0.8).astype(np.uint8)
    delta, mu, rho = 0.15, 1e5, 10.0
    x0, D, _ = _rts_stage1_numpy(f, m, dims, vs, bdir, delta, 4)
    got0 = oracle.rts(f, m, dims, vs, bdir, delta, mu, rho, 1e-2, 0, 4)
    assert np.max(np.abs(got0 - x0)) <= 1e-12 * np.max(np.abs(x0))
    # one ADMM step from g = 0: x1 = Re IFFT(FFT(x0) (M / rho) inv_a)
    kx, ky, kz = [np.fft.fftfreq(nn, v) for nn, v in zip(dims, vs)]
    L = sum(np.meshgrid(*[(2 - 2 * np.cos(2 * np.pi * np.arange(nn) / nn)) / v ** 2
                          for nn, v in zip(dims, vs)], indexing="ij"))
    M = np.where(np.abs(D) > delta, mu, 0.0)
    a = M + rho * L
    with np.errstate(divide="ignore"):
        inv = np.where(np.abs(a) > 1e-20, rho / a, 0.0)
    fh = np.where((np.abs(M) > 1e-20) & (np.abs(inv) > 1e-20), (M / rho) * inv, 0.0)
    x1 = np.real(np.fft.ifftn(np.fft.fftn(x0.reshape(dims, order="F")) * fh)).ravel(order="F")
    x1 = np.where(m > 0, x1, 0.0)
    got1 = oracle.rts(f, m, dims, vs, bdir, delta, mu, rho, 0.0, 1, 4)
    assert np.max(np.abs(got1 - x1)) <= 1e-10 * np.max(np.abs(x1))


def test_rts_reference_unit_tests(oracle):
    # src/inversion/rts.rs:219-300
    n = 8
    d3 = (n, n, n)
    ones = np.ones(n ** 3, np.uint8)
    kw = dict(delta=0.15, mu=1e5, rho=10.0, tol=1e-2, max_iter=5, lsmr_iter=2)
    assert np.max(np.abs(oracle.rts(np.zeros(n ** 3), ones, d3, (1, 1, 1), (0, 0, 1), **kw))) < 1e-6
    ramp = np.arange(n ** 3) * 0.001
    assert np.all(np.isfinite(oracle.rts(ramp, ones, d3, (1, 1, 1), (0, 0, 1), **kw)))
    m = ones.copy()
    m[0] = m[10] = 0
    chi = oracle.rts(ramp, m, d3, (1, 1, 1), (0, 0, 1), **kw)
    assert chi[0] == 0.0 and chi[10] == 0.0
    f16 = np.sin(np.arange(16 ** 3) * 0.7) * 0.01
    kw16 = dict(delta=0.15, mu=1e5, rho=10.0, tol=1e-4, max_iter=20, lsmr_iter=4)
    one = oracle.num_threads()
    oracle.set_num_threads(1)
    try:
        a = oracle.rts(f16, np.ones(16 ** 3, np.uint8), (16,) * 3, (1, 1, 1), (0, 0, 1), **kw16)
    finally:
        oracle.set_num_threads(one)
    b = oracle.rts(f16, np.ones(16 ** 3, np.uint8), (16,) * 3, (1, 1, 1), (0, 0, 1), **kw16)
    assert np.max(np.abs(a - b)) < 1e-10


def test_laplacian_unwrap_reference_tests_and_closed_form(oracle):
    # src/unwrap/laplacian.rs:170-230
    pi = np.pi
    for x, w in [(0.0, 0.0), (pi, pi), (-pi, -pi), (2 * pi, 0.0), (3 * pi, pi), (-3 * pi, -pi)]:
        assert abs(oracle.wrap(x) - w) < 1e-10
    n = 8
    ones = np.ones(n ** 3, np.uint8)
    u = oracle.laplacian_unwrap(np.ones(n ** 3), ones, (n, n, n), (1, 1, 1))
    assert np.max(np.abs(u - u.mean())) < 1e-6
    # closed form: a smooth periodic phase whose neighbour differences stay
    # below pi, wrapped into [-pi, pi], unwraps to itself minus its mean
    dims, vs = (16, 12, 10), (1.0, 0.8, 1.2)
    i, j, k = [np.arange(q) for q in dims]
    true = (6.0 * np.sin(2 * pi * i[:, None, None] / dims[0])
            + 4.0 * np.cos(2 * pi * j[None, :, None] / dims[1])
            + 3.0 * np.sin(2 * pi * k[None, None, :] / dims[2])).ravel(order="F")
    wrapped = np.array([oracle.wrap(v) for v in true])
    ones = np.ones(true.size, np.uint8)
    got = oracle.laplacian_unwrap(wrapped, ones, dims, vs)
    assert np.max(np.abs(got - (true - true.mean()))) < 1e-10


def _nltv_numpy(f, dims, vs, bdir, lam, mu, max_iter, newton_iter):
    """Independent numpy restatement of src/inversion/nltv.rs:63-133 (tol = 0,
    all-ones mask) for tiny volumes: full-spectrum FFTs, f64 throughout."""
    rho = 100 * lam
    F = lambda a: np.fft.fftn(a)
    Fi = lambda a: np.fft.ifftn(a)
    kx, ky, kz = [np.fft.fftfreq(n, v) for n, v in zip(dims, vs)]
    KX, KY, KZ = np.meshgrid(kx, ky, kz, indexing="ij")
    bb = np.asarray(bdir, float) / np.linalg.norm(bdir)
    k2 = KX ** 2 + KY ** 2 + KZ ** 2
    with np.errstate(divide="ignore", invalid="ignore"):
        D = np.where(k2 > 1e-20, 1 / 3 - (KX * bb[0] + KY * bb[1] + KZ * bb[2]) ** 2 / k2, 0.0)
    h = [1 / v ** 2 for v in vs]
    stencil = np.zeros(dims)
    stencil[0, 0, 0] = 2 * sum(h)
    for a, ha in enumerate(h):
        for s in (1, -1):
            idx = [0, 0, 0]
            idx[a] = s % dims[a]
            stencil[tuple(idx)] -= ha
    L = F(stencil).real
    A = D * D + rho * L
    inv = np.where(np.abs(A) > 1e-20, 1 / np.where(A == 0, 1, A), 0.0)
    fh = F(f.reshape(dims, order="F")) * D * inv
    grad = lambda x: [(np.roll(x, -1, axis=a) - x) / vs[a] for a in range(3)]
    bdiv = lambda g: sum((g[a] - np.roll(g[a], 1, axis=a)) / vs[a] for a in range(3))
    shrink = lambda v, t: np.sign(v) * np.maximum(np.abs(v) - t, 0.0)
    x = np.zeros(dims)
    u = [np.zeros(dims) for _ in range(3)]
    g = [np.zeros(dims) for _ in range(3)]
    w = np.ones(dims)
    t = lam / rho
    for _ in range(newton_iter):
        for _ in range(max_iter):
            x = Fi(fh - rho * F(bdiv(g)) * inv).real
            gr = grad(x)
            for a in range(3):
                v = gr[a] + u[a]
                z = shrink(v, t * w)
                u[a] = v - z
                g[a] = 2 * z - v
        g = grad(x)
        w = 1 / (np.sqrt(sum(q * q for q in g)) + mu * 1e-6)
        if w.max() > 1:
            w = w / w.max()
    return x.ravel(order="F")


def test_nltv_oracle_against_numpy_and_tv(oracle):
    dims, vs, b = (8, 6, 10), (1.0, 1.0, 1.0), (0.1, 0.0, 1.0)
    rng = np.random.default_rng(6)
    n = int(np.prod(dims))
    f = rng.standard_normal(n) * 0.02
    ones = np.ones(n, np.uint8)
    got = oracle.nltv(f, ones, dims, vs, b, lam=2e-3, mu=1.0, tol=0.0, max_iter=3, newton_iter=3)
    ref = _nltv_numpy(f, dims, vs, b, 2e-3, 1.0, 3, 3)
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref))
    # one Newton round with unit weights is tv_admm at rho = 100 lambda
    one = oracle.nltv(f, ones, dims, vs, b, lam=2e-3, mu=1.0, tol=0.0, max_iter=7, newton_iter=1)
    tv = oracle.tv_admm(f, ones, dims, vs, b, 2e-3, 0.2, 0.0, 7)
    assert np.array_equal(one, tv)
    # src/inversion/nltv.rs:138-185: zero field -> zero, finite
    assert np.max(np.abs(oracle.nltv(np.zeros(512), np.ones(512, np.uint8), (8, 8, 8), vs, (0, 0, 1),
                                     max_iter=5, newton_iter=2))) < 1e-10
    ramp = np.arange(512) * 0.001
    assert np.all(np.isfinite(oracle.nltv(ramp, np.ones(512, np.uint8), (8, 8, 8), vs, (0, 0, 1),
                                          max_iter=5, newton_iter=2)))
