"""Pins the oracle's hand-derived adjoint of the unrolled TV-ADMM loop
(orc_tv_unrolled_vjp) against central finite differences of the oracle's
forward tv_admm (tol = 0, src/inversion/tv.rs:63-97) in f64, and the
closed-form K = 1 case (x_1 = A f with A = Re IFFT(D inv_a FFT .), so
dL/df = A chi_bar, a symmetric operator)."""
import numpy as np
import pytest


def _problem(seed, dims, K):
    rng = np.random.default_rng(seed)
    n = int(np.prod(dims))
    f = rng.standard_normal(n) * 0.02
    m = (rng.random(n) < 0.85).astype(np.uint8)
    chi_true = rng.standard_normal(n) * 0.01
    return f, m, chi_true


def _loss(oracle, f, m, dims, vs, b, lam, rho, K, chi_true):
    chi = oracle.tv_admm(f, m, dims, vs, b, lam, rho, 0.0, K)
    return float(np.sum((chi - chi_true) ** 2))


@pytest.mark.parametrize("dims,b,K", [((8, 6, 10), (0.0, 0.0, 1.0), 4),
                                      ((6, 8, 6), (0.2, -0.1, 0.97), 6),
                                      ((7, 9, 5), (0.2, -0.1, 0.97), 12)])
def test_unrolled_vjp_matches_central_differences(oracle, dims, b, K):
    vs, lam, rho = (1.0, 1.0, 1.0), 2e-3, 0.1
    f, m, chi_true = _problem(11, dims, K)
    chi = oracle.tv_admm(f, m, dims, vs, b, lam, rho, 0.0, K)
    chi2, fb, lb = oracle.tv_unrolled_vjp(f, m, dims, vs, b, lam, rho, K, 2.0 * (chi - chi_true))
    assert np.array_equal(chi, chi2)
    # 16 random unit-RMS directions; step 2e-8 = 1e-6 of the field's RMS,
    # small enough that no soft-threshold kink is crossed (the loss is
    # piecewise smooth), large enough for f64 differences
    rng = np.random.default_rng(5)
    h = 2e-8
    errs = []
    for _ in range(16):
        v = rng.standard_normal(f.size)
        lp = _loss(oracle, f + h * v, m, dims, vs, b, lam, rho, K, chi_true)
        lm = _loss(oracle, f - h * v, m, dims, vs, b, lam, rho, K, chi_true)
        fd = (lp - lm) / (2 * h)
        an = float(fb @ v)
        errs.append(abs(fd - an) / max(abs(fd), 1e-12))
    print("field directions: max rel err", max(errs))
    assert max(errs) < 1e-5
    hl = 1e-6 * lam
    fd = (_loss(oracle, f, m, dims, vs, b, lam + hl, rho, K, chi_true)
          - _loss(oracle, f, m, dims, vs, b, lam - hl, rho, K, chi_true)) / (2 * hl)
    print("lambda: fd", fd, "adjoint", lb)
    assert abs(fd - lb) <= 1e-5 * max(abs(fd), 1e-12)


def test_unrolled_vjp_one_iteration_is_symmetric_operator(oracle):
    # K = 1: x_1 = A f, so field_bar = A (mask chi_bar); <A a, b> = <a, A b>
    dims, vs, b = (10, 8, 6), (1.0, 1.0, 1.0), (0.1, 0.0, 1.0)
    rng = np.random.default_rng(3)
    n = int(np.prod(dims))
    a1, a2 = rng.standard_normal(n), rng.standard_normal(n)
    ones = np.ones(n, np.uint8)
    Aa1 = oracle.tv_admm(a1, ones, dims, vs, b, 1e-3, 0.1, 0.0, 1)
    _, At_a2, lb = oracle.tv_unrolled_vjp(a1, ones, dims, vs, b, 1e-3, 0.1, 1, a2)
    assert lb == 0.0
    lhs, rhs = float(Aa1 @ a2), float(a1 @ At_a2)
    assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), 1.0)
    # and A is what Tikhonov with lambda = rho computes (admm.rs:165-176)
    tk = oracle.tikhonov_gradient(a2, ones, dims, vs, b, 0.1)
    assert np.max(np.abs(At_a2 - tk)) <= 1e-12 * np.max(np.abs(tk))


def test_unrolled_vjp_is_linear_in_chi_bar(oracle):
    dims, vs, b, K = (8, 8, 8), (1.0, 1.0, 1.0), (0.0, 0.0, 1.0), 5
    f, m, _ = _problem(2, dims, K)
    rng = np.random.default_rng(4)
    c1, c2 = rng.standard_normal(f.size), rng.standard_normal(f.size)
    _, g1, l1 = oracle.tv_unrolled_vjp(f, m, dims, vs, b, 2e-3, 0.1, K, c1)
    _, g2, l2 = oracle.tv_unrolled_vjp(f, m, dims, vs, b, 2e-3, 0.1, K, c2)
    _, g3, l3 = oracle.tv_unrolled_vjp(f, m, dims, vs, b, 2e-3, 0.1, K, 2.0 * c1 - 3.0 * c2)
    assert np.max(np.abs(g3 - (2 * g1 - 3 * g2))) <= 1e-10 * np.max(np.abs(g3))
    assert abs(l3 - (2 * l1 - 3 * l2)) <= 1e-10 * max(abs(l3), 1e-12)
