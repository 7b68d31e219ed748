"""Pin the oracle's FFT (the stand-in for rustfft 6.4.1, Cargo.lock:326-329):
the DFT definition by brute force, numpy golden fixtures, the reference's
round-trip test (src/fft.rs:692-721), Parseval, impulse/constant closed forms
and OpenMP thread-count invariance (src/fft.rs:849-932)."""
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def dft3(x3, sign=-1):
    """Direct O(N^2) DFT over each axis (brute-force definition)."""
    out = x3.astype(np.complex128)
    for ax in range(3):
        n = out.shape[ax]
        k = np.arange(n)
        W = np.exp(sign * 2j * np.pi * np.outer(k, k) / n)
        out = np.moveaxis(np.tensordot(W, np.moveaxis(out, ax, 0), axes=(1, 0)), 0, ax)
    return out


@pytest.mark.parametrize("dims", [(2, 3, 5), (6, 7, 4), (11, 2, 13), (1, 4, 3), (10, 14, 22)])
def test_fft3d_matches_dft_definition(oracle, dims):
    rng = np.random.default_rng(sum(dims))
    x = rng.standard_normal(dims) + 1j * rng.standard_normal(dims)
    got = oracle.fft3d(x.ravel(order="F"), dims)
    ref = dft3(x).ravel(order="F")
    assert np.max(np.abs(got - ref)) < 1e-12 * max(1.0, np.max(np.abs(ref)))
    got_i = oracle.fft3d(x.ravel(order="F"), dims, inverse=True)
    ref_i = dft3(x, sign=+1).ravel(order="F") / np.prod(dims)
    assert np.max(np.abs(got_i - ref_i)) < 1e-12 * max(1.0, np.max(np.abs(ref_i)))


@pytest.mark.parametrize("n", [1, 2, 3, 5, 7, 11, 13, 17, 97, 144, 160, 176, 208, 224, 240, 384])
def test_fft1d_all_radices(oracle, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    k = np.arange(n)
    ref = np.exp(-2j * np.pi * np.outer(k, k) / n) @ x
    assert np.max(np.abs(oracle.fft1d(x) - ref)) < 1e-11 * max(1.0, np.max(np.abs(ref)))


def test_golden_fixtures(oracle):
    g = np.load(os.path.join(GOLDEN, "fft_odd.npz"))
    for name in "abcd":
        dims = tuple(int(v) for v in g[f"{name}_dims"])
        x = g[f"{name}_x"]
        fwd = oracle.fft3d(x, dims)
        inv = oracle.fft3d(x, dims, inverse=True)
        assert np.max(np.abs(fwd - g[f"{name}_fwd"])) < 1e-12 * np.max(np.abs(g[f"{name}_fwd"]))
        assert np.max(np.abs(inv - g[f"{name}_inv"])) < 1e-12 * np.max(np.abs(g[f"{name}_inv"]))


def test_roundtrip_reference_case(oracle):
    # src/fft.rs:692-721: 4^3 ramp, IFFT(FFT(x)) == x to 1e-10, imag < 1e-10
    x = np.arange(64, dtype=float)
    y = oracle.fft3d(oracle.fft3d(x, (4, 4, 4)), (4, 4, 4), inverse=True)
    assert np.max(np.abs(y.real - x)) < 1e-10
    assert np.max(np.abs(y.imag)) < 1e-10


def test_parseval_and_closed_forms(oracle):
    rng = np.random.default_rng(3)
    dims = (12, 10, 14)
    x = rng.standard_normal(dims).ravel(order="F")
    X = oracle.fft3d(x, dims)
    assert abs(np.sum(np.abs(X) ** 2) / x.size - np.sum(x ** 2)) < 1e-10 * np.sum(x ** 2)
    imp = np.zeros(x.size)
    imp[0] = 1.0
    assert np.allclose(oracle.fft3d(imp, dims), 1.0, atol=1e-14)
    ones = np.ones(x.size)
    F1 = oracle.fft3d(ones, dims)
    assert abs(F1[0] - x.size) < 1e-9 and np.max(np.abs(F1[1:])) < 1e-9


def test_thread_count_invariance(oracle):
    # src/fft.rs:849-932: the parallel transform equals the 1-thread one
    rng = np.random.default_rng(7)
    dims = (16, 16, 16)
    x = (rng.standard_normal(dims) + 1j * rng.standard_normal(dims)).ravel(order="F")
    nt = oracle.num_threads()
    try:
        oracle.set_num_threads(1)
        a = oracle.fft3d(x, dims)
        oracle.set_num_threads(max(2, nt))
        b = oracle.fft3d(x, dims)
    finally:
        oracle.set_num_threads(nt)
    assert np.array_equal(a, b)


def test_fftfreq_reference_values(oracle):
    # src/fft.rs:724-739
    assert np.allclose(oracle.fftfreq(4, 1.0), [0.0, 0.25, -0.5, -0.25], atol=1e-10)
    assert np.allclose(oracle.fftfreq(5, 1.0), [0.0, 0.2, 0.4, -0.4, -0.2], atol=1e-10)
    k = np.load(os.path.join(GOLDEN, "kernels.npz"))
    for n in (4, 5, 7, 10, 13):
        assert np.array_equal(oracle.fftfreq(n, float(k[f"fftfreq_{n}_d"])), k[f"fftfreq_{n}"])
