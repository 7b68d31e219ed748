"""Write the small golden fixtures under tests/golden/ (run in the build
container; the fixtures are committed, this script documents how).

Every fixture is computed here with numpy only (pocketfft for the FFT), from
independent restatements of the reference definitions -- none calls the C++
oracle, so comparing the oracle with these files checks two separate
implementations against each other:

  fft_odd.npz        fftn / ifftn of complex volumes whose axes need radices
                     3, 5, 7, 11, 13 (and a power-of-two case)
  kernels.npz        fftfreq(n, d) values, a tilted-B0 dipole kernel
                     (src/kernels/dipole.rs:24-72), the negative-Laplacian
                     spectrum (FFT of src/kernels/laplacian.rs:20-62) and
                     SMV sphere voxel counts by brute force
                     (src/kernels/smv.rs:19-67)
  tv_traj.npz        a numpy restatement of tv_admm (src/inversion/tv.rs:63-97
                     with admm_step, src/inversion/admm.rs:76-178) on a 16^3
                     case: chi after 1, 2, 10 and 50 iterations
  vsharp16.npz       a numpy restatement of vsharp_with_radii
                     (src/bgremove/vsharp.rs:88-199) on 16^3: local field and
                     eroded mask
Usage: python tests/golden/make_golden.py
"""
from __future__ import annotations

import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def F(a):
    return np.asarray(a).reshape(-1).copy()


def vol(flat, dims):
    return np.asarray(flat).reshape(dims, order="F")


def flat(v):
    return np.asarray(v).ravel(order="F")


# --------------------------------------------------------------- kernels --
def fftfreq(n, d):
    return np.fft.fftfreq(n, d)


def dipole(dims, vs, b):
    b = np.asarray(b, float)
    b = b / np.sqrt(b @ b)
    kx, ky, kz = (fftfreq(n, s) for n, s in zip(dims, vs))
    KX, KY, KZ = np.meshgrid(kx, ky, kz, indexing="ij")
    kdb = KX * b[0] + KY * b[1] + KZ * b[2]
    k2 = KX ** 2 + KY ** 2 + KZ ** 2
    with np.errstate(divide="ignore", invalid="ignore"):
        return np.where(k2 > 1e-20, 1.0 / 3.0 - kdb ** 2 / k2, 0.0)


def neg_laplacian_stencil(dims, vs):
    L = np.zeros(dims)
    h = [1.0 / v ** 2 for v in vs]
    L[0, 0, 0] = 2 * sum(h)
    for a in range(3):
        if dims[a] > 1:
            for s in (1, dims[a] - 1):
                idx = [0, 0, 0]
                idx[a] = s
                L[tuple(idx)] -= h[a]
    return L


def smv_sphere(dims, vs, r):
    g = []
    for n, s in zip(dims, vs):
        i = np.arange(n)
        g.append(np.where(i <= n // 2, i, i - n).astype(float) * s)
    X, Y, Z = np.meshgrid(*g, indexing="ij")
    s = ((X * X + Y * Y + Z * Z) <= r * r).astype(float)
    return s / s.sum(), s.sum()


# ------------------------------------------------------------ TV-ADMM ----
def tv_admm_np(f, mask, dims, vs, b, lam, rho, iters):
    n = np.prod(dims)
    D = dipole(dims, vs, b)
    Lh = np.real(np.fft.fftn(neg_laplacian_stencil(dims, vs)))
    a = D * D + rho * Lh
    inv_a = np.where(np.abs(a) > 1e-20, 1.0 / np.where(a == 0, 1, a), 0.0)
    fh = np.fft.fftn(vol(f, dims)) * D * inv_a
    h = [1.0 / v for v in vs]
    t = lam / rho
    x = np.zeros(dims)
    u = np.zeros((3,) + dims)
    g = np.zeros((3,) + dims)
    sh = lambda v: np.where(v > t, v - t, np.where(v < -t, v + t, 0.0))
    for _ in range(iters):
        div = sum((g[k] - np.roll(g[k], 1, axis=k)) * h[k] for k in range(3))
        xh = fh - rho * np.fft.fftn(div) * inv_a
        x = np.real(np.fft.ifftn(xh))
        grad = np.stack([(np.roll(x, -1, axis=k) - x) * h[k] for k in range(3)])
        v = grad + u
        z = sh(v)
        u = v - z
        g = 2 * z - v
    return flat(np.where(vol(mask, dims) > 0, x, 0.0))


# ------------------------------------------------------------- V-SHARP ---
def vsharp_np(field, mask, dims, vs, radii, thr):
    radii = sorted(radii, reverse=True)
    fh = np.fft.fftn(vol(field, dims))
    m3 = vol(mask, dims).astype(float)
    delta = 1.0 - np.sqrt(1e-7)
    processed = np.zeros(dims, bool)
    local = np.zeros(dims)
    inv = None
    for r in radii:
        s, _ = smv_sphere(dims, vs, r)
        S = np.real(np.fft.fftn(s))
        if inv is None:
            oms = 1.0 - S
            inv = np.where(np.abs(oms) < thr, 0.0, 1.0 / np.where(oms == 0, 1, oms))
        er = np.real(np.fft.ifftn(np.fft.fftn(m3) * S)) > delta
        filt = np.real(np.fft.ifftn(fh * (1.0 - S)))
        sel = er & ~processed
        local[sel] = filt[sel]
        processed |= sel
    out = np.real(np.fft.ifftn(np.fft.fftn(local) * inv))
    out = np.where(processed, out, 0.0)
    return flat(out), flat(processed.astype(np.uint8))


def main():
    rng = np.random.default_rng(20240601)
    # FFT fixtures (Fortran-ordered flat vectors, like the oracle's arrays)
    fx = {}
    for name, dims in {"a": (8, 6, 10), "b": (22, 13, 6), "c": (7, 5, 9), "d": (16, 8, 4)}.items():
        x = rng.standard_normal(dims) + 1j * rng.standard_normal(dims)
        fx[f"{name}_dims"] = np.array(dims)
        fx[f"{name}_x"] = flat(x)
        fx[f"{name}_fwd"] = flat(np.fft.fftn(x))
        fx[f"{name}_inv"] = flat(np.fft.ifftn(x))
    np.savez_compressed(os.path.join(HERE, "fft_odd.npz"), **fx)

    k = {}
    for n, d in [(4, 1.0), (5, 1.0), (7, 0.6), (10, 0.9), (13, 1.3)]:
        k[f"fftfreq_{n}"] = fftfreq(n, d)
        k[f"fftfreq_{n}_d"] = np.array(d)
    dims, vs, b = (10, 7, 6), (0.9, 1.1, 1.3), (0.2, -0.3, 0.93)
    k["dipole_dims"], k["dipole_vs"], k["dipole_b"] = np.array(dims), np.array(vs), np.array(b)
    k["dipole"] = flat(dipole(dims, vs, b))
    k["lap_dims"], k["lap_vs"] = np.array(dims), np.array(vs)
    k["lap_spec"] = flat(np.real(np.fft.fftn(neg_laplacian_stencil(dims, vs))))
    cnt = []
    for (sd, svs, r) in [((16, 16, 16), (1, 1, 1), 3.0), ((20, 18, 16), (0.6, 0.8, 1.0), 2.4),
                         ((32, 32, 32), (0.9, 0.9, 1.0), 6.0), ((8, 8, 8), (1, 1, 1), 0.5),
                         ((16, 16, 16), (1, 1, 1), 4.0)]:
        _, c = smv_sphere(sd, svs, r)
        cnt.append([*sd, *svs, r, c])
    k["smv_counts"] = np.array(cnt)
    np.savez_compressed(os.path.join(HERE, "kernels.npz"), **k)

    dims, vs, b = (16, 16, 16), (1.0, 1.0, 1.0), (0.1, 0.2, 1.0)
    f = np.sin(np.arange(16 ** 3) * 0.7) * 0.01 + rng.standard_normal(16 ** 3) * 0.002
    m = (rng.random(16 ** 3) < 0.85).astype(np.uint8)
    tv = {"dims": np.array(dims), "vs": np.array(vs), "b": np.array(b), "f": f, "mask": m,
          "lam": np.array(1e-3), "rho": np.array(0.1)}
    for it in (1, 2, 10, 50):
        tv[f"x{it}"] = tv_admm_np(f, m, dims, vs, b, 1e-3, 0.1, it)
    np.savez_compressed(os.path.join(HERE, "tv_traj.npz"), **tv)

    n = 16
    g = np.arange(n) - n // 2
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    mask = flat(((X ** 2 + Y ** 2 + Z ** 2) <= (n // 3) ** 2).astype(np.uint8))
    field = flat(Z * 0.1 + 0.02 * X * Y / n) + rng.standard_normal(n ** 3) * 0.01
    radii = [6.0, 4.0, 3.0, 2.0]
    loc, em = vsharp_np(field, mask, (n, n, n), (1.0, 1.0, 1.0), radii, 0.05)
    np.savez_compressed(os.path.join(HERE, "vsharp16.npz"), field=field, mask=mask,
                        radii=np.array(radii), threshold=np.array(0.05), local=loc, eroded=em)
    print("wrote fixtures to", HERE)


if __name__ == "__main__":
    main()
