"""Seeded synthetic QSM inputs: ellipsoid susceptibility phantoms, their local
field and, for background-removal workloads, a harmonic background field.

The random stream is SplitMix64 (counter based, so any draw can be recomputed
from the seed alone) with Box-Muller normals.  The local field is the
reference's forward model Re IFFT(D * FFT(chi)) (tests/chisep_integration.rs:
18-62, dipole kernel of src/kernels/dipole.rs:24-72), computed here in float64
with a real-to-complex FFT (scipy.fft, multithreaded; numpy's FFT if scipy is
absent); noise is added and the field is zeroed outside the mask.

Arrays are flat float64 / uint8 vectors in Fortran voxel order
(idx = i + j*nx + k*nx*ny).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, counter: np.ndarray) -> np.ndarray:
    """SplitMix64 output number ``counter`` (0-based) of the stream ``seed``."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + (np.asarray(counter, dtype=np.uint64) + np.uint64(1)) * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def uniform(seed: int, counter: np.ndarray) -> np.ndarray:
    """Uniform doubles in [0, 1) from the top 53 bits."""
    return (splitmix64(seed, counter) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def normal(seed: int, n: int, offset: int = 0) -> np.ndarray:
    """n standard normals (Box-Muller over pairs of uniforms)."""
    m = (n + 1) // 2
    c = np.arange(m, dtype=np.uint64) * np.uint64(2) + np.uint64(offset)
    u1 = uniform(seed, c)
    u2 = uniform(seed, c + np.uint64(1))
    r = np.sqrt(-2.0 * np.log1p(-u1))
    z = np.empty(2 * m)
    z[0::2] = r * np.cos(2.0 * np.pi * u2)
    z[1::2] = r * np.sin(2.0 * np.pi * u2)
    return z[:n]


class _Stream:
    def __init__(self, seed: int):
        self.seed = seed
        self.pos = 1 << 40  # scalar draws live far away from the noise stream

    def u(self, lo=0.0, hi=1.0) -> float:
        v = float(uniform(self.seed, np.array([self.pos], dtype=np.uint64))[0])
        self.pos += 1
        return lo + (hi - lo) * v


def fftfreq_grid(dims, vs):
    return [np.fft.fftfreq(n, d) for n, d in zip(dims, vs)]


def dipole_kernel(dims, vs, bdir) -> np.ndarray:
    """D = 1/3 - (k.b)^2/|k|^2, D(0) = 0, as a (nx, ny, nz) float64 array
    (src/kernels/dipole.rs:24-72)."""
    kx, ky, kz = fftfreq_grid(dims, vs)
    b = np.asarray(bdir, dtype=np.float64)
    b = b / math.sqrt(float(b @ b))
    KX, KY, KZ = np.meshgrid(kx, ky, kz, indexing="ij")
    kdb = KX * b[0] + KY * b[1] + KZ * b[2]
    k2 = KX * KX + KY * KY + KZ * KZ
    with np.errstate(divide="ignore", invalid="ignore"):
        D = np.where(k2 > 1e-20, 1.0 / 3.0 - kdb * kdb / k2, 0.0)
    return D


def dipole_forward(chi3: np.ndarray, vs, bdir) -> np.ndarray:
    """Re IFFT(D * FFT(chi)) (tests/chisep_integration.rs:18-62), through a
    real FFT.  Re IFFT(F D) = IFFT(F D_h) with D_h(k) = (D(k) + D(k'))/2,
    k' = k with every component that sits on its axis's Nyquist index
    negated (fftfreq puts -n/2 there, so the DFT's mirror of k is -k' there);
    D_h differs from D only on those planes and only for a tilted B0."""
    dims = chi3.shape
    nx, ny, nz = dims
    try:
        import scipy.fft as sfft
        F = sfft.rfftn(chi3, workers=-1)
        irfftn = lambda a: sfft.irfftn(a, s=dims, workers=-1)  # noqa: E731
    except ImportError:  # pragma: no cover
        F = np.fft.rfftn(chi3)
        irfftn = lambda a: np.fft.irfftn(a, s=dims)  # noqa: E731
    b = np.asarray(bdir, dtype=np.float64)
    b = b / math.sqrt(float(b @ b))

    def freqs(n, d, half):
        f = np.fft.fftfreq(n, d)
        if half:
            f = f[: n // 2 + 1]
        fl = f.copy()
        if n % 2 == 0:
            fl[n // 2] = -f[n // 2]
        return f, fl

    kx, kxf = freqs(nx, vs[0], False)
    ky, kyf = freqs(ny, vs[1], False)
    kz, kzf = freqs(nz, vs[2], True)

    def dk(ax, ay, az):
        ax, ay, az = ax[:, None, None], ay[None, :, None], az[None, None, :]
        kdb = ax * b[0] + ay * b[1] + az * b[2]
        k2 = ax * ax + ay * ay + az * az
        with np.errstate(divide="ignore", invalid="ignore"):
            return np.where(k2 > 1e-20, 1.0 / 3.0 - kdb * kdb / k2, 0.0)

    D = 0.5 * (dk(kx, ky, kz) + dk(kxf, kyf, kzf))
    return irfftn(F * D)


@dataclass
class Phantom:
    dims: tuple
    voxel_size: tuple
    bdir: tuple
    chi: np.ndarray        # ground-truth susceptibility (ppm), flat
    mask: np.ndarray       # uint8, flat
    local_field: np.ndarray  # ppm, flat, zero outside the mask
    total_field: np.ndarray | None = None  # local + background (+ noise), flat


def ellipsoid_mask(dims) -> np.ndarray:
    """Ellipsoid with semi-axes 0.4*dims centred at dims/2."""
    t = [((np.arange(n, dtype=np.float64) - n / 2.0) / (0.4 * n)) ** 2 for n in dims]
    r2 = t[0][:, None, None] + t[1][None, :, None] + t[2][None, None, :]
    return (r2 <= 1.0).astype(np.uint8)


def make_phantom(dims, voxel_size=(1.0, 1.0, 1.0), bdir=(0.0, 0.0, 1.0), n_ellipsoids=12,
                 seed=1000, noise_sigma=0.005, semi_axes=(5.0, 25.0),
                 chi_range=(-0.1, 0.15)) -> Phantom:
    """Ellipsoid phantom: n axis-aligned ellipsoids, semi-axes U(5, 25) voxels,
    centres uniform inside the mask, chi ~ U(-0.1, 0.15) ppm summed on a zero
    background; local field = Re IFFT(D FFT(chi)) + N(0, sigma), zero outside
    the mask."""
    dims = tuple(int(n) for n in dims)
    mask3 = ellipsoid_mask(dims)
    st = _Stream(seed)
    g = [np.arange(n, dtype=np.float64) for n in dims]
    chi3 = np.zeros(dims)
    for _ in range(n_ellipsoids):
        while True:  # centre uniform inside the mask (rejection)
            c = [st.u(0.0, float(n)) for n in dims]
            r2 = sum(((c[a] - dims[a] / 2.0) / (0.4 * dims[a])) ** 2 for a in range(3))
            if r2 <= 1.0:
                break
        ax = [st.u(*semi_axes) for _ in range(3)]
        val = st.u(*chi_range)
        # separable evaluation (sum of three 1-D terms) over the bounding box
        lo = [max(0, int(math.floor(c[a] - ax[a]))) for a in range(3)]
        hi = [min(dims[a], int(math.ceil(c[a] + ax[a])) + 1) for a in range(3)]
        tx = ((g[0][lo[0]:hi[0]] - c[0]) / ax[0]) ** 2
        ty = ((g[1][lo[1]:hi[1]] - c[1]) / ax[1]) ** 2
        tz = ((g[2][lo[2]:hi[2]] - c[2]) / ax[2]) ** 2
        inside = (tx[:, None, None] + ty[None, :, None] + tz[None, None, :]) <= 1.0
        chi3[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]][inside] += val
    field3 = dipole_forward(chi3, voxel_size, bdir)
    n = int(np.prod(dims))
    noise = normal(seed, n).reshape(dims, order="F") * noise_sigma
    field3 = np.where(mask3 > 0, field3 + noise, 0.0)
    return Phantom(dims, tuple(voxel_size), tuple(bdir), chi3.ravel(order="F"),
                   mask3.ravel(order="F"), field3.ravel(order="F"))


def add_background(ph: Phantom, n_sources=3, seed=None, moment_range=(1.0, 5.0),
                   dist_range_mm=(50.0, 150.0), noise_sigma=0.005) -> Phantom:
    """Total field = local field + fields of external point dipoles (moment
    U(1,5) ppm*cm^3, 5-15 cm outside the mask, m (3 cos^2 t - 1) / (4 pi r^3)
    along B0) + N(0, sigma); harmonic inside the field of view."""
    dims, vs = ph.dims, ph.voxel_size
    st = _Stream((seed if seed is not None else 7777) + 17)
    b = np.asarray(ph.bdir, dtype=np.float64)
    b = b / math.sqrt(float(b @ b))
    centre = np.array([dims[a] / 2.0 * vs[a] for a in range(3)])
    mask_radius = max(0.4 * dims[a] * vs[a] for a in range(3))
    X, Y, Z = [(np.arange(n) * vs[a]).reshape([n if q == a else 1 for q in range(3)])
               for a, n in enumerate(dims)]
    bg = np.zeros(dims)
    for _ in range(n_sources):
        m = st.u(*moment_range) * 1000.0  # ppm * mm^3
        d = mask_radius + st.u(*dist_range_mm)
        th = math.acos(st.u(-1.0, 1.0))
        ph_ = st.u(0.0, 2.0 * math.pi)
        pos = centre + d * np.array([math.sin(th) * math.cos(ph_), math.sin(th) * math.sin(ph_),
                                     math.cos(th)])
        rx, ry, rz = X - pos[0], Y - pos[1], Z - pos[2]
        r = np.sqrt(rx * rx + ry * ry + rz * rz)
        cos_t = (rx * b[0] + ry * b[1] + rz * b[2]) / r
        bg += m * (3.0 * cos_t * cos_t - 1.0) / (4.0 * math.pi * r ** 3)
    n = int(np.prod(dims))
    noise = normal((seed if seed is not None else 7777) + 99, n).reshape(dims, order="F") * noise_sigma
    local3 = dipole_forward(ph.chi.reshape(dims, order="F"), vs, ph.bdir)
    total3 = local3 + bg + noise
    ph.total_field = total3.ravel(order="F")
    return ph


def tilted_bdir(seed: int, max_deg: float = 20.0):
    """B0 tilted from z by U(0, max_deg) with uniform azimuth."""
    st = _Stream(seed + 31)
    t = math.radians(st.u(0.0, max_deg))
    a = st.u(0.0, 2.0 * math.pi)
    return (math.sin(t) * math.cos(a), math.sin(t) * math.sin(a), math.cos(t))


def uneven_batch_shapes(seed: int, count: int, axes=tuple(range(144, 257, 16)),
                        vs_range=(0.6, 1.2)):
    """Per-volume dims drawn from `axes`, voxel sizes U(0.6, 1.2) mm per axis."""
    st = _Stream(seed + 53)
    out = []
    for _ in range(count):
        dims = tuple(axes[min(len(axes) - 1, int(st.u(0.0, len(axes))))] for _ in range(3))
        vs = tuple(st.u(*vs_range) for _ in range(3))
        out.append((dims, vs))
    return out
