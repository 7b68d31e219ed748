"""ctypes binding of the CPU f64 oracle (``oracle/liboracle.so``).

TEST INFRASTRUCTURE ONLY.  This module is the checker: the test-suite, the
smoke check and the CPU-baseline leg of ``bench.py`` import it; the product
package never does.

Every wrapper mirrors one function of qsm-core (the reference); the citation in
each docstring names the reference file:line.  Arrays are numpy arrays in
Fortran voxel order (``idx = i + j*nx + k*nx*ny``), i.e. a flat vector whose
reshape with ``order="F"`` to ``(nx, ny, nz)`` is the volume.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

PROGRESS_FN = C.CFUNCTYPE(None, C.c_size_t, C.c_size_t, C.c_void_p)

_D = C.POINTER(C.c_double)
_U8 = C.POINTER(C.c_uint8)
_SZ = C.c_size_t
_F = C.c_double


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (g++ -fopenmp)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "all"], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        sig = {
            "orc_fftfreq": (None, [_SZ, _F, _D]),
            "orc_fft3d": (None, [_D, _SZ, _SZ, _SZ, C.c_int]),
            "orc_fft1d": (None, [_D, _SZ, C.c_int]),
            "orc_dipole_kernel": (None, [_SZ] * 3 + [_F] * 6 + [_D]),
            "orc_laplacian_kernel": (None, [_SZ] * 3 + [_F] * 3 + [C.c_int, _D]),
            "orc_smv_kernel": (_F, [_SZ] * 3 + [_F] * 4 + [_D]),
            "orc_erode_mask_smv": (None, [_U8, _D] + [_SZ] * 3 + [_F, _U8]),
            "orc_fgrad": (None, [_D, _D, _D, _D] + [_SZ] * 3 + [_F] * 3),
            "orc_bdiv": (None, [_D, _D, _D, _D] + [_SZ] * 3 + [_F] * 3),
            "orc_shrink": (_F, [_F, _F]),
            "orc_relative_change": (_F, [_D, _D, _SZ]),
            "orc_apply_mask_zero": (None, [_D, _U8, _SZ]),
            "orc_prepare_admm_spectral": (None, [_D] + [_SZ] * 3 + [_F] * 7 + [_D, _D]),
            "orc_tv_admm": (None, [_D, _U8] + [_SZ] * 3 + [_F] * 9 + [_SZ, PROGRESS_FN,
                                   C.c_void_p, _D, C.POINTER(_SZ), _D]),
            "orc_tikhonov_gradient": (None, [_D, _U8] + [_SZ] * 3 + [_F] * 7 + [_D]),
            "orc_wrap": (_F, [_F]),
            "orc_laplacian_unwrap": (None, [_D, _U8] + [_SZ] * 3 + [_F] * 3 + [_D]),
            "orc_nltv": (None, [_D, _U8] + [_SZ] * 3 + [_F] * 9 + [_SZ, _SZ, PROGRESS_FN,
                                                                 C.c_void_p, _D]),
            "orc_tkd": (None, [_D, _U8] + [_SZ] * 3 + [_F] * 7 + [C.c_int, _D]),
            "orc_tikhonov": (None, [_D, _U8] + [_SZ] * 3 + [_F] * 7 + [C.c_int, _D]),
            "orc_rts": (None, [_D, _U8] + [_SZ] * 3 + [_F] * 10 + [_SZ, _SZ, PROGRESS_FN,
                                                                  C.c_void_p, _D, C.POINTER(_SZ)]),
            "orc_tv_unrolled_vjp": (None, [_D, _U8] + [_SZ] * 3 + [_F] * 8 + [_SZ, _D, _D, _D,
                                                                            C.POINTER(_F)]),
            "orc_vsharp_radii": (_SZ, [_F, _F, _D, _SZ]),
            "orc_vsharp_with_radii": (None, [_D, _U8] + [_SZ] * 3 + [_F] * 3 +
                                      [_D, _SZ, _F, PROGRESS_FN, C.c_void_p, _D, _U8]),
            "orc_vsharp": (None, [_D, _U8] + [_SZ] * 3 + [_F] * 6 +
                           [PROGRESS_FN, C.c_void_p, _D, _U8]),
            "orc_sharp_core": (None, [_D, _U8] + [_SZ] * 3 + [_F] * 5 + [_D, _U8]),
            "orc_dipole_forward": (None, [_D] + [_SZ] * 3 + [_F] * 6 + [_D]),
            "orc_rmse": (_F, [_D, _D, _U8, _SZ]),
            "orc_nrmse": (_F, [_D, _D, _U8, _SZ]),
            "orc_num_threads": (C.c_int, []),
            "orc_set_num_threads": (None, [C.c_int]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_D)


def _u8(a):
    return a.ctypes.data_as(_U8)


def _f64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel(order="F") if np.ndim(a) == 3
                                else np.asarray(a, dtype=np.float64).ravel())


def _mask(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint8).ravel(order="F") if np.ndim(a) == 3
                                else np.asarray(a, dtype=np.uint8).ravel())


def _cb(progress):
    if progress is None:
        return PROGRESS_FN(0)
    return PROGRESS_FN(lambda cur, tot, _u: progress(int(cur), int(tot)))


def num_threads() -> int:
    return lib().orc_num_threads()


def set_num_threads(n: int) -> None:
    lib().orc_set_num_threads(int(n))


def fftfreq(n: int, d: float) -> np.ndarray:
    """src/fft.rs:544-566"""
    out = np.empty(n)
    lib().orc_fftfreq(n, d, _d(out))
    return out


def fft3d(data: np.ndarray, dims, inverse: bool = False) -> np.ndarray:
    """src/fft.rs:97-209 / 412-487 (inverse includes 1/N)."""
    c = np.array(np.asarray(data, dtype=np.complex128).ravel(), copy=True)
    lib().orc_fft3d(c.ctypes.data_as(_D), dims[0], dims[1], dims[2], int(inverse))
    return c


def fft1d(data: np.ndarray, inverse: bool = False) -> np.ndarray:
    c = np.array(np.asarray(data, dtype=np.complex128).ravel(), copy=True)
    lib().orc_fft1d(c.ctypes.data_as(_D), c.size, int(inverse))
    return c


def dipole_kernel(dims, vs, bdir) -> np.ndarray:
    """src/kernels/dipole.rs:24-72"""
    out = np.empty(int(np.prod(dims)))
    lib().orc_dipole_kernel(*dims, *vs, *bdir, _d(out))
    return out


def laplacian_kernel(dims, vs, negative: bool) -> np.ndarray:
    """src/kernels/laplacian.rs:20-62"""
    out = np.empty(int(np.prod(dims)))
    lib().orc_laplacian_kernel(*dims, *vs, int(negative), _d(out))
    return out


def smv_kernel(dims, vs, radius):
    """src/kernels/smv.rs:19-67 -> (kernel, voxel count)."""
    out = np.empty(int(np.prod(dims)))
    cnt = lib().orc_smv_kernel(*dims, *vs, radius, _d(out))
    return out, cnt


def erode_mask_smv(mask, s_fft, dims, delta) -> np.ndarray:
    """src/kernels/smv.rs:79-105"""
    m = _mask(mask)
    s = _f64(s_fft)
    out = np.empty(m.size, dtype=np.uint8)
    lib().orc_erode_mask_smv(_u8(m), _d(s), *dims, delta, _u8(out))
    return out


def fgrad(x, dims, vs):
    """src/utils/gradient.rs:21-58"""
    x = _f64(x)
    gx, gy, gz = (np.empty(x.size) for _ in range(3))
    lib().orc_fgrad(_d(gx), _d(gy), _d(gz), _d(x), *dims, *vs)
    return gx, gy, gz


def bdiv(gx, gy, gz, dims, vs):
    """src/utils/gradient.rs:71-108"""
    gx, gy, gz = _f64(gx), _f64(gy), _f64(gz)
    out = np.empty(gx.size)
    lib().orc_bdiv(_d(out), _d(gx), _d(gy), _d(gz), *dims, *vs)
    return out


def shrink(x: float, t: float) -> float:
    """src/utils/mod.rs:44-52"""
    return lib().orc_shrink(x, t)


def relative_change(x, xp) -> float:
    """src/utils/mod.rs:69-77"""
    x, xp = _f64(x), _f64(xp)
    return lib().orc_relative_change(_d(x), _d(xp), x.size)


def apply_mask_zero(data, mask) -> np.ndarray:
    """src/utils/mask.rs:47-53"""
    d = _f64(data).copy()
    m = _mask(mask)
    lib().orc_apply_mask_zero(_d(d), _u8(m), d.size)
    return d


def prepare_admm_spectral(local_field, dims, vs, bdir, rho):
    """src/inversion/admm.rs:150-178 -> (inv_a, f_hat)."""
    f = _f64(local_field)
    inv_a = np.empty(f.size)
    f_hat = np.empty(f.size, dtype=np.complex128)
    lib().orc_prepare_admm_spectral(_d(f), *dims, *vs, *bdir, rho, _d(inv_a),
                                    f_hat.ctypes.data_as(_D))
    return inv_a, f_hat


def tv_admm(local_field, mask, dims, vs, bdir, lam, rho, tol, max_iter, progress=None,
            return_info=False):
    """src/inversion/tv.rs:63-97"""
    f = _f64(local_field)
    m = _mask(mask)
    out = np.empty(f.size)
    iters = _SZ(0)
    rel = _F(0.0)
    cb = _cb(progress)
    lib().orc_tv_admm(_d(f), _u8(m), *dims, *vs, *bdir, lam, rho, tol, max_iter, cb, None,
                      _d(out), C.byref(iters), C.byref(rel))
    if return_info:
        return out, int(iters.value), float(rel.value)
    return out


def tv_unrolled_vjp(local_field, mask, dims, vs, bdir, lam, rho, K, chi_bar):
    """VJP of K unrolled tv_admm iterations (tol = 0): returns (chi,
    field_bar, lambda_bar) for chi_bar = dL/dchi."""
    f = _f64(local_field)
    m = _mask(mask)
    cb = _f64(chi_bar)
    chi = np.empty(f.size)
    fb = np.empty(f.size)
    lb = _F(0.0)
    lib().orc_tv_unrolled_vjp(_d(f), _u8(m), *dims, *vs, *bdir, lam, rho, K, _d(cb), _d(chi),
                              _d(fb), C.byref(lb))
    return chi, fb, float(lb.value)


def wrap(x):
    """src/unwrap/laplacian.rs:41-50"""
    return float(lib().orc_wrap(float(x)))


def laplacian_unwrap(phase, mask, dims, vs):
    """src/unwrap/laplacian.rs:135-161"""
    p, m = _f64(phase), _mask(mask)
    out = np.empty(p.size)
    lib().orc_laplacian_unwrap(_d(p), _u8(m), *dims, *vs, _d(out))
    return out


def nltv(local_field, mask, dims, vs, bdir, lam=1e-3, mu=1.0, tol=1e-3, max_iter=250,
         newton_iter=10, progress=None):
    """src/inversion/nltv.rs:63-133"""
    f, m = _f64(local_field), _mask(mask)
    out = np.empty(f.size)
    cb = _cb(progress)
    lib().orc_nltv(_d(f), _u8(m), *dims, *vs, *bdir, lam, mu, tol, max_iter, newton_iter, cb, None,
                   _d(out))
    return out


def tkd(local_field, mask, dims, vs, bdir, threshold, tsvd=False):
    """src/inversion/tkd.rs:48-99 (tsvd=True: the tsvd variant after it)"""
    f, m = _f64(local_field), _mask(mask)
    out = np.empty(f.size)
    lib().orc_tkd(_d(f), _u8(m), *dims, *vs, *bdir, threshold, 1 if tsvd else 0, _d(out))
    return out


def tikhonov(local_field, mask, dims, vs, bdir, lam, reg):
    """src/inversion/tikhonov.rs:62-131; reg 0 Identity, 1 Gradient, 2 Laplacian"""
    f, m = _f64(local_field), _mask(mask)
    out = np.empty(f.size)
    lib().orc_tikhonov(_d(f), _u8(m), *dims, *vs, *bdir, lam, int(reg), _d(out))
    return out


def rts(local_field, mask, dims, vs, bdir, delta=0.15, mu=1e5, rho=10.0, tol=1e-2, max_iter=20,
        lsmr_iter=4, progress=None, return_info=False):
    """src/inversion/rts.rs:71-212"""
    f, m = _f64(local_field), _mask(mask)
    out = np.empty(f.size)
    iters = _SZ(0)
    cb = _cb(progress)
    lib().orc_rts(_d(f), _u8(m), *dims, *vs, *bdir, delta, mu, rho, tol, max_iter, lsmr_iter, cb,
                  None, _d(out), C.byref(iters))
    return (out, int(iters.value)) if return_info else out


def tikhonov_gradient(local_field, mask, dims, vs, bdir, lam):
    """src/inversion/tikhonov.rs:62-131 (Regularization::Gradient)"""
    f = _f64(local_field)
    m = _mask(mask)
    out = np.empty(f.size)
    lib().orc_tikhonov_gradient(_d(f), _u8(m), *dims, *vs, *bdir, lam, _d(out))
    return out


def vsharp_radii(max_radius, min_radius):
    """src/bgremove/vsharp.rs:70-82"""
    n = lib().orc_vsharp_radii(max_radius, min_radius, None, 0)
    out = np.empty(n)
    lib().orc_vsharp_radii(max_radius, min_radius, _d(out), n)
    return out


def vsharp_with_radii(field, mask, dims, vs, radii, threshold, progress=None):
    """src/bgremove/vsharp.rs:88-199"""
    f = _f64(field)
    m = _mask(mask)
    r = np.ascontiguousarray(np.asarray(radii, dtype=np.float64))
    local = np.empty(f.size)
    mout = np.empty(f.size, dtype=np.uint8)
    cb = _cb(progress)
    lib().orc_vsharp_with_radii(_d(f), _u8(m), *dims, *vs, _d(r), r.size, threshold, cb, None,
                                _d(local), _u8(mout))
    return local, mout


def vsharp(field, mask, dims, vs, threshold=0.05, max_radius=12.0, min_radius=1.0,
           progress=None):
    """src/bgremove/vsharp.rs:55-64"""
    f = _f64(field)
    m = _mask(mask)
    local = np.empty(f.size)
    mout = np.empty(f.size, dtype=np.uint8)
    cb = _cb(progress)
    lib().orc_vsharp(_d(f), _u8(m), *dims, *vs, threshold, max_radius, min_radius, cb, None,
                     _d(local), _u8(mout))
    return local, mout


def sharp_core(field, mask, dims, vs, threshold, radius):
    """src/bgremove/sharp.rs:64-131"""
    f = _f64(field)
    m = _mask(mask)
    local = np.empty(f.size)
    mout = np.empty(f.size, dtype=np.uint8)
    lib().orc_sharp_core(_d(f), _u8(m), *dims, *vs, threshold, radius, _d(local), _u8(mout))
    return local, mout


def dipole_forward(chi, dims, vs, bdir):
    """tests/chisep_integration.rs:18-62 (dipole convolution)"""
    c = _f64(chi)
    out = np.empty(c.size)
    lib().orc_dipole_forward(_d(c), *dims, *vs, *bdir, _d(out))
    return out


def rmse(a, b, mask) -> float:
    """tests/common/mod.rs:17-31"""
    a, b, m = _f64(a), _f64(b), _mask(mask)
    return lib().orc_rmse(_d(a), _d(b), _u8(m), a.size)


def nrmse(a, b, mask) -> float:
    """tests/common/mod.rs:34-51 (range of ``b`` inside the mask)."""
    a, b, m = _f64(a), _f64(b), _mask(mask)
    return lib().orc_nrmse(_d(a), _d(b), _u8(m), a.size)
