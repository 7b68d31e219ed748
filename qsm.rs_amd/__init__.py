"""Host-side mirror of QSM.rs's TV-ADMM / V-SHARP API over the gfx950 C ABI.

Every public name mirrors the reference crate (qsm-core):

=====================================  =========================================
this module                            reference
=====================================  =========================================
``Grid``                               ``Grid`` (src/grid.rs:9-44)
``TvParams``                           ``inversion::TvParams`` (src/inversion/tv.rs:21-43)
``VsharpParams``                       ``bgremove::VsharpParams`` (src/bgremove/vsharp.rs:20-39)
``tv_admm``                            ``inversion::tv_admm`` (src/inversion/tv.rs:63-97)
``vsharp``                             ``bgremove::vsharp`` (src/bgremove/vsharp.rs:55-64)
``vsharp_with_radii``                  ``vsharp_with_radii`` (src/bgremove/vsharp.rs:88-199)
``ScanMetadata``, ``InversionConfig``,
``BgRemovalConfig``, ``BgRemovalResult``,
``PipelineError``                      ``pipeline::config`` (src/pipeline/config.rs)
``run_dipole_inversion``               src/pipeline/inversion.rs:28-114
``run_bg_removal``                     src/pipeline/bg_removal.rs:20-92
=====================================  =========================================

Arrays are flat numpy vectors in Fortran voxel order (``idx = i + j*nx +
k*nx*ny``); a 3-D array is accepted and flattened with ``order="F"``.  All
computation runs in ``libqsm_gpu.so`` on the GPU; there is no CPU fallback:
importing this package fails if the library has not been built.
"""
from __future__ import annotations

import ctypes as C
import enum
import os
from dataclasses import dataclass, field
from typing import Callable, Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libqsm_gpu.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"{LIB_PATH} is missing: build it with `python qsm.rs_amd/build.py` "
        "(or __graft_entry__.build()); there is no CPU fallback")

_lib = C.CDLL(LIB_PATH)

# ---------------------------------------------------------------- C types --
QSM_OK, QSM_EDIM, QSM_EPARAM, QSM_EHIP, QSM_ENOMEM, QSM_ECONFIG, QSM_EUNSUPPORTED = range(7)

PROGRESS_FN = C.CFUNCTYPE(None, C.c_size_t, C.c_size_t, C.c_void_p)


class _CGrid(C.Structure):
    _fields_ = [("nx", C.c_size_t), ("ny", C.c_size_t), ("nz", C.c_size_t),
                ("vsx", C.c_double), ("vsy", C.c_double), ("vsz", C.c_double)]


class _CTvParams(C.Structure):
    _fields_ = [("lambda_", C.c_double), ("rho", C.c_double), ("tol", C.c_double),
                ("max_iter", C.c_size_t)]


class _CVsharpParams(C.Structure):
    _fields_ = [("threshold", C.c_double), ("max_radius", C.c_double),
                ("min_radius", C.c_double)]


class _CTkdParams(C.Structure):
    _fields_ = [("threshold", C.c_double)]


class _CTikhonovParams(C.Structure):
    _fields_ = [("lambda_", C.c_double), ("reg", C.c_int)]


class _CRtsParams(C.Structure):
    _fields_ = [("delta", C.c_double), ("mu", C.c_double), ("rho", C.c_double), ("tol", C.c_double),
                ("max_iter", C.c_size_t), ("lsmr_iter", C.c_size_t)]


class _CNltvParams(C.Structure):
    _fields_ = [("lambda_", C.c_double), ("mu", C.c_double), ("tol", C.c_double),
                ("max_iter", C.c_size_t), ("newton_iter", C.c_size_t)]


class _CSharpParams(C.Structure):
    _fields_ = [("threshold", C.c_double), ("radius", C.c_double)]


class _CInversionConfig(C.Structure):
    _fields_ = [("algorithm", C.c_int), ("tkd", _CTkdParams), ("tsvd", _CTkdParams),
                ("tikhonov", _CTikhonovParams), ("tv", _CTvParams), ("rts", _CRtsParams),
                ("nltv", _CNltvParams)]


class _CBgRemovalConfig(C.Structure):
    _fields_ = [("algorithm", C.c_int), ("vsharp", _CVsharpParams), ("sharp", _CSharpParams)]


class _CNiftiInfo(C.Structure):
    _fields_ = [("nx", C.c_size_t), ("ny", C.c_size_t), ("nz", C.c_size_t), ("nt", C.c_size_t),
                ("vsx", C.c_double), ("vsy", C.c_double), ("vsz", C.c_double),
                ("affine", C.c_double * 16), ("scl_slope", C.c_double), ("scl_inter", C.c_double),
                ("datatype", C.c_int)]


class _CScanMetadata(C.Structure):
    _fields_ = [("nx", C.c_size_t), ("ny", C.c_size_t), ("nz", C.c_size_t),
                ("vsx", C.c_double), ("vsy", C.c_double), ("vsz", C.c_double),
                ("echo_times", C.POINTER(C.c_double)), ("n_echo_times", C.c_size_t),
                ("field_strength", C.c_double), ("b0_direction", C.c_double * 3)]


class _CTvVolume(C.Structure):
    _fields_ = [("grid", _CGrid), ("bdir", C.c_double * 3), ("d_field", C.c_void_p),
                ("d_mask", C.c_void_p), ("d_chi", C.c_void_p)]


class _CVsharpVolume(C.Structure):
    _fields_ = [("grid", _CGrid), ("d_field", C.c_void_p), ("d_mask", C.c_void_p),
                ("d_local", C.c_void_p), ("d_mask_out", C.c_void_p)]


_D = C.POINTER(C.c_double)
_U8 = C.POINTER(C.c_uint8)
_SZ = C.c_size_t

_SIGS = {
    "qsm_tv_params_default": (None, [C.POINTER(_CTvParams)]),
    "qsm_vsharp_params_default": (None, [C.POINTER(_CVsharpParams)]),
    "qsm_last_error_message": (C.c_char_p, []),
    "qsm_backend_info": (C.c_int, [C.c_char_p, _SZ]),
    "qsm_fft_length_supported": (C.c_int, [_SZ]),
    "qsm_tv_admm": (C.c_int, [_D, _U8, C.POINTER(_CGrid), _D, C.POINTER(_CTvParams), PROGRESS_FN,
                              C.c_void_p, _D]),
    "qsm_tv_admm_ex": (C.c_int, [_D, _U8, C.POINTER(_CGrid), _D, C.POINTER(_CTvParams),
                                 PROGRESS_FN, C.c_void_p, _D, C.POINTER(_SZ), _D]),
    "qsm_vsharp": (C.c_int, [_D, _U8, C.POINTER(_CGrid), C.POINTER(_CVsharpParams), PROGRESS_FN,
                             C.c_void_p, _D, _U8]),
    "qsm_vsharp_with_radii": (C.c_int, [_D, _U8, C.POINTER(_CGrid), _D, _SZ, C.c_double,
                                        PROGRESS_FN, C.c_void_p, _D, _U8]),
    "qsm_run_dipole_inversion": (C.c_int, [_D, _SZ, _U8, _SZ, C.POINTER(_CScanMetadata), C.c_int,
                                           C.POINTER(_CTvParams), PROGRESS_FN, C.c_void_p, _D]),
    "qsm_run_bg_removal": (C.c_int, [_D, _SZ, _U8, _SZ, C.POINTER(_CScanMetadata), C.c_int,
                                     C.POINTER(_CVsharpParams), PROGRESS_FN, C.c_void_p, _D, _U8]),
    "qsm_tv_batch_create": (C.c_int, [C.POINTER(_CTvVolume), _SZ, C.POINTER(_CTvParams),
                                      C.c_void_p, C.POINTER(C.c_void_p)]),
    "qsm_tv_batch_run": (C.c_int, [C.c_void_p, PROGRESS_FN, C.c_void_p]),
    "qsm_tv_batch_set_graphs": (C.c_int, [C.c_void_p, C.c_int]),
    "qsm_set_fft_schedule": (C.c_int, [C.c_int]),
    "qsm_fft_schedule": (C.c_int, [C.c_int, C.c_int, C.POINTER(C.c_int)]),
    "qsm_tv_batch_status": (C.c_int, [C.c_void_p, _SZ, C.POINTER(_SZ), _D]),
    "qsm_tv_batch_destroy": (None, [C.c_void_p]),
    "qsm_tv_batch_profile": (C.c_int, [C.c_void_p, _SZ, _D, _SZ]),
    "qsm_tkd_params_default": (None, [C.POINTER(_CTkdParams)]),
    "qsm_tikhonov_params_default": (None, [C.POINTER(_CTikhonovParams)]),
    "qsm_rts_params_default": (None, [C.POINTER(_CRtsParams)]),
    "qsm_sharp_params_default": (None, [C.POINTER(_CSharpParams)]),
    "qsm_inversion_config_default": (None, [C.POINTER(_CInversionConfig)]),
    "qsm_bg_removal_config_default": (None, [C.POINTER(_CBgRemovalConfig)]),
    "qsm_tkd": (C.c_int, [_D, _U8, C.POINTER(_CGrid), _D, C.POINTER(_CTkdParams), _D]),
    "qsm_tsvd": (C.c_int, [_D, _U8, C.POINTER(_CGrid), _D, C.POINTER(_CTkdParams), _D]),
    "qsm_tikhonov": (C.c_int, [_D, _U8, C.POINTER(_CGrid), _D, C.POINTER(_CTikhonovParams), _D]),
    "qsm_rts": (C.c_int, [_D, _U8, C.POINTER(_CGrid), _D, C.POINTER(_CRtsParams), PROGRESS_FN,
                          C.c_void_p, _D]),
    "qsm_rts_batch_create": (C.c_int, [C.POINTER(_CTvVolume), _SZ, C.POINTER(_CRtsParams),
                                       C.c_void_p, C.POINTER(C.c_void_p)]),
    "qsm_sharp": (C.c_int, [_D, _U8, C.POINTER(_CGrid), C.POINTER(_CSharpParams), _D, _U8]),
    "qsm_laplacian_unwrap": (C.c_int, [_D, _U8, C.POINTER(_CGrid), _D]),
    "qsm_nltv_params_default": (None, [C.POINTER(_CNltvParams)]),
    "qsm_nltv": (C.c_int, [_D, _U8, C.POINTER(_CGrid), _D, C.POINTER(_CNltvParams), PROGRESS_FN,
                           C.c_void_p, _D]),
    "qsm_run_dipole_inversion_cfg": (C.c_int, [_D, _SZ, _U8, _SZ, C.POINTER(_CScanMetadata),
                                               C.POINTER(_CInversionConfig), PROGRESS_FN,
                                               C.c_void_p, _D]),
    "qsm_run_bg_removal_cfg": (C.c_int, [_D, _SZ, _U8, _SZ, C.POINTER(_CScanMetadata),
                                         C.POINTER(_CBgRemovalConfig), PROGRESS_FN, C.c_void_p,
                                         _D, _U8]),
    "qsm_nifti_is_gzip": (C.c_int, [C.c_void_p, _SZ]),
    "qsm_nifti_load": (C.c_int, [C.c_void_p, _SZ, C.POINTER(_CNiftiInfo), _D, _SZ]),
    "qsm_nifti_save": (C.c_int, [_D, _SZ, _SZ, _SZ, _D, _D, C.c_int, C.POINTER(C.c_void_p),
                                 C.POINTER(_SZ)]),
    "qsm_nifti_free": (None, [C.c_void_p]),
    "qsm_nifti_read_file": (C.c_int, [C.c_char_p, C.POINTER(_CNiftiInfo), _D, _SZ]),
    "qsm_nifti_write_file": (C.c_int, [C.c_char_p, _D, _SZ, _SZ, _SZ, _D, _D]),
    "qsm_vsharp_batch_create": (C.c_int, [C.POINTER(_CVsharpVolume), _SZ, _D, _SZ, C.c_double,
                                          C.c_void_p, C.POINTER(C.c_void_p)]),
    "qsm_vsharp_batch_run": (C.c_int, [C.c_void_p, PROGRESS_FN, C.c_void_p]),
    "qsm_vsharp_batch_destroy": (None, [C.c_void_p]),
    "qsm_tv_unrolled_create": (C.c_int, [C.POINTER(_CTvVolume), _SZ, C.POINTER(_CTvParams),
                                         C.c_void_p, C.POINTER(C.c_void_p)]),
    "qsm_tv_unrolled_forward": (C.c_int, [C.c_void_p]),
    "qsm_tv_unrolled_backward": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p),
                                           C.POINTER(C.c_void_p), _D]),
    "qsm_tv_unrolled_destroy": (None, [C.c_void_p]),
    "qsm_dipole_forward": (C.c_int, [C.c_void_p, C.POINTER(_CGrid), _D, C.c_void_p, C.c_void_p]),
    "qsm_dipole_adjoint": (C.c_int, [C.c_void_p, C.POINTER(_CGrid), _D, C.c_void_p, C.c_void_p]),
    "qsm_dipole_forward_noalloc": (C.c_int, [C.c_void_p, C.POINTER(_CGrid), _D, C.c_void_p,
                                             C.c_void_p]),
    "qsm_device_alloc": (C.c_int, [_SZ, C.POINTER(C.c_void_p)]),
    "qsm_device_free": (C.c_int, [C.c_void_p]),
    "qsm_memcpy_htod": (C.c_int, [C.c_void_p, C.c_void_p, _SZ]),
    "qsm_memcpy_dtoh": (C.c_int, [C.c_void_p, C.c_void_p, _SZ]),
    "qsm_memcpy_dtod": (C.c_int, [C.c_void_p, C.c_void_p, _SZ]),
    "qsm_device_synchronize": (C.c_int, []),
    "qsm_hbm_copy_probe": (C.c_int, [_SZ, C.c_int, _D]),
    "qsm_hbm_stream_probe": (C.c_int, [_SZ, C.c_int, C.c_int, C.c_int, _D]),
}
for _name, (_res, _args) in _SIGS.items():
    _fn = getattr(_lib, _name)
    _fn.restype = _res
    _fn.argtypes = _args


def c_library() -> C.CDLL:
    """The loaded libqsm_gpu.so."""
    return _lib


class QsmError(RuntimeError):
    """A non-zero status from the C ABI."""

    def __init__(self, code: int, message: str):
        super().__init__(f"[{code}] {message}")
        self.code = code
        self.message = message


def _check(rc: int):
    if rc != QSM_OK:
        raise QsmError(rc, _lib.qsm_last_error_message().decode())


def backend_info() -> str:
    buf = C.create_string_buffer(512)
    _check(_lib.qsm_backend_info(buf, 512))
    return buf.value.decode()


def fft_length_supported(n: int) -> bool:
    return bool(_lib.qsm_fft_length_supported(n))


# ------------------------------------------------------------------ types --
@dataclass(frozen=True)
class Grid:
    """src/grid.rs:9-44"""
    dims: tuple
    voxel_size: tuple = (1.0, 1.0, 1.0)

    @staticmethod
    def new(nx, ny, nz, vsx, vsy, vsz) -> "Grid":
        return Grid((int(nx), int(ny), int(nz)), (float(vsx), float(vsy), float(vsz)))

    def nx(self): return self.dims[0]
    def ny(self): return self.dims[1]
    def nz(self): return self.dims[2]
    def vsx(self): return self.voxel_size[0]
    def vsy(self): return self.voxel_size[1]
    def vsz(self): return self.voxel_size[2]

    def n_total(self) -> int:
        return self.dims[0] * self.dims[1] * self.dims[2]

    def _c(self) -> _CGrid:
        return _CGrid(*[int(d) for d in self.dims], *[float(v) for v in self.voxel_size])


@dataclass
class TvParams:
    """src/inversion/tv.rs:21-43 (``lambda`` is spelled ``lambda_``)."""
    lambda_: float = 2e-4
    rho: float = 2e-2
    tol: float = 1e-3
    max_iter: int = 250

    @staticmethod
    def default() -> "TvParams":
        c = _CTvParams()
        _lib.qsm_tv_params_default(C.byref(c))
        return TvParams(c.lambda_, c.rho, c.tol, int(c.max_iter))

    def _c(self) -> _CTvParams:
        return _CTvParams(float(self.lambda_), float(self.rho), float(self.tol), int(self.max_iter))


@dataclass
class VsharpParams:
    """src/bgremove/vsharp.rs:20-39"""
    threshold: float = 0.05
    max_radius: float = 12.0
    min_radius: float = 1.0

    @staticmethod
    def default() -> "VsharpParams":
        c = _CVsharpParams()
        _lib.qsm_vsharp_params_default(C.byref(c))
        return VsharpParams(c.threshold, c.max_radius, c.min_radius)

    def _c(self) -> _CVsharpParams:
        return _CVsharpParams(float(self.threshold), float(self.max_radius), float(self.min_radius))


@dataclass
class TkdParams:
    """src/inversion/tkd.rs:21-33 (also the tsvd's parameters)"""
    threshold: float = 0.15

    def _c(self):
        return _CTkdParams(float(self.threshold))


class Regularization(enum.IntEnum):
    """src/inversion/tikhonov.rs:24-33"""
    Identity = 0
    Gradient = 1
    Laplacian = 2


@dataclass
class TikhonovParams:
    """src/inversion/tikhonov.rs:36-50 (``lambda`` is spelled ``lambda_``)"""
    lambda_: float = 0.01
    reg: Regularization = Regularization.Identity

    def _c(self):
        return _CTikhonovParams(float(self.lambda_), int(self.reg))


@dataclass
class RtsParams:
    """src/inversion/rts.rs:24-52"""
    delta: float = 0.15
    mu: float = 1e5
    rho: float = 10.0
    tol: float = 1e-2
    max_iter: int = 20
    lsmr_iter: int = 4

    def _c(self):
        return _CRtsParams(float(self.delta), float(self.mu), float(self.rho), float(self.tol),
                           int(self.max_iter), int(self.lsmr_iter))


@dataclass
class NltvParams:
    """src/inversion/nltv.rs:24-49 (``lambda`` is spelled ``lambda_``)"""
    lambda_: float = 1e-3
    mu: float = 1.0
    tol: float = 1e-3
    max_iter: int = 250
    newton_iter: int = 10

    def _c(self):
        return _CNltvParams(float(self.lambda_), float(self.mu), float(self.tol), int(self.max_iter),
                            int(self.newton_iter))


@dataclass
class SharpParams:
    """src/bgremove/sharp.rs:27-45"""
    threshold: float = 0.05
    radius: float = 6.0

    def _c(self):
        return _CSharpParams(float(self.threshold), float(self.radius))


def _flat64(a) -> np.ndarray:
    a = np.asarray(a)
    if a.ndim == 3:
        a = a.ravel(order="F")
    return np.ascontiguousarray(a.ravel(), dtype=np.float64)


def _flatu8(a) -> np.ndarray:
    a = np.asarray(a)
    if a.ndim == 3:
        a = a.ravel(order="F")
    return np.ascontiguousarray(a.ravel(), dtype=np.uint8)


def _progress(cb: Optional[Callable[[int, int], None]]):
    if cb is None:
        return PROGRESS_FN(0)
    return PROGRESS_FN(lambda cur, tot, _u: cb(int(cur), int(tot)))


def _check_len(grid: Grid, *arrays):
    n = grid.n_total()
    for a in arrays:
        if a.size != n:
            # the reference panics on short slices (index out of bounds)
            raise ValueError(f"array of length {a.size} does not match grid of {n} voxels")


# ------------------------------------------------------------- algorithms --
def tv_admm(local_field, mask, grid: Grid, bdir, params: TvParams,
            progress: Optional[Callable[[int, int], None]] = None,
            return_info: bool = False):
    """``inversion::tv_admm`` (src/inversion/tv.rs:63-97) on the GPU.

    FP32 FFTs and ADMM state, FP64 dipole/inverse-operator construction and
    FP64 convergence norms.  Returns chi (float64, N); with ``return_info``
    also (iterations run, last relative change)."""
    f = _flat64(local_field)
    m = _flatu8(mask)
    _check_len(grid, f, m)
    out = np.empty(f.size)
    b = np.ascontiguousarray(np.asarray(bdir, dtype=np.float64))
    g, p = grid._c(), params._c()
    it, rel = _SZ(0), C.c_double(0.0)
    cb = _progress(progress)
    _check(_lib.qsm_tv_admm_ex(f.ctypes.data_as(_D), m.ctypes.data_as(_U8), C.byref(g),
                               b.ctypes.data_as(_D), C.byref(p), cb, None,
                               out.ctypes.data_as(_D), C.byref(it), C.byref(rel)))
    if return_info:
        return out, int(it.value), float(rel.value)
    return out


def vsharp(field, mask, grid: Grid, params: VsharpParams,
           progress: Optional[Callable[[int, int], None]] = None):
    """``bgremove::vsharp`` (src/bgremove/vsharp.rs:55-64) on the GPU.
    Returns (local_field float64 N, eroded_mask uint8 N)."""
    f = _flat64(field)
    m = _flatu8(mask)
    _check_len(grid, f, m)
    local = np.empty(f.size)
    mout = np.empty(f.size, dtype=np.uint8)
    g, p = grid._c(), params._c()
    cb = _progress(progress)
    _check(_lib.qsm_vsharp(f.ctypes.data_as(_D), m.ctypes.data_as(_U8), C.byref(g), C.byref(p), cb,
                           None, local.ctypes.data_as(_D), mout.ctypes.data_as(_U8)))
    return local, mout


def vsharp_with_radii(field, mask, grid: Grid, radii: Sequence[float], threshold: float,
                      progress: Optional[Callable[[int, int], None]] = None):
    """``vsharp_with_radii`` (src/bgremove/vsharp.rs:88-199) on the GPU."""
    f = _flat64(field)
    m = _flatu8(mask)
    _check_len(grid, f, m)
    r = np.ascontiguousarray(np.asarray(radii, dtype=np.float64).ravel())
    local = np.empty(f.size)
    mout = np.empty(f.size, dtype=np.uint8)
    g = grid._c()
    cb = _progress(progress)
    rp = r.ctypes.data_as(_D) if r.size else None
    _check(_lib.qsm_vsharp_with_radii(f.ctypes.data_as(_D), m.ctypes.data_as(_U8), C.byref(g), rp,
                                      r.size, float(threshold), cb, None,
                                      local.ctypes.data_as(_D), mout.ctypes.data_as(_U8)))
    return local, mout


def _filter(fn, local_field, mask, grid, bdir, cparams):
    f, m = _flat64(local_field), _flatu8(mask)
    _check_len(grid, f, m)
    out = np.empty(f.size)
    g = grid._c()
    b = np.ascontiguousarray(np.asarray(bdir, dtype=np.float64))
    _check(fn(f.ctypes.data_as(_D), m.ctypes.data_as(_U8), C.byref(g), b.ctypes.data_as(_D),
              C.byref(cparams), out.ctypes.data_as(_D)))
    return out


def tkd(local_field, mask, grid: Grid, bdir, params: TkdParams = None):
    """src/inversion/tkd.rs:48-99"""
    return _filter(_lib.qsm_tkd, local_field, mask, grid, bdir, (params or TkdParams())._c())


def tsvd(local_field, mask, grid: Grid, bdir, params: TkdParams = None):
    """The tsvd of src/inversion/tkd.rs (zero where |D| <= threshold)."""
    return _filter(_lib.qsm_tsvd, local_field, mask, grid, bdir, (params or TkdParams())._c())


def tikhonov(local_field, mask, grid: Grid, bdir, params: TikhonovParams = None):
    """src/inversion/tikhonov.rs:62-131"""
    return _filter(_lib.qsm_tikhonov, local_field, mask, grid, bdir,
                   (params or TikhonovParams())._c())


def rts(local_field, mask, grid: Grid, bdir, params: RtsParams = None,
        progress: Optional[Callable[[int, int], None]] = None):
    """src/inversion/rts.rs:71-212"""
    f, m = _flat64(local_field), _flatu8(mask)
    _check_len(grid, f, m)
    out = np.empty(f.size)
    g = grid._c()
    b = np.ascontiguousarray(np.asarray(bdir, dtype=np.float64))
    p = (params or RtsParams())._c()
    cb = _progress(progress)
    _check(_lib.qsm_rts(f.ctypes.data_as(_D), m.ctypes.data_as(_U8), C.byref(g),
                        b.ctypes.data_as(_D), C.byref(p), cb, None, out.ctypes.data_as(_D)))
    return out


def laplacian_unwrap(phase, mask, grid: Grid):
    """src/unwrap/laplacian.rs:135-161 (wrapped phase in radians)"""
    p, m = _flat64(phase), _flatu8(mask)
    _check_len(grid, p, m)
    out = np.empty(p.size)
    g = grid._c()
    _check(_lib.qsm_laplacian_unwrap(p.ctypes.data_as(_D), m.ctypes.data_as(_U8), C.byref(g),
                                     out.ctypes.data_as(_D)))
    return out


def nltv(local_field, mask, grid: Grid, bdir, params: NltvParams = None,
         progress: Optional[Callable[[int, int], None]] = None):
    """src/inversion/nltv.rs:63-133"""
    f, m = _flat64(local_field), _flatu8(mask)
    _check_len(grid, f, m)
    out = np.empty(f.size)
    g = grid._c()
    b = np.ascontiguousarray(np.asarray(bdir, dtype=np.float64))
    p = (params or NltvParams())._c()
    cb = _progress(progress)
    _check(_lib.qsm_nltv(f.ctypes.data_as(_D), m.ctypes.data_as(_U8), C.byref(g),
                         b.ctypes.data_as(_D), C.byref(p), cb, None, out.ctypes.data_as(_D)))
    return out


def sharp(field, mask, grid: Grid, params: SharpParams = None):
    """src/bgremove/sharp.rs:48-56 -> (local_field, eroded_mask)"""
    f, m = _flat64(field), _flatu8(mask)
    _check_len(grid, f, m)
    local = np.empty(f.size)
    mout = np.empty(f.size, dtype=np.uint8)
    g = grid._c()
    p = (params or SharpParams())._c()
    _check(_lib.qsm_sharp(f.ctypes.data_as(_D), m.ctypes.data_as(_U8), C.byref(g), C.byref(p),
                          local.ctypes.data_as(_D), mout.ctypes.data_as(_U8)))
    return local, mout


# --------------------------------------------------------------- pipeline --
class InversionAlgorithm(enum.IntEnum):
    """src/pipeline/config.rs:47-58"""
    Tkd = 0
    Tsvd = 1
    Tikhonov = 2
    Tv = 3
    Rts = 4
    Nltv = 5
    Medi = 6
    Ilsqr = 7
    Tgv = 8
    Qsmart = 9


class BgRemovalAlgorithm(enum.IntEnum):
    """src/pipeline/config.rs:33-43"""
    Vsharp = 0
    Pdf = 1
    Lbv = 2
    Ismv = 3
    Sharp = 4
    Resharp = 5
    Harperella = 6
    Iharperella = 7


@dataclass
class ScanMetadata:
    """src/pipeline/config.rs:300-318"""
    dims: tuple
    voxel_size: tuple
    echo_times: list = field(default_factory=lambda: [0.005])
    field_strength: float = 3.0
    b0_direction: tuple = (0.0, 0.0, 1.0)

    def grid(self) -> Grid:
        return Grid(tuple(self.dims), tuple(self.voxel_size))

    def _c(self):
        et = np.ascontiguousarray(np.asarray(self.echo_times, dtype=np.float64))
        c = _CScanMetadata(*[int(d) for d in self.dims], *[float(v) for v in self.voxel_size],
                           et.ctypes.data_as(_D) if et.size else None, et.size,
                           float(self.field_strength),
                           (C.c_double * 3)(*[float(v) for v in self.b0_direction]))
        return c, et


@dataclass
class InversionConfig:
    """src/pipeline/config.rs:233-264 (the arms with a GPU backend: Tkd, Tsvd,
    Tikhonov, Tv, Rts, Nltv)."""
    algorithm: InversionAlgorithm = InversionAlgorithm.Rts
    tkd: TkdParams = field(default_factory=TkdParams)
    tsvd: TkdParams = field(default_factory=TkdParams)
    tikhonov: TikhonovParams = field(default_factory=TikhonovParams)
    tv: TvParams = field(default_factory=TvParams)
    rts: RtsParams = field(default_factory=RtsParams)
    nltv: NltvParams = field(default_factory=NltvParams)

    def _c(self):
        return _CInversionConfig(int(self.algorithm), self.tkd._c(), self.tsvd._c(),
                                 self.tikhonov._c(), self.tv._c(), self.rts._c(), self.nltv._c())


@dataclass
class BgRemovalConfig:
    """src/pipeline/config.rs:202-230 (the arms with a GPU backend: Vsharp,
    Sharp)."""
    algorithm: BgRemovalAlgorithm = BgRemovalAlgorithm.Vsharp
    vsharp: VsharpParams = field(default_factory=VsharpParams)
    sharp: SharpParams = field(default_factory=SharpParams)

    def _c(self):
        return _CBgRemovalConfig(int(self.algorithm), self.vsharp._c(), self.sharp._c())


@dataclass
class BgRemovalResult:
    """src/pipeline/config.rs:336-342"""
    local_field_ppm: np.ndarray
    eroded_mask: np.ndarray


class PipelineError(Exception):
    """src/pipeline/config.rs:344-368: kind is one of InvalidConfig,
    InvalidInput, AlgorithmError, DimensionMismatch."""

    def __init__(self, kind: str, message: str, expected: int = None, got: int = None):
        super().__init__(f"{kind}: {message}")
        self.kind = kind
        self.expected = expected
        self.got = got


def _pipeline_error(e: QsmError, expected=None, got=None) -> PipelineError:
    kind = {QSM_EDIM: "DimensionMismatch", QSM_ECONFIG: "InvalidConfig",
            QSM_EPARAM: "InvalidInput"}.get(e.code, "AlgorithmError")
    return PipelineError(kind, e.message, expected, got)


def run_dipole_inversion(local_field_ppm, mask, metadata: ScanMetadata, config: InversionConfig,
                         magnitude=None, progress: Optional[Callable[[int, int], None]] = None):
    """src/pipeline/inversion.rs:28-114 -- length check, then the Tkd / Tsvd /
    Tikhonov / Tv / Rts arm on the GPU; Tgv, Qsmart and arms without a GPU
    backend raise PipelineError('InvalidConfig')."""
    f = _flat64(local_field_ppm)
    m = _flatu8(mask)
    md, _keep = metadata._c()
    c = config._c()
    out = np.empty(f.size)
    cb = _progress(progress)
    rc = _lib.qsm_run_dipole_inversion_cfg(f.ctypes.data_as(_D), f.size, m.ctypes.data_as(_U8),
                                           m.size, C.byref(md), C.byref(c), cb, None,
                                           out.ctypes.data_as(_D))
    if rc != QSM_OK:
        n = int(np.prod(metadata.dims))
        raise _pipeline_error(QsmError(rc, _lib.qsm_last_error_message().decode()), n, f.size)
    return out


def run_bg_removal(field_ppm, mask, metadata: ScanMetadata, config: BgRemovalConfig,
                   progress: Optional[Callable[[int, int], None]] = None) -> BgRemovalResult:
    """src/pipeline/bg_removal.rs:20-92 -- length check, then the Vsharp or
    Sharp arm on the GPU; other arms raise PipelineError('InvalidConfig')."""
    f = _flat64(field_ppm)
    m = _flatu8(mask)
    md, _keep = metadata._c()
    c = config._c()
    local = np.empty(f.size)
    mout = np.empty(f.size, dtype=np.uint8)
    cb = _progress(progress)
    rc = _lib.qsm_run_bg_removal_cfg(f.ctypes.data_as(_D), f.size, m.ctypes.data_as(_U8), m.size,
                                     C.byref(md), C.byref(c), cb, None,
                                     local.ctypes.data_as(_D), mout.ctypes.data_as(_U8))
    if rc != QSM_OK:
        n = int(np.prod(metadata.dims))
        raise _pipeline_error(QsmError(rc, _lib.qsm_last_error_message().decode()), n, f.size)
    return BgRemovalResult(local, mout)


# ------------------------------------------------------------------ NIfTI --
@dataclass
class NiftiData:
    """src/io.rs:13-26 (plus the volume count nt and the on-disk datatype);
    data is float64, Fortran order, volume after volume."""
    data: np.ndarray
    dims: tuple
    voxel_size: tuple
    affine: np.ndarray
    scl_slope: float
    scl_inter: float
    nt: int = 1
    datatype: int = 16


def _nifti_from(info, data):
    return NiftiData(data, (int(info.nx), int(info.ny), int(info.nz)),
                     (info.vsx, info.vsy, info.vsz), np.array(list(info.affine)),
                     info.scl_slope, info.scl_inter, int(info.nt), int(info.datatype))


def load_nifti(raw: bytes) -> NiftiData:
    """src/io.rs:62-158 / 160-229 (gzip detected by its magic bytes)"""
    b = np.frombuffer(bytes(raw), dtype=np.uint8)
    info = _CNiftiInfo()
    _check(_lib.qsm_nifti_load(b.ctypes.data_as(C.c_void_p), b.size, C.byref(info), None, 0))
    n = info.nx * info.ny * info.nz * info.nt
    out = np.empty(n)
    _check(_lib.qsm_nifti_load(b.ctypes.data_as(C.c_void_p), b.size, C.byref(info),
                               out.ctypes.data_as(_D), n))
    return _nifti_from(info, out)


def save_nifti(data, dims, voxel_size, affine, gzip: bool = False) -> bytes:
    """src/io.rs:262-357 (FLOAT32 data, sform from the affine)"""
    d = _flat64(data)
    vs = np.ascontiguousarray(np.asarray(voxel_size, dtype=np.float64))
    af = np.ascontiguousarray(np.asarray(affine, dtype=np.float64).ravel())
    p, n = C.c_void_p(), _SZ(0)
    _check(_lib.qsm_nifti_save(d.ctypes.data_as(_D), int(dims[0]), int(dims[1]), int(dims[2]),
                               vs.ctypes.data_as(_D), af.ctypes.data_as(_D), 1 if gzip else 0,
                               C.byref(p), C.byref(n)))
    try:
        return C.string_at(p, n.value)
    finally:
        _lib.qsm_nifti_free(p)


def read_nifti_file(path) -> NiftiData:
    """src/io.rs:423-427"""
    info = _CNiftiInfo()
    pb = str(path).encode()
    _check(_lib.qsm_nifti_read_file(pb, C.byref(info), None, 0))
    n = info.nx * info.ny * info.nz * info.nt
    out = np.empty(n)
    _check(_lib.qsm_nifti_read_file(pb, C.byref(info), out.ctypes.data_as(_D), n))
    return _nifti_from(info, out)


def save_nifti_to_file(path, data, dims, voxel_size, affine):
    """src/io.rs:429-452 (.nii.gz compressed by suffix)"""
    d = _flat64(data)
    vs = np.ascontiguousarray(np.asarray(voxel_size, dtype=np.float64))
    af = np.ascontiguousarray(np.asarray(affine, dtype=np.float64).ravel())
    _check(_lib.qsm_nifti_write_file(str(path).encode(), d.ctypes.data_as(_D), int(dims[0]),
                                     int(dims[1]), int(dims[2]), vs.ctypes.data_as(_D),
                                     af.ctypes.data_as(_D)))


# ---------------------------------------------------------- device memory --
class DeviceArray:
    """A device allocation owned by this library's HIP runtime, holding a
    1-D numpy dtype.  ``ptr`` is the raw device address."""

    def __init__(self, n: int, dtype):
        self.dtype = np.dtype(dtype)
        self.n = int(n)
        p = C.c_void_p()
        _check(_lib.qsm_device_alloc(self.n * self.dtype.itemsize, C.byref(p)))
        self.ptr = p.value

    @staticmethod
    def from_numpy(a) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = DeviceArray(a.size, a.dtype)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.size == self.n
        _check(_lib.qsm_memcpy_htod(C.c_void_p(self.ptr), a.ctypes.data_as(C.c_void_p), a.nbytes))

    def numpy(self) -> np.ndarray:
        out = np.empty(self.n, dtype=self.dtype)
        _check(_lib.qsm_memcpy_dtoh(out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr), out.nbytes))
        return out

    def free(self):
        if getattr(self, "ptr", None):
            _lib.qsm_device_free(C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def hbm_copy_probe(nbytes: int = 1 << 30, reps: int = 5) -> float:
    """GB/s of a plain one-element-per-thread device copy (read + write)."""
    g = C.c_double(0.0)
    _check(_lib.qsm_hbm_copy_probe(int(nbytes), int(reps), C.byref(g)))
    return g.value


def hbm_stream_probe(n: int, reads: int, writes: int, reps: int = 5) -> float:
    """GB/s of a pass shape: `reads` arrays of n floats read and `writes`
    arrays written element by element (1..4 each), all bytes counted."""
    g = C.c_double(0.0)
    _check(_lib.qsm_hbm_stream_probe(int(n), int(reads), int(writes), int(reps), C.byref(g)))
    return g.value


FFT_SCHEDULES = {0: "two-stage", 1: "staged", -1: "untuned"}


def set_fft_schedule(mode: int):
    """0: fixed per-length table (default, the same bits in every process);
    1: always the two-stage column FFT; 2: always the LDS-staged one;
    3: tuned per axis length by timing both on the first batch using it."""
    _check(_lib.qsm_set_fft_schedule(int(mode)))


def fft_schedule(axis: str, n: int) -> str:
    """Schedule used by the 'y' or 'z' pass of length n."""
    out = C.c_int(0)
    _check(_lib.qsm_fft_schedule({"y": 1, "z": 2}[axis], int(n), C.byref(out)))
    return FFT_SCHEDULES[out.value]


def synchronize():
    _check(_lib.qsm_device_synchronize())


def copy_device(dst: DeviceArray, src: DeviceArray):
    """Device-to-device copy of src into dst (same byte size)."""
    nbytes = src.n * src.dtype.itemsize
    assert dst.n * dst.dtype.itemsize == nbytes
    _check(_lib.qsm_memcpy_dtod(C.c_void_p(dst.ptr), C.c_void_p(src.ptr), nbytes))


# ------------------------------------------------ device-resident batches --
class TvBatch:
    """Batched TV-ADMM over device-resident volumes (``qsm_tv_batch_*``).

    ``volumes`` is a list of (grid, bdir, d_field_ptr, d_mask_ptr, d_chi_ptr)
    with raw device addresses (f64 field, u8 mask, f64 output); the caller
    keeps the memory alive.  ``stream`` is a raw hipStream_t or None."""

    def __init__(self, volumes, params, stream: Optional[int] = None):
        """params: TvParams (TV-ADMM) or RtsParams (the same batch running RTS)."""
        arr = (_CTvVolume * len(volumes))()
        for i, (grid, bdir, pf, pm, pc) in enumerate(volumes):
            arr[i].grid = grid._c()
            arr[i].bdir = (C.c_double * 3)(*[float(v) for v in bdir])
            arr[i].d_field, arr[i].d_mask, arr[i].d_chi = int(pf), int(pm), int(pc)
        self._p = params._c()
        h = C.c_void_p()
        create = _lib.qsm_rts_batch_create if isinstance(params, RtsParams) else _lib.qsm_tv_batch_create
        _check(create(arr, len(volumes), C.byref(self._p),
                      C.c_void_p(stream) if stream else None, C.byref(h)))
        self._h = h
        self.n = len(volumes)

    def run(self, progress: Optional[Callable[[int, int], None]] = None):
        cb = _progress(progress)
        _check(_lib.qsm_tv_batch_run(self._h, cb, None))

    def set_graphs(self, enable: bool):
        """Iterations after the first as HIP graph replays (default) or as
        eager launches; the results are identical."""
        _check(_lib.qsm_tv_batch_set_graphs(self._h, 1 if enable else 0))

    def status(self, v: int):
        it, rel = _SZ(0), C.c_double(0.0)
        _check(_lib.qsm_tv_batch_status(self._h, v, C.byref(it), C.byref(rel)))
        return int(it.value), float(rel.value)

    def profile(self, iters: int = 20):
        """Mean per-launch milliseconds of the five kernels of one iteration
        (see qsm_tv_batch_profile in include/qsm_gpu.h)."""
        out = np.zeros(5)
        _check(_lib.qsm_tv_batch_profile(self._h, iters, out.ctypes.data_as(_D), 5))
        return out

    def close(self):
        if getattr(self, "_h", None):
            _lib.qsm_tv_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class VsharpBatch:
    """Batched V-SHARP over device-resident volumes (``qsm_vsharp_batch_*``).

    ``volumes`` is a list of (grid, d_field_ptr, d_mask_ptr, d_local_ptr,
    d_mask_out_ptr): f64 total field and u8 mask in, f64 local field and u8
    eroded mask out, raw device addresses kept alive by the caller.  All
    volumes share ``radii`` (mm, any order; one radius gives SHARP) and
    ``threshold``.  The outputs can feed a TvBatch directly (the
    run_bg_removal -> run_dipole_inversion chain without host copies)."""

    def __init__(self, volumes, radii: Sequence[float], threshold: float,
                 stream: Optional[int] = None):
        arr = (_CVsharpVolume * len(volumes))()
        for i, (grid, pf, pm, pl, pmo) in enumerate(volumes):
            arr[i].grid = grid._c()
            arr[i].d_field, arr[i].d_mask = int(pf), int(pm)
            arr[i].d_local, arr[i].d_mask_out = int(pl), int(pmo)
        r = np.ascontiguousarray(np.asarray(radii, dtype=np.float64))
        h = C.c_void_p()
        _check(_lib.qsm_vsharp_batch_create(arr, len(volumes), r.ctypes.data_as(_D), r.size,
                                            float(threshold),
                                            C.c_void_p(stream) if stream else None, C.byref(h)))
        self._h = h
        self.n = len(volumes)

    def run(self, progress: Optional[Callable[[int, int], None]] = None):
        cb = _progress(progress)
        _check(_lib.qsm_vsharp_batch_run(self._h, cb, None))

    def close(self):
        if getattr(self, "_h", None):
            _lib.qsm_vsharp_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _tv_volume_array(volumes):
    arr = (_CTvVolume * len(volumes))()
    for i, (grid, bdir, pf, pm, pc) in enumerate(volumes):
        arr[i].grid = grid._c()
        arr[i].bdir = (C.c_double * 3)(*[float(v) for v in bdir])
        arr[i].d_field, arr[i].d_mask, arr[i].d_chi = int(pf), int(pm), int(pc)
    return arr


class TvUnrolled:
    """K = params.max_iter unrolled TV-ADMM iterations (tol ignored) over
    device-resident volumes, with the vector-Jacobian product of the masked
    output chi = mask * x_K w.r.t. the local field and lambda
    (``qsm_tv_unrolled_*``).  ``volumes`` as for TvBatch."""

    def __init__(self, volumes, params: TvParams, stream: Optional[int] = None):
        arr = _tv_volume_array(volumes)
        self._p = params._c()
        h = C.c_void_p()
        _check(_lib.qsm_tv_unrolled_create(arr, len(volumes), C.byref(self._p),
                                           C.c_void_p(stream) if stream else None, C.byref(h)))
        self._h = h
        self.n = len(volumes)

    def forward(self):
        _check(_lib.qsm_tv_unrolled_forward(self._h))

    def backward(self, chi_bar_ptrs, field_bar_ptrs):
        """chi_bar_ptrs / field_bar_ptrs: device addresses (N f64 each) per
        volume; returns dL/dlambda per volume."""
        cb = (C.c_void_p * self.n)(*[int(p) for p in chi_bar_ptrs])
        fb = (C.c_void_p * self.n)(*[int(p) for p in field_bar_ptrs])
        lb = np.zeros(self.n)
        _check(_lib.qsm_tv_unrolled_backward(self._h, cb, fb, lb.ctypes.data_as(_D)))
        return lb

    def close(self):
        if getattr(self, "_h", None):
            _lib.qsm_tv_unrolled_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def tv_unrolled_vjp(local_field, mask, grid: Grid, bdir, params: TvParams, chi_bar):
    """Host-array convenience over TvUnrolled for one volume: returns
    (chi, field_bar, lambda_bar)."""
    f, m, cbar = _flat64(local_field), _flatu8(mask), _flat64(chi_bar)
    _check_len(grid, f, m, cbar)
    tf, tm = DeviceArray.from_numpy(f), DeviceArray.from_numpy(m)
    tc, tcb = DeviceArray(f.size, np.float64), DeviceArray.from_numpy(cbar)
    tfb = DeviceArray(f.size, np.float64)
    u = TvUnrolled([(grid, bdir, tf.ptr, tm.ptr, tc.ptr)], params)
    try:
        u.forward()
        lb = u.backward([tcb.ptr], [tfb.ptr])
    finally:
        u.close()
    return tc.numpy(), tfb.numpy(), float(lb[0])


class UnrolledTvFunction:
    """Autograd-style wrapper (forward / backward) of the unrolled solver for
    test code: ``y = F.forward(f, lam)``, ``gf, glam = F.backward(ybar)``.
    A torch.autograd.Function is not used because torch's bundled HIP
    runtime and this library's would both be loaded in one process."""

    def __init__(self, mask, grid: Grid, bdir, rho: float, K: int):
        self.mask, self.grid, self.bdir, self.rho, self.K = mask, grid, bdir, rho, K
        self._saved = None

    def forward(self, local_field, lam: float):
        self._saved = (np.array(local_field, dtype=np.float64, copy=True), float(lam))
        p = TvParams(lambda_=lam, rho=self.rho, tol=0.0, max_iter=self.K)
        chi, _, _ = tv_unrolled_vjp(local_field, self.mask, self.grid, self.bdir, p,
                                    np.zeros(np.size(local_field)))
        return chi

    def backward(self, chi_bar):
        f, lam = self._saved
        p = TvParams(lambda_=lam, rho=self.rho, tol=0.0, max_iter=self.K)
        _, fb, lb = tv_unrolled_vjp(f, self.mask, self.grid, self.bdir, p, chi_bar)
        return fb, lb


def dipole_forward(chi, grid: Grid, bdir, adjoint: bool = False):
    """Re IFFT(D FFT(chi)) on the GPU (host arrays in and out)."""
    c = _flat64(chi)
    _check_len(grid, c)
    tin, tout = DeviceArray.from_numpy(c), DeviceArray(c.size, np.float64)
    g = grid._c()
    b = np.ascontiguousarray(np.asarray(bdir, dtype=np.float64))
    fn = _lib.qsm_dipole_adjoint if adjoint else _lib.qsm_dipole_forward
    _check(fn(C.c_void_p(tin.ptr), C.byref(g), b.ctypes.data_as(_D), C.c_void_p(tout.ptr), None))
    return tout.numpy()


def vsharp_radii(params: VsharpParams):
    """The radius list vsharp uses for ``params`` (src/bgremove/vsharp.rs:70-82):
    max_radius down to min_radius in steps of min_radius, or [max_radius]
    when that is empty.  min_radius must be > 0 (the reference would loop)."""
    if not params.min_radius > 0:
        raise ValueError("min_radius must be > 0 (it is the step between radii)")
    out = []
    r = params.max_radius
    while r >= params.min_radius:
        out.append(r)
        r -= params.min_radius
    return out or [params.max_radius]


__all__ = [
    "Grid", "TvParams", "VsharpParams", "tv_admm", "vsharp", "vsharp_with_radii",
    "ScanMetadata", "InversionAlgorithm", "BgRemovalAlgorithm", "InversionConfig",
    "BgRemovalConfig", "BgRemovalResult", "PipelineError", "run_dipole_inversion",
    "run_bg_removal", "TvBatch", "QsmError", "backend_info", "fft_length_supported",
    "c_library", "LIB_PATH", "DeviceArray", "synchronize", "VsharpBatch", "vsharp_radii",
    "TvUnrolled", "tv_unrolled_vjp", "UnrolledTvFunction", "dipole_forward",
    "TkdParams", "Regularization", "TikhonovParams", "RtsParams", "SharpParams", "tkd", "tsvd",
    "tikhonov", "rts", "sharp", "laplacian_unwrap", "NltvParams", "nltv",
    "NiftiData", "load_nifti", "save_nifti", "read_nifti_file", "save_nifti_to_file",
]
