"""qsm_dipole_forward_noalloc (the dipole operator with its workspace in a
fixed device array) against qsm_dipole_forward, both called through the C
ABI: the same kernels, so the same bits, on two calls in a row (the second
with a smaller volume, so the fixed workspace holds the first call's
larger spectrum when the second starts)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

_D = C.POINTER(C.c_double)


def _dipole(qsm, name, chi, dims, vs, bdir):
    lib = qsm.c_library()
    g = qsm.Grid(dims, vs)._c()
    b = np.ascontiguousarray(bdir, dtype=np.float64)
    tin = qsm.DeviceArray.from_numpy(np.ascontiguousarray(chi, dtype=np.float64))
    tout = qsm.DeviceArray(chi.size, np.float64)
    try:
        rc = getattr(lib, name)(C.c_void_p(tin.ptr), C.byref(g), b.ctypes.data_as(_D),
                                C.c_void_p(tout.ptr), None)
        assert rc == 0, (name, rc, lib.qsm_last_error_message().decode())
        return tout.numpy()
    finally:
        tin.free()
        tout.free()


def test_noalloc_equals_dipole_forward_twice(qsm):
    rng = np.random.default_rng(77)
    cases = [((160, 144, 96), (0.9, 0.8, 1.1), (0.2, -0.1, 1.0)),
             ((64, 48, 40), (1.0, 1.0, 1.0), (0.0, 0.0, 1.0))]
    for dims, vs, bdir in cases:
        chi = rng.standard_normal(int(np.prod(dims))) * 0.05
        ref = _dipole(qsm, "qsm_dipole_forward", chi, dims, vs, bdir)
        out = _dipole(qsm, "qsm_dipole_forward_noalloc", chi, dims, vs, bdir)
        print(f"{dims}: max |noalloc - forward| = {np.max(np.abs(out - ref)):.3e}", flush=True)
        assert np.array_equal(out, ref), dims


def test_noalloc_long_and_chirp_lengths(qsm):
    # the fixed twiddle storage at its largest: a 1000-point axis (direct,
    # 8-column z tiles) and a 509-point chirp-z axis (convolution length
    # 1024) in one volume, then a smaller call with a 263-point chirp-z axis
    rng = np.random.default_rng(78)
    cases = [((1000, 12, 509), (1.0, 0.9, 1.1), (0.1, 0.2, 1.0)),
             ((20, 263, 16), (1.0, 1.0, 1.0), (0.0, 0.3, 1.0))]
    for dims, vs, bdir in cases:
        chi = rng.standard_normal(int(np.prod(dims))) * 0.05
        ref = _dipole(qsm, "qsm_dipole_forward", chi, dims, vs, bdir)
        out = _dipole(qsm, "qsm_dipole_forward_noalloc", chi, dims, vs, bdir)
        print(f"{dims}: max |noalloc - forward| = {np.max(np.abs(out - ref)):.3e}", flush=True)
        assert np.array_equal(out, ref), dims


def test_noalloc_rejects_volumes_beyond_its_workspace(qsm):
    # 512 x 512 x 128 needs a 134 MB half spectrum, more than the fixed 72 MiB:
    # refused before any device work, with the message naming the limit
    lib = qsm.c_library()
    g = qsm.Grid((512, 512, 128), (1.0, 1.0, 1.0))._c()
    b = np.array([0.0, 0.0, 1.0])
    tiny = qsm.DeviceArray(16, np.float64)
    try:
        rc = lib.qsm_dipole_forward_noalloc(C.c_void_p(tiny.ptr), C.byref(g), b.ctypes.data_as(_D),
                                            C.c_void_p(tiny.ptr), None)
        msg = lib.qsm_last_error_message().decode()
    finally:
        tiny.free()
    assert rc == qsm.QSM_EPARAM, (rc, msg)
    assert "fixed scratch" in msg, msg
