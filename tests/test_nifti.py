"""NIfTI-1 input/output (src/io.rs) through the C ABI: the reference's own
io tests (src/io.rs:454-620) restated, the save layout checked byte for byte
against an independent construction of the reference's header
(src/io.rs:262-337), and loads of headers written here field by field
(datatypes, scaling, byte order, 4-D, sform / no sform)."""
import gzip
import struct

import numpy as np
import pytest

from conftest import load_package


@pytest.fixture(scope="module")
def q():
    return load_package()


def _reference_bytes(data, dims, vs, affine):
    h = bytearray(348)
    struct.pack_into("<i", h, 0, 348)
    struct.pack_into("<8h", h, 40, 3, dims[0], dims[1], dims[2], 1, 1, 1, 1)
    struct.pack_into("<h", h, 70, 16)
    struct.pack_into("<h", h, 72, 32)
    struct.pack_into("<8f", h, 76, 1.0, vs[0], vs[1], vs[2], 1.0, 1.0, 1.0, 1.0)
    struct.pack_into("<f", h, 108, 352.0)
    struct.pack_into("<f", h, 112, 1.0)
    struct.pack_into("<f", h, 116, 0.0)
    struct.pack_into("<h", h, 254, 1)
    for row in range(3):
        struct.pack_into("<4f", h, 280 + 16 * row, *affine[4 * row: 4 * row + 4])
    h[344:348] = b"n+1\0"
    return bytes(h) + b"\0" * 4 + np.asarray(data, dtype="<f4").tobytes()


def _header(dims, dt, bitpix, vs=(1.0, 1.0, 1.0), slope=1.0, inter=0.0, sform=None, endian="<",
            vox_offset=352.0):
    h = bytearray(348)
    struct.pack_into(endian + "i", h, 0, 348)
    d = list(dims) + [1] * (7 - len(dims))
    struct.pack_into(endian + "8h", h, 40, len(dims), *d)
    struct.pack_into(endian + "h", h, 70, dt)
    struct.pack_into(endian + "h", h, 72, bitpix)
    struct.pack_into(endian + "8f", h, 76, 1.0, vs[0], vs[1], vs[2], 1.0, 1.0, 1.0, 1.0)
    struct.pack_into(endian + "f", h, 108, vox_offset)
    struct.pack_into(endian + "f", h, 112, slope)
    struct.pack_into(endian + "f", h, 116, inter)
    if sform is not None:
        struct.pack_into(endian + "h", h, 254, 1)
        for row in range(3):
            struct.pack_into(endian + "4f", h, 280 + 16 * row, *sform[row])
    h[344:348] = b"n+1\0"
    return bytes(h) + b"\0" * (int(vox_offset) - 348)


def test_reference_io_unit_tests(q, tmp_path):
    # test_gzip_detection (io.rs:470-475)
    assert q.c_library().qsm_nifti_is_gzip(bytes([0x1F, 0x8B, 0x00]), 3) == 1
    assert q.c_library().qsm_nifti_is_gzip(bytes([0, 0, 0]), 3) == 0
    assert q.c_library().qsm_nifti_is_gzip(bytes([0x1F]), 1) == 0
    # test_save_nifti_header (io.rs:477-500)
    eye = np.eye(4).ravel()
    b = q.save_nifti(np.zeros(8), (2, 2, 2), (1.0, 1.0, 1.0), eye)
    assert len(b) == 352 + 8 * 4 and b[344:348] == b"n+1\0"
    assert struct.unpack_from("<i", b, 0)[0] == 348
    # test_save_and_read_nifti_roundtrip (io.rs:502-546)
    aff = np.array([1, 0, 0, 10, 0, 2, 0, 20, 0, 0, 3, 30, 0, 0, 0, 1.0])
    data = np.arange(64) * 0.5 + 1.0
    p = tmp_path / "rt.nii"
    q.save_nifti_to_file(p, data, (4, 4, 4), (1.0, 2.0, 3.0), aff)
    ld = q.read_nifti_file(p)
    assert ld.dims == (4, 4, 4) and np.allclose(ld.voxel_size, (1, 2, 3), atol=1e-5)
    assert np.max(np.abs(ld.data - data)) < 0.01
    assert np.array_equal(ld.affine, aff)
    # test_save_and_read_nifti_f32 (io.rs:548-581): f32 values survive exactly
    d32 = (np.arange(64, dtype=np.float32) * np.float32(0.1)).astype(np.float64)
    p2 = tmp_path / "f32.nii"
    q.save_nifti_to_file(p2, d32, (4, 4, 4), (1.5, 1.5, 1.5), np.diag([1.5, 1.5, 1.5, 1.0]).ravel())
    assert np.array_equal(q.read_nifti_file(p2).data, d32)
    # test_save_nifti_gzip (io.rs:583-620)
    p3 = tmp_path / "g.nii.gz"
    q.save_nifti_to_file(p3, np.arange(64.0), (4, 4, 4), (1.0, 1.0, 1.0), eye)
    raw = p3.read_bytes()
    assert raw[:2] == b"\x1f\x8b"
    assert np.array_equal(q.read_nifti_file(p3).data, np.arange(64.0))


def test_save_layout_is_the_references_byte_for_byte(q):
    rng = np.random.default_rng(0)
    dims, vs = (5, 4, 3), (0.9, 1.1, 1.3)
    aff = np.array([0.9, 0.01, 0, -10, 0, 1.1, 0.02, -20, 0, 0, 1.3, -30, 0, 0, 0, 1])
    data = rng.standard_normal(60)
    b = q.save_nifti(data, dims, vs, aff)
    assert b == _reference_bytes(data, dims, vs, aff)
    gz = q.save_nifti(data, dims, vs, aff, gzip=True)
    assert gzip.decompress(gz) == b


@pytest.mark.parametrize("dt,np_t,bitpix", [(2, "u1", 8), (4, "i2", 16), (8, "i4", 32), (16, "f4", 32),
                                            (64, "f8", 64), (256, "i1", 8), (512, "u2", 16)])
def test_load_datatypes_scaling_and_byte_order(q, dt, np_t, bitpix):
    dims = (3, 4, 2)
    raw = (np.arange(24) % 100).astype(np_t)
    for endian in "<>":
        hdr = _header(dims, dt, bitpix, vs=(0.5, 0.75, 2.0), slope=2.5, inter=-1.0, endian=endian)
        b = hdr + raw.astype(endian + np_t).tobytes()
        ld = q.load_nifti(b)
        assert ld.dims == dims and ld.voxel_size == (0.5, 0.75, 2.0)
        assert ld.scl_slope == 2.5 and ld.scl_inter == -1.0 and ld.datatype == dt
        assert np.array_equal(ld.data, raw.astype(np.float64) * 2.5 - 1.0)
        # no sform: affine = diag(voxel size) (io.rs:245-258, test_affine_identity)
        assert np.array_equal(ld.affine, np.diag([0.5, 0.75, 2.0, 1.0]).ravel())
        assert np.array_equal(q.load_nifti(gzip.compress(b)).data, ld.data)


def test_load_4d_sform_and_unscaled(q):
    dims = (2, 3, 2, 3)
    raw = np.arange(36, dtype="<f4") * 0.25
    sform = [(1.0, 0.0, 0.0, 5.0), (0.0, 2.0, 0.0, 6.0), (0.0, 0.0, 3.0, 7.0)]
    b = _header(dims, 16, 32, vs=(1.0, 2.0, 3.0), slope=0.0, sform=sform, vox_offset=400.0) + raw.tobytes()
    ld = q.load_nifti(b)
    assert ld.dims == (2, 3, 2) and ld.nt == 3
    assert ld.scl_slope == 1.0  # 0 on disk: no scaling, reported as 1 (io.rs:101)
    assert np.array_equal(ld.data, raw.astype(np.float64))
    assert np.array_equal(ld.affine.reshape(4, 4)[:3], np.array(sform))
    with pytest.raises(q.QsmError):
        q.load_nifti(b[:300])
