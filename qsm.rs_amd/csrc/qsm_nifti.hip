// NIfTI-1 input/output for the host side of the library (src/io.rs): the GPU
// path can take scans straight from disk or from in-memory bytes (gzip or
// plain), and write results in the reference's own output layout.  Host code
// only; no device work here.
//
//   load:  sizeof_hdr decides the byte order; dim[0] >= 3; dims dim[1..4]
//          (4th = volumes, 1 if absent); pixdim[1..3]; scl_slope 0 means no
//          scaling (reported as 1); affine = sform rows when sform_code > 0,
//          else diag(pixdim[1..3]) (src/io.rs:231-259); data converted to f64
//          with value * slope + inter, Fortran voxel order, volumes one after
//          another (src/io.rs:62-158, 160-229).
//   save:  348-byte header + 4 zero extension bytes, FLOAT32 data at offset
//          352, dim = {3, nx, ny, nz, 1, 1, 1, 1}, pixdim = {1, vs, 1, 1, 1,
//          1}, scl_slope 1, scl_inter 0, sform_code 1 with the affine's first
//          three rows, magic "n+1" (src/io.rs:262-337); gzip-wrapped for
//          .nii.gz (src/io.rs:339-357, 435-452).
#include <zlib.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/qsm_gpu.h"

namespace qsm {
int fail(int code, const char* fmt, ...);
}

namespace {

using qsm::fail;

bool is_gzip(const uint8_t* b, size_t n) { return n >= 2 && b[0] == 0x1f && b[1] == 0x8b; }

int gunzip(const uint8_t* in, size_t n, std::vector<uint8_t>& out) {
    z_stream zs;
    std::memset(&zs, 0, sizeof zs);
    if (inflateInit2(&zs, 16 + MAX_WBITS) != Z_OK) return fail(QSM_EPARAM, "zlib init failed");
    zs.next_in = const_cast<Bytef*>(in);
    zs.avail_in = (uInt)n;
    out.clear();
    std::vector<uint8_t> chunk(1 << 20);
    int rc;
    do {
        zs.next_out = chunk.data();
        zs.avail_out = (uInt)chunk.size();
        rc = inflate(&zs, Z_NO_FLUSH);
        if (rc != Z_OK && rc != Z_STREAM_END) {
            inflateEnd(&zs);
            return fail(QSM_EPARAM, "gzip stream is corrupt (zlib %d)", rc);
        }
        out.insert(out.end(), chunk.data(), chunk.data() + (chunk.size() - zs.avail_out));
    } while (rc != Z_STREAM_END);
    inflateEnd(&zs);
    return QSM_OK;
}

int gzip(const uint8_t* in, size_t n, std::vector<uint8_t>& out) {
    z_stream zs;
    std::memset(&zs, 0, sizeof zs);
    if (deflateInit2(&zs, Z_DEFAULT_COMPRESSION, Z_DEFLATED, 16 + MAX_WBITS, 8, Z_DEFAULT_STRATEGY) != Z_OK)
        return fail(QSM_EPARAM, "zlib init failed");
    out.resize(deflateBound(&zs, (uLong)n) + 32);
    zs.next_in = const_cast<Bytef*>(in);
    zs.avail_in = (uInt)n;
    zs.next_out = out.data();
    zs.avail_out = (uInt)out.size();
    const int rc = deflate(&zs, Z_FINISH);
    out.resize(out.size() - zs.avail_out);
    deflateEnd(&zs);
    if (rc != Z_STREAM_END) return fail(QSM_EPARAM, "gzip compression failed (zlib %d)", rc);
    return QSM_OK;
}

struct Reader {
    const uint8_t* b;
    bool swap;
    template <class T>
    T get(size_t off) const {
        uint8_t tmp[sizeof(T)];
        std::memcpy(tmp, b + off, sizeof(T));
        if (swap)
            for (size_t i = 0; i < sizeof(T) / 2; ++i) std::swap(tmp[i], tmp[sizeof(T) - 1 - i]);
        T v;
        std::memcpy(&v, tmp, sizeof(T));
        return v;
    }
};

template <class T>
void convert(const Reader& r, size_t off, size_t n, double slope, double inter, bool scale, double* out) {
    for (size_t i = 0; i < n; ++i) {
        const double v = (double)r.get<T>(off + i * sizeof(T));
        out[i] = scale ? v * slope + inter : v;
    }
}

size_t type_size(int dt) {
    switch (dt) {
        case 2: case 256: return 1;     // uint8, int8
        case 4: case 512: return 2;     // int16, uint16
        case 8: case 768: case 16: return 4;  // int32, uint32, float32
        case 64: case 1024: case 1280: return 8;  // float64, int64, uint64
        default: return 0;
    }
}

int parse(const uint8_t* bytes, size_t len, qsm_nifti_info* info, double* data, size_t data_len) {
    if (len < 348) return fail(QSM_EPARAM, "NIfTI data too small (%zu bytes, need at least 348)", len);
    Reader r{bytes, false};
    int32_t sizeof_hdr = r.get<int32_t>(0);
    if (sizeof_hdr != 348) {
        r.swap = true;
        sizeof_hdr = r.get<int32_t>(0);
        if (sizeof_hdr != 348) return fail(QSM_EPARAM, "not a NIfTI-1 header (sizeof_hdr != 348)");
    }
    const int ndim = r.get<int16_t>(40);
    if (ndim < 3) return fail(QSM_EDIM, "Expected at least 3D volume, got %dD", ndim);
    int16_t dim[8];
    for (int i = 0; i < 8; ++i) dim[i] = r.get<int16_t>(40 + 2 * i);
    if (dim[1] <= 0 || dim[2] <= 0 || dim[3] <= 0) return fail(QSM_EDIM, "non-positive dimension");
    const size_t nt = (ndim >= 4 && dim[4] > 0) ? (size_t)dim[4] : 1;
    qsm_nifti_info I;
    std::memset(&I, 0, sizeof I);
    I.nx = (size_t)dim[1];
    I.ny = (size_t)dim[2];
    I.nz = (size_t)dim[3];
    I.nt = nt;
    I.vsx = (double)r.get<float>(76 + 4);
    I.vsy = (double)r.get<float>(76 + 8);
    I.vsz = (double)r.get<float>(76 + 12);
    I.datatype = r.get<int16_t>(70);
    const float slope = r.get<float>(112);
    I.scl_slope = slope == 0.0f ? 1.0 : (double)slope;
    I.scl_inter = (double)r.get<float>(116);
    const int16_t sform_code = r.get<int16_t>(254);
    if (sform_code > 0) {
        for (int row = 0; row < 3; ++row)
            for (int c = 0; c < 4; ++c) I.affine[row * 4 + c] = (double)r.get<float>(280 + row * 16 + c * 4);
    } else {
        I.affine[0] = I.vsx;
        I.affine[5] = I.vsy;
        I.affine[10] = I.vsz;
    }
    I.affine[15] = 1.0;
    const size_t ts = type_size(I.datatype);
    if (ts == 0) return fail(QSM_EUNSUPPORTED, "NIfTI datatype %d not supported", I.datatype);
    const float vox = r.get<float>(108);
    const size_t off = vox >= 352.0f ? (size_t)vox : 352;
    const size_t n = I.nx * I.ny * I.nz * nt;
    if (info) *info = I;
    if (!data) return QSM_OK;
    if (data_len < n) return fail(QSM_EDIM, "data buffer holds %zu values, the file has %zu", data_len, n);
    if (off + n * ts > len) return fail(QSM_EPARAM, "NIfTI data truncated");
    const bool scale = slope != 0.0f;
    const double s = (double)slope, c = I.scl_inter;
    switch (I.datatype) {
        case 2: convert<uint8_t>(r, off, n, s, c, scale, data); break;
        case 256: convert<int8_t>(r, off, n, s, c, scale, data); break;
        case 4: convert<int16_t>(r, off, n, s, c, scale, data); break;
        case 512: convert<uint16_t>(r, off, n, s, c, scale, data); break;
        case 8: convert<int32_t>(r, off, n, s, c, scale, data); break;
        case 768: convert<uint32_t>(r, off, n, s, c, scale, data); break;
        case 16: convert<float>(r, off, n, s, c, scale, data); break;
        case 64: convert<double>(r, off, n, s, c, scale, data); break;
        case 1024: convert<int64_t>(r, off, n, s, c, scale, data); break;
        default: convert<uint64_t>(r, off, n, s, c, scale, data); break;
    }
    return QSM_OK;
}

void put_i16(uint8_t* h, size_t off, int16_t v) { std::memcpy(h + off, &v, 2); }
void put_i32(uint8_t* h, size_t off, int32_t v) { std::memcpy(h + off, &v, 4); }
void put_f32(uint8_t* h, size_t off, float v) { std::memcpy(h + off, &v, 4); }

int build(const double* data, size_t nx, size_t ny, size_t nz, const double vs[3], const double affine[16],
          std::vector<uint8_t>& out) {
    if (!data || !vs || !affine) return fail(QSM_EPARAM, "null argument");
    if (nx > 32767 || ny > 32767 || nz > 32767) return fail(QSM_EDIM, "dimension exceeds the NIfTI-1 limit");
    const size_t n = nx * ny * nz;
    out.assign(352 + n * 4, 0);
    uint8_t* h = out.data();
    put_i32(h, 0, 348);
    const int16_t dim[8] = {3, (int16_t)nx, (int16_t)ny, (int16_t)nz, 1, 1, 1, 1};
    for (int i = 0; i < 8; ++i) put_i16(h, 40 + 2 * i, dim[i]);
    put_i16(h, 70, 16);
    put_i16(h, 72, 32);
    const float pixdim[8] = {1.0f, (float)vs[0], (float)vs[1], (float)vs[2], 1.0f, 1.0f, 1.0f, 1.0f};
    for (int i = 0; i < 8; ++i) put_f32(h, 76 + 4 * i, pixdim[i]);
    put_f32(h, 108, 352.0f);
    put_f32(h, 112, 1.0f);
    put_f32(h, 116, 0.0f);
    put_i16(h, 254, 1);
    for (int row = 0; row < 3; ++row)
        for (int c = 0; c < 4; ++c) put_f32(h, 280 + row * 16 + c * 4, (float)affine[row * 4 + c]);
    std::memcpy(h + 344, "n+1\0", 4);
    for (size_t i = 0; i < n; ++i) put_f32(h, 352 + 4 * i, (float)data[i]);
    return QSM_OK;
}

bool ends_with(const std::string& s, const char* suf) {
    const size_t m = std::strlen(suf);
    return s.size() >= m && s.compare(s.size() - m, m, suf) == 0;
}

}  // namespace

extern "C" {

int qsm_nifti_is_gzip(const uint8_t* bytes, size_t len) { return bytes && is_gzip(bytes, len) ? 1 : 0; }

int qsm_nifti_load(const uint8_t* bytes, size_t len, qsm_nifti_info* info, double* data, size_t data_len) {
    if (!bytes) return fail(QSM_EPARAM, "null argument");
    if (is_gzip(bytes, len)) {
        std::vector<uint8_t> raw;
        int rc = gunzip(bytes, len, raw);
        if (rc) return rc;
        return parse(raw.data(), raw.size(), info, data, data_len);
    }
    return parse(bytes, len, info, data, data_len);
}

int qsm_nifti_save(const double* data, size_t nx, size_t ny, size_t nz, const double voxel_size[3],
                   const double affine[16], int gzip_it, uint8_t** bytes_out, size_t* len_out) {
    if (!bytes_out || !len_out) return fail(QSM_EPARAM, "null argument");
    std::vector<uint8_t> raw;
    int rc = build(data, nx, ny, nz, voxel_size, affine, raw);
    if (rc) return rc;
    std::vector<uint8_t> z;
    if (gzip_it) {
        if ((rc = gzip(raw.data(), raw.size(), z))) return rc;
        raw.swap(z);
    }
    uint8_t* p = (uint8_t*)std::malloc(raw.size());
    if (!p) return fail(QSM_ENOMEM, "out of host memory");
    std::memcpy(p, raw.data(), raw.size());
    *bytes_out = p;
    *len_out = raw.size();
    return QSM_OK;
}

void qsm_nifti_free(uint8_t* bytes) { std::free(bytes); }

int qsm_nifti_read_file(const char* path, qsm_nifti_info* info, double* data, size_t data_len) {
    if (!path) return fail(QSM_EPARAM, "null argument");
    FILE* f = std::fopen(path, "rb");
    if (!f) return fail(QSM_EPARAM, "Failed to read file '%s'", path);
    std::vector<uint8_t> bytes;
    uint8_t buf[1 << 16];
    size_t k;
    while ((k = std::fread(buf, 1, sizeof buf, f)) > 0) bytes.insert(bytes.end(), buf, buf + k);
    std::fclose(f);
    return qsm_nifti_load(bytes.data(), bytes.size(), info, data, data_len);
}

int qsm_nifti_write_file(const char* path, const double* data, size_t nx, size_t ny, size_t nz,
                         const double voxel_size[3], const double affine[16]) {
    if (!path) return fail(QSM_EPARAM, "null argument");
    uint8_t* bytes = nullptr;
    size_t len = 0;
    int rc = qsm_nifti_save(data, nx, ny, nz, voxel_size, affine, ends_with(path, ".nii.gz") ? 1 : 0, &bytes, &len);
    if (rc) return rc;
    FILE* f = std::fopen(path, "wb");
    if (!f) {
        qsm_nifti_free(bytes);
        return fail(QSM_EPARAM, "Failed to write file '%s'", path);
    }
    const size_t w = std::fwrite(bytes, 1, len, f);
    std::fclose(f);
    qsm_nifti_free(bytes);
    if (w != len) return fail(QSM_EPARAM, "Failed to write file '%s'", path);
    return QSM_OK;
}

}  // extern "C"
