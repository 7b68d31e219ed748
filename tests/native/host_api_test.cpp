// The reference's unit tests for the TV-ADMM / V-SHARP path, restated in C++
// against the GPU backend through the host mirror qsm.rs_amd/host/qsm.hpp.
// Each test names the Rust test it follows.  Built by __graft_entry__.build()
// with g++ against qsm.rs_amd/libqsm_gpu.so; exit status 0 when every test
// passes.
#include "../../qsm.rs_amd/host/qsm.hpp"

#include <cmath>
#include <cstdio>
#include <vector>

using namespace qsm;

static int g_fail = 0;
#define EXPECT(cond, msg)                                  \
    do {                                                   \
        if (!(cond)) {                                     \
            std::printf("  FAIL %s: %s\n", __func__, msg); \
            ++g_fail;                                      \
            return;                                        \
        }                                                  \
    } while (0)

static std::vector<double> ramp(size_t n, double s) {
    std::vector<double> v(n);
    for (size_t i = 0; i < n; ++i) v[i] = (double)i * s;
    return v;
}

// src/inversion/tv.rs:113-126 test_tv_admm_zero_field
static void test_tv_admm_zero_field() {
    const size_t n = 8;
    Grid g = Grid::make(n, n, n, 1, 1, 1);
    inversion::TvParams p{1e-3, 0.1, 1e-2, 10};
    auto chi = inversion::tv_admm(std::vector<double>(n * n * n, 0.0),
                                  std::vector<uint8_t>(n * n * n, 1), g, {0, 0, 1}, p);
    for (double v : chi) EXPECT(std::fabs(v) < 1e-8, "zero field should give zero chi");
}

// src/inversion/tv.rs:129-142 test_tv_admm_finite
static void test_tv_admm_finite() {
    const size_t n = 8;
    Grid g = Grid::make(n, n, n, 1, 1, 1);
    inversion::TvParams p{1e-3, 0.1, 1e-2, 10};
    auto chi = inversion::tv_admm(ramp(n * n * n, 0.001), std::vector<uint8_t>(n * n * n, 1), g,
                                  {0, 0, 1}, p);
    for (double v : chi) EXPECT(std::isfinite(v), "chi should be finite");
}

// src/inversion/tv.rs:144-168 test_tv_admm_smoother_than_tkd (TV finite)
static void test_tv_admm_smoother_than_tkd() {
    const size_t n = 8;
    std::vector<double> f(n * n * n);
    for (size_t i = 0; i < f.size(); ++i) f[i] = (i % 2 == 0) ? 0.01 : -0.01;
    Grid g = Grid::make(n, n, n, 1, 1, 1);
    inversion::TvParams p{1e-2, 1.0, 1e-2, 50};
    auto chi = inversion::tv_admm(f, std::vector<uint8_t>(n * n * n, 1), g, {0, 0, 1}, p);
    double tv = 0;
    for (size_t k = 0; k < n; ++k)
        for (size_t j = 0; j < n; ++j)
            for (size_t i = 0; i < n; ++i) {
                size_t idx = i + j * n + k * n * n;
                tv += std::fabs(chi[(i + 1) % n + j * n + k * n * n] - chi[idx]) +
                      std::fabs(chi[i + ((j + 1) % n) * n + k * n * n] - chi[idx]) +
                      std::fabs(chi[i + j * n + ((k + 1) % n) * n * n] - chi[idx]);
            }
    EXPECT(std::isfinite(tv), "TV should be finite");
}

// src/inversion/tv.rs:34-43 TvParams::default
static void test_tv_params_default() {
    auto p = inversion::TvParams::default_params();
    EXPECT(p.lambda == 2e-4 && p.rho == 2e-2 && p.tol == 1e-3 && p.max_iter == 250, "defaults");
}

// src/bgremove/vsharp.rs:206-219 test_vsharp_zero_field
static void test_vsharp_zero_field() {
    const size_t n = 8;
    Grid g = Grid::make(n, n, n, 1, 1, 1);
    auto r = bgremove::vsharp_with_radii(std::vector<double>(n * n * n, 0.0),
                                         std::vector<uint8_t>(n * n * n, 1), g, {4, 3, 2}, 0.05);
    for (double v : r.first) EXPECT(std::fabs(v) < 1e-10, "zero field -> zero local field");
}

// src/bgremove/vsharp.rs:247-265 test_vsharp_nonuniform_voxels
static void test_vsharp_nonuniform_voxels() {
    const size_t n = 8;
    Grid g = Grid::make(n, n, n, 0.5, 1.0, 2.0);
    auto r = bgremove::vsharp_with_radii(ramp(n * n * n, 0.001), std::vector<uint8_t>(n * n * n, 1),
                                         g, {4, 3, 2}, 0.05);
    size_t cnt = 0;
    for (size_t i = 0; i < r.first.size(); ++i) {
        EXPECT(std::isfinite(r.first[i]), "finite");
        cnt += r.second[i];
    }
    EXPECT(cnt > 0, "final mask should have some voxels");
}

// src/bgremove/vsharp.rs:302-318 test_vsharp_empty_radii
static void test_vsharp_empty_radii() {
    const size_t n = 8;
    Grid g = Grid::make(n, n, n, 1, 1, 1);
    std::vector<uint8_t> mask(n * n * n, 1);
    auto r = bgremove::vsharp_with_radii(ramp(n * n * n, 0.001), mask, g, {}, 0.05);
    for (double v : r.first) EXPECT(v == 0.0, "empty radii -> zeros");
    EXPECT(r.second == mask, "empty radii -> original mask");
}

// src/bgremove/vsharp.rs:321-366 test_vsharp_larger_volume
static void test_vsharp_larger_volume() {
    const size_t n = 16;
    std::vector<double> field(n * n * n);
    std::vector<uint8_t> mask(n * n * n, 0);
    const int c = (int)n / 2, rad = (int)n / 3;
    for (size_t z = 0; z < n; ++z)
        for (size_t y = 0; y < n; ++y)
            for (size_t x = 0; x < n; ++x) {
                field[x + y * n + z * n * n] = (double)z * 0.1;
                int dx = (int)x - c, dy = (int)y - c, dz = (int)z - c;
                if (dx * dx + dy * dy + dz * dz <= rad * rad) mask[x + y * n + z * n * n] = 1;
            }
    Grid g = Grid::make(n, n, n, 1, 1, 1);
    auto r = bgremove::vsharp_with_radii(field, mask, g, {6, 4, 3, 2}, 0.05);
    size_t cnt = 0;
    for (size_t i = 0; i < field.size(); ++i) {
        EXPECT(std::isfinite(r.first[i]), "finite");
        if (r.second[i] == 0) EXPECT(r.first[i] == 0.0, "outside the final mask is zero");
        cnt += r.second[i];
    }
    EXPECT(cnt > 0, "final mask non-empty");
}

// src/bgremove/vsharp.rs:369-391 test_vsharp_with_progress
static void test_vsharp_with_progress() {
    const size_t n = 8;
    Grid g = Grid::make(n, n, n, 1, 1, 1);
    std::vector<std::pair<size_t, size_t>> calls;
    bgremove::vsharp_with_radii(ramp(n * n * n, 0.001), std::vector<uint8_t>(n * n * n, 1), g,
                                {4, 3, 2}, 0.05,
                                [&](size_t a, size_t b) { calls.emplace_back(a, b); });
    const bool ok = calls.size() == 3 && calls[0].first == 1 && calls[0].second == 3 &&
                    calls[1].first == 2 && calls[2].first == 3 && calls[2].second == 3;
    EXPECT(ok, "progress (1,3),(2,3),(3,3)");
}

// src/bgremove/vsharp.rs:431-459 test_vsharp_unsorted_radii
static void test_vsharp_unsorted_radii() {
    const size_t n = 8;
    Grid g = Grid::make(n, n, n, 1, 1, 1);
    auto f = ramp(n * n * n, 0.001);
    std::vector<uint8_t> m(n * n * n, 1);
    auto a = bgremove::vsharp_with_radii(f, m, g, {4, 3, 2}, 0.05);
    auto b = bgremove::vsharp_with_radii(f, m, g, {2, 4, 3}, 0.05);
    EXPECT(a.second == b.second, "same mask");
    for (size_t i = 0; i < f.size(); ++i) EXPECT(std::fabs(a.first[i] - b.first[i]) < 1e-10, "same field");
}

// src/pipeline/bg_removal.rs:99-120 test_bg_removal_dispatches_vsharp
static void test_bg_removal_dispatches_vsharp() {
    pipeline::ScanMetadata md{{8, 8, 8}, {1, 1, 1}, {0.005}, 3.0, {0, 0, 1}};
    auto r = pipeline::run_bg_removal(std::vector<double>(512, 0.1), std::vector<uint8_t>(512, 1), md,
                                      pipeline::BgRemovalConfig{});
    EXPECT(r.ok && r.value.local_field_ppm.size() == 512 && r.value.eroded_mask.size() == 512,
           "dispatch to vsharp");
}

// src/pipeline/bg_removal.rs:198-214 test_bg_removal_validates_dims
static void test_bg_removal_validates_dims() {
    pipeline::ScanMetadata md{{4, 4, 4}, {1, 1, 1}, {0.005}, 3.0, {0, 0, 1}};
    auto r = pipeline::run_bg_removal(std::vector<double>(32, 0.0), std::vector<uint8_t>(64, 1), md,
                                      pipeline::BgRemovalConfig{});
    EXPECT(!r.ok && r.error.kind == pipeline::PipelineErrorKind::DimensionMismatch,
           "wrong length -> DimensionMismatch");
}

// src/pipeline/inversion.rs:214-217 test_inversion_tv, 254-272 test_inversion_rejects_tgv
static void test_inversion_tv_and_rejects_tgv() {
    pipeline::ScanMetadata md{{8, 8, 8}, {1, 1, 1}, {0.005}, 3.0, {0, 0, 1}};
    pipeline::InversionConfig cfg;
    cfg.algorithm = QSM_INV_TV;
    auto r = pipeline::run_dipole_inversion(std::vector<double>(512, 0.01),
                                            std::vector<uint8_t>(512, 1), md, cfg);
    EXPECT(r.ok && r.value.size() == 512, "Tv arm runs");
    pipeline::ScanMetadata md4{{4, 4, 4}, {1, 1, 1}, {0.005}, 3.0, {0, 0, 1}};
    cfg.algorithm = QSM_INV_TGV;
    auto t = pipeline::run_dipole_inversion(std::vector<double>(64, 0.0), std::vector<uint8_t>(64, 1),
                                            md4, cfg);
    EXPECT(!t.ok && t.error.kind == pipeline::PipelineErrorKind::InvalidConfig, "Tgv rejected");
}

// src/inversion/tkd.rs:160-231 (zero field, mask, finite, tsvd differs)
static void test_tkd_reference_tests() {
    const size_t n = 8;
    Grid g = Grid::make(n, n, n, 1, 1, 1);
    std::vector<uint8_t> m(n * n * n, 1);
    auto z = inversion::tkd(std::vector<double>(n * n * n, 0.0), m, g, {0, 0, 1}, {0.15});
    for (double v : z) EXPECT(std::fabs(v) < 1e-10, "zero field -> zero chi");
    std::vector<uint8_t> m2 = m;
    m2[0] = m2[1] = 0;
    auto c = inversion::tkd(std::vector<double>(n * n * n, 1.0), m2, g, {0, 0, 1}, {0.15});
    EXPECT(c[0] == 0.0 && c[1] == 0.0, "outside the mask is 0");
    std::vector<double> f16(16 * 16 * 16);
    for (size_t i = 0; i < f16.size(); ++i) f16[i] = std::sin(i * 0.1);
    Grid g16 = Grid::make(16, 16, 16, 1, 1, 1);
    std::vector<uint8_t> m16(f16.size(), 1);
    auto a = inversion::tkd(f16, m16, g16, {0, 0, 1}, {0.15});
    auto b = inversion::tsvd(f16, m16, g16, {0, 0, 1}, {0.15});
    double d = 0;
    for (size_t i = 0; i < a.size(); ++i) d += std::fabs(a[i] - b[i]);
    EXPECT(d > 1e-10, "tkd and tsvd differ");
}

// src/inversion/tikhonov.rs:173-199 (regularisation types differ)
static void test_tikhonov_regularization_types() {
    const size_t n = 8;
    Grid g = Grid::make(n, n, n, 1, 1, 1);
    std::vector<double> f(n * n * n);
    for (size_t i = 0; i < f.size(); ++i) f[i] = std::sin(i * 0.1);
    std::vector<uint8_t> m(f.size(), 1);
    using R = inversion::Regularization;
    auto i0 = inversion::tikhonov(f, m, g, {0, 0, 1}, {0.01, R::Identity});
    auto i1 = inversion::tikhonov(f, m, g, {0, 0, 1}, {0.01, R::Gradient});
    auto i2 = inversion::tikhonov(f, m, g, {0, 0, 1}, {0.01, R::Laplacian});
    double d01 = 0, d12 = 0;
    for (size_t i = 0; i < f.size(); ++i) { d01 += std::fabs(i0[i] - i1[i]); d12 += std::fabs(i1[i] - i2[i]); }
    EXPECT(d01 > 1e-10 && d12 > 1e-10, "Identity, Gradient and Laplacian differ");
}

// src/inversion/rts.rs:219-265 (zero field, finite, mask)
static void test_rts_reference_tests() {
    const size_t n = 8;
    Grid g = Grid::make(n, n, n, 1, 1, 1);
    inversion::RtsParams p;
    p.max_iter = 5;
    p.lsmr_iter = 2;
    std::vector<uint8_t> m(n * n * n, 1);
    auto z = inversion::rts(std::vector<double>(n * n * n, 0.0), m, g, {0, 0, 1}, p);
    for (double v : z) EXPECT(std::fabs(v) < 1e-6, "zero field -> near-zero chi");
    auto r = inversion::rts(ramp(n * n * n, 0.001), m, g, {0, 0, 1}, p);
    for (double v : r) EXPECT(std::isfinite(v), "finite");
    std::vector<uint8_t> m2 = m;
    m2[0] = m2[10] = 0;
    auto c = inversion::rts(ramp(n * n * n, 0.001), m2, g, {0, 0, 1}, p);
    EXPECT(c[0] == 0.0 && c[10] == 0.0, "masked voxels are 0");
}

// src/pipeline/inversion.rs: the default config (Rts) runs; Sharp arm runs
static void test_pipeline_default_rts_and_sharp_arm() {
    pipeline::ScanMetadata md{{8, 8, 8}, {1, 1, 1}, {0.005}, 3.0, {0, 0, 1}};
    auto r = pipeline::run_dipole_inversion(std::vector<double>(512, 0.01), std::vector<uint8_t>(512, 1),
                                            md, pipeline::InversionConfig{});
    EXPECT(r.ok && r.value.size() == 512, "default (Rts) arm runs");
    pipeline::BgRemovalConfig bc;
    bc.algorithm = QSM_BG_SHARP;
    bc.sharp.radius = 3.0;
    auto s = pipeline::run_bg_removal(ramp(512, 0.001), std::vector<uint8_t>(512, 1), md, bc);
    EXPECT(s.ok && s.value.eroded_mask.size() == 512, "Sharp arm runs");
}

// src/unwrap/laplacian.rs:181-230 (constant -> constant, smooth, finite)
static void test_laplacian_unwrap_reference_tests() {
    const size_t n = 8;
    Grid g = Grid::make(n, n, n, 1, 1, 1);
    std::vector<uint8_t> m(n * n * n, 1);
    auto u = unwrap::laplacian_unwrap(std::vector<double>(n * n * n, 1.0), m, g);
    double mean = 0;
    for (double v : u) mean += v;
    mean /= u.size();
    for (double v : u) EXPECT(std::fabs(v - mean) < 1e-6, "constant phase -> constant");
    const size_t n16 = 16;
    Grid g16 = Grid::make(n16, n16, n16, 1, 1, 1);
    std::vector<double> ph(n16 * n16 * n16);
    for (size_t k = 0; k < n16; ++k)
        for (size_t j = 0; j < n16; ++j)
            for (size_t i = 0; i < n16; ++i) ph[i + j * n16 + k * n16 * n16] = 0.5 * std::sin(2 * M_PI * i / n16);
    auto s = unwrap::laplacian_unwrap(ph, std::vector<uint8_t>(ph.size(), 1), g16);
    for (size_t i = 0; i < ph.size(); ++i) EXPECT(std::isfinite(s[i]) && std::fabs(ph[i] - s[i]) < 1.0, "smooth");
}

// src/inversion/nltv.rs:138-185 (zero field, finite; weighted_shrink is
// shrink(x, t w), pinned in the oracle tests)
static void test_nltv_reference_tests() {
    const size_t n = 8;
    Grid g = Grid::make(n, n, n, 1, 1, 1);
    inversion::NltvParams p;
    p.max_iter = 5;
    p.newton_iter = 2;
    std::vector<uint8_t> m(n * n * n, 1);
    auto z = inversion::nltv(std::vector<double>(n * n * n, 0.0), m, g, {0, 0, 1}, p);
    for (double v : z) EXPECT(std::fabs(v) < 1e-10, "zero field -> zero chi");
    auto r = inversion::nltv(ramp(n * n * n, 0.001), m, g, {0, 0, 1}, p);
    for (double v : r) EXPECT(std::isfinite(v), "finite");
}

// src/io.rs:477-546 (save header layout, round trip)
static void test_nifti_save_and_load() {
    std::vector<double> data(64);
    for (size_t i = 0; i < data.size(); ++i) data[i] = i * 0.5 + 1.0;
    const std::array<double, 16> aff = {1, 0, 0, 10, 0, 2, 0, 20, 0, 0, 3, 30, 0, 0, 0, 1};
    auto bytes = io::save_nifti(data, {4, 4, 4}, {1.0, 2.0, 3.0}, aff);
    EXPECT(bytes.size() == 352 + 64 * 4 && bytes[344] == 'n' && bytes[345] == '+', "header layout");
    auto d = io::load_nifti(bytes);
    EXPECT(d.dims == std::make_tuple(size_t(4), size_t(4), size_t(4)), "dims");
    EXPECT(std::get<1>(d.voxel_size) == 2.0 && d.affine[3] == 10.0 && d.affine[11] == 30.0, "geometry");
    for (size_t i = 0; i < data.size(); ++i) EXPECT(std::fabs(d.data[i] - data[i]) < 0.01, "values");
    auto gz = io::save_nifti(data, {4, 4, 4}, {1.0, 2.0, 3.0}, aff, true);
    EXPECT(gz[0] == 0x1f && gz[1] == 0x8b && io::load_nifti(gz).data == d.data, "gzip round trip");
}

int main() {
    void (*tests[])() = {test_tv_admm_zero_field, test_tv_admm_finite, test_tv_admm_smoother_than_tkd,
                         test_tv_params_default, test_vsharp_zero_field, test_vsharp_nonuniform_voxels,
                         test_vsharp_empty_radii, test_vsharp_larger_volume, test_vsharp_with_progress,
                         test_vsharp_unsorted_radii, test_bg_removal_dispatches_vsharp,
                         test_bg_removal_validates_dims, test_inversion_tv_and_rejects_tgv,
                         test_tkd_reference_tests, test_tikhonov_regularization_types,
                         test_rts_reference_tests, test_pipeline_default_rts_and_sharp_arm,
                         test_laplacian_unwrap_reference_tests, test_nltv_reference_tests,
                         test_nifti_save_and_load};
    int i = 0;
    for (auto t : tests) {
        const int before = g_fail;
        try {
            t();
        } catch (const std::exception& e) {
            std::printf("  EXCEPTION in test %d: %s\n", i, e.what());
            ++g_fail;
        }
        std::printf("test %2d %s\n", i, g_fail == before ? "ok" : "FAILED");
        ++i;
    }
    std::printf("%s (%d failures)\n", g_fail ? "FAILED" : "ALL OK", g_fail);
    return g_fail ? 1 : 0;
}
