// C++ mirror of qsm-core's TV-ADMM / V-SHARP API over the gfx950 C ABI
// (include/qsm_gpu.h).  Names, argument order and meaning follow the Rust
// crate so that call sites read the same:
//
//   qsm::Grid                       Grid                  src/grid.rs:8-44
//   qsm::inversion::TvParams        inversion::TvParams   src/inversion/tv.rs:21-43
//   qsm::inversion::tv_admm         inversion::tv_admm    src/inversion/tv.rs:63-97
//   qsm::bgremove::VsharpParams     bgremove::VsharpParams src/bgremove/vsharp.rs:20-39
//   qsm::bgremove::vsharp           bgremove::vsharp      src/bgremove/vsharp.rs:55-64
//   qsm::inversion::tkd / tsvd      src/inversion/tkd.rs:48-99
//   qsm::inversion::tikhonov        src/inversion/tikhonov.rs:62-131
//   qsm::inversion::rts             src/inversion/rts.rs:71-212
//   qsm::inversion::nltv            src/inversion/nltv.rs:63-133
//   qsm::bgremove::sharp            src/bgremove/sharp.rs:48-56
//   qsm::unwrap::laplacian_unwrap   src/unwrap/laplacian.rs:135-161
//   qsm::io::load_nifti / save_nifti src/io.rs:62-357
//   qsm::pipeline::run_dipole_inversion / run_bg_removal  src/pipeline/*.rs
//
// Algorithm functions throw qsm::Error on a non-zero status (the Rust
// functions panic on bad slices); pipeline functions return a Result-like
// struct carrying a PipelineError, as the Rust pipeline does.
#pragma once

#include <array>
#include <cstddef>
#include <cstdint>
#include <functional>
#include <stdexcept>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "../../include/qsm_gpu.h"

namespace qsm {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void check(int rc) {
    if (rc != QSM_OK) throw Error(rc, qsm_last_error_message());
}

using Progress = std::function<void(size_t, size_t)>;

namespace detail {
inline void trampoline(size_t cur, size_t total, void* user) {
    (*static_cast<Progress*>(user))(cur, total);
}
}  // namespace detail

struct Grid {
    std::tuple<size_t, size_t, size_t> dims;
    std::tuple<double, double, double> voxel_size;
    static Grid make(size_t nx, size_t ny, size_t nz, double vsx, double vsy, double vsz) {
        return Grid{{nx, ny, nz}, {vsx, vsy, vsz}};
    }
    size_t nx() const { return std::get<0>(dims); }
    size_t ny() const { return std::get<1>(dims); }
    size_t nz() const { return std::get<2>(dims); }
    double vsx() const { return std::get<0>(voxel_size); }
    double vsy() const { return std::get<1>(voxel_size); }
    double vsz() const { return std::get<2>(voxel_size); }
    size_t n_total() const { return nx() * ny() * nz(); }
    qsm_grid c() const { return qsm_grid{nx(), ny(), nz(), vsx(), vsy(), vsz()}; }
};

namespace inversion {

struct TvParams {
    double lambda, rho, tol;
    size_t max_iter;
    static TvParams default_params() {
        qsm_tv_params p;
        qsm_tv_params_default(&p);
        return TvParams{p.lambda, p.rho, p.tol, p.max_iter};
    }
    qsm_tv_params c() const { return qsm_tv_params{lambda, rho, tol, max_iter}; }
};

inline std::vector<double> tv_admm(const std::vector<double>& local_field,
                                   const std::vector<uint8_t>& mask, const Grid& grid,
                                   std::tuple<double, double, double> bdir, const TvParams& params,
                                   Progress progress = nullptr) {
    if (local_field.size() != grid.n_total() || mask.size() != grid.n_total())
        throw Error(QSM_EDIM, "array length does not match the grid");
    std::vector<double> chi(grid.n_total());
    const qsm_grid g = grid.c();
    const double b[3] = {std::get<0>(bdir), std::get<1>(bdir), std::get<2>(bdir)};
    const qsm_tv_params p = params.c();
    check(qsm_tv_admm(local_field.data(), mask.data(), &g, b, &p,
                      progress ? detail::trampoline : nullptr, progress ? &progress : nullptr,
                      chi.data()));
    return chi;
}

struct TkdParams {
    double threshold = 0.15;
    qsm_tkd_params c() const { return qsm_tkd_params{threshold}; }
};

enum class Regularization { Identity = QSM_REG_IDENTITY, Gradient = QSM_REG_GRADIENT, Laplacian = QSM_REG_LAPLACIAN };

struct TikhonovParams {
    double lambda = 0.01;
    Regularization reg = Regularization::Identity;
    qsm_tikhonov_params c() const { return qsm_tikhonov_params{lambda, (qsm_regularization)reg}; }
};

struct RtsParams {
    double delta = 0.15, mu = 1e5, rho = 10.0, tol = 1e-2;
    size_t max_iter = 20, lsmr_iter = 4;
    qsm_rts_params c() const { return qsm_rts_params{delta, mu, rho, tol, max_iter, lsmr_iter}; }
};

namespace detail {
template <class F, class P>
inline std::vector<double> closed_form(F fn, const std::vector<double>& f, const std::vector<uint8_t>& m,
                                       const Grid& grid, std::tuple<double, double, double> bdir,
                                       const P& p) {
    if (f.size() != grid.n_total() || m.size() != grid.n_total())
        throw Error(QSM_EDIM, "array length does not match the grid");
    std::vector<double> chi(grid.n_total());
    const qsm_grid g = grid.c();
    const double b[3] = {std::get<0>(bdir), std::get<1>(bdir), std::get<2>(bdir)};
    const auto pc = p.c();
    check(fn(f.data(), m.data(), &g, b, &pc, chi.data()));
    return chi;
}
}  // namespace detail

// src/inversion/tkd.rs:48-99 and the tsvd after it
inline std::vector<double> tkd(const std::vector<double>& f, const std::vector<uint8_t>& m, const Grid& g,
                               std::tuple<double, double, double> bdir, const TkdParams& p) {
    return detail::closed_form(qsm_tkd, f, m, g, bdir, p);
}
inline std::vector<double> tsvd(const std::vector<double>& f, const std::vector<uint8_t>& m, const Grid& g,
                                std::tuple<double, double, double> bdir, const TkdParams& p) {
    return detail::closed_form(qsm_tsvd, f, m, g, bdir, p);
}
// src/inversion/tikhonov.rs:62-131
inline std::vector<double> tikhonov(const std::vector<double>& f, const std::vector<uint8_t>& m,
                                    const Grid& g, std::tuple<double, double, double> bdir,
                                    const TikhonovParams& p) {
    return detail::closed_form(qsm_tikhonov, f, m, g, bdir, p);
}
// src/inversion/rts.rs:71-212
inline std::vector<double> rts(const std::vector<double>& f, const std::vector<uint8_t>& m, const Grid& grid,
                               std::tuple<double, double, double> bdir, const RtsParams& params,
                               Progress progress = nullptr) {
    if (f.size() != grid.n_total() || m.size() != grid.n_total())
        throw Error(QSM_EDIM, "array length does not match the grid");
    std::vector<double> chi(grid.n_total());
    const qsm_grid g = grid.c();
    const double b[3] = {std::get<0>(bdir), std::get<1>(bdir), std::get<2>(bdir)};
    const qsm_rts_params p = params.c();
    check(qsm_rts(f.data(), m.data(), &g, b, &p, progress ? qsm::detail::trampoline : nullptr,
                  progress ? &progress : nullptr, chi.data()));
    return chi;
}

struct NltvParams {
    double lambda = 1e-3, mu = 1.0, tol = 1e-3;
    size_t max_iter = 250, newton_iter = 10;
    qsm_nltv_params c() const { return qsm_nltv_params{lambda, mu, tol, max_iter, newton_iter}; }
};

// src/inversion/nltv.rs:63-133
inline std::vector<double> nltv(const std::vector<double>& f, const std::vector<uint8_t>& m, const Grid& grid,
                                std::tuple<double, double, double> bdir, const NltvParams& params,
                                Progress progress = nullptr) {
    if (f.size() != grid.n_total() || m.size() != grid.n_total())
        throw Error(QSM_EDIM, "array length does not match the grid");
    std::vector<double> chi(grid.n_total());
    const qsm_grid g = grid.c();
    const double b[3] = {std::get<0>(bdir), std::get<1>(bdir), std::get<2>(bdir)};
    const qsm_nltv_params p = params.c();
    check(qsm_nltv(f.data(), m.data(), &g, b, &p, progress ? qsm::detail::trampoline : nullptr,
                   progress ? &progress : nullptr, chi.data()));
    return chi;
}

}  // namespace inversion

namespace bgremove {

struct SharpParams {
    double threshold = 0.05, radius = 6.0;
};

// src/bgremove/sharp.rs:48-56
inline std::pair<std::vector<double>, std::vector<uint8_t>> sharp(const std::vector<double>& field,
                                                                   const std::vector<uint8_t>& mask,
                                                                   const Grid& grid, const SharpParams& p) {
    if (field.size() != grid.n_total() || mask.size() != grid.n_total())
        throw Error(QSM_EDIM, "array length does not match the grid");
    std::vector<double> local(grid.n_total());
    std::vector<uint8_t> eroded(grid.n_total());
    const qsm_grid g = grid.c();
    const qsm_sharp_params pc{p.threshold, p.radius};
    check(qsm_sharp(field.data(), mask.data(), &g, &pc, local.data(), eroded.data()));
    return {std::move(local), std::move(eroded)};
}

struct VsharpParams {
    double threshold, max_radius, min_radius;
    static VsharpParams default_params() {
        qsm_vsharp_params p;
        qsm_vsharp_params_default(&p);
        return VsharpParams{p.threshold, p.max_radius, p.min_radius};
    }
    qsm_vsharp_params c() const { return qsm_vsharp_params{threshold, max_radius, min_radius}; }
};

inline std::pair<std::vector<double>, std::vector<uint8_t>> vsharp(
    const std::vector<double>& field, const std::vector<uint8_t>& mask, const Grid& grid,
    const VsharpParams& params, Progress progress = nullptr) {
    if (field.size() != grid.n_total() || mask.size() != grid.n_total())
        throw Error(QSM_EDIM, "array length does not match the grid");
    std::vector<double> local(grid.n_total());
    std::vector<uint8_t> eroded(grid.n_total());
    const qsm_grid g = grid.c();
    const qsm_vsharp_params p = params.c();
    check(qsm_vsharp(field.data(), mask.data(), &g, &p, progress ? detail::trampoline : nullptr,
                     progress ? &progress : nullptr, local.data(), eroded.data()));
    return {std::move(local), std::move(eroded)};
}

inline std::pair<std::vector<double>, std::vector<uint8_t>> vsharp_with_radii(
    const std::vector<double>& field, const std::vector<uint8_t>& mask, const Grid& grid,
    const std::vector<double>& radii, double threshold, Progress progress = nullptr) {
    if (field.size() != grid.n_total() || mask.size() != grid.n_total())
        throw Error(QSM_EDIM, "array length does not match the grid");
    std::vector<double> local(grid.n_total());
    std::vector<uint8_t> eroded(grid.n_total());
    const qsm_grid g = grid.c();
    check(qsm_vsharp_with_radii(field.data(), mask.data(), &g, radii.data(), radii.size(),
                                threshold, progress ? detail::trampoline : nullptr,
                                progress ? &progress : nullptr, local.data(), eroded.data()));
    return {std::move(local), std::move(eroded)};
}

}  // namespace bgremove

namespace unwrap {
// src/unwrap/laplacian.rs:135-161
inline std::vector<double> laplacian_unwrap(const std::vector<double>& phase, const std::vector<uint8_t>& mask,
                                            const Grid& grid) {
    if (phase.size() != grid.n_total() || mask.size() != grid.n_total())
        throw Error(QSM_EDIM, "array length does not match the grid");
    std::vector<double> out(grid.n_total());
    const qsm_grid g = grid.c();
    check(qsm_laplacian_unwrap(phase.data(), mask.data(), &g, out.data()));
    return out;
}
}  // namespace unwrap

namespace io {
// NiftiData, src/io.rs:13-26
struct NiftiData {
    std::vector<double> data;
    std::tuple<size_t, size_t, size_t> dims;
    std::tuple<double, double, double> voxel_size;
    std::array<double, 16> affine;
    double scl_slope, scl_inter;
};

// load_nifti, src/io.rs:62-158 (first volume of a 4-D file)
inline NiftiData load_nifti(const std::vector<uint8_t>& bytes) {
    qsm_nifti_info info;
    check(qsm_nifti_load(bytes.data(), bytes.size(), &info, nullptr, 0));
    std::vector<double> all(info.nx * info.ny * info.nz * info.nt);
    check(qsm_nifti_load(bytes.data(), bytes.size(), &info, all.data(), all.size()));
    all.resize(info.nx * info.ny * info.nz);
    NiftiData d{std::move(all), {info.nx, info.ny, info.nz}, {info.vsx, info.vsy, info.vsz}, {},
                info.scl_slope, info.scl_inter};
    for (int i = 0; i < 16; ++i) d.affine[i] = info.affine[i];
    return d;
}

// save_nifti / save_nifti_gz, src/io.rs:262-357
inline std::vector<uint8_t> save_nifti(const std::vector<double>& data, std::tuple<size_t, size_t, size_t> dims,
                                       std::tuple<double, double, double> vs,
                                       const std::array<double, 16>& affine, bool gz = false) {
    uint8_t* p = nullptr;
    size_t n = 0;
    const double v[3] = {std::get<0>(vs), std::get<1>(vs), std::get<2>(vs)};
    check(qsm_nifti_save(data.data(), std::get<0>(dims), std::get<1>(dims), std::get<2>(dims), v,
                         affine.data(), gz ? 1 : 0, &p, &n));
    std::vector<uint8_t> out(p, p + n);
    qsm_nifti_free(p);
    return out;
}
}  // namespace io

namespace pipeline {

enum class PipelineErrorKind { InvalidConfig, InvalidInput, AlgorithmError, DimensionMismatch };

struct PipelineError {
    PipelineErrorKind kind;
    std::string message;
};

template <class T>
struct Result {
    bool ok;
    T value;
    PipelineError error;
};

struct ScanMetadata {
    std::tuple<size_t, size_t, size_t> dims;
    std::tuple<double, double, double> voxel_size;
    std::vector<double> echo_times;
    double field_strength;
    std::tuple<double, double, double> b0_direction;
    Grid grid() const { return Grid{dims, voxel_size}; }
    qsm_scan_metadata c() const {
        qsm_scan_metadata m;
        m.nx = std::get<0>(dims); m.ny = std::get<1>(dims); m.nz = std::get<2>(dims);
        m.vsx = std::get<0>(voxel_size); m.vsy = std::get<1>(voxel_size); m.vsz = std::get<2>(voxel_size);
        m.echo_times = echo_times.data();
        m.n_echo_times = echo_times.size();
        m.field_strength = field_strength;
        m.b0_direction[0] = std::get<0>(b0_direction);
        m.b0_direction[1] = std::get<1>(b0_direction);
        m.b0_direction[2] = std::get<2>(b0_direction);
        return m;
    }
};

struct InversionConfig {
    qsm_inversion_algorithm algorithm = QSM_INV_RTS;
    inversion::TkdParams tkd;
    inversion::TkdParams tsvd;
    inversion::TikhonovParams tikhonov;
    inversion::TvParams tv = inversion::TvParams::default_params();
    inversion::RtsParams rts;
    inversion::NltvParams nltv;
    qsm_inversion_config c() const {
        return qsm_inversion_config{algorithm, tkd.c(), tsvd.c(), tikhonov.c(), tv.c(), rts.c(), nltv.c()};
    }
};

struct BgRemovalConfig {
    qsm_bg_removal_algorithm algorithm = QSM_BG_VSHARP;
    bgremove::VsharpParams vsharp = bgremove::VsharpParams::default_params();
    bgremove::SharpParams sharp;
    qsm_bg_removal_config c() const {
        return qsm_bg_removal_config{algorithm, vsharp.c(), qsm_sharp_params{sharp.threshold, sharp.radius}};
    }
};

struct BgRemovalResult {
    std::vector<double> local_field_ppm;
    std::vector<uint8_t> eroded_mask;
};

inline PipelineError to_error(int rc) {
    PipelineErrorKind k = rc == QSM_EDIM ? PipelineErrorKind::DimensionMismatch
                        : rc == QSM_ECONFIG ? PipelineErrorKind::InvalidConfig
                        : rc == QSM_EPARAM ? PipelineErrorKind::InvalidInput
                                           : PipelineErrorKind::AlgorithmError;
    return PipelineError{k, qsm_last_error_message()};
}

inline Result<std::vector<double>> run_dipole_inversion(const std::vector<double>& local_field_ppm,
                                                        const std::vector<uint8_t>& mask,
                                                        const ScanMetadata& md,
                                                        const InversionConfig& config,
                                                        Progress progress = nullptr) {
    Result<std::vector<double>> r{false, std::vector<double>(local_field_ppm.size()), {}};
    const qsm_scan_metadata m = md.c();
    const qsm_inversion_config c = config.c();
    const int rc = qsm_run_dipole_inversion_cfg(
        local_field_ppm.data(), local_field_ppm.size(), mask.data(), mask.size(), &m, &c,
        progress ? detail::trampoline : nullptr, progress ? &progress : nullptr, r.value.data());
    if (rc != QSM_OK) { r.error = to_error(rc); r.value.clear(); return r; }
    r.ok = true;
    return r;
}

inline Result<BgRemovalResult> run_bg_removal(const std::vector<double>& field_ppm,
                                              const std::vector<uint8_t>& mask,
                                              const ScanMetadata& md, const BgRemovalConfig& config,
                                              Progress progress = nullptr) {
    Result<BgRemovalResult> r{false, {std::vector<double>(field_ppm.size()),
                                      std::vector<uint8_t>(field_ppm.size())}, {}};
    const qsm_scan_metadata m = md.c();
    const qsm_bg_removal_config c = config.c();
    const int rc = qsm_run_bg_removal_cfg(field_ppm.data(), field_ppm.size(), mask.data(), mask.size(),
                                          &m, &c, progress ? detail::trampoline : nullptr,
                                          progress ? &progress : nullptr,
                                          r.value.local_field_ppm.data(), r.value.eroded_mask.data());
    if (rc != QSM_OK) { r.error = to_error(rc); r.value = {}; return r; }
    r.ok = true;
    return r;
}

}  // namespace pipeline
}  // namespace qsm
