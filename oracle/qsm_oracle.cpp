// CPU f64 oracle for the TV-ADMM / V-SHARP path of QSM.rs (qsm-core).
//
// TEST INFRASTRUCTURE ONLY: the checker the HIP path is compared against.
// Nothing in the product library links, loads or calls this file.
//
// Each function below restates the reference function named in its comment
// (file:line inside the qsm-core sources), keeping the reference's order of
// floating-point operations wherever that order is observable.  The one part
// the reference takes from a third-party crate -- the 1-D FFT, rustfft 6.4.1
// (Cargo.lock:326-329) -- is replaced by a mixed-radix decimation-in-time
// Cooley-Tukey transform written here; it computes the same DFT
// (X_k = sum_n x_n exp(-2 pi i k n / N), unnormalised forward, inverse scaled
// by 1/N in fft3d), so values agree with rustfft to rounding, not bit for bit.
//
// Parallel scope mirrors the reference's `parallel` feature exactly: OpenMP
// covers the three FFT axis loops and the 1/N scale (src/fft.rs:202-209,
// 423-479); every other pass is a serial loop, as in src/inversion/admm.rs and
// src/utils/gradient.rs.

#include "qsm_oracle.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

using cd = std::complex<double>;

// ---------------------------------------------------------------------------
// 1-D mixed-radix FFT (stand-in for rustfft's FftPlanner::plan_fft,
// src/fft.rs:65-89).
// ---------------------------------------------------------------------------
struct Plan1d {
    size_t n = 0;
    std::vector<size_t> fac;  // radices, outermost decomposition first
    std::vector<cd> w;        // w[j] = exp(-2 pi i j / n)

    explicit Plan1d(size_t n_) : n(n_) {
        size_t m = n;
        // Prefer radix 4, then the small primes, then any remaining factor.
        while (m % 4 == 0 && m > 4) { fac.push_back(4); m /= 4; }
        for (size_t p = 2; m > 1;) {
            if (m % p == 0) { fac.push_back(p); m /= p; }
            else { p = (p == 2) ? 3 : p + 2; }
        }
        w.resize(n);
        for (size_t j = 0; j < n; ++j) {
            // angle reduced exactly in integer arithmetic before scaling
            long double a = -2.0L * 3.14159265358979323846264338327950288L *
                            (long double)j / (long double)n;
            w[j] = cd((double)cosl(a), (double)sinl(a));
        }
    }

    // out[0..n) = DFT of in[0], in[s], in[2s], ... (length len), level lvl.
    void dit(const cd* in, size_t s, cd* out, size_t len, size_t lvl, bool inv) const {
        if (len == 1) { out[0] = in[0]; return; }
        const size_t p = fac[lvl];
        const size_t m = len / p;
        for (size_t r = 0; r < p; ++r) dit(in + r * s, s * p, out + r * m, m, lvl + 1, inv);
        const size_t step = n / len;  // W_len^e = w[e * step]
        cd t[64];
        std::vector<cd> tv;
        cd* tp = t;
        if (p > 64) { tv.resize(p); tp = tv.data(); }
        for (size_t k = 0; k < m; ++k) {
            for (size_t r = 0; r < p; ++r) {
                cd v = out[r * m + k];
                if (r && k) {
                    cd tw = w[(r * k * step) % n];
                    if (inv) tw = std::conj(tw);
                    v *= tw;
                }
                tp[r] = v;
            }
            if (p == 2) {
                out[k] = tp[0] + tp[1];
                out[m + k] = tp[0] - tp[1];
            } else if (p == 4) {
                cd a = tp[0] + tp[2], b = tp[0] - tp[2];
                cd c = tp[1] + tp[3], d = tp[1] - tp[3];
                // -i*d for the forward transform, +i*d for the inverse
                cd jd = inv ? cd(-d.imag(), d.real()) : cd(d.imag(), -d.real());
                out[k] = a + c;
                out[m + k] = b + jd;
                out[2 * m + k] = a - c;
                out[3 * m + k] = b - jd;
            } else {
                const size_t pstep = n / p;
                for (size_t q = 0; q < p; ++q) {
                    cd acc = tp[0];
                    for (size_t r = 1; r < p; ++r) {
                        cd tw = w[((r * q) % p) * pstep];
                        if (inv) tw = std::conj(tw);
                        acc += tp[r] * tw;
                    }
                    out[q * m + k] = acc;
                }
            }
        }
    }

    // In place on a contiguous line; tmp must hold n values.
    void exec(cd* line, cd* tmp, bool inv) const {
        dit(line, 1, tmp, n, 0, inv);
        std::memcpy(line, tmp, n * sizeof(cd));
    }
};

std::mutex g_plan_mu;
std::map<size_t, std::shared_ptr<Plan1d>> g_plans;

std::shared_ptr<Plan1d> plan_for(size_t n) {
    std::lock_guard<std::mutex> lk(g_plan_mu);
    auto it = g_plans.find(n);
    if (it != g_plans.end()) return it->second;
    auto p = std::make_shared<Plan1d>(n);
    g_plans[n] = p;
    return p;
}

inline size_t idx3d(size_t i, size_t j, size_t k, size_t nx, size_t ny) {
    return i + j * nx + k * nx * ny;  // src/fft.rs:236-241
}

// src/fft.rs:423-479 (transform_3d_par) followed, for the inverse, by the
// 1/N division of src/fft.rs:202-209.
void fft3d(cd* data, size_t nx, size_t ny, size_t nz, bool inv) {
    auto px = plan_for(nx), py = plan_for(ny), pz = plan_for(nz);
    const size_t nxy = nx * ny;
    const long nrows = (long)(ny * nz);
    // x-axis: contiguous rows
#pragma omp parallel
    {
        std::vector<cd> tmp(nx);
#pragma omp for schedule(static)
        for (long r = 0; r < nrows; ++r) px->exec(data + (size_t)r * nx, tmp.data(), inv);
    }
    // y-axis: one task per z-slab
#pragma omp parallel
    {
        std::vector<cd> buf(ny), tmp(ny);
#pragma omp for schedule(static)
        for (long k = 0; k < (long)nz; ++k) {
            for (size_t i = 0; i < nx; ++i) {
                for (size_t j = 0; j < ny; ++j) buf[j] = data[i + j * nx + k * nxy];
                py->exec(buf.data(), tmp.data(), inv);
                for (size_t j = 0; j < ny; ++j) data[i + j * nx + k * nxy] = buf[j];
            }
        }
    }
    // z-axis: one task per fixed-j plane
#pragma omp parallel
    {
        std::vector<cd> buf(nz), tmp(nz);
#pragma omp for schedule(static)
        for (long j = 0; j < (long)ny; ++j) {
            for (size_t i = 0; i < nx; ++i) {
                for (size_t k = 0; k < nz; ++k) buf[k] = data[i + j * nx + k * nxy];
                pz->exec(buf.data(), tmp.data(), inv);
                for (size_t k = 0; k < nz; ++k) data[i + j * nx + k * nxy] = buf[k];
            }
        }
    }
    if (inv) {
        const double n_total = (double)(nx * ny * nz);
        const long n = (long)(nx * ny * nz);
#pragma omp parallel for schedule(static)
        for (long i = 0; i < n; ++i) data[i] /= n_total;
    }
}

std::vector<cd> to_complex(const double* x, size_t n) {
    std::vector<cd> c(n);
    for (size_t i = 0; i < n; ++i) c[i] = cd(x[i], 0.0);
    return c;
}

// src/fft.rs:514-520 (fft_real_kernel): FFT, keep real parts.
std::vector<double> fft_real_kernel(const double* kern, size_t nx, size_t ny, size_t nz) {
    size_t n = nx * ny * nz;
    auto c = to_complex(kern, n);
    fft3d(c.data(), nx, ny, nz, false);
    std::vector<double> out(n);
    for (size_t i = 0; i < n; ++i) out[i] = c[i].real();
    return out;
}

}  // namespace

extern "C" {

void orc_set_num_threads(int n) {
#ifdef _OPENMP
    omp_set_num_threads(n > 0 ? n : 1);
#else
    (void)n;
#endif
}

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

// src/fft.rs:544-566
void orc_fftfreq(size_t n, double d, double* freq) {
    double val = 1.0 / ((double)n * d);
    if (n % 2 == 0) {
        for (size_t i = 0; i < n / 2; ++i) freq[i] = (double)i * val;
        for (size_t i = n / 2; i < n; ++i) freq[i] = (double)((long long)i - (long long)n) * val;
    } else {
        for (size_t i = 0; i <= (n - 1) / 2; ++i) freq[i] = (double)i * val;
        for (size_t i = (n + 1) / 2; i < n; ++i) freq[i] = (double)((long long)i - (long long)n) * val;
    }
}

void orc_fft3d(double* data, size_t nx, size_t ny, size_t nz, int inverse) {
    fft3d(reinterpret_cast<cd*>(data), nx, ny, nz, inverse != 0);
}

void orc_fft1d(double* data, size_t n, int inverse) {
    auto p = plan_for(n);
    std::vector<cd> tmp(n);
    p->exec(reinterpret_cast<cd*>(data), tmp.data(), inverse != 0);
}

// src/kernels/dipole.rs:24-72
void orc_dipole_kernel(size_t nx, size_t ny, size_t nz, double vsx, double vsy,
                       double vsz, double bx, double by, double bz, double* d) {
    std::vector<double> kx(nx), ky(ny), kz(nz);
    orc_fftfreq(nx, vsx, kx.data());
    orc_fftfreq(ny, vsy, ky.data());
    orc_fftfreq(nz, vsz, kz.data());
    double bnorm = std::sqrt(bx * bx + by * by + bz * bz);
    bx = bx / bnorm; by = by / bnorm; bz = bz / bnorm;
    const double one_third = 1.0 / 3.0;
    for (size_t k = 0; k < nz; ++k) {
        double kz_val = kz[k];
        for (size_t j = 0; j < ny; ++j) {
            double ky_val = ky[j];
            for (size_t i = 0; i < nx; ++i) {
                double kx_val = kx[i];
                double k_dot_b = kx_val * bx + ky_val * by + kz_val * bz;
                double k_squared = kx_val * kx_val + ky_val * ky_val + kz_val * kz_val;
                size_t idx = i + j * nx + k * nx * ny;
                d[idx] = (k_squared > 1e-20) ? one_third - (k_dot_b * k_dot_b) / k_squared : 0.0;
            }
        }
    }
}

// src/kernels/laplacian.rs:20-62
void orc_laplacian_kernel(size_t nx, size_t ny, size_t nz, double vsx, double vsy,
                          double vsz, int negative, double* l) {
    size_t n = nx * ny * nz;
    std::fill(l, l + n, 0.0);
    double hx = 1.0 / (vsx * vsx), hy = 1.0 / (vsy * vsy), hz = 1.0 / (vsz * vsz);
    double center = -2.0 * (hx + hy + hz);
    double sign = negative ? -1.0 : 1.0;
    l[0] = sign * center;
    if (nx > 1) { l[1] = sign * hx; l[nx - 1] = sign * hx; }
    if (ny > 1) { l[nx] = sign * hy; l[(ny - 1) * nx] = sign * hy; }
    if (nz > 1) { l[nx * ny] = sign * hz; l[(nz - 1) * nx * ny] = sign * hz; }
}

// src/kernels/smv.rs:19-67
double orc_smv_kernel(size_t nx, size_t ny, size_t nz, double vsx, double vsy,
                      double vsz, double radius, double* s) {
    size_t n = nx * ny * nz;
    std::fill(s, s + n, 0.0);
    double r_squared = radius * radius;
    double count = 0.0;
    for (size_t k = 0; k < nz; ++k) {
        double zk = (k <= nz / 2) ? (double)k : (double)((long long)k - (long long)nz);
        double dz = zk * vsz;
        for (size_t j = 0; j < ny; ++j) {
            double yj = (j <= ny / 2) ? (double)j : (double)((long long)j - (long long)ny);
            double dy = yj * vsy;
            for (size_t i = 0; i < nx; ++i) {
                double xi = (i <= nx / 2) ? (double)i : (double)((long long)i - (long long)nx);
                double dx = xi * vsx;
                double dist_sq = dx * dx + dy * dy + dz * dz;
                if (dist_sq <= r_squared) { s[i + j * nx + k * nx * ny] = 1.0; count += 1.0; }
            }
        }
    }
    if (count > 0.0) {
        double norm = 1.0 / count;
        for (size_t i = 0; i < n; ++i) s[i] *= norm;
    }
    return count;
}

// src/kernels/smv.rs:79-105
void orc_erode_mask_smv(const uint8_t* mask, const double* s_fft, size_t nx,
                        size_t ny, size_t nz, double delta, uint8_t* out) {
    size_t n = nx * ny * nz;
    std::vector<cd> mc(n);
    for (size_t i = 0; i < n; ++i) mc[i] = cd((double)mask[i], 0.0);
    fft3d(mc.data(), nx, ny, nz, false);
    for (size_t i = 0; i < n; ++i) mc[i] *= s_fft[i];
    fft3d(mc.data(), nx, ny, nz, true);
    for (size_t i = 0; i < n; ++i) out[i] = (mc[i].real() > delta) ? 1 : 0;
}

// src/utils/gradient.rs:21-58
void orc_fgrad(double* gx, double* gy, double* gz, const double* x, size_t nx,
               size_t ny, size_t nz, double vsx, double vsy, double vsz) {
    double hx = 1.0 / vsx, hy = 1.0 / vsy, hz = 1.0 / vsz;
    for (size_t k = 0; k < nz; ++k) {
        size_t kp1 = (k + 1 < nz) ? k + 1 : 0;
        size_t k_off = k * nx * ny, kp1_off = kp1 * nx * ny;
        for (size_t j = 0; j < ny; ++j) {
            size_t jp1 = (j + 1 < ny) ? j + 1 : 0;
            size_t j_off = j * nx, jp1_off = jp1 * nx;
            for (size_t i = 0; i < nx; ++i) {
                size_t ip1 = (i + 1 < nx) ? i + 1 : 0;
                size_t idx = i + j_off + k_off;
                double x_val = x[idx];
                gx[idx] = (x[ip1 + j_off + k_off] - x_val) * hx;
                gy[idx] = (x[i + jp1_off + k_off] - x_val) * hy;
                gz[idx] = (x[i + j_off + kp1_off] - x_val) * hz;
            }
        }
    }
}

// src/utils/gradient.rs:71-108
void orc_bdiv(double* div, const double* gx, const double* gy, const double* gz,
              size_t nx, size_t ny, size_t nz, double vsx, double vsy, double vsz) {
    double hx = 1.0 / vsx, hy = 1.0 / vsy, hz = 1.0 / vsz;
    for (size_t k = 0; k < nz; ++k) {
        size_t km1 = (k == 0) ? nz - 1 : k - 1;
        size_t k_off = k * nx * ny, km1_off = km1 * nx * ny;
        for (size_t j = 0; j < ny; ++j) {
            size_t jm1 = (j == 0) ? ny - 1 : j - 1;
            size_t j_off = j * nx, jm1_off = jm1 * nx;
            for (size_t i = 0; i < nx; ++i) {
                size_t im1 = (i == 0) ? nx - 1 : i - 1;
                size_t idx = i + j_off + k_off;
                div[idx] = (gx[idx] - gx[im1 + j_off + k_off]) * hx +
                           (gy[idx] - gy[i + jm1_off + k_off]) * hy +
                           (gz[idx] - gz[i + j_off + km1_off]) * hz;
            }
        }
    }
}

// src/utils/mod.rs:44-52
double orc_shrink(double x, double threshold) {
    if (x > threshold) return x - threshold;
    if (x < -threshold) return x + threshold;
    return 0.0;
}

// src/utils/mod.rs:69-77
double orc_relative_change(const double* x, const double* x_prev, size_t n) {
    double nd = 0.0, nx2 = 0.0;
    for (size_t i = 0; i < n; ++i) {
        double diff = x[i] - x_prev[i];
        nd += diff * diff;
        nx2 += x[i] * x[i];
    }
    return std::sqrt(nd) / (std::sqrt(nx2) + 1e-20);
}

// src/utils/mask.rs:47-53
void orc_apply_mask_zero(double* data, const uint8_t* mask, size_t n) {
    for (size_t i = 0; i < n; ++i)
        if (mask[i] == 0) data[i] = 0.0;
}

// src/inversion/admm.rs:150-178
void orc_prepare_admm_spectral(const double* local_field, size_t nx, size_t ny,
                               size_t nz, double vsx, double vsy, double vsz,
                               double bx, double by, double bz, double rho,
                               double* inv_a, double* f_hat_out) {
    size_t n = nx * ny * nz;
    std::vector<double> d(n), l(n);
    orc_dipole_kernel(nx, ny, nz, vsx, vsy, vsz, bx, by, bz, d.data());
    orc_laplacian_kernel(nx, ny, nz, vsx, vsy, vsz, 1, l.data());
    auto lc = to_complex(l.data(), n);
    fft3d(lc.data(), nx, ny, nz, false);
    for (size_t i = 0; i < n; ++i) {
        double a = d[i] * d[i] + rho * lc[i].real();
        inv_a[i] = (std::fabs(a) > 1e-20) ? 1.0 / a : 0.0;
    }
    cd* f_hat = reinterpret_cast<cd*>(f_hat_out);
    for (size_t i = 0; i < n; ++i) f_hat[i] = cd(local_field[i], 0.0);
    fft3d(f_hat, nx, ny, nz, false);
    for (size_t i = 0; i < n; ++i) f_hat[i] = f_hat[i] * d[i] * inv_a[i];
}

// src/inversion/tv.rs:63-97, with admm_step inlined from
// src/inversion/admm.rs:76-145 and AdmmBuffers from admm.rs:18-53.
void orc_tv_admm(const double* local_field, const uint8_t* mask, size_t nx,
                 size_t ny, size_t nz, double vsx, double vsy, double vsz,
                 double bx, double by, double bz, double lambda, double rho,
                 double tol, size_t max_iter, orc_progress_fn progress,
                 void* user, double* chi_out, size_t* iters_out,
                 double* relchange_out) {
    const size_t n = nx * ny * nz;
    std::vector<double> inv_a(n);
    std::vector<cd> f_hat(n);
    orc_prepare_admm_spectral(local_field, nx, ny, nz, vsx, vsy, vsz, bx, by, bz,
                              rho, inv_a.data(), reinterpret_cast<double*>(f_hat.data()));
    std::vector<double> x(n, 0.0), x_prev(n, 0.0), ux(n, 0.0), uy(n, 0.0), uz(n, 0.0),
        gx(n, 0.0), gy(n, 0.0), gz(n, 0.0), div(n, 0.0);
    std::vector<cd> work(n, cd(0.0, 0.0));
    const double t = lambda / rho;
    size_t iters = 0;
    double rel = NAN;
    for (size_t iter = 0; iter < max_iter; ++iter) {
        if (progress) progress(iter + 1, max_iter, user);
        iters = iter + 1;
        std::swap(x, x_prev);
        orc_bdiv(div.data(), gx.data(), gy.data(), gz.data(), nx, ny, nz, vsx, vsy, vsz);
        for (size_t i = 0; i < n; ++i) work[i] = cd(div[i], 0.0);
        fft3d(work.data(), nx, ny, nz, false);
        for (size_t i = 0; i < n; ++i) work[i] = f_hat[i] - rho * work[i] * inv_a[i];
        fft3d(work.data(), nx, ny, nz, true);
        for (size_t i = 0; i < n; ++i) x[i] = work[i].real();
        rel = orc_relative_change(x.data(), x_prev.data(), n);
        if (rel < tol) {
            if (progress) progress(iter + 1, iter + 1, user);
            break;
        }
        orc_fgrad(gx.data(), gy.data(), gz.data(), x.data(), nx, ny, nz, vsx, vsy, vsz);
        for (size_t i = 0; i < n; ++i) {
            double vx = gx[i] + ux[i], vy = gy[i] + uy[i], vz = gz[i] + uz[i];
            double zx = orc_shrink(vx, t), zy = orc_shrink(vy, t), zz = orc_shrink(vz, t);
            ux[i] = vx - zx; uy[i] = vy - zy; uz[i] = vz - zz;
            gx[i] = 2.0 * zx - vx; gy[i] = 2.0 * zy - vy; gz[i] = 2.0 * zz - vz;
        }
    }
    orc_apply_mask_zero(x.data(), mask, n);
    std::memcpy(chi_out, x.data(), n * sizeof(double));
    if (iters_out) *iters_out = iters;
    if (relchange_out) *relchange_out = rel;
}

// Adjoint of the unrolled loop of orc_tv_admm (tol = 0).  Iteration j:
//   c_j = bdiv(g_{j-1}), g = z - u;  x_j = Re IFFT(f_hat - rho FFT(c_j) inv_a)
//   v_j = fgrad(x_j) + u_{j-1};  z_j = shrink(v_j, t);  u_j = v_j - z_j
// (src/inversion/admm.rs:97-142).  With q_j = sign(v_j) [|v_j| > t],
// A = Re IFFT(D inv_a FFT .) and M = Re IFFT(inv_a FFT .) (both symmetric
// for real even multipliers) and c-bar_j = -rho M x-bar_j:
//   x-bar_K = mask chi-bar,  s-bar_K = 0
//   z-bar_j = -fgrad(c-bar_{j+1}),  u-bar_j = s-bar_{j+1} + fgrad(c-bar_{j+1})
//   s-bar_j = |q_j| z-bar_j + (1 - |q_j|) u-bar_j,  x-bar_j = -bdiv(s-bar_j)
//   t-bar  += sum (u-bar_j - z-bar_j) q_j
//   field-bar = A sum_j x-bar_j,  lambda-bar = t-bar / rho.
void orc_tv_unrolled_vjp(const double* local_field, const uint8_t* mask, size_t nx,
                         size_t ny, size_t nz, double vsx, double vsy, double vsz,
                         double bx, double by, double bz, double lambda, double rho,
                         size_t K, const double* chi_bar, double* chi_out,
                         double* field_bar, double* lambda_bar) {
    const size_t n = nx * ny * nz;
    std::vector<double> inv_a(n), d(n);
    std::vector<cd> f_hat(n);
    orc_prepare_admm_spectral(local_field, nx, ny, nz, vsx, vsy, vsz, bx, by, bz, rho,
                              inv_a.data(), reinterpret_cast<double*>(f_hat.data()));
    orc_dipole_kernel(nx, ny, nz, vsx, vsy, vsz, bx, by, bz, d.data());
    // forward, recording q_j for j = 1..K-1
    std::vector<double> x(n, 0.0), ux(n, 0.0), uy(n, 0.0), uz(n, 0.0), gx(n, 0.0), gy(n, 0.0),
        gz(n, 0.0), div(n, 0.0);
    std::vector<cd> work(n);
    std::vector<std::vector<int8_t>> q(K > 1 ? K - 1 : 0, std::vector<int8_t>(3 * n));
    const double t = lambda / rho;
    auto sgn = [t](double v) -> int8_t { return v > t ? 1 : (v < -t ? -1 : 0); };
    for (size_t j = 1; j <= K; ++j) {
        orc_bdiv(div.data(), gx.data(), gy.data(), gz.data(), nx, ny, nz, vsx, vsy, vsz);
        for (size_t i = 0; i < n; ++i) work[i] = cd(div[i], 0.0);
        fft3d(work.data(), nx, ny, nz, false);
        for (size_t i = 0; i < n; ++i) work[i] = f_hat[i] - rho * work[i] * inv_a[i];
        fft3d(work.data(), nx, ny, nz, true);
        for (size_t i = 0; i < n; ++i) x[i] = work[i].real();
        if (j == K) break;
        orc_fgrad(gx.data(), gy.data(), gz.data(), x.data(), nx, ny, nz, vsx, vsy, vsz);
        int8_t* qj = q[j - 1].data();
        for (size_t i = 0; i < n; ++i) {
            double vx = gx[i] + ux[i], vy = gy[i] + uy[i], vz = gz[i] + uz[i];
            qj[i] = sgn(vx); qj[n + i] = sgn(vy); qj[2 * n + i] = sgn(vz);
            double zx = orc_shrink(vx, t), zy = orc_shrink(vy, t), zz = orc_shrink(vz, t);
            ux[i] = vx - zx; uy[i] = vy - zy; uz[i] = vz - zz;
            gx[i] = 2.0 * zx - vx; gy[i] = 2.0 * zy - vy; gz[i] = 2.0 * zz - vz;
        }
    }
    for (size_t i = 0; i < n; ++i) chi_out[i] = (K > 0 && mask[i]) ? x[i] : 0.0;
    // backward
    std::vector<double> xb(n), acc(n, 0.0), cb(n), sx(n, 0.0), sy(n, 0.0), sz(n, 0.0), cgx(n),
        cgy(n), cgz(n);
    double tbar = 0.0;
    auto apply = [&](const std::vector<double>& in, const double* mult, const double* mult2,
                     double scale, std::vector<double>& out) {
        for (size_t i = 0; i < n; ++i) work[i] = cd(in[i], 0.0);
        fft3d(work.data(), nx, ny, nz, false);
        for (size_t i = 0; i < n; ++i) work[i] *= scale * mult[i] * (mult2 ? mult2[i] : 1.0);
        fft3d(work.data(), nx, ny, nz, true);
        for (size_t i = 0; i < n; ++i) out[i] = work[i].real();
    };
    for (size_t j = K; j >= 1; --j) {
        if (j == K) {
            for (size_t i = 0; i < n; ++i) xb[i] = mask[i] ? chi_bar[i] : 0.0;
        } else {
            orc_fgrad(cgx.data(), cgy.data(), cgz.data(), cb.data(), nx, ny, nz, vsx, vsy, vsz);
            const int8_t* qj = q[j - 1].data();
            for (size_t i = 0; i < n; ++i) {
                const double g3[3] = {cgx[i], cgy[i], cgz[i]};
                double* s3[3] = {&sx[i], &sy[i], &sz[i]};
                for (int a = 0; a < 3; ++a) {
                    const double zb = -g3[a], ub = *s3[a] + g3[a];
                    const int qa = qj[a * n + i];
                    tbar += (ub - zb) * (double)qa;
                    *s3[a] = qa != 0 ? zb : ub;
                }
            }
            orc_bdiv(xb.data(), sx.data(), sy.data(), sz.data(), nx, ny, nz, vsx, vsy, vsz);
            for (size_t i = 0; i < n; ++i) xb[i] = -xb[i];
        }
        for (size_t i = 0; i < n; ++i) acc[i] += xb[i];
        if (j > 1) apply(xb, inv_a.data(), nullptr, -rho, cb);
    }
    if (K > 0) {
        apply(acc, d.data(), inv_a.data(), 1.0, xb);
        std::memcpy(field_bar, xb.data(), n * sizeof(double));
    } else {
        std::fill(field_bar, field_bar + n, 0.0);
    }
    if (lambda_bar) *lambda_bar = tbar / rho;
}

// src/inversion/tkd.rs:48-99: 1/D, or sign(D)/thr where |D| <= thr (tkd);
// 0 there (tsvd, the function after it).
void orc_tkd(const double* local_field, const uint8_t* mask, size_t nx, size_t ny, size_t nz,
             double vsx, double vsy, double vsz, double bx, double by, double bz,
             double threshold, int tsvd, double* chi_out) {
    const size_t n = nx * ny * nz;
    std::vector<double> d(n);
    orc_dipole_kernel(nx, ny, nz, vsx, vsy, vsz, bx, by, bz, d.data());
    const double inv_t = 1.0 / threshold;
    auto fc = to_complex(local_field, n);
    fft3d(fc.data(), nx, ny, nz, false);
    for (size_t i = 0; i < n; ++i) {
        double m;
        if (std::fabs(d[i]) <= threshold) m = tsvd ? 0.0 : (d[i] >= 0.0 ? inv_t : -inv_t);
        else m = 1.0 / d[i];
        fc[i] *= m;
    }
    fft3d(fc.data(), nx, ny, nz, true);
    for (size_t i = 0; i < n; ++i) chi_out[i] = fc[i].real();
    orc_apply_mask_zero(chi_out, mask, n);
}

// src/inversion/tikhonov.rs:62-131: D / (D^2 + lambda Gamma), Gamma = 1,
// FFT of the negative Laplacian stencil, or the square of the FFT of the
// Laplacian stencil.
void orc_tikhonov(const double* local_field, const uint8_t* mask, size_t nx, size_t ny,
                  size_t nz, double vsx, double vsy, double vsz, double bx, double by,
                  double bz, double lambda, int reg, double* chi_out) {
    const size_t n = nx * ny * nz;
    std::vector<double> d(n), gamma(n, 1.0);
    orc_dipole_kernel(nx, ny, nz, vsx, vsy, vsz, bx, by, bz, d.data());
    if (reg == 1 || reg == 2) {
        std::vector<double> l(n);
        orc_laplacian_kernel(nx, ny, nz, vsx, vsy, vsz, reg == 1 ? 1 : 0, l.data());
        auto lc = to_complex(l.data(), n);
        fft3d(lc.data(), nx, ny, nz, false);
        for (size_t i = 0; i < n; ++i)
            gamma[i] = reg == 1 ? lc[i].real() : lc[i].real() * lc[i].real();
    }
    auto fc = to_complex(local_field, n);
    fft3d(fc.data(), nx, ny, nz, false);
    for (size_t i = 0; i < n; ++i) {
        const double denom = d[i] * d[i] + lambda * gamma[i];
        fc[i] *= (std::fabs(denom) > 1e-20) ? d[i] / denom : 0.0;
    }
    fft3d(fc.data(), nx, ny, nz, true);
    for (size_t i = 0; i < n; ++i) chi_out[i] = fc[i].real();
    orc_apply_mask_zero(chi_out, mask, n);
}

// src/inversion/rts.rs:71-212.  Step 1: per-frequency refinement in the
// well-conditioned cone |D| > delta; x0 = masked Re IFFT.  Step 2: ADMM
// (admm_step with rho = 1 and inv_a = rho / (M + rho L^), M = mu on the
// cone), shrink threshold 1/rho, starting from x0 with u = g = 0.
void orc_rts(const double* local_field, const uint8_t* mask, size_t nx, size_t ny, size_t nz,
             double vsx, double vsy, double vsz, double bx, double by, double bz,
             double delta, double mu, double rho, double tol, size_t max_iter,
             size_t lsmr_iter, orc_progress_fn progress, void* user, double* chi_out,
             size_t* iters_out) {
    const size_t n = nx * ny * nz;
    std::vector<double> dk(n), lk(n);
    orc_dipole_kernel(nx, ny, nz, vsx, vsy, vsz, bx, by, bz, dk.data());
    orc_laplacian_kernel(nx, ny, nz, vsx, vsy, vsz, 1, lk.data());
    std::vector<cd> work = to_complex(lk.data(), n);
    fft3d(work.data(), nx, ny, nz, false);
    std::vector<double> m_mask(n, 0.0), inv_a(n, 0.0);
    for (size_t i = 0; i < n; ++i) {
        const double l_fft = work[i].real();
        if (std::fabs(dk[i]) > delta) m_mask[i] = mu;
        const double a = m_mask[i] + rho * l_fft;
        if (std::fabs(a) > 1e-20) inv_a[i] = rho / a;
    }
    for (size_t i = 0; i < n; ++i) work[i] = cd(local_field[i], 0.0);
    fft3d(work.data(), nx, ny, nz, false);
    const std::vector<cd> field_fft = work;
    for (size_t i = 0; i < n; ++i) {
        const double d = dk[i];
        work[i] = (std::fabs(d) > delta) ? field_fft[i] * d / (d * d + 1e-6) : cd(0.0, 0.0);
    }
    std::vector<cd> residual(n);
    for (size_t it = 0; it < lsmr_iter; ++it) {
        for (size_t i = 0; i < n; ++i) residual[i] = field_fft[i] - work[i] * dk[i];
        for (size_t i = 0; i < n; ++i) {
            const double d = dk[i];
            if (std::fabs(d) > delta) work[i] += residual[i] * d / (d * d + 1e-6);
        }
    }
    fft3d(work.data(), nx, ny, nz, true);
    std::vector<double> x(n);
    for (size_t i = 0; i < n; ++i) x[i] = mask[i] ? work[i].real() : 0.0;
    for (size_t i = 0; i < n; ++i) work[i] = cd(x[i], 0.0);
    fft3d(work.data(), nx, ny, nz, false);
    std::vector<cd> f_hat(n, cd(0.0, 0.0));
    for (size_t i = 0; i < n; ++i)
        if (std::fabs(m_mask[i]) > 1e-20 && std::fabs(inv_a[i]) > 1e-20)
            f_hat[i] = work[i] * (m_mask[i] / rho) * inv_a[i];
    std::vector<double> x_prev(n, 0.0), ux(n, 0.0), uy(n, 0.0), uz(n, 0.0), gx(n, 0.0),
        gy(n, 0.0), gz(n, 0.0), div(n, 0.0);
    const double t = 1.0 / rho;
    size_t iters = 0;
    for (size_t iter = 0; iter < max_iter; ++iter) {
        if (progress) progress(iter + 1, max_iter, user);
        iters = iter + 1;
        std::swap(x, x_prev);
        orc_bdiv(div.data(), gx.data(), gy.data(), gz.data(), nx, ny, nz, vsx, vsy, vsz);
        for (size_t i = 0; i < n; ++i) work[i] = cd(div[i], 0.0);
        fft3d(work.data(), nx, ny, nz, false);
        for (size_t i = 0; i < n; ++i) work[i] = f_hat[i] - 1.0 * work[i] * inv_a[i];
        fft3d(work.data(), nx, ny, nz, true);
        for (size_t i = 0; i < n; ++i) x[i] = work[i].real();
        if (orc_relative_change(x.data(), x_prev.data(), n) < tol) {
            if (progress) progress(iter + 1, iter + 1, user);
            break;
        }
        orc_fgrad(gx.data(), gy.data(), gz.data(), x.data(), nx, ny, nz, vsx, vsy, vsz);
        for (size_t i = 0; i < n; ++i) {
            double vx = gx[i] + ux[i], vy = gy[i] + uy[i], vz = gz[i] + uz[i];
            double zx = orc_shrink(vx, t), zy = orc_shrink(vy, t), zz = orc_shrink(vz, t);
            ux[i] = vx - zx; uy[i] = vy - zy; uz[i] = vz - zz;
            gx[i] = 2.0 * zx - vx; gy[i] = 2.0 * zy - vy; gz[i] = 2.0 * zz - vz;
        }
    }
    orc_apply_mask_zero(x.data(), mask, n);
    std::memcpy(chi_out, x.data(), n * sizeof(double));
    if (iters_out) *iters_out = iters;
}

// src/inversion/tikhonov.rs:62-131 (Regularization::Gradient arm)
void orc_tikhonov_gradient(const double* local_field, const uint8_t* mask,
                           size_t nx, size_t ny, size_t nz, double vsx,
                           double vsy, double vsz, double bx, double by,
                           double bz, double lambda, double* chi_out) {
    size_t n = nx * ny * nz;
    std::vector<double> d(n), l(n);
    orc_dipole_kernel(nx, ny, nz, vsx, vsy, vsz, bx, by, bz, d.data());
    orc_laplacian_kernel(nx, ny, nz, vsx, vsy, vsz, 1, l.data());
    auto lc = to_complex(l.data(), n);
    fft3d(lc.data(), nx, ny, nz, false);
    std::vector<double> inv_d(n);
    for (size_t i = 0; i < n; ++i) {
        double denom = d[i] * d[i] + lambda * lc[i].real();
        inv_d[i] = (std::fabs(denom) > 1e-20) ? d[i] / denom : 0.0;
    }
    auto fc = to_complex(local_field, n);
    fft3d(fc.data(), nx, ny, nz, false);
    for (size_t i = 0; i < n; ++i) fc[i] *= inv_d[i];
    fft3d(fc.data(), nx, ny, nz, true);
    for (size_t i = 0; i < n; ++i) chi_out[i] = fc[i].real();
    orc_apply_mask_zero(chi_out, mask, n);
}

// src/bgremove/vsharp.rs:70-82
size_t orc_vsharp_radii(double max_radius, double min_radius, double* out, size_t cap) {
    size_t cnt = 0;
    double r = max_radius;
    double step = min_radius;
    if (!(step > 0.0)) {  // the reference loops forever here; refuse instead
        if (cap > 0) out[0] = max_radius;
        return 1;
    }
    while (r >= step) {
        if (cnt < cap) out[cnt] = r;
        ++cnt;
        r -= step;
    }
    if (cnt == 0) {
        if (cap > 0) out[0] = max_radius;
        cnt = 1;
    }
    return cnt;
}

// src/bgremove/sharp.rs:64-131
void orc_sharp_core(const double* field, const uint8_t* mask, size_t nx,
                    size_t ny, size_t nz, double vsx, double vsy, double vsz,
                    double threshold, double radius, double* local_out,
                    uint8_t* eroded_mask) {
    size_t n = nx * ny * nz;
    std::vector<double> s(n);
    orc_smv_kernel(nx, ny, nz, vsx, vsy, vsz, radius, s.data());
    auto s_fft = fft_real_kernel(s.data(), nx, ny, nz);
    orc_erode_mask_smv(mask, s_fft.data(), nx, ny, nz, 1.0 - std::sqrt(1e-7), eroded_mask);
    auto fc = to_complex(field, n);
    fft3d(fc.data(), nx, ny, nz, false);
    for (size_t i = 0; i < n; ++i) fc[i] *= 1.0 - s_fft[i];
    fft3d(fc.data(), nx, ny, nz, true);
    for (size_t i = 0; i < n; ++i)
        if (eroded_mask[i] == 0) fc[i] = cd(0.0, 0.0);
    fft3d(fc.data(), nx, ny, nz, false);
    for (size_t i = 0; i < n; ++i) {
        double one_minus_s = 1.0 - s_fft[i];
        if (std::fabs(one_minus_s) < threshold) fc[i] = cd(0.0, 0.0);
        else fc[i] /= one_minus_s;
    }
    fft3d(fc.data(), nx, ny, nz, true);
    for (size_t i = 0; i < n; ++i) local_out[i] = (eroded_mask[i] == 1) ? fc[i].real() : 0.0;
}

// src/bgremove/vsharp.rs:88-199
void orc_vsharp_with_radii(const double* field, const uint8_t* mask, size_t nx,
                           size_t ny, size_t nz, double vsx, double vsy,
                           double vsz, const double* radii, size_t n_radii,
                           double threshold, orc_progress_fn progress,
                           void* user, double* local_out, uint8_t* mask_out) {
    const size_t n = nx * ny * nz;
    if (n_radii == 0) {
        std::fill(local_out, local_out + n, 0.0);
        std::memcpy(mask_out, mask, n);
        return;
    }
    if (n_radii == 1) {
        if (progress) progress(1, 1, user);
        orc_sharp_core(field, mask, nx, ny, nz, vsx, vsy, vsz, threshold, radii[0],
                       local_out, mask_out);
        return;
    }
    std::vector<double> sorted(radii, radii + n_radii);
    std::stable_sort(sorted.begin(), sorted.end(), [](double a, double b) { return b < a; });

    auto field_fft = to_complex(field, n);
    fft3d(field_fft.data(), nx, ny, nz, false);

    std::vector<uint8_t> processed(n, 0), final_mask(n, 0), eroded(n);
    std::vector<double> local(n, 0.0), s(n), inverse_kernel;
    const double delta = 1.0 - std::sqrt(1e-7);
    std::vector<cd> filtered(n);

    for (size_t idx = 0; idx < n_radii; ++idx) {
        if (progress) progress(idx + 1, n_radii, user);
        orc_smv_kernel(nx, ny, nz, vsx, vsy, vsz, sorted[idx], s.data());
        auto s_fft = fft_real_kernel(s.data(), nx, ny, nz);
        if (inverse_kernel.empty()) {
            inverse_kernel.resize(n);
            for (size_t i = 0; i < n; ++i) {
                double one_minus_s = 1.0 - s_fft[i];
                inverse_kernel[i] = (std::fabs(one_minus_s) < threshold) ? 0.0 : 1.0 / one_minus_s;
            }
        }
        orc_erode_mask_smv(mask, s_fft.data(), nx, ny, nz, delta, eroded.data());
        filtered = field_fft;
        for (size_t i = 0; i < n; ++i) filtered[i] *= 1.0 - s_fft[i];
        fft3d(filtered.data(), nx, ny, nz, true);
        for (size_t i = 0; i < n; ++i) {
            if (eroded[i] == 1 && !processed[i]) {
                local[i] = filtered[i].real();
                processed[i] = 1;
                final_mask[i] = 1;
            }
        }
    }
    auto lc = to_complex(local.data(), n);
    fft3d(lc.data(), nx, ny, nz, false);
    for (size_t i = 0; i < n; ++i) lc[i] *= inverse_kernel[i];
    fft3d(lc.data(), nx, ny, nz, true);
    for (size_t i = 0; i < n; ++i) local_out[i] = (final_mask[i] == 1) ? lc[i].real() : 0.0;
    std::memcpy(mask_out, final_mask.data(), n);
}

// src/bgremove/vsharp.rs:55-64
void orc_vsharp(const double* field, const uint8_t* mask, size_t nx, size_t ny,
                size_t nz, double vsx, double vsy, double vsz, double threshold,
                double max_radius, double min_radius, orc_progress_fn progress,
                void* user, double* local_out, uint8_t* mask_out) {
    size_t cnt = orc_vsharp_radii(max_radius, min_radius, nullptr, 0);
    std::vector<double> radii(cnt);
    orc_vsharp_radii(max_radius, min_radius, radii.data(), cnt);
    orc_vsharp_with_radii(field, mask, nx, ny, nz, vsx, vsy, vsz, radii.data(), cnt,
                          threshold, progress, user, local_out, mask_out);
}

// tests/chisep_integration.rs:18-62 (dipole convolution part)
void orc_dipole_forward(const double* chi, size_t nx, size_t ny, size_t nz,
                        double vsx, double vsy, double vsz, double bx,
                        double by, double bz, double* out) {
    size_t n = nx * ny * nz;
    std::vector<double> d(n);
    orc_dipole_kernel(nx, ny, nz, vsx, vsy, vsz, bx, by, bz, d.data());
    auto c = to_complex(chi, n);
    fft3d(c.data(), nx, ny, nz, false);
    for (size_t i = 0; i < n; ++i) c[i] *= d[i];
    fft3d(c.data(), nx, ny, nz, true);
    for (size_t i = 0; i < n; ++i) out[i] = c[i].real();
}

// tests/common/mod.rs:17-31
double orc_rmse(const double* a, const double* b, const uint8_t* mask, size_t n) {
    double sum_sq = 0.0;
    size_t count = 0;
    for (size_t i = 0; i < n; ++i) {
        if (mask[i] > 0) {
            double diff = a[i] - b[i];
            sum_sq += diff * diff;
            ++count;
        }
    }
    if (count == 0) return 0.0;
    return std::sqrt(sum_sq / (double)count);
}

// tests/common/mod.rs:34-51
double orc_nrmse(const double* a, const double* b, const uint8_t* mask, size_t n) {
    double r = orc_rmse(a, b, mask, n);
    double mn = INFINITY, mx = -INFINITY;
    for (size_t i = 0; i < n; ++i) {
        if (mask[i] > 0) {
            if (b[i] < mn) mn = b[i];
            if (b[i] > mx) mx = b[i];
        }
    }
    double range = mx - mn;
    if (range == 0.0) return 0.0;
    return r / range;
}

}  // extern "C"

extern "C" {
// src/unwrap/laplacian.rs:41-50
double orc_wrap(double x) {
    double y = std::fmod(x, 2.0 * M_PI);
    if (y > M_PI) y -= 2.0 * M_PI;
    else if (y < -M_PI) y += 2.0 * M_PI;
    return y;
}

// src/unwrap/laplacian.rs:135-161 with wrapped_laplacian_periodic (:67-122)
// and solve_poisson_fft (:93-133).
void orc_laplacian_unwrap(const double* phase, const uint8_t* mask, size_t nx, size_t ny, size_t nz,
                          double vsx, double vsy, double vsz, double* out) {
    const size_t n = nx * ny * nz, nxy = nx * ny;
    const double dx2 = 1.0 / (vsx * vsx), dy2 = 1.0 / (vsy * vsy), dz2 = 1.0 / (vsz * vsz);
    std::vector<cd> f(n);
    for (size_t k = 0; k < nz; ++k) {
        const size_t km1 = k == 0 ? nz - 1 : k - 1, kp1 = k + 1 >= nz ? 0 : k + 1;
        for (size_t j = 0; j < ny; ++j) {
            const size_t jm1 = j == 0 ? ny - 1 : j - 1, jp1 = j + 1 >= ny ? 0 : j + 1;
            for (size_t i = 0; i < nx; ++i) {
                const size_t im1 = i == 0 ? nx - 1 : i - 1, ip1 = i + 1 >= nx ? 0 : i + 1;
                const size_t idx = i + j * nx + k * nxy;
                const double u = phase[idx];
                const double lx = (orc_wrap(phase[ip1 + j * nx + k * nxy] - u) -
                                   orc_wrap(u - phase[im1 + j * nx + k * nxy])) * dx2;
                const double ly = (orc_wrap(phase[i + jp1 * nx + k * nxy] - u) -
                                   orc_wrap(u - phase[i + jm1 * nx + k * nxy])) * dy2;
                const double lz = (orc_wrap(phase[i + j * nx + kp1 * nxy] - u) -
                                   orc_wrap(u - phase[i + j * nx + km1 * nxy])) * dz2;
                f[idx] = cd(mask[idx] != 0 ? lx + ly + lz : 0.0, 0.0);
            }
        }
    }
    fft3d(f.data(), nx, ny, nz, false);
    for (size_t k = 0; k < nz; ++k) {
        const double fk = k <= nz / 2 ? (double)k / nz : ((double)k - (double)nz) / nz;
        const double lz = 2.0 * (std::cos(2.0 * M_PI * fk) - 1.0) * dz2;
        for (size_t j = 0; j < ny; ++j) {
            const double fj = j <= ny / 2 ? (double)j / ny : ((double)j - (double)ny) / ny;
            const double ly = 2.0 * (std::cos(2.0 * M_PI * fj) - 1.0) * dy2;
            for (size_t i = 0; i < nx; ++i) {
                const double fi = i <= nx / 2 ? (double)i / nx : ((double)i - (double)nx) / nx;
                const double lx = 2.0 * (std::cos(2.0 * M_PI * fi) - 1.0) * dx2;
                const double lam = lx + ly + lz;
                const size_t idx = i + j * nx + k * nxy;
                if (std::fabs(lam) > 1e-20) f[idx] /= lam;
                else f[idx] = cd(0.0, 0.0);
            }
        }
    }
    fft3d(f.data(), nx, ny, nz, true);
    for (size_t i = 0; i < n; ++i) out[i] = mask[i] != 0 ? f[i].real() : 0.0;
}
}

extern "C" {
// src/inversion/nltv.rs:63-133: rho = 100 lambda; inner ADMM loops
// (admm_step with weighted_shrink, src/utils/mod.rs:55-59) that stop at
// convergence; after each, g is overwritten with grad(x) (which the next
// inner loop's first step then uses) and the weights become
// 1/(|grad x| + mu 1e-6), divided by their maximum when that exceeds 1.
void orc_nltv(const double* local_field, const uint8_t* mask, size_t nx, size_t ny, size_t nz,
              double vsx, double vsy, double vsz, double bx, double by, double bz, double lambda,
              double mu, double tol, size_t max_iter, size_t newton_iter, orc_progress_fn progress,
              void* user, double* chi_out) {
    const size_t n = nx * ny * nz;
    const double eps = 1e-6;
    const double rho = 100.0 * lambda;
    std::vector<double> inv_a(n);
    std::vector<cd> f_hat(n);
    orc_prepare_admm_spectral(local_field, nx, ny, nz, vsx, vsy, vsz, bx, by, bz, rho, inv_a.data(),
                              reinterpret_cast<double*>(f_hat.data()));
    std::vector<double> x(n, 0.0), x_prev(n, 0.0), ux(n, 0.0), uy(n, 0.0), uz(n, 0.0), gx(n, 0.0),
        gy(n, 0.0), gz(n, 0.0), div(n, 0.0), w(n, 1.0);
    std::vector<cd> work(n);
    const size_t total = max_iter * newton_iter;
    size_t current = 0;
    const double t = lambda / rho;
    for (size_t newton = 0; newton < newton_iter; ++newton) {
        for (size_t it = 0; it < max_iter; ++it) {
            ++current;
            if (progress) progress(current, total, user);
            std::swap(x, x_prev);
            orc_bdiv(div.data(), gx.data(), gy.data(), gz.data(), nx, ny, nz, vsx, vsy, vsz);
            for (size_t i = 0; i < n; ++i) work[i] = cd(div[i], 0.0);
            fft3d(work.data(), nx, ny, nz, false);
            for (size_t i = 0; i < n; ++i) work[i] = f_hat[i] - rho * work[i] * inv_a[i];
            fft3d(work.data(), nx, ny, nz, true);
            for (size_t i = 0; i < n; ++i) x[i] = work[i].real();
            if (orc_relative_change(x.data(), x_prev.data(), n) < tol) break;
            orc_fgrad(gx.data(), gy.data(), gz.data(), x.data(), nx, ny, nz, vsx, vsy, vsz);
            for (size_t i = 0; i < n; ++i) {
                const double ti = t * w[i];
                double vx = gx[i] + ux[i], vy = gy[i] + uy[i], vz = gz[i] + uz[i];
                double zx = orc_shrink(vx, ti), zy = orc_shrink(vy, ti), zz = orc_shrink(vz, ti);
                ux[i] = vx - zx; uy[i] = vy - zy; uz[i] = vz - zz;
                gx[i] = 2.0 * zx - vx; gy[i] = 2.0 * zy - vy; gz[i] = 2.0 * zz - vz;
            }
        }
        orc_fgrad(gx.data(), gy.data(), gz.data(), x.data(), nx, ny, nz, vsx, vsy, vsz);
        double wmax = 0.0;
        for (size_t i = 0; i < n; ++i) {
            const double gm = std::sqrt(gx[i] * gx[i] + gy[i] * gy[i] + gz[i] * gz[i]);
            w[i] = 1.0 / (gm + mu * eps);
            wmax = std::max(wmax, w[i]);
        }
        if (wmax > 1.0)
            for (size_t i = 0; i < n; ++i) w[i] /= wmax;
    }
    orc_apply_mask_zero(x.data(), mask, n);
    std::memcpy(chi_out, x.data(), n * sizeof(double));
}
}
