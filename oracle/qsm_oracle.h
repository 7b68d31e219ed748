/*
 * CPU f64 oracle for the TV-ADMM / V-SHARP path of QSM.rs (qsm-core).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the HIP path is
 * compared against; it is never linked into, loaded by or called from the
 * product library.  Only the test-suite, the smoke check and the CPU-baseline
 * leg of the benchmark may load it.
 *
 * Every function restates one reference function; the citation beside each
 * declaration names the file:line it follows.  Arrays are Fortran-ordered
 * (idx = i + j*nx + k*nx*ny, src/fft.rs:236-241).  Complex arrays are
 * interleaved (re, im) doubles.
 */
#ifndef QSM_ORACLE_H
#define QSM_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void (*orc_progress_fn)(size_t cur, size_t total, void* user);

/* numpy-compatible fftfreq (src/fft.rs:544-566). */
void orc_fftfreq(size_t n, double d, double* out);

/* In-place 3-D C2C FFT, x then y then z; inverse divides by N
 * (src/fft.rs:97-209, 412-414, 423-487).  OpenMP covers exactly the axis
 * loops and the 1/N scale, as the reference's `parallel` feature does. */
void orc_fft3d(double* data, size_t nx, size_t ny, size_t nz, int inverse);

/* 1-D C2C FFT of one contiguous sequence (used to pin the transform). */
void orc_fft1d(double* data, size_t n, int inverse);

/* src/kernels/dipole.rs:24-72 */
void orc_dipole_kernel(size_t nx, size_t ny, size_t nz, double vsx, double vsy,
                       double vsz, double bx, double by, double bz, double* out);
/* src/kernels/laplacian.rs:20-62 */
void orc_laplacian_kernel(size_t nx, size_t ny, size_t nz, double vsx, double vsy,
                          double vsz, int negative, double* out);
/* src/kernels/smv.rs:19-67; returns the voxel count of the sphere. */
double orc_smv_kernel(size_t nx, size_t ny, size_t nz, double vsx, double vsy,
                      double vsz, double radius, double* out);
/* src/kernels/smv.rs:79-105 */
void orc_erode_mask_smv(const uint8_t* mask, const double* s_fft, size_t nx,
                        size_t ny, size_t nz, double delta, uint8_t* out);
/* src/utils/gradient.rs:21-58 (periodic forward differences) */
void orc_fgrad(double* gx, double* gy, double* gz, const double* x, size_t nx,
               size_t ny, size_t nz, double vsx, double vsy, double vsz);
/* src/utils/gradient.rs:71-108 (periodic backward divergence) */
void orc_bdiv(double* div, const double* gx, const double* gy, const double* gz,
              size_t nx, size_t ny, size_t nz, double vsx, double vsy, double vsz);
/* src/utils/mod.rs:44-52 */
double orc_shrink(double x, double t);
/* src/utils/mod.rs:69-77 */
double orc_relative_change(const double* x, const double* x_prev, size_t n);
/* src/utils/mask.rs:47-53 */
void orc_apply_mask_zero(double* data, const uint8_t* mask, size_t n);

/* src/inversion/admm.rs:150-178: inv_a (N reals) and f_hat (N complex). */
void orc_prepare_admm_spectral(const double* local_field, size_t nx, size_t ny,
                               size_t nz, double vsx, double vsy, double vsz,
                               double bx, double by, double bz, double rho,
                               double* inv_a, double* f_hat);

/* src/inversion/tv.rs:63-97 with admm_step (src/inversion/admm.rs:76-145).
 * iters_out receives the number of iterations run; relchange_out the last
 * relative change evaluated. */
void orc_tv_admm(const double* local_field, const uint8_t* mask, size_t nx,
                 size_t ny, size_t nz, double vsx, double vsy, double vsz,
                 double bx, double by, double bz, double lambda, double rho,
                 double tol, size_t max_iter, orc_progress_fn progress,
                 void* user, double* chi_out, size_t* iters_out,
                 double* relchange_out);

/* Vector-Jacobian product of K unrolled tv_admm iterations (tol = 0; the
 * forward is orc_tv_admm's loop, src/inversion/admm.rs:76-145): given
 * chi_bar = dL/dchi for chi = mask * x_K, writes chi (the forward output),
 * field_bar = dL/d(local_field) and *lambda_bar = dL/dlambda.  Derived by
 * hand from the forward (no reference counterpart); pinned by central
 * differences of orc_tv_admm in the tests. */
void orc_tv_unrolled_vjp(const double* local_field, const uint8_t* mask, size_t nx,
                         size_t ny, size_t nz, double vsx, double vsy, double vsz,
                         double bx, double by, double bz, double lambda, double rho,
                         size_t K, const double* chi_bar, double* chi_out,
                         double* field_bar, double* lambda_bar);

/* src/inversion/tkd.rs:48-99 (tkd) and the tsvd that follows it. */
void orc_tkd(const double* local_field, const uint8_t* mask, size_t nx, size_t ny, size_t nz,
             double vsx, double vsy, double vsz, double bx, double by, double bz,
             double threshold, int tsvd, double* chi_out);
/* src/inversion/tikhonov.rs:62-131; reg 0 Identity, 1 Gradient, 2 Laplacian. */
void orc_tikhonov(const double* local_field, const uint8_t* mask, size_t nx, size_t ny,
                  size_t nz, double vsx, double vsy, double vsz, double bx, double by,
                  double bz, double lambda, int reg, double* chi_out);
/* src/inversion/rts.rs:71-212 (with admm_step, src/inversion/admm.rs:76-145). */
void orc_rts(const double* local_field, const uint8_t* mask, size_t nx, size_t ny, size_t nz,
             double vsx, double vsy, double vsz, double bx, double by, double bz,
             double delta, double mu, double rho, double tol, size_t max_iter,
             size_t lsmr_iter, orc_progress_fn progress, void* user, double* chi_out,
             size_t* iters_out);

/* src/unwrap/laplacian.rs:41-50 and :135-161. */
double orc_wrap(double x);
void orc_laplacian_unwrap(const double* phase, const uint8_t* mask, size_t nx, size_t ny, size_t nz,
                          double vsx, double vsy, double vsz, double* out);

/* src/inversion/nltv.rs:63-133 */
void orc_nltv(const double* local_field, const uint8_t* mask, size_t nx, size_t ny, size_t nz,
              double vsx, double vsy, double vsz, double bx, double by, double bz, double lambda,
              double mu, double tol, size_t max_iter, size_t newton_iter, orc_progress_fn progress,
              void* user, double* chi_out);

/* src/inversion/tikhonov.rs:62-131 with Regularization::Gradient. */
void orc_tikhonov_gradient(const double* local_field, const uint8_t* mask,
                           size_t nx, size_t ny, size_t nz, double vsx,
                           double vsy, double vsz, double bx, double by,
                           double bz, double lambda, double* chi_out);

/* src/bgremove/vsharp.rs:70-82; returns the number of radii (writes at most cap). */
size_t orc_vsharp_radii(double max_radius, double min_radius, double* out, size_t cap);
/* src/bgremove/vsharp.rs:88-199 */
void orc_vsharp_with_radii(const double* field, const uint8_t* mask, size_t nx,
                           size_t ny, size_t nz, double vsx, double vsy,
                           double vsz, const double* radii, size_t n_radii,
                           double threshold, orc_progress_fn progress,
                           void* user, double* local_out, uint8_t* mask_out);
/* src/bgremove/vsharp.rs:55-64 */
void orc_vsharp(const double* field, const uint8_t* mask, size_t nx, size_t ny,
                size_t nz, double vsx, double vsy, double vsz, double threshold,
                double max_radius, double min_radius, orc_progress_fn progress,
                void* user, double* local_out, uint8_t* mask_out);
/* src/bgremove/sharp.rs:64-131 */
void orc_sharp_core(const double* field, const uint8_t* mask, size_t nx,
                    size_t ny, size_t nz, double vsx, double vsy, double vsz,
                    double threshold, double radius, double* local_out,
                    uint8_t* mask_out);

/* Forward model Re IFFT(D * FFT(chi)) (tests/chisep_integration.rs:18-62,
 * without the Hz conversion and mean removal). */
void orc_dipole_forward(const double* chi, size_t nx, size_t ny, size_t nz,
                        double vsx, double vsy, double vsz, double bx,
                        double by, double bz, double* out);

/* tests/common/mod.rs:17-51 */
double orc_rmse(const double* a, const double* b, const uint8_t* mask, size_t n);
double orc_nrmse(const double* a, const double* b, const uint8_t* mask, size_t n);

/* Number of OpenMP threads the FFT passes use, and a setter (mirrors the
 * reference's rayon pool size; src/inversion/tv.rs:181-184 runs a 1-thread
 * pool to check thread-count invariance). */
int orc_num_threads(void);
void orc_set_num_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
