/*
 * qsm_gpu.h -- C ABI of the MI355X (gfx950) backend for QSM.rs's TV-ADMM
 * dipole inversion and V-SHARP background-field removal.
 *
 * Conventions shared with the reference crate (qsm-core):
 *   - volumes are flat arrays in Fortran order, idx = i + j*nx + k*nx*ny
 *     (src/fft.rs:236-241);
 *   - voxel sizes are in mm, B0 direction is normalised internally;
 *   - masks are u8, 0 = outside; the eroded mask returned by V-SHARP is 0/1.
 * Every entry point is synchronous, returns a status code (0 = QSM_OK) and
 * never aborts the process; qsm_last_error_message() describes the last
 * failure on the calling thread.  Progress callbacks run on the calling
 * thread with exactly the reference's (current, total) sequence.
 *
 * Host-pointer entry points allocate their device memory per call.  The
 * batch entry points (qsm_tv_batch_*) take device pointers, keep their
 * workspace between runs and are the throughput path.
 */
#ifndef QSM_GPU_H
#define QSM_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define QSM_OK 0
#define QSM_EDIM 1          /* length mismatch: PipelineError::DimensionMismatch
                               (src/pipeline/config.rs:344-368) */
#define QSM_EPARAM 2        /* invalid argument (null pointer, bad parameter) */
#define QSM_EHIP 3          /* HIP runtime error */
#define QSM_ENOMEM 4        /* device allocation failed */
#define QSM_ECONFIG 5       /* PipelineError::InvalidConfig: algorithm not served here */
#define QSM_EUNSUPPORTED 6  /* axis length outside what the FFT supports */

/* ---- types ------------------------------------------------------------- */

/* Grid, src/grid.rs:9-14: dims (nx, ny, nz) and voxel size (mm). */
typedef struct qsm_grid {
    size_t nx, ny, nz;
    double vsx, vsy, vsz;
} qsm_grid;

/* TvParams, src/inversion/tv.rs:21-43 (defaults 2e-4, 2e-2, 1e-3, 250). */
typedef struct qsm_tv_params {
    double lambda;
    double rho;
    double tol;
    size_t max_iter;
} qsm_tv_params;

/* VsharpParams, src/bgremove/vsharp.rs:20-39 (defaults 0.05, 12 mm, 1 mm). */
typedef struct qsm_vsharp_params {
    double threshold;
    double max_radius;
    double min_radius;
} qsm_vsharp_params;

/* Progress callback: the reference's `impl FnMut(usize, usize)`. */
typedef void (*qsm_progress_fn)(size_t current, size_t total, void* user);

/* ScanMetadata, src/pipeline/config.rs:300-318 (the fields the TV / V-SHARP
 * arms read: dims, voxel size, B0 direction). */
typedef struct qsm_scan_metadata {
    size_t nx, ny, nz;
    double vsx, vsy, vsz;
    const double* echo_times;
    size_t n_echo_times;
    double field_strength;
    double b0_direction[3];
} qsm_scan_metadata;

/* InversionAlgorithm, src/pipeline/config.rs:47-58 (same order). */
typedef enum qsm_inversion_algorithm {
    QSM_INV_TKD = 0, QSM_INV_TSVD, QSM_INV_TIKHONOV, QSM_INV_TV, QSM_INV_RTS,
    QSM_INV_NLTV, QSM_INV_MEDI, QSM_INV_ILSQR, QSM_INV_TGV, QSM_INV_QSMART
} qsm_inversion_algorithm;

/* BgRemovalAlgorithm, src/pipeline/config.rs:33-43 (same order). */
typedef enum qsm_bg_removal_algorithm {
    QSM_BG_VSHARP = 0, QSM_BG_PDF, QSM_BG_LBV, QSM_BG_ISMV, QSM_BG_SHARP,
    QSM_BG_RESHARP, QSM_BG_HARPERELLA, QSM_BG_IHARPERELLA
} qsm_bg_removal_algorithm;

/* TkdParams, src/inversion/tkd.rs:21-33 (also the tsvd's parameters). */
typedef struct qsm_tkd_params {
    double threshold;  /* default 0.15 */
} qsm_tkd_params;

/* Regularization, src/inversion/tikhonov.rs:24-33 (same order). */
typedef enum qsm_regularization { QSM_REG_IDENTITY = 0, QSM_REG_GRADIENT, QSM_REG_LAPLACIAN } qsm_regularization;

/* TikhonovParams, src/inversion/tikhonov.rs:36-50. */
typedef struct qsm_tikhonov_params {
    double lambda;           /* default 0.01 */
    qsm_regularization reg;  /* default Identity */
} qsm_tikhonov_params;

/* RtsParams, src/inversion/rts.rs:24-52. */
typedef struct qsm_rts_params {
    double delta;      /* 0.15 */
    double mu;         /* 1e5 */
    double rho;        /* 10 */
    double tol;        /* 1e-2 */
    size_t max_iter;   /* 20 */
    size_t lsmr_iter;  /* 4 */
} qsm_rts_params;

/* NltvParams, src/inversion/nltv.rs:24-49. */
typedef struct qsm_nltv_params {
    double lambda;       /* 1e-3 (rho = 100 lambda) */
    double mu;           /* 1 */
    double tol;          /* 1e-3 */
    size_t max_iter;     /* 250 ADMM steps per Newton round */
    size_t newton_iter;  /* 10 */
} qsm_nltv_params;

/* SharpParams, src/bgremove/sharp.rs:27-45. */
typedef struct qsm_sharp_params {
    double threshold;  /* 0.05 */
    double radius;     /* 6 mm */
} qsm_sharp_params;

/* InversionConfig, src/pipeline/config.rs:233-264, for the arms served here. */
typedef struct qsm_inversion_config {
    qsm_inversion_algorithm algorithm;  /* default Rts */
    qsm_tkd_params tkd;
    qsm_tkd_params tsvd;
    qsm_tikhonov_params tikhonov;
    qsm_tv_params tv;
    qsm_rts_params rts;
    qsm_nltv_params nltv;
} qsm_inversion_config;

/* BgRemovalConfig, src/pipeline/config.rs:202-229, for the arms served here. */
typedef struct qsm_bg_removal_config {
    qsm_bg_removal_algorithm algorithm;  /* default Vsharp */
    qsm_vsharp_params vsharp;
    qsm_sharp_params sharp;
} qsm_bg_removal_config;

/* ---- defaults and diagnostics ------------------------------------------ */

void qsm_tkd_params_default(qsm_tkd_params* p);
void qsm_tikhonov_params_default(qsm_tikhonov_params* p);
void qsm_rts_params_default(qsm_rts_params* p);
void qsm_sharp_params_default(qsm_sharp_params* p);
void qsm_nltv_params_default(qsm_nltv_params* p);
void qsm_inversion_config_default(qsm_inversion_config* c);
void qsm_bg_removal_config_default(qsm_bg_removal_config* c);

/* TvParams::default(), src/inversion/tv.rs:34-43 */
void qsm_tv_params_default(qsm_tv_params* p);
/* VsharpParams::default(), src/bgremove/vsharp.rs:31-39 */
void qsm_vsharp_params_default(qsm_vsharp_params* p);
/* Message of the last failed call on this thread ("" if none). */
const char* qsm_last_error_message(void);
/* Device name, architecture and library build info into buf. */
int qsm_backend_info(char* buf, size_t len);
/* 1 if n is an axis length the GPU FFT supports, else 0: n <= 1024 factoring
 * into 2, 3, 5, 7, 11, 13, or any n <= 512 (a prime factor above 13 runs as a
 * chirp-z transform).  The reference (rustfft) transforms any length. */
int qsm_fft_length_supported(size_t n);

/* ---- drop-in algorithm entry points (host pointers) -------------------- */

/* inversion::tv_admm, src/inversion/tv.rs:63-70.
 * local_field: N f64; mask: N u8; bdir: 3 f64; chi_out: N f64 (caller-owned).
 * Progress: (iter+1, max_iter) before each iteration, (k, k) on convergence. */
int qsm_tv_admm(const double* local_field, const uint8_t* mask, const qsm_grid* grid,
                const double bdir[3], const qsm_tv_params* params, qsm_progress_fn progress,
                void* user, double* chi_out);

/* Same as qsm_tv_admm, also reporting the number of iterations run and the
 * last relative change (diagnostics; no reference counterpart). */
int qsm_tv_admm_ex(const double* local_field, const uint8_t* mask, const qsm_grid* grid,
                   const double bdir[3], const qsm_tv_params* params, qsm_progress_fn progress,
                   void* user, double* chi_out, size_t* iters_out, double* relchange_out);

/* bgremove::vsharp, src/bgremove/vsharp.rs:55-61.
 * local_out: N f64; mask_out: N u8 (eroded mask, 0/1). */
int qsm_vsharp(const double* field, const uint8_t* mask, const qsm_grid* grid,
               const qsm_vsharp_params* params, qsm_progress_fn progress, void* user,
               double* local_out, uint8_t* mask_out);

/* vsharp_with_radii, src/bgremove/vsharp.rs:88-95 (explicit radii in mm, any
 * order; empty -> zeros + copy of mask; one radius -> SHARP). */
int qsm_vsharp_with_radii(const double* field, const uint8_t* mask, const qsm_grid* grid,
                          const double* radii, size_t n_radii, double threshold,
                          qsm_progress_fn progress, void* user, double* local_out,
                          uint8_t* mask_out);

/* ---- pipeline entry points --------------------------------------------- */

/* run_dipole_inversion, src/pipeline/inversion.rs:28-35.  Serves the Tkd,
 * Tsvd, Tikhonov, Tv, Rts and Nltv arms (tv_params for Tv, defaults for the
 * others; qsm_run_dipole_inversion_cfg takes every parameter set); Tgv and
 * Qsmart, and arms without a GPU backend here, return QSM_ECONFIG.  Lengths
 * are checked as the reference does (DimensionMismatch -> QSM_EDIM). */
int qsm_run_dipole_inversion(const double* local_field_ppm, size_t field_len,
                             const uint8_t* mask, size_t mask_len,
                             const qsm_scan_metadata* metadata, qsm_inversion_algorithm algorithm,
                             const qsm_tv_params* tv_params, qsm_progress_fn progress,
                             void* user, double* chi_out);

/* ---- closed-form inversions, RTS, SHARP -------------------------------- */
/* tkd, src/inversion/tkd.rs:48-99; tsvd, the function after it. */
int qsm_tkd(const double* local_field, const uint8_t* mask, const qsm_grid* grid, const double bdir[3],
            const qsm_tkd_params* params, double* chi_out);
int qsm_tsvd(const double* local_field, const uint8_t* mask, const qsm_grid* grid, const double bdir[3],
             const qsm_tkd_params* params, double* chi_out);
/* tikhonov, src/inversion/tikhonov.rs:62-131. */
int qsm_tikhonov(const double* local_field, const uint8_t* mask, const qsm_grid* grid,
                 const double bdir[3], const qsm_tikhonov_params* params, double* chi_out);
/* rts, src/inversion/rts.rs:71-212 (progress as tv_admm's). */
int qsm_rts(const double* local_field, const uint8_t* mask, const qsm_grid* grid, const double bdir[3],
            const qsm_rts_params* params, qsm_progress_fn progress, void* user, double* chi_out);
/* nltv, src/inversion/nltv.rs:63-133 (progress (current, total) before every
 * step, total = max_iter * newton_iter; no final call on convergence). */
int qsm_nltv(const double* local_field, const uint8_t* mask, const qsm_grid* grid, const double bdir[3],
             const qsm_nltv_params* params, qsm_progress_fn progress, void* user, double* chi_out);
/* laplacian_unwrap, src/unwrap/laplacian.rs:135-161 (wrapped phase in
 * radians, f64; output masked). */
int qsm_laplacian_unwrap(const double* phase, const uint8_t* mask, const qsm_grid* grid, double* out);
/* sharp, src/bgremove/sharp.rs:48-56. */
int qsm_sharp(const double* field, const uint8_t* mask, const qsm_grid* grid,
              const qsm_sharp_params* params, double* local_out, uint8_t* mask_out);
/* run_dipole_inversion / run_bg_removal with a full config (the arms above). */
int qsm_run_dipole_inversion_cfg(const double* local_field_ppm, size_t field_len, const uint8_t* mask,
                                 size_t mask_len, const qsm_scan_metadata* metadata,
                                 const qsm_inversion_config* config, qsm_progress_fn progress,
                                 void* user, double* chi_out);
int qsm_run_bg_removal_cfg(const double* field_ppm, size_t field_len, const uint8_t* mask,
                           size_t mask_len, const qsm_scan_metadata* metadata,
                           const qsm_bg_removal_config* config, qsm_progress_fn progress,
                           void* user, double* local_out, uint8_t* mask_out);

/* run_bg_removal, src/pipeline/bg_removal.rs:20-26.  Serves the Vsharp arm
 * (vsharp_params) and the Sharp arm (default SharpParams; see
 * qsm_run_bg_removal_cfg); other algorithms return QSM_ECONFIG. */
int qsm_run_bg_removal(const double* field_ppm, size_t field_len, const uint8_t* mask,
                       size_t mask_len, const qsm_scan_metadata* metadata,
                       qsm_bg_removal_algorithm algorithm,
                       const qsm_vsharp_params* vsharp_params, qsm_progress_fn progress,
                       void* user, double* local_out, uint8_t* mask_out);

/* ---- device-resident batched TV-ADMM (no reference counterpart) -------- */

/* One volume of a batch: device pointers, caller-owned. */
typedef struct qsm_tv_volume {
    qsm_grid grid;
    double bdir[3];
    const double* d_field;  /* N f64 */
    const uint8_t* d_mask;  /* N u8 */
    double* d_chi;          /* N f64 output */
} qsm_tv_volume;

typedef struct qsm_tv_batch qsm_tv_batch;

/* Plan a batch of volumes of any (supported) shapes sharing one TvParams.
 * stream: a hipStream_t, or NULL for a stream owned by the batch.  Allocates
 * the device workspace once. */
int qsm_tv_batch_create(const qsm_tv_volume* vols, size_t n_vols, const qsm_tv_params* params,
                        void* stream, qsm_tv_batch** out);
/* The same batch running RTS (src/inversion/rts.rs:71-212) instead of TV:
 * qsm_tv_batch_run / _status / _destroy apply unchanged. */
int qsm_rts_batch_create(const qsm_tv_volume* vols, size_t n_vols, const qsm_rts_params* params,
                         void* stream, qsm_tv_batch** out);
/* Run the whole solve (spectral setup, iterations, masking) for every volume
 * in one launch per kernel.  Progress: (iter+1, max_iter) per iteration of
 * the batch, (k, k) once every volume has converged.  Blocks until done. */
int qsm_tv_batch_run(qsm_tv_batch* batch, qsm_progress_fn progress, void* user);
/* FFT schedule of the y and z passes (no reference counterpart).  Lengths
 * 16*R0 (64, 128, 144 ... 256) have two: a two-stage transform with one LDS
 * exchange (0) and the LDS-staged Stockham stages (1); they round
 * differently in the last bits.  mode 0 (default): a fixed per-length table
 * (staged for y = 144 and 176, two-stage otherwise), so every process
 * returns the same bits; 1: always two-stage; 2: always staged; 3: tuned,
 * the first batch using a length times both on its own data and the faster
 * is kept for the process (results may then differ between processes). */
int qsm_set_fft_schedule(int mode);
/* Schedule in use for the y (axis 1) or z (axis 2) pass of length n:
 * 0 two-stage, 1 staged, -1 not tuned yet (mode 3). */
int qsm_fft_schedule(int axis, int n, int* sched);
/* Iterations after the first run as replays of HIP graphs captured on the
 * batch's stream at the first run (default), or as eager launches
 * (enable = 0), e.g. for a caller that captures the run into a graph of its
 * own.  Results are identical either way.  No reference counterpart. */
int qsm_tv_batch_set_graphs(qsm_tv_batch* batch, int enable);
/* Iterations run and last relative change of volume v after a run. */
int qsm_tv_batch_status(const qsm_tv_batch* batch, size_t v, size_t* iters, double* relchange);
void qsm_tv_batch_destroy(qsm_tv_batch* batch);
/* Mean duration (ms, HIP events on the batch's stream) of each kernel of the
 * iteration over `iters` timed iterations after a fresh spectral setup:
 * ms_out[0] previous iteration's relative-change reduction + z/u update +
 * divergence + R2C x, [1] C2C y, [2] C2C z + x-update + inverse z,
 * [3] inverse C2C y, [4] C2R x + relative-change sums.  n_out >= 5.  Leaves
 * the batch ready for qsm_tv_batch_run. */
int qsm_tv_batch_profile(qsm_tv_batch* batch, size_t iters, double* ms_out, size_t n_out);

/* ---- device memory helpers (no reference counterpart) ------------------ */
/* For callers that hold no HIP runtime of their own (tests, benchmarks,
 * bindings): plain device allocations and blocking copies on device 0's
 * null stream. */
int qsm_device_alloc(size_t bytes, void** out);
int qsm_device_free(void* ptr);
int qsm_memcpy_htod(void* dst, const void* src, size_t bytes);
int qsm_memcpy_dtoh(void* dst, const void* src, size_t bytes);
int qsm_memcpy_dtod(void* dst, const void* src, size_t bytes);
int qsm_device_synchronize(void);
/* Achievable HBM rate reference: a one-element-per-thread device copy of
 * `bytes`, read + write counted, best of `reps` launches (GB/s). */
int qsm_hbm_copy_probe(size_t bytes, int reps, double* gbps);
/* Achievable rate of a pass's access shape: `reads` arrays of n floats read
 * and `writes` arrays written element by element (1..4 each), all bytes
 * counted, best of `reps` launches (GB/s).  No reference counterpart. */
int qsm_hbm_stream_probe(size_t n, int reads, int writes, int reps, double* gbps);

/* ---- differentiable unrolled TV-ADMM (no reference counterpart) ------- */
/* K = params->max_iter iterations of tv_admm (src/inversion/tv.rs:63-97)
 * with tol ignored (treated as 0), over device-resident volumes as in
 * qsm_tv_batch_create; then the vector-Jacobian product of chi = mask * x_K
 * with respect to the local field and lambda.  The forward keeps, per
 * iteration, one byte per gradient component (the soft-threshold pattern);
 * the backward runs one adjoint iteration per forward iteration. */
typedef struct qsm_tv_unrolled qsm_tv_unrolled;
int qsm_tv_unrolled_create(const qsm_tv_volume* vols, size_t n_vols, const qsm_tv_params* params,
                           void* stream, qsm_tv_unrolled** out);
/* Runs the K iterations and writes each volume's d_chi. */
int qsm_tv_unrolled_forward(qsm_tv_unrolled* h);
/* After a forward: given d_chi_bar[v] = dL/dchi (N f64, device), writes
 * d_field_bar[v] = dL/d(local field) (N f64, device) and, if lambda_bar is
 * non-null, lambda_bar[v] = dL/dlambda (host). */
int qsm_tv_unrolled_backward(qsm_tv_unrolled* h, const double* const* d_chi_bar,
                             double* const* d_field_bar, double* lambda_bar);
void qsm_tv_unrolled_destroy(qsm_tv_unrolled* h);

/* Dipole forward operator field = Re IFFT(D FFT(chi)) (the forward model of
 * tests/chisep_integration.rs:18-62 with the kernel of
 * src/kernels/dipole.rs:24-72) and its adjoint, device pointers (N f64),
 * FP32 transforms; stream may be null.  The operator is self-adjoint. */
int qsm_dipole_forward(const double* d_chi, const qsm_grid* grid, const double bdir[3], double* d_field,
                       void* stream);
int qsm_dipole_adjoint(const double* d_field, const qsm_grid* grid, const double bdir[3], double* d_chi,
                       void* stream);
/* qsm_dipole_forward without any device allocation, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  The workspace (one
 * half spectrum, at most 72 MiB: e.g. 256x256x256) is a fixed device array
 * that the library holds on each device until the process ends; larger
 * volumes return QSM_EPARAM.  Every call on a device shares that array, so
 * calls that overlap on one device race; each call blocks until its passes
 * are done.  Same kernels and results as qsm_dipole_forward; stream may be
 * null (the null stream). */
int qsm_dipole_forward_noalloc(const double* d_chi, const qsm_grid* grid, const double bdir[3],
                               double* d_field, void* stream);

/* ---- NIfTI-1 input/output, src/io.rs ---------------------------------- */
/* NiftiData, src/io.rs:13-26 (dims, voxel size, affine, scaling) plus the
 * number of volumes and the on-disk datatype. */
typedef struct qsm_nifti_info {
    size_t nx, ny, nz, nt;       /* nt = 1 for a 3-D file */
    double vsx, vsy, vsz;        /* pixdim[1..3] */
    double affine[16];           /* sform rows if sform_code > 0, else diag(voxel size) */
    double scl_slope, scl_inter; /* slope 0 on disk is reported as 1 */
    int datatype;
} qsm_nifti_info;
/* is_gzip, src/io.rs:28-30 */
int qsm_nifti_is_gzip(const uint8_t* bytes, size_t len);
/* load_nifti / load_nifti_4d, src/io.rs:62-229: plain or gzip bytes; data may
 * be null (header only), else receives nx*ny*nz*nt f64 values scaled by
 * slope/inter, Fortran order, volume after volume. */
int qsm_nifti_load(const uint8_t* bytes, size_t len, qsm_nifti_info* info, double* data, size_t data_len);
/* save_nifti / save_nifti_gz, src/io.rs:262-357: FLOAT32, sform from the
 * affine; *bytes_out is released with qsm_nifti_free. */
int qsm_nifti_save(const double* data, size_t nx, size_t ny, size_t nz, const double voxel_size[3],
                   const double affine[16], int gzip, uint8_t** bytes_out, size_t* len_out);
void qsm_nifti_free(uint8_t* bytes);
/* read_nifti_file / save_nifti_to_file, src/io.rs:423-452 (.nii.gz by suffix). */
int qsm_nifti_read_file(const char* path, qsm_nifti_info* info, double* data, size_t data_len);
int qsm_nifti_write_file(const char* path, const double* data, size_t nx, size_t ny, size_t nz,
                         const double voxel_size[3], const double affine[16]);

/* ---- device-resident batched V-SHARP (no reference counterpart) ------- */

typedef struct qsm_vsharp_volume {
    qsm_grid grid;
    const double* d_field;  /* N f64 total field */
    const uint8_t* d_mask;  /* N u8 */
    double* d_local;        /* N f64 output local field */
    uint8_t* d_mask_out;    /* N u8 output eroded mask */
} qsm_vsharp_volume;

typedef struct qsm_vsharp_batch qsm_vsharp_batch;

/* Plan V-SHARP over a batch of volumes sharing one radius list (mm, any
 * order, at least one; one radius gives SHARP, src/bgremove/sharp.rs:64-131)
 * and threshold.  Allocates the device workspace once. */
int qsm_vsharp_batch_create(const qsm_vsharp_volume* vols, size_t n_vols, const double* radii,
                            size_t n_radii, double threshold, void* stream,
                            qsm_vsharp_batch** out);
/* Run every volume; progress (r+1, R) per radius as vsharp does. */
int qsm_vsharp_batch_run(qsm_vsharp_batch* batch, qsm_progress_fn progress, void* user);
void qsm_vsharp_batch_destroy(qsm_vsharp_batch* batch);

#ifdef __cplusplus
}
#endif
#endif /* QSM_GPU_H */
