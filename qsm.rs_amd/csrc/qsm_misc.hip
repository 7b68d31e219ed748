// Small kernels outside the FFT passes: the closing relative-change
// reduction, masked f64 output, fills, and the pass dispatcher over
// specialised lengths.
#include "qsm_kernels.hpp"

namespace qsm {

// Relative change of the last iteration of a run (the earlier ones are
// reduced by the next iteration's first pass).  One block per volume.
__global__ __launch_bounds__(kThreads) void k_tv_reduce(const VolDesc* __restrict__ vols, int nvol,
                                                        VolStatus* status, VolStatus* mirror) {
    __shared__ double sh[kThreads / 64];
    const int v = blockIdx.x;
    if (v >= nvol || status[v].done) return;
    tv_reduce_block(vols[v], v, status, mirror, sh);
}

// dL/dlambda = t-bar / rho (t = lambda / rho), t-bar summed over the
// (iteration, block) partials of the adjoint R2C passes in a fixed order.
__global__ __launch_bounds__(kThreads) void k_tv_lbar(const VolDesc* __restrict__ vols, int nvol, int nrows) {
    __shared__ double sh[kThreads / 64];
    const int v = blockIdx.x;
    if (v >= nvol) return;
    const VolDesc& V = vols[v];
    const size_t n = (size_t)(nrows > 0 ? nrows : 0) * V.adj_nblk;
    double a = 0.0;
    for (size_t q = threadIdx.x; q < n; q += kThreads) a += V.adj_tpart[q];
    const double t = block_sum(a, sh);
    if (threadIdx.x == 0) V.adj_lbar[0] = t / V.rho;
}

// NLTV weights after a Newton round (src/inversion/nltv.rs:112-125):
// w = 1/(|grad x| + mu 1e-6) from x[parity] (periodic forward differences,
// src/utils/gradient.rs:21-58), in FP64, stored FP32 in nltv_w[sel]; per-block
// maxima into nltv_wpart.  Blocks of kThreads*4 voxels.
__global__ __launch_bounds__(kThreads) void k_nltv_weights(const VolDesc* __restrict__ vols, LaunchTable lt,
                                                           int parity, int sel) {
    __shared__ double sh[kThreads / 64];
    int lb;
    const int gv = locate(lt, blockIdx.x, lb);
    const VolDesc& V = vols[gv];
    const int nx = V.nx, ny = V.ny, nz = V.nz;
    const size_t nxy = (size_t)nx * ny, N = nxy * nz;
    const float* x = V.x[parity];
    float* w = V.nltv_w[sel];
    const size_t end = (size_t)(lb + 1) * kThreads * 4;
    double mx = 0.0;
    for (size_t q = (size_t)lb * kThreads * 4 + threadIdx.x; q < N && q < end; q += kThreads) {
        const int i = (int)(q % nx);
        const int j = (int)((q / nx) % ny);
        const int k = (int)(q / nxy);
        const long ox = (i + 1 < nx) ? 1L : -(long)(nx - 1);
        const long oy = (j + 1 < ny) ? (long)nx : -(long)(ny - 1) * nx;
        const long oz = (k + 1 < nz) ? (long)nxy : -(long)(nz - 1) * (long)nxy;
        const long lq = (long)q;
        const double xc = x[q];
        const double gx = ((double)x[lq + ox] - xc) / V.vsx;
        const double gy = ((double)x[lq + oy] - xc) / V.vsy;
        const double gz = ((double)x[lq + oz] - xc) / V.vsz;
        const double wv = 1.0 / (sqrt(gx * gx + gy * gy + gz * gz) + V.filt_b * 1e-6);
        w[q] = (float)wv;
        mx = fmax(mx, wv);
    }
    // block maximum
    for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, 64));
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) sh[wave] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = 0.0;
        for (int q = 0; q < kThreads / 64; ++q) m = fmax(m, sh[q]);
        V.nltv_wpart[lb] = m;
    }
}

// nltv_wnorm[sel] = max(w) if it exceeds 1, else 1 (src/inversion/nltv.rs:
// 120-125; the weights are used as w / nltv_wnorm).  One block per volume.
__global__ __launch_bounds__(kThreads) void k_nltv_wmax(const VolDesc* __restrict__ vols, int nvol, int nblk,
                                                        int sel) {
    __shared__ double sh[kThreads / 64];
    const int v = blockIdx.x;
    if (v >= nvol) return;
    const VolDesc& V = vols[v];
    const size_t N = (size_t)V.nx * V.ny * V.nz;
    const int nb = (int)((N + kThreads * 4 - 1) / (kThreads * 4));
    double mx = 0.0;
    for (int q = threadIdx.x; q < nb && q < nblk; q += kThreads) mx = fmax(mx, V.nltv_wpart[q]);
    for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, 64));
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) sh[wave] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = 0.0;
        for (int q = 0; q < kThreads / 64; ++q) m = fmax(m, sh[q]);
        V.nltv_wnorm[sel] = m > 1.0 ? m : 1.0;
    }
}

// Weights of the first round: w = 1, normalisation 1.
__global__ __launch_bounds__(kThreads) void k_nltv_init(const VolDesc* __restrict__ vols, LaunchTable lt) {
    int lb;
    const int gv = locate(lt, blockIdx.x, lb);
    const VolDesc& V = vols[gv];
    const size_t N = (size_t)V.nx * V.ny * V.nz;
    const size_t end = (size_t)(lb + 1) * kThreads * 4;
    for (size_t q = (size_t)lb * kThreads * 4 + threadIdx.x; q < N && q < end; q += kThreads) V.nltv_w[0][q] = 1.0f;
    if (lb == 0 && threadIdx.x < 2) V.nltv_wnorm[threadIdx.x] = 1.0;
}

__global__ void k_status_clear_done(VolStatus* status, int nvol) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < nvol) status[v].done = 0;
}

// One float2 per thread, read and written once: the reference point for what
// the FFT passes' HBM rates can reach on this device.
__global__ __launch_bounds__(kThreads) void k_copy_probe(const float2* __restrict__ a, float2* __restrict__ b,
                                                         size_t n) {
    const size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x;
    if (i < n) b[i] = a[i];
}

// R read and W write streams of n floats each, 16 floats per thread per
// stream, every read issued before the first write: the access shape of a
// pass that reads R and writes W arrays element by element.
template <int R, int W>
__global__ __launch_bounds__(kThreads) void k_stream_probe(const float* __restrict__ in, float* __restrict__ out,
                                                           size_t n) {
    const size_t i0 = (size_t)blockIdx.x * kThreads * 16 + threadIdx.x;
    float a[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const size_t q = min(i0 + (size_t)k * kThreads, n - 1);
        float s = 0.0f;
#pragma unroll
        for (int r = 0; r < R; ++r) s += in[r * n + q];
        a[k] = s;
    }
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const size_t q = i0 + (size_t)k * kThreads;
        if (q < n) {
#pragma unroll
            for (int w = 0; w < W; ++w) out[w * n + q] = a[k] + (float)w;
        }
    }
}

// chi = x where mask != 0, else 0 (src/utils/mask.rs:47-53), widened to f64.
__global__ __launch_bounds__(kThreads) void k_tv_output(const VolDesc* __restrict__ vols, LaunchTable lt,
                                                        const VolStatus* __restrict__ status) {
    int lb;
    const int gv = locate(lt, blockIdx.x, lb);
    const VolDesc& V = vols[gv];
    const size_t N = (size_t)V.nx * V.ny * V.nz;
    const float* x = V.x[status[gv].iters & 1];
    const size_t end = (size_t)(lb + 1) * kThreads * 4;
    for (size_t i = (size_t)lb * kThreads * 4 + threadIdx.x; i < N && i < end; i += kThreads)
        V.out[i] = V.mask[i] ? (double)x[i] : 0.0;
}

__global__ __launch_bounds__(kThreads) void k_vs_init(const VolDesc* __restrict__ vols, LaunchTable lt) {
    int lb;
    const int gv = locate(lt, blockIdx.x, lb);
    const VolDesc& V = vols[gv];
    const size_t N = (size_t)V.nx * V.ny * V.nz;
    const size_t end = (size_t)(lb + 1) * kThreads * 4;
    for (size_t i = (size_t)lb * kThreads * 4 + threadIdx.x; i < N && i < end; i += kThreads) {
        V.vs_local[i] = 0.0f;
        V.vs_final[i] = 0;
    }
}

// Zero fill (a kernel rather than a memset, so the sequence stays capturable).
__global__ __launch_bounds__(kThreads) void k_fill_zero(float* __restrict__ p, size_t n) {
    size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * kThreads;
    for (; i < n; i += stride) p[i] = 0.0f;
}

__global__ void k_status_reset(VolStatus* status, int nvol) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < nvol) {
        VolStatus s;
        s.done = 0;
        s.iters = 0;
        s.rel = 0.0;
        status[v] = s;
    }
}

void launch_tv_reduce(hipStream_t st, const VolDesc* d, int nvol, VolStatus* status, VolStatus* mirror) {
    k_tv_reduce<<<dim3(nvol), dim3(kThreads), 0, st>>>(d, nvol, status, mirror);
}
void launch_tv_lbar(hipStream_t st, const VolDesc* d, int nvol, int nrows) {
    k_tv_lbar<<<dim3(nvol), dim3(kThreads), 0, st>>>(d, nvol, nrows);
}
void launch_nltv_weights(hipStream_t st, const VolDesc* d, const LaunchTable& lt, int parity, int sel) {
    k_nltv_weights<<<dim3(lt.start[lt.nvol]), dim3(kThreads), 0, st>>>(d, lt, parity, sel);
}
void launch_nltv_wmax(hipStream_t st, const VolDesc* d, int nvol, int nblk, int sel) {
    k_nltv_wmax<<<dim3(nvol), dim3(kThreads), 0, st>>>(d, nvol, nblk, sel);
}
void launch_nltv_init(hipStream_t st, const VolDesc* d, const LaunchTable& lt) {
    k_nltv_init<<<dim3(lt.start[lt.nvol]), dim3(kThreads), 0, st>>>(d, lt);
}
void launch_status_clear_done(hipStream_t st, VolStatus* s, int nvol) {
    k_status_clear_done<<<dim3((nvol + 63) / 64), dim3(64), 0, st>>>(s, nvol);
}
void launch_copy_probe(hipStream_t st, const float2* a, float2* b, size_t n) {
    k_copy_probe<<<dim3((unsigned)((n + kThreads - 1) / kThreads)), dim3(kThreads), 0, st>>>(a, b, n);
}
int launch_stream_probe(hipStream_t st, int reads, int writes, const float* in, float* out, size_t n) {
    const dim3 g((unsigned)((n + kThreads * 16 - 1) / (kThreads * 16))), b(kThreads);
#define QSM_PROBE(r, w) \
    if (reads == r && writes == w) { k_stream_probe<r, w><<<g, b, 0, st>>>(in, out, n); return 0; }
    QSM_PROBE(1, 1) QSM_PROBE(1, 2) QSM_PROBE(1, 3) QSM_PROBE(1, 4)
    QSM_PROBE(2, 1) QSM_PROBE(2, 2) QSM_PROBE(2, 3) QSM_PROBE(2, 4)
    QSM_PROBE(3, 1) QSM_PROBE(3, 2) QSM_PROBE(3, 3) QSM_PROBE(3, 4)
    QSM_PROBE(4, 1) QSM_PROBE(4, 2) QSM_PROBE(4, 3) QSM_PROBE(4, 4)
#undef QSM_PROBE
    return -1;
}
void launch_tv_output(hipStream_t st, const VolDesc* d, const LaunchTable& lt, const VolStatus* status) {
    k_tv_output<<<dim3(lt.start[lt.nvol]), dim3(kThreads), 0, st>>>(d, lt, status);
}
void launch_vs_init(hipStream_t st, const VolDesc* d, const LaunchTable& lt) {
    k_vs_init<<<dim3(lt.start[lt.nvol]), dim3(kThreads), 0, st>>>(d, lt);
}
void launch_fill_zero(hipStream_t st, float* p, size_t n) {
    if (n == 0) return;
    size_t blocks = (n + kThreads - 1) / kThreads;
    if (blocks > 4096) blocks = 4096;
    k_fill_zero<<<dim3((unsigned)blocks), dim3(kThreads), 0, st>>>(p, n);
}
void launch_status_reset(hipStream_t st, VolStatus* s, int nvol) {
    k_status_reset<<<dim3((nvol + 63) / 64), dim3(64), 0, st>>>(s, nvol);
}

// ------------------------------------------------------------ dispatcher --
#define QSM_FIXED_LENGTHS(X) X(48) X(64) X(72) X(80) X(96) X(112) X(120) X(128) X(144) X(150) X(160) X(168) \
    X(176) X(180) X(192) X(200) X(208) X(216) X(220) X(224) X(240) X(256) X(288) X(320) X(384) X(416) X(448) \
    X(480) X(512)
#define QSM_DECLARE(n) void launch_fixed_##n(int kind, const LaunchTable& lt, const PassArgs& a);
QSM_FIXED_LENGTHS(QSM_DECLARE)
void launch_generic(int kind, int emax, const LaunchTable& lt, const PassArgs& a);

bool fixed_length(int n) {
#define QSM_IS(m) if (n == m) return true;
    QSM_FIXED_LENGTHS(QSM_IS)
    return false;
}

bool has_fixed_kernels(int n) {
    switch (n) {
#define QSM_FIXED_CASE(m) case m: return true;
        QSM_FIXED_LENGTHS(QSM_FIXED_CASE)
#undef QSM_FIXED_CASE
        default: return false;
    }
}

void launch_pass(int kind, int n, int emax, const LaunchTable& lt, const PassArgs& a) {
    if (lt.nvol == 0 || lt.start[lt.nvol] == 0) return;
    switch (n) {
#define QSM_CASE(m) case m: launch_fixed_##m(kind, lt, a); return;
        QSM_FIXED_LENGTHS(QSM_CASE)
        default: launch_generic(kind, emax, lt, a); return;
    }
}

}  // namespace qsm
