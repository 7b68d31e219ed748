// gfx950 kernels for the TV-ADMM and V-SHARP path.
//
// The 3-D transforms are real-to-complex along x (two real rows packed as one
// complex row), then complex along y and z on the half spectrum, in three
// LDS-staged passes.  Element-wise work is fused into the passes:
//   TV-ADMM iteration (src/inversion/admm.rs:76-145), one launch per line:
//     k_r2c_x<SRC_TVDIV>  z/u update of the previous iteration + bdiv + R2C x
//     k_y_pass<fwd>       C2C along y
//     k_z_tv_solve        C2C z, x^ = f^ - rho*inv_a*FFT(div), inverse z
//     k_y_pass<inv>       inverse C2C along y
//     k_c2r_tv            C2R x, write x, partial sums of relative_change
//     k_tv_reduce         relative_change in FP64, convergence flag
#include "qsm_passes.hpp"

namespace qsm {

// ---------------------------------------------------- spectral multipliers --
// D * inv_a and inv_a at one index (src/kernels/dipole.rs:24-72,
// src/inversion/admm.rs:157-168), with the Laplacian spectrum in closed form
// sum_a (2 - 2 cos(2 pi k_a / n_a)) / vs_a^2 (the DFT of the stencil of
// src/kernels/laplacian.rs:20-62, exactly 0 at DC).
__device__ inline void tv_multipliers(const VolDesc& V, int kx, int ky, int kz, double& dinv,
                                      double& inv) {
    const double fx = fftfreq_at(kx, V.nx, V.vsx);
    const double fy = fftfreq_at(ky, V.ny, V.vsy);
    const double fz = fftfreq_at(kz, V.nz, V.vsz);
    const double kdb = fx * V.bx + fy * V.by + fz * V.bz;
    const double k2 = fx * fx + fy * fy + fz * fz;
    const double D = (k2 > 1e-20) ? (1.0 / 3.0) - (kdb * kdb) / k2 : 0.0;
    const double two_pi = 6.283185307179586476925286766559;
    const double L = (2.0 - 2.0 * cos(two_pi * kx / V.nx)) / (V.vsx * V.vsx) +
                     (2.0 - 2.0 * cos(two_pi * ky / V.ny)) / (V.vsy * V.vsy) +
                     (2.0 - 2.0 * cos(two_pi * kz / V.nz)) / (V.vsz * V.vsz);
    const double a = D * D + V.rho * L;
    inv = (fabs(a) > 1e-20) ? 1.0 / a : 0.0;
    dinv = D * inv;
}

// ------------------------------------------------------------- z passes ---
enum ZOp : int { ZOP_TV_SETUP = 0, ZOP_TV_SOLVE = 1 };

template <int EMAX, int OP>
__global__ __launch_bounds__(kThreads) void k_z_pass(const VolDesc* __restrict__ vols, LaunchTable lt,
                                                     const float2* __restrict__ tw,
                                                     const VolStatus* __restrict__ status) {
    __shared__ float2 lds[kThreads * EMAX];
    const int bid = xcd_remap(blockIdx.x, gridDim.x);
    const int v = find_vol(lt, bid);
    const int lb = bid - lt.start[v];
    const VolDesc& V = vols[v];
    if (status && status[v].done) return;
    const int ky = lb / V.ntx;
    const int tx = lb - ky * V.ntx;
    const int kx0 = tx * kTile;
    const int nxh = V.nxh, nz = V.nz;
    const size_t zs = (size_t)nxh * V.ny;
    const size_t base = (size_t)ky * nxh;
    float2* buf = V.spec[SPEC_WORK] + base;
    for (int e = threadIdx.x; e < nz * kTile; e += kThreads) {
        const int kz = e >> kLog2Tile, c = e & (kTile - 1);
        const int kx = kx0 + c;
        lds[e] = (kx < nxh) ? buf[kx + kz * zs] : make_float2(0.0f, 0.0f);
    }
    __syncthreads();
    qsmfft::lds_fft<EMAX, true, kLog2Tile, false>(lds, V.pz, kTile, kTile, 1, tw + V.pz.tw_off);
    if (OP == ZOP_TV_SETUP) {
        // f^ = FFT(f) * D * inv_a and rho*inv_a, both scaled by 1/N so that the
        // inverse transform needs no separate normalisation.  The multipliers
        // are averaged over k and -k: the reference keeps Re(IFFT(.)) of a
        // full spectrum, which equals the inverse of its Hermitian part; on
        // Nyquist planes D(k) and D(-k) differ when B0 is tilted.
        const double invN = 1.0 / ((double)V.nx * V.ny * V.nz);
        float2* F = V.spec[SPEC_A] + base;
        float* Q = V.q + base;
        for (int e = threadIdx.x; e < nz * kTile; e += kThreads) {
            const int kz = e >> kLog2Tile, c = e & (kTile - 1);
            const int kx = kx0 + c;
            if (kx < nxh) {
                double d1, i1, d2, i2;
                tv_multipliers(V, kx, ky, kz, d1, i1);
                tv_multipliers(V, mirror_idx(kx, V.nx), mirror_idx(ky, V.ny), mirror_idx(kz, nz),
                               d2, i2);
                const double m = 0.5 * (d1 + d2) * invN;
                const double qv = V.rho * 0.5 * (i1 + i2) * invN;
                const float2 s = lds[e];
                F[kx + kz * zs] = make_float2((float)((double)s.x * m), (float)((double)s.y * m));
                Q[kx + kz * zs] = (float)qv;
            }
        }
        return;
    }
    // ZOP_TV_SOLVE: x^ = f^ - rho * FFT(div) * inv_a
    {
        const float2* F = V.spec[SPEC_A] + base;
        const float* Q = V.q + base;
        for (int e = threadIdx.x; e < nz * kTile; e += kThreads) {
            const int kz = e >> kLog2Tile, c = e & (kTile - 1);
            const int kx = kx0 + c;
            if (kx < nxh) {
                const float2 f = F[kx + kz * zs];
                const float q = Q[kx + kz * zs];
                const float2 s = lds[e];
                lds[e] = make_float2(f.x - q * s.x, f.y - q * s.y);
            }
        }
    }
    __syncthreads();
    qsmfft::lds_fft<EMAX, true, kLog2Tile, true>(lds, V.pz, kTile, kTile, 1, tw + V.pz.tw_off);
    for (int e = threadIdx.x; e < nz * kTile; e += kThreads) {
        const int kz = e >> kLog2Tile, c = e & (kTile - 1);
        const int kx = kx0 + c;
        if (kx < nxh) buf[kx + kz * zs] = lds[e];
    }
}

// TV: x_it = Re IFFT(x^) (src/inversion/admm.rs:111-114) and the two sums of
// relative_change (src/utils/mod.rs:69-77) accumulated in FP64.
template <int EMAX>
__global__ __launch_bounds__(kThreads) void k_c2r_tv(const VolDesc* __restrict__ vols, LaunchTable lt,
                                                     const float2* __restrict__ tw, int parity,
                                                     const VolStatus* __restrict__ status) {
    __shared__ float2 lds[kThreads * EMAX];
    __shared__ double sh[kThreads / 64];
    const int bid = xcd_remap(blockIdx.x, gridDim.x);
    const int v = find_vol(lt, bid);
    const int lb = bid - lt.start[v];
    const VolDesc& V = vols[v];
    if (status && status[v].done) return;
    const int nx = V.nx, RB = V.rows_x;
    const int nrows = V.ny * V.nz;
    const int row0 = lb * RB * 2;
    const float2* s = V.spec[SPEC_WORK];
    c2r_load<EMAX>(lds, V, s, s, row0, 2, nrows);
    __syncthreads();
    qsmfft::lds_fft<EMAX, false, 0, true>(lds, V.px, RB, 1, nx, tw + V.px.tw_off);
    const float* xo = V.x[parity];
    float* xn = V.x[parity ^ 1];
    double sd = 0.0, sx = 0.0;
    for (int e = threadIdx.x; e < RB * nx; e += kThreads) {
        const int p = (int)fastdiv((uint32_t)e, V.m_nx);
        const int i = e - p * nx;
        const int ra = row0 + 2 * p, rb = ra + 1;
        const float2 z = lds[e];
        if (ra < nrows) {
            const size_t idx = i + (size_t)ra * nx;
            const double d = (double)z.x - (double)xo[idx];
            xn[idx] = z.x;
            sd += d * d;
            sx += (double)z.x * (double)z.x;
        }
        if (rb < nrows) {
            const size_t idx = i + (size_t)rb * nx;
            const double d = (double)z.y - (double)xo[idx];
            xn[idx] = z.y;
            sd += d * d;
            sx += (double)z.y * (double)z.y;
        }
    }
    const double tsd = block_sum(sd, sh);
    const double tsx = block_sum(sx, sh);
    if (threadIdx.x == 0) {
        V.partial[2 * lb] = tsd;
        V.partial[2 * lb + 1] = tsx;
    }
}

// relative_change = ||x - x_prev|| / (||x|| + 1e-20); converged when < tol
// (src/inversion/admm.rs:117-119).  One block per volume, fixed summation
// order, so repeated runs give identical results.
__global__ __launch_bounds__(kThreads) void k_tv_reduce(const VolDesc* __restrict__ vols, int nvol,
                                                        VolStatus* __restrict__ status,
                                                        VolStatus* __restrict__ mirror) {
    __shared__ double sh[kThreads / 64];
    const int v = blockIdx.x;
    if (v >= nvol) return;
    const VolDesc& V = vols[v];
    if (status[v].done) return;
    double sd = 0.0, sx = 0.0;
    for (int b = threadIdx.x; b < V.npartial; b += kThreads) {
        sd += V.partial[2 * b];
        sx += V.partial[2 * b + 1];
    }
    const double tsd = block_sum(sd, sh);
    const double tsx = block_sum(sx, sh);
    if (threadIdx.x == 0) {
        const double rel = sqrt(tsd) / (sqrt(tsx) + 1e-20);
        VolStatus st = status[v];
        st.iters += 1;
        st.rel = rel;
        if (rel < V.tol) st.done = 1;
        status[v] = st;
        if (mirror) {
            // host-visible copy, read by the host loop after an event wait
            mirror[v] = st;
            __threadfence_system();
        }
    }
}

// chi = x where mask != 0, else 0 (src/utils/mask.rs:47-53), widened to f64.
__global__ __launch_bounds__(kThreads) void k_tv_output(const VolDesc* __restrict__ vols, LaunchTable lt,
                                                        const VolStatus* __restrict__ status) {
    const int bid = blockIdx.x;
    const int v = find_vol(lt, bid);
    const int lb = bid - lt.start[v];
    const VolDesc& V = vols[v];
    const size_t N = (size_t)V.nx * V.ny * V.nz;
    const float* x = V.x[status[v].iters & 1];
    for (size_t i = (size_t)lb * kThreads * 4 + threadIdx.x; i < N && i < (size_t)(lb + 1) * kThreads * 4;
         i += kThreads)
        V.out[i] = V.mask[i] ? (double)x[i] : 0.0;
}

// Zero-fill of float buffers (used instead of memset nodes so the sequence
// stays capturable in a HIP graph).
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

// -------------------------------------------------------- host launchers --
template <int EMAX>
struct Launch {
    static void r2c(int src, hipStream_t st, const VolDesc* d, const LaunchTable& lt, const float2* tw,
                    int out_sel, int parity, const VolStatus* status) {
        dim3 g(lt.start[lt.nvol]), b(kThreads);
        switch (src) {
            case SRC_F64: k_r2c_x<EMAX, SRC_F64><<<g, b, 0, st>>>(d, lt, tw, out_sel, parity, status); break;
            case SRC_U8: k_r2c_x<EMAX, SRC_U8><<<g, b, 0, st>>>(d, lt, tw, out_sel, parity, status); break;
            case SRC_F32: k_r2c_x<EMAX, SRC_F32><<<g, b, 0, st>>>(d, lt, tw, out_sel, parity, status); break;
            case SRC_SPHERE: k_r2c_x<EMAX, SRC_SPHERE><<<g, b, 0, st>>>(d, lt, tw, out_sel, parity, status); break;
            default: k_r2c_x<EMAX, SRC_TVDIV><<<g, b, 0, st>>>(d, lt, tw, out_sel, parity, status); break;
        }
    }
    static void ypass(bool inv, hipStream_t st, const VolDesc* d, const LaunchTable& lt, const float2* tw,
                      int sel, int sel2, const VolStatus* status) {
        dim3 g(lt.start[lt.nvol]), b(kThreads);
        if (inv) k_y_pass<EMAX, true><<<g, b, 0, st>>>(d, lt, tw, sel, sel2, status);
        else k_y_pass<EMAX, false><<<g, b, 0, st>>>(d, lt, tw, sel, sel2, status);
    }
    static void zpass(int op, hipStream_t st, const VolDesc* d, const LaunchTable& lt, const float2* tw,
                      const VolStatus* status) {
        dim3 g(lt.start[lt.nvol]), b(kThreads);
        if (op == ZOP_TV_SETUP) k_z_pass<EMAX, ZOP_TV_SETUP><<<g, b, 0, st>>>(d, lt, tw, status);
        else k_z_pass<EMAX, ZOP_TV_SOLVE><<<g, b, 0, st>>>(d, lt, tw, status);
    }
    static void c2r_tv(hipStream_t st, const VolDesc* d, const LaunchTable& lt, const float2* tw, int parity,
                       const VolStatus* status) {
        k_c2r_tv<EMAX><<<dim3(lt.start[lt.nvol]), dim3(kThreads), 0, st>>>(d, lt, tw, parity, status);
    }
};

}  // namespace qsm

// Entry points used by the host runtime (qsm_runtime.hip).
namespace qsm {

void launch_r2c(int emax, int src, hipStream_t st, const VolDesc* d, const LaunchTable& lt,
                const float2* tw, int out_sel, int parity, const VolStatus* status) {
    if (emax <= 16) Launch<16>::r2c(src, st, d, lt, tw, out_sel, parity, status);
    else if (emax <= 24) Launch<24>::r2c(src, st, d, lt, tw, out_sel, parity, status);
    else Launch<32>::r2c(src, st, d, lt, tw, out_sel, parity, status);
}
void launch_ypass(int emax, bool inv, hipStream_t st, const VolDesc* d, const LaunchTable& lt,
                  const float2* tw, int sel, int sel2, const VolStatus* status) {
    if (emax <= 16) Launch<16>::ypass(inv, st, d, lt, tw, sel, sel2, status);
    else if (emax <= 24) Launch<24>::ypass(inv, st, d, lt, tw, sel, sel2, status);
    else Launch<32>::ypass(inv, st, d, lt, tw, sel, sel2, status);
}
void launch_zpass(int emax, int op, hipStream_t st, const VolDesc* d, const LaunchTable& lt,
                  const float2* tw, const VolStatus* status) {
    if (emax <= 16) Launch<16>::zpass(op, st, d, lt, tw, status);
    else if (emax <= 24) Launch<24>::zpass(op, st, d, lt, tw, status);
    else Launch<32>::zpass(op, st, d, lt, tw, status);
}
void launch_c2r_tv(int emax, hipStream_t st, const VolDesc* d, const LaunchTable& lt, const float2* tw,
                   int parity, const VolStatus* status) {
    if (emax <= 16) Launch<16>::c2r_tv(st, d, lt, tw, parity, status);
    else if (emax <= 24) Launch<24>::c2r_tv(st, d, lt, tw, parity, status);
    else Launch<32>::c2r_tv(st, d, lt, tw, parity, status);
}
void launch_tv_reduce(hipStream_t st, const VolDesc* d, int nvol, VolStatus* status, VolStatus* mirror) {
    k_tv_reduce<<<dim3(nvol), dim3(kThreads), 0, st>>>(d, nvol, status, mirror);
}
void launch_tv_output(hipStream_t st, const VolDesc* d, const LaunchTable& lt, const VolStatus* status) {
    k_tv_output<<<dim3(lt.start[lt.nvol]), dim3(kThreads), 0, st>>>(d, lt, status);
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

}  // namespace qsm
