// Pass kernels of the TV-ADMM and V-SHARP paths, as templates on the length
// N of the transformed axis (N = 0: runtime plan with EG elements per thread).
//
// The 3-D transforms are real-to-complex along x (two real rows packed as one
// complex row), then complex along y and z on the half spectrum, in three
// LDS-staged passes; element-wise work is fused into the passes:
//   TV-ADMM iteration (src/inversion/admm.rs:76-145):
//     k_r2c_x<TVDIV>   z/u update of the previous iteration + bdiv + R2C x
//     k_y_pass<fwd>    C2C along y
//     k_z_tv<SOLVE>    C2C z, x^ = f^ - rho*inv_a*FFT(div), inverse z
//     k_y_pass<inv>    inverse C2C along y
//     k_c2r_tv         C2R x, write x, partial sums of relative_change
//   V-SHARP radius (src/bgremove/vsharp.rs:132-177):
//     k_r2c_x<SPHERE>, k_y_pass<fwd>, k_z_vs<RADIUS>, k_y_pass<inv> (x2),
//     k_c2r_vs_select
// Global phases issue all of a thread's loads before the first wait (loops
// over a compile-time count into registers); twiddles are staged in LDS.
#pragma once

#include "fft_fixed.hpp"
#include "qsm_internal.hpp"

namespace qsm {

using qsmfft::kThreads;
using qsmfft::fastdiv;

constexpr int kMaxLen = 512;

// elements per thread of an x (row) pass / a column pass
__host__ __device__ constexpr int x_emax(int n) {
    return n > 0 ? (qsmfft::fixed_rows(n) * n + kThreads - 1) / kThreads : 16;
}
__host__ __device__ constexpr int col_emax(int n, int eg) {
    return n > 0 ? (kTile * n + kThreads - 1) / kThreads : eg;
}
__host__ __device__ constexpr int x_lds(int n) { return n > 0 ? qsmfft::fixed_rows(n) * n : kThreads * 16; }
__host__ __device__ constexpr int col_lds(int n, int eg) { return n > 0 ? kTile * n : kThreads * eg; }
__host__ __device__ constexpr int tw_lds(int n) { return n > 0 ? n : kMaxLen; }

template <int N, int EMAX, bool COLFAST, int LOG2C, bool INV>
__device__ __forceinline__ void run_fft(float2* lds, const qsmfft::Plan& P, int ncols, int es, int cs,
                                        const float2* twl) {
    if constexpr (N > 0)
        qsmfft::fixed_fft<N, COLFAST ? kTile : qsmfft::fixed_rows(N), COLFAST, INV>(lds, twl);
    else
        qsmfft::lds_fft<EMAX, COLFAST, LOG2C, INV>(lds, P, ncols, es, cs, twl);
}

__device__ __forceinline__ void load_twiddles(float2* twl, const float2* __restrict__ tw, int n) {
    for (int q = threadIdx.x; q < n; q += kThreads) twl[q] = tw[q];
}

__device__ __forceinline__ float shrinkf(float x, float t) {
    // src/utils/mod.rs:44-52
    return x > t ? x - t : (x < -t ? x + t : 0.0f);
}

__device__ inline double block_sum(double v, double* sh) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[wave] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < kThreads / 64; ++w) s += sh[w];
    return s;
}

__device__ __forceinline__ double fftfreq_at(int i, int n, double d) {
    // src/fft.rs:544-566: index i of numpy fftfreq(n, d)
    const double val = 1.0 / ((double)n * d);
    const int s = (i < (n + 1) / 2) ? i : i - n;
    return (double)s * val;
}

__device__ __forceinline__ int mirror_idx(int i, int n) { return i == 0 ? 0 : n - i; }

// Binary sphere of src/kernels/smv.rs:19-67 (un-normalised: 1 inside).  The
// squared distance is formed with explicit round-to-nearest operations, in the
// reference's order and without fused multiply-adds, so boundary voxels are
// classified exactly as the reference (and the host voxel count) do.
__device__ __forceinline__ float sphere_element(const VolDesc& V, int i, int r, int ridx) {
    const int k = (int)fastdiv((uint32_t)r, V.m_ny);
    const int j = r - k * V.ny;
    const double xi = (i <= V.nx / 2) ? (double)i : (double)(i - V.nx);
    const double yj = (j <= V.ny / 2) ? (double)j : (double)(j - V.ny);
    const double zk = (k <= V.nz / 2) ? (double)k : (double)(k - V.nz);
    const double dx = __dmul_rn(xi, V.vsx), dy = __dmul_rn(yj, V.vsy), dz = __dmul_rn(zk, V.vsz);
    const double d2 = __dadd_rn(__dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)), __dmul_rn(dz, dz));
    const double rad = V.vs_radii[ridx];
    return (d2 <= __dmul_rn(rad, rad)) ? 1.0f : 0.0f;
}

// ---------------------------------------------------------- tile helpers --
// Column tile: element (row, c) at lds[row*kTile + c] <-> base[kx0 + c + row*stride];
// columns kx >= nxh read as zero and are not written.
template <int EMAX>
__device__ __forceinline__ void col_load(float2* lds, const float2* __restrict__ base, size_t stride,
                                         int n, int kx0, int nxh) {
    float2 r[EMAX];
    const int total = n * kTile;
#pragma unroll
    for (int t = 0; t < EMAX; ++t) {
        const int e = threadIdx.x + t * kThreads;
        const int row = e >> kLog2Tile;
        const int kx = kx0 + (e & (kTile - 1));
        r[t] = (e < total && kx < nxh) ? base[kx + row * stride] : make_float2(0.0f, 0.0f);
    }
#pragma unroll
    for (int t = 0; t < EMAX; ++t) {
        const int e = threadIdx.x + t * kThreads;
        if (e < total) lds[e] = r[t];
    }
}

template <int EMAX>
__device__ __forceinline__ void col_store(const float2* lds, float2* __restrict__ base, size_t stride,
                                          int n, int kx0, int nxh) {
    const int total = n * kTile;
#pragma unroll
    for (int t = 0; t < EMAX; ++t) {
        const int e = threadIdx.x + t * kThreads;
        const int row = e >> kLog2Tile;
        const int kx = kx0 + (e & (kTile - 1));
        if (e < total && kx < nxh) base[kx + row * stride] = lds[e];
    }
}

// ------------------------------------------------ x-pass source fills ----
enum SrcKind : int { SRC_F64 = 0, SRC_U8 = 1, SRC_F32 = 2, SRC_SPHERE = 3, SRC_TVDIV = 4 };

template <int SRC>
__device__ __forceinline__ float simple_src(const VolDesc& V, int i, int r, int aux) {
    if constexpr (SRC == SRC_F64) return (float)V.field[i + (size_t)r * V.nx];
    else if constexpr (SRC == SRC_U8) return (float)V.mask[i + (size_t)r * V.nx];
    else if constexpr (SRC == SRC_F32) return V.vs_local[i + (size_t)r * V.nx];
    else return sphere_element(V, i, r, aux);
}

// lds[p*nx + i] = (row 2p, row 2p+1) of the block's rows.
template <int EMAX, int SRC>
__device__ __forceinline__ void fill_simple(float2* lds, const VolDesc& V, int nx, int RB, int row0,
                                            int nrows, int aux) {
    const int total = RB * nx;
    float2 r[EMAX];
#pragma unroll
    for (int t = 0; t < EMAX; ++t) {
        const int e = threadIdx.x + t * kThreads;
        const int p = e / nx;
        const int i = e - p * nx;
        const int ra = row0 + 2 * p, rb = ra + 1;
        r[t].x = (e < total && ra < nrows) ? simple_src<SRC>(V, i, ra, aux) : 0.0f;
        r[t].y = (e < total && rb < nrows) ? simple_src<SRC>(V, i, rb, aux) : 0.0f;
    }
#pragma unroll
    for (int t = 0; t < EMAX; ++t) {
        const int e = threadIdx.x + t * kThreads;
        if (e < total) lds[e] = r[t];
    }
}

// Divergence of g = 2z - v for the z/u update fused in front of the
// x-subproblem: v = grad(x) + u (periodic forward differences,
// src/utils/gradient.rs:21-58), z = shrink(v, lambda/rho), u' = v - z,
// g = 2z - v (src/inversion/admm.rs:127-142), div = bdiv(g)
// (src/utils/gradient.rs:71-108).  The x component of g at i-1 is recomputed
// from x and u at i-1 (likewise y and z), so g never goes to memory.
// CH elements (2*CH voxels) per round, all their loads issued before the
// arithmetic; restrict-qualified so the u' stores of one round do not order
// the next round's loads.
template <int EMAX>
__device__ __forceinline__ void fill_tvdiv(float2* lds, const VolDesc& V, int nx, int RB,
                                           const float* __restrict__ x, const float* __restrict__ u,
                                           float* __restrict__ uo, int row0, int nrows) {
    constexpr int CH = 2;
    const int ny = V.ny, nz = V.nz;
    const int nxy = nx * ny, N = nxy * nz;
    const int total = RB * nx;
    const float t = V.thr, hx = V.hx, hy = V.hy, hz = V.hz;
#pragma unroll
    for (int t0 = 0; t0 < EMAX; t0 += CH) {
        float xv[CH][2][7], uv[CH][2][6];
        int idxv[CH][2];
        bool ok[CH][2];
#pragma unroll
        for (int c = 0; c < CH; ++c) {
            const int e = threadIdx.x + (t0 + c) * kThreads;
            const int p = e / nx;
            const int i = e - p * nx;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int r = row0 + 2 * p + h;
                ok[c][h] = (t0 + c < EMAX) && (e < total) && (r < nrows);
                const int rr = ok[c][h] ? r : 0;
                const int ii = ok[c][h] ? i : 0;
                const int k = (int)fastdiv((uint32_t)rr, V.m_ny);
                const int j = rr - k * ny;
                const int idx = ii + rr * nx;
                idxv[c][h] = idx;
                const int dxp = (ii + 1 < nx) ? 1 : 1 - nx;
                const int dxm = (ii > 0) ? -1 : nx - 1;
                const int dyp = (j + 1 < ny) ? nx : nx - nxy;
                const int dym = (j > 0) ? -nx : nxy - nx;
                const int dzp = (k + 1 < nz) ? nxy : nxy - N;
                const int dzm = (k > 0) ? -nxy : N - nxy;
                xv[c][h][0] = x[idx];
                xv[c][h][1] = x[idx + dxp];
                xv[c][h][2] = x[idx + dyp];
                xv[c][h][3] = x[idx + dzp];
                xv[c][h][4] = x[idx + dxm];
                xv[c][h][5] = x[idx + dym];
                xv[c][h][6] = x[idx + dzm];
                uv[c][h][0] = u[idx];
                uv[c][h][1] = u[N + idx];
                uv[c][h][2] = u[2 * N + idx];
                uv[c][h][3] = u[idx + dxm];
                uv[c][h][4] = u[N + idx + dym];
                uv[c][h][5] = u[2 * N + idx + dzm];
            }
        }
#pragma unroll
        for (int c = 0; c < CH; ++c) {
            const int e = threadIdx.x + (t0 + c) * kThreads;
            float d[2];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const float xc = xv[c][h][0];
                const float vx = (xv[c][h][1] - xc) * hx + uv[c][h][0];
                const float vy = (xv[c][h][2] - xc) * hy + uv[c][h][1];
                const float vz = (xv[c][h][3] - xc) * hz + uv[c][h][2];
                const float vxm = (xc - xv[c][h][4]) * hx + uv[c][h][3];
                const float vym = (xc - xv[c][h][5]) * hy + uv[c][h][4];
                const float vzm = (xc - xv[c][h][6]) * hz + uv[c][h][5];
                const float zx = shrinkf(vx, t), zy = shrinkf(vy, t), zz = shrinkf(vz, t);
                if (ok[c][h]) {
                    const int idx = idxv[c][h];
                    uo[idx] = vx - zx;
                    uo[N + idx] = vy - zy;
                    uo[2 * N + idx] = vz - zz;
                }
                const float gx = 2.0f * zx - vx, gy = 2.0f * zy - vy, gz = 2.0f * zz - vz;
                const float gxm = 2.0f * shrinkf(vxm, t) - vxm;
                const float gym = 2.0f * shrinkf(vym, t) - vym;
                const float gzm = 2.0f * shrinkf(vzm, t) - vzm;
                d[h] = ok[c][h] ? (gx - gxm) * hx + (gy - gym) * hy + (gz - gzm) * hz : 0.0f;
            }
            if (t0 + c < EMAX && e < total) lds[e] = make_float2(d[0], d[1]);
        }
    }
}

// ------------------------------------------------------------- R2C x pass --
// Block = RB pairs of real rows.  Pair p packs rows (2p, 2p+1) as
// z = a + i b; after a complex FFT of length nx the half spectra are
//   A[k] = (Z[k] + conj Z[-k]) / 2,   B[k] = (Z[k] - conj Z[-k]) / (2i).
// aux: TV iteration parity for SRC_TVDIV, radius index for SRC_SPHERE.
template <int N, int EG, int SRC>
__global__ __launch_bounds__(kThreads) void k_r2c_x(const VolDesc* __restrict__ vols, LaunchTable lt,
                                                    const float2* __restrict__ tw, int out_sel, int aux,
                                                    const VolStatus* __restrict__ status) {
    constexpr int EMAX = x_emax(N);
    __shared__ float2 lds[x_lds(N)];
    __shared__ float2 twl[tw_lds(N)];
    const int bid = xcd_remap(blockIdx.x, gridDim.x);
    int lb;
    const int gv = locate(lt, bid, lb);
    const VolDesc& V = vols[gv];
    if (status && status[gv].done) return;
    const int nx = N > 0 ? N : V.nx;
    const int nxh = nx / 2 + 1;
    const int RB = N > 0 ? qsmfft::fixed_rows(N) : V.rows_x;
    const int nrows = V.ny * V.nz;
    const int row0 = lb * RB * 2;
    load_twiddles(twl, tw + V.px.tw_off, nx);
    if constexpr (SRC == SRC_TVDIV)
        fill_tvdiv<EMAX>(lds, V, nx, RB, V.x[aux], V.u[aux ^ 1], V.u[aux], row0, nrows);
    else
        fill_simple<EMAX, SRC>(lds, V, nx, RB, row0, nrows, aux);
    __syncthreads();
    run_fft<N, EMAX, false, 0, false>(lds, V.px, RB, 1, nx, twl);
    float2* __restrict__ out = V.spec[out_sel];
    const int total = RB * nxh;
#pragma unroll
    for (int t = 0; t < EMAX; ++t) {
        const int e = threadIdx.x + t * kThreads;
        if (e < total) {
            const int p = e / nxh;
            const int k = e - p * nxh;
            const int ra = row0 + 2 * p, rb = ra + 1;
            const float2 zk = lds[p * nx + k];
            const float2 zm = lds[p * nx + (k == 0 ? 0 : nx - k)];
            const float2 A = make_float2(0.5f * (zk.x + zm.x), 0.5f * (zk.y - zm.y));
            const float2 B = make_float2(0.5f * (zk.y + zm.y), -0.5f * (zk.x - zm.x));
            if (ra < nrows) out[k + (size_t)ra * nxh] = A;
            if (rb < nrows) out[k + (size_t)rb * nxh] = B;
        }
    }
}

// ------------------------------------------------------------- C2C y pass --
// Block = one tile of kTile consecutive kx columns at one kz, full ny.
// sel2 >= 0 transforms a second spectrum with a second set of blocks.
template <int N, int EG, bool INV>
__global__ __launch_bounds__(kThreads) void k_y_pass(const VolDesc* __restrict__ vols, LaunchTable lt,
                                                     const float2* __restrict__ tw, int sel, int sel2,
                                                     const VolStatus* __restrict__ status) {
    constexpr int EMAX = col_emax(N, EG);
    __shared__ float2 lds[col_lds(N, EG)];
    __shared__ float2 twl[tw_lds(N)];
    const int bid = xcd_remap(blockIdx.x, gridDim.x);
    int lb;
    const int gv = locate(lt, bid, lb);
    const VolDesc& V = vols[gv];
    if (status && status[gv].done) return;
    const int per = V.ntx * V.nz;
    int s = sel;
    if (lb >= per) { lb -= per; s = sel2; }
    const int kz = lb / V.ntx;
    const int tx = lb - kz * V.ntx;
    const int kx0 = tx * kTile;
    const int nxh = V.nxh;
    const int ny = N > 0 ? N : V.ny;
    float2* buf = V.spec[s] + (size_t)kz * nxh * ny;
    load_twiddles(twl, tw + V.py.tw_off, ny);
    col_load<EMAX>(lds, buf, nxh, ny, kx0, nxh);
    __syncthreads();
    run_fft<N, EMAX, true, kLog2Tile, INV>(lds, V.py, kTile, kTile, 1, twl);
    col_store<EMAX>(lds, buf, nxh, ny, kx0, nxh);
}

// ---------------------------------------------------- TV spectral passes --
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

enum ZOp : int { ZOP_TV_SETUP = 0, ZOP_TV_SOLVE = 1, ZOP_VS_FWD = 2, ZOP_VS_RADIUS = 3, ZOP_VS_DECONV = 4 };

// TV z pass.  SETUP: forward z of FFT(f) (after x, y), then
//   f^ = FFT(f) * D * inv_a / N  -> SPEC_A,   rho * inv_a / N -> q,
// with both multipliers averaged over k and -k: the reference keeps
// Re(IFFT(.)) of a full spectrum, which equals the inverse of its Hermitian
// part, and on Nyquist planes D(k) != D(-k) when B0 is tilted.
// SOLVE: forward z, x^ = f^ - rho * FFT(div) * inv_a (src/inversion/admm.rs:
// 106-108, scaled by 1/N), inverse z.
template <int N, int EG, int OP>
__global__ __launch_bounds__(kThreads) void k_z_tv(const VolDesc* __restrict__ vols, LaunchTable lt,
                                                   const float2* __restrict__ tw,
                                                   const VolStatus* __restrict__ status) {
    constexpr int EMAX = col_emax(N, EG);
    __shared__ float2 lds[col_lds(N, EG)];
    __shared__ float2 twl[tw_lds(N)];
    const int bid = xcd_remap(blockIdx.x, gridDim.x);
    int lb;
    const int gv = locate(lt, bid, lb);
    const VolDesc& V = vols[gv];
    if (status && status[gv].done) return;
    const int ky = lb / V.ntx;
    const int tx = lb - ky * V.ntx;
    const int kx0 = tx * kTile;
    const int nxh = V.nxh;
    const int nz = N > 0 ? N : V.nz;
    const size_t zs = (size_t)nxh * V.ny;
    const size_t base = (size_t)ky * nxh;
    const int total = nz * kTile;
    float2* buf = V.spec[SPEC_WORK] + base;
    load_twiddles(twl, tw + V.pz.tw_off, nz);
    col_load<EMAX>(lds, buf, zs, nz, kx0, nxh);
    // x-update operands, in flight during the forward FFT
    float2 fr[OP == ZOP_TV_SOLVE ? EMAX : 1];
    float qr[OP == ZOP_TV_SOLVE ? EMAX : 1];
    if constexpr (OP == ZOP_TV_SOLVE) {
        const float2* __restrict__ F = V.spec[SPEC_A] + base;
        const float* __restrict__ Q = V.q + base;
#pragma unroll
        for (int t = 0; t < EMAX; ++t) {
            const int e = threadIdx.x + t * kThreads;
            const int kz = e >> kLog2Tile;
            const int kx = kx0 + (e & (kTile - 1));
            const bool in = e < total && kx < nxh;
            fr[t] = in ? F[kx + kz * zs] : make_float2(0.0f, 0.0f);
            qr[t] = in ? Q[kx + kz * zs] : 0.0f;
        }
    }
    __syncthreads();
    run_fft<N, EMAX, true, kLog2Tile, false>(lds, V.pz, kTile, kTile, 1, twl);
    if constexpr (OP == ZOP_TV_SETUP) {
        const double invN = 1.0 / ((double)V.nx * V.ny * V.nz);
        float2* F = V.spec[SPEC_A] + base;
        float* Q = V.q + base;
        for (int e = threadIdx.x; e < total; e += kThreads) {
            const int kz = e >> kLog2Tile, c = e & (kTile - 1);
            const int kx = kx0 + c;
            if (kx < nxh) {
                double d1, i1, d2, i2;
                tv_multipliers(V, kx, ky, kz, d1, i1);
                tv_multipliers(V, mirror_idx(kx, V.nx), mirror_idx(ky, V.ny), mirror_idx(kz, nz),
                               d2, i2);
                const double m = 0.5 * (d1 + d2) * invN;
                const double qv = V.rho * 0.5 * (i1 + i2) * invN;
                const float2 sv = lds[e];
                F[kx + kz * zs] = make_float2((float)((double)sv.x * m), (float)((double)sv.y * m));
                Q[kx + kz * zs] = (float)qv;
            }
        }
    } else {
#pragma unroll
        for (int t = 0; t < EMAX; ++t) {
            const int e = threadIdx.x + t * kThreads;
            if (e < total) {
                const float2 sv = lds[e];
                lds[e] = make_float2(fr[t].x - qr[t] * sv.x, fr[t].y - qr[t] * sv.y);
            }
        }
        __syncthreads();
        run_fft<N, EMAX, true, kLog2Tile, true>(lds, V.pz, kTile, kTile, 1, twl);
        col_store<EMAX>(lds, buf, zs, nz, kx0, nxh);
    }
}

// ----------------------------------------------------------- C2R x passes --
// Rebuild the full complex row Z = A + iB from two half spectra (negative
// frequencies by conjugate symmetry; the imaginary parts of the DC and
// Nyquist bins are dropped, as Re(IFFT) of a full spectrum does).
// step 2: A, B = rows 2p, 2p+1 of sa; step 1: A = row p of sa, B = row p of sb.
template <int EMAX>
__device__ __forceinline__ void c2r_load(float2* lds, int nx, int nxh, int RB,
                                         const float2* __restrict__ sa, const float2* __restrict__ sb,
                                         int row0, int step, int nrows) {
    const int total = RB * nx;
    float2 r[EMAX];
#pragma unroll
    for (int t = 0; t < EMAX; ++t) {
        const int e = threadIdx.x + t * kThreads;
        const int p = e / nx;
        const int k = e - p * nx;
        const bool neg = (2 * k > nx);
        const int kk = neg ? nx - k : k;
        const int ra = row0 + step * p;
        const int rb = (step == 2) ? ra + 1 : ra;
        float2 A = (e < total && ra < nrows) ? sa[kk + (size_t)ra * nxh] : make_float2(0.0f, 0.0f);
        float2 B = (e < total && rb < nrows) ? sb[kk + (size_t)rb * nxh] : make_float2(0.0f, 0.0f);
        if (kk == 0 || 2 * kk == nx) { A.y = 0.0f; B.y = 0.0f; }
        if (neg) { A.y = -A.y; B.y = -B.y; }
        r[t] = make_float2(A.x - B.y, A.y + B.x);
    }
#pragma unroll
    for (int t = 0; t < EMAX; ++t) {
        const int e = threadIdx.x + t * kThreads;
        if (e < total) lds[e] = r[t];
    }
}

// TV: x_it = Re IFFT(x^) (src/inversion/admm.rs:111-114) and the two sums of
// relative_change (src/utils/mod.rs:69-77) accumulated in FP64 per block.
template <int N, int EG>
__global__ __launch_bounds__(kThreads) void k_c2r_tv(const VolDesc* __restrict__ vols, LaunchTable lt,
                                                     const float2* __restrict__ tw, int parity,
                                                     const VolStatus* __restrict__ status) {
    constexpr int EMAX = x_emax(N);
    __shared__ float2 lds[x_lds(N)];
    __shared__ float2 twl[tw_lds(N)];
    __shared__ double sh[kThreads / 64];
    const int bid = xcd_remap(blockIdx.x, gridDim.x);
    int lb;
    const int gv = locate(lt, bid, lb);
    const VolDesc& V = vols[gv];
    if (status && status[gv].done) return;
    const int nx = N > 0 ? N : V.nx;
    const int nxh = nx / 2 + 1;
    const int RB = N > 0 ? qsmfft::fixed_rows(N) : V.rows_x;
    const int nrows = V.ny * V.nz;
    const int row0 = lb * RB * 2;
    const int total = RB * nx;
    const float2* s = V.spec[SPEC_WORK];
    load_twiddles(twl, tw + V.px.tw_off, nx);
    c2r_load<EMAX>(lds, nx, nxh, RB, s, s, row0, 2, nrows);
    const float* __restrict__ xo = V.x[parity];
    float* __restrict__ xn = V.x[parity ^ 1];
    float2 xp[EMAX];  // previous iterate, loaded during the inverse FFT
#pragma unroll
    for (int t = 0; t < EMAX; ++t) {
        const int e = threadIdx.x + t * kThreads;
        const int p = e / nx;
        const int i = e - p * nx;
        const int ra = row0 + 2 * p, rb = ra + 1;
        xp[t].x = (e < total && ra < nrows) ? xo[i + (size_t)ra * nx] : 0.0f;
        xp[t].y = (e < total && rb < nrows) ? xo[i + (size_t)rb * nx] : 0.0f;
    }
    __syncthreads();
    run_fft<N, EMAX, false, 0, true>(lds, V.px, RB, 1, nx, twl);
    double sd = 0.0, sx = 0.0;
#pragma unroll
    for (int t = 0; t < EMAX; ++t) {
        const int e = threadIdx.x + t * kThreads;
        if (e < total) {
            const int p = e / nx;
            const int i = e - p * nx;
            const int ra = row0 + 2 * p, rb = ra + 1;
            const float2 z = lds[e];
            if (ra < nrows) {
                const double d = (double)z.x - (double)xp[t].x;
                xn[i + (size_t)ra * nx] = z.x;
                sd += d * d;
                sx += (double)z.x * (double)z.x;
            }
            if (rb < nrows) {
                const double d = (double)z.y - (double)xp[t].y;
                xn[i + (size_t)rb * nx] = z.y;
                sd += d * d;
                sx += (double)z.y * (double)z.y;
            }
        }
    }
    const double tsd = block_sum(sd, sh);
    const double tsx = block_sum(sx, sh);
    if (threadIdx.x == 0) {
        V.partial[2 * lb] = tsd;
        V.partial[2 * lb + 1] = tsx;
    }
}

// ------------------------------------------------ V-SHARP spectral passes --
// VS_FWD: forward z only (finishes FFT(field) / FFT(mask) in spec[sel]).
// VS_RADIUS: SPEC_WORK holds the x/y-transformed sphere indicator; forward z
//   gives S^ = Re / count (the normalised kernel, src/kernels/smv.rs:57-64);
//   P = M^ S^ / N -> SPEC_WORK and Q = F^ (1 - S^) / N -> SPEC_C, each
//   inverse-transformed along z (SPEC_A = F^, SPEC_B = M^).  The first
//   (largest) radius also keeps S^ for the deconvolution
//   (src/bgremove/vsharp.rs:147-156).
// VS_DECONV: forward z of FFT(local), X / (1 - S^) with |1 - S^| < threshold
//   -> 0 (src/bgremove/vsharp.rs:180-196; division form of
//   src/bgremove/sharp.rs:113-123 when divide != 0), scaled by 1/N, inverse z.
template <int N, int EG, int OP>
__global__ __launch_bounds__(kThreads) void k_z_vs(const VolDesc* __restrict__ vols, LaunchTable lt,
                                                   const float2* __restrict__ tw, int sel, int ridx,
                                                   int flag) {
    constexpr int EMAX = col_emax(N, EG);
    __shared__ float2 lds[col_lds(N, EG)];
    __shared__ float2 twl[tw_lds(N)];
    const int bid = xcd_remap(blockIdx.x, gridDim.x);
    int lb;
    const int gv = locate(lt, bid, lb);
    const VolDesc& V = vols[gv];
    const int ky = lb / V.ntx;
    const int tx = lb - ky * V.ntx;
    const int kx0 = tx * kTile;
    const int nxh = V.nxh;
    const int nz = N > 0 ? N : V.nz;
    const size_t zs = (size_t)nxh * V.ny;
    const size_t base = (size_t)ky * nxh;
    const int total = nz * kTile;
    float2* buf = V.spec[OP == ZOP_VS_FWD ? sel : SPEC_WORK] + base;
    load_twiddles(twl, tw + V.pz.tw_off, nz);
    col_load<EMAX>(lds, buf, zs, nz, kx0, nxh);
    __syncthreads();
    run_fft<N, EMAX, true, kLog2Tile, false>(lds, V.pz, kTile, kTile, 1, twl);
    if constexpr (OP == ZOP_VS_FWD) {
        col_store<EMAX>(lds, buf, zs, nz, kx0, nxh);
    } else if constexpr (OP == ZOP_VS_RADIUS) {
        const float inv_count = (float)(1.0 / V.vs_counts[ridx]);
        const float invN = (float)(1.0 / ((double)V.nx * V.ny * V.nz));
        const float2* __restrict__ FH = V.spec[SPEC_A] + base;
        const float2* __restrict__ MH = V.spec[SPEC_B] + base;
        float sreg[EMAX];
        float2 fh[EMAX];
#pragma unroll
        for (int t = 0; t < EMAX; ++t) {
            const int e = threadIdx.x + t * kThreads;
            if (e < total) {
                const int kz = e >> kLog2Tile;
                const int kx = kx0 + (e & (kTile - 1));
                const float sv = lds[e].x * inv_count;
                sreg[t] = sv;
                if (kx < nxh) {
                    if (flag) V.vs_sfirst[base + kx + kz * zs] = sv;
                    const float2 m = MH[kx + kz * zs];
                    fh[t] = FH[kx + kz * zs];
                    lds[e] = make_float2(m.x * sv * invN, m.y * sv * invN);
                } else {
                    fh[t] = make_float2(0.0f, 0.0f);
                    lds[e] = make_float2(0.0f, 0.0f);
                }
            }
        }
        __syncthreads();
        run_fft<N, EMAX, true, kLog2Tile, true>(lds, V.pz, kTile, kTile, 1, twl);
        col_store<EMAX>(lds, buf, zs, nz, kx0, nxh);
        __syncthreads();
#pragma unroll
        for (int t = 0; t < EMAX; ++t) {
            const int e = threadIdx.x + t * kThreads;
            if (e < total) {
                const float hp = (1.0f - sreg[t]) * invN;
                lds[e] = make_float2(fh[t].x * hp, fh[t].y * hp);
            }
        }
        __syncthreads();
        run_fft<N, EMAX, true, kLog2Tile, true>(lds, V.pz, kTile, kTile, 1, twl);
        col_store<EMAX>(lds, V.spec[SPEC_C] + base, zs, nz, kx0, nxh);
    } else {  // ZOP_VS_DECONV
        const double invN = 1.0 / ((double)V.nx * V.ny * V.nz);
        for (int e = threadIdx.x; e < total; e += kThreads) {
            const int kz = e >> kLog2Tile, c = e & (kTile - 1);
            const int kx = kx0 + c;
            if (kx < nxh) {
                const double oms = 1.0 - (double)V.vs_sfirst[base + kx + kz * zs];
                const float2 xv = lds[e];
                float2 y;
                if (fabs(oms) < V.vs_threshold) {
                    y = make_float2(0.0f, 0.0f);
                } else if (flag) {
                    const double sc = invN / oms;
                    y = make_float2((float)(xv.x * sc), (float)(xv.y * sc));
                } else {
                    const double sc = (1.0 / oms) * invN;
                    y = make_float2((float)(xv.x * sc), (float)(xv.y * sc));
                }
                lds[e] = y;
            }
        }
        __syncthreads();
        run_fft<N, EMAX, true, kLog2Tile, true>(lds, V.pz, kTile, kTile, 1, twl);
        col_store<EMAX>(lds, buf, zs, nz, kx0, nxh);
    }
}

// Row pass of one radius: Z = P + iQ per row, one inverse FFT; Re = mask
// convolution (src/kernels/smv.rs:79-105), Im = high-passed field
// (src/bgremove/vsharp.rs:163-168).  Erosion keeps a voxel when
// conv > delta; conv times the sphere's voxel count is an integer
// neighbour count, so it is rounded and the test is made on that integer in
// FP64, which reproduces the reference's f64 decision.  A voxel takes the
// value of the first (largest) radius whose eroded mask contains it
// (src/bgremove/vsharp.rs:170-176).
template <int N, int EG>
__global__ __launch_bounds__(kThreads) void k_c2r_vs_select(const VolDesc* __restrict__ vols,
                                                            LaunchTable lt,
                                                            const float2* __restrict__ tw, int ridx,
                                                            double delta) {
    constexpr int EMAX = x_emax(N);
    __shared__ float2 lds[x_lds(N)];
    __shared__ float2 twl[tw_lds(N)];
    const int bid = xcd_remap(blockIdx.x, gridDim.x);
    int lb;
    const int gv = locate(lt, bid, lb);
    const VolDesc& V = vols[gv];
    const int nx = N > 0 ? N : V.nx;
    const int nxh = nx / 2 + 1;
    const int RB = N > 0 ? qsmfft::fixed_rows(N) : V.rows_x;
    const int nrows = V.ny * V.nz;
    const int row0 = lb * RB;
    load_twiddles(twl, tw + V.px.tw_off, nx);
    c2r_load<EMAX>(lds, nx, nxh, RB, V.spec[SPEC_WORK], V.spec[SPEC_C], row0, 1, nrows);
    __syncthreads();
    run_fft<N, EMAX, false, 0, true>(lds, V.px, RB, 1, nx, twl);
    const double count = V.vs_counts[ridx];
    const float countf = (float)count;
    const double norm = 1.0 / count;
    const int total = RB * nx;
#pragma unroll
    for (int t = 0; t < EMAX; ++t) {
        const int e = threadIdx.x + t * kThreads;
        const int p = e / nx;
        const int i = e - p * nx;
        const int r = row0 + p;
        if (e < total && r < nrows) {
            const size_t idx = i + (size_t)r * nx;
            const float2 z = lds[e];
            const float k = rintf(z.x * countf);
            if ((double)k * norm > delta && !V.vs_final[idx]) {
                V.vs_local[idx] = z.y;
                V.vs_final[idx] = 1;
            }
        }
    }
}

// Final rows: local = Re IFFT(.) inside the final mask, 0 outside
// (src/bgremove/vsharp.rs:194-196); also writes the 0/1 mask.
template <int N, int EG>
__global__ __launch_bounds__(kThreads) void k_c2r_vs_out(const VolDesc* __restrict__ vols, LaunchTable lt,
                                                         const float2* __restrict__ tw) {
    constexpr int EMAX = x_emax(N);
    __shared__ float2 lds[x_lds(N)];
    __shared__ float2 twl[tw_lds(N)];
    const int bid = xcd_remap(blockIdx.x, gridDim.x);
    int lb;
    const int gv = locate(lt, bid, lb);
    const VolDesc& V = vols[gv];
    const int nx = N > 0 ? N : V.nx;
    const int nxh = nx / 2 + 1;
    const int RB = N > 0 ? qsmfft::fixed_rows(N) : V.rows_x;
    const int nrows = V.ny * V.nz;
    const int row0 = lb * RB * 2;
    const float2* s = V.spec[SPEC_WORK];
    load_twiddles(twl, tw + V.px.tw_off, nx);
    c2r_load<EMAX>(lds, nx, nxh, RB, s, s, row0, 2, nrows);
    __syncthreads();
    run_fft<N, EMAX, false, 0, true>(lds, V.px, RB, 1, nx, twl);
    const int total = RB * nx;
#pragma unroll
    for (int t = 0; t < EMAX; ++t) {
        const int e = threadIdx.x + t * kThreads;
        if (e < total) {
            const int p = e / nx;
            const int i = e - p * nx;
            const int ra = row0 + 2 * p, rb = ra + 1;
            const float2 z = lds[e];
            if (ra < nrows) {
                const size_t idx = i + (size_t)ra * nx;
                const uint8_t fm = V.vs_final[idx];
                V.out[idx] = fm == 1 ? (double)z.x : 0.0;
                V.mask_out[idx] = fm;
            }
            if (rb < nrows) {
                const size_t idx = i + (size_t)rb * nx;
                const uint8_t fm = V.vs_final[idx];
                V.out[idx] = fm == 1 ? (double)z.y : 0.0;
                V.mask_out[idx] = fm;
            }
        }
    }
}

// ---------------------------------------------------------- launching ----
template <int N, int EG>
void launch_kind(int kind, const LaunchTable& lt, const PassArgs& a) {
    const dim3 g(lt.start[lt.nvol]), b(kThreads);
    hipStream_t st = a.stream;
    switch (kind) {
        case P_R2C_TVDIV: k_r2c_x<N, EG, SRC_TVDIV><<<g, b, 0, st>>>(a.vols, lt, a.tw, a.sel, a.parity, a.status); break;
        case P_R2C_F64: k_r2c_x<N, EG, SRC_F64><<<g, b, 0, st>>>(a.vols, lt, a.tw, a.sel, 0, a.status); break;
        case P_R2C_U8: k_r2c_x<N, EG, SRC_U8><<<g, b, 0, st>>>(a.vols, lt, a.tw, a.sel, 0, a.status); break;
        case P_R2C_F32: k_r2c_x<N, EG, SRC_F32><<<g, b, 0, st>>>(a.vols, lt, a.tw, a.sel, 0, a.status); break;
        case P_R2C_SPHERE: k_r2c_x<N, EG, SRC_SPHERE><<<g, b, 0, st>>>(a.vols, lt, a.tw, a.sel, a.ridx, a.status); break;
        case P_Y_FWD: k_y_pass<N, EG, false><<<g, b, 0, st>>>(a.vols, lt, a.tw, a.sel, a.sel2, a.status); break;
        case P_Y_INV: k_y_pass<N, EG, true><<<g, b, 0, st>>>(a.vols, lt, a.tw, a.sel, a.sel2, a.status); break;
        case P_Z_TV_SETUP: k_z_tv<N, EG, ZOP_TV_SETUP><<<g, b, 0, st>>>(a.vols, lt, a.tw, a.status); break;
        case P_Z_TV_SOLVE: k_z_tv<N, EG, ZOP_TV_SOLVE><<<g, b, 0, st>>>(a.vols, lt, a.tw, a.status); break;
        case P_Z_VS_FWD: k_z_vs<N, EG, ZOP_VS_FWD><<<g, b, 0, st>>>(a.vols, lt, a.tw, a.sel, 0, 0); break;
        case P_Z_VS_RADIUS: k_z_vs<N, EG, ZOP_VS_RADIUS><<<g, b, 0, st>>>(a.vols, lt, a.tw, 0, a.ridx, a.first); break;
        case P_Z_VS_DECONV: k_z_vs<N, EG, ZOP_VS_DECONV><<<g, b, 0, st>>>(a.vols, lt, a.tw, 0, 0, a.divide); break;
        case P_C2R_TV: k_c2r_tv<N, EG><<<g, b, 0, st>>>(a.vols, lt, a.tw, a.parity, a.status); break;
        case P_C2R_VS_SELECT: k_c2r_vs_select<N, EG><<<g, b, 0, st>>>(a.vols, lt, a.tw, a.ridx, a.delta); break;
        case P_C2R_VS_OUT: k_c2r_vs_out<N, EG><<<g, b, 0, st>>>(a.vols, lt, a.tw); break;
        default: break;
    }
}

}  // namespace qsm
