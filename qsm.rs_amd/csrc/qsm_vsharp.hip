// gfx950 kernels of V-SHARP / SHARP background-field removal
// (src/bgremove/vsharp.rs:88-199, src/bgremove/sharp.rs:64-131).
//
// Per radius, one chain of passes does all of the reference's per-radius
// work (which costs it four full C2C transforms):
//   k_r2c_x<SRC_SPHERE>   sphere indicator generated in-kernel, R2C x
//   k_y_pass<fwd>         C2C y
//   k_z_vs_radius         C2C z -> S^ = Re/count; P = M^ S^/N, Q = F^(1-S^)/N;
//                         inverse z of P and of Q
//   k_y_pass<inv> x2      inverse C2C y of P and Q
//   k_c2r_vs_select       one complex inverse row FFT of P + iQ gives the
//                         mask convolution (real part) and the high-passed
//                         field (imaginary part); erosion test on the integer
//                         neighbour count; first-radius-wins selection.
// F^ = FFT(field) and M^ = FFT(mask) are computed once (the reference
// re-transforms the mask for every radius, src/kernels/smv.rs:90-94).
#include "qsm_passes.hpp"

namespace qsm {

// Forward z pass only (finishes FFT(field) / FFT(mask)).
template <int EMAX>
__global__ __launch_bounds__(kThreads) void k_z_fwd(const VolDesc* __restrict__ vols, LaunchTable lt,
                                                    const float2* __restrict__ tw, int sel) {
    __shared__ float2 lds[kThreads * EMAX];
    const int bid = xcd_remap(blockIdx.x, gridDim.x);
    const int v = find_vol(lt, bid);
    const int lb = bid - lt.start[v];
    const VolDesc& V = vols[v];
    const int ky = lb / V.ntx;
    const int tx = lb - ky * V.ntx;
    const int kx0 = tx * kTile;
    const int nxh = V.nxh, nz = V.nz;
    const size_t zs = (size_t)nxh * V.ny;
    float2* buf = V.spec[sel] + (size_t)ky * nxh;
    for (int e = threadIdx.x; e < nz * kTile; e += kThreads) {
        const int kz = e >> kLog2Tile, c = e & (kTile - 1);
        const int kx = kx0 + c;
        lds[e] = (kx < nxh) ? buf[kx + kz * zs] : make_float2(0.0f, 0.0f);
    }
    __syncthreads();
    qsmfft::lds_fft<EMAX, true, kLog2Tile, false>(lds, V.pz, kTile, kTile, 1, tw + V.pz.tw_off);
    for (int e = threadIdx.x; e < nz * kTile; e += kThreads) {
        const int kz = e >> kLog2Tile, c = e & (kTile - 1);
        const int kx = kx0 + c;
        if (kx < nxh) buf[kx + kz * zs] = lds[e];
    }
}

// Radius step in the spectral domain.  SPEC_WORK holds the x/y-transformed
// sphere indicator; SPEC_A = F^, SPEC_B = M^.  Writes P (erosion product)
// back into SPEC_WORK and Q (high-pass product) into SPEC_C, both already
// inverse-transformed along z.  On the first (largest) radius also keeps
// S^ for the deconvolution (src/bgremove/vsharp.rs:147-156).
template <int EMAX>
__global__ __launch_bounds__(kThreads) void k_z_vs_radius(const VolDesc* __restrict__ vols, LaunchTable lt,
                                                          const float2* __restrict__ tw, int ridx,
                                                          int first) {
    constexpr int PER = EMAX;  // elements per thread upper bound
    __shared__ float2 lds[kThreads * EMAX];
    const int bid = xcd_remap(blockIdx.x, gridDim.x);
    const int v = find_vol(lt, bid);
    const int lb = bid - lt.start[v];
    const VolDesc& V = vols[v];
    const int ky = lb / V.ntx;
    const int tx = lb - ky * V.ntx;
    const int kx0 = tx * kTile;
    const int nxh = V.nxh, nz = V.nz;
    const size_t zs = (size_t)nxh * V.ny;
    const size_t base = (size_t)ky * nxh;
    float2* work = V.spec[SPEC_WORK] + base;
    const int total = nz * kTile;
    for (int e = threadIdx.x; e < total; e += kThreads) {
        const int kz = e >> kLog2Tile, c = e & (kTile - 1);
        const int kx = kx0 + c;
        lds[e] = (kx < nxh) ? work[kx + kz * zs] : make_float2(0.0f, 0.0f);
    }
    __syncthreads();
    qsmfft::lds_fft<EMAX, true, kLog2Tile, false>(lds, V.pz, kTile, kTile, 1, tw + V.pz.tw_off);
    const float inv_count = (float)(1.0 / V.vs_counts[ridx]);
    const float invN = (float)(1.0 / ((double)V.nx * V.ny * V.nz));
    float sreg[PER];
    const float2* FH = V.spec[SPEC_A] + base;
    const float2* MH = V.spec[SPEC_B] + base;
#pragma unroll
    for (int t = 0; t < PER; ++t) {
        const int e = threadIdx.x + t * kThreads;
        if (e < total) {
            const int kz = e >> kLog2Tile, c = e & (kTile - 1);
            const int kx = kx0 + c;
            const float s = lds[e].x * inv_count;  // Re FFT of the normalised sphere
            sreg[t] = s;
            if (kx < nxh) {
                if (first) V.vs_sfirst[base + kx + kz * zs] = s;
                const float2 m = MH[kx + kz * zs];
                lds[e] = make_float2(m.x * s * invN, m.y * s * invN);
            } else {
                lds[e] = make_float2(0.0f, 0.0f);
            }
        }
    }
    __syncthreads();
    qsmfft::lds_fft<EMAX, true, kLog2Tile, true>(lds, V.pz, kTile, kTile, 1, tw + V.pz.tw_off);
    float2 pv[PER];
#pragma unroll
    for (int t = 0; t < PER; ++t) {
        const int e = threadIdx.x + t * kThreads;
        if (e < total) pv[t] = lds[e];
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < PER; ++t) {
        const int e = threadIdx.x + t * kThreads;
        if (e < total) {
            const int kz = e >> kLog2Tile, c = e & (kTile - 1);
            const int kx = kx0 + c;
            if (kx < nxh) {
                work[kx + kz * zs] = pv[t];
                const float2 f = FH[kx + kz * zs];
                const float hp = (1.0f - sreg[t]) * invN;
                lds[e] = make_float2(f.x * hp, f.y * hp);
            } else {
                lds[e] = make_float2(0.0f, 0.0f);
            }
        }
    }
    __syncthreads();
    qsmfft::lds_fft<EMAX, true, kLog2Tile, true>(lds, V.pz, kTile, kTile, 1, tw + V.pz.tw_off);
    float2* qout = V.spec[SPEC_C] + base;
    for (int e = threadIdx.x; e < total; e += kThreads) {
        const int kz = e >> kLog2Tile, c = e & (kTile - 1);
        const int kx = kx0 + c;
        if (kx < nxh) qout[kx + kz * zs] = lds[e];
    }
}

// Row pass of one radius: Z = P + iQ per row, one inverse FFT; Re = mask
// convolution (src/kernels/smv.rs:79-105), Im = high-passed field
// (src/bgremove/vsharp.rs:163-168).  Erosion keeps a voxel when
// conv > delta; the convolution times the sphere's voxel count is an integer
// neighbour count, so it is rounded and the test is made on that integer in
// FP64, which reproduces the reference's f64 decision.  A voxel takes the
// value of the first (largest) radius whose eroded mask contains it
// (src/bgremove/vsharp.rs:170-176).
template <int EMAX>
__global__ __launch_bounds__(kThreads) void k_c2r_vs_select(const VolDesc* __restrict__ vols, LaunchTable lt,
                                                            const float2* __restrict__ tw, int ridx,
                                                            double delta) {
    __shared__ float2 lds[kThreads * EMAX];
    const int bid = xcd_remap(blockIdx.x, gridDim.x);
    const int v = find_vol(lt, bid);
    const int lb = bid - lt.start[v];
    const VolDesc& V = vols[v];
    const int nx = V.nx, RB = V.rows_x;
    const int nrows = V.ny * V.nz;
    const int row0 = lb * RB;
    c2r_load<EMAX>(lds, V, V.spec[SPEC_WORK], V.spec[SPEC_C], row0, 1, nrows);
    __syncthreads();
    qsmfft::lds_fft<EMAX, false, 0, true>(lds, V.px, RB, 1, nx, tw + V.px.tw_off);
    const double count = V.vs_counts[ridx];
    const float countf = (float)count;
    const double norm = 1.0 / count;
    for (int e = threadIdx.x; e < RB * nx; e += kThreads) {
        const int p = (int)fastdiv((uint32_t)e, V.m_nx);
        const int i = e - p * nx;
        const int r = row0 + p;
        if (r < nrows) {
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

// Deconvolution in the spectral domain (src/bgremove/vsharp.rs:180-196;
// src/bgremove/sharp.rs:113-123 when divide != 0): X * 1/(1-S^) with
// |1-S^| < threshold -> 0, S^ of the largest sphere, scaled by 1/N.
template <int EMAX>
__global__ __launch_bounds__(kThreads) void k_z_vs_deconv(const VolDesc* __restrict__ vols, LaunchTable lt,
                                                          const float2* __restrict__ tw, int divide) {
    __shared__ float2 lds[kThreads * EMAX];
    const int bid = xcd_remap(blockIdx.x, gridDim.x);
    const int v = find_vol(lt, bid);
    const int lb = bid - lt.start[v];
    const VolDesc& V = vols[v];
    const int ky = lb / V.ntx;
    const int tx = lb - ky * V.ntx;
    const int kx0 = tx * kTile;
    const int nxh = V.nxh, nz = V.nz;
    const size_t zs = (size_t)nxh * V.ny;
    const size_t base = (size_t)ky * nxh;
    float2* work = V.spec[SPEC_WORK] + base;
    const int total = nz * kTile;
    for (int e = threadIdx.x; e < total; e += kThreads) {
        const int kz = e >> kLog2Tile, c = e & (kTile - 1);
        const int kx = kx0 + c;
        lds[e] = (kx < nxh) ? work[kx + kz * zs] : make_float2(0.0f, 0.0f);
    }
    __syncthreads();
    qsmfft::lds_fft<EMAX, true, kLog2Tile, false>(lds, V.pz, kTile, kTile, 1, tw + V.pz.tw_off);
    const double invN = 1.0 / ((double)V.nx * V.ny * V.nz);
    for (int e = threadIdx.x; e < total; e += kThreads) {
        const int kz = e >> kLog2Tile, c = e & (kTile - 1);
        const int kx = kx0 + c;
        if (kx < nxh) {
            const double oms = 1.0 - (double)V.vs_sfirst[base + kx + kz * zs];
            const float2 x = lds[e];
            float2 y;
            if (fabs(oms) < V.vs_threshold) {
                y = make_float2(0.0f, 0.0f);
            } else if (divide) {
                const double sc = invN / oms;
                y = make_float2((float)(x.x * sc), (float)(x.y * sc));
            } else {
                const double sc = (1.0 / oms) * invN;
                y = make_float2((float)(x.x * sc), (float)(x.y * sc));
            }
            lds[e] = y;
        }
    }
    __syncthreads();
    qsmfft::lds_fft<EMAX, true, kLog2Tile, true>(lds, V.pz, kTile, kTile, 1, tw + V.pz.tw_off);
    for (int e = threadIdx.x; e < total; e += kThreads) {
        const int kz = e >> kLog2Tile, c = e & (kTile - 1);
        const int kx = kx0 + c;
        if (kx < nxh) work[kx + kz * zs] = lds[e];
    }
}

// Final rows: local = Re IFFT(.) inside the final mask, 0 outside
// (src/bgremove/vsharp.rs:194-196); also writes the 0/1 mask.
template <int EMAX>
__global__ __launch_bounds__(kThreads) void k_c2r_vs_out(const VolDesc* __restrict__ vols, LaunchTable lt,
                                                         const float2* __restrict__ tw) {
    __shared__ float2 lds[kThreads * EMAX];
    const int bid = xcd_remap(blockIdx.x, gridDim.x);
    const int v = find_vol(lt, bid);
    const int lb = bid - lt.start[v];
    const VolDesc& V = vols[v];
    const int nx = V.nx, RB = V.rows_x;
    const int nrows = V.ny * V.nz;
    const int row0 = lb * RB * 2;
    const float2* s = V.spec[SPEC_WORK];
    c2r_load<EMAX>(lds, V, s, s, row0, 2, nrows);
    __syncthreads();
    qsmfft::lds_fft<EMAX, false, 0, true>(lds, V.px, RB, 1, nx, tw + V.px.tw_off);
    for (int e = threadIdx.x; e < RB * nx; e += kThreads) {
        const int p = (int)fastdiv((uint32_t)e, V.m_nx);
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

__global__ __launch_bounds__(kThreads) void k_vs_init(const VolDesc* __restrict__ vols, LaunchTable lt) {
    const int bid = blockIdx.x;
    const int v = find_vol(lt, bid);
    const int lb = bid - lt.start[v];
    const VolDesc& V = vols[v];
    const size_t N = (size_t)V.nx * V.ny * V.nz;
    for (size_t i = (size_t)lb * kThreads * 4 + threadIdx.x; i < N && i < (size_t)(lb + 1) * kThreads * 4;
         i += kThreads) {
        V.vs_local[i] = 0.0f;
        V.vs_final[i] = 0;
    }
}

template <int EMAX>
struct VsLaunch {
    static void zfwd(hipStream_t st, const VolDesc* d, const LaunchTable& lt, const float2* tw, int sel) {
        k_z_fwd<EMAX><<<dim3(lt.start[lt.nvol]), dim3(kThreads), 0, st>>>(d, lt, tw, sel);
    }
    static void zradius(hipStream_t st, const VolDesc* d, const LaunchTable& lt, const float2* tw, int ridx,
                        int first) {
        k_z_vs_radius<EMAX><<<dim3(lt.start[lt.nvol]), dim3(kThreads), 0, st>>>(d, lt, tw, ridx, first);
    }
    static void zdeconv(hipStream_t st, const VolDesc* d, const LaunchTable& lt, const float2* tw, int divide) {
        k_z_vs_deconv<EMAX><<<dim3(lt.start[lt.nvol]), dim3(kThreads), 0, st>>>(d, lt, tw, divide);
    }
    static void select(hipStream_t st, const VolDesc* d, const LaunchTable& lt, const float2* tw, int ridx,
                       double delta) {
        k_c2r_vs_select<EMAX><<<dim3(lt.start[lt.nvol]), dim3(kThreads), 0, st>>>(d, lt, tw, ridx, delta);
    }
    static void out(hipStream_t st, const VolDesc* d, const LaunchTable& lt, const float2* tw) {
        k_c2r_vs_out<EMAX><<<dim3(lt.start[lt.nvol]), dim3(kThreads), 0, st>>>(d, lt, tw);
    }
};

#define QSM_EMAX_DISPATCH(emax, call) \
    do {                              \
        if ((emax) <= 16) VsLaunch<16>::call; \
        else if ((emax) <= 24) VsLaunch<24>::call; \
        else VsLaunch<32>::call;      \
    } while (0)

void launch_vs_zfwd(int emax, hipStream_t st, const VolDesc* d, const LaunchTable& lt, const float2* tw,
                    int sel) {
    QSM_EMAX_DISPATCH(emax, zfwd(st, d, lt, tw, sel));
}
void launch_vs_zradius(int emax, hipStream_t st, const VolDesc* d, const LaunchTable& lt, const float2* tw,
                       int ridx, int first) {
    QSM_EMAX_DISPATCH(emax, zradius(st, d, lt, tw, ridx, first));
}
void launch_vs_zdeconv(int emax, hipStream_t st, const VolDesc* d, const LaunchTable& lt, const float2* tw,
                       int divide) {
    QSM_EMAX_DISPATCH(emax, zdeconv(st, d, lt, tw, divide));
}
void launch_vs_select(hipStream_t st, const VolDesc* d, const LaunchTable& lt, const float2* tw, int ridx,
                      double delta) {
    VsLaunch<16>::select(st, d, lt, tw, ridx, delta);
}
void launch_vs_out(hipStream_t st, const VolDesc* d, const LaunchTable& lt, const float2* tw) {
    VsLaunch<16>::out(st, d, lt, tw);
}
void launch_vs_init(hipStream_t st, const VolDesc* d, const LaunchTable& lt) {
    k_vs_init<<<dim3(lt.start[lt.nvol]), dim3(kThreads), 0, st>>>(d, lt);
}

}  // namespace qsm
