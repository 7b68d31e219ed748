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

#include <type_traits>

#include "fft_col2.hpp"
#include "qsm_internal.hpp"

namespace qsm {

using qsmfft::kThreads;
using qsmfft::fastdiv;

constexpr int kMaxLen = 1024;

// Device pointers held in the volume descriptor are generic (flat) to the
// compiler; accessed through global-address-space types they compile to
// global loads/stores with a scalar base instead of flat operations, which
// also keeps them off the LDS wait counter.
#define QSM_GLOBAL __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ QSM_GLOBAL T* glb(T* p) { return (QSM_GLOBAL T*)p; }
template <class T>
__device__ __forceinline__ const QSM_GLOBAL T* glb(const T* p) { return (const QSM_GLOBAL T*)p; }

// 8-byte complex access through a global pointer (one dwordx2 load/store).
typedef float qsm_f2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float2 ldg2(const QSM_GLOBAL float2* p) {
    const qsm_f2v v = *(const QSM_GLOBAL qsm_f2v*)p;
    return make_float2(v.x, v.y);
}
__device__ __forceinline__ void stg2(QSM_GLOBAL float2* p, float2 a) {
    qsm_f2v v;
    v.x = a.x;
    v.y = a.y;
    *(QSM_GLOBAL qsm_f2v*)p = v;
}

// elements per thread of an x (row) pass / a column pass
__host__ __device__ constexpr int x_emax(int n) {
    return n > 0 ? (qsmfft::fixed_rows(n) * n + kThreads - 1) / kThreads : kRowEmaxRt;
}
__host__ __device__ constexpr int xc_emax(int n) {
    return n > 0 ? (qsmfft::fixed_rows_c(n) * n + kThreads - 1) / kThreads : kRowEmaxRt;
}
__host__ __device__ constexpr int col_emax(int n, int eg) {
    return n > 0 ? (kTile * n + kThreads - 1) / kThreads : eg;
}
__host__ __device__ constexpr int x_lds(int n) { return n > 0 ? qsmfft::fixed_rows(n) * n : kThreads * kRowEmaxRt; }
__host__ __device__ constexpr int col_lds(int n, int eg) { return n > 0 ? kTile * n : kThreads * eg; }
__host__ __device__ constexpr int tw_lds(int n) { return n > 0 ? n : kMaxLen; }
// column passes of the lengths fft_col2.hpp covers use its padded layout
__host__ __device__ constexpr int col_lds2(int n, int eg) {
    return qsmfft::col2_ok(n) ? 16 * n + 256 : col_lds(n, eg);
}

// The second LDS buffer of the runtime-plan transforms: one per kernel and
// size, shared by every transform the kernel runs.
template <int SIZE>
__device__ __forceinline__ float2* pp_buffer() {
    __shared__ float2 pp[SIZE];
    return pp;
}

// The M-point transform of a runtime plan (P.n == M: a chirp-z convolution
// or a direct length) on the compile-time stages (fft_fixed.hpp), in place
// in buf, when the tile holds the sequences those stages are built for:
// column tiles of 2^LOG2C columns (layout e * 2^LOG2C + c), or row tiles of
// kThreads * EMAX / M sequences at stride M.  Returns false (nothing done)
// otherwise.
template <int EMAX, bool COLFAST, int LOG2C, bool INV, int M>
__device__ __forceinline__ bool fixed_pow2_stages(float2* buf, const qsmfft::Plan& P, int ncols,
                                                  const float2* twl) {
    constexpr int NC = COLFAST ? (1 << LOG2C) : (kThreads * EMAX) / M;
    if constexpr (NC < 1 || NC * M > kThreads * EMAX) {
        return false;
    } else {
        if (P.n != M || ncols != NC) return false;
        qsmfft::fixed_fft<M, NC, COLFAST, INV>(buf, twl);
        return true;
    }
}

// twl: the pass's twiddle table in LDS; twg: the twiddle bank in global
// memory (the chirp-z tables of runtime-planned lengths).
template <int N, int EMAX, bool COLFAST, int LOG2C, bool INV, int NC = 0>
__device__ __forceinline__ void run_fft(float2* lds, const qsmfft::Plan& P, int ncols, int es, int cs,
                                        const float2* twl, const float2* twg) {
    if constexpr (N > 0) {
        qsmfft::fixed_fft<N, NC ? NC : (COLFAST ? kTile : qsmfft::fixed_rows(N)), COLFAST, INV>(lds, twl);
    } else {
        // runtime-planned lengths: out-of-place stages between the kernel's
        // tile (kThreads * EMAX elements in every such kernel) and a second
        // buffer of that size; a transform of length 128, 256, 512 or 1024 (a
        // chirp-z convolution, or a direct 1024) runs on the compile-time
        // stages instead, in place, when the tile holds exactly the sequences
        // they are built for
        auto inner = [&](float2* src, float2* dst, auto inv, int wes, int wcs) -> float2* {
            constexpr bool I = decltype(inv)::value;
            if (fixed_pow2_stages<EMAX, COLFAST, LOG2C, I, 128>(src, P, ncols, twl) ||
                fixed_pow2_stages<EMAX, COLFAST, LOG2C, I, 256>(src, P, ncols, twl) ||
                fixed_pow2_stages<EMAX, COLFAST, LOG2C, I, 512>(src, P, ncols, twl) ||
                fixed_pow2_stages<EMAX, COLFAST, LOG2C, I, 1024>(src, P, ncols, twl))
                return src;
            return qsmfft::plan_stages<COLFAST, LOG2C, I>(src, dst, P, ncols, wes, wcs, twl);
        };
        qsmfft::lds_fft_any<COLFAST, LOG2C, INV>(lds, pp_buffer<kThreads * EMAX>(), P, ncols, es, cs, twg, inner);
    }
}

// A pass's twiddle table (exp(-2 pi i q / n), q < n) is fetched into
// registers before the pass issues its tile loads and written to LDS after
// them: a table load whose value had to reach LDS first would hold every block
// for one memory round trip before its tile loads went out.
// The table of plan P (its length P.n: the axis length, or the chirp-z
// convolution length).
template <int N>
struct TwStage {
    static constexpr int K = (tw_lds(N) + kThreads - 1) / kThreads;
    float2 r[K];
    int n;
    __device__ __forceinline__ void fetch(const float2* __restrict__ tw, const qsmfft::Plan& P) {
        n = N > 0 ? N : P.n;
        const QSM_GLOBAL float2* g = glb(tw + P.tw_off);
#pragma unroll
        for (int t = 0; t < K; ++t) r[t] = ldg2(g + min((int)threadIdx.x + t * kThreads, n - 1));
    }
    // threads past n hold entry n - 1 and write it again (no branch)
    __device__ __forceinline__ void commit(float2* twl) const {
#pragma unroll
        for (int t = 0; t < K; ++t) twl[min((int)threadIdx.x + t * kThreads, n - 1)] = r[t];
    }
};

// exp(-2 pi i k / n) of a z pass: its twiddle table, except for a chirp-z
// length, whose table is that of the convolution length.
template <int N>
__device__ __forceinline__ float2 ztw(const VolDesc& V, const float2* twl, int kz, int nz) {
    if (N == 0 && V.pz.bl_n) {
        float sn, cs;
        sincospif(-2.0f * (float)kz / (float)nz, &sn, &cs);
        return make_float2(cs, sn);
    }
    return twl[kz];
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

// relative_change = ||x - x_prev|| / (||x|| + 1e-20), converged when < tol
// (src/inversion/admm.rs:117-119, src/utils/mod.rs:69-77), from the per-block
// FP64 sums (sum d^2, sum x^2) that the TV C2R pass left in V.partial.  One
// block, fixed summation order: repeated runs give identical results.
__device__ __forceinline__ void tv_reduce_block(const VolDesc& V, int gv, VolStatus* status_rw,
                                                VolStatus* mirror, double* sh) {
    typedef double d2v __attribute__((ext_vector_type(2)));
    const QSM_GLOBAL d2v* part = (const QSM_GLOBAL d2v*)glb(V.partial);
    double a = 0.0, b = 0.0;
#pragma unroll 4
    for (int q = threadIdx.x; q < V.npartial; q += kThreads) {
        const d2v v = part[q];
        a += v.x;
        b += v.y;
    }
    const double ta = block_sum(a, sh);
    const double tb = block_sum(b, sh);
    if (threadIdx.x == 0) {
        const double rel = sqrt(ta) / (sqrt(tb) + 1e-20);
        VolStatus st = status_rw[gv];
        st.iters += 1;
        st.rel = rel;
        if (rel < V.tol) st.done = 1;
        status_rw[gv] = st;
        if (mirror) {
            // host-visible copy, read by the host loop after an event wait
            mirror[gv] = st;
            __threadfence_system();
        }
    }
}

__device__ __forceinline__ double fftfreq_at(int i, int n, double d) {
    // src/fft.rs:544-566: index i of numpy fftfreq(n, d)
    const double val = 1.0 / ((double)n * d);
    const int s = (i < (n + 1) / 2) ? i : i - n;
    return (double)s * val;
}

__device__ __forceinline__ int mirror_idx(int i, int n) { return i == 0 ? 0 : n - i; }

// ------------------------------------------------ packed half spectrum --
// A half spectrum stores nxh = (nx + 1) / 2 columns kx.  For even nx the x
// transform of a real row has two real bins, DC and Nyquist (kx = nx/2); both
// go to column 0 as (DC, Nyquist), so every row holds nx/2 complex values and
// the y / z passes work on whole 16-column tiles with rows of nx/2 * 8 bytes.
// Column 0 then carries the kx = 0 plane P0 and the kx = nx/2 plane PN, both
// real in (y, z) and so Hermitian in (ky, kz).  The forward y pass
// separates them: after it, entry ky of column 0 holds
//   ky = 0, ny/2 : P0(ky) + i PN(ky)            (both real in ky, left packed)
//   0 < ky < ny/2: P0(ky)
//   ny/2 < ky    : PN(ny - ky)
// so the z pass sees column 0 at ky < ny/2 as frequency (0, ky), at
// ky > ny/2 as (nx/2, ny - ky), and at ky = 0, ny/2 a column Z = A + i B of
// two z spectra A (kx = 0) and B (kx = nx/2), each Hermitian in kz, which a
// real per-frequency multiplier pair (mA, mB) maps to
//   alpha Z(kz) + beta conj(Z(-kz)),  alpha = (mA + mB)/2, beta = (mA - mB)/2.
// The inverse y pass packs the entries back before transforming.  Odd nx has
// no Nyquist bin and no packing.
__device__ __forceinline__ bool packed_x(const VolDesc& V) { return (V.nx & 1) == 0; }

// Frequency (fx, fy) of column kx of a z-pass tile at ky (not for the two
// doubly packed columns, see pk_double).
__device__ __forceinline__ void pk_freq(const VolDesc& V, int kx, int ky, int& fx, int& fy) {
    const bool nyq = kx == 0 && packed_x(V) && 2 * ky > V.ny;
    fx = nyq ? V.nx / 2 : kx;
    fy = nyq ? V.ny - ky : ky;
}

// Tile tx at ky holds a doubly packed column 0 (block-uniform).
__device__ __forceinline__ bool pk_double(const VolDesc& V, int tx, int ky) {
    return tx == 0 && packed_x(V) && (ky == 0 || 2 * ky == V.ny);
}

// alpha z + beta conj(zm) for the multiplier pair (mA, mB)
__device__ __forceinline__ float2 pk_apply(float2 z, float2 zm, float mA, float mB) {
    const float a = 0.5f * (mA + mB), b = 0.5f * (mA - mB);
    return make_float2(a * z.x + b * zm.x, a * z.y - b * zm.y);
}
__device__ __forceinline__ float2 pk_apply_d(float2 z, float2 zm, double mA, double mB) {
    const double a = 0.5 * (mA + mB), b = 0.5 * (mA - mB);
    return make_float2((float)(a * (double)z.x + b * (double)zm.x), (float)(a * (double)z.y - b * (double)zm.y));
}

// Entry s of column 0 after the forward y transform (w = W(s), wm = W(ny-s)).
__device__ __forceinline__ float2 ypk_fwd(int s, int ny, float2 w, float2 wm) {
    if (s == 0 || 2 * s == ny) return w;
    if (2 * s < ny) return make_float2(0.5f * (w.x + wm.x), 0.5f * (w.y - wm.y));  // P0(s)
    return make_float2(0.5f * (wm.y + w.y), 0.5f * (w.x - wm.x));                  // PN(ny-s)
}
// W(s) rebuilt before the inverse y transform from entries s and ny - s.
__device__ __forceinline__ float2 ypk_inv(int s, int ny, float2 own, float2 mir) {
    if (s == 0 || 2 * s == ny) return own;
    if (2 * s < ny) return make_float2(own.x - mir.y, own.y + mir.x);  // P0(s) + i PN(s)
    return make_float2(mir.x + own.y, own.x - mir.y);                  // conj P0 + i conj PN
}

// The same on column 0 of a tile of TW columns held in LDS (lds[row * TW +
// c], ny <= kMaxLen); every thread of the block calls it.
template <bool INV, int TW = kTile>
__device__ __forceinline__ void ypk_lds(float2* lds, int ny) {
    constexpr int Q = (TW == kTile ? kWideTileMaxLen : kMaxLen) / kThreads;
    float2 own[Q], mir[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int s = min((int)threadIdx.x + q * kThreads, ny - 1);
        own[q] = lds[s * TW];
        mir[q] = lds[mirror_idx(s, ny) * TW];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int s = threadIdx.x + q * kThreads;
        if (s < ny) lds[s * TW] = INV ? ypk_inv(s, ny, own[q], mir[q]) : ypk_fwd(s, ny, own[q], mir[q]);
    }
    __syncthreads();
}

// Binary sphere of src/kernels/smv.rs:19-67 (un-normalised: 1 inside).  The
// squared distance is formed in FP64 in the reference's order with every
// product and sum rounded (contraction off), so boundary voxels are
// classified exactly as the reference (and the host voxel count) do.
__device__ __forceinline__ float sphere_element(const VolDesc& V, int i, int r, int ridx) {
    // The reference rounds every product and sum (no fused multiply-add);
    // with FMA contraction voxels at exactly the radius (e.g. (4,3,0) *
    // 0.9 mm at 4.5 mm) fall outside.  The pragma covers only operations
    // written in this scope (the __dmul_rn / __dadd_rn helpers are still
    // contracted after inlining), so the arithmetic is spelled out here.
#pragma clang fp contract(off)
    const int k = (int)fastdiv((uint32_t)r, V.m_ny);
    const int j = r - k * V.ny;
    const double xi = (i <= V.nx / 2) ? (double)i : (double)(i - V.nx);
    const double yj = (j <= V.ny / 2) ? (double)j : (double)(j - V.ny);
    const double zk = (k <= V.nz / 2) ? (double)k : (double)(k - V.nz);
    const double dx = xi * V.vsx, dy = yj * V.vsy, dz = zk * V.vsz;
    const double dxx = dx * dx, dyy = dy * dy, dzz = dz * dz;
    const double d2 = (dxx + dyy) + dzz;
    const double rad = V.vs_radii[ridx];
    return (d2 <= rad * rad) ? 1.0f : 0.0f;
}

// ---------------------------------------------------------- tile helpers --
// Column tile: element (row, c) at lds[row*TW + c] <-> base[kx0 + c + row*stride];
// columns kx >= nxh read as zero and are not written.
template <int EMAX, int TW = kTile>
__device__ __forceinline__ void col_load(float2* lds, const QSM_GLOBAL float2* __restrict__ base,
                                         size_t stride, int n, int kx0, int nxh) {
    constexpr int L2 = tile_log2(TW);
    float2 r[EMAX];
    const int total = n * TW;
    // unconditional loads: rows past n and columns past nxh read in-bounds
    // elements that are never stored
#pragma unroll
    for (int t = 0; t < EMAX; ++t) {
        const int e = threadIdx.x + t * kThreads;
        const int row = min(e >> L2, n - 1);
        const int kx = min(kx0 + (e & (TW - 1)), nxh - 1);
        r[t] = ldg2(base + kx + row * stride);
    }
#pragma unroll
    for (int t = 0; t < EMAX; ++t) {
        const int e = threadIdx.x + t * kThreads;
        if (e < total) lds[e] = r[t];
    }
}

template <int EMAX, int TW = kTile>
__device__ __forceinline__ void col_store(const float2* lds, QSM_GLOBAL float2* __restrict__ base,
                                          size_t stride, int n, int kx0, int nxh) {
    constexpr int L2 = tile_log2(TW);
    const int total = n * TW;
#pragma unroll
    for (int t = 0; t < EMAX; ++t) {
        const int e = threadIdx.x + t * kThreads;
        const int row = e >> L2;
        const int kx = kx0 + (e & (TW - 1));
        if (e < total && kx < nxh) stg2(base + kx + row * stride, lds[e]);
    }
}

// Tile I/O in the register layouts of fft_col2.hpp (thread (j, c) = (t >> 4,
// t & 15), column kx0 + c).  L16: rows j + 16 r, r < R0; LR: rows j + R0 r,
// r < 16, threads j < R0.  Columns kx >= nxh load column nxh - 1 (in bounds,
// their results are never stored) and are not written.
static_assert(kTile == 16 && kThreads == 256, "fft_col2.hpp assumes 16-column tiles and 256 threads");

// The TV / RTS z-pass operands f^ (SPEC_A) and q are written once at setup
// and read at every iteration by the same z-pass block, so they are stored
// tile-major: the tile of TW columns at (tx, ky) holds its nz rows of TW
// elements contiguously.  Element (kx, kz) of that tile is at
// zop_base(...) + kx + kz * TW.
__host__ __device__ inline size_t zop_base(int ky, int tx, int ntx, int nz, int TW) {
    return ((size_t)ky * ntx + tx) * (size_t)nz * TW - (size_t)tx * TW;
}

template <int R0>
__device__ __forceinline__ void l16_load(float2 (&v)[16], const QSM_GLOBAL float2* __restrict__ base,
                                         size_t stride, int kx0, int nxh) {
    const int c = threadIdx.x & 15, j = threadIdx.x >> 4;
    const int kx = min(kx0 + c, nxh - 1);
    const QSM_GLOBAL float2* p = base + kx + (size_t)j * stride;
#pragma unroll
    for (int r = 0; r < R0; ++r) v[r] = ldg2(p + (size_t)(16 * r) * stride);
}

template <int R0>
__device__ __forceinline__ void l16_store(const float2 (&v)[16], QSM_GLOBAL float2* __restrict__ base,
                                          size_t stride, int kx0, int nxh) {
    const int c = threadIdx.x & 15, j = threadIdx.x >> 4;
    const int kx = kx0 + c;
    if (kx >= nxh) return;
    QSM_GLOBAL float2* p = base + kx + (size_t)j * stride;
#pragma unroll
    for (int r = 0; r < R0; ++r) stg2(p + (size_t)(16 * r) * stride, v[r]);
}

template <int R0>
__device__ __forceinline__ void lr_store(const float2 (&v)[16], QSM_GLOBAL float2* __restrict__ base,
                                         size_t stride, int kx0, int nxh) {
    const int c = threadIdx.x & 15, j = threadIdx.x >> 4;
    const int kx = kx0 + c;
    if (kx >= nxh || j >= R0) return;
    QSM_GLOBAL float2* p = base + kx + (size_t)j * stride;
#pragma unroll
    for (int r = 0; r < 16; ++r) stg2(p + (size_t)(R0 * r) * stride, v[r]);
}

// ------------------------------------------------ x-pass source fills ----
enum SrcKind : int {
    SRC_F64 = 0, SRC_U8 = 1, SRC_F32 = 2, SRC_SPHERE = 3, SRC_TVDIV = 4,
    SRC_TVDIV_REC = 5,    // SRC_TVDIV that also records the shrink pattern
    SRC_TVADJ_FIRST = 6,  // adjoint source of the last iteration: x-bar = mask * chi-bar
    SRC_TVADJ = 7,        // adjoint source of an earlier iteration
    SRC_ACC = 8,          // the adjoint accumulator
    SRC_X0 = 9,           // x[0] (the RTS starting point)
    SRC_WRAPLAP = 10,     // masked wrapped Laplacian of the f64 phase in field
    SRC_TVDIV_W = 11,     // SRC_TVDIV with NLTV weights (buffer ck)
    SRC_NLTV_BND = 12     // first step of a later NLTV Newton round (previous weights ck)
};

// wrap into [-pi, pi] as src/unwrap/laplacian.rs:41-50 (f64 remainder).
__device__ __forceinline__ double wrap_pi(double x) {
    const double two_pi = 6.283185307179586476925286766559;
    const double pi = 3.14159265358979323846264338327950288;
    double y = fmod(x, two_pi);
    if (y > pi) y -= two_pi;
    else if (y < -pi) y += two_pi;
    return y;
}

// Wrapped periodic Laplacian at voxel (i, row r), zero outside the mask
// (src/unwrap/laplacian.rs:67-122 and the masking of :146-150), in FP64.
__device__ __forceinline__ float wraplap_at(const VolDesc& V, int i, int r) {
#pragma clang fp contract(off)
    const int nx = V.nx, ny = V.ny, nz = V.nz;
    const size_t idx = i + (size_t)r * nx;
    if (!glb(V.mask)[idx]) return 0.0f;
    const int k = (int)fastdiv((uint32_t)r, V.m_ny);
    const int j = r - k * ny;
    const size_t nxy = (size_t)nx * ny;
    auto at = [&](int ii, int jj, int kk) { return glb(V.field)[ii + (size_t)jj * nx + (size_t)kk * nxy]; };
    const double u = at(i, j, k);
    const double uxp = at(i + 1 < nx ? i + 1 : 0, j, k), uxm = at(i > 0 ? i - 1 : nx - 1, j, k);
    const double uyp = at(i, j + 1 < ny ? j + 1 : 0, k), uym = at(i, j > 0 ? j - 1 : ny - 1, k);
    const double uzp = at(i, j, k + 1 < nz ? k + 1 : 0), uzm = at(i, j, k > 0 ? k - 1 : nz - 1);
    const double dx2 = 1.0 / (V.vsx * V.vsx), dy2 = 1.0 / (V.vsy * V.vsy), dz2 = 1.0 / (V.vsz * V.vsz);
    const double lx = (wrap_pi(uxp - u) - wrap_pi(u - uxm)) * dx2;
    const double ly = (wrap_pi(uyp - u) - wrap_pi(u - uym)) * dy2;
    const double lz = (wrap_pi(uzp - u) - wrap_pi(u - uzm)) * dz2;
    return (float)(lx + ly + lz);
}

template <int SRC>
__device__ __forceinline__ float simple_src(const VolDesc& V, int i, int r, int aux) {
    if constexpr (SRC == SRC_F64) return (float)glb(V.field)[i + (size_t)r * V.nx];
    else if constexpr (SRC == SRC_U8) return (float)glb(V.mask)[i + (size_t)r * V.nx];
    else if constexpr (SRC == SRC_F32) return glb(V.vs_local)[i + (size_t)r * V.nx];
    else if constexpr (SRC == SRC_ACC) return glb(V.adj_acc)[i + (size_t)r * V.nx];
    else if constexpr (SRC == SRC_X0) return glb(V.x[0])[i + (size_t)r * V.nx];
    else if constexpr (SRC == SRC_WRAPLAP) return wraplap_at(V, i, r);
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
// The operands of CH elements (2*CH voxels) of the divergence.
// W: per-voxel shrink weights (NLTV, weighted_shrink of src/utils/mod.rs:
// 55-59): wv = w at i, i-e_x, i-e_y, i-e_z.
template <int CH, bool W = false>
struct TvDivRound {
    float xv[CH][2][7], uv[CH][2][6];
    float wv[CH][2][W ? 4 : 1];
    int idx[CH][2];
    bool ok[CH][2];
};

template <int CH, bool W = false>
__device__ __forceinline__ void tvdiv_load(TvDivRound<CH, W>& R, const VolDesc& V, int nx, int t0, int EMAX,
                                           int total, int row0, int nrows,
                                           const QSM_GLOBAL float* __restrict__ x,
                                           const QSM_GLOBAL float* __restrict__ u,
                                           const QSM_GLOBAL float* __restrict__ w = nullptr) {
    const int ny = V.ny, nz = V.nz;
    const int nxy = nx * ny, N = nxy * nz;
#pragma unroll
    for (int c = 0; c < CH; ++c) {
        const int e = threadIdx.x + (t0 + c) * kThreads;
        const int p = e / nx;
        const int i = e - p * nx;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int r = row0 + 2 * p + h;
            const bool ok = (t0 + c < EMAX) && (e < total) && (r < nrows);
            R.ok[c][h] = ok;
            const int rr = ok ? r : 0;
            const int ii = ok ? i : 0;
            const int k = (int)fastdiv((uint32_t)rr, V.m_ny);
            const int j = rr - k * ny;
            const int idx = ii + rr * nx;
            R.idx[c][h] = idx;
            const int dxp = (ii + 1 < nx) ? 1 : 1 - nx;
            const int dxm = (ii > 0) ? -1 : nx - 1;
            const int dyp = (j + 1 < ny) ? nx : nx - nxy;
            const int dym = (j > 0) ? -nx : nxy - nx;
            const int dzp = (k + 1 < nz) ? nxy : nxy - N;
            const int dzm = (k > 0) ? -nxy : N - nxy;
            R.xv[c][h][0] = x[idx];
            R.xv[c][h][1] = x[idx + dxp];
            R.xv[c][h][2] = x[idx + dyp];
            R.xv[c][h][3] = x[idx + dzp];
            R.xv[c][h][4] = x[idx + dxm];
            R.xv[c][h][5] = x[idx + dym];
            R.xv[c][h][6] = x[idx + dzm];
            R.uv[c][h][0] = u[idx];
            R.uv[c][h][1] = u[N + idx];
            R.uv[c][h][2] = u[2 * N + idx];
            R.uv[c][h][3] = u[idx + dxm];
            R.uv[c][h][4] = u[N + idx + dym];
            R.uv[c][h][5] = u[2 * N + idx + dzm];
            if constexpr (W) {
                R.wv[c][h][0] = w[idx];
                R.wv[c][h][1] = w[idx + dxm];
                R.wv[c][h][2] = w[idx + dym];
                R.wv[c][h][3] = w[idx + dzm];
            }
        }
    }
}

__device__ __forceinline__ int8_t shrink_sign(float v, float t) {
    return v > t ? (int8_t)1 : (v < -t ? (int8_t)-1 : (int8_t)0);
}

template <int CH, bool REC, bool W = false>
__device__ __forceinline__ void tvdiv_compute(const TvDivRound<CH, W>& R, float2* lds, const VolDesc& V,
                                              int t0, int EMAX, int total,
                                              QSM_GLOBAL float* __restrict__ uo,
                                              QSM_GLOBAL int8_t* __restrict__ rec, float wnorm = 1.0f) {
    const int N = V.nx * V.ny * V.nz;
    const float t = V.thr, hx = V.hx, hy = V.hy, hz = V.hz;
#pragma unroll
    for (int c = 0; c < CH; ++c) {
        const int e = threadIdx.x + (t0 + c) * kThreads;
        float d[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const float xc = R.xv[c][h][0];
            const float vx = (R.xv[c][h][1] - xc) * hx + R.uv[c][h][0];
            const float vy = (R.xv[c][h][2] - xc) * hy + R.uv[c][h][1];
            const float vz = (R.xv[c][h][3] - xc) * hz + R.uv[c][h][2];
            const float vxm = (xc - R.xv[c][h][4]) * hx + R.uv[c][h][3];
            const float vym = (xc - R.xv[c][h][5]) * hy + R.uv[c][h][4];
            const float vzm = (xc - R.xv[c][h][6]) * hz + R.uv[c][h][5];
            // thresholds lambda/rho, times the weight (divided by its
            // normalisation) of the voxel each component belongs to
            float ti = t, txm = t, tym = t, tzm = t;
            if constexpr (W) {
                ti = t * (R.wv[c][h][0] / wnorm);
                txm = t * (R.wv[c][h][1] / wnorm);
                tym = t * (R.wv[c][h][2] / wnorm);
                tzm = t * (R.wv[c][h][3] / wnorm);
            }
            const float zx = shrinkf(vx, ti), zy = shrinkf(vy, ti), zz = shrinkf(vz, ti);
            if (R.ok[c][h]) {
                const int idx = R.idx[c][h];
                uo[idx] = vx - zx;
                uo[N + idx] = vy - zy;
                uo[2 * N + idx] = vz - zz;
                if constexpr (REC) {
                    // d z / d v of this update, for the unrolled adjoint
                    rec[idx] = shrink_sign(vx, t);
                    rec[N + idx] = shrink_sign(vy, t);
                    rec[2 * N + idx] = shrink_sign(vz, t);
                }
            }
            const float gx = 2.0f * zx - vx, gy = 2.0f * zy - vy, gz = 2.0f * zz - vz;
            const float gxm = 2.0f * shrinkf(vxm, txm) - vxm;
            const float gym = 2.0f * shrinkf(vym, tym) - vym;
            const float gzm = 2.0f * shrinkf(vzm, tzm) - vzm;
            d[h] = R.ok[c][h] ? (gx - gxm) * hx + (gy - gym) * hy + (gz - gzm) * hz : 0.0f;
        }
        if (t0 + c < EMAX && e < total) lds[e] = make_float2(d[0], d[1]);
    }
}

// Divergence of g = 2z - v for the z/u update fused in front of the
// x-subproblem: v = grad(x) + u (periodic forward differences,
// src/utils/gradient.rs:21-58), z = shrink(v, lambda/rho), u' = v - z,
// g = 2z - v (src/inversion/admm.rs:127-142), div = bdiv(g)
// (src/utils/gradient.rs:71-108).  The x component of g at i-1 is recomputed
// from x and u at i-1 (likewise y and z), so g never goes to memory.
// Rounds of CH elements (2*CH voxels) are software-pipelined: the loads of
// round r+1 are issued before round r is computed and stored.
template <int EMAX, bool REC, bool W = false>
__device__ __forceinline__ void fill_tvdiv(float2* lds, const VolDesc& V, int nx, int RB,
                                           const QSM_GLOBAL float* __restrict__ x,
                                           const QSM_GLOBAL float* __restrict__ u,
                                           QSM_GLOBAL float* __restrict__ uo, int row0, int nrows,
                                           QSM_GLOBAL int8_t* __restrict__ rec,
                                           const QSM_GLOBAL float* __restrict__ w = nullptr,
                                           float wnorm = 1.0f) {
    constexpr int CH = (W || EMAX <= kRowEmaxRt) ? 1 : 2;
    constexpr int NR = (EMAX + CH - 1) / CH;
    const int total = RB * nx;
    TvDivRound<CH, W> A, B;
    tvdiv_load<CH, W>(A, V, nx, 0, EMAX, total, row0, nrows, x, u, w);
#pragma unroll
    for (int r = 0; r < NR; r += 2) {
        if (r + 1 < NR) tvdiv_load<CH, W>(B, V, nx, (r + 1) * CH, EMAX, total, row0, nrows, x, u, w);
        tvdiv_compute<CH, REC, W>(A, lds, V, r * CH, EMAX, total, uo, rec, wnorm);
        if (r + 1 < NR) {
            if (r + 2 < NR) tvdiv_load<CH, W>(A, V, nx, (r + 2) * CH, EMAX, total, row0, nrows, x, u, w);
            tvdiv_compute<CH, REC, W>(B, lds, V, (r + 1) * CH, EMAX, total, uo, rec, wnorm);
        }
    }
}

// Lane rotations across the 64-wide wavefront (DPP wave_rol / wave_ror):
// lane l receives the value of lane l + 1 (lane_next) or l - 1 (lane_prev),
// modulo 64.
__device__ __forceinline__ float lane_next(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x134, 0xF, 0xF,
                                                                 false));
}
__device__ __forceinline__ float lane_prev(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x13C, 0xF, 0xF,
                                                                 false));
}

// Row lengths and pair-rows per block (RB) the wave-row form takes: at
// least one pair-row per wave, at most 4 column groups per lane.  Measured
// (one column per lane): faster than the element-wise fill at 144 ... 256;
// at 384 (6 groups of 64, 222 VGPRs) slower, and with 7 or 8 groups its two
// row buffers take all 256 VGPRs (one wave per SIMD, where the element-wise
// fill keeps two).  With two columns per lane, 384 takes 178 VGPRs (2 waves
// per SIMD against the element-wise fill's 3) and its TV-ADMM ran 9.0 ->
// 8.8 solves/s; 512 takes 234.
__host__ __device__ constexpr bool tvdiv_rows_form(int n, int rb) {
    return n > 0 && n <= 256 && rb >= 4;
}

// Columns per lane of the wave-row form (VW): 4 or 2 where groups of 256 or
// 128 columns pad the row no further than groups of 64 do (no more idle
// lanes), so that every load and store moves 16 or 8 bytes per lane
// (measured at 256: 277.9 -> 274.8 us per 4-volume launch with 2, 268.2
// with 4; 151 -> 118 -> 104 VGPRs); otherwise 1.
__host__ __device__ constexpr int tv_pad(int n, int g) { return (n + g - 1) / g * g; }
__host__ __device__ constexpr int tvdiv_rows_vw(int n) {
    return (n % 4 == 0 && tv_pad(n, 256) == tv_pad(n, 64))
               ? 4
               : ((n % 2 == 0 && tv_pad(n, 128) == tv_pad(n, 64)) ? 2 : 1);
}

template <int VW>
using lanev = std::conditional_t<VW == 1, float, std::conditional_t<VW == 2, float2, float4>>;
typedef float qsm_f4v __attribute__((ext_vector_type(4)));
// element e (a compile-time constant after unrolling) of a lane's columns
template <int VW>
__device__ __forceinline__ float vel(const lanev<VW>& v, int e) {
    if constexpr (VW == 1) return v;
    else if constexpr (VW == 2) return e == 0 ? v.x : v.y;
    else return e == 0 ? v.x : (e == 1 ? v.y : (e == 2 ? v.z : v.w));
}
template <int VW>
__device__ __forceinline__ lanev<VW> vpack(const float (&a)[VW]) {
    if constexpr (VW == 1) return a[0];
    else if constexpr (VW == 2) return make_float2(a[0], a[1]);
    else return make_float4(a[0], a[1], a[2], a[3]);
}
template <int VW>
__device__ __forceinline__ float vfirst(const lanev<VW>& v) { return vel<VW>(v, 0); }
template <int VW>
__device__ __forceinline__ float vlast(const lanev<VW>& v) { return vel<VW>(v, VW - 1); }
template <int VW>
__device__ __forceinline__ lanev<VW> vzero() {
    const float z[VW] = {};
    return vpack<VW>(z);
}
template <int VW>
__device__ __forceinline__ lanev<VW> ldv(const QSM_GLOBAL float* p) {
    if constexpr (VW == 1) return *p;
    else if constexpr (VW == 2) return ldg2((const QSM_GLOBAL float2*)p);
    else {
        const qsm_f4v v = *(const QSM_GLOBAL qsm_f4v*)p;
        return make_float4(v.x, v.y, v.z, v.w);
    }
}
template <int VW>
__device__ __forceinline__ void stv(QSM_GLOBAL float* p, lanev<VW> v) {
    if constexpr (VW == 1) *p = v;
    else if constexpr (VW == 2) stg2((QSM_GLOBAL float2*)p, v);
    else {
        qsm_f4v w;
        w.x = v.x;
        w.y = v.y;
        w.z = v.z;
        w.w = v.w;
        *(QSM_GLOBAL qsm_f4v*)p = w;
    }
}

// One row of the wave-row form below: lane l holds columns VW l ... VW l +
// VW - 1 of each group of G = 64 VW columns, m < C = ceil(NX / G) (in a
// partial last group the lanes past the row's end hold its last VW columns
// again and are never stored).
template <int NX, int VW>
struct TvRow {
    static constexpr int G = 64 * VW;
    static constexpr int C = (NX + G - 1) / G;
    lanev<VW> xc[C], ux[C], uy[C], uz[C], xzp[C], xzm[C], uzm[C];
};

// first column of the lane's set in group m
template <int NX, int VW>
__device__ __forceinline__ int tv_col(int lane, int m) {
    constexpr int G = 64 * VW, C = (NX + G - 1) / G;
    return (NX % G != 0 && m == C - 1) ? min(VW * lane + G * m, NX - VW) : VW * lane + G * m;
}

// Row r of the volume (r = j + k ny, k the plane): its own x and u, x on the
// planes k +- 1 and u_z on plane k - 1 (periodic).
template <int NX, int VW>
__device__ __forceinline__ void tvrow_load(TvRow<NX, VW>& R, int r, int k, int nz, int nxy, int N, int lane,
                                           const QSM_GLOBAL float* __restrict__ x,
                                           const QSM_GLOBAL float* __restrict__ u) {
    constexpr int C = TvRow<NX, VW>::C;
    const int base = r * NX;
    const int dzp = (k + 1 < nz) ? nxy : nxy - N;
    const int dzm = (k > 0) ? -nxy : N - nxy;
    const QSM_GLOBAL float* xr = x + base;
    const QSM_GLOBAL float* ur = u + base;
#pragma unroll
    for (int m = 0; m < C; ++m) R.xc[m] = ldv<VW>(xr + tv_col<NX, VW>(lane, m));
#pragma unroll
    for (int m = 0; m < C; ++m) {
        const int i = tv_col<NX, VW>(lane, m);
        R.ux[m] = ldv<VW>(ur + i);
        R.uy[m] = ldv<VW>(ur + N + i);
        R.uz[m] = ldv<VW>(ur + 2 * N + i);
        R.xzp[m] = ldv<VW>(xr + dzp + i);
        R.xzm[m] = ldv<VW>(xr + dzm + i);
        R.uzm[m] = ldv<VW>(ur + 2 * N + dzm + i);
    }
}

// The z/u update and the divergence at one voxel (the arithmetic of
// tvdiv_compute): x and its six neighbours, u and the components of u at
// the three lower neighbours; writes u' = v - shrink(v), returns div(2z - v).
__device__ __forceinline__ float tv_voxel(float xc, float xp, float xm, float xyp, float xym, float xzp,
                                          float xzm, float ux, float uy, float uz, float uxm, float uym,
                                          float uzm, float t, float hx, float hy, float hz, float& ox,
                                          float& oy, float& oz) {
    const float vx = (xp - xc) * hx + ux;
    const float vy = (xyp - xc) * hy + uy;
    const float vz = (xzp - xc) * hz + uz;
    const float vxm = (xc - xm) * hx + uxm;
    const float vym = (xc - xym) * hy + uym;
    const float vzm = (xc - xzm) * hz + uzm;
    const float zx = shrinkf(vx, t), zy = shrinkf(vy, t), zz = shrinkf(vz, t);
    ox = vx - zx;
    oy = vy - zy;
    oz = vz - zz;
    const float gx = 2.0f * zx - vx, gy = 2.0f * zy - vy, gz = 2.0f * zz - vz;
    const float gxm = 2.0f * shrinkf(vxm, t) - vxm;
    const float gym = 2.0f * shrinkf(vym, t) - vym;
    const float gzm = 2.0f * shrinkf(vzm, t) - vzm;
    return (gx - gxm) * hx + (gy - gym) * hy + (gz - gzm) * hz;
}

// fill_tvdiv with the block's RB pair-rows split over its four waves: wave w
// takes pair-rows [w RB / 4, (w + 1) RB / 4), lane l the columns VW l ...
// VW l + VW - 1 (+ 64 VW m).  Row indices are then the same in every lane
// (scalar), the in-row neighbours come from the lane's own columns and, at
// their ends, from lane rotations (x and u_x left of the lane's first
// column from lane l - 1, x right of its last from lane l + 1; with a
// partial last column group, the periodic wrap between columns NX - 1 and 0
// from lane broadcasts) and the y neighbours from the wave's neighbouring
// rows in registers; only the rows above and below the wave's rows, and
// y-wrapping rows, are loaded again.  Loads per voxel: 7 plus 3 / (rows per
// wave), against 13.
template <int NX, int RB>
__device__ __forceinline__ void fill_tvdiv_rows(float2* lds, const VolDesc& V, int row0, int nrows,
                                                const QSM_GLOBAL float* __restrict__ x,
                                                const QSM_GLOBAL float* __restrict__ u,
                                                QSM_GLOBAL float* __restrict__ uo) {
    constexpr int VW = tvdiv_rows_vw(NX);
    using vt = lanev<VW>;
    constexpr int G = TvRow<NX, VW>::G;
    constexpr int C = TvRow<NX, VW>::C;
    constexpr int LV = (NX - G * (C - 1)) / VW;  // valid lanes of the last column group
    static_assert(kThreads == 256, "four waves per block");
    const int ny = V.ny, nz = V.nz;
    const int nxy = NX * ny, N = nxy * nz;
    const float t = V.thr, hx = V.hx, hy = V.hy, hz = V.hz;
    const int lane = (int)(threadIdx.x & 63);
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    // pair-rows [p0, p0 + nrw / 2) of this wave; the same count for every
    // wave (a compile-time constant, so the row loop unrolls) when 4 | RB
    const int p0 = (RB % 4 == 0) ? wv * (RB / 4) : (wv * RB) >> 2;
    const int nrw = (RB % 4 == 0) ? RB / 2 : 2 * ((((wv + 1) * RB) >> 2) - p0);
    const int r0 = row0 + 2 * p0;
    // plane and in-plane index of the wave's first row, then stepped
    int k = r0 / ny;
    int j = r0 - k * ny;
    TvRow<NX, VW> A, B;
    vt hym_x[C], hym_uy[C], hyp_x[C];  // rows above / below the wave's rows
    if (r0 < nrows) {
        tvrow_load<NX, VW>(A, r0, k, nz, nxy, N, lane, x, u);
        // the row before the wave's first row, periodic in y
        const int ra = j > 0 ? r0 - 1 : r0 + ny - 1;
        const QSM_GLOBAL float* xr = x + ra * NX;
        const QSM_GLOBAL float* ur = u + N + ra * NX;
#pragma unroll
        for (int m = 0; m < C; ++m) {
            hym_x[m] = ldv<VW>(xr + tv_col<NX, VW>(lane, m));
            hym_uy[m] = ldv<VW>(ur + tv_col<NX, VW>(lane, m));
        }
    } else {
#pragma unroll
        for (int m = 0; m < C; ++m) hym_x[m] = hym_uy[m] = vzero<VW>();
    }
    vt pxc[C], puy[C], dev[C];
    // row q of the wave from cur, with the next row's loads issued into nxt;
    // ODD: the second row of a pair-row (the pair goes to LDS)
    auto step = [&](TvRow<NX, VW>& cur, TvRow<NX, VW>& nxt, const int q, auto odd) {
        constexpr bool ODD = decltype(odd)::value;
        const int r = r0 + q;
        const bool ok = r < nrows;
        // plane / in-plane index of row r + 1
        int jn = j + 1, kn = k;
        if (jn == ny) {
            jn = 0;
            ++kn;
        }
        if (q + 1 < nrw && r + 1 < nrows) tvrow_load<NX, VW>(nxt, r + 1, kn, nz, nxy, N, lane, x, u);
        if (q + 2 == nrw && ok) {
            // the row after the wave's last row rl = r + 1 (in-plane index
            // jn), periodic in y; issued one row ahead
            const int rl = r + 1;
            const bool okl = rl < nrows;
            const int rb = (jn + 1 < ny) ? rl + 1 : rl - ny + 1;
            const QSM_GLOBAL float* xr = x + (okl ? rb : 0) * NX;
#pragma unroll
            for (int m = 0; m < C; ++m) hyp_x[m] = okl ? ldv<VW>(xr + tv_col<NX, VW>(lane, m)) : vzero<VW>();
        }
        vt xym[C], uym[C], xyp[C];
        if (!ok) {
            // rows past the volume (the last block): d = 0, nothing stored
#pragma unroll
            for (int m = 0; m < C; ++m) {
                xym[m] = uym[m] = xyp[m] = vzero<VW>();
                cur.xc[m] = cur.ux[m] = cur.uy[m] = cur.uz[m] = vzero<VW>();
                cur.xzp[m] = cur.xzm[m] = cur.uzm[m] = vzero<VW>();
            }
        } else {
            // y - 1
            if (q == 0) {
#pragma unroll
                for (int m = 0; m < C; ++m) {
                    xym[m] = hym_x[m];
                    uym[m] = hym_uy[m];
                }
            } else if (j == 0) {
                const int ra = r + ny - 1;
#pragma unroll
                for (int m = 0; m < C; ++m) {
                    xym[m] = ldv<VW>(x + ra * NX + tv_col<NX, VW>(lane, m));
                    uym[m] = ldv<VW>(u + N + ra * NX + tv_col<NX, VW>(lane, m));
                }
            } else {
#pragma unroll
                for (int m = 0; m < C; ++m) {
                    xym[m] = pxc[m];
                    uym[m] = puy[m];
                }
            }
            // y + 1
            if (q + 1 == nrw) {
#pragma unroll
                for (int m = 0; m < C; ++m) xyp[m] = hyp_x[m];
            } else if (j + 1 == ny) {
                const int rb = r - ny + 1;
#pragma unroll
                for (int m = 0; m < C; ++m) xyp[m] = ldv<VW>(x + rb * NX + tv_col<NX, VW>(lane, m));
            } else {
#pragma unroll
                for (int m = 0; m < C; ++m) xyp[m] = nxt.xc[m];
            }
        }
        // in-row neighbours at the ends of the lane's columns: x right of
        // its last column from lane l + 1 (lane 63: lane 0 of the next
        // column group), x and u_x left of its first column from lane l - 1
        // (lane 0: lane 63 of the previous group); columns NX - 1 and 0 are
        // neighbours (periodic), through lane broadcasts when the last group
        // is partial
        float rn[C], rp[C], rpu[C];
#pragma unroll
        for (int m = 0; m < C; ++m) {
            rn[m] = lane_next(vfirst<VW>(cur.xc[m]));
            rp[m] = lane_prev(vlast<VW>(cur.xc[m]));
            rpu[m] = lane_prev(vlast<VW>(cur.ux[m]));
        }
        // (whole groups: the rotations already wrap column NX - 1 to 0)
        float x0 = 0.0f, xl = 0.0f, uxl = 0.0f;
        if constexpr (NX % G != 0) {
            x0 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, vfirst<VW>(cur.xc[0])), 0));
            xl = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, vlast<VW>(cur.xc[C - 1])), LV - 1));
            uxl = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, vlast<VW>(cur.ux[C - 1])), LV - 1));
        }
        vt d[C];
#pragma unroll
        for (int m = 0; m < C; ++m) {
            float xp = lane == 63 ? rn[(m + 1) % C] : rn[m];
            float xm = lane == 0 ? rp[(m + C - 1) % C] : rp[m];
            float uxm = lane == 0 ? rpu[(m + C - 1) % C] : rpu[m];
            if constexpr (NX % G != 0) {
                if (m == C - 1 && lane == LV - 1) xp = x0;  // column NX - 1
                if (m == 0 && lane == 0) {                  // column 0
                    xm = xl;
                    uxm = uxl;
                }
            }
            const bool col_ok = NX % G == 0 || m < C - 1 || lane < LV;
            QSM_GLOBAL float* o = uo + r * NX + VW * lane + G * m;
            if constexpr (VW == 1) {
                // (tv_voxel written out: the same arithmetic, with the stores
                // between the two halves, in fewer registers)
                const float xc = cur.xc[m];
                const float vx = (xp - xc) * hx + cur.ux[m];
                const float vy = (xyp[m] - xc) * hy + cur.uy[m];
                const float vz = (cur.xzp[m] - xc) * hz + cur.uz[m];
                const float vxm = (xc - xm) * hx + uxm;
                const float vym = (xc - xym[m]) * hy + uym[m];
                const float vzm = (xc - cur.xzm[m]) * hz + cur.uzm[m];
                const float zx = shrinkf(vx, t), zy = shrinkf(vy, t), zz = shrinkf(vz, t);
                if (ok && col_ok) {
                    o[0] = vx - zx;
                    o[N] = vy - zy;
                    o[2 * N] = vz - zz;
                }
                const float gx = 2.0f * zx - vx, gy = 2.0f * zy - vy, gz = 2.0f * zz - vz;
                const float gxm = 2.0f * shrinkf(vxm, t) - vxm;
                const float gym = 2.0f * shrinkf(vym, t) - vym;
                const float gzm = 2.0f * shrinkf(vzm, t) - vzm;
                d[m] = ok ? (gx - gxm) * hx + (gy - gym) * hy + (gz - gzm) * hz : 0.0f;
            } else {
                // the lane's columns e = 0 ... VW - 1: in-row neighbours
                // from its own registers except at the two ends
                float ox[VW], oy[VW], oz[VW], dd[VW];
#pragma unroll
                for (int e = 0; e < VW; ++e) {
                    const float xpe = e + 1 < VW ? vel<VW>(cur.xc[m], e + 1) : xp;
                    const float xme = e > 0 ? vel<VW>(cur.xc[m], e - 1) : xm;
                    const float uxme = e > 0 ? vel<VW>(cur.ux[m], e - 1) : uxm;
                    dd[e] = tv_voxel(vel<VW>(cur.xc[m], e), xpe, xme, vel<VW>(xyp[m], e), vel<VW>(xym[m], e),
                                     vel<VW>(cur.xzp[m], e), vel<VW>(cur.xzm[m], e), vel<VW>(cur.ux[m], e),
                                     vel<VW>(cur.uy[m], e), vel<VW>(cur.uz[m], e), uxme, vel<VW>(uym[m], e),
                                     vel<VW>(cur.uzm[m], e), t, hx, hy, hz, ox[e], oy[e], oz[e]);
                }
                if (ok && col_ok) {
                    stv<VW>(o, vpack<VW>(ox));
                    stv<VW>(o + N, vpack<VW>(oy));
                    stv<VW>(o + 2 * N, vpack<VW>(oz));
                }
                d[m] = ok ? vpack<VW>(dd) : vzero<VW>();
            }
        }
        if constexpr (ODD) {
            const int p = p0 + (q >> 1);
#pragma unroll
            for (int m = 0; m < C; ++m) {
                const bool col_ok = NX % G == 0 || m < C - 1 || lane < LV;
                if (col_ok) {
                    float2* l2 = lds + p * NX + VW * lane + G * m;
#pragma unroll
                    for (int e = 0; e < VW; ++e) l2[e] = make_float2(vel<VW>(dev[m], e), vel<VW>(d[m], e));
                }
            }
        } else {
#pragma unroll
            for (int m = 0; m < C; ++m) dev[m] = d[m];
        }
#pragma unroll
        for (int m = 0; m < C; ++m) {
            pxc[m] = cur.xc[m];
            puy[m] = cur.uy[m];
        }
        j = jn;
        k = kn;
    };
    // pair-rows of this wave (nrw / 2, wave-uniform): the row
    // buffers keep fixed roles, so nothing is indexed at run time
#pragma unroll
    for (int s = 0; s < (nrw >> 1); ++s) {
        step(A, B, 2 * s, std::false_type{});
        step(B, A, 2 * s + 1, std::true_type{});
    }
}

// First step of an NLTV Newton round after the first (src/inversion/
// nltv.rs:86-120): the previous round ended either at convergence (no z/u
// update) or after max_iter steps (a z/u update with the previous weights is
// still due); then g was overwritten with grad(x).  So: u' = u or
// v - shrink(v, t w_old) with v = grad x + u, and the divergence source is
// bdiv(grad x).
template <int EMAX>
__device__ __forceinline__ void fill_nltv_boundary(float2* lds, const VolDesc& V, int nx, int RB,
                                                   const QSM_GLOBAL float* __restrict__ x,
                                                   const QSM_GLOBAL float* __restrict__ u,
                                                   QSM_GLOBAL float* __restrict__ uo, int row0,
                                                   int nrows, bool converged,
                                                   const QSM_GLOBAL float* __restrict__ w, float wnorm) {
    const int ny = V.ny, nz = V.nz;
    const int nxy = nx * ny, N = nxy * nz;
    const float t = V.thr, hx = V.hx, hy = V.hy, hz = V.hz;
    const int total = RB * nx;
#pragma unroll
    for (int tt = 0; tt < EMAX; ++tt) {
        const int e = threadIdx.x + tt * kThreads;
        const int p = e / nx;
        const int i = e - p * nx;
        float d[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int r = row0 + 2 * p + h;
            const bool ok = (e < total) && (r < nrows);
            const int rr = ok ? r : 0;
            const int ii = ok ? i : 0;
            const int k = (int)fastdiv((uint32_t)rr, V.m_ny);
            const int j = rr - k * ny;
            const int idx = ii + rr * nx;
            const int dxp = (ii + 1 < nx) ? 1 : 1 - nx;
            const int dxm = (ii > 0) ? -1 : nx - 1;
            const int dyp = (j + 1 < ny) ? nx : nx - nxy;
            const int dym = (j > 0) ? -nx : nxy - nx;
            const int dzp = (k + 1 < nz) ? nxy : nxy - N;
            const int dzm = (k > 0) ? -nxy : N - nxy;
            const float xc = x[idx];
            const float gx = (x[idx + dxp] - xc) * hx, gy = (x[idx + dyp] - xc) * hy, gz = (x[idx + dzp] - xc) * hz;
            const float gxm = (xc - x[idx + dxm]) * hx, gym = (xc - x[idx + dym]) * hy, gzm = (xc - x[idx + dzm]) * hz;
            const float ux = u[idx], uy = u[N + idx], uz = u[2 * N + idx];
            if (ok) {
                if (converged) {
                    uo[idx] = ux;
                    uo[N + idx] = uy;
                    uo[2 * N + idx] = uz;
                } else {
                    const float ti = t * (w[idx] / wnorm);
                    const float vx = gx + ux, vy = gy + uy, vz = gz + uz;
                    uo[idx] = vx - shrinkf(vx, ti);
                    uo[N + idx] = vy - shrinkf(vy, ti);
                    uo[2 * N + idx] = vz - shrinkf(vz, ti);
                }
            }
            d[h] = ok ? (gx - gxm) * hx + (gy - gym) * hy + (gz - gzm) * hz : 0.0f;
        }
        if (e < total) lds[e] = make_float2(d[0], d[1]);
    }
}

// Adjoint of one unrolled TV-ADMM iteration, fused in front of the R2C of
// x-bar.  Forward iteration j (tol = 0, src/inversion/admm.rs:76-145):
//   c_j = bdiv(z_{j-1} - u_{j-1}),  x_j = A f - rho M c_j,
//   v_j = grad(x_j) + u_{j-1},  z_j = shrink(v_j, t),  u_j = v_j - z_j,
// with A = Re IFFT(D inv_a FFT .), M = Re IFFT(inv_a FFT .) (both
// symmetric) and t = lambda / rho.  Given c-bar_{j+1} = -rho M x-bar_{j+1}
// and s-bar_{j+1} (the adjoint of v_{j+1}), with q_j = sign(v_j) [|v_j| > t]:
//   z-bar_j = -grad c-bar_{j+1},  u-bar_j = s-bar_{j+1} + grad c-bar_{j+1},
//   s-bar_j = |q_j| z-bar_j + (1 - |q_j|) u-bar_j,
//   x-bar_j = grad^T s-bar_j = -bdiv(s-bar_j),
//   t-bar  += sum (u-bar_j - z-bar_j) q_j.
// The components of s-bar at i - e_a (needed by bdiv) are recomputed from
// c-bar, s-bar_{j+1} and q_j at i - e_a.  x-bar_j is added to the
// accumulator (field-bar = A sum_j x-bar_j); t-bar's FP64 partial is
// returned to the caller.  FIRST (j = K): x-bar_K = mask * chi-bar (chi = the
// masked x_K, src/utils/mask.rs:47-53), s-bar_K = 0.
__device__ __forceinline__ float adj_sbar(float g, float so, int8_t q) {
    return q != 0 ? -g : so + g;
}

template <int EMAX, bool FIRST>
__device__ __forceinline__ void fill_tvadj(float2* lds, const VolDesc& V, int nx, int RB, int row0,
                                           int nrows, int ck, int par, double& tsum) {
    const int ny = V.ny, nz = V.nz;
    const int nxy = nx * ny, N = nxy * nz;
    const float hx = V.hx, hy = V.hy, hz = V.hz;
    const QSM_GLOBAL float* __restrict__ c = glb(V.adj_c);
    const QSM_GLOBAL float* __restrict__ so = glb(V.adj_s[par ^ 1]);
    QSM_GLOBAL float* __restrict__ sn = glb(V.adj_s[par]);
    const QSM_GLOBAL int8_t* __restrict__ q = glb(V.ckpt) + (size_t)(ck < 0 ? 0 : ck) * 3 * N;
    QSM_GLOBAL float* __restrict__ acc = glb(V.adj_acc);
    const int total = RB * nx;
    tsum = 0.0;
#pragma unroll
    for (int t = 0; t < EMAX; ++t) {
        const int e = threadIdx.x + t * kThreads;
        const int p = e / nx;
        const int i = e - p * nx;
        float d[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int r = row0 + 2 * p + h;
            const bool ok = (e < total) && (r < nrows);
            const int rr = ok ? r : 0;
            const int ii = ok ? i : 0;
            const int idx = ii + rr * nx;
            float xb;
            if constexpr (FIRST) {
                xb = (ok && glb(V.mask)[idx]) ? (float)glb(V.adj_chibar)[idx] : 0.0f;
                if (ok) {
                    sn[idx] = 0.0f;
                    sn[N + idx] = 0.0f;
                    sn[2 * N + idx] = 0.0f;
                    acc[idx] = xb;
                }
            } else {
                const int k = (int)fastdiv((uint32_t)rr, V.m_ny);
                const int j = rr - k * ny;
                const int dxp = (ii + 1 < nx) ? 1 : 1 - nx;
                const int dxm = (ii > 0) ? -1 : nx - 1;
                const int dyp = (j + 1 < ny) ? nx : nx - nxy;
                const int dym = (j > 0) ? -nx : nxy - nx;
                const int dzp = (k + 1 < nz) ? nxy : nxy - N;
                const int dzm = (k > 0) ? -nxy : N - nxy;
                const float cc = c[idx];
                const float cxp = c[idx + dxp], cyp = c[idx + dyp], czp = c[idx + dzp];
                const float cxm = c[idx + dxm], cym = c[idx + dym], czm = c[idx + dzm];
                const float sox = so[idx], soy = so[N + idx], soz = so[2 * N + idx];
                const float soxm = so[idx + dxm], soym = so[N + idx + dym], sozm = so[2 * N + idx + dzm];
                const int8_t qx = q[idx], qy = q[N + idx], qz = q[2 * N + idx];
                const int8_t qxm = q[idx + dxm], qym = q[N + idx + dym], qzm = q[2 * N + idx + dzm];
                const float gx = (cxp - cc) * hx, gy = (cyp - cc) * hy, gz = (czp - cc) * hz;
                const float gxm = (cc - cxm) * hx, gym = (cc - cym) * hy, gzm = (cc - czm) * hz;
                const float sx = adj_sbar(gx, sox, qx), sy = adj_sbar(gy, soy, qy), sz = adj_sbar(gz, soz, qz);
                const float sxm = adj_sbar(gxm, soxm, qxm), sym = adj_sbar(gym, soym, qym),
                            szm = adj_sbar(gzm, sozm, qzm);
                xb = -((sx - sxm) * hx + (sy - sym) * hy + (sz - szm) * hz);
                if (ok) {
                    sn[idx] = sx;
                    sn[N + idx] = sy;
                    sn[2 * N + idx] = sz;
                    acc[idx] += xb;
                    tsum += (double)(sox + 2.0f * gx) * (double)qx + (double)(soy + 2.0f * gy) * (double)qy +
                            (double)(soz + 2.0f * gz) * (double)qz;
                } else {
                    xb = 0.0f;
                }
            }
            d[h] = xb;
        }
        if (e < total) lds[e] = make_float2(d[0], d[1]);
    }
}

// ------------------------------------------------------------- R2C x pass --
// Row-block index of dispatch slot lb when every plane holds a whole number
// bpp = ny / R of R-row blocks: slots run plane-fastest (lb = yb nz + k), so
// blocks dispatched together (and sharing an XCD) hold the same rows of
// neighbouring planes, and the divergence's loads of x at k +- 1 and u_z at
// k - 1 find the rows their neighbours just read in L2.  Otherwise slots are
// blocks in row order.  A bijection on [0, bpp nz) either way.
__device__ __forceinline__ int plane_fast_block(int lb, int ny, int nz, int R) {
    const int bpp = ny / R;
    if (bpp * R != ny) return lb;
    const int yb = lb / nz;
    return (lb - yb * nz) * bpp + yb;
}

// Block = RB pairs of real rows.  Pair p packs rows (2p, 2p+1) as
// z = a + i b; after a complex FFT of length nx the half spectra are
//   A[k] = (Z[k] + conj Z[-k]) / 2,   B[k] = (Z[k] - conj Z[-k]) / (2i).
// aux: TV iteration parity for SRC_TVDIV, radius index for SRC_SPHERE.
// ck: checkpoint index (SRC_TVDIV_REC: the record written; SRC_TVADJ: the
// record read, and the row of adj_tpart written).
// RBS: the block takes fixed_rows(N) >> RBS pair-rows (small volumes: more,
// smaller blocks); the host sizes rows_x and the launch the same way.
template <int N, int EG, int SRC, int RBS = 0>
__global__ __launch_bounds__(kThreads) void k_r2c_x(const VolDesc* __restrict__ vols, LaunchTable lt,
                                                    const float2* __restrict__ tw, int out_sel, int aux,
                                                    const VolStatus* status, VolStatus* status_rw,
                                                    VolStatus* mirror, int ck) {
    constexpr int RBX = N > 0 ? (qsmfft::fixed_rows(N) >> RBS) : 0;
    static_assert(N == 0 || RBX >= 1, "row shift too large for this length");
    constexpr int EMAX = N > 0 ? (RBX * N + kThreads - 1) / kThreads : kRowEmaxRt;
    __shared__ float2 lds[N > 0 ? RBX * N : x_lds(0)];
    __shared__ float2 twl[tw_lds(N)];
    const int bid = xcd_remap(blockIdx.x, gridDim.x);
    int lb;
    const int gv = locate(lt, bid, lb);
    const VolDesc& V = vols[gv];
    if (status && status[gv].done) return;
    if constexpr (SRC == SRC_TVDIV || SRC == SRC_TVDIV_REC) {
        // The first block of a volume closes the previous iteration: it
        // reduces the C2R pass's sums and may mark the volume converged.
        // Blocks of this pass that still see it running do work that is
        // discarded (they write only the spectrum and the dual variable,
        // never the iterate x that is returned).
        __shared__ double sh[kThreads / 64];
        if (status_rw && lb == 0) tv_reduce_block(V, gv, status_rw, mirror, sh);
    }
    const int nx = N > 0 ? N : V.nx;
    const int nxh = (nx + 1) / 2;
    const int RB = N > 0 ? RBX : V.rows_x;
    const int nrows = V.ny * V.nz;
    const int rb = (SRC == SRC_TVDIV || SRC == SRC_TVDIV_REC || SRC == SRC_TVDIV_W || SRC == SRC_NLTV_BND)
                       ? plane_fast_block(lb, V.ny, V.nz, 2 * RB)
                       : lb;
    const int row0 = rb * RB * 2;
    TwStage<N> tws;
    tws.fetch(tw, V.px);
    if constexpr (SRC == SRC_TVDIV && tvdiv_rows_form(N, RBX)) {
        fill_tvdiv_rows<N, RBX>(lds, V, row0, nrows, glb(V.x[aux]), glb(V.u[aux ^ 1]), glb(V.u[aux]));
    } else if constexpr (SRC == SRC_TVDIV) {
        fill_tvdiv<EMAX, false>(lds, V, nx, RB, glb(V.x[aux]), glb(V.u[aux ^ 1]), glb(V.u[aux]), row0,
                                nrows, nullptr);
    } else if constexpr (SRC == SRC_TVDIV_REC) {
        const size_t N3 = (size_t)3 * V.nx * V.ny * V.nz;
        fill_tvdiv<EMAX, true>(lds, V, nx, RB, glb(V.x[aux]), glb(V.u[aux ^ 1]), glb(V.u[aux]), row0,
                               nrows, glb(V.ckpt) + (size_t)ck * N3);
    } else if constexpr (SRC == SRC_TVDIV_W) {
        fill_tvdiv<EMAX, false, true>(lds, V, nx, RB, glb(V.x[aux]), glb(V.u[aux ^ 1]), glb(V.u[aux]),
                                      row0, nrows, nullptr, glb(V.nltv_w[ck]), (float)V.nltv_wnorm[ck]);
    } else if constexpr (SRC == SRC_NLTV_BND) {
        // status_rw carries the convergence state of the round just ended
        fill_nltv_boundary<EMAX>(lds, V, nx, RB, glb(V.x[aux]), glb(V.u[aux ^ 1]), glb(V.u[aux]), row0,
                                 nrows, status_rw[gv].done != 0, glb(V.nltv_w[ck]),
                                 (float)V.nltv_wnorm[ck]);
    } else if constexpr (SRC == SRC_TVADJ_FIRST || SRC == SRC_TVADJ) {
        __shared__ double sh[kThreads / 64];
        double ts;
        fill_tvadj<EMAX, SRC == SRC_TVADJ_FIRST>(lds, V, nx, RB, row0, nrows, ck, aux, ts);
        if constexpr (SRC == SRC_TVADJ) {
            const double bt = block_sum(ts, sh);
            if (threadIdx.x == 0) glb(V.adj_tpart)[(size_t)ck * V.adj_nblk + lb] = bt;
        }
    } else {
        fill_simple<EMAX, SRC>(lds, V, nx, RB, row0, nrows, aux);
    }
    tws.commit(twl);
    __syncthreads();
    run_fft<N, EMAX, false, 0, false, RBX>(lds, V.px, RB, 1, nx, twl, tw);
    QSM_GLOBAL float2* __restrict__ out = glb(V.spec[out_sel]);
    const int total = RB * nxh;
    if constexpr (N > 0 && N % 4 == 0) {
        // two consecutive half-spectrum entries per thread and row: 16-byte
        // stores (nxh is even, so every pair is 16-byte aligned)
        constexpr int NH2 = N / 4, TOT2 = RBX * NH2;
#pragma unroll
        for (int t = 0; t < (TOT2 + kThreads - 1) / kThreads; ++t) {
            const int e = threadIdx.x + t * kThreads;
            if (e < TOT2) {
                const int p = e / NH2;
                const int k0 = 2 * (e - p * NH2);
                const int ra = row0 + 2 * p, rb = ra + 1;
                float a4[4], b4[4];
#pragma unroll
                for (int c = 0; c < 2; ++c) {
                    const int k = k0 + c;
                    const float2 zk = lds[p * N + k];
                    const float2 zm = lds[p * N + (k == 0 ? 0 : N - k)];
                    float2 A = make_float2(0.5f * (zk.x + zm.x), 0.5f * (zk.y - zm.y));
                    float2 B = make_float2(0.5f * (zk.y + zm.y), -0.5f * (zk.x - zm.x));
                    if (k == 0) {
                        // (DC, Nyquist) of both rows: real parts of Z[0] and Z[N/2]
                        const float2 zn = lds[p * N + N / 2];
                        A = make_float2(zk.x, zn.x);
                        B = make_float2(zk.y, zn.y);
                    }
                    a4[2 * c] = A.x;
                    a4[2 * c + 1] = A.y;
                    b4[2 * c] = B.x;
                    b4[2 * c + 1] = B.y;
                }
                if (ra < nrows) stv<4>((QSM_GLOBAL float*)(out + k0 + (size_t)ra * nxh), vpack<4>(a4));
                if (rb < nrows) stv<4>((QSM_GLOBAL float*)(out + k0 + (size_t)rb * nxh), vpack<4>(b4));
            }
        }
        return;
    }
#pragma unroll
    for (int t = 0; t < EMAX; ++t) {
        const int e = threadIdx.x + t * kThreads;
        if (e < total) {
            const int p = e / nxh;
            const int k = e - p * nxh;
            const int ra = row0 + 2 * p, rb = ra + 1;
            const float2 zk = lds[p * nx + k];
            const float2 zm = lds[p * nx + (k == 0 ? 0 : nx - k)];
            float2 A = make_float2(0.5f * (zk.x + zm.x), 0.5f * (zk.y - zm.y));
            float2 B = make_float2(0.5f * (zk.y + zm.y), -0.5f * (zk.x - zm.x));
            if ((nx & 1) == 0 && k == 0) {
                // (DC, Nyquist) of both rows: real parts of Z[0] and Z[nx/2]
                const float2 zn = lds[p * nx + nx / 2];
                A = make_float2(zk.x, zn.x);
                B = make_float2(zk.y, zn.y);
            }
            if (ra < nrows) stg2(out + k + (size_t)ra * nxh, A);
            if (rb < nrows) stg2(out + k + (size_t)rb * nxh, B);
        }
    }
}

// ------------------------------------------------------------- C2C y pass --
// Block = one tile of TW consecutive kx columns at one kz, full ny (TW =
// kTile; kTileLong for runtime-planned lengths above 256).
// sel2 >= 0 transforms a second spectrum with a second set of blocks.
// SCHED 0: two-stage column FFT (fft_col2.hpp) where the length allows;
// SCHED 1: the LDS-staged Stockham stages (less LDS per block).
template <int N, int EG, bool INV, int SCHED, int TW = kTile>
__global__ __launch_bounds__(kThreads) void k_y_pass(const VolDesc* __restrict__ vols, LaunchTable lt,
                                                     const float2* __restrict__ tw, int sel, int sel2,
                                                     const VolStatus* __restrict__ status) {
    static_assert(TW == kTile || N == 0, "compiled lengths use tiles of kTile columns");
    constexpr int EMAX = col_emax(N, EG);
    constexpr bool TWO = SCHED == 0 && qsmfft::col2_ok(N);
    __shared__ float2 lds[TWO ? col_lds2(N, EG) : col_lds(N, EG)];
    __shared__ float2 twl[tw_lds(N)];
    const int bid = xcd_remap(blockIdx.x, gridDim.x);
    int lb;
    const int gv = locate(lt, bid, lb);
    const VolDesc& V = vols[gv];
    if (status && status[gv].done) return;
    const int ntx = TW == kTile ? V.ntx : V.ntx_long;
    const int per = ntx * V.nz;
    int s = sel;
    if (lb >= per) { lb -= per; s = sel2; }
    const int kz = lb / ntx;
    const int tx = lb - kz * ntx;
    const int kx0 = tx * TW;
    const int nxh = V.nxh;
    const int ny = N > 0 ? N : V.ny;
    QSM_GLOBAL float2* buf = glb(V.spec[s]) + (size_t)kz * nxh * ny;
    TwStage<N> tws;
    tws.fetch(tw, V.py);
    // tile 0 of an even-nx volume: column 0 is packed (DC and Nyquist planes)
    const bool pkc = tx == 0 && packed_x(V);
    if constexpr (TWO) {
        constexpr int R0 = N / 16;
        const int c = threadIdx.x & 15, j = threadIdx.x >> 4;
        float2 v[16];
        l16_load<R0>(v, buf, nxh, kx0, nxh);
        tws.commit(twl);
        // column 0 of a packed tile goes through LDS (free before the
        // first exchange and after the last one)
        if constexpr (INV) {
            if (pkc) {  // L16 layout: thread (j, c) holds rows j + 16 r
                if (c == 0) {
#pragma unroll
                    for (int r = 0; r < R0; ++r) lds[j + 16 * r] = v[r];
                }
                __syncthreads();
                if (c == 0) {
#pragma unroll
                    for (int r = 0; r < R0; ++r) {
                        const int row = j + 16 * r;
                        v[r] = ypk_inv(row, N, v[r], lds[mirror_idx(row, N)]);
                    }
                }
                __syncthreads();  // before the transform's first exchange writes
            }
        }
        qsmfft::Col2<R0>::template l16_to_lr<INV>(v, lds, twl);
        if constexpr (!INV) {
            if (pkc) {  // LR layout: thread (j < R0, c) holds rows j + R0 r
                __syncthreads();  // the transform's last reads are done
                if (c == 0 && j < R0) {
#pragma unroll
                    for (int r = 0; r < 16; ++r) lds[j + R0 * r] = v[r];
                }
                __syncthreads();
                if (c == 0 && j < R0) {
#pragma unroll
                    for (int r = 0; r < 16; ++r) {
                        const int row = j + R0 * r;
                        v[r] = ypk_fwd(row, N, v[r], lds[mirror_idx(row, N)]);
                    }
                }
            }
        }
        lr_store<R0>(v, buf, nxh, kx0, nxh);
    } else {
        col_load<EMAX, TW>(lds, buf, nxh, ny, kx0, nxh);
        tws.commit(twl);
        __syncthreads();
        if (INV && pkc) ypk_lds<true, TW>(lds, ny);
        run_fft<N, EMAX, true, tile_log2(TW), INV>(lds, V.py, TW, TW, 1, twl, tw);
        if (!INV && pkc) ypk_lds<false, TW>(lds, ny);
        col_store<EMAX, TW>(lds, buf, nxh, ny, kx0, nxh);
    }
}

// ---------------------------------------------------- TV spectral passes --
// The periodic second difference along one axis, (2 - 2 cos(2 pi k / n)) /
// vs^2: the Poisson eigenvalue of src/unwrap/laplacian.rs:93-118 (negated),
// 0 at k = 0 and on every length-1 axis.
__device__ inline double lap_closed(int k, int n, double vs) {
#pragma clang fp contract(off)
    const double two_pi = 6.283185307179586476925286766559;
    return (2.0 - 2.0 * cos(two_pi * k / n)) / (vs * vs);
}

// One axis's share of the DFT of the negative Laplacian stencil of
// src/kernels/laplacian.rs:20-62.  That stencil is written by assignment:
// the centre holds 2/vs^2 for every axis, the two neighbours -1/vs^2 only
// where n > 1, and where n = 2 both neighbours are the same voxel, written
// once.  So the share is the periodic second difference where n > 2, but
// 2/vs^2 (at every k, DC included) where n = 1 and (2 - cos(pi k))/vs^2,
// i.e. 1/vs^2 at k = 0 and 3/vs^2 at k = 1, where n = 2.
__device__ inline double lap_stencil(int k, int n, double vs) {
#pragma clang fp contract(off)
    if (n > 2) return lap_closed(k, n, vs);
    return (n == 1 ? 2.0 : (k == 0 ? 1.0 : 3.0)) / (vs * vs);
}

// Spectral terms at one index, in FP64 with every operation rounded (no
// contraction) so threshold decisions match the reference's f64 arithmetic:
// D = 1/3 - (k.b)^2/|k|^2, 0 where |k|^2 <= 1e-20 (src/kernels/dipole.rs:
// 24-72), and L, the DFT of the negative Laplacian stencil (lap_stencil per
// axis; exactly 0 at DC unless an axis has length 1).
__device__ inline void spec_terms(const VolDesc& V, int kx, int ky, int kz, double& D, double& L) {
#pragma clang fp contract(off)
    const double fx = fftfreq_at(kx, V.nx, V.vsx);
    const double fy = fftfreq_at(ky, V.ny, V.vsy);
    const double fz = fftfreq_at(kz, V.nz, V.vsz);
    const double kdb = fx * V.bx + fy * V.by + fz * V.bz;
    const double k2 = fx * fx + fy * fy + fz * fz;
    D = (k2 > 1e-20) ? (1.0 / 3.0) - (kdb * kdb) / k2 : 0.0;
    L = lap_stencil(kx, V.nx, V.vsx) + lap_stencil(ky, V.ny, V.vsy) + lap_stencil(kz, V.nz, V.vsz);
}

// TV: D inv_a and inv_a, inv_a = 1/(D^2 + rho L) or 0 where |.| <= 1e-20
// (src/inversion/admm.rs:157-168).
__device__ inline void tv_multipliers(const VolDesc& V, int kx, int ky, int kz, double& dinv,
                                      double& inv) {
#pragma clang fp contract(off)
    double D, L;
    spec_terms(V, kx, ky, kz, D, L);
    const double a = D * D + V.rho * L;
    inv = (fabs(a) > 1e-20) ? 1.0 / a : 0.0;
    dinv = D * inv;
}

// RTS (src/inversion/rts.rs:102-183): M = mu on the cone |D| > delta, else 0;
// inv_a = rho / (M + rho L) or 0; f^ multiplier (M / rho) inv_a where both
// are non-zero; the ADMM step uses rho = 1, so its multiplier is inv_a.
__device__ inline void rts_multipliers(const VolDesc& V, int kx, int ky, int kz, double& fm,
                                       double& inv) {
#pragma clang fp contract(off)
    double D, L;
    spec_terms(V, kx, ky, kz, D, L);
    const double M = (fabs(D) > V.filt_a) ? V.filt_b : 0.0;
    const double a = M + V.rho * L;
    inv = (fabs(a) > 1e-20) ? V.rho / a : 0.0;
    fm = (fabs(M) > 1e-20 && fabs(inv) > 1e-20) ? (M / V.rho) * inv : 0.0;
}

// The dipole kernel alone (src/kernels/dipole.rs:24-72).
__device__ inline double dipole_at(const VolDesc& V, int kx, int ky, int kz) {
    double D, L;
    spec_terms(V, kx, ky, kz, D, L);
    return D;
}

// Closed-form spectral inversions, one multiplier per frequency:
//   TKD  1/D, or sign(D)/thr where |D| <= thr      (src/inversion/tkd.rs:48-99)
//   TSVD 1/D, or 0 where |D| <= thr                 (the tsvd after it)
//   Tikhonov D / (D^2 + lambda Gamma), Gamma = 1, L or L^2, 0 where the
//     denominator is <= 1e-20                        (src/inversion/tikhonov.rs:62-131)
//   RTS step 1: lsmr_iter rounds of w += (1 - w D) D/(D^2 + 1e-6) from
//     w = D/(D^2 + 1e-6) on the cone |D| > delta, 0 off it (the per-frequency
//     refinement of src/inversion/rts.rs:117-150, applied to FFT(f))
//   Poisson: 1/lambda with lambda the periodic Laplacian's eigenvalue
//     (lap_closed summed over the axes, negated), 0
//     where |lambda| <= 1e-20 (src/unwrap/laplacian.rs:93-133)
enum Filter : int { FILT_TKD = 0, FILT_TSVD = 1, FILT_TIK_ID = 2, FILT_TIK_GRAD = 3, FILT_TIK_LAP = 4,
                    FILT_RTS_X0 = 5, FILT_POISSON = 6 };

__device__ inline double filter_mult(const VolDesc& V, int kx, int ky, int kz) {
#pragma clang fp contract(off)
    double D, L;
    spec_terms(V, kx, ky, kz, D, L);
    const double p = V.filt_a;
    switch (V.filt) {
        case FILT_TKD: {
            const double inv_t = 1.0 / p;
            return fabs(D) <= p ? (D >= 0.0 ? inv_t : -inv_t) : 1.0 / D;
        }
        case FILT_TSVD: return fabs(D) <= p ? 0.0 : 1.0 / D;
        case FILT_TIK_ID:
        case FILT_TIK_GRAD:
        case FILT_TIK_LAP: {
            const double g = V.filt == FILT_TIK_ID ? 1.0 : (V.filt == FILT_TIK_GRAD ? L : L * L);
            const double denom = D * D + p * g;
            return fabs(denom) > 1e-20 ? D / denom : 0.0;
        }
        case FILT_POISSON: {
            const double lam = -(lap_closed(kx, V.nx, V.vsx) + lap_closed(ky, V.ny, V.vsy) +
                                 lap_closed(kz, V.nz, V.vsz));
            return fabs(lam) > 1e-20 ? 1.0 / lam : 0.0;
        }
        default: {  // FILT_RTS_X0
            if (!(fabs(D) > p)) return 0.0;
            const double gk = D / (D * D + 1e-6);
            double w = gk;
            for (int it = 0; it < V.lsmr_iter; ++it) w += (1.0 - w * D) * gk;
            return w;
        }
    }
}

enum ZOp : int {
    ZOP_TV_SETUP = 0, ZOP_TV_SOLVE = 1, ZOP_VS_FWD = 2, ZOP_VS_RADIUS = 3, ZOP_VS_DECONV = 4,
    ZOP_TV_ADJ = 5,     // c-bar^ = -rho inv_a x-bar^ / N (the adjoint of the x-update in c)
    ZOP_TV_APPLYA = 6,  // field-bar^ = D inv_a acc^ / N  (the adjoint in f, A is symmetric)
    ZOP_DIPOLE = 7,     // D chi^ / N  (the dipole forward operator, self-adjoint)
    ZOP_FILTER = 8,     // closed-form inversion multiplier / N (filter_mult)
    ZOP_RTS_SETUP = 9,  // as TV_SETUP with the RTS multipliers
    ZOP_TV_SOLVE0 = 10, // first ADMM step from g = 0: x^ = f^, inverse z only
    ZOP_RTS_SOLVE = 11  // ZOP_TV_SOLVE with the stored multiplier q (RTS)
};

// The TV x-update multiplier q = rho inv_a / N, inv_a = 1/(D^2 + rho L) (0 at
// DC; src/inversion/admm.rs:157-168), averaged over k and its mirror as the
// stored one is, formed in FP32 where it is used instead of being read (it
// is the same at every iteration; reading it costs 4 bytes per spectral
// element per iteration).  There is no threshold in it besides DC, so FP32
// only perturbs it by a few ulp.  k and its mirror differ only on Nyquist
// planes, where the mirror keeps the (negative) Nyquist frequency.
// init: the x and y terms of one column (FP64, once per thread); at: one kz,
// with twz = exp(-2 pi i kz / nz) from the z pass's twiddle table.
struct TvQ {
    float sxy, sxym, k2xy, lxy, rlxy, dfz, bz, hz2, rho, scale;
    bool sym;  // the mirror of (kx, ky) is (-kx, -ky): no Nyquist index in it
    // In FP32 (the multiplier is FP32): fftfreq (src/fft.rs:544-566), whose
    // mirror is its negative except at a Nyquist index, and lap_stencil with
    // 2 - 2 cos t written as 4 sin^2(t / 2) (no cancellation).
    __device__ __forceinline__ void init(const VolDesc& V, int kx, int ky) {
        const bool nyx = 2 * kx == V.nx, nyy = 2 * ky == V.ny;
        sym = !nyx && !nyy;
        // 1/(n vs) and 1/vs^2 from the FP32 1/vs (V.hx ...)
        const float fx = (float)(kx < (V.nx + 1) / 2 ? kx : kx - V.nx) * (V.hx * __builtin_amdgcn_rcpf((float)V.nx));
        const float fy = (float)(ky < (V.ny + 1) / 2 ? ky : ky - V.ny) * (V.hy * __builtin_amdgcn_rcpf((float)V.ny));
        const float fxm = nyx ? fx : -fx, fym = nyy ? fy : -fy;
        const float bx = (float)V.bx, by = (float)V.by;
        sxy = fx * bx + fy * by;
        sxym = fxm * bx + fym * by;
        k2xy = fx * fx + fy * fy;
        auto lap = [](int k, int n, float h) {
            const float h2 = h * h;
            if (n > 2) {
                const float sn = sinpif((float)k / (float)n);
                return 4.0f * (sn * sn) * h2;
            }
            return (n == 1 ? 2.0f : (k == 0 ? 1.0f : 3.0f)) * h2;
        };
        lxy = lap(kx, V.nx, V.hx) + lap(ky, V.ny, V.hy);
        rho = (float)V.rho;
        rlxy = rho * lxy;
        dfz = V.hz * __builtin_amdgcn_rcpf((float)V.nz);
        bz = (float)V.bz;
        hz2 = V.hz * V.hz;
        scale = 0.5f * rho * __builtin_amdgcn_rcpf((float)V.nx * (float)V.ny * (float)V.nz);
    }
    __device__ __forceinline__ float at(int kz, int nz, float2 twz) const {
        const int sz = kz < (nz + 1) / 2 ? kz : kz - nz;           // fftfreq index
        const int szm = (2 * kz == nz) ? sz : -sz;                  // its mirror's
        const float fz = (float)sz * dfz, fzm = (float)szm * dfz;
        const float kdb = sxy + fz * bz, kdbm = sxym + fzm * bz;
        const float k2 = k2xy + fz * fz;
        // 2 - 2 cos(t) = 2 sin^2(t) / (1 + cos(t)) where cos(t) > 0 (no
        // cancellation); lap_stencil's shares where nz <= 2 (twz.x = cos(pi kz)
        // where nz = 2)
        const float cz = twz.x;
        const float lz = (nz > 2 ? (cz > 0.0f ? 2.0f * (twz.y * twz.y) * __builtin_amdgcn_rcpf(1.0f + cz)
                                              : 2.0f - 2.0f * cz)
                                 : (nz == 1 ? 2.0f : 2.0f - cz)) * hz2;
        const float rl = rho * (lxy + lz);
        const float rk = k2 > 1e-20f ? __builtin_amdgcn_rcpf(k2) : 0.0f;  // D = 0 at DC
        const float D = k2 > 1e-20f ? (1.0f / 3.0f) - kdb * kdb * rk : 0.0f;
        const float a = D * D + rl;
        const float ia = a > 1e-20f ? __builtin_amdgcn_rcpf(a) : 0.0f;
        // where the mirror is -k (sym, kz not the Nyquist plane) sxym = -sxy
        // and fzm = -fz exactly, so kdbm = -kdb and its term is ia bit for bit
        if (sym && 2 * kz != nz) return scale * (ia + ia);
        const float Dm = k2 > 1e-20f ? (1.0f / 3.0f) - kdbm * kdbm * rk : 0.0f;
        const float am = Dm * Dm + rl;
        const float iam = am > 1e-20f ? __builtin_amdgcn_rcpf(am) : 0.0f;
        return scale * (ia + iam);
    }
    // The terms of at() that depend on kz only, shared by every column of a
    // z tile: (fz bz, fz^2, rho lz / vs_z^2) with fz the plane's frequency
    // and lz its share of the Laplacian (cos and sin of 2 pi kz / nz here).
    __device__ static float4 plane(const VolDesc& V, int kz, int nz) {
        const float dfz = V.hz * __builtin_amdgcn_rcpf((float)nz);
        const int sz = kz < (nz + 1) / 2 ? kz : kz - nz;
        const float fz = (float)sz * dfz;
        float sn, cs;
        sincospif(2.0f * (float)kz / (float)nz, &sn, &cs);
        const float lz = (nz > 2 ? (cs > 0.0f ? 2.0f * (sn * sn) * __builtin_amdgcn_rcpf(1.0f + cs)
                                              : 2.0f - 2.0f * cs)
                                 : (nz == 1 ? 2.0f : 2.0f - cs)) *
                         (V.hz * V.hz);
        return make_float4(fz * (float)V.bz, fz * fz, (float)V.rho * lz, 0.0f);
    }
    // at() from the plane's terms P = plane(V, kz, nz)
    __device__ __forceinline__ float at_plane(int kz, int nz, float4 P) const {
        const float kdb = sxy + P.x;
        const float k2 = k2xy + P.y;
        const float rl = rlxy + P.z;
        const float rk = k2 > 1e-20f ? __builtin_amdgcn_rcpf(k2) : 0.0f;  // D = 0 at DC
        const float D = k2 > 1e-20f ? (1.0f / 3.0f) - kdb * kdb * rk : 0.0f;
        const float a = D * D + rl;
        const float ia = a > 1e-20f ? __builtin_amdgcn_rcpf(a) : 0.0f;
        if (sym && 2 * kz != nz) return scale * (ia + ia);
        const float kdbm = sxym + (2 * kz == nz ? P.x : -P.x);  // the mirror's kz is -kz but on Nyquist
        const float Dm = k2 > 1e-20f ? (1.0f / 3.0f) - kdbm * kdbm * rk : 0.0f;
        const float am = Dm * Dm + rl;
        const float iam = am > 1e-20f ? __builtin_amdgcn_rcpf(am) : 0.0f;
        return scale * (ia + iam);
    }
};

// Symmetrised multipliers of the spectral setup at one frequency, averaged
// over k and its mirror (the reference keeps Re(IFFT(.)) of a full
// spectrum, the inverse of its Hermitian part; on Nyquist planes D(k) !=
// D(-k) when B0 is tilted): m for f^ and q = rho inv_a / N (TV) or
// inv_a / N (RTS).
template <int OP>
__device__ inline void setup_mult(const VolDesc& V, int kx, int ky, int kz, double& m, double& qv) {
    const double invN = 1.0 / ((double)V.nx * V.ny * V.nz);
    const int mx = mirror_idx(kx, V.nx), my = mirror_idx(ky, V.ny), mz = mirror_idx(kz, V.nz);
    double d1, i1, d2, i2;
    if constexpr (OP == ZOP_TV_SETUP) {
        tv_multipliers(V, kx, ky, kz, d1, i1);
        tv_multipliers(V, mx, my, mz, d2, i2);
        m = 0.5 * (d1 + d2) * invN;
        qv = V.rho * 0.5 * (i1 + i2) * invN;
    } else {
        rts_multipliers(V, kx, ky, kz, d1, i1);
        rts_multipliers(V, mx, my, mz, d2, i2);
        m = 0.5 * (d1 + d2) * invN;
        qv = 0.5 * (i1 + i2) * invN;
    }
}

// Symmetrised multiplier of APPLYA (D inv_a / N), DIPOLE (D / N) or FILTER.
template <int OP>
__device__ inline double apply_mult(const VolDesc& V, int kx, int ky, int kz) {
    const double invN = 1.0 / ((double)V.nx * V.ny * V.nz);
    const int mx = mirror_idx(kx, V.nx), my = mirror_idx(ky, V.ny), mz = mirror_idx(kz, V.nz);
    if constexpr (OP == ZOP_TV_APPLYA) {
        double d1, i1, d2, i2;
        tv_multipliers(V, kx, ky, kz, d1, i1);
        tv_multipliers(V, mx, my, mz, d2, i2);
        return 0.5 * (d1 + d2) * invN;
    } else if constexpr (OP == ZOP_DIPOLE) {
        return 0.5 * (dipole_at(V, kx, ky, kz) + dipole_at(V, mx, my, mz)) * invN;
    } else {
        return 0.5 * (filter_mult(V, kx, ky, kz) + filter_mult(V, mx, my, mz)) * invN;
    }
}

// RTS step multiplier q = inv_a / N (symmetrised), in FP64.
__device__ inline float rts_q(const VolDesc& V, int kx, int ky, int kz) {
    double m, qv;
    setup_mult<ZOP_RTS_SETUP>(V, kx, ky, kz, m, qv);
    return (float)qv;
}

// TV z pass.  SETUP: forward z of FFT(f) (after x, y), then
//   f^ = FFT(f) * D * inv_a / N  -> SPEC_A,   rho * inv_a / N -> q,
// with both multipliers averaged over k and -k: the reference keeps
// Re(IFFT(.)) of a full spectrum, which equals the inverse of its Hermitian
// part, and on Nyquist planes D(k) != D(-k) when B0 is tilted.
// SOLVE: forward z, x^ = f^ - q * FFT(div) (src/inversion/admm.rs:106-108,
// scaled by 1/N), inverse z, with q formed in-kernel (TvQ); RTS_SOLVE reads
// the stored q instead.  ADJ: forward z, -q * spectrum, inverse z (q from
// TvQ).  APPLYA / DIPOLE: forward z, times the Hermitian part of D inv_a / N
// or of D / N (in FP64), inverse z.
// The per-iteration ops (SOLVE, SOLVE0, ADJ, RTS_SOLVE) at the lengths of
// fft_col2.hpp: forward z from the load layout, the product in registers,
// inverse z back to the load layout; two LDS exchanges in all.
template <int N, int OP>
__device__ __forceinline__ void z_tv_col2(const VolDesc& V, float2* lds, float2* twl, const TwStage<N>& tws,
                                          int ky, int kx0, bool dpk) {
    constexpr int R0 = N / 16;
    using C = qsmfft::Col2<R0>;
    const int nxh = V.nxh;
    const size_t zs = (size_t)nxh * V.ny;
    const size_t base = (size_t)ky * nxh;
    QSM_GLOBAL float2* buf = glb(V.spec[SPEC_WORK]) + base;
    float2 v[16];
    if constexpr (OP == ZOP_TV_SOLVE0) {
        // g = 0 before the first step (AdmmBuffers::new), so div = 0 and x^ = f^
        l16_load<R0>(v, glb(V.spec[SPEC_A]) + zop_base(ky, kx0 / kTile, V.ntx, N, kTile), kTile, kx0, nxh);
        tws.commit(twl);
        C::template l16_to_lr<true>(v, lds, twl);
        lr_store<R0>(v, buf, zs, kx0, nxh);
        return;
    }
    l16_load<R0>(v, buf, zs, kx0, nxh);
    // x-update operands in the LR layout, in flight during the forward FFT
    const int c = threadIdx.x & 15, j = threadIdx.x >> 4;
    const bool act = j < R0;
    constexpr bool HAS_F = OP == ZOP_TV_SOLVE || OP == ZOP_RTS_SOLVE;
    constexpr bool Q_MEM = OP == ZOP_RTS_SOLVE;
    float2 fr[HAS_F ? 16 : 1];
    float qr[16];
    {
        // threads j >= R0 load a copy of row j mod R0 (unused): loads under a
        // branch would make the compiler wait for every load before the FFT
        const size_t off = zop_base(ky, kx0 / kTile, V.ntx, N, kTile) + min(kx0 + c, nxh - 1) +
                           (size_t)(act ? j : j % R0) * kTile;
        const QSM_GLOBAL float2* __restrict__ F = glb(V.spec[SPEC_A]) + off;
        const QSM_GLOBAL float* __restrict__ Q = glb(V.q) + off;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            if constexpr (HAS_F) fr[r] = ldg2(F + R0 * r * kTile);
            if constexpr (Q_MEM) qr[r] = Q[R0 * r * kTile];
        }
    }
    tws.commit(twl);
    // the multiplier's per-plane terms, once per block for every column
    __shared__ float4 zqt[Q_MEM ? 1 : N];
    if constexpr (!Q_MEM) {
        for (int kz = threadIdx.x; kz < N; kz += kThreads) zqt[kz] = TvQ::plane(V, kz, N);
    }
    C::template a_fwd<false>(threadIdx.x, v, lds);
    __syncthreads();
    if constexpr (!Q_MEM) {
        // the tile is in LDS and the operands are in flight: form q here
        TvQ tq;
        int fx, fy;
        pk_freq(V, min(kx0 + c, nxh - 1), ky, fx, fy);
        tq.init(V, fx, fy);
        if (act) {
            float4 pz[16];  // all LDS reads issued before the first use
#pragma unroll
            for (int r = 0; r < 16; ++r) pz[r] = zqt[j + R0 * r];
#pragma unroll
            for (int r = 0; r < 16; ++r) qr[r] = tq.at_plane(j + R0 * r, N, pz[r]);
        }
    }
    C::template b_fwd<false>(threadIdx.x, v, lds, twl);
    // doubly packed column 0 (block-uniform dpk): its entries go through LDS
    // so that each element can see Z(-kz); q at (0, ky) is qr, the Nyquist
    // plane's multiplier is formed here
    const bool dcol = dpk && c == 0 && act;
    if (dpk) {
        __syncthreads();  // every read of the first exchange is done
        if (dcol) {
#pragma unroll
            for (int r = 0; r < 16; ++r) lds[j + R0 * r] = v[r];
        }
        __syncthreads();
    }
    if (act) {
        TvQ tqB;
        if constexpr (!Q_MEM) {
            if (dcol) tqB.init(V, V.nx / 2, ky);
        }
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            float2 qs = make_float2(qr[r] * v[r].x, qr[r] * v[r].y);
            if (dcol) {
                const int kz = j + R0 * r;
                float qB;
                if constexpr (Q_MEM) qB = rts_q(V, V.nx / 2, ky, kz);
                else qB = tqB.at_plane(kz, N, zqt[kz]);
                qs = pk_apply(v[r], lds[mirror_idx(kz, N)], qr[r], qB);
            }
            if constexpr (HAS_F)
                v[r] = make_float2(fr[r].x - qs.x, fr[r].y - qs.y);
            else
                v[r] = make_float2(-qs.x, -qs.y);
        }
    }
    __syncthreads();  // every read of the first exchange precedes the second's writes
    C::template lr_to_l16<true>(v, lds, twl);
    l16_store<R0>(v, buf, zs, kx0, nxh);
}

template <int N, int EG, int OP, int SCHED, int TW = kTile>
__global__ __launch_bounds__(kThreads) void k_z_tv(const VolDesc* __restrict__ vols, LaunchTable lt,
                                                   const float2* __restrict__ tw,
                                                   const VolStatus* __restrict__ status) {
    constexpr int EMAX = col_emax(N, EG);
    constexpr bool TWO = SCHED == 0 && qsmfft::col2_ok(N) &&
                         (OP == ZOP_TV_SOLVE || OP == ZOP_TV_SOLVE0 || OP == ZOP_TV_ADJ || OP == ZOP_RTS_SOLVE);
    __shared__ float2 lds[TWO ? col_lds2(N, EG) : col_lds(N, EG)];
    __shared__ float2 twl[tw_lds(N)];
    const int bid = xcd_remap(blockIdx.x, gridDim.x);
    int lb;
    const int gv = locate(lt, bid, lb);
    const VolDesc& V = vols[gv];
    if (status && status[gv].done) return;
    static_assert(TW == kTile || N == 0, "compiled lengths use tiles of kTile columns");
    constexpr int L2 = tile_log2(TW);
    const int ntx = TW == kTile ? V.ntx : V.ntx_long;
    const int ky = lb / ntx;
    const int tx = lb - ky * ntx;
    const int kx0 = tx * TW;
    const int nxh = V.nxh;
    const int nz = N > 0 ? N : V.nz;
    const size_t zs = (size_t)nxh * V.ny;
    const size_t base = (size_t)ky * nxh;
    const int total = nz * TW;
    // f^ / q operands (tile-major): element (kx, kz) at fb + kx + kz * TW
    const size_t fb = zop_base(ky, tx, ntx, nz, TW);
    // column 0 holds two z spectra (the DC and Nyquist planes at ky = 0, ny/2)
    const bool dpk = pk_double(V, tx, ky);
    QSM_GLOBAL float2* buf = glb(V.spec[SPEC_WORK]) + base;
    TwStage<N> tws;
    tws.fetch(tw, V.pz);
    if constexpr (TWO) {
        z_tv_col2<N, OP>(V, lds, twl, tws, ky, kx0, dpk);
        return;
    }
    if constexpr (OP == ZOP_TV_SOLVE0) {
        // g = 0 before the first step (AdmmBuffers::new), so div = 0 and x^ = f^
        col_load<EMAX, TW>(lds, glb(V.spec[SPEC_A]) + fb, TW, nz, kx0, nxh);
        tws.commit(twl);
        __syncthreads();
        run_fft<N, EMAX, true, L2, true>(lds, V.pz, TW, TW, 1, twl, tw);
        col_store<EMAX, TW>(lds, buf, zs, nz, kx0, nxh);
        return;
    }
    col_load<EMAX, TW>(lds, buf, zs, nz, kx0, nxh);
    // x-update operands, in flight during the forward FFT
    constexpr bool HAS_F = OP == ZOP_TV_SOLVE || OP == ZOP_RTS_SOLVE;
    constexpr bool Q_MEM = OP == ZOP_RTS_SOLVE;
    float2 fr[HAS_F ? EMAX : 1];
    float qr[Q_MEM ? EMAX : 1];
    if constexpr (HAS_F || Q_MEM) {
        const QSM_GLOBAL float2* __restrict__ F = glb(V.spec[SPEC_A]) + fb;
        const QSM_GLOBAL float* __restrict__ Q = glb(V.q) + fb;
#pragma unroll
        for (int t = 0; t < EMAX; ++t) {
            const int e = threadIdx.x + t * kThreads;
            const int kz = e >> L2;
            const int kx = kx0 + (e & (TW - 1));
            const bool in = e < total && kx < nxh;
            if constexpr (HAS_F) fr[t] = in ? ldg2(F + kx + kz * TW) : make_float2(0.0f, 0.0f);
            if constexpr (Q_MEM) qr[t] = in ? Q[kx + kz * TW] : 0.0f;
        }
    }
    tws.commit(twl);
    __syncthreads();
    run_fft<N, EMAX, true, L2, false>(lds, V.pz, TW, TW, 1, twl, tw);
    // a thread's elements all lie in column c
    const int c = threadIdx.x & (TW - 1);
    const bool dcol = dpk && c == 0;
    // doubly packed column: Z(-kz) of each element, read before any is rewritten
    float2 zmir[EMAX];
    if (dpk) {
        if (dcol) {
#pragma unroll
            for (int t = 0; t < EMAX; ++t) {
                const int kz = min((int)(threadIdx.x + t * kThreads) >> L2, nz - 1);
                zmir[t] = lds[mirror_idx(kz, nz) * TW];
            }
        }
        __syncthreads();
    }
    if constexpr (OP == ZOP_TV_SETUP || OP == ZOP_RTS_SETUP) {
        QSM_GLOBAL float2* F = glb(V.spec[SPEC_A]) + fb;
        QSM_GLOBAL float* Q = glb(V.q) + fb;
#pragma unroll
        for (int t = 0; t < EMAX; ++t) {
            const int e = threadIdx.x + t * kThreads;
            const int kz = e >> L2;
            const int kx = kx0 + c;
            if (e < total && kx < nxh) {
                const float2 sv = lds[e];
                float2 fv;
                double qv;
                if (dcol) {
                    double mA, mB, qB;
                    setup_mult<OP>(V, 0, ky, kz, mA, qv);
                    setup_mult<OP>(V, V.nx / 2, ky, kz, mB, qB);
                    fv = pk_apply_d(sv, zmir[t], mA, mB);
                } else {
                    int fx, fy;
                    pk_freq(V, kx, ky, fx, fy);
                    double m;
                    setup_mult<OP>(V, fx, fy, kz, m, qv);
                    fv = make_float2((float)((double)sv.x * m), (float)((double)sv.y * m));
                }
                stg2(F + kx + kz * TW, fv);
                Q[kx + kz * TW] = (float)qv;  // (RTS recomputes q in the doubly packed column)
            }
        }
    } else {
        if constexpr (OP == ZOP_TV_APPLYA || OP == ZOP_DIPOLE || OP == ZOP_FILTER) {
#pragma unroll
            for (int t = 0; t < EMAX; ++t) {
                const int e = threadIdx.x + t * kThreads;
                const int kz = e >> L2;
                const int kx = kx0 + c;
                if (e < total) {
                    const float2 sv = lds[e];
                    float2 out = make_float2(0.0f, 0.0f);
                    if (kx < nxh) {
                        if (dcol) {
                            out = pk_apply_d(sv, zmir[t], apply_mult<OP>(V, 0, ky, kz),
                                             apply_mult<OP>(V, V.nx / 2, ky, kz));
                        } else {
                            int fx, fy;
                            pk_freq(V, kx, ky, fx, fy);
                            const double m = apply_mult<OP>(V, fx, fy, kz);
                            out = make_float2((float)((double)sv.x * m), (float)((double)sv.y * m));
                        }
                    }
                    lds[e] = out;
                }
            }
        } else {
            TvQ tq, tqB;
            if constexpr (!Q_MEM) {
                int fx, fy;
                pk_freq(V, min(kx0 + c, nxh - 1), ky, fx, fy);
                tq.init(V, fx, fy);
                if (dcol) tqB.init(V, V.nx / 2, ky);
            }
#pragma unroll
            for (int t = 0; t < EMAX; ++t) {
                const int e = threadIdx.x + t * kThreads;
                if (e < total) {
                    const int kz = e >> L2;
                    const float2 sv = lds[e];
                    float2 qs;  // q * spectrum
                    if (dcol) {
                        float qA, qB;
                        if constexpr (Q_MEM) {
                            qA = rts_q(V, 0, ky, kz);
                            qB = rts_q(V, V.nx / 2, ky, kz);
                        } else {
                            const float2 tz = ztw<N>(V, twl, kz, nz);
                            qA = tq.at(kz, nz, tz);
                            qB = tqB.at(kz, nz, tz);
                        }
                        qs = pk_apply(sv, zmir[t], qA, qB);
                    } else {
                        float q;
                        if constexpr (Q_MEM) q = qr[t];
                        else q = tq.at(kz, nz, ztw<N>(V, twl, kz, nz));
                        qs = make_float2(q * sv.x, q * sv.y);
                    }
                    if constexpr (HAS_F)
                        lds[e] = make_float2(fr[t].x - qs.x, fr[t].y - qs.y);
                    else
                        lds[e] = make_float2(-qs.x, -qs.y);
                }
            }
        }
        __syncthreads();
        run_fft<N, EMAX, true, L2, true>(lds, V.pz, TW, TW, 1, twl, tw);
        col_store<EMAX, TW>(lds, buf, zs, nz, kx0, nxh);
    }
}

// ----------------------------------------------------------- C2R x passes --
// Rebuild the full complex row Z = A + iB from two half spectra (negative
// frequencies by conjugate symmetry; the imaginary parts of the DC and
// Nyquist bins are dropped, as Re(IFFT) of a full spectrum does).
// step 2: A, B = rows 2p, 2p+1 of sa; step 1: A = row p of sa, B = row p of sb.
template <int EMAX>
__device__ __forceinline__ void c2r_load(float2* lds, int nx, int nxh, int RB,
                                         const QSM_GLOBAL float2* __restrict__ sa,
                                         const QSM_GLOBAL float2* __restrict__ sb, int row0, int step,
                                         int nrows) {
    const int total = RB * nx;
    float2 r[EMAX];
#pragma unroll
    for (int t = 0; t < EMAX; ++t) {
        const int e = threadIdx.x + t * kThreads;
        const int p = e / nx;
        const int k = e - p * nx;
        const bool neg = (2 * k > nx);
        const int kk = neg ? nx - k : k;
        // even nx: the Nyquist bin is the imaginary part of column 0
        const bool nyq = 2 * kk == nx;
        // rows past the volume (or past the block for e >= total) read its
        // last row: in bounds, and their outputs are never stored (each
        // half of the packed row is Hermitian, so they do not mix)
        const int ra = min(row0 + step * p, nrows - 1);
        const int rb = min((step == 2) ? row0 + step * p + 1 : row0 + step * p, nrows - 1);
        const int col = nyq ? 0 : kk;
        float2 A = ldg2(sa + col + (size_t)ra * nxh);
        float2 B = ldg2(sb + col + (size_t)rb * nxh);
        if (nyq) { A.x = A.y; B.x = B.y; }
        if (kk == 0 || nyq) { A.y = 0.0f; B.y = 0.0f; }
        if (neg) { A.y = -A.y; B.y = -B.y; }
        r[t] = make_float2(A.x - B.y, A.y + B.x);
    }
#pragma unroll
    for (int t = 0; t < EMAX; ++t) {
        const int e = threadIdx.x + t * kThreads;
        if (e < total) lds[e] = r[t];
    }
}

// c2r_load for compiled lengths divisible by 4: a thread takes two
// consecutive half-spectrum entries k0, k0 + 1 of a row pair (16-byte loads,
// every entry read once) and writes the full-spectrum elements they give:
// k and N - k for 0 < k < N / 2, and for k = 0 the DC element and the
// Nyquist element N / 2 (the imaginary parts of the packed column 0).
template <int N, int RB>
__device__ __forceinline__ void c2r_load_pairs(float2* lds, const QSM_GLOBAL float2* __restrict__ sa,
                                               const QSM_GLOBAL float2* __restrict__ sb, int row0, int step,
                                               int nrows) {
    static_assert(N % 4 == 0, "pairs of half-spectrum entries");
    constexpr int NH = N / 2, NH2 = N / 4, TOT2 = RB * NH2;
    constexpr int T = (TOT2 + kThreads - 1) / kThreads;
    float4 a[T], b[T];
#pragma unroll
    for (int t = 0; t < T; ++t) {
        const int e = threadIdx.x + t * kThreads;
        const int ec = e < TOT2 ? e : 0;  // idle threads load a valid pair (unused)
        const int p = ec / NH2;
        const int k0 = 2 * (ec - p * NH2);
        // rows past the volume read its last row: in bounds, never stored
        const int ra = min(row0 + step * p, nrows - 1);
        const int rb = min((step == 2) ? row0 + step * p + 1 : row0 + step * p, nrows - 1);
        a[t] = ldv<4>((const QSM_GLOBAL float*)(sa + k0 + (size_t)ra * NH));
        b[t] = ldv<4>((const QSM_GLOBAL float*)(sb + k0 + (size_t)rb * NH));
    }
#pragma unroll
    for (int t = 0; t < T; ++t) {
        const int e = threadIdx.x + t * kThreads;
        if (e < TOT2) {
            const int p = e / NH2;
            const int k0 = 2 * (e - p * NH2);
            float2* row = lds + p * N;
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                const int k = k0 + c;
                const float2 A = c == 0 ? make_float2(a[t].x, a[t].y) : make_float2(a[t].z, a[t].w);
                const float2 B = c == 0 ? make_float2(b[t].x, b[t].y) : make_float2(b[t].z, b[t].w);
                if (k == 0) {
                    row[0] = make_float2(A.x, B.x);
                    row[NH] = make_float2(A.y, B.y);
                } else {
                    row[k] = make_float2(A.x - B.y, A.y + B.x);
                    row[N - k] = make_float2(A.x + B.y, B.x - A.y);
                }
            }
        }
    }
}

// c2r_load_pairs where it applies, else c2r_load.
template <int N, int RBC, int EMAX>
__device__ __forceinline__ void c2r_load_rows(float2* lds, int nx, int nxh, int RB,
                                              const QSM_GLOBAL float2* __restrict__ sa,
                                              const QSM_GLOBAL float2* __restrict__ sb, int row0, int step,
                                              int nrows) {
    if constexpr (N > 0 && N % 4 == 0) c2r_load_pairs<N, RBC>(lds, sa, sb, row0, step, nrows);
    else c2r_load<EMAX>(lds, nx, nxh, RB, sa, sb, row0, step, nrows);
}

// TV: x_it = Re IFFT(x^) (src/inversion/admm.rs:111-114) and the two sums of
// relative_change (src/utils/mod.rs:69-77) accumulated in FP64 per block;
// the next pass's first block (or k_tv_reduce after the last iteration)
// reduces them.
// MODE C2R_TV: as above.  C2R_ADJ: Re IFFT to adj_c (c-bar), nothing
// else.  C2R_F64: Re IFFT widened to f64 into adj_out.
// C2R_F64_MASKED: f64 into out, 0 where mask == 0 (src/utils/mask.rs:47-53).
// C2R_X0: FP32 into x[0], 0 where mask == 0 (the RTS starting point).
enum C2RMode : int { C2R_TV = 0, C2R_ADJ = 1, C2R_F64 = 2, C2R_F64_MASKED = 3, C2R_X0 = 4 };

template <int N, int EG, int MODE, int RBS = 0>
__global__ __launch_bounds__(kThreads) void k_c2r_tv(const VolDesc* __restrict__ vols, LaunchTable lt,
                                                     const float2* __restrict__ tw, int parity,
                                                     const VolStatus* __restrict__ status) {
    // half the R2C pass's pair-rows (after its row shift RBS), at least 1
    constexpr int RBF = N > 0 ? ((qsmfft::fixed_rows(N) >> RBS) / 2 > 0 ? (qsmfft::fixed_rows(N) >> RBS) / 2 : 1) : 0;
    constexpr int EMAX = N > 0 ? (RBF * N + kThreads - 1) / kThreads : kRowEmaxRt;
    __shared__ float2 lds[N > 0 ? RBF * N : x_lds(0)];
    __shared__ float2 twl[tw_lds(N)];
    __shared__ double sh[kThreads / 64];
    const int bid = xcd_remap(blockIdx.x, gridDim.x);
    int lb;
    const int gv = locate(lt, bid, lb);
    const VolDesc& V = vols[gv];
    if (status && status[gv].done) return;
    const int nx = N > 0 ? N : V.nx;
    const int nxh = (nx + 1) / 2;
    const int RB = N > 0 ? RBF : V.rows_c;
    const int nrows = V.ny * V.nz;
    const int row0 = lb * RB * 2;
    const int total = RB * nx;
    const QSM_GLOBAL float2* s = glb(V.spec[SPEC_WORK]);
    TwStage<N> tws;
    tws.fetch(tw, V.px);
    c2r_load_rows<N, RBF, EMAX>(lds, nx, nxh, RB, s, s, row0, 2, nrows);
    if constexpr (MODE != C2R_TV) {
        tws.commit(twl);
        __syncthreads();
        run_fft<N, EMAX, false, 0, true, RBF>(lds, V.px, RB, 1, nx, twl, tw);
#pragma unroll
        for (int t = 0; t < EMAX; ++t) {
            const int e = threadIdx.x + t * kThreads;
            if (e < total) {
                const int p = e / nx;
                const int i = e - p * nx;
                const int ra = row0 + 2 * p, rb = ra + 1;
                const float2 z = lds[e];
                if constexpr (MODE == C2R_ADJ) {
                    QSM_GLOBAL float* __restrict__ o = glb(V.adj_c);
                    if (ra < nrows) o[i + (size_t)ra * nx] = z.x;
                    if (rb < nrows) o[i + (size_t)rb * nx] = z.y;
                } else if constexpr (MODE == C2R_F64_MASKED || MODE == C2R_X0) {
                    const QSM_GLOBAL uint8_t* __restrict__ m = glb(V.mask);
                    const size_t ia = i + (size_t)ra * nx, ib = i + (size_t)rb * nx;
                    if constexpr (MODE == C2R_F64_MASKED) {
                        QSM_GLOBAL double* __restrict__ o = glb(V.out);
                        if (ra < nrows) o[ia] = m[ia] ? (double)z.x : 0.0;
                        if (rb < nrows) o[ib] = m[ib] ? (double)z.y : 0.0;
                    } else {
                        QSM_GLOBAL float* __restrict__ o = glb(V.x[0]);
                        if (ra < nrows) o[ia] = m[ia] ? z.x : 0.0f;
                        if (rb < nrows) o[ib] = m[ib] ? z.y : 0.0f;
                    }
                } else {
                    QSM_GLOBAL double* __restrict__ o = glb(V.adj_out);
                    if (ra < nrows) o[i + (size_t)ra * nx] = (double)z.x;
                    if (rb < nrows) o[i + (size_t)rb * nx] = (double)z.y;
                }
            }
        }
        return;
    }
    const QSM_GLOBAL float* __restrict__ xo = glb(V.x[parity]);
    QSM_GLOBAL float* __restrict__ xn = glb(V.x[parity ^ 1]);
    if constexpr (N > 0 && N % 4 == 0) {
        // four consecutive columns per thread: 16-byte loads of the previous
        // iterate and stores of the new one
        constexpr int E4 = (RBF * N / 4 + kThreads - 1) / kThreads;
        float4 xa[E4], xb[E4];  // previous iterate, in flight during the inverse FFT
#pragma unroll
        for (int t = 0; t < E4; ++t) {
            const int e = 4 * (threadIdx.x + t * kThreads);
            const int p = e / N;
            const int i = e - p * N;
            // rows past the volume read its last row: always in bounds, never used
            const int ra = min(row0 + 2 * p, nrows - 1), rb = min(row0 + 2 * p + 1, nrows - 1);
            xa[t] = ldv<4>(xo + i + (size_t)ra * N);
            xb[t] = ldv<4>(xo + i + (size_t)rb * N);
        }
        tws.commit(twl);
        __syncthreads();
        run_fft<N, EMAX, false, 0, true, RBF>(lds, V.px, RB, 1, nx, twl, tw);
        double sd = 0.0, sx = 0.0;
#pragma unroll
        for (int t = 0; t < E4; ++t) {
            const int e = 4 * (threadIdx.x + t * kThreads);
            if (e < total) {
                const int p = e / N;
                const int i = e - p * N;
                const int ra = row0 + 2 * p, rb = ra + 1;
                float za[4], zb[4];
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const float2 z = lds[e + c];
                    za[c] = z.x;
                    zb[c] = z.y;
                }
                if (ra < nrows) {
                    stv<4>(xn + i + (size_t)ra * N, vpack<4>(za));
#pragma unroll
                    for (int c = 0; c < 4; ++c) {
                        const double d = (double)za[c] - (double)vel<4>(xa[t], c);
                        sd += d * d;
                        sx += (double)za[c] * (double)za[c];
                    }
                }
                if (rb < nrows) {
                    stv<4>(xn + i + (size_t)rb * N, vpack<4>(zb));
#pragma unroll
                    for (int c = 0; c < 4; ++c) {
                        const double d = (double)zb[c] - (double)vel<4>(xb[t], c);
                        sd += d * d;
                        sx += (double)zb[c] * (double)zb[c];
                    }
                }
            }
        }
        const double tsd = block_sum(sd, sh);
        const double tsx = block_sum(sx, sh);
        if (threadIdx.x == 0) {
            QSM_GLOBAL double* part = glb(V.partial);
            part[2 * lb] = tsd;
            part[2 * lb + 1] = tsx;
        }
        return;
    }
    float2 xp[EMAX];  // previous iterate, in flight during the inverse FFT
#pragma unroll
    for (int t = 0; t < EMAX; ++t) {
        const int e = threadIdx.x + t * kThreads;
        const int p = e / nx;
        const int i = e - p * nx;
        // rows past the volume read its last row: always in bounds, never used
        const int ra = min(row0 + 2 * p, nrows - 1), rb = min(row0 + 2 * p + 1, nrows - 1);
        xp[t].x = xo[i + (size_t)ra * nx];
        xp[t].y = xo[i + (size_t)rb * nx];
    }
    tws.commit(twl);
    __syncthreads();
    run_fft<N, EMAX, false, 0, true, RBF>(lds, V.px, RB, 1, nx, twl, tw);
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
        QSM_GLOBAL double* part = glb(V.partial);
        part[2 * lb] = tsd;
        part[2 * lb + 1] = tsx;
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
template <int N, int EG, int OP, int TW = kTile>
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
    static_assert(TW == kTile || N == 0, "compiled lengths use tiles of kTile columns");
    constexpr int L2 = tile_log2(TW);
    const int ntx = TW == kTile ? V.ntx : V.ntx_long;
    const int ky = lb / ntx;
    const int tx = lb - ky * ntx;
    const int kx0 = tx * TW;
    const int nxh = V.nxh;
    const int nz = N > 0 ? N : V.nz;
    const size_t zs = (size_t)nxh * V.ny;
    const size_t base = (size_t)ky * nxh;
    const int total = nz * TW;
    QSM_GLOBAL float2* buf = glb(V.spec[OP == ZOP_VS_FWD ? sel : SPEC_WORK]) + base;
    TwStage<N> tws;
    tws.fetch(tw, V.pz);
    col_load<EMAX, TW>(lds, buf, zs, nz, kx0, nxh);
    tws.commit(twl);
    __syncthreads();
    run_fft<N, EMAX, true, L2, false>(lds, V.pz, TW, TW, 1, twl, tw);
    if constexpr (OP == ZOP_VS_FWD) {
        col_store<EMAX, TW>(lds, buf, zs, nz, kx0, nxh);
    } else {
        // column 0 of tile 0 at ky = 0, ny/2 holds two z spectra (the DC and
        // Nyquist planes, even nx); the sphere's spectrum there is SA + i SB,
        // both real, and the S^ of the Nyquist plane is kept after the half
        // spectrum in vs_sfirst (2 nz entries)
        const bool dpk = pk_double(V, tx, ky);
        const int c = threadIdx.x & (TW - 1);
        const bool dcol = dpk && c == 0;
        QSM_GLOBAL float* sfirst = glb(V.vs_sfirst) + base;
        QSM_GLOBAL float* sfirst_nyq =
            glb(V.vs_sfirst) + (size_t)nxh * V.ny * V.nz + (ky == 0 ? 0 : (size_t)nz);
        if constexpr (OP == ZOP_VS_RADIUS) {
            const float inv_count = (float)(1.0 / V.vs_counts[ridx]);
            const float invN = (float)(1.0 / ((double)V.nx * V.ny * V.nz));
            const QSM_GLOBAL float2* __restrict__ FH = glb(V.spec[SPEC_A]) + base;
            const QSM_GLOBAL float2* __restrict__ MH = glb(V.spec[SPEC_B]) + base;
            float2 qv[EMAX];  // F^ (1 - S^) / N, transformed after P
#pragma unroll
            for (int t = 0; t < EMAX; ++t) {
                const int e = threadIdx.x + t * kThreads;
                if (e < total) {
                    const int kz = e >> L2;
                    const int kx = kx0 + c;
                    const float2 z = lds[e];
                    if (kx < nxh) {
                        const float2 m = ldg2(MH + kx + kz * zs);
                        const float2 fh = ldg2(FH + kx + kz * zs);
                        if (dcol) {
                            const int mz = mirror_idx(kz, nz);
                            const float sA = z.x * inv_count, sB = z.y * inv_count;
                            if (flag) {
                                sfirst[kx + kz * zs] = sA;
                                sfirst_nyq[kz] = sB;
                            }
                            const float2 mm = ldg2(MH + kx + mz * zs), fm = ldg2(FH + kx + mz * zs);
                            lds[e] = pk_apply(m, mm, sA * invN, sB * invN);
                            qv[t] = pk_apply(fh, fm, (1.0f - sA) * invN, (1.0f - sB) * invN);
                        } else {
                            const float sv = z.x * inv_count;
                            if (flag) sfirst[kx + kz * zs] = sv;
                            lds[e] = make_float2(m.x * sv * invN, m.y * sv * invN);
                            const float hp = (1.0f - sv) * invN;
                            qv[t] = make_float2(fh.x * hp, fh.y * hp);
                        }
                    } else {
                        lds[e] = make_float2(0.0f, 0.0f);
                        qv[t] = make_float2(0.0f, 0.0f);
                    }
                }
            }
            __syncthreads();
            run_fft<N, EMAX, true, L2, true>(lds, V.pz, TW, TW, 1, twl, tw);
            col_store<EMAX, TW>(lds, buf, zs, nz, kx0, nxh);
            __syncthreads();
#pragma unroll
            for (int t = 0; t < EMAX; ++t) {
                const int e = threadIdx.x + t * kThreads;
                if (e < total) lds[e] = qv[t];
            }
            __syncthreads();
            run_fft<N, EMAX, true, L2, true>(lds, V.pz, TW, TW, 1, twl, tw);
            col_store<EMAX, TW>(lds, glb(V.spec[SPEC_C]) + base, zs, nz, kx0, nxh);
        } else {  // ZOP_VS_DECONV
            const double invN = 1.0 / ((double)V.nx * V.ny * V.nz);
            // X / (1 - S^), 0 where |1 - S^| < threshold
            auto deconv = [&](double oms) {
                if (fabs(oms) < V.vs_threshold) return 0.0;
                return flag ? invN / oms : (1.0 / oms) * invN;
            };
            float2 zmir[EMAX];
            if (dpk) {
                if (dcol) {
#pragma unroll
                    for (int t = 0; t < EMAX; ++t) {
                        const int kz = min((int)(threadIdx.x + t * kThreads) >> L2, nz - 1);
                        zmir[t] = lds[mirror_idx(kz, nz) * TW];
                    }
                }
                __syncthreads();
            }
#pragma unroll
            for (int t = 0; t < EMAX; ++t) {
                const int e = threadIdx.x + t * kThreads;
                const int kz = e >> L2;
                const int kx = kx0 + c;
                if (e < total && kx < nxh) {
                    const float2 xv = lds[e];
                    const double oms = 1.0 - (double)sfirst[kx + kz * zs];
                    float2 y;
                    if (dcol) {
                        y = pk_apply_d(xv, zmir[t], deconv(oms), deconv(1.0 - (double)sfirst_nyq[kz]));
                    } else {
                        const double sc = deconv(oms);
                        y = make_float2((float)(xv.x * sc), (float)(xv.y * sc));
                    }
                    lds[e] = y;
                }
            }
            __syncthreads();
            run_fft<N, EMAX, true, L2, true>(lds, V.pz, TW, TW, 1, twl, tw);
            col_store<EMAX, TW>(lds, buf, zs, nz, kx0, nxh);
        }
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
    const int nxh = (nx + 1) / 2;
    const int RB = N > 0 ? qsmfft::fixed_rows(N) : V.rows_x;
    const int nrows = V.ny * V.nz;
    const int row0 = lb * RB;
    TwStage<N> tws;
    tws.fetch(tw, V.px);
    c2r_load_rows<N, (N > 0 ? qsmfft::fixed_rows(N) : 0), EMAX>(lds, nx, nxh, RB, glb(V.spec[SPEC_WORK]),
                                                            glb(V.spec[SPEC_C]), row0, 1, nrows);
    tws.commit(twl);
    QSM_GLOBAL uint8_t* fin = glb(V.vs_final);
    QSM_GLOBAL float* loc = glb(V.vs_local);
    __syncthreads();
    run_fft<N, EMAX, false, 0, true>(lds, V.px, RB, 1, nx, twl, tw);
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
            if ((double)k * norm > delta && !fin[idx]) {
                loc[idx] = z.y;
                fin[idx] = 1;
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
    const int nxh = (nx + 1) / 2;
    const int RB = N > 0 ? qsmfft::fixed_rows(N) : V.rows_x;
    const int nrows = V.ny * V.nz;
    const int row0 = lb * RB * 2;
    const QSM_GLOBAL float2* s = glb(V.spec[SPEC_WORK]);
    TwStage<N> tws;
    tws.fetch(tw, V.px);
    c2r_load_rows<N, (N > 0 ? qsmfft::fixed_rows(N) : 0), EMAX>(lds, nx, nxh, RB, s, s, row0, 2, nrows);
    tws.commit(twl);
    __syncthreads();
    run_fft<N, EMAX, false, 0, true>(lds, V.px, RB, 1, nx, twl, tw);
    const int total = RB * nx;
    const QSM_GLOBAL uint8_t* fin = glb(V.vs_final);
    QSM_GLOBAL double* out = glb(V.out);
    QSM_GLOBAL uint8_t* mout = glb(V.mask_out);
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
                const uint8_t fm = fin[idx];
                out[idx] = fm == 1 ? (double)z.x : 0.0;
                mout[idx] = fm;
            }
            if (rb < nrows) {
                const size_t idx = i + (size_t)rb * nx;
                const uint8_t fm = fin[idx];
                out[idx] = fm == 1 ? (double)z.y : 0.0;
                mout[idx] = fm;
            }
        }
    }
}

}  // namespace qsm
