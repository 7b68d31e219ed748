// Pass kernels shared by the TV-ADMM and V-SHARP translation units: the
// R2C x pass (with its input sources), the C2C y pass, the C2R row loader
// and an FP64 block reduction.
#pragma once

#include "qsm_internal.hpp"

namespace qsm {

using qsmfft::kThreads;
using qsmfft::fastdiv;

enum SrcKind : int { SRC_F64 = 0, SRC_U8 = 1, SRC_F32 = 2, SRC_SPHERE = 3, SRC_TVDIV = 4 };

__device__ inline float shrinkf(float x, float t) {
    // src/utils/mod.rs:44-52
    return x > t ? x - t : (x < -t ? x + t : 0.0f);
}

// ---------------------------------------------------------------- sources --
struct TvIter {
    const float* x;   // x_{it-1}
    const float* uin; // u_{it-2}
    float* uout;      // u_{it-1}
};

// One element of the divergence of g = 2z - v for the z/u update fused in
// front of the x-subproblem.  v = grad(x) + u (src/utils/gradient.rs:21-58,
// periodic), z = shrink(v, lambda/rho), u' = v - z, g = 2z - v
// (src/inversion/admm.rs:127-142); div = bdiv(g) (src/utils/gradient.rs:71-108).
// The x component of g at i-1 is recomputed from x and u at i-1 (likewise y,
// z), so g itself never goes to memory.
__device__ inline float tv_div_element(const VolDesc& V, const TvIter& T, int i, int r) {
    const int nx = V.nx, ny = V.ny, nz = V.nz;
    const int k = (int)fastdiv((uint32_t)r, V.m_ny);
    const int j = r - k * ny;
    const int nxy = nx * ny;
    const int N = nxy * nz;
    const int idx = i + r * nx;
    const int dxp = (i + 1 < nx) ? 1 : 1 - nx;
    const int dxm = (i > 0) ? -1 : nx - 1;
    const int dyp = (j + 1 < ny) ? nx : nx - nxy;
    const int dym = (j > 0) ? -nx : nxy - nx;
    const int dzp = (k + 1 < nz) ? nxy : nxy - N;
    const int dzm = (k > 0) ? -nxy : N - nxy;
    const float* x = T.x;
    const float* u = T.uin;
    const float xc = x[idx];
    const float t = V.thr;
    const float hx = V.hx, hy = V.hy, hz = V.hz;

    const float vx = (x[idx + dxp] - xc) * hx + u[idx];
    const float vy = (x[idx + dyp] - xc) * hy + u[N + idx];
    const float vz = (x[idx + dzp] - xc) * hz + u[2 * N + idx];
    const float vxm = (xc - x[idx + dxm]) * hx + u[idx + dxm];
    const float vym = (xc - x[idx + dym]) * hy + u[N + idx + dym];
    const float vzm = (xc - x[idx + dzm]) * hz + u[2 * N + idx + dzm];

    const float zx = shrinkf(vx, t), zy = shrinkf(vy, t), zz = shrinkf(vz, t);
    T.uout[idx] = vx - zx;
    T.uout[N + idx] = vy - zy;
    T.uout[2 * N + idx] = vz - zz;
    const float gx = 2.0f * zx - vx, gy = 2.0f * zy - vy, gz = 2.0f * zz - vz;
    const float gxm = 2.0f * shrinkf(vxm, t) - vxm;
    const float gym = 2.0f * shrinkf(vym, t) - vym;
    const float gzm = 2.0f * shrinkf(vzm, t) - vzm;
    return (gx - gxm) * hx + (gy - gym) * hy + (gz - gzm) * hz;
}

// Binary sphere of src/kernels/smv.rs:19-67 (un-normalised: 1 inside).  The
// squared distance is formed with explicit round-to-nearest operations, in the
// reference's order and without fused multiply-adds, so boundary voxels are
// classified exactly as the reference (and the host voxel count) do.
__device__ inline float sphere_element(const VolDesc& V, int i, int r, int ridx) {
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

template <int SRC>
__device__ inline float load_src(const VolDesc& V, const TvIter& T, int i, int r, int aux) {
    if (SRC == SRC_F64) return (float)V.field[i + (size_t)r * V.nx];
    if (SRC == SRC_U8) return (float)V.mask[i + (size_t)r * V.nx];
    if (SRC == SRC_F32) return V.vs_local[i + (size_t)r * V.nx];
    if (SRC == SRC_SPHERE) return sphere_element(V, i, r, aux);
    return tv_div_element(V, T, i, r);
}

// ------------------------------------------------------------- R2C x pass --
// Block = rows_x pairs of real rows.  Pair p packs rows (2p, 2p+1) as
// z = a + i b; after a complex FFT of length nx the half spectra are
//   A[k] = (Z[k] + conj Z[-k]) / 2,   B[k] = (Z[k] - conj Z[-k]) / (2i).
template <int EMAX, int SRC>
__global__ __launch_bounds__(kThreads) void k_r2c_x(const VolDesc* __restrict__ vols, LaunchTable lt,
                                                    const float2* __restrict__ tw, int out_sel,
                                                    int parity, const VolStatus* __restrict__ status) {
    // parity: TV iteration parity for SRC_TVDIV, radius index for SRC_SPHERE
    __shared__ float2 lds[kThreads * EMAX];
    const int bid = xcd_remap(blockIdx.x, gridDim.x);
    const int v = find_vol(lt, bid);
    const int lb = bid - lt.start[v];
    const VolDesc& V = vols[v];
    if (status && status[v].done) return;
    const int nx = V.nx, nxh = V.nxh, RB = V.rows_x;
    const int nrows = V.ny * V.nz;
    const int row0 = lb * RB * 2;
    TvIter T;
    T.x = V.x[parity];
    T.uin = V.u[parity ^ 1];
    T.uout = V.u[parity];
    for (int e = threadIdx.x; e < RB * nx; e += kThreads) {
        const int p = (int)fastdiv((uint32_t)e, V.m_nx);
        const int i = e - p * nx;
        const int ra = row0 + 2 * p, rb = ra + 1;
        const float a = (ra < nrows) ? load_src<SRC>(V, T, i, ra, parity) : 0.0f;
        const float b = (rb < nrows) ? load_src<SRC>(V, T, i, rb, parity) : 0.0f;
        lds[e] = make_float2(a, b);
    }
    __syncthreads();
    qsmfft::lds_fft<EMAX, false, 0, false>(lds, V.px, RB, 1, nx, tw + V.px.tw_off);
    float2* out = V.spec[out_sel];
    for (int e = threadIdx.x; e < RB * nxh; e += kThreads) {
        const int p = (int)fastdiv((uint32_t)e, V.m_nxh);
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

// ------------------------------------------------------------- C2C y pass --
// Block = one tile of kTile consecutive kx columns at one kz, full ny.
// sel2 >= 0 transforms a second spectrum with a second set of blocks.
template <int EMAX, bool INV>
__global__ __launch_bounds__(kThreads) void k_y_pass(const VolDesc* __restrict__ vols, LaunchTable lt,
                                                     const float2* __restrict__ tw, int sel, int sel2,
                                                     const VolStatus* __restrict__ status) {
    __shared__ float2 lds[kThreads * EMAX];
    const int bid = xcd_remap(blockIdx.x, gridDim.x);
    const int v = find_vol(lt, bid);
    int lb = bid - lt.start[v];
    const VolDesc& V = vols[v];
    if (status && status[v].done) return;
    const int per = V.ntx * V.nz;
    int s = sel;
    if (lb >= per) { lb -= per; s = sel2; }
    const int kz = lb / V.ntx;
    const int tx = lb - kz * V.ntx;
    const int kx0 = tx * kTile;
    const int nxh = V.nxh, ny = V.ny;
    float2* buf = V.spec[s] + (size_t)kz * nxh * ny;
    for (int e = threadIdx.x; e < ny * kTile; e += kThreads) {
        const int ky = e >> kLog2Tile, c = e & (kTile - 1);
        const int kx = kx0 + c;
        lds[e] = (kx < nxh) ? buf[kx + ky * nxh] : make_float2(0.0f, 0.0f);
    }
    __syncthreads();
    qsmfft::lds_fft<EMAX, true, kLog2Tile, INV>(lds, V.py, kTile, kTile, 1, tw + V.py.tw_off);
    for (int e = threadIdx.x; e < ny * kTile; e += kThreads) {
        const int ky = e >> kLog2Tile, c = e & (kTile - 1);
        const int kx = kx0 + c;
        if (kx < nxh) buf[kx + ky * nxh] = lds[e];
    }
}

// ----------------------------------------------------------- C2R x passes --
// Rebuild the full complex row Z = A + iB from the two half spectra
// (negative frequencies by conjugate symmetry; the imaginary parts of the DC
// and Nyquist bins are dropped, as Re(IFFT) of a full spectrum does).
template <int EMAX>
__device__ inline void c2r_load(float2* lds, const VolDesc& V, const float2* sa, const float2* sb,
                                int row0, int step, int nrows) {
    const int nx = V.nx, nxh = V.nxh, RB = V.rows_x;
    for (int e = threadIdx.x; e < RB * nx; e += kThreads) {
        const int p = (int)fastdiv((uint32_t)e, V.m_nx);
        const int k = e - p * nx;
        const bool neg = (2 * k > nx);
        const int kk = neg ? nx - k : k;
        const int ra = row0 + step * p;
        const int rb = (step == 2) ? ra + 1 : ra;
        float2 A = (ra < nrows) ? sa[kk + (size_t)ra * nxh] : make_float2(0.0f, 0.0f);
        float2 B = (rb < nrows && (step == 2 ? true : sb != nullptr))
                       ? sb[kk + (size_t)rb * nxh] : make_float2(0.0f, 0.0f);
        if (kk == 0 || 2 * kk == nx) { A.y = 0.0f; B.y = 0.0f; }
        if (neg) { A.y = -A.y; B.y = -B.y; }
        lds[e] = make_float2(A.x - B.y, A.y + B.x);
    }
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


__device__ inline double fftfreq_at(int i, int n, double d) {
    // src/fft.rs:544-566: index i of numpy fftfreq(n, d)
    const double val = 1.0 / ((double)n * d);
    const int s = (i < (n + 1) / 2) ? i : i - n;
    return (double)s * val;
}

__device__ inline int mirror_idx(int i, int n) { return i == 0 ? 0 : n - i; }

}  // namespace qsm
