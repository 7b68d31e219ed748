// Shared device/host structures of the gfx950 QSM library.
//
// Data layout in HBM (per volume, FP32):
//   real volumes   x, field, ...  : nx*ny*nz floats, Fortran order
//                                   (idx = i + j*nx + k*nx*ny)
//   dual variable  u              : 3 planes of N floats (ux | uy | uz)
//   half spectrum  spec, F, ...   : nxh*ny*nz complex (float2), nxh = (nx+1)/2
//                                   (even nx: DC and Nyquist share column 0),
//                                   idx = kx + ky*nxh + kz*nxh*ny
// Every kernel is launched once for a whole batch of volumes; a LaunchTable
// maps the flat block index to (volume, block-within-volume).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fft_core.hpp"

namespace qsm {

constexpr int kMaxVolumes = 64;
constexpr int kTile = 16;      // columns per y/z-pass tile (16 * 8 B = 128 B rows)
constexpr int kLog2Tile = 4;
// Runtime-planned column lengths above 256 run in tiles of kTileLong columns
// (two LDS buffers of kTileLong * 1024 elements fit a workgroup); compiled
// lengths, at most kWideTileMaxLen, in tiles of kTile.
constexpr int kTileLong = 8;
constexpr int kWideTileMaxLen = 512;
constexpr int tile_log2(int tw) { return tw == 16 ? 4 : (tw == 8 ? 3 : -1); }
// Added to a runtime-plan column kernel's elements per thread (EG) when it
// runs in tiles of kTileLong columns.
constexpr int kColLongTile = 256;
// Elements per thread of the runtime-plan row kernels (tiles of
// kThreads * kRowEmaxRt elements; the compiled ones take 16).
constexpr int kRowEmaxRt = 8;

struct LaunchTable {
    int nvol;
    int start[kMaxVolumes + 1];  // start[v] = first block of entry v; start[nvol] = total
    int vol[kMaxVolumes];        // global volume index of entry v
};

// Convergence state of one TV-ADMM volume, written by the reduce kernel.
struct VolStatus {
    int done;      // 1 once relative_change < tol
    int iters;     // iterations executed
    double rel;    // last relative change
};

// Spectral buffers a kernel can address by index.
enum SpecSel : int { SPEC_WORK = 0, SPEC_A = 1, SPEC_B = 2, SPEC_C = 3, SPEC_COUNT = 4 };

struct VolDesc {
    int nx, ny, nz, nxh;
    int rows_x;                  // pair-rows per block in the x passes
    int rows_c;                  // pair-rows per block in the TV C2R pass
    int ntx;                     // y/z tiles along kx: ceil(nxh / kTile)
    int ntx_long;                // the same in tiles of kTileLong columns
    uint64_t m_nx, m_nxh, m_ny;  // magic divisors
    qsmfft::Plan px, py, pz;
    float hx, hy, hz;            // 1/voxel size (gradient / divergence)
    float thr;                   // lambda / rho
    double vsx, vsy, vsz;
    double bx, by, bz;           // normalised B0 direction
    double rho, tol;
    // TV-ADMM state
    float* x[2];                 // ping-pong solution
    float* u[2];                 // ping-pong dual variable (3N each)
    float2* spec[SPEC_COUNT];    // half spectra
    float* q;                    // rho * inv_a / N (half spectrum, real; read by RTS, TV forms it in-kernel)
    double* partial;             // per-block (sum d^2, sum x^2) of the C2R pass
    int npartial;
    // inputs / outputs
    const double* field;         // f64 input (device)
    const uint8_t* mask;         // u8 mask (device)
    double* out;                 // f64 output (device)
    uint8_t* mask_out;           // u8 mask output (device, V-SHARP)
    // V-SHARP state
    float* vs_local;             // selected high-passed field (N)
    float* vs_sfirst;            // Re FFT of the largest normalised sphere (half spectrum)
    uint8_t* vs_final;           // final (eroded) mask (N)
    const double* vs_radii;      // radii in processing order (device)
    const double* vs_counts;     // voxels inside each sphere (device)
    double vs_threshold;
    // unrolled TV-ADMM adjoint (vector-Jacobian product) state
    int8_t* ckpt;                // shrink pattern q_k in {-1,0,1} of iterations 1..K-1 (3N each)
    float* adj_c;                // c-bar of the last adjoint step (N)
    float* adj_s[2];             // ping-pong s-bar (3N each)
    float* adj_acc;              // sum over iterations of x-bar (N)
    double* adj_tpart;           // per (iteration, block) FP64 partials of t-bar
    int adj_nblk;                // x-pass blocks per iteration (stride of adj_tpart)
    const double* adj_chibar;    // dL/dchi (device, f64)
    double* adj_out;             // f64 output of the C2R F64 mode (field-bar, dipole output)
    double* adj_lbar;            // dL/dlambda (device, one double)
    // closed-form inversions and RTS
    int filt;                    // Filter kind (ZOP_FILTER)
    int lsmr_iter;               // RTS step-1 refinement rounds
    double filt_a;               // threshold (TKD/TSVD), lambda (Tikhonov), delta (RTS)
    double filt_b;               // mu (RTS, NLTV)
    // NLTV shrink weights (ping-pong) and their normalisations
    float* nltv_w[2];
    double* nltv_wnorm;          // 2 doubles (device)
    double* nltv_wpart;          // per-block maxima of the weight pass
};

__device__ __host__ inline int find_vol(const LaunchTable& lt, int bid) {
    int v = 0;
    while (v + 1 < lt.nvol && bid >= lt.start[v + 1]) ++v;
    return v;
}

// Global volume index of flat block `bid`; lb receives the block index
// within that volume.
__device__ __host__ inline int locate(const LaunchTable& lt, int bid, int& lb) {
    const int v = find_vol(lt, bid);
    lb = bid - lt.start[v];
    return lt.vol[v];
}

// Flat block index -> logical block, so that the blocks one XCD runs
// (hardware dispatches consecutive block ids round-robin over 8 XCDs) are a
// contiguous range of logical blocks sharing that XCD's L2.
__device__ __host__ inline int xcd_remap(int bid, int total) {
    const int q = total / 8, r = total % 8;
    const int x = bid % 8, t = bid / 8;
    return x * q + (x < r ? x : r) + t;
}

// Everything one pass launch needs; kind selects the kernel.
enum PassKind : int {
    P_R2C_TVDIV = 0, P_R2C_F64, P_R2C_U8, P_R2C_F32, P_R2C_SPHERE,
    P_Y_FWD, P_Y_INV,
    P_Z_TV_SETUP, P_Z_TV_SOLVE, P_Z_VS_FWD, P_Z_VS_RADIUS, P_Z_VS_DECONV,
    P_C2R_TV, P_C2R_VS_SELECT, P_C2R_VS_OUT,
    // unrolled TV-ADMM adjoint and the dipole operator
    P_R2C_TVDIV_REC, P_R2C_TVADJ_FIRST, P_R2C_TVADJ, P_R2C_ACC,
    P_Z_TV_ADJ, P_Z_TV_APPLYA, P_Z_DIPOLE,
    P_C2R_ADJ, P_C2R_F64OUT,
    // closed-form inversions and RTS
    P_Z_FILTER, P_Z_RTS_SETUP, P_Z_TV_SOLVE0, P_C2R_F64_MASKED, P_C2R_X0, P_R2C_X0,
    P_R2C_WRAPLAP, P_R2C_TVDIV_W, P_R2C_NLTV_BND, P_Z_RTS_SOLVE,
    P_KIND_COUNT
};

struct PassArgs {
    hipStream_t stream;
    const VolDesc* vols;
    const float2* tw;
    const VolStatus* status;  // convergence flags checked at block entry (may be null)
    VolStatus* status_rw;     // convergence state updated by the TV R2C pass (null: no update)
    VolStatus* mirror;        // host-visible copy of status_rw (may be null)
    int sel, sel2;            // spectrum selectors (y / z passes, R2C output)
    int parity;               // TV iteration parity
    int ridx, first, divide;  // V-SHARP radius index / first radius / SHARP division
                              // (TV record / adjoint passes: ridx = checkpoint index)
    double delta;             // erosion threshold
    int sched;                // column-pass FFT schedule (y / z passes): 0 two-stage, 1 LDS-staged
    int xshift;               // x passes: 0 or 2 (fixed_rows >> xshift pair-rows per block)
};

// Axis lengths with compiled-length kernels.
bool has_fixed_kernels(int n);

// Column-pass FFT schedule of (pass kind, axis length): the two-stage
// transform where the length allows it, unless the tuning table or the
// process-wide mode says otherwise.
int column_schedule(int kind, int n);

// Launch `kind` over the volumes of `lt`, all of which share the length of
// the pass axis `n`: a compile-time-length kernel when n is one of the
// specialised lengths, else the runtime-plan kernel sized by emax.
void launch_pass(int kind, int n, int emax, const LaunchTable& lt, const PassArgs& a);

}  // namespace qsm
