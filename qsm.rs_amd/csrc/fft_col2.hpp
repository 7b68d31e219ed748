// Two-stage column FFT for a tile of 16 columns of length N = 16 * R0
// (R0 <= 16: N = 64 ... 256) with 256 threads, one LDS exchange per transform.
//
// Thread t owns column c = t & 15 and row index j = t >> 4 (0..15).  Two
// register layouts of a tile:
//   L16: thread (j, c) holds rows j + 16 r, r < R0          (every thread)
//   LR : thread (j, c) holds rows j + R0 r, r < 16          (threads j < R0)
// L16 is the layout in which a block loads a tile from global memory (16
// consecutive threads read one 128-byte row segment), LR the layout the first
// transform leaves, so a forward transform, an element-wise product and an
// inverse transform run with two LDS exchanges in all.
//
// Both transforms are the mixed-radix Stockham factorisation (see
// fft_core.hpp) with two stages:
//   l16_to_lr: radix R0 (Ns = 1) then radix 16 (Ns = R0)
//   lr_to_l16: radix 16 (Ns = 1) then radix R0 (Ns = 16)
// The first stage's inputs are the thread's registers, the last stage's
// outputs stay in registers; only the hand-over between the stages goes
// through LDS.  Twiddle of the second stage: exp(-+2 pi i j r / N), read from
// the length-N table in LDS.
//
// LDS rows are padded by 16 elements (128 bytes) per group of rows one thread
// writes, where that makes the four row groups of a wave alternate between
// the two halves of the banks (groups of R0 rows when R0 is even, of 16 rows
// in lr_to_l16); reads are contiguous across a wave either way.
//
// Each stage is a separate function taking the thread index, so the host
// emulation can run every thread of one stage before the next one.
#pragma once

#include "fft_fixed.hpp"

namespace qsmfft {

// ---------------------------------------------- compile-time roots of unity
constexpr double kPi = 3.14159265358979323846264338327950288;

// sin and cos of 2 pi m / R by Taylor series after reduction to [-pi, pi].
constexpr double cx_angle(int m, int R) {
    int mm = m % R;
    if (2 * mm > R) mm -= R;
    return 2.0 * kPi * (double)mm / (double)R;
}
constexpr double cx_sin(double x) {
    double term = x, sum = x;
    for (int n = 1; n < 16; ++n) {
        term *= -x * x / ((2.0 * n) * (2.0 * n + 1.0));
        sum += term;
    }
    return sum;
}
constexpr double cx_cos(double x) {
    double term = 1.0, sum = 1.0;
    for (int n = 1; n < 16; ++n) {
        term *= -x * x / ((2.0 * n - 1.0) * (2.0 * n));
        sum += term;
    }
    return sum;
}

template <int R>
struct RootTable {
    float c[R], s[R];  // cos, sin of 2 pi m / R
    constexpr RootTable() : c(), s() {
        for (int m = 0; m < R; ++m) {
            c[m] = (float)cx_cos(cx_angle(m, R));
            s[m] = (float)cx_sin(cx_angle(m, R));
        }
    }
};

// DFT of R = A * B points in registers (Cooley-Tukey, one split):
//   X[k1 + A k2] = sum_n2 W_B^(n2 k2) W_R^(n2 k1) sum_n1 W_A^(n1 k1) x[B n1 + n2]
// with W_m = exp(-2 pi i / m) (forward; the inverse conjugates).
template <int A, int B, bool INV>
struct DftSplit {
    __host__ __device__ static __forceinline__ void run(float2* v, const float2*, int) {
        constexpr int R = A * B;
        constexpr RootTable<R> W{};
        float2 a[B][A];
#pragma unroll
        for (int n2 = 0; n2 < B; ++n2) {
            float2 t[A];
#pragma unroll
            for (int n1 = 0; n1 < A; ++n1) t[n1] = v[B * n1 + n2];
            Dft<A, INV>::run(t, nullptr, 0);
#pragma unroll
            for (int k1 = 0; k1 < A; ++k1) {
                const int m = (n2 * k1) % R;
                if (m == 0) {
                    a[n2][k1] = t[k1];
                } else {
                    const float2 w = make_float2(W.c[m], INV ? W.s[m] : -W.s[m]);
                    a[n2][k1] = cmul(t[k1], w);
                }
            }
        }
#pragma unroll
        for (int k1 = 0; k1 < A; ++k1) {
            float2 t[B];
#pragma unroll
            for (int n2 = 0; n2 < B; ++n2) t[n2] = a[n2][k1];
            Dft<B, INV>::run(t, nullptr, 0);
#pragma unroll
            for (int k2 = 0; k2 < B; ++k2) v[k1 + A * k2] = t[k2];
        }
    }
};

template <bool INV> struct Dft<9, INV> : DftSplit<3, 3, INV> {};
template <bool INV> struct Dft<10, INV> : DftSplit<2, 5, INV> {};
template <bool INV> struct Dft<12, INV> : DftSplit<4, 3, INV> {};
template <bool INV> struct Dft<14, INV> : DftSplit<2, 7, INV> {};
template <bool INV> struct Dft<15, INV> : DftSplit<3, 5, INV> {};

// ------------------------------------------------------------ column FFT --
// Lengths 16 * R0 with a register DFT of R0 points: 64, 128, 144 ... 256.
__host__ __device__ constexpr bool col2_ok(int n) {
    if (n <= 0 || n % 16 != 0) return false;
    const int r0 = n / 16;
    return r0 == 4 || (r0 >= 8 && r0 <= 16);
}
template <int R0>
struct Col2 {
    static constexpr int N = 16 * R0;
    static constexpr int VMAX = 16;  // registers per thread (R0 <= 16)
    // padding of the l16_to_lr exchange: per group of R0 rows (even R0)
    static constexpr int PAD_A = (R0 % 2 == 0) ? 16 : 0;
    // LDS elements for either exchange
    static constexpr int LDS = N * 16 + 16 * 16;

    // --- l16_to_lr ---
    template <bool INV>
    __host__ __device__ static __forceinline__ void a_fwd(int tid, float2 (&v)[VMAX], float2* lds) {
        const int c = tid & 15, j = tid >> 4;
        Dft<R0, INV>::run(v, nullptr, 0);
        float2* p = lds + (j * R0) * 16 + c + j * PAD_A;
#pragma unroll
        for (int r = 0; r < R0; ++r) p[r * 16] = v[r];
    }
    template <bool INV>
    __host__ __device__ static __forceinline__ void b_fwd(int tid, float2 (&v)[VMAX], const float2* lds,
                                                          const float2* twl) {
        const int c = tid & 15, j = tid >> 4;
        if (j >= R0) return;
#pragma unroll
        for (int r = 0; r < 16; ++r) v[r] = lds[(j + R0 * r) * 16 + c + r * PAD_A];
        if (j != 0) {
#pragma unroll
            for (int r = 1; r < 16; ++r) {
                float2 w = twl[j * r];
                if (INV) w = cconj(w);
                v[r] = cmul(v[r], w);
            }
        }
        Dft<16, INV>::run(v, nullptr, 0);
    }

    // --- lr_to_l16 ---
    template <bool INV>
    __host__ __device__ static __forceinline__ void a_rev(int tid, float2 (&v)[VMAX], float2* lds) {
        const int c = tid & 15, j = tid >> 4;
        if (j >= R0) return;
        Dft<16, INV>::run(v, nullptr, 0);
        float2* p = lds + (j * 16) * 16 + c + j * 16;
#pragma unroll
        for (int r = 0; r < 16; ++r) p[r * 16] = v[r];
    }
    template <bool INV>
    __host__ __device__ static __forceinline__ void b_rev(int tid, float2 (&v)[VMAX], const float2* lds,
                                                          const float2* twl) {
        const int c = tid & 15, j = tid >> 4;
#pragma unroll
        for (int r = 0; r < R0; ++r) v[r] = lds[(j + 16 * r) * 16 + c + r * 16];
        if (j != 0) {
#pragma unroll
            for (int r = 1; r < R0; ++r) {
                float2 w = twl[j * r];
                if (INV) w = cconj(w);
                v[r] = cmul(v[r], w);
            }
        }
        Dft<R0, INV>::run(v, nullptr, 0);
    }

#ifdef __HIPCC__
    template <bool INV>
    __device__ static __forceinline__ void l16_to_lr(float2 (&v)[VMAX], float2* lds, const float2* twl) {
        a_fwd<INV>(threadIdx.x, v, lds);
        __syncthreads();
        b_fwd<INV>(threadIdx.x, v, lds, twl);
    }
    template <bool INV>
    __device__ static __forceinline__ void lr_to_l16(float2 (&v)[VMAX], float2* lds, const float2* twl) {
        a_rev<INV>(threadIdx.x, v, lds);
        __syncthreads();
        b_rev<INV>(threadIdx.x, v, lds, twl);
    }
#endif
};

}  // namespace qsmfft
