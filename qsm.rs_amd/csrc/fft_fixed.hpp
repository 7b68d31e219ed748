// Compile-time-length Stockham FFT for the axis lengths the workloads use.
//
// Same stage formulation as fft_core.hpp (inputs j + r*N/R, twiddle
// exp(-+2 pi i (j mod Ns) r / (Ns R)), outputs (j/Ns)*Ns*R + j mod Ns + r*Ns),
// but N, the radix sequence, the number of sequences per block and every
// stage's geometry are template constants: the stages are straight-line code,
// all index divisions fold to shifts or multiplies, and each stage's
// registers are dead before the next begins.  Radix 16 (a 4x4 split in
// registers) halves the number of LDS round trips for power-of-two lengths:
// 256 = 16*16 takes two stages.
#pragma once

#include "fft_core.hpp"

namespace qsmfft {

// ------------------------------------------------------------- radix 16 --
// X[k1 + 4 k2] = sum_n2 W4^(n2 k2) W16^(n2 k1) sum_n1 W4^(n1 k1) x[4 n1 + n2]
template <bool INV>
struct Dft<16, INV> {
    __host__ __device__ static __forceinline__ void run(float2* v, const float2*, int) {
        // W16^m = exp(-2 pi i m/16), m = 0..9 (forward); inverse conjugates
        constexpr float c[10] = {1.0f, 0.92387953251128675613f, 0.70710678118654752440f,
                                 0.38268343236508977173f, 0.0f, -0.38268343236508977173f,
                                 -0.70710678118654752440f, -0.92387953251128675613f, -1.0f,
                                 -0.92387953251128675613f};
        constexpr float s[10] = {0.0f, 0.38268343236508977173f, 0.70710678118654752440f,
                                 0.92387953251128675613f, 1.0f, 0.92387953251128675613f,
                                 0.70710678118654752440f, 0.38268343236508977173f, 0.0f,
                                 -0.38268343236508977173f};
        float2 a[4][4];
#pragma unroll
        for (int n2 = 0; n2 < 4; ++n2) {
            float2 t[4] = {v[n2], v[4 + n2], v[8 + n2], v[12 + n2]};
            Dft<4, INV>::run(t, nullptr, 0);
#pragma unroll
            for (int k1 = 0; k1 < 4; ++k1) {
                const int m = n2 * k1;
                if (m == 0) {
                    a[n2][k1] = t[k1];
                } else {
                    const float2 w = make_float2(c[m], INV ? s[m] : -s[m]);
                    a[n2][k1] = cmul(t[k1], w);
                }
            }
        }
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) {
            float2 t[4] = {a[0][k1], a[1][k1], a[2][k1], a[3][k1]};
            Dft<4, INV>::run(t, nullptr, 0);
#pragma unroll
            for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = t[k2];
        }
    }
};

// ------------------------------------------------------- radix sequences --
struct RadixSeq {
    int count;
    int r[kMaxStages];
};

// Radices for a fixed length: 16s, then 8, 4, 2, then the odd primes.
constexpr RadixSeq fixed_radices(int n) {
    RadixSeq s{0, {0, 0, 0, 0, 0, 0, 0, 0, 0, 0}};
    int m = n;
    while (m % 16 == 0 && s.count < kMaxStages) { s.r[s.count++] = 16; m /= 16; }
    while (m % 8 == 0 && s.count < kMaxStages) { s.r[s.count++] = 8; m /= 8; }
    while (m % 4 == 0 && s.count < kMaxStages) { s.r[s.count++] = 4; m /= 4; }
    while (m % 2 == 0 && s.count < kMaxStages) { s.r[s.count++] = 2; m /= 2; }
    const int odd[5] = {3, 5, 7, 11, 13};
    for (int i = 0; i < 5; ++i)
        while (m % odd[i] == 0 && s.count < kMaxStages) { s.r[s.count++] = odd[i]; m /= odd[i]; }
    if (m != 1) s.count = -1;
    return s;
}

constexpr int fixed_prefix(int n, int stage) {
    RadixSeq s = fixed_radices(n);
    int p = 1;
    for (int i = 0; i < stage; ++i) p *= s.r[i];
    return p;
}

template <int N>
struct FixedSeq {
    static constexpr RadixSeq seq = fixed_radices(N);
    static constexpr int count = seq.count;
    static_assert(seq.count > 0, "length has a prime factor the FFT does not handle");
};

// Geometry of stage S for NCOLS sequences of length N; COLFAST: sequence
// index fastest across threads (column tiles, element (e, c) at e*NCOLS + c),
// otherwise butterfly index fastest (rows, element (e, c) at e + c*N).
template <int N, int S, int NCOLS, bool COLFAST>
struct FixedStage {
    static constexpr int R = FixedSeq<N>::seq.r[S];
    static constexpr int Ns = fixed_prefix(N, S);
    static constexpr int nb = N / R;
    static constexpr int tstep = nb / Ns;
    static constexpr int total = nb * NCOLS;
    static constexpr int B = (total + kThreads - 1) / kThreads;
    static constexpr bool exact = (total % kThreads) == 0;
    static constexpr int es = COLFAST ? NCOLS : 1;
    static constexpr int cs = COLFAST ? 1 : N;
    __host__ __device__ static __forceinline__ void split(int b, int& j, int& c) {
        if (COLFAST) { c = b % NCOLS; j = b / NCOLS; }
        else { j = b % nb; c = b / nb; }
    }
};

template <int N, int S, int NCOLS, bool COLFAST>
__host__ __device__ __forceinline__ void fixed_stage_load(
    const float2* lds, int tid,
    float2 (&v)[FixedStage<N, S, NCOLS, COLFAST>::B][FixedStage<N, S, NCOLS, COLFAST>::R]) {
    using G = FixedStage<N, S, NCOLS, COLFAST>;
#pragma unroll
    for (int t = 0; t < G::B; ++t) {
        const int b = tid + t * kThreads;
        if (G::exact || b < G::total) {
            int j, c;
            G::split(b, j, c);
            const float2* p = lds + j * G::es + c * G::cs;
#pragma unroll
            for (int r = 0; r < G::R; ++r) v[t][r] = p[r * G::nb * G::es];
        }
    }
}

template <int N, int S, int NCOLS, bool COLFAST, bool INV>
__host__ __device__ __forceinline__ void fixed_stage_store(
    float2* lds, int tid,
    float2 (&v)[FixedStage<N, S, NCOLS, COLFAST>::B][FixedStage<N, S, NCOLS, COLFAST>::R],
    const float2* tw) {
    using G = FixedStage<N, S, NCOLS, COLFAST>;
#pragma unroll
    for (int t = 0; t < G::B; ++t) {
        const int b = tid + t * kThreads;
        if (G::exact || b < G::total) {
            int j, c;
            G::split(b, j, c);
            const int k = j % G::Ns;
            const int grp = j / G::Ns;
            if (G::Ns > 1 && k != 0) {
                const int ts = k * G::tstep;
#pragma unroll
                for (int r = 1; r < G::R; ++r) {
                    float2 w = tw[r * ts];
                    if (INV) w = cconj(w);
                    v[t][r] = cmul(v[t][r], w);
                }
            }
            Dft<G::R, INV>::run(v[t], tw, N);
            float2* p = lds + (grp * G::Ns * G::R + k) * G::es + c * G::cs;
#pragma unroll
            for (int r = 0; r < G::R; ++r) p[r * G::Ns * G::es] = v[t][r];
        }
    }
}

#ifdef __HIPCC__
template <int N, int NCOLS, bool COLFAST, bool INV, int S = 0>
__device__ __forceinline__ void fixed_fft(float2* lds, const float2* tw) {
    if constexpr (S < FixedSeq<N>::count) {
        using G = FixedStage<N, S, NCOLS, COLFAST>;
        float2 v[G::B][G::R];
        fixed_stage_load<N, S, NCOLS, COLFAST>(lds, threadIdx.x, v);
        __syncthreads();
        fixed_stage_store<N, S, NCOLS, COLFAST, INV>(lds, threadIdx.x, v, tw);
        __syncthreads();
        fixed_fft<N, NCOLS, COLFAST, INV, S + 1>(lds, tw);
    }
}
#endif

// Sequences per block for a row (x) pass of length N: pairs of real rows,
// at most 32 and at most 16 elements per thread.
constexpr int fixed_rows(int n) {
    int rb = (kThreads * 16) / n;
    return rb < 1 ? 1 : (rb > 32 ? 32 : rb);
}

// Pair-rows per block of the TV C2R pass: half of the other row passes, so
// the previous iterate can be held in registers across the FFT.
constexpr int fixed_rows_c(int n) {
    int rb = fixed_rows(n) / 2;
    return rb < 1 ? 1 : rb;
}

}  // namespace qsmfft
