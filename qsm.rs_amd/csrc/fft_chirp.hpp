// Host-side planning of chirp-z (Bluestein) transforms: which lengths run
// directly, the convolution length of the others and their tables (see
// lds_fft_any in fft_core.hpp).  Host code only.
#pragma once

#include <cmath>
#include <vector>

#include "fft_core.hpp"

namespace qsmfft {

constexpr int kMaxDirectLen = 1024;
// Lengths with a prime factor above 13 run as chirp-z transforms through a
// convolution of length M >= 2n - 1 <= kMaxDirectLen, so up to this length.
constexpr int kMaxChirpLen = kMaxDirectLen / 2;

inline bool direct_length(int n) {
    if (n < 1 || n > kMaxDirectLen) return false;
    int r[kMaxStages], s;
    return factorize(n, r, &s);
}

// Relative cost of one runtime-plan stage of radix r per element: every
// stage is one LDS round trip and one barrier; radices 11 and 13 also keep
// one butterfly in flight per thread and a long register DFT.
inline double stage_cost(int r) {
    switch (r) {
        case 2: case 4: return 1.0;
        case 3: case 8: return 1.1;
        case 5: return 1.2;
        case 7: return 1.5;
        default: return 2.2;  // 11, 13
    }
}

// Convolution length of the chirp-z transform of length n: of the lengths
// M >= 2n - 1 the stages handle directly (up to kMaxDirectLen), the one with
// the least M * sum of its stage costs, so 512 = 8 8 8 is taken over
// 468 = 4 3 3 13 for n = 232 (0 if there is none).  128, 256, 512 and 1024
// count half: the runtime-plan kernels run them on compile-time stages.
inline int chirp_length(int n) {
    int best = 0;
    double best_cost = 0.0;
    for (int m = 2 * n - 1; m <= kMaxDirectLen; ++m) {
        int r[kMaxStages], st;
        if (!direct_length(m) || !factorize(m, r, &st)) continue;
        double c = 0.0;
        for (int i = 0; i < st; ++i) c += stage_cost(r[i]);
        c *= m;
        if (m == 128 || m == 256 || m == 512 || m == 1024) c *= 0.5;
        if (!best || c < best_cost) {
            best = m;
            best_cost = c;
        }
    }
    return best;
}

// Chirp c[j] = exp(-i pi j^2 / n), j < n (j^2 reduced mod 2n, exact), and the
// filter B[k] / M: B = DFT_M of conj(c[m]) at m and M - m, m < n, zero
// elsewhere, evaluated in long double.
inline void chirp_tables(int n, int M, std::vector<float2>& chirp, std::vector<float2>& filt) {
    const long double pi = 3.14159265358979323846264338327950288L;
    std::vector<long double> cr(n), ci(n);
    for (int j = 0; j < n; ++j) {
        const long long q = ((long long)j * j) % (2LL * n);
        const long double a = -pi * (long double)q / n;
        cr[j] = cosl(a);
        ci[j] = sinl(a);
    }
    chirp.resize(n);
    for (int j = 0; j < n; ++j) chirp[j] = make_float2((float)cr[j], (float)ci[j]);
    std::vector<long double> br(M, 0.0L), bi(M, 0.0L), wr(M), wi(M);
    for (int m = 0; m < n; ++m) {
        br[m] = cr[m];
        bi[m] = -ci[m];
        if (m > 0) {
            br[M - m] = cr[m];
            bi[M - m] = -ci[m];
        }
    }
    for (int q = 0; q < M; ++q) {
        const long double a = -2.0L * pi * q / M;
        wr[q] = cosl(a);
        wi[q] = sinl(a);
    }
    filt.resize(M);
    for (int k = 0; k < M; ++k) {
        long double sr = 0.0L, si = 0.0L;
        for (int m = 0; m < M; ++m) {
            if (br[m] == 0.0L && bi[m] == 0.0L) continue;
            const int q = (int)(((long long)m * k) % M);
            sr += br[m] * wr[q] - bi[m] * wi[q];
            si += br[m] * wi[q] + bi[m] * wr[q];
        }
        filt[k] = make_float2((float)(sr / M), (float)(si / M));
    }
}

}  // namespace qsmfft
