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

// Convolution length of the chirp-z transform of length n: the smallest
// length >= 2n - 1 the stages handle directly (0 if none <= kMaxDirectLen).
inline int chirp_length(int n) {
    for (int m = 2 * n - 1; m <= kMaxDirectLen; ++m)
        if (direct_length(m)) return m;
    return 0;
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
