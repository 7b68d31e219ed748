// CPU emulation of the LDS Stockham FFT in fft_core.hpp.
//
// Runs the same stage_load / stage_store code the GPU kernels use, one
// simulated thread at a time with a "barrier" between the phases, for every
// length the workloads need, both layouts (row and column tiles) and both
// directions, and compares against a direct O(n^2) DFT in double precision.
// Built with hipcc as host code; prints one line per case and exits non-zero
// on any mismatch.
#include "../../qsm.rs_amd/csrc/fft_col2.hpp"
#include "../../qsm.rs_amd/csrc/fft_chirp.hpp"

#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <array>
#include <vector>

using namespace qsmfft;

// The out-of-place stages of lds_fft_any: every simulated thread runs its
// share of a stage from src to dst, then the buffers swap (the barrier).
// Returns the buffer holding the result.
template <bool COLFAST, int LOG2C, bool INV>
static std::vector<float2>* emu_stages(std::vector<float2>* src, std::vector<float2>* dst, const Plan& P,
                                       int ncols, int es, int cs, const float2* tw) {
    for (int s = 0; s < P.nstages; ++s) {
        for (int tid = 0; tid < kThreads; ++tid)
            plan_stage<COLFAST, LOG2C, INV>(src->data(), dst->data(), tid, P, s, ncols, es, cs, tw);
        std::swap(src, dst);
    }
    return src;
}

template <int EMAX, bool COLFAST, int LOG2C, bool INV>
static void emu_fft(std::vector<float2>& lds, const Plan& P, int ncols, int es, int cs,
                    const float2* tw) {
    std::vector<float2> other(lds.size(), make_float2(NAN, NAN));
    std::vector<float2>* r = emu_stages<COLFAST, LOG2C, INV>(&lds, &other, P, ncols, es, cs, tw);
    if (r != &lds) lds = *r;
}

template <int EMAX, bool COLFAST, int LOG2C, bool INV>
static double run_case(int n, int ncols) {
    Plan P{};
    if (!make_plan_stages(n, P)) return -1.0;
    std::vector<float2> tw(n);
    for (int q = 0; q < n; ++q) {
        double a = -2.0 * M_PI * q / n;
        tw[q] = make_float2((float)std::cos(a), (float)std::sin(a));
    }
    int es = COLFAST ? ncols : 1;
    int cs = COLFAST ? 1 : n;
    std::vector<float2> lds(n * ncols);
    std::vector<std::complex<double>> in(n * ncols);
    srand(n * 7919 + ncols);
    for (int c = 0; c < ncols; ++c)
        for (int e = 0; e < n; ++e) {
            double re = rand() / (double)RAND_MAX - 0.5, im = rand() / (double)RAND_MAX - 0.5;
            in[c * n + e] = {re, im};
            lds[e * es + c * cs] = make_float2((float)re, (float)im);
        }
    emu_fft<EMAX, COLFAST, LOG2C, INV>(lds, P, ncols, es, cs, tw.data());
    double maxerr = 0.0, maxref = 0.0;
    const double sgn = INV ? 1.0 : -1.0;
    for (int c = 0; c < ncols; ++c)
        for (int k = 0; k < n; ++k) {
            std::complex<double> acc = 0.0;
            for (int e = 0; e < n; ++e)
                acc += in[c * n + e] * std::polar(1.0, sgn * 2.0 * M_PI * ((long)k * e % n) / n);
            float2 g = lds[k * es + c * cs];
            maxerr = std::fmax(maxerr, std::abs(acc - std::complex<double>(g.x, g.y)));
            maxref = std::fmax(maxref, std::abs(acc));
        }
    return maxerr / maxref;
}

// ---- chirp-z (Bluestein) lengths: the device phases of lds_fft_any ----
template <int EMAX, bool COLFAST, int LOG2C, bool INV>
static void emu_bluestein(std::vector<float2>& lds, const Plan& P, int ncols, int es, int cs,
                          const float2* tw, const float2* chirp, const float2* filt) {
    const BlGeom g = bl_geom(P, ncols, es, cs, COLFAST);
    std::vector<float2> other(lds.size(), make_float2(NAN, NAN));
    for (int tid = 0; tid < kThreads; ++tid) bl_in<COLFAST, LOG2C, INV>(lds.data(), other.data(), tid, g, chirp);
    std::vector<float2>* r = emu_stages<COLFAST, LOG2C, false>(&other, &lds, P, ncols, g.wes, g.wcs, tw);
    for (int tid = 0; tid < kThreads; ++tid) bl_filter<COLFAST, LOG2C>(r->data(), tid, g, filt);
    std::vector<float2>* o = (r == &lds) ? &other : &lds;
    r = emu_stages<COLFAST, LOG2C, true>(r, o, P, ncols, g.wes, g.wcs, tw);
    o = (r == &lds) ? &other : &lds;
    for (int tid = 0; tid < kThreads; ++tid) bl_out<COLFAST, LOG2C, INV>(r->data(), o->data(), tid, g, chirp);
    if (o != &lds) lds = *o;
}

// Length n through the chirp-z path with ncols sequences; the LDS array is
// sized as the kernels size it (ncols * M elements).
template <int EMAX, bool COLFAST, int LOG2C, bool INV>
static double run_bluestein(int n, int ncols) {
    const int M = chirp_length(n);
    if (M <= 0) return -1.0;
    Plan P{};
    if (!make_plan_stages(M, P)) return -1.0;
    P.bl_n = n;
    std::vector<float2> tw(M), chirp, filt;
    for (int q = 0; q < M; ++q) {
        double a = -2.0 * M_PI * q / M;
        tw[q] = make_float2((float)std::cos(a), (float)std::sin(a));
    }
    chirp_tables(n, M, chirp, filt);
    const int es = COLFAST ? ncols : 1, cs = COLFAST ? 1 : n;
    std::vector<float2> lds((size_t)ncols * M, make_float2(NAN, NAN));
    std::vector<std::complex<double>> in(n * ncols);
    srand(n * 977 + ncols);
    for (int c = 0; c < ncols; ++c)
        for (int e = 0; e < n; ++e) {
            double re = rand() / (double)RAND_MAX - 0.5, im = rand() / (double)RAND_MAX - 0.5;
            in[c * n + e] = {re, im};
            lds[e * es + c * cs] = make_float2((float)re, (float)im);
        }
    emu_bluestein<EMAX, COLFAST, LOG2C, INV>(lds, P, ncols, es, cs, tw.data(), chirp.data(), filt.data());
    double maxerr = 0.0, maxref = 0.0;
    const double sgn = INV ? 1.0 : -1.0;
    for (int c = 0; c < ncols; ++c)
        for (int k = 0; k < n; ++k) {
            std::complex<double> acc = 0.0;
            for (int e = 0; e < n; ++e)
                acc += in[c * n + e] * std::polar(1.0, sgn * 2.0 * M_PI * ((long)k * e % n) / n);
            float2 g = lds[k * es + c * cs];
            maxerr = std::fmax(maxerr, std::abs(acc - std::complex<double>(g.x, g.y)));
            maxref = std::fmax(maxref, std::abs(acc));
        }
    return maxerr / maxref;
}

static int check_bluestein(int n) {
    const int M = chirp_length(n);
    // column tiles of 16 (EMAX 32: 16 M <= 8192), of 8 for M above 512; rows
    // as the x passes size them
    const bool wide = M <= 512;
    double e1 = wide ? run_bluestein<32, true, 4, false>(n, 16) : run_bluestein<32, true, 3, false>(n, 8);
    double e2 = wide ? run_bluestein<32, true, 4, true>(n, 16) : run_bluestein<32, true, 3, true>(n, 8);
    int rows = (kThreads * 16) / M; if (rows < 1) rows = 1; if (rows > 32) rows = 32;
    double e3 = run_bluestein<16, false, 0, false>(n, rows);
    double e4 = run_bluestein<16, false, 0, true>(n, rows);
    double worst = std::fmax(std::fmax(e1, e2), std::fmax(e3, e4));
    bool ok = e1 >= 0 && e2 >= 0 && e3 >= 0 && e4 >= 0 && worst < 4e-6;
    printf("chirp-z n=%4d M=%4d colfwd=%.2e colinv=%.2e rowfwd=%.2e rowinv=%.2e %s\n", n, M, e1, e2,
           e3, e4, ok ? "ok" : "FAIL");
    return ok ? 0 : 1;
}

// ---- compile-time-length path (fft_fixed.hpp) ----
template <int N, int S, int NCOLS, bool COLFAST, bool INV>
static void emu_fixed_stages(std::vector<float2>& lds, const float2* tw) {
    if constexpr (S < FixedSeq<N>::count) {
        using G = FixedStage<N, S, NCOLS, COLFAST>;
        std::vector<std::array<std::array<float2, G::R>, G::B>> regs(kThreads);
        for (int tid = 0; tid < kThreads; ++tid) {
            float2 v[G::B][G::R];
            fixed_stage_load<N, S, NCOLS, COLFAST>(lds.data(), tid, v);
            for (int t = 0; t < G::B; ++t)
                for (int r = 0; r < G::R; ++r) regs[tid][t][r] = v[t][r];
        }
        for (int tid = 0; tid < kThreads; ++tid) {
            float2 v[G::B][G::R];
            for (int t = 0; t < G::B; ++t)
                for (int r = 0; r < G::R; ++r) v[t][r] = regs[tid][t][r];
            fixed_stage_store<N, S, NCOLS, COLFAST, INV>(lds.data(), tid, v, tw);
        }
        emu_fixed_stages<N, S + 1, NCOLS, COLFAST, INV>(lds, tw);
    }
}

template <int N, int NCOLS, bool COLFAST, bool INV>
static double run_fixed(void) {
    std::vector<float2> tw(N);
    for (int q = 0; q < N; ++q) {
        double a = -2.0 * M_PI * q / N;
        tw[q] = make_float2((float)std::cos(a), (float)std::sin(a));
    }
    const int es = COLFAST ? NCOLS : 1, cs = COLFAST ? 1 : N;
    std::vector<float2> lds(N * NCOLS);
    std::vector<std::complex<double>> in(N * NCOLS);
    srand(N * 31 + NCOLS);
    for (int c = 0; c < NCOLS; ++c)
        for (int e = 0; e < N; ++e) {
            double re = rand() / (double)RAND_MAX - 0.5, im = rand() / (double)RAND_MAX - 0.5;
            in[c * N + e] = {re, im};
            lds[e * es + c * cs] = make_float2((float)re, (float)im);
        }
    emu_fixed_stages<N, 0, NCOLS, COLFAST, INV>(lds, tw.data());
    double maxerr = 0.0, maxref = 0.0;
    const double sgn = INV ? 1.0 : -1.0;
    for (int c = 0; c < NCOLS; ++c)
        for (int k = 0; k < N; ++k) {
            std::complex<double> acc = 0.0;
            for (int e = 0; e < N; ++e)
                acc += in[c * N + e] * std::polar(1.0, sgn * 2.0 * M_PI * ((long)k * e % N) / N);
            float2 g = lds[k * es + c * cs];
            maxerr = std::fmax(maxerr, std::abs(acc - std::complex<double>(g.x, g.y)));
            maxref = std::fmax(maxref, std::abs(acc));
        }
    return maxerr / maxref;
}

template <int N>
static int check_fixed() {
    double e1 = run_fixed<N, 16, true, false>();
    double e2 = run_fixed<N, 16, true, true>();
    double e3 = run_fixed<N, fixed_rows(N), false, false>();
    double e4 = run_fixed<N, fixed_rows(N), false, true>();
    // the row counts of the small-volume x passes: a quarter of the rows
    // (R2C) and half of that (C2R)
    constexpr int q = (fixed_rows(N) >> 2) > 0 ? (fixed_rows(N) >> 2) : 1;
    constexpr int qc = q / 2 > 0 ? q / 2 : 1;
    e3 = std::fmax(e3, run_fixed<N, q, false, false>());
    e4 = std::fmax(e4, run_fixed<N, qc, false, true>());
    double worst = std::fmax(std::fmax(e1, e2), std::fmax(e3, e4));
    bool ok = worst < 2e-6;
    printf("fixed n=%4d stages=%d colfwd=%.2e colinv=%.2e rowfwd=%.2e rowinv=%.2e %s\n", N,
           FixedSeq<N>::count, e1, e2, e3, e4, ok ? "ok" : "FAIL");
    return ok ? 0 : 1;
}

// ---- two-stage column path (fft_col2.hpp) ----
// Column c of the tile is random; the l16_to_lr transform starts from the
// load layout and must leave X[j + R0 r] in thread (j, c); the lr_to_l16
// transform starts from that layout and must leave X[j + 16 r].
template <int R0, bool INV>
static double run_col2(bool rev) {
    using C = Col2<R0>;
    constexpr int N = C::N;
    std::vector<float2> tw(N);
    for (int q = 0; q < N; ++q) {
        double a = -2.0 * M_PI * q / N;
        tw[q] = make_float2((float)std::cos(a), (float)std::sin(a));
    }
    std::vector<std::complex<double>> in(N * 16);
    srand(N * 131 + (rev ? 7 : 0) + (INV ? 3 : 0));
    for (auto& z : in) z = {rand() / (double)RAND_MAX - 0.5, rand() / (double)RAND_MAX - 0.5};
    std::vector<float2> lds(C::LDS, make_float2(NAN, NAN));
    std::vector<std::array<float2, C::VMAX>> regs(kThreads);
    // input layout: L16 (rev = false) or LR (rev = true)
    for (int tid = 0; tid < kThreads; ++tid) {
        const int c = tid & 15, j = tid >> 4;
        for (int r = 0; r < C::VMAX; ++r) regs[tid][r] = make_float2(NAN, NAN);
        if (!rev) {
            for (int r = 0; r < R0; ++r) {
                auto z = in[c * N + j + 16 * r];
                regs[tid][r] = make_float2((float)z.real(), (float)z.imag());
            }
        } else if (j < R0) {
            for (int r = 0; r < 16; ++r) {
                auto z = in[c * N + j + R0 * r];
                regs[tid][r] = make_float2((float)z.real(), (float)z.imag());
            }
        }
    }
    for (int tid = 0; tid < kThreads; ++tid) {
        float2 v[C::VMAX];
        for (int r = 0; r < C::VMAX; ++r) v[r] = regs[tid][r];
        if (!rev) C::template a_fwd<INV>(tid, v, lds.data());
        else C::template a_rev<INV>(tid, v, lds.data());
        for (int r = 0; r < C::VMAX; ++r) regs[tid][r] = v[r];
    }
    for (int tid = 0; tid < kThreads; ++tid) {
        float2 v[C::VMAX];
        for (int r = 0; r < C::VMAX; ++r) v[r] = regs[tid][r];
        if (!rev) C::template b_fwd<INV>(tid, v, lds.data(), tw.data());
        else C::template b_rev<INV>(tid, v, lds.data(), tw.data());
        for (int r = 0; r < C::VMAX; ++r) regs[tid][r] = v[r];
    }
    const double sgn = INV ? 1.0 : -1.0;
    double maxerr = 0.0, maxref = 0.0;
    int seen = 0;
    for (int tid = 0; tid < kThreads; ++tid) {
        const int c = tid & 15, j = tid >> 4;
        const int cnt = rev ? R0 : (j < R0 ? 16 : 0);
        for (int r = 0; r < cnt; ++r) {
            const int k = rev ? j + 16 * r : j + R0 * r;
            std::complex<double> acc = 0.0;
            for (int e = 0; e < N; ++e)
                acc += in[c * N + e] * std::polar(1.0, sgn * 2.0 * M_PI * ((long)k * e % N) / N);
            const float2 g = regs[tid][r];
            maxerr = std::fmax(maxerr, std::abs(acc - std::complex<double>(g.x, g.y)));
            maxref = std::fmax(maxref, std::abs(acc));
            ++seen;
        }
    }
    if (seen != N * 16) return 1.0;  // every output exactly once
    return maxerr / maxref;
}

template <int R0>
static int check_col2() {
    const double e1 = run_col2<R0, false>(false), e2 = run_col2<R0, true>(false);
    const double e3 = run_col2<R0, false>(true), e4 = run_col2<R0, true>(true);
    const double worst = std::fmax(std::fmax(e1, e2), std::fmax(e3, e4));
    const bool ok = worst < 2e-6;
    printf("col2 n=%4d l16fwd=%.2e l16inv=%.2e lrfwd=%.2e lrinv=%.2e %s\n", 16 * R0, e1, e2, e3, e4,
           ok ? "ok" : "FAIL");
    return ok ? 0 : 1;
}

int main() {
    int bad_col2 = check_col2<4>() + check_col2<8>() + check_col2<9>() + check_col2<10>() +
                   check_col2<11>() + check_col2<12>() + check_col2<13>() + check_col2<14>() +
                   check_col2<15>() + check_col2<16>();
    int bad_fixed = check_fixed<16>() + check_fixed<64>() + check_fixed<128>() + check_fixed<144>() +
                    check_fixed<160>() + check_fixed<176>() + check_fixed<192>() +
                    check_fixed<208>() + check_fixed<224>() + check_fixed<240>() +
                    check_fixed<256>() + check_fixed<384>() + check_fixed<96>() + check_fixed<112>() +
                    check_fixed<180>() + check_fixed<200>() + check_fixed<216>() + check_fixed<220>() +
                    check_fixed<320>() + check_fixed<512>() + check_fixed<288>() +
                    check_fixed<416>() + check_fixed<448>() + check_fixed<480>() + check_fixed<48>() +
                    check_fixed<72>() + check_fixed<80>() + check_fixed<120>() + check_fixed<150>() +
                    check_fixed<168>();
    const int lens[] = {1, 2, 3, 4, 5, 6, 7, 8, 10, 11, 12, 13, 14, 16, 22, 26, 32, 64,
                        72, 80, 88, 96, 104, 112, 120, 128, 144, 160, 176, 192, 208, 224,
                        240, 256, 384};
    int bad = 0;
    for (int n : lens) {
        // column tiles of 16 (EMAX 24 covers n <= 384)
        double e1 = (n * 16 <= kThreads * 24) ? run_case<24, true, 4, false>(n, 16) : 0.0;
        double e2 = (n * 16 <= kThreads * 24) ? run_case<24, true, 4, true>(n, 16) : 0.0;
        // row tiles: as many rows as fit in EMAX 16
        int rows = (kThreads * 16) / n; if (rows < 1) rows = 1; if (rows > 32) rows = 32;
        double e3 = (n <= kThreads * 16) ? run_case<16, false, 0, false>(n, rows) : 0.0;
        double e4 = (n <= kThreads * 16) ? run_case<16, false, 0, true>(n, rows) : 0.0;
        double worst = std::fmax(std::fmax(e1, e2), std::fmax(e3, e4));
        bool ok = e1 >= 0 && e2 >= 0 && e3 >= 0 && e4 >= 0 && worst < 2e-6;
        printf("n=%4d colfwd=%.2e colinv=%.2e rowfwd=%.2e rowinv=%.2e %s\n", n, e1, e2, e3, e4,
               ok ? "ok" : "FAIL");
        if (!ok) ++bad;
    }
    // lengths with a prime factor above 13 (17, 19, 23, 68 = 4 x 17, 97,
    // 164 = 4 x 41, 232 = 8 x 29, 251, 256 - 2 = 254 = 2 x 127) and a direct
    // length run through the same path
    for (int n : {17, 19, 23, 68, 97, 164, 232, 251, 254, 30, 263, 293, 509}) bad += check_bluestein(n);
    // direct lengths above 512: column tiles of 8 (EMAX 32: 8 n <= 8192), rows of 4
    for (int n : {520, 576, 640, 729, 768, 1000, 1024}) {
        double e1 = run_case<32, true, 3, false>(n, 8), e2 = run_case<32, true, 3, true>(n, 8);
        double e3 = run_case<16, false, 0, false>(n, 4), e4 = run_case<16, false, 0, true>(n, 4);
        double worst = std::fmax(std::fmax(e1, e2), std::fmax(e3, e4));
        bool ok = e1 >= 0 && e2 >= 0 && e3 >= 0 && e4 >= 0 && worst < 2e-6;
        printf("n=%4d col8fwd=%.2e col8inv=%.2e rowfwd=%.2e rowinv=%.2e %s\n", n, e1, e2, e3, e4,
               ok ? "ok" : "FAIL");
        if (!ok) ++bad;
    }
    bad += bad_fixed + bad_col2;
    printf("%s\n", bad ? "FAILED" : "ALL OK");
    return bad ? 1 : 0;
}
