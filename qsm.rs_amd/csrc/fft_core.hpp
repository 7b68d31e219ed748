// Mixed-radix Stockham FFT building blocks for gfx950.
//
// A workgroup holds `ncols` independent sequences of length n in LDS, element
// e of sequence c at lds[e*es + c*cs].  Each stage of radix R reads the R
// inputs of every butterfly into registers, waits at a barrier, then applies
// twiddles, a radix-R DFT and writes the R outputs back to their Stockham
// (self-sorting) positions, so the transform ends in natural order with no
// bit-reversal pass.  For butterfly j of a stage where Ns is the product of
// the radices already applied:
//   inputs   e = j + r*(n/R)                          r = 0..R-1
//   twiddle  w = exp(-+2 pi i * (j mod Ns) * r / (Ns*R))
//   outputs  e = (j / Ns)*Ns*R + (j mod Ns) + r*Ns
// (the Stockham autosort formulation of a decimation-in-time FFT).
//
// Everything index-related is __host__ __device__ so the same code can be
// run one simulated thread at a time on the CPU and compared against a
// reference DFT before it ever runs on the GPU.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace qsmfft {

constexpr int kMaxStages = 10;
constexpr int kThreads = 256;  // threads per workgroup for every FFT kernel

// Per-length plan.  Stored in device memory; read with uniform loads.
// A length with a prime factor above 13 runs as a chirp-z (Bluestein)
// transform: bl_n is then the axis length, n the convolution length M the
// stages run at, bl_chirp / bl_filt the offsets of its chirp and filter
// tables (see bluestein_fft); bl_n = 0 for a direct transform.
struct Plan {
    int n;
    int nstages;
    int radix[kMaxStages];
    int ns[kMaxStages];          // product of the radices before stage s
    uint64_t m_nb[kMaxStages];   // magic for division by n / radix[s]
    uint64_t m_ns[kMaxStages];   // magic for division by ns[s]
    int tw_off;  // offset (in float2) of this length's table exp(-2 pi i q/n), q<n
    int bl_n, bl_chirp, bl_filt;
};

__host__ __device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__host__ __device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__host__ __device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
__host__ __device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// multiply by -i (forward) or +i (inverse)
template <bool INV>
__host__ __device__ __forceinline__ float2 mul_mi(float2 a) {
    return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

// ---------------------------------------------------------------- DFT(R) --
// In-register DFT of R points; forward uses exp(-2 pi i/R), inverse exp(+).
template <int R, bool INV>
struct Dft;

template <bool INV>
struct Dft<2, INV> {
    __host__ __device__ static __forceinline__ void run(float2* v, const float2*, int) {
        float2 a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    }
};

template <bool INV>
struct Dft<3, INV> {
    __host__ __device__ static __forceinline__ void run(float2* v, const float2*, int) {
        const float c = -0.5f;
        const float s = 0.86602540378443864676f;  // sin(2 pi/3)
        float2 a = v[0], b = v[1], d = v[2];
        float2 t = cadd(b, d);
        float2 m = cadd(a, cscale(t, c));
        float2 q = cscale(csub(b, d), s);  // X1 = m - i q (fwd), X2 = m + i q
        v[0] = cadd(a, t);
        v[1] = cadd(m, mul_mi<INV>(q));
        v[2] = csub(m, mul_mi<INV>(q));
    }
};

template <bool INV>
struct Dft<4, INV> {
    __host__ __device__ static __forceinline__ void run(float2* v, const float2*, int) {
        float2 a = cadd(v[0], v[2]), b = csub(v[0], v[2]);
        float2 c = cadd(v[1], v[3]), d = mul_mi<INV>(csub(v[1], v[3]));
        v[0] = cadd(a, c);
        v[1] = cadd(b, d);
        v[2] = csub(a, c);
        v[3] = csub(b, d);
    }
};

template <bool INV>
struct Dft<5, INV> {
    __host__ __device__ static __forceinline__ void run(float2* v, const float2*, int) {
        const float c1 = 0.30901699437494742410f;   // cos(2 pi/5)
        const float c2 = -0.80901699437494742410f;  // cos(4 pi/5)
        const float s1 = 0.95105651629515357212f;   // sin(2 pi/5)
        const float s2 = 0.58778525229247312917f;   // sin(4 pi/5)
        float2 x0 = v[0];
        float2 t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
        float2 t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
        float2 a1 = cadd(x0, cadd(cscale(t1, c1), cscale(t2, c2)));
        float2 a2 = cadd(x0, cadd(cscale(t1, c2), cscale(t2, c1)));
        float2 b1 = cadd(cscale(t3, s1), cscale(t4, s2));
        float2 b2 = csub(cscale(t3, s2), cscale(t4, s1));
        v[0] = cadd(x0, cadd(t1, t2));
        v[1] = cadd(a1, mul_mi<INV>(b1));
        v[4] = csub(a1, mul_mi<INV>(b1));
        v[2] = cadd(a2, mul_mi<INV>(b2));
        v[3] = csub(a2, mul_mi<INV>(b2));
    }
};

template <bool INV>
struct Dft<8, INV> {
    __host__ __device__ static __forceinline__ void run(float2* v, const float2*, int) {
        const float h = 0.70710678118654752440f;
        float2 e[4] = {v[0], v[2], v[4], v[6]};
        float2 o[4] = {v[1], v[3], v[5], v[7]};
        Dft<4, INV>::run(e, nullptr, 0);
        Dft<4, INV>::run(o, nullptr, 0);
        // W8^1 = h(1 -+ i), W8^2 = -+i, W8^3 = h(-1 -+ i)
        float2 w1 = INV ? make_float2(h * (o[1].x - o[1].y), h * (o[1].x + o[1].y))
                        : make_float2(h * (o[1].x + o[1].y), h * (o[1].y - o[1].x));
        float2 w2 = mul_mi<INV>(o[2]);
        float2 w3 = INV ? make_float2(-h * (o[3].x + o[3].y), h * (o[3].x - o[3].y))
                        : make_float2(h * (o[3].y - o[3].x), -h * (o[3].x + o[3].y));
        v[0] = cadd(e[0], o[0]);
        v[4] = csub(e[0], o[0]);
        v[1] = cadd(e[1], w1);
        v[5] = csub(e[1], w1);
        v[2] = cadd(e[2], w2);
        v[6] = csub(e[2], w2);
        v[3] = cadd(e[3], w3);
        v[7] = csub(e[3], w3);
    }
};

// Odd prime radix R: pairs x_r and x_{R-r} share cos and (negated) sin, so
//   X_q     = A_q -+ i B_q,   X_{R-q} = A_q +- i B_q   (forward / inverse)
//   A_q = x0 + sum_r (x_r + x_{R-r}) cos(2 pi r q/R)
//   B_q =      sum_r (x_r - x_{R-r}) sin(2 pi r q/R),   r, q = 1..(R-1)/2,
// with the cosines and sines as compile-time constants.
template <int R>
struct PrimeTable;
template <>
struct PrimeTable<7> {
    static constexpr float c[7] = {1.0f, 0.62348980185873359439f, -0.22252093395631433737f,
                                   -0.90096886790241903498f, -0.90096886790241914600f,
                                   -0.22252093395631458717f, 0.62348980185873337234f};
    static constexpr float s[7] = {0.0f, 0.78183148246802980363f, 0.97492791218182361934f,
                                   0.43388373911755823142f, -0.43388373911755800938f,
                                   -0.97492791218182361934f, -0.78183148246802991466f};
};
template <>
struct PrimeTable<11> {
    static constexpr float c[11] = {1.0f, 0.84125353283118120551f, 0.41541501300188643508f,
                                    -0.14231483827328500480f, -0.65486073394528498959f,
                                    -0.95949297361449736865f, -0.95949297361449747967f,
                                    -0.65486073394528521163f, -0.14231483827328522684f,
                                    0.41541501300188604651f, 0.84125353283118120551f};
    static constexpr float s[11] = {0.0f, 0.54064081745559755543f, 0.90963199535451833011f,
                                    0.98982144188093279524f, 0.75574957435425826890f,
                                    0.28173255684142967104f, -0.28173255684142939348f,
                                    -0.75574957435425815788f, -0.98982144188093268422f,
                                    -0.90963199535451855215f, -0.54064081745559744441f};
};
template <>
struct PrimeTable<13> {
    static constexpr float c[13] = {1.0f, 0.88545602565320991051f, 0.56806474673115592289f,
                                    0.12053668025532300601f, -0.35460488704253545489f,
                                    -0.74851074817110119231f, -0.97094181742605201180f,
                                    -0.97094181742605212282f, -0.74851074817110130333f,
                                    -0.35460488704253589898f, 0.12053668025532320029f,
                                    0.56806474673115481266f, 0.88545602565321002153f};
    static constexpr float s[13] = {0.0f, 0.46472317204376850652f, 0.82298386589365635224f,
                                    0.99270887409805397272f, 0.93501624268541483342f,
                                    0.66312265824079519305f, 0.23931566428755768339f,
                                    -0.23931566428755743359f, -0.66312265824079497101f,
                                    -0.93501624268541472240f, -0.99270887409805397272f,
                                    -0.82298386589365701838f, -0.46472317204376839550f};
};

template <int R, bool INV>
struct DftOddPrime {
    __host__ __device__ static __forceinline__ void run(float2* v, const float2*, int) {
        constexpr int M = (R - 1) / 2;
        float2 sp[M], sm[M];
#pragma unroll
        for (int r = 1; r <= M; ++r) {
            sp[r - 1] = cadd(v[r], v[R - r]);
            sm[r - 1] = csub(v[r], v[R - r]);
        }
        const float2 x0 = v[0];
        float2 tot = x0;
#pragma unroll
        for (int r = 0; r < M; ++r) tot = cadd(tot, sp[r]);
#pragma unroll
        for (int q = 1; q <= M; ++q) {
            float2 A = x0, B = make_float2(0.0f, 0.0f);
#pragma unroll
            for (int r = 1; r <= M; ++r) {
                const float c = PrimeTable<R>::c[(r * q) % R];
                const float s = PrimeTable<R>::s[(r * q) % R];
                A = make_float2(A.x + sp[r - 1].x * c, A.y + sp[r - 1].y * c);
                B = make_float2(B.x + sm[r - 1].x * s, B.y + sm[r - 1].y * s);
            }
            // -i*B for the forward transform, +i*B for the inverse
            const float2 jb = mul_mi<INV>(B);
            v[q] = cadd(A, jb);
            v[R - q] = csub(A, jb);
        }
        v[0] = tot;
    }
};
template <bool INV> struct Dft<7, INV> : DftOddPrime<7, INV> {};
template <bool INV> struct Dft<11, INV> : DftOddPrime<11, INV> {};
template <bool INV> struct Dft<13, INV> : DftOddPrime<13, INV> {};

// --------------------------------------------------------- index helpers --
// Exact q = j / d for j, d < 2^16 via m = ceil(2^32 / d).
__host__ __device__ inline uint64_t magic_for(uint32_t d) {
    return ((uint64_t(1) << 32) + d - 1) / d;
}
__host__ __device__ __forceinline__ uint32_t fastdiv(uint32_t j, uint64_t m) {
    return (uint32_t)(((uint64_t)j * m) >> 32);
}

// Geometry of one stage.  COLFAST: consecutive threads take consecutive
// sequences (column tiles, LOG2C = log2(ncols)); otherwise consecutive
// threads take consecutive butterflies of one sequence (rows).
struct StageGeom {
    int n, ncols, es, cs;
    int nb;           // n / R
    int Ns;           // product of radices already applied
    int tstep;        // n / (Ns * R)
    uint64_t m_nb;    // magic for / nb
    uint64_t m_Ns;    // magic for / Ns
};

__host__ __device__ inline StageGeom make_geom(int n, int ncols, int es, int cs, int R, int Ns) {
    StageGeom g;
    g.n = n; g.ncols = ncols; g.es = es; g.cs = cs;
    g.nb = n / R; g.Ns = Ns; g.tstep = n / (Ns * R);
    g.m_nb = magic_for((uint32_t)g.nb);
    g.m_Ns = magic_for((uint32_t)Ns);
    return g;
}

// Stage geometry from a plan's precomputed per-stage values (no divisions).
__host__ __device__ __forceinline__ StageGeom plan_geom(const Plan& P, int s, int ncols, int es, int cs) {
    StageGeom g;
    const int R = P.radix[s];
    g.n = P.n; g.ncols = ncols; g.es = es; g.cs = cs;
    g.Ns = P.ns[s];
    g.nb = P.n / R;
    g.tstep = g.nb / g.Ns;
    g.m_nb = P.m_nb[s];
    g.m_Ns = P.m_ns[s];
    return g;
}

template <bool COLFAST, int LOG2C>
__host__ __device__ __forceinline__ void split_b(int b, const StageGeom& g, int& j, int& c) {
    if (COLFAST) {
        c = b & ((1 << LOG2C) - 1);
        j = b >> LOG2C;
    } else {
        c = (int)fastdiv((uint32_t)b, g.m_nb);
        j = b - c * g.nb;
    }
}

// One butterfly of a stage, out of place: reads its R inputs from src,
// applies the twiddles and a radix-R DFT, writes its R outputs to their
// Stockham positions in dst.  Registers: the R values only.
template <int R, bool COLFAST, int LOG2C, bool INV>
__host__ __device__ __forceinline__ void stage_bfly(const float2* src, float2* dst, int b, const StageGeom& g,
                                                    const float2* tw) {
    int j, c;
    split_b<COLFAST, LOG2C>(b, g, j, c);
    const float2* p = src + j * g.es + c * g.cs;
    float2 v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = p[r * g.nb * g.es];
    const int grp = (int)fastdiv((uint32_t)j, g.m_Ns);
    const int k = j - grp * g.Ns;
    if (k != 0) {
        const int ts = k * g.tstep;
#pragma unroll
        for (int r = 1; r < R; ++r) {
            float2 w = tw[r * ts];
            if (INV) w = cconj(w);
            v[r] = cmul(v[r], w);
        }
    }
    Dft<R, INV>::run(v, tw, g.n);
    float2* q = dst + (grp * g.Ns * R + k) * g.es + c * g.cs;
#pragma unroll
    for (int r = 0; r < R; ++r) q[r * g.Ns * g.es] = v[r];
}

// Thread tid's share of a stage: butterflies tid, tid + kThreads, ...  The
// stage reads only src and writes only dst, so the threads need no barrier
// inside it, one after it.
template <int R, bool COLFAST, int LOG2C, bool INV>
__host__ __device__ __forceinline__ void stage_pp(const float2* src, float2* dst, int tid, const StageGeom& g,
                                                  const float2* tw) {
    const int total = g.nb * g.ncols;
    // butterflies in flight per thread: 4 on column tiles, 2 on rows
    // (measured), 1 for radices 11 and 13, whose register footprint would
    // otherwise set the whole kernel's
    constexpr int U = R > 8 ? 1 : (COLFAST ? 4 : 2);
#pragma unroll U
    for (int b = tid; b < total; b += kThreads) stage_bfly<R, COLFAST, LOG2C, INV>(src, dst, b, g, tw);
}

// Stage s of plan P from src to dst (the radix chosen at run time).
template <bool COLFAST, int LOG2C, bool INV>
__host__ __device__ inline void plan_stage(const float2* src, float2* dst, int tid, const Plan& P, int s, int ncols,
                                           int es, int cs, const float2* tw) {
    const StageGeom g = plan_geom(P, s, ncols, es, cs);
    switch (P.radix[s]) {
        case 2: stage_pp<2, COLFAST, LOG2C, INV>(src, dst, tid, g, tw); break;
        case 3: stage_pp<3, COLFAST, LOG2C, INV>(src, dst, tid, g, tw); break;
        case 4: stage_pp<4, COLFAST, LOG2C, INV>(src, dst, tid, g, tw); break;
        case 5: stage_pp<5, COLFAST, LOG2C, INV>(src, dst, tid, g, tw); break;
        case 7: stage_pp<7, COLFAST, LOG2C, INV>(src, dst, tid, g, tw); break;
        case 8: stage_pp<8, COLFAST, LOG2C, INV>(src, dst, tid, g, tw); break;
        case 11: stage_pp<11, COLFAST, LOG2C, INV>(src, dst, tid, g, tw); break;
        case 13: stage_pp<13, COLFAST, LOG2C, INV>(src, dst, tid, g, tw); break;
        default: break;
    }
}

// ------------------------------------------------------------ Bluestein --
// Chirp-z transform of length n through a circular convolution of length M
// (M >= 2n - 1, a length the stages handle).  With c[j] = exp(-i pi j^2/n),
// j k = (j^2 + k^2 - (k - j)^2) / 2 gives
//   X[k] = sum_j x[j] exp(-2 pi i j k / n) = c[k] sum_j (x[j] c[j]) conj(c[k - j]),
// so X = c . IFFT_M(FFT_M(x c, zero-padded) . B) / M with B = FFT_M of
// conj(c[m]) placed at m and M - m (Bluestein 1970).  The bank holds c (n
// entries) and B / M (M entries); the inverse transform is
// conj(forward(conj(x))).  Element (e, c) is at [e*es + c*cs] in the
// caller's layout (e < n) and at [e*wes + c*wcs] while working (e < M):
// column tiles keep their layout with M rows, rows are respaced to stride M.
// Each phase is element-wise from one LDS buffer to the other (or in place),
// a per-thread function the host emulation runs as well.
struct BlGeom {
    int n, M, ncols, es, cs, wes, wcs;
};
__host__ __device__ inline BlGeom bl_geom(const Plan& P, int ncols, int es, int cs, bool colfast) {
    BlGeom g;
    g.n = P.bl_n; g.M = P.n; g.ncols = ncols; g.es = es; g.cs = cs;
    g.wes = colfast ? es : 1;
    g.wcs = colfast ? cs : P.n;
    return g;
}
// Element q of thread tid's share of an ncols x len array: (c, e) with
// consecutive threads on consecutive columns of a tile (COLFAST, ncols =
// 2^LOG2C) or on consecutive elements of a row (rows walked one by one).
// Calls f(c, e) for every element of the array once over all threads.
template <bool COLFAST, int LOG2C, class F>
__host__ __device__ __forceinline__ void bl_each(int tid, int ncols, int len, F&& f) {
    if (COLFAST) {
        for (int idx = tid; idx < (len << LOG2C); idx += kThreads) f(idx & ((1 << LOG2C) - 1), idx >> LOG2C);
    } else {
        for (int c = 0; c < ncols; ++c)
            for (int e = tid; e < len; e += kThreads) f(c, e);
    }
}
// x c (conj(x) c for the inverse) from src (caller's layout) into dst
// (working layout), zeros at n <= e < M
template <bool COLFAST, int LOG2C, bool INV>
__host__ __device__ inline void bl_in(const float2* src, float2* dst, int tid, const BlGeom& g, const float2* chirp) {
    bl_each<COLFAST, LOG2C>(tid, g.ncols, g.M, [&](int c, int e) {
        float2 v = make_float2(0.0f, 0.0f);
        if (e < g.n) {
            const float2 x = src[e * g.es + c * g.cs];
            v = cmul(INV ? cconj(x) : x, chirp[e]);
        }
        dst[e * g.wes + c * g.wcs] = v;
    });
}
// times the filter B / M, in place (working layout)
template <bool COLFAST, int LOG2C>
__host__ __device__ inline void bl_filter(float2* buf, int tid, const BlGeom& g, const float2* filt) {
    bl_each<COLFAST, LOG2C>(tid, g.ncols, g.M, [&](int c, int e) {
        float2* p = buf + e * g.wes + c * g.wcs;
        *p = cmul(*p, filt[e]);
    });
}
// c . conv (conjugated for the inverse) from src (working layout) into dst
// (caller's layout)
template <bool COLFAST, int LOG2C, bool INV>
__host__ __device__ inline void bl_out(const float2* src, float2* dst, int tid, const BlGeom& g,
                                       const float2* chirp) {
    bl_each<COLFAST, LOG2C>(tid, g.ncols, g.n, [&](int c, int e) {
        const float2 v = cmul(src[e * g.wes + c * g.wcs], chirp[e]);
        dst[e * g.es + c * g.cs] = INV ? cconj(v) : v;
    });
}

#ifdef __HIPCC__
// The runtime-plan stages of P from src to dst and back, one barrier after
// each; returns the buffer that holds the result.
template <bool COLFAST, int LOG2C, bool INV>
__device__ __forceinline__ float2* plan_stages(float2* src, float2* dst, const Plan& P, int ncols, int es, int cs,
                                               const float2* tw) {
    for (int s = 0; s < P.nstages; ++s) {
        plan_stage<COLFAST, LOG2C, INV>(src, dst, threadIdx.x, P, s, ncols, es, cs, tw);
        __syncthreads();
        float2* t = src;
        src = dst;
        dst = t;
    }
    return src;
}

// Full transform of the ncols sequences in buffer a on the device, buffer b
// as the other half of each stage (both hold ncols * n, or ncols * M for a
// chirp-z length, elements).  inner(src, dst, inv, es, cs) runs a P.n-point
// transform (inv: std::true_type / std::false_type) and returns the buffer
// that holds its result.  Caller must have synchronised after filling a;
// returns after a final barrier with the result in a.
template <bool COLFAST, int LOG2C, bool INV, class Inner>
__device__ __forceinline__ void lds_fft_any(float2* a, float2* b, const Plan& P, int ncols, int es, int cs,
                                            const float2* twg, Inner&& inner) {
    const int tid = threadIdx.x;
    float2* src;
    int n = P.n;
    if (P.bl_n) {
        const BlGeom g = bl_geom(P, ncols, es, cs, COLFAST);
        const float2* chirp = twg + P.bl_chirp;
        n = g.n;
        bl_in<COLFAST, LOG2C, INV>(a, b, tid, g, chirp);
        __syncthreads();
        float2* r = inner(b, a, std::false_type{}, g.wes, g.wcs);
        bl_filter<COLFAST, LOG2C>(r, tid, g, twg + P.bl_filt);
        __syncthreads();
        r = inner(r, r == a ? b : a, std::true_type{}, g.wes, g.wcs);
        src = r == a ? b : a;
        bl_out<COLFAST, LOG2C, INV>(r, src, tid, g, chirp);
        __syncthreads();
    } else {
        src = inner(a, b, std::integral_constant<bool, INV>{}, es, cs);
    }
    if (src != a) {
        // the caller's region is contiguous: ncols rows of n (es 1, cs n)
        // or n rows of ncols (es ncols, cs 1)
        for (int i = tid; i < n * ncols; i += kThreads) a[i] = src[i];
        __syncthreads();
    }
}
#endif

// --------------------------------------------------------- host planning --
// Radix sequence for length n: 8s first, then 4, 2, 3, 5, 7, 11, 13.
// Returns false if n has a prime factor outside {2,3,5,7,11,13} or needs
// more than kMaxStages stages.
inline bool factorize(int n, int* radix, int* nstages) {
    int m = n, s = 0;
    if (n < 1) return false;
    auto push = [&](int r) {
        if (s >= kMaxStages) return false;
        radix[s++] = r;
        m /= r;
        return true;
    };
    while (m % 8 == 0) if (!push(8)) return false;
    while (m % 4 == 0) if (!push(4)) return false;
    while (m % 2 == 0) if (!push(2)) return false;
    const int odd[] = {3, 5, 7, 11, 13};
    for (int p : odd)
        while (m % p == 0) if (!push(p)) return false;
    if (m != 1) return false;
    *nstages = s;
    return true;
}

// Fill a plan's radix sequence and per-stage geometry for length n.
inline bool make_plan_stages(int n, Plan& p) {
    p.n = n;
    if (!factorize(n, p.radix, &p.nstages)) return false;
    int ns = 1;
    for (int s = 0; s < p.nstages; ++s) {
        p.ns[s] = ns;
        p.m_nb[s] = magic_for((uint32_t)(n / p.radix[s]));
        p.m_ns[s] = magic_for((uint32_t)ns);
        ns *= p.radix[s];
    }
    return true;
}

}  // namespace qsmfft
