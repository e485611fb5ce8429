// Field solves of the xi step (lcode.py:122-364) on gfx950.
//
// Every solve is out = T_x T_y (Mat o (T_y T_x rhs)) with T_x, T_y type-1
// sine or cosine transforms along x (i) and y (j).  The transforms commute,
// so a solve runs as three passes over lines, batched over all fields:
//
//   pass A  one RHS row i per line: assemble the RHS on the fly, transform
//           along j, store transposed: T[k][i]
//   pass B  one row k of T per line: transform along i, multiply by the
//           eigenvalue table, transform along i again, store back in place
//   pass C  one output row i per line: gather the column T[.][i], transform
//           along j, then the epilogue (Bz mean, Variant A, averaging)
//
// Each type-1 transform of length n is one complex FFT of length
// L = 2 (n + 1) (sine) or 2 (n - 1) (cosine) of the odd or even extension;
// two real lines ride in one complex FFT (real and imaginary parts).  All
// six fields of a step share L = 2 (N - 1).  The FFT is a mixed-radix
// (4, 2, 3, 5) Stockham transform in LDS.
//
// The transforms, the spectra T and the eigenvalue tables are in S = double
// (plain path) or S = float (mixed-precision path); the RHS assembly, the
// Bz mean and the epilogue are FP64 in both.

#include "common.hpp"

#include <type_traits>

namespace lcode {

namespace {

template <typename S> struct Vec;
template <> struct Vec<double> { using T = double2; };
template <> struct Vec<float> { using T = float2; };

template <typename V>
__device__ __forceinline__ V mk(decltype(V::x) x, decltype(V::x) y) {
    V v;
    v.x = x;
    v.y = y;
    return v;
}
template <typename V>
__device__ __forceinline__ V cmul(V a, V b) {
    return mk<V>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename V>
__device__ __forceinline__ V cadd(V a, V b) { return mk<V>(a.x + b.x, a.y + b.y); }
template <typename V>
__device__ __forceinline__ V csub(V a, V b) { return mk<V>(a.x - b.x, a.y - b.y); }
// multiply by -i
template <typename V>
__device__ __forceinline__ V mul_mi(V a) { return mk<V>(a.y, -a.x); }

// LDS layout of a pair's L complex points.  A 16-B point p sits in slot
// p mod 16 of its 256-B group (all LDS banks once); the radix-4 stages
// with span 1 and 4 store 16 lanes at points 4b + q and 16 (b / 4) +
// (b mod 4) + 4 q, which hit only 4 slots.  XOR-ing the slot with
// 5 ((p / 16) mod 4) (the group's low bits into both halves of the slot)
// spreads both over the 16 slots, and leaves the consecutive accesses of
// the other stages, the loads and the stores distinct (their group bits
// are constant across 16 lanes).  Measured at 641^2 (FP64): pass B 43.2 us
// with the XOR layout, 41.2 us with the identity layout (0, the default):
// these conflicts are not what bounds the passes.
#ifndef LCODE_FFT_SWIZZLE
#define LCODE_FFT_SWIZZLE 0
#endif
__host__ __device__ constexpr int zi(int e) {
    return LCODE_FFT_SWIZZLE ? (e & ~15) | ((e & 15) ^ (5 * ((e >> 4) & 3))) : e;
}
__host__ __device__ constexpr int padded_len(int L) {
    return LCODE_FFT_SWIZZLE ? (L + 15) & ~15 : L;
}

// Forward DFT of R points held in registers: y_p = sum_q x_q e^{-2 pi i pq/R}.
template <int R, typename V> struct Dft;

template <typename V> struct Dft<2, V> {
    static __device__ __forceinline__ void run(V* v) {
        const V a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    }
};

template <typename V> struct Dft<4, V> {
    static __device__ __forceinline__ void run(V* v) {
        const V s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
        const V s13 = cadd(v[1], v[3]), d13 = mul_mi(csub(v[1], v[3]));
        v[0] = cadd(s02, s13);
        v[2] = csub(s02, s13);
        v[1] = cadd(d02, d13);
        v[3] = csub(d02, d13);
    }
};

// Radix 8 as two radix-4 DFTs of the even and odd points combined with
// the eighth roots of unity: y_p = e_p + w8^p o_p, y_(p+4) = e_p - w8^p o_p.
template <typename V> struct Dft<8, V> {
    static __device__ __forceinline__ void run(V* v) {
        using S = decltype(V::x);
        constexpr S r = (S)0.70710678118654752440;   // 1 / sqrt(2)
        V e[4] = {v[0], v[2], v[4], v[6]};
        V o[4] = {v[1], v[3], v[5], v[7]};
        Dft<4, V>::run(e);
        Dft<4, V>::run(o);
        const V o1 = mk<V>(r * (o[1].x + o[1].y), r * (o[1].y - o[1].x));    // (1 - i) / sqrt 2
        const V o2 = mul_mi(o[2]);                                           // -i
        const V o3 = mk<V>(r * (o[3].y - o[3].x), -r * (o[3].x + o[3].y));   // -(1 + i) / sqrt 2
        v[0] = cadd(e[0], o[0]);
        v[4] = csub(e[0], o[0]);
        v[1] = cadd(e[1], o1);
        v[5] = csub(e[1], o1);
        v[2] = cadd(e[2], o2);
        v[6] = csub(e[2], o2);
        v[3] = cadd(e[3], o3);
        v[7] = csub(e[3], o3);
    }
};

template <typename V> struct Dft<3, V> {
    static __device__ __forceinline__ void run(V* v) {
        using S = decltype(V::x);
        constexpr S s = (S)0.86602540378443864676;  // sin(2 pi / 3)
        const V t = cadd(v[1], v[2]), d = csub(v[1], v[2]);
        const V base = mk<V>(v[0].x - (S).5 * t.x, v[0].y - (S).5 * t.y);
        const V rot = mk<V>(s * d.y, -s * d.x);     // -i s d
        v[0] = cadd(v[0], t);
        v[1] = cadd(base, rot);
        v[2] = csub(base, rot);
    }
};

template <typename V> struct Dft<5, V> {
    static __device__ __forceinline__ void run(V* v) {
        using S = decltype(V::x);
        constexpr S c1 = (S)0.30901699437494742410;   // cos(2 pi / 5)
        constexpr S c2 = (S)-0.80901699437494742410;  // cos(4 pi / 5)
        constexpr S s1 = (S)0.95105651629515357212;   // sin(2 pi / 5)
        constexpr S s2 = (S)0.58778525229247312917;   // sin(4 pi / 5)
        const V x0 = v[0];
        const V t1 = cadd(v[1], v[4]), d1 = csub(v[1], v[4]);
        const V t2 = cadd(v[2], v[3]), d2 = csub(v[2], v[3]);
        const V a1 = mk<V>(x0.x + c1 * t1.x + c2 * t2.x, x0.y + c1 * t1.y + c2 * t2.y);
        const V a2 = mk<V>(x0.x + c2 * t1.x + c1 * t2.x, x0.y + c2 * t1.y + c1 * t2.y);
        const V b1 = mk<V>(s1 * d1.x + s2 * d2.x, s1 * d1.y + s2 * d2.y);
        const V b2 = mk<V>(s2 * d1.x - s1 * d2.x, s2 * d1.y - s1 * d2.y);
        v[0] = cadd(x0, cadd(t1, t2));
        v[1] = cadd(a1, mul_mi(b1));
        v[4] = csub(a1, mul_mi(b1));
        v[2] = cadd(a2, mul_mi(b2));
        v[3] = csub(a2, mul_mi(b2));
    }
};

// ---------------------------------------------------------------------------
// Runtime plan (any L = 2^a 3^b 5^c): twiddles from a quarter-wave cosine
// table ct[r] = cos(2 pi r / L), r in [0, L/4], held in LDS.
// ---------------------------------------------------------------------------

template <typename V>
__device__ __forceinline__ V twiddle(const decltype(V::x)* ct, int t, int L) {
    using S = decltype(V::x);
    const int Q = L >> 2;
    const int quad = t / Q, r = t - quad * Q;
    const S c = ct[r], s = ct[Q - r];   // cos and sin of 2 pi r / L
    S cs, sn;                           // of 2 pi t / L = quad pi / 2 + ...
    switch (quad) {
        case 0: cs = c; sn = s; break;
        case 1: cs = -s; sn = c; break;
        case 2: cs = -c; sn = -s; break;
        default: cs = s; sn = -c; break;
    }
    return mk<V>(cs, -sn);
}

// One radix-R Stockham stage over the L points of this thread's pair.
// Thread t of Tp handles butterflies t and t + Tp (Tp >= L / 4 makes two
// enough for every radix >= 2).  span = product of the previous radices.
template <int R, typename V>
__device__ __forceinline__ void fft_stage(V* z, const decltype(V::x)* ct,
                                          int L, int span, int t, int Tp) {
    const int T = L / R;
    const int tw = L / (span * R);
    V v[2][R];
#pragma unroll
    for (int it = 0; it < 2; ++it) {
        const int b = t + it * Tp;
        if (b < T) {
            const int k = b % span;
#pragma unroll
            for (int q = 0; q < R; ++q) {
                V x = z[zi(b + q * T)];
                if (q && k) x = cmul(x, twiddle<V>(ct, q * k * tw, L));
                v[it][q] = x;
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < 2; ++it) {
        const int b = t + it * Tp;
        if (b < T) {
            Dft<R, V>::run(v[it]);
            const int k = b % span;
            const int j = (b - k) * R + k;
#pragma unroll
            for (int q = 0; q < R; ++q) z[zi(j + q * span)] = v[it][q];
        }
    }
    __syncthreads();
}

// In-place forward FFT of every pair's buffer (block-wide barriers inside:
// every thread of the block must call it).
template <typename V>
__device__ void fft_lds(V* z, const FftPlan& plan, const decltype(V::x)* ct,
                        int t, int Tp) {
    int span = 1;
    for (int s = 0; s < plan.nstages; ++s) {
        const int R = plan.radix[s];
        switch (R) {
            case 4: fft_stage<4>(z, ct, plan.L, span, t, Tp); break;
            case 2: fft_stage<2>(z, ct, plan.L, span, t, Tp); break;
            case 3: fft_stage<3>(z, ct, plan.L, span, t, Tp); break;
            default: fft_stage<5>(z, ct, plan.L, span, t, Tp); break;
        }
        span *= R;
    }
}

// ---------------------------------------------------------------------------
// Compile-time plans for the common transform lengths: the radix of each
// stage and its span are constants, so the index arithmetic folds away.
// A stage with span s and radix R needs the twiddles w^q, 1 <= q < R, of
// w = exp(-2 pi i k / (s R)) for k < s: the table holds w alone, at
// tw[tw_offset(L, s) + k] (the spans of the earlier stages summed; 341
// entries for L = 1280 instead of L - 1 for every power), and the stage
// forms w^2 = w w, w^3 = w^2 w, w^4 = w^2 w^2 in registers.
// ---------------------------------------------------------------------------

// Radix of the next static stage (build option LCODE_FFT_RADIX8: radix 8
// first where it divides; measured at 641^2: pass B 44.6 us with
// 8-8-4-5, 41.8 us with 4-4-4-4-5, so off by default).
#ifndef LCODE_FFT_RADIX8
#define LCODE_FFT_RADIX8 0
#endif
__host__ __device__ constexpr int next_radix(int rem) {
    return (LCODE_FFT_RADIX8 && rem % 8 == 0) ? 8
           : rem % 4 == 0 ? 4 : rem % 2 == 0 ? 2 : rem % 3 == 0 ? 3 : 5;
}

__host__ __device__ constexpr int tw_offset(int L, int span) {
    int off = 0;
    for (int sp = 1; sp < span && sp < L; sp *= next_radix(L / sp)) off += sp;
    return off;
}

__host__ __device__ constexpr int tw_count(int L) { return tw_offset(L, L); }

// Threads of one line pair: enough for the L / 4 butterflies of a radix-4
// stage (the runtime plans need it: two butterflies per thread at most).
__host__ __device__ constexpr int threads_per_pair(int L) {
    return ((L + 3) / 4 + 63) / 64 * 64;
}

// LCODE_FFT_PAIR_WAVES_F32 / _F64 (0: threads_per_pair): waves per line
// pair of the static plans; each thread then takes up to two butterflies of
// a stage.  Kept at or above four input points per thread and three
// twiddle-table entries per thread.  Fewer waves per pair let more pairs
// share a CU's 32 wave slots (at 641^2, 5 waves per pair leave 966 blocks
// for 768 slots); measured at 641^2, six-field solve: FP64 92.6 us with
// 5 waves, 84.2 with 4; FP32 83.6 with 5, 72.0 with 4, 70.9 with 3 (two
// pairs per block); with one pair per block FP64 76.8 / 71.9 / 77.7 us with
// 5 / 4 / 3 waves, FP32 60.6 / 59.3 us with 4 / 3.
#ifndef LCODE_FFT_PAIR_WAVES_F32
#define LCODE_FFT_PAIR_WAVES_F32 3
#endif
#ifndef LCODE_FFT_PAIR_WAVES_F64
#define LCODE_FFT_PAIR_WAVES_F64 4
#endif
__host__ __device__ constexpr int static_threads_per_pair(int L, bool fp32) {
    const int d = threads_per_pair(L);
    const int w = fp32 ? LCODE_FFT_PAIR_WAVES_F32 : LCODE_FFT_PAIR_WAVES_F64;
    if (w <= 0) return d;
    const int lo = ((L / 2 + 1 + 3) / 4 + 63) / 64 * 64;   // <= 4 input points
    int t = 64 * w;
    if (t < lo) t = lo;
    if (t > d) t = d;
    if (2 * t < (L + 3) / 4 || 3 * t < tw_count(L)) t = d;
    return t;
}

// Input points per thread of a pair: a transform input has at most
// L / 2 + 1 points.
__host__ __device__ constexpr int points_per_thread(int L, bool fp32) {
    return (L / 2 + 1 + static_threads_per_pair(L, fp32) - 1) /
           static_threads_per_pair(L, fp32);
}

template <int L, int SPAN, typename V>
__device__ __forceinline__ void fft_static(V* z, const double2* __restrict__ tw,
                                           int t) {
    if constexpr (SPAN < L) {
        constexpr int R = next_radix(L / SPAN);
        constexpr int T = L / R;
        constexpr int Tp = static_threads_per_pair(L, sizeof(V) == sizeof(float2));
        constexpr int ITER = (T + Tp - 1) / Tp;
        constexpr int OFF = tw_offset(L, SPAN);
        V v[ITER][R];
#pragma unroll
        for (int it = 0; it < ITER; ++it) {
            const int b = t + it * Tp;
            if (b < T) {
                const int k = b % SPAN;
                // the root in FP64 and its powers formed in FP64, each
                // rounded once to the transform's precision
                double2 w[R];
                if constexpr (SPAN > 1) {
                    w[1] = tw[OFF + k];
                    if constexpr (R > 2) w[2] = cmul(w[1], w[1]);
                    if constexpr (R > 3) w[3] = cmul(w[2], w[1]);
                    if constexpr (R > 4) w[4] = cmul(w[2], w[2]);
                    if constexpr (R > 5) w[5] = cmul(w[4], w[1]);
                    if constexpr (R > 6) w[6] = cmul(w[4], w[2]);
                    if constexpr (R > 7) w[7] = cmul(w[4], w[3]);
                }
#pragma unroll
                for (int q = 0; q < R; ++q) {
                    V x = z[zi(b + q * T)];
                    if (SPAN > 1 && q) x = cmul(x, mk<V>(w[q].x, w[q].y));
                    v[it][q] = x;
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int it = 0; it < ITER; ++it) {
            const int b = t + it * Tp;
            if (b < T) {
                Dft<R, V>::run(v[it]);
                const int k = b % SPAN;
                const int j = (b - k) * R + k;
#pragma unroll
                for (int q = 0; q < R; ++q) z[zi(j + q * SPAN)] = v[it][q];
            }
        }
        __syncthreads();
        fft_static<L, SPAN * R>(z, tw, t);
    }
}

// Twiddle source of a block: for L == 0 (runtime plan) the quarter-wave
// cosine table, copied to LDS; for a static L the stage twiddle roots (FP64
// in both precisions), copied to LDS when TW_LDS, else read from global
// memory.
template <typename S>
struct Tw {
    const S* ct;
    const double2* tw;
};

template <typename S> struct TwSrc;
template <> struct TwSrc<double> {
    static __device__ const double* ct(const SolverGeom& g) { return g.costab; }
};
template <> struct TwSrc<float> {
    static __device__ const float* ct(const SolverGeom& g) { return g.costab32; }
};

template <int L, bool TW_LDS, typename S>
__device__ __forceinline__ Tw<S> load_twiddles(typename Vec<S>::T* lds,
                                               const SolverGeom& g) {
    using V = typename Vec<S>::T;
    Tw<S> r{nullptr, nullptr};
    if constexpr (L == 0) {
        S* ct = reinterpret_cast<S*>(lds + (size_t)g.P * padded_len(g.fft.L));
        const S* src = TwSrc<S>::ct(g);
        for (int k = threadIdx.x; k <= (g.fft.L >> 2); k += blockDim.x) ct[k] = src[k];
        r.ct = ct;
    } else if constexpr (TW_LDS) {
        // at most three entries per thread (static_threads_per_pair keeps
        // P * Tp >= tw_count / 3), all loaded before the LDS stores
        constexpr int n = tw_count(L);
        static_assert(n <= 3 * static_threads_per_pair(L, std::is_same<S, float>::value),
                      "twiddle table larger than three per thread");
        // (the data end is 16-byte aligned: an even number of points)
        double2* d = reinterpret_cast<double2*>(lds + (size_t)g.P * padded_len(L));
        const double2* src = g.stage_tw;
        const int nt = blockDim.x;
        double2 v[3];
#pragma unroll
        for (int u = 0; u < 3; ++u) v[u] = src[min((int)threadIdx.x + u * nt, n - 1)];
#pragma unroll
        for (int u = 0; u < 3; ++u) {
            const int k = threadIdx.x + u * nt;
            if (k < n) d[k] = v[u];
        }
        r.tw = d;
    } else {
        r.tw = g.stage_tw;
    }
    return r;
}

template <int L, typename S>
__device__ __forceinline__ void fft_any(typename Vec<S>::T* z,
                                        const SolverGeom& g, const Tw<S>& tw,
                                        int t) {
    if constexpr (L == 0) {
        fft_lds(z, g.fft, tw.ct, t, g.Tp);
    } else {
        fft_static<L, 1>(z, tw.tw, t);
    }
}

// Extension writers/readers.  Cosine (DCT-I, n points): z[k] and z[L-k]
// hold input k.  Sine (DST-I, n points): z[k+1] = input k, z[L-1-k] =
// -input k, z[0] = z[n+1] = 0.  For a pair (a real, b imaginary) the FFT
// gives DCT(a) = Re Z[k], DCT(b) = Im Z[k]; DST(a) = -Im Z[k+1],
// DST(b) = Re Z[k+1].
template <typename V>
__device__ __forceinline__ void put_ext(V* z, int xf, int n, int len, int L,
                                        double a, double b) {
    using S = decltype(V::x);
    const V v = mk<V>((S)a, (S)b);
    if (xf == XF_DCT) {
        z[zi(n)] = v;
        if (n >= 1 && n <= len - 2) z[zi(L - n)] = v;
    } else {
        z[zi(n + 1)] = v;
        z[zi(L - 1 - n)] = mk<V>(-v.x, -v.y);
    }
}

template <typename V>
__device__ __forceinline__ void put_ext_zeros(V* z, int xf, int len, int t) {
    if (xf == XF_DST && t == 0) {
        z[zi(0)] = mk<V>(0, 0);
        z[zi(len + 1)] = mk<V>(0, 0);
    }
}

template <typename V>
__device__ __forceinline__ void get_ext(const V* z, int xf, int k, double& a,
                                        double& b) {
    if (xf == XF_DCT) {
        const V v = z[zi(k)];
        a = v.x;
        b = v.y;
    } else {
        const V v = z[zi(k + 1)];
        a = -v.y;
        b = v.x;
    }
}

template <typename G>
__device__ __forceinline__ double rhs_value(int kind, int i, int j, int N,
                                            const RhsInputs& in) {
    return rhs_value_dev<G>(kind, i, j, N, in);
}

// LCODE_RHS_BATCH (build option, the FP64 wave pass A only; the other pass
// A kernels keep 0, which their register budgets need): 0, each
// right-hand-side value loads and combines its inputs in one expression, and
// the compiler streams the loads (about a dozen in flight per wave); B > 0,
// the inputs of B columns of both rows are loaded (rhs_load_k) before any is
// combined (rhs_combine_k, the same arithmetic as rhs_value_k), so that up
// to 18 B loads are in flight.
#ifndef LCODE_RHS_BATCH
#define LCODE_RHS_BATCH 0
#endif

// The loads of rhs_value_k<KIND>(i, j): r[0..2] and r[3..5] the source
// (density or jz, its previous-pass value, beam) at the forward and backward
// neighbour, r[6..8] the previous-slice current, the current and the trick
// field at the point (Ex, Ey, Bx, By); r[0..3] the four current neighbours
// (Ez, Bz).
template <int KIND, typename G>
__device__ __forceinline__ void rhs_load_k(int i, int j, int N, const RhsInputs& in,
                                           double* r) {
    const size_t idx = (size_t)i * N + j;
    const bool inner = i >= 1 && i <= N - 2 && j >= 1 && j <= N - 2;
    const size_t oi = inner ? (size_t)N : 0, oj = inner ? 1 : 0;
    const G *jx = gp<G>(in.jx), *jy = gp<G>(in.jy);
    if constexpr (KIND == F_EZ) {
        r[0] = (double)jx[idx + oi];
        r[1] = (double)jx[idx - oi];
        r[2] = (double)jy[idx + oj];
        r[3] = (double)jy[idx - oj];
    } else if constexpr (KIND == F_BZ) {
        r[0] = (double)jx[idx + oj];
        r[1] = (double)jx[idx - oj];
        r[2] = (double)jy[idx + oi];
        r[3] = (double)jy[idx - oi];
    } else {
        constexpr bool rho = KIND == F_EX || KIND == F_EY;
        constexpr bool along_i = KIND == F_EX || KIND == F_BY;
        const size_t o = along_i ? oi : oj;
        const G* src = gp<G>(rho ? in.ro : in.jz);
        const double* sp0 = rho ? in.ro_prev : in.jz_prev;
        const G* prev = sp0 ? gp<G>(sp0) : src;
        const double* dummy = std::is_same<G, double>::value ? in.ro : in.zero64;
        const double* b = in.beam ? in.beam : dummy;
        r[0] = (double)src[idx + o];
        r[1] = (double)prev[idx + o];
        r[2] = b[idx + o];
        r[3] = (double)src[idx - o];
        r[4] = (double)prev[idx - o];
        r[5] = b[idx - o];
        constexpr bool xcur = KIND == F_EX || KIND == F_BY;
        const G* jp = gp<G>(xcur ? in.jx_prev : in.jy_prev);
        const G* jc = xcur ? jx : jy;
        constexpr int ft = KIND == F_EX ? 0 : KIND == F_EY ? 1 : KIND == F_BX ? 2 : 3;
        r[6] = (double)jp[idx];
        r[7] = (double)jc[idx];
        r[8] = (double)gp<G>(in.Ftrick[ft])[idx];
    }
}

// rhs_value_k<KIND>(i, j) from the values rhs_load_k loaded.
template <int KIND>
__device__ __forceinline__ double rhs_combine_k(int i, int j, int N, const RhsInputs& in,
                                                const double* r) {
    const bool inner = i >= 1 && i <= N - 2 && j >= 1 && j <= N - 2;
    const double h2 = in.h * 2;
    if constexpr (KIND == F_EZ) {
        return -((r[0] - r[1]) + (r[2] - r[3])) / h2;
    } else if constexpr (KIND == F_BZ) {
        const double djx_dy = r[0] - r[1];
        const double djy_dx = r[2] - r[3];
        return -((inner ? djx_dy : 0.) - (inner ? djy_dx : 0.)) / h2;
    } else {
        auto source = [&](double v, double p, double b) {
            const double s = (v + p) / 2;
            return s + (b * (in.beam ? 1. : 0.) + (in.beam ? 0. : in.beam_scalar));
        };
        const double d = (source(r[0], r[1], r[2]) - source(r[3], r[4], r[5])) / h2;
        const double dj = (r[6] - r[7]) / in.dxi;
        if constexpr (KIND == F_BX) return +(((inner ? d : 0.) - dj) + r[8] * in.trick);
        else return -(((inner ? d : 0.) - dj) - r[8] * in.trick);
    }
}

// Right-hand sides of rows ia and ib at columns j0 + t + u Tp, u < K, as
// straight-line code: columns clamped to [j0, j0 + len), so all loads of
// the thread are issued before their values are used.
template <int KIND, int K, typename G, int B>
__device__ __forceinline__ void rhs_rows_k(int ia, int ib, int j0, int len,
                                           int t, int Tp, int N,
                                           const RhsInputs& in, double* a,
                                           double* c) {
    if constexpr (B > 0) {
#pragma unroll
        for (int u0 = 0; u0 < K; u0 += B) {
            double ra[B][9], rc[B][9];
#pragma unroll
            for (int u = u0; u < u0 + B && u < K; ++u) {
                const int j = j0 + min(t + u * Tp, len - 1);
                rhs_load_k<KIND, G>(ia, j, N, in, ra[u - u0]);
                rhs_load_k<KIND, G>(ib, j, N, in, rc[u - u0]);
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int u = u0; u < u0 + B && u < K; ++u) {
                const int j = j0 + min(t + u * Tp, len - 1);
                a[u] = rhs_combine_k<KIND>(ia, j, N, in, ra[u - u0]);
                c[u] = rhs_combine_k<KIND>(ib, j, N, in, rc[u - u0]);
            }
        }
    } else {
#pragma unroll
        for (int u = 0; u < K; ++u) {
            const int j = j0 + min(t + u * Tp, len - 1);
            a[u] = rhs_value_k<KIND, G>(ia, j, N, in);
            c[u] = rhs_value_k<KIND, G>(ib, j, N, in);
        }
    }
}

template <int K, typename G, int B = 0>
__device__ __forceinline__ void rhs_rows(int kind, int ia, int ib, int j0,
                                         int len, int t, int Tp, int N,
                                         const RhsInputs& in, double* a,
                                         double* c) {
    switch (kind) {
        case F_EX: rhs_rows_k<F_EX, K, G, B>(ia, ib, j0, len, t, Tp, N, in, a, c); break;
        case F_EY: rhs_rows_k<F_EY, K, G, B>(ia, ib, j0, len, t, Tp, N, in, a, c); break;
        case F_BX: rhs_rows_k<F_BX, K, G, B>(ia, ib, j0, len, t, Tp, N, in, a, c); break;
        case F_BY: rhs_rows_k<F_BY, K, G, B>(ia, ib, j0, len, t, Tp, N, in, a, c); break;
        case F_EZ: rhs_rows_k<F_EZ, K, G, B>(ia, ib, j0, len, t, Tp, N, in, a, c); break;
        default: rhs_rows_k<F_BZ, K, G, B>(ia, ib, j0, len, t, Tp, N, in, a, c); break;
    }
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// XCD-aware block order.  Workgroups go to the 8 XCDs round-robin by
// linear id and each XCD has its own L2; neighbouring line blocks touch the
// same 128-B lines of the transposed spectra, so give every XCD a
// contiguous range of (field, line block) pairs.  A bijection on the grid.
struct BlockPos {
    int field, block;
};

__device__ __forceinline__ BlockPos xcd_block() {
    const int gx = gridDim.x, total = gx * gridDim.y;
    const int logical = xcd_logical(blockIdx.y * gx + blockIdx.x, total);
    return BlockPos{logical / gx, logical % gx};
}

// Pass A's order: the fields of one line block get consecutive logical
// ids, so one XCD assembles all six right-hand sides of its rows and reads
// their shared sources (densities, currents, beam) into its L2 once.
#ifndef LCODE_PASS_A_FIELDS_INTERLEAVED
#define LCODE_PASS_A_FIELDS_INTERLEAVED 1
#endif
__device__ __forceinline__ BlockPos xcd_block_rows() {
    if (!LCODE_PASS_A_FIELDS_INTERLEAVED) return xcd_block();
    const int gx = gridDim.x, nf = gridDim.y, total = gx * nf;
    const int logical = xcd_logical(blockIdx.y * gx + blockIdx.x, total);
    return BlockPos{logical % nf, logical / nf};
}

// ---------------------------------------------------------------------------
// Passes
// ---------------------------------------------------------------------------

// LCODE_PASS_A_WAVES, LCODE_PASS_C_WAVES (0: the compiler's choice): waves
// per SIMD the register allocation of passes A and C must allow.
#ifndef LCODE_PASS_A_WAVES
#define LCODE_PASS_A_WAVES 0
#endif
#ifndef LCODE_PASS_C_WAVES
#define LCODE_PASS_C_WAVES 0
#endif
#if LCODE_PASS_A_WAVES > 0
#define LCODE_PASS_A_ATTR __attribute__((amdgpu_waves_per_eu(LCODE_PASS_A_WAVES)))
#else
#define LCODE_PASS_A_ATTR
#endif
#if LCODE_PASS_C_WAVES > 0
#define LCODE_PASS_C_ATTR __attribute__((amdgpu_waves_per_eu(LCODE_PASS_C_WAVES)))
#else
#define LCODE_PASS_C_ATTR
#endif

// G: storage type of the source grids (FP64 arithmetic in both).
template <int L_, bool TW_LDS, typename S, typename G>
__global__ void __launch_bounds__(1024) LCODE_PASS_A_ATTR
pass_a_kernel(SolverGeom g, SolveBatch b) {
    using V = typename Vec<S>::T;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    V* lds = reinterpret_cast<V*>(lds_raw);
    if (!gate_open(b.gate)) return;
    const BlockPos bp = xcd_block_rows();
    const FieldPlan& fp = b.f[bp.field];
    const int Tp = g.Tp, L = g.fft.L, N = g.N;
    const int lines = fp.rowsA;
    if (bp.block * g.P * 2 >= lines) return;
    const int pair = threadIdx.x / Tp, t = threadIdx.x - pair * Tp;
    V* z = lds + (size_t)pair * padded_len(L);
    const Tw<S> tw = load_twiddles<L_, TW_LDS, S>(lds, g);
    const int ra = (bp.block * g.P + pair) * 2, rb = ra + 1;
    const bool va = ra < lines, vb = rb < lines;
    if constexpr (L_ != 0) {
        // rows without data read row rowsA0 (in bounds) and are dropped
        constexpr int kPer = points_per_thread(L_, std::is_same<S, float>::value);
        double a[kPer], c[kPer];
        rhs_rows<kPer, G>(fp.kind, fp.rowsA0 + (va ? ra : 0), fp.rowsA0 + (vb ? rb : 0),
                       fp.colA0, fp.lenA, t, Tp, N, b.in, a, c);
#pragma unroll
        for (int u = 0; u < kPer; ++u) {
            const int n = t + u * Tp;
            if (n < fp.lenA) put_ext(z, fp.xfA, n, fp.lenA, L, va ? a[u] : 0., vb ? c[u] : 0.);
        }
    } else {
        for (int n = t; n < fp.lenA; n += Tp) {
            const int j = fp.colA0 + n;
            const double a = va ? rhs_value<G>(fp.kind, fp.rowsA0 + ra, j, N, b.in) : 0.;
            const double c = vb ? rhs_value<G>(fp.kind, fp.rowsA0 + rb, j, N, b.in) : 0.;
            put_ext(z, fp.xfA, n, fp.lenA, L, a, c);
        }
    }
    put_ext_zeros(z, fp.xfA, fp.lenA, t);
    __syncthreads();
    fft_any<L_, S>(z, g, tw, t);
    S* T = static_cast<S*>(fp.T);
    for (int k = t; k < fp.lenA; k += Tp) {
        double a, c;
        get_ext(z, fp.xfA, k, a, c);
        if (va) T[(size_t)k * lines + ra] = (S)a;
        if (vb) T[(size_t)k * lines + rb] = (S)c;
    }
}

constexpr int kMaxPerThread = 4;

// 1: pass B evaluates the spectral multipliers from the eigenvalues (two
// FP64 divisions per point, no table traffic); 0: it reads the
// precomputed tables.  Measured at 641^2: pass B 43.9 us evaluated, 41.7 us
// read, so 0 by default.
#ifndef LCODE_FFT_TABLE_INLINE
#define LCODE_FFT_TABLE_INLINE 0
#endif

// LCODE_PASS_B_WAVES: waves per SIMD the register allocation must allow
// (6: at most 80 VGPRs; at 641^2 a CU holds three 8-wave FP64 blocks or two
// 12-wave FP32 blocks, and at 8 the FP64 pass spills).
#ifndef LCODE_PASS_B_WAVES
#define LCODE_PASS_B_WAVES 6
#endif
template <int L_, bool TW_LDS, typename S>
__global__ void __launch_bounds__(1024)
__attribute__((amdgpu_waves_per_eu(LCODE_PASS_B_WAVES)))
pass_b_kernel(SolverGeom g, SolveBatch b) {
    using V = typename Vec<S>::T;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    V* lds = reinterpret_cast<V*>(lds_raw);
    __shared__ double red[2][16];
    if (!gate_open(b.gate)) return;
    constexpr int kPts = L_ != 0 ? points_per_thread(L_, std::is_same<S, float>::value) : kMaxPerThread;
    const BlockPos bp = xcd_block();
    const FieldPlan& fp = b.f[bp.field];
    const int Tp = g.Tp, L = g.fft.L;
    const int lines = fp.lenA, len = fp.rowsA, xf = fp.xfB;
    if (bp.block * g.P * 2 >= lines) return;
    const int pair = threadIdx.x / Tp, t = threadIdx.x - pair * Tp;
    V* z = lds + (size_t)pair * padded_len(L);
    const Tw<S> tw = load_twiddles<L_, TW_LDS, S>(lds, g);
    const int la = (bp.block * g.P + pair) * 2, lb = la + 1;
    const bool va = la < lines, vb = lb < lines;
    S* rowa = static_cast<S*>(fp.T) + (size_t)la * len;
    S* rowb = static_cast<S*>(fp.T) + (size_t)lb * len;

    if constexpr (L_ != 0) {
        constexpr int kPer = points_per_thread(L_, std::is_same<S, float>::value);
        const S* T = static_cast<const S*>(fp.T);
        const size_t oa = (size_t)(va ? la : 0) * len, ob = (size_t)(vb ? lb : 0) * len;
        S xa[kPer], xb[kPer];
#pragma unroll
        for (int u = 0; u < kPer; ++u) {
            const int nc = min(t + u * Tp, len - 1);
            xa[u] = T[oa + nc];
            xb[u] = T[ob + nc];
        }
#pragma unroll
        for (int u = 0; u < kPer; ++u) {
            const int n = t + u * Tp;
            if (n < len)
                put_ext(z, xf, n, len, L, va ? (double)xa[u] : 0., vb ? (double)xb[u] : 0.);
        }
    } else {
        for (int n = t; n < len; n += Tp)
            put_ext(z, xf, n, len, L, va ? (double)rowa[n] : 0., vb ? (double)rowb[n] : 0.);
    }
    put_ext_zeros(z, xf, len, t);
    __syncthreads();
    fft_any<L_, S>(z, g, tw, t);
    // the spectral multipliers of this thread's points, all loaded before
    // their first use (rows and points clamped in bounds, unused values
    // dropped); loading them before the first transform instead keeps 12
    // more VGPRs live across it, which spill under the 64-VGPR budget
    S ta[kPts], tb[kPts];
    if (LCODE_FFT_TABLE_INLINE) {
        // evaluated as the tables are built on the host, operation for
        // operation (the same values, none of the table's HBM traffic)
        const double lla = fp.lamL[va ? la : 0], llb = fp.lamL[vb ? lb : 0];
#pragma unroll
        for (int u = 0; u < kPts; ++u) {
            const int kc = min(t + u * Tp, len - 1);
            const double lk = fp.lamK[kc];
            const double a = 1 / (lla + lk + fp.lam_add) / fp.lam_norm;
            const double c = 1 / (llb + lk + fp.lam_add) / fp.lam_norm;
            ta[u] = (S)(fp.zero_mode && la == 0 && kc == 0 ? 0. : a);
            tb[u] = (S)c;
        }
    } else {
        const size_t oa = (size_t)(va ? la : 0) * len, ob = (size_t)(vb ? lb : 0) * len;
        const S* tab = static_cast<const S*>(fp.table);
#pragma unroll
        for (int u = 0; u < kPts; ++u) {
            const int kc = min(t + u * Tp, len - 1);
            ta[u] = tab[oa + kc];
            tb[u] = tab[ob + kc];
        }
    }

    S ha[kPts], hb[kPts];
#pragma unroll
    for (int u = 0; u < kPts; ++u) {
        const int k = t + u * Tp;
        ha[u] = hb[u] = 0;
        if (k < len) {
            double a, c;
            get_ext(z, xf, k, a, c);
            if (va) ha[u] = (S)a * ta[u];
            if (vb) hb[u] = (S)c * tb[u];
        }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < kPts; ++u) {
        const int k = t + u * Tp;
        if (k < len) put_ext(z, xf, k, len, L, (double)ha[u], (double)hb[u]);
    }
    put_ext_zeros(z, xf, len, t);
    __syncthreads();
    fft_any<L_, S>(z, g, tw, t);

    double sa = 0., sb = 0.;
    for (int k = t; k < len; k += Tp) {
        double a, c;
        get_ext(z, xf, k, a, c);
        if (va) rowa[k] = (S)a;
        if (vb) rowb[k] = (S)c;
        sa += a;
        sb += c;
    }
    if (fp.partial) {
        // Bz mean: the sum of all Bz equals sum_k colsum[k] * sum_i T[k][i].
        const int wave = threadIdx.x >> 6;
        sa = wave_sum(sa);
        sb = wave_sum(sb);
        if ((threadIdx.x & 63) == 0) {
            red[0][wave] = sa;
            red[1][wave] = sb;
        }
        __syncthreads();
        if (t == 0) {
            const int w0 = (pair * Tp) >> 6, nw = Tp >> 6;
            double ta = 0., tb = 0.;
            for (int w = 0; w < nw; ++w) {
                ta += red[0][w0 + w];
                tb += red[1][w0 + w];
            }
            if (va) fp.partial[la] = fp.colsum[la] * ta;
            if (vb) fp.partial[lb] = fp.colsum[lb] * tb;
        }
    }
}

// G: storage type of the fields, averages and guess (FP64 epilogue
// arithmetic in both).
template <int L_, bool TW_LDS, typename S, typename G>
__global__ void __launch_bounds__(1024) LCODE_PASS_C_ATTR
pass_c_kernel(SolverGeom g, SolveBatch b) {
    using V = typename Vec<S>::T;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    V* lds = reinterpret_cast<V*>(lds_raw);
    __shared__ double red[16];
    if (!gate_open(b.gate)) return;
    const BlockPos bp = xcd_block();
    const FieldPlan& fp = b.f[bp.field];
    const int Tp = g.Tp, L = g.fft.L, N = g.N;
    if (bp.block * g.P * 2 >= N) {
        if (b.resid && threadIdx.x == 0) {    // no rows: a zero partial
            const int lin = blockIdx.y * gridDim.x + blockIdx.x;
            b.resid_part[2 * lin] = b.resid_part[2 * lin + 1] = 0.;
        }
        return;
    }
    const int pair = threadIdx.x / Tp, t = threadIdx.x - pair * Tp;
    V* z = lds + (size_t)pair * padded_len(L);
    const Tw<S> tw = load_twiddles<L_, TW_LDS, S>(lds, g);
    const int ra = (bp.block * g.P + pair) * 2, rb = ra + 1;
    const int ca = ra - fp.rowsA0, cb = rb - fp.rowsA0;
    const bool ha = ra < N && ca >= 0 && ca < fp.rowsA;
    const bool hb = rb < N && cb >= 0 && cb < fp.rowsA;
    const int stride = fp.rowsA;
    const S* T = static_cast<const S*>(fp.T);
    if constexpr (L_ != 0) {
        // every column load of the thread first (addresses clamped in
        // bounds, unused values dropped), then the LDS stores
        constexpr int kPer = points_per_thread(L_, std::is_same<S, float>::value);
        const int cac = ha ? ca : 0, cbc = hb ? cb : 0;
        S xa[kPer], xb[kPer];
#pragma unroll
        for (int u = 0; u < kPer; ++u) {
            const S* col = T + (size_t)min(t + u * Tp, fp.lenA - 1) * stride;
            xa[u] = col[cac];
            xb[u] = col[cbc];
        }
#pragma unroll
        for (int u = 0; u < kPer; ++u) {
            const int n = t + u * Tp;
            if (n < fp.lenA)
                put_ext(z, fp.xfA, n, fp.lenA, L, ha ? (double)xa[u] : 0.,
                        hb ? (double)xb[u] : 0.);
        }
    } else {
        for (int n = t; n < fp.lenA; n += Tp) {
            const S* col = T + (size_t)n * stride;
            put_ext(z, fp.xfA, n, fp.lenA, L, ha ? (double)col[ca] : 0.,
                    hb ? (double)col[cb] : 0.);
        }
    }
    put_ext_zeros(z, fp.xfA, fp.lenA, t);
    __syncthreads();
    fft_any<L_, S>(z, g, tw, t);

    double mean = 0.;
    if (fp.partial) {
        double s = 0.;
        for (int k = threadIdx.x; k < fp.lenA; k += blockDim.x) s += fp.partial[k];
        s = wave_sum(s);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
        __syncthreads();
        double tot = 0.;
        const int nw = blockDim.x >> 6;
        for (int w = 0; w < nw; ++w) tot += red[w];
        mean = tot / ((double)N * N);
    }

    const bool transverse = fp.kind == F_EX || fp.kind == F_EY ||
                            fp.kind == F_BX || fp.kind == F_BY;
    const bool varA = b.variantA && transverse;
    const G* guess = b.resid ? gp<G>(b.guess[fp.kind]) : nullptr;
    const G* prevF = gp<G>(fp.prevF);
    G* const outF = gp<G>(fp.outF);
    G* const outAvg = gp<G>(fp.outAvg);
    double r_num = 0., r_den = 0.;
    // one output point (i, j) of this pair: F from the spectrum value v,
    // prev and guess values p, q of that point
    auto emit = [&](int i, int j, double v, double p, double q) {
        const size_t idx = (size_t)i * N + j;
        double F = v;
        if (fp.partial) F -= mean;                         // lcode.py:362
        if (varA) F = 2 * F - p;                           // lcode.py:869-871
        if (outF) outF[idx] = (G)F;
        if (outAvg) outAvg[idx] = (G)((F + p) / 2);        // lcode.py:876-881
        if (guess) {
            const double d = F - q;
            r_num += d * d;
            r_den += F * F;
        }
    };
    // the field value at column j: a row without data shares its FFT with
    // a partner row; keep it an exact zero instead of the partner's noise
    auto spec = [&](int j, double& va, double& vb) {
        va = vb = 0.;
        const int k = j - fp.colA0;
        if (k >= 0 && k < fp.lenA) get_ext(z, fp.xfA, k, va, vb);
        if (!ha) va = 0.;
        if (!hb) vb = 0.;
    };
    if constexpr (L_ != 0) {
        // the prev and guess loads of every point of the thread first
        // (rows and columns clamped in bounds), then the stores
        constexpr int kPer = points_per_thread(L_, std::is_same<S, float>::value);
        const size_t oa = (size_t)min(ra, N - 1) * N, ob = (size_t)min(rb, N - 1) * N;
        double pa[kPer], pb[kPer], qa[kPer], qb[kPer];
#pragma unroll
        for (int u = 0; u < kPer; ++u) pa[u] = pb[u] = qa[u] = qb[u] = 0.;
        if (prevF) {
#pragma unroll
            for (int u = 0; u < kPer; ++u) {
                const int jc = min(t + u * Tp, N - 1);
                pa[u] = prevF[oa + jc];
                pb[u] = prevF[ob + jc];
            }
        }
        if (guess) {
#pragma unroll
            for (int u = 0; u < kPer; ++u) {
                const int jc = min(t + u * Tp, N - 1);
                qa[u] = guess[oa + jc];
                qb[u] = guess[ob + jc];
            }
        }
#pragma unroll
        for (int u = 0; u < kPer; ++u) {
            const int j = t + u * Tp;
            if (j >= N) continue;
            double va, vb;
            spec(j, va, vb);
            if (ra < N) emit(ra, j, va, pa[u], qa[u]);
            if (rb < N) emit(rb, j, vb, pb[u], qb[u]);
        }
    } else {
        for (int j = t; j < N; j += Tp) {
            double va, vb;
            spec(j, va, vb);
            const size_t ia = (size_t)ra * N + j, ib = (size_t)rb * N + j;
            if (ra < N)
                emit(ra, j, va, prevF ? (double)prevF[ia] : 0., guess ? (double)guess[ia] : 0.);
            if (rb < N)
                emit(rb, j, vb, prevF ? (double)prevF[ib] : 0., guess ? (double)guess[ib] : 0.);
        }
    }
    if (b.resid) {
        // guided step: this block's share of the corrector's residual against
        // the guess, summed over its waves; one reduction kernel adds the
        // blocks' partials afterwards (no same-address atomics)
        r_num = wave_sum(r_num);
        r_den = wave_sum(r_den);
        __shared__ double rs[2][16];
        const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
        if ((threadIdx.x & 63) == 0) {
            rs[0][w] = r_num;
            rs[1][w] = r_den;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            double a = 0., d = 0.;
            for (int k = 0; k < nw; ++k) {
                a += rs[0][k];
                d += rs[1][k];
            }
            const int lin = blockIdx.y * gridDim.x + blockIdx.x;
            b.resid_part[2 * lin] = a;
            b.resid_part[2 * lin + 1] = d;
        }
    }
}

// resid[0..1] = sums of the blocks' partials (fixed order)
__global__ void __launch_bounds__(256)
resid_reduce_kernel(const double* part, int nblocks, double* resid) {
    __shared__ double red[2][256];
    double a = 0., d = 0.;
    for (int k = threadIdx.x; k < nblocks; k += 256) {
        a += part[2 * k];
        d += part[2 * k + 1];
    }
    red[0][threadIdx.x] = a;
    red[1][threadIdx.x] = d;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            red[0][threadIdx.x] += red[0][threadIdx.x + off];
            red[1][threadIdx.x] += red[1][threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        resid[0] = red[0][0];
        resid[1] = red[1][0];
    }
}

// Guided step set-up: guess = sum_t c[t] hist[t] for the six fields, the
// predictor-pass averages (guess + prev) / 2 (lcode.py:876-881 with the
// guess in place of the predicted fields), and zeroed residual sums.
template <typename G>
__global__ void __launch_bounds__(256)
guess_kernel(GuessArgs p) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (p.acc_zero && k < 2) p.acc_zero[k] = 0.;
    if (k >= p.n) return;
#pragma unroll
    for (int f = 0; f < 6; ++f) {
        double gsum = 0.;
        for (int t = 0; t < p.nt; ++t) gsum += p.c[t] * (double)gp<G>(p.h[t][f])[k];
        const G gs = (G)gsum;
        gp<G>(p.guess[f])[k] = gs;
        gp<G>(p.avg[f])[k] = (G)(((double)gs + (double)gp<G>(p.prev[f])[k]) / 2);
    }
}


// ---------------------------------------------------------------------------
// Wave FFT (L = 1280 and 2560): one line pair per T = L / 20 lanes (one
// wave at L = 1280, two at 2560), one pair per block (four in pass C at
// L = 1280).  With n = l + T m and k = k1 + 20 k2,
//   X[k1 + 20 k2] = sum_l W_T^(l k2) W_L^(l k1) sum_m x[l + T m] W_20^(m k1):
// lane l holds the 20 points x[l + T m] in registers and transforms them
// there (a Good-Thomas 4 x 5 DFT, no internal twiddles), multiplies by
// W_L^(l k1), and the 20 transforms of length T across the lanes run as
// Stockham radix-4 (and a last radix-2) stages through LDS, batched over k1
// with k1 fastest: z[row * 20 + (k1 ^ swz(row))], lanes taking butterflies
// beta = k1 + 20 b in order, so every LDS access of an instruction is a run
// of consecutive 16-B points (the XOR spreads the first exchange's
// stride-20 stores over the banks).  The last stage keeps its outputs in
// registers: lane l's output (r, q) is X[l + T r + q L / R], consecutive
// across lanes, so the passes read and write global memory coalesced.
// Against the radix-4/5 plan above: four LDS exchanges per transform instead
// of six, no barrier across waves at L = 1280, a quarter less VALU work.
// Between the passes the spectra are kept in grid-row order, rows N + 1
// long, so a pass-A or pass-C line pair (two grid rows) is one aligned
// 16-byte element per k.
// ---------------------------------------------------------------------------

// LCODE_FFT_WAVE_POINTS_1280: points per lane at L = 1280, 20 (one wave per
// line pair) or 10 (two waves: a Good-Thomas 2 x 5 DFT-10 per lane, stages
// 4, 4, 4, 2 across 128 lanes, four waves per SIMD).  L = 2560 takes 20.
#ifndef LCODE_FFT_WAVE_POINTS_1280
#define LCODE_FFT_WAVE_POINTS_1280 20
#endif

template <int L> struct WaveLanes {
    static constexpr int R1 = L == 1280 ? LCODE_FFT_WAVE_POINTS_1280 : 20;   // points per lane
    static constexpr int T = L / R1;                                          // lanes per pair
    static_assert(R1 == 20 || R1 == 10, "wave FFT: 20 or 10 points per lane");
    static_assert(T == 64 || T == 128, "wave FFT: one or two waves per line pair");
};

// LDS address of point (row, k1) of the batched transforms: rows of R1
// points, k1 XOR-swizzled within the row (20: by (row >> 1) & 3, 10: by
// (row >> 2) & 1) so that the first exchange's stride-R1 stores of eight
// consecutive lanes hit eight different 16-byte bank groups.
template <int R1>
__device__ __forceinline__ int waddr(int row, int k1) {
    const int sw = R1 == 20 ? (row >> 1) & 3 : (row >> 2) & 1;
    return row * R1 + (k1 ^ sw);
}

// In-place DFT of R1 = N1 N2 points (gcd 1) by Good-Thomas: n = (N2 n1 + N1
// n2) mod R1, k = (N2 (N2^-1 mod N1) k1 + N1 (N1^-1 mod N2) k2) mod R1, so
// W_R1^(nk) = W_N1^(n1 k1) W_N2^(n2 k2) and no twiddles are needed: 20 = 4
// x 5 (k = 5 k1 + 16 k2), 10 = 2 x 5 (k = 5 k1 + 6 k2).
template <int R1, typename V>
__device__ __forceinline__ void dft_lane(V (&x)[R1]) {
    constexpr int N1 = R1 == 20 ? 4 : 2, N2 = 5;
    constexpr int K1 = R1 == 20 ? 5 : 5, K2 = R1 == 20 ? 16 : 6;
    V a[N2][N1];
#pragma unroll
    for (int n2 = 0; n2 < N2; ++n2) {
#pragma unroll
        for (int n1 = 0; n1 < N1; ++n1) a[n2][n1] = x[(N2 * n1 + N1 * n2) % R1];
        Dft<N1, V>::run(a[n2]);
    }
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) {
        V b[N2];
#pragma unroll
        for (int n2 = 0; n2 < N2; ++n2) b[n2] = a[n2][k1];
        Dft<N2, V>::run(b);
#pragma unroll
        for (int k2 = 0; k2 < N2; ++k2) x[(K1 * k1 + K2 * k2) % R1] = b[k2];
    }
}

// Lane l's extension points x[l + T m] of the type-1 transform whose n
// inputs (pairs) are e[0 .. n) in LDS: cosine x[p] = e[p] (p <= L/2),
// e[L - p] above; sine x[p] = e[p - 1] below L/2, -e[L - 1 - p] above,
// 0 at p = 0 and L/2.
template <int L, typename V>
__device__ __forceinline__ void wext_load(const V* e, int xf, int lane,
                                          V (&x)[WaveLanes<L>::R1]) {
    constexpr int T = WaveLanes<L>::T, R1 = WaveLanes<L>::R1, H = L / 2;
#pragma unroll
    for (int m = 0; m < R1; ++m) {
        const int p = lane + T * m;        // p < H exactly for m < R1 / 2
        if (xf == XF_DCT) {
            x[m] = e[m < R1 / 2 ? p : L - p];
        } else if (m < R1 / 2) {
            const V v = e[max(p - 1, 0)];
            x[m] = p == 0 ? mk<V>(0, 0) : v;
        } else {
            const V v = e[min(L - 1 - p, H - 2)];
            x[m] = p == H ? mk<V>(0, 0) : mk<V>(-v.x, -v.y);
        }
    }
}

// One Stockham stage of radix R and span SPAN over the R1 transforms of
// length T in z: lane l takes butterflies beta = l + T r, r < ceil(R1 / R),
// of the R1 T / R (the last r partial when R does not divide R1); the last
// stage (R dividing R1) leaves output (r, q) in x[r * R + q].
template <int L, int R, int SPAN, bool LAST, typename V>
__device__ __forceinline__ void wstage(V* z, V (&x)[WaveLanes<L>::R1],
                                       const double2* __restrict__ tw, int lane) {
    constexpr int T = WaveLanes<L>::T, R1 = WaveLanes<L>::R1;
    constexpr int NB = (R1 + R - 1) / R, NBT = R1 * T / R, STRIDE = T / R;
    static_assert(!LAST || R1 % R == 0, "the last stage's radix divides R1");
    V v[NB][R];
    double2 w[NB];   // W_(SPAN R)^k = W_L^(k L / (SPAN R)) of each butterfly
#pragma unroll
    for (int r = 0; r < NB; ++r) {
        const int beta = min(lane + T * r, NBT - 1);   // a partial r repeats the last
        const int k1 = beta % R1, b = beta / R1, k = b % SPAN;
        if constexpr (SPAN > 1) w[r] = tw[k * (L / (SPAN * R))];
#pragma unroll
        for (int q = 0; q < R; ++q) v[r][q] = z[waddr<R1>(b + q * STRIDE, k1)];
    }
    __syncthreads();   // every read of the stage before its first write
#pragma unroll
    for (int r = 0; r < NB; ++r) {
        const int beta = lane + T * r;
        if (NBT % T != 0 && beta >= NBT) continue;   // (only in a partial r)
        if constexpr (SPAN > 1) {
            // the powers formed in FP64, each rounded once to the transform's
            // precision
            double2 wq = w[r];
#pragma unroll
            for (int q = 1; q < R; ++q) {
                if (q > 1) wq = cmul(wq, w[r]);
                v[r][q] = cmul(v[r][q], mk<V>(wq.x, wq.y));
            }
        }
        Dft<R, V>::run(v[r]);
        if constexpr (LAST) {
#pragma unroll
            for (int q = 0; q < R; ++q) x[r * R + q] = v[r][q];
        } else {
            const int k1 = beta % R1, b = beta / R1, k = b % SPAN;
            const int j = (b - k) * R + k;
#pragma unroll
            for (int q = 0; q < R; ++q) z[waddr<R1>(j + q * SPAN, k1)] = v[r][q];
        }
    }
    if constexpr (!LAST) __syncthreads();
}

// Forward FFT of length L of the extension points in x (from wext_load).
// z: T * R1 points of LDS, free on entry (the caller synchronised after its
// last read).  On return x[r * R + q] = X[lane + T r + q L / R], R the last
// radix (4 for one wave per pair at L = 1280, 2 otherwise), and z is free
// again.
template <int L, typename V>
__device__ __forceinline__ void wave_fft(V* z, V (&x)[WaveLanes<L>::R1],
                                         const double2* __restrict__ tw, int lane) {
    constexpr int T = WaveLanes<L>::T, R1 = WaveLanes<L>::R1;
    // W_L^(lane k1) = W_L^(lane 4a) W_L^(lane c) for k1 = 4a + c: a few
    // table entries, one FP64 product each
    constexpr int NA = (R1 + 3) / 4;
    double2 lo[4], hi[NA];
#pragma unroll
    for (int c = 1; c < 4; ++c) lo[c] = tw[lane * c];
#pragma unroll
    for (int a = 1; a < NA; ++a) hi[a] = tw[lane * 4 * a];
    dft_lane<R1>(x);
#pragma unroll
    for (int k1 = 1; k1 < R1; ++k1) {
        const int a = k1 / 4, c = k1 % 4;
        const double2 w = a == 0 ? lo[c] : c == 0 ? hi[a] : cmul(hi[a], lo[c]);
        x[k1] = cmul(x[k1], mk<V>(w.x, w.y));
    }
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) z[waddr<R1>(lane, k1)] = x[k1];
    __syncthreads();
    wstage<L, 4, 1, false>(z, x, tw, lane);
    wstage<L, 4, 4, false>(z, x, tw, lane);
    if constexpr (T == 64) {
        wstage<L, 4, 16, true>(z, x, tw, lane);
    } else {
        wstage<L, 4, 16, false>(z, x, tw, lane);
        wstage<L, 2, 64, true>(z, x, tw, lane);
    }
}

template <int L> struct WaveOut {
    static constexpr int R = WaveLanes<L>::T == 64 ? 4 : 2;   // last radix
};

#define LCODE_WAVE_R1 WaveLanes<L>::R1

// The outputs of the type-1 transform held by this lane: f(u, k) for each
// (register u, output k).
template <int L, typename F>
__device__ __forceinline__ void wext_index(int xf, int lane, F&& f) {
    constexpr int T = WaveLanes<L>::T, R = WaveOut<L>::R, H = L / 2;
#pragma unroll
    for (int u = 0; u < LCODE_WAVE_R1; ++u) {
        const int r = u / R, q = u % R;
        if (q >= R / 2) continue;
        const int idx = lane + T * r + q * (L / R);
        if (xf == XF_DCT) f(u, idx);
        else if (idx >= 1) f(u, idx - 1);
    }
    if (xf == XF_DCT && lane == 0) f(R / 2, H);
}

// Calls f(u, k, a, b) for every output k of the type-1 transform held by
// this lane (register u; wext_index gives the (u, k) pairs alone): cosine
// (L / 2 + 1 outputs) a = Re X[k], b = Im X[k]; sine (L / 2 - 1 outputs)
// a = -Im X[k + 1], b = Re X[k + 1].  The registers with q < R / 2 hold
// X[idx] for idx < L / 2; X[L / 2] is lane 0's register R / 2.
template <int L, typename V, typename F>
__device__ __forceinline__ void wext_out(const V (&x)[LCODE_WAVE_R1], int xf,
                                         int lane, F&& f) {
    wext_index<L>(xf, lane, [&](int u, int k) {
        if (xf == XF_DCT) f(u, k, (double)x[u].x, (double)x[u].y);
        else f(u, k, -(double)x[u].y, (double)x[u].x);
    });
}

// Waves per SIMD the LDS allows (160 KiB per CU, L points per block of T
// lanes): 40 / R1 with FP64 points (two at 20 points per lane, four at 10),
// twice that with FP32 (three asked for: 168 registers, where the FP32
// passes of L = 1280 spill at most 6; those of L = 2560 are not launched).
// LCODE_FFT_WAVE_PAIRS (4): line pairs per block of the wave pass C: a block
// reads the spectra of P consecutive row pairs together, P 16-byte elements
// per k in one wave-wide access (pass C 24.2 -> 22.6 us at 641^2 with 4; the
// same for pass A's stores cost it 26.4 -> 30.0 us: the block barriers of the
// transform tie its waves' long right-hand-side loads together).
// At L = 2560 one pair per block: four would take all of a CU's LDS (the
// 1281^2 step 717 -> 712 xi-steps/s).
#ifndef LCODE_FFT_WAVE_PAIRS
#define LCODE_FFT_WAVE_PAIRS 4
#endif
template <int L> struct WavePairs {
    static constexpr int P = L == 1280 ? LCODE_FFT_WAVE_PAIRS : 1;
};

#define LCODE_WAVE_ATTR(S) \
    __attribute__((amdgpu_waves_per_eu(sizeof(S) == 8 ? 40 / WaveLanes<L>::R1 : 3)))

__device__ __forceinline__ double block_sum_waves(double v, double* scratch) {
    v = wave_sum(v);
    if (blockDim.x == 64) return v;
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) scratch[w] = v;
    __syncthreads();
    double t = 0.;
    for (int k = 0; k < (int)(blockDim.x >> 6); ++k) t += scratch[k];
    __syncthreads();
    return t;
}

template <int L, typename S, typename G>
__global__ void __launch_bounds__(WaveLanes<L>::T) LCODE_WAVE_ATTR(S)
pass_a_wave_kernel(SolverGeom g, SolveBatch b) {
    using V = typename Vec<S>::T;
    constexpr int T = WaveLanes<L>::T;
    constexpr int kPer = (L / 2 + 1 + T - 1) / T;   // inputs per lane
    __shared__ V z[T * LCODE_WAVE_R1];
    if (!gate_open(b.gate)) return;
    const BlockPos bp = xcd_block_rows();
    const FieldPlan& fp = b.f[bp.field];
    const int N = g.N, lane = threadIdx.x;
    // grid rows ga (even) and ga + 1; the spectra keep grid row order with
    // rows N + 1 long, so the pair is one 16-byte element per k
    const int ga = bp.block * 2, gb = ga + 1;
    const int ta = ga - fp.rowsA0, tb = gb - fp.rowsA0;   // rows of the transform
    const bool va = ta >= 0 && ta < fp.rowsA, vb = tb >= 0 && tb < fp.rowsA;
    if (!va && !vb) return;
    {
        double a[kPer], c[kPer];
        const int gc = va ? ga : gb;   // rows without data read a valid one
        rhs_rows<kPer, G, sizeof(S) == 8 ? LCODE_RHS_BATCH : 0>(fp.kind, gc, vb ? gb : gc,
                                                                fp.colA0, fp.lenA, lane, T, N,
                                                                b.in, a, c);
#pragma unroll
        for (int u = 0; u < kPer; ++u) {
            const int n = lane + u * T;
            if (n < fp.lenA) z[n] = mk<V>(va ? (S)a[u] : (S)0, vb ? (S)c[u] : (S)0);
        }
    }
    __syncthreads();
    V x[LCODE_WAVE_R1];
    wext_load<L>(z, fp.xfA, lane, x);
    __syncthreads();
    wave_fft<L>(z, x, g.wave_tw, lane);
    V* Tsp = static_cast<V*>(fp.T);
    const int rowp = (N + 1) / 2;   // pairs per row of the spectra
    wext_out<L>(x, fp.xfA, lane, [&](int, int k, double a, double c) {
        Tsp[(size_t)k * rowp + ga / 2] = mk<V>((S)a, (S)c);
    });
}

// Waves per SIMD of the FP32 wave pass C (build option; 3 as the other
// FP32 passes: at 2 the mixed six-field solve measured the same).
#ifndef LCODE_WAVE_C_F32_WAVES
#define LCODE_WAVE_C_F32_WAVES 3
#endif
#define LCODE_WAVE_C_ATTR(S) \
    __attribute__((amdgpu_waves_per_eu(sizeof(S) == 8 ? 40 / WaveLanes<L>::R1 : LCODE_WAVE_C_F32_WAVES)))

// LCODE_WAVE_TABLE_INLINE (0): the wave pass B evaluates its spectral
// multipliers instead of reading the table.  LCODE_WAVE_TABLE_EARLY: it
// loads them before its first transform.
#ifndef LCODE_WAVE_TABLE_INLINE
#define LCODE_WAVE_TABLE_INLINE 0
#endif
#ifndef LCODE_WAVE_TABLE_EARLY
#define LCODE_WAVE_TABLE_EARLY 0
#endif

// Waves per SIMD the FP64 wave pass B at L = 2560 is compiled for (build
// option): 2 (the other passes' budget, 256 registers: 41 spill to
// scratch) or 1 (up to 512: 256 VGPRs and 45 AGPRs, no scratch, half the
// waves).  The 1281^2 plain run: 707.6-708.9 xi-steps/s with 2, 711.5-711.7
// with 1 (A/B on one box, two runs each).
#ifndef LCODE_WAVE_B2560_WAVES
#define LCODE_WAVE_B2560_WAVES 1
#endif
// Waves per SIMD of the FP32 wave pass B (build option): 3, as the other
// FP32 passes, spills 6 registers at L = 1280; 2 (184 VGPRs) spills none:
// the mixed six-field solve 61.5 -> 59.7-59.9 us at 641^2 (A/B on one
// box, two runs each).
#ifndef LCODE_WAVE_B_F32_WAVES
#define LCODE_WAVE_B_F32_WAVES 2
#endif
#define LCODE_WAVE_B_ATTR(S)                                                        \
    __attribute__((amdgpu_waves_per_eu(sizeof(S) == 8 ? (L == 2560 ? LCODE_WAVE_B2560_WAVES \
                                                                    : 40 / WaveLanes<L>::R1) \
                                                      : LCODE_WAVE_B_F32_WAVES)))

template <int L, typename S>
__global__ void __launch_bounds__(WaveLanes<L>::T) LCODE_WAVE_B_ATTR(S)
pass_b_wave_kernel(SolverGeom g, SolveBatch b) {
    using V = typename Vec<S>::T;
    constexpr int T = WaveLanes<L>::T;
    constexpr int kPer = (L / 2 + 1 + T - 1) / T;
    __shared__ V z[T * LCODE_WAVE_R1];
    if (!gate_open(b.gate)) return;
    const BlockPos bp = xcd_block();
    const FieldPlan& fp = b.f[bp.field];
    const int lines = fp.lenA, len = fp.rowsA, xf = fp.xfB, lane = threadIdx.x;
    const int la = bp.block * 2, lb = la + 1;
    if (la >= lines) return;
    const bool vb = lb < lines;
    // spectra rows N + 1 long in grid row order: line k's transform inputs
    // start at grid row rowsA0
    const size_t stride = (size_t)g.N + 1;
    S* rowa = static_cast<S*>(fp.T) + (size_t)la * stride + fp.rowsA0;
    S* rowb = static_cast<S*>(fp.T) + (size_t)(vb ? lb : la) * stride + fp.rowsA0;
    {
        S xa[kPer], xb[kPer];
#pragma unroll
        for (int u = 0; u < kPer; ++u) {
            const int n = min(lane + u * T, len - 1);
            xa[u] = rowa[n];
            xb[u] = rowb[n];
        }
#pragma unroll
        for (int u = 0; u < kPer; ++u) {
            const int n = lane + u * T;
            if (n < len) z[n] = mk<V>(xa[u], vb ? xb[u] : (S)0);
        }
    }
    // the spectral multipliers of this lane's outputs (or,
    // LCODE_WAVE_TABLE_INLINE, evaluated from the eigenvalues as the host
    // builds the table: the same correctly rounded divisions); with
    // LCODE_WAVE_TABLE_EARLY their loads are in flight during the first
    // transform
    S ma[LCODE_WAVE_R1], mb[LCODE_WAVE_R1];
    auto multipliers = [&]() {
        if (LCODE_WAVE_TABLE_INLINE) {
            const double lla = fp.lamL[la], llb = fp.lamL[vb ? lb : la];
            wext_index<L>(xf, lane, [&](int u, int k) {
                const double lk = fp.lamK[k];
                const double va = 1 / (lla + lk + fp.lam_add) / fp.lam_norm;
                ma[u] = (S)(fp.zero_mode && la == 0 && k == 0 ? 0. : va);
                mb[u] = (S)(1 / (llb + lk + fp.lam_add) / fp.lam_norm);
            });
        } else {
            const S* tab = static_cast<const S*>(fp.table);
            const S* ta = tab + (size_t)la * len;
            const S* tb = tab + (size_t)(vb ? lb : la) * len;
            wext_index<L>(xf, lane, [&](int u, int k) {
                ma[u] = ta[k];
                mb[u] = tb[k];
            });
        }
    };
    if (LCODE_WAVE_TABLE_EARLY) multipliers();
    __syncthreads();
    V x[LCODE_WAVE_R1];
    wext_load<L>(z, xf, lane, x);
    __syncthreads();
    wave_fft<L>(z, x, g.wave_tw, lane);
    {
        if (!LCODE_WAVE_TABLE_EARLY) multipliers();
        wext_out<L>(x, xf, lane, [&](int u, int k, double a, double c) {
            z[k] = mk<V>((S)a * ma[u], vb ? (S)c * mb[u] : (S)0);
        });
    }
    __syncthreads();
    wext_load<L>(z, xf, lane, x);
    __syncthreads();
    wave_fft<L>(z, x, g.wave_tw, lane);
    double sa = 0., sb = 0.;
    wext_out<L>(x, xf, lane, [&](int, int k, double a, double c) {
        rowa[k] = (S)a;
        if (vb) rowb[k] = (S)c;
        sa += a;
        sb += c;
    });
    if (fp.partial) {
        // Bz mean: the sum of all Bz equals sum_k colsum[k] * sum_i T[k][i]
        double* scratch = reinterpret_cast<double*>(z);
        sa = block_sum_waves(sa, scratch);
        sb = block_sum_waves(sb, scratch);
        if (lane == 0) {
            fp.partial[la] = fp.colsum[la] * sa;
            if (vb) fp.partial[lb] = fp.colsum[lb] * sb;
        }
    }
}

template <int L, typename S, typename G>
__global__ void __launch_bounds__(WaveLanes<L>::T * WavePairs<L>::P) LCODE_WAVE_C_ATTR(S)
pass_c_wave_kernel(SolverGeom g, SolveBatch b) {
    using V = typename Vec<S>::T;
    constexpr int T = WaveLanes<L>::T, P = WavePairs<L>::P, NT = T * P;
    constexpr int kPer = (L / 2 + 1 + T - 1) / T;
    constexpr int kZ = T * LCODE_WAVE_R1;   // LDS points per line pair
    // pair p's staged inputs start kStage p points into its LDS (64 bytes
    // apart: the four lanes storing one k hit different banks)
    constexpr int kStage = P > 1 ? 4 : 0;
    static_assert(L / 2 + 1 + kStage * (P - 1) <= kZ, "staged inputs fit the pair's LDS");
    __shared__ V zall[P * kZ];
    if (!gate_open(b.gate)) return;
    const BlockPos bp = xcd_block();
    const FieldPlan& fp = b.f[bp.field];
    const int N = g.N, w = threadIdx.x / T, lane = threadIdx.x % T;
    V* z = zall + w * kZ;
    // this wave's pair of output rows; a pair beyond the grid (the last
    // block) runs through the barriers and emits nothing
    const int ra = (bp.block * P + w) * 2, rb = ra + 1;
    const bool live = ra < N;
    auto has = [&](int r) { return r < N && r - fp.rowsA0 >= 0 && r - fp.rowsA0 < fp.rowsA; };
    const bool ha = has(ra), hb = has(rb);
    {
        // the block's P pairs of grid rows are P consecutive 16-byte
        // elements per k of the spectra (rows N + 1 long, grid row order):
        // the block reads them together, 16 k per wave-wide load, into each
        // pair's LDS; rows without data hold stale values and are dropped
        const V* Tsp = static_cast<const V*>(fp.T);
        const int rowp = (N + 1) / 2;
        V xv[kPer];
#pragma unroll
        for (int u = 0; u < kPer; ++u) {
            const int idx = threadIdx.x + u * NT;
            const int k = min(idx / P, fp.lenA - 1), pr = bp.block * P + idx % P;
            xv[u] = Tsp[(size_t)k * rowp + min(pr, rowp - 1)];
        }
#pragma unroll
        for (int u = 0; u < kPer; ++u) {
            const int idx = threadIdx.x + u * NT;
            const int k = idx / P, p = idx % P, r0 = (bp.block * P + p) * 2;
            if (k < fp.lenA && r0 < N)
                zall[p * kZ + kStage * p + k] =
                    mk<V>(has(r0) ? xv[u].x : (S)0, has(r0 + 1) ? xv[u].y : (S)0);
        }
    }
    __syncthreads();
    V x[LCODE_WAVE_R1];
    wext_load<L>(z + kStage * w, fp.xfA, lane, x);
    __syncthreads();
    wave_fft<L>(z, x, g.wave_tw, lane);

    double mean = 0.;
    if (fp.partial) {
        double s = 0.;
        for (int k = threadIdx.x; k < fp.lenA; k += NT) s += fp.partial[k];
        mean = block_sum_waves(s, reinterpret_cast<double*>(zall)) / ((double)N * N);
    }
    const bool transverse = fp.kind == F_EX || fp.kind == F_EY ||
                            fp.kind == F_BX || fp.kind == F_BY;
    const bool varA = b.variantA && transverse;
    const G* guess = b.resid ? gp<G>(b.guess[fp.kind]) : nullptr;
    const G* prevF = gp<G>(fp.prevF);
    G* const outF = gp<G>(fp.outF);
    G* const outAvg = gp<G>(fp.outAvg);
    double r_num = 0., r_den = 0.;
    auto emit = [&](int i, int j, double v, double p, double q) {
        const size_t idx = (size_t)i * N + j;
        double F = v;
        if (fp.partial) F -= mean;                         // lcode.py:362
        if (varA) F = 2 * F - p;                           // lcode.py:869-871
        if (outF) outF[idx] = (G)F;
        if (outAvg) outAvg[idx] = (G)((F + p) / 2);        // lcode.py:876-881
        if (guess) {
            const double d = F - q;
            r_num += d * d;
            r_den += F * F;
        }
    };
    // output column j of the lane's register u: j = lane + T r + q L / R
    // (the cosine transform's index, 1 + the sine transform's), below L / 2
    // for q < R / 2; column L / 2 = N - 1 is lane 0's register R / 2.
    // Columns outside [colA0, colA0 + lenA) are exact zeros.
    constexpr int R = WaveOut<L>::R, H = L / 2;
    const size_t oa = (size_t)ra * N, ob = (size_t)min(rb, N - 1) * N;
    auto value = [&](int u, int j, double& va, double& vb) {
        va = vb = 0.;
        const int k = j - fp.colA0;
        if (k >= 0 && k < fp.lenA) {
            if (fp.xfA == XF_DCT) {
                va = x[u].x;
                vb = x[u].y;
            } else {
                va = -(double)x[u].y;
                vb = x[u].x;
            }
        }
        if (!ha) va = 0.;
        if (!hb) vb = 0.;
    };
    if (live) {
        constexpr int kOut = LCODE_WAVE_R1 / 2;   // registers with q < R / 2
        G pa[kOut], pb[kOut], qa[kOut], qb[kOut];
#pragma unroll
        for (int s = 0; s < kOut; ++s) {
            const int r = s / (R / 2), q = s % (R / 2), u = r * R + q;
            (void)u;
            const int j = lane + T * r + q * (L / R);
            pa[s] = pb[s] = qa[s] = qb[s] = (G)0;
            if (prevF) {
                pa[s] = prevF[oa + j];
                pb[s] = prevF[ob + j];
            }
            if (guess) {
                qa[s] = guess[oa + j];
                qb[s] = guess[ob + j];
            }
        }
#pragma unroll
        for (int s = 0; s < kOut; ++s) {
            const int r = s / (R / 2), q = s % (R / 2), u = r * R + q;
            const int j = lane + T * r + q * (L / R);
            double va, vb;
            value(u, j, va, vb);
            emit(ra, j, va, pa[s], qa[s]);
            if (rb < N) emit(rb, j, vb, pb[s], qb[s]);
        }
    }
    if (live && lane == 0) {
        double va, vb;
        value(R / 2, H, va, vb);
        emit(ra, H, va, prevF ? (double)prevF[oa + H] : 0., guess ? (double)guess[oa + H] : 0.);
        if (rb < N)
            emit(rb, H, vb, prevF ? (double)prevF[ob + H] : 0., guess ? (double)guess[ob + H] : 0.);
    }
    if (b.resid) {
        double* scratch = reinterpret_cast<double*>(zall);
        r_num = block_sum_waves(r_num, scratch);
        r_den = block_sum_waves(r_den, scratch);
        if (threadIdx.x == 0) {
            const int lin = blockIdx.y * gridDim.x + blockIdx.x;
            b.resid_part[2 * lin] = r_num;
            b.resid_part[2 * lin + 1] = r_den;
        }
    }
}

}  // namespace

template <int L_, bool TW_LDS, typename S, typename G>
static int launch_passes(const SolverGeom& g, const SolveBatch& b,
                         hipStream_t s, Profiler* prof, size_t lds,
                         dim3 ga, dim3 gb, dim3 gc) {
    const dim3 block(g.P * g.Tp);
    {
        ProfScope ps(prof, "solve_pass_a", s, true);
        launch_timed(ps, pass_a_kernel<L_, TW_LDS, S, G>, ga, block, lds, s, g, b);
    }
    LCODE_HIP(hipGetLastError());
    {
        ProfScope ps(prof, "solve_pass_b", s, true);
        launch_timed(ps, pass_b_kernel<L_, TW_LDS, S>, gb, block, lds, s, g, b);
    }
    LCODE_HIP(hipGetLastError());
    {
        ProfScope ps(prof, "solve_pass_c", s, true);
        launch_timed(ps, pass_c_kernel<L_, TW_LDS, S, G>, gc, block, lds, s, g, b);
    }
    LCODE_HIP(hipGetLastError());
    if (b.resid) {
        ProfScope ps(prof, "resid_reduce", s);
        resid_reduce_kernel<<<1, 256, 0, s>>>(b.resid_part, (int)(gc.x * gc.y), b.resid);
        LCODE_HIP(hipGetLastError());
    }
    return LCODE_OK;
}

int solver_lds_bytes(const SolverGeom& g) {
    const size_t cplx = g.fp32 ? sizeof(float2) : sizeof(double2);
    const size_t data = (size_t)g.P * padded_len(g.fft.L) * cplx;
    if (g.static_L == 0) return (int)(data + ((size_t)g.fft.L / 4 + 1) * cplx / 2);
    if (g.tw_lds) return (int)(data + (size_t)tw_count(g.fft.L) * sizeof(double2));
    return (int)data;
}

int solver_padded_len(int L) { return padded_len(L); }

int solver_tw_count(int L) { return tw_count(L); }

int solver_threads_per_pair(int L, bool fp32, bool static_plan) {
    return static_plan ? static_threads_per_pair(L, fp32) : threads_per_pair(L);
}

int solver_static_radix(int rem) { return next_radix(rem); }

#ifndef LCODE_FFT_STATIC_PLANS
#define LCODE_FFT_STATIC_PLANS 1
#endif

// LCODE_FFT_WAVE (1): the wave FFT passes for the FP64 transforms of L = 1280
// and 2560 (grids of 641^2 and 1281^2); 0 keeps the radix-4/5 plan there too.
#ifndef LCODE_FFT_WAVE
#define LCODE_FFT_WAVE 1
#endif

bool solver_static_length(int L) {
    if (!LCODE_FFT_STATIC_PLANS) return false;
    switch (L) {
        case 32: case 64: case 128: case 256: case 512: case 640:
        case 1024: case 1280: case 2048: case 2560:
            return true;
        default:
            return false;
    }
}

template <int L, typename S, typename G>
static int launch_wave_passes(const SolverGeom& g, const SolveBatch& b,
                              hipStream_t s, Profiler* prof, int max_b) {
    const dim3 block(WaveLanes<L>::T);
    constexpr int PC = WavePairs<L>::P;
    const int nblk_c = ((g.N + 1) / 2 + PC - 1) / PC;
    {
        ProfScope ps(prof, "solve_pass_a", s, true);
        // one block per pair of grid rows
        launch_timed(ps, pass_a_wave_kernel<L, S, G>, dim3((g.N + 1) / 2, b.nfields),
                     block, 0, s, g, b);
    }
    LCODE_HIP(hipGetLastError());
    {
        ProfScope ps(prof, "solve_pass_b", s, true);
        launch_timed(ps, pass_b_wave_kernel<L, S>, dim3((max_b + 1) / 2, b.nfields),
                     block, 0, s, g, b);
    }
    LCODE_HIP(hipGetLastError());
    const dim3 gc(nblk_c, b.nfields);
    {
        ProfScope ps(prof, "solve_pass_c", s, true);
        launch_timed(ps, pass_c_wave_kernel<L, S, G>, gc, dim3(WaveLanes<L>::T * PC), 0, s,
                     g, b);
    }
    LCODE_HIP(hipGetLastError());
    if (b.resid) {
        ProfScope ps(prof, "resid_reduce", s);
        resid_reduce_kernel<<<1, 256, 0, s>>>(b.resid_part, (int)(gc.x * gc.y), b.resid);
        LCODE_HIP(hipGetLastError());
    }
    return LCODE_OK;
}

bool solver_wave_length(int L) {
    return LCODE_FFT_WAVE && (L == 1280 || L == 2560);
}

template <typename S, typename G>
static int launch_solve_t(const SolverGeom& g, const SolveBatch& b,
                          hipStream_t s, Profiler* prof) {
    const size_t lds = (size_t)solver_lds_bytes(g);
    const int per_block = 2 * g.P;
    int max_a = 0, max_b = 0;
    for (int f = 0; f < b.nfields; ++f) {
        max_a = max_a > b.f[f].rowsA ? max_a : b.f[f].rowsA;
        max_b = max_b > b.f[f].lenA ? max_b : b.f[f].lenA;
    }
    // the wave FFT runs the FP64 transforms and, with LCODE_FFT_WAVE_FP32, the
    // FP32 ones of L = 1280 (the workspace sets wave_tw only then)
    if constexpr (std::is_same<S, double>::value || LCODE_FFT_WAVE_FP32) if (g.wave_tw) {
        // the wave passes take a cosine transform of N points from column
        // (or line) 0 and a sine transform of N - 2 from 1
        for (int f = 0; f < b.nfields; ++f) {
            const FieldPlan& p = b.f[f];
            const bool ok_a = p.xfA == XF_DCT ? p.lenA == g.N && p.colA0 == 0
                                              : p.lenA == g.N - 2 && p.colA0 == 1;
            const bool ok_b = p.rowsA == (p.xfB == XF_DCT ? g.N : g.N - 2);
            if (!ok_a || !ok_b) {
                set_error("internal: field plan outside the wave FFT layout");
                return LCODE_E_INVALID;
            }
        }
        if (g.fft.L == 1280) return launch_wave_passes<1280, S, G>(g, b, s, prof, max_b);
        if (g.fft.L == 2560) return launch_wave_passes<2560, S, G>(g, b, s, prof, max_b);
    }
    const dim3 ga((max_a + per_block - 1) / per_block, b.nfields);
    const dim3 gb((max_b + per_block - 1) / per_block, b.nfields);
    const dim3 gc((g.N + per_block - 1) / per_block, b.nfields);
#define LCODE_CASE(LL)                                                          \
    case LL:                                                                    \
        return g.tw_lds                                                         \
                   ? launch_passes<LL, true, S, G>(g, b, s, prof, lds, ga, gb, gc) \
                   : launch_passes<LL, false, S, G>(g, b, s, prof, lds, ga, gb, gc);
    switch (g.static_L) {
        LCODE_CASE(32) LCODE_CASE(64) LCODE_CASE(128) LCODE_CASE(256)
        LCODE_CASE(512) LCODE_CASE(640) LCODE_CASE(1024) LCODE_CASE(1280)
        LCODE_CASE(2048) LCODE_CASE(2560)
        default:
            return launch_passes<0, false, S, G>(g, b, s, prof, lds, ga, gb, gc);
    }
#undef LCODE_CASE
}

int launch_guess(const GuessArgs& p, hipStream_t s, Profiler* prof) {
    {
        ProfScope ps(prof, "guess", s);
        if (p.grid32)
            guess_kernel<float><<<(p.n + 255) / 256, 256, 0, s>>>(p);
        else
            guess_kernel<double><<<(p.n + 255) / 256, 256, 0, s>>>(p);
    }
    LCODE_HIP(hipGetLastError());
    return LCODE_OK;
}

int launch_solve(const SolverGeom& g, const SolveBatch& b, hipStream_t s,
                 Profiler* prof) {
    if (b.nfields == 0) return LCODE_OK;
    // FP32 grids come only with FP32 transforms (the mixed form)
    if (b.grid32 && !g.fp32) {
        set_error("FP32 grids need the FP32 transforms");
        return LCODE_E_INVALID;
    }
    if (b.grid32) return launch_solve_t<float, float>(g, b, s, prof);
    return g.fp32 ? launch_solve_t<float, double>(g, b, s, prof)
                  : launch_solve_t<double, double>(g, b, s, prof);
}

}  // namespace lcode
