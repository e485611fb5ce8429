// Field solves as dense N x N sine/cosine-matrix contractions on MFMA.
//
// The same solves as fields.hip (lcode.py:122-364), written as four matrix
// products per field: with T_x, T_y the dense type-1 transform matrices
// along x and y,
//   P1 = T_x R,  P2 = (P1 T_y^T) o Mat,  P3 = T_x P2,  out = P3 T_y^T,
// batched over the six fields.  Two arithmetic forms:
//   DENSE_F64    v_mfma_f64_16x16x4_f64 on FP64 data;
//   DENSE_BF16X3 v_mfma_f32_16x16x32_bf16 on hi/lo bf16 splits of both
//                operands, three products (lo hi + hi lo + hi hi) accumulated
//                in FP32; the RHS, Bz zero mode and mean, Variant A and the
//                averages stay FP64.
// Tiles: 64 x 64 outputs per 256-thread block, four waves of 32 x 32
// (2 x 2 MFMA tiles of 16 x 16), operands staged through LDS.

#include "common.hpp"

namespace lcode {

namespace {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int kBM = 64, kBN = 64;

// Loads of the logical operands with their offsets and bounds (zero
// outside): A(r, k) = A[(r - a_off) * lda + k], B(k, c) = B[k * ldb + c - b_off].
__device__ __forceinline__ double load_a(const GemmOp& o, int r, int k) {
    const int rr = r - o.a_off;
    return (rr >= 0 && rr < o.a_rows && k < o.K) ? o.A[(size_t)rr * o.lda + k] : 0.;
}
__device__ __forceinline__ double load_b(const GemmOp& o, int k, int c) {
    const int cc = c - o.b_off;
    return (cc >= 0 && cc < o.b_cols && k < o.K) ? o.B[(size_t)k * o.ldb + cc] : 0.;
}

// Per-thread sums of the epilogues, reduced once per wave.
struct EpiAcc {
    double mean;               // mode 1, Bz: colsum^T P2 colsum
    double rn, rd;             // mode 2, guided step: sum (F - guess)^2, sum F^2
};

// Epilogues: 0 store; 1 multiply by the eigenvalue table (and, for Bz,
// accumulate the mean); 2 the final field write (Bz mean, Variant A,
// averages, guided residual) over the whole N x N grid, like the last pass
// of the FFT solver.
__device__ __forceinline__ void epilogue(const GemmOp& o, int r, int c, double v,
                                         EpiAcc& acc) {
    if (r >= o.M || c >= o.N) return;
    if (o.mode == 0) {
        o.C[(size_t)r * o.ldc + c] = v;
    } else if (o.mode == 1) {
        v *= o.tab[(size_t)r * o.ldt + c];
        o.C[(size_t)r * o.ldc + c] = v;
        if (o.mean_acc) acc.mean += o.colsum[r] * v * o.colsum[c];
    } else {
        const size_t idx = (size_t)r * o.N + c;   // output grid, M = N = grid
        double F = v;
        if (o.mean_in) F -= *o.mean_in / ((double)o.N * o.N);   // lcode.py:362
        if (o.variantA) F = 2 * F - o.prevF[idx];               // lcode.py:869-871
        if (o.outF) o.outF[idx] = F;
        if (o.outAvg) o.outAvg[idx] = (F + o.prevF[idx]) / 2;   // lcode.py:876-881
        if (o.guess) {
            const double d = F - o.guess[idx];
            acc.rn += d * d;
            acc.rd += F * F;
        }
    }
}

__device__ __forceinline__ double wave_sum64(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// DENSE_F64: FP64 operands with offsets and zero fill, staged through LDS.
__global__ void __launch_bounds__(256)
gemm_kernel(GemmBatch batch) {
    if (!gate_open(batch.gate)) return;
    const GemmOp& o = batch.op[blockIdx.z];
    const int m0 = blockIdx.y * kBM, n0 = blockIdx.x * kBN;
    if (m0 >= o.M || n0 >= o.N) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    const int li = lane & 15, lk = lane >> 4;
    EpiAcc ea{0., 0., 0.};
    {
        constexpr int BK = 16;
        __shared__ double As[kBM][BK + 1];
        __shared__ double Bs[BK][kBN + 1];
        v4d acc[2][2];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) acc[a][b] = v4d{0., 0., 0., 0.};
        for (int k0 = 0; k0 < o.K; k0 += BK) {
            for (int e = threadIdx.x; e < kBM * BK; e += 256) {
                const int r = e / BK, k = e % BK;
                As[r][k] = load_a(o, m0 + r, k0 + k);
            }
            for (int e = threadIdx.x; e < BK * kBN; e += 256) {
                const int k = e / kBN, c = e % kBN;
                Bs[k][c] = load_b(o, k0 + k, n0 + c);
            }
            __syncthreads();
#pragma unroll
            for (int kk = 0; kk < BK; kk += 4) {
                double fa[2], fb[2];
#pragma unroll
                for (int t = 0; t < 2; ++t) {
                    fa[t] = As[wm + t * 16 + li][kk + lk];
                    fb[t] = Bs[kk + lk][wn + t * 16 + li];
                }
#pragma unroll
                for (int a = 0; a < 2; ++a)
#pragma unroll
                    for (int b = 0; b < 2; ++b)
                        acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(
                            fa[a], fb[b], acc[a][b], 0, 0, 0);
            }
            __syncthreads();
        }
        // v_mfma_f64_16x16x4f64 returns D[i][j] in lane j + 16 (i % 4),
        // item i / 4 (measured on gfx950)
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    epilogue(o, m0 + wm + a * 16 + 4 * q + lk,
                             n0 + wn + b * 16 + li, acc[a][b][q], ea);
    }
    const int wlane = threadIdx.x & 63;
    if (o.mode == 1 && o.mean_acc) {
        const double m = wave_sum64(ea.mean);
        if (wlane == 0) atomicAdd(o.mean_acc, m);
    }
    if (o.mode == 2 && o.guess && batch.resid) {
        // one pair of atomics per block
        __shared__ double rs[2][4];
        const double rn = wave_sum64(ea.rn), rd = wave_sum64(ea.rd);
        if (wlane == 0) {
            rs[0][wave] = rn;
            rs[1][wave] = rd;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            atomicAdd(&batch.resid[0], rs[0][0] + rs[0][1] + rs[0][2] + rs[0][3]);
            atomicAdd(&batch.resid[1], rs[1][0] + rs[1][1] + rs[1][2] + rs[1][3]);
        }
    }
}

// RHS of the six fields into dense matrices R[f] of rowsA x lenA (rows
// i = rowsA0 + r, columns j = colA0 + c), FP64; zeroes the Bz mean sum.
__global__ void __launch_bounds__(256)
dense_rhs_kernel(DenseRhsArgs a) {
    if (!gate_open(a.gate)) return;
    const int f = blockIdx.y;
    const FieldPlan& fp = a.f[f];
    const int total = fp.rowsA * fp.lenA;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (f == 0 && e == 0 && a.mean_zero) *a.mean_zero = 0.;
    if (e >= total) return;
    const int r = e / fp.lenA, c = e - r * fp.lenA;
    a.R[f][e] = rhs_value_dev(fp.kind, fp.rowsA0 + r, fp.colA0 + c, a.N, a.in);
}


// ---------------------------------------------------------------------------
// DENSE_BF16X3 on pre-split planes.  Every operand lives in HBM as a pair of
// BF16 planes (hi, lo) of a zero-padded Np x Np array (Np = N rounded up to
// 128), in the layout its consumer reads with 16-byte loads: an A operand
// row-major (k contiguous), a B operand column by column (k contiguous).
// The products run in the order that lets the RHS be written row-major:
//   RHS kernel -> X (rows)              G1: Q = X T_A^T            (columns)
//   G2: P = (T_B Q) o table  (rows)     G3: Q = P T_A^T            (columns)
//   G4: field = T_B Q, FP64 epilogue on the N x N grid.
// Plane slots are per field kind, so each slot always holds one shape and
// every value outside it stays zero: the main loop has no bounds checks and
// no conversions.  Block tile (32 TM) x (32 TN), four waves of (16 TM) x
// (16 TN), K step 32, LDS double buffer.
// ---------------------------------------------------------------------------

typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

enum BfMode { BF_COLS = 0, BF_TAB_ROWS = 1, BF_FINAL = 2 };

#ifndef LCODE_BF_TM
#define LCODE_BF_TM 2
#endif
#ifndef LCODE_BF_TN
#define LCODE_BF_TN 2
#endif

struct BfOp {
    const uint16_t *Ah, *Al;   // A(r, k) = A[r ld + k]
    const uint16_t *Bh, *Bl;   // B(k, c) = B[c ld + k]
    uint16_t *Ch, *Cl;         // BF_COLS: C[c ld + r]; BF_TAB_ROWS: C[r ld + c]
    int ld;                    // Np
    int M, N;                  // valid output extent
    int Mp, Npad;              // padded output extent (multiples of the tile)
    int K;                     // padded reduction length (multiple of 32)
    int mode;
    const double* tab; int ldt;    // BF_TAB_ROWS: eigenvalue table [M][N]
    const double* colsum;          // ... and, for Bz, the mean sum
    double* mean_acc;
    int row0, col0, Ngrid;         // BF_FINAL: placement on the grid
    const double* mean_in;
    int variantA;
    const double* prevF;
    double *outF, *outAvg;
    const double* guess;
};

struct BfBatch {
    int n;
    BfOp op[6];
    Gate gate;
    double* resid;
    int tiles_m, tiles_n;      // block tiles per field (the largest field)
};

__device__ __forceinline__ uint16_t bf_bits(__bf16 v) {
    return __builtin_bit_cast(uint16_t, v);
}

// hi = RNE bf16 of v, lo = RNE bf16 of the remainder: |v - hi - lo| <= 2^-18 |v|
__device__ __forceinline__ void bf_split(double v, uint16_t& h, uint16_t& l) {
    const __bf16 hi = (__bf16)(float)v;
    const __bf16 lo = (__bf16)(float)(v - (double)(float)hi);
    h = bf_bits(hi);
    l = bf_bits(lo);
}

// Finished field value F at grid cell idx (the last pass of the FFT
// solver does the same): Bz mean, Variant A, outputs, guided residual.
__device__ __forceinline__ void final_cell(const BfOp& o, size_t idx, double F,
                                           EpiAcc& acc) {
    if (o.mean_in) F -= *o.mean_in / ((double)o.Ngrid * o.Ngrid);   // lcode.py:362
    if (o.variantA) F = 2 * F - o.prevF[idx];                       // lcode.py:869-871
    if (o.outF) o.outF[idx] = F;
    if (o.outAvg) o.outAvg[idx] = (F + o.prevF[idx]) / 2;           // lcode.py:876-881
    if (o.guess) {
        const double d = F - o.guess[idx];
        acc.rn += d * d;
        acc.rd += F * F;
    }
}

constexpr int kBfLds = 40;   // LDS row stride in bf16 elements (32 + 8)

template <int TM, int TN>
__global__ void __launch_bounds__(256)
bf_gemm_kernel(BfBatch batch) {
    constexpr int BM = 32 * TM, BN = 32 * TN;
    constexpr int PA = BM * 4 / 256, PB = BN * 4 / 256;   // 16-byte pieces per thread
    if (!gate_open(batch.gate)) return;
    // 1-D grid of n x tiles_m x tiles_n blocks in XCD-aware order: each XCD
    // takes a contiguous run of one field's tiles, row of tiles by row of
    // tiles, so the A rows and B columns it reads stay in its L2
    const int tiles_n = batch.tiles_n, per_field = batch.tiles_m * tiles_n;
    const int logical = xcd_logical(blockIdx.x, gridDim.x);
    const int z = logical / per_field, t = logical - z * per_field;
    const BfOp& o = batch.op[z];
    const int m0 = (t / tiles_n) * BM, n0 = (t % tiles_n) * BN;
    if (m0 >= o.Mp || n0 >= o.Npad) return;
    __shared__ __align__(16) uint16_t sA[2][2][BM][kBfLds];   // [stage][hi/lo]
    __shared__ __align__(16) uint16_t sB[2][2][BN][kBfLds];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wm = (wave >> 1) * 16 * TM, wn = (wave & 1) * 16 * TN;
    const int li = lane & 15, lk = lane >> 4;
    size_t offA[PA], offB[PB];
    int rowA[PA], rowB[PB], colP = (tid & 3) * 8;
#pragma unroll
    for (int p = 0; p < PA; ++p) {
        rowA[p] = (tid >> 2) + 64 * p;
        offA[p] = (size_t)(m0 + rowA[p]) * o.ld + colP;
    }
#pragma unroll
    for (int p = 0; p < PB; ++p) {
        rowB[p] = (tid >> 2) + 64 * p;
        offB[p] = (size_t)(n0 + rowB[p]) * o.ld + colP;
    }
    uint4 rAh[PA], rAl[PA], rBh[PB], rBl[PB];
    auto fetch = [&](int k0) {
#pragma unroll
        for (int p = 0; p < PA; ++p) {
            rAh[p] = *reinterpret_cast<const uint4*>(o.Ah + offA[p] + k0);
            rAl[p] = *reinterpret_cast<const uint4*>(o.Al + offA[p] + k0);
        }
#pragma unroll
        for (int p = 0; p < PB; ++p) {
            rBh[p] = *reinterpret_cast<const uint4*>(o.Bh + offB[p] + k0);
            rBl[p] = *reinterpret_cast<const uint4*>(o.Bl + offB[p] + k0);
        }
    };
    auto stash = [&](int st) {
#pragma unroll
        for (int p = 0; p < PA; ++p) {
            *reinterpret_cast<uint4*>(&sA[st][0][rowA[p]][colP]) = rAh[p];
            *reinterpret_cast<uint4*>(&sA[st][1][rowA[p]][colP]) = rAl[p];
        }
#pragma unroll
        for (int p = 0; p < PB; ++p) {
            *reinterpret_cast<uint4*>(&sB[st][0][rowB[p]][colP]) = rBh[p];
            *reinterpret_cast<uint4*>(&sB[st][1][rowB[p]][colP]) = rBl[p];
        }
    };
    fetch(0);
    stash(0);
    __syncthreads();
    v4f acc[TM][TN];
#pragma unroll
    for (int a = 0; a < TM; ++a)
#pragma unroll
        for (int b = 0; b < TN; ++b) acc[a][b] = v4f{0.f, 0.f, 0.f, 0.f};
    const int nks = o.K / 32;
    for (int ks = 0; ks < nks; ++ks) {
        const int cur = ks & 1;
        const bool more = ks + 1 < nks;
        if (more) fetch((ks + 1) * 32);
        bf8 bh[TN], bl[TN];
#pragma unroll
        for (int b = 0; b < TN; ++b) {
            bh[b] = *reinterpret_cast<const bf8*>(&sB[cur][0][wn + b * 16 + li][8 * lk]);
            bl[b] = *reinterpret_cast<const bf8*>(&sB[cur][1][wn + b * 16 + li][8 * lk]);
        }
        bf8 ah[TM], al[TM];
#pragma unroll
        for (int a = 0; a < TM; ++a) {
            ah[a] = *reinterpret_cast<const bf8*>(&sA[cur][0][wm + a * 16 + li][8 * lk]);
            al[a] = *reinterpret_cast<const bf8*>(&sA[cur][1][wm + a * 16 + li][8 * lk]);
        }
        // product-major order: consecutive MFMAs write different accumulators
#pragma unroll
        for (int a = 0; a < TM; ++a)
#pragma unroll
            for (int b = 0; b < TN; ++b)
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[a], bh[b], acc[a][b], 0, 0, 0);
#pragma unroll
        for (int a = 0; a < TM; ++a)
#pragma unroll
            for (int b = 0; b < TN; ++b)
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], bl[b], acc[a][b], 0, 0, 0);
#pragma unroll
        for (int a = 0; a < TM; ++a)
#pragma unroll
            for (int b = 0; b < TN; ++b)
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], bh[b], acc[a][b], 0, 0, 0);
        if (more) stash(cur ^ 1);
        __syncthreads();
    }

    // v_mfma_f32_16x16x32_bf16 returns D[i][j] in lane j + 16 (i / 4), item i % 4
    EpiAcc ea{0., 0., 0.};
#pragma unroll
    for (int a = 0; a < TM; ++a)
#pragma unroll
        for (int b = 0; b < TN; ++b) {
            const int r0 = m0 + wm + a * 16 + 4 * lk;   // rows r0 .. r0 + 3
            const int c = n0 + wn + b * 16 + li;
            if (o.mode == BF_COLS) {
                uint16_t h[4], l[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    h[q] = l[q] = 0;
                    if (r0 + q < o.M && c < o.N) bf_split((double)acc[a][b][q], h[q], l[q]);
                }
                const size_t at = (size_t)c * o.ld + r0;
                *reinterpret_cast<ushort4*>(o.Ch + at) = make_ushort4(h[0], h[1], h[2], h[3]);
                *reinterpret_cast<ushort4*>(o.Cl + at) = make_ushort4(l[0], l[1], l[2], l[3]);
            } else if (o.mode == BF_TAB_ROWS) {
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int r = r0 + q;
                    uint16_t h = 0, l = 0;
                    if (r < o.M && c < o.N) {
                        const double v = (double)acc[a][b][q] * o.tab[(size_t)r * o.ldt + c];
                        if (o.mean_acc) ea.mean += o.colsum[r] * v * o.colsum[c];
                        bf_split(v, h, l);
                    }
                    o.Ch[(size_t)r * o.ld + c] = h;
                    o.Cl[(size_t)r * o.ld + c] = l;
                }
            } else {
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int r = r0 + q;
                    if (r < o.M && c < o.N)
                        final_cell(o, (size_t)(o.row0 + r) * o.Ngrid + o.col0 + c,
                                   (double)acc[a][b][q], ea);
                }
            }
        }
    const int wlane = tid & 63;
    if (o.mode == BF_TAB_ROWS && o.mean_acc) {
        const double m = wave_sum64(ea.mean);
        if (wlane == 0) atomicAdd(o.mean_acc, m);
    }
    if (o.mode == BF_FINAL && o.guess && batch.resid) {
        // one pair of atomics per block
        __shared__ double rs[2][4];
        const double rn = wave_sum64(ea.rn), rd = wave_sum64(ea.rd);
        if (wlane == 0) {
            rs[0][wave] = rn;
            rs[1][wave] = rd;
        }
        __syncthreads();
        if (tid == 0) {
            atomicAdd(&batch.resid[0], rs[0][0] + rs[0][1] + rs[0][2] + rs[0][3]);
            atomicAdd(&batch.resid[1], rs[1][0] + rs[1][1] + rs[1][2] + rs[1][3]);
        }
    }
}

// The right-hand sides of all requested fields in one pass over the grid
// (lcode.py:166-168, :234-268, :339-343, the same arithmetic as the FFT
// solver's), written as BF16 planes X[r][c] of the value at grid
// (rowsA0 + r, colA0 + c).  A 16 x 32-cell tile stages the four stencil
// sources (rho + beam, jz + beam, jx, jy) with a one-cell halo in LDS; the
// centre-only inputs are read directly.  Extra blocks give the grid cells
// outside each field's data region (the zero rim of Ez, Ex, Ey, Bx, By)
// the final-pass treatment of F = 0.
struct BfRhsArgs {
    int nfields;
    FieldPlan f[6];
    RhsInputs in;
    int N, Np;
    uint16_t *Xh[6], *Xl[6];
    double* mean_zero;
    Gate gate;
    int variantA;
    const double* guess[6];
    double* resid;
    int need_rho, need_jz, need_j;   // which stencil sources the fields use
};

constexpr int kRhsTI = 16, kRhsTJ = 32;

constexpr int kRhsThreads = kRhsTI * kRhsTJ;   // one cell per thread

__global__ void __launch_bounds__(kRhsThreads)
bf_rhs_kernel(BfRhsArgs a) {
    if (!gate_open(a.gate)) return;
    __shared__ double s_rho[kRhsTI + 2][kRhsTJ + 2], s_jz[kRhsTI + 2][kRhsTJ + 2];
    __shared__ double s_jx[kRhsTI + 2][kRhsTJ + 2], s_jy[kRhsTI + 2][kRhsTJ + 2];
    const int N = a.N;
    const int tid = threadIdx.x;
    const int tj_n = (N + kRhsTJ - 1) / kRhsTJ, ti_n = (N + kRhsTI - 1) / kRhsTI;
    const int ntiles = tj_n * ti_n;
    const RhsInputs& in = a.in;
    if (blockIdx.x == 0 && tid == 0 && a.mean_zero) *a.mean_zero = 0.;
    if ((int)blockIdx.x < ntiles) {
        const int i0 = (blockIdx.x / tj_n) * kRhsTI, j0 = (blockIdx.x % tj_n) * kRhsTJ;
        constexpr int kStage = (kRhsTI + 2) * (kRhsTJ + 2);
        constexpr int kIter = (kStage + kRhsThreads - 1) / kRhsThreads;
        double r[kIter], z[kIter], x[kIter], y[kIter];
#pragma unroll
        for (int it = 0; it < kIter; ++it) {   // all loads issued before any use
            const int e = tid + kRhsThreads * it;
            const int ii = e / (kRhsTJ + 2), jj = e - ii * (kRhsTJ + 2);
            const int i = i0 - 1 + ii, j = j0 - 1 + jj;
            r[it] = z[it] = x[it] = y[it] = 0.;
            if (e < kStage && i >= 0 && i < N && j >= 0 && j < N) {
                const size_t idx = (size_t)i * N + j;
                if (a.need_rho) r[it] = rho_src(in, idx);
                if (a.need_jz) z[it] = jz_src(in, idx);
                if (a.need_j) { x[it] = in.jx[idx]; y[it] = in.jy[idx]; }
            }
        }
#pragma unroll
        for (int it = 0; it < kIter; ++it) {
            const int e = tid + kRhsThreads * it;
            if (e >= kStage) break;
            const int ii = e / (kRhsTJ + 2), jj = e - ii * (kRhsTJ + 2);
            s_rho[ii][jj] = r[it];
            s_jz[ii][jj] = z[it];
            s_jx[ii][jj] = x[it];
            s_jy[ii][jj] = y[it];
        }
        __syncthreads();
        const double h2 = in.h * 2;
        {
            const int ti = tid / kRhsTJ, tj = tid % kRhsTJ;
            const int i = i0 + ti, j = j0 + tj;
            if (i >= N || j >= N) return;
            const size_t idx = (size_t)i * N + j;
            const int li = ti + 1, lj = tj + 1;
            const bool inner = i >= 1 && i <= N - 2 && j >= 1 && j <= N - 2;
            for (int f = 0; f < a.nfields; ++f) {
                const FieldPlan& fp = a.f[f];
                const int r = i - fp.rowsA0, c = j - fp.colA0;
                if (r < 0 || r >= fp.rowsA || c < 0 || c >= fp.lenA) continue;
                double v;
                switch (fp.kind) {
                    case F_EX: {
                        const double d = inner ? (s_rho[li + 1][lj] - s_rho[li - 1][lj]) / h2 : 0.;
                        const double dj = (in.jx_prev[idx] - in.jx[idx]) / in.dxi;
                        v = -((d - dj) - in.Ftrick[0][idx] * in.trick);
                        break;
                    }
                    case F_EY: {
                        const double d = inner ? (s_rho[li][lj + 1] - s_rho[li][lj - 1]) / h2 : 0.;
                        const double dj = (in.jy_prev[idx] - in.jy[idx]) / in.dxi;
                        v = -((d - dj) - in.Ftrick[1][idx] * in.trick);
                        break;
                    }
                    case F_BX: {
                        const double d = inner ? (s_jz[li][lj + 1] - s_jz[li][lj - 1]) / h2 : 0.;
                        const double dj = (in.jy_prev[idx] - in.jy[idx]) / in.dxi;
                        v = +((d - dj) + in.Ftrick[2][idx] * in.trick);
                        break;
                    }
                    case F_BY: {
                        const double d = inner ? (s_jz[li + 1][lj] - s_jz[li - 1][lj]) / h2 : 0.;
                        const double dj = (in.jx_prev[idx] - in.jx[idx]) / in.dxi;
                        v = -((d - dj) - in.Ftrick[3][idx] * in.trick);
                        break;
                    }
                    case F_EZ:
                        v = -((s_jx[li + 1][lj] - s_jx[li - 1][lj]) +
                              (s_jy[li][lj + 1] - s_jy[li][lj - 1])) / h2;
                        break;
                    default: {
                        const double djx_dy = inner ? s_jx[li][lj + 1] - s_jx[li][lj - 1] : 0.;
                        const double djy_dx = inner ? s_jy[li + 1][lj] - s_jy[li - 1][lj] : 0.;
                        v = -(djx_dy - djy_dx) / h2;
                    }
                }
                uint16_t hb, lb;
                bf_split(v, hb, lb);
                a.Xh[f][(size_t)r * a.Np + c] = hb;
                a.Xl[f][(size_t)r * a.Np + c] = lb;
            }
        }
        return;
    }
    // rim: cells of the four grid edges outside each field's data region
    const int e = ((int)blockIdx.x - ntiles) * kRhsThreads + tid;
    const int side = e / N, t = e - side * N;
    int i = 0, j = 0;
    bool ok = e < 4 * N;
    if (side < 2) { i = side ? N - 1 : 0; j = t; }
    else { i = t; j = side == 3 ? N - 1 : 0; ok = ok && t > 0 && t < N - 1; }
    double rn = 0., rd = 0.;
    for (int f = 0; f < a.nfields; ++f) {
        const FieldPlan& fp = a.f[f];
        const bool inside = i >= fp.rowsA0 && i < fp.rowsA0 + fp.rowsA &&
                            j >= fp.colA0 && j < fp.colA0 + fp.lenA;
        if (!ok || inside) continue;
        const size_t idx = (size_t)i * N + j;
        const bool transverse = fp.kind == F_EX || fp.kind == F_EY ||
                                fp.kind == F_BX || fp.kind == F_BY;
        double F = 0.;
        if (a.variantA && transverse) F = -fp.prevF[idx];   // 2 * 0 - prev
        if (fp.outF) fp.outF[idx] = F;
        if (fp.outAvg) fp.outAvg[idx] = (F + fp.prevF[idx]) / 2;
        const double* g = a.resid ? a.guess[fp.kind] : nullptr;
        if (g) {
            const double d = F - g[idx];
            rn += d * d;
            rd += F * F;
        }
    }
    if (a.resid) {
        rn = wave_sum64(rn);
        rd = wave_sum64(rd);
        if ((tid & 63) == 0 && (rn != 0. || rd != 0.)) {
            atomicAdd(&a.resid[0], rn);
            atomicAdd(&a.resid[1], rd);
        }
    }
}

int launch_solve_bf16x3(int N, const DenseGeom& d, const SolveBatch& b,
                        hipStream_t s, Profiler* prof) {
    constexpr int TM = LCODE_BF_TM, TN = LCODE_BF_TN;
    constexpr int BM = 32 * TM, BN = 32 * TN;
    const int Np = d.Np;
    const size_t np2 = (size_t)Np * Np;
    auto up = [](int x, int m) { return (x + m - 1) / m * m; };
    BfRhsArgs ra{};
    ra.nfields = b.nfields;
    ra.in = b.in;
    ra.N = N;
    ra.Np = Np;
    ra.mean_zero = d.mean;
    ra.gate = b.gate;
    ra.variantA = b.variantA;
    ra.resid = b.resid;
    for (int k = 0; k < 6; ++k) ra.guess[k] = b.guess[k];
    BfBatch g1{}, g2{}, g3{}, g4{};
    g1.n = g2.n = g3.n = g4.n = b.nfields;
    g1.gate = g2.gate = g3.gate = g4.gate = b.gate;
    g4.resid = b.resid;
    int maxM = 0, maxN = 0;
    for (int f = 0; f < b.nfields; ++f) {
        const FieldPlan& fp = b.f[f];
        const int R = fp.rowsA, C = fp.lenA;
        const int kind = fp.kind;
        ra.f[f] = fp;
        if (kind == F_EX || kind == F_EY) ra.need_rho = 1;
        if (kind == F_BX || kind == F_BY) ra.need_jz = 1;
        if (kind == F_EZ || kind == F_BZ) ra.need_j = 1;
        // plane slot = field kind: one shape per slot, zeros elsewhere
        uint16_t *Xh = d.Xh + kind * np2, *Xl = d.Xl + kind * np2;
        uint16_t *Qh = d.Ph + kind * np2, *Ql = d.Pl + kind * np2;
        ra.Xh[f] = Xh;
        ra.Xl[f] = Xl;
        const bool dstB = fp.xfB == XF_DST, dstA = fp.xfA == XF_DST;
        const uint16_t* TBh = dstB ? d.Sh : d.Ch;
        const uint16_t* TBl = dstB ? d.Sl : d.Cl;
        const uint16_t* TAh = dstA ? d.Sh : d.Ch;
        const uint16_t* TAl = dstA ? d.Sl : d.Cl;
        BfOp o{};
        o.ld = Np; o.M = R; o.N = C; o.Mp = up(R, BM); o.Npad = up(C, BN);
        maxM = maxM > o.Mp ? maxM : o.Mp;
        maxN = maxN > o.Npad ? maxN : o.Npad;
        // G1: Q = X T_A^T, stored by columns
        BfOp o1 = o;
        o1.Ah = Xh; o1.Al = Xl; o1.Bh = TAh; o1.Bl = TAl; o1.K = up(C, 32);
        o1.mode = BF_COLS; o1.Ch = Qh; o1.Cl = Ql;
        g1.op[f] = o1;
        // G2: P = (T_B Q) o table, stored by rows (over X: X is consumed)
        BfOp o2 = o;
        o2.Ah = TBh; o2.Al = TBl; o2.Bh = Qh; o2.Bl = Ql; o2.K = up(R, 32);
        o2.mode = BF_TAB_ROWS; o2.Ch = Xh; o2.Cl = Xl;
        o2.tab = kind == F_EY || kind == F_BX ? d.mixed
               : kind == F_EX || kind == F_BY ? d.mixed_t
               : static_cast<const double*>(fp.table);
        o2.ldt = C;
        if (kind == F_BZ) { o2.colsum = d.colsum; o2.mean_acc = d.mean; }
        g2.op[f] = o2;
        // G3: Q = P T_A^T, stored by columns
        BfOp o3 = o1;
        g3.op[f] = o3;
        // G4: field = T_B Q on the grid
        BfOp o4 = o;
        o4.Ah = TBh; o4.Al = TBl; o4.Bh = Qh; o4.Bl = Ql; o4.K = up(R, 32);
        o4.mode = BF_FINAL;
        o4.row0 = fp.rowsA0; o4.col0 = fp.colA0; o4.Ngrid = N;
        if (kind == F_BZ) o4.mean_in = d.mean;
        const bool transverse = kind == F_EX || kind == F_EY ||
                                kind == F_BX || kind == F_BY;
        o4.variantA = b.variantA && transverse;
        o4.prevF = fp.prevF;
        o4.outF = fp.outF;
        o4.outAvg = fp.outAvg;
        o4.guess = b.resid ? b.guess[kind] : nullptr;
        g4.op[f] = o4;
    }
    const int tiles = ((N + kRhsTJ - 1) / kRhsTJ) * ((N + kRhsTI - 1) / kRhsTI);
    const dim3 grid_rhs(tiles + (4 * N + kRhsThreads - 1) / kRhsThreads);
    g1.tiles_m = g2.tiles_m = g3.tiles_m = g4.tiles_m = maxM / BM;
    g1.tiles_n = g2.tiles_n = g3.tiles_n = g4.tiles_n = maxN / BN;
    const dim3 grid_mm((maxM / BM) * (maxN / BN) * b.nfields);
    {
        ProfScope ps(prof, "dense_rhs", s, true);
        launch_timed(ps, bf_rhs_kernel, grid_rhs, dim3(kRhsThreads), 0, s, ra);
    }
    LCODE_HIP(hipGetLastError());
    const BfBatch* gs[4] = {&g1, &g2, &g3, &g4};
    const char* names[4] = {"dense_gemm_1", "dense_gemm_2", "dense_gemm_3", "dense_gemm_4"};
    for (int k = 0; k < 4; ++k) {
        {
            ProfScope ps(prof, names[k], s, true);
            launch_timed(ps, bf_gemm_kernel<TM, TN>, grid_mm, dim3(256), 0, s, *gs[k]);
        }
        LCODE_HIP(hipGetLastError());
    }
    return LCODE_OK;
}

}  // namespace

int launch_gemm(const GemmBatch& b, hipStream_t s, Profiler* prof,
                const char* name) {
    int maxM = 0, maxN = 0;
    for (int k = 0; k < b.n; ++k) {
        maxM = maxM > b.op[k].M ? maxM : b.op[k].M;
        maxN = maxN > b.op[k].N ? maxN : b.op[k].N;
    }
    const dim3 grid((maxN + kBN - 1) / kBN, (maxM + kBM - 1) / kBM, b.n);
    {
        ProfScope ps(prof, name, s, true);
        launch_timed(ps, gemm_kernel, grid, dim3(256), 0, s, b);
    }
    LCODE_HIP(hipGetLastError());
    return LCODE_OK;
}

int launch_dense_rhs(const DenseRhsArgs& a, int nfields, int max_total,
                     hipStream_t s, Profiler* prof) {
    const dim3 grid((max_total + 255) / 256, nfields);
    {
        ProfScope ps(prof, "dense_rhs", s, true);
        launch_timed(ps, dense_rhs_kernel, grid, dim3(256), 0, s, a);
    }
    LCODE_HIP(hipGetLastError());
    return LCODE_OK;
}

int launch_solve_dense(int N, const DenseGeom& d, const SolveBatch& b,
                       hipStream_t s, Profiler* prof) {
    if (b.nfields == 0) return LCODE_OK;
    if (d.prec == DENSE_BF16X3) return launch_solve_bf16x3(N, d, b, s, prof);
    const size_t n2 = (size_t)N * N;
    DenseRhsArgs ra{};
    ra.in = b.in;
    ra.N = N;
    ra.mean_zero = d.mean;
    ra.gate = b.gate;
    int max_total = 0;
    GemmBatch g1{}, g2{}, g3{}, g4{};
    g1.n = g2.n = g3.n = g4.n = b.nfields;
    g1.gate = g2.gate = g3.gate = g4.gate = b.gate;
    g4.resid = b.resid;
    for (int f = 0; f < b.nfields; ++f) {
        const FieldPlan& fp = b.f[f];
        const int R = fp.rowsA, C = fp.lenA;
        ra.f[f] = fp;
        ra.R[f] = d.X + f * n2;
        max_total = max_total > R * C ? max_total : R * C;
        // rows: P = T_B P (R x R); columns: P = P T_A^T (C x C)
        const double* TB = fp.xfB == XF_DST ? d.Sd : d.Cd;
        const double* TAt = fp.xfA == XF_DST ? d.Sd : d.CdT;
        double* X = d.X + f * n2;
        double* Y = d.Y + f * n2;
        const double* tab = fp.kind == F_EY || fp.kind == F_BX ? d.mixed
                          : fp.kind == F_EX || fp.kind == F_BY ? d.mixed_t
                          : static_cast<const double*>(fp.table);
        GemmOp o{};
        o.a_off = 0; o.b_off = 0;
        // P1 = T_B R
        o.A = TB; o.lda = R; o.a_rows = R;
        o.B = X; o.ldb = C; o.b_cols = C;
        o.C = Y; o.ldc = C; o.M = R; o.N = C; o.K = R; o.mode = 0;
        g1.op[f] = o;
        // P2 = (P1 T_A^T) o table
        o.A = Y; o.lda = C; o.a_rows = R;
        o.B = TAt; o.ldb = C; o.b_cols = C;
        o.C = X; o.ldc = C; o.M = R; o.N = C; o.K = C; o.mode = 1;
        o.tab = tab; o.ldt = C;
        if (fp.kind == F_BZ) { o.colsum = d.colsum; o.mean_acc = d.mean; }
        g2.op[f] = o;
        // P3 = T_B P2
        o = GemmOp{};
        o.A = TB; o.lda = R; o.a_rows = R;
        o.B = X; o.ldb = C; o.b_cols = C;
        o.C = Y; o.ldc = C; o.M = R; o.N = C; o.K = R; o.mode = 0;
        g3.op[f] = o;
        // field = P3 T_A^T placed at (rowsA0, colA0) of the N x N grid
        o = GemmOp{};
        o.A = Y; o.lda = C; o.a_off = fp.rowsA0; o.a_rows = R;
        o.B = TAt; o.ldb = C; o.b_off = fp.colA0; o.b_cols = C;
        o.M = N; o.N = N; o.K = C; o.mode = 2;
        if (fp.kind == F_BZ) o.mean_in = d.mean;
        const bool transverse = fp.kind == F_EX || fp.kind == F_EY ||
                                fp.kind == F_BX || fp.kind == F_BY;
        o.variantA = b.variantA && transverse;
        o.prevF = fp.prevF;
        o.outF = fp.outF;
        o.outAvg = fp.outAvg;
        o.guess = b.resid ? b.guess[fp.kind] : nullptr;
        g4.op[f] = o;
    }
    int rc;
    if ((rc = launch_dense_rhs(ra, b.nfields, max_total, s, prof))) return rc;
    if ((rc = launch_gemm(g1, s, prof, "dense_gemm_1"))) return rc;
    if ((rc = launch_gemm(g2, s, prof, "dense_gemm_2"))) return rc;
    if ((rc = launch_gemm(g3, s, prof, "dense_gemm_3"))) return rc;
    return launch_gemm(g4, s, prof, "dense_gemm_4");
}

}  // namespace lcode
