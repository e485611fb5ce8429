// Field solves as dense N x N sine/cosine-matrix contractions on MFMA.
//
// The same solves as fields.hip (lcode.py:122-364), written as four matrix
// products per field: with T_x, T_y the dense type-1 transform matrices
// along x and y,
//   P1 = T_x R,  P2 = (P1 T_y^T) o Mat,  P3 = T_x P2,  out = P3 T_y^T,
// batched over the six fields.  Two arithmetic forms:
//   DENSE_F64    v_mfma_f64_16x16x4_f64 on FP64 data (the plain path);
//   DENSE_BF16X3 v_mfma_f32_16x16x32_bf16 on hi/lo bf16 splits of both
//                operands, three products (hi hi + hi lo + lo hi) accumulated
//                in FP32; the RHS, Bz zero mode and mean, Variant A and the
//                averages stay FP64.
// Tiles: 64 x 64 outputs per 256-thread block, four waves of 32 x 32
// (2 x 2 MFMA tiles of 16 x 16), operands staged through LDS.

#include "common.hpp"

namespace lcode {

namespace {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

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

template <int PREC>
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
    if constexpr (PREC == DENSE_F64) {
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
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    epilogue(o, m0 + wm + a * 16 + 4 * lk + q,
                             n0 + wn + b * 16 + li, acc[a][b][q], ea);
    } else {
        constexpr int BK = 32;
        // hi/lo bf16 splits; B stored transposed so fragments are contiguous
        __shared__ __bf16 Ah[kBM][BK + 8], Al[kBM][BK + 8];
        __shared__ __bf16 Bh[kBN][BK + 8], Bl[kBN][BK + 8];
        v4f acc[2][2];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) acc[a][b] = v4f{0.f, 0.f, 0.f, 0.f};
        for (int k0 = 0; k0 < o.K; k0 += BK) {
            for (int e = threadIdx.x; e < kBM * BK; e += 256) {
                const int r = e / BK, k = e % BK;
                const double v = load_a(o, m0 + r, k0 + k);
                const __bf16 h = (__bf16)(float)v;
                Ah[r][k] = h;
                Al[r][k] = (__bf16)(float)(v - (double)(float)h);
            }
            for (int e = threadIdx.x; e < BK * kBN; e += 256) {
                const int k = e / kBN, c = e % kBN;
                const double v = load_b(o, k0 + k, n0 + c);
                const __bf16 h = (__bf16)(float)v;
                Bh[c][k] = h;
                Bl[c][k] = (__bf16)(float)(v - (double)(float)h);
            }
            __syncthreads();
            v8bf ah[2], al[2], bh[2], bl[2];
#pragma unroll
            for (int t = 0; t < 2; ++t)
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    ah[t][q] = Ah[wm + t * 16 + li][8 * lk + q];
                    al[t][q] = Al[wm + t * 16 + li][8 * lk + q];
                    bh[t][q] = Bh[wn + t * 16 + li][8 * lk + q];
                    bl[t][q] = Bl[wn + t * 16 + li][8 * lk + q];
                }
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b) {
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                        al[a], bh[b], acc[a][b], 0, 0, 0);
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                        ah[a], bl[b], acc[a][b], 0, 0, 0);
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                        ah[a], bh[b], acc[a][b], 0, 0, 0);
                }
            __syncthreads();
        }
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    epilogue(o, m0 + wm + a * 16 + 4 * lk + q,
                             n0 + wn + b * 16 + li, (double)acc[a][b][q], ea);
    }
    const int wlane = threadIdx.x & 63;
    if (o.mode == 1 && o.mean_acc) {
        const double m = wave_sum64(ea.mean);
        if (wlane == 0) atomicAdd(o.mean_acc, m);
    }
    if (o.mode == 2 && o.guess && batch.resid) {
        const double rn = wave_sum64(ea.rn), rd = wave_sum64(ea.rd);
        if (wlane == 0) {
            atomicAdd(&batch.resid[0], rn);
            atomicAdd(&batch.resid[1], rd);
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

}  // namespace

int launch_gemm(const GemmBatch& b, int prec, hipStream_t s, Profiler* prof,
                const char* name) {
    int maxM = 0, maxN = 0;
    for (int k = 0; k < b.n; ++k) {
        maxM = maxM > b.op[k].M ? maxM : b.op[k].M;
        maxN = maxN > b.op[k].N ? maxN : b.op[k].N;
    }
    const dim3 grid((maxN + kBN - 1) / kBN, (maxM + kBM - 1) / kBM, b.n);
    {
        ProfScope ps(prof, name, s);
        if (prec == DENSE_F64)
            gemm_kernel<DENSE_F64><<<grid, 256, 0, s>>>(b);
        else
            gemm_kernel<DENSE_BF16X3><<<grid, 256, 0, s>>>(b);
    }
    LCODE_HIP(hipGetLastError());
    return LCODE_OK;
}

int launch_dense_rhs(const DenseRhsArgs& a, int nfields, int max_total,
                     hipStream_t s, Profiler* prof) {
    const dim3 grid((max_total + 255) / 256, nfields);
    {
        ProfScope ps(prof, "dense_rhs", s);
        dense_rhs_kernel<<<grid, 256, 0, s>>>(a);
    }
    LCODE_HIP(hipGetLastError());
    return LCODE_OK;
}

int launch_solve_dense(int N, const DenseGeom& d, const SolveBatch& b,
                       hipStream_t s, Profiler* prof) {
    if (b.nfields == 0) return LCODE_OK;
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
    if ((rc = launch_gemm(g1, d.prec, s, prof, "dense_gemm_1"))) return rc;
    if ((rc = launch_gemm(g2, d.prec, s, prof, "dense_gemm_2"))) return rc;
    if ((rc = launch_gemm(g3, d.prec, s, prof, "dense_gemm_3"))) return rc;
    return launch_gemm(g4, d.prec, s, prof, "dense_gemm_4");
}

}  // namespace lcode
