// A small convolutional field predictor for the guided step (the learned
// form of the predictor hook, lcode.py:841-881): its own forward, backward
// and optimiser kernels, FP32 arithmetic on FP32 activations.
//
// Layout: channel planes [C][H][W] (row-major H x W grids, H = W = N).
// Layers are 3 x 3 convolutions with zero padding ("same" size); ReLU
// between layers is fused into the forward kernel and its derivative into
// the backward-data kernel.  At most kPredMaxC channels per layer.
//
//   conv3x3_fwd     out[co] = b[co] + sum_ci sum_k w[co][ci][k] in[ci](p + d_k)
//   conv3x3_bwd_in  gin[ci] = sum_co sum_k w[co][ci][k] gout[co](p - d_k),
//                   times (in[ci] > 0) when the input came out of a ReLU
//   conv3x3_bwd_w   gw[co][ci][k] = sum_p gout[co](p) in[ci](p + d_k),
//                   gb[co] = sum_p gout[co](p): per-tile partial sums, then
//                   a deterministic reduction over tiles
//   mse_grad        g = 2 (y - t) / count, per-tile partial sums of (y - t)^2
//   adam            the Adam update of a flat parameter buffer
//   pred_inputs / pred_targets / pred_guess: FP64 fields <-> normalised
//                   FP32 planes (the network predicts the change of the next
//                   slice's fields from their quadratic extrapolation).

#include "common.hpp"

namespace lcode {

namespace {

constexpr int kPredMaxC = 16;
constexpr int kT = 16;                 // 16 x 16 pixel tiles, 256 threads

struct ConvArgs {
    const float* in;  int cin;
    const float* w;   const float* b;    // w[cout][cin][9], b[cout]
    float* out;       int cout;
    int H, W;
    int relu;                            // forward: ReLU on the output;
                                         // backward-data: mask by in > 0
};

// Stage a (kT + 2)^2 tile of `c` planes with a one-pixel zero halo: eight
// loads per thread issued together (addresses clamped in bounds, values
// outside the planes replaced by zero), then their LDS stores.
__device__ __forceinline__ void stage_tile(float (*t)[kT + 2][kT + 2],
                                           const float* __restrict__ src, int c,
                                           int H, int W, int i0, int j0) {
    constexpr int kE = (kT + 2) * (kT + 2);
    constexpr int kU = 8;
    const int total = c * kE;
    for (int base = 0; base < total; base += 256 * kU) {
        float v[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int e = base + (int)threadIdx.x + u * 256;
            const int ch = e / kE, r = e - ch * kE;
            const int i = i0 - 1 + r / (kT + 2), j = j0 - 1 + r % (kT + 2);
            const bool ok = e < total && i >= 0 && i < H && j >= 0 && j < W;
            const float x = src[ok ? ((size_t)ch * H + i) * W + j : 0];
            v[u] = ok ? x : 0.f;
        }
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int e = base + (int)threadIdx.x + u * 256;
            if (e < total) {
                const int ch = e / kE, r = e - ch * kE;
                t[ch][r / (kT + 2)][r % (kT + 2)] = v[u];
            }
        }
    }
}

// CO = a.cout as a constant: the accumulators are unrolled registers.
template <int CO>
__global__ void __launch_bounds__(256)
conv3x3_fwd_kernel(ConvArgs a) {
    __shared__ float t[kPredMaxC][kT + 2][kT + 2];
    // the weights are the same for every thread: read through the scalar
    // path from global memory rather than once per multiply from LDS
    const float* __restrict__ w = a.w;
    const int tiles_w = (a.W + kT - 1) / kT;
    const int i0 = (blockIdx.x / tiles_w) * kT, j0 = (blockIdx.x % tiles_w) * kT;
    stage_tile(t, a.in, a.cin, a.H, a.W, i0, j0);
    __syncthreads();
    const int ti = threadIdx.x / kT, tj = threadIdx.x % kT;
    const int i = i0 + ti, j = j0 + tj;
    float acc[CO];
#pragma unroll
    for (int co = 0; co < CO; ++co) acc[co] = a.b[co];
    for (int ci = 0; ci < a.cin; ++ci) {
        float x[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) x[k] = t[ci][ti + k / 3][tj + k % 3];
#pragma unroll
        for (int co = 0; co < CO; ++co) {
            const float* wk = w + (co * a.cin + ci) * 9;
#pragma unroll
            for (int k = 0; k < 9; ++k) acc[co] = fmaf(wk[k], x[k], acc[co]);
        }
    }
    if (i >= a.H || j >= a.W) return;
#pragma unroll
    for (int co = 0; co < CO; ++co) {
        const float v = acc[co];
        a.out[((size_t)co * a.H + i) * a.W + j] = a.relu ? fmaxf(v, 0.f) : v;
    }
}

// gin[ci](p) = sum_co sum_k w[co][ci][k] gout[co](p - d_k), with d_k the
// offset (k / 3 - 1, k % 3 - 1): a forward pass over gout with the kernel
// flipped and the channel roles swapped.  a.in = gout (cin = cout of the
// layer), a.out = gin; mask: the layer's input (post-ReLU) or NULL.
struct ConvBwdInArgs {
    const float* gout; int cout;
    const float* w;                      // the layer's w[cout][cin][9]
    float* gin;       int cin;
    const float* mask;                   // layer input, gradient zero where <= 0
    int H, W;
};

// CI = a.cin as a constant, as the forward kernel's CO.
template <int CI>
__global__ void __launch_bounds__(256)
conv3x3_bwd_in_kernel(ConvBwdInArgs a) {
    __shared__ float t[kPredMaxC][kT + 2][kT + 2];
    const float* __restrict__ w = a.w;   // uniform: scalar path, as forward
    const int tiles_w = (a.W + kT - 1) / kT;
    const int i0 = (blockIdx.x / tiles_w) * kT, j0 = (blockIdx.x % tiles_w) * kT;
    stage_tile(t, a.gout, a.cout, a.H, a.W, i0, j0);
    __syncthreads();
    const int ti = threadIdx.x / kT, tj = threadIdx.x % kT;
    const int i = i0 + ti, j = j0 + tj;
    float acc[CI];
#pragma unroll
    for (int ci = 0; ci < CI; ++ci) acc[ci] = 0.f;
    for (int co = 0; co < a.cout; ++co) {
        float g[9];
        // gout at p - d_k: tile offset (1 - (k/3 - 1), 1 - (k%3 - 1))
#pragma unroll
        for (int k = 0; k < 9; ++k) g[k] = t[co][ti + 2 - k / 3][tj + 2 - k % 3];
#pragma unroll
        for (int ci = 0; ci < CI; ++ci) {
            const float* wk = w + (co * CI + ci) * 9;
#pragma unroll
            for (int k = 0; k < 9; ++k) acc[ci] = fmaf(wk[k], g[k], acc[ci]);
        }
    }
    if (i >= a.H || j >= a.W) return;
#pragma unroll
    for (int ci = 0; ci < CI; ++ci) {
        const size_t idx = ((size_t)ci * a.H + i) * a.W + j;
        float v = acc[ci];
        if (a.mask && !(a.mask[idx] > 0.f)) v = 0.f;
        a.gin[idx] = v;
    }
}

// Per-tile partial weight gradients: part[tile][cout * cin * 9 + cout].
struct ConvBwdWArgs {
    const float* in;  int cin;
    const float* gout; int cout;
    float* part;
    int H, W;
};

__global__ void __launch_bounds__(256)
conv3x3_bwd_w_kernel(ConvBwdWArgs a) {
    __shared__ float t[kPredMaxC][kT + 2][kT + 2];
    __shared__ float g[kPredMaxC][kT][kT];
    const int tiles_w = (a.W + kT - 1) / kT;
    const int i0 = (blockIdx.x / tiles_w) * kT, j0 = (blockIdx.x % tiles_w) * kT;
    stage_tile(t, a.in, a.cin, a.H, a.W, i0, j0);
    // the output-gradient tile: all loads of a thread before its stores
    constexpr int kG = kPredMaxC * kT * kT / 256;
    float gv[kG];
#pragma unroll
    for (int u = 0; u < kG; ++u) {
        const int e = threadIdx.x + u * 256;
        const int co = e / (kT * kT), r = e - co * kT * kT;
        const int i = i0 + r / kT, j = j0 + r % kT;
        const bool ok = co < a.cout && i < a.H && j < a.W;
        const float x = a.gout[ok ? ((size_t)co * a.H + i) * a.W + j : 0];
        gv[u] = ok ? x : 0.f;
    }
#pragma unroll
    for (int u = 0; u < kG; ++u) {
        const int e = threadIdx.x + u * 256;
        const int co = e / (kT * kT), r = e - co * kT * kT;
        if (co < a.cout) g[co][r / kT][r % kT] = gv[u];
    }
    __syncthreads();
    const int nw = a.cout * a.cin * 9, np = nw + a.cout;
    float* part = a.part + (size_t)blockIdx.x * np;
    for (int q = threadIdx.x; q < np; q += 256) {
        float s = 0.f;
        if (q < nw) {
            const int co = q / (a.cin * 9), r = q - co * a.cin * 9;
            const int ci = r / 9, k = r % 9;
            const int di = k / 3, dj = k % 3;
            for (int ti = 0; ti < kT; ++ti)
#pragma unroll 4
                for (int tj = 0; tj < kT; ++tj)
                    s = fmaf(g[co][ti][tj], t[ci][ti + di][tj + dj], s);
        } else {
            const int co = q - nw;
            for (int ti = 0; ti < kT; ++ti)
                for (int tj = 0; tj < kT; ++tj) s += g[co][ti][tj];
        }
        part[q] = s;
    }
}

// out[q] = sum over tiles of part[tile][q] (fixed order: deterministic)
__global__ void __launch_bounds__(256)
reduce_parts_kernel(const float* part, int ntiles, int np, float* out) {
    const int q = blockIdx.x;
    float s = 0.f;
    for (int t = threadIdx.x; t < ntiles; t += 256) s += part[(size_t)t * np + q];
    __shared__ float red[256];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[q] = red[0];
}

// g = 2 (y - t) / count; part[block] = sum (y - t)^2 over the block's values
__global__ void __launch_bounds__(256)
mse_grad_kernel(const float* y, const float* t, size_t n, float inv_count,
                float* g, float* part) {
    const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
    float d2 = 0.f;
    if (e < n) {
        const float d = y[e] - t[e];
        g[e] = 2.f * d * inv_count;
        d2 = d * d;
    }
    __shared__ float red[256];
    red[threadIdx.x] = d2;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// Adam (Kingma & Ba): m = b1 m + (1 - b1) g, v = b2 v + (1 - b2) g^2,
// p -= lr mhat / (sqrt(vhat) + eps) with bias-corrected mhat, vhat.
__global__ void __launch_bounds__(256)
adam_kernel(float* p, const float* g, float* m, float* v, int n, float lr,
            float b1, float b2, float eps, float c1, float c2) {
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= n) return;
    const float gi = g[e];
    const float mi = b1 * m[e] + (1.f - b1) * gi;
    const float vi = b2 * v[e] + (1.f - b2) * gi * gi;
    m[e] = mi;
    v[e] = vi;
    p[e] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// Network inputs of one sample, with Q = 3 F(n-1) - 3 F(n-2) + F(n-3) and
// D = F(n-1) - F(n-2): planes 0-5 = Q / s, 6-11 = D / s_d, 12 = beam / s_b,
// 13 = ro(n-1) / s_r (FP64 -> FP32).
struct PredIoArgs {
    const double* f1[6];     // fields of slice n-1
    const double* f2[6];     // fields of slice n-2
    const double* f3[6];     // fields of slice n-3
    const double* beam;      // NULL: beam_scalar
    double beam_scalar;
    const double* ro;        // plasma density of slice n-1
    const double* next[6];   // targets: fields of slice n
    float scale[14];         // input scales (planes 0-5 and 6-11 share s_f)
    float tscale[6];         // target scales
    float* planes;           // [14][N][N] or [6][N][N]
    const float* pred;       // [6][N][N] network output
    double* guess[6];        // pred_guess output
    int n;                   // N * N
};

__global__ void __launch_bounds__(256)
pred_inputs_kernel(PredIoArgs a) {
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= a.n) return;
#pragma unroll
    for (int f = 0; f < 6; ++f) {
        const double f1 = a.f1[f][e], f2 = a.f2[f][e];
        const double q = 3 * f1 - 3 * f2 + a.f3[f][e];
        a.planes[(size_t)f * a.n + e] = (float)(q / a.scale[f]);
        a.planes[(size_t)(6 + f) * a.n + e] = (float)((f1 - f2) / a.scale[6 + f]);
    }
    a.planes[(size_t)12 * a.n + e] = (float)((a.beam ? a.beam[e] : a.beam_scalar) / a.scale[12]);
    a.planes[(size_t)13 * a.n + e] = (float)(a.ro[e] / a.scale[13]);
}

// targets: (F(n) - Q) / t_f
__global__ void __launch_bounds__(256)
pred_targets_kernel(PredIoArgs a) {
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= a.n) return;
#pragma unroll
    for (int f = 0; f < 6; ++f) {
        const double q = 3 * a.f1[f][e] - 3 * a.f2[f][e] + a.f3[f][e];
        a.planes[(size_t)f * a.n + e] = (float)((a.next[f][e] - q) / a.tscale[f]);
    }
}

// guess: Q + t_f pred (FP64)
__global__ void __launch_bounds__(256)
pred_guess_kernel(PredIoArgs a) {
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= a.n) return;
#pragma unroll
    for (int f = 0; f < 6; ++f)
        a.guess[f][e] = 3 * a.f1[f][e] - 3 * a.f2[f][e] + a.f3[f][e] +
                        (double)a.tscale[f] * (double)a.pred[(size_t)f * a.n + e];
}

inline int tiles(int H, int W) { return ((H + kT - 1) / kT) * ((W + kT - 1) / kT); }

}  // namespace

}  // namespace lcode

using namespace lcode;

namespace {

inline bool channels_ok(int c) { return c >= 1 && c <= kPredMaxC; }

}  // namespace

extern "C" {

int lcode_pred_conv3x3_forward(const float* in, int32_t cin, const float* w,
                               const float* b, float* out, int32_t cout,
                               int32_t H, int32_t W, int32_t relu,
                               hipStream_t stream) {
    LCODE_REQUIRE(in && w && b && out, "NULL array");
    LCODE_REQUIRE(channels_ok(cin) && channels_ok(cout), "channels must be 1..16");
    LCODE_REQUIRE(H > 0 && W > 0, "empty plane");
    ConvArgs a{in, cin, w, b, out, cout, H, W, relu};
    switch (cout) {
#define LCODE_FWD(C) case C: conv3x3_fwd_kernel<C><<<tiles(H, W), 256, 0, stream>>>(a); break;
        LCODE_FWD(1) LCODE_FWD(2) LCODE_FWD(3) LCODE_FWD(4) LCODE_FWD(5) LCODE_FWD(6)
        LCODE_FWD(7) LCODE_FWD(8) LCODE_FWD(9) LCODE_FWD(10) LCODE_FWD(11) LCODE_FWD(12)
        LCODE_FWD(13) LCODE_FWD(14) LCODE_FWD(15) LCODE_FWD(16)
#undef LCODE_FWD
    }
    LCODE_HIP(hipGetLastError());
    return LCODE_OK;
}

int lcode_pred_conv3x3_backward_input(const float* gout, int32_t cout,
                                      const float* w, float* gin, int32_t cin,
                                      const float* relu_input, int32_t H,
                                      int32_t W, hipStream_t stream) {
    LCODE_REQUIRE(gout && w && gin, "NULL array");
    LCODE_REQUIRE(channels_ok(cin) && channels_ok(cout), "channels must be 1..16");
    LCODE_REQUIRE(H > 0 && W > 0, "empty plane");
    ConvBwdInArgs a{gout, cout, w, gin, cin, relu_input, H, W};
    switch (cin) {
#define LCODE_BWD(C) case C: conv3x3_bwd_in_kernel<C><<<tiles(H, W), 256, 0, stream>>>(a); break;
        LCODE_BWD(1) LCODE_BWD(2) LCODE_BWD(3) LCODE_BWD(4) LCODE_BWD(5) LCODE_BWD(6)
        LCODE_BWD(7) LCODE_BWD(8) LCODE_BWD(9) LCODE_BWD(10) LCODE_BWD(11) LCODE_BWD(12)
        LCODE_BWD(13) LCODE_BWD(14) LCODE_BWD(15) LCODE_BWD(16)
#undef LCODE_BWD
    }
    LCODE_HIP(hipGetLastError());
    return LCODE_OK;
}

int64_t lcode_pred_backward_weight_scratch(int32_t cin, int32_t cout,
                                           int32_t H, int32_t W) {
    return (int64_t)tiles(H, W) * (cout * cin * 9 + cout);
}

int lcode_pred_conv3x3_backward_weight(const float* in, int32_t cin,
                                       const float* gout, int32_t cout,
                                       float* scratch, float* gwb, int32_t H,
                                       int32_t W, hipStream_t stream) {
    LCODE_REQUIRE(in && gout && scratch && gwb, "NULL array");
    LCODE_REQUIRE(channels_ok(cin) && channels_ok(cout), "channels must be 1..16");
    LCODE_REQUIRE(H > 0 && W > 0, "empty plane");
    const int nt = tiles(H, W), np = cout * cin * 9 + cout;
    ConvBwdWArgs a{in, cin, gout, cout, scratch, H, W};
    conv3x3_bwd_w_kernel<<<nt, 256, 0, stream>>>(a);
    LCODE_HIP(hipGetLastError());
    reduce_parts_kernel<<<np, 256, 0, stream>>>(scratch, nt, np, gwb);
    LCODE_HIP(hipGetLastError());
    return LCODE_OK;
}

int lcode_pred_mse_grad(const float* y, const float* t, int64_t n, float* g,
                        float* scratch, float* loss_sum, hipStream_t stream) {
    LCODE_REQUIRE(y && t && g && scratch && loss_sum, "NULL array");
    LCODE_REQUIRE(n > 0 && n < (int64_t)1 << 31, "n out of range");
    const int nb = (int)((n + 255) / 256);
    mse_grad_kernel<<<nb, 256, 0, stream>>>(y, t, (size_t)n, 1.f / (float)n, g, scratch);
    LCODE_HIP(hipGetLastError());
    reduce_parts_kernel<<<1, 256, 0, stream>>>(scratch, nb, 1, loss_sum);
    LCODE_HIP(hipGetLastError());
    return LCODE_OK;
}

int lcode_pred_adam(float* p, const float* g, float* m, float* v, int32_t n,
                    float lr, float b1, float b2, float eps, int32_t t,
                    hipStream_t stream) {
    LCODE_REQUIRE(p && g && m && v, "NULL array");
    LCODE_REQUIRE(n > 0 && t >= 1, "bad size or step");
    const float c1 = 1.f - powf(b1, (float)t), c2 = 1.f - powf(b2, (float)t);
    adam_kernel<<<(n + 255) / 256, 256, 0, stream>>>(p, g, m, v, n, lr, b1, b2, eps, c1, c2);
    LCODE_HIP(hipGetLastError());
    return LCODE_OK;
}

static int pred_io(int which, const lcode_pred_io* io, hipStream_t s) {
    LCODE_REQUIRE(io && io->n > 0, "bad io");
    PredIoArgs a{};
    for (int f = 0; f < 6; ++f) {
        LCODE_REQUIRE(io->f1[f] && io->f2[f] && io->f3[f], "NULL field");
        a.f1[f] = io->f1[f];
        a.f2[f] = io->f2[f];
        a.f3[f] = io->f3[f];
        a.next[f] = io->next[f];
        a.guess[f] = io->guess[f];
        a.tscale[f] = io->tscale[f];
    }
    for (int c = 0; c < 14; ++c) a.scale[c] = io->scale[c];
    a.beam = io->beam;
    a.beam_scalar = io->beam_scalar;
    a.ro = io->ro;
    a.planes = io->planes;
    a.pred = io->pred;
    a.n = io->n;
    const int nb = (io->n + 255) / 256;
    if (which == 0) {
        LCODE_REQUIRE(io->ro && io->planes, "NULL ro/planes");
        pred_inputs_kernel<<<nb, 256, 0, s>>>(a);
    } else if (which == 1) {
        LCODE_REQUIRE(io->planes, "NULL planes");
        for (int f = 0; f < 6; ++f) LCODE_REQUIRE(io->next[f], "NULL target field");
        pred_targets_kernel<<<nb, 256, 0, s>>>(a);
    } else {
        LCODE_REQUIRE(io->pred, "NULL pred");
        for (int f = 0; f < 6; ++f) LCODE_REQUIRE(io->guess[f], "NULL guess field");
        pred_guess_kernel<<<nb, 256, 0, s>>>(a);
    }
    LCODE_HIP(hipGetLastError());
    return LCODE_OK;
}

int lcode_pred_inputs(const lcode_pred_io* io, hipStream_t s) { return pred_io(0, io, s); }
int lcode_pred_targets(const lcode_pred_io* io, hipStream_t s) { return pred_io(1, io, s); }
int lcode_pred_guess(const lcode_pred_io* io, hipStream_t s) { return pred_io(2, io, s); }

}  // extern "C"
