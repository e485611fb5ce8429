// Learned field predictor: a 4-level convolutional encoder-decoder with 32
// channels per level (no counterpart in the reference; it stands in for the
// predictor pass of lcode.py:841-881).  Activations and gradients are
// bfloat16 in HBM, laid out [H][W][32] (pixel-major, 32 channels of 64 B);
// weights are FP32 master copies packed to bfloat16 for the matrix cores;
// every sum is FP32.
//
// 3 x 3 convolutions are implicit GEMMs on v_mfma_f32_16x16x32_bf16:
// A = the weights of one tap (16 output channels x 32 input channels),
// B = the input pixels at that tap's offset (32 channels x 16 pixels), so a
// block's output tile accumulates over 9 taps x (input channels / 32)
// K steps.  The input tile (with its one-pixel halo) and the packed weights
// are staged in LDS once per block.  The same kernel computes the input
// gradient of a convolution, with the weights transposed and flipped.
// Weight gradients are GEMMs over pixels: per block, the output gradient
// and the input tile are staged transposed ([channel][pixel]) in LDS, each
// wave takes a subset of the taps, and the per-block partial sums are
// reduced in a fixed order.

#include "common.hpp"

namespace lcode {

namespace {

typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int kC = 32;            // channels per activation tensor
constexpr int kTR = 4;            // output tile rows (one wave each)
constexpr int kTC = 32;           // output tile columns (two 16-pixel tiles)
constexpr int kPix = 40;          // LDS stride of one pixel (32 ch + 8 pad)
constexpr int kHaloR = kTR + 2, kHaloC = kTC + 2;

__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);            // round to nearest even
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float bf2f(uint16_t b) {
    return __uint_as_float((uint32_t)b << 16);
}

// ---------------------------------------------------------------------------
// Convolution (forward, and input gradient with transposed weights)
// ---------------------------------------------------------------------------

// Output channels of the block: mt * 16 (mt = 1, 2 or 4).
template <int MT>
__global__ void __launch_bounds__(256)
unet_conv_kernel(UnetConvArgs a) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    uint16_t* sx = reinterpret_cast<uint16_t*>(lds_raw);          // [nin][6][34][40]
    const int cin = kC * a.nin;
    const int wrow = cin + 8;                                       // packed row stride
    uint16_t* sw = sx + (size_t)a.nin * kHaloR * kHaloC * kPix;     // [9][MT*16][cin+8]
    const int tiles_x = (a.W + kTC - 1) / kTC;
    const int r0 = (blockIdx.x / tiles_x) * kTR, c0 = (blockIdx.x % tiles_x) * kTC;

    // stage the input tile with its halo (zeros outside the plane), 16 B
    // per thread and step
    const int chunks = a.nin * kHaloR * kHaloC * 4;
    for (int k = threadIdx.x; k < chunks; k += blockDim.x) {
        const int q = k & 3, pix = k >> 2;
        const int ch = pix / (kHaloR * kHaloC), rem = pix % (kHaloR * kHaloC);
        const int rr = rem / kHaloC, cc = rem % kHaloC;
        const int gr = r0 - 1 + rr, gc = c0 - 1 + cc;
        uint4 v = make_uint4(0, 0, 0, 0);
        if (gr >= 0 && gr < a.H && gc >= 0 && gc < a.W) {
            const int mode = a.in_mode[ch];
            if (mode == 0) {
                v = *reinterpret_cast<const uint4*>(a.in[ch] + ((size_t)gr * a.W + gc) * kC + q * 8);
            } else if (mode == 2) {       // nearest upsampling of a coarser tensor
                v = *reinterpret_cast<const uint4*>(
                    a.in[ch] + ((size_t)(gr / 2) * a.in_w[ch] + gc / 2) * kC + q * 8);
            } else {                      // 2 x 2 mean of a finer tensor
                float f[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
                for (int da = 0; da < 2; ++da)
                    for (int db = 0; db < 2; ++db) {
                        const int ii = 2 * gr + da, jj = 2 * gc + db;
                        if (ii >= a.in_h[ch] || jj >= a.in_w[ch]) continue;
                        const uint4 u = *reinterpret_cast<const uint4*>(
                            a.in[ch] + ((size_t)ii * a.in_w[ch] + jj) * kC + q * 8);
                        const uint32_t w4[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
                        for (int e = 0; e < 4; ++e) {
                            f[2 * e] += bf2f((uint16_t)(w4[e] & 0xFFFFu));
                            f[2 * e + 1] += bf2f((uint16_t)(w4[e] >> 16));
                        }
                    }
                uint32_t w4[4];
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    w4[e] = (uint32_t)f2bf(f[2 * e] * .25f) | ((uint32_t)f2bf(f[2 * e + 1] * .25f) << 16);
                v = make_uint4(w4[0], w4[1], w4[2], w4[3]);
            }
        }
        *reinterpret_cast<uint4*>(sx + (size_t)pix * kPix + q * 8) = v;
    }
    // stage the packed weights [9][MT*16][cin] into rows of cin + 8
    const int wchunks = 9 * MT * 16 * (cin / 8);
    for (int k = threadIdx.x; k < wchunks; k += blockDim.x) {
        const int q = k % (cin / 8), row = k / (cin / 8);
        *reinterpret_cast<uint4*>(sw + (size_t)row * wrow + q * 8) =
            *reinterpret_cast<const uint4*>(a.w + (size_t)row * cin + q * 8);
    }
    __syncthreads();

    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int li = lane & 15, lk = lane >> 4;
    v4f acc[2][MT];
#pragma unroll
    for (int n = 0; n < 2; ++n)
#pragma unroll
        for (int m = 0; m < MT; ++m) acc[n][m] = v4f{0.f, 0.f, 0.f, 0.f};
    for (int ch = 0; ch < a.nin; ++ch) {
        const uint16_t* xc = sx + (size_t)ch * kHaloR * kHaloC * kPix;
#pragma unroll 1
        for (int tap = 0; tap < 9; ++tap) {
            const int ty = tap / 3, tx = tap % 3;
            bf8 b[2];
#pragma unroll
            for (int n = 0; n < 2; ++n) {
                const int pix = (wave + ty) * kHaloC + n * 16 + li + tx;
                b[n] = *reinterpret_cast<const bf8*>(xc + (size_t)pix * kPix + 8 * lk);
            }
#pragma unroll
            for (int m = 0; m < MT; ++m) {
                const bf8 w8 = *reinterpret_cast<const bf8*>(
                    sw + (size_t)(tap * MT * 16 + m * 16 + li) * wrow + ch * kC + 8 * lk);
#pragma unroll
                for (int n = 0; n < 2; ++n)
                    acc[n][m] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w8, b[n], acc[n][m], 0, 0, 0);
            }
        }
    }

    // D[i][j] (i: output channel within the tile, j: pixel) sits in lane
    // j + 16 (i / 4), item i % 4: each lane holds four consecutive channels
    // of one pixel
    const int r = r0 + wave;
#pragma unroll
    for (int n = 0; n < 2; ++n) {
        const int c = c0 + n * 16 + li;
        if (r >= a.H || c >= a.W) continue;
        const size_t p = (size_t)r * a.W + c;
#pragma unroll
        for (int m = 0; m < MT; ++m) {
            const int co = m * 16 + 4 * lk;          // channels co .. co + 3
            float v[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                float x = acc[n][m][q];
                if (a.bias) x += co + q < a.cout ? a.bias[co + q] : 0.f;
                if (a.relu) x = fmaxf(x, 0.f);
                if (co + q >= a.cout) x = 0.f;
                v[q] = x;
            }
            if (a.out_f32) {
                *reinterpret_cast<float4*>(a.out_f32 + p * 16 + co) =
                    make_float4(v[0], v[1], v[2], v[3]);
                continue;
            }
            const int half = co / kC, cc = co % kC;
            if (a.mask) {
                const ushort4 mk = *reinterpret_cast<const ushort4*>(a.mask + p * kC + cc);
                if (!(bf2f(mk.x) > 0.f)) v[0] = 0.f;
                if (!(bf2f(mk.y) > 0.f)) v[1] = 0.f;
                if (!(bf2f(mk.z) > 0.f)) v[2] = 0.f;
                if (!(bf2f(mk.w) > 0.f)) v[3] = 0.f;
            }
            uint16_t* o = a.out[half];
            if (!o) continue;
            *reinterpret_cast<ushort4*>(o + p * kC + cc) =
                make_ushort4(f2bf(v[0]), f2bf(v[1]), f2bf(v[2]), f2bf(v[3]));
        }
    }
}

// ---------------------------------------------------------------------------
// Weight gradient: gw[tap][co][ci] = sum_p g[p][co] x[p + off(tap)][ci]
// ---------------------------------------------------------------------------

// Per block: the output gradient of its 4 x 32 pixels transposed into
// gT[co][128] and the input tile (with halo) into xT[ci][6 * 34]; wave w
// takes taps w, w + 4 (and 8 for wave 0); MT output-channel tiles of 16,
// NN = cin / 16 input-channel tiles.
template <int MT, int NN>
__global__ void __launch_bounds__(256)
unet_wgrad_kernel(UnetWgradArgs a) {
    constexpr int kTile = kTR * kTC;                 // 128 pixels
    constexpr int kCin = NN * 16;
    constexpr int kHalo = kHaloR * kHaloC;           // 204 pixels
    __shared__ __align__(16) uint16_t gT[MT * 16][kTile + 8];
    __shared__ __align__(16) uint16_t xT[kCin][kHalo + 4];
    const int tiles_x = (a.W + kTC - 1) / kTC;
    const int r0 = (blockIdx.x / tiles_x) * kTR, c0 = (blockIdx.x % tiles_x) * kTC;
    for (int k = threadIdx.x; k < kTile * MT * 16; k += blockDim.x) {
        const int co = k % (MT * 16), px = k / (MT * 16);
        const int gr = r0 + px / kTC, gc = c0 + px % kTC;
        uint16_t v = 0;
        if (gr < a.H && gc < a.W) v = a.g[((size_t)gr * a.W + gc) * kC + co];
        gT[co][px] = v;
    }
    for (int k = threadIdx.x; k < kHalo * kCin; k += blockDim.x) {
        const int ci = k % kCin, px = k / kCin;
        const int gr = r0 - 1 + px / kHaloC, gc = c0 - 1 + px % kHaloC;
        uint16_t v = 0;
        if (gr >= 0 && gr < a.H && gc >= 0 && gc < a.W)
            v = a.in[ci / kC][((size_t)gr * a.W + gc) * kC + ci % kC];
        xT[ci][px] = v;
    }
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int li = lane & 15, lk = lane >> 4;
    float* part = a.partial + (size_t)blockIdx.x * (9 * MT * 16 * kCin + MT * 16);
    for (int tap = wave; tap < 9; tap += 4) {
        const int ty = tap / 3, tx = tap % 3;
        v4f acc[MT][NN];
#pragma unroll
        for (int m = 0; m < MT; ++m)
#pragma unroll
            for (int nn = 0; nn < NN; ++nn) acc[m][nn] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll 1
        for (int rr = 0; rr < kTR; ++rr) {
            bf8 A[MT];
#pragma unroll
            for (int m = 0; m < MT; ++m)
                A[m] = *reinterpret_cast<const bf8*>(&gT[m * 16 + li][rr * kTC + 8 * lk]);
            const int base = (rr + ty) * kHaloC + 8 * lk + tx;
#pragma unroll
            for (int nn = 0; nn < NN; ++nn) {
                const uint16_t* src = &xT[nn * 16 + li][base];
                bf8 B;
#pragma unroll
                for (int u = 0; u < 8; ++u)
                    B[u] = __builtin_bit_cast(__bf16, src[u]);
#pragma unroll
                for (int m = 0; m < MT; ++m)
                    acc[m][nn] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[m], B, acc[m][nn], 0, 0, 0);
            }
        }
        // D[i = co][j = ci]: lane j + 16 (i / 4), item i % 4
#pragma unroll
        for (int m = 0; m < MT; ++m)
#pragma unroll
            for (int nn = 0; nn < NN; ++nn)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int co = m * 16 + 4 * lk + q, ci = nn * 16 + li;
                    part[((size_t)tap * MT * 16 + co) * kCin + ci] = acc[m][nn][q];
                }
    }
    // bias gradient: the output gradient summed over the tile's pixels
    if (threadIdx.x < MT * 16) {
        float s = 0.f;
        for (int px = 0; px < kTile; ++px) s += bf2f(gT[threadIdx.x][px]);
        part[9 * MT * 16 * kCin + threadIdx.x] = s;
    }
}

// Sums the per-block partials in block order and writes the gradient in
// the master layout w[co][ci][tap] (co < cout, ci < cin real), then b[co].
__global__ void __launch_bounds__(256)
unet_wgrad_reduce_kernel(const float* partial, int nblocks, int cout_pad,
                         int cinp, int cout, int cin, float* grad) {
    const int per = 9 * cout_pad * cinp + cout_pad;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= cout * cin * 9 + cout) return;
    int src;
    if (k < cout * cin * 9) {
        const int co = k / (cin * 9), ci = (k / 9) % cin, tap = k % 9;
        src = (tap * cout_pad + co) * cinp + ci;
    } else {
        src = 9 * cout_pad * cinp + (k - cout * cin * 9);
    }
    float s = 0.f;
    for (int b = 0; b < nblocks; ++b) s += partial[(size_t)b * per + src];
    grad[k] = s;
}

// One Adam step t >= 1 (Kingma & Ba 2015, Algorithm 1, bias-corrected).
__global__ void __launch_bounds__(256)
unet_adam_kernel(float* p, const float* g, float* m, float* v, int n, float lr,
                 float b1, float b2, float eps, float c1, float c2) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const float gk = g[k];
    const float mk = b1 * m[k] + (1.f - b1) * gk;
    const float vk = b2 * v[k] + (1.f - b2) * gk * gk;
    m[k] = mk;
    v[k] = vk;
    p[k] -= lr * (mk / c1) / (sqrtf(vk / c2) + eps);
}

// ---------------------------------------------------------------------------
// Weight packing, pooling, upsampling, loss, input/target/guess planes
// ---------------------------------------------------------------------------

// master w[cout][cin][3][3] (cin real channels) -> fwd[tap][cout_pad][cinp]
// (cinp = cin padded to the input tensor's 32-channel chunks) and the
// input-gradient form bwd[tap'][cin][32] with bwd[t][ci][co] =
// w[co][ci][8 - t]; zeros for co >= cout and ci >= cin.
__global__ void __launch_bounds__(256)
unet_pack_kernel(const float* w, int cout, int cin, int cout_pad, int cinp,
                 uint16_t* fwd, uint16_t* bwd) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < 9 * cout_pad * cinp) {
        const int tap = k / (cout_pad * cinp), co = (k / cinp) % cout_pad, ci = k % cinp;
        fwd[k] = co < cout && ci < cin ? f2bf(w[((size_t)co * cin + ci) * 9 + tap]) : 0;
    }
    if (bwd && k < 9 * cin * kC) {
        const int tap = k / (cin * kC), ci = (k / kC) % cin, co = k % kC;
        bwd[k] = co < cout ? f2bf(w[((size_t)co * cin + ci) * 9 + (8 - tap)]) : 0;
    }
}

// 2 x 2 mean (zeros beyond the edge) of [H][W][32] -> [h][w][32]
__global__ void __launch_bounds__(256)
unet_pool_kernel(const uint16_t* x, int H, int W, uint16_t* y, int h, int w) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= h * w * kC) return;
    const int c = k % kC, p = k / kC, i = p / w, j = p % w;
    float s = 0.f;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const int ii = 2 * i + a, jj = 2 * j + b;
            if (ii < H && jj < W) s += bf2f(x[((size_t)ii * W + jj) * kC + c]);
        }
    y[k] = f2bf(s * .25f);
}

// nearest upsampling [h][w][32] -> [H][W][32]
__global__ void __launch_bounds__(256)
unet_up_kernel(const uint16_t* x, int h, int w, uint16_t* y, int H, int W) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= H * W * kC) return;
    const int c = k % kC, p = k / kC, i = p / W, j = p % W;
    y[k] = x[((size_t)(i / 2) * w + j / 2) * kC + c];
}

// gradient through up, masked by the coarse activation:
// out[I][J] = (sum of g over the 2 x 2 block, in bounds) (mask > 0)
__global__ void __launch_bounds__(256)
unet_up_t_kernel(const uint16_t* g, int H, int W, const uint16_t* mask,
                 uint16_t* out, int h, int w) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= h * w * kC) return;
    const int c = k % kC, p = k / kC, i = p / w, j = p % w;
    float s = 0.f;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const int ii = 2 * i + a, jj = 2 * j + b;
            if (ii < H && jj < W) s += bf2f(g[((size_t)ii * W + jj) * kC + c]);
        }
    out[k] = f2bf(bf2f(mask[k]) > 0.f ? s : 0.f);
}

// gradient through pool plus the skip gradient, masked by the activation:
// out[i][j] = (g[i/2][j/2] / 4 + skip[i][j]) (mask > 0)
__global__ void __launch_bounds__(256)
unet_pool_t_kernel(const uint16_t* g, int h, int w, const uint16_t* skip,
                   const uint16_t* mask, uint16_t* out, int H, int W) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= H * W * kC) return;
    const int c = k % kC, p = k / kC, i = p / W, j = p % W;
    const float v = bf2f(g[((size_t)(i / 2) * w + j / 2) * kC + c]) * .25f + bf2f(skip[k]);
    out[k] = f2bf(bf2f(mask[k]) > 0.f ? v : 0.f);
}

// g[p][c] = 2 (y - t) / n for c < 6 (0 above), as bfloat16 [P][32];
// per-block partial sums of (y - t)^2 / n (n = 6 P: the loss is the mean)
__global__ void __launch_bounds__(256)
unet_mse_kernel(const float* y, const float* t, int P, float inv_n,
                uint16_t* g, float* part) {
    __shared__ float red[256];
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    float sq = 0.f;
    if (k < P * kC) {
        const int c = k % kC, p = k / kC;
        float v = 0.f;
        if (c < 6) {
            const float d = y[(size_t)p * 16 + c] - t[(size_t)p * 6 + c];
            v = 2.f * d * inv_n;
            sq = d * d * inv_n;
        }
        g[k] = f2bf(v);
    }
    red[threadIdx.x] = sq;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(256)
unet_sum_kernel(const float* part, int n, float* out) {
    __shared__ float red[256];
    float s = 0.f;
    for (int k = threadIdx.x; k < n; k += 256) s += part[k];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

// Field value f of slice set `s` at (i, j) as double, from FP64 or FP32
// grids.
__device__ __forceinline__ double fld(const void* p, int f32, size_t idx) {
    return f32 ? (double)static_cast<const float*>(p)[idx]
               : static_cast<const double*>(p)[idx];
}

// Input planes at level-0 size: the 4 x 4 mean (zeros beyond the edge, /16)
// of F(n-1)/s (6), (F(n-1) - F(n-2))/s (6), beam/s, ro(n-1)/s, channels
// 14..31 zero; bfloat16 [S][S][32] (or FP32 [S][S][16]).  Thread g P + p
// takes level-0 pixel p and plane group g (field g < 6: channels g and
// 6 + g; 6: the beam; 7: the density and the zero channels), so adjacent
// threads read adjacent 4-value row segments of one plane.
__global__ void __launch_bounds__(256)
unet_inputs_kernel(UnetIoArgs a) {
    const int S = a.S, N = a.N, P = S * S;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= 8 * P) return;
    const int g = k / P, p = k % P, I = p / S, J = p % S;
    const void* src = g < 6 ? a.f1[g] : a.ro;
    double s1 = 0., s2 = 0.;
    for (int a4 = 0; a4 < 4; ++a4) {
        const int i = 4 * I + a4;
        if (i >= N) break;
        for (int b4 = 0; b4 < 4; ++b4) {
            const int j = 4 * J + b4;
            if (j >= N) break;
            const size_t idx = (size_t)i * N + j;
            if (g == 6) {
                s1 += a.beam ? a.beam[idx] : a.beam_scalar;
            } else {
                const double x1 = fld(src, a.f32, idx);
                s1 += x1;
                if (g < 6) s2 += x1 - fld(a.f2[g], a.f32, idx);
            }
        }
    }
    const int c1 = g < 6 ? g : (g == 6 ? 12 : 13);
    const float v1 = (float)(s1 / 16 / a.scale[c1]);
    const float v2 = g < 6 ? (float)(s2 / 16 / a.scale[6 + g]) : 0.f;
    if (a.planes_f32) {
        float* o = a.planes_f32 + (size_t)p * 16;
        o[c1] = v1;
        if (g < 6) o[6 + g] = v2;
        if (g == 7) o[14] = o[15] = 0.f;
        return;
    }
    uint16_t* o = a.planes + (size_t)p * kC;
    o[c1] = f2bf(v1);
    if (g < 6) o[6 + g] = f2bf(v2);
    if (g == 7) {
        o[14] = o[15] = 0;
        uint4* z = reinterpret_cast<uint4*>(o + 16);
        z[0] = z[1] = make_uint4(0, 0, 0, 0);
    }
}

// The base the network's output corrects: Q = 3 F(n-1) - 3 F(n-2) +
// F(n-3), or with slice n-4 given 2 F(n-1) - 2 F(n-3) + F(n-4) (exact for
// quadratic fields, and it carries the alternating part that Variant A
// leaves in the transverse fields on as the step does, so the correction
// the network learns is smooth from slice to slice).
__device__ __forceinline__ double unet_base(const UnetIoArgs& a, int f, size_t idx) {
    if (a.f4[f])
        return 2 * fld(a.f1[f], a.f32, idx) - 2 * fld(a.f3[f], a.f32, idx) +
               fld(a.f4[f], a.f32, idx);
    return 3 * fld(a.f1[f], a.f32, idx) - 3 * fld(a.f2[f], a.f32, idx) +
           fld(a.f3[f], a.f32, idx);
}

// Targets at level-0 size: the 4 x 4 mean of (F1(n) - Q(n)) / t; FP32
// [S][S][6].
__global__ void __launch_bounds__(256)
unet_targets_kernel(UnetIoArgs a) {
    const int S = a.S, N = a.N;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= S * S * 6) return;
    const int f = k % 6, p = k / 6, I = p / S, J = p % S;
    double s = 0.;
    for (int a4 = 0; a4 < 4; ++a4)
        for (int b4 = 0; b4 < 4; ++b4) {
            const int i = 4 * I + a4, j = 4 * J + b4;
            if (i >= N || j >= N) continue;
            const size_t idx = (size_t)i * N + j;
            s += fld(a.next[f], a.f32, idx) - unet_base(a, f, idx);
        }
    a.targets[k] = (float)(s / 16 / a.tscale[f]);
}

// Guess: Q + t y(I / 4, J / 4) over the full grid, in the grids' type.
__global__ void __launch_bounds__(256)
unet_guess_kernel(UnetIoArgs a) {
    const int N = a.N, S = a.S;
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (size_t)N * N) return;
    const int i = (int)(k / N), j = (int)(k % N);
    const size_t p = (size_t)(i / 4) * S + j / 4;
#pragma unroll
    for (int f = 0; f < 6; ++f) {
        const double g = unet_base(a, f, k) + (double)a.tscale[f] * a.y[p * 16 + f];
        if (a.f32) static_cast<float*>(a.guess[f])[k] = (float)g;
        else static_cast<double*>(a.guess[f])[k] = g;
    }
}

inline int nblk(long n) { return (int)((n + 255) / 256); }

}  // namespace

int unet_conv_lds_bytes(int nin, int mt) {
    return (nin * kHaloR * kHaloC * kPix + 9 * mt * 16 * (kC * nin + 8)) * 2;
}

int unet_tiles(int H, int W) {
    return ((H + kTR - 1) / kTR) * ((W + kTC - 1) / kTC);
}

int launch_unet_conv(const UnetConvArgs& a, hipStream_t s) {
    const int mt = a.cout_pad / 16;
    const size_t lds = (size_t)unet_conv_lds_bytes(a.nin, mt);
    const dim3 grid(unet_tiles(a.H, a.W));
    // blocks above 64 KiB of LDS (two input chunks) need the limit raised
    // once per instantiation (gfx950 has 160 KiB per CU)
    static bool raised[3] = {false, false, false};
    auto launch = [&](auto kernel, int slot) {
        if (lds > 65536 && !raised[slot]) {
            if (hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                    hipFuncAttributeMaxDynamicSharedMemorySize,
                                    160 * 1024) != hipSuccess)
                return false;
            raised[slot] = true;
        }
        kernel<<<grid, 256, lds, s>>>(a);
        return true;
    };
    bool ok;
    switch (mt) {
        case 1: ok = launch(unet_conv_kernel<1>, 0); break;
        case 2: ok = launch(unet_conv_kernel<2>, 1); break;
        case 4: ok = launch(unet_conv_kernel<4>, 2); break;
        default: set_error("unet conv: cout_pad must be 16, 32 or 64"); return LCODE_E_INVALID;
    }
    if (!ok) {
        set_error("hipFuncSetAttribute (dynamic LDS) failed for the unet convolution");
        return LCODE_E_HIP;
    }
    LCODE_HIP(hipGetLastError());
    return LCODE_OK;
}

int launch_unet_wgrad(const UnetWgradArgs& a, int cout, int cin, float* grad,
                      hipStream_t s) {
    const int mt = a.cout_pad / 16, nn = a.nin * kC / 16;
    const dim3 grid(unet_tiles(a.H, a.W));
#define LCODE_WG(M, NN) unet_wgrad_kernel<M, NN><<<grid, 256, 0, s>>>(a)
    if (mt == 1 && nn == 2) LCODE_WG(1, 2);
    else if (mt == 2 && nn == 2) LCODE_WG(2, 2);
    else if (mt == 2 && nn == 4) LCODE_WG(2, 4);
    else {
        set_error("unet weight gradient: unsupported channel counts");
        return LCODE_E_INVALID;
    }
#undef LCODE_WG
    LCODE_HIP(hipGetLastError());
    const int n = cout * cin * 9 + cout;
    unet_wgrad_reduce_kernel<<<nblk(n), 256, 0, s>>>(a.partial, (int)grid.x, a.cout_pad,
                                                      a.nin * kC, cout, cin, grad);
    LCODE_HIP(hipGetLastError());
    return LCODE_OK;
}

int launch_unet_pack(const float* w, int cout, int cin, int cout_pad, int cinp,
                     uint16_t* fwd, uint16_t* bwd, hipStream_t s) {
    const int n = max(9 * cout_pad * cinp, bwd ? 9 * cin * kC : 0);
    unet_pack_kernel<<<nblk(n), 256, 0, s>>>(w, cout, cin, cout_pad, cinp, fwd, bwd);
    LCODE_HIP(hipGetLastError());
    return LCODE_OK;
}

int launch_unet_adam(float* p, const float* g, float* m, float* v, int n,
                     float lr, int t, hipStream_t s) {
    const float b1 = .9f, b2 = .999f;
    const float c1 = 1.f - powf(b1, (float)t), c2 = 1.f - powf(b2, (float)t);
    unet_adam_kernel<<<nblk(n), 256, 0, s>>>(p, g, m, v, n, lr, b1, b2, 1e-8f, c1, c2);
    LCODE_HIP(hipGetLastError());
    return LCODE_OK;
}

int launch_unet_pool(const uint16_t* x, int H, int W, uint16_t* y, hipStream_t s) {
    const int h = (H + 1) / 2, w = (W + 1) / 2;
    unet_pool_kernel<<<nblk((long)h * w * kC), 256, 0, s>>>(x, H, W, y, h, w);
    LCODE_HIP(hipGetLastError());
    return LCODE_OK;
}

int launch_unet_up(const uint16_t* x, int h, int w, uint16_t* y, int H, int W,
                   hipStream_t s) {
    unet_up_kernel<<<nblk((long)H * W * kC), 256, 0, s>>>(x, h, w, y, H, W);
    LCODE_HIP(hipGetLastError());
    return LCODE_OK;
}

int launch_unet_up_t(const uint16_t* g, int H, int W, const uint16_t* mask,
                     uint16_t* out, int h, int w, hipStream_t s) {
    unet_up_t_kernel<<<nblk((long)h * w * kC), 256, 0, s>>>(g, H, W, mask, out, h, w);
    LCODE_HIP(hipGetLastError());
    return LCODE_OK;
}

int launch_unet_pool_t(const uint16_t* g, int h, int w, const uint16_t* skip,
                       const uint16_t* mask, uint16_t* out, int H, int W,
                       hipStream_t s) {
    unet_pool_t_kernel<<<nblk((long)H * W * kC), 256, 0, s>>>(g, h, w, skip, mask, out, H, W);
    LCODE_HIP(hipGetLastError());
    return LCODE_OK;
}

int launch_unet_mse(const float* y, const float* t, int P, uint16_t* g,
                    float* part, float* loss, hipStream_t s) {
    const int nb = nblk((long)P * kC);
    unet_mse_kernel<<<nb, 256, 0, s>>>(y, t, P, 1.f / (6.f * P), g, part);
    LCODE_HIP(hipGetLastError());
    unet_sum_kernel<<<1, 256, 0, s>>>(part, nb, loss);
    LCODE_HIP(hipGetLastError());
    return LCODE_OK;
}

int launch_unet_io(const UnetIoArgs& a, int which, hipStream_t s) {
    switch (which) {
        case 0: unet_inputs_kernel<<<nblk(8L * a.S * a.S), 256, 0, s>>>(a); break;
        case 1: unet_targets_kernel<<<nblk((long)a.S * a.S * 6), 256, 0, s>>>(a); break;
        case 2: unet_guess_kernel<<<nblk((long)a.N * a.N), 256, 0, s>>>(a); break;
        default: return LCODE_E_INVALID;
    }
    LCODE_HIP(hipGetLastError());
    return LCODE_OK;
}

}  // namespace lcode

// ---------------------------------------------------------------------------
// The network (C ABI): buffers owned by the caller, every pass enqueued on
// the caller's stream
// ---------------------------------------------------------------------------

namespace lcode {
namespace {

struct LayerShape {
    int cout, cin, cout_pad, cinp;   // real and padded channel counts
    int level;                       // 0..3
};

// C0 .. C7 (see the header): 14 input planes read as one 32-channel chunk
constexpr LayerShape kLayers[8] = {
    {32, 14, 32, 32, 0}, {32, 32, 32, 32, 1}, {32, 32, 32, 32, 2}, {32, 32, 32, 32, 3},
    {32, 64, 32, 64, 2}, {32, 64, 32, 64, 1}, {32, 64, 32, 64, 0}, {6, 32, 16, 32, 0}};

// activation and gradient tensors ([s][s][32] bfloat16 each) and levels
enum {
    T_E0, T_P0, T_E1, T_P1, T_E2, T_P2, T_E3, T_U3, T_D2, T_U2, T_D1, T_U1, T_D0,
    T_GO, T_GD0, T_GU1, T_GS0, T_GD1, T_GU2, T_GS1, T_GD2, T_GU3, T_GS2, T_GE3,
    T_GP2, T_GE2, T_GP1, T_GE1, T_GP0, T_GE0, T_COUNT
};
constexpr int kTensorLevel[T_COUNT] = {0, 1, 1, 2, 2, 3, 3, 2, 2, 1, 1, 0, 0,
                                       0, 0, 0, 0, 1, 1, 1, 2, 2, 2, 3,
                                       3, 2, 2, 1, 1, 0};

int level_size(int S, int l) {
    for (int k = 0; k < l; ++k) S = (S + 1) / 2;
    return S;
}

int64_t param_offset(int layer) {
    int64_t off = 0;
    for (int k = 0; k < layer; ++k)
        off += (int64_t)kLayers[k].cout * kLayers[k].cin * 9 + kLayers[k].cout;
    return off;
}

int64_t pack_offset(int layer, bool bwd) {
    int64_t off = 0;
    for (int k = 0; k <= layer; ++k) {
        const LayerShape& L = kLayers[k];
        const int64_t f = 9LL * L.cout_pad * L.cinp, b = 9LL * L.cin * 32;
        if (k == layer) return bwd ? off + f : off;
        off += f + b;
    }
    return off;
}

int64_t tensor_offset(int S, int t) {
    int64_t off = 0;
    for (int k = 0; k < t; ++k) {
        const int s = level_size(S, kTensorLevel[k]);
        off += (int64_t)s * s * 32;
    }
    return off;
}

int64_t max_tiles(int S) { return unet_tiles(S, S); }

}  // namespace
}  // namespace lcode

using namespace lcode;

extern "C" {

int lcode_unet_sizes(int32_t S, int64_t sizes[5]) {
    LCODE_REQUIRE(S >= 2 && S <= 4096 && sizes, "bad level-0 size or NULL");
    sizes[0] = param_offset(8);
    sizes[1] = pack_offset(8, false);
    sizes[2] = tensor_offset(S, T_COUNT);
    // per-block partials of the largest layer (9 x 32 x 64 + 32 floats)
    sizes[3] = max_tiles(S) * (9 * 32 * 64 + 32);
    sizes[4] = ((int64_t)S * S * 32 + 255) / 256;
    return LCODE_OK;
}

int64_t lcode_unet_tensor_offset(int32_t S, int32_t index) {
    if (index < 0 || index > T_COUNT || S < 2) return -1;
    return tensor_offset(S, index);
}

int64_t lcode_unet_param_offset(int32_t layer) {
    if (layer < 0 || layer > 8) return -1;
    return param_offset(layer);
}

static int check_net(const lcode_unet* n) {
    LCODE_REQUIRE(n && n->S >= 2 && n->params && n->wpack && n->act && n->y,
                  "incomplete lcode_unet");
    return LCODE_OK;
}

int lcode_unet_pack(const lcode_unet* n, hipStream_t stream) {
    int rc = check_net(n);
    if (rc) return rc;
    for (int k = 0; k < 8; ++k) {
        const LayerShape& L = kLayers[k];
        if ((rc = launch_unet_pack(n->params + param_offset(k), L.cout, L.cin, L.cout_pad,
                                   L.cinp, n->wpack + pack_offset(k, false),
                                   k ? n->wpack + pack_offset(k, true) : nullptr, stream)))
            return rc;
    }
    return LCODE_OK;
}

static uint16_t* T(const lcode_unet* n, int t) { return n->act + tensor_offset(n->S, t); }

static int conv(const lcode_unet* n, int k, const uint16_t* in0, const uint16_t* in1,
                bool relu, uint16_t* out0, float* out_f32, hipStream_t s) {
    const LayerShape& L = kLayers[k];
    UnetConvArgs a{};
    a.in[0] = in0;
    a.in[1] = in1;
    a.nin = L.cinp / 32;
    a.w = n->wpack + pack_offset(k, false);
    a.bias = n->params + param_offset(k) + (int64_t)L.cout * L.cin * 9;
    a.H = a.W = level_size(n->S, L.level);
    a.cout = L.cout;
    a.cout_pad = L.cout_pad;
    a.relu = relu;
    a.out[0] = out0;
    a.out_f32 = out_f32;
    return launch_unet_conv(a, s);
}

// input gradient of layer k from g: channels 0..31 to out0, 32..63 to out1
static int conv_t(const lcode_unet* n, int k, const uint16_t* g, const uint16_t* mask,
                  uint16_t* out0, uint16_t* out1, hipStream_t s) {
    const LayerShape& L = kLayers[k];
    UnetConvArgs a{};
    a.in[0] = g;
    a.nin = 1;
    a.w = n->wpack + pack_offset(k, true);
    a.H = a.W = level_size(n->S, L.level);
    a.cout = L.cin;
    a.cout_pad = L.cin;
    a.mask = mask;
    a.out[0] = out0;
    a.out[1] = out1;
    return launch_unet_conv(a, s);
}

static int wgrad(const lcode_unet* n, int k, const uint16_t* in0, const uint16_t* in1,
                 const uint16_t* g, hipStream_t s) {
    const LayerShape& L = kLayers[k];
    UnetWgradArgs a{};
    a.in[0] = in0;
    a.in[1] = in1;
    a.g = g;
    a.H = a.W = level_size(n->S, L.level);
    a.nin = L.cinp / 32;
    a.cout_pad = L.cout_pad;
    a.partial = n->partial;
    return launch_unet_wgrad(a, L.cout, L.cin, n->grad + param_offset(k), s);
}

int lcode_unet_forward(const lcode_unet* n, const uint16_t* x, hipStream_t s) {
    int rc = check_net(n);
    if (rc) return rc;
    LCODE_REQUIRE(x, "NULL input planes");
    const int S0 = n->S, S1 = level_size(S0, 1), S2 = level_size(S0, 2),
              S3 = level_size(S0, 3);
#define LCODE_TRY(expr) if ((rc = (expr))) return rc
    LCODE_TRY(conv(n, 0, x, nullptr, true, T(n, T_E0), nullptr, s));
    LCODE_TRY(launch_unet_pool(T(n, T_E0), S0, S0, T(n, T_P0), s));
    LCODE_TRY(conv(n, 1, T(n, T_P0), nullptr, true, T(n, T_E1), nullptr, s));
    LCODE_TRY(launch_unet_pool(T(n, T_E1), S1, S1, T(n, T_P1), s));
    LCODE_TRY(conv(n, 2, T(n, T_P1), nullptr, true, T(n, T_E2), nullptr, s));
    LCODE_TRY(launch_unet_pool(T(n, T_E2), S2, S2, T(n, T_P2), s));
    LCODE_TRY(conv(n, 3, T(n, T_P2), nullptr, true, T(n, T_E3), nullptr, s));
    LCODE_TRY(launch_unet_up(T(n, T_E3), S3, S3, T(n, T_U3), S2, S2, s));
    LCODE_TRY(conv(n, 4, T(n, T_U3), T(n, T_E2), true, T(n, T_D2), nullptr, s));
    LCODE_TRY(launch_unet_up(T(n, T_D2), S2, S2, T(n, T_U2), S1, S1, s));
    LCODE_TRY(conv(n, 5, T(n, T_U2), T(n, T_E1), true, T(n, T_D1), nullptr, s));
    LCODE_TRY(launch_unet_up(T(n, T_D1), S1, S1, T(n, T_U1), S0, S0, s));
    LCODE_TRY(conv(n, 6, T(n, T_U1), T(n, T_E0), true, T(n, T_D0), nullptr, s));
    LCODE_TRY(conv(n, 7, T(n, T_D0), nullptr, false, nullptr, n->y, s));
    return LCODE_OK;
}

// The forward without the pooled and upsampled copies: each convolution
// reads its pooled or upsampled input through its staging (inference).
int lcode_unet_infer(const lcode_unet* n, const uint16_t* x, hipStream_t s) {
    int rc = check_net(n);
    if (rc) return rc;
    LCODE_REQUIRE(x, "NULL input planes");
    const int S0 = n->S, S1 = level_size(S0, 1), S2 = level_size(S0, 2),
              S3 = level_size(S0, 3);
    const int sz[4] = {S0, S1, S2, S3};
    auto fused = [&](int k, const uint16_t* in0, int mode0, int lv0, const uint16_t* in1,
                     bool relu, uint16_t* out0, float* out_f32) {
        const LayerShape& L = kLayers[k];
        UnetConvArgs a{};
        a.in[0] = in0;
        a.in[1] = in1;
        a.in_mode[0] = mode0;
        a.in_h[0] = a.in_w[0] = sz[lv0];
        a.nin = L.cinp / 32;
        a.w = n->wpack + pack_offset(k, false);
        a.bias = n->params + param_offset(k) + (int64_t)L.cout * L.cin * 9;
        a.H = a.W = sz[L.level];
        a.cout = L.cout;
        a.cout_pad = L.cout_pad;
        a.relu = relu;
        a.out[0] = out0;
        a.out_f32 = out_f32;
        return launch_unet_conv(a, s);
    };
    LCODE_TRY(fused(0, x, 0, 0, nullptr, true, T(n, T_E0), nullptr));
    LCODE_TRY(fused(1, T(n, T_E0), 1, 0, nullptr, true, T(n, T_E1), nullptr));
    LCODE_TRY(fused(2, T(n, T_E1), 1, 1, nullptr, true, T(n, T_E2), nullptr));
    LCODE_TRY(fused(3, T(n, T_E2), 1, 2, nullptr, true, T(n, T_E3), nullptr));
    LCODE_TRY(fused(4, T(n, T_E3), 2, 3, T(n, T_E2), true, T(n, T_D2), nullptr));
    LCODE_TRY(fused(5, T(n, T_D2), 2, 2, T(n, T_E1), true, T(n, T_D1), nullptr));
    LCODE_TRY(fused(6, T(n, T_D1), 2, 1, T(n, T_E0), true, T(n, T_D0), nullptr));
    LCODE_TRY(fused(7, T(n, T_D0), 0, 0, nullptr, false, nullptr, n->y));
    return LCODE_OK;
}

int lcode_unet_backward(const lcode_unet* n, const uint16_t* x, const float* targets,
                        hipStream_t s) {
    int rc = check_net(n);
    if (rc) return rc;
    LCODE_REQUIRE(x && targets && n->grad && n->partial && n->loss_part && n->loss,
                  "NULL input, targets or training buffer");
    const int S0 = n->S, S1 = level_size(S0, 1), S2 = level_size(S0, 2),
              S3 = level_size(S0, 3);
    LCODE_TRY(launch_unet_mse(n->y, targets, S0 * S0, T(n, T_GO), n->loss_part, n->loss, s));
    LCODE_TRY(wgrad(n, 7, T(n, T_D0), nullptr, T(n, T_GO), s));
    LCODE_TRY(conv_t(n, 7, T(n, T_GO), T(n, T_D0), T(n, T_GD0), nullptr, s));
    LCODE_TRY(wgrad(n, 6, T(n, T_U1), T(n, T_E0), T(n, T_GD0), s));
    LCODE_TRY(conv_t(n, 6, T(n, T_GD0), nullptr, T(n, T_GU1), T(n, T_GS0), s));
    LCODE_TRY(launch_unet_up_t(T(n, T_GU1), S0, S0, T(n, T_D1), T(n, T_GD1), S1, S1, s));
    LCODE_TRY(wgrad(n, 5, T(n, T_U2), T(n, T_E1), T(n, T_GD1), s));
    LCODE_TRY(conv_t(n, 5, T(n, T_GD1), nullptr, T(n, T_GU2), T(n, T_GS1), s));
    LCODE_TRY(launch_unet_up_t(T(n, T_GU2), S1, S1, T(n, T_D2), T(n, T_GD2), S2, S2, s));
    LCODE_TRY(wgrad(n, 4, T(n, T_U3), T(n, T_E2), T(n, T_GD2), s));
    LCODE_TRY(conv_t(n, 4, T(n, T_GD2), nullptr, T(n, T_GU3), T(n, T_GS2), s));
    LCODE_TRY(launch_unet_up_t(T(n, T_GU3), S2, S2, T(n, T_E3), T(n, T_GE3), S3, S3, s));
    LCODE_TRY(wgrad(n, 3, T(n, T_P2), nullptr, T(n, T_GE3), s));
    LCODE_TRY(conv_t(n, 3, T(n, T_GE3), nullptr, T(n, T_GP2), nullptr, s));
    LCODE_TRY(launch_unet_pool_t(T(n, T_GP2), S3, S3, T(n, T_GS2), T(n, T_E2), T(n, T_GE2),
                                 S2, S2, s));
    LCODE_TRY(wgrad(n, 2, T(n, T_P1), nullptr, T(n, T_GE2), s));
    LCODE_TRY(conv_t(n, 2, T(n, T_GE2), nullptr, T(n, T_GP1), nullptr, s));
    LCODE_TRY(launch_unet_pool_t(T(n, T_GP1), S2, S2, T(n, T_GS1), T(n, T_E1), T(n, T_GE1),
                                 S1, S1, s));
    LCODE_TRY(wgrad(n, 1, T(n, T_P0), nullptr, T(n, T_GE1), s));
    LCODE_TRY(conv_t(n, 1, T(n, T_GE1), nullptr, T(n, T_GP0), nullptr, s));
    LCODE_TRY(launch_unet_pool_t(T(n, T_GP0), S1, S1, T(n, T_GS0), T(n, T_E0), T(n, T_GE0),
                                 S0, S0, s));
    LCODE_TRY(wgrad(n, 0, x, nullptr, T(n, T_GE0), s));
    return LCODE_OK;
}

int lcode_unet_adam(const lcode_unet* n, int32_t t, float lr, hipStream_t s) {
    int rc = check_net(n);
    if (rc) return rc;
    LCODE_REQUIRE(t >= 1 && n->grad && n->adam_m && n->adam_v, "bad Adam step or NULL moments");
    LCODE_TRY(launch_unet_adam(n->params, n->grad, n->adam_m, n->adam_v,
                               (int)param_offset(8), lr, t, s));
    return lcode_unet_pack(n, s);
}

int lcode_unet_io(const lcode_unet* n, const lcode_unet_io_args* io, int32_t which,
                  hipStream_t s) {
    int rc = check_net(n);
    if (rc) return rc;
    LCODE_REQUIRE(io && which >= 0 && which <= 2, "bad io call");
    LCODE_REQUIRE(io->grid_steps >= 5 && (io->grid_steps + 3) / 4 == n->S,
                  "the network's level-0 size must be (grid_steps + 3) / 4");
    UnetIoArgs a{};
    for (int f = 0; f < 6; ++f) {
        a.f1[f] = io->f1[f];
        a.f2[f] = io->f2[f];
        a.f3[f] = io->f3[f];
        a.f4[f] = io->f4[f];
        a.next[f] = io->next[f];
        a.guess[f] = io->guess[f];
        a.tscale[f] = io->tscale[f];
    }
    for (int c = 0; c < 14; ++c) a.scale[c] = io->scale[c];
    a.ro = io->ro;
    a.beam = io->beam;
    a.beam_scalar = io->beam_scalar;
    a.f32 = io->f32;
    a.N = io->grid_steps;
    a.S = n->S;
    a.planes = io->planes;
    a.planes_f32 = io->planes_f32;
    a.targets = io->targets;
    a.y = n->y;
    if (which == 0) {
        for (int f = 0; f < 6; ++f) LCODE_REQUIRE(a.f1[f] && a.f2[f], "NULL field");
        LCODE_REQUIRE(a.ro && (a.planes || a.planes_f32), "NULL density or planes");
    } else {
        for (int f = 0; f < 6; ++f) LCODE_REQUIRE(a.f1[f] && a.f2[f] && a.f3[f], "NULL field");
        if (which == 1) {
            for (int f = 0; f < 6; ++f) LCODE_REQUIRE(a.next[f], "NULL target field");
            LCODE_REQUIRE(a.targets, "NULL targets");
        } else {
            for (int f = 0; f < 6; ++f) LCODE_REQUIRE(a.guess[f], "NULL guess field");
        }
    }
    return launch_unet_io(a, which, s);
#undef LCODE_TRY
}

}  // extern "C"
