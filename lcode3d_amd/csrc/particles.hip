// Particle kernels: field-free estimate, fused interpolation + pusher, and
// coarse->fine interpolation + TSC deposition.  FP64 throughout.

#include "common.hpp"

namespace lcode {

namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ int clampi(int v, int n) {
    return v < 0 ? 0 : (v >= n ? n - 1 : v);
}

// Field-free half-step estimate with coordinate-only reflection using >= and
// <= (lcode.py:369-391).  Returns the estimated offsets.
__device__ __forceinline__ void estimate_one(double m, double xi0, double yi0,
                                             double xo, double yo, double px,
                                             double py, double pz, double dxi,
                                             double r, double& ex,
                                             double& ey) {
    double x = xi0 + xo, y = yi0 + yo;
    const double gamma_m = sqrt(m * m + pz * pz + px * px + py * py);
    x += px / (gamma_m - pz) * dxi;
    y += py / (gamma_m - pz) * dxi;
    if (x >= +r) x = +2 * r - x;
    if (x <= -r) x = -2 * r - x;
    if (y >= +r) y = +2 * r - y;
    if (y <= -r) y = -2 * r - y;
    ex = x - xi0;
    ey = y - yi0;
}

__global__ void __launch_bounds__(kBlock)
estimate_kernel(EstimateArgs a) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.n) return;
    double ex, ey;
    estimate_one(a.m[k], a.x_init[k], a.y_init[k], a.prev_x_offt[k],
                 a.prev_y_offt[k], a.px[k], a.py[k], a.pz[k], a.dxi,
                 a.reflect, ex, ey);
    a.out_x[k] = ex;
    a.out_y[k] = ey;
}

// move_smart_kernel (lcode.py:704-793): midpoint position, 6-field TSC
// gather, two fixed-point iterations of the half-step momentum, offset and
// momentum update, reflection with strict > / < flipping the momentum.
__global__ void __launch_bounds__(kBlock)
push_kernel(PushArgs a) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.n) return;
    const double m = a.m[k], q = a.q[k];
    const double opx = a.px[k], opy = a.py[k], opz = a.pz[k];
    const double xi0 = a.x_init[k], yi0 = a.y_init[k];
    double x_offt = a.prev_x_offt[k], y_offt = a.prev_y_offt[k];
    double ex, ey;
    if (a.est_x) {
        ex = a.est_x[k];
        ey = a.est_y[k];
    } else {
        estimate_one(m, xi0, yi0, x_offt, y_offt, opx, opy, opz, a.dxi,
                     a.reflect, ex, ey);
    }
    const double x_half = xi0 + (x_offt + ex) / 2;
    const double y_half = yi0 + (y_offt + ey) / 2;
    Tsc t;
    tsc_weights(x_half, y_half, a.N, a.h, t, a.fault);
    const double Ex = interp9(a.F[0], a.N, t);
    const double Ey = interp9(a.F[1], a.N, t);
    const double Ez = interp9(a.F[2], a.N, t);
    const double Bx = interp9(a.F[3], a.N, t);
    const double By = interp9(a.F[4], a.N, t);
    const double Bz = interp9(a.F[5], a.N, t);

    double px = opx, py = opy, pz = opz;
    double dpx = 0, dpy = 0, dpz = 0;
#pragma unroll
    for (int it = 0; it < 2; ++it) {
        const double gamma_m = sqrt(m * m + pz * pz + px * px + py * py);
        const double vx = px / gamma_m, vy = py / gamma_m, vz = pz / gamma_m;
        const double factor_1 = q * a.dxi / (1 - pz / gamma_m);
        dpx = factor_1 * (Ex + vy * Bz - vz * By);
        dpy = factor_1 * (Ey - vx * Bz + vz * Bx);
        dpz = factor_1 * (Ez + vx * By - vy * Bx);
        px = opx + dpx / 2;
        py = opy + dpy / 2;
        pz = opz + dpz / 2;
    }
    const double gamma_m = sqrt(m * m + pz * pz + px * px + py * py);
    x_offt += px / (gamma_m - pz) * a.dxi;
    y_offt += py / (gamma_m - pz) * a.dxi;
    px = opx + dpx;
    py = opy + dpy;
    pz = opz + dpz;

    const double r = a.reflect;
    double x = xi0 + x_offt, y = yi0 + y_offt;
    if (x > +r) { x = +2 * r - x; x_offt = x - xi0; px = -px; }
    if (x < -r) { x = -2 * r - x; x_offt = x - xi0; px = -px; }
    if (y > +r) { y = +2 * r - y; y_offt = y - yi0; py = -py; }
    if (y < -r) { y = -2 * r - y; y_offt = y - yi0; py = -py; }

    a.out_x[k] = x_offt;
    a.out_y[k] = y_offt;
    a.out_px[k] = px;
    a.out_py[k] = py;
    a.out_pz[k] = pz;
}

// Coarse->fine interpolation of fine particle (fi, fj) (lcode.py:591-621)
// and its deposition values (lcode.py:651-655).
struct FineParticle {
    double x, y, v[4];   // position; dro, djx, djy, djz
};

__device__ __forceinline__ void fine_particle(const DepositArgs& a, int fi,
                                              int fj, FineParticle& p) {
    const double ipi = a.ip[fi], ini = a.in[fi];
    const double ipj = a.ip[fj], inj = a.in[fj];
    const double A = ipi * ipj, B = ipi * inj, C = ini * ipj, D = ini * inj;
    const int nc = a.nc;
    // neighbour indices come from the caller: clamp so that bad ones cannot
    // read outside the coarse arrays
    const int pi = clampi((int)a.P[fi], nc), ni = clampi((int)a.Nn[fi], nc);
    const int pj = clampi((int)a.P[fj], nc), nj = clampi((int)a.Nn[fj], nc);
    const int pp = pi * nc + pj, pn = pi * nc + nj;
    const int np = ni * nc + pj, nn = ni * nc + nj;
#define LCODE_MIX(arr) (A * arr[pp] + B * arr[pn] + C * arr[np] + D * arr[nn])
    p.x = a.fg[fi] + LCODE_MIX(a.x_offt);
    p.y = a.fg[fj] + LCODE_MIX(a.y_offt);
    const double s = a.smallness;
    const double m = s * LCODE_MIX(a.m);
    const double q = s * LCODE_MIX(a.q);
    const double px = s * LCODE_MIX(a.px);
    const double py = s * LCODE_MIX(a.py);
    const double pz = s * LCODE_MIX(a.pz);
#undef LCODE_MIX
    const double gamma_m = sqrt(m * m + px * px + py * py + pz * pz);
    const double dro = q / (1 - pz / gamma_m);
    p.v[0] = dro;
    p.v[1] = px * (dro / gamma_m);
    p.v[2] = py * (dro / gamma_m);
    p.v[3] = pz * (dro / gamma_m);
}

// Tiles of the owner-computes deposition: a block owns kTile x kTile cells
// and deposits, in LDS, every fine particle whose cell moved at most
// a.halo cells from its initial cell and whose stencil touches the tile.
constexpr int kTile = 32;
constexpr int kDepBlock = 256;

// First fine index whose initial cell is >= c (fg is increasing).
__device__ int first_fine_at_or_after(const double* fg, int nf, int c, int N,
                                      double h) {
    int lo = 0, hi = nf;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (cell_of(fg[mid], N, h) < c) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// deposit_kernel (lcode.py:626-663) as owner-computes tiles: each block
// accumulates its kTile^2 cells of ro, jx, jy, jz in LDS (FP64 LDS atomics),
// then stores them once, adding the ion background last (lcode.py:687).
// Every cell of the grid is written by exactly one block, so the outputs
// need no zeroing.  Work is split by initial cell: a thread takes the
// fineness^2 fine particles that start in one cell, sums in registers the
// contributions of those that land in the same cell as the first one, and
// adds them to LDS with one set of 36 atomics.  A particle displaced by
// more than a.halo cells is left to deposit_spill_kernel, appended to the
// spill list by its home tile (the tile owning its initial cell).
constexpr int kMaxExt = kTile + 2 * 16 + 3;   // extended rows for halo <= 16

// Adds val[c][k] (or v[c] * w[k] when val is NULL) to the stencil points of
// cell (i, j) that lie in this block's tile.
__device__ __forceinline__ void lds_deposit(double (*acc)[kTile][kTile],
                                            int i, int j, int ci0, int cj0,
                                            const double* v, const double* w,
                                            const double (*val)[9]) {
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const int si = i + stencil_di(k) - ci0;
        const int sj = j + stencil_dj(k) - cj0;
        if (si >= 0 && si < kTile && sj >= 0 && sj < kTile) {
#pragma unroll
            for (int c = 0; c < 4; ++c)
                atomicAdd(&acc[c][si][sj], val ? val[c][k] : v[c] * w[k]);
        }
    }
}

__global__ void __launch_bounds__(kDepBlock)
deposit_tile_kernel(DepositArgs a) {
    __shared__ double acc[4][kTile][kTile];
    __shared__ int fsI[kMaxExt + 1], fsJ[kMaxExt + 1];
    const int N = a.N;
    const int ci0 = blockIdx.y * kTile, cj0 = blockIdx.x * kTile;
    const int ext = kTile + 2 * a.halo + 2;       // initial cells per side
    const int ci_lo = ci0 - 1 - a.halo, cj_lo = cj0 - 1 - a.halo;
    for (int k = threadIdx.x; k < 4 * kTile * kTile; k += blockDim.x)
        (&acc[0][0][0])[k] = 0.;
    for (int r = threadIdx.x; r <= ext; r += blockDim.x) {
        fsI[r] = first_fine_at_or_after(a.fg, a.nf, ci_lo + r, N, a.h);
        fsJ[r] = first_fine_at_or_after(a.fg, a.nf, cj_lo + r, N, a.h);
    }
    __syncthreads();
    for (int task = threadIdx.x; task < ext * ext; task += blockDim.x) {
        const int r = task / ext, c = task - r * ext;
        const int ci = ci_lo + r, cj = cj_lo + c;   // initial cell
        const bool home = ci >= ci0 && ci < ci0 + kTile &&
                          cj >= cj0 && cj < cj0 + kTile;
        double sum[4][9];
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int k = 0; k < 9; ++k) sum[q][k] = 0.;
        int i0 = -1, j0 = -1;
        for (int fi = fsI[r]; fi < fsI[r + 1]; ++fi) {
            for (int fj = fsJ[c]; fj < fsJ[c + 1]; ++fj) {
                FineParticle fp;
                fine_particle(a, fi, fj, fp);
                Tsc t;
                tsc_weights(fp.x, fp.y, N, a.h, t, home ? a.fault : nullptr);
                if (abs(t.i - ci) > a.halo || abs(t.j - cj) > a.halo) {
                    if (home) a.spill[atomicAdd(a.spill_count, 1)] = fi * a.nf + fj;
                    continue;
                }
                if (i0 < 0) {
                    i0 = t.i;
                    j0 = t.j;
                }
                if (t.i == i0 && t.j == j0) {
#pragma unroll
                    for (int q = 0; q < 4; ++q)
#pragma unroll
                        for (int k = 0; k < 9; ++k) sum[q][k] += fp.v[q] * t.w[k];
                } else {
                    lds_deposit(acc, t.i, t.j, ci0, cj0, fp.v, t.w, nullptr);
                }
            }
        }
        if (i0 >= 0) lds_deposit(acc, i0, j0, ci0, cj0, nullptr, nullptr, sum);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < kTile * kTile; k += blockDim.x) {
        const int r = k / kTile, c = k % kTile;
        const int i = ci0 + r, j = cj0 + c;
        if (i >= N || j >= N) continue;
        const size_t idx = (size_t)i * N + j;
        double ro = acc[0][r][c];
        if (a.ro_initial) ro += a.ro_initial[idx];
        a.ro[idx] = ro;
        a.jx[idx] = acc[1][r][c];
        a.jy[idx] = acc[2][r][c];
        a.jz[idx] = acc[3][r][c];
    }
}

// Atomic deposition of the spilled particles; the last block to finish
// resets the list for the next call (every block has read the count by
// then).
__global__ void __launch_bounds__(kBlock)
deposit_spill_kernel(DepositArgs a) {
    const int count = atomicAdd(a.spill_count, 0);
    const int N = a.N;
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < count;
         p += gridDim.x * blockDim.x) {
        const int fk = a.spill[p];
        FineParticle fp;
        fine_particle(a, fk / a.nf, fk % a.nf, fp);
        Tsc t;
        tsc_weights(fp.x, fp.y, N, a.h, t, nullptr);
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            const size_t idx = (size_t)(t.i + stencil_di(k)) * N + (t.j + stencil_dj(k));
            unsafeAtomicAdd(a.ro + idx, fp.v[0] * t.w[k]);
            unsafeAtomicAdd(a.jx + idx, fp.v[1] * t.w[k]);
            unsafeAtomicAdd(a.jy + idx, fp.v[2] * t.w[k]);
            unsafeAtomicAdd(a.jz + idx, fp.v[3] * t.w[k]);
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        if (atomicAdd(a.spill_done, 1) == (int)gridDim.x - 1) {
            atomicExch(a.spill_count, 0);
            atomicExch(a.spill_done, 0);
        }
    }
}

constexpr int kSpillBlocks = 64;

inline int blocks_for(long n) { return (int)((n + kBlock - 1) / kBlock); }

}  // namespace

int launch_estimate(const EstimateArgs& a, hipStream_t s, Profiler* prof) {
    if (a.n == 0) return LCODE_OK;
    ProfScope ps(prof, "estimate", s);
    estimate_kernel<<<blocks_for(a.n), kBlock, 0, s>>>(a);
    LCODE_HIP(hipGetLastError());
    return LCODE_OK;
}

int launch_push(const PushArgs& a, hipStream_t s, Profiler* prof) {
    if (a.n == 0) return LCODE_OK;
    ProfScope ps(prof, "push", s);
    push_kernel<<<blocks_for(a.n), kBlock, 0, s>>>(a);
    LCODE_HIP(hipGetLastError());
    return LCODE_OK;
}

int launch_deposit(const DepositArgs& a, hipStream_t s, Profiler* prof) {
    const dim3 grid((a.N + kTile - 1) / kTile, (a.N + kTile - 1) / kTile);
    {
        ProfScope ps(prof, "deposit_tile", s);
        deposit_tile_kernel<<<grid, kDepBlock, 0, s>>>(a);
    }
    LCODE_HIP(hipGetLastError());
    {
        ProfScope ps(prof, "deposit_spill", s);
        deposit_spill_kernel<<<kSpillBlocks, kBlock, 0, s>>>(a);
    }
    LCODE_HIP(hipGetLastError());
    return LCODE_OK;
}

}  // namespace lcode
