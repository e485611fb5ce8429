// Particle kernels: field-free estimate, fused interpolation + pusher,
// coarse->fine interpolation + TSC deposition, and the device-side plasma
// set-up.  Particle positions and momenta are FP64 in every form; the
// deposition's arithmetic and LDS sums are FP64 too (its output grids are
// FP32 in the mixed form).

#include "common.hpp"

namespace lcode {

namespace {

constexpr int kBlock = 256;
// Blocks of the grid-wide spill deposition (build option).
#ifndef LCODE_SPILL_BLOCKS
#define LCODE_SPILL_BLOCKS 256
#endif
constexpr int kSpillBlocks = LCODE_SPILL_BLOCKS;

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
// One-wave blocks spread the (few) coarse particles over all CUs.
#ifndef LCODE_PUSH_BLOCK
#define LCODE_PUSH_BLOCK 64
#endif
constexpr int kPushBlock = LCODE_PUSH_BLOCK;
// Deposition tiles per row that the push's per-tile bound table holds
// (more: each particle raises its tile's slot with a device atomic).
constexpr int kRegionRowMax = 128;
// A launch has about one wave per SIMD, so occupancy buys nothing: with a
// low waves-per-EU target the compiler issues the 54 field gathers of a
// particle together instead of one round trip at a time.
#ifndef LCODE_PUSH_WAVES
#define LCODE_PUSH_WAVES 2
#endif
#if LCODE_PUSH_WAVES > 0
#define LCODE_PUSH_ATTR __attribute__((amdgpu_waves_per_eu(1, LCODE_PUSH_WAVES)))
#else
#define LCODE_PUSH_ATTR
#endif

#ifndef LCODE_PUSH_BOUND_READ
#define LCODE_PUSH_BOUND_READ 1
#endif

// One coarse particle's push; returns max(|x_offt|, |y_offt|) after it and
// the particle's initial position in xi0, yi0.
// G: storage type of the field grids (the momentum update is FP64 in both).
template <typename G>
__device__ __forceinline__ double push_one(const PushArgs& a, int k, double& xi0,
                                           double& yi0) {
    const double m = a.m[k], q = a.q[k];
    const double opx = a.px[k], opy = a.py[k], opz = a.pz[k];
    xi0 = a.x_init[k];
    yi0 = a.y_init[k];
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
    const double Ex = interp9(gp<G>(a.F[0]), a.N, t);
    const double Ey = interp9(gp<G>(a.F[1]), a.N, t);
    const double Ez = interp9(gp<G>(a.F[2]), a.N, t);
    const double Bx = interp9(gp<G>(a.F[3]), a.N, t);
    const double By = interp9(gp<G>(a.F[4]), a.N, t);
    const double Bz = interp9(gp<G>(a.F[5]), a.N, t);

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
    // infinity marks a non-finite particle (fmax alone would drop a NaN)
    const bool finite = isfinite(x_offt) && isfinite(y_offt) && isfinite(px) &&
                        isfinite(py) && isfinite(pz);
    return finite ? fmax(fabs(x_offt), fabs(y_offt)) : __builtin_inf();
}

template <typename G>
__global__ void __launch_bounds__(kPushBlock) LCODE_PUSH_ATTR
push_kernel(PushArgs a) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int k0 = blockIdx.x * blockDim.x;
    // the block's first initial row, for the per-tile bound table: loaded
    // with the particle data, not after the stores (which may alias it)
    const double x_first = a.region_max ? a.x_init[k0] : 0.;
    double d = 0., xi0 = 0., yi0 = 0.;
    if (k < a.n && gate_open(a.gate)) d = push_one<G>(a, k, xi0, yi0);
    if (a.max_cells) {
        // cells the particles moved from their initial positions, at most
        // (the deposition sizes its tile halo by it); a non-finite state,
        // or one beyond kFarCells cells, is the kNonFinite mark, on which
        // the deposition writes NaN instead of depositing
        const double c = ceil(d / a.h);
        int cells = c <= kFarCells ? (int)c : kNonFinite;
        if (a.region_max) {
            // the bound of the deposition tile that holds the particle's
            // initial cell: the block's particles (consecutive, fewer than
            // a coarse row) lie in at most two rows of tiles, max-reduced
            // in an LDS table first (one LDS atomic per particle), then one
            // device atomic per tile touched (a slot takes those of the ~10
            // blocks over its particles); a shuffle loop over the wave's
            // tiles instead cost the push ~3 us of dependent latency
            __shared__ int rtab[2 * kRegionRowMax];
            const int nr = a.region_n, rn = nr - 1;
            const int ri0 = min(max(cell_of(x_first, a.N, a.h) / a.region_t, 0), rn);
            for (int e = threadIdx.x; e < 2 * kRegionRowMax; e += blockDim.x) rtab[e] = 0;
            __syncthreads();
            if (cells > 0) {   // (a particle that moved exists: k < n)
                const int ri = min(max(cell_of(xi0, a.N, a.h) / a.region_t, 0), rn);
                const int rj = min(max(cell_of(yi0, a.N, a.h) / a.region_t, 0), rn);
                const int dr = ri - ri0;
                if (dr >= 0 && dr < 2 && nr <= kRegionRowMax)
                    atomicMax(&rtab[dr * kRegionRowMax + rj], cells);
                else
                    atomicMax(a.region_max + ri * nr + rj, cells);
            }
            __syncthreads();
            if (nr <= kRegionRowMax) {
                for (int e = threadIdx.x; e < 2 * nr; e += blockDim.x) {
                    const int dr = e / nr, rj = e - dr * nr;
                    const int v = rtab[dr * kRegionRowMax + rj];
                    if (v > 0 && ri0 + dr <= rn) atomicMax(a.region_max + (ri0 + dr) * nr + rj, v);
                }
            }
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) cells = max(cells, __shfl_xor(cells, off, 64));
        // one atomic per wave, and (LCODE_PUSH_BOUND_READ) only when it raises
        // the bound: after the first waves the bound is reached and the rest
        // only read it (an atomic from every wave to the one address took the
        // push at 641^2 from 9.0 to 13.7 us)
        if ((threadIdx.x & 63) == 0) {
            if (!LCODE_PUSH_BOUND_READ ||
                cells > __hip_atomic_load(a.max_cells, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                atomicMax(a.max_cells, cells);
            if (cells == kNonFinite) {
                if (a.host_flag)
                    __hip_atomic_store(a.host_flag, 1, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_SYSTEM);
            }
        }
    }
}

// Deposition values of a fine particle with mass m, charge q, momentum p
// (lcode.py:651-655): dro = q / (1 - pz / gamma), dj = p dro / gamma,
// evaluated with one division as dro = q gamma r, dj = q p r,
// r = 1 / (gamma - pz) (the same quantities, rounded differently).
// G = float: the FP32-grid form, in FP32 arithmetic.
template <typename G>
__device__ __forceinline__ void deposit_values(G m, G q, G px, G py, G pz,
                                               G* v) {
    const G gamma_m = sqrt(m * m + px * px + py * py + pz * pz);
    const G qr = q / (gamma_m - pz);
    v[0] = gamma_m * qr;
    v[1] = px * qr;
    v[2] = py * qr;
    v[3] = pz * qr;
}

// Coarse->fine interpolation of fine particle (fi, fj) (lcode.py:591-621)
// and its deposition values, reading everything from global memory.
template <typename G>
struct FineParticle {
    double x, y;         // position (FP64 in every form)
    G v[4];              // dro, djx, djy, djz
};

template <typename G>
__device__ __forceinline__ void fine_particle(const DepositArgs& a, int fi,
                                              int fj, FineParticle<G>& p) {
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
    deposit_values<G>((G)(s * LCODE_MIX(a.m)), (G)(s * LCODE_MIX(a.q)),
                      (G)(s * LCODE_MIX(a.px)), (G)(s * LCODE_MIX(a.py)),
                      (G)(s * LCODE_MIX(a.pz)), p.v);
#undef LCODE_MIX
}

// Tiles of the owner-computes deposition: a block owns kTile x kTile cells
// and deposits, in LDS, every fine particle whose cell moved at most
// a.halo cells from its initial cell and whose stencil touches the tile.
constexpr int kTile = kDepTile;
constexpr int kDepBlock = LCODE_DEP_BLOCK;
#ifdef LCODE_DEP_WAVES
#define LCODE_DEP_ATTR __attribute__((amdgpu_waves_per_eu(LCODE_DEP_WAVES, LCODE_DEP_WAVES)))
#else
#define LCODE_DEP_ATTR
#endif

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
// accumulates its kTile^2 cells of ro, jx, jy, jz in LDS (FP64 LDS atomics,
// ds_add_f64), then stores them once, adding the ion background last
// (lcode.py:687).  Every cell of the grid is written by exactly one block,
// so the outputs need no zeroing.  A block visits the fine particles whose
// initial cell lies within halo + 1 cells of its tile; their per-index
// interpolation data and, when they fit, the coarse particles they
// interpolate from are staged in LDS first; then one thread per fine
// particle adds its 9 x 4 weighted values to the LDS tile.  A particle
// displaced by more than a.halo cells from its initial cell is appended to
// the spill list by its home tile (the tile owning its initial cell), and
// deposit_spill_kernel, launched next, adds the list with device atomics
// across the whole grid.  A non-finite particle deposits nothing: the tile
// (or, when the pushes reported one, every tile) is written as NaN.
// A split tile is deposited by two blocks, each over alternate runs of its
// particle groups: the tiles with the most work then take about half as
// long, using block slots the launch leaves free; the second block to
// finish adds the first one's sums (published in global memory) and
// stores the tile.
constexpr int kFineMax = kDepFineMax;       // fine indices per side
constexpr int kCoarseMax = kDepCoarseMax;   // coarse particles per side staged in LDS
constexpr int kExtMax = kTile + 14;   // initial cells per side: kTile + 2 halo + 2, halo <= 6

struct FineTab {                 // per fine index along one axis
    double2 w[kFineMax];         // interpolation weights (ip, in)
    double fg[kFineMax];         // initial position
    int4 idx[kFineMax];          // coarse neighbours (P, Nn), initial cell, -
};

// First fine index whose initial cell is >= c, found by one wave (every
// lane calls it): 64 samples narrow it to a stride, 64 more pin it down.
__device__ int wave_first_fine(const double* fg, int nf, int c, int N,
                               double h) {
    const int lane = threadIdx.x & 63;
    if (nf > 64 * 64) return first_fine_at_or_after(fg, nf, c, N, h);
    const int stride = (nf + 63) / 64;
    const int i1 = lane * stride;
    const bool below1 = i1 < nf && cell_of(fg[i1], N, h) < c;
    const int n1 = __popcll(__ballot(below1));
    if (n1 == 0) return 0;
    const int base = (n1 - 1) * stride;
    const int i2 = base + lane;
    const bool below2 = lane < stride && i2 < nf && cell_of(fg[i2], N, h) < c;
    return base + __popcll(__ballot(below2));
}

// Adds v[c] * w[k] (or val[c][k]) to the stencil points of cell (i, j) that
// lie in this block's tile.
template <typename G>
__device__ __forceinline__ void lds_deposit(double (*acc)[kTile][kTile],
                                            int i, int j, int ci0, int cj0,
                                            const G* v, const G* w,
                                            const G (*val)[9]) {
    const bool inside = i - 1 >= ci0 && i + 1 < ci0 + kTile &&
                        j - 1 >= cj0 && j + 1 < cj0 + kTile;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const int si = i + stencil_di(k) - ci0;
        const int sj = j + stencil_dj(k) - cj0;
        if (inside || (si >= 0 && si < kTile && sj >= 0 && sj < kTile)) {
#pragma unroll
            for (int c = 0; c < 4; ++c)
                atomicAdd(&acc[c][si][sj], (double)(val ? val[c][k] : v[c] * w[k]));
        }
    }
}

template <typename A, typename G>
__device__ __forceinline__ void deposit_spill_one(const DepositArgs& a, int fk,
                                                  int N) {
    FineParticle<A> fp;
    fine_particle<A>(a, fk / a.nf, fk % a.nf, fp);
    TscT<A> t;
    tsc_weights<A>(fp.x, fp.y, N, a.h, t, nullptr, 1. / a.h);
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const size_t idx = (size_t)(t.i + stencil_di(k)) * N + (t.j + stencil_dj(k));
        unsafeAtomicAdd(gp<G>(a.ro) + idx, (G)(fp.v[0] * t.w[k]));
        unsafeAtomicAdd(gp<G>(a.jx) + idx, (G)(fp.v[1] * t.w[k]));
        unsafeAtomicAdd(gp<G>(a.jy) + idx, (G)(fp.v[2] * t.w[k]));
        unsafeAtomicAdd(gp<G>(a.jz) + idx, (G)(fp.v[3] * t.w[k]));
    }
}

// A: type of the per-particle arithmetic and register sums, G: storage
// type of the output grids (double, double: the plain path; double,
// float: the mixed form, whose density and currents are stored in FP32).
// The LDS tile accumulates in FP64: at 641^2 FP32 LDS atomics (ds_add_f32)
// kept the LDS pipeline 8x longer busy than FP64 ones (ds_add_f64) for the
// same instruction count (148 against 35 us per launch).  FP32 arithmetic
// (A = float) took the deposition from 49.8 to 47.9 us only (the kernel is
// bound by latency, not by its VALU work) while its rounding, relative to
// the background density 1, left the density response of a small beam
// (amplitude 0.01) with relative errors up to 4e-4 in Ez; the mixed form
// keeps FP64 arithmetic.
template <bool STAGED, typename A, typename G>
__device__ __forceinline__ void deposit_tile_body(const DepositArgs& a) {
    // the count slot of the next call is zeroed here (also when gated off)
    // unless a spill kernel follows, which does it
    if (a.spill_in_tiles && blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0)
        *a.spill_next = 0;
    if (!gate_open(a.gate)) return;   // block-uniform: before any barrier
    __shared__ double acc[4][kTile][kTile];
    __shared__ int last_block, split_last;
    __shared__ FineTab ti, tj;
    // the staged coarse particles in 16-byte pairs where the arithmetic
    // type allows, so that one LDS read brings two of a corner's values:
    // (x_offt, y_offt), (m, q), (px, py), pz
    using A2 = typename std::conditional<std::is_same<A, double>::value, double2, float2>::type;
    __shared__ double2 cpos[kCoarseMax * kCoarseMax];
    __shared__ A2 cmq[kCoarseMax * kCoarseMax], cpxy[kCoarseMax * kCoarseMax];
    __shared__ A cpz[kCoarseMax * kCoarseMax];
    __shared__ int range[4];
    __shared__ int tile_nan;
    const int N = a.N;
    const double inv_h = 1. / a.h;
    // tile of this block: row-major, or the host's order in a.tile_order
    // (which keeps neighbouring tiles on one XCD and gives each CU a mix of
    // the wake's heavy tiles and light ones)
    // A split tile's two parts (the tile's own block, part 0, and one of the
    // nsplit blocks after the tiles, part 1) take alternate runs of
    // blockDim.x particle groups.
    const int ntx = a.ntx, ntiles = ntx * ntx, lin = blockIdx.x;
    int tile, slot = -1, part = 0;
    if (lin >= ntiles) {
        slot = lin - ntiles;
        tile = a.split_tile[slot];
        part = 1;
    } else {
        tile = a.tile_order ? a.tile_order[lin] : lin;
        if (a.nsplit) slot = a.split_slot[tile];
    }
    const int nparts = slot >= 0 ? 2 : 1;
    const int ty = tile / ntx, tx = tile - ty * ntx;
    const int ci0 = ty * kTile, cj0 = tx * kTile;
    // halo: the displacement bound of the last pushes when one is given.
    // A fine particle's offset is a convex mix of coarse offsets, so its
    // cell moved at most `bound` cells: with bound <= a.halo no particle
    // can leave the halo, and every block (all read the same bound) skips
    // the spill list.  Otherwise particles beyond the halo take the exact
    // spill path.
    // With per-tile bounds the tile takes the largest of its 3 x 3 tile
    // neighbourhood: the fine particles it visits (initial cells within
    // halo + 1 of it) interpolate from coarse particles whose initial cells
    // lie within coarseness + 1 more cells, inside that neighbourhood (the
    // host enables the per-tile bounds only when that holds), and every tile
    // a particle's stencil reaches has the particle's home tile in its
    // neighbourhood, so all of them visit it, or (beyond a.halo) none does
    // and its home tile lists it.
    const int global = a.halo_cells ? max(*a.halo_cells, 0) : a.halo + 1;
    // the pushes met a non-finite (or absurdly far) particle: the whole
    // deposition is NaN, with no particle visited (already reported)
    const bool nonfinite = global >= kNonFinite;
    int bound = global;
    if (a.halo_regions && !nonfinite) {
        bound = 0;
        for (int di = -1; di <= 1; ++di)
            for (int dj = -1; dj <= 1; ++dj) {
                const int r = ty + di, c = tx + dj;
                if (r >= 0 && r < ntx && c >= 0 && c < ntx)
                    bound = max(bound, a.halo_regions[r * ntx + c]);
            }
    }
    const int halo = min(bound, a.halo);
    const bool may_spill = bound > a.halo;
    if (threadIdx.x == 0 && part == 0) {
        // zero the slots the next pushes raise; nothing reads them here
        if (a.halo_next && tile == 0) *a.halo_next = 0;
        if (a.halo_regions_next) a.halo_regions_next[tile] = 0;
    }
    if (threadIdx.x == 0) tile_nan = 0;
    const int ext = kTile + 2 * halo + 2;
    const int ci_lo = ci0 - 1 - halo, cj_lo = cj0 - 1 - halo;
    for (int k = threadIdx.x; k < 4 * kTile * kTile; k += blockDim.x)
        (&acc[0][0][0])[k] = 0;
    {
        const int wave = threadIdx.x >> 6;
        if (wave < 4) {
            const int lo = (wave < 2) ? ci_lo : cj_lo;
            const int r = wave_first_fine(a.fg, a.nf, lo + ((wave & 1) ? ext : 0), N, a.h);
            if ((threadIdx.x & 63) == 0) range[wave] = r;
        }
    }
    __syncthreads();
    const int fi_a = range[0], ni = range[1] - range[0];
    const int fj_a = range[2], nj = range[3] - range[2];
    const bool overfull = ni > kFineMax || nj > kFineMax || ext > kExtMax;
    if (overfull) {
        // denser fine grid than the host-checked fineness allows: flag the
        // tile as faulted and fill it with NaN rather than drop particles
        if (threadIdx.x == 0 && part == 0) atomicAdd(a.fault, 1);
    } else if (!nonfinite) {
        for (int k = threadIdx.x; k < ni + nj; k += blockDim.x) {
            const bool isj = k >= ni;
            const int f = isj ? fj_a + k - ni : fi_a + k;
            FineTab& t = isj ? tj : ti;
            const int s = isj ? k - ni : k;
            t.w[s] = make_double2(a.ip[f], a.in[f]);
            t.fg[s] = a.fg[f];
            t.idx[s] = make_int4(clampi((int)a.P[f], a.nc), clampi((int)a.Nn[f], a.nc),
                                 cell_of(a.fg[f], N, a.h), 0);
        }
        __syncthreads();
        // coarse rectangle referenced by this block's fine particles
        int cI0 = 0, cJ0 = 0, cW = 0;
        if constexpr (STAGED) {
            // indices_prev/next grow with the fine index, so the rectangle
            // is [P[first], Nn[last]] on each axis; the host checked the fit
            cI0 = ni ? ti.idx[0].x : 0;
            cJ0 = nj ? tj.idx[0].x : 0;
            int cH = ni ? ti.idx[ni - 1].y - cI0 + 1 : 0;
            cW = nj ? tj.idx[nj - 1].y - cJ0 + 1 : 0;
            if (cH > kCoarseMax || cW > kCoarseMax || cH < 0 || cW < 0) {
                if (threadIdx.x == 0 && part == 0) atomicAdd(a.fault, 1);
                cH = cH < 0 ? 0 : (cH > kCoarseMax ? kCoarseMax : cH);
                cW = cW < 0 ? 0 : (cW > kCoarseMax ? kCoarseMax : cW);
            }
            const double* src[7] = {a.x_offt, a.y_offt, a.m, a.q, a.px, a.py, a.pz};
            for (int k = threadIdx.x; k < cH * cW; k += blockDim.x) {
                const size_t g = (size_t)(cI0 + k / cW) * a.nc + cJ0 + k % cW;
                cpos[k] = make_double2(src[0][g], src[1][g]);
                cmq[k].x = (A)src[2][g];
                cmq[k].y = (A)src[3][g];
                cpxy[k].x = (A)src[4][g];
                cpxy[k].y = (A)src[5][g];
                cpz[k] = (A)src[6][g];
            }
            // every thread has read the rectangle's corners: turn the
            // neighbour indices into offsets in it, rows scaled by its
            // width, each clamped into it, so that a row part plus a column
            // part is an in-bounds index of cs
            __syncthreads();
            const int hmax = max(cH - 1, 0), wmax = max(cW - 1, 0);
            for (int k = threadIdx.x; k < ni + nj; k += blockDim.x) {
                if (k < ni) {
                    ti.idx[k].x = min(max(ti.idx[k].x - cI0, 0), hmax) * cW;
                    ti.idx[k].y = min(max(ti.idx[k].y - cI0, 0), hmax) * cW;
                } else {
                    const int s = k - ni;
                    tj.idx[s].x = min(max(tj.idx[s].x - cJ0, 0), wmax);
                    tj.idx[s].y = min(max(tj.idx[s].y - cJ0, 0), wmax);
                }
            }
        }
        __syncthreads();
        const A sm = (A)a.smallness;
        // coarse->fine interpolation of tile-local fine particle (li, lj)
        // (lcode.py:591-621): its position, and apart from it (only for the
        // particles whose stencil reaches the tile) its deposition values
        // (lcode.py:651-655)
        auto position = [&](int li, int lj, double& x, double& y) {
            const double2 wi = ti.w[li], wj = tj.w[lj];
            const int4 xi = ti.idx[li], xj = tj.idx[lj];
            const double wA = wi.x * wj.x, wB = wi.x * wj.y;
            const double wC = wi.y * wj.x, wD = wi.y * wj.y;
            if constexpr (STAGED) {
                // row offset + column offset (both clamped into the staged
                // rectangle above)
                const int pp = xi.x + xj.x, pn = xi.x + xj.y;
                const int np = xi.y + xj.x, nn = xi.y + xj.y;
                const double2 cA = cpos[pp], cB = cpos[pn], cC = cpos[np], cD = cpos[nn];
#define LCODE_MIXP(c) (wA * cA.c + wB * cB.c + wC * cC.c + wD * cD.c)
                x = ti.fg[li] + LCODE_MIXP(x);
                y = tj.fg[lj] + LCODE_MIXP(y);
#undef LCODE_MIXP
            } else {
                const int nc = a.nc;
                const int pp = xi.x * nc + xj.x, pn = xi.x * nc + xj.y;
                const int np = xi.y * nc + xj.x, nn = xi.y * nc + xj.y;
#define LCODE_MIXP(arr) (wA * arr[pp] + wB * arr[pn] + wC * arr[np] + wD * arr[nn])
                x = ti.fg[li] + LCODE_MIXP(a.x_offt);
                y = tj.fg[lj] + LCODE_MIXP(a.y_offt);
#undef LCODE_MIXP
            }
        };
        auto values = [&](int li, int lj, A* v) {
            const double2 wi = ti.w[li], wj = tj.w[lj];
            const int4 xi = ti.idx[li], xj = tj.idx[lj];
            const A Ag = (A)(wi.x * wj.x), Bg = (A)(wi.x * wj.y);
            const A Cg = (A)(wi.y * wj.x), Dg = (A)(wi.y * wj.y);
            if constexpr (STAGED) {
                const int pp = xi.x + xj.x, pn = xi.x + xj.y;
                const int np = xi.y + xj.x, nn = xi.y + xj.y;
                const A2 mA = cmq[pp], mB = cmq[pn], mC = cmq[np], mD = cmq[nn];
                const A2 pA = cpxy[pp], pB = cpxy[pn], pC = cpxy[np], pD = cpxy[nn];
#define LCODE_MIX(a, b, c, d) (Ag * (a) + Bg * (b) + Cg * (c) + Dg * (d))
                deposit_values<A>(sm * LCODE_MIX(mA.x, mB.x, mC.x, mD.x),
                                  sm * LCODE_MIX(mA.y, mB.y, mC.y, mD.y),
                                  sm * LCODE_MIX(pA.x, pB.x, pC.x, pD.x),
                                  sm * LCODE_MIX(pA.y, pB.y, pC.y, pD.y),
                                  sm * LCODE_MIX(cpz[pp], cpz[pn], cpz[np], cpz[nn]), v);
#undef LCODE_MIX
            } else {
                const int nc = a.nc;
                const int pp = xi.x * nc + xj.x, pn = xi.x * nc + xj.y;
                const int np = xi.y * nc + xj.x, nn = xi.y * nc + xj.y;
#define LCODE_MIX(arr) (Ag * (A)arr[pp] + Bg * (A)arr[pn] + Cg * (A)arr[np] + Dg * (A)arr[nn])
                deposit_values<A>(sm * LCODE_MIX(a.m), sm * LCODE_MIX(a.q),
                                  sm * LCODE_MIX(a.px), sm * LCODE_MIX(a.py),
                                  sm * LCODE_MIX(a.pz), v);
#undef LCODE_MIX
            }
        };
        // does the 3 x 3 stencil around centre (i, j) reach this tile?
        auto reaches = [&](int i, int j) {
            return i + 1 >= ci0 && i - 1 < ci0 + kTile && j + 1 >= cj0 && j - 1 < cj0 + kTile;
        };
        // TSC centre and weights of the particle; false when it moved more
        // than the halo from its initial cell (its home tile lists it)
        auto locate = [&](int li, int lj, double x, double y, TscT<A>& t) {
            const int ci = ti.idx[li].z, cj = tj.idx[lj].z;   // initial cell
            const bool home = ci >= ci0 && ci < ci0 + kTile &&
                              cj >= cj0 && cj < cj0 + kTile;
            if (!(fabs(x) <= kFarCells * a.h && fabs(y) <= kFarCells * a.h)) {
                // non-finite or absurdly far particle (a deposition called
                // without the pushes' bound): its home tile becomes NaN
                if (home) {
                    tile_nan = 1;
                    if (a.host_flag)
                        __hip_atomic_store(a.host_flag, 1, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_SYSTEM);
                }
                return false;
            }
            tsc_weights<A>(x, y, N, a.h, t, home ? a.fault : nullptr, inv_h);
            if (abs(t.i - ci) > halo || abs(t.j - cj) > halo) {
                // (beyond a bound the pushes reported: cannot happen; counted
                // as a fault rather than listed for a pass that does not run)
                if (home) {
                    if (may_spill) {
                        a.spill[atomicAdd(a.spill_count, 1)] = (fi_a + li) * a.nf + fj_a + lj;
                        if (a.host_spill)
                            __hip_atomic_store(a.host_spill, 1, __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_SYSTEM);
                    } else {
                        atomicAdd(a.fault, 1);
                    }
                }
                return false;
            }
            return true;
        };
#if LCODE_DEP_GROUP
        // One thread per group of gs x gs fine particles with consecutive
        // indices (the tile's fine range starts at a cell edge, so with gs =
        // fineness a group is the particles of one initial cell).  The
        // thread sums the 3 x 3 x 4 stencil values of successive particles
        // with a common stencil centre in registers and adds them to LDS
        // when the centre changes and after the group's last particle.
        // Adjacent lanes hold adjacent cells, so the atomics of one
        // instruction hit distinct LDS addresses.
        const int gs = a.group > 0 ? a.group : 1;
        const int ngi = (ni + gs - 1) / gs, ngj = (nj + gs - 1) / gs;
        // With LCODE_DEP_SPREAD the lanes of a wave take groups kSpread
        // apart (g = L * stride mod n, the stride a prime that does not
        // divide n, so the map is a permutation): where the wake compresses
        // the plasma, adjacent groups' particles share cells, and adjacent
        // lanes would then add to the same LDS addresses in one instruction.
        const int ng = ngi * ngj;
        int stride = 1;
        if (LCODE_DEP_SPREAD) {
            constexpr int kPrimes[6] = {13, 17, 19, 23, 29, 31};
            stride = kPrimes[5];
#pragma unroll
            for (int u = 5; u >= 0; --u)
                if (ng % kPrimes[u] != 0) stride = kPrimes[u];
        }
        for (int L = threadIdx.x + part * blockDim.x; L < ng; L += nparts * blockDim.x) {
            const int g = LCODE_DEP_SPREAD ? (int)((long long)L * stride % ng) : L;
            const int gi = g / ngj, gj = g - gi * ngj;
            A pat[4][9];
#pragma unroll
            for (int c = 0; c < 4; ++c)
#pragma unroll
                for (int k = 0; k < 9; ++k) pat[c][k] = 0;
            int ai = -1, aj = -1;   // centre of the sums in pat (-1: none)
            int qi = 0, qj = 0;     // particle (qi, qj) of the group
            for (int q = 0; q <= gs * gs; ++q) {
                const int li = gi * gs + qi, lj = gj * gs + qj;
                if (++qj == gs) {
                    qj = 0;
                    ++qi;
                }
                A v[4] = {0, 0, 0, 0};
                TscT<A> t;
                t.i = t.j = -1;
                if (q < gs * gs && li < ni && lj < nj) {
                    double x, y;
                    position(li, lj, x, y);
                    if (!locate(li, lj, x, y, t) || !reaches(t.i, t.j)) t.i = -1;
                    else values(li, lj, v);
                }
                if (ai >= 0 && (t.i != ai || t.j != aj)) {
                    if (q == gs * gs || t.i >= 0) {
                        lds_deposit<A>(acc, ai, aj, ci0, cj0, nullptr, nullptr, pat);
#pragma unroll
                        for (int c = 0; c < 4; ++c)
#pragma unroll
                            for (int k = 0; k < 9; ++k) pat[c][k] = 0;
                        ai = -1;
                    }
                }
                if (t.i >= 0) {
                    ai = t.i;
                    aj = t.j;
#pragma unroll
                    for (int k = 0; k < 9; ++k)
#pragma unroll
                        for (int c = 0; c < 4; ++c) pat[c][k] += v[c] * t.w[k];
                }
            }
        }
#else
        for (int p = threadIdx.x + part * blockDim.x; p < ni * nj; p += nparts * blockDim.x) {
            const int li = p / nj, lj = p - li * nj;
            double x, y;
            A v[4];
            position(li, lj, x, y);
            TscT<A> t;
            if (!locate(li, lj, x, y, t) || !reaches(t.i, t.j)) continue;
            values(li, lj, v);
            lds_deposit<A>(acc, t.i, t.j, ci0, cj0, v, t.w, nullptr);
        }
#endif
    }
    __syncthreads();
    // a split tile: both parts publish their sums (all NaN from a part that
    // met a non-finite particle of the tile, so that the total is NaN); the
    // second to finish adds the other's sums to its own and stores the tile
    bool store = true;
    if (nparts == 2) {
        constexpr int kAcc = 4 * kTile * kTile;
        double* const sums = &acc[0][0][0];
        double* const mine = a.split_part + ((size_t)slot * 2 + part) * kAcc;
        const double* const other = a.split_part + ((size_t)slot * 2 + 1 - part) * kAcc;
        // The sums travel as device-scope atomic stores and loads (written
        // through to, and read from, the device-coherent level), and every
        // wave waits for its stores to complete before the barrier that
        // precedes the count (the barrier alone does not wait for global
        // stores).  A device-scope fence instead writes back and invalidates
        // the whole L2 of the XCD, under every other block there (measured:
        // 40.7 -> 58.6 us per launch with 39 split tiles).
        for (int k = threadIdx.x; k < kAcc; k += blockDim.x)
            __hip_atomic_store(mine + k, tile_nan ? __builtin_nan("") : sums[k],
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __builtin_amdgcn_s_waitcnt(0);   // vmcnt, expcnt and lgkmcnt all zero
        __syncthreads();
        if (threadIdx.x == 0) split_last = atomicAdd(a.split_count + slot, 1) == 1;
        __syncthreads();
        store = split_last;
        if (store) {
            for (int k = threadIdx.x; k < kAcc; k += blockDim.x)
                sums[k] += __hip_atomic_load(other + k, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT);
            if (threadIdx.x == 0) a.split_count[slot] = 0;
            __syncthreads();
        }
    }
    // store the tile, adding the ion background to ro last (lcode.py:687);
    // all background loads are issued before the first store
    constexpr int kOut = (kTile * kTile + kDepBlock - 1) / kDepBlock;
    if (store) {
        double base[kOut];
#pragma unroll
        for (int u = 0; u < kOut; ++u) {
            const int k = threadIdx.x + u * kDepBlock;
            const int r = k / kTile, c = k % kTile;
            const int i = ci0 + r, j = cj0 + c;
            base[u] = 0.;
            if (k < kTile * kTile && i < N && j < N && a.ro_initial)
                base[u] = a.ro_initial[(size_t)i * N + j];
        }
#pragma unroll
        for (int u = 0; u < kOut; ++u) {
            const int k = threadIdx.x + u * kDepBlock;
            const int r = k / kTile, c = k % kTile;
            const int i = ci0 + r, j = cj0 + c;
            if (k >= kTile * kTile || i >= N || j >= N) continue;
            const size_t idx = (size_t)i * N + j;
            G* const ro = gp<G>(a.ro);
            G* const jx = gp<G>(a.jx);
            G* const jy = gp<G>(a.jy);
            G* const jz = gp<G>(a.jz);
            if (overfull || nonfinite || tile_nan) {
                ro[idx] = jx[idx] = jy[idx] = jz[idx] = (G)__builtin_nan("");
                continue;
            }
            // the FP64 ion background is added in FP64, then rounded once
            ro[idx] = (G)(a.ro_initial ? acc[0][r][c] + base[u] : acc[0][r][c]);
            jx[idx] = (G)acc[1][r][c];
            jy[idx] = (G)acc[2][r][c];
            jz[idx] = (G)acc[3][r][c];
        }
    }
    // a particle may have spilled somewhere (the bound every block read
    // exceeds the largest halo) and no spill kernel follows: the last block
    // to finish, after every block's stores are visible, deposits the list
    if (a.spill_in_tiles && global > a.halo && !nonfinite) {
        __threadfence();   // every thread's stores visible at device scope
        __syncthreads();
        if (threadIdx.x == 0)
            last_block = atomicAdd(a.done, 1) == (int)(gridDim.x * gridDim.y) - 1;
        __syncthreads();
        if (last_block) {
            __threadfence();
            const int count = __hip_atomic_load(a.spill_count, __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_AGENT);
            for (int p = threadIdx.x; p < count; p += blockDim.x)
                deposit_spill_one<A, G>(a, a.spill[p], N);
            if (threadIdx.x == 0) *a.done = 0;
        }
    }
}

// Deposits the particles the tile kernel listed as displaced beyond its
// halo, with device-scope atomics over the whole grid (after the tile
// kernel's stores, which it follows on the stream), and zeroes the count
// slot of the next call.  An empty list costs one load per block.
template <typename A, typename G>
__global__ void __launch_bounds__(kBlock)
deposit_spill_kernel(DepositArgs a) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *a.spill_next = 0;
    if (!gate_open(a.gate)) return;
    const int count = *a.spill_count;
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < count;
         p += gridDim.x * blockDim.x)
        deposit_spill_one<A, G>(a, a.spill[p], a.N);
}

template <bool STAGED, typename A, typename G>
__global__ void __launch_bounds__(kDepBlock) LCODE_DEP_ATTR
deposit_tile_kernel(DepositArgs a) {
    deposit_tile_body<STAGED, A, G>(a);
}

__global__ void __launch_bounds__(64) zero_ints_kernel(int* p, int n) {
    for (int k = threadIdx.x; k < n; k += blockDim.x) p[k] = 0;
}

// ---------------------------------------------------------------------------
// Plasma set-up (make_plasma, lcode.py:452-573) on the device, in the
// reference's FP64 arithmetic (integer counters converted exactly)
// ---------------------------------------------------------------------------

// Coarse grid point i of 2 hc - 1: -(hc - 1 - i) s for the left half,
// (i - hc + 1) s for the right one (lcode.py:452-465).
__device__ __forceinline__ double coarse_at(int i, int hc, double s) {
#pragma clang fp contract(off)
    const int k = i - (hc - 1);
    return k < 0 ? -((double)(-k) * s) : (double)k * s;
}

// Fine grid point i (lcode.py:468-487): odd fineness, 2 hf - 1 points on
// k s; even fineness, 2 hf points on (.5 + k) s.
__device__ __forceinline__ double fine_at(int i, int hf, int odd, double s) {
#pragma clang fp contract(off)
    if (odd) {
        const int k = i - (hf - 1);
        return k < 0 ? -((double)(-k) * s) : (double)k * s;
    }
    return i < hf ? -((.5 + (double)(hf - 1 - i)) * s) : (.5 + (double)(i - hf)) * s;
}

__global__ void __launch_bounds__(kBlock)
plasma_coarse_kernel(PlasmaArgs a) {
#pragma clang fp contract(off)
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.nc * a.nc) return;
    const int i = k / a.nc, j = k % a.nc;
    const double s = a.cell * a.coarseness;
    a.x_init[k] = coarse_at(i, a.hc, s);
    a.y_init[k] = coarse_at(j, a.hc, s);
    const double c2 = (double)(a.coarseness * a.coarseness);
    a.m[k] = 1. * kElectronMass * c2;        // ones * ELECTRON_MASS * c**2
    a.q[k] = 1. * kElectronCharge * c2;
}

// virt_params (lcode.py:526-570): searchsorted of each fine point in the
// coarse grid (first index with coarse >= fine), the clipped neighbour
// indices and the interpolation weights with the boundary fixes.
__global__ void __launch_bounds__(kBlock)
plasma_fine_kernel(PlasmaArgs a) {
    // no fused multiply-adds: every operation rounds as NumPy's does
#pragma clang fp contract(off)
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= a.nf) return;
    const double sc = a.cell * a.coarseness;
    const double sf = a.cell / a.fineness;
    const double x = fine_at(f, a.hf, a.fineness % 2, sf);
    int lo = 0, hi = a.nc;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (coarse_at(mid, a.hc, sc) < x) lo = mid + 1; else hi = mid;
    }
    const int nxt = lo < 0 ? 0 : (lo > a.nc - 1 ? a.nc - 1 : lo);
    const int prv = lo - 1 < 0 ? 0 : (lo - 1 > a.nc - 1 ? a.nc - 1 : lo - 1);
    double wp = (coarse_at(nxt, a.hc, sc) - x) / sc;
    double wn = (x - coarse_at(prv, a.hc, sc)) / sc;
    if (nxt == 0) { wp = 0.; wn = 1.; }
    if (prv == a.nc - 1) { wn = 0.; wp = 1.; }
    a.fg[f] = x;
    a.ip[f] = wp;
    a.in[f] = wn;
    a.P[f] = prv;
    a.Nn[f] = nxt;
}

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
    ProfScope ps(prof, "push", s, true);
    const dim3 blocks((a.n + kPushBlock - 1) / kPushBlock);
    if (a.grid32)
        launch_timed(ps, push_kernel<float>, blocks, dim3(kPushBlock), 0, s, a);
    else
        launch_timed(ps, push_kernel<double>, blocks, dim3(kPushBlock), 0, s, a);
    LCODE_HIP(hipGetLastError());
    return LCODE_OK;
}

int launch_make_plasma(const PlasmaArgs& a, hipStream_t s) {
    if (a.nc > 0) {
        plasma_coarse_kernel<<<blocks_for((long)a.nc * a.nc), kBlock, 0, s>>>(a);
        LCODE_HIP(hipGetLastError());
    }
    if (a.nf > 0) {
        plasma_fine_kernel<<<blocks_for(a.nf), kBlock, 0, s>>>(a);
        LCODE_HIP(hipGetLastError());
    }
    return LCODE_OK;
}

int launch_zero_ints(int* p, int n, hipStream_t s) {
    if (n <= 0) return LCODE_OK;
    zero_ints_kernel<<<1, 64, 0, s>>>(p, n);
    LCODE_HIP(hipGetLastError());
    return LCODE_OK;
}

int deposit_blocks_per_cu(bool staged) {
    int a = 0, b = 0;
    hipError_t e1, e2;
    if (staged) {
        e1 = hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &a, deposit_tile_kernel<true, double, double>, kDepBlock, 0);
        e2 = hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &b, deposit_tile_kernel<true, double, float>, kDepBlock, 0);
    } else {
        e1 = hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &a, deposit_tile_kernel<false, double, double>, kDepBlock, 0);
        e2 = hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &b, deposit_tile_kernel<false, double, float>, kDepBlock, 0);
    }
    if (e1 != hipSuccess || e2 != hipSuccess) {
        (void)hipGetLastError();   // leave no error behind for later checks
        return 0;
    }
    return std::min(a, b);
}

int launch_deposit(const DepositArgs& a, hipStream_t s, Profiler* prof) {
    LCODE_REQUIRE(a.ntx == (a.N + kTile - 1) / kTile, "deposition: tiles per side do not match N");
    LCODE_REQUIRE(a.nsplit >= 0 && a.nsplit <= a.ntx * a.ntx &&
                      (!a.nsplit || (a.split_tile && a.split_slot && a.split_part && a.split_count)),
                  "deposition: inconsistent split tiles");
    const dim3 grid(a.ntx * a.ntx + a.nsplit);
    {
        ProfScope ps(prof, "deposit_tile", s, true);
        const dim3 block(kDepBlock);
        if (a.grid32 && a.staged)
            launch_timed(ps, deposit_tile_kernel<true, double, float>, grid, block, 0, s, a);
        else if (a.grid32)
            launch_timed(ps, deposit_tile_kernel<false, double, float>, grid, block, 0, s, a);
        else if (a.staged)
            launch_timed(ps, deposit_tile_kernel<true, double, double>, grid, block, 0, s, a);
        else
            launch_timed(ps, deposit_tile_kernel<false, double, double>, grid, block, 0, s, a);
    }
    LCODE_HIP(hipGetLastError());
    if (!a.spill_in_tiles) {
        ProfScope ps(prof, "deposit_spill", s, true);
        if (a.grid32)
            launch_timed(ps, deposit_spill_kernel<double, float>, dim3(kSpillBlocks),
                         dim3(kBlock), 0, s, a);
        else
            launch_timed(ps, deposit_spill_kernel<double, double>, dim3(kSpillBlocks),
                         dim3(kBlock), 0, s, a);
    }
    LCODE_HIP(hipGetLastError());
    return LCODE_OK;
}

}  // namespace lcode
