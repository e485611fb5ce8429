// Shared device helpers and internal launcher declarations of the gfx950
// LCODE 3D step library.  Everything here is FP64: positions, momenta,
// grids and transforms.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <type_traits>
#include <cstdint>
#include <cstring>
#include <string>

#include "../../include/lcode3d_hip.h"

namespace lcode {

// ---------------------------------------------------------------------------
// Error reporting
// ---------------------------------------------------------------------------

void set_error(const std::string& msg);

#define LCODE_HIP(expr)                                                     \
    do {                                                                    \
        hipError_t e_ = (expr);                                             \
        if (e_ != hipSuccess) {                                             \
            ::lcode::set_error(std::string(#expr) + ": " +                  \
                               hipGetErrorString(e_));                      \
            return LCODE_E_HIP;                                             \
        }                                                                   \
    } while (0)

#define LCODE_REQUIRE(cond, msg)                                            \
    do {                                                                    \
        if (!(cond)) {                                                      \
            ::lcode::set_error(msg);                                        \
            return LCODE_E_INVALID;                                         \
        }                                                                   \
    } while (0)

// ---------------------------------------------------------------------------
// Kernel timing with HIP events (enabled by lcode_profile_begin)
// ---------------------------------------------------------------------------

struct Profiler {
    bool on = false;
    int capacity = 0;                 // event pairs
    int used = 0;
    hipEvent_t* ev = nullptr;         // 2 * capacity events
    const char** names = nullptr;     // name of pair k
    char only[32] = {0};              // time only this kernel name ("": all)
    bool paused = false;
};

// Brackets one kernel launch; a no-op when timing is off or full, or when
// the launch is not the one kernel being timed.  By default it records an
// event before and after the launch (marker packets: the pair then also
// spans the launch's dispatch, a few us).  With `attached` it records
// nothing itself: the caller launches through hipExtLaunchKernelGGL with
// start() and stop(), and the runtime stamps the two events with the
// kernel's own start and end, as a kernel trace does.
struct ProfScope {
    Profiler* p;
    hipStream_t s;
    int k = -1;
    bool attached = false;
    ProfScope(Profiler* prof, const char* name, hipStream_t stream,
              bool attach = false)
        : p(prof), s(stream), attached(attach) {
        if (p && p->on && !p->paused && p->used < p->capacity &&
            (!p->only[0] || std::strcmp(p->only, name) == 0)) {
            k = p->used++;
            p->names[k] = name;
            if (!attached) (void)hipEventRecord(p->ev[2 * k], s);
        }
    }
    bool timed() const { return k >= 0; }
    hipEvent_t start() const { return p->ev[2 * k]; }
    hipEvent_t stop() const { return p->ev[2 * k + 1]; }
    ~ProfScope() {
        if (k >= 0 && !attached) (void)hipEventRecord(p->ev[2 * k + 1], s);
    }
};

// Launches kernel<<<grid, block, lds, stream>>>(args...), timed with
// events attached to the dispatch when `ps` is timing this launch.
template <typename K, typename... Args>
inline void launch_timed(const ProfScope& ps, K kernel, dim3 grid, dim3 block,
                         size_t lds, hipStream_t stream, Args... args) {
    if (ps.timed())
        hipExtLaunchKernelGGL(kernel, grid, block, (std::uint32_t)lds, stream,
                              ps.start(), ps.stop(), 0u, args...);
    else
        hipLaunchKernelGGL(kernel, grid, block, lds, stream, args...);
}

// ---------------------------------------------------------------------------
// Device-side gate of the guided step's fallback kernels
// ---------------------------------------------------------------------------

// acc = {sum (F - guess)^2, sum F^2} of the guided corrector; a gated kernel
// runs only when the residual exceeds the threshold (or is not a number).
struct Gate {
    const double* acc;        // NULL: always run
    double thr2;              // threshold squared
};

__device__ __forceinline__ bool gate_open(const Gate& g) {
    if (!g.acc) return true;
    const double num = g.acc[0], den = g.acc[1];
    return !(num <= g.thr2 * den);
}

// Deposition tile side in cells and threads per tile block (build options).
// At 641^2, coarseness 3, with grouped particles, 256 threads per block
// (165 VGPRs: three blocks per CU), in the developed wake (slices 420-719,
// HIP events on a sample of the launches, three boxes): 24 cells (729
// tiles, all resident at once) 43.0-43.2 us per launch, 28 cells (529
// tiles) 43.7-47.8 us, 30-36 cells 49-58 us, 16-22 cells 44-51 us (more
// tiles than block slots); 512 threads 54-59 us.  With the central tiles
// split over the free block slots: 24 cells 39.3 us, 25 cells 39.8, 26
// cells (143 split tiles) 42.5, 27 cells 43.3.
#ifndef LCODE_DEP_TILE
#define LCODE_DEP_TILE 24
#endif
#ifndef LCODE_DEP_BLOCK
#define LCODE_DEP_BLOCK 256
#endif
constexpr int kDepTile = LCODE_DEP_TILE;
// LDS sizes of a deposition tile: fine indices per side (at least fineness
// (kDepTile + 2 halo + 2) for fineness <= 3, halo <= 6) and coarse
// particles per side of the staged coarse rectangle (a tile whose rectangle
// is larger takes the unstaged kernel).
#ifndef LCODE_DEP_FINE_MAX
#define LCODE_DEP_FINE_MAX 128
#endif
#ifndef LCODE_DEP_COARSE_MAX
#define LCODE_DEP_COARSE_MAX 18
#endif
constexpr int kDepFineMax = LCODE_DEP_FINE_MAX;
constexpr int kDepCoarseMax = LCODE_DEP_COARSE_MAX;
// 1: each deposition thread takes a group of fine particles (fineness x
// fineness, one initial cell) and sums successive particles with a common
// stencil centre in registers before the LDS atomics (43.8 us per launch at
// 641^2); 0: one thread per fine particle (57.7 us).
#ifndef LCODE_DEP_GROUP
#define LCODE_DEP_GROUP 1
#endif
// 1: the lanes of a deposition wave take groups a prime stride apart
// instead of adjacent ones (build option; at 641^2 39.6 -> 44.7 us per
// launch: adjacent lanes' LDS reads of the staged tables lose their
// locality)
#ifndef LCODE_DEP_SPREAD
#define LCODE_DEP_SPREAD 0
#endif

// XCD-aware block order.  Workgroups go to the 8 XCDs round-robin by linear
// id and each XCD has its own L2: map linear id `lin` of `total` to a
// logical id such that every XCD gets one contiguous range of logical ids.
__device__ __forceinline__ int xcd_logical(int lin, int total) {
    constexpr int kXcd = 8;
    const int q = total / kXcd, r = total % kXcd;
    const int xcd = lin % kXcd, slot = lin / kXcd;
    return xcd < r ? xcd * (q + 1) + slot : r * (q + 1) + (xcd - r) * q + slot;
}

// ---------------------------------------------------------------------------
// TSC-2D shape (lcode.py:396-418)
// ---------------------------------------------------------------------------

// The 9 stencil points in the reference's weight order (lcode.py:427-429):
// (-1,+1) (0,+1) (+1,+1) (-1,0) (0,0) (+1,0) (-1,-1) (0,-1) (+1,-1).
__device__ __forceinline__ int stencil_di(int k) { return (k % 3) - 1; }
__device__ __forceinline__ int stencil_dj(int k) { return 1 - (k / 3); }

template <typename W>
struct TscT {
    int i, j;      // centre cell, clamped to [1, N-2]
    W w[9];        // wMP w0P wPP wM0 w00 wP0 wMM w0M wPM
};
using Tsc = TscT<double>;

// Cell index of the coordinate x: floor(x / h + .5) + N / 2.
__device__ __forceinline__ int cell_of(double x, int N, double h) {
    return (int)floor(x / h + .5) + N / 2;
}

// Cell index and quadratic-B-spline weights of the point (x, y).  A centre
// cell whose 3x3 stencil would leave the grid is clamped into it and, when
// fault is not NULL, counted there (the reference has no check: there it
// is an out-of-bounds access).  x_h = x / h + .5 exactly as the reference
// (lcode.py:404), or x * inv_h + .5 when inv_h is given (used by the
// deposition; a 1-ulp change of x_h moves no weight measurably: the TSC
// weights are continuous across cell edges).
// The nine weights from the in-cell offsets (x_loc, y_loc) in [-.5, .5).
template <typename W>
__device__ __forceinline__ void tsc_w(W x_loc, W y_loc, W* w) {
    const W h = W(.5);
    const W wx0 = W(.75) - x_loc * x_loc, wy0 = W(.75) - y_loc * y_loc;
    const W wxP = (h + x_loc) * (h + x_loc) / 2;
    const W wyP = (h + y_loc) * (h + y_loc) / 2;
    const W wxM = (h - x_loc) * (h - x_loc) / 2;
    const W wyM = (h - y_loc) * (h - y_loc) / 2;
    w[0] = wxM * wyP; w[1] = wx0 * wyP; w[2] = wxP * wyP;
    w[3] = wxM * wy0; w[4] = wx0 * wy0; w[5] = wxP * wy0;
    w[6] = wxM * wyM; w[7] = wx0 * wyM; w[8] = wxP * wyM;
}

// The cell and the in-cell offsets are found in FP64 in every precision
// (W = float only rounds the offsets before the weights are formed).
template <typename W>
__device__ __forceinline__ void tsc_weights(double x, double y, int N,
                                            double h, TscT<W>& t, int* fault,
                                            double inv_h = 0.,
                                            double* xl = nullptr,
                                            double* yl = nullptr) {
    const double x_h = inv_h != 0. ? x * inv_h + .5 : x / h + .5;
    const double y_h = inv_h != 0. ? y * inv_h + .5 : y / h + .5;
    const double fx = floor(x_h), fy = floor(y_h);
    int i = (int)fx + N / 2, j = (int)fy + N / 2;
    const double x_loc = x_h - fx - .5, y_loc = y_h - fy - .5;
    if (!(i >= 1 && i <= N - 2 && j >= 1 && j <= N - 2)) {
        if (fault) atomicAdd(fault, 1);
        i = i < 1 ? 1 : (i > N - 2 ? N - 2 : i);
        j = j < 1 ? 1 : (j > N - 2 ? N - 2 : j);
    }
    t.i = i;
    t.j = j;
    tsc_w<W>((W)x_loc, (W)y_loc, t.w);
    if (xl) *xl = x_loc;
    if (yl) *yl = y_loc;
}

// 9-point gather in the reference's summation order (lcode.py:421-430).
// G is the grid's storage type (float grids: each value converts exactly
// to double; the sum is FP64 in both).
template <typename G>
__device__ __forceinline__ double interp9(const G* __restrict__ a, int N,
                                          const Tsc& t) {
    const G* c = a + (size_t)t.i * N + t.j;
    return (double)c[-N + 1] * t.w[0] + (double)c[1] * t.w[1] +
           (double)c[N + 1] * t.w[2] + (double)c[-N] * t.w[3] +
           (double)c[0] * t.w[4] + (double)c[N] * t.w[5] +
           (double)c[-N - 1] * t.w[6] + (double)c[-1] * t.w[7] +
           (double)c[N - 1] * t.w[8];
}

// A grid pointer of the argument structs (declared double) as the grid's
// storage type: the FP32-grid form of the step passes float grids through
// the same members.
template <typename G>
__device__ __forceinline__ G* gp(double* p) { return reinterpret_cast<G*>(p); }
template <typename G>
__device__ __forceinline__ const G* gp(const double* p) {
    return reinterpret_cast<const G*>(p);
}

// ---------------------------------------------------------------------------
// Internal launchers (particles.hip, fields.hip)
// ---------------------------------------------------------------------------

struct PushArgs {
    int n;                 // particles
    int N;                 // grid_steps
    double h, dxi, reflect;
    const double *m, *q, *x_init, *y_init, *prev_x_offt, *prev_y_offt;
    const double *est_x, *est_y;   // NULL: estimate inline (lcode.py:369)
    const double *px, *py, *pz;
    const double *F[6];            // Ex Ey Ez Bx By Bz (averaged)
    double *out_x, *out_y, *out_px, *out_py, *out_pz;
    int* fault;
    int* max_cells;                // atomic max of ceil(|offset| / h), or NULL
    // the same bound per deposition tile of the particles' initial cells
    // ([region_n][region_n] of region_t-cell squares), or NULL
    int* region_max;
    int region_n, region_t;
    // non-finite (or beyond kFarCells) particle state: flagged in
    // host-visible memory (NULL: not reported)
    int* host_flag;
    int grid32;                    // the six field grids are float arrays
    Gate gate;
};

// Displacement bound (in cells) that marks a non-finite or absurd particle
// state; the deposition fills its tiles with NaN instead of depositing.
constexpr int kNonFinite = 1 << 30;
constexpr double kFarCells = 1e6;

struct EstimateArgs {
    int n;
    double dxi, reflect;
    const double *m, *x_init, *y_init, *prev_x_offt, *prev_y_offt;
    const double *px, *py, *pz;
    double *out_x, *out_y;
};

struct DepositArgs {
    int nf, nc, N;
    double h, smallness;
    const double *x_offt, *y_offt, *m, *q, *px, *py, *pz;
    const double *ip, *in, *fg;
    const int64_t *P, *Nn;
    const double* ro_initial;      // may be NULL
    double *ro, *jx, *jy, *jz;
    int* fault;
    int* host_flag;                // host-visible flag of a non-finite
                                   // particle met (or NULL)
    // particles displaced by more than the tile halo: appended to spill[]
    // by their home tile, deposited with atomics by a second, grid-wide
    // kernel; the two count slots are used in turn
    int* spill;                    // nf * nf capacity
    int* spill_count;              // this call's count (zero on entry)
    int* spill_next;               // the next call's count: zeroed here
    int halo;                      // tile halo in cells (maximum)
    int* halo_cells;               // NULL, or the pushes' displacement bound
    int* halo_next;                // NULL, or the bound slot of the next
                                   // pushes: set to 0 by the deposition
    // NULL, or the pushes' bound per tile ([ntx][ntx]): a tile's
    // halo is the largest bound of its 3 x 3 tile neighbourhood; and the
    // next pushes' per-tile slots, each zeroed by its tile
    const int* halo_regions;
    int* halo_regions_next;
    // NULL, or the tile each of the first ntx * ntx blocks takes (a
    // permutation of the row-major tile indices)
    const int* tile_order;
    int ntx;                       // tiles per side; the tile kernel runs
                                   // ntx * ntx + nsplit one-dimensional blocks
    // split tiles (nsplit 0: none).  Block ntx * ntx + s is the second part
    // of tile split_tile[s], and split_slot[tile] is s for that tile (-1 for
    // the others).  The two parts of a split tile each deposit half of its
    // particle groups and write their sums to split_part (2 x 4 x tile^2
    // doubles per slot); the second to finish, counted in split_count[s]
    // (zero between calls: that block resets it), adds the other's sums to
    // its own and stores the tile.
    int nsplit;
    const int* split_tile;
    const int* split_slot;
    double* split_part;
    int* split_count;
    // 1: no spill kernel follows; when a particle may spill (the pushes'
    // bound exceeds a.halo) every tile block ends with a fence and a count
    // in *done, and the last one deposits the spill list and resets *done
    int spill_in_tiles;
    int* done;
    // host-visible flag set when a tile lists a particle (or NULL)
    int* host_spill;
    int staged;                    // coarse rectangle of a tile fits LDS
    int group;                     // fine particles per group side (fineness)
    int grid32;                    // ro jx jy jz are float arrays (the
                                   // arithmetic and ro_initial stay FP64)
    Gate gate;
};

int launch_estimate(const EstimateArgs& a, hipStream_t s, Profiler* prof);
int launch_push(const PushArgs& a, hipStream_t s, Profiler* prof);
int launch_deposit(const DepositArgs& a, hipStream_t s, Profiler* prof);
// Tile blocks of the deposition one CU holds at once (the smaller of its
// FP64- and FP32-grid forms, staged or not), or 0 when the runtime cannot
// say.
int deposit_blocks_per_cu(bool staged);
// make_plasma (lcode.py:452-573) on the device.
constexpr double kElectronMass = 1.;      // lcode.py:47
constexpr double kElectronCharge = -1.;   // lcode.py:46
struct PlasmaArgs {
    int coarseness, fineness;
    int hc, hf;                    // half-grid counts of the two grids
    int nc, nf;                    // coarse and fine points per axis
    double cell;                   // cell size
    double *x_init, *y_init, *m, *q;          // nc * nc
    double *ip, *in, *fg;                     // nf each
    int64_t *P, *Nn;                          // nf each
};
int launch_make_plasma(const PlasmaArgs& a, hipStream_t s);

// Sets n ints at p to 0 with a one-block kernel (no memset node).
int launch_zero_ints(int* p, int n, hipStream_t s);

// Field solves ---------------------------------------------------------------

enum FieldKind { F_EX = 0, F_EY = 1, F_EZ = 2, F_BX = 3, F_BY = 4, F_BZ = 5 };
enum Xform { XF_DCT = 0, XF_DST = 1 };

constexpr int kMaxStages = 24;

struct FftPlan {
    int L;                       // 2 (N - 1)
    int nstages;
    int radix[kMaxStages];
};

// Sources of the right-hand sides (lcode.py:166-168, :258-268, :339-343).
struct RhsInputs {
    const double *ro, *jx, *jy, *jz;      // this pass's densities
    const double *ro_prev, *jz_prev;      // non-NULL: Variant A averaging
    const double *jx_prev, *jy_prev;      // previous slice (dj/dxi)
    const double *Ftrick[4];              // Ex Ey Bx By in the trick term
    const double* beam;                   // NULL: beam_scalar everywhere
    double beam_scalar;
    double trick, h, dxi;
    // FP32-grid form: an FP64 grid of N * N zeros read in place of a
    // missing beam array (the grid members above then point to floats)
    const double* zero64;
};

// Right-hand sides (lcode.py:166-168, :234-268, :339-343), FP64.

// Without averaging a source is averaged with itself, (r + r) / 2 == r
// exactly, and without a beam array the beam term is x * 0 + beam_scalar:
// the same values as the conditional forms, with no branch between loads.
// G: storage type of the grid members (float in the FP32-grid form: the
// values convert exactly and the arithmetic stays FP64).  The beam is FP64.
template <typename G = double>
__device__ __forceinline__ double beam_at(const RhsInputs& in, size_t idx) {
    const double* dummy = std::is_same<G, double>::value ? in.ro : in.zero64;
    const double* b = in.beam ? in.beam : dummy;
    return b[idx] * (in.beam ? 1. : 0.) + (in.beam ? 0. : in.beam_scalar);
}

template <typename G = double>
__device__ __forceinline__ double rho_src(const RhsInputs& in, size_t idx) {
    const G* ro = gp<G>(in.ro);
    const G* prev = in.ro_prev ? gp<G>(in.ro_prev) : ro;
    const double r = ((double)ro[idx] + (double)prev[idx]) / 2;   // lcode.py:862
    return r + beam_at<G>(in, idx);
}

template <typename G = double>
__device__ __forceinline__ double jz_src(const RhsInputs& in, size_t idx) {
    const G* jz = gp<G>(in.jz);
    const G* prev = in.jz_prev ? gp<G>(in.jz_prev) : jz;
    const double r = ((double)jz[idx] + (double)prev[idx]) / 2;   // lcode.py:863
    return r + beam_at<G>(in, idx);
}

// The right-hand side of field `kind` at (i, j).  Outside the interior the
// derivative terms are zero; their neighbour offsets then collapse to the
// point itself, so every load stays in bounds for any (i, j) of the grid.
template <int KIND, typename G = double>
__device__ __forceinline__ double rhs_value_k(int i, int j, int N,
                                              const RhsInputs& in) {
    const size_t idx = (size_t)i * N + j;
    const bool inner = i >= 1 && i <= N - 2 && j >= 1 && j <= N - 2;
    const size_t oi = inner ? (size_t)N : 0, oj = inner ? 1 : 0;
    const double h2 = in.h * 2;
    const G *jx = gp<G>(in.jx), *jy = gp<G>(in.jy);
    const G *jxp = gp<G>(in.jx_prev), *jyp = gp<G>(in.jy_prev);
    if constexpr (KIND == F_EX) {
        const double d = (rho_src<G>(in, idx + oi) - rho_src<G>(in, idx - oi)) / h2;
        const double dj = ((double)jxp[idx] - (double)jx[idx]) / in.dxi;
        return -(((inner ? d : 0.) - dj) - (double)gp<G>(in.Ftrick[0])[idx] * in.trick);
    } else if constexpr (KIND == F_EY) {
        const double d = (rho_src<G>(in, idx + oj) - rho_src<G>(in, idx - oj)) / h2;
        const double dj = ((double)jyp[idx] - (double)jy[idx]) / in.dxi;
        return -(((inner ? d : 0.) - dj) - (double)gp<G>(in.Ftrick[1])[idx] * in.trick);
    } else if constexpr (KIND == F_BX) {
        const double d = (jz_src<G>(in, idx + oj) - jz_src<G>(in, idx - oj)) / h2;
        const double dj = ((double)jyp[idx] - (double)jy[idx]) / in.dxi;
        return +(((inner ? d : 0.) - dj) + (double)gp<G>(in.Ftrick[2])[idx] * in.trick);
    } else if constexpr (KIND == F_BY) {
        const double d = (jz_src<G>(in, idx + oi) - jz_src<G>(in, idx - oi)) / h2;
        const double dj = ((double)jxp[idx] - (double)jx[idx]) / in.dxi;
        return -(((inner ? d : 0.) - dj) - (double)gp<G>(in.Ftrick[3])[idx] * in.trick);
    } else if constexpr (KIND == F_EZ) {
        // the interior only carries this right-hand side (zero elsewhere)
        return -(((double)jx[idx + oi] - (double)jx[idx - oi]) +
                 ((double)jy[idx + oj] - (double)jy[idx - oj])) / h2;
    } else {  // F_BZ
        const double djx_dy = (double)jx[idx + oj] - (double)jx[idx - oj];
        const double djy_dx = (double)jy[idx + oi] - (double)jy[idx - oi];
        return -((inner ? djx_dy : 0.) - (inner ? djy_dx : 0.)) / h2;
    }
}

template <typename G = double>
__device__ __forceinline__ double rhs_value_dev(int kind, int i, int j, int N,
                                                const RhsInputs& in) {
    switch (kind) {
        case F_EX: return rhs_value_k<F_EX, G>(i, j, N, in);
        case F_EY: return rhs_value_k<F_EY, G>(i, j, N, in);
        case F_BX: return rhs_value_k<F_BX, G>(i, j, N, in);
        case F_BY: return rhs_value_k<F_BY, G>(i, j, N, in);
        case F_EZ: return rhs_value_k<F_EZ, G>(i, j, N, in);
        default: return rhs_value_k<F_BZ, G>(i, j, N, in);
    }
}

struct FieldPlan {
    int kind;                 // FieldKind
    int rowsA0, rowsA;        // rows i of the RHS transformed along j
    int xfA, lenA, colA0;     // transform along j: type, length, first j
    int xfB;                  // transform along i (length rowsA)
    const void* table;        // [lenA][rowsA] spectral multiplier (S)
    // the same multiplier as 1 / (lamL[line] + lamK[k] + lam_add) / lam_norm
    // (0 at line = k = 0 when zero_mode): the eigenvalues of the two
    // transforms, the trick term and the normalisation
    const double *lamL, *lamK;
    double lam_add, lam_norm;
    int zero_mode;
    void* T;                  // scratch spectra (S): [lenA][rowsA]; in the wave
                              // FFT [lenA][N + 1], indexed by grid row
    const double* colsum;     // Bz only: column sums of the last DCT
    double* partial;          // Bz only: [lenA] mean partials
    const double* prevF;      // Variant A / averaging source (or NULL)
    double* outF;             // solved field (or NULL)
    double* outAvg;           // (F + prevF) / 2 (or NULL)
};

struct SolveBatch {
    int nfields;
    FieldPlan f[6];
    RhsInputs in;
    int variantA;             // 2 F - prev for Ex Ey Bx By
    Gate gate;                // the whole solve is skipped when closed
    const double* guess[6];   // guided step: the caller's field guess ...
    double* resid;            // ... and {sum (F - guess)^2, sum F^2}
    double* resid_part;       // FFT solver: 2 partial sums per last-pass block
    int grid32;               // every N x N grid above (sources, prev, out,
                              // averages, guess) is a float array
};

// guess[f] = sum_t c[t] h[t][f]; avg[f] = (guess[f] + prev[f]) / 2 over n
// values of the six fields; acc_zero[0..1] = 0.
struct GuessArgs {
    const double* h[4][6];
    double c[4];
    int nt;
    const double* prev[6];
    double* guess[6];
    double* avg[6];
    double* acc_zero;
    int n;
    int grid32;                // the field arrays are float arrays
};

int launch_guess(const GuessArgs& p, hipStream_t s, Profiler* prof);

// Dense (MFMA) form of the field solves -------------------------------------

enum DensePrec { DENSE_F64 = 0, DENSE_BF16X3 = 1 };


struct GemmOp {                // C = A B over an M x N x K iteration space
    const double* A; int lda, a_off, a_rows;   // A(r, k) = A[(r - a_off) lda + k]
    const double* B; int ldb, b_off, b_cols;   // B(k, c) = B[k ldb + c - b_off]
    double* C; int ldc;
    int M, N, K;
    int mode;                  // 0 store, 1 x table (+ Bz mean sum), 2 final
    const double* tab; int ldt;
    const double* colsum;      // mode 1, Bz: column sums of the DCT-I
    double* mean_acc;
    const double* mean_in;     // mode 2, Bz: the sum to remove (/ N^2)
    int variantA;
    const double* prevF;
    double *outF, *outAvg;
    const double* guess;       // mode 2, guided step: residual against it
};

struct GemmBatch {
    int n;
    GemmOp op[6];
    Gate gate;
    double* resid;             // {sum (F - guess)^2, sum F^2}
};

struct DenseRhsArgs {
    FieldPlan f[6];
    RhsInputs in;
    int N;
    double* R[6];
    double* mean_zero;
    Gate gate;
};

struct DenseGeom {
    int prec;                  // DensePrec
    const double* Sd;          // DST-I, (N-2) x (N-2), symmetric
    const double* Cd;          // DCT-I, N x N: Cd[k][n] = w_n cos(pi k n / (N-1))
    const double* CdT;         // its transpose
    const double* colsum;      // column sums of Cd
    const double *mixed, *mixed_t;   // eigenvalue tables [M][N] and [N][M]
    double *X, *Y;             // 6 N^2 scratch each
    double* mean;              // Bz sum
    // DENSE_BF16X3: BF16 hi/lo planes, [Np][Np] each, zero padded
    int Np;                    // N rounded up to a multiple of 128
    const uint16_t *Sh, *Sl;   // DST-I matrix
    const uint16_t *Ch, *Cl;   // DCT-I matrix
    uint16_t *Xh, *Xl;         // 6 plane pairs (one per field kind): RHS, then P, by rows
    uint16_t *Ph, *Pl;         // 6 plane pairs: Q, by columns
};

int launch_gemm(const GemmBatch& b, hipStream_t s, Profiler* prof,
                const char* name);
int launch_dense_rhs(const DenseRhsArgs& a, int nfields, int max_total,
                     hipStream_t s, Profiler* prof);
int launch_solve_dense(int N, const DenseGeom& d, const SolveBatch& b,
                       hipStream_t s, Profiler* prof);

// zn diagnostic: Gaussian taps w[0..2r], scratch grid tmp, sum scratch.
int launch_diag_zn(const double* ro, const double* w, int r, int N, double* tmp,
                   double* sum, double* out, hipStream_t s, Profiler* prof);

struct SolverGeom {
    int N;
    FftPlan fft;
    const double* costab;     // cos(2 pi r / L), r in [0, L / 4]
    const double2* stage_tw;  // per-stage twiddles of the static plan
    const float* costab32;    // the same tables in FP32
    int fp32;                 // transforms, spectra and tables in FP32
    int static_L;             // L with a compile-time plan, else 0
    int tw_lds;               // static plan: twiddles copied to LDS
    int P, Tp;                // pairs of lines per block, threads per pair
    const double2* wave_tw;   // wave FFT (L = 1280, 2560): W_L^t, t < L; else NULL
};

int launch_solve(const SolverGeom& g, const SolveBatch& b, hipStream_t s,
                 Profiler* prof);
bool solver_static_length(int L);
bool solver_wave_length(int L);   // L has the wave FFT passes
// LCODE_FFT_WAVE_FP32 (1): the FP32 transforms (FP32 and mixed solvers) of
// L = 1280 run the wave FFT too (mixed six-field solve at 641^2 67.5 ->
// 61.9 us); 0 keeps them on the radix-4/5 plan, as at L = 2560 always (there
// the FP32 wave passes spill: the 1281^2 mixed step 871 -> 826 xi-steps/s).
#ifndef LCODE_FFT_WAVE_FP32
#define LCODE_FFT_WAVE_FP32 1
#endif
int solver_padded_len(int L);
int solver_tw_count(int L);   // static plan: twiddle table entries
int solver_threads_per_pair(int L, bool fp32, bool static_plan);   // threads of one line pair
int solver_static_radix(int rem);   // static plan: radix of the next stage
int solver_lds_bytes(const SolverGeom& g);


// Learned predictor (unet.hip) ----------------------------------------------

// 3 x 3 convolution over bfloat16 [H][W][32] tensors: nin input chunks of
// 32 channels, cout_pad = 16, 32 or 64 output channels (cout real).
struct UnetConvArgs {
    const uint16_t* in[2];    // [H][W][32] each (in[1] with nin == 2)
    const uint16_t* w;        // packed [9][cout_pad][32 nin] bfloat16
    const float* bias;        // [cout] or NULL
    int H, W, nin, cout, cout_pad;
    int relu;
    const uint16_t* mask;     // zero where mask <= 0 ([H][W][32]) or NULL
    uint16_t* out[2];         // channels 0..31, 32..63 ([H][W][32] bf16)
    float* out_f32;           // or FP32 [H][W][16] (cout_pad == 16)
    // how input chunk k is read: 0 as is; 1 the 2 x 2 mean of a finer
    // [in_h][in_w][32] tensor; 2 nearest upsampling of a coarser one
    int in_mode[2], in_h[2], in_w[2];
};

// Weight gradient of one convolution: per-block partials of
// sum_p g[p][co] x[p + off][ci] and sum_p g[p][co].
struct UnetWgradArgs {
    const uint16_t* in[2];    // the layer input, nin chunks
    const uint16_t* g;        // output gradient [H][W][32] (cout_pad used)
    int H, W, nin, cout_pad;
    float* partial;           // [tiles][9 cout_pad 32 nin + cout_pad]
};

// Planes of the network from the step's fields (FP64 or FP32 grids).
struct UnetIoArgs {
    const void *f1[6], *f2[6], *f3[6];   // fields of slices n-1, n-2, n-3
    const void* f4[6];                   // slice n-4 (the Variant-A-aware
                                         // base), or NULLs
    const void* next[6];                 // targets: the fields to predict
    const void* ro;                      // plasma density of slice n-1
    const double* beam;                  // beam of slice n, or NULL
    double beam_scalar;
    int f32;                             // the grids are float arrays
    int N, S;                            // grid and level-0 sizes
    float scale[14], tscale[6];
    uint16_t* planes;                    // inputs [S][S][32] bf16
    float* planes_f32;                   // or FP32 [S][S][16] (14 used)
    float* targets;                      // targets [S][S][6]
    const float* y;                      // network output [S][S][16]
    void* guess[6];                      // guess output (grid type)
};

int unet_conv_lds_bytes(int nin, int mt);
int unet_tiles(int H, int W);
int launch_unet_conv(const UnetConvArgs& a, hipStream_t s);
int launch_unet_wgrad(const UnetWgradArgs& a, int cout, int cin, float* grad,
                      hipStream_t s);
int launch_unet_pack(const float* w, int cout, int cin, int cout_pad, int cinp,
                     uint16_t* fwd, uint16_t* bwd, hipStream_t s);
int launch_unet_adam(float* p, const float* g, float* m, float* v, int n,
                     float lr, int t, hipStream_t s);
int launch_unet_pool(const uint16_t* x, int H, int W, uint16_t* y, hipStream_t s);
int launch_unet_up(const uint16_t* x, int h, int w, uint16_t* y, int H, int W,
                   hipStream_t s);
int launch_unet_up_t(const uint16_t* g, int H, int W, const uint16_t* mask,
                     uint16_t* out, int h, int w, hipStream_t s);
int launch_unet_pool_t(const uint16_t* g, int h, int w, const uint16_t* skip,
                       const uint16_t* mask, uint16_t* out, int H, int W,
                       hipStream_t s);
int launch_unet_mse(const float* y, const float* t, int P, uint16_t* g,
                    float* part, float* loss, hipStream_t s);
int launch_unet_io(const UnetIoArgs& a, int which, hipStream_t s);

}  // namespace lcode
