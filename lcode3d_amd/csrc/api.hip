// C ABI of the gfx950 LCODE 3D step: workspace (the eigenvalue tables the
// reference memoizes with cp.memoize, FFT twiddles, scratch), the
// standalone operators and the fused xi step (lcode.py:828-931).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "common.hpp"
#include "tile_order.hpp"

namespace lcode {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

}  // namespace lcode

using namespace lcode;

struct lcode_workspace {
    int device = 0;
    int N = 0, M = 0;
    double h = 0;
    int trick_on = 0;
    int nc = 0, nf = 0;
    SolverGeom geom{};
    double* costab = nullptr;      // quarter-wave cosine table
    double2* stage_tw = nullptr;   // static-plan stage twiddles
    double2* wave_tw = nullptr;    // wave FFT twiddles W_L^t
    float* costab32 = nullptr;
    int fp32 = 0;                  // FP32 transforms (FFT32, MIXED)
    int solver = 0;                // LCODE_SOLVER_*
    int grid32 = 0;                // set for the duration of a mixed step:
                                   // the state grids are float arrays
    double* zero64 = nullptr;      // N * N FP64 zeros (MIXED: the RHS's
                                   // stand-in for a missing beam array)
    DenseGeom dense{};             // LCODE_SOLVER_DENSE*: matrices, scratch
    void* tab_mixed = nullptr;     // [M][N]   (DST index, DCT index)
    void* tab_mixed_t = nullptr;   // [N][M]
    void* tab_dir = nullptr;       // [M][M]
    void* tab_neu = nullptr;       // [N][N]
    double* colsum = nullptr;      // [N]
    double* lam_i = nullptr;       // [M] eigenvalues k = 1 .. N - 2 (DST)
    double* lam_j = nullptr;       // [N] eigenvalues k = 0 .. N - 1 (DCT)
    double lam_add = 0, lam_norm = 1;
    void* T = nullptr;             // 6 * N * (N + 1) transform scratch (S)
    double* partial = nullptr;     // [N]
    double* p1[5] = {};            // predictor particles (nc * nc)
    double* p2[5] = {};            // corrector particles
    double* dep[4] = {};           // ro jx jy jz of passes 1 and 2
    double* favg[6] = {};          // field averages of the current pass
    int* fault = nullptr;
    int* spill = nullptr;          // nf * nf spilled fine particle indices
    int* spill_ctr = nullptr;      // [0], [1] spill counts, [2], [3] the
                                   // pushes' displacement bound in cells,
                                   // each pair of slots used in turn; [4]
                                   // tile blocks done (the in-kernel spill)
    int* regions = nullptr;        // the same bound per deposition tile:
                                   // two slots of region_n^2 ints
    int region_n = 0;              // tiles per side (0: no per-tile bounds)
    int* tile_order = nullptr;     // deposition tiles, centre first
    // deposition tiles taken by two blocks: split_tile[s] for s < nsplit,
    // then the slot of each tile (-1: not split); the two parts' sums
    // (nsplit x 2 x 4 x tile^2 doubles) and a count per slot
    int nsplit = 0;
    int* split_tile = nullptr;
    double* split_part = nullptr;
    int* split_count = nullptr;
    int bound_slot = 0;            // slot the next push raises
    int spill_slot = 0;            // count slot of the next deposition
    int grid_spill_left = 0;       // depositions still followed by the spill
                                   // kernel after spills were last seen
    int* host_flag = nullptr;      // pinned host ints: [0] a non-finite particle
                                   // was met, [1] a deposition listed a
                                   // particle beyond its halo (written by
                                   // the kernels)
    double* guess[6] = {};         // guided step: the field guess (after a
                                   // step: its guess, or F1 when rejected,
                                   // or a plain step's F1 with keep_f1)
    int keep_f1 = 0;               // plain steps keep F1 in guess[]
    double* resid = nullptr;       // {sum (F - guess)^2, sum F^2}
    double* resid_part = nullptr;  // 2 per block of the last FFT pass
    double* diag_w = nullptr;      // zn diagnostic: Gaussian taps [2 r + 1]
    int diag_r = 0;
    double* diag_tmp = nullptr;    // [N * N]
    double* diag_sum = nullptr;
    Profiler prof;
    std::vector<void*> allocs;
};

namespace {

#ifndef LCODE_TW_LDS_OK
#define LCODE_TW_LDS_OK 1
#endif
// Line pairs per FFT block (at most).  One pair per block measured best at
// 641^2 (six-field solve FP64 79.6 -> 74.2 us, FP32 66.1 -> 61.3 us against
// up to four): the same pairs per CU, a quarter or half the waves behind
// each barrier, and a finer-grained last round of blocks.
#ifndef LCODE_FFT_MAX_PAIRS_F64
#define LCODE_FFT_MAX_PAIRS_F64 1
#endif
#ifndef LCODE_FFT_MAX_PAIRS_F32
#define LCODE_FFT_MAX_PAIRS_F32 1
#endif
// Static-plan twiddle roots go to LDS when the block has at least this many
// line pairs (else they are read through the caches).  With one pair per
// block, LDS roots measured 74.0 -> 72.4 us (FP64) and 62.3 -> 60.2 us (FP32).
#ifndef LCODE_TW_LDS_MIN_PAIRS
#define LCODE_TW_LDS_MIN_PAIRS 1
#endif
#ifndef LCODE_SOLVER_MAX_LDS
#define LCODE_SOLVER_MAX_LDS 65536
#endif

// Cells a fine particle may move from its initial cell and still be
// deposited through the LDS tiles (beyond: the spill list, deposited with
// device atomics by a grid-wide kernel).  The adaptive halo is the pushes'
// displacement bound, so small displacements keep small tiles; 6 covers the
// nonlinear 1281^2 case (up to ~4.4 cells) within the LDS tables.
constexpr int kDepositHalo = 6;
// 1: per-tile displacement bounds size each deposition tile's halo; 0: the
// largest bound over all particles sizes every tile's halo (build option)
#ifndef LCODE_DEP_REGIONS
#define LCODE_DEP_REGIONS 1
#endif
constexpr bool kDepRegions = LCODE_DEP_REGIONS;
// Order of the deposition tiles over the blocks (build option).  0:
// row-major.  1: by Chebyshev distance from the grid centre (the wake of an
// axial beam makes those the slowest, so they would start first when there
// are more tiles than block slots).  Measured at 641^2 in the developed wake: no
// change with 24-cell tiles (43.03 us either way: every tile is resident
// at once), slower with 20-cell ones (51.2 against 46.9 us).
// 2: XCD clusters: workgroups go to the 8 XCDs round-robin by linear id and
// each XCD has its own L2, so row-major tiles spread every tile's
// neighbours (which share its staged coarse particles, fine-index tables
// and the 128-byte lines at its edges) over all eight; here the tiles come
// in 3 x 3 clusters (LCODE_DEP_CLUSTER), each cluster on one XCD, the clusters dealt to the
// XCDs round-robin in row-major order (adjacent clusters on different
// XCDs, so the wake's slow central tiles still spread over them).
// Measured at 641^2 (24-cell tiles, developed wake, A/B on one box): order 2
// took the deposition from 41.6 to 41.0 us and its HBM traffic from 31.2 to
// 22.7 MB per launch (reads 16.4 -> 9.0 MB, writes 14.8 -> 13.7 MB); 4 x 4
// and 5 x 5 clusters 41.1-41.3 and 41.5 us at about the same traffic.
// Default 2.
#ifndef LCODE_DEP_ORDER
#define LCODE_DEP_ORDER 2
#endif
constexpr int kDepOrder = LCODE_DEP_ORDER;

// Deposition tiles taken by two blocks each (build option; -1: as many as
// the launch leaves block slots free when all tile blocks are resident at
// once, at the tile blocks per CU the runtime reports for the kernel, or
// kDepResident when it reports none; 0: none).  The tiles nearest the grid centre
// are split: an axisymmetric beam's wake compresses and displaces the plasma
// there, so those tiles have the largest halos and the most particles.
#ifndef LCODE_DEP_SPLIT
#define LCODE_DEP_SPLIT -1
#endif
constexpr int kDepResident = 3;   // tile blocks per CU (158 VGPRs, 4 waves) when
                                  // the runtime reports none

// The deposition stages each tile's coarse rectangle in LDS: it has at
// most ext / coarseness + 3 rows (ext: the tile and its largest halo).
static bool deposit_staged(const lcode_config* cfg) {
    return (kDepTile + 2 * kDepositHalo + 2) / cfg->plasma_coarseness + 3 <= kDepCoarseMax;
}

// Side of the XCD clusters of the tile order (tile_order.hpp).
#ifndef LCODE_DEP_CLUSTER
#define LCODE_DEP_CLUSTER 3
#endif

// 1: each XCD's tiles spread over its CUs in one-round launches
// (balance_xcd_slots, tile_order.hpp); 2: in every launch
#ifndef LCODE_DEP_BALANCE
#define LCODE_DEP_BALANCE 1
#endif

// Round-to-nearest-even BF16 bits of a float (finite inputs).
uint16_t bf16_bits(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

float bf16_value(uint16_t b) {
    const uint32_t u = (uint32_t)b << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

// hi + lo BF16 split of v: hi = bf16(v), lo = bf16(v - hi).
void split_bf16(double v, uint16_t& hi, uint16_t& lo) {
    hi = bf16_bits((float)v);
    lo = bf16_bits((float)(v - (double)bf16_value(hi)));
}

int dev_alloc(lcode_workspace* ws, void** p, size_t bytes) {
    LCODE_HIP(hipMalloc(p, bytes));
    ws->allocs.push_back(*p);
    return LCODE_OK;
}

template <typename T>
int upload(lcode_workspace* ws, T** dst, const std::vector<T>& src) {
    int rc = dev_alloc(ws, (void**)dst, src.size() * sizeof(T));
    if (rc) return rc;
    LCODE_HIP(hipMemcpy(*dst, src.data(), src.size() * sizeof(T),
                        hipMemcpyHostToDevice));
    return LCODE_OK;
}

// Eigenvalue table in the workspace's transform precision.
int upload_table(lcode_workspace* ws, void** dst, const std::vector<double>& v) {
    if (!ws->fp32) return upload(ws, reinterpret_cast<double**>(dst), v);
    std::vector<float> f(v.begin(), v.end());
    return upload(ws, reinterpret_cast<float**>(dst), f);
}

bool factor_radices(int L, FftPlan& plan) {
    plan.L = L;
    plan.nstages = 0;
    int rest = L;
    const int radices[] = {4, 2, 3, 5};
    for (int r : radices) {
        while (rest % r == 0) {
            if (plan.nstages == kMaxStages) return false;
            plan.radix[plan.nstages++] = r;
            rest /= r;
        }
    }
    return rest == 1;
}

// Second-difference eigenvalues 4/h^2 sin^2(k pi / (2 (N - 1)))
// (lcode.py:152-154, :227-228, :326), in the reference's FP64 arithmetic.
double lam(int k, int N, double h) {
    const double s = std::sin(k * M_PI / (2 * (N - 1)));
    return 4 / (h * h) * (s * s);
}

int check_cfg(const lcode_config* c) {
    LCODE_REQUIRE(c, "config is NULL");
    LCODE_REQUIRE(c->grid_steps >= 5 && c->grid_steps % 2 == 1,
                  "grid_steps must be odd and >= 5");
    LCODE_REQUIRE(c->grid_step_size > 0 && c->xi_step_size > 0,
                  "grid_step_size and xi_step_size must be positive");
    LCODE_REQUIRE(c->nc >= 0 && c->nf >= 0, "negative particle counts");
    LCODE_REQUIRE(c->plasma_coarseness > 0 && c->plasma_fineness > 0,
                  "plasma_coarseness and plasma_fineness must be positive");
    LCODE_REQUIRE(c->plasma_fineness <= 3,
                  "plasma_fineness > 3 is not supported (deposition tile tables)");
    LCODE_REQUIRE(c->plasma_fineness * (kDepTile + 2 * kDepositHalo + 2) <= kDepFineMax,
                  "plasma_fineness too large for this build's deposition tiles");
    LCODE_REQUIRE(c->solver == LCODE_SOLVER_FFT64 || c->solver == LCODE_SOLVER_FFT32 ||
                      c->solver == LCODE_SOLVER_DENSE64 ||
                      c->solver == LCODE_SOLVER_DENSE_BF16X3 ||
                      c->solver == LCODE_SOLVER_MIXED,
                  "unknown solver");
    return LCODE_OK;
}

int check_ws(const lcode_workspace* ws, const lcode_config* c) {
    int rc = check_cfg(c);
    if (rc) return rc;
    LCODE_REQUIRE(ws, "workspace is NULL");
    LCODE_REQUIRE(ws->N == c->grid_steps && ws->h == c->grid_step_size &&
                      ws->trick_on == (c->subtraction_trick != 0),
                  "workspace was built for another grid/trick");
    LCODE_REQUIRE(ws->solver == c->solver,
                  "workspace was built for another solver");
    return LCODE_OK;
}

// Field plans ------------------------------------------------------------

FieldPlan plan_for(lcode_workspace* ws, int kind, int slot) {
    const int N = ws->N, M = ws->M;
    FieldPlan p{};
    p.kind = kind;
    p.T = static_cast<char*>(ws->T) +
          (size_t)slot * N * (N + 1) * (ws->fp32 ? sizeof(float) : sizeof(double));
    switch (kind) {
        case F_EY:
        case F_BX:  // DST along x (rows 1..N-2), DCT along y (lcode.py:275-286)
            p.rowsA0 = 1; p.rowsA = M;
            p.xfA = XF_DCT; p.lenA = N; p.colA0 = 0;
            p.xfB = XF_DST; p.table = ws->tab_mixed_t;
            p.lamL = ws->lam_j; p.lamK = ws->lam_i; p.lam_add = ws->lam_add;
            break;
        case F_EX:
        case F_BY:  // solved on the transposed RHS (lcode.py:287-288)
            p.rowsA0 = 0; p.rowsA = N;
            p.xfA = XF_DST; p.lenA = M; p.colA0 = 1;
            p.xfB = XF_DCT; p.table = ws->tab_mixed;
            p.lamL = ws->lam_i; p.lamK = ws->lam_j; p.lam_add = ws->lam_add;
            break;
        case F_EZ:  // DST-I both ways on the interior (lcode.py:166-182)
            p.rowsA0 = 1; p.rowsA = M;
            p.xfA = XF_DST; p.lenA = M; p.colA0 = 1;
            p.xfB = XF_DST; p.table = ws->tab_dir;
            p.lamL = ws->lam_i; p.lamK = ws->lam_i;
            break;
        default:    // F_BZ: DCT-I both ways, zero mode dropped (:339-362)
            p.rowsA0 = 0; p.rowsA = N;
            p.xfA = XF_DCT; p.lenA = N; p.colA0 = 0;
            p.xfB = XF_DCT; p.table = ws->tab_neu;
            p.lamL = ws->lam_j; p.lamK = ws->lam_j; p.zero_mode = 1;
            p.colsum = ws->colsum;
            p.partial = ws->partial;
            break;
    }
    p.lam_norm = ws->lam_norm;
    return p;
}

int run_solve(lcode_workspace* ws, const SolveBatch& b, hipStream_t stream) {
    if (ws->solver == LCODE_SOLVER_DENSE64 || ws->solver == LCODE_SOLVER_DENSE_BF16X3)
        return launch_solve_dense(ws->N, ws->dense, b, stream, &ws->prof);
    return launch_solve(ws->geom, b, stream, &ws->prof);
}

RhsInputs rhs_base(const lcode_config* c) {
    RhsInputs in{};
    in.trick = c->subtraction_trick;
    in.h = c->grid_step_size;
    in.dxi = c->xi_step_size;
    return in;
}

}  // namespace

extern "C" {

const char* lcode_last_error(void) { return g_last_error.c_str(); }

const char* lcode_version(void) { return "lcode3d_hip 0.1 (gfx950, FP64)"; }

int lcode_workspace_create(const lcode_config* cfg, lcode_workspace** out) {
    int rc = check_cfg(cfg);
    if (rc) return rc;
    LCODE_REQUIRE(out, "out is NULL");
    *out = nullptr;
    auto* ws = new lcode_workspace();
    auto fail = [&](int code) {
        lcode_workspace_destroy(ws);
        return code;
    };
    const int N = cfg->grid_steps, M = N - 2;
    ws->N = N;
    ws->M = M;
    ws->h = cfg->grid_step_size;
    ws->trick_on = cfg->subtraction_trick != 0;
    ws->nc = cfg->nc;
    ws->nf = cfg->nf;
    if (hipGetDevice(&ws->device) != hipSuccess) {
        set_error("hipGetDevice failed");
        return fail(LCODE_E_HIP);
    }

    // FFT geometry: one complex FFT of L = 2 (N - 1) points per line pair.
    SolverGeom& g = ws->geom;
    g.N = N;
    ws->solver = cfg->solver;
    ws->fp32 = cfg->solver == LCODE_SOLVER_FFT32 || cfg->solver == LCODE_SOLVER_MIXED;
    g.fp32 = ws->fp32;
    const int cplx = ws->fp32 ? (int)sizeof(float2) : (int)sizeof(double2);
    const int L = 2 * (N - 1);
    if (!factor_radices(L, g.fft)) {
        set_error("2 (grid_steps - 1) = " + std::to_string(L) +
                  " has a prime factor other than 2, 3, 5");
        return fail(LCODE_E_UNSUPPORTED);
    }
    g.static_L = solver_static_length(L) ? L : 0;
    g.Tp = solver_threads_per_pair(L, ws->fp32, g.static_L != 0);
    const int max_lds = LCODE_SOLVER_MAX_LDS;   // per block, bytes
    // line pairs per block: as many as 1024 threads and the LDS limit allow,
    // at most LCODE_FFT_MAX_PAIRS_F64 / _F32
    const int max_pairs = ws->fp32 ? LCODE_FFT_MAX_PAIRS_F32 : LCODE_FFT_MAX_PAIRS_F64;
    auto pairs_for = [&](int fixed_bytes) {
        int P = 1024 / g.Tp;
        const int by_lds = (max_lds - fixed_bytes) / (solver_padded_len(L) * cplx);
        if (by_lds < P) P = by_lds;
        return P > max_pairs ? max_pairs : P;
    };
    int P;
    if (g.static_L) {
        // stage twiddles in LDS when the block keeps LCODE_TW_LDS_MIN_PAIRS pairs
        P = pairs_for(solver_tw_count(L) * (int)sizeof(double2));   // FP64 roots
        g.tw_lds = LCODE_TW_LDS_OK && P >= LCODE_TW_LDS_MIN_PAIRS;
        if (!g.tw_lds) P = pairs_for(0);
    } else {
        g.tw_lds = 0;
        P = pairs_for((L / 4 + 1) * cplx / 2);
    }
    if (P < 1) {
        set_error("grid too large for one FFT line pair in 64 KiB of LDS");
        return fail(LCODE_E_UNSUPPORTED);
    }
    g.P = P;
    if ((N + g.Tp - 1) / g.Tp > 4) {
        set_error("internal: too many points per thread");
        return fail(LCODE_E_UNSUPPORTED);
    }
    if (g.static_L) {
        // stage with span sp and radix R: exp(-2 pi i k / (sp R)), k < sp,
        // after the entries of the earlier stages (the kernels form the
        // powers q = 2 .. R - 1)
        std::vector<double2> tw(solver_tw_count(L));
        int off = 0;
        for (int sp = 1, R = 1; sp < L; sp *= R) {
            R = solver_static_radix(L / sp);
            for (int k = 0; k < sp; ++k) {
                const long double a = 2.0L * 3.14159265358979323846264338327950288L *
                                      k / ((long double)sp * R);
                tw[off + k] = make_double2((double)std::cos(a), (double)-std::sin(a));
            }
            off += sp;
        }
        if ((rc = upload(ws, &ws->stage_tw, tw))) return fail(rc);
        g.stage_tw = ws->stage_tw;
    }

    g.wave_tw = nullptr;
    if (solver_wave_length(L) && (!ws->fp32 || (LCODE_FFT_WAVE_FP32 && L == 1280))) {
        // W_L^t = exp(-2 pi i t / L) for every t < L
        std::vector<double2> wt(L);
        for (int t = 0; t < L; ++t) {
            const long double a = 2.0L * 3.14159265358979323846264338327950288L * t / L;
            wt[t] = make_double2((double)std::cos(a), (double)-std::sin(a));
        }
        if ((rc = upload(ws, &ws->wave_tw, wt))) return fail(rc);
        g.wave_tw = ws->wave_tw;
    }

    std::vector<double> costab(L / 4 + 1);
    for (int r = 0; r <= L / 4; ++r) {
        const long double a = 2.0L * 3.14159265358979323846264338327950288L * r / L;
        costab[r] = (double)std::cos(a);
    }
    if ((rc = upload(ws, &ws->costab, costab))) return fail(rc);
    g.costab = ws->costab;
    if (ws->fp32) {
        std::vector<float> c32(costab.begin(), costab.end());
        if ((rc = upload(ws, &ws->costab32, c32))) return fail(rc);
        g.costab32 = ws->costab32;
    }

    // Eigenvalue tables, normalisation 1 / (2 (N - 1))^2 folded in
    // (lcode.py:144-157, :215-231, :316-330).
    const double norm = (double)(2 * (N - 1)) * (2 * (N - 1));
    std::vector<double> li(M), lj(N);
    for (int k = 1; k <= M; ++k) li[k - 1] = lam(k, N, ws->h);
    for (int k = 0; k < N; ++k) lj[k] = lam(k, N, ws->h);
    const double add = ws->trick_on ? 1 : 0;
    std::vector<double> mixed((size_t)M * N), mixed_t((size_t)N * M);
    for (int a = 0; a < M; ++a)
        for (int b = 0; b < N; ++b) {
            const double v = 1 / (li[a] + lj[b] + add) / norm;
            mixed[(size_t)a * N + b] = v;
            mixed_t[(size_t)b * M + a] = v;
        }
    std::vector<double> dir((size_t)M * M), neu((size_t)N * N);
    for (int a = 0; a < M; ++a)
        for (int b = 0; b < M; ++b)
            dir[(size_t)a * M + b] = 1 / (li[a] + li[b]) / norm;
    for (int a = 0; a < N; ++a)
        for (int b = 0; b < N; ++b)
            neu[(size_t)a * N + b] = (a == 0 && b == 0) ? 0. : 1 / (lj[a] + lj[b]) / norm;
    // Column sums of the final DCT-I of Bz: the mean of Bz is
    // sum_k colsum[k] sum_i T[k][i] / N^2.
    std::vector<double> colsum(N);
    for (int k = 0; k < N; ++k) {
        long double s = 0;
        for (int j = 0; j < N; ++j)
            s += std::cos((long double)3.14159265358979323846264338327950288L * k * j / (N - 1));
        const long double c = (k == 0 || k == N - 1) ? 1 : 2;
        colsum[k] = (double)(c * s);
    }
    if ((rc = upload_table(ws, &ws->tab_mixed, mixed))) return fail(rc);
    if ((rc = upload_table(ws, &ws->tab_mixed_t, mixed_t))) return fail(rc);
    if ((rc = upload_table(ws, &ws->tab_dir, dir))) return fail(rc);
    if ((rc = upload_table(ws, &ws->tab_neu, neu))) return fail(rc);
    if ((rc = upload(ws, &ws->colsum, colsum))) return fail(rc);
    if ((rc = upload(ws, &ws->lam_i, li))) return fail(rc);
    if ((rc = upload(ws, &ws->lam_j, lj))) return fail(rc);
    ws->lam_add = add;
    ws->lam_norm = norm;

    {
        // zn diagnostic taps (scipy.ndimage.gaussian_filter1d: truncate 4)
        const double sigma = 0.25 / ws->h;
        const int r = (int)(4.0 * sigma + 0.5);
        std::vector<double> w(2 * r + 1);
        double tot = 0.;
        for (int k = -r; k <= r; ++k) tot += (w[k + r] = std::exp(-0.5 / (sigma * sigma) * k * k));
        for (double& v : w) v /= tot;
        ws->diag_r = r;
        if ((rc = upload(ws, &ws->diag_w, w))) return fail(rc);
        if ((rc = dev_alloc(ws, (void**)&ws->diag_tmp, (size_t)N * N * sizeof(double))))
            return fail(rc);
        if ((rc = dev_alloc(ws, (void**)&ws->diag_sum, sizeof(double)))) return fail(rc);
    }

    if (ws->solver == LCODE_SOLVER_DENSE64 || ws->solver == LCODE_SOLVER_DENSE_BF16X3) {
        // Dense type-1 transform matrices, unnormalised like the FFT form:
        // DST-I Sd[k][n] = 2 sin(pi (k+1)(n+1) / (N-1)), DCT-I
        // Cd[k][n] = w_n cos(pi k n / (N-1)) with w = 1 at both ends, else 2.
        const long double pi = 3.14159265358979323846264338327950288L;
        std::vector<double> Sd((size_t)M * M), Cd((size_t)N * N), CdT((size_t)N * N);
        for (int k = 0; k < M; ++k)
            for (int n = 0; n < M; ++n)
                Sd[(size_t)k * M + n] = (double)(2 * std::sin(pi * ((k + 1) * (long long)(n + 1) %
                                                                 (2 * (N - 1))) / (N - 1)));
        for (int k = 0; k < N; ++k)
            for (int n = 0; n < N; ++n) {
                const long double w = (n == 0 || n == N - 1) ? 1 : 2;
                const double v = (double)(w * std::cos(pi * ((long long)k * n % (2 * (N - 1))) /
                                                       (N - 1)));
                Cd[(size_t)k * N + n] = v;
                CdT[(size_t)n * N + k] = v;
            }
        DenseGeom& d = ws->dense;
        d.prec = ws->solver == LCODE_SOLVER_DENSE64 ? DENSE_F64 : DENSE_BF16X3;
        double *pS, *pC, *pCT, *pm, *pmt;
        if ((rc = upload(ws, &pS, Sd))) return fail(rc);
        if ((rc = upload(ws, &pC, Cd))) return fail(rc);
        if ((rc = upload(ws, &pCT, CdT))) return fail(rc);
        if ((rc = upload(ws, &pm, mixed))) return fail(rc);
        if ((rc = upload(ws, &pmt, mixed_t))) return fail(rc);
        d.Sd = pS; d.Cd = pC; d.CdT = pCT; d.mixed = pm; d.mixed_t = pmt;
        d.colsum = ws->colsum;
        if ((rc = dev_alloc(ws, (void**)&d.mean, sizeof(double)))) return fail(rc);
        if (d.prec == DENSE_F64) {
            const size_t n2d = (size_t)N * N;
            if ((rc = dev_alloc(ws, (void**)&d.X, 6 * n2d * sizeof(double)))) return fail(rc);
            if ((rc = dev_alloc(ws, (void**)&d.Y, 6 * n2d * sizeof(double)))) return fail(rc);
        } else {
            // hi/lo BF16 planes of the two matrices, zero padded to Np x Np
            const int Np = (N + 127) / 128 * 128;
            d.Np = Np;
            const size_t np2 = (size_t)Np * Np;
            std::vector<uint16_t> sh(np2, 0), sl(np2, 0), ch(np2, 0), cl(np2, 0);
            for (int k = 0; k < M; ++k)
                for (int n = 0; n < M; ++n)
                    split_bf16(Sd[(size_t)k * M + n], sh[(size_t)k * Np + n], sl[(size_t)k * Np + n]);
            for (int k = 0; k < N; ++k)
                for (int n = 0; n < N; ++n)
                    split_bf16(Cd[(size_t)k * N + n], ch[(size_t)k * Np + n], cl[(size_t)k * Np + n]);
            uint16_t *p0, *p1, *p2, *p3;
            if ((rc = upload(ws, &p0, sh)) || (rc = upload(ws, &p1, sl)) ||
                (rc = upload(ws, &p2, ch)) || (rc = upload(ws, &p3, cl)))
                return fail(rc);
            d.Sh = p0; d.Sl = p1; d.Ch = p2; d.Cl = p3;
            uint16_t** planes[4] = {&d.Xh, &d.Xl, &d.Ph, &d.Pl};
            for (uint16_t** pl : planes) {
                if ((rc = dev_alloc(ws, (void**)pl, 6 * np2 * sizeof(uint16_t)))) return fail(rc);
                LCODE_HIP(hipMemset(*pl, 0, 6 * np2 * sizeof(uint16_t)));
            }
        }
    }

    const size_t n2 = (size_t)N * N, np = (size_t)cfg->nc * cfg->nc;
    // six N x (N + 1) slots: the wave FFT keeps rows N + 1 long (even)
    if ((rc = dev_alloc(ws, &ws->T, 6 * (n2 + N) * (ws->fp32 ? sizeof(float) : sizeof(double)))))
        return fail(rc);
    if ((rc = dev_alloc(ws, (void**)&ws->partial, N * sizeof(double)))) return fail(rc);
    for (int k = 0; k < 5; ++k) {
        if ((rc = dev_alloc(ws, (void**)&ws->p1[k], (np ? np : 1) * sizeof(double)))) return fail(rc);
        if ((rc = dev_alloc(ws, (void**)&ws->p2[k], (np ? np : 1) * sizeof(double)))) return fail(rc);
    }
    for (int k = 0; k < 4; ++k)
        if ((rc = dev_alloc(ws, (void**)&ws->dep[k], n2 * sizeof(double)))) return fail(rc);
    for (int k = 0; k < 6; ++k) {
        if ((rc = dev_alloc(ws, (void**)&ws->favg[k], n2 * sizeof(double)))) return fail(rc);
        if ((rc = dev_alloc(ws, (void**)&ws->guess[k], n2 * sizeof(double)))) return fail(rc);
    }
    if ((rc = dev_alloc(ws, (void**)&ws->resid, 2 * sizeof(double)))) return fail(rc);
    if (cfg->solver == LCODE_SOLVER_MIXED) {
        if ((rc = dev_alloc(ws, (void**)&ws->zero64, n2 * sizeof(double)))) return fail(rc);
        LCODE_HIP(hipMemset(ws->zero64, 0, n2 * sizeof(double)));
    }
    {
        // last FFT pass: ceil(N / (2 P)) blocks per field (one line pair per
        // block in the wave FFT), six fields: sized for P = 1, the most
        const int per_block = 2;
        const size_t nblk = (size_t)6 * ((N + per_block - 1) / per_block);
        if ((rc = dev_alloc(ws, (void**)&ws->resid_part, 2 * nblk * sizeof(double))))
            return fail(rc);
    }
    if ((rc = dev_alloc(ws, (void**)&ws->fault, sizeof(int)))) return fail(rc);
    const size_t nfp = (size_t)cfg->nf * cfg->nf;
    if ((rc = dev_alloc(ws, (void**)&ws->spill, (nfp ? nfp : 1) * sizeof(int)))) return fail(rc);
    if ((rc = dev_alloc(ws, (void**)&ws->spill_ctr, 5 * sizeof(int)))) return fail(rc);
    // per-tile displacement bounds: a deposition tile takes the largest of
    // its 3 x 3 neighbourhood, which holds every coarse particle a fine
    // particle within the largest halo + 1 cells of the tile interpolates
    // from when 2 coarseness + 10 <= tile side (a fine particle lies within
    // 2 coarseness + 2 cells of its coarse neighbours, at the plasma edge)
    const int nt = (N + kDepTile - 1) / kDepTile;
    if (kDepRegions && 2 * cfg->plasma_coarseness + kDepositHalo + 4 <= kDepTile) {
        ws->region_n = nt;
        if ((rc = dev_alloc(ws, (void**)&ws->regions, 2 * (size_t)nt * nt * sizeof(int))))
            return fail(rc);
    }
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ws->device) !=
        hipSuccess) {
        set_error("hipDeviceGetAttribute (multiprocessor count) failed");
        return fail(LCODE_E_HIP);
    }
    int resident = deposit_blocks_per_cu(deposit_staged(cfg));
    if (resident <= 0) resident = kDepResident;
    if (kDepOrder) {
        std::vector<int> order((size_t)nt * nt);
        for (int k = 0; k < nt * nt; ++k) order[k] = k;
        if (kDepOrder == 2) {
            order = xcd_cluster_order(nt, LCODE_DEP_CLUSTER);
            // (when every tile block is resident at once: one round)
            if (LCODE_DEP_BALANCE && cus >= 8 && cus % 8 == 0 &&
                (LCODE_DEP_BALANCE == 2 || nt * nt <= resident * cus))
                balance_xcd_slots(order, nt, cus / 8);
        } else {
            const double mid = (nt - 1) / 2.;
            auto dist = [&](int k) {
                return std::max(std::fabs(k / nt - mid), std::fabs(k % nt - mid));
            };
            std::stable_sort(order.begin(), order.end(),
                             [&](int u, int v) { return dist(u) < dist(v); });
        }
        if ((rc = dev_alloc(ws, (void**)&ws->tile_order, order.size() * sizeof(int))) ||
            (rc = hipMemcpy(ws->tile_order, order.data(), order.size() * sizeof(int),
                            hipMemcpyHostToDevice) == hipSuccess ? LCODE_OK : LCODE_E_HIP)) {
            if (rc == LCODE_E_HIP) set_error("hipMemcpy of the tile order failed");
            return fail(rc);
        }
    }
    {
        const int free_slots = std::max(resident * cus - nt * nt, 0);
        const int ns = std::min(LCODE_DEP_SPLIT < 0 ? free_slots : LCODE_DEP_SPLIT, nt * nt);
        if (ns > 0) {
            const std::vector<int> tab = split_tile_table(nt, ns);
            ws->nsplit = ns;
            const size_t part = 2 * 4 * (size_t)kDepTile * kDepTile;
            if ((rc = upload(ws, &ws->split_tile, tab)) ||
                (rc = dev_alloc(ws, (void**)&ws->split_part, ns * part * sizeof(double))) ||
                (rc = dev_alloc(ws, (void**)&ws->split_count, ns * sizeof(int))))
                return fail(rc);
            if (hipMemset(ws->split_count, 0, ns * sizeof(int)) != hipSuccess) {
                set_error("hipMemset of the split counts failed");
                return fail(LCODE_E_HIP);
            }
        }
    }
    if (hipHostMalloc((void**)&ws->host_flag, 2 * sizeof(int), hipHostMallocDefault) != hipSuccess) {
        ws->host_flag = nullptr;
        set_error("hipHostMalloc of the host flags failed");
        return fail(LCODE_E_HIP);
    }
    ws->host_flag[0] = ws->host_flag[1] = 0;
    if (hipMemset(ws->fault, 0, sizeof(int)) != hipSuccess ||
        hipMemset(ws->spill_ctr, 0, 5 * sizeof(int)) != hipSuccess ||
        (ws->regions &&
         hipMemset(ws->regions, 0, 2 * (size_t)nt * nt * sizeof(int)) != hipSuccess) ||
        hipDeviceSynchronize() != hipSuccess) {
        set_error("hipMemset of the workspace counters failed");
        return fail(LCODE_E_HIP);
    }
    *out = ws;
    return LCODE_OK;
}

static void profiler_free(Profiler& p) {
    for (int k = 0; k < 2 * p.capacity; ++k) (void)hipEventDestroy(p.ev[k]);
    delete[] p.ev;
    delete[] p.names;
    p = Profiler();
}

int lcode_workspace_destroy(lcode_workspace* ws) {
    if (!ws) return LCODE_OK;
    profiler_free(ws->prof);
    int rc = LCODE_OK;
    for (void* p : ws->allocs)
        if (hipFree(p) != hipSuccess) rc = LCODE_E_HIP;
    if (ws->host_flag && hipHostFree(ws->host_flag) != hipSuccess) rc = LCODE_E_HIP;
    delete ws;
    if (rc) set_error("hipFree failed in lcode_workspace_destroy");
    return rc;
}

int lcode_profile_begin(lcode_workspace* ws, int32_t max_launches) {
    LCODE_REQUIRE(ws && max_launches > 0, "bad profile_begin arguments");
    profiler_free(ws->prof);
    Profiler& p = ws->prof;
    p.ev = new hipEvent_t[2 * (size_t)max_launches];
    p.names = new const char*[max_launches];
    for (int k = 0; k < 2 * max_launches; ++k) {
        if (hipEventCreate(&p.ev[k]) != hipSuccess) {
            p.capacity = k / 2;
            profiler_free(p);
            set_error("hipEventCreate failed");
            return LCODE_E_HIP;
        }
        p.capacity = (k + 1) / 2;
    }
    p.capacity = max_launches;
    p.used = 0;
    p.on = true;
    return LCODE_OK;
}

int lcode_profile_only(lcode_workspace* ws, const char* kernel_name) {
    LCODE_REQUIRE(ws, "NULL workspace");
    const char* n = kernel_name ? kernel_name : "";
    LCODE_REQUIRE(std::strlen(n) < sizeof(ws->prof.only), "kernel name too long");
    std::strcpy(ws->prof.only, n);
    return LCODE_OK;
}

int lcode_profile_pause(lcode_workspace* ws, int32_t paused) {
    LCODE_REQUIRE(ws, "NULL workspace");
    ws->prof.paused = paused != 0;
    return LCODE_OK;
}

int lcode_profile_end(lcode_workspace* ws, hipStream_t stream,
                      lcode_kernel_time* out, int32_t max_out,
                      int32_t* n_out) {
    LCODE_REQUIRE(ws && out && n_out, "bad profile_end arguments");
    Profiler& p = ws->prof;
    p.on = false;
    LCODE_HIP(hipStreamSynchronize(stream));
    int n = 0;
    for (int k = 0; k < p.used; ++k) {
        float ms = 0;
        LCODE_HIP(hipEventElapsedTime(&ms, p.ev[2 * k], p.ev[2 * k + 1]));
        int slot = -1;
        for (int m = 0; m < n; ++m)
            if (std::strcmp(out[m].name, p.names[k]) == 0) slot = m;
        if (slot < 0) {
            if (n == max_out) continue;
            slot = n++;
            std::memset(&out[slot], 0, sizeof(out[slot]));
            std::strncpy(out[slot].name, p.names[k], sizeof(out[slot].name) - 1);
        }
        out[slot].calls += 1;
        out[slot].total_ms += ms;
    }
    *n_out = n;
    profiler_free(p);
    return LCODE_OK;
}

int lcode_diag_zn(lcode_workspace* ws, const lcode_config* cfg, const double* ro,
                  double* zn_out, hipStream_t stream) {
    int rc = check_ws(ws, cfg);
    if (rc) return rc;
    LCODE_REQUIRE(ro && zn_out, "NULL array");
    return launch_diag_zn(ro, ws->diag_w, ws->diag_r, ws->N, ws->diag_tmp,
                          ws->diag_sum, zn_out, stream, &ws->prof);
}

int lcode_workspace_nonfinite(lcode_workspace* ws, int32_t reset,
                              int32_t* flag) {
    LCODE_REQUIRE(ws && flag, "NULL argument");
    volatile int* f = ws->host_flag;
    *flag = *f;
    if (reset) *f = 0;
    return LCODE_OK;
}

int lcode_workspace_fault_count(lcode_workspace* ws, hipStream_t stream,
                                int64_t* count) {
    LCODE_REQUIRE(ws && count, "NULL argument");
    int host = 0;
    LCODE_HIP(hipMemcpyAsync(&host, ws->fault, sizeof(int),
                             hipMemcpyDeviceToHost, stream));
    LCODE_HIP(hipStreamSynchronize(stream));
    LCODE_HIP(hipMemsetAsync(ws->fault, 0, sizeof(int), stream));
    *count = host;
    return LCODE_OK;
}

int lcode_move_estimate_wo_fields(lcode_workspace* ws, const lcode_config* cfg,
                                  const double* m, const double* x_init,
                                  const double* y_init,
                                  const double* prev_x_offt,
                                  const double* prev_y_offt, const double* px,
                                  const double* py, const double* pz,
                                  double* x_offt_out, double* y_offt_out,
                                  hipStream_t stream) {
    int rc = check_ws(ws, cfg);
    if (rc) return rc;
    EstimateArgs a{};
    a.n = cfg->nc * cfg->nc;
    a.dxi = cfg->xi_step_size;
    a.reflect = cfg->reflect_boundary;
    a.m = m; a.x_init = x_init; a.y_init = y_init;
    a.prev_x_offt = prev_x_offt; a.prev_y_offt = prev_y_offt;
    a.px = px; a.py = py; a.pz = pz;
    a.out_x = x_offt_out; a.out_y = y_offt_out;
    LCODE_REQUIRE(a.n == 0 || (m && x_init && y_init && prev_x_offt && prev_y_offt &&
                               px && py && pz && x_offt_out && y_offt_out),
                  "NULL array");
    return launch_estimate(a, stream, &ws->prof);
}

static int do_push(lcode_workspace* ws, const lcode_config* cfg,
                   const double* m, const double* q, const double* x_init,
                   const double* y_init, const double* xo, const double* yo,
                   const double* ex, const double* ey, const double* px,
                   const double* py, const double* pz,
                   const double* const F[6], double* const out[5],
                   hipStream_t stream, Gate gate = Gate{nullptr, 0.}) {
    PushArgs a{};
    a.n = cfg->nc * cfg->nc;
    a.N = cfg->grid_steps;
    a.h = cfg->grid_step_size;
    a.dxi = cfg->xi_step_size;
    a.reflect = cfg->reflect_boundary;
    a.m = m; a.q = q; a.x_init = x_init; a.y_init = y_init;
    a.prev_x_offt = xo; a.prev_y_offt = yo;
    a.est_x = ex; a.est_y = ey;
    a.px = px; a.py = py; a.pz = pz;
    for (int k = 0; k < 6; ++k) a.F[k] = F[k];
    a.out_x = out[0]; a.out_y = out[1];
    a.out_px = out[2]; a.out_py = out[3]; a.out_pz = out[4];
    a.fault = ws->fault;
    a.max_cells = ws->spill_ctr + 2 + ws->bound_slot;
    if (ws->regions) {
        a.region_max = ws->regions + (size_t)ws->bound_slot * ws->region_n * ws->region_n;
        a.region_n = ws->region_n;
        a.region_t = kDepTile;
    }
    a.host_flag = ws->host_flag;
    a.grid32 = ws->grid32;
    a.gate = gate;
    return launch_push(a, stream, &ws->prof);
}

int lcode_move_smart(lcode_workspace* ws, const lcode_config* cfg,
                     const double* m, const double* q, const double* x_init,
                     const double* y_init, const double* x_prev_offt,
                     const double* y_prev_offt,
                     const double* estimated_x_offt,
                     const double* estimated_y_offt, const double* px_prev,
                     const double* py_prev, const double* pz_prev,
                     const double* const fields[6], double* const out[5],
                     hipStream_t stream) {
    int rc = check_ws(ws, cfg);
    if (rc) return rc;
    LCODE_REQUIRE(fields && out, "NULL fields/out");
    LCODE_REQUIRE(estimated_x_offt && estimated_y_offt,
                  "estimated offsets are required");
    for (int k = 0; k < 6; ++k) LCODE_REQUIRE(fields[k], "NULL field");
    for (int k = 0; k < 5; ++k) LCODE_REQUIRE(out[k], "NULL output");
    return do_push(ws, cfg, m, q, x_init, y_init, x_prev_offt, y_prev_offt,
                   estimated_x_offt, estimated_y_offt, px_prev, py_prev,
                   pz_prev, fields, out, stream);
}

// Deposition arguments that do not depend on where the scratch lives
// (spill list, counters and flags are set by the caller).
static DepositArgs deposit_args(const lcode_config* cfg,
                                const double* ro_initial, const double* xo,
                                const double* yo, const double* m,
                                const double* q, const double* px,
                                const double* py, const double* pz,
                                const lcode_virt_params* vp,
                                double* const out[4]) {
    DepositArgs a{};
    a.nf = cfg->nf;
    a.nc = cfg->nc;
    a.N = cfg->grid_steps;
    a.h = cfg->grid_step_size;
    const double cf = (double)cfg->plasma_coarseness * cfg->plasma_fineness;
    a.smallness = 1 / (cf * cf);               // lcode.py:672-673
    a.x_offt = xo; a.y_offt = yo; a.m = m; a.q = q;
    a.px = px; a.py = py; a.pz = pz;
    a.ip = vp->influence_prev; a.in = vp->influence_next; a.fg = vp->fine_grid;
    a.P = vp->indices_prev; a.Nn = vp->indices_next;
    a.ro_initial = ro_initial;
    a.ro = out[0]; a.jx = out[1]; a.jy = out[2]; a.jz = out[3];
    a.halo = kDepositHalo;
    a.ntx = (a.N + kDepTile - 1) / kDepTile;
    a.staged = deposit_staged(cfg);
    a.group = cfg->plasma_fineness;
    return a;
}

static int do_deposit(lcode_workspace* ws, const lcode_config* cfg,
                      const double* ro_initial, const double* xo,
                      const double* yo, const double* m, const double* q,
                      const double* px, const double* py, const double* pz,
                      const lcode_virt_params* vp, double* const out[4],
                      hipStream_t stream, Gate gate = Gate{nullptr, 0.},
                      bool adaptive_halo = false) {
    DepositArgs a = deposit_args(cfg, ro_initial, xo, yo, m, q, px, py, pz,
                                 vp, out);
    a.grid32 = ws->grid32;
    a.fault = ws->fault;
    a.spill = ws->spill;
    a.host_flag = ws->host_flag;
    a.spill_count = ws->spill_ctr + ws->spill_slot;
    a.spill_next = ws->spill_ctr + (1 - ws->spill_slot);
    a.tile_order = ws->tile_order;
    a.nsplit = ws->nsplit;
    a.split_tile = ws->split_tile;
    a.split_slot = ws->split_tile ? ws->split_tile + ws->nsplit : nullptr;
    a.split_part = ws->split_part;
    a.split_count = ws->split_count;
    // the spill list is deposited by the tile kernel's last block, unless
    // a deposition listed particles recently: then for the next 64 the
    // grid-wide spill kernel follows (a long list is many blocks' work)
    {
        volatile int* seen = ws->host_flag + 1;
        if (*seen) {
            *seen = 0;
            ws->grid_spill_left = 64;
        }
        a.spill_in_tiles = ws->grid_spill_left > 0 ? 0 : 1;
        if (ws->grid_spill_left > 0) --ws->grid_spill_left;
        a.done = ws->spill_ctr + 4;
        a.host_spill = ws->host_flag + 1;
    }
    // after a push of this workspace: a halo of the pushes' displacement
    // bound, read from the slot the pushes raised; the deposition zeroes
    // the other slot, which the next pushes raise
    if (adaptive_halo) {
        a.halo_cells = ws->spill_ctr + 2 + ws->bound_slot;
        a.halo_next = ws->spill_ctr + 2 + (1 - ws->bound_slot);
        if (ws->regions) {
            const size_t nr = (size_t)ws->region_n * ws->region_n;
            a.halo_regions = ws->regions + ws->bound_slot * nr;
            a.halo_regions_next = ws->regions + (1 - ws->bound_slot) * nr;
        }
    }
    a.gate = gate;
    LCODE_REQUIRE(cfg->nf <= ws->nf, "workspace was built for fewer fine particles");
    const int rc = launch_deposit(a, stream, &ws->prof);
    if (rc == LCODE_OK) ws->spill_slot = 1 - ws->spill_slot;
    if (rc == LCODE_OK && adaptive_halo) ws->bound_slot = 1 - ws->bound_slot;
    return rc;
}

int lcode_plasma_sizes(int32_t steps, int32_t coarseness, int32_t fineness,
                       int32_t* nc, int32_t* nf) {
    LCODE_REQUIRE(nc && nf, "NULL output");
    LCODE_REQUIRE(steps > 0 && coarseness > 0 && fineness > 0,
                  "steps, coarseness and fineness must be positive");
    const int hc = steps / (coarseness * 2), hf = steps / 2 * fineness;
    LCODE_REQUIRE(hc >= 1 && hf >= 1, "too few plasma cells for this coarseness");
    *nc = 2 * hc - 1;
    *nf = fineness % 2 ? 2 * hf - 1 : 2 * hf;
    return LCODE_OK;
}

int lcode_make_plasma(int32_t steps, double cell_size, int32_t coarseness,
                      int32_t fineness, double* x_init, double* y_init,
                      double* m, double* q, double* influence_prev,
                      double* influence_next, double* fine_grid,
                      int64_t* indices_prev, int64_t* indices_next,
                      hipStream_t stream) {
    int nc = 0, nf = 0;
    int rc = lcode_plasma_sizes(steps, coarseness, fineness, &nc, &nf);
    if (rc) return rc;
    LCODE_REQUIRE(cell_size > 0, "cell_size must be positive");
    LCODE_REQUIRE(x_init && y_init && m && q && influence_prev && influence_next &&
                      fine_grid && indices_prev && indices_next,
                  "NULL output array");
    PlasmaArgs a{};
    a.coarseness = coarseness;
    a.fineness = fineness;
    a.hc = steps / (coarseness * 2);
    a.hf = steps / 2 * fineness;
    a.nc = nc;
    a.nf = nf;
    a.cell = cell_size;
    a.x_init = x_init; a.y_init = y_init; a.m = m; a.q = q;
    a.ip = influence_prev; a.in = influence_next; a.fg = fine_grid;
    a.P = indices_prev; a.Nn = indices_next;
    return launch_make_plasma(a, stream);
}

int lcode_deposit(lcode_workspace* ws, const lcode_config* cfg,
                  const double* ro_initial, const double* x_offt,
                  const double* y_offt, const double* m, const double* q,
                  const double* px, const double* py, const double* pz,
                  const lcode_virt_params* virt_params, double* const out[4],
                  hipStream_t stream) {
    int rc = check_ws(ws, cfg);
    if (rc) return rc;
    LCODE_REQUIRE(virt_params && out, "NULL virt_params/out");
    for (int k = 0; k < 4; ++k) LCODE_REQUIRE(out[k], "NULL output");
    return do_deposit(ws, cfg, ro_initial, x_offt, y_offt, m, q, px, py, pz,
                      virt_params, out, stream);
}

}  // extern "C"

// Scratch of lcode_deposit_noalloc: the spill list (kNoAllocMaxFine^2
// fine-particle indices) followed by its counters.  One copy per device,
// held until the process ends; in host code `scratch` names a host-side
// stand-in, so only the address hipGetSymbolAddress returns for the
// current device is ever used.
namespace {
constexpr int kNoAllocMaxFine = 2048;
constexpr size_t kNoAllocSpillCap = (size_t)kNoAllocMaxFine * kNoAllocMaxFine;
constexpr int kNoAllocCounters = 64;  // [0] spill count, [1] the spill
                                      // kernel's next-count slot, [2] faults
constexpr int kNoAllocMaxDevices = 64;
}  // namespace

static __device__ __align__(256) int scratch[kNoAllocSpillCap + kNoAllocCounters];

static std::mutex g_scratch_mu;
static int* g_scratch_ptr[kNoAllocMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per
// device (with that device current) and kept.
static int scratch_on_current_device(int** out) {
    int dev = 0;
    LCODE_HIP(hipGetDevice(&dev));
    LCODE_REQUIRE(dev >= 0 && dev < kNoAllocMaxDevices, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        int* p = nullptr;
        const hipError_t e = hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error(std::string("hipGetSymbolAddress(scratch) failed: ") +
                      hipGetErrorString(e));
            return LCODE_E_HIP;
        }
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return LCODE_OK;
}

extern "C" {

int lcode_deposit_noalloc(const lcode_config* cfg, const double* ro_initial,
                          const double* x_offt, const double* y_offt,
                          const double* m, const double* q, const double* px,
                          const double* py, const double* pz,
                          const lcode_virt_params* virt_params,
                          double* const out[4], hipStream_t stream) {
    int rc = check_cfg(cfg);
    if (rc) return rc;
    LCODE_REQUIRE(virt_params && out, "NULL virt_params/out");
    LCODE_REQUIRE(x_offt && y_offt && m && q && px && py && pz, "NULL particle array");
    LCODE_REQUIRE(virt_params->influence_prev && virt_params->influence_next &&
                      virt_params->fine_grid && virt_params->indices_prev &&
                      virt_params->indices_next,
                  "NULL virt_params array");
    for (int k = 0; k < 4; ++k) LCODE_REQUIRE(out[k], "NULL output");
    LCODE_REQUIRE(cfg->nf <= kNoAllocMaxFine,
                  "lcode_deposit_noalloc: more than 2048 fine particles per axis "
                  "(the spill list of its fixed scratch holds 2048^2)");
    int* s = nullptr;
    if ((rc = scratch_on_current_device(&s))) return rc;
    DepositArgs a = deposit_args(cfg, ro_initial, x_offt, y_offt, m, q, px, py,
                                 pz, virt_params, out);
    int* ctr = s + kNoAllocSpillCap;
    a.spill = s;
    a.spill_count = ctr;
    a.spill_next = ctr + 1;
    a.fault = ctr + 2;
    // the previous call left its count in ctr[0]: clear it with a kernel
    // (no memset node, so the call can be captured in a graph)
    if ((rc = launch_zero_ints(ctr, 3, stream))) return rc;
    return launch_deposit(a, stream, nullptr);
}

int lcode_calculate_Ex_Ey_Bx_By(lcode_workspace* ws, const lcode_config* cfg,
                                const double* Ex_avg, const double* Ey_avg,
                                const double* Bx_avg, const double* By_avg,
                                const double* beam_ro, double beam_scalar,
                                const double* ro, const double* jx,
                                const double* jy, const double* jz,
                                const double* jx_prev, const double* jy_prev,
                                double* const out[4], hipStream_t stream) {
    int rc = check_ws(ws, cfg);
    if (rc) return rc;
    LCODE_REQUIRE(Ex_avg && Ey_avg && Bx_avg && By_avg && ro && jx && jy &&
                      jz && jx_prev && jy_prev && out,
                  "NULL array");
    SolveBatch b{};
    b.nfields = 4;
    const int kinds[4] = {F_EX, F_EY, F_BX, F_BY};
    for (int k = 0; k < 4; ++k) {
        LCODE_REQUIRE(out[k], "NULL output");
        b.f[k] = plan_for(ws, kinds[k], k);
        b.f[k].outF = out[k];
    }
    b.in = rhs_base(cfg);
    b.in.ro = ro; b.in.jx = jx; b.in.jy = jy; b.in.jz = jz;
    b.in.jx_prev = jx_prev; b.in.jy_prev = jy_prev;
    b.in.Ftrick[0] = Ex_avg; b.in.Ftrick[1] = Ey_avg;
    b.in.Ftrick[2] = Bx_avg; b.in.Ftrick[3] = By_avg;
    b.in.beam = beam_ro;
    b.in.beam_scalar = beam_scalar;
    b.variantA = 0;
    return run_solve(ws, b, stream);
}

static int solve_single(lcode_workspace* ws, const lcode_config* cfg,
                        int kind, const double* jx, const double* jy,
                        double* outF, hipStream_t stream) {
    int rc = check_ws(ws, cfg);
    if (rc) return rc;
    LCODE_REQUIRE(jx && jy && outF, "NULL array");
    SolveBatch b{};
    b.nfields = 1;
    b.f[0] = plan_for(ws, kind, 0);
    b.f[0].outF = outF;
    b.in = rhs_base(cfg);
    b.in.jx = jx;
    b.in.jy = jy;
    return run_solve(ws, b, stream);
}

int lcode_calculate_Ez(lcode_workspace* ws, const lcode_config* cfg,
                       const double* jx, const double* jy, double* Ez,
                       hipStream_t stream) {
    return solve_single(ws, cfg, F_EZ, jx, jy, Ez, stream);
}

int lcode_calculate_Bz(lcode_workspace* ws, const lcode_config* cfg,
                       const double* jx, const double* jy, double* Bz,
                       hipStream_t stream) {
    return solve_single(ws, cfg, F_BZ, jx, jy, Bz, stream);
}

// Field half of a predictor/corrector pass (lcode.py:862-881, :893-911):
// the six solves in one batch; F = solved fields (or NULL), avg = (F + prev)/2.
static int field_pass(lcode_workspace* ws, const lcode_config* cfg,
                      const lcode_state* prev, const double* const trick[4],
                      const double* beam_ro, double beam_scalar,
                      double* const dep[4], double* const F[6],
                      double* const avg[6], hipStream_t stream,
                      Gate gate = Gate{nullptr, 0.},
                      const double* const guess[6] = nullptr,
                      double* resid = nullptr) {
    SolveBatch b{};
    b.gate = gate;
    if (guess && resid) {
        for (int k = 0; k < 6; ++k) b.guess[k] = guess[k];
        b.resid = resid;
        b.resid_part = ws->resid_part;
    }
    b.nfields = 6;
    const double* prevF[6] = {prev->Ex, prev->Ey, prev->Ez,
                              prev->Bx, prev->By, prev->Bz};
    for (int k = 0; k < 6; ++k) {
        b.f[k] = plan_for(ws, k, k);   // slot k, kind k (F_EX..F_BZ order)
        b.f[k].prevF = prevF[k];
        b.f[k].outF = F ? F[k] : nullptr;
        b.f[k].outAvg = avg[k];
    }
    b.in = rhs_base(cfg);
    b.in.ro = dep[0]; b.in.jx = dep[1]; b.in.jy = dep[2]; b.in.jz = dep[3];
    if (cfg->variant_A) {
        b.in.ro_prev = prev->ro;
        b.in.jz_prev = prev->jz;
    }
    b.in.jx_prev = prev->jx;
    b.in.jy_prev = prev->jy;
    for (int k = 0; k < 4; ++k) b.in.Ftrick[k] = trick[k];
    b.in.beam = beam_ro;
    b.in.beam_scalar = beam_scalar;
    b.in.zero64 = ws->zero64;
    b.variantA = cfg->variant_A ? 1 : 0;
    b.grid32 = ws->grid32;
    return run_solve(ws, b, stream);
}

}  // extern "C"

// The composition of the xi step shared by lcode_step (FP64 grids) and
// lcode_step_mixed (float grids, ws->grid32 set by the caller).
static int step_impl(lcode_workspace* ws, const lcode_config* cfg,
                     const lcode_const* c, const lcode_virt_params* vp,
                     const lcode_state* prev, const double* beam_ro,
                     double beam_scalar, const lcode_state* out,
                     hipStream_t stream) {
    int rc = check_ws(ws, cfg);
    if (rc) return rc;
    LCODE_REQUIRE(c && vp && prev && out, "NULL struct");
    LCODE_REQUIRE(cfg->nc == ws->nc, "workspace was built for another nc");
    LCODE_REQUIRE(prev->Ex != out->Ex && prev->x_offt != out->x_offt,
                  "prev and out must not alias");
    const double* prevF[6] = {prev->Ex, prev->Ey, prev->Ez,
                              prev->Bx, prev->By, prev->Bz};

    // 1-2: field-free estimate fused into the predictor push with prev
    // fields (lcode.py:843-854).
    if ((rc = do_push(ws, cfg, c->m, c->q, c->x_init, c->y_init,
                      prev->x_offt, prev->y_offt, nullptr, nullptr,
                      prev->px, prev->py, prev->pz, prevF, ws->p1, stream)))
        return rc;
    // 3: deposit (lcode.py:856-859)
    if ((rc = do_deposit(ws, cfg, c->ro_initial, ws->p1[0], ws->p1[1], c->m,
                         c->q, ws->p1[2], ws->p1[3], ws->p1[4], vp, ws->dep,
                         stream, Gate{nullptr, 0.}, true)))
        return rc;
    // 4: fields with prev fields in the trick term; only the averages are
    // used afterwards (lcode.py:862-881)
    // (kept in the workspace's guess grids when asked: the next slice's
    // guess can start from them)
    const double* trick_prev[4] = {prev->Ex, prev->Ey, prev->Bx, prev->By};
    double* const keepF[6] = {ws->guess[0], ws->guess[1], ws->guess[2],
                              ws->guess[3], ws->guess[4], ws->guess[5]};
    if ((rc = field_pass(ws, cfg, prev, trick_prev, beam_ro, beam_scalar,
                         ws->dep, ws->keep_f1 ? keepF : nullptr, ws->favg, stream)))
        return rc;
    // 5: corrector push with the averaged fields from the pass-1 positions
    // (lcode.py:884-889)
    const double* avg[6] = {ws->favg[0], ws->favg[1], ws->favg[2],
                            ws->favg[3], ws->favg[4], ws->favg[5]};
    if ((rc = do_push(ws, cfg, c->m, c->q, c->x_init, c->y_init,
                      prev->x_offt, prev->y_offt, ws->p1[0], ws->p1[1],
                      prev->px, prev->py, prev->pz, avg, ws->p2, stream)))
        return rc;
    // 6: deposit (lcode.py:890-891)
    if ((rc = do_deposit(ws, cfg, c->ro_initial, ws->p2[0], ws->p2[1], c->m,
                         c->q, ws->p2[2], ws->p2[3], ws->p2[4], vp, ws->dep,
                         stream, Gate{nullptr, 0.}, true)))
        return rc;
    // 7: fields with the pass-1 averages in the trick term; the solved
    // fields are the step's output (lcode.py:893-911, :927-928).  The
    // trick inputs are read in pass A, before pass C overwrites favg.
    const double* trick_avg[4] = {ws->favg[0], ws->favg[1], ws->favg[3],
                                  ws->favg[4]};
    double* const outF[6] = {out->Ex, out->Ey, out->Ez,
                             out->Bx, out->By, out->Bz};
    if ((rc = field_pass(ws, cfg, prev, trick_avg, beam_ro, beam_scalar,
                         ws->dep, outF, ws->favg, stream)))
        return rc;
    // 8: final push with the pass-2 averages from the pass-2 positions
    // (lcode.py:914-919)
    double* const outP[5] = {out->x_offt, out->y_offt, out->px, out->py,
                             out->pz};
    if ((rc = do_push(ws, cfg, c->m, c->q, c->x_init, c->y_init,
                      prev->x_offt, prev->y_offt, ws->p2[0], ws->p2[1],
                      prev->px, prev->py, prev->pz, avg, outP, stream)))
        return rc;
    // 9: final deposit into the output (lcode.py:920-921)
    double* const outD[4] = {out->ro, out->jx, out->jy, out->jz};
    return do_deposit(ws, cfg, c->ro_initial, out->x_offt, out->y_offt, c->m,
                      c->q, out->px, out->py, out->pz, vp, outD, stream,
                      Gate{nullptr, 0.}, true);
}

static int guided_impl(lcode_workspace* ws, const lcode_config* cfg,
                       const lcode_const* c, const lcode_virt_params* vp,
                       const lcode_state* prev, const double* beam_ro,
                       double beam_scalar, const lcode_guess* guess,
                       double threshold, const lcode_state* out,
                       hipStream_t stream) {
    int rc = check_ws(ws, cfg);
    if (rc) return rc;
    LCODE_REQUIRE(c && vp && prev && out && guess, "NULL struct");
    LCODE_REQUIRE(cfg->nc == ws->nc, "workspace was built for another nc");
    LCODE_REQUIRE(prev->Ex != out->Ex && prev->x_offt != out->x_offt,
                  "prev and out must not alias");
    LCODE_REQUIRE(guess->nterms >= 1 && guess->nterms <= 4, "guess needs 1-4 terms");
    LCODE_REQUIRE(threshold >= 0, "negative threshold");
    for (int t = 0; t < guess->nterms; ++t)
        for (int f = 0; f < 6; ++f) LCODE_REQUIRE(guess->f[t][f], "NULL guess field");
    const double* prevF[6] = {prev->Ex, prev->Ey, prev->Ez,
                              prev->Bx, prev->By, prev->Bz};
    const double* avg[6] = {ws->favg[0], ws->favg[1], ws->favg[2],
                            ws->favg[3], ws->favg[4], ws->favg[5]};
    const double* trick_prev[4] = {prev->Ex, prev->Ey, prev->Bx, prev->By};
    const double* trick_avg[4] = {ws->favg[0], ws->favg[1], ws->favg[3],
                                  ws->favg[4]};
    double* const outF[6] = {out->Ex, out->Ey, out->Ez, out->Bx, out->By, out->Bz};
    const double* g6[6] = {ws->guess[0], ws->guess[1], ws->guess[2],
                           ws->guess[3], ws->guess[4], ws->guess[5]};
    const Gate fallback{ws->resid, threshold * threshold};

    // predictor push with prev fields (lcode.py:843-854)
    if ((rc = do_push(ws, cfg, c->m, c->q, c->x_init, c->y_init,
                      prev->x_offt, prev->y_offt, nullptr, nullptr,
                      prev->px, prev->py, prev->pz, prevF, ws->p1, stream)))
        return rc;
    // the guess replaces the predictor's deposition and field solve
    GuessArgs ga{};
    for (int t = 0; t < guess->nterms; ++t) {
        ga.c[t] = guess->coef[t];
        for (int f = 0; f < 6; ++f) ga.h[t][f] = guess->f[t][f];
    }
    ga.nt = guess->nterms;
    for (int f = 0; f < 6; ++f) {
        ga.prev[f] = prevF[f];
        ga.guess[f] = ws->guess[f];
        ga.avg[f] = ws->favg[f];
    }
    ga.acc_zero = ws->resid;
    ga.n = cfg->grid_steps * cfg->grid_steps;
    ga.grid32 = ws->grid32;
    if ((rc = launch_guess(ga, stream, &ws->prof))) return rc;
    // corrector (lcode.py:884-911) on the guess, residual accumulated
    if ((rc = do_push(ws, cfg, c->m, c->q, c->x_init, c->y_init,
                      prev->x_offt, prev->y_offt, ws->p1[0], ws->p1[1],
                      prev->px, prev->py, prev->pz, avg, ws->p2, stream)))
        return rc;
    if ((rc = do_deposit(ws, cfg, c->ro_initial, ws->p2[0], ws->p2[1], c->m,
                         c->q, ws->p2[2], ws->p2[3], ws->p2[4], vp, ws->dep,
                         stream, Gate{nullptr, 0.}, true)))
        return rc;
    if ((rc = field_pass(ws, cfg, prev, trick_avg, beam_ro, beam_scalar,
                         ws->dep, outF, ws->favg, stream, Gate{nullptr, 0.},
                         g6, ws->resid)))
        return rc;
    // fallback: the plain predictor and corrector passes, gated on the
    // residual (lcode.py:856-911).  The predictor particles' displacement
    // bound was consumed by the corrector deposition above, so their
    // deposition uses the full static halo.
    if ((rc = do_deposit(ws, cfg, c->ro_initial, ws->p1[0], ws->p1[1], c->m,
                         c->q, ws->p1[2], ws->p1[3], ws->p1[4], vp, ws->dep,
                         stream, fallback)))
        return rc;
    // (the predictor's fields replace the rejected guess in the guess grids)
    double* const keepF[6] = {ws->guess[0], ws->guess[1], ws->guess[2],
                              ws->guess[3], ws->guess[4], ws->guess[5]};
    if ((rc = field_pass(ws, cfg, prev, trick_prev, beam_ro, beam_scalar,
                         ws->dep, keepF, ws->favg, stream, fallback)))
        return rc;
    if ((rc = do_push(ws, cfg, c->m, c->q, c->x_init, c->y_init,
                      prev->x_offt, prev->y_offt, ws->p1[0], ws->p1[1],
                      prev->px, prev->py, prev->pz, avg, ws->p2, stream,
                      fallback)))
        return rc;
    if ((rc = do_deposit(ws, cfg, c->ro_initial, ws->p2[0], ws->p2[1], c->m,
                         c->q, ws->p2[2], ws->p2[3], ws->p2[4], vp, ws->dep,
                         stream, fallback, true)))
        return rc;
    if ((rc = field_pass(ws, cfg, prev, trick_avg, beam_ro, beam_scalar,
                         ws->dep, outF, ws->favg, stream, fallback)))
        return rc;
    // final push and deposition (lcode.py:914-921)
    double* const outP[5] = {out->x_offt, out->y_offt, out->px, out->py,
                             out->pz};
    if ((rc = do_push(ws, cfg, c->m, c->q, c->x_init, c->y_init,
                      prev->x_offt, prev->y_offt, ws->p2[0], ws->p2[1],
                      prev->px, prev->py, prev->pz, avg, outP, stream)))
        return rc;
    double* const outD[4] = {out->ro, out->jx, out->jy, out->jz};
    return do_deposit(ws, cfg, c->ro_initial, out->x_offt, out->y_offt, c->m,
                      c->q, out->px, out->py, out->pz, vp, outD, stream,
                      Gate{nullptr, 0.}, true);
}

// ws->grid32 for the duration of one call
struct Grid32Scope {
    lcode_workspace* ws;
    Grid32Scope(lcode_workspace* w, int on) : ws(w) { if (ws) ws->grid32 = on; }
    ~Grid32Scope() { if (ws) ws->grid32 = 0; }
};

// A mixed state as the FP64-typed struct the shared composition takes
// (the grid pointers keep pointing to floats; ws->grid32 says so).
static lcode_state as_state(const lcode_state_mixed* m) {
    lcode_state s{};
    s.x_offt = m->x_offt; s.y_offt = m->y_offt;
    s.px = m->px; s.py = m->py; s.pz = m->pz;
    s.Ex = reinterpret_cast<double*>(m->Ex); s.Ey = reinterpret_cast<double*>(m->Ey);
    s.Ez = reinterpret_cast<double*>(m->Ez); s.Bx = reinterpret_cast<double*>(m->Bx);
    s.By = reinterpret_cast<double*>(m->By); s.Bz = reinterpret_cast<double*>(m->Bz);
    s.ro = reinterpret_cast<double*>(m->ro); s.jx = reinterpret_cast<double*>(m->jx);
    s.jy = reinterpret_cast<double*>(m->jy); s.jz = reinterpret_cast<double*>(m->jz);
    return s;
}

extern "C" {

int lcode_step(lcode_workspace* ws, const lcode_config* cfg,
               const lcode_const* c, const lcode_virt_params* vp,
               const lcode_state* prev, const double* beam_ro,
               double beam_scalar, const lcode_state* out,
               hipStream_t stream) {
    LCODE_REQUIRE(cfg && cfg->solver != LCODE_SOLVER_MIXED,
                  "LCODE_SOLVER_MIXED takes FP32 grids: use lcode_step_mixed");
    Grid32Scope g(ws, 0);
    return step_impl(ws, cfg, c, vp, prev, beam_ro, beam_scalar, out, stream);
}

int lcode_step_guided(lcode_workspace* ws, const lcode_config* cfg,
                      const lcode_const* c, const lcode_virt_params* vp,
                      const lcode_state* prev, const double* beam_ro,
                      double beam_scalar, const lcode_guess* guess,
                      double threshold, const lcode_state* out,
                      hipStream_t stream) {
    LCODE_REQUIRE(cfg && cfg->solver != LCODE_SOLVER_MIXED,
                  "LCODE_SOLVER_MIXED takes FP32 grids: use lcode_step_guided_mixed");
    Grid32Scope g(ws, 0);
    return guided_impl(ws, cfg, c, vp, prev, beam_ro, beam_scalar, guess,
                       threshold, out, stream);
}

int lcode_step_mixed(lcode_workspace* ws, const lcode_config* cfg,
                     const lcode_const* c, const lcode_virt_params* vp,
                     const lcode_state_mixed* prev, const double* beam_ro,
                     double beam_scalar, const lcode_state_mixed* out,
                     hipStream_t stream) {
    LCODE_REQUIRE(cfg && cfg->solver == LCODE_SOLVER_MIXED,
                  "lcode_step_mixed needs cfg->solver == LCODE_SOLVER_MIXED");
    LCODE_REQUIRE(prev && out, "NULL state");
    const lcode_state p = as_state(prev), o = as_state(out);
    Grid32Scope g(ws, 1);
    return step_impl(ws, cfg, c, vp, &p, beam_ro, beam_scalar, &o, stream);
}

int lcode_step_guided_mixed(lcode_workspace* ws, const lcode_config* cfg,
                            const lcode_const* c, const lcode_virt_params* vp,
                            const lcode_state_mixed* prev, const double* beam_ro,
                            double beam_scalar, const lcode_guess_mixed* guess,
                            double threshold, const lcode_state_mixed* out,
                            hipStream_t stream) {
    LCODE_REQUIRE(cfg && cfg->solver == LCODE_SOLVER_MIXED,
                  "lcode_step_guided_mixed needs cfg->solver == LCODE_SOLVER_MIXED");
    LCODE_REQUIRE(prev && out && guess, "NULL state or guess");
    const lcode_state p = as_state(prev), o = as_state(out);
    lcode_guess gd{};
    gd.nterms = guess->nterms;
    for (int t = 0; t < 4; ++t) {
        gd.coef[t] = guess->coef[t];
        for (int f = 0; f < 6; ++f)
            gd.f[t][f] = reinterpret_cast<const double*>(guess->f[t][f]);
    }
    Grid32Scope g(ws, 1);
    return guided_impl(ws, cfg, c, vp, &p, beam_ro, beam_scalar, &gd, threshold,
                       &o, stream);
}

int lcode_workspace_keep_predictor_fields(lcode_workspace* ws, int32_t keep) {
    LCODE_REQUIRE(ws, "NULL workspace");
    ws->keep_f1 = keep != 0;
    return LCODE_OK;
}

int lcode_workspace_predictor_fields(lcode_workspace* ws, void* fields[6]) {
    LCODE_REQUIRE(ws && fields, "NULL argument");
    for (int f = 0; f < 6; ++f) fields[f] = ws->guess[f];
    return LCODE_OK;
}

int lcode_step_guided_residual(lcode_workspace* ws, hipStream_t stream,
                               double* residual) {
    LCODE_REQUIRE(ws && residual, "NULL argument");
    double host[2] = {0., 0.};
    LCODE_HIP(hipMemcpyAsync(host, ws->resid, sizeof(host),
                             hipMemcpyDeviceToHost, stream));
    LCODE_HIP(hipStreamSynchronize(stream));
    *residual = host[1] > 0 ? std::sqrt(host[0] / host[1])
                            : (host[0] > 0 ? INFINITY : 0.);
    return LCODE_OK;
}

int lcode_step_guided_residual_async(lcode_workspace* ws, double* dst,
                                     hipStream_t stream) {
    LCODE_REQUIRE(ws && dst, "NULL argument");
    LCODE_HIP(hipMemcpyAsync(dst, ws->resid, 2 * sizeof(double),
                             hipMemcpyDefault, stream));
    return LCODE_OK;
}

}  // extern "C"
