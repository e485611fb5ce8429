/*
 * lcode3d_hip.h - C ABI of the MI355X (gfx950) implementation of LCODE 3D's
 * per-xi-slice step.
 *
 * Every entry point takes plain device pointers (FP64 arrays in C order,
 * int64 index arrays), sizes in a config struct and a HIP stream, and
 * returns an int status: LCODE_OK (0) or a negative LCODE_E_* code, with a
 * human-readable message from lcode_last_error().  Nothing here allocates
 * the caller's outputs: the caller passes them in.  All work is enqueued on
 * `stream`; no entry point synchronises the device except
 * lcode_workspace_fault_count().
 *
 * Interface map onto the reference (lotov/lcode3d, lcode.py):
 *   lcode_workspace_create/destroy  - the cp.memoize'd eigenvalue tables
 *                                      (lcode.py:144, :215, :316) plus the
 *                                      scratch the fused step needs
 *   lcode_move_estimate_wo_fields   - move_estimate_wo_fields (lcode.py:369)
 *   lcode_move_smart                - move_smart + move_smart_kernel
 *                                      (lcode.py:796, :704)
 *   lcode_deposit                   - deposit + deposit_kernel
 *                                      (lcode.py:666, :626)
 *   lcode_plasma_sizes/make_plasma  - make_plasma (lcode.py:490)
 *   lcode_calculate_Ex_Ey_Bx_By     - calculate_Ex_Ey_Bx_By (lcode.py:246)
 *   lcode_calculate_Ez              - calculate_Ez (lcode.py:160)
 *   lcode_calculate_Bz              - calculate_Bz (lcode.py:333)
 *   lcode_step                      - step (lcode.py:828)
 *   lcode_step_mixed                - the same step in mixed precision
 *   lcode_step_guided(_residual)    - step with its predictor pass
 *                                      (lcode.py:841-881) replaced by a
 *                                      supplied field guess
 *   lcode_diag_zn                   - diags_ro_zn (lcode.py:981)
 *
 * Grids are (N, N) with [i, j] <-> (x, y) = ((i - N/2) h, (j - N/2) h)
 * (lcode.py:951-953).  Coarse particle arrays hold nc*nc values, the
 * virt_params arrays nf values each (lcode.py:566-570).
 *
 * Threading: a workspace may be used by one stream at a time; calls on the
 * same workspace from two streams race on its scratch buffers.
 */
#ifndef LCODE3D_HIP_H
#define LCODE3D_HIP_H

#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    LCODE_OK = 0,
    LCODE_E_INVALID = -1,   /* bad argument or config */
    LCODE_E_HIP = -2,       /* a HIP runtime call failed */
    LCODE_E_UNSUPPORTED = -3 /* e.g. a transform length 2 (N - 1) with a
                                prime factor other than 2, 3 and 5 */
};

/* Transform back-end of the field solves. */
enum {
    LCODE_SOLVER_FFT64 = 0, /* hand-written mixed-radix FFT in LDS, FP64 */
    LCODE_SOLVER_FFT32 = 1, /* the same in FP32 (mixed precision): the
                               transforms, spectra and eigenvalue tables;
                               RHS, Bz mean and zero mode, Variant A and
                               averages stay FP64 */
    LCODE_SOLVER_DENSE64 = 2,      /* dense N x N sine/cosine matrices,
                                      FP64 MFMA (v_mfma_f64_16x16x4) */
    LCODE_SOLVER_DENSE_BF16X3 = 3, /* the same matrices on BF16 MFMA with
                                      hi/lo splits of both operands, three
                                      products, FP32 accumulation; RHS,
                                      tables, Bz mean, Variant A and
                                      averages stay FP64 */
    LCODE_SOLVER_MIXED = 4         /* the mixed-precision step
                                      (lcode_step_mixed): the ten N x N
                                      state grids (fields, densities,
                                      currents, averages, guess) stored in
                                      FP32 and the FFT32 transforms; FP64
                                      stays in the particle positions and
                                      momenta and their updates, the
                                      deposition arithmetic, the RHS
                                      arithmetic (trick term included), the
                                      Bz zero mode and mean, and the ion
                                      background ro_initial.  The standalone
                                      operators treat it as FFT32 on FP64
                                      grids. */
};

/* The config attributes the step reads (config_example.py:1-17), with
 * reflect_boundary as init() computes it (lcode.py:947-949). */
typedef struct {
    int32_t grid_steps;            /* N, odd */
    double  grid_step_size;        /* h */
    double  xi_step_size;          /* dxi */
    double  subtraction_trick;     /* RHS multiplier; the matrix uses != 0 */
    int32_t variant_A;             /* field_solver_variant_A */
    double  reflect_boundary;
    int32_t plasma_coarseness;
    int32_t plasma_fineness;
    int32_t nc;                    /* coarse particles per axis */
    int32_t nf;                    /* fine particles per axis */
    int32_t solver;                /* LCODE_SOLVER_* */
} lcode_config;

/* const of init() (lcode.py:964-965): nc*nc arrays, ro_initial N*N. */
typedef struct {
    const double *m, *q, *x_init, *y_init, *ro_initial;
} lcode_const;

/* virt_params of make_plasma() (lcode.py:566-570): nf entries each. */
typedef struct {
    const double  *influence_prev, *influence_next, *fine_grid;
    const int64_t *indices_prev, *indices_next;
} lcode_virt_params;

/* The state step() consumes and returns (lcode.py:926-929). */
typedef struct {
    double *x_offt, *y_offt, *px, *py, *pz;      /* nc*nc */
    double *Ex, *Ey, *Ez, *Bx, *By, *Bz;          /* N*N */
    double *ro, *jx, *jy, *jz;                    /* N*N */
} lcode_state;

/* The state of the mixed-precision step (LCODE_SOLVER_MIXED): FP64
 * particles, FP32 grids, the same 15 arrays as lcode_state. */
typedef struct {
    double *x_offt, *y_offt, *px, *py, *pz;      /* nc*nc, FP64 */
    float *Ex, *Ey, *Ez, *Bx, *By, *Bz;           /* N*N, FP32 */
    float *ro, *jx, *jy, *jz;                     /* N*N, FP32 */
} lcode_state_mixed;

typedef struct lcode_workspace lcode_workspace;

/* Last error message of the calling thread ("" if none). */
const char *lcode_last_error(void);

/* Library version string. */
const char *lcode_version(void);

/* Builds the eigenvalue and twiddle tables for cfg (N, h, trick, solver)
 * on the current device and allocates the step's scratch. */
int lcode_workspace_create(const lcode_config *cfg, lcode_workspace **out);
int lcode_workspace_destroy(lcode_workspace *ws);

/* Number of particle cell lookups that fell outside the grid (and were
 * clamped to it) since the last call; resets the counter.  The reference
 * has no such check: an escaped particle is a memory violation there
 * (docs/overview/window_and_grids.rst).  Synchronises `stream`. */
int lcode_workspace_fault_count(lcode_workspace *ws, hipStream_t stream,
                                int64_t *count);

/* Whether a particle with a non-finite (or more than 1e6 cells) offset or
 * momentum was met since the flag was last reset: 1 or 0 in *flag.  Does
 * not synchronise: it reads a host-visible flag that the pushes and the
 * deposition set, so a kernel still running has not reported yet (after a
 * stream synchronisation every finished kernel has).  When the flag is set
 * the depositions that met it wrote NaN grids instead of depositing, and
 * everything computed from them is NaN.  reset != 0 clears it.  The
 * reference has no such check: a NaN particle there deposits at an
 * undefined cell or faults. */
int lcode_workspace_nonfinite(lcode_workspace *ws, int32_t reset,
                              int32_t *flag);

/* Per-kernel timing (instrumentation; no counterpart in the reference).
 * Between begin and end every kernel the library launches is timed with a
 * pair of HIP events on its stream: for the step's kernels (push,
 * deposition, spill, the solve passes and the dense products) events
 * attached to the dispatch (hipExtLaunchKernel), which the runtime stamps
 * with the kernel's own start and end; for the others events recorded
 * before and after the launch.  end waits for the stream and returns, per
 * kernel name, the launch count and the summed event time.  max_launches
 * bounds the number of timed launches (later ones run untimed).
 * lcode_profile_only, called after begin, restricts the timing to the
 * launches of one kernel name (NULL or "": every kernel again).
 * lcode_profile_pause(ws, 1) stops timing until lcode_profile_pause(ws, 0)
 * (launches in between run untimed), so that a long loop can time a
 * sample of its launches. */
typedef struct {
    char name[32];
    int64_t calls;
    double total_ms;
} lcode_kernel_time;

int lcode_profile_begin(lcode_workspace *ws, int32_t max_launches);
int lcode_profile_only(lcode_workspace *ws, const char *kernel_name);
int lcode_profile_pause(lcode_workspace *ws, int32_t paused);
int lcode_profile_end(lcode_workspace *ws, hipStream_t stream,
                      lcode_kernel_time *out, int32_t max_out,
                      int32_t *n_out);

/* lcode.py:369-391.  n = nc*nc particles. */
int lcode_move_estimate_wo_fields(lcode_workspace *ws, const lcode_config *cfg,
                                  const double *m, const double *x_init,
                                  const double *y_init,
                                  const double *prev_x_offt,
                                  const double *prev_y_offt,
                                  const double *px, const double *py,
                                  const double *pz,
                                  double *x_offt_out, double *y_offt_out,
                                  hipStream_t stream);

/* lcode.py:796-823.  fields = {Ex, Ey, Ez, Bx, By, Bz}_avg (N*N each);
 * out = {x_offt, y_offt, px, py, pz}. */
int lcode_move_smart(lcode_workspace *ws, const lcode_config *cfg,
                     const double *m, const double *q,
                     const double *x_init, const double *y_init,
                     const double *x_prev_offt, const double *y_prev_offt,
                     const double *estimated_x_offt,
                     const double *estimated_y_offt,
                     const double *px_prev, const double *py_prev,
                     const double *pz_prev,
                     const double *const fields[6],
                     double *const out[5],
                     hipStream_t stream);

/* Sizes of make_plasma's arrays (lcode.py:452-487) for a plasma of `steps`
 * cells: nc coarse particles per axis, nf fine ones. */
int lcode_plasma_sizes(int32_t steps, int32_t coarseness, int32_t fineness,
                       int32_t *nc, int32_t *nf);

/* make_plasma (lcode.py:490-573) on the device: x_init, y_init, m, q of
 * the coarse particles (nc*nc each; offsets and momenta start at zero) and
 * virt_params (nf each): the same FP64 values as the reference's NumPy. */
int lcode_make_plasma(int32_t steps, double cell_size, int32_t coarseness,
                      int32_t fineness, double *x_init, double *y_init,
                      double *m, double *q, double *influence_prev,
                      double *influence_next, double *fine_grid,
                      int64_t *indices_prev, int64_t *indices_next,
                      hipStream_t stream);

/* lcode.py:666-689.  ro_initial may be NULL (the reference's scalar 0 in
 * initial_deposition, lcode.py:697).  out = {ro, jx, jy, jz}. */
int lcode_deposit(lcode_workspace *ws, const lcode_config *cfg,
                  const double *ro_initial,
                  const double *x_offt, const double *y_offt,
                  const double *m, const double *q,
                  const double *px, const double *py, const double *pz,
                  const lcode_virt_params *virt_params,
                  double *const out[4], hipStream_t stream);

/* lcode_deposit without a workspace and without any allocation.  No
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments and kernels as
 * lcode_deposit and the same results up to the order of the floating-point
 * sums (it takes every tile with one block, where lcode_deposit may split
 * tiles over two), except that clamped lookups are not counted and a
 * non-finite particle is not flagged to the host (its tile is still written
 * as NaN).  Its scratch - the spill list of particles displaced beyond the
 * deposition tiles' halo and its counters - is a fixed device array of
 * 2048^2 + 64 ints (16 MiB) per device, held until the process ends, so
 * cfg->nf must be at most 2048 (LCODE_E_INVALID otherwise).  Every call on
 * a device shares that array: calls that overlap on one device (two streams
 * or two host threads) race and give wrong results.  Each call clears the
 * counters it needs with a small kernel (no memset node).  Launches a
 * zero-fill kernel, the tile kernel and the spill kernel on `stream`. */
int lcode_deposit_noalloc(const lcode_config *cfg, const double *ro_initial,
                          const double *x_offt, const double *y_offt,
                          const double *m, const double *q,
                          const double *px, const double *py,
                          const double *pz,
                          const lcode_virt_params *virt_params,
                          double *const out[4], hipStream_t stream);

/* lcode.py:246-290.  beam_ro is an N*N array or NULL, in which case the
 * scalar beam_scalar is used everywhere (the reference's beam returns a
 * scalar 0 outside the bunch, config_example.py:24-25). out = {Ex,Ey,Bx,By}.*/
int lcode_calculate_Ex_Ey_Bx_By(lcode_workspace *ws, const lcode_config *cfg,
                                const double *Ex_avg, const double *Ey_avg,
                                const double *Bx_avg, const double *By_avg,
                                const double *beam_ro, double beam_scalar,
                                const double *ro, const double *jx,
                                const double *jy, const double *jz,
                                const double *jx_prev, const double *jy_prev,
                                double *const out[4], hipStream_t stream);

/* lcode.py:160-184. */
int lcode_calculate_Ez(lcode_workspace *ws, const lcode_config *cfg,
                       const double *jx, const double *jy, double *Ez,
                       hipStream_t stream);

/* lcode.py:333-364. */
int lcode_calculate_Bz(lcode_workspace *ws, const lcode_config *cfg,
                       const double *jx, const double *jy, double *Bz,
                       hipStream_t stream);

/* lcode.py:828-931.  `prev` is read, `out` written; they must not alias.
 * beam_ro / beam_scalar as in lcode_calculate_Ex_Ey_Bx_By. */
int lcode_step(lcode_workspace *ws, const lcode_config *cfg,
               const lcode_const *cnst, const lcode_virt_params *virt_params,
               const lcode_state *prev, const double *beam_ro,
               double beam_scalar, const lcode_state *out,
               hipStream_t stream);

/* A field guess for the guided step: guess = sum_t coef[t] f[t], the six
 * fields Ex Ey Ez Bx By Bz of up to four N*N field sets (e.g. a predictor
 * network's output with coef {1}, the quadratic extrapolation
 * 3 F(n-1) - 3 F(n-2) + F(n-3) with coef {3, -3, 1}, or a term taken from
 * lcode_workspace_predictor_fields). */
typedef struct {
    const double *f[4][6];
    double coef[4];
    int32_t nterms;
} lcode_guess;

/* The predictor hook of lcode.py:841-881: the predictor pass's deposition
 * and field solve are replaced by `guess`.  The step pushes with prev
 * fields (as the predictor does), then runs the corrector pass with the
 * averages (guess + prev)/2, and measures the residual
 * ||F_corr - guess|| / ||F_corr|| over the six fields on the device.  When
 * it exceeds `threshold` (or is not a number) the plain predictor and
 * corrector passes run after all; the final push and deposition follow
 * either way.  No host synchronisation: the fallback kernels are always
 * enqueued and return at once when the residual is small. */
int lcode_step_guided(lcode_workspace *ws, const lcode_config *cfg,
                      const lcode_const *cnst,
                      const lcode_virt_params *virt_params,
                      const lcode_state *prev, const double *beam_ro,
                      double beam_scalar, const lcode_guess *guess,
                      double threshold, const lcode_state *out,
                      hipStream_t stream);

/* lcode_step in the mixed precision of LCODE_SOLVER_MIXED (cfg->solver
 * must be LCODE_SOLVER_MIXED; the const arrays, ro_initial included, stay
 * FP64): the same passes, with the state's grids, the deposition and the
 * transforms in FP32. */
int lcode_step_mixed(lcode_workspace *ws, const lcode_config *cfg,
                     const lcode_const *cnst,
                     const lcode_virt_params *virt_params,
                     const lcode_state_mixed *prev, const double *beam_ro,
                     double beam_scalar, const lcode_state_mixed *out,
                     hipStream_t stream);

/* A field guess over FP32 field sets (the mixed step's fields). */
typedef struct {
    const float *f[4][6];
    double coef[4];
    int32_t nterms;
} lcode_guess_mixed;

/* lcode_step_guided in the mixed precision of LCODE_SOLVER_MIXED. */
int lcode_step_guided_mixed(lcode_workspace *ws, const lcode_config *cfg,
                            const lcode_const *cnst,
                            const lcode_virt_params *virt_params,
                            const lcode_state_mixed *prev,
                            const double *beam_ro, double beam_scalar,
                            const lcode_guess_mixed *guess, double threshold,
                            const lcode_state_mixed *out, hipStream_t stream);

/* Residual of the last guided step on this workspace (synchronises). */
int lcode_step_guided_residual(lcode_workspace *ws, hipStream_t stream,
                               double *residual);

/* The same without synchronising: enqueues on `stream` a copy of the last
 * guided step's two sums {sum (F_corr - guess)^2, sum F_corr^2} to dst
 * (2 doubles, device or pinned host memory); the residual is
 * sqrt(dst[0] / dst[1]) once the stream has reached the copy. */
int lcode_step_guided_residual_async(lcode_workspace *ws, double *dst,
                                     hipStream_t stream);

/* The predictor pass's fields of the last step on this workspace, kept for
 * a guess of the next slice (no counterpart in the reference; the
 * reference's predictor pass is lcode.py:841-874).  With keep != 0 every
 * plain step (lcode_step, lcode_step_mixed) writes its predictor-pass
 * fields F1 (after Variant A) into six workspace grids; a guided step
 * leaves its guess there, or F1 when its guess was rejected, in either
 * case whether or not keep is set.  lcode_workspace_predictor_fields
 * returns the six device pointers (FP64 grids, or FP32 ones after a mixed
 * step), valid until the workspace is destroyed; they may be passed as a
 * term of the next guess, an lcode_guess: the guess kernel reads each
 * element before it overwrites it.  A guess of the next slice's F1 from them: the
 * extrapolation of the corrector's fields minus the last slice's one-pass
 * change F - F1. */
int lcode_workspace_keep_predictor_fields(lcode_workspace *ws, int32_t keep);
int lcode_workspace_predictor_fields(lcode_workspace *ws, void *fields[6]);

/* Plasma-noise diagnostic of one density grid (lcode.py:980-989,
 * diags_ro_zn before its running maximum): zn = mean |ro - G ro| /
 * 4.23045376e-04, where G is the Gaussian blur of sigma =
 * 0.25 / grid_step_size cells with scipy.ndimage.gaussian_filter's
 * conventions (taps truncated at 4 sigma, half-sample 'reflect'
 * boundaries, axis 0 then axis 1).  `ro` is an N*N device grid; the value
 * is written to the device double *zn_out; no host synchronisation. */
int lcode_diag_zn(lcode_workspace *ws, const lcode_config *cfg,
                  const double *ro, double *zn_out, hipStream_t stream);



/* ---------------------------------------------------------------------------
 * Learned predictor: a 4-level convolutional encoder-decoder with 32
 * channels per level.  No counterpart in the reference: it trains on slices
 * the plain step generates and its output is a field guess for
 * lcode_step_guided (the learned form of the predictor pass,
 * lcode.py:841-881).  Level 0 has S = (N + 3) / 4 pixels per side (the
 * input planes are 4 x 4 means of the grids), levels 1..3 have
 * S_{l+1} = (S_l + 1) / 2:
 *   e0 = relu(C0(x))  e1 = relu(C1(pool e0))  e2 = relu(C2(pool e1))
 *   e3 = relu(C3(pool e2))  d2 = relu(C4[up e3, e2])  d1 = relu(C5[up d2, e1])
 *   d0 = relu(C6[up d1, e0])  y = C7(d0)
 * Ck: 3 x 3 convolutions, zero padding, bias; channels 14->32, 32->32 (x3),
 * 64->32 (x3), 32->6.  pool: 2 x 2 mean (zeros beyond the edge); up:
 * nearest neighbour.  BF16 weights (packed from FP32 master copies) and
 * activations on v_mfma_f32_16x16x32_bf16, FP32 accumulation.  The caller
 * owns every buffer (sizes from lcode_unet_sizes).
 * -------------------------------------------------------------------------*/
typedef struct {
    int32_t S;                 /* level-0 size */
    float *params;             /* FP32 master parameters: per layer
                                  w[cout][cin][3][3] then b[cout] (sizes[0]) */
    float *grad;               /* their gradient (sizes[0]) */
    float *adam_m, *adam_v;    /* Adam moments (sizes[0] each) */
    uint16_t *wpack;           /* packed bfloat16 weights (sizes[1]) */
    uint16_t *act;             /* activations and gradients, bfloat16
                                  [s][s][32] tensors (sizes[2]) */
    float *y;                  /* network output [S][S][16], channels 0..5 */
    float *partial;            /* weight-gradient partials (sizes[3]) */
    float *loss_part;          /* loss partial sums (sizes[4]) */
    float *loss;               /* 1 float: the last training loss */
} lcode_unet;

/* Element counts of the buffers: params, wpack, act, partial, loss_part. */
int lcode_unet_sizes(int32_t S, int64_t sizes[5]);
/* Element offset of activation/gradient tensor `index` in act, and of
 * layer `layer`'s parameters in params (index/layer == count: the total). */
int64_t lcode_unet_tensor_offset(int32_t S, int32_t index);
int64_t lcode_unet_param_offset(int32_t layer);

/* params -> wpack (after the caller writes params). */
int lcode_unet_pack(const lcode_unet *net, hipStream_t stream);
/* y = net(x); x = input planes [S][S][32] bfloat16 (lcode_unet_io 0). */
int lcode_unet_forward(const lcode_unet *net, const uint16_t *x,
                       hipStream_t stream);
/* The same y without keeping the pooled and upsampled copies that
 * lcode_unet_backward needs (each convolution pools or upsamples its input
 * while staging it): the inference form. */
int lcode_unet_infer(const lcode_unet *net, const uint16_t *x,
                     hipStream_t stream);
/* After lcode_unet_forward(x): the mean squared error against targets
 * [S][S][6] (FP32, into *loss) and its gradient into grad. */
int lcode_unet_backward(const lcode_unet *net, const uint16_t *x,
                        const float *targets, hipStream_t stream);
/* One Adam step t >= 1 on params from grad, then lcode_unet_pack. */
int lcode_unet_adam(const lcode_unet *net, int32_t t, float lr,
                    hipStream_t stream);

/* Planes of the network from the step's fields (FP64, or FP32 with
 * f32 != 0).  which = 0: input planes[S][S][32] = 4 x 4 means of
 * F(n-1)/scale (6), (F(n-1) - F(n-2))/scale (6), beam(n)/scale,
 * ro(n-1)/scale, zeros; 1: targets[S][S][6] = 4 x 4 means of
 * (next - Q)/tscale; 2: guess = Q + tscale y(i / 4, j / 4) on the grid
 * (after lcode_unet_forward).  Q = 3 F(n-1) - 3 F(n-2) + F(n-3), or with
 * f4 set (slice n-4) the Variant-A-aware 2 F(n-1) - 2 F(n-3) + F(n-4). */
typedef struct {
    const void *f1[6], *f2[6], *f3[6];   /* slices n-1, n-2, n-3 */
    const void *f4[6];                   /* slice n-4, or NULLs */
    const void *next[6];                 /* targets: the fields to learn */
    const void *ro;                      /* plasma density of slice n-1 */
    const double *beam;                  /* beam of slice n, or NULL */
    double beam_scalar;
    int32_t f32;                         /* grids are float arrays */
    int32_t grid_steps;                  /* N; S = (N + 3) / 4 */
    float scale[14], tscale[6];
    uint16_t *planes;                    /* which 0 */
    float *planes_f32;                   /* which 0 instead of planes: FP32
                                            [S][S][16] (14 used) */
    float *targets;                      /* which 1 */
    void *guess[6];                      /* which 2 (the grids' type) */
} lcode_unet_io_args;

int lcode_unet_io(const lcode_unet *net, const lcode_unet_io_args *io,
                  int32_t which, hipStream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* LCODE3D_HIP_H */
