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
 *   lcode_calculate_Ex_Ey_Bx_By     - calculate_Ex_Ey_Bx_By (lcode.py:246)
 *   lcode_calculate_Ez              - calculate_Ez (lcode.py:160)
 *   lcode_calculate_Bz              - calculate_Bz (lcode.py:333)
 *   lcode_step                      - step (lcode.py:828)
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
    LCODE_SOLVER_DENSE_BF16X3 = 3  /* the same matrices on BF16 MFMA with
                                      hi/lo splits of both operands, three
                                      products, FP32 accumulation; RHS,
                                      tables, Bz mean, Variant A and
                                      averages stay FP64 */
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
 * Between begin and end every kernel the library launches is bracketed by
 * HIP events on its stream; end waits for the stream and returns, per
 * kernel name, the launch count and the summed event time.  max_launches
 * bounds the number of timed launches (later ones run untimed). */
typedef struct {
    char name[32];
    int64_t calls;
    double total_ms;
} lcode_kernel_time;

int lcode_profile_begin(lcode_workspace *ws, int32_t max_launches);
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
 * harnesses that forbid hipMalloc).  Same arguments, kernels and results as
 * lcode_deposit, except that clamped lookups are not counted and a
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
 * fields Ex Ey Ez Bx By Bz of up to three N*N field sets (e.g. a predictor
 * network's output with coef {1}, or the quadratic extrapolation
 * 3 F(n-1) - 3 F(n-2) + F(n-3) with coef {3, -3, 1}). */
typedef struct {
    const double *f[3][6];
    double coef[3];
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

/* Residual of the last guided step on this workspace (synchronises). */
int lcode_step_guided_residual(lcode_workspace *ws, hipStream_t stream,
                               double *residual);

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
 * Learned predictor kernels.  No counterpart in the reference: they train
 * and run a small convolutional network whose output is a field guess for
 * lcode_step_guided (the learned form of the predictor pass,
 * lcode.py:841-881).  FP32 planes [C][H][W]; 3 x 3 convolutions with zero
 * padding; weights w[cout][cin][3][3], biases b[cout]; 1 <= cin, cout <= 16.
 * -------------------------------------------------------------------------*/

/* out = conv(in, w) + b, then max(., 0) when relu != 0. */
int lcode_pred_conv3x3_forward(const float *in, int32_t cin, const float *w,
                               const float *b, float *out, int32_t cout,
                               int32_t H, int32_t W, int32_t relu,
                               hipStream_t stream);

/* Gradient with respect to the layer input; when relu_input is not NULL
 * (the layer's input came out of a ReLU) it is zeroed where relu_input <= 0. */
int lcode_pred_conv3x3_backward_input(const float *grad_out, int32_t cout,
                                      const float *w, float *grad_in,
                                      int32_t cin, const float *relu_input,
                                      int32_t H, int32_t W, hipStream_t stream);

/* Gradients of w and b, written as grad_wb[cout*cin*9 + cout] (weights,
 * then biases); `scratch` holds lcode_pred_backward_weight_scratch floats. */
int lcode_pred_conv3x3_backward_weight(const float *in, int32_t cin,
                                       const float *grad_out, int32_t cout,
                                       float *scratch, float *grad_wb,
                                       int32_t H, int32_t W, hipStream_t stream);
int64_t lcode_pred_backward_weight_scratch(int32_t cin, int32_t cout,
                                           int32_t H, int32_t W);

/* g = 2 (y - t) / n and *loss_sum = sum (y - t)^2 (device float);
 * `scratch` holds (n + 255) / 256 floats. */
int lcode_pred_mse_grad(const float *y, const float *t, int64_t n, float *g,
                        float *scratch, float *loss_sum, hipStream_t stream);

/* One Adam step (step number t >= 1) on n parameters. */
int lcode_pred_adam(float *p, const float *g, float *m, float *v, int32_t n,
                    float lr, float beta1, float beta2, float eps, int32_t t,
                    hipStream_t stream);

/* FP64 fields <-> the network's FP32 planes.  With Q = 3 F(n-1) - 3 F(n-2)
 * + F(n-3) (the quadratic extrapolation of slice n) and D = F(n-1) - F(n-2):
 * inputs: planes[14][N][N] = Q/s (6), D/s_d (6), beam/s_b, ro(n-1)/s_r;
 * targets: planes[6][N][N] = (F(n) - Q) / t; guess: guess[f] = Q + t_f pred[f]. */
typedef struct {
    const double *f1[6];      /* fields of slice n-1 (Ex Ey Ez Bx By Bz) */
    const double *f2[6];      /* fields of slice n-2 */
    const double *f3[6];      /* fields of slice n-3 */
    const double *beam;       /* beam of slice n (NULL: beam_scalar) */
    double beam_scalar;
    const double *ro;         /* plasma density of slice n-1 */
    const double *next[6];    /* targets only: fields of slice n */
    float scale[14];          /* input scales */
    float tscale[6];          /* target scales */
    float *planes;            /* inputs / targets output */
    const float *pred;        /* guess only: network output [6][N][N] */
    double *guess[6];         /* guess only: output fields */
    int32_t n;                /* N * N */
} lcode_pred_io;

int lcode_pred_inputs(const lcode_pred_io *io, hipStream_t stream);
int lcode_pred_targets(const lcode_pred_io *io, hipStream_t stream);
int lcode_pred_guess(const lcode_pred_io *io, hipStream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* LCODE3D_HIP_H */
