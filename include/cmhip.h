/*
 * cmhip.h -- C ABI of the gfx950 contact solver (libcmhip.so).
 *
 * Plain C: device pointers are `void*` (HIP device memory, e.g. the data
 * pointer of a torch-ROCm tensor), sizes are ints, the stream is a
 * `hipStream_t` passed as `void*` (NULL = the context's own stream).  No torch
 * types appear here.  Every entry point returns an `int` status: 0 (or
 * CM_CONVERGED) on success, a negative CM_E_* code on error; the message of the
 * last error is available from cm_last_error().
 *
 * Array layout: real fields on the computational grid are C-order (Mx, My)
 * with My contiguous: Mx = nx, My = ny for periodic substrates and
 * Mx = 2 nx, My = 2 ny for free ones (the topography occupies the top-left
 * nx x ny block).  Heights are C-order (nx, ny) FP64.  Forces are per-pixel
 * forces (not pressures) with compressive force negative, as in the reference.
 *
 * Precision: CM_FP64 keeps every field and transform in double.  CM_FP32
 * keeps forces, conjugate directions and the transforms of the iteration in
 * float; the gap u - h is kept in double, updated by the FP32 transform of
 * the force change and recomputed with FP64 transforms every refresh_every
 * iterations; all global reductions, the CG scalars and the convergence
 * logic run in double.
 */
#ifndef CMHIP_H
#define CMHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ----------------------------------------------------- */
#define CM_OK 0
#define CM_CONVERGED 0
#define CM_MAXITER 1          /* reference: success=False, "Reached maxiter = N" */
#define CM_RUNNING 2          /* progress callback only: solve still iterating */
#define CM_E_ARG (-1)         /* bad argument (reference: ValueError) */
#define CM_E_SHAPE (-2)       /* shape/size mismatch (reference: Substrate.Error) */
#define CM_E_NAN (-3)         /* reference: RuntimeError('nan encountered.') */
#define CM_E_OOM (-4)         /* device memory budget exceeded */
#define CM_E_HIP (-5)         /* HIP runtime error */
#define CM_E_UNSUPPORTED (-6) /* grid not supported (odd My, or a line too long for one block) */
#define CM_E_BISECT (-7)      /* bearing-area bisection failed (reference: scipy bisect raising) */

#define CM_PERIODIC 0
#define CM_FREE 1
#define CM_SEMI 2
#define CM_FP64 0
#define CM_FP32 1

typedef struct cm_ctx cm_ctx;
typedef struct cm_plan cm_plan;
typedef struct cm_session cm_session;

/* Substrate description.  Replaces the constructor arguments of
 * PeriodicFFTElasticHalfSpace (FFTElasticHalfSpace.py:169-306),
 * FreeFFTElasticHalfSpace (FFTElasticHalfSpace.py:912-964) and
 * SemiPeriodicFFTElasticHalfSpace (FFTElasticHalfSpace.py:1286-1352).
 * Computational grid: periodic nx x ny; free 2nx x 2ny; semi-periodic n along
 * a periodic direction, 2n - 1 along a free one.  My must be even; the host
 * layer computes other grids on the transposed grid or (periodic y) on two
 * periods with tile_y = 2. */
typedef struct cm_substrate_desc {
  int nx, ny;              /* topography grid */
  int bc;                  /* CM_PERIODIC, CM_FREE or CM_SEMI */
  int precision;           /* CM_FP64 or CM_FP32 */
  double sx, sy;           /* physical sizes */
  double young;            /* Young's modulus (contact modulus if poisson == 0) */
  double poisson;          /* Poisson ratio (periodic only) */
  int q0_mode;             /* periodic: 0 = stiffness_q0 None, 1 = stiffness_q0 == 0, 2 = given value */
  double stiffness_q0;     /* used when q0_mode == 2 */
  double thickness;        /* periodic finite thickness; <= 0 or NaN = infinite half-space */
  int periodic_x, periodic_y; /* CM_SEMI: periodicity per direction (reference default 0, 1) */
  int n_images;            /* CM_SEMI: periodic images per side in the kernel sum (reference default 10) */
  int tile_y;              /* 0 or 1; 2: ny (and sy) hold two periods along a periodic y, and the
                              Green's function is that of one period (zero at odd ky) */
} cm_substrate_desc;

/* Keyword arguments of constrained_conjugate_gradients
 * (ConstrainedConjugateGradients.py:47-60). */
typedef struct cm_ccg_params {
  int load_control;        /* 1: external_force given, 0: offset given */
  double external_force;
  double offset;
  int has_hardness;        /* hardness is not None */
  double hardness;         /* per-pixel force cap (pressure * area_per_pt) */
  double pentol, forcetol, thermotol;
  double mixfac, mixdecfac;
  int minmixsteps;
  int maxiter;
  int init_mode;           /* 0: forces from initial displacements, 1: initial forces given */
  int iters_per_sync;      /* 0 = adaptive; 1 = host sees every iteration */
  int use_graph;           /* 1 = replay chunks of iterations from a captured hipGraph */
  int refresh_every;       /* FP32 plans: the displacements are recomputed from the forces with FP64
                              transforms every refresh_every iterations (0 = default 8) and updated
                              by FP32 transforms of the force change in between */
} cm_ccg_params;

/* Per-solve statistics; the scalar fields of the reference OptimizeResult
 * and of its per-iteration log (ConstrainedConjugateGradients.py:436-509). */
typedef struct cm_ccg_stats {
  int status;              /* CM_CONVERGED, CM_MAXITER or CM_E_* */
  int nit, nfev;
  int last_mode;           /* 0 = 'reset', 1 = 'sd', 2 = 'cg', 3 = 'mix' (delta_str) */
  double offset, fun;
  double area, cg_area, plastic_area;
  double total_force, rms_pen, max_pen, max_pres, pad_pres, maxdu, tau, mixfac;
  double nc_count;
} cm_ccg_stats;

/* Called by cm_ccg_solve with the statistics of instance 0: with
 * iters_per_sync > 0, once per host synchronisation (the latest iteration);
 * with iters_per_sync == 0 (adaptive chunks), once for every iteration of
 * each chunk, from the device's log ring of the last 64 iterations. */
typedef void (*cm_progress_fn)(void* user, const cm_ccg_stats* stats);

const char* cm_last_error(cm_ctx* ctx);
const char* cm_version(void);

/* Context: device, default stream, memory budget (bytes, 0 = unlimited). */
int cm_ctx_create(int device, size_t mem_budget, cm_ctx** out);
int cm_ctx_destroy(cm_ctx* ctx);
size_t cm_ctx_mem_used(cm_ctx* ctx);
/* The context's own non-blocking stream (what a NULL stream argument means). */
int cm_ctx_stream(cm_ctx* ctx, void** stream);

/* Plan: Green's function, FFT twiddles and launch shapes for one substrate.
 * Replaces PeriodicFFTElasticHalfSpace._compute_greens_function /
 * _compute_surface_stiffness (FFTElasticHalfSpace.py:445-571) and
 * FreeFFTElasticHalfSpace._compute_greens_function (:1033-1139). */
int cm_plan_create(cm_ctx* ctx, const cm_substrate_desc* desc, cm_plan** out);
int cm_plan_destroy(cm_plan* plan);
/* Computational grid (Mx, My) of the plan. */
int cm_plan_grid(cm_plan* plan, int* Mx, int* My);
/* Copy the plan's real Green's function (FP64, C-order (My/2+1, Mx), already
 * including the -1/(area_per_pt * Mx * My) factor of evaluate_disp) to host. */
int cm_plan_green_host(cm_plan* plan, double* host_out);

/* Autotuner: time evaluate_disp over the candidate launch shapes of the FFT
 * passes (columns per column block, rows per row block, register elements per
 * thread) and keep the fastest; `out` receives a JSON description.  Must not
 * run while a session of this plan is in flight.  cm_plan_set_shape applies a
 * cached choice (0 = automatic), cm_plan_shape reads the current one.  Explicit
 * shapes exist for FP64 plans; FP32 plans keep the automatic shape and both
 * calls return CM_E_UNSUPPORTED for them.  Register elements are 8, 16 or 32.
 * No reference counterpart (muGrid picks FFTW plans internally). */
int cm_plan_autotune(cm_plan* plan, int reps, char* out, size_t out_len);
int cm_plan_set_shape(cm_plan* plan, int Cb, int R, int erow, int ecol);
int cm_plan_shape(cm_plan* plan, int* Cb, int* R, int* erow, int* ecol);

/* u = evaluate_disp(f) on the (Mx, My) grid (FFTElasticHalfSpace.py:573-589,
 * padded branch of 1141-1164).  f, u: device arrays of the plan precision;
 * in-place (u == f) is allowed. */
int cm_evaluate_disp(cm_plan* plan, const void* f, void* u, void* stream);
/* f = evaluate_force(u) (FFTElasticHalfSpace.py:591-611). */
int cm_evaluate_force(cm_plan* plan, const void* u, void* f, void* stream);

/* ---- k-space helpers (FFTElasticHalfSpace.py:613-885) -----------------
 * Fourier fields use the reference's layout: (Mx/2 + 1) x My complex, C
 * order, the real-to-complex axis along x (FFTElasticHalfSpace.py:119-128),
 * unnormalised forward transform; complex values are (re, im) pairs in the
 * plan precision.  The helpers share the plan's spectrum buffer and
 * reduction workspace: calls on one plan must not overlap. */
/* Forward transform of a real (Mx, My) field: kind 0 the plain DFT (the
 * fourier buffer of fftengine.fft); kind 1 evaluate_k_disp (:613-633),
 * G DFT(-f) / area_per_pt; kind 2 evaluate_k_force (:635-648),
 * -K DFT(u) area_per_pt. */
int cm_fourier(cm_plan* plan, const void* field, void* out_k, int kind, void* stream);
/* evaluate_k_force_k (:650-666): out = -K disp_k area_per_pt (in place allowed). */
int cm_k_force_k(cm_plan* plan, const void* disp_k, void* out_k, void* stream);
/* evaluate_scalar_product_k_space (:686-805): the Hermitian-weighted sum of
 * Re(conj(ka) kb) / (Mx My), compensated FP64, synchronously to host *out. */
int cm_scalar_product_k(cm_plan* plan, const void* ka, const void* kb, double* out, void* stream);
/* sum of a b over the (Mx, My) grid, compensated FP64, to host *out
 * (evaluate_elastic_energy, :676-684, is 0.5 * dot(u, -f)). */
int cm_dot(cm_plan* plan, const void* a, const void* b, double* out, void* stream);

/* Constrained conjugate gradients (ConstrainedConjugateGradients.py:47-509).
 *  heights   device FP64 (nx, ny)
 *  surf_mask device uint8 (nx, ny), 1 = valid pixel, NULL = no mask
 *  u         device (Mx, My) plan precision: in = initial displacements
 *            (zeros if none), out = result.x
 *  f         device (Mx, My): in = initial forces when init_mode == 1,
 *            out = final forces (result.jac = -f[:nx, :ny])
 *  cmask     device uint8 (Mx, My) out: bit0 = result.active_set,
 *            bit1 = gap < 0 at the start of the last iteration (result.plastic,
 *            and the active set on the topography after a mixing step)
 * Returns CM_CONVERGED, CM_MAXITER or a negative error. */
int cm_ccg_solve(cm_plan* plan, const cm_ccg_params* params, const double* heights,
                 const uint8_t* surf_mask, void* u, void* f, uint8_t* cmask,
                 cm_ccg_stats* stats, cm_progress_fn progress, void* user, void* stream);

/* cm_ccg_solve without any allocation.  No counterpart in the reference: for
 * callers that cannot allocate (serving loops with a fixed memory footprint,
 * harnesses that forbid hipMalloc).  The whole workspace of the solve lives in
 * a static device array of 136 MiB that every GPU the library runs on holds
 * until the process ends; a solve whose workspace (cm_ccg_session_bytes(plan,
 * 1)) exceeds it returns CM_E_OOM.  Calls on one GPU share that array, which
 * keeps the last call's contents: calls that overlap in time on the same GPU
 * (from any thread or context) race.  No progress callback; use_graph is
 * ignored (the iterations are launched directly).  Same kernels and results
 * as cm_ccg_solve. */
int cm_ccg_solve_noalloc(cm_plan* plan, const cm_ccg_params* params, const double* heights,
                         const uint8_t* surf_mask, void* u, void* f, uint8_t* cmask,
                         cm_ccg_stats* stats, void* stream);

/* Batched solve of B independent instances on the same plan (same grid and
 * elastic constants), one launch per kernel for all instances.  Arrays are
 * host arrays of B device pointers; stats is a host array of B records. */
int cm_ccg_solve_batch(cm_plan* plan, int B, const cm_ccg_params* params,
                       const double* const* heights, const uint8_t* const* surf_masks,
                       void* const* u, void* const* f, uint8_t* const* cmask,
                       cm_ccg_stats* stats, void* stream);

/* ---- resident solves (serving, batching, benchmarking) ----------------
 * A session keeps B instances' solver state (conjugate directions, spectra,
 * scalar state) resident on the device.  cm_ccg_begin runs the set-up
 * (ConstrainedConjugateGradients.py:179-219) asynchronously on `stream`;
 * cm_ccg_step enqueues n iterations (finished instances turn into no-ops);
 * cm_ccg_poll synchronises and reads the statistics; cm_ccg_finish computes
 * the elastic energy; cm_ccg_end frees the workspace.  Between begin and
 * finish the u buffers hold u - h on the topography points (the heights are
 * read once per iteration that way); cm_ccg_finish restores u, after which
 * cm_ccg_step and cm_ccg_profile return CM_E_ARG.  Sessions on different
 * streams run concurrently.  No counterpart in the reference, which only has
 * the blocking call replaced by cm_ccg_solve. */
size_t cm_ccg_session_bytes(cm_plan* plan, int B);
int cm_ccg_begin(cm_plan* plan, int B, const cm_ccg_params* params, const double* const* heights,
                 const uint8_t* const* surf_masks, void* const* u, void* const* f, uint8_t* const* cmask,
                 void* stream, cm_session** out);
int cm_ccg_step(cm_session* s, int n_iters);
int cm_ccg_poll(cm_session* s, cm_ccg_stats* stats, int* all_done);
int cm_ccg_finish(cm_session* s, cm_ccg_stats* stats);
int cm_ccg_end(cm_session* s);
/* Run n iterations launch by launch with HIP events between the kernels of
 * an iteration; kernel_us[9] receives the mean time of each slot (phase A,
 * bisection, direction r2c, direction column pass with the step length,
 * tau c2r, phase B, force r2c, column, displacement c2r).  Phase A is fused
 * into the displacement c2r and the step length into the direction column
 * pass, so slots 0 and 4 measure only the gap between two launches.
 * Measurement hook, no reference counterpart. */
int cm_ccg_profile(cm_session* s, int n_iters, float* kernel_us);

#ifdef __cplusplus
}
#endif
#endif /* CMHIP_H */
