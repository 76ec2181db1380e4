/*
 * swampe_capi.h -- C ABI of the MI355X spectral shallow-water time-step core.
 *
 * The reference (SWAMPE-JAX, package my_swampe) is pure Python/JAX with no FFI;
 * its drop-in boundary is the Python API in src/my_swampe/model.py.  This
 * library is what that boundary calls: the Python host module
 * my_swampe_amd.model (run_model_scan / run_model_scan_final / run_model)
 * binds these entry points with ctypes.  Each entry point cites the reference
 * function whose device work it replaces.
 *
 * Conventions
 *   - extern "C", plain pointers and sizes; status 0 = ok, < 0 = error (see
 *     swx_last_error()).  Nothing throws across the ABI.
 *   - Device arrays are C-order, FP64 unless stated.  Complex arrays are
 *     complex128 stored as interleaved (re, im) doubles.
 *   - Grid (J, I): j = Gauss node, mu ascending; i = longitude from -pi.
 *     Fourier (J, M+1); spectral (M+1, N+1) [m][n], zero for n < m.
 *   - A plan owns its device buffers and is bound to the device that was
 *     current when it was created.  It is not re-entrant: calls on one plan
 *     must be serialized by the caller (one stream at a time).
 *   - `stream` is a hipStream_t passed as void* (NULL = default stream).
 */
#ifndef SWAMPE_CAPI_H
#define SWAMPE_CAPI_H

#ifdef __cplusplus
extern "C" {
#endif

#define SWX_OK 0
#define SWX_ERR_ARG (-1)
#define SWX_ERR_HIP (-2)
#define SWX_ERR_STATE (-3)

/* Transform precision.  The spectral state update and the diffusion filter
 * are FP64 in every mode; only the DFT/Legendre contractions change. */
#define SWX_PREC_FP64 0
#define SWX_PREC_FP32 1
#define SWX_PREC_BF16X3 2

/* Per-member physical parameters: SWX_NPARAM doubles per member, in this
 * order.  Mirrors the scalar arguments of run_model_scan
 * (src/my_swampe/model.py:1345-1388); K6PHI < 0 means "same as K6"
 * (model.py:513).  SI_ALPHA (semi-implicit off-centering) and WILLIAMS_ALPHA
 * (RAW filter) are read only by plans with semi_implicit / raw_filter set
 * (model.py:156-187 RunFlags.si_alpha, williams_alpha). */
enum {
  SWX_P_DT = 0,
  SWX_P_A,
  SWX_P_OMEGA,
  SWX_P_G,
  SWX_P_PHIBAR,
  SWX_P_TAURAD,
  SWX_P_TAUDRAG,
  SWX_P_DPHIEQ,
  SWX_P_K6,
  SWX_P_K6PHI,
  SWX_P_ALPHA,
  SWX_P_BLOWUP,
  SWX_P_SI_ALPHA,
  SWX_P_WILLIAMS_ALPHA,
  SWX_NPARAM
};

typedef struct swx_plan swx_plan;

typedef struct swx_plan_desc {
  int M;           /* truncation, 42, 63 or 106 (N = M); initial_conditions.py:50-85 */
  int n_members;   /* ensemble members sharing this truncation (>= 1) */
  int precision;   /* SWX_PREC_* */
  int expflag;     /* 0 = modified Euler, 1 = explicit (time_stepping.py:384) */
  int forcflag;    /* model.py:177-181 RunFlags */
  int diffflag;
  int modalflag;
  int diagnostics; /* per-step RMS/min/max reductions and blow-up gate */
  int test;        /* -1 = None (forced), 1 or 2 (Williamson tests) */
  int exact_basis; /* 0 = reference normalization (default; see swx_geometry),
                      1 = exactly orthonormal basis at every (m, n) */
  int semi_implicit; /* 1 = semi-implicit leapfrog + exponential hyperdiffusion
                        (semi_implicit_tdiff.py:79-264); not with expflag */
  int raw_filter;    /* 1 = Robert-Asselin-Williams filter (model.py:1007-1087);
                        not with expflag */
  /* Optional host arrays [J] (NULL = built-in): Gauss nodes mu (ascending, in
   * (-1, 1)) and weights to use instead of the built-in quadrature, e.g. the
   * caller's scipy.special.roots_legendre(J) as the reference takes it
   * (spectral_transform.py:80-114).  Those weights carry rounding of up to
   * ~3e-11 relative at J = 160; passing them reproduces the reference's
   * quadrature exactly.  The basis is evaluated at the given nodes. */
  const double *quad_mus;
  const double *quad_w;
} swx_plan_desc;

/* Pointers (device) receiving the reference's 27-field scan carry
 * (src/my_swampe/model.py:380-425), each for all members.  Any pointer may
 * be NULL to skip that field. */
typedef struct swx_state_out {
  /* complex128 [B][J][M+1] */
  double *etam_prev, *etam_curr, *deltam_prev, *deltam_curr, *Phim_prev, *Phim_curr;
  /* float64 [B][J][I] */
  double *eta_prev, *eta_curr, *delta_prev, *delta_curr, *Phi_prev, *Phi_curr, *U_curr, *V_curr;
  /* complex128 [B][J][M+1] */
  double *Um_curr, *Vm_curr, *Am_curr, *Bm_curr, *Cm_curr, *Dm_curr, *Em_curr;
  double *PhiFm_curr, *Fm_curr, *Gm_curr, *Rum_lag, *Rvm_lag;
  /* uint8 [B] */
  unsigned char *dead;
  /* complex128 [B][3][M+1][N+1]: spectral eta, delta, Phi of the current
   * level.  No reference counterpart (the reference never keeps them). */
  double *spec;
} swx_state_out;

/* Host-only FP64 geometry for truncation M (no GPU needed).
 * Replaces spectral_params + gauss_legendre + PmnHmn
 * (initial_conditions.py:50-85; spectral_transform.py:54-114, :219-287).
 * Outputs (any may be NULL): mus[J], w[J], lambdas[I], Pmn[J][M+1][N+1],
 * Hmn[J][M+1][N+1].  With exact_basis = 0 the basis reproduces the
 * reference's factorial normalization, which rounds to a subnormal or zero
 * scale for n + m >= ~170 (only reachable at M = 106); exact_basis = 1 gives
 * the exactly orthonormal basis everywhere.  SWX_ERR_ARG for an unsupported M. */
int swx_geometry(int M, int exact_basis, double *mus, double *w, double *lambdas, double *Pmn, double *Hmn);

/* Host-only Gauss-Legendre nodes (ascending) and weights of order J >= 1: replaces
 * gauss_legendre (src/my_swampe/spectral_transform.py:80-114).  mus[J], w[J]. */
int swx_gauss_legendre(int J, double *mus, double *w);

/* Host-only longitudes linspace(-pi, pi, I, endpoint=False): replaces build_lambdas
 * (src/my_swampe/spectral_transform.py:54-77).  lambdas[I], I >= 1. */
int swx_lambdas(int I, double *lambdas);

/* Host-only normalized associated Legendre basis and its derivative term at J arbitrary nodes
 * mus (each in [-1, 1]): replaces PmnHmn (src/my_swampe/spectral_transform.py:219-287).
 * Pmn, Hmn [J][M+1][N+1]; exact_basis as in swx_geometry. */
int swx_legendre_basis(int J, int M, int N, const double *mus, int exact_basis, double *Pmn, double *Hmn);

/* Grid sizes for truncation M: *J, *I (either may be NULL). */
int swx_grid_size(int M, int *J, int *I);

/* Create a plan: precompute the FP64 geometry, upload the basis and twiddle
 * tables in the requested precision, and allocate the device carry.
 * params: host [n_members][SWX_NPARAM].  Replaces build_static
 * (model.py:457-557). */
int swx_plan_create(const swx_plan_desc *desc, const double *params, swx_plan **out);

/* Replace the per-member parameters (host [n_members][SWX_NPARAM]). */
int swx_plan_set_params(swx_plan *plan, const double *params, void *stream);

void swx_plan_destroy(swx_plan *plan);

/* Two-level initialization from physical fields, device [B][J][I] each.
 * U0 and V0 must both be given or both NULL; NULL diagnoses the winds
 * spectrally from eta0/delta0.  Replaces _init_state_from_fields and
 * _diagnose_winds (model.py:694-772, :644-666). */
int swx_init(swx_plan *plan, const double *eta0, const double *delta0, const double *Phi0,
             const double *U0, const double *V0, void *stream);

/* After swx_init from a carry's current-level grids (eta, delta, Phi, U, V
 * _curr): set the physical prev level of eta, delta, Phi (device [B][J][I]
 * each, all three or none) and the per-member dead flags (device [B] bytes,
 * or NULL), so that the plan holds a mid-run carry of the modified-Euler
 * scheme, whose step reads no other prev-level field.  Replaces passing a
 * State pytree into _step_once / simulate_scan(_last) (model.py:795-803,
 * :1237-1342); other schemes return SWX_ERR_ARG. */
int swx_set_prev_level(swx_plan *plan, const double *eta_prev, const double *delta_prev, const double *Phi_prev,
                       const unsigned char *dead, void *stream);

/* Advance every member by nsteps steps with step indices t0, t0+1, ...
 * Replaces the lax.scan of simulate_scan / simulate_scan_last
 * (model.py:1237-1342) over _step_once (model.py:795-1194).
 *
 * History (optional, all device, may be NULL):
 *   hist_fields float64 [B][5][hist_T][J][I]: eta, delta, Phi, U, V after
 *     each step, written at time index hist_k0 + s for step s.
 *   hist_scalars float64 [B][hist_T][4]: rms, spin_min, 0, 0 (diagnostics).
 *   hist_phiext float64 [B][hist_T][J][2]: per-latitude min/max of the
 *     Robert-Asselin filtered Phi (reduce over J for phi_min/phi_max).
 *   hist_dead uint8 [B][hist_T].
 * Steps run on `stream`; the call returns without synchronizing. */
int swx_run(swx_plan *plan, int t0, int nsteps, double *hist_fields, double *hist_scalars,
            double *hist_phiext, unsigned char *hist_dead, int hist_T, int hist_k0, void *stream);

/* Copy the scan carry out to caller-provided device buffers (the State the reference's
 * run_model_scan returns as last_state, model.py:1735-1742).  When the last call carried the
 * physical prev level as the spectral lineage (SWX_OPT_PHYS_LINEAGE), the grid fields are rebuilt
 * first, into rows the plan allocates on its first rebuild. */
int swx_export_state(swx_plan *plan, const swx_state_out *out, void *stream);

/* swx_export_state that allocates nothing: the rebuild's rows live in a file-scope device array
 * of the library (64 MiB, one per GPU, held until the process ends), so the call makes no hipMalloc
 * or other allocation.  No counterpart in the reference; for callers that cannot allocate (code
 * under HIP graph capture, serving loops).  Plans whose rows exceed the array (B * 5 * J * (M+1) *
 * 2 * transform-precision bytes > 64 MiB: more than 609 FP32 / 304 FP64 members at T42, 273 / 136
 * at T63, 97 / 48 at T106) get SWX_ERR_ARG.  Calls that overlap on one GPU (concurrent streams or
 * threads) race on the array. */
int swx_export_state_noalloc(swx_plan *plan, const swx_state_out *out, void *stream);

/* Forward-only driver: run nsteps steps with no history and no export as
 * captured HIP graphs (graph_steps <= 0 disables graphs): the steps with
 * t <= 2 eagerly, then blocks of `graph_steps` steps while more than that
 * many remain, then the rest of the call (at most graph_steps steps) as one
 * graph.  The plan keeps up to 8 instantiated graphs keyed by the shape of
 * their step sequence (least recently replayed replaced), so repeated calls
 * of one length capture once.  Same semantics as swx_run with all history
 * pointers NULL.  No reference counterpart beyond the jitted lax.scan it
 * replaces (model.py:1281-1342). */
int swx_run_graph(swx_plan *plan, int t0, int nsteps, int graph_steps, void *stream);

/* Mixed truncations in one launch: an ensemble of two initialized FP32 plans of
 * different truncations (42/63, 42/106 or 63/106; plans keep their own state,
 * parameters and tables).  swx_ensemble_run_graph advances both plans by
 * nsteps steps from t0 exactly as swx_run_graph does for each, with one
 * Legendre launch and one grid launch per step covering both plans' members,
 * replayed from captured HIP graphs as swx_run_graph replays them.  The plans must
 * outlive the ensemble and are not modified otherwise.  Replaces batching
 * members of different truncations, which the reference cannot do in one
 * vmapped scan (one Static per jit, model.py:1281-1342). */
typedef struct swx_ensemble swx_ensemble;
int swx_ensemble_create(swx_plan *a, swx_plan *b, swx_ensemble **out);
int swx_ensemble_run_graph(swx_ensemble *ens, int t0, int nsteps, int graph_steps, void *stream);
void swx_ensemble_destroy(swx_ensemble *ens);

/* Concurrent stepping of 2 to 4 initialized plans (for example one ensemble's
 * members split in halves): plan k runs on internal stream k (default priority,
 * created with the pipeline so that the streams sit on distinct hardware queues),
 * forked after the caller's prior work on `stream` (the null stream too) and
 * joined before its later work, so that the step kernels of different
 * plans overlap on the device (one plan's Legendre launch beside another's
 * grid launch).  swx_pipeline_run_graph advances every plan as swx_run_graph
 * does; results are those of each plan run alone.  The plans must outlive the
 * pipeline and must not be stepped through another stream meanwhile.
 * Replaces the single vmapped scan over all members (model.py:1281-1342). */
typedef struct swx_pipeline swx_pipeline;
int swx_pipeline_create(swx_plan *const *plans, int n, swx_pipeline **out);
int swx_pipeline_run_graph(swx_pipeline *pipe, int t0, int nsteps, int graph_steps, void *stream);
void swx_pipeline_destroy(swx_pipeline *pipe);

/* Device memory (bytes) the plan holds: basis and DFT tables, carry, work
 * buffers, and the adjoint workspace once allocated.  Caller-owned tensors
 * (checkpoints, outputs) are not included.  No reference counterpart. */
int swx_plan_device_bytes(const swx_plan *plan, size_t *bytes);
/* Steps the plan has run since its creation, by how: counts[0] replayed from a captured graph,
 * counts[1] launched eagerly, counts[2] (of either) advancing the spectral prev lineage,
 * counts[3] writing the physical grid carry.  Bookkeeping of this build (the reference's jitted
 * scan has no counterpart): it lets a caller or a benchmark state which path its steps took. */
int swx_plan_step_counts(const swx_plan *plan, long long *counts);

/* Measurement hook (HIP events on `stream`): run nsteps steps eagerly
 * (t0 .. t0+nsteps-1, the steady physical-carry mode of swx_run_graph) timed
 * as one block, then nsteps back-to-back launches of the grid kernel on the
 * final state (it reads the Fourier rows and rewrites the same Legendre
 * inputs, so the state is unchanged).  *ms_grid = grid block / nsteps,
 * *ms_leg = step block / nsteps - *ms_grid.  When the physical carry is
 * written every step (|alpha| >= 0.5, diagnostics, or the full-carry option)
 * each launch is bracketed by events instead.  Synchronizes the stream.  No
 * reference counterpart (the reference times whole scans,
 * scripts/benchmark_scan.py:116-141). */
int swx_time_kernels(swx_plan *plan, int t0, int nsteps, float *ms_leg, float *ms_grid, void *stream);

/* Measurement hook: ceilings of the current GPU measured with the instructions and access
 * patterns the step kernels use.  out[0..2] = dense matrix-core throughput (TFLOP/s) of
 * v_mfma_f32_16x16x4_f32, v_mfma_f64_16x16x4_f64 and v_mfma_f32_16x16x16_bf16 (8 independent
 * accumulators per wave, 8 waves per SIMD); out[3..5] = HBM bandwidth (TB/s) of a device copy
 * (bytes read + written), a read-only pass and a write-only pass over `bytes` bytes (a multiple of
 * 16, at least 256 MiB; two such buffers are allocated and freed here).  Shortest of 5 timed
 * launches after 2 untimed ones, HIP events on `stream`, which is synchronized.  No reference
 * counterpart: the measured ceilings the step kernels' roofline fractions are quoted against. */
int swx_measure_ceilings(double *out, size_t bytes, void *stream);

/* ---- Spectral initial state and checkpoints ------------------------------
 * State arrays are device complex128 [B][3][M+1][N+1] (eta, delta, Phi
 * spectral coefficients, n >= m used). */

/* Two-level start from spectral coefficients c: the grid fields are their
 * inverse transforms (winds by the spectral inversion), so the first step is
 * the same map as every later step.  The differentiable entry point of the
 * adjoint (the reference differentiates the scan through jax.grad,
 * unit_tests/test_autodiff.py:53-185). */
int swx_init_spectral(swx_plan *plan, const double *c, void *stream);

/* Save / restore the current spectral level (checkpointing).  state_set also
 * recomputes the Fourier rows and Legendre inputs of the restored level. */
int swx_state_get(swx_plan *plan, double *out, void *stream);
int swx_state_set(swx_plan *plan, const double *in, void *stream);

/* ---- Hand-written adjoint (reverse mode) of the modified-Euler step -----
 * lam = dL/dS is held by the plan.  Usage: begin; terminal (cotangents of
 * the final grid fields, device float64 [B][J][I], any may be NULL); then for
 * t = T-1 .. 0: step(S_t) with the forward state S_t before step t; result
 * gives dL/dS_0 (device complex128 [B][3][M+1][N+1]) and the accumulated
 * parameter gradients (device float64 [B][3]: taurad, taudrag, DPhieq).
 * Replaces the reverse-mode transpose of the lax.scan that jax.grad builds
 * (model.py:1281-1342).  The forward carry of the plan
 * (Fourier rows) is used as scratch. */
int swx_adjoint_begin(swx_plan *plan, void *stream);
/* Add the cotangents of the grid fields observed at the sweep's current
 * position (the state whose dL/dS the plan holds) to it: the time-series
 * counterpart of swx_adjoint_terminal (which assigns), for losses that read
 * the trajectory at several steps (the phase-curve likelihood accumulated
 * over the scan, DEVELOPMENT.md:587-616).  No counterpart in the reference's
 * code (jax.grad differentiates such a scan by itself). */
int swx_adjoint_observe(swx_plan *plan, const double *g_eta, const double *g_delta, const double *g_Phi,
                        const double *g_U, const double *g_V, void *stream);
int swx_adjoint_terminal(swx_plan *plan, const double *g_eta, const double *g_delta, const double *g_Phi,
                         const double *g_U, const double *g_V, void *stream);
int swx_adjoint_step(swx_plan *plan, const double *state_t, void *stream);
/* Forward sweep that records its trajectory for the reverse sweep: for k = 0 .. nsteps-1,
 * copy the current spectral state into tape[k] (device, [B][3][M+1][N+1] complex128 per k, as
 * swx_state_get), then advance one step (step index t0 + k); only the last step writes the
 * physical grid carry.  swx_adjoint_sweep then runs swx_adjoint_step(tape[k]) for
 * k = nsteps-1 .. 0 (after swx_adjoint_begin and the terminal cotangents).  One call per sweep
 * instead of one per step: the reverse-mode scan of model.py:1281-1342 without a
 * rematerialized segment.  Not available with semi_implicit / raw_filter (as swx_state_get). */
int swx_run_tape(swx_plan *plan, int t0, int nsteps, double *tape, void *stream);
int swx_adjoint_sweep(swx_plan *plan, int nsteps, const double *tape, void *stream);
int swx_adjoint_result(swx_plan *plan, double *lam_out, double *pgrad_out, void *stream);
/* Gradients in the hyperdiffusion coefficients K6 (eta, delta) and K6Phi (Phi), the reverse
 * mode of the sigma6 / sigma6Phi factors (filters.py:95-151) that jax.grad reaches through the
 * scan (unit_tests/test_autodiff.py:64-86).  A step ends with S_{t+1} = sigma(K6) * P, so with
 * the plan holding lam = dL/dS_{t+1}, swx_adjoint_diffusion(S_{t+1}) adds
 * sum Re(conj(lam) dS_{t+1}/dK6) per member: call it for every state a step produced, before
 * the reverse step that consumes that lam (swx_adjoint_sweep does so for the states inside its
 * tape under SWX_OPT_ADJ_DIFFUSION; the caller passes the state past the tape's end, e.g. the
 * final state after swx_adjoint_terminal).  swx_adjoint_diffusion_result copies the accumulated
 * device float64 [B][2] (dL/dK6 over eta and delta, dL/dK6Phi over Phi; with K6Phi following
 * K6, dL/dK6 is their sum).  Zero without diffflag.  State layout as swx_state_get. */
int swx_adjoint_diffusion(swx_plan *plan, const double *state, void *stream);
int swx_adjoint_diffusion_result(swx_plan *plan, double *out, void *stream);

/* Phase-curve observable (replaces the projection of
 * unit_tests/test_parity_reference_regression.py:67-82, the retrieval's
 * visibility-weighted disc integral of Phi):
 *   flux[b][k] = sum_j sqrt(max(1 - mu_j^2, 0)) sum_i Phi_b[j][i] max(cos(lambda_i - phases[k]), 0)
 * with lambda_i = -pi + 2 pi i / I, then divided by mean_k |flux[b][k]| when
 * normalize != 0 and that mean is > 0.  Phi_b = Phi + b * member_stride, b < n,
 * each a row-major J x I grid; mus: the J Gauss nodes; phases: K angles;
 * flux: (n, K); flux_raw: (n, K) before normalization, or null.  All device
 * pointers.  1 <= J, I <= 1024, 1 <= K <= 2048; n = 0 does nothing. */
int swx_phase_curve(const double *Phi, long long member_stride, int n, int J, int I, const double *mus,
                    const double *phases, int K, int normalize, double *flux, double *flux_raw, void *stream);
/* Vector-Jacobian product of swx_phase_curve: g_flux (n, K), the cotangent of
 * flux, and flux_raw from the forward call -> g_Phi (n, J, I), dense,
 * overwritten.  No counterpart in the reference's code (jax.grad
 * differentiates the projection by itself). */
int swx_phase_curve_vjp(const double *g_flux, const double *flux_raw, int n, int J, int I, const double *mus,
                        const double *phases, int K, int normalize, double *g_Phi, void *stream);
/* swx_phase_curve of every member's current Phi grid (the Phi_curr that
 * swx_export_state returns) with the plan's Gauss nodes: flux (B, K). */
int swx_plan_phase_curve(swx_plan *plan, const double *phases, int K, int normalize, double *flux,
                         double *flux_raw, void *stream);

/* Vorticity and divergence from Fourier winds, spectral and on the grid: replaces
 * diagnostic_eta_delta (src/my_swampe/spectral_transform.py:599-650) with the same arguments,
 * as device FP64 arrays: Um, Vm (n, J, M+1) complex (n members, contiguous); fmn (M+1, N+1)
 * complex; Pmn, Hmn (J, M+1, N+1); w (J); tstepcoeff, mJarray (J, M+1); dt.  Outputs (the
 * reference's return tuple): eta, delta (n, J, I) complex, the grids of invrs_fft(invrs_leg(.))
 * (either may be null), etamn (= zeta_mn + fmn), deltamn (n, M+1, N+1) complex.  Errors:
 * SWX_ERR_ARG for negative n or n above 65535, N < M, I < 2 M + 1 or I above 2048 with a grid
 * requested, dt = 0, null inputs. */
int swx_diagnostic_eta_delta(const double *Um, const double *Vm, const double *fmn, int n, int I, int J, int M, int N,
                             const double *Pmn, const double *Hmn, const double *w, const double *tstepcoeff,
                             const double *mJarray, double dt, double *eta, double *delta, double *etamn,
                             double *deltamn, void *stream);

/* The reference's standalone transforms on device FP64 arrays (complex = interleaved re, im),
 * K independent batch rows each (K <= 65535), on `stream`.  Errors: SWX_ERR_ARG for negative
 * or oversized counts, a longitude count I outside [1, 2048] where a DFT runs, null pointers.
 *
 * fwd_fft_trunc / fwd_fft_trunc_batch (src/my_swampe/spectral_transform.py:290-328): data
 * (K, J, C) real, or complex with is_complex = 1 -> out (K, J, min(M+1, I)) complex,
 * fft(data / I, n=I) truncated to m = 0..M (rows zero-padded or cut to I columns). */
int swx_fwd_fft_trunc(const double *data, int is_complex, int K, int J, int C, int I, int M, double *out,
                      void *stream);

/* invrs_fft (src/my_swampe/spectral_transform.py:331-348): X (K, J, C) complex ->
 * ifft(I X, n=I) (K, J, I) complex. */
int swx_invrs_fft(const double *X, int K, int J, int C, int I, double *out, void *stream);

/* fwd_leg (src/my_swampe/spectral_transform.py:405-435) with w (J), or fwd_leg_w /
 * fwd_leg_w_batch (:370-402) with w = NULL and a preweighted basis: data (K, J, M+1) complex,
 * basis (J, M+1, N+1) -> out (K, M+1, N+1) complex. */
int swx_fwd_leg(const double *data, const double *basis, const double *w, int K, int J, int M, int N, double *out,
                void *stream);

/* invrs_leg (src/my_swampe/spectral_transform.py:438-481): S (K, M+1, N+1) complex ->
 * (K, J, I) complex, +m at column m and the conjugates at column I - m; I >= M + 1,
 * M + 1 <= 4096. */
int swx_invrs_leg(const double *S, const double *basis, int K, int I, int J, int M, int N, double *out,
                  void *stream);

/* invrsUV / invrsUV_with_coeffs (src/my_swampe/spectral_transform.py:484-596): deltamn, etamn
 * (K, M+1, N+1) complex, fmn (M+1, N+1) complex, Pmn, Hmn (J, M+1, N+1), tstepcoeffmn,
 * marray (M+1, N+1) -> U, V (K, J, I) complex and the positive-m coefficients Um, Vm
 * (K, J, M+1) complex; any output may be NULL (not all four).  Needs N >= M, I >= 2 M + 1. */
int swx_invrs_uv(const double *deltamn, const double *etamn, const double *fmn, int K, int I, int J, int M, int N,
                 const double *Pmn, const double *Hmn, const double *tstepcoeffmn, const double *marray, double *U,
                 double *V, double *Um, double *Vm, void *stream);

/* Plan options (swx_plan_set_option).
 *   SWX_OPT_GRID_KERNEL: 0 = automatic (MFMA grid kernel for FP32 forward
 *     stepping when it fits), 1 = always the plain grid kernel.
 *   SWX_OPT_FULL_PHYS_CARRY: 1 = write the physical (grid) carry at every
 *     step; 0 (default) = only in the final steps whose Robert-Asselin
 *     contribution can reach the exported state (identical results).
 *   SWX_OPT_LEG_KERNEL: 0 = automatic (MFMA Legendre kernel for FP32 forward
 *     stepping), 1 = always the plain Legendre kernel.
 *   SWX_OPT_PHYS_STEPS: k > 0 = write the physical carry only in the last k
 *     steps of each swx_run / swx_run_graph call without history; 0 (default)
 *     = automatic as above.  The current-level grid fields (eta, delta, Phi,
 *     U, V _curr) are exact for any k >= 1; the physical prev level, which
 *     carries the Robert-Asselin filter's memory, is exact only when k covers
 *     that memory.  k = -1: never (the exported grid fields are not updated).
 *     For callers that need only the spectral state and the final
 *     current-level fields (the differentiable rollout and its recompute). */
#define SWX_OPT_GRID_KERNEL 1
#define SWX_OPT_FULL_PHYS_CARRY 2
#define SWX_OPT_LEG_KERNEL 3
#define SWX_OPT_PHYS_STEPS 4
/*   SWX_OPT_LEG_MB: members per workgroup of the MFMA Legendre kernel; set at
 *     plan creation to the fastest of the instantiated sizes (4, 8 at T42;
 *     2, 4, 8 at T63 and T106), timed on the plan's own buffers; setting it
 *     forces another available size.  Results do not depend on it. */
#define SWX_OPT_LEG_MB 5
/*   SWX_OPT_GRID_FOLD: which FP32 matrix-core grid kernel runs: 1 = the full
 *     longitude DFTs, 2 = the half-period folded DFTs (x(i + I/2) from the
 *     even/odd wavenumber sums: half the matrix-core work).  Set at plan
 *     creation to 2 wherever the folded kernel fits (a fixed rule, so plans of
 *     one shape round alike); setting it forces one.  Results agree to FP32
 *     rounding. */
#define SWX_OPT_GRID_FOLD 6
/*   SWX_OPT_CONCURRENT: 1 = the plan is stepped beside other plans' kernels (swx_pipeline_create
 *     sets it for its plans; set it when driving plans concurrently on several streams
 *     yourself): the steady T42 grid launch then keeps a lower occupancy that leaves room for
 *     the other kernels.  Results do not depend on it. */
#define SWX_OPT_CONCURRENT 7
/*   SWX_OPT_PHYS_LINEAGE: how a modified-Euler plan without diagnostics (test != 1) carries the
 *     physical prev level through a swx_run / swx_run_graph call without history that is longer
 *     than the Robert-Asselin window: 0 (default) = as a spectral lineage advanced by the
 *     Legendre kernel in the window steps (no FP64 grid planes streamed), the grid fields rebuilt
 *     from it and from the current Fourier rows when they are read (export, phase curve, a call
 *     that writes the grid carry); -1 = the grid carry of SWX_OPT_FULL_PHYS_CARRY /
 *     SWX_OPT_PHYS_STEPS; k > 0 = the lineage over the last k steps of such a call (the default
 *     window is exact; shorter ones are for tests).  The two carries give the same fields to the
 *     rounding of the transforms (FP64 plans ~1e-15 relative, FP32 ~1e-7). */
#define SWX_OPT_PHYS_LINEAGE 8
/*   SWX_OPT_LEG_FOLD: which FP32 matrix-core Legendre contractions run: 2 (default) = folded by
 *     the equatorial symmetry of the Gauss nodes and the basis (sums over the southern half of
 *     the latitudes on T(j) +- T(J-1-j); half the matrix-core work), 1 = over all latitudes.
 *     Results agree to FP32 rounding. */
#define SWX_OPT_LEG_FOLD 9
/*   SWX_OPT_ADJ_DIFFUSION: 1 = swx_adjoint_sweep also accumulates the diffusion partials
 *     (swx_adjoint_diffusion) of the pairs inside its tape; 0 (default) = it does not. */
#define SWX_OPT_ADJ_DIFFUSION 10
int swx_plan_set_option(swx_plan *plan, int option, int value);
/* Current value of an option (for SWX_OPT_LEG_MB and SWX_OPT_GRID_FOLD: the
 * choice in use). */
int swx_plan_get_option(const swx_plan *plan, int option, int *value);

/* Last error message of the calling thread. */
const char *swx_last_error(void);

/* Library version string. */
const char *swx_version(void);

#ifdef __cplusplus
}
#endif

#endif /* SWAMPE_CAPI_H */
