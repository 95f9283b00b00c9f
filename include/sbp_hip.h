/*
 * sbp_hip.h -- C-ABI of the MI355X (gfx950) SBP-SAT hot path.
 *
 * The reference drives its matrix-free right-hand side through PETSc's TS:
 *   PetscErrorCode rhs(TS ts, PetscReal t, Vec v_src, Vec v_dst, void *ctx)
 * registered with TSSetRHSFunction (code/src/time_stepping/ts_rk.cpp:21) and
 * advanced by ts_rk4(da, t_end, dt, v, rhs, ctx) (code/include/time_stepping/ts_rk.h:26,
 * code/src/time_stepping/ts_rk.cpp:54-65).  Every entry point below replaces one
 * of those pieces; the reference interface each one stands in for is cited at
 * the declaration.  Plain pointers and sizes only: no HIP, PETSc or C++ types.
 *
 * Arrays crossing the boundary use the reference's own single-rank layout,
 * AoS v[j][i][c] at flat index C*(i + nx*j) + c (code/include/grids/layout.h:118-127,
 * code/src/grids/create_layout.cpp:39-41), in double precision.
 *
 * Return codes mirror PetscErrorCode: 0 on success, a negative SBP_E* code on
 * failure; sbp_last_error() then describes the failure (thread-local string).
 */
#ifndef SBP_HIP_H
#define SBP_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- problem kinds (one per reference demo family) -------------------- */
typedef enum {
  SBP_ADVECTION_1D  = 1, /* u_t + a u_x = 0          demo/advection/advection_1D_sim.cpp  */
  SBP_ADVECTION_2D  = 2, /* u_t + a u_x + b u_y = 0  demo/advection/advection_2D_sim.cpp  */
  SBP_ACOUSTIC_2D   = 3, /* [u,v,p] first-order acoustics  demo/wave, demo/wave_hom       */
  SBP_REFLECTION_1D = 4  /* [u;v]_t = [v_x;u_x]      demo/reflection/reflection_sim.cpp   */
} sbp_physics;

/* Material model of SBP_ACOUSTIC_2D. */
typedef enum {
  SBP_MAT_UNIT    = 0, /* rho = K = 1, no forcing: demo/wave_hom/wave_eq_hom_rhs.h          */
  SBP_MAT_REF_RHO = 1, /* rho = 2 + x*y, K = 1, manufactured forcing: demo/wave/wave_eq_rhs.h:42-64 */
  SBP_MAT_ARRAYS  = 2  /* caller-supplied rho^-1(x,y) and K(x,y) per grid point              */
} sbp_material;

/* Boundary treatment. SBP_BC_DEFAULT selects the reference's own choice:
 * upwind inflow SAT for advection (advection_rhs.h:94-119,460-528), projection
 * u=0 for reflection (reflection_rhs.h:108-120), free-surface SAT without rho^-1
 * for acoustics (wave_eq_rhs.h:417-468). */
typedef enum {
  SBP_BC_DEFAULT           = 0,
  SBP_BC_FREE_SURFACE_REF  = 1, /* p = 0 SAT exactly as wave_eq_rhs.h:425,438,452,466 (no rho^-1) */
  SBP_BC_FREE_SURFACE      = 2, /* p = 0 SAT carrying rho^-1: energy-conserving for any rho      */
  SBP_BC_WALL              = 3  /* v.n = 0 SAT on the p equation carrying K: energy-conserving    */
} sbp_bc;

/* Arithmetic form of the device state. */
typedef enum {
  SBP_FP64       = 0, /* everything in FP64 (the reference's PetscScalar)                     */
  SBP_FP32       = 1, /* stencils, RK4 stages and state in FP32                               */
  SBP_FP32_KAHAN = 2  /* FP32 bulk; u^{n+1} = u^n + du accumulated with Kahan compensation  */
} sbp_precision;

typedef struct sbp_problem {
  int physics;          /* sbp_physics                                                   */
  int order;            /* SBP interior order 2, 4 or 6 (code/include/sbpops/op_defs.h:18-30) */
  int nx, ny;           /* grid points per direction (ny = 1 for the 1D problems)        */
  double xl, xr;        /* domain [xl,xr] x [yl,yr]; hi = (N-1)/(xr-xl) as in the drivers */
  double yl, yr;
  double ax, ay;        /* advection velocities a, b (advection_2D_sim.cpp:74-75)        */
  int material;         /* sbp_material (acoustics only)                                 */
  int bc;               /* sbp_bc                                                         */
  int precision;        /* sbp_precision                                                  */
  const double *rho_inv;/* nx*ny values [j][i], SBP_MAT_ARRAYS only; copied at create    */
  const double *bulk;   /* nx*ny values [j][i] of K, SBP_MAT_ARRAYS only; copied         */
  /* Advection velocity per grid point, nx*ny values [j][i] (1D: nx values), copied
   * at create; NULL = the constant ax / ay above.  These are the AppCtx velocity
   * functions a(i) (advection_1D_sim.cpp:19,70) and a(i,j), b(i,j)
   * (advection_2D_sim.cpp:21,74-75) sampled on the grid: the RHS multiplies the
   * derivatives by them point by point (advection_rhs.h:22,36,49,166-393) and
   * each inflow SAT takes tau from the velocity at its own boundary point
   * (advection_rhs.h:102,116,461-528).  Advection only. */
  const double *vel_x;
  const double *vel_y;  /* 2D only */
} sbp_problem;

/* Error codes. */
enum {
  SBP_OK          = 0,
  SBP_EINVAL      = -1, /* bad argument / unsupported combination      */
  SBP_EHIP        = -2, /* HIP runtime error                           */
  SBP_ENOMEM      = -3, /* device or host allocation failed            */
  SBP_ENODEV      = -4, /* no usable gfx950 device                     */
  SBP_ESTATE      = -5, /* call out of sequence (e.g. no state loaded) */
  SBP_EDIVERGED   = -6  /* adaptive integrator failed: a step rejected more than 10
                           times (PETSc's TS_DIVERGED_STEP_REJECTED) or a non-finite
                           error estimate (PETSc's TSErrorWeightedNorm, PETSC_ERR_FP) */
};

typedef struct sbp_ctx sbp_ctx; /* opaque: device buffers, stream, graphs */

/* ---- context ------------------------------------------------------------
 * Replaces the per-demo AppCtx (e.g. demo/wave/wave_eq_sim.cpp:36-45,119-127)
 * together with the DMDA that sizes its vectors (wave_eq_sim.cpp:99-106). */
int  sbp_create(const sbp_problem *desc, sbp_ctx **out);
int  sbp_destroy(sbp_ctx *ctx);
/* Number of components C and total values nx*ny*C of one AoS vector. */
int  sbp_components(const sbp_ctx *ctx);
long sbp_vector_length(const sbp_ctx *ctx);

/* Interior tile height in rows (default: chosen from the grid size).  A
 * tuning knob with no reference counterpart; results do not depend on it. */
int  sbp_set_tile_rows(sbp_ctx *ctx, int rows);

/* Two-stage stepping (on = 1): each RK4 step runs stages 1-2 and stages 3-4
 * in one launch each, the inner stage kept on chip: interior tiles, and
 * patches of the (CLS+R)-deep boundary frame that evaluate both stages
 * themselves; 5 array passes per step instead of 13, same arithmetic per
 * point.  2D acoustic contexts (every material model) and 2D advection
 * contexts (constant or per-point velocity); SBP_EINVAL otherwise.  On by
 * default for grids of at least 450^2 interior points (where it was
 * measured faster).  rows: tile height of the two-stage tiles
 * (0 = chosen from the grid).  A performance knob with no reference
 * counterpart: it changes how ts_rk4's steps are scheduled on the device
 * (ts_rk.cpp:54-65), not what they compute. */
int  sbp_set_fused(sbp_ctx *ctx, int on, int rows);
int  sbp_get_fused(const sbp_ctx *ctx);
/* Time the two-stage schedule's candidate tile heights (8 .. 136 rows) on
 * this grid for step dt and keep the fastest; *rows (optional) receives it
 * (0 when the context runs one launch per stage).  The resident state is
 * saved and restored, so tuning leaves it untouched; a tuning knob with no
 * reference counterpart, results do not depend on it. */
int  sbp_tune_fused(sbp_ctx *ctx, double dt, int *rows);
/* Store hint of the two-stage tiles' outputs: 0 (default) chooses from the
 * bytes a step streams (cached stores while U, Y2, Y3 and the materials fit
 * the 256 MB Infinity Cache, nontemporal ones beyond), 1 forces cached and 2
 * nontemporal stores.  A performance knob with no reference counterpart;
 * results do not depend on it (it selects between two builds of the same
 * kernel, which is how tests run both on small grids). */
int  sbp_set_store_policy(sbp_ctx *ctx, int policy);

/* ---- right-hand side ----------------------------------------------------
 * dst = F(t, src) on host AoS arrays.  Replaces the rhs_serial callbacks
 * (advection_1D_sim.cpp:212-228, advection_2D_sim.cpp:232-249,
 *  wave_eq_sim.cpp:291-308, wave_eq_hom_sim.cpp:291-308, reflection_sim.cpp:249-265),
 * i.e. the region kernels of partitioned_rhs/rhs.h:363-398 plus the SAT / BC
 * functors of partitioned_rhs/boundary_conditions.h:69-88. */
int  sbp_rhs(sbp_ctx *ctx, double t, const double *src, double *dst);

/* ---- time stepping ------------------------------------------------------
 * Classical RK4 with fixed dt for nsteps steps, v (host AoS) in/out.
 * Replaces ts_rk4 with an integer step count (ts_rk.cpp:54-65: TSRK4,
 * TSADAPTNONE); the stage times t+c_s*dt, c = (0,1/2,1/2,1), and the update
 * v += dt/6 (k1 + 2(k2+k3) + k4) follow RK4_custom
 * (code/include/time_stepping/time_stepping.h:127-142,267-269). */
int  sbp_rk4_advance(sbp_ctx *ctx, double *v, double t0, double dt, long nsteps);

/* ts_rk4 on the resident state: TSRK4 with TSADAPTNONE under
 * TS_EXACTFINALTIME_MATCHSTEP from t0 to t_end, first step dt
 * (ts_rk.cpp:15-39,54-65).  The step schedule is PETSc 3.12's: TSSolve clips a
 * first step that would overshoot; after each step ending at t with
 * hmax = t_end - t, a next step that would overshoot becomes hmax, one with
 * 2 dt > hmax becomes hmax/2, one with 1.01 dt > hmax becomes hmax (so
 * t_end = 10.5 dt runs 9 steps of dt and 2 of 0.75 dt; 10.005 dt runs 9 of dt
 * and one of 1.005 dt).  Step times accumulate as t = t + h.  nsteps and
 * t_final (optional) receive the steps taken and the final time. */
int  sbp_ts_rk4(sbp_ctx *ctx, double t0, double t_end, double dt, long *nsteps, double *t_final);

/* That schedule without a context (host only, no GPU): runs of equal steps,
 * run_dt[k] repeated run_len[k] times, k < *nruns <= max_runs (SBP_EINVAL if
 * more would be needed; the arrays may be null to count). */
int  sbp_matchstep_schedule(double t0, double t_end, double dt, int max_runs, double *run_dt,
                            long *run_len, int *nruns, double *t_final);

/* RK4_custom(da, Tend, dt, v, rhs, ctx) on the resident state
 * (code/include/time_stepping/time_stepping.h:98-152): tlen = round(Tend/dt)
 * steps from t = 0; when tlen*dt misses Tend by more than 1e-14 the step
 * becomes Tend/tlen.  tlen, dt_used and t_final (the values the reference
 * prints as "tlen: %d, dt: %e" and "Final t: %.8e") are optional. */
int  sbp_rk4_custom(sbp_ctx *ctx, double t_end, double dt, long *tlen, double *dt_used, double *t_final);

/* Adaptive Runge-Kutta-Fehlberg 4(5) on the resident state (FP64 contexts)
 * from t0 to t_end, first attempted step dt0.  Replaces ts_rk45
 * (ts_rk.cpp:41-52: TSRK5F, the fifth-order solution propagated and the
 * embedded fourth-order one estimating the error; TSADAPTBASIC;
 * TS_EXACTFINALTIME_MATCHSTEP).  The step loop is PETSc 3.12's TSStep_RK with
 * TSAdaptChoose_Basic at its defaults: an attempt is accepted when
 *   enorm = sqrt(mean(((y - y_hat) / (atol + rtol*max(|y|, |y_hat|)))^2)) <= 1,
 * the next step is h * clip(s * enorm^(-1/5), 0.1, 10) with s = 0.9, or 0.45
 * when the previous attempt of the same step was rejected too; after an
 * accepted step the MATCHSTEP rule of sbp_ts_rk4 adjusts it near t_end (its
 * defaults for the tolerances are atol = rtol = 1e-4).  Fails with
 * SBP_EDIVERGED when one step is rejected more than 10 times or enorm is not
 * finite, and with SBP_EINVAL after max_steps attempts (a safety bound with
 * no PETSc counterpart).  accepted, rejected and dt_next (the step the
 * controller would take next) are optional and also set on failure. */
int  sbp_rk45(sbp_ctx *ctx, double t0, double t_end, double dt0, double atol, double rtol,
              long max_steps, long *accepted, long *rejected, double *dt_next);

/* The basic adaptor's decision for one attempt (host only, no GPU):
 * TSAdaptChoose_Basic with the defaults above for an embedded pair of the
 * given order; prev_rejected: the previous attempt of the same step was
 * rejected.  SBP_EDIVERGED for a non-finite enorm. */
int  sbp_adapt_basic(double enorm, double h, int order, int prev_rejected, int *accept, double *next_h);

/* ---- device-resident state (the time loop stays on the GPU) -------------- */
int  sbp_set_state(sbp_ctx *ctx, const double *v);  /* host AoS -> device state */
int  sbp_get_state(sbp_ctx *ctx, double *v);        /* device state -> host AoS */
/* Every stride-th point in x and y (injection onto the nested grid of
 * ((nx-1)/stride+1) x ((ny-1)/stride+1) points), host AoS. */
int  sbp_get_state_strided(sbp_ctx *ctx, int stride, double *v);
/* Advance the resident state nsteps RK4 steps from time t0 (asynchronous). */
int  sbp_step(sbp_ctx *ctx, double t0, double dt, long nsteps);
int  sbp_synchronize(sbp_ctx *ctx);
/* Run nsteps resident steps bracketed by HIP events on the context's stream;
 * *ms receives the elapsed event time.  Used by the benchmark. */
int  sbp_time_steps(sbp_ctx *ctx, double t0, double dt, long nsteps, double *ms);
/* Time each launch of a step separately (HIP events around every launch,
 * nsteps steps); ms_stage[4] receives the mean duration per launch: one per
 * stage, or with two-stage stepping the launches of stages 1-2 and 3-4 in
 * ms_stage[0] and ms_stage[2] (ms_stage[1] = ms_stage[3] = 0). */
int  sbp_time_stages(sbp_ctx *ctx, double dt, long nsteps, double ms_stage[4]);

/* ---- norms ----------------------------------------------------------------
 * Discrete energy in the SBP norm H = Hx (x) Hy (code/include/sbpops/H_central.h:39-103
 * gives the 1D weights; the reference's 2D norm is a stub, H_central.h:177-183):
 *   E = sum_{j,i} Hx_i Hy_j sum_c w_c(i,j) v_c(i,j)^2
 * with w = (rho, rho, 1/K) for acoustics and w = 1 otherwise.  FP64. */
int  sbp_energy(sbp_ctx *ctx, const double *v, double *E);      /* host AoS vector  */
int  sbp_state_energy(sbp_ctx *ctx, double *E);                  /* resident state   */
/* sqrt(sum_c v_c^T H v_c): get_norm_1D (H_central.h:70-103) and its 2D analogue. */
int  sbp_hnorm(sbp_ctx *ctx, const double *v, double *norm);

/* Error norms of vec_util.h:47-61: l2 = sqrt(hx*hy)*||v-w||_2, max = ||v-w||_inf. */
int  sbp_error_norms(sbp_ctx *ctx, const double *v, const double *w,
                     double *l2, double *maxerr);  /* on the device, FP64 */

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): sbp_error_norms on n FP64 values v, w already in device
 * memory, with no context, writing out[0] = sqrt(hx*hy)*||v-w||_2 (hy = 1 for a
 * 1D grid) and out[1] = ||v-w||_inf to device memory `out` on `stream` (a
 * hipStream_t; null = the null stream), asynchronously, with the same block
 * partition and summation order as sbp_error_norms.  It allocates nothing:
 * its partial sums live in a fixed 8 KB device array per GPU, held until the
 * process ends and shared by every call on that GPU, so calls that overlap on
 * one GPU race.  n is 1 .. SBP_NOALLOC_MAX_VALUES (SBP_EINVAL beyond). */
#define SBP_NOALLOC_MAX_VALUES (1L << 40)
int  sbp_error_norms_noalloc(const double *v, const double *w, long n, double hx, double hy,
                             double *out, void *stream);

/* ---- batched ensemble ---------------------------------------------------
 * Many independent 2D acoustic problems (uneven sizes, own materials, own dt)
 * advanced together: one persistent kernel per RK4 stage (or, in the
 * two-stage configurations, per pair of stages) walks a flattened list of
 * every instance's tiles and boundary blocks, and nsteps are replayed from a
 * captured hipGraph.  No counterpart in the reference, which runs one problem per MPI
 * job (code/batch.sh:1-7); each instance's arithmetic is the single-problem
 * path's (sbp_rk4_advance).  All instances share order, precision and
 * material model (SBP_MAT_UNIT or SBP_MAT_ARRAYS). */
typedef struct sbp_batch sbp_batch;
int  sbp_batch_create(const sbp_problem *descs, int count, sbp_batch **out);
int  sbp_batch_destroy(sbp_batch *b);
int  sbp_batch_count(const sbp_batch *b);
int  sbp_batch_set_state(sbp_batch *b, int k, const double *v);   /* host AoS of instance k */
int  sbp_batch_get_state(sbp_batch *b, int k, double *v);
int  sbp_batch_get_state_strided(sbp_batch *b, int k, int stride, double *v);
/* nsteps RK4 steps of every instance, dt[k] per instance (asynchronous). */
int  sbp_batch_step(sbp_batch *b, const double *dt, long nsteps);
int  sbp_batch_synchronize(sbp_batch *b);
int  sbp_batch_time_steps(sbp_batch *b, const double *dt, long nsteps, double *ms);
int  sbp_batch_energy(sbp_batch *b, int k, double *E);
/* Launch-shape candidates (tile width, tile rows, persistent blocks per CU,
 * and whether the step runs two RK4 stages per launch: the schedule of
 * sbp_set_fused, which needs every instance larger than 2(CLS+R) points;
 * sbp_batch_set_config fails with SBP_EINVAL otherwise). */
int  sbp_batch_num_configs(void);
int  sbp_batch_set_config(sbp_batch *b, int config);
int  sbp_batch_get_config(const sbp_batch *b, int *config, int *tile_width, int *tile_rows,
                          int *blocks_per_cu, int *two_stage);
/* Time every candidate on this ensemble (resident states are saved and
 * restored) and keep the fastest; *choice receives its index. */
int  sbp_batch_autotune(sbp_batch *b, const double *dt, int *choice);
/* sbp_set_store_policy for the ensemble's two-stage configurations. */
int  sbp_batch_set_store_policy(sbp_batch *b, int policy);

/* ---- learned coarse-grid corrector ----------------------------------------
 * A 3-layer 3x3 convolutional network (8 -> 32 -> 32 -> 3 channels, ReLU)
 * applied after each coarse RK4 step of a 2D acoustic context; output: the
 * correction of u, v, p.  Inputs at each point (indices clamped into the
 * grid, d4x f = f[i-2]-4f[i-1]+6f[i]-4f[i+1]+f[i+2], Dx f = (f[i+1]-f[i-1])/2,
 * likewise in y):
 *   rho^-1 Dx(d4x p), rho^-1 Dy(d4y p), K (Dx(d4x u) + Dy(d4y v)),
 *   d4x u + d4y u, d4x v + d4y v, d4x p + d4y p, rho, K.
 * bf16 MFMA forward, backward and inference with FP32 accumulation, FP32
 * master weights and Adam.  No counterpart in the reference.  Parameters are
 * a flat FP32 vector:
 * W1[32][8][3][3], b1[32], W2[32][32][3][3], b2[32], W3[3][32][3][3], b3[3].
 * Inputs are normalised as (x - shift) * scale; the network predicts the
 * correction divided by out_scale. */
enum { SBP_CORR_INPUTS = 8, SBP_CORR_OUTPUTS = 3 };
typedef struct sbp_corrector sbp_corrector;
int  sbp_corr_param_count(void);
int  sbp_corr_create(int nx, int ny, int batch, sbp_corrector **out);   /* image size, batch capacity */
int  sbp_corr_destroy(sbp_corrector *c);
int  sbp_corr_set_params(sbp_corrector *c, const float *params, int n); /* also resets Adam */
int  sbp_corr_get_params(sbp_corrector *c, float *params, int n);
int  sbp_corr_set_normalization(sbp_corrector *c, const float *in_shift, const float *in_scale,
                                const float *out_scale);   /* 8, 8 and 3 values */
/* Restrict the correction (and the training loss) to the points at least
 * `band` grid points from every wall; the points nearer a wall (the boundary
 * closure rows) keep the plain RK4 values.  Default 0: every point. */
int  sbp_corr_set_band(sbp_corrector *c, int band);
int  sbp_corr_get_band(const sbp_corrector *c, int *band);
/* raw inputs [nb][ny][nx][8] -> out [nb][ny][nx][3] (normalised units). */
int  sbp_corr_forward(sbp_corrector *c, const float *raw, int nb, float *out);
/* MSE loss of (net - target/out_scale) and its parameter gradients. */
int  sbp_corr_gradients(sbp_corrector *c, const float *raw, const float *target, int nb,
                        float *grads, double *loss);
int  sbp_corr_load_dataset(sbp_corrector *c, const float *raw, const float *target, int n);
int  sbp_corr_train_step(sbp_corrector *c, const int *idx, int nb, float lr, double *loss);
int  sbp_corr_synchronize(sbp_corrector *c);
/* nsteps coarse RK4 steps of ctx, each followed by U += out_scale*net(U);
 * a step whose corrected energy exceeds the state's energy at the start of
 * the call times (1+bound) keeps the plain RK4 result instead (with wall or
 * free-surface SATs and no forcing the exact energy is constant, so this is
 * growth); *rejected counts them. */
int  sbp_corr_advance(sbp_corrector *c, sbp_ctx *ctx, double t0, double dt, long nsteps,
                      double bound, int *rejected);

/* The same over every instance of a batch in one network pass per step;
 * all instances must have the corrector's grid; rejected[k] optional. */
int  sbp_corr_advance_batch(sbp_corrector *c, sbp_batch *b, const double *dt, long nsteps,
                            double bound, int *rejected);

/* ---- diagnostics ---------------------------------------------------------- */
const char *sbp_last_error(void);
/* Name of the device the library runs on (e.g. "gfx950"), or "" without one. */
const char *sbp_device_arch(void);
int  sbp_version(void);
/* Best-of-reps bandwidth (read + write bytes / time) of a plain device copy
 * of `bytes` bytes: the achievable streaming ceiling, and the known-size
 * kernel that calibrates HBM traffic counters.  No reference counterpart. */
int  sbp_copy_bandwidth(long bytes, int reps, double *gbs);
/* Best-of-reps bandwidth of a flat elementwise kernel that reads n_read
 * (1..12) and writes n_write (1..6) planes of plane_bytes each: the rate a
 * launch's byte mix reaches without its stencils, halos or tiles (the
 * dominant launch of the two-stage step reads 6 FP64 planes and writes 3).
 * No reference counterpart. */
int  sbp_stream_bandwidth(long plane_bytes, int n_read, int n_write, int reps, double *gbs);

#ifdef __cplusplus
}
#endif
#endif /* SBP_HIP_H */
