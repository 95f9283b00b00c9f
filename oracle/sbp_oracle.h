/*
 * sbp_oracle.h -- CPU restatement of the reference SBP-SAT solver's hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, loads or
 * calls this code; it exists so that tests, the smoke check and the CPU leg
 * of the benchmark can compare the HIP path against an independent FP64
 * computation.  Problem descriptions use the public sbp_problem struct of
 * include/sbp_hip.h (data layout only).
 *
 * Arrays are the reference's single-rank AoS layout v[j][i][c] at
 * C*(i + nx*j) + c (code/include/grids/layout.h:118-127).
 */
#ifndef SBP_ORACLE_H
#define SBP_ORACLE_H
#include "../include/sbp_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Operator tables for one order (code/include/sbpops/D1_central.h:243-273,
 * H_central.h:193-212, HI_central.h:128-147). */
typedef struct {
  int order, r, cls, w;        /* radius, closure rows, closure width          */
  double inner[7];             /* interior stencil, 2r+1 taps                 */
  double Q[6][9];              /* left closure rows cls x w                   */
  double hq[6];                /* closure quadrature weights                  */
  double hqinv[6];             /* their inverses as tabulated by the reference */
} orc_ops;

int    orc_get_ops(int order, orc_ops *out);
int    orc_components(const sbp_problem *p);
int    orc_validate(const sbp_problem *p);

/* 1D derivative of a strided line v[k*stride], k = 0..n-1, at index i. */
double orc_d1(const orc_ops *op, const double *v, long stride, int n, int i, double hi);

/* Semi-discrete right-hand side, rhs_serial of each demo. */
int    orc_rhs(const sbp_problem *p, double t, const double *src, double *dst);
/* Classical RK4 per RK4_custom (time_stepping.h:127-142), fixed dt. */
int    orc_rk4(const sbp_problem *p, double *v, double t0, double dt, long nsteps);
/* The same step on nthreads threads over a near-square block decomposition
 * (the DMDA-like parallel path of wave_eq_sim.cpp:246-268); nthreads<=1 is serial. */
int    orc_rk4_par(const sbp_problem *p, double *v, double t0, double dt, long nsteps,
                   int nthreads);

/* ts_rk4 (ts_rk.cpp:54-65): TSRK4 with TSADAPTNONE under
 * TS_EXACTFINALTIME_MATCHSTEP from t0 to t_end, first step dt. */
int    orc_ts_rk4(const sbp_problem *p, double *v, double t0, double t_end, double dt, long *nsteps,
                  double *t_final, int nthreads);
/* The step sizes of that run (at most max_steps; returns the count or -1). */
int    orc_matchstep_steps(double t0, double t_end, double dt, long max_steps, double *steps,
                           double *t_final);
/* TSAdaptChoose_Basic's decision for one attempt. */
void   orc_adapt_basic(double enorm, double h, int order, int prev_rejected, int *accept,
                       double *next_h);
/* Adaptive Runge-Kutta-Fehlberg 4(5) from t0 to t_end (ts_rk45, ts_rk.cpp:41-52:
 * TSRK5F + TSADAPTBASIC + MATCHSTEP).  Returns 0, -1 on bad input, -2 when
 * max_steps attempts did not reach t_end, -3 when a step was rejected more
 * than 10 times or the error estimate was not finite. */
int    orc_rk45(const sbp_problem *p, double *v, double t0, double t_end, double dt0, double atol,
                double rtol, long max_steps, long *accepted, long *rejected, double *dt_next);
void   orc_rkf45_tableau(double c[6], double a[6][5], double b5[6], double b4[6]);

double orc_energy(const sbp_problem *p, const double *v);
double orc_hnorm(const sbp_problem *p, const double *v);
void   orc_error_norms(const sbp_problem *p, const double *v, const double *w,
                       double *l2, double *maxerr);
/* Grid spacing h and inverse hi of one direction (drivers: hi = (N-1)/(xr-xl)). */
double orc_hi(int n, double lo, double hi_end);

#ifdef __cplusplus
}
#endif
#endif
