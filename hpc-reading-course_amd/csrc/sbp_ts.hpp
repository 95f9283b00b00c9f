// sbp_ts.hpp -- host-side step-size logic of the time integrators that
// ts_rk4 / ts_rk45 (code/src/time_stepping/ts_rk.cpp:15-65) take from PETSc
// 3.12 TS, and of the reference's own RK4_custom
// (code/include/time_stepping/time_stepping.h:98-152).
//
// PETSc's source is not part of the reference (it is the cluster module
// PETSc/3.12.4, code/batch.sh:6); the rules below restate what its TS does
// with the options ts_rk_setup sets (TS_EXACTFINALTIME_MATCHSTEP, TSADAPTNONE
// for TSRK4, TSADAPTBASIC for TSRK5F) and its defaults:
//   TSSolve          the first step is clipped so it does not overshoot
//                    max_time (PetscIsCloseAtTol(dt, maxdt, 10 eps, 0) counts
//                    as landing on it);
//   TSAdaptChoose    after an ACCEPTED step that ends at t, with the next step
//                    h proposed by the adaptor and hmax = t_max - t:
//                      t + h > t_max                  -> next = hmax
//                      t + h < t_max, 2 h > hmax       -> next = hmax / 2
//                      t + h < t_max, 1.01 h > hmax    -> next = hmax
//                    (matchstepfac {0.01, 2}; later rules override earlier)
//   TSADAPTNONE      next = h (the step just taken)
//   TSADAPTBASIC     enorm > 1 rejects (unless h <= dt_min); the factor is
//                    safety * enorm^(-1/order) with safety 0.9, times the
//                    reject safety 0.5 when the previous attempt of the same
//                    step was rejected too; clipped to [0.1, 10], then the
//                    step to [dt_min, dt_max] = [1e-20, 1e50]
//   TSStep_RK        more than max_reject = 10 rejections of one step end the
//                    solve (TS_DIVERGED_STEP_REJECTED, an error under the
//                    default errorifstepfailed); a non-finite error norm is an
//                    error in TSErrorWeightedNorm (PETSC_ERR_FP).
// Header only, no device code.
#pragma once

#include <cmath>
#include <limits>

namespace sbp {
namespace ts {

constexpr double kSafety = 0.9;
constexpr double kRejectSafety = 0.5;
constexpr double kClipLo = 0.1, kClipHi = 10.0;
constexpr double kDtMin = 1e-20, kDtMax = 1e50;
constexpr double kMatchStepA = 0.01, kMatchStepB = 2.0;
constexpr int kMaxReject = 10;
constexpr double kEps = std::numeric_limits<double>::epsilon();

inline double clip(double x, double lo, double hi) { return x < lo ? lo : (x > hi ? hi : x); }

// TSSolve: the first step does not overshoot t_max.
inline double first_step(double ptime, double dt, double t_max)
{
  const double maxdt = t_max - ptime;
  if (dt >= maxdt) return maxdt;
  // PetscIsCloseAtTol(dt, maxdt, 10 eps, 0): |dt - maxdt| <= 10 eps max(|dt|, |maxdt|)
  const double d = std::fabs(dt - maxdt);
  if (d <= 10 * kEps * std::fmax(std::fabs(dt), std::fabs(maxdt))) return maxdt;
  return dt;
}

// TSAdaptChoose's MATCHSTEP adjustment of the step after an accepted one that
// ended at t (h = the adaptor's proposal).
inline double matchstep(double t, double h, double t_max)
{
  const double tend = t + h, hmax = t_max - t;
  const double a = 1.0 + kMatchStepA, b = kMatchStepB;
  double next = h;
  if (t < t_max && tend > t_max) next = hmax;
  if (t < t_max && tend < t_max && h * b > hmax) next = hmax / 2;
  if (t < t_max && tend < t_max && h * a > hmax) next = hmax;
  return next;
}

// TSAdaptChoose_Basic: accept / reject and the next step from the weighted
// error norm of an attempt of step h with an embedded pair of the given
// order; prev_rejected: the previous attempt of this same step was rejected.
struct Choice {
  bool accept;
  double next_h;
};

inline Choice adapt_basic(double enorm, double h, int order, bool prev_rejected)
{
  double safety = kSafety;
  bool accept = true;
  if (enorm > 1) {
    if (prev_rejected) safety *= kRejectSafety;
    accept = h < (1 + std::sqrt(kEps)) * kDtMin;
  }
  const double hfac = enorm > 0 ? safety * std::pow(enorm, -1.0 / order)
                                : std::numeric_limits<double>::infinity();
  const double h_lte = h * clip(hfac, kClipLo, kClipHi);
  return Choice{accept, clip(h_lte, kDtMin, kDtMax)};
}

// Step sizes of a fixed-step (TSADAPTNONE) MATCHSTEP run from t0 to t_max
// with first step dt, as runs of equal steps.  Returns the number of runs, or
// -1 if more than max_runs would be needed; run_dt / run_len may be null to
// count only.  *t_final (optional) receives the end time ptime reaches.
inline int matchstep_runs(double t0, double t_max, double dt, int max_runs, double *run_dt, long *run_len,
                          double *t_final)
{
  double t = t0, h = first_step(t0, dt, t_max);
  int nr = 0;
  double cur = 0;
  long len = 0;
  while (t < t_max) {
    if (len > 0 && h != cur) {
      if (nr >= max_runs) return -1;
      if (run_dt) run_dt[nr] = cur;
      if (run_len) run_len[nr] = len;
      nr++;
      len = 0;
    }
    cur = h;
    len++;
    const double next = matchstep(t + h, h, t_max);
    t = t + h;
    h = next;
  }
  if (len > 0) {
    if (nr >= max_runs) return -1;
    if (run_dt) run_dt[nr] = cur;
    if (run_len) run_len[nr] = len;
    nr++;
  }
  if (t_final) *t_final = t;
  return nr;
}

// RK4_custom's step count and step (time_stepping.h:104-109): tlen =
// round(Tend/dt); when tlen*dt misses Tend by more than 1e-14 the step
// becomes Tend/tlen.  Returns tlen; *dt_used receives the step.
inline long rk4_custom_steps(double t_end, double dt, double *dt_used, bool *changed)
{
  const long tlen = (long)std::round(t_end / dt);
  bool ch = false;
  if (std::fabs(tlen * dt - t_end) > 1e-14) {
    dt = t_end / tlen;
    ch = true;
  }
  if (dt_used) *dt_used = dt;
  if (changed) *changed = ch;
  return tlen;
}

}  // namespace ts
}  // namespace sbp
