// sbp_appctx.hpp -- C++ adapter that keeps the reference's driver-side shape
// over the C-ABI of sbp_hip.h.
//
// The reference's drivers own an AppCtx on the stack (e.g.
// code/demo/wave/wave_eq_sim.cpp:36-45,119-127), hand PETSc TS a callback
//   PetscErrorCode rhs(TS ts, PetscReal t, Vec v_src, Vec v_dst, void *ctx)
// (wave_eq_sim.cpp:291-308) and advance with
//   ts_rk4(da, t_end, dt, v, rhs, ctx)          (code/src/time_stepping/ts_rk.cpp:54-65).
// Here sbp::AppCtx owns the device context, sbp::rhs has the callback's
// argument order and meaning (vectors are the AoS arrays a Vec holds on one
// rank), sbp::ts_rk4 runs TSRK4 with TS_EXACTFINALTIME_MATCHSTEP
// (ts_rk.cpp:28) and sbp::RK4_custom the reference's own loop
// (time_stepping.h:98-152), with the state kept on the GPU for the whole solve: the
// right-hand side is fused into the RK4 stage kernels, so the rhs callback
// argument of the reference's ts_rk4 has no counterpart.
//
// Header only; link against libsbp_hip.so.  Errors are the library's negative
// codes, returned like PetscErrorCode (0 = success).
#pragma once

#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <string>

#include "sbp_hip.h"

namespace sbp {

// Error check in the style of CHKERRQ: throws with the library's message.
inline void check(int rc)
{
  if (rc != SBP_OK) throw std::runtime_error(std::string("sbp error ") + std::to_string(rc) + ": " + sbp_last_error());
}

class AppCtx {
 public:
  explicit AppCtx(const sbp_problem &p) : problem_(p) { check(sbp_create(&problem_, &ctx_)); }
  ~AppCtx()
  {
    if (ctx_) sbp_destroy(ctx_);
  }
  AppCtx(const AppCtx &) = delete;
  AppCtx &operator=(const AppCtx &) = delete;

  sbp_ctx *get() const { return ctx_; }
  const sbp_problem &problem() const { return problem_; }
  int dofs() const { return sbp_components(ctx_); }
  long length() const { return sbp_vector_length(ctx_); }
  bool is_1d() const { return problem_.physics == SBP_ADVECTION_1D || problem_.physics == SBP_REFLECTION_1D; }
  // hi = (N-1)/(xr-xl) and h = 1/hi as the drivers set them (wave_eq_sim.cpp:86-87,120-121).
  double hix() const { return (problem_.nx - 1) / (problem_.xr - problem_.xl); }
  double hiy() const { return is_1d() ? 0.0 : (problem_.ny - 1) / (problem_.yr - problem_.yl); }

 private:
  sbp_problem problem_;
  sbp_ctx *ctx_ = nullptr;
};

// The rhs / rhs_serial callback: v_dst = F(t, v_src).  The first argument
// stands for the TS handle and is ignored.
inline int rhs(void * /*ts*/, double t, const double *v_src, double *v_dst, void *ctx)
{
  return sbp_rhs(static_cast<AppCtx *>(ctx)->get(), t, v_src, v_dst);
}

// ts_rk4: advance v (AoS, in/out) from t0 to t_end with classical RK4 and
// PETSc's TS_EXACTFINALTIME_MATCHSTEP step schedule (sbp_ts_rk4).
inline int ts_rk4(double t_end, double dt, double *v, AppCtx &ctx, double t0 = 0.0)
{
  int rc = sbp_set_state(ctx.get(), v);
  if (!rc) rc = sbp_ts_rk4(ctx.get(), t0, t_end, dt, nullptr, nullptr);
  if (!rc) rc = sbp_get_state(ctx.get(), v);
  return rc;
}

// RK4_custom(da, Tend, dt, v, rhs, ctx) (time_stepping.h:98-152): round(Tend/dt)
// steps from t = 0, the step rescaled to land on Tend, with the reference's
// prints (PetscPrintf on rank 0 there, stdout here).
inline int RK4_custom(double t_end, double dt, double *v, AppCtx &ctx)
{
  const long tlen = std::lround(t_end / dt);
  if (std::fabs(tlen * dt - t_end) > 1e-14)
    std::printf("Warning: Non-matching time step. dt*tlen = %.12f.\nChanging timestep from %e to %e.\n",
                dt * tlen, dt, t_end / tlen);
  long n = 0;
  double h = dt, t = 0;
  int rc = sbp_set_state(ctx.get(), v);
  if (!rc) rc = sbp_rk4_custom(ctx.get(), t_end, dt, &n, &h, &t);
  if (rc) return rc;
  std::printf("tlen: %ld, dt: %e\n", n, h);
  std::printf("Final t: %.8e\n", t);
  return sbp_get_state(ctx.get(), v);
}

// ts_rk45 (ts_rk.cpp:41-52): adaptive Runge-Kutta-Fehlberg 4(5) of v from t0
// to t_end, first step dt, PETSc's default tolerances unless given.
inline int ts_rk45(double t_end, double dt, double *v, AppCtx &ctx, double t0 = 0.0, double atol = 1e-4,
                   double rtol = 1e-4, long *accepted = nullptr, long *rejected = nullptr)
{
  int rc = sbp_set_state(ctx.get(), v);
  if (!rc) rc = sbp_rk45(ctx.get(), t0, t_end, dt, atol, rtol, 100000000L, accepted, rejected, nullptr);
  if (!rc) rc = sbp_get_state(ctx.get(), v);
  return rc;
}

// error_l2 / error_max of code/include/util/vec_util.h:47-61.
inline int error_norms(const AppCtx &ctx, const double *v, const double *v_ref, double *l2, double *maxerr)
{
  return sbp_error_norms(ctx.get(), v, v_ref, l2, maxerr);
}

}  // namespace sbp
