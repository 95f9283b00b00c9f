// sbp_rk45_f64.hip -- adaptive Runge-Kutta-Fehlberg 4(5) on the resident
// FP64 state: what ts_rk45 (code/src/time_stepping/ts_rk.cpp:41-52) asks of
// PETSc -- TSRK5F (the fifth-order solution is propagated, the embedded
// fourth-order one estimates the error), TSADAPTBASIC, and
// TS_EXACTFINALTIME_MATCHSTEP.  PETSc's source is not part of the reference;
// the step loop restates TSSolve / TSStep_RK / TSAdaptChoose_Basic of PETSc
// 3.12 with their defaults (sbp_ts.hpp):
//   enorm = sqrt(mean((y - y_hat)^2 / tol^2)), tol = atol + rtol*max(|y|, |y_hat|)
//   accept iff enorm <= 1;  next = h * clip(0.9 [* 0.5 after a rejected attempt
//   of the same step] * enorm^(-1/5), 0.1, 10); after an accepted step the
//   MATCHSTEP rule stretches or halves the next step near t_end; more than 10
//   rejections of one step, or a non-finite enorm, fail with SBP_EDIVERGED.
//
// Every attempt is six right-hand-side launches (the RHS mode of the stage
// kernels) plus combination kernels and one weighted-norm reduction, whose
// scalar the host reads to decide.  Built without FMA contraction (file
// suffix _f64) with the oracle's operation order, so the stage values match
// the CPU restatement bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <utility>

#include "../../include/sbp_hip.h"
#include "sbp_ctx.hpp"
#include "sbp_host.hpp"
#include "sbp_ts.hpp"

using sbp::host::fail;
#define HIP_OR_FAIL(expr) SBP_HIP_OR_FAIL(expr)

namespace {

// Fehlberg's coefficients (E. Fehlberg, NASA TR R-315, 1969).
constexpr double kC[6] = {0.0, 1.0 / 4.0, 3.0 / 8.0, 12.0 / 13.0, 1.0, 1.0 / 2.0};
constexpr double kA[6][5] = {{0, 0, 0, 0, 0},
                             {1.0 / 4.0, 0, 0, 0, 0},
                             {3.0 / 32.0, 9.0 / 32.0, 0, 0, 0},
                             {1932.0 / 2197.0, -7200.0 / 2197.0, 7296.0 / 2197.0, 0, 0},
                             {439.0 / 216.0, -8.0, 3680.0 / 513.0, -845.0 / 4104.0, 0},
                             {-8.0 / 27.0, 2.0, -3544.0 / 2565.0, 1859.0 / 4104.0, -11.0 / 40.0}};
constexpr double kB5[6] = {16.0 / 135.0, 0.0, 6656.0 / 12825.0, 28561.0 / 56430.0, -9.0 / 50.0, 2.0 / 55.0};
constexpr double kB4[6] = {25.0 / 216.0, 0.0, 1408.0 / 2565.0, 2197.0 / 4104.0, -1.0 / 5.0, 0.0};

struct Regs {
  const double *K[6];
  long plane;
  int nx, ny, C, pitch, xoff;
};

struct Row {
  double c[6];
};

// Y = U + h * (a[0] K1 + ... + a[s-1] Ks) on the grid points of every plane.
__global__ __launch_bounds__(256) void rkf_stage_kernel(Regs r, int s, Row a, double h, const double *__restrict__ U,
                                                        double *__restrict__ Y)
{
  const long k = blockIdx.x * 256L + threadIdx.x;
  if (k >= (long)r.nx * r.ny) return;
  const int j = (int)(k / r.nx), i = (int)(k - (long)j * r.nx);
  for (int c = 0; c < r.C; c++) {
    const long o = c * r.plane + (long)j * r.pitch + r.xoff + i;
    double sum = a.c[0] * r.K[0][o];
    for (int q = 1; q < s; q++) sum = sum + a.c[q] * r.K[q][o];
    Y[o] = U[o] + h * sum;
  }
}

// Candidate U5 = U + h * (b5 . K), error E = h * ((b5 - b4) . K), and the
// per-block sums of (E / (atol + rtol * max(|U5|, |U5 - E|)))^2.
__global__ __launch_bounds__(256) void rkf_finish_kernel(Regs r, Row b, Row e, double h, double atol, double rtol,
                                                         const double *__restrict__ U, double *__restrict__ U5,
                                                         double *__restrict__ partial)
{
  __shared__ double sh[4];
  double acc = 0;
  const long npts = (long)r.nx * r.ny;
  for (long k = blockIdx.x * 256L + threadIdx.x; k < npts; k += (long)gridDim.x * 256) {
    const int j = (int)(k / r.nx), i = (int)(k - (long)j * r.nx);
    for (int c = 0; c < r.C; c++) {
      const long o = c * r.plane + (long)j * r.pitch + r.xoff + i;
      double sb = b.c[0] * r.K[0][o], se = e.c[0] * r.K[0][o];
      for (int q = 1; q < 6; q++) {
        sb = sb + b.c[q] * r.K[q][o];
        se = se + e.c[q] * r.K[q][o];
      }
      const double u5 = U[o] + h * sb, err = h * se, y4 = u5 - err;
      const double w = err / (atol + rtol * fmax(fabs(u5), fabs(y4)));
      acc += w * w;
      U5[o] = u5;
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

__global__ __launch_bounds__(256) void rkf_sum_kernel(const double *__restrict__ partial, int n, double *__restrict__ out)
{
  __shared__ double sh[4];
  double acc = 0;
  for (int k = threadIdx.x; k < n; k += 256) acc += partial[k];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) out[0] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

int ensure_workspace(sbp_ctx *c)
{
  const size_t bytes = c->esz * (size_t)c->C * (size_t)c->plane;
  for (void *&p : c->rkw) {
    if (p) continue;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
      p = nullptr;
      return fail(SBP_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    // zeroed like the context's own registers (a workspace register can
    // become the state register, padding included)
    HIP_OR_FAIL(hipMemsetAsync(p, 0, bytes, c->stream));
  }
  return SBP_OK;
}

}  // namespace

extern "C" {

int sbp_rk45(sbp_ctx *ctx, double t0, double t_end, double dt0, double atol, double rtol, long max_steps,
             long *accepted, long *rejected, double *dt_next)
{
  if (!ctx) return fail(SBP_EINVAL, "null context");
  if (ctx->prec != SBP_FP64) return fail(SBP_EINVAL, "the adaptive integrator runs on FP64 contexts");
  if (!ctx->have_state) return fail(SBP_ESTATE, "no state loaded");
  if (!(dt0 > 0) || !(t_end >= t0) || !std::isfinite(t_end) || !(atol > 0) || !(rtol >= 0) || max_steps < 1)
    return fail(SBP_EINVAL, "bad time span, step or tolerance");
  int rc = ensure_workspace(ctx);
  if (rc) return rc;
  double *K[6], *Y = static_cast<double *>(ctx->rkw[6]), *U5 = static_cast<double *>(ctx->rkw[7]);
  for (int q = 0; q < 6; q++) K[q] = static_cast<double *>(ctx->rkw[q]);
  Regs r{};
  for (int q = 0; q < 6; q++) r.K[q] = K[q];
  r.plane = ctx->plane;
  r.nx = ctx->nx;
  r.ny = ctx->ny;
  r.C = ctx->C;
  r.pitch = ctx->pitch;
  r.xoff = ctx->xoff;
  Row b{}, e{};
  for (int q = 0; q < 6; q++) {
    b.c[q] = kB5[q];
    e.c[q] = kB5[q] - kB4[q];
  }
  const long npts = (long)ctx->nx * ctx->ny;
  const int pgrid = (int)((npts + 255) / 256);
  const int fgrid = std::min(pgrid, ctx->npartials);
  hipStream_t s = ctx->stream;
  namespace ts = sbp::ts;
  // TSSolve: the first step does not overshoot t_end
  double t = t0, h = ts::first_step(t0, dt0, t_end);
  long acc = 0, rej = 0;
  auto report = [&]() {
    if (accepted) *accepted = acc;
    if (rejected) *rejected = rej;
    if (dt_next) *dt_next = h;
  };
  while (t < t_end) {  // TS_CONVERGED_TIME once ptime >= max_time
    bool prev_rejected = false;  // TSStep_RK's accept flag, reset every step
    int rejections = 0;
    for (;;) {
      if (acc + rej >= max_steps) {
        report();
        return fail(SBP_EINVAL, "step budget exhausted before t_end");
      }
      double *U = static_cast<double *>(ctx->U);
      HIP_OR_FAIL(sbp::host::launch_rhs(ctx, t, U, K[0]));
      for (int st = 1; st < 6; st++) {
        Row a{};
        for (int q = 0; q < st; q++) a.c[q] = kA[st][q];
        hipLaunchKernelGGL(rkf_stage_kernel, dim3(pgrid), dim3(256), 0, s, r, st, a, h, U, Y);
        HIP_OR_FAIL(hipGetLastError());
        HIP_OR_FAIL(sbp::host::launch_rhs(ctx, t + kC[st] * h, Y, K[st]));
      }
      hipLaunchKernelGGL(rkf_finish_kernel, dim3(fgrid), dim3(256), 0, s, r, b, e, h, atol, rtol, U, U5,
                         ctx->partials);
      hipLaunchKernelGGL(rkf_sum_kernel, dim3(1), dim3(256), 0, s, ctx->partials, fgrid, ctx->scalar);
      HIP_OR_FAIL(hipGetLastError());
      double ss = 0;
      HIP_OR_FAIL(hipMemcpyAsync(&ss, ctx->scalar, sizeof(double), hipMemcpyDeviceToHost, s));
      HIP_OR_FAIL(hipStreamSynchronize(s));
      const double enorm = std::sqrt(ss / (double)(npts * ctx->C));
      if (!std::isfinite(enorm)) {  // TSErrorWeightedNorm: PETSC_ERR_FP
        report();
        return fail(SBP_EDIVERGED, "non-finite error estimate (the state is not finite)");
      }
      const ts::Choice ch = ts::adapt_basic(enorm, h, 5, prev_rejected);
      if (ch.accept) {
        // the candidate becomes the state: swap the registers (the captured
        // RK4 step graph refers to the old one)
        std::swap(ctx->U, ctx->rkw[7]);
        U5 = static_cast<double *>(ctx->rkw[7]);
        sbp::host::drop_step_graph(ctx);
        const double next = ts::matchstep(t + h, ch.next_h, t_end);
        t = t + h;
        h = next;
        acc++;
        break;
      }
      rej++;
      h = ch.next_h;
      prev_rejected = true;
      if (++rejections > ts::kMaxReject) {  // TS_DIVERGED_STEP_REJECTED
        report();
        return fail(SBP_EDIVERGED, "step rejected more than 10 times (TS_DIVERGED_STEP_REJECTED)");
      }
    }
  }
  report();
  return SBP_OK;
}

}  // extern "C"
