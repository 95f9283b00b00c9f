// sbp_capi.hip -- host side of the C-ABI declared in include/sbp_hip.h.
//
// A context owns the device state of one problem: the RK4 registers U, A and
// the Y ping-pong pair (planar, padded rows), optional Kahan compensation and
// material planes, the 1D tables of the reference's manufactured forcing, an
// AoS staging buffer for the reference-layout vectors that cross the
// boundary, a stream and a captured multi-step hipGraph.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/sbp_hip.h"
#include "sbp_ctx.hpp"
#include "sbp_host.hpp"
#include "sbp_kernels.hpp"
#include "sbp_ts.hpp"

using namespace sbp;

namespace {
thread_local std::string g_err;
}

int sbp::host::fail(int code, const std::string &msg)
{
  g_err = msg;
  return code;
}

namespace {

using sbp::host::fail;
#define HIP_OR_FAIL(expr) SBP_HIP_OR_FAIL(expr)

constexpr int kAlign = 64;   // row alignment in elements
constexpr int kBX = 256;     // interior tile width (threads per block)
constexpr int kGraphSteps = 16;

int closure_of(int order) { return order == 2 ? 1 : (order == 4 ? 4 : 6); }
int taps_of(int order) { return order == 2 ? 3 : (order == 4 ? 6 : 9); }
int min_points(int order) { return std::max(2 * closure_of(order), taps_of(order)); }

bool is1d(int phys) { return phys == SBP_ADVECTION_1D || phys == SBP_REFLECTION_1D; }

int components(int phys)
{
  switch (phys) {
    case SBP_ADVECTION_1D:
    case SBP_ADVECTION_2D: return 1;
    case SBP_ACOUSTIC_2D: return 3;
    case SBP_REFLECTION_1D: return 2;
  }
  return 0;
}

}  // namespace


namespace {

void compute_geometry(sbp_ctx *c)
{
  if (is1d(c->phys)) {
    c->geom.grid = (c->nx + 255) / 256;
    c->geom.block = 256;
    return;
  }
  const int wi = c->nx - 2 * c->cls, hi = c->ny - 2 * c->cls;
  c->tiles_x = wi > 0 ? (wi + kBX - 1) / kBX : 0;
  c->tiles_y = hi > 0 ? (hi + c->ty_rows - 1) / c->ty_rows : 0;
  c->n_tiles = c->tiles_x * c->tiles_y;
  c->n_edge = 2 * c->cls * c->nx + 2 * c->cls * (c->ny - 2 * c->cls);
  int eb = (c->n_edge + kBX - 1) / kBX;
  c->n_edge_blocks = (eb + kNumXcd - 1) / kNumXcd * kNumXcd;
  c->geom.grid = c->n_edge_blocks + c->n_tiles;
  c->geom.block = kBX;
}

// Default tile height: the tallest of 4..64 rows that still gives about four
// tiles per CU (>= 960 tiles on 256 CUs).  Tall tiles amortise the 2R-row y
// halo; on grids too small for that, occupancy wins (measured, order 4:
// 4097^2 -> 64 rows, 2049^2 -> 16, 1025^2 -> 4 rows at 1.25x the 16-row rate).
// Two-stage interior tiles: pair_nt mid columns, pair_nt - 2R output
// columns, inside the D = CLS + R deep frame.

void compute_pair_geometry(sbp_ctx *c)
{
  const int R = c->order / 2;
  const int D = c->cls + R;
  c->p_depth = D;
  const int wi = c->nx - 2 * D, hi = c->ny - 2 * D;
  const int nt = pair_nt((int)c->esz, c->order);
  const int bxo = nt - 2 * R;
  c->p_tiles_x = wi > 0 ? (wi + bxo - 1) / bxo : 0;
  c->p_n_tiles = (wi > 0 && hi > 0) ? c->p_tiles_x * ((hi + c->p_ty_rows - 1) / c->p_ty_rows) : 0;
  c->p_n_edge = 2 * D * c->nx + 2 * D * (c->ny - 2 * D);
  // one block per frame patch, rounded up to whole XCD rounds so the tiles'
  // block index stays XCD-aligned
  const int eb = frame_patch_count(c->nx, c->ny, D);
  c->p_n_edge_blocks = (eb + kNumXcd - 1) / kNumXcd * kNumXcd;
}

// Default pair-tile height, from sweeps on gfx950 (ms per two-stage step):
// FP64 orders 2 and 4 take about ny/100 rows, at most 40 (4097^2: 40 rows
// 0.460 against 32 rows 0.465 and 48 rows 0.468; 1025^2: 10 rows 0.063
// against 8 rows 0.071);
// FP64 order 6 136 rows (4097^2 with rho/K arrays: 0.70 against 68 rows
// 0.75 and 48 rows 0.80; rho = K = 1: 0.59 against 0.61), its 4R = 12 halo
// rows per tile favouring tall tiles; FP32 16 rows (order 6: 48).  Halved
// while the grid would have fewer than 400 tiles.
// sbp_tune_fused measures the candidates on the real grid instead.
int default_pair_rows(const sbp_ctx *c)
{
  const int R = c->order / 2, D = c->cls + R;
  const int wi = c->nx - 2 * D, hi = c->ny - 2 * D;
  if (wi <= 0 || hi <= 0) return 16;
  const int nt = pair_nt((int)c->esz, c->order);
  const int tx = (wi + nt - 2 * R - 1) / (nt - 2 * R);
  int t;
  if (c->esz == 8)
    t = c->order == 6 ? 136 : std::min(40, std::max(8, (hi + 50) / 100 / 2 * 2));
  else
    t = c->order == 6 ? 48 : 16;
  while (t > 8 && (long)tx * ((hi + t - 1) / t) < 400) t /= 2;
  return t;
}

bool pair_supported(const sbp_ctx *c)
{
  const int D = c->cls + c->order / 2;
  return c->pair_launch && (c->phys == SBP_ACOUSTIC_2D || c->phys == SBP_ADVECTION_2D) && c->nx > 2 * D &&
         c->ny > 2 * D;
}

// Two-stage stepping is the default where it was measured faster: grids of
// at least 450^2 interior points.  2D acoustics, order 4 FP64, one-stage /
// two-stage ms per step with the frame patches (rho = K = 1, then rho/K
// arrays): 385^2 0.0201 / 0.0236 and 0.0213 / 0.0237; 513^2 0.0278 /
// 0.0244 and 0.0342 / 0.0246; 769^2 0.053 / 0.035 and 0.059 / 0.037; 1025^2
// 0.080 / 0.040 and 0.091 / 0.051; 4097^2 0.98 / 0.45.  2D advection, order
// 4, one-stage / two-stage ms, constant velocity and
// per-point fields: 1025^2 FP64 0.033 / 0.025 and 0.045 / 0.027, FP32 0.019 /
// 0.018 and 0.026 / 0.020; 2049^2 FP64 0.096 / 0.052 and 0.131 / 0.063, FP32
// 0.064 / 0.030 and 0.078 / 0.036; 4097^2 FP64 0.316 / 0.173 and 0.449 /
// 0.261, FP32 0.220 / 0.079 and 0.264 / 0.116.
bool pair_default(const sbp_ctx *c)
{
  const int D = c->cls + c->order / 2;
  return pair_supported(c) && (long)(c->nx - 2 * D) * (c->ny - 2 * D) >= 450L * 450L;
}

int default_ty_rows(int nx, int ny, int cls)
{
  const int wi = nx - 2 * cls, hi = ny - 2 * cls;
  if (wi <= 0 || hi <= 0) return 16;
  const int tx = (wi + kBX - 1) / kBX;
  int best = 4;
  for (int t = 4; t <= 64; t *= 2)
    if ((long)tx * ((hi + t - 1) / t) >= 960) best = t;
  return best;
}

template <typename T>
StageArgs<T> make_args(const sbp_ctx *c, int mode, const void *yin, double dt, double t)
{
  StageArgs<T> a{};
  a.yin = static_cast<const T *>(yin);
  a.uin = static_cast<const T *>(c->U);
  a.y2 = static_cast<const T *>(c->Y0);
  a.y3 = static_cast<const T *>(c->Y1);
  a.third = T(1.0 / 3.0);
  a.aio = static_cast<T *>(c->A);
  a.uout = static_cast<T *>(c->U);
  a.comp = static_cast<T *>(c->KC);
  a.tk = static_cast<T *>(c->TK);
  a.rinv = static_cast<const T *>(c->rinv);
  a.bulk = static_cast<const T *>(c->bulk);
  a.velx = static_cast<const T *>(c->velx);
  a.vely = static_cast<const T *>(c->vely);
  a.xs = c->xs;
  a.ys = c->ys;
  a.c3x = c->c3x;
  a.s3x = c->s3x;
  a.c4y = c->c4y;
  a.s4y = c->s4y;
  a.c5t = std::cos(5 * kPi * t);
  a.plane = c->plane;
  a.pitch = c->pitch;
  a.xoff = c->xoff;
  a.nx = c->nx;
  a.ny = c->ny;
  a.hix = T(c->hix);
  a.hiy = T(c->hiy);
  a.ax = T(c->d.ax);
  a.ay = T(c->d.ay);
  a.bc = c->bc;
  a.tiles_x = c->tiles_x;
  a.ty_rows = c->ty_rows;
  a.n_tiles = c->n_tiles;
  a.n_edge_blocks = c->n_edge_blocks;
  a.n_edge = c->n_edge;
  switch (mode) {
    case MODE_S1: a.cy = T(0.5 * dt); a.ca = T(dt / 6); break;
    case MODE_S2: a.cy = T(0.5 * dt); a.ca = T(dt / 3); break;
    case MODE_S3: a.cy = T(dt); a.ca = T(dt / 3); break;
    case MODE_S4: a.cy = T(0); a.ca = T(dt / 6); break;
    default: a.cy = a.ca = T(0);
  }
  return a;
}

// Stage operand and output of stage s (1..4).
//   2D: U -> Y0 (=Y2), Y0 -> Y1 (=Y3), Y1 -> A (=Y4), A -> U (reads Y0, Y1)
//   1D: U -> Y0, Y0 -> Y1, Y1 -> Y0, Y0 -> U (stage vectors in A and TK)
void stage_io(const sbp_ctx *c, int s, const void **yin, void **yout)
{
  if (is1d(c->phys)) {
    const void *in[4] = {c->U, c->Y0, c->Y1, c->Y0};
    void *out[4] = {c->Y0, c->Y1, c->Y0, nullptr};
    *yin = in[s - 1];
    *yout = out[s - 1];
  } else {
    const void *in[4] = {c->U, c->Y0, c->Y1, c->A};
    void *out[4] = {c->Y0, c->Y1, c->A, nullptr};
    *yin = in[s - 1];
    *yout = out[s - 1];
  }
}

// Launch one stage (mode 1..4 of the RK4 step, roles from stage_io) or the
// RHS mode (yin -> yout).
hipError_t launch_stage(sbp_ctx *c, int mode, double dt, double t, const void *yin, void *yout)
{
  if (c->esz == 8) {
    StageArgs<double> a = make_args<double>(c, mode, yin, dt, t);
    a.yout = static_cast<double *>(yout);
    return c->launch(mode, &a, c->geom, c->stream);
  }
  StageArgs<float> a = make_args<float>(c, mode, yin, dt, t);
  a.yout = static_cast<float *>(yout);
  return c->launch(mode, &a, c->geom, c->stream);
}

template <typename T> PairArgs<T> make_pair_args(const sbp_ctx *c, int pair, double t, double dt)
{
  PairArgs<T> a{};
  // rho = 2 + xy and its forcing (SBP_MAT_REF_RHO): tables and cos(5 pi t)
  // at the stage times t, t + dt/2 (stages 2, 3) and t + dt, as make_args forms them
  a.xs = c->xs;
  a.ys = c->ys;
  a.c3x = c->c3x;
  a.s3x = c->s3x;
  a.c4y = c->c4y;
  a.s4y = c->s4y;
  const double c5_1 = std::cos(5 * kPi * t), c5_23 = std::cos(5 * kPi * (t + 0.5 * dt)),
               c5_4 = std::cos(5 * kPi * (t + dt));
  a.c5_mid = pair == 1 ? c5_1 : c5_23;
  a.c5_out = pair == 1 ? c5_23 : c5_4;
  a.c5_k1 = c5_1;
  a.rinv = static_cast<const T *>(c->rinv);
  a.bulk = static_cast<const T *>(c->bulk);
  a.velx = static_cast<const T *>(c->velx);
  a.vely = static_cast<const T *>(c->vely);
  a.ax = T(c->d.ax);
  a.ay = T(c->d.ay);
  a.plane = c->plane;
  a.pitch = c->pitch;
  a.xoff = c->xoff;
  a.nx = c->nx;
  a.ny = c->ny;
  a.hix = T(c->hix);
  a.hiy = T(c->hiy);
  a.third = T(1.0 / 3.0);
  a.tiles_x = c->p_tiles_x;
  a.ty_rows = c->p_ty_rows;
  a.n_tiles = c->p_n_tiles;
  a.n_edge_blocks = c->p_n_edge_blocks;
  // bytes a two-stage step streams: U, Y3, U' (in the Y2 array), materials or
  // velocity fields, compensation
  const bool fields = c->rinv || c->velx;
  const double streamed = (double)c->plane * c->esz * (3 * c->C + (fields ? 2 : 0) + (c->KC ? c->C : 0));
  a.nts = c->store_policy == 0 ? (pair_cached_stores(streamed) ? 0 : 1) : (c->store_policy == 2 ? 1 : 0);
  if (pair == 1) {  // stages 1-2: U -> Y3 (Y1), Y2 (Y0) next to the frame
    a.yin = static_cast<const T *>(c->U);
    a.o1 = static_cast<T *>(c->Y0);
    a.o2 = static_cast<T *>(c->Y1);
    a.c1 = T(0.5 * dt);
    a.c2 = T(0.5 * dt);
  } else {  // stages 3-4: Y3, U -> U' (into Y0), Y4 (A) next to the frame
    a.yin = static_cast<const T *>(c->Y1);
    a.u = static_cast<const T *>(c->U);
    a.y2 = static_cast<const T *>(c->Y0);
    a.o1 = static_cast<T *>(c->Y0);
    a.o2 = static_cast<T *>(c->A);
    a.comp = static_cast<T *>(c->KC);
    a.c1 = T(dt);
    a.c2 = T(dt / 6);
    a.c0 = T(0.5 * dt);
  }
  return a;
}

// Launch `part` of a two-stage step: 1 = stages 1-2 (frame patches and
// interior tiles), 3 = stages 3-4, after which U' sits in Y0 and the two
// arrays swap roles.
template <typename T> hipError_t launch_pair_part_t(sbp_ctx *c, int part, double t, double dt)
{
  static const int mode[4] = {MODE_S1, MODE_S2, MODE_S3, MODE_S4};
  const double ts[4] = {t, t + 0.5 * dt, t + 0.5 * dt, t + dt};
  const void *yin[4] = {c->U, c->Y0, c->Y1, c->A};
  void *yout[4] = {c->Y0, c->Y1, c->A, nullptr};
  StageArgs<T> e[2];
  for (int k = 0; k < 2; k++) {
    const int s = part - 1 + k;  // stage index 0..3
    e[k] = make_args<T>(c, mode[s], yin[s], dt, ts[s]);
    e[k].yout = static_cast<T *>(yout[s]);
    e[k].depth = c->p_depth;
    e[k].n_tiles = 0;
  }
  if (part == 3) e[1].uout = static_cast<T *>(c->Y0);  // U' into the Y2 array; U read from uin
  const PairArgs<T> p = make_pair_args<T>(c, part == 1 ? 1 : 2, t, dt);
  const hipError_t err = c->pair_launch(part, &e[0], &e[1], &p, c->p_n_edge_blocks + c->p_n_tiles, c->stream);
  // the step is complete (or captured) only when its last launch went in
  if (part == 3 && err == hipSuccess) std::swap(c->U, c->Y0);
  return err;
}

hipError_t launch_pair_part(sbp_ctx *c, int part, double t, double dt)
{
  return c->esz == 8 ? launch_pair_part_t<double>(c, part, t, dt) : launch_pair_part_t<float>(c, part, t, dt);
}

hipError_t launch_step(sbp_ctx *c, double t, double dt)
{
  if (c->fused) {
    hipError_t e = launch_pair_part(c, 1, t, dt);
    if (e == hipSuccess) e = launch_pair_part(c, 3, t, dt);
    return e;
  }
  const double h = 0.5 * dt;
  const double ts[4] = {t, t + h, t + h, t + dt};  // stage times (RK4_custom)
  hipError_t e = hipSuccess;
  for (int s = 1; s <= 4 && e == hipSuccess; s++) {
    const void *yin;
    void *yout;
    stage_io(c, s, &yin, &yout);
    e = launch_stage(c, s, dt, ts[s - 1], yin, yout);
  }
  return e;
}

bool time_dependent(const sbp_ctx *c) { return c->phys == SBP_ACOUSTIC_2D && c->mat == SBP_MAT_REF_RHO; }

void drop_graph(sbp_ctx *c)
{
  for (auto &g : c->gslot) {
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    g = sbp_ctx::GraphSlot{};
  }
  c->g_next = 0;
}

// The graph of kGraphSteps steps for this dt and the current U array
// (captured on first use; kGraphSteps is even, so a two-stage capture ends
// with the arrays back in their starting roles).
int ensure_graph(sbp_ctx *c, double dt, hipGraphExec_t *out)
{
  for (auto &g : c->gslot)
    if (g.exec && g.dt == dt && g.u == c->U) {
      *out = g.exec;
      return SBP_OK;
    }
  sbp_ctx::GraphSlot &slot = c->gslot[c->g_next];
  c->g_next ^= 1;
  if (slot.exec) (void)hipGraphExecDestroy(slot.exec);
  slot = sbp_ctx::GraphSlot{};
  // Captured two-stage steps swap the U and Y0 roles as if they had run; a
  // capture that fails part-way must leave the roles as they were.
  void *u0 = c->U, *y0 = c->Y0;
  hipGraph_t g = nullptr;
  HIP_OR_FAIL(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  hipError_t le = hipSuccess;
  for (int s = 0; s < kGraphSteps && le == hipSuccess; s++) le = launch_step(c, 0.0, dt);
  hipError_t ce = hipStreamEndCapture(c->stream, &g);
  if (le != hipSuccess || ce != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    c->U = u0;
    c->Y0 = y0;
    return fail(SBP_EHIP, std::string("graph capture: ") + hipGetErrorString(le != hipSuccess ? le : ce));
  }
  hipError_t ie = hipGraphInstantiate(&slot.exec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (ie != hipSuccess) {
    slot.exec = nullptr;
    return fail(SBP_EHIP, std::string("graph instantiate: ") + hipGetErrorString(ie));
  }
  if (c->U != u0) {
    c->U = u0;
    c->Y0 = y0;
    return fail(SBP_EHIP, "graph capture left the state in another array");
  }
  slot.dt = dt;
  slot.u = u0;
  *out = slot.exec;
  return SBP_OK;
}

// nsteps steps of dt from t0; *t_end (optional) receives the time the steps
// accumulate to, t = t + dt per step as both integrators of the reference
// advance it (RK4_custom time_stepping.h:141; PETSc's ptime += time_step).
int run_steps(sbp_ctx *c, double t0, double dt, long nsteps, double *t_end = nullptr)
{
  long n = 0;
  if (!time_dependent(c) && nsteps >= kGraphSteps) {
    hipGraphExec_t ge = nullptr;
    int rc = ensure_graph(c, dt, &ge);
    if (rc) return rc;
    for (; n + kGraphSteps <= nsteps; n += kGraphSteps) HIP_OR_FAIL(hipGraphLaunch(ge, c->stream));
  }
  // Step times accumulate as t = t + dt (RK4_custom, time_stepping.h:141).
  double t = t0;
  for (long k = 0; k < n; k++) t = t + dt;
  for (; n < nsteps; n++) {
    HIP_OR_FAIL(launch_step(c, t, dt));
    t = t + dt;
  }
  if (t_end) *t_end = t;
  return SBP_OK;
}

int upload_aos(sbp_ctx *c, const double *v)
{
  const size_t bytes = sizeof(double) * (size_t)c->C * c->nx * c->ny;
  HIP_OR_FAIL(hipMemcpyAsync(c->aos, v, bytes, hipMemcpyHostToDevice, c->stream));
  return SBP_OK;
}

int download_aos(sbp_ctx *c, double *v)
{
  const size_t bytes = sizeof(double) * (size_t)c->C * c->nx * c->ny;
  HIP_OR_FAIL(hipMemcpyAsync(v, c->aos, bytes, hipMemcpyDeviceToHost, c->stream));
  HIP_OR_FAIL(hipStreamSynchronize(c->stream));
  return SBP_OK;
}

int aos_to_planar(sbp_ctx *c, void *dst, void *comp)
{
  const long n = (long)c->nx * c->ny;
  const int grid = (int)((n + 255) / 256);
  if (c->esz == 8)
    hipLaunchKernelGGL(aos_to_planar_kernel<double>, dim3(grid), dim3(256), 0, c->stream, c->aos,
                       (double *)dst, (double *)comp, c->nx, c->ny, c->C, c->pitch, c->xoff, c->plane);
  else
    hipLaunchKernelGGL(aos_to_planar_kernel<float>, dim3(grid), dim3(256), 0, c->stream, c->aos,
                       (float *)dst, (float *)comp, c->nx, c->ny, c->C, c->pitch, c->xoff, c->plane);
  HIP_OR_FAIL(hipGetLastError());
  return SBP_OK;
}

int planar_to_aos(sbp_ctx *c, const void *src, const void *comp)
{
  const long n = (long)c->nx * c->ny;
  const int grid = (int)((n + 255) / 256);
  if (c->esz == 8)
    hipLaunchKernelGGL(planar_to_aos_kernel<double>, dim3(grid), dim3(256), 0, c->stream,
                       (const double *)src, (const double *)comp, c->aos, c->nx, c->ny, c->C, c->pitch,
                       c->xoff, c->plane);
  else
    hipLaunchKernelGGL(planar_to_aos_kernel<float>, dim3(grid), dim3(256), 0, c->stream,
                       (const float *)src, (const float *)comp, c->aos, c->nx, c->ny, c->C, c->pitch,
                       c->xoff, c->plane);
  HIP_OR_FAIL(hipGetLastError());
  return SBP_OK;
}

// H-weighted sum of squares of either the AoS staging vector (aos=true) or a
// planar state (u, comp) laid out like the registers.  The result lands in
// c->scalar on the device; with out != nullptr it is also copied back.
int norm_reduce_planar(sbp_ctx *c, bool aos, bool weighted, const void *u, const void *comp, double *out)
{
  const long n = (long)c->nx * c->ny;
  int grid = (int)std::min<long>((n + 255) / 256, c->npartials);
  if (grid < 1) grid = 1;
  auto fill = [&](auto *tag) {
    using T = std::remove_pointer_t<decltype(tag)>;
    NormArgs<T> a{};
    a.aos = aos ? c->aos : nullptr;
    a.u = static_cast<const T *>(u);
    a.comp = static_cast<const T *>(comp);
    a.rinv = static_cast<const T *>(c->rinv);
    a.bulk = static_cast<const T *>(c->bulk);
    a.xs = c->xs;
    a.ys = c->ys;
    a.plane = c->plane;
    a.pitch = c->pitch;
    a.xoff = c->xoff;
    a.nx = c->nx;
    a.ny = c->ny;
    a.nc = c->C;
    a.order = c->order;
    a.is1d = is1d(c->phys);
    a.material = c->phys == SBP_ACOUSTIC_2D ? c->mat : 0;
    a.weighted = weighted;
    a.hx = 1.0 / c->hix;
    a.hy = is1d(c->phys) ? 1.0 : 1.0 / c->hiy;
    hipLaunchKernelGGL(norm_partial_kernel<T>, dim3(grid), dim3(256), 0, c->stream, a, c->partials);
  };
  if (c->esz == 8) fill((double *)nullptr);
  else fill((float *)nullptr);
  HIP_OR_FAIL(hipGetLastError());
  hipLaunchKernelGGL(norm_final_kernel, dim3(1), dim3(256), 0, c->stream, c->partials, grid, c->scalar);
  HIP_OR_FAIL(hipGetLastError());
  if (out) {
    HIP_OR_FAIL(hipMemcpyAsync(out, c->scalar, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_OR_FAIL(hipStreamSynchronize(c->stream));
  }
  return SBP_OK;
}

int norm_reduce(sbp_ctx *c, bool aos, bool weighted, double *out)
{
  return norm_reduce_planar(c, aos, weighted, c->U, c->KC, out);
}

int validate(const sbp_problem *d)
{
  if (!d) return fail(SBP_EINVAL, "null problem");
  if (components(d->physics) == 0) return fail(SBP_EINVAL, "unknown physics");
  if (d->order != 2 && d->order != 4 && d->order != 6) return fail(SBP_EINVAL, "order must be 2, 4 or 6");
  const int nmin = min_points(d->order);
  if (d->nx < nmin) return fail(SBP_EINVAL, "nx below the closure size of this order");
  if (is1d(d->physics)) {
    if (d->ny != 1) return fail(SBP_EINVAL, "1D problems need ny = 1");
  } else if (d->ny < nmin) {
    return fail(SBP_EINVAL, "ny below the closure size of this order");
  }
  if (!(d->xr > d->xl) || (!is1d(d->physics) && !(d->yr > d->yl))) return fail(SBP_EINVAL, "empty domain");
  if (d->precision < SBP_FP64 || d->precision > SBP_FP32_KAHAN) return fail(SBP_EINVAL, "bad precision");
  if (d->bc < SBP_BC_DEFAULT || d->bc > SBP_BC_WALL) return fail(SBP_EINVAL, "bad boundary condition");
  if (d->physics == SBP_ACOUSTIC_2D) {
    if (d->material < SBP_MAT_UNIT || d->material > SBP_MAT_ARRAYS) return fail(SBP_EINVAL, "bad material");
    if (d->material == SBP_MAT_ARRAYS && (!d->rho_inv || !d->bulk))
      return fail(SBP_EINVAL, "SBP_MAT_ARRAYS needs rho_inv and bulk");
  } else if (d->bc != SBP_BC_DEFAULT) {
    return fail(SBP_EINVAL, "boundary kinds other than default are acoustic only");
  }
  if ((d->vel_x || d->vel_y) && d->physics != SBP_ADVECTION_1D && d->physics != SBP_ADVECTION_2D)
    return fail(SBP_EINVAL, "velocity fields are for advection problems only");
  if (d->vel_y && d->physics == SBP_ADVECTION_1D) return fail(SBP_EINVAL, "1D advection has no y-velocity");
  if ((long)d->nx * d->ny > (1L << 30)) return fail(SBP_EINVAL, "grid too large");
  return SBP_OK;
}

template <typename P> int dalloc(P **p, size_t bytes)
{
  void *q = nullptr;
  hipError_t e = hipMalloc(&q, bytes);
  if (e != hipSuccess) return fail(SBP_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  *p = static_cast<P *>(q);
  return SBP_OK;
}

int upload_plane(sbp_ctx *c, const double *m, void **dst)
{
  const size_t n = (size_t)c->nx * c->ny;
  int rc = dalloc(dst, c->esz * (size_t)c->plane);
  if (rc) return rc;
  HIP_OR_FAIL(hipMemcpyAsync(c->aos, m, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
  const int grid = (int)((n + 255) / 256);
  if (c->esz == 8)
    hipLaunchKernelGGL(matrix_to_plane_kernel<double>, dim3(grid), dim3(256), 0, c->stream, c->aos,
                       (double *)*dst, c->nx, c->ny, c->pitch, c->xoff);
  else
    hipLaunchKernelGGL(matrix_to_plane_kernel<float>, dim3(grid), dim3(256), 0, c->stream, c->aos,
                       (float *)*dst, c->nx, c->ny, c->pitch, c->xoff);
  HIP_OR_FAIL(hipGetLastError());
  HIP_OR_FAIL(hipStreamSynchronize(c->stream));
  return SBP_OK;
}

int upload_table(sbp_ctx *c, const std::vector<double> &v, double **dst)
{
  int rc = dalloc(dst, sizeof(double) * v.size());
  if (rc) return rc;
  HIP_OR_FAIL(hipMemcpy(*dst, v.data(), sizeof(double) * v.size(), hipMemcpyHostToDevice));
  return SBP_OK;
}

void release(sbp_ctx *c)
{
  if (!c) return;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  drop_graph(c);
  void *ptrs[] = {c->U,   c->A,   c->Y0,  c->Y1,  c->KC,  c->TK,  c->rinv,     c->bulk,     c->velx, c->vely,
                  c->xs,  c->ys,  c->c3x, c->s3x, c->c4y, c->s4y, c->aos,      c->aos2,     c->partials, c->scalar};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  for (void *p : c->rkw)
    if (p) (void)hipFree(p);
  for (hipEvent_t &e : c->ev)
    if (e) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

int build_ctx(const sbp_problem *desc, sbp_ctx *c)
{
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(SBP_ENODEV, "no HIP device");
  HIP_OR_FAIL(hipGetDevice(&c->device));
  c->d = *desc;
  c->d.rho_inv = nullptr;
  c->d.bulk = nullptr;
  c->d.vel_x = nullptr;
  c->d.vel_y = nullptr;
  c->phys = desc->physics;
  c->order = desc->order;
  c->cls = closure_of(c->order);
  c->C = components(c->phys);
  c->mat = c->phys == SBP_ACOUSTIC_2D ? desc->material : SBP_MAT_UNIT;
  // 2D advection with velocity fields runs the kernels' per-point (M_ARR) form;
  // 1D advection reads the field whenever its plane exists
  const bool vel_fields = desc->vel_x || desc->vel_y;
  if (c->phys == SBP_ADVECTION_2D && vel_fields) c->mat = SBP_MAT_ARRAYS;
  c->bc = desc->bc;
  c->prec = desc->precision;
  c->esz = c->prec == SBP_FP64 ? 8 : 4;
  c->nx = desc->nx;
  c->ny = desc->ny;
  c->hix = (c->nx - 1) / (desc->xr - desc->xl);
  c->hiy = is1d(c->phys) ? 0.0 : (c->ny - 1) / (desc->yr - desc->yl);
  c->xoff = (kAlign - c->cls % kAlign) % kAlign;
  c->pitch = (c->xoff + c->nx + kAlign - 1) / kAlign * kAlign;
  c->plane = (long)c->pitch * c->ny;
  c->ty_rows = default_ty_rows(c->nx, c->ny, c->cls);
  compute_geometry(c);
  c->p_ty_rows = default_pair_rows(c);
  compute_pair_geometry(c);

  if (c->prec == SBP_FP64) {
    StageLaunchFn (*reg)(int, int, int) =
        c->order == 2 ? launcher_o2_f64 : (c->order == 4 ? launcher_o4_f64 : launcher_o6_f64);
    c->launch = reg(c->phys, c->mat, kBX);
  } else {
    StageLaunchFn (*reg)(int, int, bool, int) =
        c->order == 2 ? launcher_o2_f32 : (c->order == 4 ? launcher_o4_f32 : launcher_o6_f32);
    c->launch = reg(c->phys, c->mat, c->prec == SBP_FP32_KAHAN, kBX);
  }
  if (!c->launch) return fail(SBP_EINVAL, "no kernel for this physics/material/precision");
  if (c->phys == SBP_ACOUSTIC_2D || c->phys == SBP_ADVECTION_2D) {
    if (c->prec == SBP_FP64)
      c->pair_launch = (c->order == 2 ? pair_launcher_o2_f64
                                      : (c->order == 4 ? pair_launcher_o4_f64 : pair_launcher_o6_f64))(c->phys,
                                                                                                        c->mat);
    else
      c->pair_launch = (c->order == 2 ? pair_launcher_o2_f32
                                      : (c->order == 4 ? pair_launcher_o4_f32 : pair_launcher_o6_f32))(
          c->phys, c->mat, c->prec == SBP_FP32_KAHAN);
  }
  c->fused = pair_default(c);

  HIP_OR_FAIL(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  for (hipEvent_t &e : c->ev) HIP_OR_FAIL(hipEventCreate(&e));
  const size_t reg_bytes = c->esz * (size_t)c->C * c->plane;
  int rc;
  if ((rc = dalloc(&c->U, reg_bytes)) || (rc = dalloc(&c->A, reg_bytes)) || (rc = dalloc(&c->Y0, reg_bytes)) ||
      (rc = dalloc(&c->Y1, reg_bytes)))
    return rc;
  if (c->prec == SBP_FP32_KAHAN && (rc = dalloc(&c->KC, reg_bytes))) return rc;
  if (is1d(c->phys) && (rc = dalloc(&c->TK, reg_bytes))) return rc;
  // Zero the registers once so padding never holds stale bit patterns.
  for (void *p : {c->U, c->A, c->Y0, c->Y1, c->KC, c->TK})
    if (p) HIP_OR_FAIL(hipMemsetAsync(p, 0, reg_bytes, c->stream));
  if ((rc = dalloc(&c->aos, sizeof(double) * (size_t)c->C * c->nx * c->ny))) return rc;
  c->npartials = 1024;
  if ((rc = dalloc(&c->partials, sizeof(double) * c->npartials)) || (rc = dalloc(&c->scalar, 2 * sizeof(double))))
    return rc;
  if (c->phys == SBP_ACOUSTIC_2D && c->mat == SBP_MAT_ARRAYS) {
    if ((rc = upload_plane(c, desc->rho_inv, &c->rinv)) || (rc = upload_plane(c, desc->bulk, &c->bulk))) return rc;
  }
  if (vel_fields) {
    // a field not given is the constant ax / ay on every point
    const size_t npts = (size_t)c->nx * c->ny;
    std::vector<double> fx, fy;
    const double *vx = desc->vel_x, *vy = desc->vel_y;
    if (!vx) {
      fx.assign(npts, desc->ax);
      vx = fx.data();
    }
    if ((rc = upload_plane(c, vx, &c->velx))) return rc;
    if (c->phys == SBP_ADVECTION_2D) {
      if (!vy) {
        fy.assign(npts, desc->ay);
        vy = fy.data();
      }
      if ((rc = upload_plane(c, vy, &c->vely))) return rc;
    }
  }
  if (c->phys == SBP_ACOUSTIC_2D && c->mat == SBP_MAT_REF_RHO) {
    // x = xl + i/hi as rho_inv and forcing_* compute it (wave_eq_rhs.h:43,52).
    std::vector<double> xs(c->nx), ys(c->ny), c3(c->nx), s3(c->nx), c4(c->ny), s4(c->ny);
    for (int i = 0; i < c->nx; i++) {
      xs[i] = desc->xl + i / c->hix;
      c3[i] = std::cos(3 * kPi * xs[i]);
      s3[i] = std::sin(3 * kPi * xs[i]);
    }
    for (int j = 0; j < c->ny; j++) {
      ys[j] = desc->yl + j / c->hiy;
      c4[j] = std::cos(4 * kPi * ys[j]);
      s4[j] = std::sin(4 * kPi * ys[j]);
    }
    if ((rc = upload_table(c, xs, &c->xs)) || (rc = upload_table(c, ys, &c->ys)) ||
        (rc = upload_table(c, c3, &c->c3x)) || (rc = upload_table(c, s3, &c->s3x)) ||
        (rc = upload_table(c, c4, &c->c4y)) || (rc = upload_table(c, s4, &c->s4y)))
      return rc;
  }
  HIP_OR_FAIL(hipStreamSynchronize(c->stream));
  return SBP_OK;
}

}  // namespace

// Sum of squares and maximum of |v - w| over two FP64 vectors: per-block
// partials (sum, max) reduced with wavefront shuffles, then one block.
static __global__ __launch_bounds__(256) void diff_norms_kernel(const double *__restrict__ v,
                                                                 const double *__restrict__ w, long n,
                                                                 double *__restrict__ partial)
{
  __shared__ double ss[4], mm[4];
  double s2 = 0, mx = 0;
  for (long k = blockIdx.x * 256L + threadIdx.x; k < n; k += (long)gridDim.x * 256) {
    const double d = fabs(v[k] - w[k]);
    s2 += d * d;
    mx = fmax(mx, d);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    s2 += __shfl_down(s2, off, 64);
    mx = fmax(mx, __shfl_down(mx, off, 64));
  }
  if ((threadIdx.x & 63) == 0) {
    ss[threadIdx.x >> 6] = s2;
    mm[threadIdx.x >> 6] = mx;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = (ss[0] + ss[1]) + (ss[2] + ss[3]);
    partial[2 * blockIdx.x + 1] = fmax(fmax(mm[0], mm[1]), fmax(mm[2], mm[3]));
  }
}

static __global__ __launch_bounds__(256) void diff_norms_final_kernel(const double *__restrict__ partial, int nb,
                                                                       double *__restrict__ out)
{
  __shared__ double ss[4], mm[4];
  double s2 = 0, mx = 0;
  for (int k = threadIdx.x; k < nb; k += 256) {
    s2 += partial[2 * k];
    mx = fmax(mx, partial[2 * k + 1]);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    s2 += __shfl_down(s2, off, 64);
    mx = fmax(mx, __shfl_down(mx, off, 64));
  }
  if ((threadIdx.x & 63) == 0) {
    ss[threadIdx.x >> 6] = s2;
    mm[threadIdx.x >> 6] = mx;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    out[0] = (ss[0] + ss[1]) + (ss[2] + ss[3]);
    out[1] = fmax(fmax(mm[0], mm[1]), fmax(mm[2], mm[3]));
  }
}

// ---- no-allocation error norms (sbp_error_norms_noalloc) -----------------
// The per-block partials of diff_norms_kernel live in this file-scope device
// array instead of memory the library allocates: one (sum of squares, max)
// pair per block, at most kNoAllocBlocks blocks (8 KB), so any vector length
// reduces into it.  Each GPU holds its own copy until the process ends; every
// call on a GPU shares it (calls that overlap on one GPU race).  Every block
// of a launch writes its pair before the final kernel reads it, so nothing
// has to start at zero.
constexpr int kNoAllocBlocks = 512;
static __device__ __align__(256) double scratch[2 * kNoAllocBlocks];

static __global__ __launch_bounds__(256) void error_norms_out_kernel(const double *__restrict__ partial, int nb,
                                                                      double hprod, double *__restrict__ out)
{
  __shared__ double ss[4], mm[4];
  double s2 = 0, mx = 0;
  for (int k = threadIdx.x; k < nb; k += 256) {
    s2 += partial[2 * k];
    mx = fmax(mx, partial[2 * k + 1]);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    s2 += __shfl_down(s2, off, 64);
    mx = fmax(mx, __shfl_down(mx, off, 64));
  }
  if ((threadIdx.x & 63) == 0) {
    ss[threadIdx.x >> 6] = s2;
    mm[threadIdx.x >> 6] = mx;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    out[0] = sqrt(hprod) * sqrt((ss[0] + ss[1]) + (ss[2] + ss[3]));
    out[1] = fmax(fmax(mm[0], mm[1]), fmax(mm[2], mm[3]));
  }
}

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch[kMaxDevices] = {};

// The device address of `scratch` on the current GPU, looked up once per GPU
// with that GPU current (host code must never pass the symbol itself: on the
// host it names a stand-in, not device memory).
int scratch_on_current_device(double **out)
{
  int dev = 0;
  HIP_OR_FAIL(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(SBP_ENODEV, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(SBP_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch[dev] = static_cast<double *>(p);
  }
  *out = g_scratch[dev];
  return SBP_OK;
}
}  // namespace

hipError_t sbp::host::launch_step(sbp_ctx *c, double t, double dt) { return ::launch_step(c, t, dt); }
hipError_t sbp::host::launch_rhs(sbp_ctx *c, double t, const void *src, void *dst)
{
  return ::launch_stage(c, MODE_RHS, 0.0, t, src, dst);
}
void sbp::host::drop_step_graph(sbp_ctx *c) { ::drop_graph(c); }

int sbp::host::energy_of(sbp_ctx *c, const void *u, const void *comp, double *out)
{
  return norm_reduce_planar(c, false, true, u, comp, out);
}

extern "C" {

int sbp_create(const sbp_problem *desc, sbp_ctx **out)
{
  if (!out) return fail(SBP_EINVAL, "null output pointer");
  *out = nullptr;
  int rc = validate(desc);
  if (rc) return rc;
  sbp_ctx *c = new (std::nothrow) sbp_ctx();
  if (!c) return fail(SBP_ENOMEM, "host allocation failed");
  rc = build_ctx(desc, c);
  if (rc) {
    std::string msg = g_err;
    release(c);
    g_err = msg;
    return rc;
  }
  *out = c;
  return SBP_OK;
}

int sbp_destroy(sbp_ctx *ctx)
{
  release(ctx);
  return SBP_OK;
}

int sbp_components(const sbp_ctx *ctx) { return ctx ? ctx->C : 0; }

long sbp_vector_length(const sbp_ctx *ctx) { return ctx ? (long)ctx->C * ctx->nx * ctx->ny : 0; }

int sbp_set_tile_rows(sbp_ctx *ctx, int rows)
{
  if (!ctx || rows < 1 || rows > 4096) return fail(SBP_EINVAL, "tile rows must be 1..4096");
  HIP_OR_FAIL(hipStreamSynchronize(ctx->stream));
  drop_graph(ctx);
  ctx->ty_rows = rows;
  compute_geometry(ctx);
  return SBP_OK;
}

int sbp_set_fused(sbp_ctx *ctx, int on, int rows)
{
  if (!ctx || rows < 0 || rows > 4096) return fail(SBP_EINVAL, "bad argument");
  if (on && !pair_supported(ctx))
    return fail(SBP_EINVAL, "two-stage stepping needs a 2D acoustic or 2D advection context with more than "
                            "2(CLS+R) points per direction");
  HIP_OR_FAIL(hipStreamSynchronize(ctx->stream));
  drop_graph(ctx);
  ctx->fused = on != 0;
  ctx->p_ty_rows = rows > 0 ? rows : default_pair_rows(ctx);
  compute_pair_geometry(ctx);
  return SBP_OK;
}

int sbp_get_fused(const sbp_ctx *ctx) { return ctx && ctx->fused ? 1 : 0; }

int sbp_tune_fused(sbp_ctx *ctx, double dt, int *rows)
{
  if (!ctx || !(dt > 0)) return fail(SBP_EINVAL, "bad argument");
  if (!ctx->have_state) return fail(SBP_ESTATE, "no state loaded");
  if (!ctx->fused) {
    if (rows) *rows = 0;
    return SBP_OK;
  }
  const int R = ctx->order / 2, D = ctx->cls + R;
  const int hi = ctx->ny - 2 * D;
  static const int cand[] = {8, 10, 12, 16, 20, 24, 32, 40, 48, 68, 96, 136};
  // the trial steps advance the state: save it (and its compensation) first
  const size_t reg = ctx->esz * (size_t)ctx->C * ctx->plane;
  void *save_u = nullptr, *save_c = nullptr;
  int rc = dalloc(&save_u, reg);
  if (!rc && ctx->KC) rc = dalloc(&save_c, reg);
  auto cleanup = [&]() {
    if (save_u) (void)hipFree(save_u);
    if (save_c) (void)hipFree(save_c);
  };
  if (rc) {
    cleanup();
    return rc;
  }
  hipError_t e = hipMemcpyAsync(save_u, ctx->U, reg, hipMemcpyDeviceToDevice, ctx->stream);
  if (e == hipSuccess && ctx->KC) e = hipMemcpyAsync(save_c, ctx->KC, reg, hipMemcpyDeviceToDevice, ctx->stream);
  const int start_rows = ctx->p_ty_rows;
  int best = start_rows;
  float best_ms = 1e30f;
  for (int t : cand) {
    if (e != hipSuccess) break;
    if (t > hi && t != cand[0]) continue;
    drop_graph(ctx);
    ctx->p_ty_rows = t;
    compute_pair_geometry(ctx);
    e = launch_step(ctx, 0.0, dt);  // warm-up
    float total = 0;
    for (int rep = 0; rep < 3 && e == hipSuccess; rep++) {
      e = hipEventRecord(ctx->ev[0], ctx->stream);
      if (e == hipSuccess) e = launch_step(ctx, 0.0, dt);
      if (e == hipSuccess) e = hipEventRecord(ctx->ev[1], ctx->stream);
      if (e == hipSuccess) e = hipEventSynchronize(ctx->ev[1]);
      float ms = 0;
      if (e == hipSuccess) e = hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]);
      total += ms;
    }
    if (e == hipSuccess && total < best_ms) {
      best_ms = total;
      best = t;
    }
  }
  if (e == hipSuccess) e = hipMemcpyAsync(ctx->U, save_u, reg, hipMemcpyDeviceToDevice, ctx->stream);
  if (e == hipSuccess && ctx->KC) e = hipMemcpyAsync(ctx->KC, save_c, reg, hipMemcpyDeviceToDevice, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  cleanup();
  drop_graph(ctx);
  ctx->p_ty_rows = e == hipSuccess ? best : start_rows;
  compute_pair_geometry(ctx);
  if (e != hipSuccess) return fail(SBP_EHIP, std::string("tile tuning: ") + hipGetErrorString(e));
  if (rows) *rows = best;
  return SBP_OK;
}

int sbp_set_store_policy(sbp_ctx *ctx, int policy)
{
  if (!ctx || policy < 0 || policy > 2) return fail(SBP_EINVAL, "store policy must be 0, 1 or 2");
  HIP_OR_FAIL(hipStreamSynchronize(ctx->stream));
  drop_graph(ctx);
  ctx->store_policy = policy;
  return SBP_OK;
}

int sbp_rhs(sbp_ctx *ctx, double t, const double *src, double *dst)
{
  if (!ctx || !src || !dst) return fail(SBP_EINVAL, "null argument");
  int rc;
  if ((rc = upload_aos(ctx, src)) || (rc = aos_to_planar(ctx, ctx->Y1, nullptr))) return rc;
  HIP_OR_FAIL(launch_stage(ctx, MODE_RHS, 0.0, t, ctx->Y1, ctx->Y0));
  if ((rc = planar_to_aos(ctx, ctx->Y0, nullptr))) return rc;
  return download_aos(ctx, dst);
}

int sbp_set_state(sbp_ctx *ctx, const double *v)
{
  if (!ctx || !v) return fail(SBP_EINVAL, "null argument");
  int rc;
  if ((rc = upload_aos(ctx, v)) || (rc = aos_to_planar(ctx, ctx->U, ctx->KC))) return rc;
  HIP_OR_FAIL(hipStreamSynchronize(ctx->stream));
  ctx->have_state = true;
  return SBP_OK;
}

int sbp_get_state(sbp_ctx *ctx, double *v)
{
  if (!ctx || !v) return fail(SBP_EINVAL, "null argument");
  if (!ctx->have_state) return fail(SBP_ESTATE, "no state loaded");
  int rc = planar_to_aos(ctx, ctx->U, ctx->KC);
  if (rc) return rc;
  return download_aos(ctx, v);
}

int sbp_get_state_strided(sbp_ctx *ctx, int stride, double *v)
{
  if (!ctx || !v || stride < 1) return fail(SBP_EINVAL, "bad argument");
  if (!ctx->have_state) return fail(SBP_ESTATE, "no state loaded");
  const int cx = (ctx->nx - 1) / stride + 1, cy = (ctx->ny - 1) / stride + 1;
  const long n = (long)cx * cy;
  const int grid = (int)((n + 255) / 256);
  if (ctx->esz == 8)
    hipLaunchKernelGGL(planar_to_aos_strided_kernel<double>, dim3(grid), dim3(256), 0, ctx->stream,
                       (const double *)ctx->U, (const double *)ctx->KC, ctx->aos, ctx->nx, ctx->ny, ctx->C, ctx->pitch,
                       ctx->xoff, ctx->plane, stride);
  else
    hipLaunchKernelGGL(planar_to_aos_strided_kernel<float>, dim3(grid), dim3(256), 0, ctx->stream,
                       (const float *)ctx->U, (const float *)ctx->KC, ctx->aos, ctx->nx, ctx->ny, ctx->C, ctx->pitch,
                       ctx->xoff, ctx->plane, stride);
  HIP_OR_FAIL(hipGetLastError());
  HIP_OR_FAIL(hipMemcpyAsync(v, ctx->aos, sizeof(double) * ctx->C * n, hipMemcpyDeviceToHost, ctx->stream));
  HIP_OR_FAIL(hipStreamSynchronize(ctx->stream));
  return SBP_OK;
}

int sbp_step(sbp_ctx *ctx, double t0, double dt, long nsteps)
{
  if (!ctx || nsteps < 0) return fail(SBP_EINVAL, "bad argument");
  if (!ctx->have_state) return fail(SBP_ESTATE, "no state loaded");
  return run_steps(ctx, t0, dt, nsteps);
}

int sbp_matchstep_schedule(double t0, double t_end, double dt, int max_runs, double *run_dt, long *run_len,
                           int *nruns, double *t_final)
{
  if (!nruns || max_runs < 0 || !(dt > 0) || !std::isfinite(t0) || !std::isfinite(t_end))
    return fail(SBP_EINVAL, "bad time span or step");
  const int n = ts::matchstep_runs(t0, t_end, dt, max_runs, run_dt, run_len, t_final);
  if (n < 0) return fail(SBP_EINVAL, "the schedule needs more than max_runs runs");
  *nruns = n;
  return SBP_OK;
}

int sbp_ts_rk4(sbp_ctx *ctx, double t0, double t_end, double dt, long *nsteps, double *t_final)
{
  if (!ctx) return fail(SBP_EINVAL, "null context");
  if (!ctx->have_state) return fail(SBP_ESTATE, "no state loaded");
  if (!(dt > 0) || !std::isfinite(t0) || !std::isfinite(t_end)) return fail(SBP_EINVAL, "bad time span or step");
  // A fixed-step MATCHSTEP run changes its step at most twice, at the end.
  constexpr int kMaxRuns = 8;
  double rdt[kMaxRuns];
  long rlen[kMaxRuns];
  const int nr = ts::matchstep_runs(t0, t_end, dt, kMaxRuns, rdt, rlen, nullptr);
  if (nr < 0) return fail(SBP_EINVAL, "unexpected step schedule");
  double t = t0;
  long total = 0;
  for (int k = 0; k < nr; k++) {
    int rc = run_steps(ctx, t, rdt[k], rlen[k], &t);
    if (rc) return rc;
    total += rlen[k];
  }
  if (nsteps) *nsteps = total;
  if (t_final) *t_final = t;
  return SBP_OK;
}

int sbp_rk4_custom(sbp_ctx *ctx, double t_end, double dt, long *tlen, double *dt_used, double *t_final)
{
  if (!ctx) return fail(SBP_EINVAL, "null context");
  if (!ctx->have_state) return fail(SBP_ESTATE, "no state loaded");
  if (!(dt > 0) || !(t_end > 0) || !std::isfinite(t_end)) return fail(SBP_EINVAL, "bad time span or step");
  double h = dt;
  const long n = ts::rk4_custom_steps(t_end, dt, &h, nullptr);
  if (n < 1) return fail(SBP_EINVAL, "round(Tend/dt) is zero");
  double t = 0.0;
  int rc = run_steps(ctx, 0.0, h, n, &t);
  if (rc) return rc;
  if (tlen) *tlen = n;
  if (dt_used) *dt_used = h;
  if (t_final) *t_final = t;
  return SBP_OK;
}

int sbp_adapt_basic(double enorm, double h, int order, int prev_rejected, int *accept, double *next_h)
{
  if (!accept || !next_h || order < 1 || !(h > 0)) return fail(SBP_EINVAL, "bad argument");
  if (!std::isfinite(enorm)) return fail(SBP_EDIVERGED, "non-finite error estimate");
  const ts::Choice ch = ts::adapt_basic(enorm, h, order, prev_rejected != 0);
  *accept = ch.accept ? 1 : 0;
  *next_h = ch.next_h;
  return SBP_OK;
}

int sbp_synchronize(sbp_ctx *ctx)
{
  if (!ctx) return fail(SBP_EINVAL, "null context");
  HIP_OR_FAIL(hipStreamSynchronize(ctx->stream));
  return SBP_OK;
}

int sbp_rk4_advance(sbp_ctx *ctx, double *v, double t0, double dt, long nsteps)
{
  int rc;
  if ((rc = sbp_set_state(ctx, v)) || (rc = sbp_step(ctx, t0, dt, nsteps))) return rc;
  return sbp_get_state(ctx, v);
}

int sbp_time_steps(sbp_ctx *ctx, double t0, double dt, long nsteps, double *ms)
{
  if (!ctx || !ms) return fail(SBP_EINVAL, "null argument");
  if (!ctx->have_state) return fail(SBP_ESTATE, "no state loaded");
  if (!time_dependent(ctx) && nsteps >= kGraphSteps) {
    hipGraphExec_t ge = nullptr;
    int rc = ensure_graph(ctx, dt, &ge);  // capture outside the timed region
    if (rc) return rc;
  }
  HIP_OR_FAIL(hipEventRecord(ctx->ev[0], ctx->stream));
  int rc = run_steps(ctx, t0, dt, nsteps);
  if (rc) return rc;
  HIP_OR_FAIL(hipEventRecord(ctx->ev[1], ctx->stream));
  HIP_OR_FAIL(hipEventSynchronize(ctx->ev[1]));
  float f = 0;
  HIP_OR_FAIL(hipEventElapsedTime(&f, ctx->ev[0], ctx->ev[1]));
  *ms = f;
  return SBP_OK;
}

int sbp_time_stages(sbp_ctx *ctx, double dt, long nsteps, double ms_stage[4])
{
  if (!ctx || !ms_stage) return fail(SBP_EINVAL, "null argument");
  if (!ctx->have_state) return fail(SBP_ESTATE, "no state loaded");
  double acc[4] = {0, 0, 0, 0};
  const double ts[4] = {0, 0.5 * dt, 0.5 * dt, dt};
  // one launch per stage: four timed launches; two-stage: the launches of
  // stages 1-2 and 3-4 are timed into slots 0 and 2, slots 1 and 3 stay 0
  const int nl = ctx->fused ? 2 : 4;
  for (long n = 0; n < nsteps; n++) {
    HIP_OR_FAIL(hipEventRecord(ctx->ev[0], ctx->stream));
    for (int s = 0; s < nl; s++) {
      if (ctx->fused) {
        HIP_OR_FAIL(launch_pair_part(ctx, 2 * s + 1, n * dt, dt));
      } else {
        const void *yin;
        void *yout;
        stage_io(ctx, s + 1, &yin, &yout);
        HIP_OR_FAIL(launch_stage(ctx, MODE_S1 + s, dt, n * dt + ts[s], yin, yout));
      }
      HIP_OR_FAIL(hipEventRecord(ctx->ev[s + 1], ctx->stream));
    }
    HIP_OR_FAIL(hipEventSynchronize(ctx->ev[nl]));
    for (int s = 0; s < nl; s++) {
      float f = 0;
      HIP_OR_FAIL(hipEventElapsedTime(&f, ctx->ev[s], ctx->ev[s + 1]));
      acc[ctx->fused ? 2 * s : s] += f;
    }
  }
  for (int s = 0; s < 4; s++) ms_stage[s] = nsteps > 0 ? acc[s] / nsteps : 0.0;
  return SBP_OK;
}

int sbp_energy(sbp_ctx *ctx, const double *v, double *E)
{
  if (!ctx || !v || !E) return fail(SBP_EINVAL, "null argument");
  int rc = upload_aos(ctx, v);
  if (rc) return rc;
  return norm_reduce(ctx, true, true, E);
}

int sbp_state_energy(sbp_ctx *ctx, double *E)
{
  if (!ctx || !E) return fail(SBP_EINVAL, "null argument");
  if (!ctx->have_state) return fail(SBP_ESTATE, "no state loaded");
  return norm_reduce(ctx, false, true, E);
}

int sbp_hnorm(sbp_ctx *ctx, const double *v, double *norm)
{
  if (!ctx || !v || !norm) return fail(SBP_EINVAL, "null argument");
  int rc = upload_aos(ctx, v);
  if (rc) return rc;
  double s = 0;
  if ((rc = norm_reduce(ctx, true, false, &s))) return rc;
  *norm = std::sqrt(s);
  return SBP_OK;
}

int sbp_error_norms(sbp_ctx *ctx, const double *v, const double *w, double *l2, double *maxerr)
{
  if (!ctx || !v || !w) return fail(SBP_EINVAL, "null argument");
  const long n = (long)ctx->C * ctx->nx * ctx->ny;
  if (!ctx->aos2) {
    int rc = dalloc(&ctx->aos2, sizeof(double) * (size_t)n);
    if (rc) return rc;
  }
  HIP_OR_FAIL(hipMemcpyAsync(ctx->aos, v, sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream));
  HIP_OR_FAIL(hipMemcpyAsync(ctx->aos2, w, sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream));
  const int grid = (int)std::max<long>(1, std::min<long>((n + 255) / 256, ctx->npartials / 2));
  hipLaunchKernelGGL(diff_norms_kernel, dim3(grid), dim3(256), 0, ctx->stream, ctx->aos, ctx->aos2, n, ctx->partials);
  hipLaunchKernelGGL(diff_norms_final_kernel, dim3(1), dim3(256), 0, ctx->stream, ctx->partials, grid, ctx->scalar);
  HIP_OR_FAIL(hipGetLastError());
  double r[2] = {0, 0};
  HIP_OR_FAIL(hipMemcpyAsync(r, ctx->scalar, sizeof(r), hipMemcpyDeviceToHost, ctx->stream));
  HIP_OR_FAIL(hipStreamSynchronize(ctx->stream));
  double hprod = 1.0 / ctx->hix;
  if (!is1d(ctx->phys)) hprod *= 1.0 / ctx->hiy;
  if (l2) *l2 = std::sqrt(hprod) * std::sqrt(r[0]);
  if (maxerr) *maxerr = r[1];
  return SBP_OK;
}

int sbp_error_norms_noalloc(const double *v, const double *w, long n, double hx, double hy, double *out,
                            void *stream)
{
  if (!v || !w || !out) return fail(SBP_EINVAL, "null argument");
  if (n < 1 || n > SBP_NOALLOC_MAX_VALUES)
    return fail(SBP_EINVAL, "vector length must be 1 .. SBP_NOALLOC_MAX_VALUES (2^40)");
  if (!(hx > 0) || !(hy > 0) || !std::isfinite(hx) || !std::isfinite(hy))
    return fail(SBP_EINVAL, "grid spacings must be positive");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(SBP_ENODEV, "no HIP device");
  double *part = nullptr;
  int rc = scratch_on_current_device(&part);
  if (rc) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  // the block count of sbp_error_norms, so both reduce in the same order
  const int grid = (int)std::max<long>(1, std::min<long>((n + 255) / 256, kNoAllocBlocks));
  hipLaunchKernelGGL(diff_norms_kernel, dim3(grid), dim3(256), 0, s, v, w, n, part);
  hipLaunchKernelGGL(error_norms_out_kernel, dim3(1), dim3(256), 0, s, part, grid, hx * hy, out);
  HIP_OR_FAIL(hipGetLastError());
  return SBP_OK;
}

const char *sbp_last_error(void) { return g_err.c_str(); }

const char *sbp_device_arch(void)
{
  static thread_local char name[64];
  name[0] = 0;
  int ndev = 0, dev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return name;
  if (hipGetDevice(&dev) != hipSuccess) return name;
  hipDeviceProp_t p;
  if (hipGetDeviceProperties(&p, dev) != hipSuccess) return name;
  std::snprintf(name, sizeof(name), "%s", p.gcnArchName);
  return name;
}

// Plain copy kernel (vectorised, one 16-byte element per thread) used to
// calibrate HBM counters and to report the achievable streaming bandwidth.
static __global__ __launch_bounds__(256) void copy_probe_kernel(const double2 *__restrict__ src,
                                                                double2 *__restrict__ dst, long n)
{
  const long i = blockIdx.x * 256L + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

int sbp_copy_bandwidth(long bytes, int reps, double *gbs)
{
  if (bytes < 1024 || reps < 1 || !gbs) return fail(SBP_EINVAL, "bad argument");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(SBP_ENODEV, "no HIP device");
  const long n = bytes / (long)sizeof(double2);
  double2 *a = nullptr, *b = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = SBP_OK;
  float best = 1e30f;
  if (hipMalloc(&a, n * sizeof(double2)) != hipSuccess || hipMalloc(&b, n * sizeof(double2)) != hipSuccess) {
    rc = fail(SBP_ENOMEM, "copy probe allocation failed");
  } else if (hipMemset(a, 0, n * sizeof(double2)) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
             hipEventCreate(&e1) != hipSuccess) {
    rc = fail(SBP_EHIP, "copy probe setup failed");
  } else {
    const int grid = (int)((n + 255) / 256);
    for (int r = 0; r <= reps && rc == SBP_OK; r++) {
      (void)hipEventRecord(e0, nullptr);
      hipLaunchKernelGGL(copy_probe_kernel, dim3(grid), dim3(256), 0, nullptr, a, b, n);
      (void)hipEventRecord(e1, nullptr);
      if (hipEventSynchronize(e1) != hipSuccess) rc = fail(SBP_EHIP, "copy probe failed");
      float ms = 0;
      (void)hipEventElapsedTime(&ms, e0, e1);
      if (r > 0 && ms < best) best = ms;  // first launch is a warm-up
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  if (rc) return rc;
  *gbs = 2.0 * n * sizeof(double2) / (best * 1e-3) / 1e9;
  return SBP_OK;
}

}  // extern "C"

// Flat streaming kernel: every point reads NR planes and writes NW planes
// (16 bytes per plane and thread per pass, grid-stride, nontemporal stores),
// the byte mix of a launch without its stencils, halos or tiles.
template <int NR, int NW>
static __global__ __launch_bounds__(256) void stream_probe_kernel(const double2 *__restrict__ src,
                                                                  double2 *__restrict__ dst, long n2)
{
  for (long i = blockIdx.x * 256L + threadIdx.x; i < n2; i += (long)gridDim.x * 256) {
    double sx = 0, sy = 0;
#pragma unroll
    for (int c = 0; c < NR; c++) {
      const double2 v = src[c * n2 + i];
      sx += v.x;
      sy += v.y;
    }
#pragma unroll
    for (int c = 0; c < NW; c++) {
      double *d = reinterpret_cast<double *>(dst + c * n2 + i);
      __builtin_nontemporal_store(sx + c, d);
      __builtin_nontemporal_store(sy + c, d + 1);
    }
  }
}

typedef void (*StreamProbeFn)(const double2 *, double2 *, long);
template <int NR, int NW> static void *stream_probe_ptr() { return (void *)stream_probe_kernel<NR, NW>; }
template <int NR> static void *stream_probe_nw(int nw)
{
  switch (nw) {
  case 1: return stream_probe_ptr<NR, 1>();
  case 2: return stream_probe_ptr<NR, 2>();
  case 3: return stream_probe_ptr<NR, 3>();
  case 4: return stream_probe_ptr<NR, 4>();
  case 5: return stream_probe_ptr<NR, 5>();
  case 6: return stream_probe_ptr<NR, 6>();
  }
  return nullptr;
}
static void *stream_probe_fn(int nr, int nw)
{
  switch (nr) {
  case 1: return stream_probe_nw<1>(nw);
  case 2: return stream_probe_nw<2>(nw);
  case 3: return stream_probe_nw<3>(nw);
  case 4: return stream_probe_nw<4>(nw);
  case 5: return stream_probe_nw<5>(nw);
  case 6: return stream_probe_nw<6>(nw);
  case 7: return stream_probe_nw<7>(nw);
  case 8: return stream_probe_nw<8>(nw);
  case 9: return stream_probe_nw<9>(nw);
  case 10: return stream_probe_nw<10>(nw);
  case 11: return stream_probe_nw<11>(nw);
  case 12: return stream_probe_nw<12>(nw);
  }
  return nullptr;
}

extern "C" {

int sbp_stream_bandwidth(long plane_bytes, int n_read, int n_write, int reps, double *gbs)
{
  if (plane_bytes < 1024 || n_read < 1 || n_read > 12 || n_write < 1 || n_write > 6 || reps < 1 || !gbs)
    return fail(SBP_EINVAL, "bad argument");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(SBP_ENODEV, "no HIP device");
  int dev = 0, ncu = 0;
  if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev))
    return fail(SBP_EHIP, "device query failed");
  const long n2 = plane_bytes / (long)sizeof(double2);
  void *fn = stream_probe_fn(n_read, n_write);
  double2 *a = nullptr, *b = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = SBP_OK;
  float best = 1e30f;
  if (hipMalloc(&a, n_read * n2 * sizeof(double2)) != hipSuccess ||
      hipMalloc(&b, n_write * n2 * sizeof(double2)) != hipSuccess) {
    rc = fail(SBP_ENOMEM, "stream probe allocation failed");
  } else if (hipMemset(a, 0, n_read * n2 * sizeof(double2)) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
             hipEventCreate(&e1) != hipSuccess) {
    rc = fail(SBP_EHIP, "stream probe setup failed");
  } else {
    // 4, 8 and 16 blocks per CU; the best rate counts
    for (int bpc = 4; bpc <= 16 && rc == SBP_OK; bpc *= 2) {
      const int grid = (int)std::min<long>((long)ncu * bpc, (n2 + 255) / 256);
      for (int r = 0; r <= reps && rc == SBP_OK; r++) {
        (void)hipEventRecord(e0, nullptr);
        hipLaunchKernelGGL((StreamProbeFn)fn, dim3(grid), dim3(256), 0, nullptr, (const double2 *)a, b, n2);
        (void)hipEventRecord(e1, nullptr);
        if (hipEventSynchronize(e1) != hipSuccess) rc = fail(SBP_EHIP, "stream probe failed");
        float ms = 0;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (r > 0 && ms < best) best = ms;  // first launch is a warm-up
      }
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  if (rc) return rc;
  *gbs = (double)(n_read + n_write) * n2 * sizeof(double2) / (best * 1e-3) / 1e9;
  return SBP_OK;
}

int sbp_version(void) { return 100; }

}  // extern "C"
