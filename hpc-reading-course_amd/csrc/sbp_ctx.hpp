// sbp_ctx.hpp -- the single-problem context shared by the host translation
// units (C-ABI, learned corrector).
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/sbp_hip.h"
#include "sbp_kernels.hpp"

struct sbp_ctx {
  sbp_problem d{};
  int C = 0, order = 0, cls = 0, phys = 0, mat = 0, bc = 0, prec = 0;
  int nx = 0, ny = 0, pitch = 0, xoff = 0;
  long plane = 0;
  size_t esz = 8;
  double hix = 0, hiy = 0;
  void *U = nullptr, *A = nullptr, *Y0 = nullptr, *Y1 = nullptr, *KC = nullptr, *TK = nullptr;
  void *rinv = nullptr, *bulk = nullptr;
  void *velx = nullptr, *vely = nullptr;  // advection velocity fields (per-point a, b)
  double *xs = nullptr, *ys = nullptr, *c3x = nullptr, *s3x = nullptr, *c4y = nullptr, *s4y = nullptr;
  double *aos = nullptr, *aos2 = nullptr;  // AoS FP64 staging (aos2: second operand, on first use)
  double *partials = nullptr, *scalar = nullptr;
  int npartials = 0;
  hipStream_t stream = nullptr;
  sbp::StageLaunchFn launch = nullptr;
  sbp::LaunchGeom geom{};
  int ty_rows = 32, tiles_x = 0, tiles_y = 0, n_tiles = 0, n_edge = 0, n_edge_blocks = 0;
  bool have_state = false;
  // Two-stage interior path (2D acoustics, unit or array materials): pair
  // launcher, D = CLS + R deep frame, tile geometry of the pair tiles.
  sbp::PairLaunchFn pair_launch = nullptr;
  bool fused = false;
  int store_policy = 0;  // two-stage tile stores: 0 by streamed size, 1 cached, 2 nontemporal
  int p_depth = 0, p_ty_rows = 0, p_tiles_x = 0, p_n_tiles = 0, p_n_edge = 0, p_n_edge_blocks = 0;
  // Captured multi-step graphs, keyed by dt and by the array holding U when
  // the capture started (the two-stage path swaps U and Y2 every step).
  struct GraphSlot {
    hipGraphExec_t exec = nullptr;
    double dt = 0;
    void *u = nullptr;
  };
  GraphSlot gslot[2];
  int g_next = 0;
  hipEvent_t ev[5] = {};
  int device = 0;
  void *rkw[8] = {};  // adaptive RK workspace (K1..K6, stage input, candidate), allocated on first use
};


namespace sbp {
namespace host {
// Launch one RK4 stage / one whole step of a context (sbp_capi.hip).
hipError_t launch_step(sbp_ctx *c, double t, double dt);
// H-weighted energy of a planar state laid out like the context's registers
// (U and optional compensation); result left in c->scalar (device) and, when
// out != nullptr, copied to the host.
int energy_of(sbp_ctx *c, const void *u, const void *comp, double *out);
// dst = F(t, src) on planar registers of the context (the RHS stage kernels).
hipError_t launch_rhs(sbp_ctx *c, double t, const void *src, void *dst);
// Forget the captured step graph (after the state register moved).
void drop_step_graph(sbp_ctx *c);
}  // namespace host
}  // namespace sbp
