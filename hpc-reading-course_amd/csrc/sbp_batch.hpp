// sbp_batch.hpp -- persistent kernels of the batched acoustic ensemble.
//
// Many independent 2D acoustic problems of different sizes share one launch
// per RK4 stage (batch_stage_kernel) or, in the two-stage configurations, one
// per pair of stages (batch_pair_kernel).  The host flattens the work of every
// instance into one list of (instance, item) pairs -- the boundary-frame
// blocks (frame patches) and the interior tiles (pair tiles) of
// sbp_kernels.hpp -- and a fixed grid of persistent blocks walks that list.  The list is cut into one chunk per XCD: whole instances,
// bin-packed by work, so every XCD gets the same share of the heavy tiles and
// an instance's halo rows stay in one L2.  The per-instance descriptors live
// in device memory, so a captured graph stays valid when the time steps
// change.
#pragma once
#include "sbp_kernels.hpp"

namespace sbp {

template <typename T> struct BatchInst {
  T *U, *A, *Y0, *Y1, *comp;
  const T *rinv, *bulk;
  long plane;
  int pitch, xoff, nx, ny;
  T hix, hiy;
  double dt;
  int bc, tiles_x, ty_rows, n_tiles, n_edge_blocks, n_edge;
  // two-stage configurations: the (CLS+R)-deep frame and the pair tiles
  int p_depth, p_tiles_x, p_ty_rows, p_n_tiles, p_n_edge, p_n_edge_blocks;
  int p_nts;  // pair tiles' store policy (PairArgs::nts), one for the whole ensemble
};

template <typename T, int MODE>
__device__ __forceinline__ StageArgs<T> batch_args(const BatchInst<T> &d)
{
  StageArgs<T> a{};
  a.uin = d.U;
  a.y2 = d.Y0;
  a.y3 = d.Y1;
  a.third = T(1.0 / 3.0);
  a.aio = d.A;
  a.uout = d.U;
  a.comp = d.comp;
  a.rinv = d.rinv;
  a.bulk = d.bulk;
  a.plane = d.plane;
  a.pitch = d.pitch;
  a.xoff = d.xoff;
  a.nx = d.nx;
  a.ny = d.ny;
  a.hix = d.hix;
  a.hiy = d.hiy;
  a.bc = d.bc;
  a.tiles_x = d.tiles_x;
  a.ty_rows = d.ty_rows;
  a.n_tiles = d.n_tiles;
  a.n_edge_blocks = d.n_edge_blocks;
  a.n_edge = d.n_edge;
  const double dt = d.dt;
  if constexpr (MODE == MODE_S1) {
    a.yin = d.U;
    a.yout = d.Y0;
    a.cy = T(0.5 * dt);
    a.ca = T(dt / 6);
  } else if constexpr (MODE == MODE_S2) {
    a.yin = d.Y0;
    a.yout = d.Y1;
    a.cy = T(0.5 * dt);
    a.ca = T(dt / 3);
  } else if constexpr (MODE == MODE_S3) {
    a.yin = d.Y1;
    a.yout = d.A;  // Y4
    a.cy = T(dt);
    a.ca = T(dt / 3);
  } else {
    a.yin = d.A;
    a.yout = nullptr;
    a.ca = T(dt / 6);
  }
  return a;
}

// PF: rows of loads an interior tile keeps in flight ahead of the row it
// computes (TilePolicy); deep prefetch pays off for the short tiles of
// ensembles of small grids, where a tile is a few latency-bound rows.
template <typename T, int ORDER, int MAT, int MODE, bool KAHAN, int BX, int PF>
__global__ __launch_bounds__(BX) void batch_stage_kernel(const BatchInst<T> *__restrict__ inst,
                                                         const int2 *__restrict__ work, int nwork)
{
  // The list opens with one {offset, count} per XCD; gridDim.x is a
  // multiple of kNumXcd and block b runs on XCD b % 8, whose blocks stride
  // through that XCD's chunk.
  const int2 chunk = work[blockIdx.x % kNumXcd];
  const int2 *items = work + kNumXcd + chunk.x;
  for (int w = blockIdx.x / kNumXcd; w < chunk.y; w += gridDim.x / kNumXcd) {
    const int2 it = items[w];
    const StageArgs<T> a = batch_args<T, MODE>(inst[it.x]);
    if (it.y < a.n_edge_blocks)
      edge_point<T, ORDER, PH_ACOUSTIC, MAT, MODE, KAHAN>(a, it.y * BX + threadIdx.x);
    else
      interior_tile<T, ORDER, PH_ACOUSTIC, MAT, MODE, KAHAN, BX, TilePolicy<PF, true, true, false>>(
          a, it.y - a.n_edge_blocks);
    __syncthreads();  // the next item reuses the LDS row buffers
  }
}

// flag or-ed into a two-stage launch's part number: cached pair-tile stores
constexpr int kBatchCachedStores = 16;

using BatchLaunchFn = hipError_t (*)(int mode, const void *inst, const int2 *work, int nwork, int grid,
                                     hipStream_t s);

// --------------------------------------------------------------------------
// Two-stage configurations (the schedule of sbp_kernels.hpp pair2d_kernel):
// two persistent launches per step over a work list of frame patches and
// pair tiles,
//   part 1: stages 1-2 (frame_patch_pair and pair-1 tiles)
//   part 3: stages 3-4 (frame_patch_pair and pair-2 tiles)
// A step leaves U' in each instance's Y0 array; the host then passes the
// second descriptor set, in which U and Y0 trade places.
// --------------------------------------------------------------------------
template <typename T, int MODE>
__device__ __forceinline__ StageArgs<T> batch_frame_args(const BatchInst<T> &d)
{
  StageArgs<T> a = batch_args<T, MODE>(d);
  a.depth = d.p_depth;
  a.n_edge = d.p_n_edge;
  a.n_edge_blocks = d.p_n_edge_blocks;
  a.n_tiles = 0;
  if constexpr (MODE == MODE_S4) a.uout = d.Y0;  // U' beside the interior's (U is read through uin)
  return a;
}

template <typename T, int PAIR>
__device__ __forceinline__ PairArgs<T> batch_pair_args(const BatchInst<T> &d)
{
  PairArgs<T> a{};
  a.rinv = d.rinv;
  a.bulk = d.bulk;
  a.plane = d.plane;
  a.pitch = d.pitch;
  a.xoff = d.xoff;
  a.nx = d.nx;
  a.ny = d.ny;
  a.hix = d.hix;
  a.hiy = d.hiy;
  a.third = T(1.0 / 3.0);
  a.tiles_x = d.p_tiles_x;
  a.ty_rows = d.p_ty_rows;
  a.n_tiles = d.p_n_tiles;
  a.n_edge_blocks = d.p_n_edge_blocks;
  a.nts = d.p_nts;
  const double dt = d.dt;
  if constexpr (PAIR == 1) {
    a.yin = d.U;
    a.o1 = d.Y0;
    a.o2 = d.Y1;
    a.c1 = T(0.5 * dt);
    a.c2 = T(0.5 * dt);
  } else {
    a.yin = d.Y1;
    a.u = d.U;
    a.y2 = d.Y0;
    a.o1 = d.Y0;
    a.o2 = d.A;
    a.comp = d.comp;
    a.c1 = T(dt);
    a.c2 = T(dt / 6);
    a.c0 = T(0.5 * dt);
  }
  return a;
}

template <typename T, int ORDER, int MAT, int PART, bool KAHAN, int NT, int PF, bool NTS>
__global__ __launch_bounds__(NT) void batch_pair_kernel(const BatchInst<T> *__restrict__ inst,
                                                        const int2 *__restrict__ work, int nwork)
{
  constexpr int PAIR = PART == 1 ? 1 : 2;
  constexpr int MA = PAIR == 1 ? MODE_S1 : MODE_S3, MB = PAIR == 1 ? MODE_S2 : MODE_S4;
  const int2 chunk = work[blockIdx.x % kNumXcd];  // as in batch_stage_kernel
  const int2 *items = work + kNumXcd + chunk.x;
  for (int w = blockIdx.x / kNumXcd; w < chunk.y; w += gridDim.x / kNumXcd) {
    const int2 it = items[w];
    const BatchInst<T> &d = inst[it.x];
    if (it.y < d.p_n_edge_blocks)
      frame_patch_pair<T, ORDER, PH_ACOUSTIC, MAT, PAIR, KAHAN, NT>(batch_frame_args<T, MA>(d),
                                                                   batch_frame_args<T, MB>(d), it.y);
    else
      pair_tile_rw<T, ORDER, PH_ACOUSTIC, MAT, PAIR, KAHAN, NT, PF, NTS>(batch_pair_args<T, PAIR>(d),
                                                                        it.y - d.p_n_edge_blocks);
    __syncthreads();  // the next item reuses the LDS rows
  }
}

// part: 1 or 3; the pair launches take the prefetch depths of sbp_kernels.hpp
template <typename T, int ORDER, int MAT, bool KAHAN> struct BatchPairLauncher {
  static constexpr int NT = pair_nt(sizeof(T), ORDER);
  template <int PART, int PF, bool NTS>
  static hipError_t go(const void *inst, const int2 *work, int nwork, int grid, hipStream_t s)
  {
    hipLaunchKernelGGL((batch_pair_kernel<T, ORDER, MAT, PART, KAHAN, NT, PF, NTS>), dim3(grid), dim3(NT), 0, s,
                       static_cast<const BatchInst<T> *>(inst), work, nwork);
    return hipGetLastError();
  }
  // part | kBatchCachedStores: the pair tiles store cached (pair_cached_stores)
  static hipError_t launch(int part, const void *inst, const int2 *work, int nwork, int grid, hipStream_t s)
  {
    constexpr int PF1 = pair_pf(sizeof(T), 1), PF2 = pair_pf(sizeof(T), 2, MAT == M_ARR);
    switch (part) {
      case 1: return go<1, PF1, true>(inst, work, nwork, grid, s);
      case 1 | kBatchCachedStores: return go<1, PF1, false>(inst, work, nwork, grid, s);
      case 3: return go<3, PF2, true>(inst, work, nwork, grid, s);
      default: return go<3, PF2, false>(inst, work, nwork, grid, s);
    }
  }
};

template <typename T, bool KAHAN> BatchLaunchFn pick_batch_pair(int order, int mat)
{
  const bool arr = mat == M_ARR;
  if (order == 2)
    return arr ? &BatchPairLauncher<T, 2, M_ARR, KAHAN>::launch : &BatchPairLauncher<T, 2, M_UNIT, KAHAN>::launch;
  if (order == 4)
    return arr ? &BatchPairLauncher<T, 4, M_ARR, KAHAN>::launch : &BatchPairLauncher<T, 4, M_UNIT, KAHAN>::launch;
  return arr ? &BatchPairLauncher<T, 6, M_ARR, KAHAN>::launch : &BatchPairLauncher<T, 6, M_UNIT, KAHAN>::launch;
}

template <typename T, int ORDER, int MAT, bool KAHAN, int BX, int PF> struct BatchLauncher {
  template <int MODE>
  static hipError_t go(const void *inst, const int2 *work, int nwork, int grid, hipStream_t s)
  {
    hipLaunchKernelGGL((batch_stage_kernel<T, ORDER, MAT, MODE, KAHAN, BX, PF>), dim3(grid), dim3(BX), 0, s,
                       static_cast<const BatchInst<T> *>(inst), work, nwork);
    return hipGetLastError();
  }
  static hipError_t launch(int mode, const void *inst, const int2 *work, int nwork, int grid, hipStream_t s)
  {
    switch (mode) {
      case MODE_S1: return go<MODE_S1>(inst, work, nwork, grid, s);
      case MODE_S2: return go<MODE_S2>(inst, work, nwork, grid, s);
      case MODE_S3: return go<MODE_S3>(inst, work, nwork, grid, s);
      default: return go<MODE_S4>(inst, work, nwork, grid, s);
    }
  }
};

template <typename T, bool KAHAN, int BX, int PF> BatchLaunchFn pick_batch(int order, int mat)
{
  const bool arr = mat == M_ARR;
  if (order == 2)
    return arr ? &BatchLauncher<T, 2, M_ARR, KAHAN, BX, PF>::launch : &BatchLauncher<T, 2, M_UNIT, KAHAN, BX, PF>::launch;
  if (order == 4)
    return arr ? &BatchLauncher<T, 4, M_ARR, KAHAN, BX, PF>::launch : &BatchLauncher<T, 4, M_UNIT, KAHAN, BX, PF>::launch;
  return arr ? &BatchLauncher<T, 6, M_ARR, KAHAN, BX, PF>::launch : &BatchLauncher<T, 6, M_UNIT, KAHAN, BX, PF>::launch;
}

// sbp_batch_f64.hip / sbp_batch_f32.hip; pf is 1 or 4
BatchLaunchFn batch_launcher_f64(int order, int mat, int bx, int pf);
BatchLaunchFn batch_launcher_f32(int order, int mat, bool kahan, int bx, int pf);
// two-stage configurations (the launch's `mode` argument is the part, 1..4)
BatchLaunchFn batch_pair_launcher_f64(int order, int mat);
BatchLaunchFn batch_pair_launcher_f32(int order, int mat, bool kahan);

}  // namespace sbp
