// sbp_inst.hpp -- maps (physics, material, precision, block width) to the
// stage-kernel launcher of one SBP order.  Included by sbp_inst_o{2,4,6}_{f64,f32}.hip
// so orders and precisions compile in parallel (and the FP64 units can be
// built without FMA contraction).
#pragma once
#include "sbp_kernels.hpp"

namespace sbp {

template <typename T, int ORDER, bool KAHAN, int BX>
StageLaunchFn pick_launcher(int phys, int mat)
{
  if (phys == PH_ADV1D) return &Launcher1D<T, ORDER, PH_ADV1D, KAHAN>::launch;
  if (phys == PH_REFL1D) return &Launcher1D<T, ORDER, PH_REFL1D, KAHAN>::launch;
  if (phys == PH_ADV2D) {  // M_ARR: per-point velocity fields
    if (mat == M_ARR) return &Launcher2D<T, ORDER, PH_ADV2D, M_ARR, KAHAN, BX>::launch;
    return &Launcher2D<T, ORDER, PH_ADV2D, M_UNIT, KAHAN, BX>::launch;
  }
  if (phys == PH_ACOUSTIC) {
    if (mat == M_UNIT) return &Launcher2D<T, ORDER, PH_ACOUSTIC, M_UNIT, KAHAN, BX>::launch;
    if (mat == M_ARR) return &Launcher2D<T, ORDER, PH_ACOUSTIC, M_ARR, KAHAN, BX>::launch;
    // rho = 2 + xy and the manufactured forcing: evaluated in FP64 from FP64
    // tables in every precision, then rounded to T
    if (mat == M_REF) return &Launcher2D<T, ORDER, PH_ACOUSTIC, M_REF, KAHAN, BX>::launch;
  }
  return nullptr;
}

// Two-stage interior tiles: pair_nt threads; prefetch depths (rows of loads in
// flight per wave) from a sweep on gfx950 (pair_pf).  2D acoustics with unit,
// array or reference materials; 2D advection with constant or per-point
// velocity (M_ARR).
template <typename T, int ORDER, bool KAHAN>
PairLaunchFn pick_pair(int phys, int mat)
{
  constexpr int PF1 = pair_pf(sizeof(T), 1), PF2 = pair_pf(sizeof(T), 2);
  constexpr int PF2A = pair_pf(sizeof(T), 2, true);
  constexpr int NT = pair_nt(sizeof(T), ORDER);
  if (phys == PH_ACOUSTIC) {
    if (mat == M_UNIT) return &PairLauncher<T, ORDER, PH_ACOUSTIC, M_UNIT, KAHAN, NT, PF1, PF2>::launch;
    if (mat == M_ARR) return &PairLauncher<T, ORDER, PH_ACOUSTIC, M_ARR, KAHAN, NT, PF1, PF2A>::launch;
    if (mat == M_REF) return &PairLauncher<T, ORDER, PH_ACOUSTIC, M_REF, KAHAN, NT, PF1, PF2>::launch;
  }
  if (phys == PH_ADV2D) {
    if (mat == M_ARR) return &PairLauncher<T, ORDER, PH_ADV2D, M_ARR, KAHAN, NT, PF1, PF2>::launch;
    return &PairLauncher<T, ORDER, PH_ADV2D, M_UNIT, KAHAN, NT, PF1, PF2>::launch;
  }
  return nullptr;
}

}  // namespace sbp

#define SBP_DEFINE_LAUNCHER_F64(ORDER)                                          \
  namespace sbp {                                                               \
  StageLaunchFn launcher_o##ORDER##_f64(int phys, int mat, int bx)              \
  {                                                                             \
    if (bx == 256) return pick_launcher<double, ORDER, false, 256>(phys, mat);  \
    return nullptr;                                                             \
  }                                                                             \
  PairLaunchFn pair_launcher_o##ORDER##_f64(int phys, int mat)                  \
  {                                                                             \
    return pick_pair<double, ORDER, false>(phys, mat);                          \
  }                                                                             \
  }

#define SBP_DEFINE_LAUNCHER_F32(ORDER)                                          \
  namespace sbp {                                                               \
  StageLaunchFn launcher_o##ORDER##_f32(int phys, int mat, bool kahan, int bx)  \
  {                                                                             \
    if (bx != 256) return nullptr;                                              \
    return kahan ? pick_launcher<float, ORDER, true, 256>(phys, mat)            \
                 : pick_launcher<float, ORDER, false, 256>(phys, mat);          \
  }                                                                             \
  PairLaunchFn pair_launcher_o##ORDER##_f32(int phys, int mat, bool kahan)      \
  {                                                                             \
    return kahan ? pick_pair<float, ORDER, true>(phys, mat)                     \
                 : pick_pair<float, ORDER, false>(phys, mat);                   \
  }                                                                             \
  }
