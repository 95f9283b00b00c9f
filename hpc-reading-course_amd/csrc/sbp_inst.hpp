// sbp_inst.hpp -- maps (physics, material, precision, block width) to the
// stage-kernel launcher of one SBP order.  Included once per order by
// sbp_inst_o{2,4,6}.hip so the three orders compile in parallel.
#pragma once
#include "sbp_kernels.hpp"

namespace sbp {

template <int ORDER, int BX>
StageLaunchFn pick_launcher(int phys, int mat, int prec)
{
  if (phys == PH_ADV1D) {
    if (prec == 0) return &Launcher1D<double, ORDER, PH_ADV1D, false>::launch;
    if (prec == 1) return &Launcher1D<float, ORDER, PH_ADV1D, false>::launch;
    return &Launcher1D<float, ORDER, PH_ADV1D, true>::launch;
  }
  if (phys == PH_REFL1D) {
    if (prec == 0) return &Launcher1D<double, ORDER, PH_REFL1D, false>::launch;
    if (prec == 1) return &Launcher1D<float, ORDER, PH_REFL1D, false>::launch;
    return &Launcher1D<float, ORDER, PH_REFL1D, true>::launch;
  }
  if (phys == PH_ADV2D) {
    if (prec == 0) return &Launcher2D<double, ORDER, PH_ADV2D, M_UNIT, false, BX>::launch;
    if (prec == 1) return &Launcher2D<float, ORDER, PH_ADV2D, M_UNIT, false, BX>::launch;
    return &Launcher2D<float, ORDER, PH_ADV2D, M_UNIT, true, BX>::launch;
  }
  if (phys == PH_ACOUSTIC) {
    if (mat == M_REF) return prec == 0 ? &Launcher2D<double, ORDER, PH_ACOUSTIC, M_REF, false, BX>::launch : nullptr;
    if (mat == M_UNIT) {
      if (prec == 0) return &Launcher2D<double, ORDER, PH_ACOUSTIC, M_UNIT, false, BX>::launch;
      if (prec == 1) return &Launcher2D<float, ORDER, PH_ACOUSTIC, M_UNIT, false, BX>::launch;
      return &Launcher2D<float, ORDER, PH_ACOUSTIC, M_UNIT, true, BX>::launch;
    }
    if (mat == M_ARR) {
      if (prec == 0) return &Launcher2D<double, ORDER, PH_ACOUSTIC, M_ARR, false, BX>::launch;
      if (prec == 1) return &Launcher2D<float, ORDER, PH_ACOUSTIC, M_ARR, false, BX>::launch;
      return &Launcher2D<float, ORDER, PH_ACOUSTIC, M_ARR, true, BX>::launch;
    }
  }
  return nullptr;
}

}  // namespace sbp

#define SBP_DEFINE_LAUNCHER(ORDER)                                              \
  namespace sbp {                                                               \
  StageLaunchFn launcher_o##ORDER(int phys, int mat, int prec, int bx)          \
  {                                                                             \
    if (bx == 256) return pick_launcher<ORDER, 256>(phys, mat, prec);           \
    return nullptr;                                                             \
  }                                                                             \
  }
