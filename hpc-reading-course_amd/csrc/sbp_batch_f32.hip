// FP32 and FP32+Kahan persistent stage kernels of the batched ensemble.
#include "sbp_batch.hpp"

namespace sbp {
template <int PF> BatchLaunchFn batch_f32_pf(int order, int mat, bool kahan, int bx)
{
  if (bx == 256) return kahan ? pick_batch<float, true, 256, PF>(order, mat) : pick_batch<float, false, 256, PF>(order, mat);
  if (bx == 128) return kahan ? pick_batch<float, true, 128, PF>(order, mat) : pick_batch<float, false, 128, PF>(order, mat);
  return nullptr;
}

BatchLaunchFn batch_launcher_f32(int order, int mat, bool kahan, int bx, int pf)
{
  if (pf == 1) return batch_f32_pf<1>(order, mat, kahan, bx);
  if (pf == 4) return batch_f32_pf<4>(order, mat, kahan, bx);
  return nullptr;
}
}  // namespace sbp
