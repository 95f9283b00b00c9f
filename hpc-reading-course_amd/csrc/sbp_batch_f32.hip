// FP32 and FP32+Kahan persistent stage kernels of the batched ensemble.
#include "sbp_batch.hpp"

namespace sbp {
BatchLaunchFn batch_launcher_f32(int order, int mat, bool kahan, int bx)
{
  if (bx == 256) return kahan ? pick_batch<float, true, 256>(order, mat) : pick_batch<float, false, 256>(order, mat);
  if (bx == 128) return kahan ? pick_batch<float, true, 128>(order, mat) : pick_batch<float, false, 128>(order, mat);
  return nullptr;
}
}  // namespace sbp
