// FP32 and FP32+Kahan two-stage persistent kernels of the batched ensemble.
#include "sbp_batch.hpp"

namespace sbp {
BatchLaunchFn batch_pair_launcher_f32(int order, int mat, bool kahan)
{
  return kahan ? pick_batch_pair<float, true>(order, mat) : pick_batch_pair<float, false>(order, mat);
}
}  // namespace sbp
