// FP64 persistent stage kernels of the batched ensemble (built without FMA contraction).
#include "sbp_batch.hpp"

namespace sbp {
BatchLaunchFn batch_launcher_f64(int order, int mat, int bx)
{
  if (bx == 256) return pick_batch<double, false, 256>(order, mat);
  if (bx == 128) return pick_batch<double, false, 128>(order, mat);
  return nullptr;
}
}  // namespace sbp
