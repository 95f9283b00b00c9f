// FP64 persistent stage kernels of the batched ensemble (built without FMA contraction).
#include "sbp_batch.hpp"

namespace sbp {
template <int PF> BatchLaunchFn batch_f64_pf(int order, int mat, int bx)
{
  if (bx == 256) return pick_batch<double, false, 256, PF>(order, mat);
  if (bx == 128) return pick_batch<double, false, 128, PF>(order, mat);
  return nullptr;
}

BatchLaunchFn batch_launcher_f64(int order, int mat, int bx, int pf)
{
  if (pf == 1) return batch_f64_pf<1>(order, mat, bx);
  if (pf == 4) return batch_f64_pf<4>(order, mat, bx);
  return nullptr;
}
}  // namespace sbp
