// FP64 two-stage persistent kernels of the batched ensemble (built without FMA contraction).
#include "sbp_batch.hpp"

namespace sbp {
BatchLaunchFn batch_pair_launcher_f64(int order, int mat) { return pick_batch_pair<double, false>(order, mat); }
}  // namespace sbp
