// FP64 stage kernels of the order-6 SBP operators (built without FMA contraction).
#include "sbp_inst.hpp"
SBP_DEFINE_LAUNCHER_F64(6)
