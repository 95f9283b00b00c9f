// FP32 and FP32+Kahan stage kernels of the order-4 SBP operators.
#include "sbp_inst.hpp"
SBP_DEFINE_LAUNCHER_F32(4)
