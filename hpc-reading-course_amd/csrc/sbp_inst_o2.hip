// Stage kernels of the order-2 SBP operators.
#include "sbp_inst.hpp"
SBP_DEFINE_LAUNCHER(2)
