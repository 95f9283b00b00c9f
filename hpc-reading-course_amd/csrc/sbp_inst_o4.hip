// Stage kernels of the order-4 SBP operators.
#include "sbp_inst.hpp"
SBP_DEFINE_LAUNCHER(4)
