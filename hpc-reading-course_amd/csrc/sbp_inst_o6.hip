// Stage kernels of the order-6 SBP operators.
#include "sbp_inst.hpp"
SBP_DEFINE_LAUNCHER(6)
