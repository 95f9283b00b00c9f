// sbp_batch_impl.hpp -- the batched-ensemble object shared by the host
// translation units (batch C-ABI, learned corrector).
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/sbp_hip.h"
#include "sbp_batch.hpp"

namespace sbp {
namespace host {
struct InstHost {
  int nx, ny, pitch, xoff;
  long plane, off;  // element offset of the instance's region in every pool
  double hix, hiy, dt;
  int bc;
};
}  // namespace host
}  // namespace sbp

struct sbp_batch {
  int count = 0, order = 0, cls = 0, prec = 0, mat = 0;
  size_t esz = 4;
  std::vector<sbp::host::InstHost> inst;
  long pool = 0;       // elements per register pool
  long max_pts = 0;    // largest nx*ny
  void *U = nullptr, *A = nullptr, *Y0 = nullptr, *Y1 = nullptr, *KC = nullptr;
  void *RI = nullptr, *BK = nullptr;
  void *inst_dev = nullptr;  // two descriptor sets: as allocated, and U / Y0 traded
  size_t desc_bytes = 0;     // bytes of one set
  int2 *work_dev = nullptr;
  int nwork = 0;             // items of the stage launches (two-stage: frame + pair tiles)
  int pair_nts = 1;          // two-stage: nontemporal (1) or cached (0) pair-tile stores
  int auto_nts = 1;          // the same chosen from the streamed bytes
  int store_policy = 0;      // 0: auto_nts, 1: cached, 2: nontemporal
  bool pair = false;         // the configuration runs two RK4 stages per interior launch
  int par = 0;               // 1: a two-stage step left the state in the Y0 pool
  double *aos = nullptr, *partials = nullptr, *scalar = nullptr;
  hipStream_t stream = nullptr;
  hipGraphExec_t gexec[2] = {nullptr, nullptr};  // captured steps, by starting parity
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int config = 0, grid = 0, num_cus = 256;
  sbp::BatchLaunchFn launch = nullptr;
  std::vector<char> have_state;
};

namespace sbp {
namespace host {
// Device descriptors of the current step parity (in the second set U and Y0
// trade places), and the pool that holds the state now.
inline const void *batch_inst(const sbp_batch *b)
{
  return static_cast<const char *>(b->inst_dev) + (b->par ? b->desc_bytes : 0);
}
inline void *batch_state(const sbp_batch *b) { return b->par ? b->Y0 : b->U; }
// One RK4 step of every instance (four persistent-kernel launches).
hipError_t batch_launch_step(sbp_batch *b);
// Upload per-instance time steps into the device descriptors.
int batch_upload_dt(sbp_batch *b, const double *dt);
}  // namespace host
}  // namespace sbp
