// sbp_batch_capi.hip -- host side of the batched-ensemble C-ABI.
//
// Every instance owns a slice of six pooled allocations (U, A, Y0, Y1, the
// optional Kahan compensation and the rho^-1 / K planes), laid out exactly as
// a single-problem context lays out its registers.  A step is four launches
// of the persistent stage kernel over the flattened work list (one-stage
// configurations) or two of the persistent pair kernel (two-stage
// configurations: stages 1-2 and 3-4, each over every instance's frame
// patches and pair tiles); nsteps are replayed from a captured hipGraph.  The
// schedule, tile shape and number of persistent blocks come from a small
// table of candidates, chosen by sbp_batch_autotune on the ensemble at hand.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <new>
#include <string>
#include <vector>

#include "../../include/sbp_hip.h"
#include "sbp_batch.hpp"
#include "sbp_batch_impl.hpp"
#include "sbp_host.hpp"

using namespace sbp;
using sbp::host::fail;
using sbp::host::InstHost;
#define HIP_OR_FAIL(expr) SBP_HIP_OR_FAIL(expr)

namespace {

struct BatchConfig {
  int bx, ty, blocks_per_cu, pf;  // pf: rows of loads in flight per tile
  int pair;                        // 1: two RK4 stages per interior launch (bx from pair_nt)
};

// Candidate launch shapes of the autotuner.
// Short tiles (4, 8 rows) are for ensembles of small grids, where tall tiles
// leave most of the persistent blocks without work.
// The last nine are two-stage configurations (sbp_kernels.hpp pair2d_kernel);
// they need every instance to have an interior beyond its (CLS+R)-deep frame.
// Their tiles are pair_nt wide; bx only sizes the unused one-stage fields.
constexpr BatchConfig kConfigs[] = {
    {256, 32, 4, 1, 0}, {256, 64, 4, 1, 0}, {256, 16, 4, 1, 0}, {256, 32, 2, 1, 0}, {256, 32, 8, 1, 0},
    {256, 64, 2, 1, 0}, {128, 32, 8, 1, 0}, {128, 64, 8, 1, 0}, {128, 16, 8, 1, 0}, {128, 32, 4, 1, 0},
    {128, 64, 4, 1, 0}, {128, 32, 16, 1, 0}, {256, 8, 4, 1, 0},  {256, 4, 4, 1, 0},  {128, 8, 8, 1, 0},
    {128, 4, 8, 1, 0},  {256, 8, 4, 4, 0},  {256, 4, 4, 4, 0},  {128, 8, 8, 4, 0},  {128, 4, 8, 4, 0},
    {256, 16, 4, 0, 1}, {256, 8, 4, 0, 1},  {256, 32, 4, 0, 1}, {256, 16, 2, 0, 1}, {256, 16, 3, 0, 1},
    {256, 24, 3, 0, 1}, {256, 12, 3, 0, 1}, {256, 24, 2, 0, 1}, {256, 32, 2, 0, 1},
};
constexpr int kNumConfigs = sizeof(kConfigs) / sizeof(kConfigs[0]);
constexpr int kAlignB = 64;
constexpr int kBatchGraphSteps = 10;

int closure_of(int order) { return order == 2 ? 1 : (order == 4 ? 4 : 6); }


}  // namespace


namespace {

void drop_graph(sbp_batch *b)
{
  for (hipGraphExec_t &g : b->gexec) {
    if (g) (void)hipGraphExecDestroy(g);
    g = nullptr;
  }
}

// Bring the state back to the U pool (after an odd number of two-stage steps
// it sits in Y0); the one-stage kernels, tuning and configuration changes
// start from there.
int normalize_parity(sbp_batch *b)
{
  if (!b->par) return SBP_OK;
  HIP_OR_FAIL(hipMemcpyAsync(b->U, b->Y0, b->esz * 3 * (size_t)b->pool, hipMemcpyDeviceToDevice, b->stream));
  b->par = 0;
  return SBP_OK;
}

template <typename T> bool fill_inst(const sbp_batch *b, std::vector<BatchInst<T>> &out, int bx, int ty, bool pair)
{
  bool ok = true;
  out.resize(b->count);
  // bytes a two-stage step streams over the ensemble (U, Y2, Y3, materials,
  // compensation): decides the pair tiles' store policy
  double streamed = 0;
  for (int k = 0; k < b->count; k++)
    streamed += (double)b->inst[k].plane * b->esz * (9 + (b->RI ? 2 : 0) + (b->KC ? 3 : 0));
  const int nts = pair_cached_stores(streamed) ? 0 : 1;
  for (int k = 0; k < b->count; k++) {
    const InstHost &h = b->inst[k];
    BatchInst<T> &d = out[k];
    d.p_nts = nts;
    d.U = static_cast<T *>(b->U) + 3 * h.off;
    d.A = static_cast<T *>(b->A) + 3 * h.off;
    d.Y0 = static_cast<T *>(b->Y0) + 3 * h.off;
    d.Y1 = static_cast<T *>(b->Y1) + 3 * h.off;
    d.comp = b->KC ? static_cast<T *>(b->KC) + 3 * h.off : nullptr;
    d.rinv = b->RI ? static_cast<const T *>(b->RI) + h.off : nullptr;
    d.bulk = b->BK ? static_cast<const T *>(b->BK) + h.off : nullptr;
    d.plane = h.plane;
    d.pitch = h.pitch;
    d.xoff = h.xoff;
    d.nx = h.nx;
    d.ny = h.ny;
    d.hix = T(h.hix);
    d.hiy = T(h.hiy);
    d.dt = h.dt;
    d.bc = h.bc;
    const int wi = h.nx - 2 * b->cls, hi = h.ny - 2 * b->cls;
    d.tiles_x = wi > 0 ? (wi + bx - 1) / bx : 0;
    d.ty_rows = ty;
    d.n_tiles = d.tiles_x * (hi > 0 ? (hi + ty - 1) / ty : 0);
    d.n_edge = 2 * b->cls * h.nx + 2 * b->cls * (h.ny - 2 * b->cls);
    d.n_edge_blocks = (d.n_edge + bx - 1) / bx;
    // two-stage geometry: pair_nt-thread tiles inside a (CLS+R)-deep frame
    const int nt = pair_nt((int)b->esz, b->order), R = b->order / 2, D = b->cls + R;
    const int pw = h.nx - 2 * D, ph = h.ny - 2 * D;
    if (pair && (pw <= 0 || ph <= 0)) ok = false;
    d.p_depth = D;
    d.p_tiles_x = pw > 0 ? (pw + nt - 2 * R - 1) / (nt - 2 * R) : 0;
    d.p_ty_rows = ty;
    d.p_n_tiles = (pw > 0 && ph > 0) ? d.p_tiles_x * ((ph + ty - 1) / ty) : 0;
    d.p_n_edge = 2 * D * h.nx + 2 * D * (h.ny - 2 * D);
    d.p_n_edge_blocks = pw > 0 && ph > 0 ? frame_patch_count(h.nx, h.ny, D) : 0;  // frame patches
  }
  return ok;
}

// Upload descriptors and the work list for configuration cfg.
int apply_config(sbp_batch *b, int cfg)
{
  const BatchConfig c = kConfigs[cfg];
  const bool kahan = b->prec == SBP_FP32_KAHAN;
  BatchLaunchFn fn;
  if (c.pair)
    fn = b->prec == SBP_FP64 ? batch_pair_launcher_f64(b->order, b->mat)
                             : batch_pair_launcher_f32(b->order, b->mat, kahan);
  else
    fn = b->prec == SBP_FP64 ? batch_launcher_f64(b->order, b->mat, c.bx, c.pf)
                             : batch_launcher_f32(b->order, b->mat, kahan, c.bx, c.pf);
  if (!fn) return fail(SBP_EINVAL, "no batched kernel for this order/material/precision");
  // Work lists: kNumXcd {offset, count} headers, then the items of each
  // XCD's chunk.  Whole instances are bin-packed onto the XCDs by work
  // (heaviest first onto the least loaded XCD); a chunk holds the tiles of
  // its instances (largest first), then their frame blocks (two-stage: frame
  // patches), so long items start early and short ones fill the tail on every
  // XCD alike.
  std::vector<int2> work;
  auto build = [&](auto &desc) {
    const int R = b->order / 2;
    auto tiles = [&](int k) { return c.pair ? desc[k].p_n_tiles : desc[k].n_tiles; };
    auto eblocks = [&](int k) { return c.pair ? desc[k].p_n_edge_blocks : desc[k].n_edge_blocks; };
    std::vector<double> weight(b->count);  // in 256-point rows (a frame patch: about 4, two stages on its grown points)
    for (int k = 0; k < b->count; k++)
      weight[k] = c.pair ? (double)tiles(k) * (c.ty + 4 * R) + 4.0 * eblocks(k)
                         : (double)tiles(k) * (c.ty + 2 * R) * c.bx / 256.0 + eblocks(k);
    std::vector<int> order(b->count);
    for (int k = 0; k < b->count; k++) order[k] = k;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return weight[x] > weight[y]; });
    std::vector<std::vector<int>> on(kNumXcd);
    std::vector<double> load(kNumXcd, 0.0);
    for (int k : order) {
      const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
      on[x].push_back(k);
      load[x] += weight[k];
    }
    auto emit = [&](std::vector<int2> &list, bool with_tiles) {
      std::vector<int2> items;
      list.assign(kNumXcd, make_int2(0, 0));
      for (int x = 0; x < kNumXcd; x++) {
        const int start = (int)items.size();
        if (with_tiles)
          for (int k : on[x])
            for (int t = 0; t < tiles(k); t++) items.push_back(make_int2(k, eblocks(k) + t));
        for (int k : on[x])
          for (int e = 0; e < eblocks(k); e++) items.push_back(make_int2(k, e));
        list[x] = make_int2(start, (int)items.size() - start);
      }
      list.insert(list.end(), items.begin(), items.end());
    };
    emit(work, true);  // every stage (one-stage), parts 1 and 3 (two-stage)
  };
  // the second descriptor set trades U and Y0 (state of an odd two-stage step)
  auto swapped = [](auto v) {
    for (auto &d : v) std::swap(d.U, d.Y0);
    return v;
  };
  size_t desc_bytes;
  std::vector<BatchInst<double>> dd;
  std::vector<BatchInst<float>> df;
  bool fits;
  if (b->esz == 8) {
    fits = fill_inst(b, dd, c.bx, c.ty, c.pair);
    build(dd);
    desc_bytes = dd.size() * sizeof(BatchInst<double>);
    const auto sw = swapped(dd);
    dd.insert(dd.end(), sw.begin(), sw.end());
  } else {
    fits = fill_inst(b, df, c.bx, c.ty, c.pair);
    build(df);
    desc_bytes = df.size() * sizeof(BatchInst<float>);
    const auto sw = swapped(df);
    df.insert(df.end(), sw.begin(), sw.end());
  }
  if (!fits) return fail(SBP_EINVAL, "two-stage configurations need every instance larger than 2(CLS+R) points");
  int rc = normalize_parity(b);
  if (rc) return rc;
  HIP_OR_FAIL(hipStreamSynchronize(b->stream));
  drop_graph(b);
  if (b->work_dev) (void)hipFree(b->work_dev);
  if (b->inst_dev) (void)hipFree(b->inst_dev);
  b->work_dev = nullptr;
  b->inst_dev = nullptr;
  HIP_OR_FAIL(hipMalloc(&b->work_dev, sizeof(int2) * std::max<size_t>(1, work.size())));
  HIP_OR_FAIL(hipMalloc(&b->inst_dev, 2 * desc_bytes));
  HIP_OR_FAIL(hipMemcpy(b->work_dev, work.data(), sizeof(int2) * work.size(), hipMemcpyHostToDevice));
  HIP_OR_FAIL(hipMemcpy(b->inst_dev, b->esz == 8 ? (const void *)dd.data() : (const void *)df.data(), 2 * desc_bytes,
                        hipMemcpyHostToDevice));
  b->desc_bytes = desc_bytes;
  b->nwork = (int)work.size();
  b->pair = c.pair != 0;
  b->auto_nts = b->esz == 8 ? dd[0].p_nts : df[0].p_nts;
  b->pair_nts = b->store_policy == 0 ? b->auto_nts : (b->store_policy == 2 ? 1 : 0);
  b->grid = std::max(kNumXcd, (b->num_cus * c.blocks_per_cu + kNumXcd - 1) / kNumXcd * kNumXcd);
  b->launch = fn;
  b->config = cfg;
  return SBP_OK;
}

// Rewrite the per-instance time steps in the device descriptors (graph-safe:
// the graph reads them from memory).
int upload_dt(sbp_batch *b, const double *dt)
{
  bool same = true;
  for (int k = 0; k < b->count; k++)
    if (b->inst[k].dt != dt[k]) same = false;
  if (same) return SBP_OK;
  for (int k = 0; k < b->count; k++) {
    if (!(dt[k] > 0) || !std::isfinite(dt[k])) return fail(SBP_EINVAL, "time steps must be positive");
    b->inst[k].dt = dt[k];
  }
  const size_t stride = b->esz == 8 ? sizeof(BatchInst<double>) : sizeof(BatchInst<float>);
  const size_t off = b->esz == 8 ? offsetof(BatchInst<double>, dt) : offsetof(BatchInst<float>, dt);
  HIP_OR_FAIL(hipStreamSynchronize(b->stream));
  for (int set = 0; set < 2; set++)
    for (int k = 0; k < b->count; k++)
      HIP_OR_FAIL(hipMemcpy(static_cast<char *>(b->inst_dev) + set * b->desc_bytes + k * stride + off,
                            &b->inst[k].dt, sizeof(double), hipMemcpyHostToDevice));
  return SBP_OK;
}

hipError_t launch_step(sbp_batch *b)
{
  hipError_t e = hipSuccess;
  if (b->pair) {  // parts 1 and 3 on the descriptors of the current parity
    const void *inst = sbp::host::batch_inst(b);
    for (int p = 1; p <= 3 && e == hipSuccess; p += 2)
      e = b->launch(p | (b->pair_nts ? 0 : kBatchCachedStores), inst, b->work_dev, b->nwork, b->grid, b->stream);
    if (e == hipSuccess) b->par ^= 1;  // the new states sit in the other pool only if both parts went in
    return e;
  }
  for (int m = MODE_S1; m <= MODE_S4 && e == hipSuccess; m++)
    e = b->launch(m, b->inst_dev, b->work_dev, b->nwork, b->grid, b->stream);
  return e;
}

// kBatchGraphSteps is even: a two-stage capture ends at the parity it began.
int ensure_graph(sbp_batch *b)
{
  if (b->gexec[b->par]) return SBP_OK;
  const int par0 = b->par;  // captured steps flip the parity as if they had run
  hipGraph_t g = nullptr;
  HIP_OR_FAIL(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  hipError_t le = hipSuccess;
  for (int s = 0; s < kBatchGraphSteps && le == hipSuccess; s++) le = launch_step(b);
  hipError_t ce = hipStreamEndCapture(b->stream, &g);
  b->par = par0;
  if (le != hipSuccess || ce != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return fail(SBP_EHIP, std::string("batch graph capture: ") + hipGetErrorString(le != hipSuccess ? le : ce));
  }
  hipError_t ie = hipGraphInstantiate(&b->gexec[b->par], g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (ie != hipSuccess) {
    b->gexec[b->par] = nullptr;
    return fail(SBP_EHIP, std::string("batch graph instantiate: ") + hipGetErrorString(ie));
  }
  return SBP_OK;
}

int run(sbp_batch *b, long nsteps)
{
  long n = 0;
  if (nsteps >= kBatchGraphSteps) {
    int rc = ensure_graph(b);
    if (rc) return rc;
    for (; n + kBatchGraphSteps <= nsteps; n += kBatchGraphSteps) HIP_OR_FAIL(hipGraphLaunch(b->gexec[b->par], b->stream));
  }
  for (; n < nsteps; n++) HIP_OR_FAIL(launch_step(b));
  return SBP_OK;
}

void release(sbp_batch *b)
{
  if (!b) return;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  drop_graph(b);
  for (void *p : {b->U, b->A, b->Y0, b->Y1, b->KC, b->RI, b->BK, b->inst_dev, (void *)b->work_dev, (void *)b->aos,
                  (void *)b->partials, (void *)b->scalar})
    if (p) (void)hipFree(p);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

int check_index(const sbp_batch *b, int k)
{
  if (!b) return fail(SBP_EINVAL, "null batch");
  if (k < 0 || k >= b->count) return fail(SBP_EINVAL, "instance index out of range");
  return SBP_OK;
}

int build_batch(const sbp_problem *descs, int count, sbp_batch *b)
{
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(SBP_ENODEV, "no HIP device");
  const sbp_problem &d0 = descs[0];
  b->count = count;
  b->order = d0.order;
  b->cls = closure_of(d0.order);
  b->prec = d0.precision;
  b->mat = d0.material;
  b->esz = b->prec == SBP_FP64 ? 8 : 4;
  int dev = 0;
  HIP_OR_FAIL(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  HIP_OR_FAIL(hipGetDeviceProperties(&prop, dev));
  b->num_cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
  const int xoff = (kAlignB - b->cls % kAlignB) % kAlignB;
  long off = 0;
  b->inst.resize(count);
  for (int k = 0; k < count; k++) {
    const sbp_problem &d = descs[k];
    InstHost &h = b->inst[k];
    h.nx = d.nx;
    h.ny = d.ny;
    h.xoff = xoff;
    h.pitch = (xoff + d.nx + kAlignB - 1) / kAlignB * kAlignB;
    h.plane = (long)h.pitch * d.ny;
    h.off = off;
    off += (h.plane + kAlignB - 1) / kAlignB * kAlignB;
    h.hix = (d.nx - 1) / (d.xr - d.xl);
    h.hiy = (d.ny - 1) / (d.yr - d.yl);
    h.dt = 0;
    h.bc = d.bc;
    b->max_pts = std::max(b->max_pts, (long)d.nx * d.ny);
  }
  b->pool = off;  // planes per component; a register holds 3 * pool elements
  b->have_state.assign(count, 0);
  HIP_OR_FAIL(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
  HIP_OR_FAIL(hipEventCreate(&b->ev0));
  HIP_OR_FAIL(hipEventCreate(&b->ev1));
  const size_t reg = b->esz * 3 * (size_t)b->pool;
  for (void **p : {&b->U, &b->A, &b->Y0, &b->Y1}) {
    HIP_OR_FAIL(hipMalloc(p, reg));
    HIP_OR_FAIL(hipMemsetAsync(*p, 0, reg, b->stream));
  }
  if (b->prec == SBP_FP32_KAHAN) {
    HIP_OR_FAIL(hipMalloc(&b->KC, reg));
    HIP_OR_FAIL(hipMemsetAsync(b->KC, 0, reg, b->stream));
  }
  HIP_OR_FAIL(hipMalloc(&b->aos, sizeof(double) * 3 * (size_t)b->max_pts));
  HIP_OR_FAIL(hipMalloc(&b->partials, sizeof(double) * 1024));
  HIP_OR_FAIL(hipMalloc(&b->scalar, sizeof(double)));
  if (b->mat == SBP_MAT_ARRAYS) {
    HIP_OR_FAIL(hipMalloc(&b->RI, b->esz * (size_t)b->pool));
    HIP_OR_FAIL(hipMalloc(&b->BK, b->esz * (size_t)b->pool));
    for (int k = 0; k < count; k++) {
      const InstHost &h = b->inst[k];
      const long n = (long)h.nx * h.ny;
      const int grid = (int)((n + 255) / 256);
      for (int which = 0; which < 2; which++) {
        const double *src = which == 0 ? descs[k].rho_inv : descs[k].bulk;
        void *dst = which == 0 ? b->RI : b->BK;
        HIP_OR_FAIL(hipMemcpyAsync(b->aos, src, sizeof(double) * n, hipMemcpyHostToDevice, b->stream));
        if (b->esz == 8)
          hipLaunchKernelGGL(matrix_to_plane_kernel<double>, dim3(grid), dim3(256), 0, b->stream, b->aos,
                             static_cast<double *>(dst) + h.off, h.nx, h.ny, h.pitch, h.xoff);
        else
          hipLaunchKernelGGL(matrix_to_plane_kernel<float>, dim3(grid), dim3(256), 0, b->stream, b->aos,
                             static_cast<float *>(dst) + h.off, h.nx, h.ny, h.pitch, h.xoff);
        HIP_OR_FAIL(hipGetLastError());
        HIP_OR_FAIL(hipStreamSynchronize(b->stream));
      }
    }
  }
  HIP_OR_FAIL(hipStreamSynchronize(b->stream));
  return apply_config(b, 0);
}

int validate_batch(const sbp_problem *descs, int count)
{
  if (!descs || count < 1) return fail(SBP_EINVAL, "empty batch");
  const sbp_problem &d0 = descs[0];
  if (d0.order != 2 && d0.order != 4 && d0.order != 6) return fail(SBP_EINVAL, "order must be 2, 4 or 6");
  const int nmin = d0.order == 2 ? 3 : (d0.order == 4 ? 8 : 12);
  for (int k = 0; k < count; k++) {
    const sbp_problem &d = descs[k];
    if (d.physics != SBP_ACOUSTIC_2D) return fail(SBP_EINVAL, "batches hold 2D acoustic problems");
    if (d.order != d0.order || d.precision != d0.precision || d.material != d0.material)
      return fail(SBP_EINVAL, "all instances need the same order, precision and material model");
    if (d.material != SBP_MAT_UNIT && d.material != SBP_MAT_ARRAYS)
      return fail(SBP_EINVAL, "batched instances use unit or array materials");
    if (d.material == SBP_MAT_ARRAYS && (!d.rho_inv || !d.bulk)) return fail(SBP_EINVAL, "missing material arrays");
    if (d.nx < nmin || d.ny < nmin) return fail(SBP_EINVAL, "instance grid below the closure size");
    if (!(d.xr > d.xl) || !(d.yr > d.yl)) return fail(SBP_EINVAL, "empty domain");
    if (d.bc < SBP_BC_DEFAULT || d.bc > SBP_BC_WALL) return fail(SBP_EINVAL, "bad boundary condition");
    if (d.precision < SBP_FP64 || d.precision > SBP_FP32_KAHAN) return fail(SBP_EINVAL, "bad precision");
  }
  return SBP_OK;
}

int norm_instance(sbp_batch *b, int k, double *out)
{
  const InstHost &h = b->inst[k];
  const long n = (long)h.nx * h.ny;
  const int grid = (int)std::min<long>((n + 255) / 256, 1024);
  auto go = [&](auto *tag) {
    using T = std::remove_pointer_t<decltype(tag)>;
    NormArgs<T> a{};
    a.aos = nullptr;
    a.u = static_cast<const T *>(sbp::host::batch_state(b)) + 3 * h.off;
    a.comp = b->KC ? static_cast<const T *>(b->KC) + 3 * h.off : nullptr;
    a.rinv = b->RI ? static_cast<const T *>(b->RI) + h.off : nullptr;
    a.bulk = b->BK ? static_cast<const T *>(b->BK) + h.off : nullptr;
    a.plane = h.plane;
    a.pitch = h.pitch;
    a.xoff = h.xoff;
    a.nx = h.nx;
    a.ny = h.ny;
    a.nc = 3;
    a.order = b->order;
    a.is1d = 0;
    a.material = b->mat;
    a.weighted = 1;
    a.hx = 1.0 / h.hix;
    a.hy = 1.0 / h.hiy;
    hipLaunchKernelGGL(norm_partial_kernel<T>, dim3(grid), dim3(256), 0, b->stream, a, b->partials);
  };
  if (b->esz == 8) go((double *)nullptr);
  else go((float *)nullptr);
  HIP_OR_FAIL(hipGetLastError());
  hipLaunchKernelGGL(norm_final_kernel, dim3(1), dim3(256), 0, b->stream, b->partials, grid, b->scalar);
  HIP_OR_FAIL(hipGetLastError());
  HIP_OR_FAIL(hipMemcpyAsync(out, b->scalar, sizeof(double), hipMemcpyDeviceToHost, b->stream));
  HIP_OR_FAIL(hipStreamSynchronize(b->stream));
  return SBP_OK;
}

}  // namespace

hipError_t sbp::host::batch_launch_step(sbp_batch *b) { return ::launch_step(b); }
int sbp::host::batch_upload_dt(sbp_batch *b, const double *dt) { return ::upload_dt(b, dt); }

extern "C" {

int sbp_batch_create(const sbp_problem *descs, int count, sbp_batch **out)
{
  if (!out) return fail(SBP_EINVAL, "null output pointer");
  *out = nullptr;
  int rc = validate_batch(descs, count);
  if (rc) return rc;
  sbp_batch *b = new (std::nothrow) sbp_batch();
  if (!b) return fail(SBP_ENOMEM, "host allocation failed");
  rc = build_batch(descs, count, b);
  if (rc) {
    std::string msg = sbp_last_error();
    release(b);
    fail(rc, msg);
    return rc;
  }
  *out = b;
  return SBP_OK;
}

int sbp_batch_destroy(sbp_batch *b)
{
  release(b);
  return SBP_OK;
}

int sbp_batch_count(const sbp_batch *b) { return b ? b->count : 0; }

int sbp_batch_set_state(sbp_batch *b, int k, const double *v)
{
  int rc = check_index(b, k);
  if (rc) return rc;
  if (!v) return fail(SBP_EINVAL, "null vector");
  const InstHost &h = b->inst[k];
  const long n = (long)h.nx * h.ny;
  HIP_OR_FAIL(hipMemcpyAsync(b->aos, v, sizeof(double) * 3 * n, hipMemcpyHostToDevice, b->stream));
  const int grid = (int)((n + 255) / 256);
  if (b->esz == 8)
    hipLaunchKernelGGL(aos_to_planar_kernel<double>, dim3(grid), dim3(256), 0, b->stream, b->aos,
                       static_cast<double *>(sbp::host::batch_state(b)) + 3 * h.off, (double *)nullptr, h.nx, h.ny, 3,
                       h.pitch, h.xoff, h.plane);
  else
    hipLaunchKernelGGL(aos_to_planar_kernel<float>, dim3(grid), dim3(256), 0, b->stream, b->aos,
                       static_cast<float *>(sbp::host::batch_state(b)) + 3 * h.off,
                       b->KC ? static_cast<float *>(b->KC) + 3 * h.off : nullptr, h.nx, h.ny, 3, h.pitch, h.xoff,
                       h.plane);
  HIP_OR_FAIL(hipGetLastError());
  HIP_OR_FAIL(hipStreamSynchronize(b->stream));
  b->have_state[k] = 1;
  return SBP_OK;
}

int sbp_batch_get_state(sbp_batch *b, int k, double *v)
{
  int rc = check_index(b, k);
  if (rc) return rc;
  if (!v) return fail(SBP_EINVAL, "null vector");
  if (!b->have_state[k]) return fail(SBP_ESTATE, "no state loaded for this instance");
  const InstHost &h = b->inst[k];
  const long n = (long)h.nx * h.ny;
  const int grid = (int)((n + 255) / 256);
  if (b->esz == 8)
    hipLaunchKernelGGL(planar_to_aos_kernel<double>, dim3(grid), dim3(256), 0, b->stream,
                       static_cast<const double *>(sbp::host::batch_state(b)) + 3 * h.off, (const double *)nullptr, b->aos, h.nx, h.ny,
                       3, h.pitch, h.xoff, h.plane);
  else
    hipLaunchKernelGGL(planar_to_aos_kernel<float>, dim3(grid), dim3(256), 0, b->stream,
                       static_cast<const float *>(sbp::host::batch_state(b)) + 3 * h.off,
                       b->KC ? static_cast<const float *>(b->KC) + 3 * h.off : nullptr, b->aos, h.nx, h.ny, 3,
                       h.pitch, h.xoff, h.plane);
  HIP_OR_FAIL(hipGetLastError());
  HIP_OR_FAIL(hipMemcpyAsync(v, b->aos, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, b->stream));
  HIP_OR_FAIL(hipStreamSynchronize(b->stream));
  return SBP_OK;
}

int sbp_batch_get_state_strided(sbp_batch *b, int k, int stride, double *v)
{
  int rc = check_index(b, k);
  if (rc) return rc;
  if (!v || stride < 1) return fail(SBP_EINVAL, "bad argument");
  if (!b->have_state[k]) return fail(SBP_ESTATE, "no state loaded for this instance");
  const InstHost &h = b->inst[k];
  const int cx = (h.nx - 1) / stride + 1, cy = (h.ny - 1) / stride + 1;
  const long n = (long)cx * cy;
  const int grid = (int)((n + 255) / 256);
  if (b->esz == 8)
    hipLaunchKernelGGL(planar_to_aos_strided_kernel<double>, dim3(grid), dim3(256), 0, b->stream,
                       static_cast<const double *>(sbp::host::batch_state(b)) + 3 * h.off, (const double *)nullptr, b->aos, h.nx, h.ny, 3,
                       h.pitch, h.xoff, h.plane, stride);
  else
    hipLaunchKernelGGL(planar_to_aos_strided_kernel<float>, dim3(grid), dim3(256), 0, b->stream,
                       static_cast<const float *>(sbp::host::batch_state(b)) + 3 * h.off,
                       b->KC ? static_cast<const float *>(b->KC) + 3 * h.off : nullptr, b->aos, h.nx, h.ny, 3, h.pitch,
                       h.xoff, h.plane, stride);
  HIP_OR_FAIL(hipGetLastError());
  HIP_OR_FAIL(hipMemcpyAsync(v, b->aos, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, b->stream));
  HIP_OR_FAIL(hipStreamSynchronize(b->stream));
  return SBP_OK;
}

int sbp_batch_step(sbp_batch *b, const double *dt, long nsteps)
{
  if (!b || !dt || nsteps < 0) return fail(SBP_EINVAL, "bad argument");
  for (int k = 0; k < b->count; k++)
    if (!b->have_state[k]) return fail(SBP_ESTATE, "every instance needs a state before stepping");
  int rc = upload_dt(b, dt);
  if (rc) return rc;
  return run(b, nsteps);
}

int sbp_batch_synchronize(sbp_batch *b)
{
  if (!b) return fail(SBP_EINVAL, "null batch");
  HIP_OR_FAIL(hipStreamSynchronize(b->stream));
  return SBP_OK;
}

int sbp_batch_time_steps(sbp_batch *b, const double *dt, long nsteps, double *ms)
{
  if (!b || !dt || !ms) return fail(SBP_EINVAL, "null argument");
  for (int k = 0; k < b->count; k++)
    if (!b->have_state[k]) return fail(SBP_ESTATE, "every instance needs a state before stepping");
  int rc = upload_dt(b, dt);
  if (rc) return rc;
  if (nsteps >= kBatchGraphSteps && (rc = ensure_graph(b))) return rc;
  HIP_OR_FAIL(hipEventRecord(b->ev0, b->stream));
  if ((rc = run(b, nsteps))) return rc;
  HIP_OR_FAIL(hipEventRecord(b->ev1, b->stream));
  HIP_OR_FAIL(hipEventSynchronize(b->ev1));
  float f = 0;
  HIP_OR_FAIL(hipEventElapsedTime(&f, b->ev0, b->ev1));
  *ms = f;
  return SBP_OK;
}

int sbp_batch_num_configs(void) { return kNumConfigs; }

int sbp_batch_set_config(sbp_batch *b, int config)
{
  if (!b) return fail(SBP_EINVAL, "null batch");
  if (config < 0 || config >= kNumConfigs) return fail(SBP_EINVAL, "configuration index out of range");
  return apply_config(b, config);
}

int sbp_batch_set_store_policy(sbp_batch *b, int policy)
{
  if (!b || policy < 0 || policy > 2) return fail(SBP_EINVAL, "store policy must be 0, 1 or 2");
  HIP_OR_FAIL(hipStreamSynchronize(b->stream));
  drop_graph(b);
  b->store_policy = policy;
  b->pair_nts = policy == 0 ? b->auto_nts : (policy == 2 ? 1 : 0);
  return SBP_OK;
}

int sbp_batch_get_config(const sbp_batch *b, int *config, int *tile_width, int *tile_rows, int *blocks_per_cu,
                         int *two_stage)
{
  if (!b) return fail(SBP_EINVAL, "null batch");
  const BatchConfig c = kConfigs[b->config];
  if (config) *config = b->config;
  if (tile_width) *tile_width = c.pair ? pair_nt((int)b->esz, b->order) : c.bx;
  if (tile_rows) *tile_rows = c.ty;
  if (blocks_per_cu) *blocks_per_cu = c.blocks_per_cu;
  if (two_stage) *two_stage = c.pair;
  return SBP_OK;
}

int sbp_batch_autotune(sbp_batch *b, const double *dt, int *choice)
{
  if (!b || !dt) return fail(SBP_EINVAL, "null argument");
  int rc = upload_dt(b, dt);
  if (rc) return rc;
  // Time every candidate on this ensemble; the resident state is saved first
  // and restored afterwards, so tuning leaves the instances untouched.
  const size_t reg = b->esz * 3 * (size_t)b->pool;
  if ((rc = normalize_parity(b))) return rc;
  void *save_u = nullptr, *save_c = nullptr;
  HIP_OR_FAIL(hipMalloc(&save_u, reg));
  if (b->KC && hipMalloc(&save_c, reg) != hipSuccess) {
    (void)hipFree(save_u);
    return fail(SBP_ENOMEM, "autotune backup allocation failed");
  }
  HIP_OR_FAIL(hipMemcpyAsync(save_u, b->U, reg, hipMemcpyDeviceToDevice, b->stream));
  if (save_c) HIP_OR_FAIL(hipMemcpyAsync(save_c, b->KC, reg, hipMemcpyDeviceToDevice, b->stream));
  int best = b->config;
  float best_ms = 1e30f;
  for (int cfg = 0; cfg < kNumConfigs && rc == SBP_OK; cfg++) {
    const int arc = apply_config(b, cfg);
    if (arc == SBP_EINVAL && kConfigs[cfg].pair) continue;  // instances too small for a two-stage tile
    if ((rc = arc)) break;
    for (int s = 0; s < 2 && rc == SBP_OK; s++)
      if (launch_step(b) != hipSuccess) rc = fail(SBP_EHIP, "autotune warm-up launch failed");
    if (rc) break;
    HIP_OR_FAIL(hipEventRecord(b->ev0, b->stream));
    for (int s = 0; s < 4 && rc == SBP_OK; s++)
      if (launch_step(b) != hipSuccess) rc = fail(SBP_EHIP, "autotune launch failed");
    if (rc) break;
    HIP_OR_FAIL(hipEventRecord(b->ev1, b->stream));
    HIP_OR_FAIL(hipEventSynchronize(b->ev1));
    float f = 0;
    HIP_OR_FAIL(hipEventElapsedTime(&f, b->ev0, b->ev1));
    if (f < best_ms) {
      best_ms = f;
      best = cfg;
    }
  }
  b->par = 0;  // the trial steps' state is discarded: U comes back from the copy
  hipError_t e1 = hipMemcpyAsync(b->U, save_u, reg, hipMemcpyDeviceToDevice, b->stream);
  hipError_t e2 = save_c ? hipMemcpyAsync(b->KC, save_c, reg, hipMemcpyDeviceToDevice, b->stream) : hipSuccess;
  hipError_t e3 = hipStreamSynchronize(b->stream);
  (void)hipFree(save_u);
  if (save_c) (void)hipFree(save_c);
  if (rc) return rc;
  if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess) return fail(SBP_EHIP, "autotune state restore failed");
  if ((rc = apply_config(b, best))) return rc;
  if (choice) *choice = best;
  return SBP_OK;
}

int sbp_batch_energy(sbp_batch *b, int k, double *E)
{
  int rc = check_index(b, k);
  if (rc) return rc;
  if (!E) return fail(SBP_EINVAL, "null output");
  if (!b->have_state[k]) return fail(SBP_ESTATE, "no state loaded for this instance");
  return norm_instance(b, k, E);
}

}  // extern "C"
