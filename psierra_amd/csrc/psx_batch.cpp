// Batched Tet4 assembly, host side: a batch is a list of plans (meshes of
// uneven size, or one plan repeated for parameter samples) whose tiles form
// one global list, assembled by one launch of the tile kernels
// (launch_tet4_batch in psx_tet4.hip).
//
// The batch keeps a device table of the plans' views, refreshed whenever a
// plan's tiling or bound material changed since the last call (plan
// version), and its own displacement buffer per instance, so the same plan
// can appear several times with different displacements.
#include <algorithm>
#include <climits>
#include <string>
#include <vector>

#include "psx_internal.h"

struct psx_batch {
  int device = 0;
  int cus = 256;
  bool persistent = false;
  bool all_bound = false;
  std::vector<psx_plan*> plans;
  std::vector<int64_t> versions;
  psx::BatchDev d{};
  psx::DevPlan* d_plans = nullptr;
  int32_t* d_tile_start = nullptr;
  int32_t* d_node_start = nullptr;
  const double** d_bound = nullptr;   // 2n: bound mu of each plan, then bound lambda
  double4* ui = nullptr;              // the instances' displacements, internal node order
  double4* wi = nullptr;              // the instances' adjoint cotangents, internal node order (lazy)
  std::vector<void*> allocations;
};

namespace psx {

static int32_t batch_alloc(psx_batch* b, size_t bytes, void** out) {
  PSX_HIP_TRY(hipMalloc(out, std::max<size_t>(bytes, 16)));
  b->allocations.push_back(*out);
  return PSX_OK;
}

// Rebuild the device table from the plans' current views.  Refuses what the
// batched kernels do not take: wide plans, plans of another element type,
// device or tile size.
static int32_t batch_refresh(psx_batch* b, bool wait_for_device) {
  const int n = int(b->plans.size());
  std::vector<DevPlan> views(n);
  std::vector<int32_t> tile_start(n + 1), node_start(n + 1);
  std::vector<const double*> bound(2 * n);
  int64_t tiles = 0, nodes = 0;
  int32_t tile = 0, max_inc = 0;
  bool all_bound = true;
  const int npe = b->plans[0]->hp.npe;
  for (int i = 0; i < n; ++i) {
    const psx_plan* p = b->plans[i];
    const std::string which = "plan " + std::to_string(i) + " of the batch";
    if (p->hp.npe != npe)
      return fail(PSX_EINVAL, which + " has another element type than plan 0 (Tet4 and Tet10 plans batch apart)");
    if (p->d.wide)
      return fail(PSX_EUNSUPPORTED, which + " is wide (a node beyond 64 incidences or 255 neighbours): "
                                            "assemble it with psx_assemble");
    if (p->device != b->device) return fail(PSX_EINVAL, which + " lives on another device");
    if (i == 0) tile = p->d.tile;
    if (p->d.tile != tile) return fail(PSX_EINVAL, which + " has another tile size than plan 0");
    views[i] = p->d;
    views[i].ui = b->ui + nodes;
    views[i].wi = b->wi ? b->wi + nodes : nullptr;
    tile_start[i] = int32_t(tiles);
    node_start[i] = int32_t(nodes);
    tiles += p->d.n_tiles;
    nodes += p->d.N;
    if (tiles > INT_MAX || nodes > INT_MAX) return fail(PSX_EUNSUPPORTED, "batch beyond 2^31 tiles or nodes");
    max_inc = std::max(max_inc, p->d.max_inc);
    bound[i] = p->mu_bound;
    bound[n + i] = p->lam_bound;
    all_bound = all_bound && p->bound;
  }
  tile_start[n] = int32_t(tiles);
  node_start[n] = int32_t(nodes);
  // a launch of this batch may still read the table
  if (wait_for_device) PSX_HIP_TRY(hipDeviceSynchronize());
  PSX_HIP_TRY(hipMemcpy(b->d_plans, views.data(), n * sizeof(DevPlan), hipMemcpyHostToDevice));
  PSX_HIP_TRY(hipMemcpy(b->d_tile_start, tile_start.data(), (n + 1) * sizeof(int32_t), hipMemcpyHostToDevice));
  PSX_HIP_TRY(hipMemcpy(b->d_node_start, node_start.data(), (n + 1) * sizeof(int32_t), hipMemcpyHostToDevice));
  PSX_HIP_TRY(hipMemcpy(b->d_bound, bound.data(), 2 * n * sizeof(const double*), hipMemcpyHostToDevice));
  b->d.plans = b->d_plans;
  b->d.tile_start = b->d_tile_start;
  b->d.node_start = b->d_node_start;
  b->d.bound_mu = b->d_bound;
  b->d.bound_lam = b->d_bound + n;
  b->d.n = n;
  b->d.n_tiles = int32_t(tiles);
  b->d.n_nodes = int32_t(nodes);
  b->d.tile = tile;
  b->d.max_inc = max_inc;
  b->d.npe = npe;
  b->all_bound = all_bound;
  for (int i = 0; i < n; ++i) b->versions[i] = b->plans[i]->version;
  return PSX_OK;
}

static int32_t batch_create(psx_plan* const* plans, int32_t n, psx_batch** out) {
  if (!plans || !out || n < 1) return fail(PSX_EINVAL, "psx_batch_create needs n >= 1 plans and an output");
  *out = nullptr;
  for (int i = 0; i < n; ++i)
    if (!plans[i]) return fail(PSX_EINVAL, "null plan in psx_batch_create");
  auto* b = new psx_batch();
  b->plans.assign(plans, plans + n);
  b->versions.assign(n, -1);
  b->device = plans[0]->device;
  int dev = 0;
  (void)hipGetDevice(&dev);
  (void)hipSetDevice(b->device);
  int64_t nodes = 0;
  for (int i = 0; i < n; ++i) nodes += plans[i]->d.N;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, b->device) == hipSuccess && prop.multiProcessorCount > 0)
    b->cus = prop.multiProcessorCount;
  void *a = nullptr, *t = nullptr, *s = nullptr, *m = nullptr, *u = nullptr;
  int32_t rc = PSX_OK;
  if (nodes > INT_MAX) rc = fail(PSX_EUNSUPPORTED, "batch beyond 2^31 nodes");
  if (!rc) rc = batch_alloc(b, n * sizeof(DevPlan), &a);
  if (!rc) rc = batch_alloc(b, (n + 1) * sizeof(int32_t), &t);
  if (!rc) rc = batch_alloc(b, (n + 1) * sizeof(int32_t), &s);
  if (!rc) rc = batch_alloc(b, 2 * n * sizeof(const double*), &m);
  if (!rc) rc = batch_alloc(b, size_t(nodes) * sizeof(double4), &u);
  if (!rc) {
    b->d_plans = static_cast<DevPlan*>(a);
    b->d_tile_start = static_cast<int32_t*>(t);
    b->d_node_start = static_cast<int32_t*>(s);
    b->d_bound = static_cast<const double**>(m);
    b->ui = static_cast<double4*>(u);
    rc = batch_refresh(b, false);
  }
  (void)hipSetDevice(dev);
  if (rc) {
    for (void* p : b->allocations) (void)hipFree(p);
    delete b;
    return rc;
  }
  *out = b;
  return PSX_OK;
}

}  // namespace psx

extern "C" {

int32_t psx_batch_create(psx_plan* const* plans, int32_t n_plans, psx_batch** out) {
  psx::clear_error();
  return psx::batch_create(plans, n_plans, out);
}

int32_t psx_batch_destroy(psx_batch* batch) {
  psx::clear_error();
  if (!batch) return PSX_OK;
  int dev = 0;
  (void)hipGetDevice(&dev);
  (void)hipSetDevice(batch->device);
  for (void* p : batch->allocations) (void)hipFree(p);
  (void)hipSetDevice(dev);
  delete batch;
  return PSX_OK;
}

int32_t psx_vjp_batch(psx_batch* batch, const double* const* u, const double* const* mu, const double* const* lam,
                      const double* const* w, double* const* g_u, double* const* g_mu, double* const* g_lam,
                      int32_t precision, void* stream, int32_t* status) {
  psx::clear_error();
  if (!batch || !u || !mu || !lam || !w || !g_u || !status)
    return psx::fail(PSX_EINVAL, "null argument to psx_vjp_batch");
  if ((g_mu == nullptr) != (g_lam == nullptr))
    return psx::fail(PSX_EINVAL, "g_mu and g_lam must both be given or both be null");
  if (precision != PSX_FP64 && precision != PSX_FP32)
    return psx::fail(PSX_EINVAL, "precision must be PSX_FP64 or PSX_FP32");
  if (batch->d.npe != 4) return psx::fail(PSX_EUNSUPPORTED, "psx_vjp_batch supports batches of Tet4 plans");
  bool stale = !batch->wi;
  if (!batch->wi) {
    int dev = 0;
    (void)hipGetDevice(&dev);
    (void)hipSetDevice(batch->device);
    void* w_buf = nullptr;
    const int32_t rc = psx::batch_alloc(batch, size_t(batch->d.n_nodes) * sizeof(double4), &w_buf);
    (void)hipSetDevice(dev);
    if (rc) return rc;
    batch->wi = static_cast<double4*>(w_buf);
  }
  for (size_t i = 0; i < batch->plans.size(); ++i) stale = stale || batch->plans[i]->version != batch->versions[i];
  if (stale) {
    const int32_t rc = psx::batch_refresh(batch, true);
    if (rc) return rc;
  }
  return psx::launch_tet4_batch_vjp(batch->d, u, mu, lam, w, g_u, g_mu, g_lam, precision,
                                    static_cast<hipStream_t>(stream), status);
}

int32_t psx_batch_set_persistent(psx_batch* batch, int32_t persistent) {
  psx::clear_error();
  if (!batch) return psx::fail(PSX_EINVAL, "batch is null");
  batch->persistent = persistent != 0;
  return PSX_OK;
}

int32_t psx_batch_info(const psx_batch* batch, int64_t* info, int32_t n_info) {
  psx::clear_error();
  if (!batch || !info) return psx::fail(PSX_EINVAL, "null argument to psx_batch_info");
  int64_t nnz = 0;
  for (const psx_plan* p : batch->plans) nnz += p->hp.nnz();
  const int64_t v[] = {batch->d.n, batch->d.n_tiles, batch->d.n_nodes, batch->d.tile,
                       batch->persistent ? 1 : 0, nnz};
  const int32_t m = std::min<int32_t>(n_info, int32_t(sizeof(v) / sizeof(v[0])));
  for (int32_t i = 0; i < m; ++i) info[i] = v[i];
  return PSX_OK;
}

int32_t psx_assemble_batch(psx_batch* batch, const double* const* u, const double* const* mu,
                           const double* const* lam, double* const* f, double* const* kval, int32_t precision,
                           void* stream, int32_t* status) {
  psx::clear_error();
  if (!batch || !u || !f || !status || !mu != !lam)
    return psx::fail(PSX_EINVAL, "null argument to psx_assemble_batch");
  if (precision != PSX_FP64 && precision != PSX_FP32)
    return psx::fail(PSX_EINVAL, "precision must be PSX_FP64 or PSX_FP32");
  bool stale = false;
  for (size_t i = 0; i < batch->plans.size(); ++i) stale = stale || batch->plans[i]->version != batch->versions[i];
  if (stale) {
    const int32_t rc = psx::batch_refresh(batch, true);
    if (rc) return rc;
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (batch->d.npe == 10)
    return psx::launch_tet10_batch(batch->d, u, mu, lam, f, kval, precision, batch->cus, batch->persistent, s,
                                   status);
  if (!mu && !batch->all_bound)
    return psx::fail(PSX_EINVAL, "psx_assemble_batch without mu and lam needs psx_plan_bind_material on every plan");
  return psx::launch_tet4_batch(batch->d, u, mu, lam, f, kval, precision, batch->cus, batch->persistent, s, status);
}

}  // extern "C"
