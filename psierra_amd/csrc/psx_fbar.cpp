// Host side of the F-bar Tet4 assembly: the distance-2 pattern of
// K_dev + K_vol and the device layout the F-bar kernels read.
//
// Tet4Fbar (src/psierra/element/tet4_fbar.py:86-109) adds, for every node a,
// the block K V0_a g_a g_a^T over the dofs of a's nodal patch (the nodes of
// the elements around a).  Two nodes are coupled when they share an element
// (K_dev) or a patch (K_vol), i.e. when they are at most two element edges
// apart; that is the row set built here, per node, in ascending user id.
#include <algorithm>
#include <thread>

#include "psx_internal.h"

namespace psx {

namespace {

template <typename Fn>
void parallel_nodes(int64_t n, Fn fn) {
  unsigned T = std::min(std::max(1u, std::thread::hardware_concurrency()), 16u);
  if (n < 4096 || T == 1) {
    fn(int64_t(0), n);
    return;
  }
  std::vector<std::thread> th;
  const int64_t chunk = (n + T - 1) / T;
  for (unsigned t = 0; t < T; ++t) {
    const int64_t b = t * chunk, e = std::min<int64_t>(n, b + chunk);
    if (b >= e) break;
    th.emplace_back([=] { fn(b, e); });
  }
  for (auto& x : th) x.join();
}

template <typename T>
int32_t to_device(psx_fbar* fb, const T* host, size_t count, T** dev) {
  const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
  void* p = nullptr;
  PSX_HIP_TRY(hipMalloc(&p, bytes));
  fb->allocations.push_back(p);
  fb->device_bytes += int64_t(bytes);
  if (host && count) PSX_HIP_TRY(hipMemcpy(p, host, count * sizeof(T), hipMemcpyHostToDevice));
  *dev = static_cast<T*>(p);
  return PSX_OK;
}

void free_all(psx_fbar* fb) {
  for (void* p : fb->allocations) (void)hipFree(p);
  fb->allocations.clear();
}

// Distance-2 node adjacency (user order): for each node, the sorted union of
// its element neighbours' element neighbours.
int32_t build_distance2(const HostPattern& hp, std::vector<int32_t>& adj2_ptr, std::vector<int32_t>& adj2,
                        std::vector<int32_t>& deg2) {
  const int64_t N = hp.N;
  deg2.assign(N, 0);
  auto collect = [&](int64_t a, std::vector<int32_t>& buf) {
    buf.clear();
    for (int32_t i = hp.adj_ptr[a]; i < hp.adj_ptr[a + 1]; ++i) {
      const int32_t b = hp.adj[i];
      buf.insert(buf.end(), hp.adj.begin() + hp.adj_ptr[b], hp.adj.begin() + hp.adj_ptr[b + 1]);
    }
    std::sort(buf.begin(), buf.end());
    buf.erase(std::unique(buf.begin(), buf.end()), buf.end());
  };
  parallel_nodes(N, [&](int64_t b, int64_t e) {
    std::vector<int32_t> buf;
    for (int64_t a = b; a < e; ++a) {
      collect(a, buf);
      deg2[a] = int32_t(buf.size());
    }
  });
  adj2_ptr.assign(N + 1, 0);
  for (int64_t a = 0; a < N; ++a) adj2_ptr[a + 1] = adj2_ptr[a] + deg2[a];
  if (9 * int64_t(adj2_ptr[N]) >= (int64_t(1) << 31))
    return fail(PSX_EUNSUPPORTED, "F-bar nnz >= 2^31 is not supported (int32 CSR)");
  adj2.resize(adj2_ptr[N]);
  parallel_nodes(N, [&](int64_t b, int64_t e) {
    std::vector<int32_t> buf;
    for (int64_t a = b; a < e; ++a) {
      collect(a, buf);
      std::copy(buf.begin(), buf.end(), adj2.begin() + adj2_ptr[a]);
    }
  });
  return PSX_OK;
}

void fill_csr2(int64_t N, const std::vector<int32_t>& adj2_ptr, const std::vector<int32_t>& adj2, int32_t* indptr,
               int32_t* indices) {
  parallel_nodes(N, [&](int64_t b, int64_t e) {
    for (int64_t a = b; a < e; ++a) {
      const int32_t d = adj2_ptr[a + 1] - adj2_ptr[a];
      const int64_t base = 9 * int64_t(adj2_ptr[a]);
      for (int r = 0; r < 3; ++r) {
        indptr[3 * a + r] = int32_t(base + 3 * r * d);
        int32_t* o = indices + base + 3 * r * d;
        for (int32_t s = 0; s < d; ++s)
          for (int c = 0; c < 3; ++c) o[3 * s + c] = 3 * adj2[adj2_ptr[a] + s] + c;
      }
    }
  });
  indptr[3 * N] = int32_t(9 * int64_t(adj2_ptr[N]));
}

int32_t build(psx_plan* plan, psx_fbar* fb) {
  const HostPattern& hp = plan->hp;
  const int64_t N = hp.N;
  std::vector<int32_t> deg2;
  int32_t rc0 = build_distance2(hp, fb->adj2_ptr, fb->adj2, deg2);
  if (rc0) return rc0;
  fb->nnz2 = 9 * int64_t(fb->adj2_ptr[N]);

  // internal-order layout
  const std::vector<int32_t>& perm = plan->perm_h;   // internal -> user
  std::vector<int32_t> inv(N);
  for (int64_t i = 0; i < N; ++i) inv[perm[i]] = int32_t(i);
  std::vector<int32_t> uoff(N + 1, 0);
  for (int64_t i = 0; i < N; ++i) {
    const int32_t a = perm[i];
    uoff[i + 1] = uoff[i] + (hp.adj_ptr[a + 1] - hp.adj_ptr[a]);
  }
  std::vector<int32_t> slot2(uoff[N]);
  std::vector<int4> rec2(N);
  parallel_nodes(N, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; ++i) {
      const int32_t a = perm[i];
      const int32_t* l1 = hp.adj.data() + hp.adj_ptr[a];
      const int32_t n1 = hp.adj_ptr[a + 1] - hp.adj_ptr[a];
      const int32_t* l2 = fb->adj2.data() + fb->adj2_ptr[a];
      int32_t self2 = 0;
      for (int32_t s = 0, t = 0; s < n1; ++s) {
        while (l2[t] < l1[s]) ++t;   // l1 is a subset of l2, both ascending
        slot2[uoff[i] + s] = t;
        if (l1[s] == a) self2 = t;
      }
      rec2[i] = make_int4(int32_t(9 * int64_t(fb->adj2_ptr[a])), deg2[a], self2, 0);
    }
  });
  int32_t max_deg2 = 0;
  for (int64_t a = 0; a < N; ++a) max_deg2 = std::max(max_deg2, deg2[a]);
  if (max_deg2 > 65535) return fail(PSX_EUNSUPPORTED, "F-bar row with more than 65535 node blocks");

  // K_vol[p, q] = K sum_a V0_a g_a[p] g_a[q]^T over the common element
  // neighbours a: per (p, a) the units of a (contiguous), the unit (a, p)
  // and, for each neighbour q of a, q's slot in p's distance-2 row.
  std::vector<int64_t> moff(uoff[N] + 1, 0);
  for (int64_t i = 0; i < N; ++i) {
    const int32_t a0 = perm[i];
    for (int32_t t1 = hp.adj_ptr[a0], s = 0; t1 < hp.adj_ptr[a0 + 1]; ++t1, ++s) {
      const int32_t a = hp.adj[t1];
      moff[uoff[i] + s + 1] = moff[uoff[i] + s] + (hp.adj_ptr[a + 1] - hp.adj_ptr[a]);
    }
  }
  const int64_t n_map = moff[uoff[N]];
  if (n_map >= (int64_t(1) << 31)) return fail(PSX_EUNSUPPORTED, "F-bar volumetric slot map exceeds 2^31 entries");
  std::vector<int4> kv_unit(uoff[N]);
  std::vector<uint16_t> kv_slot(n_map);
  parallel_nodes(N, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; ++i) {
      const int32_t a0 = perm[i];
      const int32_t* l2 = fb->adj2.data() + fb->adj2_ptr[a0];
      for (int32_t t1 = hp.adj_ptr[a0], s = 0; t1 < hp.adj_ptr[a0 + 1]; ++t1, ++s) {
        const int32_t a = hp.adj[t1];
        const int32_t* la = hp.adj.data() + hp.adj_ptr[a];
        const int32_t na = hp.adj_ptr[a + 1] - hp.adj_ptr[a];
        const int32_t ub = uoff[inv[a]];
        const int32_t ap = ub + int32_t(std::lower_bound(la, la + na, a0) - la);
        const int64_t m0 = moff[uoff[i] + s];
        kv_unit[uoff[i] + s] = make_int4(ub, na, ap, int32_t(m0));
        for (int32_t q = 0, t = 0; q < na; ++q) {
          while (l2[t] < la[q]) ++t;
          kv_slot[m0 + q] = uint16_t(t);
        }
      }
    }
  });

  // K_vol by slot: node p's terms (a, q) sorted by q's slot in p's row, a
  // ascending within a slot (a stable bucket pass over the (a, q) order), so
  // each slot sums its terms in the order the unit-step kernel uses; per
  // term the unit index of (p, a) and the rank of q among a's neighbours
  // (g_a[q] is g[first unit of a + rank]); per node the first term of each
  // slot.
  std::vector<int4> kv_node(N);
  std::vector<uint16_t> kv_term(n_map);
  std::vector<uint16_t> kv_soff(static_cast<size_t>(fb->adj2_ptr[N]) + static_cast<size_t>(N));
  parallel_nodes(N, [&](int64_t b, int64_t e) {
    std::vector<int32_t> cur;
    for (int64_t i = b; i < e; ++i) {
      const int32_t a0 = perm[i];
      const int32_t nu = uoff[i + 1] - uoff[i], d2 = deg2[a0];
      const int64_t P = moff[uoff[i]], F = moff[uoff[i + 1]] - P;
      const int32_t sbase = fb->adj2_ptr[a0] + a0;
      kv_node[i] = make_int4(int32_t(P), sbase, int32_t(F), nu);
      if (nu > kKvolSlotMaxUnits || F > kKvolSlotMaxTerms) continue;   // the plan keeps the unit-step kernel
      cur.assign(size_t(d2) + 1, 0);
      for (int64_t m = P; m < P + F; ++m) ++cur[size_t(kv_slot[m]) + 1];
      for (int32_t t = 0; t < d2; ++t) cur[size_t(t) + 1] += cur[size_t(t)];
      for (int32_t t = 0; t <= d2; ++t) kv_soff[size_t(sbase) + size_t(t)] = uint16_t(cur[size_t(t)]);
      for (int32_t su = 0; su < nu; ++su) {
        const int4 ku = kv_unit[uoff[i] + su];
        for (int32_t q = 0; q < ku.y; ++q) {
          const int64_t pos = P + cur[kv_slot[size_t(ku.w) + size_t(q)]]++;
          kv_term[size_t(pos)] = uint16_t(su | (q << 8));   // q < deg a <= kMaxRankDegree
        }
      }
    }
  });
  int32_t max_terms = 0, max_units = 0;
  for (int64_t i = 0; i < N; ++i) {
    max_terms = std::max(max_terms, kv_node[i].z);
    max_units = std::max(max_units, kv_node[i].w);
  }
  const bool slots = max_terms <= kKvolSlotMaxTerms && max_units <= kKvolSlotMaxUnits && max_deg2 <= kKvolSlotMaxSlots;

  FbarDev& d = fb->d;
  if (slots) {
    int4* pkn = nullptr;
    uint16_t* pkt = nullptr;
    uint16_t* pso = nullptr;
    int32_t rc1;
    if ((rc1 = to_device(fb, kv_node.data(), kv_node.size(), &pkn))) return rc1;
    if ((rc1 = to_device(fb, kv_term.data(), kv_term.size(), &pkt))) return rc1;
    if ((rc1 = to_device(fb, kv_soff.data(), kv_soff.size(), &pso))) return rc1;
    d.kv_slots = 1;
    d.kv_max_terms = max_terms;
    d.kv_max_units = max_units;
    d.kv_node = pkn;
    d.kv_term = pkt;
    d.kv_soff = pso;
  }
  int32_t* ps2 = nullptr;
  int4* pr2 = nullptr;
  int4* pku = nullptr;
  uint16_t* pks = nullptr;
  int32_t rc;
  if ((rc = to_device(fb, slot2.data(), slot2.size(), &ps2))) return rc;
  if ((rc = to_device(fb, rec2.data(), rec2.size(), &pr2))) return rc;
  if ((rc = to_device(fb, kv_unit.data(), kv_unit.size(), &pku))) return rc;
  if ((rc = to_device(fb, kv_slot.data(), kv_slot.size(), &pks))) return rc;
  d.slot2 = ps2;
  d.node_rec2 = pr2;
  d.kv_unit = pku;
  d.kv_slot = pks;
  d.max_deg2 = max_deg2;
  const size_t E = size_t(hp.E);
  if ((rc = to_device<double>(fb, nullptr, E, &d.qv))) return rc;
  if ((rc = to_device<double>(fb, nullptr, E, &d.vjm1))) return rc;
  if ((rc = to_device<double>(fb, nullptr, 12 * E, &d.wq))) return rc;
  if ((rc = to_device<double>(fb, nullptr, size_t(N), &d.jm1))) return rc;
  if ((rc = to_device<double4>(fb, nullptr, size_t(uoff[N]), &d.g))) return rc;
  return PSX_OK;
}

}  // namespace

}  // namespace psx

extern "C" {

int32_t psx_fbar_create(psx_plan* plan, psx_fbar** out) {
  psx::clear_error();
  if (!plan || !out) return psx::fail(PSX_EINVAL, "plan/out is null");
  if (plan->hp.npe != 4) return psx::fail(PSX_EUNSUPPORTED, "F-bar needs a Tet4 plan");
  if (plan->d.wide)
    return psx::fail(PSX_EUNSUPPORTED, "F-bar needs every node to have at most 64 incident elements and "
                                       "255 neighbours");
  psx_fbar* fb = new psx_fbar();
  fb->plan = plan;
  int dev = 0;
  (void)hipGetDevice(&dev);
  (void)hipSetDevice(plan->device);
  const int32_t rc = psx::build(plan, fb);
  (void)hipSetDevice(dev);
  if (rc) {
    psx::free_all(fb);
    delete fb;
    return rc;
  }
  *out = fb;
  return PSX_OK;
}

int32_t psx_fbar_destroy(psx_fbar* fbar) {
  psx::clear_error();
  if (!fbar) return PSX_OK;
  int dev = 0;
  (void)hipGetDevice(&dev);
  (void)hipSetDevice(fbar->plan->device);
  psx::free_all(fbar);
  (void)hipSetDevice(dev);
  delete fbar;
  return PSX_OK;
}

int32_t psx_fbar_nnz(const psx_fbar* fbar, int64_t* nnz) {
  psx::clear_error();
  if (!fbar || !nnz) return psx::fail(PSX_EINVAL, "null argument");
  *nnz = fbar->nnz2;
  return PSX_OK;
}

int32_t psx_fbar_pattern(const psx_fbar* fbar, int32_t* indptr, int32_t* indices) {
  psx::clear_error();
  if (!fbar || !indptr || !indices) return psx::fail(PSX_EINVAL, "null argument");
  psx::fill_csr2(fbar->plan->hp.N, fbar->adj2_ptr, fbar->adj2, indptr, indices);
  return PSX_OK;
}

int32_t psx_fbar_pattern_nnz(const int64_t* conn, int64_t n_elements, int64_t n_nodes, int64_t* nnz) {
  psx::clear_error();
  if (!nnz) return psx::fail(PSX_EINVAL, "nnz is null");
  psx::HostPattern hp;
  std::vector<int32_t> ptr, adj2, deg2;
  int32_t rc = psx::build_host_pattern(conn, n_elements, 4, n_nodes, hp);
  if (rc || (rc = psx::build_distance2(hp, ptr, adj2, deg2))) return rc;
  *nnz = 9 * int64_t(ptr[n_nodes]);
  return PSX_OK;
}

int32_t psx_fbar_pattern_build(const int64_t* conn, int64_t n_elements, int64_t n_nodes, int32_t* indptr,
                               int32_t* indices) {
  psx::clear_error();
  if (!indptr || !indices) return psx::fail(PSX_EINVAL, "indptr/indices is null");
  psx::HostPattern hp;
  std::vector<int32_t> ptr, adj2, deg2;
  int32_t rc = psx::build_host_pattern(conn, n_elements, 4, n_nodes, hp);
  if (rc || (rc = psx::build_distance2(hp, ptr, adj2, deg2))) return rc;
  psx::fill_csr2(n_nodes, ptr, adj2, indptr, indices);
  return PSX_OK;
}

int32_t psx_fbar_csc_create(const psx_fbar* fbar, const int64_t* free_dofs, int64_t n_free, psx_csc** out) {
  psx::clear_error();
  if (!fbar) return psx::fail(PSX_EINVAL, "fbar is null");
  return psx::create_csc_adj(fbar->adj2_ptr, fbar->adj2, fbar->plan->perm_h, fbar->plan->hp.N, free_dofs, n_free,
                             out);
}

int32_t psx_fbar_assemble(psx_fbar* fbar, const double* u, const double* mu, const double* lam, double bulk,
                          double* f, double* kval, int32_t precision, void* stream, int32_t* status) {
  psx::clear_error();
  if (!fbar || !u || !mu || !lam || !f || !status) return psx::fail(PSX_EINVAL, "null argument to psx_fbar_assemble");
  if (precision != PSX_FP64 && precision != PSX_FP32)
    return psx::fail(PSX_EINVAL, "precision must be PSX_FP64 or PSX_FP32");
  return psx::launch_tet4_fbar(fbar->plan->d, fbar->d, u, mu, lam, bulk, f, kval, precision,
                               static_cast<hipStream_t>(stream), status);
}

}  // extern "C"
