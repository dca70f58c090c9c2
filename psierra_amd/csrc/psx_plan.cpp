// Host side of the engine: error state, structural pattern, plan construction.
//
// The structural pattern is the CSR that pSierra obtains from
// create_coo_sparsity_indices + coo_matrix(...).tocsr()
// (src/psierra/utility/assembly.py:6-34, 75-76).  It is built here from the
// node adjacency instead of sorting E * nde^2 triplets: node a's three rows
// hold, for every node b sharing an element with a (sorted by id), the three
// columns 3b, 3b+1, 3b+2.  Hence
//   indptr[3a + r] = 9 * adj_ptr[a] + 3 * r * deg(a),
//   value(3a + r, 3 b_s + c) = kofs(a) + 3 r deg(a) + 3 s + c,
// with b_s the s-th neighbour of a ("slot rank" s).
#include <algorithm>
#include <atomic>
#include <climits>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <thread>

#include "psx_internal.h"

namespace psx {

namespace {
thread_local std::string g_error;

unsigned worker_count() {
  unsigned n = std::thread::hardware_concurrency();
  if (n == 0) n = 1;
  return std::min(n, 16u);
}

template <typename Fn>
void parallel_for(int64_t n, Fn fn) {
  unsigned T = worker_count();
  if (n < 4096 || T == 1) {
    fn(int64_t(0), n);
    return;
  }
  std::vector<std::thread> th;
  int64_t chunk = (n + T - 1) / T;
  for (unsigned t = 0; t < T; ++t) {
    int64_t b = t * chunk, e = std::min<int64_t>(n, b + chunk);
    if (b >= e) break;
    th.emplace_back([=] { fn(b, e); });
  }
  for (auto& x : th) x.join();
}

std::string fmt_e(double v) {
  char buf[64];
  std::snprintf(buf, sizeof buf, "%.4e", v);
  return buf;
}
}  // namespace

void set_error(const std::string& msg) { g_error = msg; }
void clear_error() { g_error.clear(); }
int32_t fail(int32_t code, const std::string& msg) {
  g_error = msg;
  return code;
}

int32_t build_host_pattern(const int64_t* conn, int64_t E, int npe, int64_t N, HostPattern& hp) {
  if (!conn && E > 0) return fail(PSX_EINVAL, "conn is null");
  if (npe != 4 && npe != 10) return fail(PSX_EINVAL, "npe must be 4 or 10");
  if (E < 0 || N <= 0) return fail(PSX_EINVAL, "need n_elements >= 0 and n_nodes > 0");
  if (E >= (int64_t(1) << kIncElemBits))
    return fail(PSX_EUNSUPPORTED, "more than 2^28 elements");
  if (N * 3 >= (int64_t(1) << 31) || E * npe >= (int64_t(1) << 31))
    return fail(PSX_EUNSUPPORTED, "mesh too large for 32-bit indices");
  for (int64_t i = 0; i < E * npe; ++i)
    if (conn[i] < 0 || conn[i] >= N)
      return fail(PSX_EINVAL, "connectivity entry " + std::to_string(i) + " = " +
                                  std::to_string(conn[i]) + " outside [0, n_nodes)");
  hp.E = E;
  hp.N = N;
  hp.npe = npe;
  // node -> (element, local) incidence, in element order
  hp.n2e_ptr.assign(N + 1, 0);
  for (int64_t i = 0; i < E * npe; ++i) hp.n2e_ptr[conn[i] + 1]++;
  for (int64_t a = 0; a < N; ++a) hp.n2e_ptr[a + 1] += hp.n2e_ptr[a];
  hp.n2e.resize(E * npe);
  {
    std::vector<int32_t> fillp(hp.n2e_ptr.begin(), hp.n2e_ptr.end() - 1);
    for (int64_t e = 0; e < E; ++e)
      for (int p = 0; p < npe; ++p) hp.n2e[fillp[conn[e * npe + p]]++] = int32_t(e * npe + p);
  }
  // adjacency: sorted unique nodes of the incident elements
  std::vector<int32_t> deg(N);
  auto collect = [&](int64_t a, std::vector<int32_t>& buf) {
    buf.clear();
    for (int32_t k = hp.n2e_ptr[a]; k < hp.n2e_ptr[a + 1]; ++k) {
      int64_t e = hp.n2e[k] / npe;
      for (int q = 0; q < npe; ++q) buf.push_back(int32_t(conn[e * npe + q]));
    }
    std::sort(buf.begin(), buf.end());
    buf.erase(std::unique(buf.begin(), buf.end()), buf.end());
  };
  parallel_for(N, [&](int64_t b, int64_t e) {
    std::vector<int32_t> buf;
    for (int64_t a = b; a < e; ++a) {
      collect(a, buf);
      deg[a] = int32_t(buf.size());
    }
  });
  hp.adj_ptr.assign(N + 1, 0);
  for (int64_t a = 0; a < N; ++a) hp.adj_ptr[a + 1] = hp.adj_ptr[a] + deg[a];
  if (int64_t(hp.adj_ptr[N]) * 9 >= (int64_t(1) << 31))
    return fail(PSX_EUNSUPPORTED, "nnz >= 2^31 is not supported (int32 CSR)");
  hp.adj.resize(hp.adj_ptr[N]);
  parallel_for(N, [&](int64_t b, int64_t e) {
    std::vector<int32_t> buf;
    for (int64_t a = b; a < e; ++a) {
      collect(a, buf);
      std::copy(buf.begin(), buf.end(), hp.adj.begin() + hp.adj_ptr[a]);
    }
  });
  return PSX_OK;
}

void fill_csr(const HostPattern& hp, int32_t* indptr, int32_t* indices) {
  const int64_t N = hp.N;
  parallel_for(N, [&](int64_t b, int64_t e) {
    for (int64_t a = b; a < e; ++a) {
      const int32_t d = hp.adj_ptr[a + 1] - hp.adj_ptr[a];
      const int64_t base = 9 * int64_t(hp.adj_ptr[a]);
      for (int r = 0; r < 3; ++r) {
        indptr[3 * a + r] = int32_t(base + 3 * r * d);
        int32_t* out = indices + base + 3 * r * d;
        for (int s = 0; s < d; ++s)
          for (int c = 0; c < 3; ++c) out[3 * s + c] = 3 * hp.adj[hp.adj_ptr[a] + s] + c;
      }
    }
  });
  indptr[3 * N] = int32_t(hp.nnz());
}

namespace {

// Morton (Z-order) key of a point quantised to 21 bits per axis.
uint64_t spread3(uint64_t v) {
  v &= 0x1fffff;
  v = (v | v << 32) & 0x1f00000000ffffULL;
  v = (v | v << 16) & 0x1f0000ff0000ffULL;
  v = (v | v << 8) & 0x100f00f00f00f00fULL;
  v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
  v = (v | v << 2) & 0x1249249249249249ULL;
  return v;
}

std::vector<int32_t> locality_order(const double* X, int64_t N) {
  double lo[3] = {X[0], X[1], X[2]}, hi[3] = {X[0], X[1], X[2]};
  for (int64_t i = 0; i < N; ++i)
    for (int c = 0; c < 3; ++c) {
      lo[c] = std::min(lo[c], X[3 * i + c]);
      hi[c] = std::max(hi[c], X[3 * i + c]);
    }
  double ext = std::max({hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2], 1e-300});
  std::vector<std::pair<uint64_t, int32_t>> key(N);
  parallel_for(N, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; ++i) {
      uint64_t k = 0;
      for (int c = 0; c < 3; ++c) {
        double t = (X[3 * i + c] - lo[c]) / ext;
        uint64_t q = uint64_t(std::min(std::max(t, 0.0), 1.0) * double((1 << 21) - 1));
        k |= spread3(q) << c;
      }
      key[i] = {k, int32_t(i)};
    }
  });
  std::sort(key.begin(), key.end());
  std::vector<int32_t> perm(N);
  for (int64_t i = 0; i < N; ++i) perm[i] = key[i].second;
  return perm;
}

double det3(const double a[3][3]) {
  return a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) -
         a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
         a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
}

// Tet10 reference shape-function gradients (src/psierra/element/tet10_base.py:117-179).
void tet10_dN(double xi, double eta, double zeta, double dN[10][3]) {
  double L1 = xi, L2 = eta, L3 = zeta, L0 = 1.0 - L1 - L2 - L3;
  std::memset(dN, 0, sizeof(double) * 30);
  for (int c = 0; c < 3; ++c) dN[0][c] = -(4.0 * L0 - 1.0);
  dN[1][0] = 4.0 * L1 - 1.0;
  dN[2][1] = 4.0 * L2 - 1.0;
  dN[3][2] = 4.0 * L3 - 1.0;
  dN[4][0] = 4.0 * (L0 - L1); dN[4][1] = -4.0 * L1; dN[4][2] = -4.0 * L1;
  dN[5][0] = 4.0 * L2; dN[5][1] = 4.0 * L1;
  dN[6][0] = -4.0 * L2; dN[6][1] = 4.0 * (L0 - L2); dN[6][2] = -4.0 * L2;
  dN[7][0] = -4.0 * L3; dN[7][1] = -4.0 * L3; dN[7][2] = 4.0 * (L0 - L3);
  dN[8][0] = 4.0 * L3; dN[8][2] = 4.0 * L1;
  dN[9][1] = 4.0 * L3; dN[9][2] = 4.0 * L2;
}

int32_t check_volumes(const int64_t* conn, int64_t E, int npe, const double* X) {
  if (npe == 4) {
    for (int64_t e = 0; e < E; ++e) {
      const int64_t* c = conn + 4 * e;
      double a[3][3];
      for (int b = 0; b < 3; ++b)
        for (int r = 0; r < 3; ++r) a[b][r] = X[3 * c[b + 1] + r] - X[3 * c[0] + r];
      double d = det3(a);
      if (!(d > 0.0))
        return fail(PSX_EVOLUME, "Negative or zero element volume detected (det(A) = " + fmt_e(d) +
                                     ") at element " + std::to_string(e) +
                                     ". Check element vertex ordering.");
    }
    return PSX_OK;
  }
  const double s5 = std::sqrt(5.0);
  const double qa = (5.0 - s5) / 20.0, qb = (5.0 + 3.0 * s5) / 20.0;
  const double qp[4][3] = {{qa, qa, qa}, {qb, qa, qa}, {qa, qb, qa}, {qa, qa, qb}};
  double dN[4][10][3];
  for (int q = 0; q < 4; ++q) tet10_dN(qp[q][0], qp[q][1], qp[q][2], dN[q]);
  for (int64_t e = 0; e < E; ++e) {
    const int64_t* c = conn + 10 * e;
    for (int q = 0; q < 4; ++q) {
      double J[3][3] = {};
      for (int a = 0; a < 10; ++a)
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) J[i][j] += X[3 * c[a] + i] * dN[q][a][j];
      double d = det3(J);
      if (!(d > 0.0))
        return fail(PSX_EVOLUME, "Negative or zero Jacobian determinant: det(J) = " + fmt_e(d) +
                                     " at element " + std::to_string(e) + ".");
    }
  }
  return PSX_OK;
}

// Reference mass tensor c[p][q] = (1/V) * integral over the element of
// N_p N_q, from the exact simplex integral of barycentric monomials,
// (1/V) int L0^a L1^b L2^c L3^d = 6 a! b! c! d! / (a + b + c + d + 3)!.
// Tet4: N_p = L_p.  Tet10 (VTK order): corners L_p (2 L_p - 1), mid-edge
// nodes 4 L_i L_j on edges 01, 12, 20, 03, 13, 23.
std::vector<double> reference_mass_tensor(int npe) {
  using Mono = std::array<int, 4>;
  using Poly = std::vector<std::pair<Mono, double>>;
  auto L = [](int i, int pw) {
    Mono m{0, 0, 0, 0};
    m[i] = pw;
    return m;
  };
  std::vector<Poly> N;
  for (int i = 0; i < 4; ++i) {
    if (npe == 4)
      N.push_back({{L(i, 1), 1.0}});
    else
      N.push_back({{L(i, 2), 2.0}, {L(i, 1), -1.0}});
  }
  if (npe == 10) {
    const int edges[6][2] = {{0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 3}, {2, 3}};
    for (const auto& ed : edges) {
      Mono m{0, 0, 0, 0};
      m[ed[0]] = 1;
      m[ed[1]] = 1;
      N.push_back({{m, 4.0}});
    }
  }
  auto fact = [](int n) {
    double f = 1.0;
    for (int i = 2; i <= n; ++i) f *= i;
    return f;
  };
  std::vector<double> c(size_t(npe) * npe, 0.0);
  for (int a = 0; a < npe; ++a)
    for (int b = 0; b < npe; ++b) {
      double sum = 0.0;
      for (const auto& ta : N[a])
        for (const auto& tb : N[b]) {
          int tot = 0;
          double num = 6.0;
          for (int i = 0; i < 4; ++i) {
            const int e = ta.first[i] + tb.first[i];
            tot += e;
            num *= fact(e);
          }
          sum += ta.second * tb.second * num / fact(tot + 3);
        }
      c[size_t(a) * npe + b] = sum;
    }
  return c;
}

template <typename T>
int32_t upload(psx_plan* plan, const T* host, size_t count, T** dev) {
  size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
  void* p = nullptr;
  PSX_HIP_TRY(hipMalloc(&p, bytes));
  plan->allocations.push_back(p);
  plan->device_bytes += int64_t(bytes);
  if (count && host) PSX_HIP_TRY(hipMemcpy(p, host, count * sizeof(T), hipMemcpyHostToDevice));
  *dev = static_cast<T*>(p);
  return PSX_OK;
}

// Tet4 row-kernel records of one tiling (tet4_rows): an incidence stages
// its first off-diagonal block (the other node of lowest local index) in
// area A and its other two in area B; each area is unit-major, a (node,
// slot) unit's feeds contiguous and in run order (increasing element id).
// Per incidence: the A position and the two B positions (10 bits each).
// Per unit: its first A and first B positions and its tile-local node; a
// unit's feed counts are the differences to the next unit's (the tile's last
// unit ends at nk in A and 2 nk in B).  A unit sums its A feeds, then its B
// feeds: an order fixed by the mesh alone.
int32_t make_tet4_positions(psx_plan* plan, const std::vector<int32_t>& tiles) {
  const int64_t N = plan->hp.N;
  std::vector<int32_t> uoff(N + 1, 0);
  for (int64_t i = 0; i < N; ++i) uoff[i + 1] = uoff[i] + plan->deg_h[i];
  std::vector<uint32_t> unit_rec(size_t(std::max<int64_t>(uoff[N], 1)), 0u);
  std::vector<uint32_t> inc_pos(size_t(std::max<int64_t>(plan->n_inc, 1)), 0u);
  const int64_t n_tiles = int64_t(tiles.size()) - 1;
  std::atomic<int> bad{0};
  parallel_for(n_tiles, [&](int64_t tb, int64_t te) {
    std::vector<int32_t> fa, fb, ca, cb;
    for (int64_t t = tb; t < te; ++t) {
      const int64_t n0 = tiles[t], n1 = tiles[t + 1];
      int32_t pa = 0, pb = 0;
      for (int64_t i = n0; i < n1; ++i) {
        const int32_t d = plan->deg_h[i];
        fa.assign(size_t(d), 0);
        fb.assign(size_t(d), 0);
        ca.assign(size_t(d), 0);
        cb.assign(size_t(d), 0);
        const int64_t kb = plan->inc_ptr_h[i], ke = plan->inc_ptr_h[i + 1];
        for (int64_t k = kb; k < ke; ++k) {
          ++ca[plan->inc_slot_h[3 * k]];
          ++cb[plan->inc_slot_h[3 * k + 1]];
          ++cb[plan->inc_slot_h[3 * k + 2]];
        }
        for (int32_t sl = 0; sl < d; ++sl) {
          fa[sl] = pa;
          fb[sl] = pb;
          pa += ca[sl];
          pb += cb[sl];
          if (pa > int32_t(kPosMask) + 1 || pb > int32_t(kPosMask) + 1 || i - n0 > 511) bad = 1;
          unit_rec[uoff[i] + sl] = uint32_t(fa[sl]) | (uint32_t(fb[sl]) << 10) | (uint32_t(i - n0) << 20);
          ca[sl] = cb[sl] = 0;
        }
        for (int64_t k = kb; k < ke; ++k) {
          const int s0 = plan->inc_slot_h[3 * k], s1 = plan->inc_slot_h[3 * k + 1], s2 = plan->inc_slot_h[3 * k + 2];
          const uint32_t a0 = uint32_t(fa[s0] + ca[s0]++);
          const uint32_t b1 = uint32_t(fb[s1] + cb[s1]++);
          const uint32_t b2 = uint32_t(fb[s2] + cb[s2]++);
          inc_pos[k] = a0 | (b1 << kPosBits) | (b2 << (2 * kPosBits));
        }
      }
    }
  });
  if (bad) return fail(PSX_EUNSUPPORTED, "a tile's Tet4 staging positions exceed their 10-bit fields");
  uint32_t* dpos = nullptr;
  uint32_t* drec = nullptr;
  int32_t rc = upload(plan, inc_pos.data(), inc_pos.size(), &dpos);
  if (rc) return rc;
  if ((rc = upload(plan, unit_rec.data(), unit_rec.size(), &drec))) return rc;
  plan->d.inc_pos = dpos;
  plan->d.unit_rec = drec;
  return PSX_OK;
}

// Contribution lists of a wide Tet10 plan's general path: for every CSR
// value the element-tangent entries (e 900 + 30 row + column) that sum into
// it, and for every f entry its element-force entries (e 30 + 3 p + c), each
// in increasing element order; plus the element scratch.
int32_t make_general_maps(psx_plan* plan, const int64_t* conn, int64_t E) {
  const int npe = plan->hp.npe, nd = 3 * npe;
  const int64_t N = plan->hp.N, nnz = plan->hp.nnz();
  if (E * nd * nd >= (int64_t(1) << 31))
    return fail(PSX_EUNSUPPORTED, "the general path of a wide Tet10 plan takes at most " +
                                      std::to_string(((int64_t(1) << 31) - 1) / (nd * nd)) + " elements");
  const auto& ap = plan->hp.adj_ptr;
  const auto& adj = plan->hp.adj;
  auto pos_of = [&](int64_t a, int i, int64_t b, int c) -> int64_t {
    const int32_t d = ap[a + 1] - ap[a];
    const int32_t s = int32_t(std::lower_bound(adj.begin() + ap[a], adj.begin() + ap[a + 1], int32_t(b)) -
                              (adj.begin() + ap[a]));
    return 9 * int64_t(ap[a]) + 3 * int64_t(i) * d + 3 * s + c;
  };
  std::vector<int32_t> kptr(nnz + 1, 0), fptr(3 * N + 1, 0);
  for (int64_t e = 0; e < E; ++e)
    for (int p = 0; p < npe; ++p) {
      const int64_t a = conn[e * npe + p];
      for (int i = 0; i < 3; ++i) {
        ++fptr[3 * a + i + 1];
        for (int q = 0; q < npe; ++q)
          for (int c = 0; c < 3; ++c) ++kptr[pos_of(a, i, conn[e * npe + q], c) + 1];
      }
    }
  for (int64_t v = 0; v < nnz; ++v) kptr[v + 1] += kptr[v];
  for (int64_t v = 0; v < 3 * N; ++v) fptr[v + 1] += fptr[v];
  std::vector<int32_t> kidx(size_t(std::max<int64_t>(E * nd * nd, 1))), fidx(size_t(std::max<int64_t>(E * nd, 1)));
  std::vector<int32_t> kfill(kptr.begin(), kptr.end() - 1), ffill(fptr.begin(), fptr.end() - 1);
  for (int64_t e = 0; e < E; ++e)
    for (int p = 0; p < npe; ++p) {
      const int64_t a = conn[e * npe + p];
      for (int i = 0; i < 3; ++i) {
        fidx[ffill[3 * a + i]++] = int32_t(e * nd + 3 * p + i);
        for (int q = 0; q < npe; ++q)
          for (int c = 0; c < 3; ++c)
            kidx[kfill[pos_of(a, i, conn[e * npe + q], c)]++] = int32_t(e * nd * nd + nd * (3 * p + i) + 3 * q + c);
      }
    }
  int32_t *dkp, *dki, *dfp, *dfi;
  double *dke, *dfe;
  int32_t rc;
  if ((rc = upload(plan, kptr.data(), kptr.size(), &dkp)) || (rc = upload(plan, kidx.data(), kidx.size(), &dki)) ||
      (rc = upload(plan, fptr.data(), fptr.size(), &dfp)) || (rc = upload(plan, fidx.data(), fidx.size(), &dfi)) ||
      (rc = upload<double>(plan, nullptr, size_t(std::max<int64_t>(E * nd * nd, 1)), &dke)) ||
      (rc = upload<double>(plan, nullptr, size_t(std::max<int64_t>(E * nd, 1)), &dfe)))
    return rc;
  plan->d.gen_nnz = nnz;
  plan->d.gen_kptr = dkp;
  plan->d.gen_kidx = dki;
  plan->d.gen_fptr = dfp;
  plan->d.gen_fidx = dfi;
  plan->d.gen_ke = dke;
  plan->d.gen_fe = dfe;
  return PSX_OK;
}

// Halo of each Tet4 tile (a plan that is not wide): the nodes of the tile's
// elements outside the tile, ascending, and per incidence the tile-local
// index (the tile's own nodes 0 .. nodes-1, then its halo) of the element's
// four nodes in the row's order -- the row node, then the others by local
// index -- one byte each.  A tile with more than kTileLdsNodes own + halo
// nodes gets halo count -1 (its rows read the device copy).
int32_t make_tile_halos(psx_plan* plan, const std::vector<int32_t>& tiles) {
  const HostPattern& hp = plan->hp;
  const int64_t N = hp.N, E = hp.E;
  std::vector<int32_t> inv(static_cast<size_t>(N));
  for (int64_t i = 0; i < N; ++i) inv[plan->perm_h[i]] = int32_t(i);
  std::vector<int32_t> conn_i(static_cast<size_t>(E) * 4);
  for (int64_t a = 0; a < N; ++a)
    for (int32_t t = hp.n2e_ptr[a]; t < hp.n2e_ptr[a + 1]; ++t) conn_i[hp.n2e[t]] = inv[a];
  const int64_t T = int64_t(tiles.size()) - 1;
  std::vector<std::vector<int32_t>> halos(static_cast<size_t>(T));
  std::vector<uint32_t> loc(size_t(std::max<int64_t>(plan->n_inc, 1)), 0u);
  std::vector<int2> th(static_cast<size_t>(std::max<int64_t>(T, 1)));
  parallel_for(T, [&](int64_t tb, int64_t te) {
    for (int64_t t = tb; t < te; ++t) {
      const int32_t n0 = tiles[t], n1 = tiles[t + 1];
      std::vector<int32_t>& h = halos[t];
      for (int32_t i = n0; i < n1; ++i) {
        const int32_t a = plan->perm_h[i];
        for (int32_t r = hp.n2e_ptr[a]; r < hp.n2e_ptr[a + 1]; ++r) {
          const int64_t el = hp.n2e[r] / 4;
          for (int q = 0; q < 4; ++q) {
            const int32_t n = conn_i[4 * el + q];
            if (n < n0 || n >= n1) h.push_back(n);
          }
        }
      }
      std::sort(h.begin(), h.end());
      h.erase(std::unique(h.begin(), h.end()), h.end());
      if (int64_t(n1 - n0) + int64_t(h.size()) > kTileLdsNodes) {
        h.clear();
        th[t] = make_int2(0, -1);
        continue;
      }
      th[t] = make_int2(0, int32_t(h.size()));
      auto local = [&](int32_t n) -> uint32_t {
        if (n >= n0 && n < n1) return uint32_t(n - n0);
        return uint32_t((n1 - n0) + (std::lower_bound(h.begin(), h.end(), n) - h.begin()));
      };
      for (int32_t i = n0; i < n1; ++i) {
        const int32_t a = plan->perm_h[i];
        int64_t k = plan->inc_ptr_h[i];
        for (int32_t r = hp.n2e_ptr[a]; r < hp.n2e_ptr[a + 1]; ++r, ++k) {
          const int64_t el = hp.n2e[r] / 4;
          const int p = hp.n2e[r] % 4;
          uint32_t w = local(conn_i[4 * el + p]);
          int q = 1;
          for (int b = 0; b < 4; ++b) {
            if (b == p) continue;
            w |= local(conn_i[4 * el + b]) << (8 * q);
            ++q;
          }
          loc[k] = w;
        }
      }
    }
  });
  int64_t nh = 0;
  for (int64_t t = 0; t < T; ++t) {
    th[t].x = int32_t(nh);
    nh += int64_t(halos[t].size());
  }
  if (nh > INT32_MAX) return fail(PSX_EUNSUPPORTED, "tile halos exceed 32-bit offsets");
  std::vector<int32_t> hall(size_t(std::max<int64_t>(nh, 1)));
  for (int64_t t = 0; t < T; ++t) std::copy(halos[t].begin(), halos[t].end(), hall.begin() + th[t].x);
  int2* dth = nullptr;
  int32_t* dh = nullptr;
  uint32_t* dloc = nullptr;
  int32_t rc;
  if ((rc = upload(plan, th.data(), th.size(), &dth)) || (rc = upload(plan, hall.data(), hall.size(), &dh)) ||
      (rc = upload(plan, loc.data(), loc.size(), &dloc)))
    return rc;
  plan->d.tile_halo = dth;
  plan->d.halo = dh;
  plan->d.inc_loc = dloc;
  return PSX_OK;
}

int32_t make_tiles(psx_plan* plan, int32_t tile) {
  if (tile == 0) tile = plan->hp.npe == 4 ? 256 : 64;
  if (tile != 64 && tile != 128 && tile != 256)
    return fail(PSX_EINVAL, "tile must be 64, 128 or 256");
  const int64_t N = plan->hp.N;
  const bool wide = plan->d.wide != 0;
  if (wide && plan->hp.npe != 4) {
    // a wide Tet10 plan assembles through its general path: no tiles
    plan->d.n_tiles = 0;
    plan->d.tile = tile;
    plan->d.max_inc = plan->max_run;
    return PSX_OK;
  }
  if (plan->max_run > tile)
    return fail(PSX_EUNSUPPORTED, "node with " + std::to_string(plan->max_run) +
                                      " incident elements exceeds the tile of " + std::to_string(tile));
  // Greedy packing of consecutive nodes: at most `tile` incidences and at
  // most `cap` reduction lanes (the diagonal lanes of each node plus one
  // per (node, slot) unit) per tile.  Tet10 tiles also hold at most 64
  // nodes (their node lanes are one wave).  A wide plan packs by incidences
  // only (its Tet4 row kernel loops its reduction lanes).
  std::vector<int32_t> tiles{0};
  int64_t start_inc = 0, lanes = 0;
  const int64_t max_nodes = plan->hp.npe == 4 ? std::min<int64_t>(tile, 255) : 64;
  for (int64_t i = 0; i < N; ++i) {
    const int64_t need = (plan->hp.npe == 4 ? kTet4DiagLanes : kTet10DiagLanes) + plan->deg_h[i];
    const int64_t cap = plan->hp.npe == 4 ? tile : int64_t(kTet10LanePasses) * tile;
    if (!wide && need > cap) return fail(PSX_EUNSUPPORTED, "node degree exceeds the tile's reduction lanes");
    if (plan->inc_ptr_h[i + 1] - start_inc > tile || (!wide && lanes + need > cap) ||
        i - tiles.back() >= max_nodes) {
      tiles.push_back(int32_t(i));
      start_inc = plan->inc_ptr_h[i];
      lanes = 0;
    }
    lanes += need;
  }
  tiles.push_back(int32_t(N));
  if (tiles.size() >= 2 && tiles[tiles.size() - 2] == tiles.back()) tiles.pop_back();
  // per tile {first node, nodes, first incidence, incidences}
  std::vector<int4> info(tiles.size() - 1);
  for (size_t t = 0; t + 1 < tiles.size(); ++t) {
    const int32_t k0 = plan->inc_ptr_h[tiles[t]];
    info[t] = make_int4(tiles[t], tiles[t + 1] - tiles[t], k0, plan->inc_ptr_h[tiles[t + 1]] - k0);
  }
  int32_t* dtiles = nullptr;
  int4* dinfo = nullptr;
  int32_t rc = upload(plan, tiles.data(), tiles.size(), &dtiles);
  if (rc) return rc;
  if ((rc = upload(plan, info.data(), info.size(), &dinfo))) return rc;
  if (plan->hp.npe == 4 && wide && (rc = make_tet4_positions(plan, tiles))) return rc;
  if (plan->hp.npe == 4 && !wide && (rc = make_tile_halos(plan, tiles))) return rc;
  plan->d.tile_node = dtiles;
  plan->d.tile_info = dinfo;
  plan->d.n_tiles = int32_t(tiles.size() - 1);
  plan->d.tile = tile;
  plan->d.max_inc = plan->max_run;
  return PSX_OK;
}

}  // namespace

int32_t create_plan(const int64_t* conn, int64_t E, int npe, const double* X, int64_t N, int32_t tile,
                    psx_plan** out) {
  if (!out) return fail(PSX_EINVAL, "out is null");
  *out = nullptr;
  if (!X) return fail(PSX_EINVAL, "X is null");
  if (E <= 0) return fail(PSX_EINVAL, "plan needs at least one element");
  auto* plan = new psx_plan();
  auto bail = [&](int32_t rc) {
    for (void* p : plan->allocations) (void)hipFree(p);
    delete plan;
    return rc;
  };
  int32_t rc = build_host_pattern(conn, E, npe, N, plan->hp);
  if (rc) return bail(rc);
  rc = check_volumes(conn, E, npe, X);
  if (rc) return bail(rc);
  if (hipGetDevice(&plan->device) != hipSuccess)
    return bail(fail(PSX_ENOTREADY, "no HIP device available"));

  const HostPattern& hp = plan->hp;
  // locality order of the nodes (internal id i -> user id perm[i])
  plan->perm_h = locality_order(X, N);
  std::vector<int32_t> inv(N);
  for (int64_t i = 0; i < N; ++i) inv[plan->perm_h[i]] = int32_t(i);

  // internal connectivity (AoS, 4 ints for Tet4, 12 for Tet10)
  const int cs = npe == 4 ? 4 : 12;
  std::vector<int32_t> conn_i(size_t(E) * cs, 0);
  parallel_for(E, [&](int64_t b, int64_t e) {
    for (int64_t k = b; k < e; ++k)
      for (int p = 0; p < npe; ++p) conn_i[k * cs + p] = inv[conn[k * npe + p]];
  });
  // Tet4: per-element shape-function gradients and volume, computed in FP64
  // and stored rounded to FP32 for the FP32 element path: one float4 per
  // local node, (h_b, V), so a row can load its four gradients in any order.
  std::vector<float4> geo(npe == 4 ? size_t(E) * 4 : 0);
  if (npe == 4) {
    parallel_for(E, [&](int64_t b, int64_t e) {
      for (int64_t k = b; k < e; ++k) {
        const int64_t* c = conn + 4 * k;
        double D[3][3];
        for (int q = 0; q < 3; ++q)
          for (int r = 0; r < 3; ++r) D[r][q] = X[3 * c[q + 1] + r] - X[3 * c[0] + r];
        const double a00 = D[1][1] * D[2][2] - D[1][2] * D[2][1], a01 = D[0][2] * D[2][1] - D[0][1] * D[2][2],
                     a02 = D[0][1] * D[1][2] - D[0][2] * D[1][1], a10 = D[1][2] * D[2][0] - D[1][0] * D[2][2],
                     a11 = D[0][0] * D[2][2] - D[0][2] * D[2][0], a12 = D[0][2] * D[1][0] - D[0][0] * D[1][2],
                     a20 = D[1][0] * D[2][1] - D[1][1] * D[2][0], a21 = D[0][1] * D[2][0] - D[0][0] * D[2][1],
                     a22 = D[0][0] * D[1][1] - D[0][1] * D[1][0];
        const double det = D[0][0] * a00 + D[0][1] * a10 + D[0][2] * a20;
        const double id = 1.0 / det;
        const double h[3][3] = {{a00 * id, a01 * id, a02 * id}, {a10 * id, a11 * id, a12 * id},
                                {a20 * id, a21 * id, a22 * id}};
        const float V = float(det / 6.0);
        geo[4 * k] = make_float4(float(-(h[0][0] + h[1][0] + h[2][0])), float(-(h[0][1] + h[1][1] + h[2][1])),
                                 float(-(h[0][2] + h[1][2] + h[2][2])), V);
        for (int b = 0; b < 3; ++b)
          geo[4 * k + 1 + b] = make_float4(float(h[b][0]), float(h[b][1]), float(h[b][2]), V);
      }
    });
  }
  std::vector<double4> Xi(N);
  for (int64_t i = 0; i < N; ++i) {
    int64_t a = plan->perm_h[i];
    Xi[i] = make_double4(X[3 * a], X[3 * a + 1], X[3 * a + 2], 0.0);
  }

  // incidences in internal node order, with slot ranks
  const int rw = npe == 4 ? 1 : 3;
  plan->inc_ptr_h.assign(N + 1, 0);
  plan->deg_h.assign(N, 0);
  std::vector<int32_t> kofs(N), uoff(N + 1, 0);
  for (int64_t i = 0; i < N; ++i) {
    int64_t a = plan->perm_h[i];
    plan->inc_ptr_h[i + 1] = plan->inc_ptr_h[i] + (hp.n2e_ptr[a + 1] - hp.n2e_ptr[a]);
    int32_t d = hp.adj_ptr[a + 1] - hp.adj_ptr[a];
    plan->deg_h[i] = d;
    plan->max_deg = std::max(plan->max_deg, d);
    kofs[i] = 9 * hp.adj_ptr[a];
    uoff[i + 1] = uoff[i] + d;
  }
  for (int64_t i = 0; i < N; ++i)
    plan->max_run = std::max(plan->max_run, plan->inc_ptr_h[i + 1] - plan->inc_ptr_h[i]);
  if (plan->max_deg > 65535)
    return bail(fail(PSX_EUNSUPPORTED, "node degree " + std::to_string(plan->max_deg) +
                                           " above 65535 is not supported"));
  // beyond the bounds of the feed masks and slot-rank bytes: Tet4 row kernels only
  plan->d.wide = plan->max_deg > kMaxRankDegree || plan->max_run > kMaxNodeIncidences ? 1 : 0;
  plan->n_inc = plan->inc_ptr_h[N];
  if (npe == 4 && plan->d.wide) plan->inc_slot_h.assign(size_t(plan->n_inc) * 3, 0);
  std::vector<uint32_t> inc_elem(plan->n_inc), inc_rank(plan->n_inc * rw, 0);
  parallel_for(N, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; ++i) {
      int64_t a = plan->perm_h[i];
      const int32_t* nb = hp.adj.data() + hp.adj_ptr[a];
      const int32_t d = hp.adj_ptr[a + 1] - hp.adj_ptr[a];
      int64_t k = plan->inc_ptr_h[i];
      for (int32_t t = hp.n2e_ptr[a]; t < hp.n2e_ptr[a + 1]; ++t, ++k) {
        int64_t el = hp.n2e[t] / npe;
        int p = hp.n2e[t] % npe;
        inc_elem[k] = uint32_t(el) | (uint32_t(p) << kIncElemBits);
        auto rank_of = [&](int64_t node) -> uint32_t {
          return uint32_t(std::lower_bound(nb, nb + d, int32_t(node)) - nb);
        };
        if (npe == 4) {
          // Tet4 incidences are evaluated with the row node first, then the
          // other three in increasing local order (see tet4 kernel).  Byte 0
          // is the incidence's position in its node's run (< 64); bytes 1..3
          // are the slot ranks of the other three nodes in a's row (bytes
          // only meaningful for a plan that is not wide).
          uint32_t r = uint32_t(k - plan->inc_ptr_h[i]) & 0xffu;
          int slot = 1;
          for (int q = 0; q < 4; ++q) {
            if (q == p) continue;
            const uint32_t rk = rank_of(conn[el * 4 + q]);
            if (plan->d.wide) plan->inc_slot_h[3 * k + slot - 1] = uint16_t(rk);
            r |= (rk & 0xffu) << (8 * slot);
            ++slot;
          }
          inc_rank[k] = r;
        } else {
          // bytes 0..9: slot ranks of the element's nodes in a's row (byte p
          // is a's own); byte 10: the incidence's position in a's run
          for (int q = 0; q < 10; ++q)
            inc_rank[k * rw + q / 4] |= (rank_of(conn[el * 10 + q]) & 0xffu) << (8 * (q % 4));
          inc_rank[k * rw + 2] |= uint32_t(k - plan->inc_ptr_h[i]) << 16;
        }
      }
    }
  });

  // Node records: {first incidence, first unit, first CSR value, slot rank
  // of the node in its own row}; record N closes the last node.
  std::vector<int4> node_rec;
  {
    node_rec.resize(N + 1);
    parallel_for(N, [&](int64_t b, int64_t e) {
      for (int64_t i = b; i < e; ++i) {
        const int64_t a = plan->perm_h[i];
        const int32_t* nb = hp.adj.data() + hp.adj_ptr[a];
        const int32_t d = hp.adj_ptr[a + 1] - hp.adj_ptr[a];
        const int32_t self = int32_t(std::lower_bound(nb, nb + d, int32_t(a)) - nb);
        node_rec[i] = make_int4(plan->inc_ptr_h[i], uoff[i], kofs[i], self);
      }
    });
    node_rec[N] = make_int4(plan->inc_ptr_h[N], uoff[N], 0, 0);
  }

  psx::DevPlan& d = plan->d;
  d.E = int32_t(E);
  d.N = int32_t(N);
  d.npe = npe;
  d.conn_stride = cs;
  d.rank_words = rw;
  int32_t *dconn, *dperm, *dincp, *dkofs, *ddeg, *duoff;
  uint32_t *dince, *dincr;
  double4* dX;
  float4* dgeo = nullptr;
  int4* dnrec = nullptr;
  if ((npe == 4 && (rc = upload(plan, geo.data(), geo.size(), &dgeo))) ||
      (rc = upload(plan, node_rec.data(), node_rec.size(), &dnrec)))
    return bail(rc);
  d.geo = dgeo;
  d.node_rec = dnrec;
  if ((rc = upload(plan, conn_i.data(), conn_i.size(), &dconn)) ||
      (rc = upload(plan, Xi.data(), Xi.size(), &dX)) ||
      (rc = upload(plan, plan->perm_h.data(), plan->perm_h.size(), &dperm)) ||
      (rc = upload(plan, plan->inc_ptr_h.data(), plan->inc_ptr_h.size(), &dincp)) ||
      (rc = upload(plan, inc_elem.data(), inc_elem.size(), &dince)) ||
      (rc = upload(plan, inc_rank.data(), inc_rank.size(), &dincr)) ||
      (rc = upload(plan, kofs.data(), kofs.size(), &dkofs)) ||
      (rc = upload(plan, plan->deg_h.data(), plan->deg_h.size(), &ddeg)) ||
      (rc = upload(plan, uoff.data(), uoff.size(), &duoff)))
    return bail(rc);
  d.conn = dconn;
  d.Xi = dX;
  d.perm = dperm;
  d.inc_ptr = dincp;
  d.inc_elem = dince;
  d.inc_rank = dincr;
  d.kofs = dkofs;
  d.deg = ddeg;
  d.uoff = duoff;
  void* ui = nullptr;
  if (hipMalloc(&ui, size_t(N) * sizeof(double4)) != hipSuccess)
    return bail(fail(PSX_EINVAL, "hipMalloc of the displacement scratch failed"));
  plan->allocations.push_back(ui);
  plan->device_bytes += int64_t(N) * int64_t(sizeof(double4));
  d.ui = static_cast<double4*>(ui);
  void* wi = nullptr;
  if (hipMalloc(&wi, size_t(N) * sizeof(double4)) != hipSuccess)
    return bail(fail(PSX_EINVAL, "hipMalloc of the cotangent scratch failed"));
  plan->allocations.push_back(wi);
  plan->device_bytes += int64_t(N) * int64_t(sizeof(double4));
  d.wi = static_cast<double4*>(wi);
  if ((rc = make_tiles(plan, tile))) return bail(rc);
  if (npe == 10 && plan->d.wide && (rc = make_general_maps(plan, conn, E))) return bail(rc);
  *out = plan;
  return PSX_OK;
}

int32_t create_csc_adj(const std::vector<int32_t>& adj_ptr, const std::vector<int32_t>& adj,
                       const std::vector<int32_t>& order, int64_t N, const int64_t* free_dofs, int64_t n_free,
                       psx_csc** out) {
  if (!out) return fail(PSX_EINVAL, "out is null");
  *out = nullptr;
  struct {
    const std::vector<int32_t>& adj_ptr;
    const std::vector<int32_t>& adj;
  } hp{adj_ptr, adj};
  const int64_t n = 3 * N;
  if (n_free < 0 || n_free > n || (n_free > 0 && !free_dofs))
    return fail(PSX_EINVAL, "bad free dof list");
  std::vector<int32_t> newidx(n, -1), freev(n_free);
  for (int64_t k = 0; k < n_free; ++k) {
    const int64_t j = free_dofs[k];
    if (j < 0 || j >= n) return fail(PSX_EINVAL, "free dof " + std::to_string(j) + " out of range");
    if (k > 0 && j <= free_dofs[k - 1])
      return fail(PSX_EINVAL, "free dofs must be strictly increasing");
    newidx[j] = int32_t(k);
    freev[k] = int32_t(j);
  }
  auto deg = [&](int64_t a) { return hp.adj_ptr[a + 1] - hp.adj_ptr[a]; };
  auto rank = [&](int64_t a, int64_t b) {
    const int32_t* nb = hp.adj.data() + hp.adj_ptr[a];
    return int32_t(std::lower_bound(nb, nb + deg(a), int32_t(b)) - nb);
  };
  // Each node's free mask and free base (free dofs below 3n).
  std::vector<int32_t> fmask(static_cast<size_t>(N)), fbase(static_cast<size_t>(N) + 1, 0);
  for (int64_t a = 0; a < N; ++a) {
    int32_t m = 0;
    for (int r = 0; r < 3; ++r) m |= int32_t(newidx[3 * a + r] >= 0) << r;
    fmask[a] = m;
    fbase[a + 1] = fbase[a] + __builtin_popcount(unsigned(m));
  }
  // Column node b's unit (b, a) for each neighbour a: where block K[3a.., 3b..]
  // starts in the CSR values, a's row stride, a's free base and mask.  The
  // structural count of K[free][:, free] is kept for the output sizes.
  const int64_t n_units = int64_t(hp.adj.size());
  std::vector<int4> unit(size_t(std::max<int64_t>(n_units, 1)));
  int64_t nnz_struct = 0;
  for (int64_t b = 0; b < N; ++b) {
    const int64_t free_cols = __builtin_popcount(unsigned(fmask[b]));
    for (int32_t t = hp.adj_ptr[b]; t < hp.adj_ptr[b + 1]; ++t) {
      const int64_t a = hp.adj[t];
      unit[t] = make_int4(int32_t(9 * int64_t(hp.adj_ptr[a]) + 3 * rank(a, b)), int32_t(3 * deg(a)), fbase[a],
                          fmask[a]);
      nnz_struct += free_cols * __builtin_popcount(unsigned(fmask[a]));
    }
  }
  if (nnz_struct >= (int64_t(1) << 31)) return fail(PSX_EINVAL, "free-dof matrix has 2^31 or more entries");
  // Structural column starts: free columns in dof order, each with the free
  // rows of its node's neighbours.
  std::vector<int32_t> sstart(static_cast<size_t>(n));
  {
    int64_t pos = 0;
    for (int64_t b = 0; b < N; ++b) {
      int64_t rows = 0;
      for (int32_t t = hp.adj_ptr[b]; t < hp.adj_ptr[b + 1]; ++t) rows += __builtin_popcount(unsigned(fmask[hp.adj[t]]));
      for (int c = 0; c < 3; ++c) {
        if (newidx[3 * b + c] < 0) continue;
        sstart[3 * b + c] = int32_t(pos);
        pos += rows;
      }
    }
  }
  if (int64_t(order.size()) != N) return fail(PSX_EINVAL, "node order has the wrong length");
  // Column node records in locality order, nodes with no free dof left out.
  std::vector<int4> node;
  node.reserve(2 * size_t(N));
  for (int64_t i = 0; i < N; ++i) {
    const int32_t b = order[i];
    if (b < 0 || b >= N) return fail(PSX_EINVAL, "node order holds an id out of range");
    if (!fmask[b]) continue;
    node.push_back(make_int4(hp.adj_ptr[b], hp.adj_ptr[b + 1] - hp.adj_ptr[b], fbase[b], fmask[b]));
    node.push_back(make_int4(sstart[3 * b], sstart[3 * b + 1], sstart[3 * b + 2], 0));
  }
  auto* csc = new psx_csc();
  auto bail = [&](int32_t rc) {
    for (void* p : csc->allocations) (void)hipFree(p);
    delete csc;
    return rc;
  };
  (void)hipGetDevice(&csc->device);
  csc->n_free = n_free;
  csc->nnz_struct = nnz_struct;
  auto up = [&](const void* h, size_t bytes, void** d) -> int32_t {
    PSX_HIP_TRY(hipMalloc(d, std::max<size_t>(bytes, 16)));
    csc->allocations.push_back(*d);
    if (bytes && h) PSX_HIP_TRY(hipMemcpy(*d, h, bytes, hipMemcpyHostToDevice));
    return PSX_OK;
  };
  void *dnode, *dunit, *dfree, *dscr, *dpart;
  const size_t nb = size_t((n_free + 255) / 256) + 1;
  int32_t rc;
  if ((rc = up(node.data(), node.size() * sizeof(int4), &dnode)) ||
      (rc = up(unit.data(), size_t(n_units) * sizeof(int4), &dunit)) ||
      (rc = up(freev.data(), freev.size() * 4, &dfree)) || (rc = up(nullptr, nb * 4, &dscr)) ||
      (rc = up(nullptr, nb * 8, &dpart)))
    return bail(rc);
  csc->d.n_free = int32_t(n_free);
  csc->d.n_nodes = int32_t(node.size() / 2);
  csc->d.node = static_cast<int4*>(dnode);
  csc->d.unit = static_cast<int4*>(dunit);
  csc->d.free = static_cast<int32_t*>(dfree);
  csc->d.scratch = static_cast<int32_t*>(dscr);
  csc->d.partial = static_cast<double*>(dpart);
  csc->d.nnz_struct = int32_t(nnz_struct);
  *out = csc;
  return PSX_OK;
}

int32_t create_csc(const psx_plan* plan, const int64_t* free_dofs, int64_t n_free, psx_csc** out) {
  return create_csc_adj(plan->hp.adj_ptr, plan->hp.adj, plan->perm_h, plan->hp.N, free_dofs, n_free, out);
}

int32_t retile(psx_plan* plan, int32_t tile) {
  // The old tile array stays allocated until the plan is destroyed.
  return make_tiles(plan, tile);
}

}  // namespace psx

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

int32_t psx_abi_version(void) { return 7; }

const char* psx_last_error(void) { return psx::g_error.c_str(); }

int32_t psx_pattern_nnz(const int64_t* conn, int64_t n_elements, int32_t npe, int64_t n_nodes,
                        int64_t* nnz) {
  psx::clear_error();
  if (!nnz) return psx::fail(PSX_EINVAL, "nnz is null");
  psx::HostPattern hp;
  int32_t rc = psx::build_host_pattern(conn, n_elements, npe, n_nodes, hp);
  if (rc) return rc;
  *nnz = hp.nnz();
  return PSX_OK;
}

int32_t psx_pattern_build(const int64_t* conn, int64_t n_elements, int32_t npe, int64_t n_nodes,
                          int32_t* indptr, int32_t* indices) {
  psx::clear_error();
  if (!indptr || !indices) return psx::fail(PSX_EINVAL, "indptr/indices is null");
  psx::HostPattern hp;
  int32_t rc = psx::build_host_pattern(conn, n_elements, npe, n_nodes, hp);
  if (rc) return rc;
  psx::fill_csr(hp, indptr, indices);
  return PSX_OK;
}

int32_t psx_plan_create(const int64_t* conn, int64_t n_elements, int32_t npe, const double* X,
                        int64_t n_nodes, int32_t tile, psx_plan** out) {
  psx::clear_error();
  return psx::create_plan(conn, n_elements, npe, X, n_nodes, tile, out);
}

int32_t psx_plan_destroy(psx_plan* plan) {
  psx::clear_error();
  if (!plan) return PSX_OK;
  int dev = 0;
  (void)hipGetDevice(&dev);
  (void)hipSetDevice(plan->device);
  for (void* p : plan->allocations) (void)hipFree(p);
  (void)hipSetDevice(dev);
  delete plan;
  return PSX_OK;
}

int32_t psx_plan_info(const psx_plan* plan, int64_t* info, int32_t n_info) {
  psx::clear_error();
  if (!plan || !info) return psx::fail(PSX_EINVAL, "plan/info is null");
  int64_t v[PSX_INFO_COUNT] = {plan->hp.E,     plan->hp.N,      plan->hp.npe,
                               plan->hp.nnz(), plan->d.n_tiles, plan->d.tile,
                               plan->max_deg,  plan->device_bytes, plan->n_inc,
                               plan->max_run,  plan->d.wide};
  for (int32_t i = 0; i < n_info && i < PSX_INFO_COUNT; ++i) info[i] = v[i];
  return PSX_OK;
}

int32_t psx_plan_pattern(const psx_plan* plan, int32_t* indptr, int32_t* indices) {
  psx::clear_error();
  if (!plan || !indptr || !indices) return psx::fail(PSX_EINVAL, "null argument");
  psx::fill_csr(plan->hp, indptr, indices);
  return PSX_OK;
}

int32_t psx_plan_set_tile(psx_plan* plan, int32_t tile) {
  psx::clear_error();
  if (!plan) return psx::fail(PSX_EINVAL, "plan is null");
  const int32_t rc = psx::retile(plan, tile);
  ++plan->version;
  return rc;
}

int32_t psx_assemble(psx_plan* plan, const double* u, const double* mu, const double* lam, double* f,
                     double* kval, int32_t precision, void* stream, int32_t* status) {
  psx::clear_error();
  if (!plan || !u || !f || !status || !mu != !lam)
    return psx::fail(PSX_EINVAL, "null argument to psx_assemble");
  if (precision != PSX_FP64 && precision != PSX_FP32)
    return psx::fail(PSX_EINVAL, "precision must be PSX_FP64 or PSX_FP32");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (!mu) {
    if (!plan->bound)
      return psx::fail(PSX_EINVAL, "psx_assemble without mu and lam needs psx_plan_bind_material first");
    return psx::launch_tet4_assemble_bound(plan->d, u, plan->mu_bound, plan->lam_bound, f, kval, precision, s,
                                           status);
  }
  if (plan->hp.npe == 4)
    return psx::launch_tet4_assemble(plan->d, u, mu, lam, f, kval, precision, s, status);
  return psx::launch_tet10_assemble(plan->d, u, mu, lam, f, kval, precision, s, status);
}

int32_t psx_assemble_hyperelastic(psx_plan* plan, int32_t model, const double* const* params, const double* u,
                                  double* f, double* kval, int32_t precision, void* stream, int32_t* status) {
  psx::clear_error();
  if (!plan || !params || !u || !f || !status)
    return psx::fail(PSX_EINVAL, "null argument to psx_assemble_hyperelastic");
  if (precision != PSX_FP64 && precision != PSX_FP32)
    return psx::fail(PSX_EINVAL, "precision must be PSX_FP64 or PSX_FP32");
  const int n = model == PSX_NEO_HOOKEAN ? 2 : model == PSX_MOONEY_RIVLIN ? 3 : model == PSX_YEOH ? 4 : -1;
  if (n < 0) return psx::fail(PSX_EINVAL, "unknown material model");
  for (int i = 0; i < n; ++i)
    if (!params[i]) return psx::fail(PSX_EINVAL, "null material parameter array");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const double* prm[4] = {params[0], params[1], n > 2 ? params[2] : nullptr, n > 3 ? params[3] : nullptr};
  if (plan->hp.npe != 4) return psx::launch_tet10_hyper(plan->d, model, prm, u, f, kval, precision, s, status);
  return psx::launch_tet4_hyper(plan->d, model, prm, u, f, kval, precision, s, status);
}

int32_t psx_plan_bind_material(psx_plan* plan, const double* mu, const double* lam, void* stream) {
  psx::clear_error();
  if (!plan || !mu || !lam) return psx::fail(PSX_EINVAL, "null argument to psx_plan_bind_material");
  if (plan->hp.npe != 4) return psx::fail(PSX_EUNSUPPORTED, "psx_plan_bind_material binds Tet4 plans");
  const size_t E = size_t(std::max<int64_t>(plan->d.E, 1));
  if (!plan->d.geo_bound) {
    void* g = nullptr;
    void* m = nullptr;
    void* l = nullptr;
    PSX_HIP_TRY(hipMalloc(&g, 4 * E * sizeof(float4)));
    plan->allocations.push_back(g);
    PSX_HIP_TRY(hipMalloc(&m, E * sizeof(double)));
    plan->allocations.push_back(m);
    PSX_HIP_TRY(hipMalloc(&l, E * sizeof(double)));
    plan->allocations.push_back(l);
    plan->device_bytes += int64_t(E * (4 * sizeof(float4) + 2 * sizeof(double)));
    plan->d.geo_bound = static_cast<const float4*>(g);
    plan->mu_bound = static_cast<double*>(m);
    plan->lam_bound = static_cast<double*>(l);
  }
  plan->bound = false;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (plan->d.E > 0) {
    PSX_HIP_TRY(hipMemcpyAsync(plan->mu_bound, mu, size_t(plan->d.E) * sizeof(double), hipMemcpyDeviceToDevice, s));
    PSX_HIP_TRY(hipMemcpyAsync(plan->lam_bound, lam, size_t(plan->d.E) * sizeof(double), hipMemcpyDeviceToDevice, s));
  }
  const int32_t rc = psx::launch_bind_material(plan->d, mu, lam, const_cast<float4*>(plan->d.geo_bound), s);
  if (rc) return rc;
  plan->bound = true;
  ++plan->version;
  return PSX_OK;
}

int32_t psx_assemble_mass(psx_plan* plan, const double* rho, double* kval, void* stream) {
  psx::clear_error();
  if (!plan || !rho || !kval) return psx::fail(PSX_EINVAL, "null argument to psx_assemble_mass");
  if (plan->d.wide)
    return psx::fail(PSX_EUNSUPPORTED, "psx_assemble_mass needs every node to have at most 64 incident elements "
                                       "and 255 neighbours");
  if (!plan->mass_coef) {
    const std::vector<double> c = psx::reference_mass_tensor(plan->hp.npe);
    double* d = nullptr;
    int32_t rc = psx::upload(plan, c.data(), c.size(), &d);
    if (rc) return rc;
    plan->mass_coef = d;
  }
  if (!plan->elem_vol) {
    double* v = nullptr;
    int32_t rc = psx::upload<double>(plan, nullptr, size_t(plan->hp.E), &v);
    if (rc) return rc;
    rc = psx::launch_element_volumes(plan->d, v, static_cast<hipStream_t>(stream));
    if (rc) return rc;
    plan->elem_vol = v;
  }
  return psx::launch_mass(plan->d, rho, plan->elem_vol, plan->mass_coef, kval, static_cast<hipStream_t>(stream),
                          plan->hp.adj_ptr.back());
}

int32_t psx_vjp(psx_plan* plan, const double* u, const double* mu, const double* lam, const double* w,
                double* g_u, double* g_mu, double* g_lam, int32_t precision, void* stream,
                int32_t* status) {
  psx::clear_error();
  if (!plan || !u || !mu || !lam || !w || !g_u || !status)
    return psx::fail(PSX_EINVAL, "null argument to psx_vjp");
  if ((g_mu == nullptr) != (g_lam == nullptr))
    return psx::fail(PSX_EINVAL, "g_mu and g_lam must both be given or both be null");
  if (precision != PSX_FP64 && precision != PSX_FP32)
    return psx::fail(PSX_EINVAL, "precision must be PSX_FP64 or PSX_FP32");
  if (plan->hp.npe != 4) return psx::fail(PSX_EUNSUPPORTED, "psx_vjp supports Tet4 plans");
  return psx::launch_tet4_vjp(plan->d, u, mu, lam, w, g_u, g_mu, g_lam, precision,
                              static_cast<hipStream_t>(stream), status);
}

int32_t psx_ecsw_training(psx_plan* plan, const double* phi, int32_t rank, const double* snapshots,
                          int32_t n_snapshots, const double* mu, const double* lam, int32_t include_stiffness,
                          double* A, double* work, int64_t work_doubles, double* sumsq, void* stream,
                          int32_t* status) {
  psx::clear_error();
  if (!plan || !phi || !snapshots || !mu || !lam || !A || !status)
    return psx::fail(PSX_EINVAL, "null argument to psx_ecsw_training");
  if (plan->hp.npe != 4) return psx::fail(PSX_EUNSUPPORTED, "psx_ecsw_training supports Tet4 plans");
  if (rank < 1 || rank > 64) return psx::fail(PSX_EUNSUPPORTED, "psx_ecsw_training supports ranks 1..64");
  if (n_snapshots < 1) return psx::fail(PSX_EINVAL, "need at least one snapshot");
  if (sumsq) {
    const int64_t need = psx_ecsw_training_work_doubles(plan);
    if (!work || work_doubles < need)
      return psx::fail(PSX_EINVAL, "psx_ecsw_training with sumsq needs a work array of " + std::to_string(need) +
                                       " doubles");
  }
  return psx::launch_ecsw_training(plan->d, phi, rank, snapshots, n_snapshots, mu, lam, include_stiffness != 0,
                                   A, work, sumsq, static_cast<hipStream_t>(stream), status);
}

int32_t psx_csc_create(const psx_plan* plan, const int64_t* free_dofs, int64_t n_free, psx_csc** out) {
  psx::clear_error();
  if (!plan) return psx::fail(PSX_EINVAL, "plan is null");
  return psx::create_csc(plan, free_dofs, n_free, out);
}

int32_t psx_csc_destroy(psx_csc* csc) {
  psx::clear_error();
  if (!csc) return PSX_OK;
  int dev = 0;
  (void)hipGetDevice(&dev);
  (void)hipSetDevice(csc->device);
  for (void* p : csc->allocations) (void)hipFree(p);
  (void)hipSetDevice(dev);
  delete csc;
  return PSX_OK;
}

int32_t psx_csc_info(const psx_csc* csc, int64_t* n_free, int64_t* nnz_struct) {
  psx::clear_error();
  if (!csc || !n_free || !nnz_struct) return psx::fail(PSX_EINVAL, "null argument");
  *n_free = csc->n_free;
  *nnz_struct = csc->nnz_struct;
  return PSX_OK;
}

int32_t psx_csc_extract(psx_csc* csc, const double* kval, int32_t drop_zeros, int32_t* colptr,
                        int32_t* rowind, double* vals, void* stream) {
  psx::clear_error();
  if (!csc || !kval || !colptr || !rowind || !vals) return psx::fail(PSX_EINVAL, "null argument");
  return psx::launch_csc_extract(csc->d, kval, drop_zeros ? 1 : 0, colptr, rowind, vals,
                                 static_cast<hipStream_t>(stream));
}

int32_t psx_free_residual(psx_csc* csc, const double* f_int, const double* f_ext_free, double* r_free,
                          double* sumsq, void* stream) {
  psx::clear_error();
  if (!csc || !f_int || !r_free || !sumsq) return psx::fail(PSX_EINVAL, "null argument");
  return psx::launch_free_residual(csc->d, f_int, f_ext_free, r_free, sumsq,
                                   static_cast<hipStream_t>(stream));
}

int32_t psx_element_batch(int32_t npe, int64_t n, const double* X, const double* u, const double* mu,
                          const double* lam, double* k, double* f, int32_t precision, void* stream,
                          int32_t* status) {
  psx::clear_error();
  if (npe != 4 && npe != 10) return psx::fail(PSX_EINVAL, "npe must be 4 or 10");
  if (n < 0 || n >= (int64_t(1) << 31)) return psx::fail(PSX_EINVAL, "bad element count");
  if (n == 0) return PSX_OK;
  if (!X || !u || !mu || !lam || !f || !status) return psx::fail(PSX_EINVAL, "null argument");
  if (precision != PSX_FP64 && precision != PSX_FP32)
    return psx::fail(PSX_EINVAL, "precision must be PSX_FP64 or PSX_FP32");
  return psx::launch_element_batch(npe, n, X, u, mu, lam, k, f, precision,
                                   static_cast<hipStream_t>(stream), status);
}

int32_t psx_nh_material_batch(int64_t n, const double* F, const double* mu, const double* bulk,
                              double* S, double* C, int32_t precision, void* stream,
                              int32_t* status) {
  psx::clear_error();
  if (n < 0 || n >= (int64_t(1) << 31)) return psx::fail(PSX_EINVAL, "bad count");
  if (n == 0) return PSX_OK;
  if (!F || !mu || !bulk || !S || !C || !status) return psx::fail(PSX_EINVAL, "null argument");
  if (precision != PSX_FP64 && precision != PSX_FP32)
    return psx::fail(PSX_EINVAL, "precision must be PSX_FP64 or PSX_FP32");
  return psx::launch_material_batch(n, F, mu, bulk, S, C, precision, static_cast<hipStream_t>(stream),
                                    status);
}

}  // extern "C"
