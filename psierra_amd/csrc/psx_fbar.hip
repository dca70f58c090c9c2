// Tet4 F-bar (nodal-averaged volume) assembly for gfx950.
//
// Restates Tet4Fbar's domain assembly (src/psierra/element/tet4_fbar.py:
// 226-459) as four stream-ordered steps after the displacement permutation:
//   1. fbar_element_kernel, one lane per element: J_e - 1 (from grad u, so
//      small strains stay accurate), V_e / 4, V_e (J_e - 1) / 4 and
//      w_e q_b = V_e J_e / 4 F h_b for the element's four nodes (q_b is the
//      node's column of div_e = B_0 + B_1 + B_2, the volumetric row of B);
//   2. fbar_node_kernel, a wave (or half a wave) per node a, its incidences staged in LDS:
//      V0_a = sum V_e / 4, J_node - 1 = sum V_e (J_e - 1) / 4 / V0_a and,
//      with K, the volumetric gradient g_a over a's element neighbours,
//      g_a[b] = sum_{e in a, b} w_e q_b / V0_a (one lane per neighbour b,
//      incidences in run order);
//   3. fbar_kvol_kernel, half a wave per node p: its distance-2 row of
//      K_vol[p, q] = K_bulk sum_a V0_a g_a[p] g_a[q]^T over the common
//      neighbours a (ascending; q's slots in p's row listed with the plan),
//      accumulated in LDS and written to every value of the row; p's unit
//      records are read in one round and each unit's feeds one unit ahead;
//   4. the Tet4 row kernel with the F-bar material point (J_bar from step 2,
//      deviatoric tangent and f), adding K_dev onto the K_vol values.
// Every sum runs in an order fixed by the mesh (wave sums in a fixed
// butterfly order); nothing uses atomics except the inversion flag.
#include <algorithm>
#include <climits>

#include "psx_device.h"
#include "psx_internal.h"
#include "psx_tet4_elem.h"

namespace psx {

__global__ __launch_bounds__(256) void fbar_element_kernel(DevPlan p, FbarDev fb, int32_t* status) {
  const int e = xcd_tile() * 256 + int(threadIdx.x);
  if (e >= p.E) return;
  const int4 c = *reinterpret_cast<const int4*>(p.conn + 4 * e);
  const double4 X[4] = {p.Xi[c.x], p.Xi[c.y], p.Xi[c.z], p.Xi[c.w]};
  const double4 U[4] = {p.ui[c.x], p.ui[c.y], p.ui[c.z], p.ui[c.w]};
  double h[4][3], V, A[3][3];
  tet4_geometry<double>(X, U, h, V, A);
  const double trA = A[0][0] + A[1][1] + A[2][2];
  double trA2 = 0.0;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) trA2 += A[i][j] * A[j][i];
  const double Jm1 = trA + 0.5 * (trA * trA - trA2) + det3(A);
  const double J = 1.0 + Jm1;
  if (J <= 0.0) {
    atomicMin(&status[1], e);
    status[0] = 1;
  }
  const double v4 = 0.25 * V;
  fb.qv[e] = v4;
  fb.vjm1[e] = v4 * Jm1;
  const double w = v4 * J;
#pragma unroll
  for (int b = 0; b < 4; ++b)
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const double q = (i == 0 ? 1.0 + A[0][0] : A[i][0]) * h[b][0] + (i == 1 ? 1.0 + A[1][1] : A[i][1]) * h[b][1] +
                       (i == 2 ? 1.0 + A[2][2] : A[i][2]) * h[b][2];
      fb.wq[12 * int64_t(e) + 3 * b + i] = w * q;
    }
}

// Local index (0..3) of the q-th (1..3) node other than local node lp, in
// local order (the order of the rank word's bytes 1..3).
__device__ __forceinline__ int other_local(int lp, int q) { return (q - 1) + ((q - 1) >= lp ? 1 : 0); }

// NPW nodes per wave, W = 64 / NPW lanes each (a node has at most W
// incidences: NPW = 2 when no node has more than 32, else 1): a node's
// incidences' element quantities are loaded in one round, one lane each, and
// staged in LDS; V0_a and J_node - 1 are sums over the node's lanes in a
// fixed butterfly order, and each lane of a's element neighbours b walks the
// staged incidences in run order for g_a[b].
template <int NPW>
__global__ __launch_bounds__(256) void fbar_node_kernel(DevPlan p, FbarDev fb, int want_g) {
  constexpr int W = 64 / NPW;
  __shared__ double s_wq[4 * NPW][W][12];
  __shared__ uint32_t s_rank[4 * NPW][W];
  __shared__ uint8_t s_lp[4 * NPW][W];
  const int w = int(threadIdx.x >> 6);
  const int sub = int(threadIdx.x & 63) / W, lane = int(threadIdx.x & 63) % W;
  const int g = w * NPW + sub;   // this node's LDS slot
  const int a = (xcd_tile() * 4 + w) * NPW + sub;
  if (a >= p.N) return;   // the lanes below only exchange values within a node's group
  const int4 rec = p.node_rec[a];
  const int4 rec1 = p.node_rec[a + 1];
  const int kb = rec.x, L = rec1.x - rec.x;
  double qv = 0.0, vj = 0.0;
  if (lane < L) {
    const uint32_t pe = p.inc_elem[kb + lane];
    const int64_t e = int64_t(pe & kIncElemMask);
    qv = fb.qv[e];
    vj = fb.vjm1[e];
    if (want_g) {
      s_rank[g][lane] = p.inc_rank[kb + lane];
      s_lp[g][lane] = uint8_t(pe >> kIncElemBits);
      const double2* src = reinterpret_cast<const double2*>(fb.wq + 12 * e);
#pragma unroll
      for (int i = 0; i < 6; ++i) {
        const double2 v = src[i];
        s_wq[g][lane][2 * i] = v.x;
        s_wq[g][lane][2 * i + 1] = v.y;
      }
    }
  }
  // sums over the node's W lanes (butterfly, the same bits in every lane)
  double sv = qv, sj = vj;
#pragma unroll
  for (int m = W / 2; m >= 1; m >>= 1) {
    sv += __shfl_xor(sv, m, 64);
    sj += __shfl_xor(sj, m, 64);
  }
  if (lane == 0) fb.jm1[a] = sv > 0.0 ? sj / sv : 0.0;
  if (!want_g) return;
  __builtin_amdgcn_wave_barrier();
  const int ub = rec.y, deg = rec1.y - ub, self = rec.w;
  const double iv = sv > 0.0 ? 1.0 / sv : 0.0;
  // four lanes per neighbour, each over a quarter of the run, combined as
  // (q0 + q1) + (q2 + q3)
  const int part = lane & 3;
  const int k0 = (part * L) >> 2, k1 = ((part + 1) * L) >> 2;
  for (int base = 0; base < deg; base += W / 4) {
    const int s = base + (lane >> 2);
    double g0 = 0.0, g1 = 0.0, g2 = 0.0;
    if (s < deg) {
      for (int k = k0; k < k1; ++k) {
        const uint32_t r = s_rank[g][k];
        const int lp = s_lp[g][k];
        int loc = -1;
        if (s == self)
          loc = lp;
        else
#pragma unroll
          for (int q = 1; q < 4; ++q)
            if (int((r >> (8 * q)) & 0xffu) == s) loc = other_local(lp, q);
        if (loc >= 0) {
          g0 += s_wq[g][k][3 * loc];
          g1 += s_wq[g][k][3 * loc + 1];
          g2 += s_wq[g][k][3 * loc + 2];
        }
      }
    }
    g0 = quad_sum(g0);
    g1 = quad_sum(g1);
    g2 = quad_sum(g2);
    if (part == 0 && s < deg) fb.g[ub + s] = make_double4(g0 * iv, g1 * iv, g2 * iv, sv);
  }
}

// Half a wave per node p: K_vol[p, q] = K sum_a V0_a g_a[p] g_a[q]^T for the
// whole distance-2 row of p, accumulated in LDS.  The common neighbours a
// are taken in ascending order (p's element-pattern units); for each, the
// lanes read a's units and their slots in p's row contiguously (one lane per
// neighbour q of a) and add the 3x3 term into q's slot -- distinct q, distinct
// slots, so each slot sums its terms in a-order.  The row is then written
// with consecutive lanes on consecutive values.
constexpr int kKvolBatch = 8;
__global__ __launch_bounds__(256) void fbar_kvol_kernel(DevPlan p, FbarDev fb, int waves, double bulk, double* kval) {
  extern __shared__ double s_acc[];
  // two nodes per wave, 32 lanes each (a node's neighbours rarely exceed 32)
  const int w = int(threadIdx.x >> 6);
  const int half = int(threadIdx.x >> 5) & 1;
  const int lane = threadIdx.x & 31;
  const int pn = 2 * (xcd_tile() * waves + w) + half;
  const bool live = pn < p.N;
  if (__ballot(live) == 0ull) return;   // the wave's lanes read each other's records below
  int4 rec2 = make_int4(0, 0, 0, 0);
  int u1b = 0, u1e = 0;
  if (live) {
    rec2 = fb.node_rec2[pn];
    u1b = p.uoff[pn];
    u1e = p.uoff[pn + 1];
  }
  const int deg2 = rec2.y, n9 = 9 * deg2;
  double* acc = s_acc + size_t(2 * w + half) * 9 * fb.max_deg2;
  for (int i = lane; i < n9; i += 32) acc[i] = 0.0;
  __builtin_amdgcn_wave_barrier();
  const int nu = u1e - u1b;   // p's element neighbours a (units (p, a))
  // both halves step through max(nu) units together
  const int nmax = __builtin_amdgcn_readfirstlane(max(__builtin_amdgcn_readlane(nu, 0), __builtin_amdgcn_readlane(nu, 32)));
  for (int cb = 0; cb < nmax; cb += 32) {
    // the chunk's unit records {first unit of a, deg a, unit (a, p), first
    // slot entry} and g_a[p], one lane each, read in one round
    int4 kl = make_int4(0, 0, 0, 0);
    double4 gl = make_double4(0.0, 0.0, 0.0, 0.0);
    if (cb + lane < nu) {
      kl = fb.kv_unit[u1b + cb + lane];
      gl = fb.g[kl.z];
    }
    const int cnt = min(32, nmax - cb);
    // V0_a g_a[p] formed by the unit's lane; unit i's fields and this product
    // reach the lanes of its half by a 32-wide shuffle from lane i (a select
    // between readlanes of lanes i and 32 + i compiles to a divergent-index
    // readlane, i.e. branches).  Past the half's units kl is zero: no feeds.
    const double wl0 = gl.w * gl.x, wl1 = gl.w * gl.y, wl2 = gl.w * gl.z;
    // feeds of units i0 .. i0 + kKvolBatch - 1 (lane j < deg a: g_a[q_j] and
    // q_j's slot in p's row), all loaded before the batch's first sum: the
    // loads are unconditional (lanes past deg a, and units past the half's
    // last, read the unit's first feed or feed 0, in bounds), so the
    // compiler issues them back to back and a node waits on the memory once
    // per batch instead of once per unit
    for (int i0 = 0; i0 < cnt; i0 += kKvolBatch) {
      int deg_a[kKvolBatch], sl[kKvolBatch];
      double gx[kKvolBatch], gy[kKvolBatch], gz[kKvolBatch];
#pragma unroll
      for (int t = 0; t < kKvolBatch; ++t) {
        const int i = min(i0 + t, 31);
        const int kx = __shfl(kl.x, i, 32), kw = __shfl(kl.w, i, 32);
        const int ky = i0 + t < cnt ? __shfl(kl.y, i, 32) : 0;
        deg_a[t] = ky;
        const int j = min(lane, max(ky - 1, 0));
        const double4 g = fb.g[kx + j];
        gx[t] = g.x;
        gy[t] = g.y;
        gz[t] = g.z;
        sl[t] = fb.kv_slot[kw + j];
      }
#pragma unroll
      for (int t = 0; t < kKvolBatch; ++t) {
        if (i0 + t >= cnt) break;
        const int i = i0 + t;
        const double wp[3] = {__shfl(wl0, i, 32), __shfl(wl1, i, 32), __shfl(wl2, i, 32)};
        for (int j = lane; j < deg_a[t]; j += 32) {
          double q3[3] = {gx[t], gy[t], gz[t]};
          int s = sl[t];
          if (j != lane) {   // a neighbour with more than 32 element neighbours
            const int kx = __shfl(kl.x, i, 32), kw = __shfl(kl.w, i, 32);
            const double4 g = fb.g[kx + j];
            q3[0] = g.x;
            q3[1] = g.y;
            q3[2] = g.z;
            s = fb.kv_slot[kw + j];
          }
          double* a9 = acc + 9 * s;
#pragma unroll
          for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) a9[3 * r + c] += wp[r] * q3[c];
        }
        __builtin_amdgcn_wave_barrier();
      }
    }
  }
  // CSR order of p's rows: value r * 3 deg2 + 3 s + c
  const int row = 3 * deg2;
#pragma unroll
  for (int r = 0; r < 3; ++r)
    for (int k = lane; k < row; k += 32) {
      const int s2 = k / 3, c = k - 3 * s2;
      kval[rec2.x + r * row + k] = bulk * acc[9 * s2 + 3 * r + c];
    }
}

// One wave per node p, K_vol by slot: the node's terms (a, q) are staged in
// LDS in slot order (g_a[q] and the unit index of (p, a)), with V0_a g_a[p]
// per unit; then lane s sums the terms of slot s, a ascending -- the fused
// multiply-adds of fbar_kvol_kernel in the same order from 0.0, so the same
// bits -- in registers and writes the slot's nine values.  No slot is
// updated by more than one lane, so the sums need no read-modify-write of
// LDS.  Loads: the node's records, then its unit, term and slot records,
// then the g values they index.  The row's 9 deg2 values go through LDS
// (over the terms, once every lane is done with them) so that consecutive
// lanes store consecutive values.
__global__ __launch_bounds__(256) void fbar_kvol_slots_kernel(DevPlan p, FbarDev fb, int waves, int wave_bytes,
                                                              double bulk, double* kval) {
  extern __shared__ double s_dyn[];
  const int w = __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6));
  const int lane = int(threadIdx.x & 63);
  const int pn = xcd_tile() * waves + w;
  if (pn >= p.N) return;
  char* base = reinterpret_cast<char*>(s_dyn) + size_t(w) * size_t(wave_bytes);
  double* s_g = reinterpret_cast<double*>(base);                                  // 3 per term, then the row
  double* s_u = s_g + kvol_slot_area(fb.kv_max_terms, fb.max_deg2);               // 3 per unit
  uint8_t* s_unit = reinterpret_cast<uint8_t*>(s_u + 3 * fb.kv_max_units);        // 1 per term
  uint16_t* s_soff = reinterpret_cast<uint16_t*>(s_unit + ((fb.kv_max_terms + 7) & ~7));   // deg2 + 1
  const int4 nd = load_uniform(fb.kv_node, pn);   // {first term, first slot offset, terms, units}
  const int4 rec2 = load_uniform(fb.node_rec2, pn);
  const int u1b = p.uoff[pn];
  const int P = nd.x, F = nd.z, nu = nd.w, deg2 = rec2.y;
  constexpr int R = kKvolSlotMaxTerms / 64;
  // unit lane's unit record (first unit of a, unit (a, p)); the terms'
  // unit indices and ranks; the loads are unconditional from clamped indices
  // (every node has a term: it is its own element neighbour), so the
  // compiler issues them back to back
  const int4 ku = fb.kv_unit[u1b + min(lane, nu - 1)];
  int tm[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    if (64 * r >= F) break;
    tm[r] = fb.kv_term[P + min(lane + 64 * r, F - 1)];
  }
  for (int t = lane; t <= deg2; t += 64) s_soff[t] = fb.kv_soff[nd.y + t];
  const double4 gl = fb.g[ku.z];
  double4 gt[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    if (64 * r >= F) break;
    gt[r] = fb.g[__shfl(ku.x, tm[r] & 0xff) + (tm[r] >> 8)];   // g_a[q]: a's first unit from a's unit lane
  }
  if (lane < nu) {
    s_u[3 * lane] = gl.w * gl.x;
    s_u[3 * lane + 1] = gl.w * gl.y;
    s_u[3 * lane + 2] = gl.w * gl.z;
  }
#pragma unroll
  for (int r = 0; r < R; ++r) {
    if (64 * r >= F) break;
    const int k = lane + 64 * r;
    if (k < F) {
      s_g[3 * k] = gt[r].x;
      s_g[3 * k + 1] = gt[r].y;
      s_g[3 * k + 2] = gt[r].z;
      s_unit[k] = uint8_t(tm[r] & 0xff);
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  // CSR order of p's rows: value r * 3 deg2 + 3 s + c.  The sums go to LDS
  // in that order (over the terms, once every lane is done with them), then
  // consecutive lanes store consecutive values.
  const int row = 3 * deg2;
  double acc[2][9];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int s = lane + 64 * h;
#pragma unroll
    for (int v = 0; v < 9; ++v) acc[h][v] = 0.0;
    if (s < deg2) {
      const int k0 = s_soff[s], k1 = s_soff[s + 1];
      for (int k = k0; k < k1; ++k) {
        const int j = s_unit[k];
        const double g3[3] = {s_g[3 * k], s_g[3 * k + 1], s_g[3 * k + 2]};
        const double u3[3] = {s_u[3 * j], s_u[3 * j + 1], s_u[3 * j + 2]};
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
          for (int c = 0; c < 3; ++c) acc[h][3 * r + c] = fma(u3[r], g3[c], acc[h][3 * r + c]);
      }
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  double* s_row = s_g;   // 9 deg2 values: the term area, no longer read
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int s = lane + 64 * h;
    if (s < deg2) {
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) s_row[r * row + 3 * s + c] = acc[h][3 * r + c];
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  for (int v = lane; v < 3 * row; v += 64) kval[rec2.x + v] = bulk * s_row[v];
}

int32_t launch_tet4_fbar(const DevPlan& p, const FbarDev& fb, const double* u, const double* mu,
                         const double* lam, double bulk, double* f, double* kval, int precision, hipStream_t s,
                         int32_t* status) {
  int32_t rc = launch_permute_u(p, u, s, status);
  if (rc) return rc;
  // grids in multiples of 8 work-groups for the XCD-contiguous mapping
  auto grid = [](int64_t items, int per_group) { return dim3(unsigned(8 * ((items + 8 * per_group - 1) / (8 * per_group)))); };
  if (p.E > 0) hipLaunchKernelGGL(fbar_element_kernel, grid(p.E, 256), dim3(256), 0, s, p, fb, status);
  if (p.max_inc <= 32)
    hipLaunchKernelGGL(fbar_node_kernel<2>, grid(p.N, 8), dim3(256), 0, s, p, fb, kval ? 1 : 0);
  else
    hipLaunchKernelGGL(fbar_node_kernel<1>, grid(p.N, 4), dim3(256), 0, s, p, fb, kval ? 1 : 0);
  if (kval && fb.max_deg2 > 0 && fb.kv_slots) {
    // one node per wave: its terms, units and slot offsets in LDS
    const int wave_bytes = 8 * kvol_slot_area(fb.kv_max_terms, fb.max_deg2) + 24 * fb.kv_max_units +
                           ((fb.kv_max_terms + 7) & ~7) + ((2 * (fb.max_deg2 + 1) + 7) & ~7);
    const int waves = std::max(1, std::min(4, (64 << 10) / wave_bytes));
    hipLaunchKernelGGL(fbar_kvol_slots_kernel, grid(p.N, waves), dim3(64 * waves), size_t(wave_bytes) * waves, s,
                       p, fb, waves, wave_bytes, bulk, kval);
  } else if (kval && fb.max_deg2 > 0) {
    // up to four waves (two nodes each) per work-group within 64 KiB of LDS
    const size_t per_wave = size_t(2 * 9) * fb.max_deg2 * sizeof(double);
    const int waves = int(std::max<size_t>(1, std::min<size_t>(4, (size_t(64) << 10) / per_wave)));
    hipLaunchKernelGGL(fbar_kvol_kernel, grid(p.N, 2 * waves), dim3(64 * waves), per_wave * waves, s, p, fb, waves,
                       bulk, kval);
  }
  PSX_HIP_TRY(hipGetLastError());
  const FbarView fv{fb.node_rec2, fb.slot2};
  return launch_t4_fbar_rows(p, mu, lam, fb.jm1, fv, f, kval, precision, s, status);
}

}  // namespace psx
