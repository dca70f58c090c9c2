// Tet4 F-bar (nodal-averaged volume) assembly for gfx950.
//
// Restates Tet4Fbar's domain assembly (src/psierra/element/tet4_fbar.py:
// 226-459) as four stream-ordered steps after the displacement permutation:
//   1. fbar_element_kernel, one lane per element: J_e - 1 (from grad u, so
//      small strains stay accurate), V_e / 4, V_e (J_e - 1) / 4 and
//      w_e q_b = V_e J_e / 4 F h_b for the element's four nodes (q_b is the
//      node's column of div_e = B_0 + B_1 + B_2, the volumetric row of B);
//   2. fbar_node_kernel, one wave per node a, walking its incidences in
//      order: V0_a = sum V_e / 4, J_node - 1 = sum V_e (J_e - 1) / 4 / V0_a
//      and, with K, the volumetric gradient g_a over a's element neighbours,
//      g_a[b] = sum_{e in a, b} w_e q_b / V0_a (one lane per neighbour b);
//   3. fbar_kvol_kernel, one wave per node p, one lane per slot q of p's
//      distance-2 row: K_vol[p, q] = K_bulk sum_a V0_a g_a[p] g_a[q]^T over
//      the element neighbours a of p (ascending), written to every value of
//      the row;
//   4. the Tet4 row kernel with the F-bar material point (J_bar from step 2,
//      deviatoric tangent and f), adding K_dev onto the K_vol values.
// Every sum runs in an order fixed by the mesh; nothing uses atomics except
// the inversion flag.
#include <climits>

#include "psx_internal.h"
#include "psx_tet4_elem.h"

namespace psx {

__global__ __launch_bounds__(256) void fbar_element_kernel(DevPlan p, FbarDev fb, int32_t* status) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
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

__global__ __launch_bounds__(256) void fbar_node_kernel(DevPlan p, FbarDev fb, int want_g) {
  const int a = blockIdx.x * 4 + int(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (a >= p.N) return;
  const int4 rec = p.node_rec[a];
  const int kb = rec.x, ke = p.node_rec[a + 1].x;
  double sv = 0.0, sj = 0.0;
  for (int k = kb; k < ke; ++k) {
    const int e = int(p.inc_elem[k] & kIncElemMask);
    sv += fb.qv[e];
    sj += fb.vjm1[e];
  }
  if (lane == 0) {
    fb.v0[a] = sv;
    fb.jm1[a] = sv > 0.0 ? sj / sv : 0.0;
  }
  if (!want_g) return;
  const int ub = rec.y, deg = p.node_rec[a + 1].y - ub, self = rec.w;
  const double iv = sv > 0.0 ? 1.0 / sv : 0.0;
  for (int s = lane; s < deg; s += 64) {
    double g0 = 0.0, g1 = 0.0, g2 = 0.0;
    for (int k = kb; k < ke; ++k) {
      const uint32_t pe = p.inc_elem[k], r = p.inc_rank[k];
      const int lp = int(pe >> kIncElemBits);
      int loc = -1;
      if (s == self)
        loc = lp;
      else
#pragma unroll
        for (int q = 1; q < 4; ++q)
          if (int((r >> (8 * q)) & 0xffu) == s) loc = other_local(lp, q);
      if (loc >= 0) {
        const double* w = fb.wq + 12 * int64_t(pe & kIncElemMask) + 3 * loc;
        g0 += w[0];
        g1 += w[1];
        g2 += w[2];
      }
    }
    double* g = fb.g + 3 * int64_t(ub + s);
    g[0] = g0 * iv;
    g[1] = g1 * iv;
    g[2] = g2 * iv;
  }
}

// Position of `key` in the ascending list v[0..n), or n if absent.
__device__ __forceinline__ int find_sorted(const int32_t* v, int n, int32_t key) {
  int lo = 0, len = n;
  while (len > 0) {
    const int half = len >> 1;
    if (v[lo + half] < key) {
      lo += half + 1;
      len -= half + 1;
    } else {
      len = half;
    }
  }
  return (lo < n && v[lo] == key) ? lo : n;
}

__global__ __launch_bounds__(256) void fbar_kvol_kernel(DevPlan p, FbarDev fb, double bulk, double* kval) {
  const int pn = blockIdx.x * 4 + int(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (pn >= p.N) return;
  const int p_user = p.perm[pn];
  const int u1b = p.uoff[pn], u1e = p.uoff[pn + 1];
  const int u2b = fb.uoff2[pn];
  const int4 rec2 = fb.node_rec2[pn];
  const int deg2 = rec2.y;
  for (int s2 = lane; s2 - lane < deg2; s2 += 64) {
    const bool active = s2 < deg2;
    const int32_t q_user = active ? fb.nbr2_user[u2b + s2] : -1;
    double acc[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) acc[i] = 0.0;
    for (int u = u1b; u < u1e; ++u) {
      const int an = fb.nbr1_int[u];
      const int ab = p.uoff[an], n = p.uoff[an + 1] - ab;
      const int32_t* list = fb.nbr1_user + ab;
      const int pp = find_sorted(list, n, p_user);   // p is a neighbour of a: always found
      if (pp >= n) continue;
      const double w = fb.v0[an];
      const double* gp = fb.g + 3 * int64_t(ab + pp);
      const double wp[3] = {w * gp[0], w * gp[1], w * gp[2]};
      if (active) {
        const int pq = find_sorted(list, n, q_user);
        if (pq < n) {
          const double* gq = fb.g + 3 * int64_t(ab + pq);
#pragma unroll
          for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) acc[3 * r + c] += wp[r] * gq[c];
        }
      }
    }
    if (active) {
      double* dst = kval + rec2.x + 3 * s2;
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) dst[r * 3 * deg2 + c] = bulk * acc[3 * r + c];
    }
  }
}

int32_t launch_tet4_fbar(const DevPlan& p, const FbarDev& fb, const double* u, const double* mu,
                         const double* lam, double bulk, double* f, double* kval, int precision, hipStream_t s,
                         int32_t* status) {
  int32_t rc = launch_permute_u(p, u, s, status);
  if (rc) return rc;
  if (p.E > 0) hipLaunchKernelGGL(fbar_element_kernel, dim3((p.E + 255) / 256), dim3(256), 0, s, p, fb, status);
  const int node_blocks = (p.N + 3) / 4;
  hipLaunchKernelGGL(fbar_node_kernel, dim3(node_blocks), dim3(256), 0, s, p, fb, kval ? 1 : 0);
  if (kval) hipLaunchKernelGGL(fbar_kvol_kernel, dim3(node_blocks), dim3(256), 0, s, p, fb, bulk, kval);
  PSX_HIP_TRY(hipGetLastError());
  const FbarView fv{fb.node_rec2, fb.slot2};
  return launch_t4_fbar_rows(p, mu, lam, fb.jm1, fv, f, kval, precision, s, status);
}

}  // namespace psx
