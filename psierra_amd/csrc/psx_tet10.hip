// Tet10 Total Lagrangian Neo-Hookean kernels for gfx950 (4-point rule).
//
// Same owner-computes scheme as the Tet4 kernels: a work-group takes a tile
// of consecutive (Morton-ordered) nodes, lane t evaluates element incidence t
// (the 10 tangent blocks of the row node and its force, summed over the 4
// quadrature points) into LDS, then lanes (node, slot) reduce in FP64 and
// write the CSR values and f once each.  Follows
// Tet10TotalLagrangian.compute_nonlinear_force_and_tangent_stiffness
// (src/psierra/element/tet10_total_lagrangian.py:63-136) with the shape
// gradients of src/psierra/element/tet10_base.py:117-179 and the rule of
// src/psierra/element/quadrature.py:36-50.
#include <climits>

#include "psx_internal.h"
#include "psx_nh.h"

namespace psx {

// Quadrature point coordinates: a = (5 - sqrt5)/20, b = (5 + 3 sqrt5)/20.
constexpr double kQa = 0.1381966011250105151795413165634361882280;
constexpr double kQb = 0.5854101966249684544613760503096914353161;

__host__ __device__ constexpr double qp_coord(int q, int c) {
  return (q >= 1 && c == q - 1) ? kQb : kQa;
}

// d N_a / d xi_j at quadrature point q.
__host__ __device__ constexpr double tet10_dn(int q, int a, int j) {
  const double L1 = qp_coord(q, 0), L2 = qp_coord(q, 1), L3 = qp_coord(q, 2);
  const double L0 = 1.0 - L1 - L2 - L3;
  switch (a) {
    case 0: return -(4.0 * L0 - 1.0);
    case 1: return j == 0 ? 4.0 * L1 - 1.0 : 0.0;
    case 2: return j == 1 ? 4.0 * L2 - 1.0 : 0.0;
    case 3: return j == 2 ? 4.0 * L3 - 1.0 : 0.0;
    case 4: return j == 0 ? 4.0 * (L0 - L1) : -4.0 * L1;
    case 5: return j == 0 ? 4.0 * L2 : j == 1 ? 4.0 * L1 : 0.0;
    case 6: return j == 1 ? 4.0 * (L0 - L2) : -4.0 * L2;
    case 7: return j == 2 ? 4.0 * (L0 - L3) : -4.0 * L3;
    case 8: return j == 0 ? 4.0 * L3 : j == 2 ? 4.0 * L1 : 0.0;
    default: return j == 1 ? 4.0 * L3 : j == 2 ? 4.0 * L2 : 0.0;
  }
}

__device__ __forceinline__ int xcd_tile10() {
  const int per = gridDim.x >> 3;
  const int b = blockIdx.x;
  return (b & 7) * per + (b >> 3);
}

// Row p of a Tet10 element.  dX, dU: node coordinates/displacements minus
// node 0's (index 0 is zero).  sink(v, value) receives the 90 block values
// (v = 9 b + 3 i + k) then the 3 force values (v = 90 + i) once per
// quadrature point; the caller sums them.  Returns min over points of det F
// and sets jbad when det J <= 0 at some point.
template <typename R, bool WANT_K, class Sink>
__device__ __forceinline__ R tet10_row(const R dX[10][3], const R dU[10][3], R mu, R K, int p,
                                       Sink&& sink) {
  R Jmin = R(1);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    R Jm[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        R s = R(0);
#pragma unroll
        for (int a = 1; a < 10; ++a) s += dX[a][i] * R(tet10_dn(q, a, j));
        Jm[i][j] = s;
      }
    R adj[3][3];
    adj3(Jm, adj);
    const R det = Jm[0][0] * adj[0][0] + Jm[0][1] * adj[1][0] + Jm[0][2] * adj[2][0];
    const R idet = R(1) / det;
    const R w = det * R(1.0 / 24.0);
    // h_a[j] = sum_k dN(q,a,k) Jinv[k][j]; A = sum_a dU_a (x) h_a
    R A[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) A[i][j] = R(0);
    R hp[3] = {R(0), R(0), R(0)};
#pragma unroll
    for (int a = 0; a < 10; ++a) {
      R h[3];
#pragma unroll
      for (int j = 0; j < 3; ++j)
        h[j] = (R(tet10_dn(q, a, 0)) * adj[0][j] + R(tet10_dn(q, a, 1)) * adj[1][j] +
                R(tet10_dn(q, a, 2)) * adj[2][j]) * idet;
      if (a > 0) {
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int j = 0; j < 3; ++j) A[i][j] += dU[a][i] * h[j];
      }
      if (a == p) {
        hp[0] = h[0];
        hp[1] = h[1];
        hp[2] = h[2];
      }
    }
    NHPoint<R> m;
    nh_point(A, mu, K, m);
    Jmin = fmin(Jmin, m.J);
    R Shp[3], FSh[3];
    mat3vec(m.S, hp, Shp);
    mat3vec(m.F, Shp, FSh);
#pragma unroll
    for (int i = 0; i < 3; ++i) sink(90 + i, w * FSh[i]);
    if (WANT_K) {
      R gp[3], qp[3];
      mat3vec(m.Gm, hp, gp);
      mat3vec(m.F, hp, qp);
#pragma unroll
      for (int b = 0; b < 10; ++b) {
        R h[3], gb[3], qb[3], blk[9];
#pragma unroll
        for (int j = 0; j < 3; ++j)
          h[j] = (R(tet10_dn(q, b, 0)) * adj[0][j] + R(tet10_dn(q, b, 1)) * adj[1][j] +
                  R(tet10_dn(q, b, 2)) * adj[2][j]) * idet;
        mat3vec(m.Gm, h, gb);
        mat3vec(m.F, h, qb);
        nh_block(m, w, gp, qp, Shp, gb, qb, h, blk);
#pragma unroll
        for (int c = 0; c < 9; ++c) sink(9 * b + c, blk[c]);
      }
    }
  }
  return Jmin;
}

template <typename R, int TILE, bool WANT_K>
__global__ __launch_bounds__(TILE) void tet10_assemble_kernel(DevPlan p, const double* __restrict__ mu,
                                                              const double* __restrict__ lam,
                                                              double* __restrict__ f,
                                                              double* __restrict__ kval,
                                                              int32_t* status) {
  constexpr int NV = 93;
  __shared__ R contrib[NV * TILE];
  __shared__ uint32_t rk[3 * TILE];
  __shared__ int32_t s_incp[TILE + 1];
  __shared__ int32_t s_uoff[TILE + 1];

  const int tile = xcd_tile10();
  if (tile >= p.n_tiles) return;
  const int n0 = p.tile_node[tile], n1 = p.tile_node[tile + 1];
  const int nn = n1 - n0;
  const int k0 = p.inc_ptr[n0];
  const int nk = p.inc_ptr[n1] - k0;
  const int t = threadIdx.x;
  for (int i = t; i <= nn; i += TILE) {
    s_incp[i] = p.inc_ptr[n0 + i] - k0;
    s_uoff[i] = p.uoff[n0 + i] - p.uoff[n0];
  }
  if (t < nk) {
    const uint32_t pe = p.inc_elem[k0 + t];
    const int e = int(pe & kIncElemMask);
    const int lp = int(pe >> kIncElemBits);
#pragma unroll
    for (int w = 0; w < 3; ++w) rk[w * TILE + t] = p.inc_rank[3 * int64_t(k0 + t) + w];
    const int32_t* c = p.conn + 12 * int64_t(e);
    const int4 c0 = *reinterpret_cast<const int4*>(c);
    const int4 c1 = *reinterpret_cast<const int4*>(c + 4);
    const int2 c2 = *reinterpret_cast<const int2*>(c + 8);
    const int ids[10] = {c0.x, c0.y, c0.z, c0.w, c1.x, c1.y, c1.z, c1.w, c2.x, c2.y};
    const double4 X0 = p.Xi[ids[0]];
    const double4 U0 = p.ui[ids[0]];
    R dX[10][3], dU[10][3];
#pragma unroll
    for (int a = 0; a < 10; ++a) {
      if (a == 0) {
        dX[0][0] = dX[0][1] = dX[0][2] = R(0);
        dU[0][0] = dU[0][1] = dU[0][2] = R(0);
      } else {
        const double4 Xa = p.Xi[ids[a]];
        const double4 Ua = p.ui[ids[a]];
        dX[a][0] = R(Xa.x - X0.x);
        dX[a][1] = R(Xa.y - X0.y);
        dX[a][2] = R(Xa.z - X0.z);
        dU[a][0] = R(Ua.x - U0.x);
        dU[a][1] = R(Ua.y - U0.y);
        dU[a][2] = R(Ua.z - U0.z);
      }
    }
#pragma unroll
    for (int v = 0; v < NV; ++v) contrib[v * TILE + t] = R(0);
    const double m = mu[e], l = lam[e];
    const R J = tet10_row<R, WANT_K>(dX, dU, R(m), R(l + (2.0 / 3.0) * m), lp,
                                     [&](int v, R val) { contrib[v * TILE + t] += val; });
    if (J <= R(0)) {
      atomicMin(&status[1], e);
      status[0] = 1;
    }
  }
  __syncthreads();
  if (WANT_K) {
    const int units = s_uoff[nn];
    for (int w = t; w < units; w += TILE) {
      int lo = 0, hi = nn;
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (s_uoff[mid] <= w) lo = mid; else hi = mid;
      }
      const int j = lo;
      const uint32_t s = uint32_t(w - s_uoff[j]);
      const int node = n0 + j;
      double acc[9];
#pragma unroll
      for (int c = 0; c < 9; ++c) acc[c] = 0.0;
      double fa[3] = {0.0, 0.0, 0.0};
      for (int kk = s_incp[j]; kk < s_incp[j + 1]; ++kk) {
#pragma unroll
        for (int q = 0; q < 10; ++q) {
          const uint32_t r = (rk[(q >> 2) * TILE + kk] >> (8 * (q & 3))) & 0xffu;
          if (r == s) {
#pragma unroll
            for (int c = 0; c < 9; ++c) acc[c] += double(contrib[(9 * q + c) * TILE + kk]);
          }
        }
        if (s == 0) {
#pragma unroll
          for (int c = 0; c < 3; ++c) fa[c] += double(contrib[(90 + c) * TILE + kk]);
        }
      }
      const int d = p.deg[node];
      double* kb_out = kval + p.kofs[node] + 3 * int(s);
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) kb_out[r * 3 * d + c] = acc[3 * r + c];
      if (s == 0) {
        const int64_t a = p.perm[node];
#pragma unroll
        for (int c = 0; c < 3; ++c) f[3 * a + c] = fa[c];
      }
    }
    for (int j = t; j < nn; j += TILE) {
      if (s_uoff[j + 1] == s_uoff[j]) {
        const int64_t a = p.perm[n0 + j];
#pragma unroll
        for (int c = 0; c < 3; ++c) f[3 * a + c] = 0.0;
      }
    }
  } else {
    for (int j = t; j < nn; j += TILE) {
      double fa[3] = {0.0, 0.0, 0.0};
      for (int kk = s_incp[j]; kk < s_incp[j + 1]; ++kk)
#pragma unroll
        for (int c = 0; c < 3; ++c) fa[c] += double(contrib[(90 + c) * TILE + kk]);
      const int64_t a = p.perm[n0 + j];
#pragma unroll
      for (int c = 0; c < 3; ++c) f[3 * a + c] = fa[c];
    }
  }
}

int32_t launch_permute_u(const DevPlan& p, const double* u, hipStream_t s, int32_t* status);

template <typename R, int TILE>
static void launch_t10(const DevPlan& p, const double* mu, const double* lam, double* f, double* kval,
                       hipStream_t s, int32_t* status) {
  const int blocks = 8 * ((p.n_tiles + 7) / 8);
  if (kval)
    hipLaunchKernelGGL((tet10_assemble_kernel<R, TILE, true>), dim3(blocks), dim3(TILE), 0, s, p, mu,
                       lam, f, kval, status);
  else
    hipLaunchKernelGGL((tet10_assemble_kernel<R, TILE, false>), dim3(blocks), dim3(TILE), 0, s, p, mu,
                       lam, f, kval, status);
}

int32_t launch_tet10_assemble(const DevPlan& p, const double* u, const double* mu, const double* lam,
                              double* f, double* kval, int precision, hipStream_t s, int32_t* status) {
  int32_t rc = launch_permute_u(p, u, s, status);
  if (rc) return rc;
  if (p.n_tiles <= 0) return PSX_OK;
  const bool fp32 = precision == PSX_FP32;
  switch (p.tile) {
    case 64:
      fp32 ? launch_t10<float, 64>(p, mu, lam, f, kval, s, status)
           : launch_t10<double, 64>(p, mu, lam, f, kval, s, status);
      break;
    case 128:
      fp32 ? launch_t10<float, 128>(p, mu, lam, f, kval, s, status)
           : launch_t10<double, 128>(p, mu, lam, f, kval, s, status);
      break;
    default:
      return fail(PSX_EINVAL, "Tet10 supports tiles of 64 or 128 incidences");
  }
  PSX_HIP_TRY(hipGetLastError());
  return PSX_OK;
}

// Element batch: one lane per element, the 30x30 tangent and f.
template <typename R>
__global__ __launch_bounds__(64) void tet10_element_batch_kernel(int n, const double* __restrict__ X,
                                                                 const double* __restrict__ u,
                                                                 const double* __restrict__ mu,
                                                                 const double* __restrict__ lam,
                                                                 double* __restrict__ k,
                                                                 double* __restrict__ f,
                                                                 int32_t* status) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const double* Xe = X + 30 * int64_t(e);
  const double* ue = u + 30 * int64_t(e);
  R dX[10][3], dU[10][3];
  for (int a = 0; a < 10; ++a)
    for (int c = 0; c < 3; ++c) {
      dX[a][c] = R(Xe[3 * a + c] - Xe[c]);
      dU[a][c] = R(ue[3 * a + c] - ue[c]);
    }
  // det J at the quadrature points in FP64 (tet10_total_lagrangian.py:104-108)
  for (int q = 0; q < 4; ++q) {
    double Jm[3][3];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double s = 0.0;
        for (int a = 1; a < 10; ++a) s += (Xe[3 * a + i] - Xe[i]) * tet10_dn(q, a, j);
        Jm[i][j] = s;
      }
    if (!(det3(Jm) > 0.0)) {
      atomicMin(&status[1], e);
      status[0] = 2;
      return;
    }
  }
  const R m = R(mu[e]);
  const R Kb = R(lam[e] + (2.0 / 3.0) * mu[e]);
  double* ko = k ? k + 900 * int64_t(e) : nullptr;
  double* fo = f + 30 * int64_t(e);
  for (int p = 0; p < 10; ++p) {
    for (int i = 0; i < 3; ++i) fo[3 * p + i] = 0.0;
    if (ko)
      for (int i = 0; i < 3; ++i)
        for (int c = 0; c < 30; ++c) ko[30 * (3 * p + i) + c] = 0.0;
    R J;
    if (ko)
      J = tet10_row<R, true>(dX, dU, m, Kb, p, [&](int v, R val) {
        if (v >= 90) {
          fo[3 * p + (v - 90)] += double(val);
        } else {
          const int b = v / 9, i = (v % 9) / 3, c = v % 3;
          ko[30 * (3 * p + i) + 3 * b + c] += double(val);
        }
      });
    else
      J = tet10_row<R, false>(dX, dU, m, Kb, p, [&](int v, R val) {
        if (v >= 90) fo[3 * p + (v - 90)] += double(val);
      });
    if (p == 0 && J <= R(0)) {
      atomicMin(&status[1], e);
      status[0] = 1;
    }
  }
}

int32_t launch_tet10_element_batch(int64_t n, const double* X, const double* u, const double* mu,
                                   const double* lam, double* k, double* f, int precision,
                                   hipStream_t s, int32_t* status) {
  const int blocks = int((n + 63) / 64);
  if (precision == PSX_FP32)
    hipLaunchKernelGGL(tet10_element_batch_kernel<float>, dim3(blocks), dim3(64), 0, s, int(n), X, u,
                       mu, lam, k, f, status);
  else
    hipLaunchKernelGGL(tet10_element_batch_kernel<double>, dim3(blocks), dim3(64), 0, s, int(n), X, u,
                       mu, lam, k, f, status);
  PSX_HIP_TRY(hipGetLastError());
  return PSX_OK;
}

}  // namespace psx
