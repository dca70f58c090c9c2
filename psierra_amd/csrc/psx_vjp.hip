// Adjoint (vector-Jacobian product) of the Tet4 internal force.
//
// For a cotangent w (one value per DOF) it returns, with f = f_int(u; mu, lambda):
//   g_u      = d(w^T f)/du      = K_T w   (K_T symmetric),
//   g_mu[e]  = d(w^T f)/d mu_e  = V_e dS/dmu : (F^T grad w),
//   g_lam[e] = d(w^T f)/d lam_e = V_e ln J C^-1 : (F^T grad w).
// pSierra has no analytic gradient: its parameter identification uses a
// 2-point finite-difference Jacobian of a global parameter through
// scipy.optimize.least_squares (src/psierra/inverse/parameter_identification.py:370-379);
// by the chain rule that derivative is sum_e g_mu[e] (or g_lam).
//
// Matrix-free and owner-computes like the assembly: a work-group takes a tile
// of nodes, lane t evaluates incidence t's row of K_e w (3 values) into LDS,
// then node lanes sum in FP64 in incidence order.  The per-element outputs
// are written by the single incidence with local node 0.  Deterministic,
// no atomics.
#include <climits>

#include "psx_device.h"
#include "psx_internal.h"
#include "psx_tet4_elem.h"
#include "psx_batch.h"

namespace psx {

// u and w to internal order in one pass (one user id load per node); the
// status words are reset here, as the displacement permutation of the other
// paths does.
__global__ void permute_uw_kernel(DevPlan p, const double* __restrict__ u, const double* __restrict__ w,
                                  int32_t* status) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) {
    status[0] = 0;
    status[1] = INT_MAX;
  }
  if (i < p.N) {
    const int64_t a = p.perm[i];
    const double4 ua = make_double4(u[3 * a], u[3 * a + 1], u[3 * a + 2], 0.0);
    const double4 wa = make_double4(w[3 * a], w[3 * a + 1], w[3 * a + 2], 0.0);
    p.ui[i] = ua;
    p.wi[i] = wa;
  }
}

// One tile of node rows (P: DevPlan, or a batch's plan view in the constant
// address space).
template <typename R, int TILE, typename P>
__device__ __forceinline__ void vjp_tile(const P& p, const double* __restrict__ mu, const double* __restrict__ lam,
                                         double* __restrict__ gu, double* __restrict__ gmu,
                                         double* __restrict__ glam, int32_t* status, int tile, int t) {
  __shared__ R contrib[3 * TILE];
  __shared__ int32_t s_incp[TILE + 1];
  // the coordinates, displacements and cotangents of the tile's nodes and
  // halo staged in LDS (plans that are not wide, tiles of at most
  // kTileLdsNodes nodes): one gather per node instead of twelve per incidence
  constexpr int NLU = kTileLdsNodes;
  __shared__ double s_n[9 * NLU];
  const int n0 = p.tile_node[tile], n1 = p.tile_node[tile + 1];
  const int nn = n1 - n0;
  const int k0 = p.inc_ptr[n0];
  const int nk = p.inc_ptr[n1] - k0;
  int2 th = make_int2(0, -1);
  if (p.tile_halo) th = load_uniform(p.tile_halo, tile);
  const bool lds = th.y >= 0;
  for (int i = t; i <= nn; i += TILE) s_incp[i] = p.inc_ptr[n0 + i] - k0;
  uint32_t pe = 0u, loc = 0u;
  if (t < nk) {
    pe = p.inc_elem[k0 + t];
    if (lds) loc = p.inc_loc[k0 + t];
  }
  if (lds) {
    for (int i = t; i < nn + th.y; i += TILE) {
      const int node = i < nn ? n0 + i : p.halo[th.x + i - nn];
      const double4 x = p.Xi[node], u = p.ui[node], w = p.wi[node];
      s_n[i] = x.x;
      s_n[NLU + i] = x.y;
      s_n[2 * NLU + i] = x.z;
      s_n[3 * NLU + i] = u.x;
      s_n[4 * NLU + i] = u.y;
      s_n[5 * NLU + i] = u.z;
      s_n[6 * NLU + i] = w.x;
      s_n[7 * NLU + i] = w.y;
      s_n[8 * NLU + i] = w.z;
    }
    __syncthreads();
  }
  if (t < nk) {
    const int e = int(pe & kIncElemMask);
    const int lp = int(pe >> kIncElemBits);
    double4 X[4], U[4], W[4];
    if (lds) {
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const int l = int((loc >> (8 * b)) & 0xffu);
        X[b] = make_double4(s_n[l], s_n[NLU + l], s_n[2 * NLU + l], 0.0);
        U[b] = make_double4(s_n[3 * NLU + l], s_n[4 * NLU + l], s_n[5 * NLU + l], 0.0);
        W[b] = make_double4(s_n[6 * NLU + l], s_n[7 * NLU + l], s_n[8 * NLU + l], 0.0);
      }
    } else {
      const int4 c = *reinterpret_cast<const int4*>(p.conn + 4 * int64_t(e));
      const int ids[4] = {lp == 0 ? c.x : lp == 1 ? c.y : lp == 2 ? c.z : c.w, lp == 0 ? c.y : c.x,
                          lp <= 1 ? c.z : c.y, lp <= 2 ? c.w : c.z};
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        X[b] = p.Xi[ids[b]];
        U[b] = p.ui[ids[b]];
        W[b] = p.wi[ids[b]];
      }
    }
    const double m = mu[e], l = lam[e];
    R out[3], dmu, dlam;
    const R J = tet4_vjp_row<R>(X, U, W, R(m), R(l + (2.0 / 3.0) * m), out, dmu, dlam);
    if (J <= R(0)) {
      atomicMin(&status[1], e);
      status[0] = 1;
    }
#pragma unroll
    for (int v = 0; v < 3; ++v) contrib[v * TILE + t] = out[v];
    if (lp == 0 && gmu) {
      gmu[e] = double(dmu);
      glam[e] = double(dlam);
    }
  }
  __syncthreads();
  for (int j = t; j < nn; j += TILE) {
    double acc[3] = {0.0, 0.0, 0.0};
    for (int kk = s_incp[j]; kk < s_incp[j + 1]; ++kk)
#pragma unroll
      for (int v = 0; v < 3; ++v) acc[v] += double(contrib[v * TILE + kk]);
    const int64_t a = p.perm[n0 + j];
#pragma unroll
    for (int v = 0; v < 3; ++v) gu[3 * a + v] = acc[v];
  }
}

template <typename R, int TILE>
__global__ __launch_bounds__(TILE) void tet4_vjp_kernel(DevPlan p, const double* __restrict__ mu,
                                                        const double* __restrict__ lam,
                                                        double* __restrict__ gu,
                                                        double* __restrict__ gmu,
                                                        double* __restrict__ glam, int32_t* status) {
  const int tile = xcd_tile();
  if (tile >= p.n_tiles) return;
  vjp_tile<R, TILE>(p, mu, lam, gu, gmu, glam, status, tile, int(threadIdx.x));
}

// The adjoint of many plans in one launch (psx_batch.h): instance i's
// cotangent in its plan's internal order, then every tile of every plan.
__global__ __launch_bounds__(256) void batch_permute_w_kernel(BatchDev b, const double* const* __restrict__ w) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= b.n_nodes) return;
  const int inst = batch_owner<false>(b.node_start, b.n, i);
  const int k = i - b.node_start[inst];
  const DevPlan& p = b.plans[inst];
  const double* wi = w[inst];
  const int64_t a = p.perm[k];
  p.wi[k] = make_double4(wi[3 * a], wi[3 * a + 1], wi[3 * a + 2], 0.0);
}

template <typename R, int TILE>
__global__ __launch_bounds__(TILE) void tet4_batch_vjp(BatchDev b, const double* const* mu, const double* const* lam,
                                                       double* const* gu, double* const* gmu,
                                                       double* const* glam, int32_t* status) {
  batch_tiles<false>(b, [&](int inst, int tile, int t) {
    const ConstDevPlan& p = *(const ConstDevPlan*)(b.plans + inst);
    double* gm = gmu ? load_uniform_val(gmu + inst) : nullptr;
    double* gl = glam ? load_uniform_val(glam + inst) : nullptr;
    vjp_tile<R, TILE>(p, load_uniform_val(mu + inst), load_uniform_val(lam + inst), load_uniform_val(gu + inst), gm,
                      gl, status + 2 * inst, tile, t);
  });
}

template <typename R, int TILE>
static void launch_batch_vjp_t(const BatchDev& b, const double* const* mu, const double* const* lam,
                               double* const* gu, double* const* gmu, double* const* glam, hipStream_t s,
                               int32_t* status) {
  const dim3 grid(unsigned(8 * ((int64_t(b.n_tiles) + 7) / 8)));
  hipLaunchKernelGGL((tet4_batch_vjp<R, TILE>), grid, dim3(TILE), 0, s, b, mu, lam, gu, gmu, glam, status);
}

int32_t launch_tet4_batch_vjp(const BatchDev& b, const double* const* u, const double* const* mu,
                              const double* const* lam, const double* const* w, double* const* gu,
                              double* const* gmu, double* const* glam, int precision, hipStream_t s,
                              int32_t* status) {
  int32_t rc = launch_batch_permute(b, u, s, status);
  if (rc) return rc;
  if (b.n_nodes > 0)
    hipLaunchKernelGGL(batch_permute_w_kernel, dim3((b.n_nodes + 255) / 256), dim3(256), 0, s, b, w);
  PSX_HIP_TRY(hipGetLastError());
  if (b.n_tiles <= 0) return PSX_OK;
  const bool fp32 = precision == PSX_FP32;
  switch (b.tile) {
    case 64:
      fp32 ? launch_batch_vjp_t<float, 64>(b, mu, lam, gu, gmu, glam, s, status)
           : launch_batch_vjp_t<double, 64>(b, mu, lam, gu, gmu, glam, s, status);
      break;
    case 128:
      fp32 ? launch_batch_vjp_t<float, 128>(b, mu, lam, gu, gmu, glam, s, status)
           : launch_batch_vjp_t<double, 128>(b, mu, lam, gu, gmu, glam, s, status);
      break;
    case 256:
      fp32 ? launch_batch_vjp_t<float, 256>(b, mu, lam, gu, gmu, glam, s, status)
           : launch_batch_vjp_t<double, 256>(b, mu, lam, gu, gmu, glam, s, status);
      break;
    default:
      return fail(PSX_EINVAL, "unsupported tile");
  }
  PSX_HIP_TRY(hipGetLastError());
  return PSX_OK;
}


template <typename R, int TILE>
static void launch_vjp_t(const DevPlan& p, const double* mu, const double* lam, double* gu, double* gmu,
                         double* glam, hipStream_t s, int32_t* status) {
  const int blocks = 8 * ((p.n_tiles + 7) / 8);
  hipLaunchKernelGGL((tet4_vjp_kernel<R, TILE>), dim3(blocks), dim3(TILE), 0, s, p, mu, lam, gu, gmu,
                     glam, status);
}

int32_t launch_tet4_vjp(const DevPlan& p, const double* u, const double* mu, const double* lam,
                        const double* w, double* gu, double* gmu, double* glam, int precision,
                        hipStream_t s, int32_t* status) {
  hipLaunchKernelGGL(permute_uw_kernel, dim3((p.N + 255) / 256), dim3(256), 0, s, p, u, w, status);
  PSX_HIP_TRY(hipGetLastError());
  if (p.n_tiles <= 0) return PSX_OK;
  const bool fp32 = precision == PSX_FP32;
  switch (p.tile) {
    case 64:
      fp32 ? launch_vjp_t<float, 64>(p, mu, lam, gu, gmu, glam, s, status)
           : launch_vjp_t<double, 64>(p, mu, lam, gu, gmu, glam, s, status);
      break;
    case 128:
      fp32 ? launch_vjp_t<float, 128>(p, mu, lam, gu, gmu, glam, s, status)
           : launch_vjp_t<double, 128>(p, mu, lam, gu, gmu, glam, s, status);
      break;
    case 256:
      fp32 ? launch_vjp_t<float, 256>(p, mu, lam, gu, gmu, glam, s, status)
           : launch_vjp_t<double, 256>(p, mu, lam, gu, gmu, glam, s, status);
      break;
    default:
      return fail(PSX_EINVAL, "unsupported tile");
  }
  PSX_HIP_TRY(hipGetLastError());
  return PSX_OK;
}

}  // namespace psx
