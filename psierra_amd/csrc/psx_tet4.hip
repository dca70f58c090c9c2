// Tet4 Total Lagrangian Neo-Hookean kernels for gfx950.
//
// Global assembly is owner-computes over node rows.  The plan orders nodes
// along a Morton curve (locality) and packs consecutive nodes into tiles of
// at most TILE element incidences.  One work-group handles one tile:
//   phase 1: lane t takes incidence t (element e seen from one of its nodes
//            a): it gathers the four nodes' X and u, evaluates the element in
//            registers (FP64 or FP32) and leaves the 4 tangent blocks of row
//            a (36 values) and f_a (3 values) in LDS;
//   phase 2: lane (a, s) sums, in FP64 and in incidence order, the blocks
//            whose column node has slot rank s in a's CSR row, and writes the
//            3x3 block straight into the CSR values; f_a likewise.
// Every CSR value and every f entry is written exactly once by one lane, in
// a fixed order: the scatter is deterministic and uses no atomics.
// This replaces SolidDomain's element loops + np.add.at + coo.tocsr()
// (src/psierra/domain/solid.py:230-325, src/psierra/utility/assembly.py:37-108).
#include <climits>

#include "psx_internal.h"
#include "psx_nh.h"
#include "psx_tet4_elem.h"

namespace psx {

// Work-group -> tile mapping that gives each of the 8 XCDs a contiguous run
// of tiles (consecutive work-groups are dispatched round-robin over XCDs),
// so neighbouring tiles share an L2.
__device__ __forceinline__ int xcd_tile() {
  const int per = gridDim.x >> 3;
  const int b = blockIdx.x;
  return (b & 7) * per + (b >> 3);
}

__global__ void permute_u_kernel(DevPlan p, const double* __restrict__ u, int32_t* status) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) {
    status[0] = 0;
    status[1] = INT_MAX;
  }
  if (i < p.N) {
    const int64_t a = p.perm[i];
    p.ui[i] = make_double4(u[3 * a], u[3 * a + 1], u[3 * a + 2], 0.0);
  }
}

__device__ __forceinline__ void flag_inversion(int32_t* status, int e) {
  atomicMin(&status[1], e);
  status[0] = 1;
}

template <typename R, int TILE, bool WANT_K>
__global__ __launch_bounds__(TILE) void tet4_assemble_kernel(DevPlan p, const double* __restrict__ mu,
                                                             const double* __restrict__ lam,
                                                             double* __restrict__ f,
                                                             double* __restrict__ kval,
                                                             int32_t* status) {
  constexpr int NV = WANT_K ? 39 : 3;
  constexpr int UMAX = kTileUnitsPerIncidence * TILE;
  __shared__ R contrib[NV * TILE];
  __shared__ uint32_t rk[TILE];
  __shared__ int32_t s_incp[TILE + 1];
  __shared__ int32_t s_uoff[TILE + 1];
  __shared__ unsigned long long s_mask[WANT_K ? UMAX : 1];

  const int tile = xcd_tile();
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
  if (WANT_K)
    for (int i = t; i < UMAX; i += TILE) s_mask[i] = 0ull;
  __syncthreads();
  if (t < nk) {
    const uint32_t pe = p.inc_elem[k0 + t];
    const int e = int(pe & kIncElemMask);
    const int lp = int(pe >> kIncElemBits);
    const int4 c = *reinterpret_cast<const int4*>(p.conn + 4 * int64_t(e));
    const int i0 = lp == 0 ? c.x : lp == 1 ? c.y : lp == 2 ? c.z : c.w;
    const int i1 = lp == 0 ? c.y : c.x;
    const int i2 = lp <= 1 ? c.z : c.y;
    const int i3 = lp <= 2 ? c.w : c.z;
    double4 X[4], U[4];
    X[0] = p.Xi[i0];
    X[1] = p.Xi[i1];
    X[2] = p.Xi[i2];
    X[3] = p.Xi[i3];
    U[0] = p.ui[i0];
    U[1] = p.ui[i1];
    U[2] = p.ui[i2];
    U[3] = p.ui[i3];
    const double m = mu[e], l = lam[e];
    if (WANT_K) {
      // Mark this incidence in the masks of the three off-diagonal slots it
      // feeds (bit = position within its node's incidence run).
      const uint32_t r = p.inc_rank[k0 + t];
      rk[t] = r;
      int lo = 0, hi = nn;
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (s_incp[mid] <= t) lo = mid; else hi = mid;
      }
      const unsigned long long bit = 1ull << (t - s_incp[lo]);
#pragma unroll
      for (int q = 1; q < 4; ++q) atomicOr(&s_mask[s_uoff[lo] + ((r >> (8 * q)) & 0xffu)], bit);
    }
    R out[NV];
    const R J = tet4_row<R, WANT_K>(X, U, R(m), R(l + (2.0 / 3.0) * m), out);
    if (J <= R(0)) flag_inversion(status, e);
#pragma unroll
    for (int v = 0; v < NV; ++v) contrib[v * TILE + t] = out[v];
  }
  __syncthreads();
  if (WANT_K) {
    // (i) lanes (node, value): the diagonal block (value 0..8) and f (9..11)
    //     are plain sums over the node's contiguous incidence run.
    // (ii) lanes (off-diagonal slot): sum the incidences flagged in its mask,
    //     in increasing incidence order.
    const int units = s_uoff[nn];
    const int nd = nn * 12;
    for (int w = t; w < nd + units; w += TILE) {
      if (w < nd) {
        const int j = w / 12, v = w - 12 * (w / 12);
        const int kb = s_incp[j], ke = s_incp[j + 1];
        const int src = v < 9 ? v : 27 + v;
        double acc = 0.0;
        for (int kk = kb; kk < ke; ++kk) acc += double(contrib[src * TILE + kk]);
        const int node = n0 + j;
        if (v < 9) {
          if (ke > kb) {
            const int d = p.deg[node];
            const int self = int(rk[kb] & 0xffu);
            kval[p.kofs[node] + (v / 3) * 3 * d + 3 * self + (v % 3)] = acc;
          }
        } else {
          f[3 * int64_t(p.perm[node]) + (v - 9)] = acc;
        }
      } else {
        const int u = w - nd;
        int lo = 0, hi = nn;
        while (hi - lo > 1) {
          const int mid = (lo + hi) >> 1;
          if (s_uoff[mid] <= u) lo = mid; else hi = mid;
        }
        const uint32_t s = uint32_t(u - s_uoff[lo]);
        const int kb = s_incp[lo];
        if ((rk[kb] & 0xffu) == s) continue;  // diagonal slot: done in (i)
        double acc[9];
#pragma unroll
        for (int c = 0; c < 9; ++c) acc[c] = 0.0;
        unsigned long long msk = s_mask[u];
        while (msk) {
          const int kk = kb + __ffsll(static_cast<long long>(msk)) - 1;
          msk &= msk - 1ull;
          const uint32_t r = rk[kk];
          const int q = ((r >> 8) & 0xffu) == s ? 1 : ((r >> 16) & 0xffu) == s ? 2 : 3;
#pragma unroll
          for (int c = 0; c < 9; ++c) acc[c] += double(contrib[(9 * q + c) * TILE + kk]);
        }
        const int node = n0 + lo;
        const int d = p.deg[node];
        double* kb_out = kval + p.kofs[node] + 3 * int(s);
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
          for (int c = 0; c < 3; ++c) kb_out[r * 3 * d + c] = acc[3 * r + c];
      }
    }
  } else {
    for (int j = t; j < nn; j += TILE) {
      double fa[3] = {0.0, 0.0, 0.0};
      for (int kk = s_incp[j]; kk < s_incp[j + 1]; ++kk)
#pragma unroll
        for (int c = 0; c < 3; ++c) fa[c] += double(contrib[c * TILE + kk]);
      const int64_t a = p.perm[n0 + j];
#pragma unroll
      for (int c = 0; c < 3; ++c) f[3 * a + c] = fa[c];
    }
  }
}

template <typename R, int TILE>
static void launch_t4(const DevPlan& p, const double* mu, const double* lam, double* f, double* kval,
                      hipStream_t s, int32_t* status) {
  const int blocks = 8 * ((p.n_tiles + 7) / 8);
  if (kval)
    hipLaunchKernelGGL((tet4_assemble_kernel<R, TILE, true>), dim3(blocks), dim3(TILE), 0, s, p, mu,
                       lam, f, kval, status);
  else
    hipLaunchKernelGGL((tet4_assemble_kernel<R, TILE, false>), dim3(blocks), dim3(TILE), 0, s, p, mu,
                       lam, f, kval, status);
}

int32_t launch_permute_u(const DevPlan& p, const double* u, hipStream_t s, int32_t* status) {
  hipLaunchKernelGGL(permute_u_kernel, dim3((p.N + 255) / 256), dim3(256), 0, s, p, u, status);
  PSX_HIP_TRY(hipGetLastError());
  return PSX_OK;
}

int32_t launch_tet4_assemble(const DevPlan& p, const double* u, const double* mu, const double* lam,
                             double* f, double* kval, int precision, hipStream_t s, int32_t* status) {
  int32_t rc = launch_permute_u(p, u, s, status);
  if (rc) return rc;
  if (p.n_tiles <= 0) return PSX_OK;
  const bool fp32 = precision == PSX_FP32;
  switch (p.tile) {
    case 64:
      fp32 ? launch_t4<float, 64>(p, mu, lam, f, kval, s, status)
           : launch_t4<double, 64>(p, mu, lam, f, kval, s, status);
      break;
    case 128:
      fp32 ? launch_t4<float, 128>(p, mu, lam, f, kval, s, status)
           : launch_t4<double, 128>(p, mu, lam, f, kval, s, status);
      break;
    case 256:
      fp32 ? launch_t4<float, 256>(p, mu, lam, f, kval, s, status)
           : launch_t4<double, 256>(p, mu, lam, f, kval, s, status);
      break;
    default:
      return fail(PSX_EINVAL, "unsupported tile");
  }
  PSX_HIP_TRY(hipGetLastError());
  return PSX_OK;
}

// ---------------------------------------------------------------------------
// Element batch: one lane per element, full 12x12 and f (no assembly).
// ---------------------------------------------------------------------------
template <typename R>
__global__ __launch_bounds__(128) void tet4_element_batch_kernel(int n, const double* __restrict__ X,
                                                                 const double* __restrict__ u,
                                                                 const double* __restrict__ mu,
                                                                 const double* __restrict__ lam,
                                                                 double* __restrict__ k,
                                                                 double* __restrict__ f,
                                                                 int32_t* status) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const double* Xe = X + 12 * int64_t(e);
  const double* ue = u + 12 * int64_t(e);
  // det([1 X]) in the element's own vertex order (src/psierra/element/tet4_base.py:51-61)
  double D[3][3];
  for (int b = 0; b < 3; ++b)
    for (int r = 0; r < 3; ++r) D[r][b] = Xe[3 * (b + 1) + r] - Xe[r];
  if (!(det3(D) > 0.0)) {
    atomicMin(&status[1], e);
    status[0] = 2;
    return;
  }
  const R m = R(mu[e]);
  const R Kb = R(lam[e] + (2.0 / 3.0) * mu[e]);
  for (int p = 0; p < 4; ++p) {
    int ord[4];
    ord[0] = p;
    for (int q = 0, s = 1; q < 4; ++q)
      if (q != p) ord[s++] = q;
    double4 Xs[4], Us[4];
    for (int b = 0; b < 4; ++b) {
      Xs[b] = make_double4(Xe[3 * ord[b]], Xe[3 * ord[b] + 1], Xe[3 * ord[b] + 2], 0.0);
      Us[b] = make_double4(ue[3 * ord[b]], ue[3 * ord[b] + 1], ue[3 * ord[b] + 2], 0.0);
    }
    R out[39];
    R J;
    if (k) {
      J = tet4_row<R, true>(Xs, Us, m, Kb, out);
      for (int b = 0; b < 4; ++b)
        for (int i = 0; i < 3; ++i)
          for (int c = 0; c < 3; ++c)
            k[144 * int64_t(e) + 12 * (3 * p + i) + 3 * ord[b] + c] = double(out[9 * b + 3 * i + c]);
      for (int i = 0; i < 3; ++i) f[12 * int64_t(e) + 3 * p + i] = double(out[36 + i]);
    } else {
      J = tet4_row<R, false>(Xs, Us, m, Kb, out);
      for (int i = 0; i < 3; ++i) f[12 * int64_t(e) + 3 * p + i] = double(out[i]);
    }
    if (p == 0 && J <= R(0)) flag_inversion(status, e);
  }
}

__global__ void status_reset_kernel(int32_t* status) {
  status[0] = 0;
  status[1] = INT_MAX;
}

int32_t launch_status_reset(hipStream_t s, int32_t* status) {
  hipLaunchKernelGGL(status_reset_kernel, dim3(1), dim3(1), 0, s, status);
  PSX_HIP_TRY(hipGetLastError());
  return PSX_OK;
}

int32_t launch_tet10_element_batch(int64_t n, const double* X, const double* u, const double* mu,
                                   const double* lam, double* k, double* f, int precision,
                                   hipStream_t s, int32_t* status);

int32_t launch_element_batch(int npe, int64_t n, const double* X, const double* u, const double* mu,
                             const double* lam, double* k, double* f, int precision, hipStream_t s,
                             int32_t* status) {
  int32_t rc = launch_status_reset(s, status);
  if (rc) return rc;
  if (npe == 10) return launch_tet10_element_batch(n, X, u, mu, lam, k, f, precision, s, status);
  const int blocks = int((n + 127) / 128);
  if (precision == PSX_FP32)
    hipLaunchKernelGGL(tet4_element_batch_kernel<float>, dim3(blocks), dim3(128), 0, s, int(n), X, u,
                       mu, lam, k, f, status);
  else
    hipLaunchKernelGGL(tet4_element_batch_kernel<double>, dim3(blocks), dim3(128), 0, s, int(n), X, u,
                       mu, lam, k, f, status);
  PSX_HIP_TRY(hipGetLastError());
  return PSX_OK;
}

// ---------------------------------------------------------------------------
// Material batch: S and the 6x6 Voigt tangent for n deformation gradients.
// ---------------------------------------------------------------------------
template <typename R>
__global__ __launch_bounds__(128) void nh_material_kernel(int n, const double* __restrict__ F,
                                                          const double* __restrict__ mu,
                                                          const double* __restrict__ bulk,
                                                          double* __restrict__ S,
                                                          double* __restrict__ C, int32_t* status) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  R A[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) A[i][j] = R(F[9 * int64_t(e) + 3 * i + j] - (i == j ? 1.0 : 0.0));
  NHPoint<R> m;
  nh_point(A, R(mu[e]), R(bulk[e]), m);
  if (m.J <= R(0)) flag_inversion(status, e);
  for (int i = 0; i < 9; ++i) S[9 * int64_t(e) + i] = double(m.S[i / 3][i % 3]);
  const int VI[6] = {0, 1, 2, 0, 1, 2}, VJ[6] = {0, 1, 2, 1, 2, 0};
  for (int a = 0; a < 6; ++a)
    for (int b = 0; b < 6; ++b) {
      const int i = VI[a], j = VJ[a], kk = VI[b], l = VJ[b];
      const R v = m.al * m.Ci[i][j] * m.Ci[kk][l] +
                  R(0.5) * m.be * (m.Ci[i][kk] * m.Ci[j][l] + m.Ci[i][l] * m.Ci[j][kk]) +
                  m.ga * (m.Ci[i][j] * (kk == l ? R(1) : R(0)) + (i == j ? R(1) : R(0)) * m.Ci[kk][l]);
      C[36 * int64_t(e) + 6 * a + b] = double(v);
    }
}

int32_t launch_material_batch(int64_t n, const double* F, const double* mu, const double* bulk,
                              double* S, double* C, int precision, hipStream_t s, int32_t* status) {
  int32_t rc = launch_status_reset(s, status);
  if (rc) return rc;
  const int blocks = int((n + 127) / 128);
  if (precision == PSX_FP32)
    hipLaunchKernelGGL(nh_material_kernel<float>, dim3(blocks), dim3(128), 0, s, int(n), F, mu, bulk,
                       S, C, status);
  else
    hipLaunchKernelGGL(nh_material_kernel<double>, dim3(blocks), dim3(128), 0, s, int(n), F, mu, bulk,
                       S, C, status);
  PSX_HIP_TRY(hipGetLastError());
  return PSX_OK;
}

}  // namespace psx
