// Consistent mass matrix assembly into the plan's CSR pattern (FP64).
//
// Replaces SolidDomain.assemble_mass_matrix (src/psierra/domain/solid.py:
// 208-228) with the element matrices of Tet4Base.compute_mass_matrix
// (src/psierra/element/tet4_base.py:138-151) and
// Tet10Base.compute_mass_matrix (src/psierra/element/tet10_base.py:344-372):
// m_e[a][b] = rho_e V_e c_ab I3, c the reference mass tensor of the element
// (integral of N_a N_b over the element divided by its volume).
//
// One lane per CSR block (node a, slot s): it walks a's incidence run in
// order and adds rho_e V_e c[p][q] for every element that contains the
// slot's node (p = a's local index, q = the slot node's local index, found
// from the incidence's rank bytes); the 3x3 block is written once with that
// sum on its diagonal and explicit zeros elsewhere (the pattern keeps them,
// as coo.tocsr does).  Deterministic, no atomics.
#include "psx_device.h"
#include "psx_internal.h"
#include "psx_nh.h"

namespace psx {

// Volume of the tetrahedron on the element's four corner nodes (FP64).
__device__ __forceinline__ double corner_volume(const DevPlan& p, int64_t e) {
  const int32_t* c = p.conn + int64_t(p.conn_stride) * e;
  const double4 X0 = p.Xi[c[0]], X1 = p.Xi[c[1]], X2 = p.Xi[c[2]], X3 = p.Xi[c[3]];
  double D[3][3];
  const double4 Xs[3] = {X1, X2, X3};
#pragma unroll
  for (int b = 0; b < 3; ++b) {
    D[0][b] = Xs[b].x - X0.x;
    D[1][b] = Xs[b].y - X0.y;
    D[2][b] = Xs[b].z - X0.z;
  }
  return det3(D) * (1.0 / 6.0);
}

template <int NPE>
__global__ __launch_bounds__(256) void mass_rows_kernel(DevPlan p, const double* __restrict__ rho,
                                                        const double* __restrict__ coef,
                                                        double* __restrict__ kval) {
  const int64_t u = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t n_units = p.uoff[p.N];
  if (u >= n_units) return;
  // node owning unit u
  int lo = 0, hi = p.N;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (p.uoff[mid] <= u) lo = mid; else hi = mid;
  }
  const int a = lo;
  const int4 rec = p.node_rec[a];
  const int d = p.uoff[a + 1] - rec.y;
  const int s = int(u - rec.y);
  const bool self = s == rec.w;
  const int k1 = p.node_rec[a + 1].x;
  double sum = 0.0;
  for (int k = rec.x; k < k1; ++k) {
    const uint32_t pe = p.inc_elem[k];
    const int64_t e = int64_t(pe & kIncElemMask);
    const int lp = int(pe >> kIncElemBits);
    int q = -1;
    if (self) {
      q = lp;
    } else if (NPE == 4) {
      // bytes 1..3: the other three nodes in increasing local index
      const uint32_t r = p.inc_rank[k];
#pragma unroll
      for (int j = 1; j < 4; ++j)
        if (((r >> (8 * j)) & 0xffu) == uint32_t(s)) q = (j - 1 < lp) ? j - 1 : j;
    } else {
#pragma unroll
      for (int j = 0; j < 10; ++j)
        if (j != lp && ((p.inc_rank[3 * int64_t(k) + (j >> 2)] >> (8 * (j & 3))) & 0xffu) == uint32_t(s)) q = j;
    }
    if (q >= 0) sum += rho[e] * corner_volume(p, e) * coef[lp * NPE + q];
  }
  double* dst = kval + rec.z + 3 * s;
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) dst[r * 3 * d + c] = r == c ? sum : 0.0;
}

int32_t launch_mass(const DevPlan& p, const double* rho, const double* coef, double* kval, hipStream_t s,
                    int64_t n_units) {
  if (n_units <= 0) return PSX_OK;
  const unsigned blocks = unsigned((n_units + 255) / 256);
  if (p.npe == 4)
    hipLaunchKernelGGL(mass_rows_kernel<4>, dim3(blocks), dim3(256), 0, s, p, rho, coef, kval);
  else
    hipLaunchKernelGGL(mass_rows_kernel<10>, dim3(blocks), dim3(256), 0, s, p, rho, coef, kval);
  PSX_HIP_TRY(hipGetLastError());
  return PSX_OK;
}

}  // namespace psx
