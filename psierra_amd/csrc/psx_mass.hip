// Consistent mass matrix assembly into the plan's CSR pattern (FP64).
//
// Replaces SolidDomain.assemble_mass_matrix (src/psierra/domain/solid.py:
// 208-228) with the element matrices of Tet4Base.compute_mass_matrix
// (src/psierra/element/tet4_base.py:138-151) and
// Tet10Base.compute_mass_matrix (src/psierra/element/tet10_base.py:344-372):
// m_e[a][b] = rho_e V_e c_ab I3, c the reference mass tensor of the element
// (integral of N_a N_b over the element divided by its volume).
//
// Sixteen lanes per node a (four nodes per wave): a's incidences (at most
// 64) leave rho_e V_e (V_e from the plan's element volumes, computed once by
// element_volume_kernel), a's local index and the rank bytes in LDS (the
// mass tensor is in LDS too); then
// one lane per CSR block (a, slot s) walks the staged run in order and adds
// rho_e V_e c[p][q] for every element that contains the slot's node (p = a's
// local index, q = the slot node's local index, found from the rank bytes);
// the 3x3 block is written once with that sum on its diagonal and explicit
// zeros elsewhere (the pattern keeps them, as coo.tocsr does).
// Deterministic, no atomics.
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

// V_e of every element (its corner tetrahedron), once per plan.
__global__ __launch_bounds__(256) void element_volume_kernel(DevPlan p, double* __restrict__ vol) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e < p.E) vol[e] = corner_volume(p, e);
}

template <int NPE, int NPW>
__global__ __launch_bounds__(256) void mass_node_kernel(DevPlan p, const double* __restrict__ rho,
                                                        const double* __restrict__ vol,
                                                        const double* __restrict__ coef,
                                                        double* __restrict__ kval) {
  constexpr int NRW = NPE == 4 ? 1 : 3;   // rank words per incidence
  constexpr int W = 64 / NPW;             // lanes per node
  constexpr int KMAX = 64;                // incidences per node at most
  __shared__ double s_w[4 * NPW][KMAX];
  __shared__ uint32_t s_rank[4 * NPW][KMAX][NRW];
  __shared__ uint8_t s_lp[4 * NPW][KMAX];
  __shared__ double s_coef[NPE * NPE];
  // Tet4 nodes of at most W slots: per slot, the incidences that feed it
  __shared__ unsigned long long s_fm[NPE == 4 ? 4 * NPW : 1][NPE == 4 ? W : 1];
  const int w = int(threadIdx.x >> 6);
  const int sub = int(threadIdx.x & 63) / W, lane = int(threadIdx.x & 63) % W;
  const int g = w * NPW + sub;
  const int a = (xcd_tile() * 4 + w) * NPW + sub;
  // the node's records are loaded before the mass tensor's barrier
  int4 rec = make_int4(0, 0, 0, 0), rec1 = make_int4(0, 0, 0, 0);
  if (a < p.N) {
    rec = p.node_rec[a];
    rec1 = p.node_rec[a + 1];
  }
  for (int i = int(threadIdx.x); i < NPE * NPE; i += int(blockDim.x)) s_coef[i] = coef[i];
  __syncthreads();
  if (a >= p.N) return;
  const int kb = rec.x, L = rec1.x - rec.x;
  const int d = rec1.y - rec.y;
  const bool narrow = d <= W;
  constexpr bool MASKS = NPE == 4;
  if (MASKS) s_fm[MASKS ? g : 0][MASKS ? lane : 0] = 0ull;
  // the node's lanes are this wave's own: its LDS writes land in order
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
#pragma unroll
  for (int kk = 0; kk < KMAX / W; ++kk) {
    const int k = lane + kk * W;
    if (k < L) {
      const uint32_t pe = p.inc_elem[kb + k];
      const int64_t e = int64_t(pe & kIncElemMask);
      s_w[g][k] = rho[e] * vol[e];
      s_lp[g][k] = uint8_t(pe >> kIncElemBits);
      uint32_t rk[NRW];
#pragma unroll
      for (int i = 0; i < NRW; ++i) {
        rk[i] = p.inc_rank[NRW * int64_t(kb + k) + i];
        s_rank[g][k][i] = rk[i];
      }
      if (MASKS && narrow) {
        // bytes 1..3: the slots of the element's other three nodes
#pragma unroll
        for (int j = 1; j < 4; ++j)
          atomicOr(&s_fm[MASKS ? g : 0][MASKS ? int((rk[0] >> (8 * j)) & 0xffu) : 0], 1ull << k);
      }
    }
  }
  // the run is this wave's own: its LDS writes land before its reads
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  if (MASKS && narrow) {
    // diagonal slot: the node's lanes sum every W-th incidence in order, then
    // combine in a fixed butterfly; off-diagonal slots walk their feeds in
    // increasing incidence order
    double part = 0.0;
#pragma unroll
    for (int kk = 0; kk < KMAX / W; ++kk) {
      const int k = lane + kk * W;
      if (k < L) {
        const int lp = s_lp[g][k];
        part += s_w[g][k] * s_coef[lp * NPE + lp];
      }
    }
#pragma unroll
    for (int off = W / 2; off > 0; off >>= 1) part += __shfl_xor(part, off, W);
    const int s = lane;
    double sum = 0.0;
    if (s == rec.w) {
      sum = part;
    } else if (s < d) {
      unsigned long long m = s_fm[MASKS ? g : 0][MASKS ? s : 0];
      while (m) {
        const int k = low_bit(m);
        m &= m - 1;
        const int lp = s_lp[g][k];
        const uint32_t r = s_rank[g][k][0];
        int q = 0;
#pragma unroll
        for (int j = 1; j < 4; ++j)
          if (((r >> (8 * j)) & 0xffu) == uint32_t(s)) q = (j - 1 < lp) ? j - 1 : j;
        sum += s_w[g][k] * s_coef[lp * NPE + q];
      }
    }
    double* dst = kval + rec.z;
#pragma unroll
    for (int r = 0; r < 3; ++r)
      for (int v = lane; v < 3 * W; v += W) {
        const int sv = v / 3, c = v - 3 * sv;
        const double x = __shfl(sum, sv, W);   // every lane of the node takes part
        if (v < 3 * d) dst[r * 3 * d + v] = c == r ? x : 0.0;
      }
    return;
  }
  // a node of at most W slots: one slot per lane, then each row stored with
  // consecutive lanes on consecutive values (the slot sums shuffled within
  // the node's lanes); otherwise each lane stores its slots' blocks
  for (int s = lane; s < (narrow ? W : d); s += W) {
    const bool self = s == rec.w;
    double sum = 0.0;
#pragma unroll 4
    for (int k = 0; k < L; ++k) {
      const int lp = s_lp[g][k];
      int q = -1;
      if (self) {
        q = lp;
      } else if (NPE == 4) {
        // bytes 1..3: the other three nodes in increasing local index
        const uint32_t r = s_rank[g][k][0];
#pragma unroll
        for (int j = 1; j < 4; ++j)
          if (((r >> (8 * j)) & 0xffu) == uint32_t(s)) q = (j - 1 < lp) ? j - 1 : j;
      } else {
#pragma unroll
        for (int j = 0; j < 10; ++j)
          if (j != lp && ((s_rank[g][k][j >> 2] >> (8 * (j & 3))) & 0xffu) == uint32_t(s)) q = j;
      }
      if (q >= 0) sum += s_w[g][k] * s_coef[lp * NPE + q];
    }
    if (narrow) {
      double* dst = kval + rec.z;
#pragma unroll
      for (int r = 0; r < 3; ++r)
        for (int v = lane; v < 3 * W; v += W) {
          const int sv = v / 3, c = v - 3 * sv;
          const double x = __shfl(sum, sv, W);   // every lane of the node takes part
          if (v < 3 * d) dst[r * 3 * d + v] = c == r ? x : 0.0;
        }
    } else {
      double* dst = kval + rec.z + 3 * s;
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) dst[r * 3 * d + c] = r == c ? sum : 0.0;
    }
  }
}

int32_t launch_element_volumes(const DevPlan& p, double* vol, hipStream_t s) {
  if (p.E <= 0) return PSX_OK;
  hipLaunchKernelGGL(element_volume_kernel, dim3(unsigned((int64_t(p.E) + 255) / 256)), dim3(256), 0, s, p, vol);
  PSX_HIP_TRY(hipGetLastError());
  return PSX_OK;
}

int32_t launch_mass(const DevPlan& p, const double* rho, const double* vol, const double* coef, double* kval,
                    hipStream_t s, int64_t n_units) {
  if (n_units <= 0 || p.N <= 0) return PSX_OK;
  // four nodes per wave (16 lanes each); groups in multiples of 8 for the
  // XCD-contiguous mapping
  const unsigned blocks = unsigned(8 * ((int64_t(p.N) + 127) / 128));
  if (p.npe == 4)
    hipLaunchKernelGGL((mass_node_kernel<4, 4>), dim3(blocks), dim3(256), 0, s, p, rho, vol, coef, kval);
  else
    hipLaunchKernelGGL((mass_node_kernel<10, 4>), dim3(blocks), dim3(256), 0, s, p, rho, vol, coef, kval);
  PSX_HIP_TRY(hipGetLastError());
  return PSX_OK;
}

}  // namespace psx
