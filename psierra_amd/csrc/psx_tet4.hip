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
#include <algorithm>
#include <type_traits>

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
  // contributions in LDS; reused as the FP64 staging buffer of the tile's
  // CSR values (9 * units <= 9 * TILE doubles fit in 39 * TILE floats).
  __shared__ R contrib[NV * TILE];
  __shared__ uint32_t rk[TILE];
  __shared__ int32_t s_incp[TILE + 1];
  __shared__ int32_t s_uoff[TILE + 1];
  __shared__ int32_t s_kofs[WANT_K ? TILE : 1];
  __shared__ unsigned long long s_mask[WANT_K ? TILE : 1];

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
    if (WANT_K && i < nn) s_kofs[i] = p.kofs[n0 + i];
  }
  if (WANT_K) s_mask[t] = 0ull;
  __syncthreads();
  if (t < nk) {
    const uint32_t pe = p.inc_elem[k0 + t];
    const int e = int(pe & kIncElemMask);
    const int lp = int(pe >> kIncElemBits);
    const int4 c = *reinterpret_cast<const int4*>(p.conn + 4 * int64_t(e));
    const double m = mu[e], l = lam[e];
    R out[NV];
    R J;
    if constexpr (std::is_same<R, float>::value) {
      // FP32: precomputed gradients, displacements in the element's order
      const float4 g0 = p.geo[3 * int64_t(e)], g1 = p.geo[3 * int64_t(e) + 1], g2 = p.geo[3 * int64_t(e) + 2];
      double4 U[4];
      U[0] = p.ui[c.x];
      U[1] = p.ui[c.y];
      U[2] = p.ui[c.z];
      U[3] = p.ui[c.w];
      J = tet4_row_geo<WANT_K>(g0, g1, g2, U, lp, float(m), float(l + (2.0 / 3.0) * m), out);
    } else {
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
      J = tet4_row<R, WANT_K>(X, U, R(m), R(l + (2.0 / 3.0) * m), out);
    }
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
    if (J <= R(0)) flag_inversion(status, e);
#pragma unroll
    for (int v = 0; v < NV; ++v) contrib[v * TILE + t] = out[v];
  }
  __syncthreads();
  if (WANT_K) {
    // Lanes [0, 4 nn): (node, part) -- parts 0..2 sum row `part` of the
    // diagonal block over the node's contiguous incidence run, part 3 sums
    // f.  Lanes [4 nn, 4 nn + units): one (node, slot) unit each; an
    // off-diagonal slot sums the incidences flagged in its mask in
    // increasing order.  Every sum is FP64 in a fixed order.
    const int units = s_uoff[nn];
    const int ni = 4 * nn;
    double acc[9];
#pragma unroll
    for (int c = 0; c < 9; ++c) acc[c] = 0.0;
    int dst = -1, nrow = 0, rstride = 0;
    if (t < ni) {
      const int j = t >> 2, part = t & 3;
      const int kb = s_incp[j], ke = s_incp[j + 1];
      const int src = part < 3 ? 3 * part : 36;
      for (int kk = kb; kk < ke; ++kk)
#pragma unroll
        for (int c = 0; c < 3; ++c) acc[c] += double(contrib[(src + c) * TILE + kk]);
      if (part == 3) {
        const int64_t a = p.perm[n0 + j];
#pragma unroll
        for (int c = 0; c < 3; ++c) f[3 * a + c] = acc[c];
      } else if (ke > kb) {
        const int d = s_uoff[j + 1] - s_uoff[j];
        dst = 9 * s_uoff[j] + part * 3 * d + 3 * int(rk[kb] & 0xffu);
        nrow = 1;
      }
    } else if (t < ni + units) {
      const int u = t - ni;
      int lo = 0, hi = nn;
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (s_uoff[mid] <= u) lo = mid; else hi = mid;
      }
      const uint32_t s = uint32_t(u - s_uoff[lo]);
      const int kb = s_incp[lo];
      if ((rk[kb] & 0xffu) != s) {
        unsigned long long msk = s_mask[u];
        while (msk) {
          const int kk = kb + __ffsll(static_cast<long long>(msk)) - 1;
          msk &= msk - 1ull;
          const uint32_t r = rk[kk];
          const int q = ((r >> 8) & 0xffu) == s ? 1 : ((r >> 16) & 0xffu) == s ? 2 : 3;
#pragma unroll
          for (int c = 0; c < 9; ++c) acc[c] += double(contrib[(9 * q + c) * TILE + kk]);
        }
        const int d = s_uoff[lo + 1] - s_uoff[lo];
        dst = 9 * s_uoff[lo] + 3 * int(s);
        rstride = 3 * d;
        nrow = 3;
      }
    }
    __syncthreads();  // every read of contrib is done: reuse it for staging
    double* stage = reinterpret_cast<double*>(contrib);
    if (nrow == 1) {
#pragma unroll
      for (int c = 0; c < 3; ++c) stage[dst + c] = acc[c];
    } else if (nrow == 3) {
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) stage[dst + r * rstride + c] = acc[3 * r + c];
    }
    __syncthreads();
    // Copy-out: the tile's values, node block by node block, with
    // consecutive lanes on consecutive CSR positions.
    for (int w = t; w < 9 * units; w += TILE) {
      int lo = 0, hi = nn;
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (9 * s_uoff[mid] <= w) lo = mid; else hi = mid;
      }
      kval[s_kofs[lo] + (w - 9 * s_uoff[lo])] = stage[w];
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

// Inputs of one incidence, gathered ahead of use (software pipelining).
template <typename R>
struct IncInputs {
  uint32_t pe = 0, rank = 0;
  int4 c{};
  double m = 0.0, l = 0.0;
  float4 g0{}, g1{}, g2{};   // FP32 path: precomputed gradients
  double4 U[4];
  double4 X[4];               // FP64 path: reference coordinates
};

template <typename R>
__device__ __forceinline__ void load_stage_a(const DevPlan& p, int k, IncInputs<R>& in) {
  in.pe = p.inc_elem[k];
  in.rank = p.inc_rank[k];
}

template <typename R>
__device__ __forceinline__ void load_stage_b(const DevPlan& p, const double* mu, const double* lam,
                                             IncInputs<R>& in) {
  const int64_t e = int64_t(in.pe & kIncElemMask);
  in.c = *reinterpret_cast<const int4*>(p.conn + 4 * e);
  in.m = mu[e];
  in.l = lam[e];
  if constexpr (std::is_same<R, float>::value) {
    in.g0 = p.geo[3 * e];
    in.g1 = p.geo[3 * e + 1];
    in.g2 = p.geo[3 * e + 2];
  }
}

template <typename R>
__device__ __forceinline__ void load_stage_c(const DevPlan& p, IncInputs<R>& in) {
  const int4 c = in.c;
  if constexpr (std::is_same<R, float>::value) {
    in.U[0] = p.ui[c.x];
    in.U[1] = p.ui[c.y];
    in.U[2] = p.ui[c.z];
    in.U[3] = p.ui[c.w];
  } else {
    const int lp = int(in.pe >> kIncElemBits);
    const int i0 = lp == 0 ? c.x : lp == 1 ? c.y : lp == 2 ? c.z : c.w;
    const int i1 = lp == 0 ? c.y : c.x;
    const int i2 = lp <= 1 ? c.z : c.y;
    const int i3 = lp <= 2 ? c.w : c.z;
    in.X[0] = p.Xi[i0];
    in.X[1] = p.Xi[i1];
    in.X[2] = p.Xi[i2];
    in.X[3] = p.Xi[i3];
    in.U[0] = p.ui[i0];
    in.U[1] = p.ui[i1];
    in.U[2] = p.ui[i2];
    in.U[3] = p.ui[i3];
  }
}

// Node-table entries of one tile for one lane (lane j holds node j's
// incidence range, unit range and first CSR value).
struct TileRow {
  int32_t inc0 = 0, inc1 = 0, u0 = 0, u1 = 0, kofs = 0;
};

// Persistent f_int + K_T assembly.  Work-group j of XCD x walks the XCD's
// contiguous tile range with stride (work-groups per XCD).  The gathers are
// software-pipelined three tiles deep: at the top of the iteration for tile
// T_i the work-group issues, all at once and independent of each other, the
// nodal displacements of T_{i+1} (its connectivity arrived one iteration
// earlier), the connectivity / parameters / geometry of T_{i+2} (its
// incidence records arrived one iteration earlier), the incidence records of
// T_{i+3} and the node tables of T_{i+1}; phase 1 of T_i then runs on data
// already in registers while those loads are in flight.
//
// LDS per tile: contributions (39 values per incidence, row stride TILE + 1
// so the four lanes of a node read different banks), the incidence ->
// node and unit -> node maps, and for every (node, slot) unit three 64-bit
// masks, one per block position q = 1..3, of the incidences that feed it.
// The reduction lanes then issue their LDS loads without dependent lookups;
// the tile's CSR values are staged in LDS and written node block by node
// block, one wave per node, 64 consecutive doubles per store.
template <typename R, int TILE>
__global__ __launch_bounds__(TILE) void tet4_assemble_persistent(DevPlan p, const double* __restrict__ mu,
                                                                 const double* __restrict__ lam,
                                                                 double* __restrict__ f,
                                                                 double* __restrict__ kval,
                                                                 int32_t* status) {
  constexpr int NV = 39;
  constexpr int LD = TILE + 1;
  __shared__ R contrib[NV * LD];
  __shared__ uint8_t s_inode[TILE];
  __shared__ uint8_t s_unode[TILE];
  __shared__ int32_t s_incp[TILE / 4 + 1];
  __shared__ int32_t s_uoff[TILE / 4 + 1];
  __shared__ int32_t s_kofs[TILE / 4];
  __shared__ uint8_t s_self[TILE / 4];
  __shared__ unsigned long long s_mask[3 * TILE];

  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const int x = blockIdx.x & 7;
  const int nwg = gridDim.x >> 3;
  const int per = (p.n_tiles + 7) >> 3;
  const int tile_end = min(p.n_tiles, (x + 1) * per);
  const int tile0 = x * per + (blockIdx.x >> 3);
  if (tile0 >= tile_end) return;

  auto tile_range = [&](int tl, int& n0, int& nn, int& k0, int& nk) {
    if (tl < tile_end) {
      n0 = p.tile_node[tl];
      nn = p.tile_node[tl + 1] - n0;
      k0 = p.inc_ptr[n0];
      nk = p.inc_ptr[n0 + nn] - k0;
    } else {
      n0 = nn = k0 = nk = 0;
    }
  };
  auto load_rows = [&](int n0, int nn, TileRow& row) {
    if (t < nn) {
      row.inc0 = p.inc_ptr[n0 + t];
      row.inc1 = p.inc_ptr[n0 + t + 1];
      row.u0 = p.uoff[n0 + t];
      row.u1 = p.uoff[n0 + t + 1];
      row.kofs = p.kofs[n0 + t];
    }
  };

  IncInputs<R> cur, nx1, nx2, nx3;
  TileRow row_cur, row_nx;
  int n0c, nnc, k0c, nkc, n01, nn1, k01, nk1, n02, nn2, k02, nk2;
  tile_range(tile0, n0c, nnc, k0c, nkc);
  tile_range(tile0 + nwg, n01, nn1, k01, nk1);
  tile_range(tile0 + 2 * nwg, n02, nn2, k02, nk2);
  load_rows(n0c, nnc, row_cur);
  if (t < nkc) {
    load_stage_a(p, k0c + t, cur);
    load_stage_b(p, mu, lam, cur);
    load_stage_c(p, cur);
  }
  if (t < nk1) {
    load_stage_a(p, k01 + t, nx1);
    load_stage_b(p, mu, lam, nx1);
  }
  if (t < nk2) load_stage_a(p, k02 + t, nx2);

  for (int tile = tile0; tile < tile_end; tile += nwg) {
    int n03, nn3, k03, nk3;
    tile_range(tile + 3 * nwg, n03, nn3, k03, nk3);
    const int n0 = n0c, nn = nnc, k0 = k0c, nk = nkc;
    // node tables and the incidence -> node / unit -> node maps
    const int uoff0 = p.uoff[n0];
    if (t < nn) {
      s_incp[t] = row_cur.inc0 - k0;
      s_uoff[t] = row_cur.u0 - uoff0;
      s_kofs[t] = row_cur.kofs;
      if (t == nn - 1) {
        s_incp[nn] = row_cur.inc1 - k0;
        s_uoff[nn] = row_cur.u1 - uoff0;
      }
      for (int k = row_cur.inc0 - k0; k < row_cur.inc1 - k0; ++k) s_inode[k] = uint8_t(t);
      for (int u = row_cur.u0 - uoff0; u < row_cur.u1 - uoff0; ++u) s_unode[u] = uint8_t(t);
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) s_mask[q * TILE + t] = 0ull;
    // pipeline: issue every next-tile load before this tile's compute
    if (t < nk1) load_stage_c(p, nx1);
    if (t < nk2) load_stage_b(p, mu, lam, nx2);
    if (t < nk3) load_stage_a(p, k03 + t, nx3);
    load_rows(n01, nn1, row_nx);
    __syncthreads();
    // ---- phase 1: incidence t of this tile
    if (t < nk) {
      const int e = int(cur.pe & kIncElemMask);
      const int lp = int(cur.pe >> kIncElemBits);
      R out[NV];
      R J;
      if constexpr (std::is_same<R, float>::value)
        J = tet4_row_geo<true>(cur.g0, cur.g1, cur.g2, cur.U, lp, float(cur.m),
                               float(cur.l + (2.0 / 3.0) * cur.m), out);
      else
        J = tet4_row<R, true>(cur.X, cur.U, R(cur.m), R(cur.l + (2.0 / 3.0) * cur.m), out);
      if (J <= R(0)) flag_inversion(status, e);
      const uint32_t r = cur.rank;
      const int j = s_inode[t];
      const int kb = s_incp[j];
      if (t == kb) s_self[j] = uint8_t(r & 0xffu);
      const unsigned long long bit = 1ull << (t - kb);
      const int ubase = s_uoff[j];
#pragma unroll
      for (int q = 1; q < 4; ++q) atomicOr(&s_mask[(q - 1) * TILE + ubase + ((r >> (8 * q)) & 0xffu)], bit);
#pragma unroll
      for (int v = 0; v < NV; ++v) contrib[v * LD + t] = out[v];
    }
    __syncthreads();
    // ---- phase 2: FP64 sums in a fixed order
    const int units = s_uoff[nn];
    const int ni = 4 * nn;
    double acc[9];
#pragma unroll
    for (int c = 0; c < 9; ++c) acc[c] = 0.0;
    int dst = -1, nrow = 0, rstride = 0;
    if (t < ni) {
      // (node, part): part 0..2 = row of the diagonal block, 3 = f
      const int j = t >> 2, part = t & 3;
      const int kb = s_incp[j], ke = s_incp[j + 1];
      const R* src = contrib + (part < 3 ? 3 * part : 36) * LD;
      int kk = kb;
      for (; kk + 4 <= ke; kk += 4) {
        R v[4][3];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int c = 0; c < 3; ++c) v[i][c] = src[c * LD + kk + i];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int c = 0; c < 3; ++c) acc[c] += double(v[i][c]);
      }
      for (; kk < ke; ++kk)
#pragma unroll
        for (int c = 0; c < 3; ++c) acc[c] += double(src[c * LD + kk]);
      if (part == 3) {
        const int64_t a = p.perm[n0 + j];
#pragma unroll
        for (int c = 0; c < 3; ++c) f[3 * a + c] = acc[c];
      } else if (ke > kb) {
        const int d = s_uoff[j + 1] - s_uoff[j];
        dst = 9 * s_uoff[j] + part * 3 * d + 3 * int(s_self[j]);
        nrow = 1;
      }
    } else if (t < ni + units) {
      const int u = t - ni;
      const int j = s_unode[u];
      const uint32_t s = uint32_t(u - s_uoff[j]);
      if (s != s_self[j]) {
        const int kb = s_incp[j];
#pragma unroll
        for (int q = 1; q < 4; ++q) {
          unsigned long long msk = s_mask[(q - 1) * TILE + u];
          const R* src = contrib + 9 * q * LD + kb;
          while (msk) {
            const int a = __ffsll(static_cast<long long>(msk)) - 1;
            msk &= msk - 1ull;
            const int b = msk ? __ffsll(static_cast<long long>(msk)) - 1 : a;
            const bool two = msk != 0ull;
            msk &= msk - 1ull;
            R va[9], vb[9];
#pragma unroll
            for (int c = 0; c < 9; ++c) {
              va[c] = src[c * LD + a];
              vb[c] = src[c * LD + b];
            }
#pragma unroll
            for (int c = 0; c < 9; ++c) acc[c] += double(va[c]);
            if (two)
#pragma unroll
              for (int c = 0; c < 9; ++c) acc[c] += double(vb[c]);
          }
        }
        const int d = s_uoff[j + 1] - s_uoff[j];
        dst = 9 * s_uoff[j] + 3 * int(s);
        rstride = 3 * d;
        nrow = 3;
      }
    }
    __syncthreads();
    double* stage = reinterpret_cast<double*>(contrib);
    if (nrow == 1) {
#pragma unroll
      for (int c = 0; c < 3; ++c) stage[dst + c] = acc[c];
    } else if (nrow == 3) {
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) stage[dst + r * rstride + c] = acc[3 * r + c];
    }
    __syncthreads();
    // copy-out: one wave per node block, consecutive lanes on consecutive values
    for (int j = wave; j < nn; j += TILE / 64) {
      const int base = 9 * s_uoff[j], len = 9 * (s_uoff[j + 1] - s_uoff[j]);
      double* dstp = kval + s_kofs[j];
      for (int i = lane; i < len; i += 64) dstp[i] = stage[base + i];
    }
    // rotate the pipeline
    cur = nx1;
    nx1 = nx2;
    nx2 = nx3;
    row_cur = row_nx;
    n0c = n01; nnc = nn1; k0c = k01; nkc = nk1;
    n01 = n02; nn1 = nn2; k01 = k02; nk1 = nk2;
    n02 = n03; nn2 = nn3; k02 = k03; nk2 = nk3;
    __syncthreads();
  }
}

template <typename R, int TILE>
static int persistent_grid(const DevPlan& p) {
  static int per_cu = 0;
  static int cus = 0;
  if (per_cu == 0) {
    int dev = 0, n = 0;
    (void)hipGetDevice(&dev);
    hipDeviceProp_t prop;
    cus = hipGetDeviceProperties(&prop, dev) == hipSuccess ? prop.multiProcessorCount : 256;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, tet4_assemble_persistent<R, TILE>, TILE, 0) !=
            hipSuccess ||
        n <= 0)
      n = 1;
    per_cu = n;
  }
  const int per_xcd_tiles = (p.n_tiles + 7) / 8;
  const int wg_per_xcd = std::max(1, std::min(per_xcd_tiles, per_cu * cus / 8));
  return 8 * wg_per_xcd;
}

template <typename R, int TILE>
static void launch_t4(const DevPlan& p, const double* mu, const double* lam, double* f, double* kval,
                      hipStream_t s, int32_t* status) {
  const int blocks = 8 * ((p.n_tiles + 7) / 8);
  if (kval)
    hipLaunchKernelGGL((tet4_assemble_persistent<R, TILE>), dim3(persistent_grid<R, TILE>(p)), dim3(TILE),
                       0, s, p, mu, lam, f, kval, status);
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
