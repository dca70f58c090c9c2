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
#include <type_traits>

#include "psx_device.h"
#include "psx_internal.h"
#include "psx_nh.h"
#include "psx_batch.h"

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

// d N_a / d xi_j at the point with barycentric coordinates L (L[0] =
// 1 - L[1] - L[2] - L[3]): the expressions of tet10_dn with the point given
// at run time.
__device__ __forceinline__ double tet10_dn_at(const double L[4], int a, int j) {
  switch (a) {
    case 0: return -(4.0 * L[0] - 1.0);
    case 1: return j == 0 ? 4.0 * L[1] - 1.0 : 0.0;
    case 2: return j == 1 ? 4.0 * L[2] - 1.0 : 0.0;
    case 3: return j == 2 ? 4.0 * L[3] - 1.0 : 0.0;
    case 4: return j == 0 ? 4.0 * (L[0] - L[1]) : -4.0 * L[1];
    case 5: return j == 0 ? 4.0 * L[2] : j == 1 ? 4.0 * L[1] : 0.0;
    case 6: return j == 1 ? 4.0 * (L[0] - L[2]) : -4.0 * L[2];
    case 7: return j == 2 ? 4.0 * (L[0] - L[3]) : -4.0 * L[3];
    case 8: return j == 0 ? 4.0 * L[3] : j == 2 ? 4.0 * L[1] : 0.0;
    default: return j == 1 ? 4.0 * L[3] : j == 2 ? 4.0 * L[2] : 0.0;
  }
}

// Barycentric coordinates of quadrature point q (0..3).
__device__ __forceinline__ void tet10_qp_coords(int q, double L[4]) {
  L[1] = q == 1 ? kQb : kQa;
  L[2] = q == 2 ? kQb : kQa;
  L[3] = q == 3 ? kQb : kQa;
  L[0] = 1.0 - L[1] - L[2] - L[3];
}

// Row p of a Tet10 element with material model M (prm as load_params
// gives them) at one quadrature point L.  dX, dU: node coordinates /
// displacements minus node 0's (index 0 is zero).  sink(v, value) receives
// the 90 block values (v = 9 b + 3 i + k) then the 3 force values
// (v = 90 + i) of the point.  Returns det F at the point.
template <typename R, int M, bool WANT_K, class Sink>
__device__ __forceinline__ R tet10_qp(const R dX[10][3], const R dU[10][3], const R prm[4], int p,
                                      const double L[4], Sink&& sink) {
  R Jm[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      R s = R(0);
#pragma unroll
      for (int a = 1; a < 10; ++a) s += dX[a][i] * R(tet10_dn_at(L, a, j));
      Jm[i][j] = s;
    }
  R adj[3][3];
  adj3(Jm, adj);
  const R det = Jm[0][0] * adj[0][0] + Jm[0][1] * adj[1][0] + Jm[0][2] * adj[2][0];
  const R idet = R(1) / det;
  const R w = det * R(1.0 / 24.0);
  // h_a[j] = sum_k dN(a,k) Jinv[k][j]; A = sum_a dU_a (x) h_a
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
      h[j] = (R(tet10_dn_at(L, a, 0)) * adj[0][j] + R(tet10_dn_at(L, a, 1)) * adj[1][j] +
              R(tet10_dn_at(L, a, 2)) * adj[2][j]) * idet;
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
  HyperExtra<R> x;
  hyper_point<R, M>(A, prm, m, x);
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
        h[j] = (R(tet10_dn_at(L, b, 0)) * adj[0][j] + R(tet10_dn_at(L, b, 1)) * adj[1][j] +
                R(tet10_dn_at(L, b, 2)) * adj[2][j]) * idet;
      mat3vec(m.Gm, h, gb);
      mat3vec(m.F, h, qb);
      hyper_block<R, M>(m, x, w, gp, qp, hp, Shp, gb, qb, h, blk);
#pragma unroll
      for (int c = 0; c < 9; ++c) sink(9 * b + c, blk[c]);
    }
  }
  return m.J;
}

// Row p summed over the four quadrature points (sink receives each point's
// values in turn; the caller sums them).  Returns min over points of det F.
template <typename R, int M, bool WANT_K, class Sink>
__device__ __forceinline__ R tet10_row(const R dX[10][3], const R dU[10][3], const R prm[4], int p,
                                       Sink&& sink) {
  R Jmin = R(1);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    double L[4];
    tet10_qp_coords(q, L);
    Jmin = fmin(Jmin, tet10_qp<R, M, WANT_K>(dX, dU, prm, p, L, sink));
  }
  return Jmin;
}

// Gathered inputs of one Tet10 incidence: node coordinates and displacements
// relative to node 0 (differences in FP64, then rounded to R).
template <typename R, typename P>
__device__ __forceinline__ void tet10_gather(const P& p, int e, R dX[10][3], R dU[10][3]) {
  const int32_t* c = p.conn + 12 * int64_t(e);
  const int4 c0 = *reinterpret_cast<const int4*>(c);
  const int4 c1 = *reinterpret_cast<const int4*>(c + 4);
  const int2 c2 = *reinterpret_cast<const int2*>(c + 8);
  const int ids[10] = {c0.x, c0.y, c0.z, c0.w, c1.x, c1.y, c1.z, c1.w, c2.x, c2.y};
  const double4 X0 = p.Xi[ids[0]];
  const double4 U0 = p.ui[ids[0]];
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
}

// DPP broadcast of lane SRC of each quad to the quad's four lanes.
template <int SRC>
__device__ __forceinline__ float quad_bcast(float x) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), SRC * 0x55, 0xF, 0xF, false));
}
template <int SRC>
__device__ __forceinline__ double quad_bcast(double x) {
  const unsigned long long v = static_cast<unsigned long long>(__double_as_longlong(x));
  const unsigned lo = unsigned(__builtin_amdgcn_mov_dpp(int(unsigned(v)), SRC * 0x55, 0xF, 0xF, false));
  const unsigned hi = unsigned(__builtin_amdgcn_mov_dpp(int(unsigned(v >> 32)), SRC * 0x55, 0xF, 0xF, false));
  return __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(hi) << 32) | lo));
}

// tet10_gather for the quad of lanes that shares one incidence (all four
// lanes active): lane qp loads nodes qp, qp + 4, qp + 8 (besides node 0),
// forms their differences, and the quad exchanges them.
template <typename R, typename P>
__device__ __forceinline__ void tet10_gather_quad_ids(const P& p, int4 c0, int4 c1, int2 c2, int qp, R dX[10][3],
                                                      R dU[10][3]);

template <typename R, typename P>
__device__ __forceinline__ void tet10_gather_quad(const P& p, int e, int qp, R dX[10][3], R dU[10][3]) {
  const int32_t* c = p.conn + 12 * int64_t(e);
  const int4 c0 = *reinterpret_cast<const int4*>(c);
  const int4 c1 = *reinterpret_cast<const int4*>(c + 4);
  const int2 c2 = *reinterpret_cast<const int2*>(c + 8);
  tet10_gather_quad_ids<R>(p, c0, c1, c2, qp, dX, dU);
}

// The same from the element's node ids (loaded by the caller).
template <typename R, typename P>
__device__ __forceinline__ void tet10_gather_quad_ids(const P& p, int4 c0, int4 c1, int2 c2, int qp, R dX[10][3],
                                                      R dU[10][3]) {
  const int ids[10] = {c0.x, c0.y, c0.z, c0.w, c1.x, c1.y, c1.z, c1.w, c2.x, c2.y};
  const double4 X0 = p.Xi[ids[0]];
  const double4 U0 = p.ui[ids[0]];
  R mine[3][6];
#pragma unroll
  for (int s = 0; s < 3; ++s) {
    // node qp + 4 s of this lane (node 0 and nodes past 9 give zeros)
    const int a = qp + 4 * s;
    const int id = s == 0 ? (qp == 1 ? ids[1] : qp == 2 ? ids[2] : ids[3])
                          : s == 1 ? (qp == 0 ? ids[4] : qp == 1 ? ids[5] : qp == 2 ? ids[6] : ids[7])
                                   : (qp == 0 ? ids[8] : ids[9]);
    if (a >= 1 && a < 10) {
      const double4 Xa = p.Xi[id];
      const double4 Ua = p.ui[id];
      mine[s][0] = R(Xa.x - X0.x);
      mine[s][1] = R(Xa.y - X0.y);
      mine[s][2] = R(Xa.z - X0.z);
      mine[s][3] = R(Ua.x - U0.x);
      mine[s][4] = R(Ua.y - U0.y);
      mine[s][5] = R(Ua.z - U0.z);
    } else {
#pragma unroll
      for (int k = 0; k < 6; ++k) mine[s][k] = R(0);
    }
  }
  dX[0][0] = dX[0][1] = dX[0][2] = R(0);
  dU[0][0] = dU[0][1] = dU[0][2] = R(0);
  auto take = [&](auto src, int a, const R* v) {
    constexpr int S = decltype(src)::value;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      dX[a][k] = quad_bcast<S>(v[k]);
      dU[a][k] = quad_bcast<S>(v[3 + k]);
    }
  };
  using L0 = std::integral_constant<int, 0>;
  using L1 = std::integral_constant<int, 1>;
  using L2 = std::integral_constant<int, 2>;
  using L3 = std::integral_constant<int, 3>;
  take(L1{}, 1, mine[0]);
  take(L2{}, 2, mine[0]);
  take(L3{}, 3, mine[0]);
  take(L0{}, 4, mine[1]);
  take(L1{}, 5, mine[1]);
  take(L2{}, 6, mine[1]);
  take(L3{}, 7, mine[1]);
  take(L0{}, 8, mine[2]);
  take(L1{}, 9, mine[2]);
}

// f_int only, one lane per incidence over the four quadrature points (the
// FP32 form: its registers and LDS leave room for enough waves), f_a summed
// in FP64 over the node's run.
template <typename R, int TILE, int M, typename P>
__device__ __forceinline__ void tet10_force_tile(const P& p, const MatParams& mp, double* __restrict__ f,
                                                 int32_t* status, int tile, int t) {
  __shared__ R contrib[3 * TILE];
  __shared__ int32_t s_incp[TILE + 1];
  const int4 ti = load_uniform(p.tile_info, tile);
  const int n0 = ti.x, nn = ti.y, k0 = ti.z, nk = ti.w;
  for (int i = t; i <= nn; i += TILE) s_incp[i] = p.node_rec[n0 + i].x - k0;
  if (t < nk) {
    const uint32_t pe = p.inc_elem[k0 + t];
    const int e = int(pe & kIncElemMask);
    const int lp = int(pe >> kIncElemBits);
    R dX[10][3], dU[10][3];
    tet10_gather<R>(p, e, dX, dU);
    double pr[4];
    load_params<M>(mp, e, pr);
    const R prm[4] = {R(pr[0]), R(pr[1]), R(pr[2]), R(pr[3])};
    R fa[3] = {R(0), R(0), R(0)};
    const R J = tet10_row<R, M, false>(dX, dU, prm, lp, [&](int v, R val) { fa[v - 90] += val; });
    if (J <= R(0)) {
      atomicMin(&status[1], e);
      status[0] = 1;
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) contrib[c * TILE + t] = fa[c];
  }
  __syncthreads();
  for (int j = t; j < nn; j += TILE) {
    double fs[3] = {0.0, 0.0, 0.0};
    for (int kk = s_incp[j]; kk < s_incp[j + 1]; ++kk)
#pragma unroll
      for (int c = 0; c < 3; ++c) fs[c] += double(contrib[c * TILE + kk]);
    const int64_t a = p.perm[n0 + j];
#pragma unroll
    for (int c = 0; c < 3; ++c) f[3 * a + c] = fs[c];
  }
}

template <typename R, int TILE, int M>
__global__ __launch_bounds__(TILE) void tet10_force_kernel(DevPlan p, MatParams mp, double* __restrict__ f,
                                                           int32_t* status) {
  const int tile = xcd_tile();
  if (tile >= p.n_tiles) return;
  tet10_force_tile<R, TILE, M>(p, mp, f, status, tile, int(threadIdx.x));
}

// f_int only, four lanes per incidence (one per quadrature point, the quad
// summing f_a with DPP; the FP64 form, whose one-lane version is held to few
// waves by its registers), f_a summed in FP64 over the node's run by one lane
// per (node, component).
template <typename R, int TILE, int M, typename P>
__device__ __forceinline__ void tet10_force_quad_tile(const P& p, const MatParams& mp, double* __restrict__ f,
                                                      int32_t* status, int tile, int t) {
  constexpr int NT = 4 * TILE;
  __shared__ R contrib[3 * TILE];
  __shared__ int32_t s_incp[TILE + 1];
  const int4 ti = load_uniform(p.tile_info, tile);
  const int n0 = ti.x, nn = ti.y, k0 = ti.z, nk = ti.w;
  for (int i = t; i <= nn; i += NT) s_incp[i] = p.node_rec[n0 + i].x - k0;
  const int it = t >> 2, qp = t & 3;
  if (it < nk) {   // the same for the four lanes of a quad
    const uint32_t pe = p.inc_elem[k0 + it];
    const int e = int(pe & kIncElemMask);
    const int lp = int(pe >> kIncElemBits);
    R dX[10][3], dU[10][3];
    tet10_gather_quad<R>(p, e, qp, dX, dU);
    double pr[4];
    load_params<M>(mp, e, pr);
    const R prm[4] = {R(pr[0]), R(pr[1]), R(pr[2]), R(pr[3])};
    double L[4];
    tet10_qp_coords(qp, L);
    R fa[3] = {R(0), R(0), R(0)};
    const R J = tet10_qp<R, M, false>(dX, dU, prm, lp, L, [&](int v, R val) { fa[v - 90] = val; });
    if (J <= R(0)) {
      atomicMin(&status[1], e);
      status[0] = 1;
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) contrib[c * TILE + it] = quad_sum(fa[c]);
  }
  __syncthreads();
  for (int jc = t; jc < 3 * nn; jc += NT) {
    const int j = jc / 3, c = jc - 3 * j;
    double fs = 0.0;
    for (int kk = s_incp[j]; kk < s_incp[j + 1]; ++kk) fs += double(contrib[c * TILE + kk]);
    f[3 * int64_t(p.perm[n0 + j]) + c] = fs;
  }
}

template <typename R, int TILE, int M>
__global__ __launch_bounds__(4 * TILE) void tet10_force_quad_kernel(DevPlan p, MatParams mp, double* __restrict__ f,
                                                                    int32_t* status) {
  const int tile = xcd_tile();
  if (tile >= p.n_tiles) return;
  tet10_force_quad_tile<R, TILE, M>(p, mp, f, status, tile, int(threadIdx.x));
}

// Element position q (0..9) of the node whose slot rank in the row is s,
// from the three rank words of an incidence (bytes 0..9).
__device__ __forceinline__ int tet10_rank_pos(uint32_t w0, uint32_t w1, uint32_t w2, uint32_t s) {
  const uint32_t s4 = s * 0x01010101u;
  // lowest zero byte of x: flagged exactly (higher flags may be spurious)
  auto zero_bytes = [](uint32_t x) { return (x - 0x01010101u) & ~x & 0x80808080u; };
  const uint32_t z0 = zero_bytes(w0 ^ s4);
  const uint32_t z1 = zero_bytes(w1 ^ s4);
  const uint32_t z2 = zero_bytes(w2 ^ s4) & 0x8080u;
  return z0 ? (__ffs(int(z0)) - 1) >> 3 : z1 ? 4 + ((__ffs(int(z1)) - 1) >> 3) : 8 + ((__ffs(int(z2)) - 1) >> 3);
}

// f_int + K_T over one tile of node rows (same scheme as the Tet4 kernel):
// phase 1, lane t = incidence t: the row's 10 blocks and f summed over the
// 4 quadrature points in registers, then 93 values to LDS, and the
// incidence's bit set in the feed mask of each of the 10 units it feeds;
// phase 2, reduction lanes -- four per node (quarters of the node's run:
// K_aa and f_a, combined across the quad in a fixed order) and one per
// off-diagonal (node, slot) unit -- sum their feeds in FP64 in increasing
// incidence order, passing over the lanes up to kTet10LanePasses times, and
// write each 3x3 block once, three consecutive values per row.  A tile has
// at most 64 nodes (the plan's rule), so the node lanes are one wave.
// f_int + K_T: four threads per incidence of the tile, one per quadrature
// point; each value of the row is summed over the four points across the
// quad ((q0 + q1) + (q2 + q3), DPP) and written to LDS once, so the LDS per
// incidence (93 values) is shared by four waves' worth of threads.
template <typename R, int TILE, int M, typename P>
__device__ __forceinline__ void tet10_rows_tile(const P& p, const MatParams& mp, double* __restrict__ f,
                                                double* __restrict__ kval, int32_t* status, int tile, int t) {
  constexpr int NV = 93;
  constexpr int NT = 4 * TILE;                     // threads
  constexpr int UMAX = kTet10LanePasses * TILE;   // reduction lanes per tile
  constexpr int NN = 64;
  constexpr int NW = UMAX / 64;
  __shared__ R contrib[NV * TILE];
  __shared__ uint32_t rk[3 * TILE];
  __shared__ unsigned long long s_mask[UMAX];
  __shared__ unsigned long long s_ustart[NW];
  __shared__ int32_t s_incp[NN + 1];
  __shared__ int32_t s_uoff[NN + 1];
  __shared__ int32_t s_kofs[NN];
  __shared__ uint8_t s_self[NN];
  __shared__ uint8_t s_nz[NN];
  __shared__ uint8_t s_first[TILE];

  const int4 ti = load_uniform(p.tile_info, tile);
  const int n0 = ti.x, nn = ti.y, k0 = ti.z, nk = ti.w;
  const int lane = t & 63;

  // the incidence's records and its element's node ids, loaded before the
  // node table so their latency overlaps it and the barrier
  const int it = t >> 2, qp = t & 3;
  uint32_t pe = 0u;
  int4 c0 = make_int4(0, 0, 0, 0), c1 = make_int4(0, 0, 0, 0);
  int2 c2 = make_int2(0, 0);
  uint32_t w[3] = {0u, 0u, 0u};
  double pr[4] = {0.0, 0.0, 0.0, 0.0};
  if (it < nk) {
    pe = p.inc_elem[k0 + it];
    if (qp == 0) {
#pragma unroll
      for (int i = 0; i < 3; ++i) w[i] = p.inc_rank[3 * int64_t(k0 + it) + i];
    }
    const int e = int(pe & kIncElemMask);
    const int32_t* c = p.conn + 12 * int64_t(e);
    c0 = *reinterpret_cast<const int4*>(c);
    c1 = *reinterpret_cast<const int4*>(c + 4);
    c2 = *reinterpret_cast<const int2*>(c + 8);
    load_params<M>(mp, e, pr);
  }

  int ub = 0, ue = 0;
  if (t < nn) {
    const int4 r0 = p.node_rec[n0 + t], r1 = p.node_rec[n0 + t + 1];
    const int kb = r0.x - k0, ke = r1.x - k0;
    s_incp[t] = kb;
    s_uoff[t] = r0.y;
    s_kofs[t] = r0.z;
    s_self[t] = uint8_t(r0.w);
    if (t == nn - 1) {
      s_incp[nn] = ke;
      s_uoff[nn] = r1.y;
    }
    if (ke > kb) s_first[kb] = uint8_t(t);
    ub = r0.y;
    ue = r1.y;
  }
  if (t < 64) {
    const bool has = t < nn && ue > ub;
    const unsigned long long bal = __ballot(has);
    if (has) s_nz[__popcll(bal & ((1ull << lane) - 1ull))] = uint8_t(t);
  }
  for (int i = t; i < NW; i += NT) s_ustart[i] = 0ull;
  for (int i = t; i < UMAX; i += NT) s_mask[i] = 0ull;
  __syncthreads();

  // ---- phase 1: incidence it, quadrature point qp (a quad of lanes per incidence)
  const int u0 = s_uoff[0];
  if (t < nn && ue > ub) atomicOr(&s_ustart[(ub - u0) >> 6], 1ull << ((ub - u0) & 63));
  if (it < nk) {   // the same for the four lanes of a quad
    const int e = int(pe & kIncElemMask);
    const int lp = int(pe >> kIncElemBits);
    R dX[10][3], dU[10][3];
    tet10_gather_quad_ids<R>(p, c0, c1, c2, qp, dX, dU);
    const R prm[4] = {R(pr[0]), R(pr[1]), R(pr[2]), R(pr[3])};
    double L[4];
    tet10_qp_coords(qp, L);
    // the quad's four lanes hold the same sum and store it to the same place
    // (no per-value branch)
    const R J = tet10_qp<R, M, true>(dX, dU, prm, lp, L,
                                     [&](int v, R val) { contrib[v * TILE + it] = quad_sum(val); });
    if (J <= R(0)) {
      atomicMin(&status[1], e);
      status[0] = 1;
    }
    if (qp == 0) {
#pragma unroll
      for (int i = 0; i < 3; ++i) rk[i * TILE + it] = w[i];
      const int off = int((w[2] >> 16) & 0xffu);
      const int j = s_first[it - off];
      const int ubase = s_uoff[j] - u0;
      const unsigned long long bit = 1ull << off;
#pragma unroll
      for (int q = 0; q < 10; ++q)
        if (q != lp) atomicOr(&s_mask[ubase + int((w[q >> 2] >> (8 * (q & 3))) & 0xffu)], bit);
    }
  }
  __syncthreads();

  // ---- phase 2
  const int ni = 4 * nn;
  const int lanes = ni + s_uoff[nn] - u0;
  for (int base = 0; base < lanes; base += NT) {
    const int w = base + t;
    int j = 0, kb = 0, slot = 0;
    bool diag = false, offd = false;
    unsigned long long msk = 0ull;
    if (w < ni) {
      j = w >> 2;
      const int qq = w & 3;
      kb = s_incp[j];
      const int L = s_incp[j + 1] - kb;
      msk = bit_range((qq * L) >> 2, ((qq + 1) * L) >> 2);   // <= 16 bits
      diag = true;
    } else if (w < lanes) {
      const int u = w - ni;
      j = unit_owner<NW>(s_ustart, s_nz, u);
      kb = s_incp[j];
      slot = u - (s_uoff[j] - u0);
      if (slot != int(s_self[j])) {
        offd = true;
        msk = s_mask[u];
      }
    }
    const uint32_t s = diag ? uint32_t(s_self[j]) : uint32_t(slot);
    double acc[12];
#pragma unroll
    for (int c = 0; c < 12; ++c) acc[c] = 0.0;
    while (msk) {
      const int a = __ffsll(static_cast<long long>(msk)) - 1;
      msk &= msk - 1ull;
      const int k = kb + a;
      const int q = tet10_rank_pos(rk[k], rk[TILE + k], rk[2 * TILE + k], s);
      const R* src = contrib + 9 * q * TILE + k;
      R v[9];
#pragma unroll
      for (int c = 0; c < 9; ++c) v[c] = src[c * TILE];
#pragma unroll
      for (int c = 0; c < 9; ++c) acc[c] += double(v[c]);
      if (diag) {
#pragma unroll
        for (int c = 0; c < 3; ++c) acc[9 + c] += double(contrib[(90 + c) * TILE + k]);
      }
    }
    const int d = s_uoff[j + 1] - s_uoff[j];
    if (diag) {
#pragma unroll
      for (int c = 0; c < 12; ++c) acc[c] = quad_sum(acc[c]);
      const int qq = w & 3;
      if (qq == 3) {
        const int64_t a = p.perm[n0 + j];
#pragma unroll
        for (int c = 0; c < 3; ++c) f[3 * a + c] = acc[9 + c];
      } else if (d > 0) {
        double* dst = kval + s_kofs[j] + qq * 3 * d + 3 * int(s);
        const double e0 = qq == 0 ? acc[0] : qq == 1 ? acc[3] : acc[6];
        const double e1 = qq == 0 ? acc[1] : qq == 1 ? acc[4] : acc[7];
        const double e2 = qq == 0 ? acc[2] : qq == 1 ? acc[5] : acc[8];
        dst[0] = e0;
        dst[1] = e1;
        dst[2] = e2;
      }
    } else if (offd) {
      double* dst = kval + s_kofs[j] + 3 * slot;
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) dst[r * 3 * d + c] = acc[3 * r + c];
    }
  }
}

template <typename R, int TILE, int M>
__global__ __launch_bounds__(4 * TILE) void tet10_assemble_rows(DevPlan p, MatParams mp, double* __restrict__ f,
                                                                double* __restrict__ kval, int32_t* status) {
  const int tile = xcd_tile();
  if (tile >= p.n_tiles) return;
  tet10_rows_tile<R, TILE, M>(p, mp, f, kval, status, tile, int(threadIdx.x));
}


template <typename R, int TILE, int M>
static void launch_t10(const DevPlan& p, const MatParams& mp, double* f, double* kval, hipStream_t s,
                       int32_t* status) {
  const int blocks = 8 * ((p.n_tiles + 7) / 8);
  if (kval)
    hipLaunchKernelGGL((tet10_assemble_rows<R, TILE, M>), dim3(blocks), dim3(4 * TILE), 0, s, p, mp, f, kval,
                       status);
  else if constexpr (std::is_same<R, double>::value)
    hipLaunchKernelGGL((tet10_force_quad_kernel<R, TILE, M>), dim3(blocks), dim3(4 * TILE), 0, s, p, mp, f,
                       status);
  else
    hipLaunchKernelGGL((tet10_force_kernel<R, TILE, M>), dim3(blocks), dim3(TILE), 0, s, p, mp, f, status);
}

template <int M>
static int32_t launch_t10_tiles(const DevPlan& p, const MatParams& mp, double* f, double* kval, bool fp32,
                                hipStream_t s, int32_t* status) {
  switch (p.tile) {
    case 64:
      fp32 ? launch_t10<float, 64, M>(p, mp, f, kval, s, status) : launch_t10<double, 64, M>(p, mp, f, kval, s, status);
      break;
    case 128:
      fp32 ? launch_t10<float, 128, M>(p, mp, f, kval, s, status)
           : launch_t10<double, 128, M>(p, mp, f, kval, s, status);
      break;
    default:
      return fail(PSX_EINVAL, "Tet10 supports tiles of 64 or 128 incidences");
  }
  return PSX_OK;
}

// ---------------------------------------------------------------------------
// Batched Tet10 assembly: the tiles of many plans in one launch (the scheme
// of the Tet4 batch, psx_batch.h), each tile by the same tile routine as a
// single-plan launch, so every instance's f and K are the single-plan bits.
// ---------------------------------------------------------------------------
template <typename R, int TILE, int M, bool PERSISTENT>
__global__ __launch_bounds__(4 * TILE) void tet10_batch_rows(BatchDev b, const double* const* mu,
                                                             const double* const* lam, double* const* f,
                                                             double* const* kval, int32_t* status) {
  batch_tiles<PERSISTENT>(b, [&](int inst, int tile, int t) {
    const ConstDevPlan& p = *(const ConstDevPlan*)(b.plans + inst);
    const MatParams mp{{load_uniform_val(mu + inst), load_uniform_val(lam + inst), nullptr, nullptr}};
    tet10_rows_tile<R, TILE, M>(p, mp, load_uniform_val(f + inst), load_uniform_val(kval + inst),
                                status + 2 * inst, tile, t);
  });
}

// f only: one lane per incidence in FP32, a quad per incidence in FP64 (as
// the single-plan launches).
template <typename R, int TILE, int M, bool PERSISTENT>
__global__ __launch_bounds__((std::is_same<R, double>::value ? 4 * TILE : TILE)) void tet10_batch_force(
    BatchDev b, const double* const* mu, const double* const* lam, double* const* f, int32_t* status) {
  batch_tiles<PERSISTENT>(b, [&](int inst, int tile, int t) {
    const ConstDevPlan& p = *(const ConstDevPlan*)(b.plans + inst);
    const MatParams mp{{load_uniform_val(mu + inst), load_uniform_val(lam + inst), nullptr, nullptr}};
    if constexpr (std::is_same<R, double>::value)
      tet10_force_quad_tile<R, TILE, M>(p, mp, load_uniform_val(f + inst), status + 2 * inst, tile, t);
    else
      tet10_force_tile<R, TILE, M>(p, mp, load_uniform_val(f + inst), status + 2 * inst, tile, t);
  });
}

template <typename R, int TILE, bool PERSISTENT>
static void launch_batch_t10(const BatchDev& b, const double* const* mu, const double* const* lam,
                             double* const* f, double* const* kval, int cus, hipStream_t s, int32_t* status) {
  if (kval) {
    const auto k = tet10_batch_rows<R, TILE, kNeoHookean, PERSISTENT>;
    const int grid = batch_grid(reinterpret_cast<const void*>(k), 4 * TILE, cus, b.n_tiles, PERSISTENT);
    hipLaunchKernelGGL(k, dim3(grid), dim3(4 * TILE), 0, s, b, mu, lam, f, kval, status);
  } else {
    constexpr int threads = std::is_same<R, double>::value ? 4 * TILE : TILE;
    const auto k = tet10_batch_force<R, TILE, kNeoHookean, PERSISTENT>;
    const int grid = batch_grid(reinterpret_cast<const void*>(k), threads, cus, b.n_tiles, PERSISTENT);
    hipLaunchKernelGGL(k, dim3(grid), dim3(threads), 0, s, b, mu, lam, f, status);
  }
}

// Neo-Hookean with per-call (mu, lambda) (Tet10 plans bind no material).
int32_t launch_tet10_batch(const BatchDev& b, const double* const* u, const double* const* mu,
                           const double* const* lam, double* const* f, double* const* kval, int precision,
                           int cus, bool persistent, hipStream_t s, int32_t* status) {
  if (!mu || !lam) return fail(PSX_EINVAL, "a batch of Tet10 plans takes mu and lam (Tet10 plans bind no material)");
  int32_t rc = launch_batch_permute(b, u, s, status);
  if (rc) return rc;
  if (b.n_tiles <= 0) return PSX_OK;
  const bool fp32 = precision == PSX_FP32;
  switch (b.tile * 4 + 2 * int(fp32) + int(persistent)) {
    case 256: launch_batch_t10<double, 64, false>(b, mu, lam, f, kval, cus, s, status); break;
    case 257: launch_batch_t10<double, 64, true>(b, mu, lam, f, kval, cus, s, status); break;
    case 258: launch_batch_t10<float, 64, false>(b, mu, lam, f, kval, cus, s, status); break;
    case 259: launch_batch_t10<float, 64, true>(b, mu, lam, f, kval, cus, s, status); break;
    case 512: launch_batch_t10<double, 128, false>(b, mu, lam, f, kval, cus, s, status); break;
    case 513: launch_batch_t10<double, 128, true>(b, mu, lam, f, kval, cus, s, status); break;
    case 514: launch_batch_t10<float, 128, false>(b, mu, lam, f, kval, cus, s, status); break;
    case 515: launch_batch_t10<float, 128, true>(b, mu, lam, f, kval, cus, s, status); break;
    default: return fail(PSX_EINVAL, "Tet10 supports tiles of 64 or 128 incidences");
  }
  PSX_HIP_TRY(hipGetLastError());
  return PSX_OK;
}

// ---------------------------------------------------------------------------
// General path of a wide plan (a node with more than 64 incident elements or
// more than 255 neighbours, beyond the feed masks and rank bytes of
// tet10_assemble_rows): one lane per (element, row node) evaluates the row
// over the four quadrature points (FP64 sums of the points) into the plan's
// element scratch, then one lane per CSR value (and per f entry) adds its
// contributions in increasing element order, from a list the plan built.
// Deterministic, no atomics, no limit on incidences or neighbours.
// ---------------------------------------------------------------------------
template <typename R, int M, bool WANT_K>
__global__ __launch_bounds__(64) void tet10_elements_kernel(DevPlan p, MatParams mp, int32_t* status) {
  const int64_t w = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (w >= int64_t(p.E) * 10) return;
  const int e = int(w / 10), pr = int(w % 10);
  R dX[10][3], dU[10][3];
  tet10_gather<R>(p, e, dX, dU);
  double pd[4];
  load_params<M>(mp, e, pd);
  const R prm[4] = {R(pd[0]), R(pd[1]), R(pd[2]), R(pd[3])};
  double acc[93];
#pragma unroll
  for (int v = 0; v < 93; ++v) acc[v] = 0.0;
  const R J = tet10_row<R, M, WANT_K>(dX, dU, prm, pr, [&](int v, R val) { acc[v] += double(val); });
  if (J <= R(0)) {
    atomicMin(&status[1], e);
    status[0] = 1;
  }
  if (WANT_K) {
    double* ko = p.gen_ke + 900 * int64_t(e) + 90 * pr;   // rows 3 pr .. 3 pr + 2 of the 30 x 30 tangent
#pragma unroll
    for (int b = 0; b < 10; ++b)
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int c = 0; c < 3; ++c) ko[30 * i + 3 * b + c] = acc[9 * b + 3 * i + c];
  }
  double* fo = p.gen_fe + 30 * int64_t(e) + 3 * pr;
#pragma unroll
  for (int i = 0; i < 3; ++i) fo[i] = acc[90 + i];
}

// out[v] = sum of src[idx[k]] for k in ptr[v] .. ptr[v + 1], in order.
__global__ __launch_bounds__(256) void segment_sum_kernel(const int32_t* __restrict__ ptr,
                                                          const int32_t* __restrict__ idx,
                                                          const double* __restrict__ src, int64_t n,
                                                          double* __restrict__ out) {
  const int64_t v = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v >= n) return;
  double s = 0.0;
  for (int k = ptr[v]; k < ptr[v + 1]; ++k) s += src[idx[k]];
  out[v] = s;
}

template <int M>
static int32_t launch_tet10_general(const DevPlan& p, const MatParams& mp, double* f, double* kval, bool fp32,
                                    hipStream_t s, int32_t* status) {
  const int64_t lanes = int64_t(p.E) * 10;
  const dim3 g(unsigned((lanes + 63) / 64));
  if (kval) {
    if (fp32)
      hipLaunchKernelGGL((tet10_elements_kernel<float, M, true>), g, dim3(64), 0, s, p, mp, status);
    else
      hipLaunchKernelGGL((tet10_elements_kernel<double, M, true>), g, dim3(64), 0, s, p, mp, status);
    if (p.gen_nnz > 0)
      hipLaunchKernelGGL(segment_sum_kernel, dim3(unsigned((p.gen_nnz + 255) / 256)), dim3(256), 0, s, p.gen_kptr,
                         p.gen_kidx, p.gen_ke, p.gen_nnz, kval);
  } else {
    if (fp32)
      hipLaunchKernelGGL((tet10_elements_kernel<float, M, false>), g, dim3(64), 0, s, p, mp, status);
    else
      hipLaunchKernelGGL((tet10_elements_kernel<double, M, false>), g, dim3(64), 0, s, p, mp, status);
  }
  const int64_t n = 3 * int64_t(p.N);
  hipLaunchKernelGGL(segment_sum_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, p.gen_fptr, p.gen_fidx,
                     p.gen_fe, n, f);
  return PSX_OK;
}

int32_t launch_tet10_hyper(const DevPlan& p, int model, const double* const* params, const double* u, double* f,
                           double* kval, int precision, hipStream_t s, int32_t* status) {
  int32_t rc = launch_permute_u(p, u, s, status);
  if (rc) return rc;
  if (p.wide) {
    const MatParams mp{{params[0], params[1], params[2], params[3]}};
    const bool fp32 = precision == PSX_FP32;
    switch (model) {
      case kNeoHookean: rc = launch_tet10_general<kNeoHookean>(p, mp, f, kval, fp32, s, status); break;
      case kMooneyRivlin: rc = launch_tet10_general<kMooneyRivlin>(p, mp, f, kval, fp32, s, status); break;
      case kYeoh: rc = launch_tet10_general<kYeoh>(p, mp, f, kval, fp32, s, status); break;
      default: return fail(PSX_EINVAL, "unknown material model");
    }
    if (rc) return rc;
    PSX_HIP_TRY(hipGetLastError());
    return PSX_OK;
  }
  if (p.n_tiles <= 0) return PSX_OK;
  const bool fp32 = precision == PSX_FP32;
  const MatParams mp{{params[0], params[1], params[2], params[3]}};
  switch (model) {
    case kNeoHookean: rc = launch_t10_tiles<kNeoHookean>(p, mp, f, kval, fp32, s, status); break;
    case kMooneyRivlin: rc = launch_t10_tiles<kMooneyRivlin>(p, mp, f, kval, fp32, s, status); break;
    case kYeoh: rc = launch_t10_tiles<kYeoh>(p, mp, f, kval, fp32, s, status); break;
    default: return fail(PSX_EINVAL, "unknown material model");
  }
  if (rc) return rc;
  PSX_HIP_TRY(hipGetLastError());
  return PSX_OK;
}

int32_t launch_tet10_assemble(const DevPlan& p, const double* u, const double* mu, const double* lam,
                              double* f, double* kval, int precision, hipStream_t s, int32_t* status) {
  const double* params[4] = {mu, lam, nullptr, nullptr};
  return launch_tet10_hyper(p, kNeoHookean, params, u, f, kval, precision, s, status);
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
  const R prm[4] = {R(mu[e]), R(lam[e] + (2.0 / 3.0) * mu[e]), R(0), R(0)};
  double* ko = k ? k + 900 * int64_t(e) : nullptr;
  double* fo = f + 30 * int64_t(e);
  for (int p = 0; p < 10; ++p) {
    for (int i = 0; i < 3; ++i) fo[3 * p + i] = 0.0;
    if (ko)
      for (int i = 0; i < 3; ++i)
        for (int c = 0; c < 30; ++c) ko[30 * (3 * p + i) + c] = 0.0;
    R J;
    if (ko)
      J = tet10_row<R, kNeoHookean, true>(dX, dU, prm, p, [&](int v, R val) {
        if (v >= 90) {
          fo[3 * p + (v - 90)] += double(val);
        } else {
          const int b = v / 9, i = (v % 9) / 3, c = v % 3;
          ko[30 * (3 * p + i) + 3 * b + c] += double(val);
        }
      });
    else
      J = tet10_row<R, kNeoHookean, false>(dX, dU, prm, p, [&](int v, R val) {
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
