// Tet4 Total Lagrangian Neo-Hookean kernels for gfx950.
//
// Global assembly is owner-computes over node rows.  The plan orders nodes
// along a Morton curve (locality) and packs consecutive nodes into tiles of
// at most TILE element incidences and at most TILE reduction lanes
// (8 per node + one per (node, slot) unit).  One work-group handles one tile:
//   phase 1: lane t takes incidence t (element e seen from one of its nodes
//            a): it gathers the element's nodal displacements, evaluates the
//            element in registers (FP64, or FP32 from precomputed gradients)
//            and leaves row a's contributions in LDS -- the diagonal block
//            K_aa (6 values, symmetric), f_a (3) and the three off-diagonal
//            blocks K_ab (27);
//   phase 2: every reduction lane sums, in FP64 and in increasing incidence
//            order, one list of 9-value contributions: eight lanes per node
//            take an eighth of the node's incidence run each (K_aa and f_a,
//            combined across the eight lanes in a fixed order), one lane per
//            off-diagonal CSR block walks the incidences that feed it.  The
//            lists have similar lengths, so the lanes of a wave stay in step.
//            Each lane then writes its 3x3 block (or f_a) into the CSR
//            values; the lanes of one node are its consecutive slots, so a
//            wave's stores of one row form contiguous runs.
// Every CSR value and every f entry is written exactly once by one lane, and
// every sum runs in an order fixed by the mesh alone (not by the tile size):
// the scatter is deterministic and uses no floating-point atomics.
// This replaces SolidDomain's element loops + np.add.at + coo.tocsr()
// (src/psierra/domain/solid.py:230-325, src/psierra/utility/assembly.py:37-108).
#include <climits>
#include <algorithm>
#include <type_traits>

#include "psx_device.h"
#include "psx_internal.h"
#include "psx_nh.h"
#include "psx_tet4_elem.h"
#include "psx_batch.h"

namespace psx {

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

// Inputs of one incidence, in the row's node order: the row node first,
// then the element's other three nodes in increasing local index.
template <typename R>
struct IncInputs {
  uint32_t pe = 0, rank = 0;
  double pr[4] = {0.0, 0.0, 0.0, 0.0};   // material parameters
  float4 G[4];               // FP32 path: precomputed (h_b, V)
  double4 U[4];
  double4 X[4];              // FP64 path: reference coordinates
};

// P is DevPlan (a kernel argument) or a constant-address-space DevPlan (a
// batch's plan table): either way each field is read through the scalar unit
// where it is used.
// WITH_U = false (FP32 rows of a tile that stages its displacements in
// LDS): no connectivity and no displacement loads; the caller fills in.U.
// The element part of an FP32 incidence of a tile that stages its nodes
// (in.pe already loaded): the material parameters and the four (h_b, .)
// records in the row's order.
template <int M, typename P>
__device__ __forceinline__ void load_incidence_elem(const P& p, const MatParams& mp, IncInputs<float>& in) {
  const int64_t e = int64_t(in.pe & kIncElemMask);
  load_params<M>(mp, e, in.pr);
  const int lp = int(in.pe >> kIncElemBits);
  const int r1 = lp == 0 ? 1 : 0, r2 = lp <= 1 ? 2 : 1, r3 = lp <= 2 ? 3 : 2;
  const float4* geo = M == kNeoHookeanBound ? p.geo_bound : p.geo;
  in.G[0] = geo[4 * e + lp];
  in.G[1] = geo[4 * e + r1];
  in.G[2] = geo[4 * e + r2];
  in.G[3] = geo[4 * e + r3];
}

template <typename R, int M, typename P, bool WITH_U = true>
__device__ __forceinline__ void load_incidence(const P& p, const MatParams& mp, int k, IncInputs<R>& in) {
  in.pe = p.inc_elem[k];
  in.rank = p.inc_rank[k];
  const int64_t e = int64_t(in.pe & kIncElemMask);
  if constexpr (!WITH_U && std::is_same<R, float>::value) {
    load_incidence_elem<M>(p, mp, in);   // F-bar: J_bar - 1 from the staged nodal values
    return;
  }
  const int4 c = *reinterpret_cast<const int4*>(p.conn + 4 * e);
  load_params<M>(mp, e, in.pr);
  if constexpr (M == kFbarNeoHookean) {
    // J_bar - 1: mean over the element's nodes in element order
    // (np.mean(J_node[elements], axis=1), tet4_fbar.py:396)
    const double* jn = mp.v[2];
    in.pr[2] = (((jn[c.x] + jn[c.y]) + jn[c.z]) + jn[c.w]) * 0.25;
  }
  const int lp = int(in.pe >> kIncElemBits);
  const int r1 = lp == 0 ? 1 : 0, r2 = lp <= 1 ? 2 : 1, r3 = lp <= 2 ? 3 : 2;
  const int i0 = lp == 0 ? c.x : lp == 1 ? c.y : lp == 2 ? c.z : c.w;
  const int i1 = lp == 0 ? c.y : c.x;
  const int i2 = lp <= 1 ? c.z : c.y;
  const int i3 = lp <= 2 ? c.w : c.z;
  if constexpr (std::is_same<R, float>::value) {
    const float4* geo = M == kNeoHookeanBound ? p.geo_bound : p.geo;
    in.G[0] = geo[4 * e + lp];
    in.G[1] = geo[4 * e + r1];
    in.G[2] = geo[4 * e + r2];
    in.G[3] = geo[4 * e + r3];
  } else {
    in.X[0] = p.Xi[i0];
    in.X[1] = p.Xi[i1];
    in.X[2] = p.Xi[i2];
    in.X[3] = p.Xi[i3];
  }
  in.U[0] = p.ui[i0];
  in.U[1] = p.ui[i1];
  in.U[2] = p.ui[i2];
  in.U[3] = p.ui[i3];
}

// Row of one incidence: out[0..35] = K_{a b} for b = a, then the other three
// nodes (row-major 3x3 each), out[36..38] = f_a.  Returns det F.
template <typename R, int M, bool WANT_K>
__device__ __forceinline__ R incidence_row(const IncInputs<R>& in, R* out) {
  R prm[4] = {R(in.pr[0]), R(in.pr[1]), R(in.pr[2]), R(in.pr[3])};
  if constexpr (std::is_same<R, float>::value && M == kNeoHookeanBound) {
    // the bound material: a row's second node is local node 0 or 1 (V mu),
    // its last node local node 2 or 3 (V K), whatever the row node
    prm[0] = in.G[1].w;
    prm[1] = in.G[3].w;
  }
  if constexpr (std::is_same<R, float>::value)
    return tet4_row_grad_m<M, WANT_K>(in.G, in.U, prm, out);
  else
    return tet4_row_m<R, M, WANT_K>(in.X, in.U, prm, out);
}

// f_int only (line-search trials): one lane per incidence, f_a summed in FP64
// over the node's incidence run by one lane per (node, component).
template <typename R, int TILE, int M, typename P>
__device__ __forceinline__ void force_tile(const P& p, const MatParams& mp, double* __restrict__ f,
                                           int32_t* status, int tile, int t) {
  __shared__ R contrib[3 * TILE];
  __shared__ int32_t s_incp[TILE + 1];
  // FP32: the displacements of the tile's nodes and halo staged in LDS, as
  // in rows_tile (plans that are not wide have the tile halos)
  constexpr bool LDSU = std::is_same<R, float>::value && M != kFbarNeoHookean;
  constexpr int NLU = kTileLdsNodes;
  __shared__ double s_u[LDSU ? 3 * NLU : 1];
  const int4 ti = load_uniform(p.tile_info, tile);
  const int n0 = ti.x, nn = ti.y, k0 = ti.z, nk = ti.w;
  int2 th = make_int2(0, -1);
  if constexpr (LDSU) {
    if (p.tile_halo) th = load_uniform(p.tile_halo, tile);
  }
  const bool lds_u = LDSU && th.y >= 0;
  IncInputs<R> in;
  uint32_t loc = 0u;
  int incp = 0;
  if (t <= nn) incp = p.inc_ptr[n0 + t];
  if (lds_u && TILE >= NLU) {
    // two rounds of loads, as in rows_tile
    if (t < nk) {
      in.pe = p.inc_elem[k0 + t];
      loc = p.inc_loc[k0 + t];
    }
    int node = -1;
    if (t < nn + th.y) node = t < nn ? n0 + t : p.halo[th.x + t - nn];
    double4 u = make_double4(0.0, 0.0, 0.0, 0.0);
    if (node >= 0) u = p.ui[node];
    if constexpr (LDSU) {
      if (t < nk) load_incidence_elem<M>(p, mp, in);
      if (node >= 0) {
        s_u[t] = u.x;
        s_u[NLU + t] = u.y;
        s_u[2 * NLU + t] = u.z;
      }
    }
  } else if (t < nk) {
    if (lds_u) {
      load_incidence<R, M, P, false>(p, mp, k0 + t, in);
      loc = p.inc_loc[k0 + t];
    } else {
      load_incidence<R, M>(p, mp, k0 + t, in);
    }
  }
  if (t <= nn) s_incp[t] = incp - k0;
  for (int i = t + TILE; i <= nn; i += TILE) s_incp[i] = p.inc_ptr[n0 + i] - k0;
  if constexpr (LDSU) {
    if (lds_u) {
      if (TILE < NLU) {
        for (int i = t; i < nn + th.y; i += TILE) {
          const int node = i < nn ? n0 + i : p.halo[th.x + i - nn];
          const double4 u = p.ui[node];
          s_u[i] = u.x;
          s_u[NLU + i] = u.y;
          s_u[2 * NLU + i] = u.z;
        }
      }
      __syncthreads();
      if (t < nk) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int l = int((loc >> (8 * q)) & 0xffu);
          in.U[q] = make_double4(s_u[l], s_u[NLU + l], s_u[2 * NLU + l], 0.0);
        }
      }
    }
  }
  if (t < nk) {
    R out[3];
    const R J = incidence_row<R, M, false>(in, out);
    if (J <= R(0)) flag_inversion(status, int(in.pe & kIncElemMask));
#pragma unroll
    for (int v = 0; v < 3; ++v) contrib[v * TILE + t] = out[v];
  }
  __syncthreads();
  // one lane per (node, component): the node's run summed in incidence order
  for (int jc = t; jc < 3 * nn; jc += TILE) {
    const int j = jc / 3, c = jc - 3 * j;
    double fa = 0.0;
    for (int kk = s_incp[j]; kk < s_incp[j + 1]; ++kk) fa += double(contrib[c * TILE + kk]);
    f[3 * int64_t(p.perm[n0 + j]) + c] = fa;
  }
}

template <typename R, int TILE, int M>
__global__ __launch_bounds__(TILE) void tet4_force_kernel(DevPlan p, MatParams mp, double* __restrict__ f,
                                                          int32_t* status) {
  const int tile = xcd_tile();
  if (tile >= p.n_tiles) return;
  force_tile<R, TILE, M>(p, mp, f, status, tile, int(threadIdx.x));
}

// Position in incidence_row's output of LDS value v: 0..5 the upper
// triangle of K_aa, 6..8 f_a, 9..35 the three off-diagonal blocks.
__host__ __device__ constexpr int lds_source(int v) {
  return v < 3 ? v : v == 3 ? 4 : v == 4 ? 5 : v == 5 ? 8 : v < 9 ? 30 + v : v;
}

// Sum of one value over the eight lanes of an aligned group (quads first,
// then the two quads), the same bits in every lane of the group.
// After the quad sums, lanes 0..3 of the group hold q0 and lanes 4..7 hold
// q1; row_half_mirror gives each lane the other quad's sum, and q0 + q1 and
// q1 + q0 are the same bits (IEEE addition commutes).
__device__ __forceinline__ double oct_sum(double x) {
  x = quad_sum(x);
  return x + quad_dpp<0x141>(x);   // row_half_mirror: lane i <- lane 7 - i of its 8
}

// f_int + K_T, one tile of node rows per work-group (see the file comment).
//
// LDS holds 18 contribution values per incidence at a time, in two passes:
// pass A the upper triangle of K_aa, f_a and the first off-diagonal block,
// pass B the other two off-diagonal blocks (kept in registers meanwhile).
// Per off-diagonal unit three 64-bit masks (one per block position q =
// 1..3) mark the incidences that feed it.  Incidence t finds its node
// through the run offset kept in byte 0 of its rank word; a unit lane finds
// its node by counting unit-run starts (one bit per node with units).
//
// Feed masks hold one bit per incidence of the node's run: MaskT is uint32_t
// when no node has more than 32 incidences (smaller LDS and cheaper bit
// scans), unsigned long long otherwise.
//
// F-bar (M = kFbarNeoHookean) writes into the wider pattern of K_dev + K_vol:
// fv gives each node's first value, row length and diagonal slot there, and
// each (node, slot) unit its slot in that row; the blocks are added to the
// K_vol values already stored.
template <typename R, int TILE, int M, typename MaskT, typename P>
__device__ __forceinline__ void rows_tile(const P& p, const MatParams& mp, const FbarView& fv,
                                          double* __restrict__ f, double* __restrict__ kval, int32_t* status,
                                          int tile, int t) {
  constexpr bool FBAR = M == kFbarNeoHookean;
  constexpr int NV = 18;
  constexpr int ND = kTet4DiagLanes;
  constexpr int NN = TILE / ND;
  constexpr int NW = TILE / 64;
  // store of one CSR value: a plain write, or an add onto K_vol (F-bar)
  auto put = [](double* dst, double v) {
    if constexpr (FBAR)
      *dst += v;
    else
      *dst = v;
  };
  __shared__ R contrib[NV * TILE];
  __shared__ MaskT s_mask[3 * TILE];
  __shared__ unsigned long long s_ustart[NW];
  __shared__ int32_t s_incp[NN + 1];
  __shared__ int32_t s_uoff[NN + 1];
  __shared__ int32_t s_kofs[NN];
  __shared__ uint8_t s_self[NN];
  __shared__ uint8_t s_nz[NN];
  __shared__ uint8_t s_first[TILE];
  __shared__ int32_t s_f2[FBAR ? 2 * NN : 1];   // F-bar: diagonal block offset, row length
  // user id of each node (where its f goes); the F-bar rows read it at the
  // store instead (loaded with the node records it costs their FP64
  // instantiation 36 VGPRs)
  __shared__ int32_t s_user[FBAR ? 1 : NN];
  // FP32 rows stage the displacements of the tile's nodes and its halo in
  // LDS (one gather per node instead of four per incidence), unless the
  // tile has more than kTileLdsNodes of them
  constexpr bool LDSU = std::is_same<R, float>::value;
  constexpr int NLU = kTileLdsNodes;
  __shared__ double s_u[LDSU ? 3 * NLU : 1];
  __shared__ double s_j[LDSU && FBAR ? NLU : 1];   // F-bar: J_node - 1 of the staged nodes

  const int4 ti = load_uniform(p.tile_info, tile);
  const int n0 = ti.x, nn = ti.y, k0 = ti.z, nk = ti.w;
  const int lane = t & 63, wave = t >> 6;
  int2 th = make_int2(0, -1);
  if constexpr (LDSU) th = load_uniform(p.tile_halo, tile);
  const bool lds_u = LDSU && th.y >= 0;
  // the node records first: their loads join the first round below
  int4 nr0 = make_int4(0, 0, 0, 0), nr1 = make_int4(0, 0, 0, 0), nr2 = make_int4(0, 0, 0, 0);
  int user = 0;
  if (t < nn) {
    nr0 = p.node_rec[n0 + t];
    nr1 = p.node_rec[n0 + t + 1];
    if constexpr (!FBAR) user = p.perm[n0 + t];
    if constexpr (FBAR) nr2 = fv.node_rec2[n0 + t];
  }

  IncInputs<R> in;
  uint32_t loc = 0u;
  if constexpr (LDSU && TILE >= NLU) {
    if (lds_u) {
      // Two rounds of loads, each issued before anything waits on the
      // other: the incidence records and this thread's staging node, then
      // the element records and the node's displacement.
      if (t < nk) {
        in.pe = p.inc_elem[k0 + t];
        in.rank = p.inc_rank[k0 + t];
        loc = p.inc_loc[k0 + t];
      }
      int node = -1;
      if (t < nn + th.y) node = t < nn ? n0 + t : p.halo[th.x + t - nn];
      double4 u = make_double4(0.0, 0.0, 0.0, 0.0);
      double jm = 0.0;
      if (node >= 0) {
        u = p.ui[node];
        if constexpr (FBAR) jm = mp.v[2][node];
      }
      if (t < nk) load_incidence_elem<M>(p, mp, in);
      if (node >= 0) {
        s_u[t] = u.x;
        s_u[NLU + t] = u.y;
        s_u[2 * NLU + t] = u.z;
        if constexpr (FBAR) s_j[t] = jm;
      }
    } else if (t < nk) {
      load_incidence<R, M>(p, mp, k0 + t, in);
    }
  } else {
    if (t < nk) {
      if (lds_u) {
        load_incidence<R, M, P, false>(p, mp, k0 + t, in);
        loc = p.inc_loc[k0 + t];
      } else {
        load_incidence<R, M>(p, mp, k0 + t, in);
      }
    }
    if constexpr (LDSU) {
      if (lds_u) {
        for (int i = t; i < nn + th.y; i += TILE) {
          const int node = i < nn ? n0 + i : p.halo[th.x + i - nn];
          const double4 u = p.ui[node];
          s_u[i] = u.x;
          s_u[NLU + i] = u.y;
          s_u[2 * NLU + i] = u.z;
          if constexpr (FBAR) s_j[i] = mp.v[2][node];
        }
      }
    }
  }
  // node table
  int ub = 0, ue = 0;
  if (t < nn) {
    const int4 r0 = nr0, r1 = nr1;
    const int kb = r0.x - k0, ke = r1.x - k0;
    s_incp[t] = kb;
    s_uoff[t] = r0.y;
    s_kofs[t] = r0.z;
    s_self[t] = uint8_t(r0.w);
    if constexpr (!FBAR) s_user[t] = user;
    if constexpr (FBAR) {
      const int4 r2 = nr2;
      s_kofs[t] = r2.x;
      s_f2[t] = r2.x + 3 * r2.z;
      s_f2[NN + t] = r2.y;
    }
    if (t == nn - 1) {
      s_incp[nn] = ke;
      s_uoff[nn] = r1.y;
    }
    if (ke > kb) s_first[kb] = uint8_t(t);
    ub = r0.y;
    ue = r1.y;
  }
  if (wave == 0) {
    // nodes with units, in order (node lanes all lie in wave 0: nn <= 64)
    const bool has = t < nn && ue > ub;
    const unsigned long long bal = __ballot(has);
    if (has) s_nz[__popcll(bal & ((1ull << lane) - 1ull))] = uint8_t(t);
  }
  if (t < NW) s_ustart[t] = 0ull;
#pragma unroll
  for (int q = 0; q < 3; ++q) s_mask[q * TILE + t] = MaskT(0);
  __syncthreads();

  // ---- phase 1: incidence t
  const int u0 = s_uoff[0];
  if (t < nn && ue > ub) atomicOr(&s_ustart[(ub - u0) >> 6], 1ull << ((ub - u0) & 63));
  if constexpr (LDSU) {
    if (lds_u && t < nk) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int l = int((loc >> (8 * q)) & 0xffu);
        in.U[q] = make_double4(s_u[l], s_u[NLU + l], s_u[2 * NLU + l], 0.0);
      }
      if constexpr (FBAR) {
        // J_bar - 1: mean of the nodal J - 1 in element order (local node L
        // sits at row position 0 when L is the row node, else L + 1 below it
        // and L above), as the gathering path forms it
        const int lp = int(in.pe >> kIncElemBits);
        double jv[4];
#pragma unroll
        for (int L = 0; L < 4; ++L) {
          const int q = L == lp ? 0 : (L < lp ? L + 1 : L);
          jv[L] = s_j[int((loc >> (8 * q)) & 0xffu)];
        }
        in.pr[2] = (((jv[0] + jv[1]) + jv[2]) + jv[3]) * 0.25;
      }
    }
  }
  R keep[NV];   // pass B values: the blocks at q = 2 and 3
  if (t < nk) {
    R out[39];
    const R J = incidence_row<R, M, true>(in, out);
    if (J <= R(0)) flag_inversion(status, int(in.pe & kIncElemMask));
    const uint32_t r = in.rank;
    const int off = int(r & 0xffu);
    const int j = s_first[t - off];
    const int ubase = s_uoff[j] - u0;
    const MaskT bit = MaskT(1) << off;
#pragma unroll
    for (int q = 1; q < 4; ++q) atomicOr(&s_mask[(q - 1) * TILE + ubase + int((r >> (8 * q)) & 0xffu)], bit);
#pragma unroll
    for (int v = 0; v < NV; ++v) contrib[v * TILE + t] = out[lds_source(v)];
#pragma unroll
    for (int v = 0; v < NV; ++v) keep[v] = out[18 + v];
  }
  __syncthreads();

  // ---- phase 2: FP64 sums of 9-value contributions in incidence order
  const int nd = ND * nn;
  const int units = s_uoff[nn] - u0;
  int j = 0, kb = 0, slot = 0;
  bool diag = false, offd = false;
  MaskT m1 = 0, m2 = 0, m3 = 0;
  if (t < nd) {
    // (node, eighth of its run): K_aa upper triangle and f_a
    j = t / ND;
    const unsigned part = unsigned(t) % ND;
    kb = s_incp[j];
    const unsigned L = unsigned(s_incp[j + 1] - kb);
    m1 = MaskT(bit_range(int((part * L) / ND), int(((part + 1) * L) / ND)));   // <= 8 bits
    diag = true;
  } else if (t < nd + units) {
    // (node, off-diagonal slot)
    const int u = t - nd;
    j = unit_owner<NW>(s_ustart, s_nz, u);
    kb = s_incp[j];
    slot = u - (s_uoff[j] - u0);
    if (slot != int(s_self[j])) {
      offd = true;
      m1 = s_mask[u];
      m2 = s_mask[TILE + u];
      m3 = s_mask[2 * TILE + u];
    }
  }
  // F-bar: the K_vol values this unit adds onto, loaded now so their latency
  // hides behind the sums
  double old[FBAR ? 9 : 1];
  double* odst = kval;
  int od = 0;
  if constexpr (FBAR) {
    if (offd) {
      od = s_f2[NN + j];
      odst = kval + s_kofs[j] + 3 * fv.slot2[u0 + t - nd];
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) old[3 * r + c] = odst[r * 3 * od + c];
    }
  }
  double acc[9];
#pragma unroll
  for (int c = 0; c < 9; ++c) acc[c] = 0.0;
  // Sum the 9 values at rows row0.. of the incidences in `msk`, two at a
  // time, in increasing incidence order; rowb(a) picks the rows per bit.
  auto walk = [&](MaskT msk, auto rowb) {
    const R* base = contrib + kb;
    while (msk) {
      const int a = low_bit(msk);
      msk &= msk - 1;
      const bool two = msk != 0;
      const int b = two ? low_bit(msk) : a;
      msk &= msk - 1;
      const R* sa = base + rowb(a) * TILE + a;
      const R* sb = base + rowb(b) * TILE + b;
      R va[9], vb[9];
#pragma unroll
      for (int c = 0; c < 9; ++c) {
        va[c] = sa[c * TILE];
        vb[c] = sb[c * TILE];
      }
#pragma unroll
      for (int c = 0; c < 9; ++c) acc[c] += double(va[c]);
      if (two)
#pragma unroll
        for (int c = 0; c < 9; ++c) acc[c] += double(vb[c]);
    }
  };
  // pass A: K_aa and f_a (rows 0..8), blocks at q = 1 (rows 9..17)
  {
    const int row = diag ? 0 : 9;
    walk(m1, [&](int) { return row; });
  }
  if (diag) {
    // the eight partial sums of the node, combined in a fixed order
#pragma unroll
    for (int c = 0; c < 9; ++c) acc[c] = oct_sum(acc[c]);
    const int part = int(unsigned(t) % ND);
    const int d = FBAR ? s_f2[NN + j] : s_uoff[j + 1] - s_uoff[j];
    if (part == 3) {
      const int64_t a = FBAR ? p.perm[n0 + j] : s_user[j];
#pragma unroll
      for (int c = 0; c < 3; ++c) f[3 * a + c] = acc[6 + c];
    } else if (part < 3 && d > 0) {
      // row `part` of K_aa from its upper triangle (xx xy xz yy yz zz)
      const double e0 = part == 0 ? acc[0] : part == 1 ? acc[1] : acc[2];
      const double e1 = part == 0 ? acc[1] : part == 1 ? acc[3] : acc[4];
      const double e2 = part == 0 ? acc[2] : part == 1 ? acc[4] : acc[5];
      double* dst = kval + (FBAR ? s_f2[j] : s_kofs[j] + 3 * int(s_self[j])) + part * 3 * d;
      put(dst, e0);
      put(dst + 1, e1);
      put(dst + 2, e2);
    }
  }
  __syncthreads();   // pass A reads are done
  if (t < nk) {
#pragma unroll
    for (int v = 0; v < NV; ++v) contrib[v * TILE + t] = keep[v];
  }
  __syncthreads();
  // pass B: blocks at q = 2 (rows 0..8) and q = 3 (rows 9..17)
  if (offd) {
    walk(m2 | m3, [&](int a) { return ((m2 >> a) & 1) ? 0 : 9; });
    // each 3x3 block once; a unit lane's rows are three consecutive values,
    // so neighbouring slots give contiguous runs
    if constexpr (FBAR) {
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) odst[r * 3 * od + c] = old[3 * r + c] + acc[3 * r + c];
    } else {
      const int d = s_uoff[j + 1] - s_uoff[j];
      double* dst = kval + s_kofs[j] + 3 * slot;
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) put(dst + r * 3 * d + c, acc[3 * r + c]);
    }
  }
}

template <typename R, int TILE, int M, typename MaskT>
__global__ __launch_bounds__(TILE) void tet4_assemble_rows(DevPlan p, MatParams mp, FbarView fv,
                                                           double* __restrict__ f, double* __restrict__ kval,
                                                           int32_t* status) {
  const int tile = xcd_tile();
  if (tile >= p.n_tiles) return;
  rows_tile<R, TILE, M, MaskT>(p, mp, fv, f, kval, status, tile, int(threadIdx.x));
}

// f_int + K_T for a wide plan (a node with more than 64 incident elements or
// more than 255 neighbours, beyond the feed masks and slot-rank bytes of
// tet4_assemble_rows): one tile of node rows per work-group, the reduction
// through unit-major staging areas laid out by the plan.
//
// Incidence t (element e seen from node a) evaluates its row in registers
// and stages four 9-value blocks: its diagonal block (the upper triangle of
// K_aa) with f_a at block t -- so a node's diagonal feeds are its incidence
// run, contiguous -- and its three off-diagonal blocks K_ab at the positions
// the plan gave it (inc_pos): the first (the other node of lowest local
// index) in area A, the other two in area B.  In each area every (node,
// slot) unit finds its feeds contiguous and in run order (unit_rec: first A
// and B positions, tile-local node).  Then the lanes loop over the reduction
// items, each summing one list in FP64: eight items per node take an eighth
// of the node's run each (combined across the aligned group of eight lanes
// in a fixed order, the same bits in all eight; three write the rows of
// K_aa, one writes f_a), one item per unit sums its A feeds and then its B
// feeds.  Every CSR value and f entry is written once, every sum has an
// order fixed by the mesh alone: deterministic, the same for every tile
// size, no masks, no atomics; runs of up to TILE incidences and any number
// of neighbours.
template <typename R, int TILE, int M>
__global__ __launch_bounds__(TILE) void tet4_rows(DevPlan p, MatParams mp, double* __restrict__ f,
                                                  double* __restrict__ kval, int32_t* status) {
  constexpr int NB = 4 * TILE;       // staged blocks, value-major: value v of block b at v NB + b
  constexpr int A0 = TILE;           // first block of area A
  constexpr int B0 = 2 * TILE;       // first block of area B
  constexpr int ND = kTet4DiagLanes; // items per node for K_aa and f_a
  __shared__ R s_blk[9 * NB];
  __shared__ int32_t s_kb[TILE + 1];    // first incidence of each node (tile-relative)
  __shared__ int32_t s_ub[TILE + 1];    // first unit of each node (tile-relative)
  __shared__ int32_t s_kofs[TILE];      // first CSR value of the node's three rows
  __shared__ int32_t s_self[TILE];      // slot of the node in its own row
  __shared__ int32_t s_user[TILE];      // user id of the node
  const int tile = xcd_tile();
  if (tile >= p.n_tiles) return;
  const int4 ti = load_uniform(p.tile_info, tile);
  const int n0 = ti.x, nn = ti.y, k0 = ti.z, nk = ti.w;
  const int t = threadIdx.x;
  const int u0 = load_uniform(p.node_rec, n0).y;
  const int units = load_uniform(p.node_rec, n0 + nn).y - u0;
  const int nd = ND * nn;
  IncInputs<R> in;
  uint32_t pos = 0;
  if (t < nk) {
    load_incidence<R, M>(p, mp, k0 + t, in);
    pos = p.inc_pos[k0 + t];
  }
  for (int i = t; i <= nn; i += TILE) {
    const int4 r = p.node_rec[n0 + i];
    s_kb[i] = r.x - k0;
    s_ub[i] = r.y - u0;
    if (i < nn) {
      s_kofs[i] = r.z;
      s_self[i] = r.w;
      s_user[i] = p.perm[n0 + i];
    }
  }
  if (t < nk) {
    R out[39];
    const R J = incidence_row<R, M, true>(in, out);
    if (J <= R(0)) flag_inversion(status, int(in.pe & kIncElemMask));
    // diagonal block: K_aa upper triangle (xx xy xz yy yz zz) and f_a
    const R dv[9] = {out[0], out[1], out[2], out[4], out[5], out[8], out[36], out[37], out[38]};
#pragma unroll
    for (int c = 0; c < 9; ++c) s_blk[c * NB + t] = dv[c];
#pragma unroll
    for (int q = 1; q < 4; ++q) {
      const int b = (q == 1 ? A0 : B0) + int((pos >> (kPosBits * (q - 1))) & kPosMask);
#pragma unroll
      for (int c = 0; c < 9; ++c) s_blk[c * NB + b] = out[9 * q + c];
    }
  }
  __syncthreads();
  // Sum of `cnt` staged blocks from block `b`, in order.
  auto sum_blocks = [&](double* acc, int b, int cnt) {
    for (int k = 0; k < cnt; ++k) {
      R v[9];
#pragma unroll
      for (int c = 0; c < 9; ++c) v[c] = s_blk[c * NB + b + k];
#pragma unroll
      for (int c = 0; c < 9; ++c) acc[c] += double(v[c]);
    }
  };
  // the lanes loop over the items; a node's eight diagonal items stay in one
  // aligned group of eight lanes (nd and TILE are multiples of 8)
  for (int w = t; w < nd + units; w += TILE) {
    double acc[9];
#pragma unroll
    for (int c = 0; c < 9; ++c) acc[c] = 0.0;
    if (w < nd) {
      // (node, eighth of its run): K_aa upper triangle and f_a
      const int j = w / ND, part = w % ND;
      const int kb = s_kb[j], L = s_kb[j + 1] - kb;
      const int b = kb + (part * L) / ND, e = kb + ((part + 1) * L) / ND;
      sum_blocks(acc, b, e - b);
#pragma unroll
      for (int c = 0; c < 9; ++c) acc[c] = oct_sum(acc[c]);
      if (part == 3) {
        const int64_t a = s_user[j];
#pragma unroll
        for (int c = 0; c < 3; ++c) f[3 * a + c] = acc[6 + c];
      } else if (part < 3) {
        const int d = s_ub[j + 1] - s_ub[j];
        if (d > 0) {
          // row `part` of K_aa from its upper triangle
          const double e0 = part == 0 ? acc[0] : part == 1 ? acc[1] : acc[2];
          const double e1 = part == 0 ? acc[1] : part == 1 ? acc[3] : acc[4];
          const double e2 = part == 0 ? acc[2] : part == 1 ? acc[4] : acc[5];
          double* dst = kval + s_kofs[j] + 3 * s_self[j] + part * 3 * d;
          dst[0] = e0;
          dst[1] = e1;
          dst[2] = e2;
        }
      }
    } else {
      // (node, off-diagonal slot): first A and B positions from the unit's
      // record, counts from the next unit's (the tile's ends for the last)
      const int u = w - nd;
      const uint32_t r = p.unit_rec[u0 + u];
      const uint32_t r2 = u + 1 < units ? p.unit_rec[u0 + u + 1] : uint32_t(nk) | (uint32_t(2 * nk) << 10);
      const int fa = int(r & kPosMask), fb = int((r >> 10) & kPosMask);
      const int ca = int(r2 & kPosMask) - fa, cb = int((r2 >> 10) & kPosMask) - fb;
      if (ca + cb > 0) {
        sum_blocks(acc, A0 + fa, ca);
        sum_blocks(acc, B0 + fb, cb);
        const int j = int(r >> 20);
        const int d = s_ub[j + 1] - s_ub[j];
        double* dst = kval + s_kofs[j] + 3 * (u - s_ub[j]);
#pragma unroll
        for (int rr = 0; rr < 3; ++rr)
#pragma unroll
          for (int c = 0; c < 3; ++c) dst[rr * 3 * d + c] = acc[3 * rr + c];
      }
    }
  }
}

template <typename R, int TILE, int M>
static void launch_t4(const DevPlan& p, const MatParams& mp, const FbarView& fv, double* f, double* kval,
                      hipStream_t s, int32_t* status) {
  const int blocks = 8 * ((p.n_tiles + 7) / 8);
  if constexpr (M != kFbarNeoHookean) {
    if (kval && p.wide) {
      hipLaunchKernelGGL((tet4_rows<R, TILE, M>), dim3(blocks), dim3(TILE), 0, s, p, mp, f, kval, status);
      return;
    }
  }
  if (kval && p.max_inc <= 32)
    hipLaunchKernelGGL((tet4_assemble_rows<R, TILE, M, uint32_t>), dim3(blocks), dim3(TILE), 0, s, p, mp, fv, f,
                       kval, status);
  else if (kval)
    hipLaunchKernelGGL((tet4_assemble_rows<R, TILE, M, unsigned long long>), dim3(blocks), dim3(TILE), 0, s, p, mp,
                       fv, f, kval, status);
  else
    hipLaunchKernelGGL((tet4_force_kernel<R, TILE, M>), dim3(blocks), dim3(TILE), 0, s, p, mp, f, status);
}

template <int M>
static int32_t launch_t4_tiles(const DevPlan& p, const MatParams& mp, const FbarView& fv, double* f, double* kval,
                               bool fp32, hipStream_t s, int32_t* status) {
  switch (p.tile) {
    case 64:
      fp32 ? launch_t4<float, 64, M>(p, mp, fv, f, kval, s, status)
           : launch_t4<double, 64, M>(p, mp, fv, f, kval, s, status);
      break;
    case 128:
      fp32 ? launch_t4<float, 128, M>(p, mp, fv, f, kval, s, status)
           : launch_t4<double, 128, M>(p, mp, fv, f, kval, s, status);
      break;
    case 256:
      fp32 ? launch_t4<float, 256, M>(p, mp, fv, f, kval, s, status)
           : launch_t4<double, 256, M>(p, mp, fv, f, kval, s, status);
      break;
    default:
      return fail(PSX_EINVAL, "unsupported tile");
  }
  return PSX_OK;
}

// F-bar rows (K_dev added onto the K_vol values in kval, and f) or f only.
int32_t launch_t4_fbar_rows(const DevPlan& p, const double* mu, const double* lam, const double* jm1_node,
                            const FbarView& fv, double* f, double* kval, int precision, hipStream_t s,
                            int32_t* status) {
  if (p.n_tiles <= 0) return PSX_OK;
  const MatParams mp{{mu, lam, jm1_node, nullptr}};
  const int32_t rc = launch_t4_tiles<kFbarNeoHookean>(p, mp, fv, f, kval, precision == PSX_FP32, s, status);
  if (rc) return rc;
  PSX_HIP_TRY(hipGetLastError());
  return PSX_OK;
}

int32_t launch_permute_u(const DevPlan& p, const double* u, hipStream_t s, int32_t* status) {
  hipLaunchKernelGGL(permute_u_kernel, dim3((p.N + 255) / 256), dim3(256), 0, s, p, u, status);
  PSX_HIP_TRY(hipGetLastError());
  return PSX_OK;
}

int32_t launch_tet4_hyper(const DevPlan& p, int model, const double* const* params, const double* u,
                          double* f, double* kval, int precision, hipStream_t s, int32_t* status) {
  int32_t rc = launch_permute_u(p, u, s, status);
  if (rc) return rc;
  if (p.n_tiles <= 0) return PSX_OK;
  const bool fp32 = precision == PSX_FP32;
  const MatParams mp{{params[0], params[1], params[2], params[3]}};
  switch (model) {
    case kNeoHookean: rc = launch_t4_tiles<kNeoHookean>(p, mp, FbarView{}, f, kval, fp32, s, status); break;
    case kMooneyRivlin: rc = launch_t4_tiles<kMooneyRivlin>(p, mp, FbarView{}, f, kval, fp32, s, status); break;
    case kYeoh: rc = launch_t4_tiles<kYeoh>(p, mp, FbarView{}, f, kval, fp32, s, status); break;
    default: return fail(PSX_EINVAL, "unknown material model");
  }
  if (rc) return rc;
  PSX_HIP_TRY(hipGetLastError());
  return PSX_OK;
}

// Geometry records of a bound material (FP32 Neo-Hookean rows): record b of
// element e is (h_b, V mu_e) for local nodes 0 and 1 and (h_b, V K_e) for 2
// and 3, K = lambda + 2 mu / 3.  V is recomputed here in FP64 from the
// reference coordinates (the same determinant as the plan's), so each
// product V mu and V K is formed in FP64 and rounded to FP32 once.
// Every row's node order (row node, then the others by local index) puts a
// local node 0 or 1 second and a local node 2 or 3 last, so a row finds
// V mu in G[1].w and V K in G[3].w.  The Neo-Hookean S, its tangent
// coefficients and so f and K of a row are linear in (mu, K), so the row with
// V = 1 and (V mu, V K) is the row with V and (mu, K).
__global__ __launch_bounds__(256) void bind_material_kernel(DevPlan p, const double* __restrict__ mu,
                                                            const double* __restrict__ lam,
                                                            float4* __restrict__ out) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= p.E) return;
  const int4 c = *reinterpret_cast<const int4*>(p.conn + 4 * e);
  const double4 x0 = p.Xi[c.x], x1 = p.Xi[c.y], x2 = p.Xi[c.z], x3 = p.Xi[c.w];
  const double D[3][3] = {{x1.x - x0.x, x2.x - x0.x, x3.x - x0.x},
                          {x1.y - x0.y, x2.y - x0.y, x3.y - x0.y},
                          {x1.z - x0.z, x2.z - x0.z, x3.z - x0.z}};
  const double V = det3(D) / 6.0;
  const double m = mu[e];
  const double K = lam[e] + (2.0 / 3.0) * m;
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    float4 g = p.geo[4 * e + b];
    g.w = float(V * (b < 2 ? m : K));
    out[4 * e + b] = g;
  }
}

int32_t launch_bind_material(const DevPlan& p, const double* mu, const double* lam, float4* out, hipStream_t s) {
  if (p.E <= 0) return PSX_OK;
  hipLaunchKernelGGL(bind_material_kernel, dim3(unsigned((int64_t(p.E) + 255) / 256)), dim3(256), 0, s, p, mu,
                     lam, out);
  PSX_HIP_TRY(hipGetLastError());
  return PSX_OK;
}

// Neo-Hookean with the plan's bound material: FP32 rows read (V mu, V K) from
// the bound geometry records and load no parameters; FP64 rows use the
// plan's copies of mu and lambda as an ordinary call.
int32_t launch_tet4_assemble_bound(const DevPlan& p, const double* u, const double* mu, const double* lam,
                                   double* f, double* kval, int precision, hipStream_t s, int32_t* status) {
  if (precision != PSX_FP32) return launch_tet4_assemble(p, u, mu, lam, f, kval, precision, s, status);
  int32_t rc = launch_permute_u(p, u, s, status);
  if (rc) return rc;
  if (p.n_tiles <= 0) return PSX_OK;
  const MatParams mp{{nullptr, nullptr, nullptr, nullptr}};
  switch (p.tile) {
    case 64: launch_t4<float, 64, kNeoHookeanBound>(p, mp, FbarView{}, f, kval, s, status); break;
    case 128: launch_t4<float, 128, kNeoHookeanBound>(p, mp, FbarView{}, f, kval, s, status); break;
    case 256: launch_t4<float, 256, kNeoHookeanBound>(p, mp, FbarView{}, f, kval, s, status); break;
    default: return fail(PSX_EINVAL, "unsupported tile");
  }
  PSX_HIP_TRY(hipGetLastError());
  return PSX_OK;
}

int32_t launch_tet4_assemble(const DevPlan& p, const double* u, const double* mu, const double* lam,
                             double* f, double* kval, int precision, hipStream_t s, int32_t* status) {
  const double* params[4] = {mu, lam, nullptr, nullptr};
  return launch_tet4_hyper(p, kNeoHookean, params, u, f, kval, precision, s, status);
}

// ---------------------------------------------------------------------------
// Batched assembly: many Tet4 plans (meshes of uneven size, or one mesh
// repeated for parameter samples) in one persistent launch.  The tiles of
// all plans form one list; a fixed grid of work-groups (as many as stay
// resident) walks it, XCD x taking the x-th eighth of the list so
// neighbouring tiles of one mesh share an L2.  Each tile is assembled by the
// same tile routine as a single-plan launch, so every instance's f and K
// are bit-identical to psx_assemble's.  Tiles hold at most TILE incidences
// whatever the mesh size, so uneven instances balance by construction.

// Displacements of every instance into its plan's internal node order, and
// the per-instance status words reset.
__global__ __launch_bounds__(256) void batch_permute_kernel(BatchDev b, const double* const* __restrict__ u,
                                                            int32_t* status) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < b.n) {
    status[2 * i] = 0;
    status[2 * i + 1] = INT_MAX;
  }
  if (i >= b.n_nodes) return;
  const int inst = batch_owner<false>(b.node_start, b.n, i);
  const int k = i - b.node_start[inst];
  const DevPlan& p = b.plans[inst];
  const double* ui = u[inst];
  const int64_t a = p.perm[k];
  p.ui[k] = make_double4(ui[3 * a], ui[3 * a + 1], ui[3 * a + 2], 0.0);
}

template <typename R, int TILE, int M, typename MaskT, bool PERSISTENT>
__global__ __launch_bounds__(TILE) void tet4_batch_rows(BatchDev b, const double* const* mu,
                                                        const double* const* lam, double* const* f,
                                                        double* const* kval, int32_t* status) {
  batch_tiles<PERSISTENT>(b, [&](int inst, int tile, int t) {
    const ConstDevPlan& p = *(const ConstDevPlan*)(b.plans + inst);
    MatParams mp{{nullptr, nullptr, nullptr, nullptr}};
    if constexpr (M != kNeoHookeanBound) {
      mp.v[0] = load_uniform_val(mu + inst);
      mp.v[1] = load_uniform_val(lam + inst);
    }
    rows_tile<R, TILE, M, MaskT>(p, mp, FbarView{}, load_uniform_val(f + inst), load_uniform_val(kval + inst),
                                 status + 2 * inst, tile, t);
  });
}

template <typename R, int TILE, int M, bool PERSISTENT>
__global__ __launch_bounds__(TILE) void tet4_batch_force(BatchDev b, const double* const* mu,
                                                         const double* const* lam, double* const* f,
                                                         int32_t* status) {
  batch_tiles<PERSISTENT>(b, [&](int inst, int tile, int t) {
    const ConstDevPlan& p = *(const ConstDevPlan*)(b.plans + inst);
    MatParams mp{{nullptr, nullptr, nullptr, nullptr}};
    if constexpr (M != kNeoHookeanBound) {
      mp.v[0] = load_uniform_val(mu + inst);
      mp.v[1] = load_uniform_val(lam + inst);
    }
    force_tile<R, TILE, M>(p, mp, load_uniform_val(f + inst), status + 2 * inst, tile, t);
  });
}

template <typename R, int TILE, int M, bool PERSISTENT>
static void launch_batch_t4(const BatchDev& b, const double* const* mu, const double* const* lam,
                            double* const* f, double* const* kval, int cus, hipStream_t s, int32_t* status) {
  if (kval && b.max_inc <= 32) {
    const auto k = tet4_batch_rows<R, TILE, M, uint32_t, PERSISTENT>;
    const int grid = batch_grid(reinterpret_cast<const void*>(k), TILE, cus, b.n_tiles, PERSISTENT);
    hipLaunchKernelGGL(k, dim3(grid), dim3(TILE), 0, s, b, mu, lam, f, kval, status);
  } else if (kval) {
    const auto k = tet4_batch_rows<R, TILE, M, unsigned long long, PERSISTENT>;
    const int grid = batch_grid(reinterpret_cast<const void*>(k), TILE, cus, b.n_tiles, PERSISTENT);
    hipLaunchKernelGGL(k, dim3(grid), dim3(TILE), 0, s, b, mu, lam, f, kval, status);
  } else {
    const auto k = tet4_batch_force<R, TILE, M, PERSISTENT>;
    const int grid = batch_grid(reinterpret_cast<const void*>(k), TILE, cus, b.n_tiles, PERSISTENT);
    hipLaunchKernelGGL(k, dim3(grid), dim3(TILE), 0, s, b, mu, lam, f, status);
  }
}

template <typename R, int M>
static int32_t launch_batch_tiles(const BatchDev& b, const double* const* mu, const double* const* lam,
                                  double* const* f, double* const* kval, int cus, bool persistent,
                                  hipStream_t s, int32_t* status) {
  switch (b.tile * 2 + int(persistent)) {
    case 128: launch_batch_t4<R, 64, M, false>(b, mu, lam, f, kval, cus, s, status); break;
    case 129: launch_batch_t4<R, 64, M, true>(b, mu, lam, f, kval, cus, s, status); break;
    case 256: launch_batch_t4<R, 128, M, false>(b, mu, lam, f, kval, cus, s, status); break;
    case 257: launch_batch_t4<R, 128, M, true>(b, mu, lam, f, kval, cus, s, status); break;
    case 512: launch_batch_t4<R, 256, M, false>(b, mu, lam, f, kval, cus, s, status); break;
    case 513: launch_batch_t4<R, 256, M, true>(b, mu, lam, f, kval, cus, s, status); break;
    default: return fail(PSX_EINVAL, "unsupported tile");
  }
  return PSX_OK;
}

// u, mu, lam, f, kval: device arrays of n device pointers; mu = lam = NULL
// for the plans' bound material (FP32: the bound geometry records; FP64:
// the plans' copies of mu and lambda); kval = NULL for f only.
int32_t launch_batch_permute(const BatchDev& b, const double* const* u, hipStream_t s, int32_t* status) {
  const int total = std::max(b.n_nodes, b.n);
  hipLaunchKernelGGL(batch_permute_kernel, dim3((total + 255) / 256), dim3(256), 0, s, b, u, status);
  PSX_HIP_TRY(hipGetLastError());
  return PSX_OK;
}

int32_t launch_tet4_batch(const BatchDev& b, const double* const* u, const double* const* mu,
                          const double* const* lam, double* const* f, double* const* kval, int precision,
                          int cus, bool persistent, hipStream_t s, int32_t* status) {
  int32_t rc = launch_batch_permute(b, u, s, status);
  if (rc) return rc;
  if (b.n_tiles <= 0) return PSX_OK;
  if (!mu && precision == PSX_FP32)
    rc = launch_batch_tiles<float, kNeoHookeanBound>(b, nullptr, nullptr, f, kval, cus, persistent, s, status);
  else if (!mu)
    rc = launch_batch_tiles<double, kNeoHookean>(b, b.bound_mu, b.bound_lam, f, kval, cus, persistent, s, status);
  else if (precision == PSX_FP32)
    rc = launch_batch_tiles<float, kNeoHookean>(b, mu, lam, f, kval, cus, persistent, s, status);
  else
    rc = launch_batch_tiles<double, kNeoHookean>(b, mu, lam, f, kval, cus, persistent, s, status);
  if (rc) return rc;
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
