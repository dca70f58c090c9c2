// Free-DOF CSC extraction and the free-DOF residual for the Newton solve.
//
// pSierra builds the solver matrix as
//   K_total = K_tan + K_contact            (CSR '+', drops exact-zero sums)
//   K_tan_free = K_total[free][:, free].tocsc()
// (src/psierra/analysis/nonlinear_static.py:354-355) and hands it to SuperLU
// (src/psierra/solver/sparse_linear_solver.py:16-19).  Here the columns of
// node b (3b, 3b+1, 3b+2) are built by one wave: lane (s, r) takes row 3a + r
// of b's s-th neighbour a (ascending), whose three values K[3a+r][3b..3b+2]
// are consecutive in the CSR values (one 24-byte read serves the three
// columns).  A ballot per column marks the kept entries (free row, and
// nonzero when exact zeros are dropped, as the CSR '+' does).  The first
// pass writes every structural entry at its structural position (column
// starts fixed at creation) and counts the kept entries per column; a scan
// gives the column pointers; only when the scanned total is below the
// structural count (an exact zero was dropped) does a second pass rewrite
// each column's kept rows in ascending order at the positions the ballots
// give.  Everything is in a fixed order: the output is the same on every
// run.
//
// psx_free_residual forms r = f_ext_free - f_int[free]
// (nonlinear_static.py:322-323) and sum(r^2) in FP64 with a fixed-order
// two-level reduction.
#include "psx_device.h"
#include "psx_internal.h"

namespace psx {

constexpr int kCscBlock = 256;

// Columns per thread in the scan of the kept counts.
constexpr int kScanItems = 8;
constexpr int kScanColumns = kCscBlock * kScanItems;  // per work-group of the local scan
constexpr int kScanTopBlock = 1024;

// Exclusive prefix of v over the work-group (blockDim.x a multiple of 64, at
// most 1024): a shuffle scan per wave, then the wave totals through sh;
// *total receives the work-group's sum.
__device__ __forceinline__ int group_exclusive_scan(int v, int* sh, int* total) {
  const int lane = int(threadIdx.x & 63), w = int(threadIdx.x >> 6), nw = int(blockDim.x >> 6);
  int x = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int y = __shfl_up(x, off);
    if (lane >= off) x += y;
  }
  if (lane == 63) sh[w] = x;
  __syncthreads();
  int before = 0, all = 0;
  for (int i = 0; i < nw; ++i) {
    const int t = sh[i];
    before += i < w ? t : 0;
    all += t;
  }
  *total = all;
  return before + x - v;
}

// Column node b's free columns (one wave): kStructFill writes every
// structural entry (free row) at its structural position (the column starts
// with no entry dropped, in b's record) and colptr[column] receives the count
// of kept entries (nonzero, when exact zeros are dropped); when nothing is
// dropped -- the scan of the counts then ends at the structural count -- this
// is the whole output.  kCompactFill (only when something was dropped)
// rewrites the kept rows and values from colptr[column], the final start.
// The records carry the free indices as base + mask, so a wave's loads form
// a chain of three: b's record, its unit records, the K values.
enum CscPass { kStructFill = 0, kCompactFill = 1 };

__device__ __forceinline__ int free_index(int base, int mask, int r) {
  return (mask >> r) & 1 ? base + __popc(unsigned(mask) & ((1u << r) - 1u)) : -1;
}

template <int PASS>
__device__ __forceinline__ void csc_node(const DevCsc& c, int bi, int lane, const double* __restrict__ kval,
                                         int drop, int* __restrict__ colptr, int* __restrict__ rowind,
                                         double* __restrict__ vals) {
  const int4 h = load_uniform(c.node, 2 * bi), st = load_uniform(c.node, 2 * bi + 1);
  const int u0 = h.x, items = 3 * h.y;
  int jc[3], start[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) jc[k] = free_index(h.z, h.w, k);
  if (PASS == kStructFill) {
    start[0] = st.x;
    start[1] = st.y;
    start[2] = st.z;
  } else {
#pragma unroll
    for (int k = 0; k < 3; ++k) start[k] = jc[k] >= 0 ? colptr[jc[k]] : 0;
  }
  int kept[3] = {0, 0, 0}, placed[3] = {0, 0, 0};
  const unsigned long long below = (1ull << lane) - 1ull;
  for (int base = 0; base < items; base += 64) {
    const int l = base + lane;
    // The loads are unconditional (lanes past the units read the last unit
    // and its K row, in bounds): the compiler then issues them back to back.
    const int s = l / 3, r = l - 3 * s;
    const int4 un = c.unit[u0 + min(s, h.y - 1)];
    const int ri = free_index(un.z, l < items ? un.w : 0, r);  // a select: the load stays whole
    const double* src = kval + un.x + r * un.y;
    const double v[3] = {src[0], src[1], src[2]};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const bool structural = ri >= 0 && jc[k] >= 0;
      const bool keep = structural && (!drop || v[k] != 0.0);
      const unsigned long long m = __ballot(keep);
      if (PASS == kStructFill) {
        const unsigned long long ms = __ballot(structural);
        if (structural) {
          const int o = start[k] + placed[k] + __popcll(ms & below);
          rowind[o] = ri;
          vals[o] = v[k];
        }
        placed[k] += __popcll(ms);
      } else if (keep) {
        const int o = start[k] + kept[k] + __popcll(m & below);
        rowind[o] = ri;
        vals[o] = v[k];
      }
      kept[k] += __popcll(m);
    }
  }
  if (PASS == kStructFill && lane == 0) {
#pragma unroll
    for (int k = 0; k < 3; ++k)
      if (jc[k] >= 0) colptr[jc[k]] = kept[k];
  }
}

// kCscStructNodes consecutive column nodes (of those with a free dof, in
// locality order) per wave; each XCD takes a contiguous run of the order, so
// the waves that read the rows of one node share that XCD's L2 (the grid is a
// multiple of 8 work-groups).  When every node of the wave has at most 21
// neighbours (one 64-lane chunk each), the nodes' unit records are loaded
// together, then their K values, then the entries are placed: each wave keeps
// kCscStructNodes times the loads of one node in flight.  Otherwise the nodes
// go through csc_node one after the other.
constexpr int kCscStructNodes = 4;
__global__ __launch_bounds__(kCscBlock) void csc_struct_kernel(DevCsc c, const double* __restrict__ kval, int drop,
                                                               int* __restrict__ colptr, int* __restrict__ rowind,
                                                               double* __restrict__ vals) {
  constexpr int NPW = kCscStructNodes;
  const int lane = int(threadIdx.x & 63);
  const int b0 = (xcd_tile() * (kCscBlock / 64) + __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6))) * NPW;
  if (b0 >= c.n_nodes) return;
  int4 h[NPW], st[NPW];
  bool one_chunk = true;
#pragma unroll
  for (int n = 0; n < NPW; ++n) {
    if (b0 + n < c.n_nodes) {
      h[n] = load_uniform(c.node, 2 * (b0 + n));
      st[n] = load_uniform(c.node, 2 * (b0 + n) + 1);
    } else {
      h[n] = make_int4(0, 0, 0, 0);  // no units, no free column
      st[n] = make_int4(0, 0, 0, 0);
    }
    one_chunk = one_chunk && 3 * h[n].y <= 64;
  }
  if (!one_chunk) {
    for (int n = 0; n < NPW && b0 + n < c.n_nodes; ++n)
      csc_node<kStructFill>(c, b0 + n, lane, kval, drop, colptr, rowind, vals);
    return;
  }
  // Unconditional loads (lanes past a node's units read its last unit and
  // that unit's K row; a node past the end reads unit 0), so that the
  // compiler issues every node's loads back to back.
  const int s = lane / 3, r = lane - 3 * s;
  int4 un[NPW];
#pragma unroll
  for (int n = 0; n < NPW; ++n) un[n] = c.unit[h[n].x + max(0, min(s, h[n].y - 1))];
  int ri[NPW];
  double v[NPW][3];
#pragma unroll
  for (int n = 0; n < NPW; ++n) {
    ri[n] = free_index(un[n].z, lane < 3 * h[n].y ? un[n].w : 0, r);  // a select: the load stays whole
    const double* src = kval + un[n].x + r * un[n].y;
    v[n][0] = src[0];
    v[n][1] = src[1];
    v[n][2] = src[2];
  }
  const unsigned long long below = (1ull << lane) - 1ull;
#pragma unroll
  for (int n = 0; n < NPW; ++n) {
    const int start[3] = {st[n].x, st[n].y, st[n].z};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const int jc = free_index(h[n].z, h[n].w, k);
      const bool structural = ri[n] >= 0 && jc >= 0;
      const bool keep = structural && (!drop || v[n][k] != 0.0);
      const unsigned long long ms = __ballot(structural), m = __ballot(keep);
      if (structural) {
        const int o = start[k] + __popcll(ms & below);
        rowind[o] = ri[n];
        vals[o] = v[n][k];
      }
      if (lane == 0 && jc >= 0) colptr[jc] = __popcll(m);
    }
  }
}

// kCscCompactNodes column nodes per wave; every wave returns at once when
// nothing was dropped.
constexpr int kCscCompactNodes = 16;
__global__ __launch_bounds__(kCscBlock) void csc_compact_kernel(DevCsc c, const double* __restrict__ kval, int drop,
                                                                int* __restrict__ colptr, int* __restrict__ rowind,
                                                                double* __restrict__ vals) {
  if (colptr[c.n_free] == c.nnz_struct) return;
  const int w = blockIdx.x * (kCscBlock / 64) + __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6));
  const int b0 = w * kCscCompactNodes, b1 = min(c.n_nodes, b0 + kCscCompactNodes);
  for (int bi = b0; bi < b1; ++bi) csc_node<kCompactFill>(c, bi, threadIdx.x & 63, kval, drop, colptr, rowind, vals);
}

// colptr[j] += the offset of its kScanColumns-column group (the group-local
// prefix becomes the column's start).
__global__ __launch_bounds__(kCscBlock) void csc_scan_add_kernel(int n_free, const int* __restrict__ blockoff,
                                                                 int* __restrict__ colptr) {
  const int j = blockIdx.x * kCscBlock + threadIdx.x;
  if (j < n_free) colptr[j] += blockoff[j / kScanColumns];
}

// colptr[j] (the column counts) -> the exclusive prefix within each group of
// kScanColumns columns, kScanItems consecutive columns per thread;
// blocksum[group] = the group's total.
__global__ __launch_bounds__(kCscBlock) void csc_scan_local_kernel(int n_free, int* __restrict__ colptr,
                                                                   int* __restrict__ blocksum) {
  __shared__ int sh[kCscBlock / 64];
  const int j0 = (blockIdx.x * kCscBlock + int(threadIdx.x)) * kScanItems;
  int v[kScanItems], sum = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    v[i] = j0 + i < n_free ? colptr[j0 + i] : 0;
    sum += v[i];
  }
  int total;
  int ex = group_exclusive_scan(sum, sh, &total);
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    if (j0 + i < n_free) colptr[j0 + i] = ex;
    ex += v[i];
  }
  if (threadIdx.x == 0) blocksum[blockIdx.x] = total;
}

// One work-group of kScanTopBlock threads: the group totals -> their
// exclusive prefix (each thread a consecutive run), and colptr[n_free] = the
// kept total.
__global__ __launch_bounds__(kScanTopBlock) void csc_scan_blocks_kernel(int nblocks, int* __restrict__ blocksum,
                                                                        int* __restrict__ colptr, int n_free) {
  __shared__ int sh[kScanTopBlock / 64];
  const int per = (nblocks + kScanTopBlock - 1) / kScanTopBlock;
  const int i0 = int(threadIdx.x) * per, i1 = min(nblocks, i0 + per);
  int sum = 0;
  for (int i = i0; i < i1; ++i) sum += blocksum[i];
  int total;
  int ex = group_exclusive_scan(sum, sh, &total);
  for (int i = i0; i < i1; ++i) {
    const int v = blocksum[i];
    blocksum[i] = ex;
    ex += v;
  }
  if (threadIdx.x == 0) colptr[n_free] = total;
}

int32_t launch_csc_extract(const DevCsc& c, const double* kval, int drop, int* colptr, int* rowind,
                           double* vals, hipStream_t s) {
  const int nb = (c.n_free + kScanColumns - 1) / kScanColumns;
  if (nb == 0) {
    // no free dof: only colptr[0] = 0 (the scan of zero groups writes it)
    hipLaunchKernelGGL(csc_scan_blocks_kernel, dim3(1), dim3(kScanTopBlock), 0, s, 0, c.scratch, colptr, 0);
    PSX_HIP_TRY(hipGetLastError());
    return PSX_OK;
  }
  // structural fill and kept counts; scan of the counts; the compacting fill
  // only if the scanned total shows that entries were dropped
  constexpr int waves = kCscBlock / 64;
  const int node_blocks = ((c.n_nodes + 8 * waves * kCscStructNodes - 1) / (8 * waves * kCscStructNodes)) * 8;
  const int compact_blocks = (c.n_nodes + waves * kCscCompactNodes - 1) / (waves * kCscCompactNodes);
  hipLaunchKernelGGL(csc_struct_kernel, dim3(node_blocks), dim3(kCscBlock), 0, s, c, kval, drop, colptr, rowind,
                     vals);
  hipLaunchKernelGGL(csc_scan_local_kernel, dim3(nb), dim3(kCscBlock), 0, s, c.n_free, colptr, c.scratch);
  hipLaunchKernelGGL(csc_scan_blocks_kernel, dim3(1), dim3(kScanTopBlock), 0, s, nb, c.scratch, colptr, c.n_free);
  hipLaunchKernelGGL(csc_scan_add_kernel, dim3((c.n_free + kCscBlock - 1) / kCscBlock), dim3(kCscBlock), 0, s,
                     c.n_free, c.scratch, colptr);
  hipLaunchKernelGGL(csc_compact_kernel, dim3(compact_blocks), dim3(kCscBlock), 0, s, c, kval, drop, colptr,
                     rowind, vals);
  PSX_HIP_TRY(hipGetLastError());
  return PSX_OK;
}

// r = f_ext_free - f_int[free]; partial sums of r^2 per block, then one block
// adds the partials in index order.
__global__ __launch_bounds__(kCscBlock) void residual_kernel(DevCsc c, const double* __restrict__ f_int,
                                                             const double* __restrict__ f_ext_free,
                                                             double* __restrict__ r,
                                                             double* __restrict__ partial) {
  __shared__ double sh[kCscBlock];
  const int j = blockIdx.x * kCscBlock + threadIdx.x;
  double v = 0.0;
  if (j < c.n_free) {
    const double x = (f_ext_free ? f_ext_free[j] : 0.0) - f_int[c.free[j]];
    r[j] = x;
    v = x * x;
  }
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int off = kCscBlock / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

__global__ __launch_bounds__(kCscBlock) void sum_partials_kernel(int n, const double* __restrict__ partial,
                                                                 double* __restrict__ out) {
  __shared__ double sh[kCscBlock];
  double v = 0.0;
  for (int i = threadIdx.x; i < n; i += kCscBlock) v += partial[i];
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int off = kCscBlock / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = sh[0];
}

int32_t launch_free_residual(const DevCsc& c, const double* f_int, const double* f_ext_free, double* r,
                             double* sumsq, hipStream_t s) {
  const int nb = (c.n_free + kCscBlock - 1) / kCscBlock;
  if (nb == 0) {
    // no free dof: sumsq = 0 (the sum of zero partials)
    hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(kCscBlock), 0, s, 0, c.partial, sumsq);
    PSX_HIP_TRY(hipGetLastError());
    return PSX_OK;
  }
  hipLaunchKernelGGL(residual_kernel, dim3(nb), dim3(kCscBlock), 0, s, c, f_int, f_ext_free, r,
                     c.partial);
  hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(kCscBlock), 0, s, nb, c.partial, sumsq);
  PSX_HIP_TRY(hipGetLastError());
  return PSX_OK;
}

}  // namespace psx
