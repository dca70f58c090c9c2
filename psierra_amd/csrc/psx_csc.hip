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
#include "psx_internal.h"

namespace psx {

constexpr int kCscBlock = 256;

__device__ __forceinline__ int block_exclusive_scan(int v, int* sh) {
  // Hillis-Steele over kCscBlock lanes; returns the exclusive prefix of v.
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int off = 1; off < kCscBlock; off <<= 1) {
    const int x = t >= off ? sh[t - off] : 0;
    __syncthreads();
    sh[t] += x;
    __syncthreads();
  }
  const int incl = sh[t];
  __syncthreads();
  return incl - v;
}

// Column node b's free columns (one wave): kStructFill writes every
// structural entry (free row) at its structural position (c.sstart: the
// column starts with no entry dropped) and colptr[column] receives the count
// of kept entries (nonzero, when exact zeros are dropped); when nothing is
// dropped -- the scan of the counts then ends at the structural count -- this
// is the whole output.  kCompactFill (only when something was dropped)
// rewrites the kept rows and values from colptr[column], the final start.
enum CscPass { kStructFill = 0, kCompactFill = 1 };

template <int PASS>
__device__ __forceinline__ void csc_node(const DevCsc& c, int b, int lane, const double* __restrict__ kval,
                                         int drop, int* __restrict__ colptr, int* __restrict__ rowind,
                                         double* __restrict__ vals) {
  int jc[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) jc[k] = c.newidx[3 * b + k];
  if (jc[0] < 0 && jc[1] < 0 && jc[2] < 0) return;
  const int u0 = c.adj_ptr[b], items = 3 * (c.adj_ptr[b + 1] - u0);
  int start[3] = {0, 0, 0};
#pragma unroll
  for (int k = 0; k < 3; ++k)
    if (jc[k] >= 0) start[k] = PASS == kStructFill ? c.sstart[jc[k]] : colptr[jc[k]];
  int kept[3] = {0, 0, 0}, placed[3] = {0, 0, 0};
  const unsigned long long below = (1ull << lane) - 1ull;
  for (int base = 0; base < items; base += 64) {
    const int l = base + lane;
    int ri = -1;
    double v[3] = {0.0, 0.0, 0.0};
    if (l < items) {
      const int s = l / 3, r = l - 3 * s;
      const int4 un = c.unit[u0 + s];
      ri = c.newidx[3 * un.z + r];
      if (ri >= 0) {
        const double* src = kval + un.x + r * un.y;
        v[0] = src[0];
        v[1] = src[1];
        v[2] = src[2];
      }
    }
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

// One wave per column node, taken in locality order.
__global__ __launch_bounds__(kCscBlock) void csc_struct_kernel(DevCsc c, const double* __restrict__ kval, int drop,
                                                               int* __restrict__ colptr, int* __restrict__ rowind,
                                                               double* __restrict__ vals) {
  const int bi = blockIdx.x * (kCscBlock / 64) + int(threadIdx.x >> 6);
  if (bi >= c.N) return;
  csc_node<kStructFill>(c, c.order[bi], threadIdx.x & 63, kval, drop, colptr, rowind, vals);
}

// kCscCompactNodes column nodes per wave; every wave returns at once when
// nothing was dropped.
constexpr int kCscCompactNodes = 16;
__global__ __launch_bounds__(kCscBlock) void csc_compact_kernel(DevCsc c, const double* __restrict__ kval, int drop,
                                                                int* __restrict__ colptr, int* __restrict__ rowind,
                                                                double* __restrict__ vals) {
  if (colptr[c.n_free] == c.nnz_struct) return;
  const int w = blockIdx.x * (kCscBlock / 64) + int(threadIdx.x >> 6);
  const int b0 = w * kCscCompactNodes, b1 = min(c.N, b0 + kCscCompactNodes);
  for (int bi = b0; bi < b1; ++bi) csc_node<kCompactFill>(c, c.order[bi], threadIdx.x & 63, kval, drop, colptr, rowind, vals);
}

// colptr[j] += the offset of its 256-column block (the block-local prefix
// becomes the column's start).
__global__ __launch_bounds__(kCscBlock) void csc_scan_add_kernel(int n_free, const int* __restrict__ blockoff,
                                                                 int* __restrict__ colptr) {
  const int j = blockIdx.x * kCscBlock + threadIdx.x;
  if (j < n_free) colptr[j] += blockoff[j / kCscBlock];
}

// colptr[j] (the column counts) -> the exclusive prefix within each block of
// 256 columns; blocksum[block] = the block's total.
__global__ __launch_bounds__(kCscBlock) void csc_scan_local_kernel(int n_free, int* __restrict__ colptr,
                                                                   int* __restrict__ blocksum) {
  __shared__ int sh[kCscBlock];
  const int j = blockIdx.x * kCscBlock + threadIdx.x;
  const int v = j < n_free ? colptr[j] : 0;
  const int ex = block_exclusive_scan(v, sh);
  if (j < n_free) colptr[j] = ex;
  if (threadIdx.x == kCscBlock - 1) blocksum[blockIdx.x] = ex + v;
}

__global__ __launch_bounds__(kCscBlock) void csc_scan_blocks_kernel(int nblocks, int* __restrict__ blocksum,
                                                                    int* __restrict__ colptr, int n_free) {
  __shared__ int sh[kCscBlock];
  int carry = 0;
  for (int base = 0; base < nblocks; base += kCscBlock) {
    const int i = base + threadIdx.x;
    const int v = i < nblocks ? blocksum[i] : 0;
    const int ex = block_exclusive_scan(v, sh);
    if (i < nblocks) blocksum[i] = carry + ex;
    __shared__ int total;
    if (threadIdx.x == kCscBlock - 1) total = ex + v;
    __syncthreads();
    carry += total;
    __syncthreads();
  }
  if (threadIdx.x == 0) colptr[n_free] = carry;
}

int32_t launch_csc_extract(const DevCsc& c, const double* kval, int drop, int* colptr, int* rowind,
                           double* vals, hipStream_t s) {
  const int nb = (c.n_free + kCscBlock - 1) / kCscBlock;
  if (nb == 0) {
    // no free dof: only colptr[0] = 0 (the scan of zero blocks writes it)
    hipLaunchKernelGGL(csc_scan_blocks_kernel, dim3(1), dim3(kCscBlock), 0, s, 0, c.scratch, colptr, 0);
    PSX_HIP_TRY(hipGetLastError());
    return PSX_OK;
  }
  // structural fill and kept counts; scan of the counts; the compacting fill
  // only if the scanned total shows that entries were dropped
  constexpr int waves = kCscBlock / 64;
  const int node_blocks = (c.N + waves - 1) / waves;
  const int compact_blocks = (c.N + waves * kCscCompactNodes - 1) / (waves * kCscCompactNodes);
  hipLaunchKernelGGL(csc_struct_kernel, dim3(node_blocks), dim3(kCscBlock), 0, s, c, kval, drop, colptr, rowind,
                     vals);
  hipLaunchKernelGGL(csc_scan_local_kernel, dim3(nb), dim3(kCscBlock), 0, s, c.n_free, colptr, c.scratch);
  hipLaunchKernelGGL(csc_scan_blocks_kernel, dim3(1), dim3(kCscBlock), 0, s, nb, c.scratch, colptr, c.n_free);
  hipLaunchKernelGGL(csc_scan_add_kernel, dim3(nb), dim3(kCscBlock), 0, s, c.n_free, c.scratch, colptr);
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
