// Free-DOF CSC extraction and the free-DOF residual for the Newton solve.
//
// pSierra builds the solver matrix as
//   K_total = K_tan + K_contact            (CSR '+', drops exact-zero sums)
//   K_tan_free = K_total[free][:, free].tocsc()
// (src/psierra/analysis/nonlinear_static.py:354-355) and hands it to SuperLU
// (src/psierra/solver/sparse_linear_solver.py:16-19).  Here the structural
// CSC of K[free][:, free] is built once on the host (psx_csc_create): column
// j' = free dof j holds, in ascending order, the free rows i that share an
// element with j, and the CSR position of K[i][j].  Per Newton iteration
// psx_csc_extract gathers the values, drops exact zeros (as the CSR '+'
// does) and compacts the columns with a deterministic three-pass scan.
//
// psx_free_residual forms r = f_ext_free - f_int[free]
// (nonlinear_static.py:322-323) and sum(r^2) in FP64 with a fixed-order
// two-level reduction.
#include "psx_internal.h"

namespace psx {

constexpr int kCscBlock = 256;

__device__ __forceinline__ int col_count(const DevCsc& c, const double* __restrict__ kval, int j,
                                         int drop) {
  const int b = c.colptr_s[j], e = c.colptr_s[j + 1];
  if (!drop) return e - b;
  int n = 0;
  for (int q = b; q < e; ++q) n += kval[c.src[q]] != 0.0;
  return n;
}

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

__global__ __launch_bounds__(kCscBlock) void csc_count_kernel(DevCsc c, const double* __restrict__ kval,
                                                              int drop, int* __restrict__ blocksum) {
  __shared__ int sh[kCscBlock];
  const int j = blockIdx.x * kCscBlock + threadIdx.x;
  const int v = j < c.n_free ? col_count(c, kval, j, drop) : 0;
  const int ex = block_exclusive_scan(v, sh);
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

__global__ __launch_bounds__(kCscBlock) void csc_fill_kernel(DevCsc c, const double* __restrict__ kval,
                                                             int drop, const int* __restrict__ blockoff,
                                                             int* __restrict__ colptr,
                                                             int* __restrict__ rowind,
                                                             double* __restrict__ vals) {
  __shared__ int sh[kCscBlock];
  const int j = blockIdx.x * kCscBlock + threadIdx.x;
  const int v = j < c.n_free ? col_count(c, kval, j, drop) : 0;
  const int start = blockoff[blockIdx.x] + block_exclusive_scan(v, sh);
  if (j >= c.n_free) return;
  colptr[j] = start;
  int o = start;
  for (int q = c.colptr_s[j]; q < c.colptr_s[j + 1]; ++q) {
    const double x = kval[c.src[q]];
    if (drop && x == 0.0) continue;
    rowind[o] = c.row_s[q];
    vals[o] = x;
    ++o;
  }
}

int32_t launch_csc_extract(const DevCsc& c, const double* kval, int drop, int* colptr, int* rowind,
                           double* vals, hipStream_t s) {
  const int nb = (c.n_free + kCscBlock - 1) / kCscBlock;
  if (nb == 0) return PSX_OK;
  hipLaunchKernelGGL(csc_count_kernel, dim3(nb), dim3(kCscBlock), 0, s, c, kval, drop, c.scratch);
  hipLaunchKernelGGL(csc_scan_blocks_kernel, dim3(1), dim3(kCscBlock), 0, s, nb, c.scratch, colptr,
                     c.n_free);
  hipLaunchKernelGGL(csc_fill_kernel, dim3(nb), dim3(kCscBlock), 0, s, c, kval, drop, c.scratch, colptr,
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
  if (nb == 0) return PSX_OK;
  hipLaunchKernelGGL(residual_kernel, dim3(nb), dim3(kCscBlock), 0, s, c, f_int, f_ext_free, r,
                     c.partial);
  hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(kCscBlock), 0, s, nb, c.partial, sumsq);
  PSX_HIP_TRY(hipGetLastError());
  return PSX_OK;
}

}  // namespace psx
