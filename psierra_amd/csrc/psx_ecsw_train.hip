// ECSW training matrix on the GPU (FP64), for Tet4 Neo-Hookean meshes.
//
// ECSWTrainingMatrix.build, single-domain path
// (src/psierra/hyperreduction/ecsw.py:171-365, loops at :319-365):
//   A[k r + i][e] = (Phi_e^T f_e(u_k))_i            snapshot k, mode i < r,
// and, on request, the upper triangle (numpy triu_indices order) of
// Phi_e^T K0_e Phi_e, K0_e the element tangent at u = 0, in rows m r .. .
// Phi_e = Phi[dofs of e, :] with Phi (3N x r, user DOF order, row-major) and
// snapshots (3N x m, row-major) as the reference's arrays.  The stiffness
// block is returned unscaled (the caller applies the reference's norm
// balancing, :350-357, which needs global norms of both blocks).
//
// Layout of the work: one wave per element.  Lanes = snapshots evaluate the
// element force in FP64 and leave it in LDS; lanes = modes then hold the
// element's 12 basis rows (one coalesced 8-byte load per row) and form
// Phi_e^T f for every snapshot, writing an element-major row of the work
// array (consecutive lanes, consecutive values).  The stiffness block: four
// lanes evaluate the rows of K0_e, the wave stages Phi_e in LDS, lane j forms
// T = K0_e Phi_e[:, j] and the column j of Phi_e^T T above the diagonal.  A
// tiled transpose through LDS turns the element-major work array into the
// reference's (rows x E) layout.
#include <algorithm>
#include <climits>
#include <mutex>
#include <string>
#include <vector>

#include "psx_device.h"
#include "psx_internal.h"
#include "psx_nh.h"
#include "psx_tet4_elem.h"

namespace psx {

constexpr int kTrainMaxRank = 64;
constexpr int kTrainSnapPass = 64;

__device__ __forceinline__ void train_flag(int32_t* status, int e) {
  atomicMin(&status[1], e);
  status[0] = 1;
}

__global__ __launch_bounds__(256) void ecsw_train_force_kernel(DevPlan p, const double* __restrict__ phi,
                                                               int r, const double* __restrict__ snaps, int m,
                                                               const double* __restrict__ mu,
                                                               const double* __restrict__ lam, int rows,
                                                               int64_t e0, int64_t e1, double* __restrict__ At,
                                                               int32_t* status) {
  __shared__ double s_f[4][kTrainSnapPass][13];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int64_t base = e0 + int64_t(blockIdx.x) * 4; base < e1; base += int64_t(gridDim.x) * 4) {
    const int64_t e = base + wave;
    const bool valid = e < e1;
    int64_t un[4] = {0, 0, 0, 0};
    double4 X[4];
    double ph[12];
    double mue = 0.0, Ke = 0.0;
    if (valid) {
      const int4 c = *reinterpret_cast<const int4*>(p.conn + 4 * e);
      const int ci[4] = {c.x, c.y, c.z, c.w};
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        un[b] = p.perm[ci[b]];
        X[b] = p.Xi[ci[b]];
      }
      if (lane < r)
#pragma unroll
        for (int d = 0; d < 12; ++d) ph[d] = phi[(3 * un[d / 3] + d % 3) * r + lane];
      mue = mu[e];
      Ke = lam[e] + (2.0 / 3.0) * mue;
    }
    for (int k0 = 0; k0 < m; k0 += kTrainSnapPass) {
      const int nk = min(kTrainSnapPass, m - k0);
      if (valid && lane < nk) {
        const int64_t k = k0 + lane;
        double4 U[4];
#pragma unroll
        for (int b = 0; b < 4; ++b)
          U[b] = make_double4(snaps[(3 * un[b]) * m + k], snaps[(3 * un[b] + 1) * m + k],
                              snaps[(3 * un[b] + 2) * m + k], 0.0);
        double fe[12];
        const double J = tet4_element_force<double>(X, U, mue, Ke, fe);
        if (J <= 0.0) train_flag(status, int(e));
#pragma unroll
        for (int d = 0; d < 12; ++d) s_f[wave][lane][d] = fe[d];
      }
      __syncthreads();
      if (valid && lane < r) {
        double* dst = At + (e - e0) * rows + int64_t(k0) * r + lane;
        for (int kk = 0; kk < nk; ++kk) {
          double cv = 0.0;
#pragma unroll
          for (int d = 0; d < 12; ++d) cv += ph[d] * s_f[wave][kk][d];
          dst[int64_t(kk) * r] = cv;
        }
      }
      __syncthreads();
    }
  }
}

__global__ __launch_bounds__(256) void ecsw_train_stiff_kernel(DevPlan p, const double* __restrict__ phi,
                                                               int r, const double* __restrict__ mu,
                                                               const double* __restrict__ lam, int rows,
                                                               int row0, int64_t e0, int64_t e1,
                                                               double* __restrict__ At, int32_t* status) {
  __shared__ double s_phi[4][12][kTrainMaxRank];
  __shared__ double s_K[4][12][13];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int64_t base = e0 + int64_t(blockIdx.x) * 4; base < e1; base += int64_t(gridDim.x) * 4) {
    const int64_t e = base + wave;
    const bool valid = e < e1;
    if (valid) {
      const int4 c = *reinterpret_cast<const int4*>(p.conn + 4 * e);
      const int ci[4] = {c.x, c.y, c.z, c.w};
      if (lane < r)
#pragma unroll
        for (int d = 0; d < 12; ++d) s_phi[wave][d][lane] = phi[(3 * int64_t(p.perm[ci[d / 3]]) + d % 3) * r + lane];
      if (lane < 4) {
        // row `a` of K0_e: blocks in the order [a, others ascending]
        const int a = lane;
        const int rot[4] = {a, a == 0 ? 1 : 0, a <= 1 ? 2 : 1, a <= 2 ? 3 : 2};
        double4 Xr[4], Ur[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int cq = rot[q] == 0 ? ci[0] : rot[q] == 1 ? ci[1] : rot[q] == 2 ? ci[2] : ci[3];
          Xr[q] = p.Xi[cq];
          Ur[q] = make_double4(0.0, 0.0, 0.0, 0.0);
        }
        const double mue = mu[e];
        double out[39];
        const double J = tet4_row<double, true>(Xr, Ur, mue, lam[e] + (2.0 / 3.0) * mue, out);
        if (J <= 0.0) train_flag(status, int(e));
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
          for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int cc = 0; cc < 3; ++cc) s_K[wave][3 * a + i][3 * rot[q] + cc] = out[9 * q + 3 * i + cc];
      }
    }
    __syncthreads();
    if (valid && lane < r) {
      const int j = lane;
      double T[12];
#pragma unroll
      for (int d = 0; d < 12; ++d) {
        double s = 0.0;
#pragma unroll
        for (int cc = 0; cc < 12; ++cc) s += s_K[wave][d][cc] * s_phi[wave][cc][j];
        T[d] = s;
      }
      double* dst = At + (e - e0) * rows + row0;
      for (int i = 0; i <= j; ++i) {
        double s = 0.0;
#pragma unroll
        for (int d = 0; d < 12; ++d) s += s_phi[wave][d][i] * T[d];
        dst[i * r - (i * (i - 1)) / 2 + (j - i)] = s;
      }
    }
    __syncthreads();
  }
}

// out[c][x] = in[x][c] for x < n, c < cols (out rows of leading dimension
// ldo), 32 x 32 tiles through LDS.
__global__ __launch_bounds__(256) void transpose_f64_kernel(const double* __restrict__ in, int64_t n, int cols,
                                                            double* __restrict__ out, int64_t ldo) {
  __shared__ double tile[32][33];
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const int64_t r0 = int64_t(blockIdx.x) * 32;
  const int c0 = blockIdx.y * 32;
  for (int y = ty; y < 32; y += 8) {
    const int64_t rr = r0 + y;
    const int cc = c0 + tx;
    if (rr < n && cc < cols) tile[y][tx] = in[rr * cols + cc];
  }
  __syncthreads();
  for (int y = ty; y < 32; y += 8) {
    const int cc = c0 + y;
    const int64_t rr = r0 + tx;
    if (rr < n && cc < cols) out[int64_t(cc) * ldo + rr] = tile[tx][y];
  }
}

int32_t launch_status_reset(hipStream_t s, int32_t* status);

int32_t launch_ecsw_training(const DevPlan& p, const double* phi, int r, const double* snaps, int m,
                             const double* mu, const double* lam, int stiff, double* A, double* work,
                             int64_t work_elements, hipStream_t s, int32_t* status) {
  const int64_t rows = int64_t(m) * r + (stiff ? int64_t(r) * (r + 1) / 2 : 0);
  if (rows > INT_MAX / 2) return fail(PSX_EINVAL, "too many training rows");
  if (work_elements < 1) return fail(PSX_EINVAL, "work_elements must be at least 1");
  int32_t rc = launch_status_reset(s, status);
  if (rc) return rc;
  // Elements in chunks of work_elements columns: each chunk's element-major
  // block is transposed into its columns of A before the next chunk reuses
  // it, so the scratch is rows x work_elements instead of rows x E.
  // Chunks of a multiple of 64 elements start A's columns on a 512-byte
  // boundary; a chunk that starts inside a cache line makes every row
  // segment of the transpose straddle one more line (3.3 instead of 5 TB/s).
  const int64_t E = p.E;
  const int64_t aligned = work_elements >= 64 ? work_elements & ~int64_t(63) : work_elements;
  const int64_t chunk = std::min<int64_t>(aligned, std::max<int64_t>(E, 1));
  for (int64_t e0 = 0; e0 < E; e0 += chunk) {
    const int64_t e1 = std::min(E, e0 + chunk), n = e1 - e0;
    const int64_t groups = std::min<int64_t>((n + 3) / 4, 1 << 16);
    hipLaunchKernelGGL(ecsw_train_force_kernel, dim3(unsigned(groups)), dim3(256), 0, s, p, phi, r, snaps, m, mu,
                       lam, int(rows), e0, e1, work, status);
    if (stiff)
      hipLaunchKernelGGL(ecsw_train_stiff_kernel, dim3(unsigned(groups)), dim3(256), 0, s, p, phi, r, mu, lam,
                         int(rows), m * r, e0, e1, work, status);
    hipLaunchKernelGGL(transpose_f64_kernel, dim3(unsigned((n + 31) / 32), unsigned((rows + 31) / 32)), dim3(256),
                       0, s, work, n, int(rows), A + e0, E);
    PSX_HIP_TRY(hipGetLastError());
  }
  return PSX_OK;
}

}  // namespace psx

namespace psx {

// Scratch of psx_ecsw_training_noalloc: a file-scope device array (one copy
// per GPU, held until the process ends) in place of the caller's work block.
// Host code never passes `scratch` itself to a kernel (on the host it names
// a stand-in, not device memory): each device's address comes from
// hipGetSymbolAddress on that device's first call and is kept here.
constexpr int64_t kNoAllocScratchDoubles = int64_t(4) << 20;   // 32 MiB
static __device__ __align__(256) double scratch[kNoAllocScratchDoubles];

namespace {
std::mutex g_scratch_mutex;
std::vector<double*> g_scratch_ptr;   // per device id, nullptr until looked up

int32_t scratch_on_current_device(double** out) {
  int dev = 0;
  PSX_HIP_TRY(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (size_t(dev) >= g_scratch_ptr.size()) g_scratch_ptr.resize(size_t(dev) + 1, nullptr);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    PSX_HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return fail(PSX_ENOTREADY, "hipGetSymbolAddress returned no address for the scratch array");
    g_scratch_ptr[dev] = static_cast<double*>(ptr);
  }
  *out = g_scratch_ptr[dev];
  return PSX_OK;
}
}  // namespace

}  // namespace psx

extern "C" int32_t psx_ecsw_training_noalloc(psx_plan* plan, const double* phi, int32_t rank,
                                             const double* snapshots, int32_t n_snapshots, const double* mu,
                                             const double* lam, int32_t include_stiffness, double* A,
                                             void* stream, int32_t* status) {
  psx::clear_error();
  if (!plan || !phi || !snapshots || !mu || !lam || !A || !status)
    return psx::fail(PSX_EINVAL, "null argument to psx_ecsw_training_noalloc");
  if (plan->hp.npe != 4) return psx::fail(PSX_EUNSUPPORTED, "psx_ecsw_training_noalloc supports Tet4 plans");
  if (rank < 1 || rank > psx::kTrainMaxRank)
    return psx::fail(PSX_EUNSUPPORTED, "psx_ecsw_training_noalloc supports ranks 1..64");
  if (n_snapshots < 1) return psx::fail(PSX_EINVAL, "need at least one snapshot");
  const int64_t rows = int64_t(n_snapshots) * rank + (include_stiffness ? int64_t(rank) * (rank + 1) / 2 : 0);
  if (rows > psx::kNoAllocScratchDoubles)
    return psx::fail(PSX_EUNSUPPORTED, "psx_ecsw_training_noalloc: " + std::to_string(rows) +
                                           " training rows exceed its 32 MiB scratch (at most " +
                                           std::to_string(psx::kNoAllocScratchDoubles) + " rows)");
  int dev = 0;
  PSX_HIP_TRY(hipGetDevice(&dev));
  if (dev != plan->device)
    return psx::fail(PSX_EINVAL, "psx_ecsw_training_noalloc: the plan's device is not the current device");
  double* work = nullptr;
  const int32_t rc = psx::scratch_on_current_device(&work);
  if (rc) return rc;
  // whole elements of the scratch, in multiples of 64 when at least 64 fit
  const int64_t work_elements = psx::kNoAllocScratchDoubles / rows;
  return psx::launch_ecsw_training(plan->d, phi, rank, snapshots, n_snapshots, mu, lam, include_stiffness != 0, A,
                                   work, work_elements, static_cast<hipStream_t>(stream), status);
}
