// ECSW training matrix on the GPU (FP64), for Tet4 Neo-Hookean meshes.
//
// ECSWTrainingMatrix.build, single-domain path
// (src/psierra/hyperreduction/ecsw.py:171-365, loops at :319-365):
//   A[k r + i][e] = (Phi_e^T f_e(u_k))_i            snapshot k, mode i < r,
// and, on request, the upper triangle (numpy triu_indices order) of
// Phi_e^T K0_e Phi_e, K0_e the element tangent at u = 0, in rows m r .. .
// Phi_e = Phi[dofs of e, :] with Phi (3N x r, user DOF order, row-major) and
// snapshots (3N x m, row-major) as the reference's arrays.  The stiffness
// block is returned unscaled; the reference balances it by the norms of
// both blocks (:350-357), so the kernel also leaves the sum of squares of
// each block, and the caller scales.
//
// Layout of the work: a work-group of 256 threads takes 16 consecutive
// elements; thread t works for element t % 16, so the 16 lanes of each
// quarter-wave hold 16 consecutive elements and every store instruction
// writes four rows of A as four 128-byte runs.  Forces: (element, snapshot)
// pairs evaluate f_e(u_k) in FP64 into LDS; each thread then takes four
// modes i of its element, reads Phi_e[:, i] and writes (Phi_e^T f_k)_i for
// every snapshot.  Stiffness: K0_e, the tangent at u = 0, is Hooke's law
// with the element's (lambda, mu) and S = 0, so with G_i = sum_b phi_{i,b}
// h_b^T (the gradient of mode i over the element) and eps_i = sym(G_i),
//   (Phi_e^T K0_e Phi_e)_ij = V (lambda tr(G_i) tr(G_j) + 2 mu eps_i : eps_j),
// a short inner product: each thread forms eps of its four modes (6
// components) into LDS, then writes whole rows of the upper triangle (rows p
// and r - 1 - p for p = sub, sub + 16), element fastest, keeping eps_i in
// registers along row i.  Measured on the 48^3 box (rank 64, 8 snapshots +
// stiffness): 4.2 ms for A's 13.8 GB (3.3 TB/s), against 11.5 ms for an
// element-major work block and a transpose; a write-only kernel of the same
// layout takes 2.1 ms.  Each thread keeps the sums of squares of its force
// and stiffness values; the work-group adds them in a fixed tree into two
// partials per group, and a second kernel adds the partials in a fixed
// order, so the norms are deterministic.
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
constexpr int kTrainGroup = 16;       // elements per work-group
constexpr int kTrainThreads = 256;
constexpr int kTrainSnapPass = 16;    // snapshots whose forces are staged together

// Position of strain component q of mode i of element el in LDS: two modes
// share a row of 32 doubles (mode 2m at 0..15, mode 2m + 1 at 16..31), so the
// quarter-waves of one wave, reading consecutive modes, use different banks.
__device__ __forceinline__ int strain_pos(int i, int q, int el) {
  return ((i >> 1) * 6 + q) * (2 * kTrainGroup) + (i & 1) * kTrainGroup + el;
}

__device__ __forceinline__ void train_flag(int32_t* status, int e) {
  atomicMin(&status[1], e);
  status[0] = 1;
}

// Element groups of the training matrix (and partial-sum pairs per call).
__host__ __device__ constexpr int64_t train_groups(int64_t E) { return (E + kTrainGroup - 1) / kTrainGroup; }

__global__ __launch_bounds__(kTrainThreads) __attribute__((amdgpu_waves_per_eu(3))) void ecsw_train_kernel(DevPlan p, const double* __restrict__ phi, int r,
                                                                   const double* __restrict__ snaps, int m,
                                                                   const double* __restrict__ mu,
                                                                   const double* __restrict__ lam, int stiff,
                                                                   double* __restrict__ A, int64_t groups,
                                                                   double* __restrict__ partial,
                                                                   int32_t* status) {
  // forces of a snapshot pass, element fastest: [snapshot][dof][element], so
  // the 16 elements of a quarter-wave read 16 consecutive doubles; afterwards
  // the strain vectors of every mode at strain_pos(mode, component, element)
  constexpr int kForce = kTrainGroup * kTrainSnapPass * 12;
  constexpr int kStrain = kTrainMaxRank * 6 * kTrainGroup;
  __shared__ double s_buf[kForce > kStrain ? kForce : kStrain];
  __shared__ double s_red[2][kTrainThreads];
  const int t = threadIdx.x;
  const int el = t % kTrainGroup, sub = t / kTrainGroup;   // element, and which of the 16 sub-lanes
  // XCD-contiguous groups: the groups of one XCD are consecutive
  const int64_t per = int64_t(gridDim.x) >> 3;
  const int64_t g = (int64_t(blockIdx.x) & 7) * per + (int64_t(blockIdx.x) >> 3);
  const int64_t E = p.E;
  const int64_t e0 = g * kTrainGroup;
  const int64_t e = e0 + el;
  const bool valid = g < groups && e < E;
  int64_t un[4] = {0, 0, 0, 0};
  double4 X[4];
  double mue = 0.0, Ke = 0.0, lame = 0.0;
  if (valid) {
    const int4 c = *reinterpret_cast<const int4*>(p.conn + 4 * e);
    const int ci[4] = {c.x, c.y, c.z, c.w};
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      un[b] = p.perm[ci[b]];
      X[b] = p.Xi[ci[b]];
    }
    mue = mu[e];
    lame = lam[e];
    Ke = lame + (2.0 / 3.0) * mue;
  }
  double sq_f = 0.0, sq_s = 0.0;
  // ---- force rows: A[k r + i][e] = (Phi_e^T f_e(u_k))_i
  for (int k0 = 0; k0 < m; k0 += kTrainSnapPass) {
    const int nk = min(kTrainSnapPass, m - k0);
    __syncthreads();   // the previous pass's forces are read
    for (int kk = sub; kk < nk; kk += kTrainThreads / kTrainGroup) {
      if (!valid) continue;
      const int64_t k = k0 + kk;
      double4 U[4];
#pragma unroll
      for (int b = 0; b < 4; ++b)
        U[b] = make_double4(snaps[(3 * un[b]) * m + k], snaps[(3 * un[b] + 1) * m + k],
                            snaps[(3 * un[b] + 2) * m + k], 0.0);
      double fe[12];
      const double J = tet4_element_force<double>(X, U, mue, Ke, fe);
      if (J <= 0.0) train_flag(status, int(e));
#pragma unroll
      for (int d = 0; d < 12; ++d) s_buf[(kk * 12 + d) * kTrainGroup + el] = fe[d];
    }
    __syncthreads();
    if (valid) {
      for (int i = sub; i < r; i += kTrainThreads / kTrainGroup) {
        double ph[12];
#pragma unroll
        for (int d = 0; d < 12; ++d) ph[d] = phi[(3 * un[d / 3] + d % 3) * r + i];
        double* dst = A + (int64_t(k0) * r + i) * E + e;
        for (int kk = 0; kk < nk; ++kk, dst += int64_t(r) * E) {
          const double* fk = s_buf + kk * 12 * kTrainGroup + el;
          double cv = 0.0;
#pragma unroll
          for (int d = 0; d < 12; ++d) cv += ph[d] * fk[d * kTrainGroup];
          *dst = cv;
          sq_f += cv * cv;
        }
      }
    }
  }
  // ---- stiffness rows: upper triangle of Phi_e^T K0_e Phi_e, triu order
  if (stiff) {
    __syncthreads();   // the forces are read; s_buf takes the strain vectors
    double V = 0.0;
    if (valid) {
      const double4 Z[4] = {make_double4(0.0, 0.0, 0.0, 0.0), make_double4(0.0, 0.0, 0.0, 0.0),
                            make_double4(0.0, 0.0, 0.0, 0.0), make_double4(0.0, 0.0, 0.0, 0.0)};
      double h[4][3], A0[3][3];
      tet4_geometry<double>(X, Z, h, V, A0);
      for (int i = sub; i < r; i += kTrainThreads / kTrainGroup) {
        double ph[12];
#pragma unroll
        for (int d = 0; d < 12; ++d) ph[d] = phi[(3 * un[d / 3] + d % 3) * r + i];
        double G[3][3];
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
          for (int k = 0; k < 3; ++k)
            G[a][k] = ph[a] * h[0][k] + ph[3 + a] * h[1][k] + ph[6 + a] * h[2][k] + ph[9 + a] * h[3][k];
        // (eps_xx, eps_yy, eps_zz, eps_xy, eps_yz, eps_zx)
        const double v[6] = {G[0][0], G[1][1], G[2][2], 0.5 * (G[0][1] + G[1][0]), 0.5 * (G[1][2] + G[2][1]),
                             0.5 * (G[2][0] + G[0][2])};
#pragma unroll
        for (int q = 0; q < 6; ++q) s_buf[strain_pos(i, q, el)] = v[q];
      }
    }
    __syncthreads();
    const double wl = V * lame, wm = 2.0 * V * mue;
    if (valid) {
      // whole triangle rows per thread: rows p and r - 1 - p (together r + 1
      // entries) for the pairs p = sub, sub + 16, ..., so every thread has
      // the same number of entries and each row is a plain loop over j
      // (row i of the row-major upper triangle starts at entry
      // i r - i (i - 1) / 2)
      constexpr int kStep = kTrainThreads / kTrainGroup;
      for (int pidx = sub; pidx < (r + 1) / 2; pidx += kStep) {
#pragma unroll
        for (int side = 0; side < 2; ++side) {
          const int i = side == 0 ? pidx : r - 1 - pidx;
          if (side == 1 && i == pidx) break;   // odd rank: the middle row once
          double a[6];
#pragma unroll
          for (int q = 0; q < 6; ++q) a[q] = s_buf[strain_pos(i, q, el)];
          const double ta = (a[0] + a[1]) + a[2];
          const int64_t t0 = int64_t(i) * r - int64_t(i) * (i - 1) / 2;
          double* dst = A + (int64_t(m) * r + t0) * E + e;
#pragma unroll 4
          for (int j = i; j < r; ++j, dst += E) {
            double b[6];
#pragma unroll
            for (int q = 0; q < 6; ++q) b[q] = s_buf[strain_pos(j, q, el)];
            double ee = a[0] * b[0];
            ee += a[1] * b[1];
            ee += a[2] * b[2];
            double oo = a[3] * b[3];
            oo += a[4] * b[4];
            oo += a[5] * b[5];
            const double sv = wl * ta * ((b[0] + b[1]) + b[2]) + wm * (ee + 2.0 * oo);
            *dst = sv;
            sq_s += sv * sv;
          }
        }
      }
    }
  }
  // ---- sums of squares of the group, fixed tree over the threads
  if (partial) {
    s_red[0][t] = sq_f;
    s_red[1][t] = sq_s;
    __syncthreads();
    for (int stride = kTrainThreads / 2; stride > 0; stride >>= 1) {
      if (t < stride) {
        s_red[0][t] += s_red[0][t + stride];
        s_red[1][t] += s_red[1][t + stride];
      }
      __syncthreads();
    }
    if (t == 0 && g < groups) {
      partial[2 * g] = s_red[0][0];
      partial[2 * g + 1] = s_red[1][0];
    }
  }
}

// sumsq[0..1] = sums of the groups' partials, each lane over a strided
// subset in order, then a fixed tree.
__global__ __launch_bounds__(1024) void ecsw_train_norms_kernel(const double* __restrict__ partial, int64_t groups,
                                                                 double* __restrict__ sumsq) {
  __shared__ double red[2][1024];
  const int t = threadIdx.x;
  double a = 0.0, b = 0.0;
  for (int64_t g = t; g < groups; g += 1024) {
    a += partial[2 * g];
    b += partial[2 * g + 1];
  }
  red[0][t] = a;
  red[1][t] = b;
  __syncthreads();
  for (int stride = 512; stride > 0; stride >>= 1) {
    if (t < stride) {
      red[0][t] += red[0][t + stride];
      red[1][t] += red[1][t + stride];
    }
    __syncthreads();
  }
  if (t == 0) {
    sumsq[0] = red[0][0];
    sumsq[1] = red[1][0];
  }
}

int32_t launch_status_reset(hipStream_t s, int32_t* status);

int32_t launch_ecsw_training(const DevPlan& p, const double* phi, int r, const double* snaps, int m,
                             const double* mu, const double* lam, int stiff, double* A, double* partial,
                             double* sumsq, hipStream_t s, int32_t* status) {
  const int64_t rows = int64_t(m) * r + (stiff ? int64_t(r) * (r + 1) / 2 : 0);
  if (rows > INT_MAX / 2) return fail(PSX_EINVAL, "too many training rows");
  int32_t rc = launch_status_reset(s, status);
  if (rc) return rc;
  const int64_t groups = train_groups(p.E);
  if (groups > INT_MAX) return fail(PSX_EINVAL, "too many elements");
  // a multiple of 8 work-groups for the XCD-contiguous group map
  const int64_t blocks = 8 * ((groups + 7) / 8);
  if (groups > 0)
    hipLaunchKernelGGL(ecsw_train_kernel, dim3(unsigned(blocks)), dim3(kTrainThreads), 0, s, p, phi, r, snaps, m,
                       mu, lam, stiff, A, groups, sumsq ? partial : nullptr, status);
  if (sumsq && groups > 0) hipLaunchKernelGGL(ecsw_train_norms_kernel, dim3(1), dim3(1024), 0, s, partial, groups, sumsq);
  PSX_HIP_TRY(hipGetLastError());
  return PSX_OK;
}

}  // namespace psx

namespace psx {

// Scratch of psx_ecsw_training_noalloc: a file-scope device array (one copy
// per GPU, held until the process ends) in place of the caller's work array
// (the per-group partial sums of squares).
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

extern "C" int64_t psx_ecsw_training_work_doubles(const psx_plan* plan) {
  return plan ? 2 * psx::train_groups(plan->d.E) : 0;
}

extern "C" int32_t psx_ecsw_training_noalloc(psx_plan* plan, const double* phi, int32_t rank,
                                             const double* snapshots, int32_t n_snapshots, const double* mu,
                                             const double* lam, int32_t include_stiffness, double* A,
                                             double* sumsq, void* stream, int32_t* status) {
  psx::clear_error();
  if (!plan || !phi || !snapshots || !mu || !lam || !A || !status)
    return psx::fail(PSX_EINVAL, "null argument to psx_ecsw_training_noalloc");
  if (plan->hp.npe != 4) return psx::fail(PSX_EUNSUPPORTED, "psx_ecsw_training_noalloc supports Tet4 plans");
  if (rank < 1 || rank > psx::kTrainMaxRank)
    return psx::fail(PSX_EUNSUPPORTED, "psx_ecsw_training_noalloc supports ranks 1..64");
  if (n_snapshots < 1) return psx::fail(PSX_EINVAL, "need at least one snapshot");
  const int64_t need = 2 * psx::train_groups(plan->d.E);
  if (need > psx::kNoAllocScratchDoubles)
    return psx::fail(PSX_EUNSUPPORTED, "psx_ecsw_training_noalloc: " + std::to_string(plan->d.E) +
                                           " elements exceed its 32 MiB scratch of partial sums (at most " +
                                           std::to_string(psx::kNoAllocScratchDoubles / 2 * psx::kTrainGroup) +
                                           " elements)");
  int dev = 0;
  PSX_HIP_TRY(hipGetDevice(&dev));
  if (dev != plan->device)
    return psx::fail(PSX_EINVAL, "psx_ecsw_training_noalloc: the plan's device is not the current device");
  double* partial = nullptr;
  const int32_t rc = psx::scratch_on_current_device(&partial);
  if (rc) return rc;
  return psx::launch_ecsw_training(plan->d, phi, rank, snapshots, n_snapshots, mu, lam, include_stiffness != 0, A,
                                   partial, sumsq, static_cast<hipStream_t>(stream), status);
}
