// Batched ECSW hyper-reduced force and tangent on FP32 MFMA.
//
// For each instance (a mesh, a reduced basis V of rank r <= 64 and a reduced
// integration domain RID with weights w_e) and reduced coordinates u_r:
//   f_r = sum_{e in RID} w_e Phi_e^T f_e(Phi_e u_r)
//   K_r = sym( sum_{e in RID} w_e Phi_e^T k_e(Phi_e u_r) Phi_e )
// as HyperReducedECSWModel.evaluate_forces_and_stiffness does per element
// (src/psierra/hyperreduction/model.py:254-318, Phi_e cached at
// construction as in model.py:185-213).
//
// The tangent projection is one GEMM per instance: with A = [Phi_e] and
// B = [w_e k_e Phi_e] stacked over the RID (12 |RID| x 64 each),
// K_r = A^T B.  Every work item (an instance and a chunk of its RID) runs in
// a persistent launch.  Per super-batch of SB elements a work-group forms
// u_e = Phi_e u_r and the element rows of k_e and f_e (FP32, one lane per
// (element, row node), all SB elements at once); then per batch of EB
// elements it stages Phi_e in LDS (the next batch's Phi_e is already in
// flight into registers) and each of its four waves accumulates a 64x16
// column strip of A^T B with v_mfma_f32_16x16x4f32 (FP32 in, FP32
// accumulate), forming its own 16 columns of w k_e Phi_e in registers as
// the matrix-core operand.  Each work item writes its 64x64 partial; a
// second kernel adds the partials of an instance in a fixed order, 16x16
// tile pairs per work-group, symmetrises (model.py:317) and writes FP64.
// A basis of rank r < 64 is stored with its columns padded by zeros to the
// 64 of the matrix-core tiles: the padded rows and columns of K_r and f_r
// are exact zeros and are not written.
#include <algorithm>
#include <climits>
#include <cmath>
#include <string>
#include <vector>

#include "psx_internal.h"
#include "psx_tet4_elem.h"

namespace psx {

constexpr int kRank = 64;
constexpr int kEB = 8;        // elements staged per matrix-core batch
constexpr int kSB = 16;       // elements whose rows are evaluated together (4 lanes each)
constexpr int kChunk = 128;   // RID elements per work item
constexpr int kPhi = 12 * kRank;

typedef float floatx4 __attribute__((ext_vector_type(4)));

struct DevEcsw {
  int32_t n_inst, n_elem, n_items;
  int32_t rank;              // basis rank r (u_r, f_r, K_r strides); phi is padded to 64
  const float4* geom;        // [n_elem][5]: (h_b, 0) for b = 0..3, then (V w, mu, K, 0)
  const float* phi;          // [n_elem][12][64]
  const int32_t* item_inst;  // [n_items]
  const int32_t* item_e0;    // [n_items]
  const int32_t* item_cnt;   // [n_items]
  const int32_t* inst_item0; // [n_inst + 1]
  float* partK;              // [n_items][64][64]
  float* partF;              // [n_items][64]
};

// 35.6 KB of LDS per work-group lets four groups share a CU; the register
// budget is held to match (four waves per SIMD).
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 8))) void ecsw_project_kernel(
    DevEcsw d, const double* __restrict__ ur_all, int32_t* status) {
  constexpr int kPre = kEB * kPhi / 4 / 256;   // float4 of Phi per thread per batch
  __shared__ __align__(16) float s_phi[kEB * kPhi];
  __shared__ __align__(16) float s_k[kSB * 144];
  __shared__ float s_f[kSB * 12];
  __shared__ float s_ue[kSB * 12];
  __shared__ float s_ur[kRank];
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int wv = t >> 6;
  const int kq = lane >> 4, x = lane & 15;   // matrix-core operand row (k) and column
  for (int item = blockIdx.x; item < d.n_items; item += gridDim.x) {
    const int inst = d.item_inst[item];
    const int e0 = d.item_e0[item];
    const int cnt = d.item_cnt[item];
    float4 pre[kPre];
    {
      const int n4 = min(kEB, cnt) * (kPhi / 4);
      const float4* src = reinterpret_cast<const float4*>(d.phi + int64_t(e0) * kPhi);
#pragma unroll
      for (int i = 0; i < kPre; ++i) pre[i] = t + 256 * i < n4 ? src[t + 256 * i] : make_float4(0.f, 0.f, 0.f, 0.f);
    }
    __syncthreads();   // the previous item's reads of s_ur are done
    if (t < kRank) s_ur[t] = t < d.rank ? float(ur_all[int64_t(inst) * d.rank + t]) : 0.0f;
    floatx4 acc[4];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[mt][r] = 0.0f;
    float fr = 0.0f;
    for (int s0 = 0; s0 < cnt; s0 += kSB) {
      const int ns = min(kSB, cnt - s0);
      __syncthreads();   // s_ur is written; the previous super-batch's reads of s_k / s_f are done
      // u_e = Phi_e u_r for the super-batch, one (element, dof) per thread,
      // Phi rows read from global memory (they are staged again, from L2,
      // for the matrix-core batches below)
      for (int o = t; o < ns * 12; o += 256) {
        const float4* row = reinterpret_cast<const float4*>(d.phi + int64_t(e0 + s0) * kPhi + int64_t(o) * kRank);
        float s = 0.0f;
#pragma unroll 4
        for (int j4 = 0; j4 < kRank / 4; ++j4) {
          const float4 v = row[j4];
          s += v.x * s_ur[4 * j4];
          s += v.y * s_ur[4 * j4 + 1];
          s += v.z * s_ur[4 * j4 + 2];
          s += v.w * s_ur[4 * j4 + 3];
        }
        s_ue[o] = s;
      }
      __syncthreads();
      // the element rows of k_e and f_e, one (element, row node) per thread
      if (t < ns * 4) {
        const int el = t >> 2, a = t & 3;
        const int e = e0 + s0 + el;
        const int r1 = a == 0 ? 1 : 0, r2 = a <= 1 ? 2 : 1, r3 = a <= 2 ? 3 : 2;
        const int rot[4] = {a, r1, r2, r3};
        const float4 prm = d.geom[5 * int64_t(e) + 4];
        float h[4][3], U[4][3];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const float4 g = d.geom[5 * int64_t(e) + rot[r]];
          h[r][0] = g.x;
          h[r][1] = g.y;
          h[r][2] = g.z;
#pragma unroll
          for (int c = 0; c < 3; ++c) U[r][c] = s_ue[12 * el + 3 * rot[r] + c];
        }
        float A[3][3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int j = 0; j < 3; ++j)
            A[i][j] = (U[1][i] - U[0][i]) * h[1][j] + (U[2][i] - U[0][i]) * h[2][j] + (U[3][i] - U[0][i]) * h[3][j];
        float out[39];
        const float J = tet4_row_core<float, true>(h, prm.x, A, prm.y, prm.z, out);
        if (J <= 0.0f && a == 0) {
          atomicMin(&status[1], e);
          status[0] = 1;
        }
        float* krow = s_k + 144 * el + 12 * 3 * a;
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
          for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int c = 0; c < 3; ++c) krow[12 * i + 3 * rot[r] + c] = out[9 * r + 3 * i + c];
#pragma unroll
        for (int i = 0; i < 3; ++i) s_f[12 * el + 3 * a + i] = out[36 + i];
      }
      for (int b0 = s0; b0 < s0 + ns; b0 += kEB) {
        const int nb = min(kEB, cnt - b0);
        const int lb = b0 - s0;   // first element of the batch within the super-batch
        __syncthreads();   // the rows are written; the previous batch's reads of s_phi are done
#pragma unroll
        for (int i = 0; i < kPre; ++i)
          if (t + 256 * i < nb * (kPhi / 4)) reinterpret_cast<float4*>(s_phi)[t + 256 * i] = pre[i];
        __syncthreads();
        if (b0 + kEB < cnt) {
          // the next batch's Phi_e, in flight during this batch
          const int n4 = min(kEB, cnt - b0 - kEB) * (kPhi / 4);
          const float4* src = reinterpret_cast<const float4*>(d.phi + int64_t(e0 + b0 + kEB) * kPhi);
#pragma unroll
          for (int i = 0; i < kPre; ++i)
            pre[i] = t + 256 * i < n4 ? src[t + 256 * i] : make_float4(0.f, 0.f, 0.f, 0.f);
        }
        if (t < kRank) {
          for (int e = 0; e < nb; ++e)
#pragma unroll
            for (int r = 0; r < 12; ++r) fr += s_f[12 * (lb + e) + r] * s_phi[e * kPhi + r * kRank + t];
        }
        // A^T B on the matrix cores, wave wv owning columns 16 wv .. 16 wv + 15.
        // Per element, B's 12 x 16 block (w k_e) Phi_e[:, 16 wv ..] comes from
        // the matrix cores too: A1 = (w k_e) with its rows ordered so that
        // output row 4 q + s is k_e row 4 s + q (rows 12..15 zero), hence
        // register s of lane (kq, x) holds B[4 s + kq][16 wv + x] -- the B
        // operand of K-step s below.  Then per K-step ks (dof rows 4 ks ..
        // 4 ks + 3) and M-tile mt, A = Phi_e[4 ks + kq][4 x + mt] (one 16-byte
        // LDS read for the four tiles): tile mt accumulates
        // K_r[4 m + mt][16 wv + n].
        const int sx = x & 3, qx = x >> 2;   // A1 row x is k_e row 4 sx + qx
        for (int e = 0; e < nb; ++e) {
          const float* ph = s_phi + e * kPhi;
          const float* ke = s_k + 144 * (lb + e);
          floatx4 bq;
#pragma unroll
          for (int r = 0; r < 4; ++r) bq[r] = 0.0f;
#pragma unroll
          for (int j = 0; j < 3; ++j) {
            const float a1 = sx < 3 ? ke[12 * (4 * sx + qx) + 4 * j + kq] : 0.0f;
            const float b1 = ph[(4 * j + kq) * kRank + 16 * wv + x];
            bq = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b1, bq, 0, 0, 0);
          }
#pragma unroll
          for (int ks = 0; ks < 3; ++ks) {
            const float4 av = reinterpret_cast<const float4*>(ph + (4 * ks + kq) * kRank)[x];
            acc[0] = __builtin_amdgcn_mfma_f32_16x16x4f32(av.x, bq[ks], acc[0], 0, 0, 0);
            acc[1] = __builtin_amdgcn_mfma_f32_16x16x4f32(av.y, bq[ks], acc[1], 0, 0, 0);
            acc[2] = __builtin_amdgcn_mfma_f32_16x16x4f32(av.z, bq[ks], acc[2], 0, 0, 0);
            acc[3] = __builtin_amdgcn_mfma_f32_16x16x4f32(av.w, bq[ks], acc[3], 0, 0, 0);
          }
        }
      }
    }
    // 16x16 accumulator layout: register r of lane l holds row
    // m = 4 (l / 16) + r, column n = l % 16 of the tile, i.e. K_r row 4 m + mt.
    float* out = d.partK + int64_t(item) * kRank * kRank;
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
      for (int r = 0; r < 4; ++r) out[(16 * kq + 4 * r + mt) * kRank + 16 * wv + x] = acc[mt][r];
    if (t < kRank) d.partF[int64_t(item) * kRank + t] = fr;
  }
}

// K_r = 1/2 (S + S^T), S = sum of an instance's partials in item order, in
// FP64.  One work-group per (instance, pair of 16x16 tiles (I, J), I <= J):
// it sums tile (I, J) and tile (J, I) with coalesced row reads, transposes
// the second through LDS and writes both tiles.  The (0, 0) group also sums
// f_r.
__global__ __launch_bounds__(256) void ecsw_reduce_kernel(DevEcsw d, double* __restrict__ f_r,
                                                          double* __restrict__ K_r) {
  __shared__ double s_t[16][17];
  const int inst = blockIdx.x / 10;
  int pair = blockIdx.x % 10, I = 0;
  while (pair >= 4 - I) {
    pair -= 4 - I;
    ++I;
  }
  const int J = I + pair;
  const int t = threadIdx.x;
  const int i = t >> 4, j = t & 15;
  const int i0 = d.inst_item0[inst], i1 = d.inst_item0[inst + 1];
  double a = 0.0, b = 0.0;
  // the loads run ahead of the in-order FP64 sums
#pragma unroll 8
  for (int it = i0; it < i1; ++it) {
    const float* P = d.partK + int64_t(it) * kRank * kRank;
    a += double(P[(16 * I + i) * kRank + 16 * J + j]);
    b += double(P[(16 * J + i) * kRank + 16 * I + j]);
  }
  s_t[i][j] = b;   // S[16 J + i][16 I + j]
  __syncthreads();
  // element (16 I + i, 16 J + j): S_ij from a, S_ji = s_t[j][i]
  const double v = 0.5 * (a + s_t[j][i]);
  const int r = d.rank;
  double* K = K_r + int64_t(inst) * r * r;
  if (16 * I + i < r && 16 * J + j < r) {
    K[(16 * I + i) * r + 16 * J + j] = v;
    K[(16 * J + j) * r + 16 * I + i] = v;
  }
  if (I == 0 && J == 0 && t < r) {
    double s = 0.0;
#pragma unroll 8
    for (int it = i0; it < i1; ++it) s += double(d.partF[int64_t(it) * kRank + t]);
    f_r[int64_t(inst) * r + t] = s;
  }
}

}  // namespace psx

struct psx_ecsw {
  int device = 0;
  psx::DevEcsw d{};
  int64_t device_bytes = 0;
  int grid = 0;
  std::vector<void*> allocations;
};

namespace psx {
namespace {

int32_t ecsw_create(int32_t n_inst, const int64_t* rid_offsets, const double* Xe, const double* phi,
                    const double* weights, const double* mu, const double* lam, int32_t rank,
                    psx_ecsw** out) {
  if (!out) return fail(PSX_EINVAL, "out is null");
  *out = nullptr;
  if (rank < 1 || rank > kRank) return fail(PSX_EUNSUPPORTED, "the MFMA projection takes a basis rank of 1..64");
  if (n_inst <= 0 || !rid_offsets || !Xe || !phi || !weights || !mu || !lam)
    return fail(PSX_EINVAL, "null or empty ECSW input");
  if (rid_offsets[0] != 0) return fail(PSX_EINVAL, "rid_offsets[0] must be 0");
  for (int i = 0; i < n_inst; ++i)
    if (rid_offsets[i + 1] < rid_offsets[i]) return fail(PSX_EINVAL, "rid_offsets must be non-decreasing");
  const int64_t n_elem = rid_offsets[n_inst];
  if (n_elem <= 0 || n_elem >= (int64_t(1) << 31) / kPhi * 4)
    return fail(PSX_EINVAL, "bad RID size");
  std::vector<float4> geom(size_t(n_elem) * 5);
  for (int64_t e = 0; e < n_elem; ++e) {
    const double* X = Xe + 12 * e;
    double D[3][3];
    for (int b = 0; b < 3; ++b)
      for (int r = 0; r < 3; ++r) D[r][b] = X[3 * (b + 1) + r] - X[r];
    double adj[3][3];
    adj[0][0] = D[1][1] * D[2][2] - D[1][2] * D[2][1];
    adj[0][1] = D[0][2] * D[2][1] - D[0][1] * D[2][2];
    adj[0][2] = D[0][1] * D[1][2] - D[0][2] * D[1][1];
    adj[1][0] = D[1][2] * D[2][0] - D[1][0] * D[2][2];
    adj[1][1] = D[0][0] * D[2][2] - D[0][2] * D[2][0];
    adj[1][2] = D[0][2] * D[1][0] - D[0][0] * D[1][2];
    adj[2][0] = D[1][0] * D[2][1] - D[1][1] * D[2][0];
    adj[2][1] = D[0][1] * D[2][0] - D[0][0] * D[2][1];
    adj[2][2] = D[0][0] * D[1][1] - D[0][1] * D[1][0];
    const double det = D[0][0] * adj[0][0] + D[0][1] * adj[1][0] + D[0][2] * adj[2][0];
    if (!(det > 0.0))
      return fail(PSX_EVOLUME, "Negative or zero element volume detected in RID element " +
                                   std::to_string(e) + ". Check element vertex ordering.");
    if (!(weights[e] > 0.0)) return fail(PSX_EINVAL, "All ECSW element weights must be strictly positive (w_e > 0).");
    double h[4][3];
    for (int b = 0; b < 3; ++b)
      for (int c = 0; c < 3; ++c) h[b + 1][c] = adj[b][c] / det;
    for (int c = 0; c < 3; ++c) h[0][c] = -(h[1][c] + h[2][c] + h[3][c]);
    const double K = lam[e] + (2.0 / 3.0) * mu[e];
    for (int b = 0; b < 4; ++b) geom[5 * e + b] = make_float4(float(h[b][0]), float(h[b][1]), float(h[b][2]), 0.f);
    geom[5 * e + 4] = make_float4(float(det / 6.0 * weights[e]), float(mu[e]), float(K), 0.f);
  }
  // Phi_e (12 x rank per element) rounded to FP32 and padded to 64 columns
  std::vector<float> phif(size_t(n_elem) * kPhi, 0.0f);
  for (int64_t e = 0; e < n_elem; ++e)
    for (int d = 0; d < 12; ++d)
      for (int c = 0; c < rank; ++c)
        phif[size_t(e) * kPhi + size_t(d) * kRank + c] = float(phi[(size_t(e) * 12 + d) * rank + c]);
  std::vector<int32_t> item_inst, item_e0, item_cnt, inst_item0{0};
  for (int i = 0; i < n_inst; ++i) {
    for (int64_t e = rid_offsets[i]; e < rid_offsets[i + 1]; e += kChunk) {
      item_inst.push_back(i);
      item_e0.push_back(int32_t(e));
      item_cnt.push_back(int32_t(std::min<int64_t>(kChunk, rid_offsets[i + 1] - e)));
    }
    inst_item0.push_back(int32_t(item_inst.size()));
  }
  auto* h = new psx_ecsw();
  auto bail = [&](int32_t rc) {
    for (void* p : h->allocations) (void)hipFree(p);
    delete h;
    return rc;
  };
  (void)hipGetDevice(&h->device);
  auto up = [&](const void* src, size_t bytes, void** dst) -> int32_t {
    PSX_HIP_TRY(hipMalloc(dst, std::max<size_t>(bytes, 16)));
    h->allocations.push_back(*dst);
    h->device_bytes += int64_t(bytes);
    if (src && bytes) PSX_HIP_TRY(hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice));
    return PSX_OK;
  };
  const size_t n_items = item_inst.size();
  void *dg, *dp, *di, *de, *dc, *d0, *dk, *df;
  int32_t rc;
  if ((rc = up(geom.data(), geom.size() * sizeof(float4), &dg)) ||
      (rc = up(phif.data(), phif.size() * 4, &dp)) || (rc = up(item_inst.data(), n_items * 4, &di)) ||
      (rc = up(item_e0.data(), n_items * 4, &de)) || (rc = up(item_cnt.data(), n_items * 4, &dc)) ||
      (rc = up(inst_item0.data(), inst_item0.size() * 4, &d0)) ||
      (rc = up(nullptr, n_items * kRank * kRank * 4, &dk)) || (rc = up(nullptr, n_items * kRank * 4, &df)))
    return bail(rc);
  h->d.n_inst = n_inst;
  h->d.rank = rank;
  h->d.n_elem = int32_t(n_elem);
  h->d.n_items = int32_t(n_items);
  h->d.geom = static_cast<const float4*>(dg);
  h->d.phi = static_cast<const float*>(dp);
  h->d.item_inst = static_cast<const int32_t*>(di);
  h->d.item_e0 = static_cast<const int32_t*>(de);
  h->d.item_cnt = static_cast<const int32_t*>(dc);
  h->d.inst_item0 = static_cast<const int32_t*>(d0);
  h->d.partK = static_cast<float*>(dk);
  h->d.partF = static_cast<float*>(df);
  int cus = 256;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, h->device) == hipSuccess && prop.multiProcessorCount > 0)
    cus = prop.multiProcessorCount;
  // as many resident work-groups as the kernel's LDS and registers allow
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, ecsw_project_kernel, 256, 0) != hipSuccess || per_cu < 1)
    per_cu = 2;
  h->grid = int(std::min<size_t>(n_items, size_t(cus) * size_t(per_cu)));
  *out = h;
  return PSX_OK;
}

}  // namespace

int32_t launch_status_reset(hipStream_t s, int32_t* status);

}  // namespace psx

extern "C" {

int32_t psx_ecsw_create(int32_t n_instances, const int64_t* rid_offsets, const double* X_e,
                        const double* phi_e, const double* weights, const double* mu, const double* lam,
                        int32_t rank, psx_ecsw** out) {
  psx::clear_error();
  return psx::ecsw_create(n_instances, rid_offsets, X_e, phi_e, weights, mu, lam, rank, out);
}

int32_t psx_ecsw_destroy(psx_ecsw* h) {
  psx::clear_error();
  if (!h) return PSX_OK;
  for (void* p : h->allocations) (void)hipFree(p);
  delete h;
  return PSX_OK;
}

int32_t psx_ecsw_info(const psx_ecsw* h, int64_t* info, int32_t n_info) {
  psx::clear_error();
  if (!h || !info) return psx::fail(PSX_EINVAL, "null argument");
  const int64_t v[5] = {h->d.n_inst, h->d.n_elem, h->d.n_items, h->grid, h->device_bytes};
  for (int i = 0; i < n_info && i < 5; ++i) info[i] = v[i];
  return PSX_OK;
}

int32_t psx_ecsw_evaluate(psx_ecsw* h, const double* u_r, double* f_r, double* K_r, void* stream,
                          int32_t* status) {
  psx::clear_error();
  if (!h || !u_r || !f_r || !K_r || !status) return psx::fail(PSX_EINVAL, "null argument");
  hipStream_t s = static_cast<hipStream_t>(stream);
  int32_t rc = psx::launch_status_reset(s, status);
  if (rc) return rc;
  hipLaunchKernelGGL(psx::ecsw_project_kernel, dim3(h->grid), dim3(256), 0, s, h->d, u_r, status);
  hipLaunchKernelGGL(psx::ecsw_reduce_kernel, dim3(10 * h->d.n_inst), dim3(256), 0, s, h->d, f_r, K_r);
  PSX_HIP_TRY(hipGetLastError());
  return PSX_OK;
}

}  // extern "C"
