// Device arithmetic for the compressible Neo-Hookean Total Lagrangian path.
//
// Constitutive law (src/psierra/material/neohookean.py:13-15, 55-122):
//   Psi = mu/2 (J^-2/3 I1 - 3) + K/2 (ln J)^2,  K = lambda + 2 mu / 3,
//   S   = f (I - I1/3 C^-1) + K ln J C^-1,      f = mu J^-2/3,
//   C4  = a C^-1 (x) C^-1 + b I_{C^-1} + c (C^-1 (x) I + I (x) C^-1),
//   a = K + 2/9 f I1,  b = -2 K ln J + 2/3 f I1,  c = -2/3 f,
//   I_{C^-1}_ijkl = (C^-1_ik C^-1_jl + C^-1_il C^-1_jk) / 2.
//
// The element tangent B^T C B V + G^T (I3 x S) G V
// (src/psierra/element/tet4_total_lagrangian.py:127-131) is evaluated in
// closed form per node pair (a, b).  With h = grad_X N, g = F^-T h (the
// spatial gradient) and q = F h, the contraction of C4 with F h_a and F h_b
// collapses, because F C^-1 F^T = I, to
//   K_ab = V [ a g_a g_b^T + b/2 ((g_a . g_b) I + g_b g_a^T)
//            + c (g_a q_b^T + q_a g_b^T) + (h_a . S h_b) I ],
// and f_a = V F S h_a (= B^T S_voigt V, tet4_total_lagrangian.py:136-137).
//
// Small-strain accuracy: everything is formed from A = grad u, never from
// F - I: J - 1 = tr A + ((tr A)^2 - tr(A^2))/2 + det A, ln J = log1p(J - 1),
// E = (A + A^T + A^T A)/2 and S = C^-1 (2 f dev E + K ln J I) (which equals
// the form above because I - I1/3 C^-1 = 2 C^-1 dev E).  This keeps FP32
// element arithmetic within a few ulps of the stress itself even when the
// strain is small.
#pragma once

#include <hip/hip_runtime.h>

namespace psx {

// log1p in FP32 (a few ulp for x > -1): with u = fl(1 + x) and the rounding
// correction c (1 + x = u + c to first order, as in fdlibm's log1p),
// log1p(x) = e ln 2 + log(m) + c / u, where u = m 2^e, m in [sqrt(1/2),
// sqrt(2)), and log(m) = 2 atanh(s), s = (m - 1) / (m + 1), |s| <= 0.172,
// by its odd series through s^11 (truncation < 2e-9 relative).  Inputs of
// -1 or less (an inverted element) give a meaningless value; the caller
// flags det F <= 0 separately.
__device__ __forceinline__ float log1p_f32(float x) {
  const float u = 1.0f + x;
  const float c = u > 2.0f ? 1.0f - (u - x) : x - (u - 1.0f);
  int e = __builtin_amdgcn_frexp_expf(u);
  float m = __builtin_amdgcn_frexp_mantf(u);   // [0.5, 1)
  if (m < 0.70710678f) {
    m *= 2.0f;
    e -= 1;
  }
  const float s = (m - 1.0f) * __builtin_amdgcn_rcpf(m + 1.0f);
  const float z = s * s;
  float r = 1.0f / 11.0f;
  r = r * z + 1.0f / 9.0f;
  r = r * z + 1.0f / 7.0f;
  r = r * z + 1.0f / 5.0f;
  r = r * z + 1.0f / 3.0f;
  const float lm = 2.0f * s + (2.0f * s) * (z * r);
  return float(e) * 0.69314718f + (lm + c * __builtin_amdgcn_rcpf(u));
}

template <typename R>
__device__ __forceinline__ R d_log1p(R x);
template <>
__device__ __forceinline__ float d_log1p<float>(float x) { return log1p_f32(x); }
template <>
__device__ __forceinline__ double d_log1p<double>(double x) { return log1p(x); }

template <typename R>
__device__ __forceinline__ R d_exp(R x);
template <>
__device__ __forceinline__ float d_exp<float>(float x) { return __expf(x); }
template <>
__device__ __forceinline__ double d_exp<double>(double x) { return exp(x); }

// Reciprocal: IEEE division in FP64; the hardware v_rcp_f32 (1 ulp) in FP32,
// where the element arithmetic is only required to ~1e-7 relative anyway.
template <typename R>
__device__ __forceinline__ R d_rcp(R x);
template <>
__device__ __forceinline__ float d_rcp<float>(float x) { return __builtin_amdgcn_rcpf(x); }
template <>
__device__ __forceinline__ double d_rcp<double>(double x) { return 1.0 / x; }

template <typename R>
__device__ __forceinline__ R det3(const R M[3][3]) {
  return M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1]) -
         M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0]) +
         M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0]);
}

// adj(M): M^-1 = adj(M) / det(M).
template <typename R>
__device__ __forceinline__ void adj3(const R M[3][3], R o[3][3]) {
  o[0][0] = M[1][1] * M[2][2] - M[1][2] * M[2][1];
  o[0][1] = M[0][2] * M[2][1] - M[0][1] * M[2][2];
  o[0][2] = M[0][1] * M[1][2] - M[0][2] * M[1][1];
  o[1][0] = M[1][2] * M[2][0] - M[1][0] * M[2][2];
  o[1][1] = M[0][0] * M[2][2] - M[0][2] * M[2][0];
  o[1][2] = M[0][2] * M[1][0] - M[0][0] * M[1][2];
  o[2][0] = M[1][0] * M[2][1] - M[1][1] * M[2][0];
  o[2][1] = M[0][1] * M[2][0] - M[0][0] * M[2][1];
  o[2][2] = M[0][0] * M[1][1] - M[0][1] * M[1][0];
}

// Material state at one integration point.
template <typename R>
struct NHPoint {
  R F[3][3];   // deformation gradient I + A
  R Gm[3][3];  // F^-T
  R Ci[3][3];  // C^-1
  R S[3][3];   // 2nd Piola-Kirchhoff stress (symmetric)
  R al, be, ga;
  R J, lnJ;
  R jm23;      // J^-2/3
  R E[3][3];   // Green-Lagrange strain
  R trE;
};

// A = grad_X u (A[i][j] = du_i/dX_j); mu, K = shear and bulk modulus.
template <typename R>
__device__ __forceinline__ void nh_point(const R A[3][3], R mu, R K, NHPoint<R>& o) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) o.F[i][j] = i == j ? A[i][j] + R(1) : A[i][j];
  const R trA = A[0][0] + A[1][1] + A[2][2];
  R trA2 = R(0);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) trA2 += A[i][j] * A[j][i];
  const R Jm1 = trA + R(0.5) * (trA * trA - trA2) + det3(A);
  o.J = R(1) + Jm1;
  o.lnJ = d_log1p(Jm1);
  R adj[3][3];
  adj3(o.F, adj);
  const R iJ = d_rcp(o.J);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) o.Gm[i][j] = adj[j][i] * iJ;  // F^-T = adj(F)^T / J
  // C^-1 = F^-1 F^-T = Gm^T Gm
#pragma unroll
  for (int I = 0; I < 3; ++I)
#pragma unroll
    for (int L = I; L < 3; ++L) {
      R s = R(0);
#pragma unroll
      for (int i = 0; i < 3; ++i) s += o.Gm[i][I] * o.Gm[i][L];
      o.Ci[I][L] = s;
      o.Ci[L][I] = s;
    }
  // Green-Lagrange strain from A only
  R E[3][3];
#pragma unroll
  for (int I = 0; I < 3; ++I)
#pragma unroll
    for (int L = I; L < 3; ++L) {
      R s = A[I][L] + A[L][I];
#pragma unroll
      for (int k = 0; k < 3; ++k) s += A[k][I] * A[k][L];
      E[I][L] = R(0.5) * s;
      E[L][I] = E[I][L];
    }
  const R trE = E[0][0] + E[1][1] + E[2][2];
  o.jm23 = d_exp(R(-2.0 / 3.0) * o.lnJ);
  const R fac = mu * o.jm23;
#pragma unroll
  for (int I = 0; I < 3; ++I)
#pragma unroll
    for (int L = 0; L < 3; ++L) o.E[I][L] = E[I][L];
  o.trE = trE;
  const R I1 = R(3) + R(2) * trE;
  // Sbar = 2 f dev(E) + K lnJ I ;  S = sym(C^-1 Sbar)
  R Sb[3][3];
#pragma unroll
  for (int I = 0; I < 3; ++I)
#pragma unroll
    for (int L = 0; L < 3; ++L)
      Sb[I][L] = R(2) * fac * (E[I][L] - (I == L ? trE * R(1.0 / 3.0) : R(0))) +
                 (I == L ? K * o.lnJ : R(0));
  // C^-1 and Sbar commute, so C^-1 Sbar is symmetric: form its upper
  // triangle and mirror it.
#pragma unroll
  for (int I = 0; I < 3; ++I)
#pragma unroll
    for (int L = I; L < 3; ++L) {
      const R v = o.Ci[I][0] * Sb[0][L] + o.Ci[I][1] * Sb[1][L] + o.Ci[I][2] * Sb[2][L];
      o.S[I][L] = v;
      o.S[L][I] = v;
    }
  o.al = K + R(2.0 / 9.0) * fac * I1;
  o.be = R(-2) * K * o.lnJ + R(2.0 / 3.0) * fac * I1;
  o.ga = R(-2.0 / 3.0) * fac;
}

// Hyperelastic models on the same kinematics (all with the volumetric part
// K/2 (ln J)^2 and the isochoric invariants of C_bar = J^-2/3 C):
//   kNeoHookean    Psi = mu/2 (I1bar - 3)                      params (mu, K)
//   kMooneyRivlin  Psi = c10 (I1bar - 3) + c01 (I2bar - 3)     params (c10, c01, K)
//   kYeoh          Psi = sum_i ci0 (I1bar - 3)^i, i = 1..3     params (c10, c20, c30, K)
// (src/psierra/material/neohookean.py, mooney_rivlin.py, yeoh.py).  With
// f = 2 psi_1 J^-2/3 (psi_1 = dPsi/dI1bar) the Neo-Hookean terms above hold
// unchanged; the extra terms of the tangent push forward, with
// b = F F^T, p = I1 q - b q and z = 2 J^-2/3 (q - I1/3 g), to
//   Mooney-Rivlin (c' = c01 J^-4/3): alpha += 16/9 I2 c', beta += 8/3 I2 c',
//     K_ab += V c' [4 q_a q_b^T - 2 (b (h_a . h_b) + q_b q_a^T)
//                   - 8/3 (g_a p_b^T + p_a g_b^T)],
//   Yeoh (psi_11 = d2Psi/dI1bar2): K_ab += V psi_11 z_a z_b^T,
// and the Mooney-Rivlin stress adds 2 c' C^-1 (2 dev E + 4 tr(E) E - 4 E^2
// - 4/3 (tr(E)^2 - tr(E^2)) I), all formed from E so small strains keep
// their relative accuracy.
// kNeoHookeanBound: Neo-Hookean with (V mu, V K) carried in the FP32
// geometry records of a plan whose material was bound beforehand.
enum HyperModel { kNeoHookean = 0, kMooneyRivlin = 1, kYeoh = 2, kFbarNeoHookean = 3, kNeoHookeanBound = 4 };

template <typename R>
struct HyperExtra {
  R cq = R(0), cb = R(0), cp = R(0);   // Mooney-Rivlin block coefficients
  R psi11 = R(0), zs = R(0);           // Yeoh: psi_11 and 2 J^-2/3
  R I1 = R(3);
  R b[3][3];                           // F F^T (Mooney-Rivlin)
};

// Material point of model M from A = grad u; prm as listed above.
template <typename R, int M>
__device__ __forceinline__ void hyper_point(const R A[3][3], const R prm[4], NHPoint<R>& o, HyperExtra<R>& x) {
  if constexpr (M == kNeoHookean) {
    nh_point(A, prm[0], prm[1], o);
    return;
  } else {
    const R K = M == kMooneyRivlin ? prm[2] : prm[3];
    // mu placeholder: nh_point's isochoric part is rescaled below
    nh_point(A, R(1), K, o);
    const R t = o.trE;
    const R I1 = R(3) + R(2) * t;
    x.I1 = I1;
    R f;
    if constexpr (M == kMooneyRivlin) {
      f = R(2) * prm[0] * o.jm23;
    } else {
      const R d = R(3) * (o.jm23 - R(1)) + R(2) * t * o.jm23;   // I1bar - 3
      const R psi1 = prm[0] + R(2) * prm[1] * d + R(3) * prm[2] * d * d;
      x.psi11 = R(2) * prm[1] + R(6) * prm[2] * d;
      x.zs = R(2) * o.jm23;
      f = R(2) * psi1 * o.jm23;
    }
    // isochoric part of S for f (nh_point used f = J^-2/3): S = C^-1 Sbar
    R Sb[3][3];
#pragma unroll
    for (int I = 0; I < 3; ++I)
#pragma unroll
      for (int L = 0; L < 3; ++L)
        Sb[I][L] = R(2) * f * (o.E[I][L] - (I == L ? t * R(1.0 / 3.0) : R(0))) + (I == L ? K * o.lnJ : R(0));
    R al = K + R(2.0 / 9.0) * f * I1;
    R be = R(-2) * K * o.lnJ + R(2.0 / 3.0) * f * I1;
    if constexpr (M == kMooneyRivlin) {
      const R cpr = prm[1] * o.jm23 * o.jm23;
      R E2[3][3];
      R trE2 = R(0);
#pragma unroll
      for (int I = 0; I < 3; ++I)
#pragma unroll
        for (int L = 0; L < 3; ++L) {
          E2[I][L] = o.E[I][0] * o.E[0][L] + o.E[I][1] * o.E[1][L] + o.E[I][2] * o.E[2][L];
          if (I == L) trE2 += E2[I][L];
        }
      const R I2 = R(3) + R(4) * t + R(2) * t * t - R(2) * trE2;
#pragma unroll
      for (int I = 0; I < 3; ++I)
#pragma unroll
        for (int L = 0; L < 3; ++L)
          Sb[I][L] += R(2) * cpr *
                      (R(2) * (o.E[I][L] - (I == L ? t * R(1.0 / 3.0) : R(0))) + R(4) * t * o.E[I][L] -
                       R(4) * E2[I][L] - (I == L ? R(4.0 / 3.0) * (t * t - trE2) : R(0)));
      al += R(16.0 / 9.0) * I2 * cpr;
      be += R(8.0 / 3.0) * I2 * cpr;
      x.cq = R(4) * cpr;
      x.cb = R(-2) * cpr;
      x.cp = R(-8.0 / 3.0) * cpr;
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int k = 0; k < 3; ++k)
          x.b[i][k] = o.F[i][0] * o.F[k][0] + o.F[i][1] * o.F[k][1] + o.F[i][2] * o.F[k][2];
    }
#pragma unroll
    for (int I = 0; I < 3; ++I)
#pragma unroll
      for (int L = I; L < 3; ++L) {
        const R v = o.Ci[I][0] * Sb[0][L] + o.Ci[I][1] * Sb[1][L] + o.Ci[I][2] * Sb[2][L];
        o.S[I][L] = v;
        o.S[L][I] = v;
      }
    o.al = al;
    o.be = be;
    o.ga = R(-2.0 / 3.0) * f;
  }
}

// Per-element material parameter arrays (device, one value per element):
// Neo-Hookean (mu, lambda); Mooney-Rivlin (c10, c01, K); Yeoh (c10, c20,
// c30, K); F-bar Neo-Hookean (mu, lambda, and J_node - 1 per internal node).
struct MatParams {
  const double* v[4];
};

// Parameters of element e as the material point takes them (Neo-Hookean:
// mu and K = lambda + 2 mu / 3).
template <int M>
__device__ __forceinline__ void load_params(const MatParams& mp, int64_t e, double pr[4]) {
  if constexpr (M == kNeoHookeanBound) {
    pr[0] = pr[1] = pr[2] = pr[3] = 0.0;   // taken from the geometry records
  } else if constexpr (M == kNeoHookean || M == kFbarNeoHookean) {
    pr[0] = mp.v[0][e];
    pr[1] = mp.v[1][e] + (2.0 / 3.0) * pr[0];
    pr[2] = pr[3] = 0.0;
  } else {
    constexpr int n = M == kMooneyRivlin ? 3 : 4;
#pragma unroll
    for (int i = 0; i < 4; ++i) pr[i] = i < n ? mp.v[i][e] : 0.0;
  }
}

template <typename R>
__device__ __forceinline__ void mat3vec(const R M[3][3], const R v[3], R o[3]) {
#pragma unroll
  for (int i = 0; i < 3; ++i) o[i] = M[i][0] * v[0] + M[i][1] * v[1] + M[i][2] * v[2];
}

// One 3x3 tangent block K_ab (row-major, scaled by the weight w = V or
// w_q det J) and nothing else.  ga_, qa_ belong to the row node; gb, qb to
// the column node; Sha = S h_a; hb = h_b.
template <typename R>
__device__ __forceinline__ void nh_block(const NHPoint<R>& m, R w, const R ga_[3], const R qa_[3],
                                         const R Sha[3], const R gb[3], const R qb[3],
                                         const R hb[3], R out[9]) {
  R P[3], Rv[3], T[3];
  const R hb2 = R(0.5) * m.be * w;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    P[i] = w * (m.al * ga_[i] + m.ga * qa_[i]);
    Rv[i] = hb2 * ga_[i];
    T[i] = w * m.ga * ga_[i];
  }
  const R d = hb2 * (ga_[0] * gb[0] + ga_[1] * gb[1] + ga_[2] * gb[2]) +
              w * (Sha[0] * hb[0] + Sha[1] * hb[1] + Sha[2] * hb[2]);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int k = 0; k < 3; ++k)
      out[3 * i + k] = P[i] * gb[k] + gb[i] * Rv[k] + T[i] * qb[k] + (i == k ? d : R(0));
}

// One 3x3 tangent block K_ab of model M (weight w): the Neo-Hookean terms of
// nh_block plus the model's extra terms (see hyper_point); ha = h_a.
template <typename R, int M>
__device__ __forceinline__ void hyper_block(const NHPoint<R>& m, const HyperExtra<R>& x, R w, const R ga_[3],
                                            const R qa_[3], const R ha[3], const R Sha[3], const R gb[3],
                                            const R qb[3], const R hb[3], R out[9]) {
  nh_block(m, w, ga_, qa_, Sha, gb, qb, hb, out);
  if constexpr (M == kMooneyRivlin) {
    R bqa[3], bqb[3];
    mat3vec(x.b, qa_, bqa);
    mat3vec(x.b, qb, bqb);
    const R hh = w * x.cb * (ha[0] * hb[0] + ha[1] * hb[1] + ha[2] * hb[2]);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int k = 0; k < 3; ++k)
        out[3 * i + k] += w * x.cq * qa_[i] * qb[k] + hh * x.b[i][k] + w * x.cb * qb[i] * qa_[k] +
                          w * x.cp * (ga_[i] * (x.I1 * qb[k] - bqb[k]) + (x.I1 * qa_[i] - bqa[i]) * gb[k]);
  } else if constexpr (M == kYeoh) {
    R za[3], zb[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      za[i] = x.zs * (qa_[i] - x.I1 * R(1.0 / 3.0) * ga_[i]);
      zb[i] = x.zs * (qb[i] - x.I1 * R(1.0 / 3.0) * gb[i]);
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int k = 0; k < 3; ++k) out[3 * i + k] += w * x.psi11 * za[i] * zb[k];
  }
}

// dS/dmu at fixed lambda (K = lambda + 2 mu / 3):
//   C^-1 (2 J^-2/3 dev E + 2/3 ln J I);   dS/dlambda = ln J C^-1.
template <typename R>
__device__ __forceinline__ void nh_dS_dmu(const NHPoint<R>& m, R o[3][3]) {
  R B[3][3];
#pragma unroll
  for (int I = 0; I < 3; ++I)
#pragma unroll
    for (int L = 0; L < 3; ++L)
      B[I][L] = R(2) * m.jm23 * (m.E[I][L] - (I == L ? m.trE * R(1.0 / 3.0) : R(0))) +
                (I == L ? R(2.0 / 3.0) * m.lnJ : R(0));
  R M[3][3];
#pragma unroll
  for (int I = 0; I < 3; ++I)
#pragma unroll
    for (int L = 0; L < 3; ++L) M[I][L] = m.Ci[I][0] * B[0][L] + m.Ci[I][1] * B[1][L] + m.Ci[I][2] * B[2][L];
#pragma unroll
  for (int I = 0; I < 3; ++I)
#pragma unroll
    for (int L = 0; L < 3; ++L) o[I][L] = R(0.5) * (M[I][L] + M[L][I]);
}

}  // namespace psx
