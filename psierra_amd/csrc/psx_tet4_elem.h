// Tet4 element routines shared by the assembly and adjoint kernels.
#pragma once

#include <type_traits>

#include "psx_nh.h"

namespace psx {

// Geometry and displacement gradient of a Tet4 from its four nodes
// (reference coordinates and displacements as double4, differences taken in
// FP64 before rounding to R).  h[b] = grad_X N_b, V = |det D| / 6 with D the
// edge matrix (det([1 X]) of src/psierra/element/tet4_base.py:51-55 up to
// the vertex order), A = grad_X u.
template <typename R>
__device__ __forceinline__ void tet4_geometry(const double4 X[4], const double4 U[4], R h[4][3], R& V,
                                              R A[3][3]) {
  R D[3][3], dU[3][3];
#pragma unroll
  for (int b = 0; b < 3; ++b) {
    D[0][b] = R(X[b + 1].x - X[0].x);
    D[1][b] = R(X[b + 1].y - X[0].y);
    D[2][b] = R(X[b + 1].z - X[0].z);
    dU[0][b] = R(U[b + 1].x - U[0].x);
    dU[1][b] = R(U[b + 1].y - U[0].y);
    dU[2][b] = R(U[b + 1].z - U[0].z);
  }
  R adj[3][3];
  adj3(D, adj);
  const R det = D[0][0] * adj[0][0] + D[0][1] * adj[1][0] + D[0][2] * adj[2][0];
  V = fabs(det) * R(1.0 / 6.0);
  const R idet = R(1) / det;
#pragma unroll
  for (int b = 0; b < 3; ++b)
#pragma unroll
    for (int c = 0; c < 3; ++c) h[b + 1][c] = adj[b][c] * idet;
#pragma unroll
  for (int c = 0; c < 3; ++c) h[0][c] = -(h[1][c] + h[2][c] + h[3][c]);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) A[i][j] = dU[i][0] * h[1][j] + dU[i][1] * h[2][j] + dU[i][2] * h[3][j];
}

// F-bar row (Tet4Fbar's domain assembly, src/psierra/element/tet4_fbar.py:
// 264-459): Neo-Hookean at F_bar = r F, r = (J_bar / J)^(1/3), with
// prm = (mu, K, J_bar - 1).  With g = F^-T h, q = F h, binv = F^-T F^-1,
// T = binv g, g1 = g - tr(C^-1)/3 q and the F_bar material point
// (fb = mu J_bar^-2/3, I1_bar = r^2 I1, c1 = K + 2/9 fb I1_bar,
//  c2 = -2 K ln J_bar + 2/3 fb I1_bar, S_bar = C^-1 (2 fb dev E + K ln J_bar / r^2 I)):
//   f_a  = V [ r^2 F S_bar h_a + (1 - r^2) tr(S_bar)/3 q_a ]
//   K_ab = V [ c1 g1_a g1_b^T + c2 ((g_b g_a^T + (g_a . g_b) I)/2
//              - (T_a q_b^T + q_a T_b^T)/3 + tr(C^-2)/9 q_a q_b^T)
//              + (h_a . S_bar h_b) I ].
// This is the reference's (r^2 B_dev)^T (C(F_bar) - K_bulk m m^T) (r^2 B_dev) V
// + G^T (I x S_bar) G V: the deviatoric strain rows have zero trace, so the
// K_bulk m m^T term and the C^-1 (x) I terms of C(F_bar) drop out, and the r^4
// of B_bar cancels the r^-4 that C(F_bar) carries through C_bar^-1 = C^-1 / r^2.
// f_a is B_bar^T S_bar V with B_bar = r^2 B_dev + m b_vol^T.  The nodal
// volumetric stiffness K_vol is assembled by its own kernel.  S_bar is formed
// from E (dev E_bar = r^2 dev E), so small strains keep their relative accuracy.
template <typename R, bool WANT_K>
__device__ __forceinline__ R tet4_fbar_row_core(const R h[4][3], R V, const R A[3][3], const R prm[4], R* out) {
  NHPoint<R> m;
  nh_point(A, prm[0], prm[1], m);   // kinematics (F, F^-T, C^-1, E, J, ln J)
  const R mu = prm[0], K = prm[1];
  const R lnJb = d_log1p(prm[2]);
  const R r2 = d_exp(R(2.0 / 3.0) * (lnJb - m.lnJ));
  const R fb = mu * d_exp(R(-2.0 / 3.0) * lnJb);
  const R vol = K * lnJb * d_rcp(r2);
  R Sb[3][3], S[3][3];
#pragma unroll
  for (int I = 0; I < 3; ++I)
#pragma unroll
    for (int L = 0; L < 3; ++L)
      Sb[I][L] = R(2) * fb * (m.E[I][L] - (I == L ? m.trE * R(1.0 / 3.0) : R(0))) + (I == L ? vol : R(0));
#pragma unroll
  for (int I = 0; I < 3; ++I)
#pragma unroll
    for (int L = I; L < 3; ++L) {
      const R v = m.Ci[I][0] * Sb[0][L] + m.Ci[I][1] * Sb[1][L] + m.Ci[I][2] * Sb[2][L];
      S[I][L] = v;
      S[L][I] = v;
    }
  const R trS3 = (S[0][0] + S[1][1] + S[2][2]) * R(1.0 / 3.0);
  R Sh0[3], FSh[3], q0[3];
  mat3vec(S, h[0], Sh0);
  mat3vec(m.F, Sh0, FSh);
  mat3vec(m.F, h[0], q0);
#pragma unroll
  for (int i = 0; i < 3; ++i) out[(WANT_K ? 36 : 0) + i] = V * (r2 * FSh[i] + (R(1) - r2) * trS3 * q0[i]);
  if (WANT_K) {
    const R I1b = r2 * (R(3) + R(2) * m.trE);
    const R c1 = K + R(2.0 / 9.0) * fb * I1b;
    const R c2 = R(-2) * K * lnJb + R(2.0 / 3.0) * fb * I1b;
    const R trCi3 = (m.Ci[0][0] + m.Ci[1][1] + m.Ci[2][2]) * R(1.0 / 3.0);
    R trC2 = R(0);
#pragma unroll
    for (int I = 0; I < 3; ++I)
#pragma unroll
      for (int L = 0; L < 3; ++L) trC2 += m.Ci[I][L] * m.Ci[I][L];
    R binv[3][3];   // F^-T F^-1 = Gm Gm^T
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int k = i; k < 3; ++k) {
        const R v = m.Gm[i][0] * m.Gm[k][0] + m.Gm[i][1] * m.Gm[k][1] + m.Gm[i][2] * m.Gm[k][2];
        binv[i][k] = v;
        binv[k][i] = v;
      }
    R g[4][3], q[4][3], T[4][3], g1[4][3];
#pragma unroll
    for (int b = 1; b < 4; ++b) {
      mat3vec(m.Gm, h[b], g[b]);
      mat3vec(m.F, h[b], q[b]);
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      g[0][c] = -(g[1][c] + g[2][c] + g[3][c]);
      q[0][c] = -(q[1][c] + q[2][c] + q[3][c]);
    }
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      mat3vec(binv, g[b], T[b]);
#pragma unroll
      for (int c = 0; c < 3; ++c) g1[b][c] = g[b][c] - trCi3 * q[b][c];
    }
    // K_0b = P1 g1_b^T + g_b Rv^T + P2 q_b^T + Q0 T_b^T + (g_b . Rv + h_b . W) I
    R P1[3], Rv[3], P2[3], Q0[3], W[3];
    const R v1 = V * c1, v2 = V * c2;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      P1[i] = v1 * g1[0][i];
      Rv[i] = R(0.5) * v2 * g[0][i];
      P2[i] = v2 * (trC2 * R(1.0 / 9.0) * q[0][i] - R(1.0 / 3.0) * T[0][i]);
      Q0[i] = R(-1.0 / 3.0) * v2 * q[0][i];
      W[i] = V * Sh0[i];
    }
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const R sb = g[b][0] * Rv[0] + g[b][1] * Rv[1] + g[b][2] * Rv[2] + h[b][0] * W[0] + h[b][1] * W[1] +
                   h[b][2] * W[2];
      R* o = out + 9 * b;
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int k = (b == 0 ? i : 0); k < 3; ++k) {
          const R v = P1[i] * g1[b][k] + g[b][i] * Rv[k] + P2[i] * q[b][k] + Q0[i] * T[b][k] + (i == k ? sb : R(0));
          o[3 * i + k] = v;
          if (b == 0 && k != i) o[3 * k + i] = v;   // K_00 is symmetric
        }
    }
  }
  return m.J;
}

// Row 0 of an element, given the shape-function gradients in the row's node
// order (h[0] = row node), the volume V and grad u: the four 3x3 blocks
// K_{0b} (b = 0..3, row-major, 36 values) and f_0 (3 values), or f_0 only.
// Returns J = det F.
// prm: the model's parameters (kNeoHookean: mu, K = lambda + 2 mu / 3;
// kFbarNeoHookean: mu, K, J_bar - 1).
template <typename R, int M, bool WANT_K>
__device__ __forceinline__ R tet4_row_core_m(const R h[4][3], R V, const R A[3][3], const R prm[4], R* out) {
  if constexpr (M == kFbarNeoHookean) {
    return tet4_fbar_row_core<R, WANT_K>(h, V, A, prm, out);
  } else {
    NHPoint<R> m;
    HyperExtra<R> x;
    hyper_point<R, M>(A, prm, m, x);
    R Sh0[3];
    mat3vec(m.S, h[0], Sh0);
    // f_0 = V F S h_0
    R FSh[3];
    mat3vec(m.F, Sh0, FSh);
#pragma unroll
    for (int i = 0; i < 3; ++i) out[(WANT_K ? 36 : 0) + i] = V * FSh[i];
    if (WANT_K) {
      R g[4][3], q[4][3];
#pragma unroll
      for (int b = 1; b < 4; ++b) {
        mat3vec(m.Gm, h[b], g[b]);
        mat3vec(m.F, h[b], q[b]);
      }
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        g[0][c] = -(g[1][c] + g[2][c] + g[3][c]);
        q[0][c] = -(q[1][c] + q[2][c] + q[3][c]);
      }
      // K_0b = P g_b^T + g_b Rv^T + T q_b^T + (g_b . Rv + h_b . W) I
      R P[3], Rv[3], T[3], W[3];
      const R va = V * m.al, vb = R(0.5) * V * m.be, vg = V * m.ga;
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        P[i] = va * g[0][i] + vg * q[0][i];
        Rv[i] = vb * g[0][i];
        T[i] = vg * g[0][i];
        W[i] = V * Sh0[i];
      }
      // model-specific vectors of the row node (see psx_nh.h)
      R Qm[3] = {R(0), R(0), R(0)}, Pm[3] = {R(0), R(0), R(0)}, Z0[3] = {R(0), R(0), R(0)};
      if constexpr (M == kMooneyRivlin) {
        R bq0[3];
        mat3vec(x.b, q[0], bq0);
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          Qm[i] = V * x.cq * q[0][i];                          // times q_b^T
          Pm[i] = V * x.cp * (x.I1 * q[0][i] - bq0[i]);        // p_0, times g_b^T
        }
      } else if constexpr (M == kYeoh) {
#pragma unroll
        for (int i = 0; i < 3; ++i) Z0[i] = V * x.psi11 * x.zs * (q[0][i] - x.I1 * R(1.0 / 3.0) * g[0][i]);
      }
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const R sb = g[b][0] * Rv[0] + g[b][1] * Rv[1] + g[b][2] * Rv[2] + h[b][0] * W[0] + h[b][1] * W[1] +
                     h[b][2] * W[2];
        R pb[3] = {R(0), R(0), R(0)}, zb[3] = {R(0), R(0), R(0)};
        R hh = R(0);
        if constexpr (M == kMooneyRivlin) {
          R bqb[3];
          mat3vec(x.b, q[b], bqb);
#pragma unroll
          for (int i = 0; i < 3; ++i) pb[i] = x.I1 * q[b][i] - bqb[i];
          hh = V * x.cb * (h[0][0] * h[b][0] + h[0][1] * h[b][1] + h[0][2] * h[b][2]);
        } else if constexpr (M == kYeoh) {
#pragma unroll
          for (int i = 0; i < 3; ++i) zb[i] = x.zs * (q[b][i] - x.I1 * R(1.0 / 3.0) * g[b][i]);
        }
        R* o = out + 9 * b;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int k = (b == 0 ? i : 0); k < 3; ++k) {
            // FP32: the diagonal term first, so a diagonal value is three
            // fused multiply-adds on sb (an off-diagonal one a product and
            // two); FP64 keeps the order that holds it to 4 waves per SIMD
            R v;
            if constexpr (std::is_same<R, float>::value)
              v = (i == k ? sb : R(0)) + P[i] * g[b][k] + g[b][i] * Rv[k] + T[i] * q[b][k];
            else
              v = P[i] * g[b][k] + g[b][i] * Rv[k] + T[i] * q[b][k] + (i == k ? sb : R(0));
            if constexpr (M == kMooneyRivlin)
              v += Qm[i] * q[b][k] + hh * x.b[i][k] + V * x.cb * q[b][i] * q[0][k] +
                   V * x.cp * g[0][i] * pb[k] + Pm[i] * g[b][k];
            else if constexpr (M == kYeoh)
              v += Z0[i] * zb[k];
            o[3 * i + k] = v;
            if (b == 0 && k != i) o[3 * k + i] = v;  // K_00 is symmetric
          }
      }
    }
    return m.J;
  }
}

template <typename R, bool WANT_K>
__device__ __forceinline__ R tet4_row_core(const R h[4][3], R V, const R A[3][3], R mu, R K, R* out) {
  const R prm[4] = {mu, K, R(0), R(0)};
  return tet4_row_core_m<R, kNeoHookean, WANT_K>(h, V, A, prm, out);
}

// Element seen from its local node 0 (callers rotate the node order so the
// row node comes first), geometry from the reference coordinates.
template <typename R, int M, bool WANT_K>
__device__ __forceinline__ R tet4_row_m(const double4 X[4], const double4 U[4], const R prm[4], R* out) {
  R h[4][3], V, A[3][3];
  tet4_geometry(X, U, h, V, A);
  return tet4_row_core_m<R, M, WANT_K>(h, V, A, prm, out);
}

template <typename R, bool WANT_K>
__device__ __forceinline__ R tet4_row(const double4 X[4], const double4 U[4], R mu, R K, R* out) {
  const R prm[4] = {mu, K, R(0), R(0)};
  return tet4_row_m<R, kNeoHookean, WANT_K>(X, U, prm, out);
}

// Same row from precomputed shape-function gradients already in the row's
// node order: G[b] = (h_b, V) (b = 0 the row node), displacements U in the
// same order (differences taken in FP64 before rounding).
template <int M, bool WANT_K>
__device__ __forceinline__ float tet4_row_grad_m(const float4 G[4], const double4 U[4], const float prm[4],
                                                 float* out) {
  float h[4][3];
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    h[b][0] = G[b].x;
    h[b][1] = G[b].y;
    h[b][2] = G[b].z;
  }
  float A[3][3];
  {
    float dU[3][3];
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      dU[0][b] = float(U[b + 1].x - U[0].x);
      dU[1][b] = float(U[b + 1].y - U[0].y);
      dU[2][b] = float(U[b + 1].z - U[0].z);
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) A[i][j] = dU[i][0] * h[1][j] + dU[i][1] * h[2][j] + dU[i][2] * h[3][j];
  }
  if constexpr (M == kNeoHookeanBound)
    return tet4_row_core_m<float, kNeoHookean, WANT_K>(h, 1.0f, A, prm, out);   // prm = (V mu, V K)
  else
    return tet4_row_core_m<float, M, WANT_K>(h, G[0].w, A, prm, out);
}

template <bool WANT_K>
__device__ __forceinline__ float tet4_row_grad(const float4 G[4], const double4 U[4], float mu, float K,
                                               float* out) {
  const float prm[4] = {mu, K, 0.f, 0.f};
  return tet4_row_grad_m<kNeoHookean, WANT_K>(G, U, prm, out);
}

// Internal force of a whole element in its own node order: f[3 b + i] =
// V (F S h_b)_i (the B^T S_voigt V of tet4_total_lagrangian.py:136-137).
// Returns det F.
template <typename R>
__device__ __forceinline__ R tet4_element_force(const double4 X[4], const double4 U[4], R mu, R K, R f[12]) {
  R h[4][3], V, A[3][3];
  tet4_geometry(X, U, h, V, A);
  NHPoint<R> m;
  nh_point(A, mu, K, m);
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    R Sh[3], FSh[3];
    mat3vec(m.S, h[b], Sh);
    mat3vec(m.F, Sh, FSh);
#pragma unroll
    for (int i = 0; i < 3; ++i) f[3 * b + i] = V * FSh[i];
  }
  return m.J;
}

// Adjoint of f_int for one incidence: the row-0 part of K_e w,
//   (K_e w)_0 = V [ a tr(L) g_0 + b/2 (L + L^T) g_0 + c ((F : grad w) g_0
//               + tr(L) q_0) + grad w S h_0 ],
// with grad w = sum_b w_b (x) h_b and L = grad w F^-1 (the spatial gradient
// of w); the closed form of sum_b K_0b w_b with the blocks above.  Also the
// element's parameter sensitivities of w^T f_e (f is linear in mu, lambda):
//   d/dmu = V dS/dmu : (F^T grad w),  d/dlambda = V ln J C^-1 : (F^T grad w).
template <typename R>
__device__ __forceinline__ R tet4_vjp_row(const double4 X[4], const double4 U[4], const double4 W[4],
                                          R mu, R K, R out[3], R& gmu, R& glam) {
  R h[4][3], V, A[3][3];
  tet4_geometry(X, U, h, V, A);
  NHPoint<R> m;
  nh_point(A, mu, K, m);
  R dW[3][3];
#pragma unroll
  for (int b = 0; b < 3; ++b) {
    dW[0][b] = R(W[b + 1].x - W[0].x);
    dW[1][b] = R(W[b + 1].y - W[0].y);
    dW[2][b] = R(W[b + 1].z - W[0].z);
  }
  R Gw[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) Gw[i][j] = dW[i][0] * h[1][j] + dW[i][1] * h[2][j] + dW[i][2] * h[3][j];
  // L = Gw F^-1 ; F^-1[j][k] = Gm[k][j]
  R L[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int k = 0; k < 3; ++k) L[i][k] = Gw[i][0] * m.Gm[k][0] + Gw[i][1] * m.Gm[k][1] + Gw[i][2] * m.Gm[k][2];
  const R trL = L[0][0] + L[1][1] + L[2][2];
  R FdW = R(0);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) FdW += m.F[i][j] * Gw[i][j];
  R g0[3], q0[3], Sh0[3], GSh[3];
  mat3vec(m.Gm, h[0], g0);
  mat3vec(m.F, h[0], q0);
  mat3vec(m.S, h[0], Sh0);
  mat3vec(Gw, Sh0, GSh);
  const R c1 = m.al * trL + m.ga * FdW;
  const R c2 = m.ga * trL;
  const R hb = R(0.5) * m.be;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const R Lg = (L[i][0] + L[0][i]) * g0[0] + (L[i][1] + L[1][i]) * g0[1] + (L[i][2] + L[2][i]) * g0[2];
    out[i] = V * (c1 * g0[i] + hb * Lg + c2 * q0[i] + GSh[i]);
  }
  // M = F^T grad w
  R M[3][3];
#pragma unroll
  for (int I = 0; I < 3; ++I)
#pragma unroll
    for (int J = 0; J < 3; ++J) M[I][J] = m.F[0][I] * Gw[0][J] + m.F[1][I] * Gw[1][J] + m.F[2][I] * Gw[2][J];
  R Smu[3][3];
  nh_dS_dmu(m, Smu);
  R smu = R(0), slam = R(0);
#pragma unroll
  for (int I = 0; I < 3; ++I)
#pragma unroll
    for (int J = 0; J < 3; ++J) {
      smu += Smu[I][J] * M[I][J];
      slam += m.Ci[I][J] * M[I][J];
    }
  gmu = V * smu;
  glam = V * m.lnJ * slam;
  return m.J;
}


}  // namespace psx
