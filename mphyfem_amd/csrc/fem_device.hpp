// fem_device.hpp -- device helpers shared by the element kernels: 3x3 algebra, volumetric energy
// derivatives, the Neo-Hookean Gauss-point state, and the Hex8 / Tet10 shape-function derivatives.
//
// Reference (chennachaos/mphyfem, src/):
//   computeStressAndTangent_NeoHooke        utilitiesmaterialHyperelastic.cpp:186-295
//   volumetricfunctions_get*Derivative      utilitiesmaterialHyperelastic.cpp:45-140
//   LagrangeBasisFunsHexa / Tetra           BasisFunctionsLagrange.cpp:1202-1279, 870-956
//   getGaussPointsHexa / Tetra              QuadratureUtil.cpp:453-492, 670-712
#pragma once

#include "mpf_internal.hpp"

namespace mpf {
namespace dev {

constexpr double kGP2 = 0.577350269189626;  // QuadratureUtil.cpp:24-25
constexpr double kTetA = 0.1381966011250105, kTetB = 0.5854101966249685;  // QuadratureUtil.cpp:692-703

__device__ __forceinline__ double det3(const double A[3][3]) {
  return A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1]) - A[0][1] * (A[1][0] * A[2][2] - A[1][2] * A[2][0]) +
         A[0][2] * (A[1][0] * A[2][1] - A[1][1] * A[2][0]);
}

// Inverse by the adjugate; returns det.
__device__ __forceinline__ double inv3(const double A[3][3], double R[3][3]) {
  const double c00 = A[1][1] * A[2][2] - A[1][2] * A[2][1];
  const double c01 = A[1][2] * A[2][0] - A[1][0] * A[2][2];
  const double c02 = A[1][0] * A[2][1] - A[1][1] * A[2][0];
  const double det = A[0][0] * c00 + A[0][1] * c01 + A[0][2] * c02;
  const double id = 1.0 / det;
  R[0][0] = c00 * id;
  R[1][0] = c01 * id;
  R[2][0] = c02 * id;
  R[0][1] = (A[0][2] * A[2][1] - A[0][1] * A[2][2]) * id;
  R[1][1] = (A[0][0] * A[2][2] - A[0][2] * A[2][0]) * id;
  R[2][1] = (A[0][1] * A[2][0] - A[0][0] * A[2][1]) * id;
  R[0][2] = (A[0][1] * A[1][2] - A[0][2] * A[1][1]) * id;
  R[1][2] = (A[0][2] * A[1][0] - A[0][0] * A[1][2]) * id;
  R[2][2] = (A[0][0] * A[1][1] - A[0][1] * A[1][0]) * id;
  return det;
}

// volumetricfunctions_getFirstDerivative / getSecondDerivative (utilitiesmaterialHyperelastic.cpp:45-140)
__device__ __forceinline__ void vol_derivs(int type, double K, double J, double& d1, double& d2) {
  switch (type) {
    case 1: d1 = 0.1 * K * (J * J * J * J - 1.0 / (J * J * J * J * J * J));
            d2 = 0.2 * K * (2.0 * J * J * J + 3.0 / (J * J * J * J * J * J * J)); break;
    case 2: d1 = 0.125 * K * (J * J - 1.0 / J / J) * (J + 1.0 / J / J / J);
            d2 = 0.125 * K * (3.0 * J * J + 5.0 / (J * J * J * J * J * J)); break;
    case 3: d1 = K * (J - 1.0); d2 = K; break;
    case 4: d1 = 0.5 * K * (J - 1.0 / J); d2 = 0.5 * K * (1.0 + 1.0 / J / J); break;
    case 5: d1 = 0.5 * K * (J - 1.0 + log(J) / J); d2 = 0.5 * K * (J * J + 1.0 - log(J)) / J / J; break;
    case 6: d1 = K * log(J); d2 = K / log(J); break;
    case 7: d1 = K * (1.0 - 1.0 / J); d2 = K / J / J; break;
    case 8: d1 = K * log(J) / J; d2 = K * (1.0 - log(J)) / J / J; break;
    default: d1 = 0.0; d2 = 0.0; break;
  }
}

// Records 1 + e in *bad, keeping the smallest element id (integer CAS; off the FP64 path).
__device__ inline void flag_bad(int32_t* bad, int32_t e) {
  if (!bad) return;
  const int32_t v = e + 1;
  int32_t old = __hip_atomic_load(bad, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  while (old == 0 || v < old) {
    const int32_t prev = atomicCAS(bad, old, v);
    if (prev == old) break;
    old = prev;
  }
}

// Neo-Hookean state at one Gauss point: Cauchy stress, b = F F^T and the five scalars of the
// closed-form tangent contraction (multiplied by dv):
//   K_ab^{ik} += s1 g^a_k g^b_i + (s2 g^a_i + s3 (bg^a)_i) g^b_k + s3 g^a_i (bg^b)_k
//                + delta_ik (s4 bg^a + s5 g^a) . g^b
// which equals sum_{m,n} g^a_m C[(i,m),(k,n)] g^b_n for the reference's 9x9 spatial tangent
// (utilitiesmaterialHyperelastic.cpp:235-291) with
//   disp : s1 = dv (f1 - U'), s2 = dv (f3 + J U'' + U'), s3 = dv f2, s4 = dv f4, s5 = -dv U'
//          (s5 = 0 for the consistent tangent, which adds the volumetric stress term)
//   mixed: s1 = dv (f1 - p),  s2 = dv (f3 + p),          s3 = dv f2, s4 = dv f4, s5 = 0
//   f1 = mu J^-5/3 I1/3, f2 = -2/3 mu J^-5/3, f3 = 2/9 mu J^-5/3 I1, f4 = mu J^-5/3.
struct NHPoint {
  double sig[3][3];
  double s1, s2, s3, s4, s5;
  double b[3][3];
};

__device__ __forceinline__ void neohooke_point(const double F[3][3], double J, double mu, double K, int Utype,
                                               bool mixed, double pres, double dv, bool consistent, NHPoint& P) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) P.b[i][j] = F[i][0] * F[j][0] + F[i][1] * F[j][1] + F[i][2] * F[j][2];
  const double r1d3 = 1.0 / 3.0, r2d3 = 2.0 / 3.0;
  const double Jm2d3 = pow(J, -r2d3);
  const double Jm5d3 = Jm2d3 / J;
  const double Ib = P.b[0][0] + P.b[1][1] + P.b[2][2];
  double U1 = 0.0, U2 = 0.0;
  if (!mixed) vol_derivs(Utype, K, J, U1, U2);
  const double fs1 = mu * Jm5d3;
  const double fs2 = -fs1 * r1d3 * Ib + (mixed ? pres : U1);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) P.sig[i][j] = fs1 * P.b[i][j] + (i == j ? fs2 : 0.0);
  const double f1 = r1d3 * mu * Jm5d3 * Ib;
  const double f2 = -r2d3 * mu * Jm5d3;
  const double f3 = r2d3 * r1d3 * mu * Jm5d3 * Ib;
  const double f4 = mu * Jm5d3;
  if (mixed) {
    P.s1 = dv * (f1 - pres);
    P.s2 = dv * (f3 + pres);
    P.s5 = 0.0;
  } else {
    P.s1 = dv * (f1 - U1);
    P.s2 = dv * (f3 + (J * U2 + U1));
    P.s5 = consistent ? 0.0 : -dv * U1;
  }
  P.s3 = dv * f2;
  P.s4 = dv * f4;
}

// dN/dxi of the 8 Hex8 nodes at Gauss point gp (LagrangeBasisFunsHexa, BasisFunctionsLagrange.cpp:1202-1279;
// getGaussPointsHexa ordering QuadratureUtil.cpp:453-492).
__device__ __forceinline__ void hex8_dN(int gp, double d[3][8]) {
  const int p0 = (gp == 1 || gp == 2 || gp == 5 || gp == 6) ? 1 : 0;
  const int p1 = (gp == 2 || gp == 3 || gp == 6 || gp == 7) ? 1 : 0;
  const int p2 = (gp >= 4) ? 1 : 0;
  const double x1 = p0 ? kGP2 : -kGP2, x2 = p1 ? kGP2 : -kGP2, x3 = p2 ? kGP2 : -kGP2;
  const double Nu[2] = {0.5 * (1.0 - x1), 0.5 * (1.0 + x1)};
  const double Nv[2] = {0.5 * (1.0 - x2), 0.5 * (1.0 + x2)};
  const double Nw[2] = {0.5 * (1.0 - x3), 0.5 * (1.0 + x3)};
  const double dn[2] = {-0.5, 0.5};
  const int iu[8] = {0, 1, 1, 0, 0, 1, 1, 0}, iv[8] = {0, 0, 1, 1, 0, 0, 1, 1}, iw[8] = {0, 0, 0, 0, 1, 1, 1, 1};
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    d[0][a] = dn[iu[a]] * Nv[iv[a]] * Nw[iw[a]];
    d[1][a] = Nu[iu[a]] * dn[iv[a]] * Nw[iw[a]];
    d[2][a] = Nu[iu[a]] * Nv[iv[a]] * dn[iw[a]];
  }
}

// dN/dxi of one Hex8 node a at Gauss point gp: the same products as hex8_dN.
__device__ __forceinline__ void hex8_dN_node(int gp, int a, double d[3]) {
  const double x1 = (gp == 1 || gp == 2 || gp == 5 || gp == 6) ? kGP2 : -kGP2;
  const double x2 = (gp == 2 || gp == 3 || gp == 6 || gp == 7) ? kGP2 : -kGP2;
  const double x3 = (gp >= 4) ? kGP2 : -kGP2;
  const int iu = (a == 1 || a == 2 || a == 5 || a == 6) ? 1 : 0;
  const int iv = (a == 2 || a == 3 || a == 6 || a == 7) ? 1 : 0;
  const int iw = (a >= 4) ? 1 : 0;
  const double Nu = iu ? 0.5 * (1.0 + x1) : 0.5 * (1.0 - x1);
  const double Nv = iv ? 0.5 * (1.0 + x2) : 0.5 * (1.0 - x2);
  const double Nw = iw ? 0.5 * (1.0 + x3) : 0.5 * (1.0 - x3);
  const double du = iu ? 0.5 : -0.5, dv = iv ? 0.5 : -0.5, dw = iw ? 0.5 : -0.5;
  d[0] = du * Nv * Nw;
  d[1] = Nu * dv * Nw;
  d[2] = Nu * Nv * dw;
}

// g = grad_x N from dN/dxi (d) and the inverse Jacobian Bi (row-major, Bi[r][c] = dxi_c/dx_r in the
// element_point convention), with explicit fused multiply-adds so that every kernel that forms g
// from a stored Bi gets the same bits as the stage kernel.
__device__ __forceinline__ void grad_jinv(const double* Bi, const double d[3], double g[3]) {
  g[0] = fma(d[2], Bi[2], fma(d[1], Bi[1], d[0] * Bi[0]));
  g[1] = fma(d[2], Bi[5], fma(d[1], Bi[4], d[0] * Bi[3]));
  g[2] = fma(d[2], Bi[8], fma(d[1], Bi[7], d[0] * Bi[6]));
}

// Gauss point gp of the 4-point tet rule (QuadratureUtil.cpp:690-712).
__device__ __forceinline__ void tet_gp(int gp, double& x1, double& x2, double& x3) {
  x1 = (gp == 1) ? kTetB : kTetA;
  x2 = (gp == 2) ? kTetB : kTetA;
  x3 = (gp == 3) ? kTetB : kTetA;
}

// Tet10 dN/dxi (LagrangeBasisFunsTetra, BasisFunctionsLagrange.cpp:896-948).
__device__ __forceinline__ void tet10_dN(double xi1, double xi2, double xi3, double d1[10], double d2[10],
                                         double d3[10]) {
  const double xi4 = 1.0 - xi1 - xi2 - xi3;
  d1[0] = -4 * xi4 + 1; d1[1] = 4 * xi1 - 1; d1[2] = 0; d1[3] = 0; d1[4] = 4 * (xi4 - xi1);
  d1[5] = 4 * xi2; d1[6] = -4 * xi2; d1[7] = -4 * xi3; d1[8] = 0; d1[9] = 4 * xi3;
  d2[0] = -4 * xi4 + 1; d2[1] = 0; d2[2] = 4 * xi2 - 1; d2[3] = 0; d2[4] = -4 * xi1;
  d2[5] = 4 * xi1; d2[6] = 4 * (xi4 - xi2); d2[7] = -4 * xi3; d2[8] = 4 * xi3; d2[9] = 0;
  d3[0] = -4 * xi4 + 1; d3[1] = 0; d3[2] = 0; d3[3] = 4 * xi3 - 1; d3[4] = -4 * xi1;
  d3[5] = 0; d3[6] = -4 * xi2; d3[7] = 4 * (xi4 - xi3); d3[8] = 4 * xi2; d3[9] = 4 * xi1;
}

// Tet4 (P1 pressure) basis at the Gauss points: N0 = xi4, N1 = xi1, N2 = xi2, N3 = xi3.
__device__ __forceinline__ void tet_Np(int gp, double Np[4]) {
  double x1, x2, x3;
  tet_gp(gp, x1, x2, x3);
  Np[0] = 1.0 - x1 - x2 - x3; Np[1] = x1; Np[2] = x2; Np[3] = x3;
}

__device__ __forceinline__ double shfl_xor_d(double v, int m) { return __shfl_xor(v, m, 64); }

// A double moved across lanes by a DPP lane pattern (two 32-bit v_mov_dpp; no LDS round trip).
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

// Sum over each aligned group of NG (2, 4 or 8) lanes, every lane receiving it: bitwise the same as
// v += shfl_xor(v, 1); v += shfl_xor(v, 2); v += shfl_xor(v, 4).  The last step takes lane 7 - l of
// the group (row_half_mirror) instead of l ^ 4; after the first two steps both hold the other
// quad's sum, and the additions are the same.
template <int NG>
__device__ __forceinline__ double group_sum(double v) {
  static_assert(NG == 2 || NG == 4 || NG == 8, "DPP group");
  v += dpp_d<0xB1>(v);                   // quad_perm [1, 0, 3, 2]: lane ^ 1
  if (NG >= 4) v += dpp_d<0x4E>(v);      // quad_perm [2, 3, 0, 1]: lane ^ 2
  if (NG >= 8) v += dpp_d<0x141>(v);     // row_half_mirror
  return v;
}

__device__ __forceinline__ int rank3(int mask, int i) { return __popc(mask & ((1 << i) - 1)); }

// One Gauss point of one element: reference Jacobian from X, F from u (and Fn from u_prev for the
// mixed element), current Jacobian from x = X + u, Neo-Hookean state.  Writes g_a, (b g)_a for all
// nodes into G[a*6 + c] and returns the Gauss-point scalars; the residual contribution of node a,
// -dv sum_m g_a,m sigma_i,m, is reduced over the NGP lanes of the element (consecutive lanes,
// lane & (NGP-1) == gp) and stored by the lane gp == a % NGP into R[a*3 + i].
// Elem_Solid_3D_Disp.cpp:250-330; ElementBase2.cpp:350-470.
// COMPACT output (the gather path's Hex8 staging records): the inverse Jacobian Bi (9 doubles,
// row-major) at G, from which g of every node is formed with grad_jinv, and the six components
// b00 b11 b22 b01 b02 b12 of the symmetric b at Bv, instead of g and b g of every node at G + a * 6.
// GPAD output (the gather path's Tet10 staging records): g of every node at G + a * 4 (g0 g1 g2 0)
// and the six components of b at Bv.
template <int NPE, int NGP, bool MIXED, bool COMPACT = false, bool GPAD = false>
__device__ __forceinline__ void element_point(const double (*X)[3], const double (*U)[3], const double (*UP)[3],
                                              const double* Pn, const double (*d)[NPE], int gp, double wgp,
                                              double mu, double K, int Utype, bool consistent, bool writer,
                                              double* G, double* S, double* R, bool& negJ, double* Bv = nullptr) {
  double B[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, Bi[3][3];
#pragma unroll
  for (int a = 0; a < NPE; ++a)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double xc = X[a][c];
      B[0][c] += xc * d[0][a]; B[1][c] += xc * d[1][a]; B[2][c] += xc * d[2][a];
    }
  const double dv0 = wgp * inv3(B, Bi);
  double F[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}, Fn[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
#pragma unroll
  for (int a = 0; a < NPE; ++a) {
    const double gx = d[0][a] * Bi[0][0] + d[1][a] * Bi[0][1] + d[2][a] * Bi[0][2];
    const double gy = d[0][a] * Bi[1][0] + d[1][a] * Bi[1][1] + d[2][a] * Bi[1][2];
    const double gz = d[0][a] * Bi[2][0] + d[1][a] * Bi[2][1] + d[2][a] * Bi[2][2];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const double ui = U[a][i];
      F[i][0] += ui * gx; F[i][1] += ui * gy; F[i][2] += ui * gz;
      if (MIXED) {
        const double upi = UP[a][i];
        Fn[i][0] += upi * gx; Fn[i][1] += upi * gy; Fn[i][2] += upi * gz;
      }
    }
  }
  double Jhat = 1.0, thetahat = 0.0;
  if (MIXED && Utype != -1) {  // compute_constants_volumetric_functions (utilitiesmaterialHyperelastic.cpp:144-173)
    const double Jn = det3(Fn);
    double d1, d2;
    vol_derivs(Utype, K, Jn, d1, d2);
    Jhat = Jn - d1 / d2;
    thetahat = 1.0 / d2;
  }
  const double J = det3(F);
  negJ = J < 0.0;
#pragma unroll
  for (int r = 0; r < 3; ++r) B[r][0] = B[r][1] = B[r][2] = 0.0;
#pragma unroll
  for (int a = 0; a < NPE; ++a)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double xc = X[a][c] + U[a][c];
      B[0][c] += xc * d[0][a]; B[1][c] += xc * d[1][a]; B[2][c] += xc * d[2][a];
    }
  const double dv = wgp * inv3(B, Bi);
  double pres = 0.0;
  if (MIXED) {
    if constexpr (NPE == 8) {
      pres = Pn[0];  // Q1/P0: one pressure per element, Np = 1 (ElementBase2.cpp:389-393)
    } else {
      double Np[4];
      tet_Np(gp, Np);
#pragma unroll
      for (int q = 0; q < 4; ++q) pres += Np[q] * Pn[q];
    }
  }
  NHPoint P;
  neohooke_point(F, J, mu, K, Utype, MIXED, pres, dv, consistent, P);
  if (writer && COMPACT) {
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) G[3 * r + c] = Bi[r][c];
  }
#pragma unroll
  for (int a = 0; a < NPE; ++a) {
    double ga[3];
    const double da[3] = {d[0][a], d[1][a], d[2][a]};
    grad_jinv(&Bi[0][0], da, ga);
    const double g0 = ga[0], g1 = ga[1], g2 = ga[2];
    if (writer && GPAD) {
      double* Ga = G + a * 4;
      Ga[0] = g0; Ga[1] = g1; Ga[2] = g2; Ga[3] = 0.0;
    } else if (writer && !COMPACT) {
      double* Ga = G + a * 6;
      Ga[0] = g0; Ga[1] = g1; Ga[2] = g2;
      Ga[3] = P.b[0][0] * g0 + P.b[0][1] * g1 + P.b[0][2] * g2;
      Ga[4] = P.b[1][0] * g0 + P.b[1][1] * g1 + P.b[1][2] * g2;
      Ga[5] = P.b[2][0] * g0 + P.b[2][1] * g1 + P.b[2][2] * g2;
    }
    const double b0 = dv * g0, b1 = dv * g1, b2 = dv * g2;
    double r[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      double v = -(b0 * P.sig[i][0] + b1 * P.sig[i][1] + b2 * P.sig[i][2]);
      v = group_sum<NGP>(v);
      r[i] = v;
    }
    if (writer && (a % NGP) == gp) { R[a * 3] = r[0]; R[a * 3 + 1] = r[1]; R[a * 3 + 2] = r[2]; }
  }
  if (writer) {
    S[0] = P.s1; S[1] = P.s2; S[2] = P.s3; S[3] = P.s4; S[4] = P.s5;
    if (MIXED) { S[4] = dv; S[5] = dv0; S[6] = thetahat; S[7] = J - Jhat - thetahat * pres; }
    if (COMPACT || GPAD) {
      Bv[0] = P.b[0][0]; Bv[1] = P.b[1][1]; Bv[2] = P.b[2][2];
      Bv[3] = P.b[0][1]; Bv[4] = P.b[0][2]; Bv[5] = P.b[1][2];
    }
  }
}

// b g from the six stored components of the symmetric b (same products and order as element_point).
__device__ __forceinline__ void bsym_times(const double* B6, const double g[3], double bg[3]) {
  bg[0] = B6[0] * g[0] + B6[3] * g[1] + B6[4] * g[2];
  bg[1] = B6[3] * g[0] + B6[1] * g[1] + B6[5] * g[2];
  bg[2] = B6[4] * g[0] + B6[5] * g[1] + B6[2] * g[2];
}

// Material parameters of element e: (mu, K) from the per-element table when present.
__device__ __forceinline__ void elem_material(const ElemArgs& A, int e, double& mu, double& K) {
  if (A.emat) {
    const int t = A.emat[e];
    mu = A.mtab[2 * t];
    K = A.mtab[2 * t + 1];
  } else {
    mu = A.mu;
    K = 1.0 / A.Kinv;
  }
}

// One Gauss point's share of the 3x3 tangent block (a,b):
//   K_ab^{ik} += s1 g^a_k g^b_i + (s2 g^a_i + s3 (b g^a)_i) g^b_k + s3 g^a_i (b g^b)_k + delta_ik (s4 b g^a + s5 g^a).g^b
// with g = grad_x N and b the left Cauchy-Green tensor (the closed form of the reference's
// Gc = grad N . Cmat contraction, Elem_Solid_3D_Disp.cpp:315-366).
template <bool HAS_S5>
__device__ __forceinline__ void tangent_gp(const double ga[3], const double ba[3], const double gb[3], const double bb[3],
                                           double s1, double s2, double s3, double s4, double s5, double Kb[9]) {
  const double y = HAS_S5 ? (s4 * ba[0] + s5 * ga[0]) * gb[0] + (s4 * ba[1] + s5 * ga[1]) * gb[1] +
                                (s4 * ba[2] + s5 * ga[2]) * gb[2]
                          : s4 * (ba[0] * gb[0] + ba[1] * gb[1] + ba[2] * gb[2]);
  const double s1ga[3] = {s1 * ga[0], s1 * ga[1], s1 * ga[2]};
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const double Qi = s2 * ga[i] + s3 * ba[i], Ti = s3 * ga[i];
#pragma unroll
    for (int kk = 0; kk < 3; ++kk) {  // three fused multiply-adds into the running block entry
      double t = Kb[3 * i + kk];
      t = fma(s1ga[kk], gb[i], t);
      t = fma(Qi, gb[kk], t);
      t = fma(Ti, bb[kk], t);
      Kb[3 * i + kk] = i == kk ? t + y : t;
    }
  }
}

// 3x3 tangent block (a,b) from the staged Gauss-point data of one element: Ga = G of node a,
// Gb = G of node b (6 doubles each per Gauss point, stride gs between points), S scalars (stride ss).
template <int NGP, bool HAS_S5>
__device__ __forceinline__ void tangent_block(const double* Ga, const double* Gb, int gs, const double* S, int ss,
                                              double Kb[9]) {
#pragma unroll
  for (int t = 0; t < 9; ++t) Kb[t] = 0.0;
#pragma unroll 2
  for (int gp = 0; gp < NGP; ++gp) {
    const double* ga_ = Ga + gp * gs;
    const double* gb_ = Gb + gp * gs;
    const double* s = S + gp * ss;
    const double ga[3] = {ga_[0], ga_[1], ga_[2]}, ba[3] = {ga_[3], ga_[4], ga_[5]};
    const double gb[3] = {gb_[0], gb_[1], gb_[2]}, bb[3] = {gb_[3], gb_[4], gb_[5]};
    tangent_gp<HAS_S5>(ga, ba, gb, bb, s[0], s[1], s[2], s[3], HAS_S5 ? s[4] : 0.0, Kb);
  }
}

}  // namespace dev
}  // namespace mpf
