// elem.hip -- FP64 Neo-Hookean element kernels for gfx950 (Hex8 displacement, Tet10/P1 mixed).
//
// Reference path (chennachaos/mphyfem, src/):
//   Elem_Solid_3D_Disp::calcStiffnessAndResidual      Elem_Solid_3D_Disp.cpp:196-374
//   ElementBase::calcStiffnessAndResidualMixed3D      ElementBase2.cpp:278-552
//   GeomDataLagrange::computeBasisFunctions3D          GeomDataLagrange.cpp:349-458
//   computeStressAndTangent_NeoHooke                   utilitiesmaterialHyperelastic.cpp:186-295
//   applyDirichletBCs / applyDirichletBCs2field        ElementBase.cpp:567-594, 717-782
//   SolverPetsc::assembleMatrixAndVector(2field)       SolverPetsc.cpp:357-378, 434-502
//
// Work split.  One 64-lane wavefront owns EPW elements.
//   Phase 1: lane = (element slot k, Gauss point gp).  The lane builds dN/dxi at gp, the
//            reference-configuration Jacobian, F = I + sum u_a (x) grad_X N_a, the current
//            Jacobian (x = X + u), grad_x N for all nodes, and the Neo-Hookean state, and
//            stages g_a = grad_x N_a, (b g)_a and five tangent scalars in LDS.  Residual
//            contributions are reduced over Gauss points with lane shuffles.
//   Phase 2: lane = node pair (a,b) (Tet10: two passes plus u-p lanes).  The 3x3 block
//            K_ab = sum_gp [ s1 g^a_k g^b_i + (s2 g^a_i + s3 (bg^a)_i) g^b_k
//                            + s3 g^a_i (bg^b)_k + delta_ik (s4 bg^a + s5 g^a).g^b ]
//            is the closed-form contraction  sum_{m,n} g^a_m C[(i,m),(k,n)] g^b_n  of the
//            reference's 9x9 spatial tangent (same terms, fewer flops), with
//              disp : s1 = dv (f1 - U'), s2 = dv (f3 + J U'' + U'), s3 = dv f2, s4 = dv f4, s5 = -dv U'
//                     (s5 = 0 for the consistent tangent, which adds the volumetric stress term)
//              mixed: s1 = dv (f1 - p),  s2 = dv (f3 + p),          s3 = dv f2, s4 = dv f4, s5 = 0
//            where f1 = mu J^-5/3 I1/3, f2 = -2/3 mu J^-5/3, f3 = 2/9 mu J^-5/3 I1, f4 = mu J^-5/3.
//            The block is either written densely (element operator) or added into the CSR
//            values through the per-element slot map (one colour per launch: no two elements
//            of a launch share a node, so the FP64 read-modify-writes never race).

#include "mpf_internal.hpp"

namespace mpf {

namespace {

constexpr double kGP2 = 0.577350269189626;  // QuadratureUtil.cpp:24-25

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

__device__ void flag_bad(int32_t* bad, int32_t e) {
  if (!bad) return;
  const int32_t v = e + 1;
  int32_t old = __hip_atomic_load(bad, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  while (old == 0 || v < old) {
    const int32_t prev = atomicCAS(bad, old, v);
    if (prev == old) break;
    old = prev;
  }
}

// Neo-Hookean state at one Gauss point: Cauchy stress (symmetric, 6 stored) and the five
// tangent scalars of the closed-form contraction (multiplied by dv).  See file header.
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

// Tet Gauss point gp of the 4-point rule (QuadratureUtil.cpp:690-712) and Tet10 / Tet4 bases
// (BasisFunctionsLagrange.cpp:870-956).
__device__ __forceinline__ void tet_gp(int gp, double& x1, double& x2, double& x3) {
  const double a = 0.1381966011250105, b = 0.5854101966249685;
  x1 = (gp == 1) ? b : a;
  x2 = (gp == 2) ? b : a;
  x3 = (gp == 3) ? b : a;
}

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

__device__ __forceinline__ double shfl_xor_d(double v, int m) { return __shfl_xor(v, m, 64); }

__device__ __forceinline__ int rank3(int mask, int i) { return __popc(mask & ((1 << i) - 1)); }

// ============================================================================================
// Hex8 displacement element
// ============================================================================================
template <int EPW>
struct Hex8Smem {
  double X[EPW][8][3];
  double U[EPW][8][3];
  double G[EPW][8][8][6];   // [k][gp][a]: g_a (3), (b g)_a (3)
  double S[EPW][8][5];      // [k][gp]: s1..s5
  double R[EPW][8][3];      // [k][a]: element residual
};

template <int EPW, bool DENSE>
__global__ void __launch_bounds__(64) hex8_kernel(ElemArgs A, int32_t first, int32_t count, double* __restrict__ Ke,
                                                  double* __restrict__ Re, ScatterArgs sc) {
  __shared__ Hex8Smem<EPW> sm;
  const int lane = threadIdx.x;
  const int base_idx = blockIdx.x * EPW;
  const double K = 1.0 / A.Kinv;

  // ---- stage X and u of the EPW elements (lane loads node (k, a) triples) ----
  for (int t = lane; t < EPW * 8; t += 64) {
    const int k = t >> 3, a = t & 7;
    const int idx = base_idx + k;
    double x0 = 0, x1 = 0, x2 = 0, u0 = 0, u1 = 0, u2 = 0;
    if (idx < count) {
      const int e = DENSE ? first + idx : sc.elems[idx];
      const int nd = A.conn[(size_t)e * 8 + a];
      x0 = A.X[3 * (size_t)nd]; x1 = A.X[3 * (size_t)nd + 1]; x2 = A.X[3 * (size_t)nd + 2];
      u0 = A.u[3 * (size_t)nd]; u1 = A.u[3 * (size_t)nd + 1]; u2 = A.u[3 * (size_t)nd + 2];
    }
    sm.X[k][a][0] = x0; sm.X[k][a][1] = x1; sm.X[k][a][2] = x2;
    sm.U[k][a][0] = u0; sm.U[k][a][1] = u1; sm.U[k][a][2] = u2;
  }
  __syncthreads();

  // ---- phase 1: lane = (k, gp) ----
  {
    const int k = lane >> 3, gp = lane & 7;
    const bool act = (k < EPW) && (base_idx + k < count);
    double R[8][3];
#pragma unroll
    for (int a = 0; a < 8; ++a) R[a][0] = R[a][1] = R[a][2] = 0.0;
    if (k < EPW) {
      double d[3][8];
      hex8_dN(gp, d);
      // reference Jacobian B(r,c) = sum_a dN_a/dxi_r X_a[c]
      double B[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, Bi[3][3];
#pragma unroll
      for (int a = 0; a < 8; ++a)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const double xc = sm.X[k][a][c];
          B[0][c] += xc * d[0][a]; B[1][c] += xc * d[1][a]; B[2][c] += xc * d[2][a];
        }
      inv3(B, Bi);
      // F = I + sum_a u_a (x) grad_X N_a
      double F[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
#pragma unroll
      for (int a = 0; a < 8; ++a) {
        const double gx = d[0][a] * Bi[0][0] + d[1][a] * Bi[0][1] + d[2][a] * Bi[0][2];
        const double gy = d[0][a] * Bi[1][0] + d[1][a] * Bi[1][1] + d[2][a] * Bi[1][2];
        const double gz = d[0][a] * Bi[2][0] + d[1][a] * Bi[2][1] + d[2][a] * Bi[2][2];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          const double ui = sm.U[k][a][i];
          F[i][0] += ui * gx; F[i][1] += ui * gy; F[i][2] += ui * gz;
        }
      }
      const double J = det3(F);
      if (act && J < 0.0) flag_bad(A.bad, DENSE ? first + base_idx + k : sc.elems[base_idx + k]);
      // current Jacobian with x = X + u
      double Bc[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll
      for (int a = 0; a < 8; ++a)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const double xc = sm.X[k][a][c] + sm.U[k][a][c];
          Bc[0][c] += xc * d[0][a]; Bc[1][c] += xc * d[1][a]; Bc[2][c] += xc * d[2][a];
        }
      const double Jac = inv3(Bc, Bi);
      const double dv = 1.0 * Jac;  // Hex8 weights are 1
      NHPoint P;
      neohooke_point(F, J, A.mu, K, A.Utype, false, 0.0, dv, A.consistent != 0, P);
#pragma unroll
      for (int a = 0; a < 8; ++a) {
        const double g0 = d[0][a] * Bi[0][0] + d[1][a] * Bi[0][1] + d[2][a] * Bi[0][2];
        const double g1 = d[0][a] * Bi[1][0] + d[1][a] * Bi[1][1] + d[2][a] * Bi[1][2];
        const double g2 = d[0][a] * Bi[2][0] + d[1][a] * Bi[2][1] + d[2][a] * Bi[2][2];
        double* G = sm.G[k][gp][a];
        G[0] = g0; G[1] = g1; G[2] = g2;
        G[3] = P.b[0][0] * g0 + P.b[0][1] * g1 + P.b[0][2] * g2;
        G[4] = P.b[1][0] * g0 + P.b[1][1] * g1 + P.b[1][2] * g2;
        G[5] = P.b[2][0] * g0 + P.b[2][1] * g1 + P.b[2][2] * g2;
        const double b0 = dv * g0, b1 = dv * g1, b2 = dv * g2;
#pragma unroll
        for (int i = 0; i < 3; ++i) R[a][i] = -(b0 * P.sig[i][0] + b1 * P.sig[i][1] + b2 * P.sig[i][2]);
      }
      double* S = sm.S[k][gp];
      S[0] = P.s1; S[1] = P.s2; S[2] = P.s3; S[3] = P.s4; S[4] = P.s5;
    }
    // reduce R over the 8 Gauss points of element k (lanes k*8 .. k*8+7)
#pragma unroll
    for (int a = 0; a < 8; ++a)
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        double v = R[a][i];
        v += shfl_xor_d(v, 1);
        v += shfl_xor_d(v, 2);
        v += shfl_xor_d(v, 4);
        R[a][i] = v;
      }
    if (k < EPW) {
#pragma unroll
      for (int a = 0; a < 8; ++a)
        if (a == gp) { sm.R[k][a][0] = R[a][0]; sm.R[k][a][1] = R[a][1]; sm.R[k][a][2] = R[a][2]; }
    }
  }
  __syncthreads();

  // ---- phase 2: lane = (a, b) ----
  const int a = lane >> 3, b = lane & 7;
  for (int k = 0; k < EPW; ++k) {
    const int idx = base_idx + k;
    if (idx >= count) break;  // uniform across the wave
    double Kb[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll 2
    for (int gp = 0; gp < 8; ++gp) {
      const double* Ga = sm.G[k][gp][a];
      const double* Gb = sm.G[k][gp][b];
      const double* S = sm.S[k][gp];
      const double s1 = S[0], s2 = S[1], s3 = S[2], s4 = S[3], s5 = S[4];
      const double ga0 = Ga[0], ga1 = Ga[1], ga2 = Ga[2], ba0 = Ga[3], ba1 = Ga[4], ba2 = Ga[5];
      const double gb0 = Gb[0], gb1 = Gb[1], gb2 = Gb[2], bb0 = Gb[3], bb1 = Gb[4], bb2 = Gb[5];
      const double P0 = s1 * ga0, P1 = s1 * ga1, P2 = s1 * ga2;
      const double Q0 = s2 * ga0 + s3 * ba0, Q1 = s2 * ga1 + s3 * ba1, Q2 = s2 * ga2 + s3 * ba2;
      const double T0 = s3 * ga0, T1 = s3 * ga1, T2 = s3 * ga2;
      const double y = (s4 * ba0 + s5 * ga0) * gb0 + (s4 * ba1 + s5 * ga1) * gb1 + (s4 * ba2 + s5 * ga2) * gb2;
      const double Pv[3] = {P0, P1, P2}, Qv[3] = {Q0, Q1, Q2}, Tv[3] = {T0, T1, T2};
      const double gbv[3] = {gb0, gb1, gb2}, bbv[3] = {bb0, bb1, bb2};
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int kk = 0; kk < 3; ++kk)
          Kb[i][kk] += Pv[kk] * gbv[i] + Qv[i] * gbv[kk] + Tv[i] * bbv[kk] + (i == kk ? y : 0.0);
    }
    if (DENSE) {
      double* Ko = Ke + (size_t)idx * 576;
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int kk = 0; kk < 3; ++kk) Ko[(3 * a + i) * 24 + 3 * b + kk] = Kb[i][kk];
      if (b == 0) {
        double* Ro = Re + (size_t)idx * 24;
        Ro[3 * a] = sm.R[k][a][0]; Ro[3 * a + 1] = sm.R[k][a][1]; Ro[3 * a + 2] = sm.R[k][a][2];
      }
    } else {
      const int e = sc.elems[idx];
      const int na = A.conn[(size_t)e * 8 + a], nb = A.conn[(size_t)e * 8 + b];
      const int infa = sc.nodeinfo[na], infb = sc.nodeinfo[nb];
      const int ma = infa & 7, mb = infb & 7, La = infa >> 4;
      // Dirichlet lift: F_a -= K_ab * dubar_b over prescribed components of b (first iteration)
      double c0 = 0.0, c1 = 0.0, c2 = 0.0;
      if (sc.dispApplied) {
        double db[3];
#pragma unroll
        for (int kk = 0; kk < 3; ++kk) db[kk] = ((mb >> kk) & 1) ? 0.0 : sc.dispApplied[3 * (size_t)nb + kk];
        c0 = Kb[0][0] * db[0] + Kb[0][1] * db[1] + Kb[0][2] * db[2];
        c1 = Kb[1][0] * db[0] + Kb[1][1] * db[1] + Kb[1][2] * db[2];
        c2 = Kb[2][0] * db[0] + Kb[2][1] * db[1] + Kb[2][2] * db[2];
        c0 += shfl_xor_d(c0, 1); c0 += shfl_xor_d(c0, 2); c0 += shfl_xor_d(c0, 4);
        c1 += shfl_xor_d(c1, 1); c1 += shfl_xor_d(c1, 2); c1 += shfl_xor_d(c1, 4);
        c2 += shfl_xor_d(c2, 1); c2 += shfl_xor_d(c2, 2); c2 += shfl_xor_d(c2, 4);
      }
      const int s0 = (sc.scatterK && ma && mb) ? sc.slot[(size_t)e * kHex8Slots + lane] : -1;
      if (s0 >= 0) {
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          if (!((ma >> i) & 1)) continue;
          const int rowoff = s0 + rank3(ma, i) * La;
#pragma unroll
          for (int kk = 0; kk < 3; ++kk) {
            if (!((mb >> kk) & 1)) continue;
            double* v = sc.vals + rowoff + rank3(mb, kk);
            *v += Kb[i][kk];
          }
        }
      }
      if (b == 0 && ma) {
        const double cv[3] = {c0, c1, c2};
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          if (!((ma >> i) & 1)) continue;
          const int r = sc.nda[3 * (size_t)na + i];
          sc.rhs[r] += sm.R[k][a][i] - cv[i];
        }
      }
    }
  }
}

// ============================================================================================
// Tet10 / P1 mixed element
// ============================================================================================
template <int EPW>
struct Tet10Smem {
  double X[EPW][10][3];
  double U[EPW][10][3];
  double UP[EPW][10][3];
  double Pn[EPW][4];        // corner pressures
  double G[EPW][4][10][6];  // [k][gp][a]: g_a, (b g)_a
  double S[EPW][4][8];      // [k][gp]: s1..s4, dv, dv0, thetahat, resip
  double R[EPW][10][3];     // Fu
  double FpS[EPW][4];       // Fp
  double LP[EPW][10][4];    // first-iteration lift of Fp from the prescribed dofs of node a: Kpu(q, a) dubar_a
};

// Tet4 basis at the Gauss points of the 4-point rule: N0 = xi4, N1 = xi1, N2 = xi2, N3 = xi3.
__device__ __forceinline__ void tet_Np(int gp, double Np[4]) {
  double x1, x2, x3;
  tet_gp(gp, x1, x2, x3);
  Np[0] = 1.0 - x1 - x2 - x3; Np[1] = x1; Np[2] = x2; Np[3] = x3;
}

// Kuu block (a,b) of element slot k, summed over the 4 Gauss points.
template <int EPW>
__device__ __forceinline__ void tet10_uu_block(const Tet10Smem<EPW>& sm, int k, int a, int b, double Kb[3][3]) {
#pragma unroll
  for (int i = 0; i < 3; ++i) Kb[i][0] = Kb[i][1] = Kb[i][2] = 0.0;
#pragma unroll
  for (int gp = 0; gp < 4; ++gp) {
    const double* Ga = sm.G[k][gp][a];
    const double* Gb = sm.G[k][gp][b];
    const double* S = sm.S[k][gp];
    const double s1 = S[0], s2 = S[1], s3 = S[2], s4 = S[3];
    const double ga[3] = {Ga[0], Ga[1], Ga[2]}, ba[3] = {Ga[3], Ga[4], Ga[5]};
    const double gb[3] = {Gb[0], Gb[1], Gb[2]}, bb[3] = {Gb[3], Gb[4], Gb[5]};
    const double y = s4 * (ba[0] * gb[0] + ba[1] * gb[1] + ba[2] * gb[2]);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int kk = 0; kk < 3; ++kk)
        Kb[i][kk] += (s1 * ga[kk]) * gb[i] + (s2 * ga[i] + s3 * ba[i]) * gb[kk] + (s3 * ga[i]) * bb[kk] +
                     (i == kk ? y : 0.0);
  }
}

template <int EPW, bool DENSE>
__global__ void __launch_bounds__(64) tet10_kernel(ElemArgs A, int32_t first, int32_t count, double* __restrict__ Kuu,
                                                   double* __restrict__ Kup, double* __restrict__ Kpu,
                                                   double* __restrict__ Kpp, double* __restrict__ Fu,
                                                   double* __restrict__ Fp, ScatterArgs sc) {
  __shared__ Tet10Smem<EPW> sm;
  const int lane = threadIdx.x;
  const int base_idx = blockIdx.x * EPW;
  const double K = 1.0 / A.Kinv;
  const double wq = 0.2500000000000000 * (1.0 / 6.0);  // QuadratureUtil.cpp:706-709

  for (int t = lane; t < EPW * 10; t += 64) {
    const int k = t / 10, a = t % 10;
    const int idx = base_idx + k;
    double x[3] = {0, 0, 0}, u[3] = {0, 0, 0}, up[3] = {0, 0, 0}, pn = 0.0;
    if (idx < count) {
      const int e = DENSE ? first + idx : sc.elems[idx];
      const size_t nd = (size_t)A.conn[(size_t)e * 10 + a];
#pragma unroll
      for (int d = 0; d < 3; ++d) { x[d] = A.X[3 * nd + d]; u[d] = A.u[3 * nd + d]; up[d] = A.uprev[3 * nd + d]; }
      if (a < 4) pn = A.pres[nd];
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) { sm.X[k][a][d] = x[d]; sm.U[k][a][d] = u[d]; sm.UP[k][a][d] = up[d]; }
    if (a < 4) sm.Pn[k][a] = pn;
  }
  __syncthreads();

  // ---- phase 1: lane = (k, gp), 4 Gauss points per element ----
  {
    const int k = lane >> 2, gp = lane & 3;
    const bool act = (k < EPW) && (base_idx + k < count);
    double R[10][3];
#pragma unroll
    for (int a = 0; a < 10; ++a) R[a][0] = R[a][1] = R[a][2] = 0.0;
    if (k < EPW) {
      double x1, x2, x3;
      tet_gp(gp, x1, x2, x3);
      double d[3][10];
      tet10_dN(x1, x2, x3, d[0], d[1], d[2]);
      double B[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, Bi[3][3];
#pragma unroll
      for (int a = 0; a < 10; ++a)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const double xc = sm.X[k][a][c];
          B[0][c] += xc * d[0][a]; B[1][c] += xc * d[1][a]; B[2][c] += xc * d[2][a];
        }
      const double Jac0 = inv3(B, Bi);
      const double dv0 = wq * Jac0;
      double F[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}, Fn[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
#pragma unroll
      for (int a = 0; a < 10; ++a) {
        const double gx = d[0][a] * Bi[0][0] + d[1][a] * Bi[0][1] + d[2][a] * Bi[0][2];
        const double gy = d[0][a] * Bi[1][0] + d[1][a] * Bi[1][1] + d[2][a] * Bi[1][2];
        const double gz = d[0][a] * Bi[2][0] + d[1][a] * Bi[2][1] + d[2][a] * Bi[2][2];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          const double ui = sm.U[k][a][i], upi = sm.UP[k][a][i];
          F[i][0] += ui * gx; F[i][1] += ui * gy; F[i][2] += ui * gz;
          Fn[i][0] += upi * gx; Fn[i][1] += upi * gy; Fn[i][2] += upi * gz;
        }
      }
      // compute_constants_volumetric_functions, finite strain (utilitiesmaterialHyperelastic.cpp:144-173)
      const double Jn = det3(Fn);
      double Jhat = 1.0, thetahat = 0.0;
      if (A.Utype != -1) {
        double d1, d2;
        vol_derivs(A.Utype, K, Jn, d1, d2);
        Jhat = Jn - d1 / d2;
        thetahat = 1.0 / d2;
      }
      const double J = det3(F);
      if (act && J < 0.0) flag_bad(A.bad, DENSE ? first + base_idx + k : sc.elems[base_idx + k]);
      double Bc[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll
      for (int a = 0; a < 10; ++a)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const double xc = sm.X[k][a][c] + sm.U[k][a][c];
          Bc[0][c] += xc * d[0][a]; Bc[1][c] += xc * d[1][a]; Bc[2][c] += xc * d[2][a];
        }
      const double Jac = inv3(Bc, Bi);
      const double dv = wq * Jac;
      double Np[4];
      tet_Np(gp, Np);
      double pres = 0.0;
#pragma unroll
      for (int q = 0; q < 4; ++q) pres += Np[q] * sm.Pn[k][q];
      NHPoint P;
      neohooke_point(F, J, A.mu, K, A.Utype, true, pres, dv, false, P);
#pragma unroll
      for (int a = 0; a < 10; ++a) {
        const double g0 = d[0][a] * Bi[0][0] + d[1][a] * Bi[0][1] + d[2][a] * Bi[0][2];
        const double g1 = d[0][a] * Bi[1][0] + d[1][a] * Bi[1][1] + d[2][a] * Bi[1][2];
        const double g2 = d[0][a] * Bi[2][0] + d[1][a] * Bi[2][1] + d[2][a] * Bi[2][2];
        double* G = sm.G[k][gp][a];
        G[0] = g0; G[1] = g1; G[2] = g2;
        G[3] = P.b[0][0] * g0 + P.b[0][1] * g1 + P.b[0][2] * g2;
        G[4] = P.b[1][0] * g0 + P.b[1][1] * g1 + P.b[1][2] * g2;
        G[5] = P.b[2][0] * g0 + P.b[2][1] * g1 + P.b[2][2] * g2;
        const double b0 = dv * g0, b1 = dv * g1, b2 = dv * g2;
#pragma unroll
        for (int i = 0; i < 3; ++i) R[a][i] = -(b0 * P.sig[i][0] + b1 * P.sig[i][1] + b2 * P.sig[i][2]);
      }
      double* S = sm.S[k][gp];
      S[0] = P.s1; S[1] = P.s2; S[2] = P.s3; S[3] = P.s4;
      S[4] = dv; S[5] = dv0; S[6] = thetahat; S[7] = J - Jhat - thetahat * pres;
    }
#pragma unroll
    for (int a = 0; a < 10; ++a)
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        double v = R[a][i];
        v += shfl_xor_d(v, 1);
        v += shfl_xor_d(v, 2);
        R[a][i] = v;
      }
    if (k < EPW) {
#pragma unroll
      for (int a = 0; a < 10; ++a)
        if ((a & 3) == gp) { sm.R[k][a][0] = R[a][0]; sm.R[k][a][1] = R[a][1]; sm.R[k][a][2] = R[a][2]; }
    }
  }
  __syncthreads();

  // ---- phase 2: items 0..99 = Kuu blocks (a,b); 100..109 = Kup/Kpu of node a; 110..113 = Kpp row q ----
  for (int k = 0; k < EPW; ++k) {
    const int idx = base_idx + k;
    if (idx >= count) break;  // uniform across the wave
    const int e = DENSE ? first + idx : sc.elems[idx];
    const bool lift = !DENSE && sc.dispApplied != nullptr;
    for (int pass = 0; pass < 2; ++pass) {
      const int item = pass * 64 + lane;
      if (item < 100) {
        const int a = item / 10, b = item % 10;
        double Kb[3][3];
        tet10_uu_block(sm, k, a, b, Kb);
        if (DENSE) {
          double* Ko = Kuu + (size_t)idx * 900;
#pragma unroll
          for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int kk = 0; kk < 3; ++kk) Ko[(3 * a + i) * 30 + 3 * b + kk] = Kb[i][kk];
        } else if (sc.scatterK) {
          const int na = A.conn[(size_t)e * 10 + a], nb = A.conn[(size_t)e * 10 + b];
          const int infa = sc.nodeinfo[na], infb = sc.nodeinfo[nb];
          const int ma = infa & 7, mb = infb & 7, La = infa >> 4;
          const int s0 = (ma && mb) ? sc.slot[(size_t)e * kTet10Slots + a * 10 + b] : -1;
          if (s0 >= 0) {
#pragma unroll
            for (int i = 0; i < 3; ++i) {
              if (!((ma >> i) & 1)) continue;
              const int rowoff = s0 + rank3(ma, i) * La;
#pragma unroll
              for (int kk = 0; kk < 3; ++kk) {
                if (!((mb >> kk) & 1)) continue;
                sc.vals[rowoff + rank3(mb, kk)] += Kb[i][kk];
              }
            }
          }
        }
      } else if (item < 110) {
        // Kup(3a+i, q) += (dv g_a,i) Np_q ;  Kpu(q, 3a+j) += (dv Np_q) g_a,j   (ElementBase2.cpp:500-505, 524-533)
        const int a = item - 100;
        double up[3][4], pu[4][3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int q = 0; q < 4; ++q) { up[i][q] = 0.0; pu[q][i] = 0.0; }
#pragma unroll
        for (int gp = 0; gp < 4; ++gp) {
          const double* Ga = sm.G[k][gp][a];
          const double dv = sm.S[k][gp][4];
          double Np[4];
          tet_Np(gp, Np);
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const double bb4 = dv * Np[q];
#pragma unroll
            for (int i = 0; i < 3; ++i) {
              up[i][q] += (dv * Ga[i]) * Np[q];
              pu[q][i] += bb4 * Ga[i];
            }
          }
        }
        if (DENSE) {
#pragma unroll
          for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              Kup[(size_t)idx * 120 + (3 * a + i) * 4 + q] = up[i][q];
              Kpu[(size_t)idx * 120 + q * 30 + 3 * a + i] = pu[q][i];
            }
        } else {
          const int na = A.conn[(size_t)e * 10 + a];
          const int infa = sc.nodeinfo[na];
          const int ma = infa & 7, La = infa >> 4;
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            double c = 0.0;
            if (lift) {
#pragma unroll
              for (int j = 0; j < 3; ++j)
                if (!((ma >> j) & 1)) c += pu[q][j] * sc.dispApplied[3 * (size_t)na + j];
            }
            sm.LP[k][a][q] = c;
          }
          if (sc.scatterK && ma) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              const int sup = sc.slot[(size_t)e * kTet10Slots + 100 + a * 4 + q];
              const int spu = sc.slot[(size_t)e * kTet10Slots + 140 + q * 10 + a];
              if (sup < 0 || spu < 0) continue;
#pragma unroll
              for (int i = 0; i < 3; ++i) {
                if (!((ma >> i) & 1)) continue;
                sc.vals[sup + rank3(ma, i) * La] += up[i][q];
                sc.vals[spu + rank3(ma, i)] += pu[q][i];
              }
            }
          }
        }
      } else if (item < 114) {
        // Kpp(q, q2) -= (dv0 Np_q thetahat) Np_q2 ;  Fp_q -= (dv0 Np_q) resip   (ElementBase2.cpp:510-539)
        const int q = item - 110;
        double pp[4] = {0, 0, 0, 0}, fp = 0.0;
#pragma unroll
        for (int gp = 0; gp < 4; ++gp) {
          double Np[4];
          tet_Np(gp, Np);
          const double* S = sm.S[k][gp];
          double bb5 = S[5] * Np[q];
          fp -= bb5 * S[7];
          bb5 *= S[6];
#pragma unroll
          for (int q2 = 0; q2 < 4; ++q2) pp[q2] -= bb5 * Np[q2];
        }
        if (DENSE) {
#pragma unroll
          for (int q2 = 0; q2 < 4; ++q2) Kpp[(size_t)idx * 16 + q * 4 + q2] = pp[q2];
          Fp[(size_t)idx * 4 + q] = fp;
        } else {
          if (sc.scatterK) {
#pragma unroll
            for (int q2 = 0; q2 < 4; ++q2) {
              const int sp = sc.slot[(size_t)e * kTet10Slots + 180 + q * 4 + q2];
              if (sp >= 0) sc.vals[sp] += pp[q2];
            }
          }
          sm.FpS[k][q] = fp;
        }
      }
    }
    if (DENSE) {
      if (lane < 30) Fu[(size_t)idx * 30 + lane] = sm.R[k][lane / 3][lane % 3];
      continue;
    }
    __syncthreads();
    // residual rows with the first-iteration lift (applyDirichletBCs2field, ElementBase.cpp:717-782)
    if (lane < 10) {
      const int a = lane;
      const int na = A.conn[(size_t)e * 10 + a];
      const int ma = sc.nodeinfo[na] & 7;
      double c[3] = {0, 0, 0};
      if (lift && ma) {
        for (int b = 0; b < 10; ++b) {
          const int nb = A.conn[(size_t)e * 10 + b];
          const int mb = sc.nodeinfo[nb] & 7;
          if (mb == 7) continue;
          double db[3];
#pragma unroll
          for (int kk = 0; kk < 3; ++kk) db[kk] = ((mb >> kk) & 1) ? 0.0 : sc.dispApplied[3 * (size_t)nb + kk];
          double Kb[3][3];
          tet10_uu_block(sm, k, a, b, Kb);
#pragma unroll
          for (int i = 0; i < 3; ++i) c[i] += Kb[i][0] * db[0] + Kb[i][1] * db[1] + Kb[i][2] * db[2];
        }
      }
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        if (!((ma >> i) & 1)) continue;
        sc.rhs[sc.nda[3 * (size_t)na + i]] += sm.R[k][a][i] - c[i];
      }
    } else if (lane < 14) {
      const int q = lane - 10;
      const int nq = A.conn[(size_t)e * 10 + q];
      double c = 0.0;
      for (int a = 0; a < 10; ++a) c += sm.LP[k][a][q];
      sc.rhs[sc.dispDOF + sc.idp[nq]] += sm.FpS[k][q] - c;
    }
    __syncthreads();
  }
}

}  // namespace

// --------------------------------------------------------------------------------------------
// Launchers
// --------------------------------------------------------------------------------------------
constexpr int kHexEPW = 8;
constexpr int kTetEPW = 8;

mpf_status launch_hex8_dense(const ElemArgs& a, int32_t first, int32_t count, double* Ke, double* Re,
                             hipStream_t s) {
  if (count <= 0) return MPF_OK;
  ScatterArgs sc{};
  hipLaunchKernelGGL((hex8_kernel<kHexEPW, true>), dim3(grid_for(count, kHexEPW)), dim3(64), 0, s, a, first, count,
                     Ke, Re, sc);
  MPF_HIP(hipGetLastError());
  return MPF_OK;
}

mpf_status launch_hex8_scatter(const ElemArgs& a, const ScatterArgs& sc, hipStream_t s) {
  if (sc.count <= 0) return MPF_OK;
  hipLaunchKernelGGL((hex8_kernel<kHexEPW, false>), dim3(grid_for(sc.count, kHexEPW)), dim3(64), 0, s, a, 0,
                     sc.count, nullptr, nullptr, sc);
  MPF_HIP(hipGetLastError());
  return MPF_OK;
}

mpf_status launch_tet10_dense(const ElemArgs& a, int32_t first, int32_t count, double* Kuu, double* Kup,
                              double* Kpu, double* Kpp, double* Fu, double* Fp, hipStream_t s) {
  if (count <= 0) return MPF_OK;
  ScatterArgs sc{};
  hipLaunchKernelGGL((tet10_kernel<kTetEPW, true>), dim3(grid_for(count, kTetEPW)), dim3(64), 0, s, a, first, count,
                     Kuu, Kup, Kpu, Kpp, Fu, Fp, sc);
  MPF_HIP(hipGetLastError());
  return MPF_OK;
}

mpf_status launch_tet10_scatter(const ElemArgs& a, const ScatterArgs& sc, hipStream_t s) {
  if (sc.count <= 0) return MPF_OK;
  hipLaunchKernelGGL((tet10_kernel<kTetEPW, false>), dim3(grid_for(sc.count, kTetEPW)), dim3(64), 0, s, a, 0,
                     sc.count, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, sc);
  MPF_HIP(hipGetLastError());
  return MPF_OK;
}

}  // namespace mpf
