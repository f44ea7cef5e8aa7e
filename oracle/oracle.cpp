// oracle.cpp -- CPU FP64 restatement of chennachaos/mphyfem's femSolidmechanics
// assembly + Newton-Krylov path.  TEST INFRASTRUCTURE ONLY (the test oracle): only
// tests/, the smoke entry point and bench.py's CPU-baseline leg load it.
//
// The code follows the reference's loops and operation order where they fix
// the arithmetic (element, material, Dirichlet lift, assembly), using plain
// arrays instead of Eigen.  Citations are relative to the reference's src/.
//
// Parity status per function (how each is pinned) is listed in the module
// docstring of tests/test_oracle.py; functions whose bit-level behaviour
// depends on a third-party library (Eigen's LU, PETSc's KSP) are pinned only
// to rounding level, never bit for bit.

#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <vector>

namespace {

// ---------------------------------------------------------------------------
// Quadrature  (QuadratureUtil.cpp)
// ---------------------------------------------------------------------------

// getGaussPoints1D, 2-point rule literal (QuadratureUtil.cpp:24-25)
const double GP2 = 0.577350269189626;

// getGaussPointsHexa, ngp = 8 (QuadratureUtil.cpp:453-492): (--,-),(+-,-),(++,-),(-+,-), then zeta=+.
int gauss_hexa8(double xi[][3], double w[]) {
  const double g[2] = {-GP2, GP2};
  const int p[8][3] = {{0,0,0},{1,0,0},{1,1,0},{0,1,0},{0,0,1},{1,0,1},{1,1,1},{0,1,1}};
  for (int q = 0; q < 8; ++q) {
    xi[q][0] = g[p[q][0]]; xi[q][1] = g[p[q][1]]; xi[q][2] = g[p[q][2]]; w[q] = 1.0;
  }
  return 8;
}

// getGaussPointsTetra (QuadratureUtil.cpp:670-712); ngp 1 and 4.
int gauss_tetra(int ngp, double xi[][3], double w[]) {
  const double fact = 1.0 / 6.0;
  if (ngp == 1) {
    xi[0][0] = xi[0][1] = xi[0][2] = 0.25; w[0] = fact; return 1;
  }
  const double a = 0.1381966011250105, b = 0.5854101966249685;
  const double g1[4] = {a, b, a, a}, g2[4] = {a, a, b, a}, g3[4] = {a, a, a, b};
  for (int q = 0; q < 4; ++q) { xi[q][0] = g1[q]; xi[q][1] = g2[q]; xi[q][2] = g3[q]; w[q] = 0.2500000000000000 * fact; }
  return 4;
}

// getGaussPoints3D (QuadratureUtil.cpp:268-318) for the element types on this path.
int gauss3d(int npElem, double xi[][3], double w[]) {
  if (npElem == 8) return gauss_hexa8(xi, w);
  if (npElem == 10) return gauss_tetra(4, xi, w);
  if (npElem == 4) return gauss_tetra(1, xi, w);
  throw std::runtime_error("getGaussPoints3D ... Element type not defined...");
}

// getGaussPointsQuad, ngp = 4 (QuadratureUtil.cpp:374-381)
int gauss_quad4(double xi[][2], double w[]) {
  const double s[4][2] = {{-GP2,-GP2},{GP2,-GP2},{GP2,GP2},{-GP2,GP2}};
  for (int q = 0; q < 4; ++q) { xi[q][0] = s[q][0]; xi[q][1] = s[q][1]; w[q] = 1.0; }
  return 4;
}

// ---------------------------------------------------------------------------
// Basis functions  (BasisFunctionsLagrange.cpp)
// ---------------------------------------------------------------------------

// Lagrange_BasisFuns1D, npElem = 2 (BasisFunctionsLagrange.cpp:132-155)
void lagrange1d_2(double xi, double N[2], double dN[2]) {
  N[0] = 0.5 * (1.0 - xi); N[1] = 0.5 * (1.0 + xi);
  dN[0] = -0.5; dN[1] = 0.5;
}

// LagrangeBasisFunsHexa, npElem = 8 (BasisFunctionsLagrange.cpp:1202-1279)
void basis_hexa8(double xi1, double xi2, double xi3, double N[8], double d1[8], double d2[8], double d3[8]) {
  double Nu[2], dNu[2], Nv[2], dNv[2], Nw[2], dNw[2];
  lagrange1d_2(xi1, Nu, dNu); lagrange1d_2(xi2, Nv, dNv); lagrange1d_2(xi3, Nw, dNw);
  const int iu[8] = {0,1,1,0,0,1,1,0}, iv[8] = {0,0,1,1,0,0,1,1}, iw[8] = {0,0,0,0,1,1,1,1};
  for (int a = 0; a < 8; ++a) {
    N[a]  = Nu[iu[a]] * Nv[iv[a]] * Nw[iw[a]];
    d1[a] = dNu[iu[a]] * Nv[iv[a]] * Nw[iw[a]];
    d2[a] = Nu[iu[a]] * dNv[iv[a]] * Nw[iw[a]];
    d3[a] = Nu[iu[a]] * Nv[iv[a]] * dNw[iw[a]];
  }
}

// LagrangeBasisFunsTetra, npElem = 4 and 10 (BasisFunctionsLagrange.cpp:870-956)
void basis_tetra(int npe, double xi1, double xi2, double xi3, double* N, double* d1, double* d2, double* d3) {
  const double xi4 = 1.0 - xi1 - xi2 - xi3;
  if (npe == 4) {
    N[0] = xi4; N[1] = xi1; N[2] = xi2; N[3] = xi3;
    if (d1) {
      d1[0] = -1.0; d1[1] = 1.0; d1[2] = 0.0; d1[3] = 0.0;
      d2[0] = -1.0; d2[1] = 0.0; d2[2] = 1.0; d2[3] = 0.0;
      d3[0] = -1.0; d3[1] = 0.0; d3[2] = 0.0; d3[3] = 1.0;
    }
    return;
  }
  N[0] = xi4 * (2 * xi4 - 1); N[1] = xi1 * (2 * xi1 - 1); N[2] = xi2 * (2 * xi2 - 1); N[3] = xi3 * (2 * xi3 - 1);
  N[4] = 4 * xi1 * xi4; N[5] = 4 * xi1 * xi2; N[6] = 4 * xi2 * xi4; N[7] = 4 * xi3 * xi4;
  N[8] = 4 * xi2 * xi3; N[9] = 4 * xi1 * xi3;
  d1[0] = -4 * xi4 + 1; d1[1] = 4 * xi1 - 1; d1[2] = 0; d1[3] = 0; d1[4] = 4 * (xi4 - xi1);
  d1[5] = 4 * xi2; d1[6] = -4 * xi2; d1[7] = -4 * xi3; d1[8] = 0; d1[9] = 4 * xi3;
  d2[0] = -4 * xi4 + 1; d2[1] = 0; d2[2] = 4 * xi2 - 1; d2[3] = 0; d2[4] = -4 * xi1;
  d2[5] = 4 * xi1; d2[6] = 4 * (xi4 - xi2); d2[7] = -4 * xi3; d2[8] = 4 * xi3; d2[9] = 0;
  d3[0] = -4 * xi4 + 1; d3[1] = 0; d3[2] = 0; d3[3] = 4 * xi3 - 1; d3[4] = -4 * xi1;
  d3[5] = 0; d3[6] = -4 * xi2; d3[7] = 4 * (xi4 - xi3); d3[8] = 4 * xi2; d3[9] = 4 * xi1;
}

// LagrangeBasisFunsQuad, npElem = 4 (BasisFunctionsLagrange.cpp:602-650)
void basis_quad4(double xi1, double xi2, double N[4], double d1[4], double d2[4]) {
  double Nu[2], dNu[2], Nv[2], dNv[2];
  lagrange1d_2(xi1, Nu, dNu); lagrange1d_2(xi2, Nv, dNv);
  N[0] = Nv[0] * Nu[0]; N[1] = Nv[0] * Nu[1]; N[3] = Nv[1] * Nu[0]; N[2] = Nv[1] * Nu[1];
  d1[0] = Nv[0] * dNu[0]; d1[1] = Nv[0] * dNu[1]; d1[3] = Nv[1] * dNu[0]; d1[2] = Nv[1] * dNu[1];
  d2[0] = dNv[0] * Nu[0]; d2[1] = dNv[0] * Nu[1]; d2[3] = dNv[1] * Nu[0]; d2[2] = dNv[1] * Nu[1];
}

// ---------------------------------------------------------------------------
// 3x3 determinant / inverse.  The reference holds B, F in dynamic-size Eigen
// MatrixXd (GeomDataLagrange.cpp:386,444-445; Elem_Solid_3D_Disp.cpp:273),
// whose determinant()/inverse() go through a partial-pivot LU.  This is the
// textbook Doolittle LU with row partial pivoting (first maximal |pivot|),
// unit-lower L, and column-wise triangular solves for the inverse.
// ---------------------------------------------------------------------------
void det_inv3(const double A[3][3], double& det, double Ainv[3][3]) {
  double lu[3][3];
  int perm[3] = {0, 1, 2};
  int nswap = 0;
  std::memcpy(lu, A, sizeof(lu));
  for (int k = 0; k < 3; ++k) {
    int piv = k;
    double big = std::fabs(lu[k][k]);
    for (int r = k + 1; r < 3; ++r)
      if (std::fabs(lu[r][k]) > big) { big = std::fabs(lu[r][k]); piv = r; }
    if (big != 0.0) {
      if (piv != k) {
        for (int c = 0; c < 3; ++c) std::swap(lu[k][c], lu[piv][c]);
        std::swap(perm[k], perm[piv]);
        ++nswap;
      }
      for (int r = k + 1; r < 3; ++r) lu[r][k] /= lu[k][k];
    }
    for (int r = k + 1; r < 3; ++r)
      for (int c = k + 1; c < 3; ++c) lu[r][c] -= lu[r][k] * lu[k][c];
  }
  det = ((nswap & 1) ? -1.0 : 1.0) * (lu[0][0] * lu[1][1] * lu[2][2]);
  for (int col = 0; col < 3; ++col) {
    double y[3];
    for (int i = 0; i < 3; ++i) {  // P*e_col, then forward substitution with unit L
      double s = (perm[i] == col) ? 1.0 : 0.0;
      for (int j = 0; j < i; ++j) s -= lu[i][j] * y[j];
      y[i] = s;
    }
    for (int i = 2; i >= 0; --i) {
      double s = y[i];
      for (int j = i + 1; j < 3; ++j) s -= lu[i][j] * y[j];
      y[i] = s / lu[i][i];
    }
    for (int i = 0; i < 3; ++i) Ainv[i][col] = y[i];
  }
}

// GeomDataLagrange::computeBasisFunctions3D (GeomDataLagrange.cpp:349-458):
// B(r,c) = sum_a dN_a/dxi_r * x_a[c];  Jac = det B;  dN/dx_c = sum_r dN/dxi_r * Binv(c,r).
void basis3d(int npe, const double xi[3], const double (*x)[3], double* N, double* dNdx, double* dNdy,
             double* dNdz, double& Jac) {
  double d1[10], d2[10], d3[10];
  if (npe == 8) basis_hexa8(xi[0], xi[1], xi[2], N, d1, d2, d3);
  else basis_tetra(npe, xi[0], xi[1], xi[2], N, d1, d2, d3);
  double B[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, Binv[3][3];
  for (int a = 0; a < npe; ++a) {
    const double xx = x[a][0], yy = x[a][1], zz = x[a][2];
    B[0][0] += (xx * d1[a]); B[1][0] += (xx * d2[a]); B[2][0] += (xx * d3[a]);
    B[0][1] += (yy * d1[a]); B[1][1] += (yy * d2[a]); B[2][1] += (yy * d3[a]);
    B[0][2] += (zz * d1[a]); B[1][2] += (zz * d2[a]); B[2][2] += (zz * d3[a]);
  }
  det_inv3(B, Jac, Binv);
  for (int a = 0; a < npe; ++a) {
    dNdx[a] = d1[a] * Binv[0][0] + d2[a] * Binv[0][1] + d3[a] * Binv[0][2];
    dNdy[a] = d1[a] * Binv[1][0] + d2[a] * Binv[1][1] + d3[a] * Binv[1][2];
    dNdz[a] = d1[a] * Binv[2][0] + d2[a] * Binv[2][1] + d3[a] * Binv[2][2];
  }
}

// ElementBase::computeDefGradCur / computeDefGradPrev (ElementBase.cpp:302-361)
void defgrad(int npe, const double (*u)[3], const double* dx, const double* dy, const double* dz, double F[3][3]) {
  F[0][0] = 1; F[0][1] = 0; F[0][2] = 0; F[1][0] = 0; F[1][1] = 1; F[1][2] = 0; F[2][0] = 0; F[2][1] = 0; F[2][2] = 1;
  for (int a = 0; a < npe; ++a) {
    const double ux = u[a][0], uy = u[a][1], uz = u[a][2];
    F[0][0] += (ux * dx[a]); F[0][1] += (ux * dy[a]); F[0][2] += (ux * dz[a]);
    F[1][0] += (uy * dx[a]); F[1][1] += (uy * dy[a]); F[1][2] += (uy * dz[a]);
    F[2][0] += (uz * dx[a]); F[2][1] += (uz * dy[a]); F[2][2] += (uz * dz[a]);
  }
}

double det3(const double F[3][3]) {
  double d, inv[3][3];
  det_inv3(F, d, inv);
  return d;
}

// ---------------------------------------------------------------------------
// Material (utilitiesmaterialHyperelastic.cpp)
// ---------------------------------------------------------------------------

// volumetricfunctions_getFirstDerivative (:45-93)
double dUdJ(int type, double K, double J) {
  switch (type) {
    case -1: return 0.0;
    case 1: return 0.1 * K * (J * J * J * J - 1.0 / (J * J * J * J * J * J));
    case 2: return 0.125 * K * (J * J - 1.0 / J / J) * (J + 1.0 / J / J / J);
    case 3: return K * (J - 1.0);
    case 4: return 0.5 * K * (J - 1.0 / J);
    case 5: return 0.5 * K * (J - 1.0 + std::log(J) / J);
    case 6: return K * std::log(J);
    case 7: return K * (1.0 - 1.0 / J);
    case 8: return K * std::log(J) / J;
    default: return 0.0;
  }
}

// volumetricfunctions_getSecondDerivative (:96-140)
double d2UdJ2(int type, double K, double J) {
  switch (type) {
    case -1: return 0.0;
    case 1: return 0.2 * K * (2.0 * J * J * J + 3.0 / (J * J * J * J * J * J * J));
    case 2: return 0.125 * K * (3.0 * J * J + 5.0 / (J * J * J * J * J * J));
    case 3: return K;
    case 4: return 0.5 * K * (1.0 + 1.0 / J / J);
    case 5: return 0.5 * K * (J * J + 1.0 - std::log(J)) / J / J;
    case 6: return K / std::log(J);
    case 7: return K / J / J;
    case 8: return K * (1.0 - std::log(J)) / J / J;
    default: return 0.0;
  }
}

// compute_constants_volumetric_functions (:144-173)
void vol_constants(bool finite, int Utype, double Jn, double BULK, double& Jhat, double& thetahat) {
  if (finite) {
    if (Utype == -1) { Jhat = 1.0; thetahat = 0.0; }
    else {
      const double d1 = dUdJ(Utype, BULK, Jn), d2 = d2UdJ2(Utype, BULK, Jn);
      Jhat = Jn - d1 / d2;
      thetahat = 1.0 / d2;
    }
  } else {
    Jhat = 1.0;
    thetahat = (Utype == -1) ? 0.0 : 1.0 / BULK;
  }
}

inline double kd(int i, int j) { return i == j ? 1.0 : 0.0; }

// computeStressAndTangent_NeoHooke (:186-295), reached from
// Matl_NeoHookean::computeStressAndTangent (Matl_NeoHookean.cpp:45-53).
void neohooke(bool mixed, int Utype, double Kinv, double mu, const double F[3][3], double pres,
              double stre[9], double Cmat[9][9], bool consistent = false) {
  const double r1d3 = 1.0 / 3.0, r2d3 = 2.0 / 3.0;
  const int map4[9][2] = {{0,0},{1,0},{2,0},{0,1},{1,1},{2,1},{0,2},{1,2},{2,2}};
  const double K = 1.0 / Kinv;
  double b[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) b[i][j] = F[i][0] * F[j][0] + F[i][1] * F[j][1] + F[i][2] * F[j][2];
  const double J = det3(F);
  const double Jm2d3 = std::pow(J, -r2d3);
  const double Jm5d3 = Jm2d3 / J;
  const double Ib = b[0][0] + b[1][1] + b[2][2];
  const double U1 = dUdJ(Utype, K, J), U2 = d2UdJ2(Utype, K, J);

  double fact1 = mu * Jm5d3;
  double fact2 = -fact1 * r1d3 * Ib;
  fact2 += mixed ? pres : U1;
  stre[0] = fact1 * b[0][0]; stre[3] = fact1 * b[0][1]; stre[6] = fact1 * b[0][2];
  stre[1] = fact1 * b[1][0]; stre[4] = fact1 * b[1][1]; stre[7] = fact1 * b[1][2];
  stre[2] = fact1 * b[2][0]; stre[5] = fact1 * b[2][1]; stre[8] = fact1 * b[2][2];
  stre[0] += fact2; stre[4] += fact2; stre[8] += fact2;

  // "with stress terms" deviatoric part; volumetric part as coded in the
  // reference (without its stress term) for the displacement element.
  fact1 = r1d3 * mu * Jm5d3 * Ib;
  fact2 = -r2d3 * mu * Jm5d3;
  const double fact3 = r2d3 * r1d3 * mu * Jm5d3 * Ib;
  const double fact4 = mu * Jm5d3;
  // consistent: the volumetric term "with stress terms" that the reference leaves commented out
  // (utilitiesmaterialHyperelastic.cpp:284-285), i.e. without the fact9 term.
  const double fact7 = J * U2 + U1, fact8 = -U1, fact9 = consistent ? 0.0 : -U1;
  for (int i1 = 0; i1 < 9; ++i1) {
    const int i = map4[i1][0], j = map4[i1][1];
    const double dij = kd(i, j);
    for (int j1 = 0; j1 < 9; ++j1) {
      const int k = map4[j1][0], l = map4[j1][1];
      const double dik = kd(i, k), dil = kd(i, l), djk = kd(j, k), djl = kd(j, l), dkl = kd(k, l);
      double c = 0.0;
      c += fact1 * (dil * djk);
      c += fact2 * (b[i][j] * dkl + dij * b[k][l]);
      c += fact3 * (dij * dkl);
      c += fact4 * (b[j][l] * dik);
      if (mixed) c += pres * (dij * dkl - djk * dil);
      else c += fact7 * (dij * dkl) + fact8 * (dil * djk) + fact9 * (djl * dik);
      Cmat[i1][j1] = c;
    }
  }
}

// ---------------------------------------------------------------------------
// Elements
// ---------------------------------------------------------------------------

// Elem_Solid_3D_Disp::calcStiffnessAndResidual (Elem_Solid_3D_Disp.cpp:196-374), STATIC:
// af = td(2) = 1, acceleration factor td(5) = 0, body force 0, rho0 irrelevant.
int elem_disp(int npe, const double (*X)[3], const double (*u)[3], double mu, double Kinv, int Utype,
              double* Ke, double* Re, bool consistent = false) {
  const int nsize = 3 * npe;
  std::fill(Ke, Ke + nsize * nsize, 0.0);
  std::fill(Re, Re + nsize, 0.0);
  double xi[10][3], w[10];
  const int nGP = gauss3d(npe, xi, w);
  double xcur[10][3];
  for (int a = 0; a < npe; ++a)
    for (int d = 0; d < 3; ++d) xcur[a][d] = X[a][d] + u[a][d];  // NodePosCur = X + dispCur
  const double af = 1.0;
  double N[10], dx[10], dy[10], dz[10], Jac, F[3][3], stre[9], C[9][9], Gc[3][9];
  for (int gp = 0; gp < nGP; ++gp) {
    basis3d(npe, xi[gp], X, N, dx, dy, dz, Jac);
    defgrad(npe, u, dx, dy, dz, F);
    const double detF = det3(F);
    if (detF < 0.0) return 1;  // runtime_error("Negative Jacobian in the element")
    basis3d(npe, xi[gp], xcur, N, dx, dy, dz, Jac);
    const double dvol = w[gp] * Jac;
    neohooke(false, Utype, Kinv, mu, F, 0.0, stre, C, consistent);
    for (int ii = 0; ii < npe; ++ii) {
      const double bb1 = dvol * dx[ii], bb2 = dvol * dy[ii], bb3 = dvol * dz[ii];
      for (int kk = 0; kk < 9; ++kk) {
        Gc[0][kk] = bb1 * C[0][kk] + bb2 * C[3][kk] + bb3 * C[6][kk];
        Gc[1][kk] = bb1 * C[1][kk] + bb2 * C[4][kk] + bb3 * C[7][kk];
        Gc[2][kk] = bb1 * C[2][kk] + bb2 * C[5][kk] + bb3 * C[8][kk];
      }
      const int TI = 3 * ii;
      const double s0 = bb1 * stre[0] + bb2 * stre[3] + bb3 * stre[6];
      const double s1 = bb1 * stre[1] + bb2 * stre[4] + bb3 * stre[7];
      const double s2 = bb1 * stre[2] + bb2 * stre[5] + bb3 * stre[8];
      Re[TI] += (0.0 - s0); Re[TI + 1] += (0.0 - s1); Re[TI + 2] += (0.0 - s2);
      for (int jj = 0; jj < npe; ++jj) {
        const double cc1 = dx[jj], cc2 = dy[jj], cc3 = dz[jj];
        const int TJ = 3 * jj;
        double* K0 = Ke + (TI + 0) * nsize + TJ;
        double* K1 = Ke + (TI + 1) * nsize + TJ;
        double* K2 = Ke + (TI + 2) * nsize + TJ;
        // mass term: fact = bb5*acceFact1*cc5*rho0 = 0 for STATIC
        K0[0] += 0.0; K1[1] += 0.0; K2[2] += 0.0;
        K0[0] += af * (Gc[0][0] * cc1 + Gc[0][3] * cc2 + Gc[0][6] * cc3);
        K0[1] += af * (Gc[0][1] * cc1 + Gc[0][4] * cc2 + Gc[0][7] * cc3);
        K0[2] += af * (Gc[0][2] * cc1 + Gc[0][5] * cc2 + Gc[0][8] * cc3);
        K1[0] += af * (Gc[1][0] * cc1 + Gc[1][3] * cc2 + Gc[1][6] * cc3);
        K1[1] += af * (Gc[1][1] * cc1 + Gc[1][4] * cc2 + Gc[1][7] * cc3);
        K1[2] += af * (Gc[1][2] * cc1 + Gc[1][5] * cc2 + Gc[1][8] * cc3);
        K2[0] += af * (Gc[2][0] * cc1 + Gc[2][3] * cc2 + Gc[2][6] * cc3);
        K2[1] += af * (Gc[2][1] * cc1 + Gc[2][4] * cc2 + Gc[2][7] * cc3);
        K2[2] += af * (Gc[2][2] * cc1 + Gc[2][5] * cc2 + Gc[2][8] * cc3);
      }
    }
  }
  return 0;
}

// ElementBase::calcStiffnessAndResidualMixed3D (ElementBase2.cpp:278-552) for Tet10/P1
// (Elem_Solid_3D_Mixed1.cpp:34-76: nGP = 4, nlbfP = 4, pressure nodes = first 4 nodes).
// ElementBase::calcStiffnessAndResidualMixed3D (ElementBase2.cpp:278-552) for npe displacement nodes
// and nP pressure functions: nP = 4 the P1 corner functions of a Tet10 (Elem_Solid_3D_Mixed1), nP = 1
// the element-constant pressure of Elem_Solid_3D_Mixed0 (Np = 1, pres = presDOF[0], :389-393).
static int elem_mixed(int npe, int nP, const double (*X)[3], const double (*u)[3], const double (*up)[3],
                      const double* presn, double mu, double Kinv, int Utype, double* Kuu, double* Kup, double* Kpu,
                      double* Kpp, double* Fu, double* Fp) {
  const int nsize = 3 * npe;
  std::fill(Kuu, Kuu + nsize * nsize, 0.0);
  std::fill(Kup, Kup + nsize * nP, 0.0);
  std::fill(Kpu, Kpu + nP * nsize, 0.0);
  std::fill(Kpp, Kpp + nP * nP, 0.0);
  std::fill(Fu, Fu + nsize, 0.0);
  std::fill(Fp, Fp + nP, 0.0);
  const double BULK = 1.0 / Kinv;
  const double af = 1.0;
  double xi[10][3], w[10];
  const int nGP = gauss3d(npe, xi, w);
  double xcur[10][3];
  for (int a = 0; a < npe; ++a)
    for (int d = 0; d < 3; ++d) xcur[a][d] = X[a][d] + u[a][d];
  double N[10], dx[10], dy[10], dz[10], Jac, F[3][3], Fn[3][3], stre[9], C[9][9], Gc[3][9], Np[4];
  for (int gp = 0; gp < nGP; ++gp) {
    basis3d(npe, xi[gp], X, N, dx, dy, dz, Jac);
    const double dvol0 = w[gp] * Jac;
    defgrad(npe, up, dx, dy, dz, Fn);
    const double detFn = det3(Fn);
    double Jhat, thetahat;
    vol_constants(true, Utype, detFn, BULK, Jhat, thetahat);
    defgrad(npe, u, dx, dy, dz, F);
    const double detF = det3(F);
    if (detF < 0.0) return 1;
    basis3d(npe, xi[gp], xcur, N, dx, dy, dz, Jac);
    const double dvol = w[gp] * Jac;
    if (nP == 1) Np[0] = 1.0;
    else basis_tetra(4, xi[gp][0], xi[gp][1], xi[gp][2], Np, nullptr, nullptr, nullptr);
    double pres = 0.0;
    for (int q = 0; q < nP; ++q) pres += (Np[q] * presn[q]);
    neohooke(true, Utype, Kinv, mu, F, pres, stre, C);
    for (int ii = 0; ii < npe; ++ii) {
      const double bb1 = dvol * dx[ii], bb2 = dvol * dy[ii], bb3 = dvol * dz[ii];
      for (int kk = 0; kk < 9; ++kk) {
        Gc[0][kk] = bb1 * C[0][kk] + bb2 * C[3][kk] + bb3 * C[6][kk];
        Gc[1][kk] = bb1 * C[1][kk] + bb2 * C[4][kk] + bb3 * C[7][kk];
        Gc[2][kk] = bb1 * C[2][kk] + bb2 * C[5][kk] + bb3 * C[8][kk];
      }
      const int TI = 3 * ii;
      const double s0 = bb1 * stre[0] + bb2 * stre[3] + bb3 * stre[6];
      const double s1 = bb1 * stre[1] + bb2 * stre[4] + bb3 * stre[7];
      const double s2 = bb1 * stre[2] + bb2 * stre[5] + bb3 * stre[8];
      Fu[TI] += (0.0 - s0); Fu[TI + 1] += (0.0 - s1); Fu[TI + 2] += (0.0 - s2);
      for (int jj = 0; jj < npe; ++jj) {
        const double cc1 = dx[jj], cc2 = dy[jj], cc3 = dz[jj];
        const int TJ = 3 * jj;
        double* K0 = Kuu + (TI + 0) * nsize + TJ;
        double* K1 = Kuu + (TI + 1) * nsize + TJ;
        double* K2 = Kuu + (TI + 2) * nsize + TJ;
        K0[0] += 0.0; K1[1] += 0.0; K2[2] += 0.0;
        K0[0] += af * (Gc[0][0] * cc1 + Gc[0][3] * cc2 + Gc[0][6] * cc3);
        K0[1] += af * (Gc[0][1] * cc1 + Gc[0][4] * cc2 + Gc[0][7] * cc3);
        K0[2] += af * (Gc[0][2] * cc1 + Gc[0][5] * cc2 + Gc[0][8] * cc3);
        K1[0] += af * (Gc[1][0] * cc1 + Gc[1][3] * cc2 + Gc[1][6] * cc3);
        K1[1] += af * (Gc[1][1] * cc1 + Gc[1][4] * cc2 + Gc[1][7] * cc3);
        K1[2] += af * (Gc[1][2] * cc1 + Gc[1][5] * cc2 + Gc[1][8] * cc3);
        K2[0] += af * (Gc[2][0] * cc1 + Gc[2][3] * cc2 + Gc[2][6] * cc3);
        K2[1] += af * (Gc[2][1] * cc1 + Gc[2][4] * cc2 + Gc[2][7] * cc3);
        K2[2] += af * (Gc[2][2] * cc1 + Gc[2][5] * cc2 + Gc[2][8] * cc3);
      }
      for (int jj = 0; jj < nP; ++jj) {
        Kup[(TI + 0) * nP + jj] += (bb1 * Np[jj]);
        Kup[(TI + 1) * nP + jj] += (bb2 * Np[jj]);
        Kup[(TI + 2) * nP + jj] += (bb3 * Np[jj]);
      }
    }
    const double resip = detF - Jhat - thetahat * pres;  // finite strain
    for (int ii = 0; ii < nP; ++ii) {
      const double bb4 = dvol * Np[ii] * af;
      double bb5 = dvol0 * Np[ii];
      Fp[ii] -= bb5 * resip;
      for (int jj = 0; jj < npe; ++jj) {
        const int TJ = 3 * jj;
        Kpu[ii * nsize + TJ] += bb4 * dx[jj];
        Kpu[ii * nsize + TJ + 1] += bb4 * dy[jj];
        Kpu[ii * nsize + TJ + 2] += bb4 * dz[jj];
      }
      bb5 *= thetahat * af;
      for (int jj = 0; jj < nP; ++jj) Kpp[ii * nP + jj] -= bb5 * Np[jj];
    }
  }
  return 0;
}

int elem_tet10p1(const double (*X)[3], const double (*u)[3], const double (*up)[3], const double* presn,
                 double mu, double Kinv, int Utype, double* Kuu, double* Kup, double* Kpu, double* Kpp,
                 double* Fu, double* Fp) {
  return elem_mixed(10, 4, X, u, up, presn, mu, Kinv, Utype, Kuu, Kup, Kpu, Kpp, Fu, Fp);
}

// Elem_Solid_3D_Mixed0::calcStiffnessAndResidual (Elem_Solid_3D_Mixed0.cpp:93-104): the Hex8 mixed
// element with one element pressure, condensed:  K = Kuu - Kup (Kpu / Kpp),  F = Fu - Kup (Fp / Kpp).
// The uncondensed Kuu, Fu and Kup, Kpu, Kpp, Fp are returned too (solveForPressure, :112-199:
// p += (Fp - Kpu du_e) / Kpp).
int elem_hex8p0(const double (*X)[3], const double (*u)[3], const double (*up)[3], double pe, double mu,
                double Kinv, int Utype, double* Kc, double* Fc, double* Kuu, double* Fu, double* Kup, double* Kpu,
                double* Kpp, double* Fp) {
  const int nsize = 24;
  if (elem_mixed(8, 1, X, u, up, &pe, mu, Kinv, Utype, Kuu, Kup, Kpu, Kpp, Fu, Fp)) return 1;
  double kpu_s[24];
  for (int j = 0; j < nsize; ++j) kpu_s[j] = Kpu[j] / Kpp[0];
  const double fp_s = Fp[0] / Kpp[0];
  for (int i = 0; i < nsize; ++i) {
    for (int j = 0; j < nsize; ++j) Kc[i * nsize + j] = Kuu[i * nsize + j] - Kup[i] * kpu_s[j];
    Fc[i] = Fu[i] - Kup[i] * fp_s;
  }
  return 0;
}

// ---------------------------------------------------------------------------
// Problem
// ---------------------------------------------------------------------------

// findUnique (util.h:41-45)
void find_unique(std::vector<int>& v) {
  std::sort(v.begin(), v.end());
  v.erase(std::unique(v.begin(), v.end()), v.end());
}

struct DirPatch { std::vector<int> nodes; int dof; std::vector<double> value; };
struct TracPatch { std::vector<int> faces; double v0, v1; int npf = 4; };  // npf: 4 (Quad4) or 6 (Tria6)

}  // namespace

struct orc_problem {
  int etype, npe, nNode, nElem;
  std::vector<double> X;       // [nNode][3]
  std::vector<int> conn;       // [nElem][npe]
  double mu, Kinv; int Utype;
  bool mixed;
  bool p0 = false;          // Hex8 Q1/P0 (elem_type 9): pressure per element, condensed out of the system
  std::vector<double> qKpu, qKpp, qFp;  // Q1/P0: the last assembly's Kpu [nElem][24], Kpp, Fp per element
  bool consistent = false;  // tangent mode (see neohooke)
  std::vector<DirPatch> dir;
  std::vector<TracPatch> trac;
  // numbering
  std::vector<int> NodeDofArray;   // [nNode][3]
  std::vector<int> assyForSoln;    // [dispDOF]
  std::vector<int> ID_pres;        // [nNode]
  std::vector<int> assyForSolnPres;
  std::vector<int> dofs_specified;  // sorted unique
  int dispDOF = 0, presDOF = 0, ntot = 0;
  std::vector<int> rowptr, colidx;
  // state
  std::vector<double> disp, dispPrev, pres, presPrev, dispApplied, dispTarget;
  bool prepared = false;
};

namespace {

// forAssyVec for one element (femSolidmechanics2.cpp:248-280)
void elem_forassy(const orc_problem* p, int e, int* fav) {
  for (int a = 0; a < p->npe; ++a) {
    const int nd = p->conn[(size_t)e * p->npe + a];
    for (int d = 0; d < 3; ++d) fav[3 * a + d] = p->NodeDofArray[3 * nd + d];
  }
}

// MatSetValue(..., ADD_VALUES) into an AIJ row whose columns are sorted.
inline void csr_add(const orc_problem* p, double* vals, int r, int c, double v) {
  const int* b = p->colidx.data() + p->rowptr[r];
  const int* e = p->colidx.data() + p->rowptr[r + 1];
  const int* it = std::lower_bound(b, e, c);
  if (it == e || *it != c) throw std::runtime_error("csr_add: entry outside pattern");
  vals[it - p->colidx.data()] += v;
}

// Quad4 face dead load (femSolidmechanics1.cpp:1026-1082; LagrangeBasisFunsFace
// BasisFunctionsLagrange.cpp:1688-1732): n = (dx/deta x dx/dxi)/|.|, tangent1 = dx/dxi.
// LagrangeBasisFunsTria, npElem = 6 (BasisFunctionsLagrange.cpp:364-418): xi3 = 1 - xi1 - xi2,
// nodes 0..2 at the corners, 3 on (0,1), 4 on (1,2), 5 on (2,0)
void basis_tria6(double xi1, double xi2, double N[6], double d1[6], double d2[6]) {
  const double xi3 = 1.0 - xi1 - xi2;
  N[0] = xi3 * (2.0 * xi3 - 1.0); N[1] = xi1 * (2.0 * xi1 - 1.0); N[2] = xi2 * (2.0 * xi2 - 1.0);
  N[3] = 4.0 * xi1 * xi3; N[4] = 4.0 * xi1 * xi2; N[5] = 4.0 * xi2 * xi3;
  d1[0] = -4.0 * xi3 + 1.0; d1[1] = 4.0 * xi1 - 1.0; d1[2] = 0.0;
  d1[3] = 4.0 * (xi3 - xi1); d1[4] = 4.0 * xi2; d1[5] = -4.0 * xi2;
  d2[0] = -4.0 * xi3 + 1.0; d2[1] = 0.0; d2[2] = 4.0 * xi2 - 1.0;
  d2[3] = -4.0 * xi1; d2[4] = 4.0 * xi1; d2[5] = 4.0 * (xi3 - xi2);
}

// getGaussPoints2D for npElem = 6: the 3-point triangle rule (QuadratureUtil.cpp:223-236, 562-570)
int gauss_tria3(double xi[][2], double w[]) {
  const double a1 = 1.0 / 3.0, fact = 0.5;
  xi[0][0] = 1.0 / 6.0; xi[0][1] = 1.0 / 6.0; w[0] = fact * a1;
  xi[1][0] = 4.0 / 6.0; xi[1][1] = 1.0 / 6.0; w[1] = fact * a1;
  xi[2][0] = 1.0 / 6.0; xi[2][1] = 4.0 / 6.0; w[2] = fact * a1;
  return 3;
}

// addExternalForces, 3D branch (femSolidmechanics1.cpp:1026-1082) with LagrangeBasisFunsFace
// (BasisFunctionsLagrange.cpp:1688-1732) for Quad4 and Tria6 faces.
void external_forces(const orc_problem* p, double lam, std::vector<double>& Fext) {
  Fext.assign((size_t)p->nNode * 3, 0.0);
  double gq[4][2], wq[4], gt[3][2], wt[3];
  gauss_quad4(gq, wq);
  gauss_tria3(gt, wt);
  for (const auto& tp : p->trac) {
    const double sv0 = tp.v0 * lam, sv1 = tp.v1 * lam;
    const int npf = tp.npf;
    const int nf = (int)tp.faces.size() / npf;
    const int ngp = npf == 4 ? 4 : 3;
    for (int f = 0; f < nf; ++f) {
      const int* nn = &tp.faces[(size_t)npf * f];
      double xN[6], yN[6], zN[6];
      for (int a = 0; a < npf; ++a) { xN[a] = p->X[3 * nn[a]]; yN[a] = p->X[3 * nn[a] + 1]; zN[a] = p->X[3 * nn[a] + 2]; }
      double Fl[18] = {0};
      for (int gp = 0; gp < ngp; ++gp) {
        double N[6], d1[6], d2[6], w;
        if (npf == 4) {
          basis_quad4(gq[gp][0], gq[gp][1], N, d1, d2);
          w = wq[gp];
        } else {
          basis_tria6(gt[gp][0], gt[gp][1], N, d1, d2);
          w = wt[gp];
        }
        double du[3] = {0, 0, 0}, dv[3] = {0, 0, 0};
        for (int a = 0; a < npf; ++a) {
          du[0] += (xN[a] * d1[a]); du[1] += (yN[a] * d1[a]); du[2] += (zN[a] * d1[a]);
          dv[0] += (xN[a] * d2[a]); dv[1] += (yN[a] * d2[a]); dv[2] += (zN[a] * d2[a]);
        }
        double n1[3] = {dv[1] * du[2] - dv[2] * du[1], dv[2] * du[0] - dv[0] * du[2], dv[0] * du[1] - dv[1] * du[0]};
        const double Jac = std::sqrt(n1[0] * n1[0] + n1[1] * n1[1] + n1[2] * n1[2]);
        n1[0] /= Jac; n1[1] /= Jac; n1[2] /= Jac;
        const double dvol = w * Jac;
        double tr[3];
        for (int d = 0; d < 3; ++d) tr[d] = (sv0 * n1[d] + sv1 * du[d]) * dvol;
        for (int a = 0; a < npf; ++a)
          for (int d = 0; d < 3; ++d) Fl[3 * a + d] += N[a] * tr[d];
      }
      for (int a = 0; a < npf; ++a)
        for (int d = 0; d < 3; ++d) Fext[3 * nn[a] + d] += Fl[3 * a + d];
    }
  }
}

double norm2(const double* v, int n) {
  double s = 0.0;
  for (int i = 0; i < n; ++i) s += v[i] * v[i];
  return std::sqrt(s);
}

// Assembly loop of femSolidmechanics::calcStiffnessAndResidual (femSolidmechanics2.cpp:1047-1120)
// over elements [first, first+count); lift with applyDirichletBCs (ElementBase.cpp:567-594) /
// applyDirichletBCs2field (ElementBase.cpp:717-782); scatter as
// SolverPetsc::assembleMatrixAndVector (SolverPetsc.cpp:357-378) / ...2field (:434-502).
// Per-element scratch of the assembly loop.
struct AsmWork {
  std::vector<double> Ke, Re, Kup, Kpu, Kpp, Fp, Kuu, Fu;
  explicit AsmWork(int nsize)
      : Ke(nsize * nsize), Re(nsize), Kup(nsize * 4), Kpu(4 * nsize), Kpp(16), Fp(4), Kuu(nsize * nsize), Fu(nsize) {}
};

// One element of calcStiffnessAndResidual's loop (femSolidmechanics2.cpp:1061-1111): element
// matrices, the first-iteration lift, and the scatter into the CSR values and rhs.
static int assemble_elem(orc_problem* p, int e, bool firstIter, double* vals, double* rhs, AsmWork& W) {
  const int npe = p->npe, nsize = 3 * npe;
  std::vector<double>&Ke = W.Ke, &Re = W.Re, &Kup = W.Kup, &Kpu = W.Kpu, &Kpp = W.Kpp, &Fp = W.Fp;
  int fav[30], favP[4];
  double Xe[10][3], ue[10][3], upe[10][3], pe[4];
  const int* cn = &p->conn[(size_t)e * npe];
  for (int a = 0; a < npe; ++a)
    for (int d = 0; d < 3; ++d) {
      Xe[a][d] = p->X[3 * cn[a] + d];
      ue[a][d] = p->disp[3 * cn[a] + d];
      upe[a][d] = p->dispPrev[3 * cn[a] + d];
    }
  elem_forassy(p, e, fav);
  if (!p->mixed) {
    if (p->p0) {  // condensed Q1/P0 element; Kpu, Kpp, Fp kept for solveForPressure
      if (elem_hex8p0(Xe, ue, upe, p->pres[e], p->mu, p->Kinv, p->Utype, Ke.data(), Re.data(), W.Kuu.data(),
                      W.Fu.data(), Kup.data(), Kpu.data(), Kpp.data(), Fp.data()))
        return 1;
      std::copy(Kpu.begin(), Kpu.begin() + 24, p->qKpu.begin() + (size_t)e * 24);
      p->qKpp[e] = Kpp[0];
      p->qFp[e] = Fp[0];
    } else if (elem_disp(npe, Xe, ue, p->mu, p->Kinv, p->Utype, Ke.data(), Re.data(), p->consistent)) {
      return 1;
    }
    if (firstIter) {
      for (int ii = 0; ii < nsize; ++ii) {
        if (fav[ii] != -1) continue;
        const double fact = p->dispApplied[3 * cn[ii / 3] + ii % 3];
        for (int jj = 0; jj < nsize; ++jj)
          if (fav[jj] != -1) Re[jj] -= Ke[jj * nsize + ii] * fact;
      }
    }
    for (int ii = 0; ii < nsize; ++ii) {
      const int r = fav[ii];
      if (r < 0) continue;  // VEC_IGNORE_NEGATIVE_INDICES (SolverPetsc.cpp:52)
      rhs[r] += Re[ii];
      for (int jj = 0; jj < nsize; ++jj) {
        const int c = fav[jj];
        if (c < 0) continue;
        csr_add(p, vals, r, c, Ke[ii * nsize + jj]);
      }
    }
  } else {
    for (int q = 0; q < 4; ++q) { pe[q] = p->pres[cn[q]]; favP[q] = p->ID_pres[cn[q]]; }
    if (elem_tet10p1(Xe, ue, upe, pe, p->mu, p->Kinv, p->Utype, Ke.data(), Kup.data(), Kpu.data(), Kpp.data(),
                     Re.data(), Fp.data()))
      return 1;
    if (firstIter) {
      for (int ii = 0; ii < nsize; ++ii) {
        if (fav[ii] != -1) continue;
        const double fact = p->dispApplied[3 * cn[ii / 3] + ii % 3];
        for (int jj = 0; jj < nsize; ++jj)
          if (fav[jj] != -1) Re[jj] -= Ke[jj * nsize + ii] * fact;
        for (int jj = 0; jj < 4; ++jj)
          if (favP[jj] != -1) Fp[jj] -= Kpu[jj * nsize + ii] * fact;
      }
      // pressure dofs are never prescribed on this path (every corner node is a pressure dof)
    }
    const int off = p->dispDOF;
    for (int ii = 0; ii < nsize; ++ii) {
      const int r = fav[ii];
      if (r == -1) continue;
      rhs[r] += Re[ii];
      for (int jj = 0; jj < nsize; ++jj) {
        const int c = fav[jj];
        if (c != -1) csr_add(p, vals, r, c, Ke[ii * nsize + jj]);
      }
      for (int jj = 0; jj < 4; ++jj)
        if (favP[jj] != -1) csr_add(p, vals, r, favP[jj] + off, Kup[ii * 4 + jj]);
    }
    for (int ii = 0; ii < 4; ++ii) {
      const int r0 = favP[ii];
      if (r0 == -1) continue;
      const int r = r0 + off;
      rhs[r] += Fp[ii];
      for (int jj = 0; jj < nsize; ++jj) {
        const int c = fav[jj];
        if (c != -1) csr_add(p, vals, r, c, Kpu[ii * nsize + jj]);
      }
      for (int jj = 0; jj < 4; ++jj)
        if (favP[jj] != -1) csr_add(p, vals, r, favP[jj] + off, Kpp[ii * 4 + jj]);
    }
  }
  return 0;
}

int assemble_range(orc_problem* p, int first, int count, bool firstIter, double* vals, double* rhs) {
  AsmWork W(3 * p->npe);
  for (int e = first; e < first + count; ++e)
    if (assemble_elem(p, e, firstIter, vals, rhs, W)) return 1;
  return 0;
}

// ---------------------------------------------------------------------------
// CPU Krylov solvers and a banded direct solver (stand-ins for PETSc KSP,
// SolverPetsc.cpp:250-330; the reference's KSP iterates are not reproducible
// without PETSc, solution parity is checked through the residual).
// ---------------------------------------------------------------------------

void spmv(const orc_problem* p, const double* vals, const double* x, double* y) {
  for (int r = 0; r < p->ntot; ++r) {
    double s = 0.0;
    for (int k = p->rowptr[r]; k < p->rowptr[r + 1]; ++k) s += vals[k] * x[p->colidx[k]];
    y[r] = s;
  }
}

double diag_of(const orc_problem* p, const double* vals, int r) {
  for (int k = p->rowptr[r]; k < p->rowptr[r + 1]; ++k)
    if (p->colidx[k] == r) return vals[k];
  return 0.0;
}

double entry_of(const orc_problem* p, const double* vals, int r, int c) {
  const int* b = p->colidx.data() + p->rowptr[r];
  const int* e = p->colidx.data() + p->rowptr[r + 1];
  const int* it = std::lower_bound(b, e, c);
  return (it != e && *it == c) ? vals[it - p->colidx.data()] : 0.0;
}

// Preconditioner: 0 none, 1 Jacobi, 2 node 3x3 block-Jacobi (free dofs of a node form a block;
// pressure rows are 1x1).
struct Precond {
  int kind = 0;
  std::vector<double> dinv;            // Jacobi
  std::vector<int> blk_start, blk_size;  // block-Jacobi blocks (contiguous rows)
  std::vector<double> blk_inv;         // 9 per block
  void setup(const orc_problem* p, const double* vals, int k) {
    kind = k;
    const int n = p->ntot;
    if (k == 1) {
      dinv.resize(n);
      for (int r = 0; r < n; ++r) dinv[r] = 1.0 / diag_of(p, vals, r);
    } else if (k == 2) {
      blk_start.clear(); blk_size.clear(); blk_inv.clear();
      for (int nd = 0; nd < p->nNode; ++nd) {
        int f[3], m = 0;
        for (int d = 0; d < 3; ++d) if (p->NodeDofArray[3 * nd + d] >= 0) f[m++] = p->NodeDofArray[3 * nd + d];
        if (m == 0) continue;
        blk_start.push_back(f[0]); blk_size.push_back(m);
        double A[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}, Ai[3][3], det;
        for (int i = 0; i < m; ++i)
          for (int j = 0; j < m; ++j) A[i][j] = entry_of(p, vals, f[i], f[j]);
        det_inv3(A, det, Ai);
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) blk_inv.push_back(Ai[i][j]);
      }
      for (int r = p->dispDOF; r < n; ++r) {
        blk_start.push_back(r); blk_size.push_back(1);
        const double dv = diag_of(p, vals, r);
        const double v[9] = {1.0 / dv, 0, 0, 0, 1, 0, 0, 0, 1};
        blk_inv.insert(blk_inv.end(), v, v + 9);
      }
    }
  }
  void apply(const orc_problem* p, const double* r, double* z) const {
    const int n = p->ntot;
    if (kind == 0) { std::memcpy(z, r, sizeof(double) * n); return; }
    if (kind == 1) { for (int i = 0; i < n; ++i) z[i] = dinv[i] * r[i]; return; }
    for (size_t b = 0; b < blk_start.size(); ++b) {
      const int s = blk_start[b], m = blk_size[b];
      const double* Ai = &blk_inv[9 * b];
      for (int i = 0; i < m; ++i) {
        double acc = 0.0;
        for (int j = 0; j < m; ++j) acc += Ai[3 * i + j] * r[s + j];
        z[s + i] = acc;
      }
    }
  }
};

int pcg(const orc_problem* p, const double* vals, const double* b, double* x, int pk, double rtol, int maxit,
        double* relres) {
  const int n = p->ntot;
  Precond M; M.setup(p, vals, pk);
  std::vector<double> r(b, b + n), z(n), q(n), d(n);
  std::fill(x, x + n, 0.0);
  const double bn = norm2(b, n);
  if (bn == 0.0) { *relres = 0.0; return 0; }
  M.apply(p, r.data(), z.data());
  d = z;
  double rz = 0.0;
  for (int i = 0; i < n; ++i) rz += r[i] * z[i];
  int it = 0;
  double rn = bn;
  while (it < maxit) {
    spmv(p, vals, d.data(), q.data());
    double dq = 0.0;
    for (int i = 0; i < n; ++i) dq += d[i] * q[i];
    const double alpha = rz / dq;
    for (int i = 0; i < n; ++i) { x[i] += alpha * d[i]; r[i] -= alpha * q[i]; }
    ++it;
    rn = norm2(r.data(), n);
    if (rn / bn < rtol) break;
    M.apply(p, r.data(), z.data());
    double rz2 = 0.0;
    for (int i = 0; i < n; ++i) rz2 += r[i] * z[i];
    const double beta = rz2 / rz;
    rz = rz2;
    for (int i = 0; i < n; ++i) d[i] = z[i] + beta * d[i];
  }
  *relres = rn / bn;
  return it;
}

// Right-preconditioned restarted GMRES(m) with modified Gram-Schmidt and Givens rotations.
int gmres(const orc_problem* p, const double* vals, const double* b, double* x, int pk, int m, double rtol,
          int maxit, double* relres) {
  const int n = p->ntot;
  Precond M; M.setup(p, vals, pk);
  std::fill(x, x + n, 0.0);
  const double bn = norm2(b, n);
  if (bn == 0.0) { *relres = 0.0; return 0; }
  std::vector<double> V((size_t)(m + 1) * n), Z((size_t)m * n), H((size_t)(m + 1) * m), cs(m), sn(m), g(m + 1), w(n), r(n);
  int it = 0;
  double rn = bn;
  while (it < maxit) {
    spmv(p, vals, x, w.data());
    for (int i = 0; i < n; ++i) r[i] = b[i] - w[i];
    double beta = norm2(r.data(), n);
    rn = beta;
    if (beta / bn < rtol) break;
    for (int i = 0; i < n; ++i) V[i] = r[i] / beta;
    std::fill(g.begin(), g.end(), 0.0); g[0] = beta;
    int j = 0;
    for (; j < m && it < maxit; ++j, ++it) {
      M.apply(p, &V[(size_t)j * n], &Z[(size_t)j * n]);
      spmv(p, vals, &Z[(size_t)j * n], w.data());
      for (int i = 0; i <= j; ++i) {
        double h = 0.0;
        for (int k = 0; k < n; ++k) h += w[k] * V[(size_t)i * n + k];
        H[i * m + j] = h;
        for (int k = 0; k < n; ++k) w[k] -= h * V[(size_t)i * n + k];
      }
      const double hn = norm2(w.data(), n);
      H[(j + 1) * m + j] = hn;
      for (int k = 0; k < n; ++k) V[(size_t)(j + 1) * n + k] = (hn > 0) ? w[k] / hn : 0.0;
      for (int i = 0; i < j; ++i) {
        const double t = cs[i] * H[i * m + j] + sn[i] * H[(i + 1) * m + j];
        H[(i + 1) * m + j] = -sn[i] * H[i * m + j] + cs[i] * H[(i + 1) * m + j];
        H[i * m + j] = t;
      }
      const double a = H[j * m + j], bb = H[(j + 1) * m + j];
      const double den = std::sqrt(a * a + bb * bb);
      cs[j] = a / den; sn[j] = bb / den;
      H[j * m + j] = den; H[(j + 1) * m + j] = 0.0;
      g[j + 1] = -sn[j] * g[j]; g[j] = cs[j] * g[j];
      rn = std::fabs(g[j + 1]);
      if (rn / bn < rtol) { ++j; ++it; break; }
    }
    // back substitution y = H^{-1} g, x += Z y
    std::vector<double> y(j);
    for (int i = j - 1; i >= 0; --i) {
      double s = g[i];
      for (int k = i + 1; k < j; ++k) s -= H[i * m + k] * y[k];
      y[i] = s / H[i * m + i];
    }
    for (int i = 0; i < j; ++i)
      for (int k = 0; k < n; ++k) x[k] += y[i] * Z[(size_t)i * n + k];
    if (rn / bn < rtol) break;
  }
  spmv(p, vals, x, w.data());
  for (int i = 0; i < n; ++i) r[i] = b[i] - w[i];
  *relres = norm2(r.data(), n) / bn;
  return it;
}

// Banded LDL^T (no pivoting) for the symmetric positive-definite CSR system.
int direct_solve(const orc_problem* p, const double* vals, const double* b, double* x) {
  const int n = p->ntot;
  int bw = 0;
  for (int r = 0; r < n; ++r)
    for (int k = p->rowptr[r]; k < p->rowptr[r + 1]; ++k) bw = std::max(bw, r - p->colidx[k]);
  // lower band storage L[r][bw + c - r], c in [r-bw, r]
  const int W = bw + 1;
  std::vector<double> L((size_t)n * W, 0.0), D(n);
  for (int r = 0; r < n; ++r)
    for (int k = p->rowptr[r]; k < p->rowptr[r + 1]; ++k) {
      const int c = p->colidx[k];
      if (c <= r) L[(size_t)r * W + bw + c - r] = vals[k];
    }
  for (int j = 0; j < n; ++j) {
    // D_j = A_jj - sum_k L_jk^2 D_k
    double dj = L[(size_t)j * W + bw];
    const int k0 = std::max(0, j - bw);
    for (int k = k0; k < j; ++k) { const double l = L[(size_t)j * W + bw + k - j]; dj -= l * l * D[k]; }
    if (dj == 0.0 || !std::isfinite(dj)) return 1;  // LDL^T without pivoting: negative pivots are fine
    D[j] = dj;
    const int iend = std::min(n - 1, j + bw);
    for (int i = j + 1; i <= iend; ++i) {
      double s = L[(size_t)i * W + bw + j - i];
      const int kk0 = std::max(k0, i - bw);
      for (int k = kk0; k < j; ++k) s -= L[(size_t)i * W + bw + k - i] * L[(size_t)j * W + bw + k - j] * D[k];
      L[(size_t)i * W + bw + j - i] = s / dj;
    }
  }
  std::vector<double> y(b, b + n);
  for (int i = 0; i < n; ++i) {
    const int k0 = std::max(0, i - bw);
    for (int k = k0; k < i; ++k) y[i] -= L[(size_t)i * W + bw + k - i] * y[k];
  }
  for (int i = 0; i < n; ++i) y[i] /= D[i];
  for (int i = n - 1; i >= 0; --i) {
    const int kend = std::min(n - 1, i + bw);
    for (int k = i + 1; k <= kend; ++k) y[i] -= L[(size_t)k * W + bw + i - k] * y[k];
  }
  std::memcpy(x, y.data(), sizeof(double) * n);
  return 0;
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

int orc_gauss3d(int npElem, double* xi, double* w) {
  double g[27][3], ww[27];
  const int n = gauss3d(npElem, g, ww);
  for (int q = 0; q < n; ++q) { xi[3 * q] = g[q][0]; xi[3 * q + 1] = g[q][1]; xi[3 * q + 2] = g[q][2]; w[q] = ww[q]; }
  return n;
}

int orc_gauss_quad4(double* xi, double* w) {
  double g[4][2];
  gauss_quad4(g, w);
  for (int q = 0; q < 4; ++q) { xi[2 * q] = g[q][0]; xi[2 * q + 1] = g[q][1]; }
  return 4;
}

void orc_basis_hex8(const double xi[3], double* N, double* dN) {
  basis_hexa8(xi[0], xi[1], xi[2], N, dN, dN + 8, dN + 16);
}

void orc_basis_tet(int npe, const double xi[3], double* N, double* dN) {
  basis_tetra(npe, xi[0], xi[1], xi[2], N, dN, dN + npe, dN + 2 * npe);
}

void orc_basis_quad4(const double xi[2], double* N, double* dN) { basis_quad4(xi[0], xi[1], N, dN, dN + 4); }

void orc_det_inv3(const double A[9], double* det, double Ainv[9]) {
  double a[3][3], ai[3][3];
  std::memcpy(a, A, sizeof(a));
  det_inv3(a, *det, ai);
  std::memcpy(Ainv, ai, sizeof(ai));
}

double orc_dUdJ(int t, double K, double J) { return dUdJ(t, K, J); }
double orc_d2UdJ2(int t, double K, double J) { return d2UdJ2(t, K, J); }
void orc_vol_constants(int finite, int Utype, double Jn, double K, double* Jhat, double* th) {
  vol_constants(finite != 0, Utype, Jn, K, *Jhat, *th);
}

int orc_neohooke(int mixed, int Utype, double Kinv, double mu, const double F[9], double pres, double* stre,
                 double* Cmat) {
  double f[3][3], C[9][9];
  std::memcpy(f, F, sizeof(f));
  neohooke(mixed != 0, Utype, Kinv, mu, f, pres, stre, C);
  std::memcpy(Cmat, C, sizeof(C));
  return 0;
}

int orc_elem_disp(int npe, const double* X, const double* u, double mu, double Kinv, int Utype, double* Ke,
                  double* Re) {
  return elem_disp(npe, (const double(*)[3])X, (const double(*)[3])u, mu, Kinv, Utype, Ke, Re);
}

int orc_elem_disp_consistent(int npe, const double* X, const double* u, double mu, double Kinv, int Utype, double* Ke,
                             double* Re) {
  return elem_disp(npe, (const double(*)[3])X, (const double(*)[3])u, mu, Kinv, Utype, Ke, Re, true);
}

void orc_set_tangent(orc_problem* p, int consistent) { p->consistent = consistent != 0; }

int orc_elem_hex8p0(const double* X, const double* u, const double* uprev, double pe, double mu, double Kinv,
                    int Utype, double* Kc, double* Fc, double* Kuu, double* Fu, double* Kup, double* Kpu, double* Kpp,
                    double* Fp) {
  return elem_hex8p0((const double(*)[3])X, (const double(*)[3])u, (const double(*)[3])uprev, pe, mu, Kinv, Utype,
                     Kc, Fc, Kuu, Fu, Kup, Kpu, Kpp, Fp);
}

int orc_elem_tet10p1(const double* X, const double* u, const double* uprev, const double* pres, double mu,
                     double Kinv, int Utype, double* Kuu, double* Kup, double* Kpu, double* Kpp, double* Fu,
                     double* Fp) {
  return elem_tet10p1((const double(*)[3])X, (const double(*)[3])u, (const double(*)[3])uprev, pres, mu, Kinv,
                      Utype, Kuu, Kup, Kpu, Kpp, Fu, Fp);
}

orc_problem* orc_problem_create(int elem_type, int nNode, int nElem, const double* X, const int32_t* conn,
                                double mu, double Kinv, int Utype) {
  auto* p = new orc_problem();
  p->p0 = (elem_type == 9);
  p->etype = p->p0 ? 8 : elem_type;  // Q1/P0 shares the Hex8 numbering and pattern
  // 11: Tet10 displacement element (Elem_Solid_3D_Disp with npElem = 10, Elem_Solid_3D_Disp.cpp:47-51)
  p->npe = (elem_type == 10 || elem_type == 11) ? 10 : 8;
  p->mixed = (elem_type == 10);
  if (p->p0) {
    p->qKpu.assign((size_t)nElem * 24, 0.0);
    p->qKpp.assign(nElem, 1.0);
    p->qFp.assign(nElem, 0.0);
  }
  p->nNode = nNode; p->nElem = nElem;
  p->X.assign(X, X + (size_t)nNode * 3);
  p->conn.assign(conn, conn + (size_t)nElem * p->npe);
  p->mu = mu; p->Kinv = Kinv; p->Utype = Utype;
  p->disp.assign((size_t)nNode * 3, 0.0);
  p->dispPrev = p->disp; p->dispApplied = p->disp; p->dispTarget = p->disp;
  // SolnData.pres sized max(nElem, nNode) and indexed by node id (femSolidmechanics1.cpp:379)
  p->pres.assign(std::max(nElem, nNode), 0.0);
  p->presPrev = p->pres;
  return p;
}

void orc_problem_destroy(orc_problem* p) { delete p; }

int orc_add_dirichlet(orc_problem* p, int n, const int32_t* nodes, int dof, const double* value) {
  DirPatch dp;
  dp.nodes.assign(nodes, nodes + n);
  dp.dof = dof;
  if (value) dp.value.assign(value, value + n); else dp.value.assign(n, 0.0);
  p->dir.push_back(dp);
  return 0;
}

int orc_add_traction(orc_problem* p, int nf, const int32_t* faces, double v0, double v1) {
  TracPatch tp;
  tp.faces.assign(faces, faces + 4 * nf);
  tp.v0 = v0; tp.v1 = v1;
  p->trac.push_back(tp);
  return 0;
}

// addExternalForces alone: the external force vector at load factor lam, [nNode * 3]
void orc_external_force(const orc_problem* p, double lam, double* F) {
  std::vector<double> f;
  external_forces(p, lam, f);
  std::copy(f.begin(), f.end(), F);
}

int orc_add_traction_tri6(orc_problem* p, int nf, const int32_t* faces, double v0, double v1) {
  TracPatch tp;
  tp.faces.assign(faces, faces + 6 * nf);
  tp.v0 = v0; tp.v1 = v1; tp.npf = 6;
  p->trac.push_back(tp);
  return 0;
}

// prepareMatrixPattern (femSolidmechanics2.cpp:121-478) with
// setSpecifiedDOFs_Displacement (femSolidmechanics1.cpp:1479-1543) and
// prepareDataForMixedElements (femSolidmechanics1.cpp:274-387).
int orc_prepare_pattern(orc_problem* p) {
  const int nNode = p->nNode, nElem = p->nElem, npe = p->npe;
  std::vector<char> type((size_t)nNode * 3, 0);
  p->dofs_specified.clear();
  for (const auto& dp : p->dir) {
    for (size_t i = 0; i < dp.nodes.size(); ++i) {
      const int nn = dp.nodes[i];
      if (dp.dof >= 0) { type[3 * nn + dp.dof] = 1; p->dofs_specified.push_back(3 * nn + dp.dof); }
      else for (int d = 0; d < 3; ++d) { type[3 * nn + d] = 1; p->dofs_specified.push_back(3 * nn + d); }
    }
  }
  find_unique(p->dofs_specified);
  p->NodeDofArray.assign((size_t)nNode * 3, -1);
  p->dispDOF = 0;
  for (int i = 0; i < nNode; ++i)
    for (int d = 0; d < 3; ++d)
      if (!type[3 * i + d]) p->NodeDofArray[3 * i + d] = p->dispDOF++;
  p->assyForSoln.clear();
  for (int i = 0; i < nNode; ++i)
    for (int d = 0; d < 3; ++d)
      if (p->NodeDofArray[3 * i + d] != -1) p->assyForSoln.push_back(3 * i + d);
  p->ID_pres.assign(nNode, -1);
  p->presDOF = 0;
  p->assyForSolnPres.clear();
  if (p->mixed) {
    std::vector<int> pn;
    for (int e = 0; e < nElem; ++e)
      for (int q = 0; q < 4; ++q) pn.push_back(p->conn[(size_t)e * npe + q]);
    find_unique(pn);
    std::vector<char> isP(nNode, 0);
    for (int v : pn) isP[v] = 1;
    for (int i = 0; i < nNode; ++i)
      if (isP[i]) { p->ID_pres[i] = p->presDOF++; p->assyForSolnPres.push_back(i); }
  }
  p->ntot = p->dispDOF + p->presDOF;
  // forAssyMat push_back + findUnique
  std::vector<std::vector<int>> fam(p->ntot);
  int fav[30];
  for (int e = 0; e < nElem; ++e) {
    elem_forassy(p, e, fav);
    const int nsize = 3 * npe;
    for (int ii = 0; ii < nsize; ++ii) {
      const int r = fav[ii];
      if (r == -1) continue;
      for (int jj = 0; jj < nsize; ++jj)
        if (fav[jj] != -1) fam[r].push_back(fav[jj]);
    }
  }
  if (p->presDOF > 0) {
    const int off = p->dispDOF;
    for (int e = 0; e < nElem; ++e) {
      elem_forassy(p, e, fav);
      int ttP[4];
      for (int q = 0; q < 4; ++q) ttP[q] = p->ID_pres[p->conn[(size_t)e * npe + q]];
      for (int ii = 0; ii < 4; ++ii) {
        int row = ttP[ii];
        if (row == -1) continue;
        row += off;
        for (int jj = 0; jj < 3 * npe; ++jj) {
          const int col = fav[jj];
          if (col != -1) { fam[row].push_back(col); fam[col].push_back(row); }
        }
        for (int jj = 0; jj < 4; ++jj)
          if (ttP[jj] != -1) fam[row].push_back(ttP[jj] + off);
      }
    }
  }
  p->rowptr.assign(p->ntot + 1, 0);
  for (int r = 0; r < p->ntot; ++r) { find_unique(fam[r]); p->rowptr[r + 1] = p->rowptr[r] + (int)fam[r].size(); }
  p->colidx.resize(p->rowptr[p->ntot]);
  for (int r = 0; r < p->ntot; ++r) std::copy(fam[r].begin(), fam[r].end(), p->colidx.begin() + p->rowptr[r]);
  // per-node prescribed target values (value per unit load factor)
  p->dispTarget.assign((size_t)nNode * 3, 0.0);
  p->prepared = true;
  return 0;
}

void orc_sizes(const orc_problem* p, int32_t* dispDOF, int32_t* presDOF, int64_t* nnz) {
  *dispDOF = p->dispDOF; *presDOF = p->presDOF; *nnz = (int64_t)p->colidx.size();
}

void orc_get_pattern(const orc_problem* p, int32_t* rowptr, int32_t* colidx, int32_t* nda, int32_t* idp) {
  if (rowptr) std::copy(p->rowptr.begin(), p->rowptr.end(), rowptr);
  if (colidx) std::copy(p->colidx.begin(), p->colidx.end(), colidx);
  if (nda) std::copy(p->NodeDofArray.begin(), p->NodeDofArray.end(), nda);
  if (idp) std::copy(p->ID_pres.begin(), p->ID_pres.end(), idp);
}

uint64_t orc_color_priority(int64_t e) {
  uint64_t z = (uint64_t)e + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Element colouring (no reference counterpart): greedy first-fit over the
// element-node conflict graph in decreasing priority order.
int orc_color(const orc_problem* p, int32_t* color) {
  const int nElem = p->nElem, npe = p->npe, nNode = p->nNode;
  std::vector<int> nptr(nNode + 1, 0), nel;
  for (int e = 0; e < nElem; ++e)
    for (int a = 0; a < npe; ++a) nptr[p->conn[(size_t)e * npe + a] + 1]++;
  for (int i = 0; i < nNode; ++i) nptr[i + 1] += nptr[i];
  nel.resize(nptr[nNode]);
  std::vector<int> fill(nptr.begin(), nptr.end() - 1);
  for (int e = 0; e < nElem; ++e)
    for (int a = 0; a < npe; ++a) nel[fill[p->conn[(size_t)e * npe + a]]++] = e;
  std::vector<int> order(nElem);
  for (int e = 0; e < nElem; ++e) order[e] = e;
  std::sort(order.begin(), order.end(),
            [](int a, int b) { return orc_color_priority(a) > orc_color_priority(b); });
  std::fill(color, color + nElem, -1);
  int ncol = 0;
  std::vector<char> used;
  for (int e : order) {
    used.assign(ncol + 1, 0);
    for (int a = 0; a < npe; ++a) {
      const int nd = p->conn[(size_t)e * npe + a];
      for (int k = nptr[nd]; k < nptr[nd + 1]; ++k) {
        const int c = color[nel[k]];
        if (c >= 0) used[c] = 1;
      }
    }
    int c = 0;
    while (used[c]) ++c;
    color[e] = c;
    ncol = std::max(ncol, c + 1);
  }
  return ncol;
}

void orc_set_state(orc_problem* p, const double* disp, const double* dispPrev, const double* pres,
                   const double* presPrev) {
  const size_t n3 = (size_t)p->nNode * 3;
  const int np = p->p0 ? p->nElem : p->nNode;  // pressures per element (Q1/P0) or per node
  if (disp) std::copy(disp, disp + n3, p->disp.begin());
  if (dispPrev) std::copy(dispPrev, dispPrev + n3, p->dispPrev.begin());
  if (pres) std::copy(pres, pres + np, p->pres.begin());
  if (presPrev) std::copy(presPrev, presPrev + np, p->presPrev.begin());
}

void orc_get_state(const orc_problem* p, double* disp, double* pres) {
  if (disp) std::copy(p->disp.begin(), p->disp.end(), disp);
  if (pres) std::copy(p->pres.begin(), p->pres.begin() + (p->p0 ? p->nElem : p->nNode), pres);
}

// setBoundaryConditions (femSolidmechanics1.cpp:813-915): dispApplied = value*lambda - disp.
void orc_set_bc(orc_problem* p, double lam) {
  std::fill(p->dispApplied.begin(), p->dispApplied.end(), 0.0);
  for (const auto& dp : p->dir)
    for (size_t i = 0; i < dp.nodes.size(); ++i) {
      const int nn = dp.nodes[i];
      if (dp.dof >= 0) p->dispApplied[3 * nn + dp.dof] = dp.value[i] * lam;
      else for (int d = 0; d < 3; ++d) p->dispApplied[3 * nn + d] = 0.0;
    }
  for (size_t i = 0; i < p->dispApplied.size(); ++i) p->dispApplied[i] -= p->disp[i];
}

int orc_assemble(orc_problem* p, int first_iter, double lam, double* vals, double* rhs) {
  std::fill(vals, vals + p->colidx.size(), 0.0);  // SolverPetsc::zeroMtx (SolverPetsc.cpp:145-163)
  std::fill(rhs, rhs + p->ntot, 0.0);
  if (assemble_range(p, 0, p->nElem, first_iter != 0, vals, rhs)) return 1;
  if (!p->trac.empty()) {
    std::vector<double> Fext;
    external_forces(p, lam, Fext);
    for (int i = 0; i < p->dispDOF; ++i) rhs[i] += Fext[p->assyForSoln[i]];  // femSolidmechanics2.cpp:785-789
  }
  return 0;
}

int orc_assemble_range(orc_problem* p, int first, int count, double* vals, double* rhs) {
  return assemble_range(p, first, count, false, vals, rhs);
}

// The same element loop over all elements, colour by colour with OpenMP threads inside a colour
// (elements of one colour share no node, so their scatters touch disjoint rows).  CPU-baseline
// timing only: the CSR sums come in another order than the serial loop's.
int orc_assemble_colored(orc_problem* p, const int32_t* color, int nthreads, double* vals, double* rhs) {
  int ncol = 0;
  for (int e = 0; e < p->nElem; ++e) ncol = std::max(ncol, color[e] + 1);
  std::vector<std::vector<int>> lists(ncol);
  for (int e = 0; e < p->nElem; ++e) lists[color[e]].push_back(e);
  int bad = 0;
#pragma omp parallel num_threads(nthreads) reduction(| : bad)
  {
    AsmWork W(3 * p->npe);
    for (int c = 0; c < ncol; ++c) {
      const std::vector<int>& L = lists[c];
#pragma omp for schedule(dynamic, 16)
      for (int i = 0; i < (int)L.size(); ++i) bad |= assemble_elem(p, L[i], false, vals, rhs, W);
    }
  }
  return bad;
}

int orc_pcg(const orc_problem* p, const double* vals, const double* b, double* x, int precond, double rtol, int maxit,
            double* relres) {
  return pcg(p, vals, b, x, precond, rtol, maxit, relres);
}

int orc_gmres(const orc_problem* p, const double* vals, const double* b, double* x, int precond, int restart,
              double rtol, int maxit, double* relres) {
  return gmres(p, vals, b, x, precond, restart, rtol, maxit, relres);
}

int orc_direct_solve(const orc_problem* p, const double* vals, const double* b, double* x) {
  return direct_solve(p, vals, b, x);
}

void orc_spmv(const orc_problem* p, const double* vals, const double* x, double* y) { spmv(p, vals, x, y); }

// solveWithNewtonRaphson (femSolidmechanics2.cpp:730-851) with timeUpdate / updateIterStep /
// factoriseSolveAndUpdate (femSolidmechanics1.cpp:1126-1191; femSolidmechanics2.cpp:1124-1217),
// STATIC scheme, load factor lambda(t) = t, MyTime::update cur += dt (MyTime.cpp:26-37).
// Stress at the Gauss points (Elem_Solid_3D_Disp::projectStress, Elem_Solid_3D_Disp.cpp:480-557):
// reference-configuration gradients, F from the current displacement, the material's Cauchy stress
// (the pressure field enters it for the mixed elements), per point 11 values: the 9 components in
// the column-major map, von Mises (varindex 9, :550-551) and the mean stress tr(sigma)/3 (the
// pressure field itself for the mixed elements, whose deviatoric part is traceless).
int orc_stress_gauss(const orc_problem* p, double* out) {
  double xi[10][3], w[10];
  const int nGP = gauss3d(p->npe, xi, w);
  const bool mixedP1 = p->mixed, mixed = p->mixed || p->p0;
  for (int e = 0; e < p->nElem; ++e) {
    double X[10][3], u[10][3];
    const int* cn = &p->conn[(size_t)e * p->npe];
    for (int a = 0; a < p->npe; ++a)
      for (int d = 0; d < 3; ++d) {
        X[a][d] = p->X[3 * (size_t)cn[a] + d];
        u[a][d] = p->disp[3 * (size_t)cn[a] + d];
      }
    for (int gp = 0; gp < nGP; ++gp) {
      double N[10], dx[10], dy[10], dz[10], Jac, F[3][3], stre[9], C[9][9];
      basis3d(p->npe, xi[gp], X, N, dx, dy, dz, Jac);
      defgrad(p->npe, u, dx, dy, dz, F);
      double pres = 0.0;
      if (mixedP1) {
        double Np[4];
        basis_tetra(4, xi[gp][0], xi[gp][1], xi[gp][2], Np, nullptr, nullptr, nullptr);
        for (int q = 0; q < 4; ++q) pres += (Np[q] * p->pres[cn[q]]);
      } else if (p->p0) {
        pres = p->pres[e];
      }
      neohooke(mixed, p->Utype, p->Kinv, p->mu, F, pres, stre, C);
      double* o = out + ((size_t)e * nGP + gp) * 11;
      for (int k = 0; k < 9; ++k) o[k] = stre[k];
      o[9] = std::sqrt((std::pow(stre[0] - stre[4], 2.0) + std::pow(stre[4] - stre[8], 2.0) +
                        std::pow(stre[8] - stre[0], 2.0) +
                        6.0 * (stre[1] * stre[1] + stre[2] * stre[2] + stre[5] * stre[5])) / 2.0);
      o[10] = (stre[0] + stre[4] + stre[8]) / 3.0;
    }
  }
  return nGP;
}

// Element value = mean over its Gauss points (Elem_Solid_3D_Disp::elementContourplot,
// Elem_Solid_3D_Disp.cpp:416-421); nodal value = mean of the element values over the elements
// around the node in increasing element id (femSolidmechanics::projectStresses,
// femSolidmechanics3.cpp:918-960, with each element's value taken at all of its nodes).
void orc_stress_project(const orc_problem* p, double* elem, double* nodal) {
  double xi[10][3], w[10];
  const int nGP = gauss3d(p->npe, xi, w);
  std::vector<double> g((size_t)p->nElem * nGP * 11), ev((size_t)p->nElem * 11);
  orc_stress_gauss(p, g.data());
  for (int e = 0; e < p->nElem; ++e)
    for (int k = 0; k < 11; ++k) {
      double s = 0.0;
      for (int gp = 0; gp < nGP; ++gp) s += g[((size_t)e * nGP + gp) * 11 + k];
      ev[(size_t)e * 11 + k] = s / nGP;
    }
  if (elem) std::copy(ev.begin(), ev.end(), elem);
  if (!nodal) return;
  std::vector<double> acc((size_t)p->nNode * 11, 0.0);
  std::vector<int> cnt(p->nNode, 0);
  for (int e = 0; e < p->nElem; ++e)
    for (int a = 0; a < p->npe; ++a) {
      const int nd = p->conn[(size_t)e * p->npe + a];
      for (int k = 0; k < 11; ++k) acc[(size_t)nd * 11 + k] += ev[(size_t)e * 11 + k];
      ++cnt[nd];
    }
  for (int n = 0; n < p->nNode; ++n)
    for (int k = 0; k < 11; ++k) nodal[(size_t)n * 11 + k] = cnt[n] ? acc[(size_t)n * 11 + k] / cnt[n] : 0.0;
}

// Arc-length continuation: femSolidmechanics::solveWithArclength (femSolidmechanics2.cpp:855-1039)
// with SolverPetsc::solveArclengthSystem (SolverPetsc.cpp:671-757), timeUpdate's saveSolution
// (SolutionDataSolid.cpp:170-184: dispPrev2 = dispPrev, dispPrev = disp at the start of every attempt,
// so the displacement predictor reduces to the last state and only the load factor is extrapolated)
// and the direct solve.  Fext is the traction patches' force at load factor 1; prescribed dofs stay
// at zero (the loop never adds prescribed increments).  Per converged step: lam_out, the arc length
// in force at that step ds_out (0 for the first), the displacement u_out [n3] (may be NULL); the
// state after the last step stays in p->disp.
// Returns the number of converged steps, or -1 when a solve fails.
int orc_arclength(orc_problem* p, double dt, double timeFinal, int stepsMax, int iterationsMax, double conv_tol,
                  double* lam_out, double* ds_out, double* u_out, int32_t* attempts_out) {
  const double EPS = 1e-10;
  const int n = p->ntot, n3 = 3 * p->nNode;
  std::vector<double> Fext;
  external_forces(p, 1.0, Fext);
  double FtF = 0.0;
  for (int i = 0; i < n3; ++i) FtF += Fext[i] * Fext[i];
  std::vector<double> Fshort(n, 0.0), vals(p->colidx.size()), rhs(n), du1(n), du2(n), Dus(n);
  for (int i = 0; i < p->dispDOF; ++i) Fshort[i] = Fext[p->assyForSoln[i]];
  std::vector<double> dispPrev = p->disp, dispPrev2 = p->disp, DuFull(n3, 0.0);
  double loadFactor = dt, loadFactorPrev = 0.0, loadFactorPrev2 = 0.0;
  double arclenIncr = 0.0, arclenIncrPrev = 0.0, arclenIncrMax = 0.0, arclenIncrMin = 0.0;
  bool convergedFlag = false, convergedFlagPrev = false;
  double cur = 0.0;
  int stepsCompleted = 1, attempts = 0;
  std::fill(p->dispApplied.begin(), p->dispApplied.end(), 0.0);
  while (cur <= timeFinal - EPS && stepsCompleted <= stepsMax) {
    cur += dt;  // MyTime::update
    ++attempts;
    dispPrev2 = dispPrev;  // saveSolution
    dispPrev = p->disp;
    convergedFlagPrev = convergedFlag;
    convergedFlag = false;
    if (stepsCompleted > 1) {
      const double DsFactor = arclenIncr / arclenIncrPrev;
      for (int i = 0; i < n3; ++i) p->disp[i] = (1.0 + DsFactor) * dispPrev[i] - DsFactor * dispPrev2[i];
      loadFactor = (1.0 + DsFactor) * loadFactorPrev - DsFactor * loadFactorPrev2;
    }
    for (int i = 0; i < n3; ++i) DuFull[i] = p->disp[i] - dispPrev[i];
    double Dl = loadFactor - loadFactorPrev, dl = 0.0;
    for (int iter = 1; iter < iterationsMax; ++iter) {
      std::fill(vals.begin(), vals.end(), 0.0);
      std::fill(rhs.begin(), rhs.end(), 0.0);
      if (assemble_range(p, 0, p->nElem, iter == 1, vals.data(), rhs.data())) break;  // negative Jacobian
      for (int i = 0; i < p->dispDOF; ++i) rhs[i] += loadFactor * Fext[p->assyForSoln[i]];
      if (norm2(rhs.data(), n) < conv_tol) { convergedFlag = true; break; }
      // solveArclengthSystem (psi = 1)
      for (int i = 0; i < p->dispDOF; ++i) Dus[i] = DuFull[p->assyForSoln[i]];
      double A = 0.0, b = 1.0;
      const bool later = stepsCompleted > 1;
      if (later) {
        double dd = 0.0;
        for (int i = 0; i < n; ++i) dd += Dus[i] * Dus[i];
        A = dd + Dl * Dl * FtF - arclenIncr * arclenIncr;
        b = 2.0 * Dl * FtF;
      }
      if (direct_solve(p, vals.data(), rhs.data(), du2.data())) return -1;
      for (int i = 0; i < n; ++i) du2[i] = -du2[i];
      if (direct_solve(p, vals.data(), Fshort.data(), du1.data())) return -1;
      double adu2 = 0.0, adu1 = 0.0;
      if (later)
        for (int i = 0; i < n; ++i) { adu2 += 2.0 * Dus[i] * du2[i]; adu1 += 2.0 * Dus[i] * du1[i]; }
      dl = (adu2 - A) / (b + adu1);
      for (int i = 0; i < p->dispDOF; ++i) {
        const double v = -du2[i] + dl * du1[i];
        p->disp[p->assyForSoln[i]] += v;
        DuFull[p->assyForSoln[i]] += v;
      }
      loadFactor += dl;
      Dl += dl;
    }
    if (convergedFlag) {
      if (stepsCompleted == 1) {
        double dd = 0.0;
        for (int i = 0; i < n3; ++i) dd += DuFull[i] * DuFull[i];
        arclenIncr = std::sqrt(dd + loadFactor * loadFactor * FtF);
        arclenIncrMax = arclenIncr;
        arclenIncrMin = arclenIncr / 1024.0;
      }
      if (lam_out) lam_out[stepsCompleted - 1] = loadFactor;
      if (ds_out) ds_out[stepsCompleted - 1] = stepsCompleted > 1 ? arclenIncr : 0.0;
      if (u_out) std::copy(p->disp.begin(), p->disp.end(), u_out + (size_t)(stepsCompleted - 1) * n3);
      loadFactorPrev2 = loadFactorPrev;
      loadFactorPrev = loadFactor;
      dispPrev2 = dispPrev;
      dispPrev = p->disp;
      arclenIncrPrev = arclenIncr;
      if (convergedFlagPrev) arclenIncr = std::min(std::max(2.0 * arclenIncr, arclenIncrMin), arclenIncrMax);
      ++stepsCompleted;
    } else {
      arclenIncr = std::max(arclenIncr * (convergedFlagPrev ? 0.5 : 0.25), arclenIncrMin);
    }
  }
  if (attempts_out) *attempts_out = attempts;
  return stepsCompleted - 1;
}

int orc_newton(orc_problem* p, int nsteps, double dt, int solver, int precond, double lin_rtol, double rtol, int itmax,
               int32_t* iters_out, double* final_relres) {
  const int n = p->ntot;
  std::vector<double> vals(p->colidx.size()), rhs(n), du(n);
  double cur = 0.0;
  double lastrel = 0.0;
  for (int step = 0; step < nsteps; ++step) {
    cur += dt;
    const double lam = cur;
    p->dispPrev = p->disp;  // saveSolution (SolutionDataSolid.cpp:170-184)
    p->presPrev = p->pres;
    orc_set_bc(p, lam);
    double R0 = -1.0;
    bool conv = false;
    int iter = 1;
    for (; iter <= itmax; ++iter) {
      const bool first = (iter == 1);
      if (orc_assemble(p, first, lam, vals.data(), rhs.data())) return 2;  // negative Jacobian
      const double rn = norm2(rhs.data(), n);
      if (first) R0 = rn;
      lastrel = (R0 > 0) ? rn / R0 : 0.0;
      if (R0 == 0.0 || rn / R0 < rtol) { conv = true; break; }
      if (std::isnan(rn)) return 3;
      int err = 0;
      double rr;
      if (solver == 0) err = direct_solve(p, vals.data(), rhs.data(), du.data());
      else if (solver == 1) pcg(p, vals.data(), rhs.data(), du.data(), precond, lin_rtol, 100000, &rr);
      else gmres(p, vals.data(), rhs.data(), du.data(), precond, 50, lin_rtol, 100000, &rr);
      if (err) return 4;
      if (first)
        for (int d : p->dofs_specified) p->disp[d] += p->dispApplied[d];
      for (int i = 0; i < p->dispDOF; ++i) p->disp[p->assyForSoln[i]] += du[i];
      for (int i = 0; i < p->presDOF; ++i) p->pres[p->assyForSolnPres[i]] += du[p->dispDOF + i];
      if (p->p0) {
        // solveForPressure (Elem_Solid_3D_Mixed0.cpp:112-199) with dispIncr: the solved correction on
        // free dofs and, in the first iteration, dispApplied on prescribed ones (femSolidmechanics2.cpp:1137-1186)
        std::vector<double> inc((size_t)p->nNode * 3, 0.0);
        if (first)
          for (int d : p->dofs_specified) inc[d] = p->dispApplied[d];
        for (int i = 0; i < p->dispDOF; ++i) inc[p->assyForSoln[i]] = du[i];
        for (int e = 0; e < p->nElem; ++e) {
          double v = 0.0;
          for (int a = 0; a < 8; ++a)
            for (int d = 0; d < 3; ++d) v += p->qKpu[(size_t)e * 24 + 3 * a + d] * inc[3 * p->conn[(size_t)e * 8 + a] + d];
          p->pres[e] += (p->qFp[e] - v) / p->qKpp[e];
        }
      }
    }
    if (iters_out) iters_out[step] = iter;
    if (!conv) return 5;
  }
  if (final_relres) *final_relres = lastrel;
  return 0;
}

}  // extern "C"
