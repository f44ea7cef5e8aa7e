// pmg.hip -- two-level p-multigrid preconditioner for Tet10 and Tet10/P1 systems on gfx950.
//
// The reference solves the assembled system with KSPSolve behind SolverPetsc::factoriseAndSolve
// (SolverPetsc.cpp:250-330), the Krylov method and preconditioner coming from petsc_options
// (code defaults CG + ILU, :117-140).  For the nearly incompressible Tet10/P1 system a node-block
// Jacobi GMRES stalls; this preconditioner, used inside flexible GMRES (krylov.hip), reaches
// ||b - A x|| / ||b|| < 1e-8 with iteration counts that grow slowly with the mesh.
//
// Levels: the fine level is the Tet10 displacement dofs, the coarse level the vertex dofs (the
// linear Tet4 space inside the quadratic one), reached by the interpolation P: a vertex keeps its
// value, the mid-edge node of edge (a, b) takes (u_a + u_b) / 2, with the element's edge table
// 4:(0,1) 5:(1,2) 6:(0,2) 7:(0,3) 8:(2,3) 9:(1,3) (BasisFunctionsLagrange.cpp:870-956).  Only free
// dofs take part: a fixed vertex dof has no coarse unknown, so P has no entry towards it.
//
// Application z = M^-1 v (block upper-triangular in the displacement / pressure split):
//   z_p = S^-1 v_p                 S = diag(Kpp) (1 + 1 / (mu Kinv)): the pressure mass matrix
//                                  (Kpp = -M / K) scaled to the Schur complement -M (1/K + 1/mu)
//   z_u = V(v_u - Kup z_p)         V: one V-cycle on Kuu
// V-cycle with B the 3x3 node blocks of Kuu and Kc = P^T Kuu P:
//   x = Cheb(t, 0);  r = t - Kuu x;  x += P CG(Kc, P^T r);  x = Cheb(t, x)
// Cheb: the Chebyshev iteration of the given degree for Kuu x = t preconditioned by B, over the
// interval [beta / ratio, beta] of B^-1 Kuu's spectrum (Saad, Iterative Methods for Sparse Linear
// Systems, 2nd ed., Alg. 12.1), beta = 1.1 lambda, lambda the Rayleigh-quotient estimate of the
// largest eigenvalue after 20 power steps from the ones vector (the factors of PETSc's
// KSPChebyshev default for multigrid smoothing).  Damped Jacobi would need the damping matched to
// lambda, which ranges from about 2 (Kuu of Tet10/P1) to above 4 (nearly incompressible Tet10).
// The coarse CG (node-block Jacobi, krylov.hip) stops at a loose relative tolerance, which makes
// M change from one application to the next; the outer method is therefore flexible GMRES.
//
// Device layout: P as two (coarse index, weight) slots per fine displacement row; P^T as CSR over
// the coarse rows (fine rows ascending); Kc in CSR over the vertex-adjacency pattern (vertices
// that share an element), rows of one vertex consecutive, so the coarse CG uses the row-group
// SpMV.  Kc is formed on the device from the assembled CSR each solve, one thread per coarse row,
// in a fixed order (bitwise reproducible).

#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>

#include "mpf_internal.hpp"

namespace mpf {

struct PmgWork {
  // parameters
  int degree = 1;          // Chebyshev degree before and after the coarse correction
  double ratio = 10.0;     // beta / alpha of the smoothing interval
  double crtol = 0.1;
  int cmaxit = 1000;
  // tables (built once per pattern)
  bool built = false;
  const int32_t* rowptr_seen = nullptr;
  int64_t d = 0, pn = 0, nc = 0, nnzc = 0, ncb = 0;
  DevBuf<int32_t> pidx;               // [2 d] coarse parents of each fine displacement row (-1: none)
  DevBuf<double> pw;                  // [2 d] their weights
  DevBuf<int32_t> rptr, rfine;        // P^T: [nc + 1], [d_entries]
  DevBuf<double> rw;
  DevBuf<int32_t> crow, ccol;         // Kc pattern
  DevBuf<double> cval;
  DevBuf<float> cval32;  // FP32 copy of cval read by the coarse CG's products
  DevBuf<float> fval32;  // FP32 copy of the system's values read by the smoothing sweeps
  DevBuf<int32_t> cbs, cbm, crblk;    // coarse node blocks
  DevBuf<double> cbinv;
  bool cgrp = false;
  DevBuf<double> sinv;                // [pn]
  DevBuf<double> t, xa, xb, dd, res;  // [d]
  DevBuf<int32_t> ulen;               // [nub]: displacement columns of each node block's rows
  bool ulen_ready = false;
  double lambda = 0.0;                // estimate of the largest eigenvalue of B^-1 Kuu
  DevBuf<double> rc, xc;              // [nc]
  DevBuf<double> cr, cr2, cz, cp, cq, cpart, csc;  // coarse CG (two levels) / aggregate CG (three)
  DevBuf<int32_t> bad;
  // third level (levels == 3): the vertex level gets a V-cycle of its own -- Chebyshev smoothing
  // with its node blocks, a correction from aggregates of 8 vertices (consecutive in Morton order,
  // one dof per aggregate and component, unsmoothed: Ka = Pa^T Kc Pa), Ka solved by block-Jacobi
  // CG to crtol -- in place of the CG on Kc
  int levels = 3;
  int64_t na = 0, nka = 0, nab = 0;        // aggregate dofs, nonzeros of Ka, aggregate blocks
  DevBuf<int32_t> aoff;                    // [nc] aggregate dof of each vertex-level dof
  DevBuf<int32_t> amptr, amem;             // [na + 1], [nc]: vertex-level dofs of each aggregate dof
  DevBuf<int32_t> karow, kacol;            // Ka pattern
  DevBuf<double> kaval;
  DevBuf<float> kaval32;                   // FP32 copy read by the aggregate CG's products
  DevBuf<int32_t> kabs, kabm, karblk;      // aggregate blocks (an aggregate's components)
  DevBuf<double> kabinv;
  bool kagrp = false;
  DevBuf<int32_t> culen;                   // [ncb] row length of each vertex block's rows
  DevBuf<double> cxa, cxb, cdd, cres;      // [nc] vertex-level V-cycle
  DevBuf<double> ra, xa3;                  // [na] aggregate right-hand side and solution
  double lambda_c = 0.0;                   // largest eigenvalue of Bc^-1 Kc
  // statistics of the current solve
  int64_t vcycles = 0, coarse_its = 0;
  double setup_seconds = 0.0;
};

namespace {

constexpr int kT = 256;

// Kc(I, J) = sum over f in row I of P^T of w_fI sum over g with Kuu(f, g) != 0 of Kuu(f, g) w_gJ,
// one thread per coarse row, visited in (f, g, parent slot) order (measured on the 60^3 Tet10/P1
// system: 188 ms with the row's pattern and sums in global memory).  bad[0] = 1 if some (I, J) is
// not in the pattern (a mesh that is not conforming).
constexpr int kGT = 64;    // threads (coarse rows) per workgroup of k_galerkin
constexpr int kGCap = 80;  // coarse row lengths kept in LDS (longer rows work in global memory)
__global__ void __launch_bounds__(kGT) k_galerkin(int64_t nc, const int32_t* __restrict__ rptr,
                                                  const int32_t* __restrict__ rfine, const double* __restrict__ rw,
                                                  const int32_t* __restrict__ rowptr,
                                                  const int32_t* __restrict__ colidx, const double* __restrict__ vals,
                                                  int32_t d, const int32_t* __restrict__ pidx,
                                                  const double* __restrict__ pw, const int32_t* __restrict__ crow,
                                                  const int32_t* __restrict__ ccol, double* __restrict__ cval,
                                                  int32_t* __restrict__ bad) {
  // each thread's coarse row: its column pattern and running sums in its own LDS column
  __shared__ int32_t sCol[kGCap][kGT];
  __shared__ double sVal[kGCap][kGT];
  const int th = threadIdx.x;
  const int64_t I = (int64_t)blockIdx.x * kGT + th;
  if (I >= nc) return;  // no barrier below
  const int cb = crow[I], ce = crow[I + 1], len = ce - cb;
  const bool inl = len <= kGCap;
  if (inl) {
    for (int j = 0; j < len; ++j) {
      sCol[j][th] = ccol[cb + j];
      sVal[j][th] = 0.0;
    }
  } else {
    for (int k = cb; k < ce; ++k) cval[k] = 0.0;
  }
  for (int e = rptr[I]; e < rptr[I + 1]; ++e) {
    const int f = rfine[e];
    const double wf = rw[e];
    const int k1 = rowptr[f + 1];
    // the fine row in chunks of 8 entries: columns and values loaded together, then their parents
    // and weights together (a round trip per stage, not per entry); summed in entry order
    bool done = false;
    for (int k0 = rowptr[f]; k0 < k1 && !done; k0 += 8) {
      int gc[8], jc[16];
      double vc[8], wc[16];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        gc[u] = k0 + u < k1 ? colidx[k0 + u] : d;
        vc[u] = k0 + u < k1 ? vals[k0 + u] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < 8; ++u)
#pragma unroll
        for (int t = 0; t < 2; ++t) {
          jc[2 * u + t] = gc[u] < d ? pidx[2 * (size_t)gc[u] + t] : -1;
          wc[2 * u + t] = gc[u] < d ? pw[2 * (size_t)gc[u] + t] : 0.0;
        }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int g = gc[u];
        if (g >= d) {  // columns ascend: the pressure columns (or the row's end) close the row
          done = true;
          break;
        }
        const double kv = wf * vc[u];
        for (int t = 0; t < 2; ++t) {
          const int J = jc[2 * u + t];
          if (J < 0) continue;
          const double pwt = wc[2 * u + t];
          if (inl) {
            int lo = 0, hi = len;
            while (lo < hi) {
              const int mid = (lo + hi) >> 1;
              if (sCol[mid][th] < J) lo = mid + 1; else hi = mid;
            }
            if (lo < len && sCol[lo][th] == J) sVal[lo][th] += kv * pwt;
            else bad[0] = 1;
          } else {
            int lo = cb, hi = ce;
            while (lo < hi) {
              const int mid = (lo + hi) >> 1;
              if (ccol[mid] < J) lo = mid + 1; else hi = mid;
            }
            if (lo < ce && ccol[lo] == J) cval[lo] += kv * pwt;
            else bad[0] = 1;
          }
        }
      }
    }
  }
  if (inl)
    for (int j = 0; j < len; ++j) cval[cb + j] = sVal[j][th];
}

// FP32 copies of the matrices the preconditioner applies (the coarse matrix in the coarse CG's
// products, the system's values in the smoothing sweeps): the V-cycle is a preconditioner inside
// flexible GMRES, whose residual is formed with the FP64 system, so its operators can be rounded to
// FP32 -- they stay symmetric, the products accumulate in FP64, and the matrix bytes each product
// reads drop from 8 to 4 per value.
__global__ void k_to_f32(int64_t n, const double* __restrict__ a, float* __restrict__ b) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}

// Ka = Pa^T Kc Pa with Pa the aggregation (0/1): one thread per aggregate dof row A, which sums the
// Kc rows of its member dofs (ascending) into the columns of their aggregate dofs, in a fixed order.
// bad[0] = 1 if some column is not in the pattern.
__global__ void __launch_bounds__(kT) k_agg_values(int64_t na, const int32_t* __restrict__ amptr,
                                                   const int32_t* __restrict__ amem, const int32_t* __restrict__ crow,
                                                   const int32_t* __restrict__ ccol, const double* __restrict__ cval,
                                                   const int32_t* __restrict__ aoff, const int32_t* __restrict__ karow,
                                                   const int32_t* __restrict__ kacol, double* __restrict__ kaval,
                                                   int32_t* __restrict__ bad) {
  const int64_t A = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (A >= na) return;
  const int kb = karow[A], ke = karow[A + 1];
  for (int k = kb; k < ke; ++k) kaval[k] = 0.0;
  for (int t = amptr[A]; t < amptr[A + 1]; ++t) {
    const int i = amem[t];
    for (int k = crow[i]; k < crow[i + 1]; ++k) {
      const int J = aoff[ccol[k]];
      int lo = kb, hi = ke;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (kacol[mid] < J) lo = mid + 1; else hi = mid;
      }
      if (lo < ke && kacol[lo] == J) kaval[lo] += cval[k];
      else *bad = 1;
    }
  }
}

// ra = Pa^T r (each aggregate dof sums its members in ascending order)
__global__ void k_agg_restrict(int64_t na, const int32_t* __restrict__ amptr, const int32_t* __restrict__ amem,
                               const double* __restrict__ r, double* __restrict__ ra) {
  const int64_t A = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (A >= na) return;
  double s = 0.0;
  for (int t = amptr[A]; t < amptr[A + 1]; ++t) s += r[amem[t]];
  ra[A] = s;
}

// x += Pa xa
__global__ void k_agg_prolong(int64_t nc, const int32_t* __restrict__ aoff, const double* __restrict__ xa,
                              double* __restrict__ x) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nc) x[i] += xa[aoff[i]];
}

__global__ void k_zero(int64_t n, double* __restrict__ x) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = 0.0;
}

// Pressure block: 1 / (Kpp_qq scale), or the node-block Schur diagonal's inverse (binv_p[9 q])
// when scale is 0 or the diagonal is zero.
__global__ void __launch_bounds__(kT) k_sinv(int64_t pn, int32_t d, const int32_t* __restrict__ rowptr,
                                             const int32_t* __restrict__ colidx, const double* __restrict__ vals,
                                             double scale, const double* __restrict__ binv_p,
                                             double* __restrict__ sinv) {
  const int64_t q = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (q >= pn) return;
  const int r = d + (int)q;
  int lo = rowptr[r], hi = rowptr[r + 1];
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (colidx[mid] < r) lo = mid + 1; else hi = mid;
  }
  const double dg = (lo < rowptr[r + 1] && colidx[lo] == r) ? vals[lo] : 0.0;
  sinv[q] = (scale > 0.0 && dg != 0.0) ? 1.0 / (dg * scale) : binv_p[9 * q];
}

// z_p = S^-1 v_p
__global__ void __launch_bounds__(kT) k_pres(int64_t pn, const double* __restrict__ vp, const double* __restrict__ sinv,
                                             double* __restrict__ zp) {
  const int64_t q = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (q < pn) zp[q] = vp[q] * sinv[q];
}

// ulen[k]: the number of displacement columns (< d) of node block k's rows, which share one column
// list with the pressure columns (>= d) after them
__global__ void __launch_bounds__(kT) k_ulen(int64_t nub, const int32_t* __restrict__ bs,
                                             const int32_t* __restrict__ rowptr, const int32_t* __restrict__ colidx,
                                             int32_t d, int32_t* __restrict__ ulen) {
  const int64_t k = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (k >= nub) return;
  const int r = bs[k], b0 = rowptr[r];
  int lo = b0, hi = rowptr[r + 1];
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (colidx[mid] < d) lo = mid + 1; else hi = mid;
  }
  ulen[k] = lo - b0;
}

// First Chebyshev step from x = 0 over the node blocks, LPR lanes per block: t = v - Kup z_p on the
// block's rows (zp non-null: the pressure columns, shared by the block's rows, each read once),
// x = dd = c2 B^-1 t.
template <int LPR>
__global__ void __launch_bounds__(kT) k_jac0(int64_t nub, const int32_t* __restrict__ bs,
                                             const int32_t* __restrict__ bm, const double* __restrict__ binv,
                                             const int32_t* __restrict__ rowptr, const int32_t* __restrict__ colidx,
                                             const double* __restrict__ vals, const int32_t* __restrict__ ulen,
                                             int32_t d, const double* __restrict__ zp, double c2,
                                             const double* __restrict__ v, double* __restrict__ t,
                                             double* __restrict__ x, double* __restrict__ dd) {
  const int g = threadIdx.x / LPR, l = threadIdx.x % LPR;
  constexpr int GPB = kT / LPR;
  for (int64_t k0 = (int64_t)blockIdx.x * GPB; k0 < nub; k0 += (int64_t)gridDim.x * GPB) {
    const int64_t k = k0 + g;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    int st = 0, m = 0;
    if (k < nub) {
      st = bs[k];
      m = bm[k];
      if (zp) {  // the block's pressure columns over the lanes
        const int b0 = rowptr[st], len = rowptr[st + 1] - b0;
        const int b1 = m > 1 ? rowptr[st + 1] : b0, b2 = m > 2 ? rowptr[st + 2] : b0;
        for (int j = ulen[k] + l; j < len; j += LPR) {
          const double zv = zp[colidx[b0 + j] - d];
          s0 += vals[b0 + j] * zv;
          if (m > 1) s1 += vals[b1 + j] * zv;
          if (m > 2) s2 += vals[b2 + j] * zv;
        }
      }
    }
#pragma unroll
    for (int off = LPR / 2; off > 0; off >>= 1) {
      s0 += __shfl_xor(s0, off, LPR);
      s1 += __shfl_xor(s1, off, LPR);
      s2 += __shfl_xor(s2, off, LPR);
    }
    if (k < nub && l < m) {  // lane l finishes row l of the block
      double r[3] = {0.0, 0.0, 0.0};
      for (int i = 0; i < m; ++i) r[i] = v[st + i] - (i == 0 ? s0 : (i == 1 ? s1 : s2));
      if (zp) t[st + l] = r[l];
      const double* B = binv + 9 * k + 3 * l;
      double z = 0.0;
      for (int j = 0; j < m; ++j) z += B[j] * r[j];
      x[st + l] = c2 * z;
      dd[st + l] = c2 * z;
    }
  }
}

// One node-block sweep over the displacement rows, LPR lanes per block (the block's rows share
// one column list, so each column index and x value is read once for its up to three rows):
//   MODE 1 (Chebyshev step): dd = c1 dd + c2 B^-1 (t - Kuu x_in);  x_out = x_in + dd
//   MODE 2 (residual):       x_out = t - Kuu x_in
template <int LPR, int MODE, typename V = double>
__global__ void __launch_bounds__(kT) k_sweep(int64_t nub, const int32_t* __restrict__ bs,
                                              const int32_t* __restrict__ bm, const int32_t* __restrict__ rowptr,
                                              const int32_t* __restrict__ colidx, const V* __restrict__ vals,
                                              const int32_t* __restrict__ ulen, const double* __restrict__ binv,
                                              double c1, double c2, const double* __restrict__ t,
                                              const double* __restrict__ xin, double* __restrict__ xout,
                                              double* __restrict__ dd) {
  const int g = threadIdx.x / LPR, l = threadIdx.x % LPR;
  constexpr int GPB = kT / LPR;
  for (int64_t k0 = (int64_t)blockIdx.x * GPB; k0 < nub; k0 += (int64_t)gridDim.x * GPB) {
    const int64_t k = k0 + g;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    int st = 0, m = 0;
    // lane l < m finishes row l of the block: its operands are loaded before the row sums so that
    // their latency overlaps the sums
    double tr[3] = {0.0, 0.0, 0.0}, Bl[3] = {0.0, 0.0, 0.0}, xo = 0.0, dq = 0.0;
    if (k < nub) {
      st = bs[k];
      m = bm[k];
      if (l < m) {
        for (int j = 0; j < m; ++j) tr[j] = t[st + j];
        if (MODE == 1) {
          for (int j = 0; j < m; ++j) Bl[j] = binv[9 * k + 3 * l + j];
          xo = xin[st + l];
          if (c1 != 0.0) dq = dd[st + l];
        }
      }
      const int b0 = rowptr[st], len = ulen[k];
      const int b1 = m > 1 ? rowptr[st + 1] : b0, b2 = m > 2 ? rowptr[st + 2] : b0;
#pragma unroll 4
      for (int j = l; j < len; j += LPR) {
        const double xv = xin[colidx[b0 + j]];
        s0 += (double)vals[b0 + j] * xv;
        if (m > 1) s1 += (double)vals[b1 + j] * xv;
        if (m > 2) s2 += (double)vals[b2 + j] * xv;
      }
    }
#pragma unroll
    for (int off = LPR / 2; off > 0; off >>= 1) {  // every lane of the group ends with the sums
      s0 += __shfl_xor(s0, off, LPR);
      s1 += __shfl_xor(s1, off, LPR);
      s2 += __shfl_xor(s2, off, LPR);
    }
    if (k < nub && l < m) {
      const double r[3] = {tr[0] - s0, m > 1 ? tr[1] - s1 : 0.0, m > 2 ? tr[2] - s2 : 0.0};
      if (MODE == 2) {
        xout[st + l] = l == 0 ? r[0] : (l == 1 ? r[1] : r[2]);
      } else {
        double z = 0.0;
        for (int j = 0; j < m; ++j) z += Bl[j] * r[j];
        const double di = (c1 != 0.0 ? c1 * dq : 0.0) + c2 * z;
        dd[st + l] = di;
        xout[st + l] = xo + di;
      }
    }
  }
}

// One power step for B^-1 Kuu with x = c xin: y = B^-1 Kuu x, and per workgroup the partial sums
// of s.y, x.s and y.y (s = Kuu x) into part[0 / 1 / 2][blockIdx.x].
template <int LPR>
__global__ void __launch_bounds__(kT) k_power(int64_t nub, const int32_t* __restrict__ bs,
                                              const int32_t* __restrict__ bm, const int32_t* __restrict__ rowptr,
                                              const int32_t* __restrict__ colidx, const double* __restrict__ vals,
                                              const int32_t* __restrict__ ulen, const double* __restrict__ binv,
                                              double c, const double* __restrict__ xin, double* __restrict__ y,
                                              double* __restrict__ part) {
  __shared__ double sh[3][kT / 64];
  const int g = threadIdx.x / LPR, l = threadIdx.x % LPR;
  constexpr int GPB = kT / LPR;
  double a_sy = 0.0, a_xs = 0.0, a_yy = 0.0;
  for (int64_t k0 = (int64_t)blockIdx.x * GPB; k0 < nub; k0 += (int64_t)gridDim.x * GPB) {
    const int64_t k = k0 + g;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    int st = 0, m = 0;
    if (k < nub) {
      st = bs[k];
      m = bm[k];
      const int b0 = rowptr[st], len = ulen[k];
      const int b1 = m > 1 ? rowptr[st + 1] : b0, b2 = m > 2 ? rowptr[st + 2] : b0;
      for (int j = l; j < len; j += LPR) {
        const double xv = c * xin[colidx[b0 + j]];
        s0 += vals[b0 + j] * xv;
        if (m > 1) s1 += vals[b1 + j] * xv;
        if (m > 2) s2 += vals[b2 + j] * xv;
      }
    }
#pragma unroll
    for (int off = LPR / 2; off > 0; off >>= 1) {
      s0 += __shfl_xor(s0, off, LPR);
      s1 += __shfl_xor(s1, off, LPR);
      s2 += __shfl_xor(s2, off, LPR);
    }
    if (k < nub && l == 0) {
      const double sv[3] = {s0, s1, s2};
      const double* B = binv + 9 * k;
      for (int i = 0; i < m; ++i) {
        double z = 0.0;
        for (int j = 0; j < m; ++j) z += B[3 * i + j] * sv[j];
        y[st + i] = z;
        a_sy += sv[i] * z;
        a_xs += c * xin[st + i] * sv[i];
        a_yy += z * z;
      }
    }
  }
  double v[3] = {a_sy, a_xs, a_yy};
  for (int q = 0; q < 3; ++q) {
    for (int off = 32; off > 0; off >>= 1) v[q] += __shfl_xor(v[q], off, 64);
    if ((threadIdx.x & 63) == 0) sh[q][threadIdx.x >> 6] = v[q];
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    double t = 0.0;
    for (int w = 0; w < kT / 64; ++w) t += sh[threadIdx.x][w];
    part[(size_t)threadIdx.x * gridDim.x + blockIdx.x] = t;
  }
}

__global__ void __launch_bounds__(kT) k_ones(int64_t n, double* __restrict__ x) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i < n) x[i] = 1.0;
}

// rc = P^T r
__global__ void __launch_bounds__(kT) k_restrict(int64_t nc, const int32_t* __restrict__ rptr,
                                                 const int32_t* __restrict__ rfine, const double* __restrict__ rw,
                                                 const double* __restrict__ r, double* __restrict__ rc) {
  const int64_t I = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (I >= nc) return;
  double s = 0.0;
  const int e1 = rptr[I + 1];
  for (int e0 = rptr[I]; e0 < e1; e0 += 8) {  // 8 entries per round trip, summed in order
    int fi[8];
    double wv[8], rv[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      fi[u] = e0 + u < e1 ? rfine[e0 + u] : -1;
      wv[u] = e0 + u < e1 ? rw[e0 + u] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) rv[u] = fi[u] >= 0 ? r[fi[u]] : 0.0;
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (fi[u] >= 0) s += wv[u] * rv[u];
  }
  rc[I] = s;
}

// x += P xc
__global__ void __launch_bounds__(kT) k_prolong(int32_t d, const int32_t* __restrict__ pidx,
                                                const double* __restrict__ pw, const double* __restrict__ xc,
                                                double* __restrict__ x) {
  const int f = blockIdx.x * kT + threadIdx.x;
  if (f >= d) return;
  const int j0 = pidx[2 * (size_t)f], j1 = pidx[2 * (size_t)f + 1];
  double s = x[f];
  if (j0 >= 0) s += pw[2 * (size_t)f] * xc[j0];
  if (j1 >= 0) s += pw[2 * (size_t)f + 1] * xc[j1];
  x[f] = s;
}

// The sweeps read the FP32 copy of the matrix values (vals32): the smoother is part of the
// preconditioner, and its products still sum in FP64.
template <int MODE>
void launch_sweep(int lpr, int64_t nub, const BlockPcView& F, const CsrView& A, const float* vals32,
                  const int32_t* ulen, double c1, double c2, const double* t, const double* xin, double* xout,
                  double* dd, hipStream_t s) {
  if (lpr == 32)
    hipLaunchKernelGGL((k_sweep<32, MODE, float>), dim3(grid_for(nub, kT / 32)), dim3(kT), 0, s, nub, F.bs, F.bm,
                       A.rowptr, A.colidx, vals32, ulen, F.binv, c1, c2, t, xin, xout, dd);
  else if (lpr == 16)
    hipLaunchKernelGGL((k_sweep<16, MODE, float>), dim3(grid_for(nub, kT / 16)), dim3(kT), 0, s, nub, F.bs, F.bm,
                       A.rowptr, A.colidx, vals32, ulen, F.binv, c1, c2, t, xin, xout, dd);
  else
    hipLaunchKernelGGL((k_sweep<4, MODE, float>), dim3(grid_for(nub, kT / 4)), dim3(kT), 0, s, nub, F.bs, F.bm,
                       A.rowptr, A.colidx, vals32, ulen, F.binv, c1, c2, t, xin, xout, dd);
}

CsrView fine_csr(const mpf_problem* p);

// Largest eigenvalue of B^-1 Kuu: Rayleigh quotient (s.y) / (x.s) of 20 power steps from the ones
// vector (partial sums added on the host in workgroup order).
mpf_status estimate_lambda(hipStream_t s, const CsrView& A, const BlockPcView& F, const int32_t* ulen, int64_t nub,
                           int64_t nrows, int lpr, double* xbuf, double* ybuf, double* part, double* out) {
  const unsigned G = std::min<unsigned>(grid_for(nub, kT / lpr), kMaxPartials);
  hipLaunchKernelGGL(k_ones, dim3(grid_for(nrows, kT)), dim3(kT), 0, s, nrows, xbuf);
  double* x = xbuf;
  double* y = ybuf;
  double c = 1.0, lam = 0.0;
  std::vector<double> h(3 * (size_t)G);
  for (int it = 0; it < 20; ++it) {
    if (lpr == 32)
      hipLaunchKernelGGL(k_power<32>, dim3(G), dim3(kT), 0, s, nub, F.bs, F.bm, A.rowptr, A.colidx, A.vals, ulen, F.binv,
                         c, x, y, part);
    else if (lpr == 16)
      hipLaunchKernelGGL(k_power<16>, dim3(G), dim3(kT), 0, s, nub, F.bs, F.bm, A.rowptr, A.colidx, A.vals, ulen, F.binv,
                         c, x, y, part);
    else
      hipLaunchKernelGGL(k_power<4>, dim3(G), dim3(kT), 0, s, nub, F.bs, F.bm, A.rowptr, A.colidx, A.vals, ulen, F.binv,
                         c, x, y, part);
    MPF_HIP(hipGetLastError());
    MPF_HIP(hipMemcpyAsync(h.data(), part, h.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    MPF_HIP(hipStreamSynchronize(s));
    double sy = 0.0, xs = 0.0, yy = 0.0;
    for (unsigned q = 0; q < G; ++q) {
      sy += h[q];
      xs += h[G + q];
      yy += h[2 * (size_t)G + q];
    }
    if (!(xs > 0.0) || !(yy > 0.0)) break;
    lam = sy / xs;
    c = 1.0 / std::sqrt(yy);
    std::swap(x, y);
  }
  if (!(lam > 0.0) || !std::isfinite(lam)) return MPF_ERR_DIVERGED;
  *out = lam;
  return MPF_OK;
}

CsrView fine_csr(const mpf_problem* p) {
  CsrView A;
  A.n = (int64_t)p->pat.dispDOF + p->pat.presDOF;
  A.nnz = p->pat.nnz;
  A.rowptr = p->pat.rowptr.ptr;
  A.colidx = p->pat.colidx.ptr;
  A.vals = p->vals.ptr;
  return A;
}

// Transfer tables and the coarse pattern from the mesh (host, once per pattern).
mpf_status build_tables(mpf_problem* p, PmgWork& W) {
  hipStream_t s = p->stream;
  const int32_t nN = p->nNode, nE = p->nElem, npe = p->npe;
  const int32_t d = p->pat.dispDOF;
  if (npe != 10) return MPF_ERR_ARG;
  std::vector<int32_t> conn((size_t)nE * npe), nda((size_t)nN * 3);
  MPF_TRY(p->conn.download(conn.data(), conn.size(), s));
  MPF_TRY(p->pat.nda.download(nda.data(), nda.size(), s));
  MPF_HIP(hipStreamSynchronize(s));
  static const int kEdge[6][2] = {{0, 1}, {1, 2}, {0, 2}, {0, 3}, {2, 3}, {1, 3}};
  std::vector<char> isv(nN, 0);
  std::vector<int32_t> par((size_t)nN * 2, -1);
  for (int32_t e = 0; e < nE; ++e)
    for (int a = 0; a < 4; ++a) isv[conn[(size_t)e * npe + a]] = 1;
  for (int32_t e = 0; e < nE; ++e)
    for (int q = 0; q < 6; ++q) {
      const int32_t mnode = conn[(size_t)e * npe + 4 + q];
      if (par[2 * (size_t)mnode] < 0) {
        par[2 * (size_t)mnode] = conn[(size_t)e * npe + kEdge[q][0]];
        par[2 * (size_t)mnode + 1] = conn[(size_t)e * npe + kEdge[q][1]];
      }
    }
  // coarse numbering: free vertex dofs, vertex-major in node order
  std::vector<int32_t> cnum((size_t)nN * 3, -1), cbs, cbm, bvert;
  int64_t nc = 0;
  for (int32_t v = 0; v < nN; ++v) {
    if (!isv[v]) continue;
    const int64_t st = nc;
    for (int a = 0; a < 3; ++a)
      if (nda[3 * (size_t)v + a] >= 0) cnum[3 * (size_t)v + a] = (int32_t)nc++;
    if (nc > st) {
      cbs.push_back((int32_t)st);
      cbm.push_back((int32_t)(nc - st));
      bvert.push_back(v);
    }
  }
  if (nc == 0 || nc > INT32_MAX / 2) return MPF_ERR_ARG;
  // P: two (coarse dof, weight) slots per fine displacement row
  std::vector<int32_t> pidx((size_t)d * 2, -1);
  std::vector<double> pw((size_t)d * 2, 0.0);
  for (int32_t nd = 0; nd < nN; ++nd)
    for (int a = 0; a < 3; ++a) {
      const int32_t f = nda[3 * (size_t)nd + a];
      if (f < 0 || f >= d) continue;
      if (isv[nd]) {
        pidx[2 * (size_t)f] = cnum[3 * (size_t)nd + a];
        pw[2 * (size_t)f] = 1.0;
      } else if (par[2 * (size_t)nd] >= 0) {
        for (int t = 0; t < 2; ++t) {
          const int32_t c = cnum[3 * (size_t)par[2 * (size_t)nd + t] + a];
          if (c >= 0) {
            pidx[2 * (size_t)f + t] = c;
            pw[2 * (size_t)f + t] = 0.5;
          }
        }
      }
    }
  // P^T as CSR over the coarse rows (fine rows ascending within a row)
  std::vector<int32_t> rptr(nc + 1, 0);
  for (size_t k = 0; k < pidx.size(); ++k)
    if (pidx[k] >= 0) ++rptr[pidx[k] + 1];
  for (int64_t i = 0; i < nc; ++i) rptr[i + 1] += rptr[i];
  std::vector<int32_t> rfine(rptr[nc]), fill(rptr.begin(), rptr.end() - 1);
  std::vector<double> rw(rptr[nc]);
  for (int32_t f = 0; f < d; ++f)
    for (int t = 0; t < 2; ++t) {
      const int32_t c = pidx[2 * (size_t)f + t];
      if (c < 0) continue;
      rfine[fill[c]] = f;
      rw[fill[c]++] = pw[2 * (size_t)f + t];
    }
  // coarse pattern: vertices that share an element
  std::vector<int32_t> vptr(nN + 1, 0);
  for (int32_t e = 0; e < nE; ++e)
    for (int a = 0; a < 4; ++a) ++vptr[conn[(size_t)e * npe + a] + 1];
  for (int32_t i = 0; i < nN; ++i) vptr[i + 1] += vptr[i];
  std::vector<int32_t> vel(vptr[nN]), vfill(vptr.begin(), vptr.end() - 1);
  for (int32_t e = 0; e < nE; ++e)
    for (int a = 0; a < 4; ++a) vel[vfill[conn[(size_t)e * npe + a]]++] = e;
  // (each list is gathered without repeats through a marker per vertex, then sorted: the same sorted
  // sets as sorting the gathered lists with their repeats, at a fraction of the host time)
  std::vector<int32_t> crow(1, 0), ccol, nbr, row, mark(nN, -1);
  crow.reserve(nc + 1);
  for (int32_t v = 0; v < nN; ++v) {
    if (!isv[v]) continue;
    int nfree = 0;
    for (int a = 0; a < 3; ++a) nfree += cnum[3 * (size_t)v + a] >= 0;
    if (!nfree) continue;
    nbr.clear();
    for (int32_t k = vptr[v]; k < vptr[v + 1]; ++k)
      for (int a = 0; a < 4; ++a) {
        const int32_t w = conn[(size_t)vel[k] * npe + a];
        if (mark[w] != v) {
          mark[w] = v;
          nbr.push_back(w);
        }
      }
    std::sort(nbr.begin(), nbr.end());
    row.clear();
    for (int32_t w : nbr)
      for (int a = 0; a < 3; ++a)
        if (cnum[3 * (size_t)w + a] >= 0) row.push_back(cnum[3 * (size_t)w + a]);
    for (int i = 0; i < nfree; ++i) {
      ccol.insert(ccol.end(), row.begin(), row.end());
      crow.push_back((int32_t)ccol.size());
    }
    if (ccol.size() > (size_t)INT32_MAX) return MPF_ERR_ARG;
  }
  // third level: aggregates of 8 vertex blocks consecutive in Morton order of the vertices (on a
  // lattice of about the vertex spacing), one aggregate dof per aggregate and component present,
  // numbered aggregate by aggregate
  const int64_t ncb = (int64_t)cbs.size();
  std::vector<int32_t> aoff(nc), amptr, amem(nc), karow(1, 0), kacol, kabs, kabm;
  {
    std::vector<double> Xh((size_t)nN * 3);
    MPF_TRY(p->X.download(Xh.data(), Xh.size(), s));
    MPF_HIP(hipStreamSynchronize(s));
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int64_t b = 0; b < ncb; ++b)
      for (int c = 0; c < 3; ++c) {
        lo[c] = std::min(lo[c], Xh[3 * (size_t)bvert[b] + c]);
        hi[c] = std::max(hi[c], Xh[3 * (size_t)bvert[b] + c]);
      }
    const double ext = std::max(hi[0] - lo[0], std::max(hi[1] - lo[1], hi[2] - lo[2]));
    double vol = 1.0;
    for (int c = 0; c < 3; ++c) vol *= std::max(hi[c] - lo[c], ext * 1e-3);
    double h = ncb > 0 ? std::cbrt(vol / (double)ncb) : 1.0;
    if (!(h > 0.0)) h = 1.0;
    h = std::max(h, ext / 2097151.0);
    auto spread = [](uint64_t v) {
      uint64_t o = 0;
      for (int k = 0; k < 21; ++k) o |= ((v >> k) & 1ull) << (3 * k);
      return o;
    };
    std::vector<uint64_t> code(ncb);
    for (int64_t b = 0; b < ncb; ++b) {
      uint64_t q[3];
      for (int c = 0; c < 3; ++c) {
        const double t = std::floor((Xh[3 * (size_t)bvert[b] + c] - lo[c]) / h);
        q[c] = (uint64_t)std::min(std::max(t, 0.0), 2097151.0);
      }
      code[b] = spread(q[0]) | (spread(q[1]) << 1) | (spread(q[2]) << 2);
    }
    std::vector<int64_t> order(ncb);
    for (int64_t b = 0; b < ncb; ++b) order[b] = b;
    std::stable_sort(order.begin(), order.end(), [&](int64_t x, int64_t y) { return code[x] < code[y]; });
    std::vector<int64_t> agg(ncb);
    for (int64_t r = 0; r < ncb; ++r) agg[order[r]] = r / 8;
    const int64_t nagg = (ncb + 7) / 8;
    std::vector<char> present((size_t)nagg * 3, 0);
    for (int64_t b = 0; b < ncb; ++b)
      for (int a = 0; a < 3; ++a)
        if (cnum[3 * (size_t)bvert[b] + a] >= 0) present[3 * (size_t)agg[b] + a] = 1;
    std::vector<int32_t> adof((size_t)nagg * 3, -1);
    int64_t na = 0;
    for (int64_t g = 0; g < nagg; ++g) {
      const int64_t st = na;
      for (int a = 0; a < 3; ++a)
        if (present[3 * (size_t)g + a]) adof[3 * (size_t)g + a] = (int32_t)na++;
      if (na > st) {
        kabs.push_back((int32_t)st);
        kabm.push_back((int32_t)(na - st));
      }
    }
    for (int64_t b = 0; b < ncb; ++b)
      for (int a = 0; a < 3; ++a) {
        const int32_t i = cnum[3 * (size_t)bvert[b] + a];
        if (i >= 0) aoff[i] = adof[3 * (size_t)agg[b] + a];
      }
    amptr.assign(na + 1, 0);
    for (int64_t i = 0; i < nc; ++i) ++amptr[aoff[i] + 1];
    for (int64_t A = 0; A < na; ++A) amptr[A + 1] += amptr[A];
    std::vector<int32_t> afill(amptr.begin(), amptr.end() - 1);
    for (int64_t i = 0; i < nc; ++i) amem[afill[aoff[i]]++] = (int32_t)i;
    // a vertex dof's Kc row depends on its vertex only, so an aggregate dof whose members sit on
    // the same vertices as the previous one's (another component of the aggregate) has its columns
    std::vector<int32_t> cols, amark(na, -1), dvert(nc), mv, mvp;
    for (int64_t bb = 0; bb < ncb; ++bb)
      for (int t = 0; t < cbm[bb]; ++t) dvert[cbs[bb] + t] = bvert[bb];
    for (int64_t A = 0; A < na; ++A) {
      mv.clear();
      for (int32_t t = amptr[A]; t < amptr[A + 1]; ++t) mv.push_back(dvert[amem[t]]);
      if (A > 0 && mv == mvp) {
        kacol.insert(kacol.end(), cols.begin(), cols.end());
        karow.push_back((int32_t)kacol.size());
        if (kacol.size() > (size_t)INT32_MAX) return MPF_ERR_ARG;
        continue;
      }
      mvp.swap(mv);
      cols.clear();
      for (int32_t t = amptr[A]; t < amptr[A + 1]; ++t) {
        const int32_t i = amem[t];
        for (int32_t k = crow[i]; k < crow[i + 1]; ++k) {
          const int32_t c = aoff[ccol[k]];
          if (amark[c] != (int32_t)A) {
            amark[c] = (int32_t)A;
            cols.push_back(c);
          }
        }
      }
      std::sort(cols.begin(), cols.end());
      kacol.insert(kacol.end(), cols.begin(), cols.end());
      karow.push_back((int32_t)kacol.size());
      if (kacol.size() > (size_t)INT32_MAX) return MPF_ERR_ARG;
    }
    W.na = na;
    W.nka = (int64_t)kacol.size();
    W.nab = (int64_t)kabs.size();
  }
  std::vector<int32_t> culen(ncb);
  for (int64_t b = 0; b < ncb; ++b) culen[b] = crow[cbs[b] + 1] - crow[cbs[b]];
  W.d = d;
  W.pn = p->pat.presDOF;
  W.nc = nc;
  W.nnzc = (int64_t)ccol.size();
  W.ncb = ncb;
  MPF_TRY(W.pidx.upload(pidx.data(), pidx.size(), s));
  MPF_TRY(W.pw.upload(pw.data(), pw.size(), s));
  MPF_TRY(W.rptr.upload(rptr.data(), rptr.size(), s));
  MPF_TRY(W.rfine.upload(rfine.data(), rfine.size(), s));
  MPF_TRY(W.rw.upload(rw.data(), rw.size(), s));
  MPF_TRY(W.crow.upload(crow.data(), crow.size(), s));
  MPF_TRY(W.ccol.upload(ccol.data(), ccol.size(), s));
  MPF_TRY(W.cbs.upload(cbs.data(), cbs.size(), s));
  MPF_TRY(W.cbm.upload(cbm.data(), cbm.size(), s));
  MPF_TRY(W.cval.alloc(W.nnzc));
  MPF_TRY(W.cval32.alloc(W.nnzc));
  MPF_TRY(W.crblk.alloc(nc));
  MPF_TRY(W.cbinv.alloc(9 * (size_t)W.ncb));
  MPF_TRY(W.sinv.alloc(W.pn > 0 ? W.pn : 1));
  for (DevBuf<double>* b : {&W.t, &W.xa, &W.xb, &W.dd, &W.res}) MPF_TRY(b->alloc(d > 0 ? d : 1));
  for (DevBuf<double>* b : {&W.rc, &W.xc, &W.cr, &W.cr2, &W.cz, &W.cp, &W.cq}) MPF_TRY(b->alloc(nc));
  MPF_TRY(W.aoff.upload(aoff.data(), aoff.size(), s));
  MPF_TRY(W.amptr.upload(amptr.data(), amptr.size(), s));
  MPF_TRY(W.amem.upload(amem.data(), amem.size(), s));
  MPF_TRY(W.karow.upload(karow.data(), karow.size(), s));
  MPF_TRY(W.kacol.upload(kacol.data(), kacol.size(), s));
  MPF_TRY(W.kabs.upload(kabs.data(), kabs.size(), s));
  MPF_TRY(W.kabm.upload(kabm.data(), kabm.size(), s));
  MPF_TRY(W.culen.upload(culen.data(), culen.size(), s));
  MPF_TRY(W.kaval.alloc(W.nka > 0 ? W.nka : 1));
  MPF_TRY(W.kaval32.alloc(W.nka > 0 ? W.nka : 1));
  MPF_TRY(W.karblk.alloc(W.na > 0 ? W.na : 1));
  MPF_TRY(W.kabinv.alloc(9 * (size_t)(W.nab > 0 ? W.nab : 1)));
  for (DevBuf<double>* b : {&W.cxa, &W.cxb, &W.cdd, &W.cres}) MPF_TRY(b->alloc(nc));
  for (DevBuf<double>* b : {&W.ra, &W.xa3}) MPF_TRY(b->alloc(W.na > 0 ? W.na : 1));
  MPF_TRY(W.cpart.alloc(3 * (size_t)kMaxPartials));  // coarse CG: 2 per workgroup; power steps: 3
  MPF_TRY(W.csc.alloc(kCgScalars));
  MPF_TRY(W.bad.alloc(1));
  CsrView C;
  C.n = nc;
  C.nnz = W.nnzc;
  C.rowptr = W.crow.ptr;
  C.colidx = W.ccol.ptr;
  MPF_TRY(block_pc_structure(C, W.ncb, W.cbs.ptr, W.cbm.ptr, W.crblk.ptr, &W.cgrp, s));  // synchronises
  CsrView Ka;
  Ka.n = W.na;
  Ka.nnz = W.nka;
  Ka.rowptr = W.karow.ptr;
  Ka.colidx = W.kacol.ptr;
  if (W.na > 0) MPF_TRY(block_pc_structure(Ka, W.nab, W.kabs.ptr, W.kabm.ptr, W.karblk.ptr, &W.kagrp, s));
  W.rowptr_seen = p->pat.rowptr.ptr;
  W.built = true;
  W.ulen_ready = false;
  return MPF_OK;
}

CsrView coarse_csr(const PmgWork& W) {
  CsrView C;
  C.n = W.nc;
  C.nnz = W.nnzc;
  C.rowptr = W.crow.ptr;
  C.colidx = W.ccol.ptr;
  C.vals = W.cval.ptr;
  C.vals32 = W.cval32.ptr;
  return C;
}

BlockPcView coarse_blocks(const PmgWork& W) {
  BlockPcView C;
  C.nblk = W.ncb;
  C.bs = W.cbs.ptr;
  C.bm = W.cbm.ptr;
  C.rblk = W.crblk.ptr;
  C.binv = W.cbinv.ptr;
  C.grp = W.cgrp;
  return C;
}

CsrView agg_csr(const PmgWork& W) {
  CsrView K;
  K.n = W.na;
  K.nnz = W.nka;
  K.rowptr = W.karow.ptr;
  K.colidx = W.kacol.ptr;
  K.vals = W.kaval.ptr;
  K.vals32 = W.kaval32.ptr;
  return K;
}

BlockPcView agg_blocks(const PmgWork& W) {
  BlockPcView K;
  K.nblk = W.nab;
  K.bs = W.kabs.ptr;
  K.bm = W.kabm.ptr;
  K.rblk = W.karblk.ptr;
  K.binv = W.kabinv.ptr;
  K.grp = W.kagrp;
  return K;
}

int lanes_coarse(const PmgWork& W) {
  const double avg = W.nc > 0 ? (double)W.nnzc / (double)W.nc : 0.0;
  return avg > 48 ? 32 : (avg > 12 ? 16 : 4);
}

// Chebyshev step coefficients (three-term recurrence of the corrections) on [lambda / ratio,
// lambda] widened by 10 %: step k takes dd = c1[k] dd + c2[k] B^-1 (t - K x), x += dd.
void cheb_coefficients(double lambda, double ratio, int degree, double* c1, double* c2) {
  const double beta = 1.1 * lambda, alpha = beta / ratio;
  const double theta = 0.5 * (beta + alpha), delta = 0.5 * (beta - alpha), sigma = theta / delta;
  double rho_old = 1.0 / sigma;
  c1[0] = 0.0;
  c2[0] = 1.0 / theta;
  for (int k = 1; k < degree; ++k) {
    const double rho = 1.0 / (2.0 * sigma - rho_old);
    c1[k] = rho * rho_old;
    c2[k] = 2.0 * rho / delta;
    rho_old = rho;
  }
}

int lanes_for(const mpf_problem* p) {
  const int64_t n = (int64_t)p->pat.dispDOF + p->pat.presDOF;
  const double avg = n > 0 ? (double)p->pat.nnz / (double)n : 0.0;
  return avg > 48 ? 32 : (avg > 12 ? 16 : 4);
}

}  // namespace

mpf_status pmg_setup(mpf_problem* p, const BlockPcView& F, int64_t nub) {
  const auto t0 = std::chrono::steady_clock::now();
  hipStream_t s = p->stream;
  if (!p->pmg) p->pmg = new PmgWork();
  PmgWork& W = *p->pmg;
  if (!W.built || W.rowptr_seen != p->pat.rowptr.ptr || W.d != p->pat.dispDOF || W.pn != p->pat.presDOF)
    MPF_TRY(build_tables(p, W));
  W.vcycles = 0;
  W.coarse_its = 0;
  const CsrView A = fine_csr(p);
  int32_t hbad = 0;
  MPF_HIP(hipMemcpyAsync(W.bad.ptr, &hbad, sizeof(int32_t), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_galerkin, dim3(grid_for(W.nc, kGT)), dim3(kGT), 0, s, W.nc, W.rptr.ptr, W.rfine.ptr, W.rw.ptr,
                     A.rowptr, A.colidx, A.vals, (int32_t)W.d, W.pidx.ptr, W.pw.ptr, W.crow.ptr, W.ccol.ptr,
                     W.cval.ptr, W.bad.ptr);
  MPF_TRY(block_pc_values(coarse_csr(W), W.ncb, W.cbs.ptr, W.cbm.ptr, W.cbinv.ptr, s));
  hipLaunchKernelGGL(k_to_f32, dim3(grid_for(W.nnzc, kT)), dim3(kT), 0, s, W.nnzc, W.cval.ptr, W.cval32.ptr);
  if (W.fval32.n != (size_t)A.nnz) MPF_TRY(W.fval32.alloc(A.nnz > 0 ? (size_t)A.nnz : 1));
  hipLaunchKernelGGL(k_to_f32, dim3(grid_for(A.nnz, kT)), dim3(kT), 0, s, A.nnz, A.vals, W.fval32.ptr);
  if (!W.ulen_ready || W.ulen.n != (size_t)nub) {
    MPF_TRY(W.ulen.alloc(nub > 0 ? nub : 1));
    hipLaunchKernelGGL(k_ulen, dim3(grid_for(nub, kT)), dim3(kT), 0, s, nub, F.bs, A.rowptr, A.colidx, (int32_t)W.d,
                       W.ulen.ptr);
    W.ulen_ready = true;
  }
  MPF_TRY(estimate_lambda(s, A, F, W.ulen.ptr, nub, W.d, lanes_for(p), W.xa.ptr, W.xb.ptr, W.cpart.ptr, &W.lambda));
  if (W.levels == 3 && W.na > 0) {  // aggregate matrix, its blocks, and the vertex level's smoothing interval
    hipLaunchKernelGGL(k_agg_values, dim3(grid_for(W.na, kT)), dim3(kT), 0, s, W.na, W.amptr.ptr, W.amem.ptr,
                       W.crow.ptr, W.ccol.ptr, W.cval.ptr, W.aoff.ptr, W.karow.ptr, W.kacol.ptr, W.kaval.ptr,
                       W.bad.ptr);
    MPF_TRY(block_pc_values(agg_csr(W), W.nab, W.kabs.ptr, W.kabm.ptr, W.kabinv.ptr, s));
    hipLaunchKernelGGL(k_to_f32, dim3(grid_for(W.nka, kT)), dim3(kT), 0, s, W.nka, W.kaval.ptr, W.kaval32.ptr);
    MPF_TRY(estimate_lambda(s, coarse_csr(W), coarse_blocks(W), W.culen.ptr, W.ncb, W.nc, lanes_coarse(W), W.cxa.ptr,
                            W.cxb.ptr, W.cpart.ptr, &W.lambda_c));
  }
  if (W.pn > 0) {
    const double scale = (p->m.Kinv > 0.0 && p->m.mu > 0.0) ? 1.0 + 1.0 / (p->m.mu * p->m.Kinv) : 0.0;
    hipLaunchKernelGGL(k_sinv, dim3(grid_for(W.pn, kT)), dim3(kT), 0, s, W.pn, (int32_t)W.d, A.rowptr, A.colidx,
                       A.vals, scale, F.binv + 9 * nub, W.sinv.ptr);
  }
  MPF_HIP(hipGetLastError());
  MPF_HIP(hipMemcpyAsync(&hbad, W.bad.ptr, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  MPF_HIP(hipStreamSynchronize(s));
  if (hbad) return MPF_ERR_ARG;
  W.setup_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return MPF_OK;
}

mpf_status pmg_apply(mpf_problem* p, const BlockPcView& F, int64_t nub, const double* v, double* z) {
  if (!p->pmg || !p->pmg->built) return MPF_ERR_STATE;
  PmgWork& W = *p->pmg;
  hipStream_t s = p->stream;
  const int32_t d = (int32_t)W.d;
  const CsrView A = fine_csr(p);
  const int lpr = lanes_for(p);
  const double* t = v;  // right-hand side of the V-cycle: v_u - Kup z_p (written by k_jac0)
  if (W.pn > 0) {
    hipLaunchKernelGGL(k_pres, dim3(grid_for(W.pn, kT)), dim3(kT), 0, s, W.pn, v + d, W.sinv.ptr, z + d);
    t = W.t.ptr;
  }
  double c1[8], c2[8];
  cheb_coefficients(W.lambda, W.ratio, W.degree, c1, c2);
  // pre-smoothing from x = 0
  double* cur = W.xa.ptr;
  double* oth = W.xb.ptr;
  hipLaunchKernelGGL(k_jac0<16>, dim3(grid_for(nub, kT / 16)), dim3(kT), 0, s, nub, F.bs, F.bm, F.binv, A.rowptr,
                     A.colidx, A.vals, W.ulen.ptr, d, W.pn > 0 ? z + d : nullptr, c2[0], v, W.t.ptr, cur, W.dd.ptr);
  for (int k = 1; k < W.degree; ++k) {
    launch_sweep<1>(lpr, nub, F, A, W.fval32.ptr, W.ulen.ptr, c1[k], c2[k], t, cur, oth, W.dd.ptr, s);
    std::swap(cur, oth);
  }
  // coarse correction
  launch_sweep<2>(lpr, nub, F, A, W.fval32.ptr, W.ulen.ptr, 0.0, 0.0, t, cur, W.res.ptr, nullptr, s);
  hipLaunchKernelGGL(k_restrict, dim3(grid_for(W.nc, kT)), dim3(kT), 0, s, W.nc, W.rptr.ptr, W.rfine.ptr, W.rw.ptr,
                     W.res.ptr, W.rc.ptr);
  const CgScratch cw{W.cr.ptr, W.cr2.ptr, W.cz.ptr, W.cp.ptr, W.cq.ptr, W.cpart.ptr, W.csc.ptr};
  mpf_krylov_report cr{};
  if (W.levels == 3 && W.na > 0) {
    // the vertex level's own V-cycle: Chebyshev smoothing from 0, the aggregates' correction (Ka
    // by block-Jacobi CG to crtol), Chebyshev smoothing again, the result in xc
    const BlockPcView Cb = coarse_blocks(W);
    const CsrView Cv = coarse_csr(W);
    const int lc = lanes_coarse(W);
    const int64_t nc = W.nc, ncb = W.ncb;
    double e1[8], e2[8];
    cheb_coefficients(W.lambda_c, W.ratio, W.degree, e1, e2);
    double* xc0 = W.cxa.ptr;
    double* xc1 = W.cxb.ptr;
    hipLaunchKernelGGL(k_zero, dim3(grid_for(nc, kT)), dim3(kT), 0, s, nc, xc0);
    for (int k = 0; k < W.degree; ++k) {
      launch_sweep<1>(lc, ncb, Cb, Cv, W.cval32.ptr, W.culen.ptr, e1[k], e2[k], W.rc.ptr, xc0, xc1, W.cdd.ptr, s);
      std::swap(xc0, xc1);
    }
    launch_sweep<2>(lc, ncb, Cb, Cv, W.cval32.ptr, W.culen.ptr, 0.0, 0.0, W.rc.ptr, xc0, W.cres.ptr, nullptr, s);
    hipLaunchKernelGGL(k_agg_restrict, dim3(grid_for(W.na, kT)), dim3(kT), 0, s, W.na, W.amptr.ptr, W.amem.ptr,
                       W.cres.ptr, W.ra.ptr);
    MPF_TRY(pcg_csr(agg_csr(W), agg_blocks(W), cw, W.ra.ptr, W.xa3.ptr, W.crtol, W.cmaxit, s, &cr));
    hipLaunchKernelGGL(k_agg_prolong, dim3(grid_for(nc, kT)), dim3(kT), 0, s, nc, W.aoff.ptr, W.xa3.ptr, xc0);
    for (int k = 0; k < W.degree; ++k) {
      double* out = (k == W.degree - 1) ? W.xc.ptr : xc1;
      launch_sweep<1>(lc, ncb, Cb, Cv, W.cval32.ptr, W.culen.ptr, e1[k], e2[k], W.rc.ptr, xc0, out, W.cdd.ptr, s);
      xc1 = xc0;
      xc0 = out;
    }
  } else {
    MPF_TRY(pcg_csr(coarse_csr(W), coarse_blocks(W), cw, W.rc.ptr, W.xc.ptr, W.crtol, W.cmaxit, s, &cr));
  }
  W.coarse_its += cr.iterations;
  hipLaunchKernelGGL(k_prolong, dim3(grid_for(d, kT)), dim3(kT), 0, s, d, W.pidx.ptr, W.pw.ptr, W.xc.ptr, cur);
  // post-smoothing from the corrected x, the last step into z
  for (int k = 0; k < W.degree; ++k) {
    double* out = (k == W.degree - 1) ? z : oth;
    launch_sweep<1>(lpr, nub, F, A, W.fval32.ptr, W.ulen.ptr, c1[k], c2[k], t, cur, out, W.dd.ptr, s);
    oth = cur;
    cur = out;
  }
  MPF_HIP(hipGetLastError());
  ++W.vcycles;
  return MPF_OK;
}

void pmg_free(mpf_problem* p) {
  delete p->pmg;
  p->pmg = nullptr;
}

}  // namespace mpf

using namespace mpf;

extern "C" mpf_status mpf_set_pmg(mpf_problem* p, int32_t degree, double ratio, double coarse_rtol,
                                  int32_t coarse_maxit) {
  if (!p || degree < 1 || degree > 8 || !(ratio > 1.0 && ratio < 1e6) || !(coarse_rtol > 0.0 && coarse_rtol < 1.0) ||
      coarse_maxit < 1)
    return MPF_ERR_ARG;
  if (!p->pmg) p->pmg = new PmgWork();
  p->pmg->degree = degree;
  p->pmg->ratio = ratio;
  p->pmg->crtol = coarse_rtol;
  p->pmg->cmaxit = coarse_maxit;
  return MPF_OK;
}

extern "C" mpf_status mpf_pmg_stats(const mpf_problem* p, mpf_pmg_report* rep) {
  if (!p || !rep) return MPF_ERR_ARG;
  if (!p->pmg || !p->pmg->built) return MPF_ERR_STATE;
  rep->coarse_dofs = p->pmg->nc;
  rep->coarse_nnz = p->pmg->nnzc;
  rep->vcycles = p->pmg->vcycles;
  rep->coarse_iterations = p->pmg->coarse_its;
  rep->setup_seconds = p->pmg->setup_seconds;
  rep->lambda_max = p->pmg->lambda;
  const bool three = p->pmg->levels == 3 && p->pmg->na > 0;
  rep->lambda_vertex = three ? p->pmg->lambda_c : 0.0;
  rep->agg_dofs = p->pmg->na;
  rep->agg_nnz = p->pmg->nka;
  rep->levels = p->pmg->levels;
  rep->degree = p->pmg->degree;
  return MPF_OK;
}

extern "C" mpf_status mpf_set_pmg_levels(mpf_problem* p, int32_t levels) {
  if (!p || (levels != 2 && levels != 3)) return MPF_ERR_ARG;
  if (!p->pmg) p->pmg = new PmgWork();
  p->pmg->levels = levels;
  return MPF_OK;
}
