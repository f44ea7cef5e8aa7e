// krylov.hip -- FP64 CSR SpMV, deterministic reductions, Jacobi / node-block-Jacobi
// preconditioners, PCG, GMRES(m), and flexible GMRES(m) around the two-level preconditioner of
// pmg.hip, on gfx950.
//
// Stands in for PETSc's KSPSolve behind SolverPetsc::factoriseAndSolve (SolverPetsc.cpp:250-330):
// zero initial guess (VecZeroEntries(solnVec), :290), stop on ||b - A x|| / ||b|| < rtol.
// The reference selects KSP/PC from petsc_options (code default KSPCG + PCILU, :117-140); the
// device path offers CG and GMRES with Jacobi or 3x3 node block-Jacobi.
//
// Reductions are two-stage (per-block partials, then one block sums them in a fixed order), so
// results are bitwise reproducible run to run.  PCG keeps all scalars on the device; the host
// enqueues iterations in chunks and reads the convergence flag between chunks.

#include <cmath>
#include <chrono>

#include "mpf_internal.hpp"

namespace mpf {

constexpr int kRB = 256;        // threads per block for vector kernels
constexpr int kMaxParts = kMaxPartials;  // partial sums per reduction (4096 workgroups: a full GPU of waves)
constexpr int kMaxRestart = 200;  // GMRES(m) upper bound (Hessenberg column scratch on the device)

struct KrylovWork {
  int64_t n = 0;
  DevBuf<double> r, z, p, q, w;
  DevBuf<double> r2;          // PCG residual, the other half of the pair (alternates by iteration)
  DevBuf<double> part;        // partial sums [kMaxParts * 64]
  DevBuf<double> sc;          // device scalars
  DevBuf<double> V;           // GMRES basis [(m+1) * n]
  int vcap = 0;
  DevBuf<double> Z;           // flexible GMRES: preconditioned basis [m * n]
  int zcap = 0;
  // preconditioner blocks
  DevBuf<int32_t> blk_start;
  DevBuf<int32_t> blk_size;
  DevBuf<int32_t> row_blk;    // block of each row
  DevBuf<double> blk_inv;
  int64_t nblk = 0;
  int pc_kind = -1;
  bool grp_spmv = false;  // every node block's rows share one column list (checked once)
};

namespace {

// scalar slots
enum { S_RZ0 = 0, S_RZ1 = 1, S_PQ = 2, S_RR = 3, S_BB = 4, S_DONE = 5, S_ITS = 6, S_TOL2 = 7, S_MAXIT = 8, S_N = 16 };
static_assert(S_N == kCgScalars, "CG scalar slots");

__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// block-wide sum of one value; result valid in thread 0
__device__ __forceinline__ double block_sum(double v, double* sh) {
  v = wave_sum(v);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += sh[i];
  __syncthreads();
  return s;
}

unsigned nparts_for(int64_t n) {
  int64_t g = (n + kRB - 1) / kRB;
  if (g > kMaxParts) g = kMaxParts;
  if (g < 1) g = 1;
  return (unsigned)g;
}

// y = A x ; partial[b] = sum over the block's rows of x_r y_r (when part != null).  V: the type the
// matrix values are stored in (double, or float for a preconditioner's operator); products and
// sums in double.
template <int LPR, typename V = double>
__global__ void __launch_bounds__(kRB) k_spmv(int64_t n, const int32_t* __restrict__ rowptr,
                                              const int32_t* __restrict__ colidx, const V* __restrict__ vals,
                                              const double* __restrict__ x, double* __restrict__ y,
                                              double* __restrict__ part, const double* __restrict__ sc) {
  if (sc && sc[S_DONE] != 0.0) return;
  __shared__ double sh[kRB / 64];
  const int g = threadIdx.x / LPR, l = threadIdx.x % LPR;
  constexpr int RPB = kRB / LPR;
  double dot = 0.0;
  for (int64_t r0 = (int64_t)blockIdx.x * RPB; r0 < n; r0 += (int64_t)gridDim.x * RPB) {
    const int64_t r = r0 + g;
    double s = 0.0;
    if (r < n) {
      const int b = rowptr[r], e = rowptr[r + 1];
#pragma unroll 4
      for (int k = b + l; k < e; k += LPR) s += (double)vals[k] * x[colidx[k]];
    }
#pragma unroll
    for (int off = LPR / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, LPR);
    if (r < n && l == 0) {
      y[r] = s;
      if (part) dot += x[r] * s;
    }
  }
  if (part) {
    const double t = block_sum(dot, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
  }
}

// y = A x over row groups: block k holds rows bs[k] .. bs[k] + bm[k] - 1 (a node's free dofs, or one
// pressure row), which share one column list, so each column index and x value is read once for
// up to three rows.  Entries go to the LPR lanes as in k_spmv (same per-row summation order);
// part[b] = sum over the block's rows of x_r y_r (when part != null).
template <int LPR, typename V = double>
__global__ void __launch_bounds__(kRB) k_spmv_grp(int64_t nblk, const int32_t* __restrict__ bs,
                                                  const int32_t* __restrict__ bm, const int32_t* __restrict__ rowptr,
                                                  const int32_t* __restrict__ colidx, const V* __restrict__ vals,
                                                  const double* __restrict__ x, double* __restrict__ y,
                                                  double* __restrict__ part, const double* __restrict__ sc) {
  if (sc && sc[S_DONE] != 0.0) return;
  __shared__ double sh[kRB / 64];
  const int g = threadIdx.x / LPR, l = threadIdx.x % LPR;
  constexpr int GPB = kRB / LPR;
  double dot = 0.0;
  for (int64_t k0 = (int64_t)blockIdx.x * GPB; k0 < nblk; k0 += (int64_t)gridDim.x * GPB) {
    const int64_t k = k0 + g;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    int st = 0, m = 0;
    if (k < nblk) {
      st = bs[k];
      m = bm[k];
      const int b0 = rowptr[st], len = rowptr[st + 1] - b0;
      const int b1 = m > 1 ? rowptr[st + 1] : b0, b2 = m > 2 ? rowptr[st + 2] : b0;
#pragma unroll 4
      for (int j = l; j < len; j += LPR) {
        const double xv = x[colidx[b0 + j]];
        s0 += (double)vals[b0 + j] * xv;
        if (m > 1) s1 += (double)vals[b1 + j] * xv;
        if (m > 2) s2 += (double)vals[b2 + j] * xv;
      }
    }
#pragma unroll
    for (int off = LPR / 2; off > 0; off >>= 1) {
      s0 += __shfl_xor(s0, off, LPR);
      s1 += __shfl_xor(s1, off, LPR);
      s2 += __shfl_xor(s2, off, LPR);
    }
    if (k < nblk && l == 0) {
      y[st] = s0;
      if (m > 1) y[st + 1] = s1;
      if (m > 2) y[st + 2] = s2;
      if (part) {
        dot += x[st] * s0;
        if (m > 1) dot += x[st + 1] * s1;
        if (m > 2) dot += x[st + 2] * s2;
      }
    }
  }
  if (part) {
    const double t = block_sum(dot, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
  }
}

// bad[0] = 1 if some block's rows do not share one column list (one thread per block)
__global__ void k_check_groups(int64_t nblk, const int32_t* __restrict__ bs, const int32_t* __restrict__ bm,
                               const int32_t* __restrict__ rowptr, const int32_t* __restrict__ colidx,
                               int32_t* __restrict__ bad) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nblk) return;
  const int st = bs[k], m = bm[k];
  const int b0 = rowptr[st], len = rowptr[st + 1] - b0;
  bool same = m >= 1 && m <= 3;
  for (int i = 1; same && i < m; ++i) {
    const int bi = rowptr[st + i];
    if (rowptr[st + i + 1] - bi != len) { same = false; break; }
    for (int j = 0; j < len; ++j)
      if (colidx[bi + j] != colidx[b0 + j]) { same = false; break; }
  }
  if (!same) bad[0] = 1;
}

// out[j] = sum_i part[j*stride + i], i < np (fixed order)
__global__ void __launch_bounds__(kRB) k_finish(const double* __restrict__ part, int np, int stride, int nout,
                                                double* __restrict__ out) {
  __shared__ double sh[kRB / 64];
  for (int j = blockIdx.x; j < nout; j += gridDim.x) {  // one workgroup per output
    double s = 0.0;
    for (int i = threadIdx.x; i < np; i += kRB) s += part[(size_t)j * stride + i];
    const double t = block_sum(s, sh);
    if (threadIdx.x == 0) out[j] = t;
  }
}

// sum of squares partials
__global__ void __launch_bounds__(kRB) k_dot_part(int64_t n, const double* __restrict__ a, const double* __restrict__ b,
                                                  double* __restrict__ part) {
  __shared__ double sh[kRB / 64];
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kRB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kRB) s += a[i] * b[i];
  const double t = block_sum(s, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// preconditioner apply over blocks: z[s..s+m) = Binv r[s..s+m)
__device__ __forceinline__ void apply_block(int s, int m, const double* __restrict__ Bi, const double* __restrict__ r,
                                            double* __restrict__ z, double& rz, double& rr) {
  double rv[3] = {0, 0, 0};
  for (int i = 0; i < m; ++i) rv[i] = r[s + i];
  for (int i = 0; i < m; ++i) {
    double acc = 0.0;
    for (int j = 0; j < m; ++j) acc += Bi[3 * i + j] * rv[j];
    z[s + i] = acc;
    rz += rv[i] * acc;
    rr += rv[i] * rv[i];
  }
}

// PCG start: r = b, x = 0, z = M r, p = z, partials of r.z and b.b
__global__ void __launch_bounds__(kRB) k_cg_start(int64_t nblk, const int32_t* __restrict__ bs,
                                                  const int32_t* __restrict__ bm, const double* __restrict__ binv,
                                                  const double* __restrict__ b, double* __restrict__ x,
                                                  double* __restrict__ r, double* __restrict__ z,
                                                  double* __restrict__ p, double* __restrict__ part, int np) {
  __shared__ double sh[kRB / 64];
  double rz = 0.0, rr = 0.0;
  for (int64_t k = (int64_t)blockIdx.x * kRB + threadIdx.x; k < nblk; k += (int64_t)gridDim.x * kRB) {
    const int s = bs[k], m = bm[k];
    for (int i = 0; i < m; ++i) { r[s + i] = b[s + i]; x[s + i] = 0.0; }
    apply_block(s, m, binv + 9 * k, r, z, rz, rr);
    for (int i = 0; i < m; ++i) p[s + i] = z[s + i];
  }
  const double t1 = block_sum(rz, sh);
  const double t2 = block_sum(rr, sh);
  if (threadIdx.x == 0) { part[blockIdx.x] = t1; part[np + blockIdx.x] = t2; }
}

// One thread per row: x += alpha p ; r_out = r_in - alpha q ; z = M r_out with the row's 3x3 block
// (its block's r_in and q are read, only the row itself is written, so rows of one block never
// race); partials of r.z and r.r.   alpha = rz[par] / pq
__global__ void __launch_bounds__(kRB) k_cg_update_rows(int64_t n, const int32_t* __restrict__ rblk,
                                                        const int32_t* __restrict__ bs, const int32_t* __restrict__ bm,
                                                        const double* __restrict__ binv, double* __restrict__ x,
                                                        const double* __restrict__ rin, double* __restrict__ rout,
                                                        double* __restrict__ z, const double* __restrict__ p,
                                                        const double* __restrict__ q, double* __restrict__ part, int np,
                                                        const double* __restrict__ sc, int par) {
  if (sc[S_DONE] != 0.0) return;
  __shared__ double sh[kRB / 64];
  const double alpha = sc[S_RZ0 + par] / sc[S_PQ];
  double rz = 0.0, rr = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kRB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kRB) {
    x[i] += alpha * p[i];
    const int k = rblk[i], s0 = bs[k], m = bm[k], li = (int)(i - s0);
    if (li < 0 || li >= m || m > 3) {  // not reached: every row lies in its block
      rout[i] = rin[i] - alpha * q[i];
      z[i] = 0.0;
      continue;
    }
    double rv[3] = {0.0, 0.0, 0.0};
    for (int j = 0; j < m; ++j) rv[j] = rin[s0 + j] - alpha * q[s0 + j];
    const double* Bi = binv + 9 * (size_t)k + 3 * li;
    double zi = 0.0;
    for (int j = 0; j < m; ++j) zi += Bi[j] * rv[j];
    rout[i] = rv[li];
    z[i] = zi;
    rz += rv[li] * zi;
    rr += rv[li] * rv[li];
  }
  const double t1 = block_sum(rz, sh);
  const double t2 = block_sum(rr, sh);
  if (threadIdx.x == 0) { part[blockIdx.x] = t1; part[np + blockIdx.x] = t2; }
}

// row -> block of the preconditioner's block list
__global__ void k_row_blk(int64_t nblk, const int32_t* __restrict__ bs, const int32_t* __restrict__ bm,
                          int32_t* __restrict__ rblk) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nblk) return;
  for (int j = 0; j < bm[k]; ++j) rblk[bs[k] + j] = (int32_t)k;
}

// reduce the two partial arrays into rz[1-par], rr; convergence test
__global__ void __launch_bounds__(kRB) k_cg_check(const double* __restrict__ part, int np, double* __restrict__ sc,
                                                  int par, int it) {
  if (sc[S_DONE] != 0.0) return;
  __shared__ double sh[kRB / 64];
  double s1 = 0.0, s2 = 0.0;
  for (int i = threadIdx.x; i < np; i += kRB) { s1 += part[i]; s2 += part[np + i]; }
  const double t1 = block_sum(s1, sh);
  const double t2 = block_sum(s2, sh);
  if (threadIdx.x == 0) {
    sc[S_RZ0 + (1 - par)] = t1;
    sc[S_RR] = t2;
    sc[S_ITS] = (double)(it + 1);
    if (t2 <= sc[S_TOL2] || (double)(it + 1) >= sc[S_MAXIT] || !(t2 == t2)) sc[S_DONE] = 1.0;
  }
}

// p = z + beta p,  beta = rz[1-par] / rz[par]
__global__ void __launch_bounds__(kRB) k_cg_p(int64_t n, const double* __restrict__ z, double* __restrict__ p,
                                              const double* __restrict__ sc, int par) {
  if (sc[S_DONE] != 0.0) return;
  const double beta = sc[S_RZ0 + (1 - par)] / sc[S_RZ0 + par];
  for (int64_t i = (int64_t)blockIdx.x * kRB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kRB)
    p[i] = z[i] + beta * p[i];
}

__global__ void k_cg_init_scalars(double* __restrict__ sc, double rtol, int maxit) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    const double bb = sc[S_RR];
    sc[S_BB] = bb;
    sc[S_TOL2] = rtol * rtol * bb;
    sc[S_MAXIT] = (double)maxit;
    sc[S_DONE] = (bb == 0.0 || maxit <= 0) ? 1.0 : 0.0;
    sc[S_ITS] = 0.0;
  }
}

// Block-Jacobi setup: one thread per block; extracts the diagonal block from CSR and inverts it.
__device__ __forceinline__ double csr_entry(const int32_t* __restrict__ rowptr, const int32_t* __restrict__ colidx,
                                            const double* __restrict__ vals, int r, int c) {
  int lo = rowptr[r], hi = rowptr[r + 1];
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (colidx[mid] < c) lo = mid + 1; else hi = mid;
  }
  return (lo < rowptr[r + 1] && colidx[lo] == c) ? vals[lo] : 0.0;
}

__global__ void k_block_setup(int64_t nblk, const int32_t* __restrict__ bs, const int32_t* __restrict__ bm,
                              const int32_t* __restrict__ rowptr, const int32_t* __restrict__ colidx,
                              const double* __restrict__ vals, double* __restrict__ binv, int identity) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nblk) return;
  const int s = bs[k], m = bm[k];
  double A[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  double* Bi = binv + 9 * k;
  if (identity) {
    for (int i = 0; i < 9; ++i) Bi[i] = (i % 4 == 0) ? 1.0 : 0.0;
    return;
  }
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < m; ++j) A[i][j] = csr_entry(rowptr, colidx, vals, s + i, s + j);
  const double c00 = A[1][1] * A[2][2] - A[1][2] * A[2][1];
  const double c01 = A[1][2] * A[2][0] - A[1][0] * A[2][2];
  const double c02 = A[1][0] * A[2][1] - A[1][1] * A[2][0];
  const double det = A[0][0] * c00 + A[0][1] * c01 + A[0][2] * c02;
  const double id = 1.0 / det;
  Bi[0] = c00 * id;
  Bi[3] = c01 * id;
  Bi[6] = c02 * id;
  Bi[1] = (A[0][2] * A[2][1] - A[0][1] * A[2][2]) * id;
  Bi[4] = (A[0][0] * A[2][2] - A[0][2] * A[2][0]) * id;
  Bi[7] = (A[0][1] * A[2][0] - A[0][0] * A[2][1]) * id;
  Bi[2] = (A[0][1] * A[1][2] - A[0][2] * A[1][1]) * id;
  Bi[5] = (A[0][2] * A[1][0] - A[0][0] * A[1][2]) * id;
  Bi[8] = (A[0][0] * A[1][1] - A[0][1] * A[1][0]) * id;
}

// Pressure rows of the block-Jacobi preconditioner: instead of the bare Kpp diagonal (tiny for a
// nearly incompressible material), the diagonal of the Schur complement
//   S_qq = Kpp_qq - sum over u-node blocks n of  Kpu(q, n) A_n^-1 Kup(n, q),
// A_n the node's 3x3 Kuu block (already inverted in binv), Kup(n, q) taken as Kpu(q, n)^T (the
// mixed element's coupling blocks are transposes of each other).  With it, the preconditioner is
// blockdiag(A, S), whose preconditioned saddle-point operator has clustered eigenvalues.
__global__ void k_schur_pressure(int32_t presDOF, int32_t dispDOF, const int32_t* __restrict__ rowptr,
                                 const int32_t* __restrict__ colidx, const double* __restrict__ vals,
                                 const int32_t* __restrict__ bs, const int32_t* __restrict__ bm, int64_t nub,
                                 double* __restrict__ binv) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= presDOF) return;
  const int r = dispDOF + q;
  double S = 0.0;
  int k = rowptr[r];
  const int kend = rowptr[r + 1];
  while (k < kend) {
    const int c = colidx[k];
    if (c >= dispDOF) {
      if (c == r) S += vals[k];
      ++k;
      continue;
    }
    // node block holding column c (blocks sorted by first row)
    int64_t lo = 0, hi = nub - 1;
    while (lo < hi) {
      const int64_t mid = (lo + hi + 1) >> 1;
      if (bs[mid] <= c) lo = mid; else hi = mid - 1;
    }
    const int b0 = bs[lo], m = bm[lo];
    double b[3] = {0.0, 0.0, 0.0};
    int kk = k;
    while (kk < kend && colidx[kk] < b0 + m) {
      b[colidx[kk] - b0] = vals[kk];
      ++kk;
    }
    const double* Ai = binv + 9 * lo;
    double t = 0.0;
    for (int i = 0; i < m; ++i)
      for (int j = 0; j < m; ++j) t += b[i] * Ai[3 * i + j] * b[j];
    S -= t;
    k = kk;
  }
  double* Bi = binv + 9 * (nub + q);
  for (int i = 0; i < 9; ++i) Bi[i] = (i % 4 == 0) ? 1.0 : 0.0;
  Bi[0] = S != 0.0 ? 1.0 / S : 1.0;
}

// Block list: node blocks of free u dofs (consecutive rows), then 1x1 pressure rows.
__global__ void k_node_blocks(int32_t nNode, const int32_t* __restrict__ nda, const int32_t* __restrict__ first_blk,
                              int32_t* __restrict__ bs, int32_t* __restrict__ bm) {
  const int nd = blockIdx.x * blockDim.x + threadIdx.x;
  if (nd >= nNode) return;
  int m = 0, s = -1;
  for (int d = 0; d < 3; ++d) {
    const int r = nda[3 * (size_t)nd + d];
    if (r >= 0) { if (s < 0) s = r; ++m; }
  }
  if (m == 0) return;
  const int k = first_blk[nd];
  bs[k] = s;
  bm[k] = m;
}

__global__ void k_has_free(int32_t nNode, const int32_t* __restrict__ nda, int32_t* __restrict__ f) {
  const int nd = blockIdx.x * blockDim.x + threadIdx.x;
  if (nd >= nNode) return;
  f[nd] = (nda[3 * (size_t)nd] >= 0 || nda[3 * (size_t)nd + 1] >= 0 || nda[3 * (size_t)nd + 2] >= 0) ? 1 : 0;
}

__global__ void k_row_blocks(int64_t nb0, int64_t count, int32_t row0, int32_t* __restrict__ bs,
                             int32_t* __restrict__ bm) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  bs[nb0 + i] = (int32_t)(row0 + i);
  bm[nb0 + i] = 1;
}

// ---- GMRES helpers ----
// h[j] partials: part[j*np + b] = sum_i V_j[i] w[i] for j <= jmax
__global__ void __launch_bounds__(kRB) k_multidot(int64_t n, int nv, const double* __restrict__ V,
                                                  const double* __restrict__ w, double* __restrict__ part, int np) {
  __shared__ double sh[kRB / 64];
  for (int j = 0; j < nv; ++j) {
    const double* v = V + (size_t)j * n;
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kRB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kRB) s += v[i] * w[i];
    const double t = block_sum(s, sh);
    if (threadIdx.x == 0) part[(size_t)j * np + blockIdx.x] = t;
  }
}

// w -= sum_j h[j] V_j
__global__ void __launch_bounds__(kRB) k_multiaxpy(int64_t n, int nv, const double* __restrict__ V,
                                                   const double* __restrict__ h, double* __restrict__ w) {
  for (int64_t i = (int64_t)blockIdx.x * kRB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kRB) {
    double s = w[i];
    for (int j = 0; j < nv; ++j) s -= h[j] * V[(size_t)j * n + i];
    w[i] = s;
  }
}

__global__ void __launch_bounds__(kRB) k_scale(int64_t n, const double* __restrict__ a, double s, double* __restrict__ b) {
  for (int64_t i = (int64_t)blockIdx.x * kRB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kRB) b[i] = a[i] * s;
}

// out = sum_j y[j] V_j  (y on device)
__global__ void __launch_bounds__(kRB) k_combine(int64_t n, int nv, const double* __restrict__ V,
                                                 const double* __restrict__ y, double* __restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * kRB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kRB) {
    double s = 0.0;
    for (int j = 0; j < nv; ++j) s += y[j] * V[(size_t)j * n + i];
    out[i] = s;
  }
}

__global__ void __launch_bounds__(kRB) k_axpy(int64_t n, double a, const double* __restrict__ x, double* __restrict__ y) {
  for (int64_t i = (int64_t)blockIdx.x * kRB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kRB) y[i] += a * x[i];
}

__global__ void __launch_bounds__(kRB) k_sub(int64_t n, const double* __restrict__ a, const double* __restrict__ b,
                                             double* __restrict__ c) {
  for (int64_t i = (int64_t)blockIdx.x * kRB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kRB) c[i] = a[i] - b[i];
}

// z = M r over blocks (no reductions)
__global__ void __launch_bounds__(kRB) k_pc_apply(int64_t nblk, const int32_t* __restrict__ bs,
                                                  const int32_t* __restrict__ bm, const double* __restrict__ binv,
                                                  const double* __restrict__ r, double* __restrict__ z) {
  for (int64_t k = (int64_t)blockIdx.x * kRB + threadIdx.x; k < nblk; k += (int64_t)gridDim.x * kRB) {
    double a = 0.0, b = 0.0;
    apply_block(bs[k], bm[k], binv + 9 * k, r, z, a, b);
  }
}

}  // namespace

// SpMV dispatcher: lanes per row from the mean row length.
template <typename V>
static mpf_status spmv_t(const int32_t* rowptr, const int32_t* colidx, const V* vals, int64_t n, int64_t nnz,
                         const double* x, double* y, double* part, unsigned np, const double* sc, hipStream_t s) {
  const double avg = n > 0 ? (double)nnz / (double)n : 0.0;
  if (avg > 48) {
    constexpr int L = 32;
    const unsigned g = part ? np : grid_for(n, kRB / L);
    hipLaunchKernelGGL((k_spmv<L, V>), dim3(g), dim3(kRB), 0, s, n, rowptr, colidx, vals, x, y, part, sc);
  } else if (avg > 12) {
    constexpr int L = 16;
    const unsigned g = part ? np : grid_for(n, kRB / L);
    hipLaunchKernelGGL((k_spmv<L, V>), dim3(g), dim3(kRB), 0, s, n, rowptr, colidx, vals, x, y, part, sc);
  } else {
    constexpr int L = 4;
    const unsigned g = part ? np : grid_for(n, kRB / L);
    hipLaunchKernelGGL((k_spmv<L, V>), dim3(g), dim3(kRB), 0, s, n, rowptr, colidx, vals, x, y, part, sc);
  }
  MPF_HIP(hipGetLastError());
  return MPF_OK;
}

mpf_status spmv(const int32_t* rowptr, const int32_t* colidx, const double* vals, int64_t n, int64_t nnz,
                const double* x, double* y, double* part, unsigned np, const double* sc, hipStream_t s) {
  return spmv_t<double>(rowptr, colidx, vals, n, nnz, x, y, part, np, sc, s);
}

template <typename V>
static mpf_status spmv_grp_t(const CsrView& A, const BlockPcView& B, const V* vals, const double* x, double* y,
                             double* part, unsigned np, const double* sc, hipStream_t s) {
  const int64_t n = A.n, nnz = A.nnz;
  const double avg = n > 0 ? (double)nnz / (double)n : 0.0;
  const int32_t *bs = B.bs, *bm = B.bm, *rp = A.rowptr, *ci = A.colidx;
  if (avg > 48) {
    const unsigned g = part ? np : grid_for(B.nblk, kRB / 32);
    hipLaunchKernelGGL((k_spmv_grp<32, V>), dim3(g), dim3(kRB), 0, s, B.nblk, bs, bm, rp, ci, vals, x, y, part, sc);
  } else if (avg > 12) {
    const unsigned g = part ? np : grid_for(B.nblk, kRB / 16);
    hipLaunchKernelGGL((k_spmv_grp<16, V>), dim3(g), dim3(kRB), 0, s, B.nblk, bs, bm, rp, ci, vals, x, y, part, sc);
  } else {
    const unsigned g = part ? np : grid_for(B.nblk, kRB / 4);
    hipLaunchKernelGGL((k_spmv_grp<4, V>), dim3(g), dim3(kRB), 0, s, B.nblk, bs, bm, rp, ci, vals, x, y, part, sc);
  }
  MPF_HIP(hipGetLastError());
  return MPF_OK;
}

// y = A x: the row-group kernel when the blocks share column lists, else the per-row kernel; the
// FP32 values when the view has them.  part: partial x.y per workgroup (np of them) or null.
static mpf_status spmv_view(const CsrView& A, const BlockPcView& B, const double* x, double* y, double* part,
                            unsigned np, const double* sc, hipStream_t s) {
  if (A.vals32) {
    if (!B.grp) return spmv_t<float>(A.rowptr, A.colidx, A.vals32, A.n, A.nnz, x, y, part, np, sc, s);
    return spmv_grp_t<float>(A, B, A.vals32, x, y, part, np, sc, s);
  }
  if (!B.grp) return spmv(A.rowptr, A.colidx, A.vals, A.n, A.nnz, x, y, part, np, sc, s);
  return spmv_grp_t<double>(A, B, A.vals, x, y, part, np, sc, s);
}

// Workgroups (and so p.q partials) of spmv_view's product with partials: as many as the product
// runs without them (its lanes per row decide), at most kMaxParts -- a small system's product
// then spreads over the GPU instead of looping in a few workgroups.
static unsigned spmv_parts(const CsrView& A, const BlockPcView& B) {
  const double avg = A.n > 0 ? (double)A.nnz / (double)A.n : 0.0;
  const int lpr = avg > 48 ? 32 : (avg > 12 ? 16 : 4);  // spmv_view's choice for either kernel
  const int64_t units = B.grp ? B.nblk : A.n;
  int64_t g = (units + kRB / lpr - 1) / (kRB / lpr);
  if (g > kMaxParts) g = kMaxParts;
  if (g < 1) g = 1;
  return (unsigned)g;
}

static CsrView csr_of(const mpf_problem* p) {
  CsrView A;
  A.n = (int64_t)p->pat.dispDOF + p->pat.presDOF;
  A.nnz = p->pat.nnz;
  A.rowptr = p->pat.rowptr.ptr;
  A.colidx = p->pat.colidx.ptr;
  A.vals = p->vals.ptr;
  return A;
}

static BlockPcView blocks_of(const KrylovWork& W) {
  BlockPcView B;
  B.nblk = W.nblk;
  B.bs = W.blk_start.ptr;
  B.bm = W.blk_size.ptr;
  B.rblk = W.row_blk.ptr;
  B.binv = W.blk_inv.ptr;
  B.grp = W.grp_spmv;
  return B;
}

static mpf_status spmv_k(const mpf_problem* p, const KrylovWork& W, const double* x, double* y, double* part,
                         unsigned np, const double* sc, hipStream_t s) {
  return spmv_view(csr_of(p), blocks_of(W), x, y, part, np, sc, s);
}

mpf_status dev_norm2(const double* x, int64_t n, double* out_host, hipStream_t s, double* scratch) {
  const unsigned np = nparts_for(n);
  hipLaunchKernelGGL(k_dot_part, dim3(np), dim3(kRB), 0, s, n, x, x, scratch);
  hipLaunchKernelGGL(k_finish, dim3(1), dim3(kRB), 0, s, scratch, (int)np, 0, 1, scratch + kMaxParts);
  double h = 0.0;
  MPF_HIP(hipMemcpyAsync(&h, scratch + kMaxParts, sizeof(double), hipMemcpyDeviceToHost, s));
  MPF_HIP(hipStreamSynchronize(s));
  *out_host = std::sqrt(h);
  return MPF_OK;
}

mpf_status block_pc_structure(const CsrView& A, int64_t nblk, const int32_t* bs, const int32_t* bm, int32_t* rblk,
                              bool* grp, hipStream_t s) {
  const int B = 256;
  if (A.n > 0) MPF_HIP(hipMemsetAsync(rblk, 0, (size_t)A.n * sizeof(int32_t), s));
  if (nblk > 0) hipLaunchKernelGGL(k_row_blk, dim3(grid_for(nblk, B)), dim3(B), 0, s, nblk, bs, bm, rblk);
  *grp = false;
  if (nblk > 0) {
    DevBuf<int32_t> bad;
    MPF_TRY(bad.alloc(1));
    int32_t hbad = 0;
    MPF_HIP(hipMemcpyAsync(bad.ptr, &hbad, sizeof(int32_t), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_check_groups, dim3(grid_for(nblk, B)), dim3(B), 0, s, nblk, bs, bm, A.rowptr, A.colidx,
                       bad.ptr);
    MPF_HIP(hipMemcpyAsync(&hbad, bad.ptr, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    MPF_HIP(hipStreamSynchronize(s));
    *grp = hbad == 0;
  }
  MPF_HIP(hipGetLastError());
  return MPF_OK;
}

mpf_status block_pc_values(const CsrView& A, int64_t nblk, const int32_t* bs, const int32_t* bm, double* binv,
                           hipStream_t s) {
  if (nblk > 0)
    hipLaunchKernelGGL(k_block_setup, dim3(grid_for(nblk, 256)), dim3(256), 0, s, nblk, bs, bm, A.rowptr, A.colidx,
                       A.vals, binv, 0);
  MPF_HIP(hipGetLastError());
  return MPF_OK;
}

static mpf_status setup_work(mpf_problem* p, int pc, int restart) {
  if (!p->kw) p->kw = new KrylovWork();
  KrylovWork& W = *p->kw;
  hipStream_t s = p->stream;
  const int64_t n = (int64_t)p->pat.dispDOF + p->pat.presDOF;
  if (W.n != n) {
    W.n = n;
    MPF_TRY(W.r.alloc(n)); MPF_TRY(W.z.alloc(n)); MPF_TRY(W.p.alloc(n)); MPF_TRY(W.q.alloc(n)); MPF_TRY(W.w.alloc(n));
    MPF_TRY(W.r2.alloc(n));
    MPF_TRY(W.sc.alloc(S_N + kMaxRestart + 2));
    W.pc_kind = -1;
  }
  // partial sums: 2 arrays for CG, (restart + 1) for the GMRES multi-dot, +1 slot for dev_norm2's result
  const size_t nparts_needed = (size_t)kMaxParts * (size_t)(std::max(2, restart + 1) + 1);
  if (W.part.n < nparts_needed) MPF_TRY(W.part.alloc(nparts_needed));
  if (restart > 0 && W.vcap < restart + 1) {
    MPF_TRY(W.V.alloc((size_t)(restart + 1) * n));
    W.vcap = restart + 1;
  }
  if (pc == MPF_PC_PMG && W.zcap < restart) {
    MPF_TRY(W.Z.alloc((size_t)restart * n));
    W.zcap = restart;
  }
  // block list
  const int kind = (pc == MPF_PC_BJACOBI3 || pc == MPF_PC_PMG) ? 2 : 1;
  if (W.pc_kind != kind) {
    const int B = 256;
    if (kind == 1) {
      W.nblk = n;
      MPF_TRY(W.blk_start.alloc(n)); MPF_TRY(W.blk_size.alloc(n));
      hipLaunchKernelGGL(k_row_blocks, dim3(grid_for(n, B)), dim3(B), 0, s, (int64_t)0, n, 0, W.blk_start.ptr,
                         W.blk_size.ptr);
    } else {
      DevBuf<int32_t> f, first;
      MPF_TRY(f.alloc(p->nNode + 1)); MPF_TRY(first.alloc(p->nNode + 1));
      MPF_HIP(hipMemsetAsync(f.ptr, 0, (p->nNode + 1) * sizeof(int32_t), s));
      hipLaunchKernelGGL(k_has_free, dim3(grid_for(p->nNode, B)), dim3(B), 0, s, p->nNode, p->pat.nda.ptr, f.ptr);
      // exclusive scan on host (one-time setup, nNode ints)
      std::vector<int32_t> hf(p->nNode + 1), hfirst(p->nNode + 1);
      MPF_HIP(hipMemcpyAsync(hf.data(), f.ptr, (p->nNode + 1) * sizeof(int32_t), hipMemcpyDeviceToHost, s));
      MPF_HIP(hipStreamSynchronize(s));
      int64_t acc = 0;
      for (int i = 0; i <= p->nNode; ++i) { hfirst[i] = (int32_t)acc; acc += hf[i]; }
      const int64_t nub = hfirst[p->nNode];
      MPF_HIP(hipMemcpyAsync(first.ptr, hfirst.data(), (p->nNode + 1) * sizeof(int32_t), hipMemcpyHostToDevice, s));
      W.nblk = nub + p->pat.presDOF;
      MPF_TRY(W.blk_start.alloc(W.nblk)); MPF_TRY(W.blk_size.alloc(W.nblk));
      hipLaunchKernelGGL(k_node_blocks, dim3(grid_for(p->nNode, B)), dim3(B), 0, s, p->nNode, p->pat.nda.ptr,
                         first.ptr, W.blk_start.ptr, W.blk_size.ptr);
      if (p->pat.presDOF > 0)
        hipLaunchKernelGGL(k_row_blocks, dim3(grid_for(p->pat.presDOF, B)), dim3(B), 0, s, nub,
                           (int64_t)p->pat.presDOF, p->pat.dispDOF, W.blk_start.ptr, W.blk_size.ptr);
      MPF_HIP(hipStreamSynchronize(s));
    }
    MPF_TRY(W.blk_inv.alloc((size_t)W.nblk * 9));
    MPF_TRY(W.row_blk.alloc(n > 0 ? (size_t)n : 1));
    MPF_HIP(hipMemsetAsync(W.row_blk.ptr, 0, (n > 0 ? (size_t)n : 1) * sizeof(int32_t), s));
    hipLaunchKernelGGL(k_row_blk, dim3(grid_for(W.nblk, B)), dim3(B), 0, s, W.nblk, W.blk_start.ptr, W.blk_size.ptr,
                       W.row_blk.ptr);
    W.pc_kind = kind;
    W.grp_spmv = false;
    if (kind == 2 && W.nblk > 0) {
      DevBuf<int32_t> bad;
      MPF_TRY(bad.alloc(1));
      int32_t hbad = 0;
      MPF_HIP(hipMemcpyAsync(bad.ptr, &hbad, sizeof(int32_t), hipMemcpyHostToDevice, s));
      hipLaunchKernelGGL(k_check_groups, dim3(grid_for(W.nblk, B)), dim3(B), 0, s, W.nblk, W.blk_start.ptr,
                         W.blk_size.ptr, p->pat.rowptr.ptr, p->pat.colidx.ptr, bad.ptr);
      MPF_HIP(hipMemcpyAsync(&hbad, bad.ptr, sizeof(int32_t), hipMemcpyDeviceToHost, s));
      MPF_HIP(hipStreamSynchronize(s));
      W.grp_spmv = hbad == 0;
    }
  }
  hipLaunchKernelGGL(k_block_setup, dim3(grid_for(W.nblk, 256)), dim3(256), 0, s, W.nblk, W.blk_start.ptr,
                     W.blk_size.ptr, p->pat.rowptr.ptr, p->pat.colidx.ptr, p->vals.ptr, W.blk_inv.ptr,
                     pc == MPF_PC_NONE ? 1 : 0);
  if (kind == 2 && p->pat.presDOF > 0) {
    const int64_t nub = W.nblk - p->pat.presDOF;
    hipLaunchKernelGGL(k_schur_pressure, dim3(grid_for(p->pat.presDOF, 256)), dim3(256), 0, s, p->pat.presDOF,
                       p->pat.dispDOF, p->pat.rowptr.ptr, p->pat.colidx.ptr, p->vals.ptr, W.blk_start.ptr,
                       W.blk_size.ptr, nub, W.blk_inv.ptr);
  }
  MPF_HIP(hipGetLastError());
  if (pc == MPF_PC_PMG) {
    if (!W.grp_spmv) return MPF_ERR_ARG;  // node rows that do not share a column list: not an element pattern
    MPF_TRY(pmg_setup(p, blocks_of(W), W.nblk - p->pat.presDOF));
  }
  return MPF_OK;
}

mpf_status pcg_csr(const CsrView& A, const BlockPcView& B, const CgScratch& w, const double* b, double* x,
                   double rtol, int maxit, hipStream_t s, mpf_krylov_report* rep) {
  const int64_t n = A.n;
  const unsigned np = nparts_for(B.nblk);
  const unsigned nps = nparts_for(n);
  double* sc = w.sc;
  hipLaunchKernelGGL(k_cg_start, dim3(np), dim3(kRB), 0, s, B.nblk, B.bs, B.bm, B.binv, b, x, w.r, w.z, w.p, w.part,
                     (int)np);
  // rz -> sc[S_RZ0], bb -> sc[S_RR]
  hipLaunchKernelGGL(k_finish, dim3(1), dim3(kRB), 0, s, w.part, (int)np, (int)np, 1, sc + S_RZ0);
  hipLaunchKernelGGL(k_finish, dim3(1), dim3(kRB), 0, s, w.part + np, (int)np, (int)np, 1, sc + S_RR);
  hipLaunchKernelGGL(k_cg_init_scalars, dim3(1), dim3(64), 0, s, sc, rtol, maxit);
  int it = 0;
  int chunk = 8;
  double hs[S_N];
  const unsigned nsp = spmv_parts(A, B);
  for (;;) {
    for (int c = 0; c < chunk; ++c, ++it) {
      const int par = it & 1;
      MPF_TRY(spmv_view(A, B, w.p, w.q, w.part, nsp, sc, s));
      hipLaunchKernelGGL(k_finish, dim3(1), dim3(kRB), 0, s, w.part, (int)nsp, (int)nsp, 1, sc + S_PQ);
      double* rb[2] = {w.r, w.r2};
      hipLaunchKernelGGL(k_cg_update_rows, dim3(nps), dim3(kRB), 0, s, n, B.rblk, B.bs, B.bm, B.binv, x, rb[par],
                         rb[1 - par], w.z, w.p, w.q, w.part, (int)nps, sc, par);
      hipLaunchKernelGGL(k_cg_check, dim3(1), dim3(kRB), 0, s, w.part, (int)nps, sc, par, it);
      hipLaunchKernelGGL(k_cg_p, dim3(nps), dim3(kRB), 0, s, n, w.z, w.p, sc, par);
    }
    MPF_HIP(hipMemcpyAsync(hs, sc, sizeof(hs), hipMemcpyDeviceToHost, s));
    MPF_HIP(hipStreamSynchronize(s));
    if (hs[S_DONE] != 0.0) break;
    if (chunk < 64) chunk *= 2;
  }
  MPF_HIP(hipGetLastError());
  const double bb = hs[S_BB];
  rep->iterations = (int32_t)hs[S_ITS];
  rep->relres = bb > 0 ? std::sqrt(hs[S_RR] / bb) : 0.0;
  rep->converged = (bb == 0.0) || (hs[S_RR] <= hs[S_TOL2]);
  return MPF_OK;
}

static mpf_status pcg(mpf_problem* p, const mpf_krylov_opts* o, mpf_krylov_report* rep) {
  KrylovWork& W = *p->kw;
  const CgScratch w{W.r.ptr, W.r2.ptr, W.z.ptr, W.p.ptr, W.q.ptr, W.part.ptr, W.sc.ptr};
  return pcg_csr(csr_of(p), blocks_of(W), w, p->rhs.ptr, p->du.ptr, o->rtol, o->maxit, p->stream, rep);
}

// Right-preconditioned GMRES(m), classical Gram-Schmidt with one re-orthogonalisation;
// Hessenberg least squares (Givens) on the host.  With MPF_PC_PMG the preconditioner changes from
// one application to the next (its coarse solve is an inner CG to a loose tolerance), so the
// preconditioned vectors z_j are kept and the update is x += Z y (flexible GMRES).
static mpf_status gmres(mpf_problem* p, const mpf_krylov_opts* o, mpf_krylov_report* rep) {
  KrylovWork& W = *p->kw;
  hipStream_t s = p->stream;
  const int64_t n = W.n;
  const int m = o->restart > 0 ? o->restart : 30;
  const bool flex = o->pc == MPF_PC_PMG;
  const int64_t nub = W.nblk - p->pat.presDOF;
  const unsigned nps = nparts_for(n);
  const unsigned npb = nparts_for(W.nblk);
  double* x = p->du.ptr;
  const double* b = p->rhs.ptr;
  double* V = W.V.ptr;
  double* d_h = W.sc.ptr + S_N;  // device scratch for h / y
  MPF_HIP(hipMemsetAsync(x, 0, n * sizeof(double), s));
  double bn = 0.0;
  MPF_TRY(dev_norm2(b, n, &bn, s, W.part.ptr));
  rep->iterations = 0;
  rep->converged = 1;
  rep->relres = 0.0;
  if (bn == 0.0) return MPF_OK;
  std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), g(m + 1), h(m + 2), h2(m + 2), y(m);
  int it = 0;
  double rn = bn;
  bool first_cycle = true;
  while (it < o->maxit) {
    // r = b - A x  (x = 0 on the first cycle)
    if (first_cycle) {
      MPF_HIP(hipMemcpyAsync(W.r.ptr, b, n * sizeof(double), hipMemcpyDeviceToDevice, s));
      rn = bn;
    } else {
      MPF_TRY(spmv_k(p, W, x, W.w.ptr, nullptr, 0, nullptr, s));
      hipLaunchKernelGGL(k_sub, dim3(nps), dim3(kRB), 0, s, n, b, W.w.ptr, W.r.ptr);
      MPF_TRY(dev_norm2(W.r.ptr, n, &rn, s, W.part.ptr));
    }
    first_cycle = false;
    if (rn / bn < o->rtol) break;
    hipLaunchKernelGGL(k_scale, dim3(nps), dim3(kRB), 0, s, n, W.r.ptr, 1.0 / rn, V);
    std::fill(g.begin(), g.end(), 0.0);
    g[0] = rn;
    int j = 0;
    for (; j < m && it < o->maxit; ++j) {
      // w = A M^-1 v_j
      if (flex) {
        double* zj = W.Z.ptr + (size_t)j * n;
        MPF_TRY(pmg_apply(p, blocks_of(W), nub, V + (size_t)j * n, zj));
        MPF_TRY(spmv_k(p, W, zj, W.w.ptr, nullptr, 0, nullptr, s));
      } else {
        hipLaunchKernelGGL(k_pc_apply, dim3(npb), dim3(kRB), 0, s, W.nblk, W.blk_start.ptr, W.blk_size.ptr,
                           W.blk_inv.ptr, V + (size_t)j * n, W.z.ptr);
        MPF_TRY(spmv_k(p, W, W.z.ptr, W.w.ptr, nullptr, 0, nullptr, s));
      }
      // CGS2
      for (int pass = 0; pass < 2; ++pass) {
        hipLaunchKernelGGL(k_multidot, dim3(nps), dim3(kRB), 0, s, n, j + 1, V, W.w.ptr, W.part.ptr, (int)nps);
        hipLaunchKernelGGL(k_finish, dim3(j + 1), dim3(kRB), 0, s, W.part.ptr, (int)nps, (int)nps, j + 1, d_h);
        hipLaunchKernelGGL(k_multiaxpy, dim3(nps), dim3(kRB), 0, s, n, j + 1, V, d_h, W.w.ptr);
        MPF_HIP(hipMemcpyAsync(pass == 0 ? h.data() : h2.data(), d_h, (j + 1) * sizeof(double),
                               hipMemcpyDeviceToHost, s));
      }
      double hn = 0.0;
      MPF_TRY(dev_norm2(W.w.ptr, n, &hn, s, W.part.ptr));  // synchronises: h, h2 valid
      for (int i = 0; i <= j; ++i) H[(size_t)i * m + j] = h[i] + h2[i];
      H[(size_t)(j + 1) * m + j] = hn;
      if (hn > 0) hipLaunchKernelGGL(k_scale, dim3(nps), dim3(kRB), 0, s, n, W.w.ptr, 1.0 / hn, V + (size_t)(j + 1) * n);
      for (int i = 0; i < j; ++i) {
        const double t = cs[i] * H[(size_t)i * m + j] + sn[i] * H[(size_t)(i + 1) * m + j];
        H[(size_t)(i + 1) * m + j] = -sn[i] * H[(size_t)i * m + j] + cs[i] * H[(size_t)(i + 1) * m + j];
        H[(size_t)i * m + j] = t;
      }
      const double a = H[(size_t)j * m + j], bb = H[(size_t)(j + 1) * m + j];
      const double den = std::sqrt(a * a + bb * bb);
      cs[j] = a / den;
      sn[j] = bb / den;
      H[(size_t)j * m + j] = den;
      H[(size_t)(j + 1) * m + j] = 0.0;
      g[j + 1] = -sn[j] * g[j];
      g[j] = cs[j] * g[j];
      rn = std::fabs(g[j + 1]);
      ++it;
      if (rn / bn < o->rtol || hn == 0.0) { ++j; break; }
    }
    for (int i = j - 1; i >= 0; --i) {
      double sacc = g[i];
      for (int k = i + 1; k < j; ++k) sacc -= H[(size_t)i * m + k] * y[k];
      y[i] = sacc / H[(size_t)i * m + i];
    }
    // x += M^-1 (V y)   (flexible: x += Z y)
    MPF_HIP(hipMemcpyAsync(d_h, y.data(), j * sizeof(double), hipMemcpyHostToDevice, s));
    if (flex) {
      hipLaunchKernelGGL(k_combine, dim3(nps), dim3(kRB), 0, s, n, j, W.Z.ptr, d_h, W.z.ptr);
    } else {
      hipLaunchKernelGGL(k_combine, dim3(nps), dim3(kRB), 0, s, n, j, V, d_h, W.w.ptr);
      hipLaunchKernelGGL(k_pc_apply, dim3(npb), dim3(kRB), 0, s, W.nblk, W.blk_start.ptr, W.blk_size.ptr,
                         W.blk_inv.ptr, W.w.ptr, W.z.ptr);
    }
    hipLaunchKernelGGL(k_axpy, dim3(nps), dim3(kRB), 0, s, n, 1.0, W.z.ptr, x);
    if (rn / bn < o->rtol) break;
  }
  // true residual
  MPF_TRY(spmv_k(p, W, x, W.w.ptr, nullptr, 0, nullptr, s));
  hipLaunchKernelGGL(k_sub, dim3(nps), dim3(kRB), 0, s, n, b, W.w.ptr, W.r.ptr);
  double tr = 0.0;
  MPF_TRY(dev_norm2(W.r.ptr, n, &tr, s, W.part.ptr));
  rep->iterations = it;
  rep->relres = tr / bn;
  rep->converged = rep->relres < o->rtol * 10.0 ? 1 : 0;
  MPF_HIP(hipGetLastError());
  return MPF_OK;
}

mpf_status krylov_solve(mpf_problem* p, const mpf_krylov_opts* o, mpf_krylov_report* rep) {
  if (o->ksp == MPF_KSP_BAND) return band_solve(p, rep);
  const auto t0 = std::chrono::steady_clock::now();
  const int64_t n = (int64_t)p->pat.dispDOF + p->pat.presDOF;
  if (o->ksp == MPF_KSP_GMRES && o->restart > kMaxRestart) return MPF_ERR_ARG;
  // the two-level preconditioner needs the Tet10 mesh and a flexible outer method
  if (o->pc == MPF_PC_PMG && (o->ksp != MPF_KSP_GMRES || (p->etype != MPF_TET10 && p->etype != MPF_TET10P1)))
    return MPF_ERR_ARG;
  MPF_TRY(p->du.alloc(n > 0 ? n : 1));
  MPF_TRY(setup_work(p, o->pc, o->ksp == MPF_KSP_GMRES ? (o->restart > 0 ? o->restart : 30) : 0));
  mpf_status st;
  if (o->ksp == MPF_KSP_GMRES) st = gmres(p, o, rep);
  else st = pcg(p, o, rep);
  MPF_HIP(hipStreamSynchronize(p->stream));
  rep->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return st;
}

void krylov_free(mpf_problem* p) {
  delete p->kw;
  p->kw = nullptr;
}

}  // namespace mpf
