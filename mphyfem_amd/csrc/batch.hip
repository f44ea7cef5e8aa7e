// batch.hip -- many independent meshes solved as one problem: batched Newton with batched
// Jacobi-PCG, replayed from HIP graphs, with tile shapes chosen by a measuring autotuner.
//
// The meshes are concatenated into one problem (disjoint node and element ranges, a material per
// mesh), so pattern, colouring and assembly are the single-mesh ones (the CSR matrix is block
// diagonal).  What is per mesh is the solver: every mesh runs its own conjugate-gradient recurrence
// (its own alpha, beta and stopping test: the same iterates as a CG on that mesh alone) and its own
// Newton convergence test ||R_m|| / ||R0_m|| < rtol, with R0_m taken at the plain start of the load
// step (femSolidmechanics2.cpp:730-851 per mesh; SolverPetsc.cpp:250-330 with KSPCG + PCJACOBI).
// Converged meshes are masked out of the remaining work.
//
// Vector work is organised in tiles: contiguous dof ranges inside one mesh (meshes own contiguous dof
// ranges because free dofs are numbered node-major).  One workgroup per tile; per-tile partial dot
// products are reduced per mesh in tile order, so every reduction is deterministic.
//
// One Newton iteration of the whole batch is one graph launch and one 16-byte read-back: the update
// of the previous iteration's solve, the assembly, residual norms, Newton test and CG start, then K
// CG iterations (or the banded direct solve of every mesh).  K follows the slowest mesh's CG count
// in the previous solve, rounded up to a power of two; the iterations after every mesh's CG has
// stopped cost only their launch (each tile checks its mesh's flag first), and a solve that needs
// more than K goes on with graphs of `chunk` iterations.  The load factor is read from device
// memory, so the graphs are captured once (per K) and replayed across load steps.

#include <algorithm>
#include <chrono>
#include <cmath>
#include <functional>
#include <vector>

#include "mpf_internal.hpp"

// problem.hip
mpf_status enqueue_assembly(mpf_problem* p, bool first_iter, double lam, int scatterK,
                            const double* lam_dev = nullptr);
mpf_status do_update(mpf_problem* p, bool first_iter, const int32_t* emat = nullptr, const int32_t* mflag = nullptr,
                     int32_t stride = 0);

namespace mpf {

// Batched-solver state (owned by mpf_problem::batch).
struct BatchWork {
  int32_t nmesh = 0;
  std::vector<int32_t> node_off, elem_off, dof_off;  // host, [nmesh + 1]
  DevBuf<int32_t> emat;                              // element -> mesh
  DevBuf<double> mtab;                               // [nmesh][2]: mu, K
  // tiles
  int32_t T = 1024, LPR = 8, chunk = 16, ntiles = 0;
  bool tuned = false;
  // element batch picked by the autotuner: assembly algorithm, chunk kernel's staged-element cap
  int32_t asm_mode = -1, asm_cap = -1;
  double asm_ms = 0.0;
  DevBuf<int32_t> tlo, thi, tmesh, mtile;  // tile ranges and mesh, mesh -> first tile [nmesh + 1]
  // row groups: the free rows of one node, [gs[g], gs[g + 1]); tile t holds groups [tg[t], tg[t + 1])
  std::vector<int32_t> gs_h;
  DevBuf<int32_t> gs, tg;
  bool grp_spmv = false;  // every group's rows share one column list (checked once)
  // CG vectors (x is the problem's du)
  DevBuf<double> dinv, r, z, pv, q;
  DevBuf<double> part;  // [ntiles][2]
  // per-mesh scalars: see the k_b_* kernels for the slots
  DevBuf<double> ms;    // [nmesh][kMS]
  DevBuf<int32_t> mf;   // [nmesh][kMF]
  DevBuf<int32_t> cnt;  // [4]: CG-active meshes, Newton-active meshes, diverged meshes, most CG
                        // iterations of one mesh in the current solve
  int32_t* cnt_h = nullptr;  // pinned host mirror
  DevBuf<double> lam_dev;    // load factor read by the captured assembly
  // Newton-iteration graphs [variant][K index]: variant 0 / 1: head of the first iteration (plain /
  // guessed start); 2 / 3: update of a first / later solve, then the head.  K index 0: no solve
  // (the last test); k > 0: kKs[k] CG iterations, or the direct solve.
  static constexpr int kNV = 4, kNK = 9;
  hipGraphExec_t g_it[kNV][kNK] = {};
  hipGraphExec_t g_chunk = nullptr, g_r0 = nullptr;
  // the graphs hold these as kernel arguments: any change re-captures them
  double g_rtol = -1.0, g_cg_rtol = -1.0;
  int32_t g_cg_maxit = -1;
  int32_t g_solver = -1;            // linear solver the graphs were captured for
  int32_t k_pred = 4;               // K index of the next first-iteration solve (256 iterations)
  ~BatchWork() { release_graphs(); if (cnt_h) (void)hipHostFree(cnt_h); }
  void release_graphs() {
    for (auto& row : g_it)
      for (auto& g : row) if (g) { (void)hipGraphExecDestroy(g); g = nullptr; }
    if (g_chunk) { (void)hipGraphExecDestroy(g_chunk); g_chunk = nullptr; }
    if (g_r0) { (void)hipGraphExecDestroy(g_r0); g_r0 = nullptr; }
    g_solver = -1;
  }
};

// CG iterations per Newton-iteration graph, by K index
constexpr int kKs[BatchWork::kNK] = {0, 32, 64, 128, 256, 512, 1024, 2048, 4096};

// per-mesh double slots
//   kETA forcing term of the last solve, kRNP residual norm before the last solve
enum { kRZ = 0, kPQ, kRR, kRR0, kAL, kBE, kTOL2, kR0, kRN, kREL, kETA, kRNP, kMS };
// per-mesh int slots
//   kCGA CG running, kNA Newton running, kCGIT CG iterations of the current solve, kNIT Newton
//   iterations (solves) of the current load step, kDIV diverged, kCGT CG iterations of the load step
enum { kCGA = 0, kNA, kCGIT, kNIT, kDIV, kCGT, kMF };

namespace {

constexpr int kB = 256;  // threads per tile workgroup

__device__ __forceinline__ double block_sum(double v, double* sh) {
  const int t = threadIdx.x;
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  if ((t & 63) == 0) sh[t >> 6] = v;
  __syncthreads();
  double s = 0.0;
  if (t == 0)
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += sh[w];
  __syncthreads();
  return s;  // valid in thread 0
}

// Jacobi: inverse of the diagonal of every row.
__global__ void k_b_diag(int64_t n, const int32_t* __restrict__ rowptr, const int32_t* __restrict__ colidx,
                         const double* __restrict__ vals, double* __restrict__ dinv) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int lo = rowptr[i], hi = rowptr[i + 1] - 1;
  double d = 0.0;
  while (lo <= hi) {
    const int mid = (lo + hi) >> 1;
    const int c = colidx[mid];
    if (c == i) { d = vals[mid]; break; }
    if (c < i) lo = mid + 1; else hi = mid - 1;
  }
  dinv[i] = d != 0.0 ? 1.0 / d : 0.0;
}

// partial[tile][comp] = x . y over the tile (meshes with flag[m*kMF + slot] == 0 contribute 0)
__global__ void __launch_bounds__(kB) k_b_dot(const int32_t* __restrict__ tlo, const int32_t* __restrict__ thi,
                                              const double* __restrict__ x, const double* __restrict__ y,
                                              double* __restrict__ part, int comp) {
  __shared__ double sh[kB / 64];
  const int b = blockIdx.x;
  double v = 0.0;
  for (int i = tlo[b] + threadIdx.x; i < thi[b]; i += kB) v += x[i] * y[i];
  v = block_sum(v, sh);
  if (threadIdx.x == 0) part[2 * b + comp] = v;
}

// y = A x on the tile's rows (LPR lanes per row) and partial[tile][0] = x . y; skipped for meshes
// whose CG has stopped.
template <int LPR>
__global__ void __launch_bounds__(kB) k_b_spmv(const int32_t* __restrict__ tlo, const int32_t* __restrict__ thi,
                                               const int32_t* __restrict__ tmesh, const int32_t* __restrict__ mf,
                                               const int32_t* __restrict__ rowptr, const int32_t* __restrict__ colidx,
                                               const double* __restrict__ vals, const double* __restrict__ x,
                                               double* __restrict__ y, double* __restrict__ part) {
  __shared__ double sh[kB / 64];
  const int b = blockIdx.x;
  if (!mf[tmesh[b] * kMF + kCGA]) {
    if (threadIdx.x == 0) part[2 * b] = 0.0;
    return;
  }
  const int g = threadIdx.x / LPR, l = threadIdx.x % LPR;
  double dot = 0.0;
  for (int row = tlo[b] + g; row < thi[b]; row += kB / LPR) {
    double s = 0.0;
    for (int k = rowptr[row] + l; k < rowptr[row + 1]; k += LPR) s += vals[k] * x[colidx[k]];
    for (int off = LPR / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, LPR);
    if (l == 0) {
      y[row] = s;
      dot += x[row] * s;
    }
  }
  dot = block_sum(dot, sh);
  if (threadIdx.x == 0) part[2 * b] = dot;
}

// q = A p over the tile's row groups (a node's free rows share one column list, so each column
// index and p value is read once for up to three rows); same per-row summation order as k_b_spmv.
template <int LPR>
__global__ void __launch_bounds__(kB) k_b_spmv_grp(const int32_t* __restrict__ tg, const int32_t* __restrict__ gs,
                                                   const int32_t* __restrict__ tmesh, const int32_t* __restrict__ mf,
                                                   const int32_t* __restrict__ rowptr, const int32_t* __restrict__ colidx,
                                                   const double* __restrict__ vals, const double* __restrict__ x,
                                                   double* __restrict__ y, double* __restrict__ part) {
  __shared__ double sh[kB / 64];
  const int b = blockIdx.x;
  if (!mf[tmesh[b] * kMF + kCGA]) {
    if (threadIdx.x == 0) part[2 * b] = 0.0;
    return;
  }
  const int gi = threadIdx.x / LPR, l = threadIdx.x % LPR;
  double dot = 0.0;
  for (int g = tg[b] + gi; g < tg[b + 1]; g += kB / LPR) {
    const int st = gs[g], m = gs[g + 1] - st;
    const int b0 = rowptr[st], len = rowptr[st + 1] - b0;
    const int b1 = m > 1 ? rowptr[st + 1] : b0, b2 = m > 2 ? rowptr[st + 2] : b0;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
#pragma unroll 4
    for (int j = l; j < len; j += LPR) {
      const double xv = x[colidx[b0 + j]];
      s0 += vals[b0 + j] * xv;
      if (m > 1) s1 += vals[b1 + j] * xv;
      if (m > 2) s2 += vals[b2 + j] * xv;
    }
    for (int off = LPR / 2; off > 0; off >>= 1) {
      s0 += __shfl_xor(s0, off, LPR);
      s1 += __shfl_xor(s1, off, LPR);
      s2 += __shfl_xor(s2, off, LPR);
    }
    if (l == 0) {
      y[st] = s0;
      dot += x[st] * s0;
      if (m > 1) { y[st + 1] = s1; dot += x[st + 1] * s1; }
      if (m > 2) { y[st + 2] = s2; dot += x[st + 2] * s2; }
    }
  }
  dot = block_sum(dot, sh);
  if (threadIdx.x == 0) part[2 * b] = dot;
}

// bad[0] = 1 if some group's rows do not share one column list, or a group has more than 3 rows
__global__ void k_b_check_groups(int32_t ng, const int32_t* __restrict__ gs, const int32_t* __restrict__ rowptr,
                                 const int32_t* __restrict__ colidx, int32_t* __restrict__ bad) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ng) return;
  const int st = gs[g], m = gs[g + 1] - st;
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

// CG start: x = 0, r = b, z = D^-1 r, p = z; partials r.r and r.z (meshes not Newton-active: r = 0).
__global__ void __launch_bounds__(kB) k_b_start(const int32_t* __restrict__ tlo, const int32_t* __restrict__ thi,
                                                const int32_t* __restrict__ tmesh, const int32_t* __restrict__ mf,
                                                const double* __restrict__ rhs, const double* __restrict__ dinv,
                                                double* __restrict__ x, double* __restrict__ r, double* __restrict__ z,
                                                double* __restrict__ p, double* __restrict__ part) {
  __shared__ double sh[kB / 64];
  const int b = blockIdx.x;
  const bool on = mf[tmesh[b] * kMF + kNA] != 0;
  double rr = 0.0, rz = 0.0;
  for (int i = tlo[b] + threadIdx.x; i < thi[b]; i += kB) {
    const double ri = on ? rhs[i] : 0.0;
    const double zi = dinv[i] * ri;
    x[i] = 0.0;
    r[i] = ri;
    z[i] = zi;
    p[i] = zi;
    rr += ri * ri;
    rz += ri * zi;
  }
  rr = block_sum(rr, sh);
  rz = block_sum(rz, sh);
  if (threadIdx.x == 0) { part[2 * b] = rr; part[2 * b + 1] = rz; }
}

// x += a p, r -= a q, z = D^-1 r; partials r.r and r.z.
__global__ void __launch_bounds__(kB) k_b_xrz(const int32_t* __restrict__ tlo, const int32_t* __restrict__ thi,
                                              const int32_t* __restrict__ tmesh, const int32_t* __restrict__ mf,
                                              const double* __restrict__ ms, const double* __restrict__ dinv,
                                              double* __restrict__ x, double* __restrict__ r, double* __restrict__ z,
                                              const double* __restrict__ p, const double* __restrict__ q,
                                              double* __restrict__ part) {
  __shared__ double sh[kB / 64];
  const int b = blockIdx.x;
  const int m = tmesh[b];
  if (!mf[m * kMF + kCGA]) {
    if (threadIdx.x == 0) { part[2 * b] = 0.0; part[2 * b + 1] = 0.0; }
    return;
  }
  const double a = ms[m * kMS + kAL];
  double rr = 0.0, rz = 0.0;
  for (int i = tlo[b] + threadIdx.x; i < thi[b]; i += kB) {
    x[i] += a * p[i];
    const double ri = r[i] - a * q[i];
    const double zi = dinv[i] * ri;
    r[i] = ri;
    z[i] = zi;
    rr += ri * ri;
    rz += ri * zi;
  }
  rr = block_sum(rr, sh);
  rz = block_sum(rz, sh);
  if (threadIdx.x == 0) { part[2 * b] = rr; part[2 * b + 1] = rz; }
}

__global__ void __launch_bounds__(kB) k_b_p(const int32_t* __restrict__ tlo, const int32_t* __restrict__ thi,
                                            const int32_t* __restrict__ tmesh, const int32_t* __restrict__ mf,
                                            const double* __restrict__ ms, const double* __restrict__ z,
                                            double* __restrict__ p) {
  const int b = blockIdx.x;
  const int m = tmesh[b];
  if (!mf[m * kMF + kCGA]) return;
  const double be = ms[m * kMS + kBE];
  for (int i = tlo[b] + threadIdx.x; i < thi[b]; i += kB) p[i] = z[i] + be * p[i];
}

// Per-mesh scalar steps (one workgroup; thread per mesh; partials summed in tile order).
__device__ __forceinline__ double mesh_sum(const double* part, const int32_t* mtile, int m, int comp) {
  double s = 0.0;
  for (int t = mtile[m]; t < mtile[m + 1]; ++t) s += part[2 * t + comp];
  return s;
}

__device__ __forceinline__ void count_into(int32_t v, int32_t* out) {
  __shared__ int32_t sc[kB / 64];
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  if ((threadIdx.x & 63) == 0) sc[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    int32_t s = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += sc[w];
    *out = s;
  }
  __syncthreads();
}

__device__ __forceinline__ void max_into(int32_t v, int32_t* out) {
  __shared__ int32_t sm[kB / 64];
  for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    int32_t s = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s = max(s, sm[w]);
    *out = s;
  }
  __syncthreads();
}

// after the residual norm partials (comp 0): Newton test per mesh.  first: R0 = ||R||.
__global__ void __launch_bounds__(kB) k_b_newton(int32_t nmesh, const int32_t* __restrict__ mtile,
                                                 const double* __restrict__ part, double* __restrict__ ms,
                                                 int32_t* __restrict__ mf, int32_t first, double rtol, double eta_min,
                                                 int32_t* __restrict__ cnt) {
  int32_t act = 0, div = 0;
  for (int m = threadIdx.x; m < nmesh; m += blockDim.x) {
    double* S = ms + m * kMS;
    int32_t* F = mf + m * kMF;
    const double rn = sqrt(mesh_sum(part, mtile, m, 0));
    if (first) S[kR0] = rn;
    S[kRN] = rn;
    const double R0 = S[kR0];
    const double rel = R0 > 0.0 ? rn / R0 : 0.0;
    S[kREL] = rel;
    if (F[kNA]) {
      if (rn != rn) { F[kDIV] = 1; F[kNA] = 0; }
      else if (R0 == 0.0 || rel < rtol) F[kNA] = 0;
      else F[kNIT] += 1;
    }
    // Inexact Newton: CG to eta ||R||.  First solve of the step: eta = 0.1 min(1, rel); later solves
    // follow the observed nonlinear reduction (Eisenstat-Walker, choice 2: 0.9 (|R_k|/|R_k-1|)^2,
    // kept >= 0.9 eta_prev^2 while that exceeds 0.1); never tighter than the step's target needs
    // (0.5 rtol / rel) nor than eta_min, never looser than 0.5.
    double eta;
    if (F[kNIT] <= 1) {
      eta = 0.1 * fmin(1.0, rel);
    } else {
      const double ratio = S[kRNP] > 0.0 ? rn / S[kRNP] : 1.0;
      eta = 0.9 * ratio * ratio;
      const double sg = 0.9 * S[kETA] * S[kETA];
      if (sg > 0.1) eta = fmax(eta, sg);
    }
    if (rel > 0.0) eta = fmax(eta, 0.5 * rtol / rel);
    eta = fmin(0.5, fmax(eta_min, eta));
    if (F[kNA]) { S[kETA] = eta; S[kRNP] = rn; }
    S[kTOL2] = eta * eta;
    act += F[kNA] != 0;
    div += F[kDIV] != 0;
  }
  count_into(act, cnt + 1);
  count_into(div, cnt + 2);
}

__global__ void __launch_bounds__(kB) k_b_start_scalars(int32_t nmesh, const int32_t* __restrict__ mtile,
                                                        const double* __restrict__ part, double* __restrict__ ms,
                                                        int32_t* __restrict__ mf, int32_t* __restrict__ cnt) {
  int32_t act = 0;
  for (int m = threadIdx.x; m < nmesh; m += blockDim.x) {
    double* S = ms + m * kMS;
    int32_t* F = mf + m * kMF;
    const double rr = mesh_sum(part, mtile, m, 0), rz = mesh_sum(part, mtile, m, 1);
    S[kRR0] = rr;
    S[kRR] = rr;
    S[kRZ] = rz;
    F[kCGT] += F[kCGIT];
    F[kCGIT] = 0;
    F[kCGA] = (F[kNA] && rr > 0.0) ? 1 : 0;
    act += F[kCGA];
  }
  count_into(act, cnt);
  if (threadIdx.x == 0) cnt[3] = 0;
}

__global__ void __launch_bounds__(kB) k_b_alpha(int32_t nmesh, const int32_t* __restrict__ mtile,
                                                const double* __restrict__ part, double* __restrict__ ms,
                                                const int32_t* __restrict__ mf) {
  for (int m = threadIdx.x; m < nmesh; m += blockDim.x) {
    double* S = ms + m * kMS;
    const double pq = mesh_sum(part, mtile, m, 0);
    S[kPQ] = pq;
    S[kAL] = (mf[m * kMF + kCGA] && pq != 0.0) ? S[kRZ] / pq : 0.0;
  }
}

__global__ void __launch_bounds__(kB) k_b_beta(int32_t nmesh, const int32_t* __restrict__ mtile,
                                               const double* __restrict__ part, double* __restrict__ ms,
                                               int32_t* __restrict__ mf, int32_t maxit, int32_t* __restrict__ cnt) {
  int32_t act = 0, most = 0;
  for (int m = threadIdx.x; m < nmesh; m += blockDim.x) {
    double* S = ms + m * kMS;
    int32_t* F = mf + m * kMF;
    if (F[kCGA]) {
      const double rr = mesh_sum(part, mtile, m, 0), rz = mesh_sum(part, mtile, m, 1);
      F[kCGIT] += 1;
      S[kRR] = rr;
      S[kBE] = S[kRZ] != 0.0 ? rz / S[kRZ] : 0.0;
      S[kRZ] = rz;
      if (rr <= S[kTOL2] * S[kRR0] || F[kCGIT] >= maxit || rr != rr) F[kCGA] = 0;
    }
    act += F[kCGA];
    most = max(most, F[kNA] ? F[kCGIT] : 0);
  }
  count_into(act, cnt);
  max_into(most, cnt + 3);
}

__global__ void k_b_set_scalar(double* __restrict__ dst, double v) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *dst = v;
}

// prescribed dofs at their values for load factor lam (the start point of a guessed step)
__global__ void k_b_set_prescribed(int32_t nbc, const int32_t* __restrict__ dof, const double* __restrict__ val,
                                   double lam, double* __restrict__ disp) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nbc) disp[dof[i]] = val[i] * lam;
}

__global__ void k_b_zero_i32(int32_t* a, int32_t n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0;
}

// Tiles of about T dofs inside each mesh, cut only where a node's rows begin (so every row group
// lies in one tile).
mpf_status build_tiles(BatchWork& B, int T, hipStream_t s) {
  std::vector<int32_t> lo, hi, tm, tgh, mt(B.nmesh + 1, 0);
  const std::vector<int32_t>& gs = B.gs_h;
  const int32_t ng = (int32_t)gs.size() - 1;
  int32_t g = 0;
  for (int m = 0; m < B.nmesh; ++m) {
    mt[m] = (int32_t)lo.size();
    while (g < ng && gs[g] < B.dof_off[m + 1]) {  // groups of mesh m, in dof order
      const int32_t a = gs[g];
      lo.push_back(a);
      tgh.push_back(g);
      tm.push_back(m);
      while (g < ng && gs[g] < B.dof_off[m + 1] && gs[g] - a < T) ++g;
      hi.push_back(gs[g] < B.dof_off[m + 1] ? gs[g] : B.dof_off[m + 1]);
    }
  }
  tgh.push_back(g);
  mt[B.nmesh] = (int32_t)lo.size();
  MPF_TRY(B.tg.upload(tgh.data(), tgh.size(), s));
  B.T = T;
  B.ntiles = (int32_t)lo.size();
  MPF_TRY(B.tlo.upload(lo.data(), lo.size(), s));
  MPF_TRY(B.thi.upload(hi.data(), hi.size(), s));
  MPF_TRY(B.tmesh.upload(tm.data(), tm.size(), s));
  MPF_TRY(B.mtile.upload(mt.data(), mt.size(), s));
  MPF_TRY(B.part.alloc((size_t)std::max(B.ntiles, 1) * 2));
  MPF_HIP(hipStreamSynchronize(s));  // the host vectors go out of scope
  return MPF_OK;
}

// One CG iteration of every active mesh.
mpf_status enqueue_cg_iter(mpf_problem* p, BatchWork& B, int32_t maxit) {
  hipStream_t s = p->stream;
  const PatternBufs& P = p->pat;
  const dim3 gt(std::max(B.ntiles, 1)), bt(kB);
  if (B.grp_spmv) {
    switch (B.LPR) {
      case 4:
        hipLaunchKernelGGL(k_b_spmv_grp<4>, gt, bt, 0, s, B.tg.ptr, B.gs.ptr, B.tmesh.ptr, B.mf.ptr, P.rowptr.ptr,
                           P.colidx.ptr, p->vals.ptr, B.pv.ptr, B.q.ptr, B.part.ptr);
        break;
      case 16:
        hipLaunchKernelGGL(k_b_spmv_grp<16>, gt, bt, 0, s, B.tg.ptr, B.gs.ptr, B.tmesh.ptr, B.mf.ptr, P.rowptr.ptr,
                           P.colidx.ptr, p->vals.ptr, B.pv.ptr, B.q.ptr, B.part.ptr);
        break;
      default:
        hipLaunchKernelGGL(k_b_spmv_grp<8>, gt, bt, 0, s, B.tg.ptr, B.gs.ptr, B.tmesh.ptr, B.mf.ptr, P.rowptr.ptr,
                           P.colidx.ptr, p->vals.ptr, B.pv.ptr, B.q.ptr, B.part.ptr);
    }
  } else {
    switch (B.LPR) {
      case 4:
        hipLaunchKernelGGL(k_b_spmv<4>, gt, bt, 0, s, B.tlo.ptr, B.thi.ptr, B.tmesh.ptr, B.mf.ptr, P.rowptr.ptr,
                           P.colidx.ptr, p->vals.ptr, B.pv.ptr, B.q.ptr, B.part.ptr);
        break;
      case 16:
        hipLaunchKernelGGL(k_b_spmv<16>, gt, bt, 0, s, B.tlo.ptr, B.thi.ptr, B.tmesh.ptr, B.mf.ptr, P.rowptr.ptr,
                           P.colidx.ptr, p->vals.ptr, B.pv.ptr, B.q.ptr, B.part.ptr);
        break;
      default:
        hipLaunchKernelGGL(k_b_spmv<8>, gt, bt, 0, s, B.tlo.ptr, B.thi.ptr, B.tmesh.ptr, B.mf.ptr, P.rowptr.ptr,
                           P.colidx.ptr, p->vals.ptr, B.pv.ptr, B.q.ptr, B.part.ptr);
    }
  }
  hipLaunchKernelGGL(k_b_alpha, dim3(1), bt, 0, s, B.nmesh, B.mtile.ptr, B.part.ptr, B.ms.ptr, B.mf.ptr);
  hipLaunchKernelGGL(k_b_xrz, gt, bt, 0, s, B.tlo.ptr, B.thi.ptr, B.tmesh.ptr, B.mf.ptr, B.ms.ptr, B.dinv.ptr,
                     p->du.ptr, B.r.ptr, B.z.ptr, B.pv.ptr, B.q.ptr, B.part.ptr);
  hipLaunchKernelGGL(k_b_beta, dim3(1), bt, 0, s, B.nmesh, B.mtile.ptr, B.part.ptr, B.ms.ptr, B.mf.ptr, maxit,
                     B.cnt.ptr);
  hipLaunchKernelGGL(k_b_p, gt, bt, 0, s, B.tlo.ptr, B.thi.ptr, B.tmesh.ptr, B.mf.ptr, B.ms.ptr, B.z.ptr, B.pv.ptr);
  MPF_HIP(hipGetLastError());
  return MPF_OK;
}

// assembly + residual norms + Newton test (+ CG start)
mpf_status enqueue_newton_head(mpf_problem* p, BatchWork& B, bool first, double lam, double rtol, double eta_min,
                               bool cg_start = true) {
  hipStream_t s = p->stream;
  const PatternBufs& P = p->pat;
  const int64_t n = P.dispDOF;
  const dim3 gt(std::max(B.ntiles, 1)), bt(kB);
  hipLaunchKernelGGL(k_b_zero_i32, dim3(1), dim3(64), 0, s, p->bad.ptr, 1);
  MPF_TRY(enqueue_assembly(p, first, lam, 1, B.lam_dev.ptr));
  hipLaunchKernelGGL(k_b_dot, gt, bt, 0, s, B.tlo.ptr, B.thi.ptr, p->rhs.ptr, p->rhs.ptr, B.part.ptr, 0);
  hipLaunchKernelGGL(k_b_newton, dim3(1), bt, 0, s, B.nmesh, B.mtile.ptr, B.part.ptr, B.ms.ptr, B.mf.ptr,
                     (int32_t)first, rtol, eta_min, B.cnt.ptr);
  if (!cg_start) {
    MPF_HIP(hipGetLastError());
    return MPF_OK;
  }
  hipLaunchKernelGGL(k_b_diag, dim3(grid_for(n, 256)), dim3(256), 0, s, n, P.rowptr.ptr, P.colidx.ptr, p->vals.ptr,
                     B.dinv.ptr);
  hipLaunchKernelGGL(k_b_start, gt, bt, 0, s, B.tlo.ptr, B.thi.ptr, B.tmesh.ptr, B.mf.ptr, p->rhs.ptr, B.dinv.ptr,
                     p->du.ptr, B.r.ptr, B.z.ptr, B.pv.ptr, B.part.ptr);
  hipLaunchKernelGGL(k_b_start_scalars, dim3(1), bt, 0, s, B.nmesh, B.mtile.ptr, B.part.ptr, B.ms.ptr, B.mf.ptr,
                     B.cnt.ptr);
  MPF_HIP(hipGetLastError());
  return MPF_OK;
}

// capture the work enqueued by f on the problem's stream into an executable graph
template <typename Fn>
mpf_status capture(mpf_problem* p, hipGraphExec_t* out, Fn f) {
  hipStream_t s = p->stream;
  if (*out) { (void)hipGraphExecDestroy(*out); *out = nullptr; }
  MPF_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  mpf_status st = f();
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(s, &g);
  if (st != MPF_OK) { if (g) (void)hipGraphDestroy(g); return st; }
  MPF_HIP(e);
  const hipError_t e2 = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  MPF_HIP(e2);
  return MPF_OK;
}

mpf_status launch(mpf_problem* p, hipGraphExec_t g, bool use_graph, const std::function<mpf_status()>& eager) {
  if (use_graph) {
    MPF_HIP(hipGraphLaunch(g, p->stream));
    return MPF_OK;
  }
  return eager();
}

mpf_status read_counts(mpf_problem* p, BatchWork& B) {
  MPF_HIP(hipMemcpyAsync(B.cnt_h, B.cnt.ptr, 4 * sizeof(int32_t), hipMemcpyDeviceToHost, p->stream));
  MPF_HIP(hipStreamSynchronize(p->stream));
  return MPF_OK;
}

double secs(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

mpf_status batch_node_offsets(const mpf_problem* p, std::vector<int32_t>& off) {
  if (!p->batch || p->batch->nmesh <= 0) return MPF_ERR_STATE;
  off = p->batch->node_off;
  return MPF_OK;
}

mpf_status batch_elem_offsets(const mpf_problem* p, std::vector<int32_t>& off) {
  if (!p->batch || p->batch->nmesh <= 0) return MPF_ERR_STATE;
  off = p->batch->elem_off;
  return MPF_OK;
}

void batch_free(mpf_problem* p) {
  delete p->batch;
  p->batch = nullptr;
}

mpf_status batch_set(mpf_problem* p, int32_t nmesh, const int32_t* node_off, const int32_t* elem_off,
                     const double* mu, const double* Kinv) {
  if (p->etype != MPF_HEX8 || nmesh <= 0 || !node_off || !elem_off || !mu || !Kinv) return MPF_ERR_ARG;
  // the layout, tiles, per-mesh flags, band key and captured graphs are built for one batch: a
  // second set_batch would leave them sized for the first
  if (p->batch) return MPF_ERR_STATE;
  if (node_off[0] != 0 || elem_off[0] != 0 || node_off[nmesh] != p->nNode || elem_off[nmesh] != p->nElem)
    return MPF_ERR_ARG;
  for (int m = 0; m < nmesh; ++m)
    if (node_off[m + 1] < node_off[m] || elem_off[m + 1] < elem_off[m] || Kinv[m] <= 0.0) return MPF_ERR_ARG;
  // meshes must not share nodes: element nodes of mesh m lie in its node range
  std::vector<int32_t> conn((size_t)p->nElem * p->npe);
  MPF_TRY(p->conn.download(conn.data(), conn.size(), p->stream));
  MPF_HIP(hipStreamSynchronize(p->stream));
  std::vector<int32_t> emat(p->nElem);
  std::vector<double> tab(2 * (size_t)nmesh);
  for (int m = 0; m < nmesh; ++m) {
    tab[2 * m] = mu[m];
    tab[2 * m + 1] = 1.0 / Kinv[m];
    for (int32_t e = elem_off[m]; e < elem_off[m + 1]; ++e) {
      emat[e] = m;
      for (int a = 0; a < p->npe; ++a) {
        const int32_t v = conn[(size_t)e * p->npe + a];
        if (v < node_off[m] || v >= node_off[m + 1]) return MPF_ERR_ARG;
      }
    }
  }
  p->batch = new BatchWork();
  BatchWork& B = *p->batch;
  B.nmesh = nmesh;
  B.node_off.assign(node_off, node_off + nmesh + 1);
  B.elem_off.assign(elem_off, elem_off + nmesh + 1);
  MPF_TRY(B.emat.upload(emat.data(), emat.size(), p->stream));
  MPF_TRY(B.mtab.upload(tab.data(), tab.size(), p->stream));
  MPF_HIP(hipStreamSynchronize(p->stream));
  return MPF_OK;
}

// after the pattern: dof ranges, vectors, tiles
static mpf_status batch_prepare(mpf_problem* p) {
  BatchWork& B = *p->batch;
  if (!B.dof_off.empty()) return MPF_OK;
  std::vector<int32_t> nda((size_t)p->nNode * 3);
  MPF_TRY(p->pat.nda.download(nda.data(), nda.size(), p->stream));
  MPF_HIP(hipStreamSynchronize(p->stream));
  B.dof_off.assign(B.nmesh + 1, 0);
  B.gs_h.clear();
  int32_t c = 0;
  for (int m = 0; m < B.nmesh; ++m) {
    B.dof_off[m] = c;
    for (int32_t v = B.node_off[m]; v < B.node_off[m + 1]; ++v) {
      const int32_t c0 = c;
      for (int d = 0; d < 3; ++d)
        if (nda[3 * (size_t)v + d] >= 0) {
          if (nda[3 * (size_t)v + d] != c) return MPF_ERR_STATE;  // free dofs must be node-major
          ++c;
        }
      if (c > c0) B.gs_h.push_back(c0);
    }
  }
  B.dof_off[B.nmesh] = c;
  B.gs_h.push_back(c);
  if (c == 0) return MPF_ERR_ARG;  // nothing to solve
  {  // row groups share their column lists? (the node-to-node pattern makes them; checked once)
    const int32_t ng = (int32_t)B.gs_h.size() - 1;
    MPF_TRY(B.gs.upload(B.gs_h.data(), B.gs_h.size(), p->stream));
    hipLaunchKernelGGL(k_b_zero_i32, dim3(1), dim3(64), 0, p->stream, p->bad.ptr, 1);
    hipLaunchKernelGGL(k_b_check_groups, dim3(grid_for(ng, 256)), dim3(256), 0, p->stream, ng, B.gs.ptr,
                       p->pat.rowptr.ptr, p->pat.colidx.ptr, p->bad.ptr);
    int32_t bad = 1;
    MPF_HIP(hipMemcpyAsync(&bad, p->bad.ptr, sizeof(int32_t), hipMemcpyDeviceToHost, p->stream));
    MPF_HIP(hipStreamSynchronize(p->stream));
    B.grp_spmv = bad == 0;
  }
  const size_t n = c;
  MPF_TRY(B.dinv.alloc(n));
  MPF_TRY(B.r.alloc(n));
  MPF_TRY(B.z.alloc(n));
  MPF_TRY(B.pv.alloc(n));
  MPF_TRY(B.q.alloc(n));
  MPF_TRY(B.ms.alloc((size_t)B.nmesh * kMS));
  MPF_TRY(B.mf.alloc((size_t)B.nmesh * kMF));
  MPF_TRY(B.cnt.alloc(4));
  MPF_TRY(B.lam_dev.alloc(1));
  hipLaunchKernelGGL(k_b_set_scalar, dim3(1), dim3(1), 0, p->stream, B.lam_dev.ptr, 1.0);  // autotuner's system
  if (!B.cnt_h) MPF_HIP(hipHostMalloc((void**)&B.cnt_h, 4 * sizeof(int32_t), hipHostMallocDefault));
  return build_tiles(B, B.T, p->stream);
}

ElemArgs batch_elem_args(mpf_problem* p, ElemArgs a) {
  if (p->batch && p->batch->nmesh > 0) {
    a.emat = p->batch->emat.ptr;
    a.mtab = p->batch->mtab.ptr;
  }
  return a;
}

mpf_status batch_autotune(mpf_problem* p, int32_t* T_out, int32_t* LPR_out, double* ms_per_iter) {
  if (!p->batch || !p->prepared) return MPF_ERR_STATE;
  MPF_TRY(batch_prepare(p));
  BatchWork& B = *p->batch;
  hipStream_t s = p->stream;
  // a realistic system: assemble at the current state with the lift, start CG on every mesh
  {
    std::vector<int32_t> init((size_t)B.nmesh * kMF, 0);
    for (int m = 0; m < B.nmesh; ++m) init[(size_t)m * kMF + kNA] = 1;
    MPF_HIP(hipMemcpyAsync(B.mf.ptr, init.data(), init.size() * sizeof(int32_t), hipMemcpyHostToDevice, s));
    MPF_HIP(hipStreamSynchronize(s));
  }
  hipEvent_t e0, e1;
  MPF_HIP(hipEventCreate(&e0));
  MPF_HIP(hipEventCreate(&e1));
  // element batch: the row-owner gather with its chunk kernel staging at most cap distinct elements
  // per chunk in LDS (cap 8 .. 32), the gather reading every record directly, or the coloured scatter,
  // each timed over three first-iteration assemblies
  {
    struct Cand { int32_t mode, cap; };
    std::vector<Cand> cands;
    for (int32_t cap : {8, 16, 24, 32}) cands.push_back({MPF_ASSEMBLY_GATHER, cap});
    cands.push_back({MPF_ASSEMBLY_GATHER_DIRECT, p->chunk_cap});
    cands.push_back({MPF_ASSEMBLY_COLORED_SCATTER, p->chunk_cap});
    double abest = 1e300;
    Cand pick = cands[0];
    for (const Cand& c : cands) {
      p->asm_mode = c.mode;
      p->chunk_cap = c.cap;
      if (c.mode != MPF_ASSEMBLY_COLORED_SCATTER) MPF_TRY(ensure_gather_tables(p));
      MPF_TRY(enqueue_assembly(p, true, 1.0, 1));  // warm-up
      MPF_HIP(hipEventRecord(e0, s));
      const int reps = 3;
      for (int k = 0; k < reps; ++k) MPF_TRY(enqueue_assembly(p, true, 1.0, 1));
      MPF_HIP(hipEventRecord(e1, s));
      MPF_HIP(hipEventSynchronize(e1));
      float ms = 0.f;
      MPF_HIP(hipEventElapsedTime(&ms, e0, e1));
      if (ms / reps < abest) { abest = ms / reps; pick = c; }
    }
    p->asm_mode = pick.mode;
    p->chunk_cap = pick.cap;
    B.asm_mode = pick.mode;
    B.asm_cap = pick.cap;
    B.asm_ms = abest;
  }
  const int Ts[] = {512, 1024, 2048, 4096};
  const int Ls[] = {4, 8, 16};
  double best = 1e300;
  int bT = B.T, bL = B.LPR;
  for (int T : Ts) {
    MPF_TRY(build_tiles(B, T, s));
    MPF_TRY(enqueue_newton_head(p, B, true, 1.0, 0.0, 1e-300));
    for (int L : Ls) {
      B.LPR = L;
      MPF_TRY(enqueue_cg_iter(p, B, 1 << 30));  // warm-up
      MPF_HIP(hipEventRecord(e0, s));
      const int reps = 5;
      for (int k = 0; k < reps; ++k) MPF_TRY(enqueue_cg_iter(p, B, 1 << 30));
      MPF_HIP(hipEventRecord(e1, s));
      MPF_HIP(hipEventSynchronize(e1));
      float ms = 0.f;
      MPF_HIP(hipEventElapsedTime(&ms, e0, e1));
      if (ms / reps < best) { best = ms / reps; bT = T; bL = L; }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  MPF_TRY(build_tiles(B, bT, s));
  B.LPR = bL;
  B.tuned = true;
  B.release_graphs();
  if (T_out) *T_out = bT;
  if (LPR_out) *LPR_out = bL;
  if (ms_per_iter) *ms_per_iter = best;
  return MPF_OK;
}

static mpf_status batch_configure(mpf_problem* p, const mpf_batch_opts* o, mpf_batch_report* rep) {
  if (!p->batch || !p->prepared || !o || !rep) return MPF_ERR_STATE;
  if (o->itmax <= 0 || o->cg_maxit <= 0 || o->solver < 0 || o->solver > 1) return MPF_ERR_ARG;
  MPF_TRY(batch_prepare(p));
  BatchWork& B = *p->batch;
  if (o->solver == 1) MPF_TRY(band_prepare(p, B.node_off));
  if (o->chunk > 0 && o->chunk != B.chunk) {
    B.chunk = o->chunk;
    B.release_graphs();
  }
  if (o->tile > 0 && o->tile != B.T) {
    MPF_TRY(build_tiles(B, o->tile, p->stream));
    B.release_graphs();
  }
  if (o->lanes_per_row == 4 || o->lanes_per_row == 8 || o->lanes_per_row == 16) {
    if (o->lanes_per_row != B.LPR) B.release_graphs();
    B.LPR = o->lanes_per_row;
  }
  rep->tile = B.T;
  rep->lanes_per_row = B.LPR;
  rep->chunk = B.chunk;
  return MPF_OK;
}

// residual at the plain start of a step (with the lift) -> R0 of every mesh
static mpf_status enqueue_r0(mpf_problem* p, BatchWork& B, double lam) {
  hipStream_t s = p->stream;
  const dim3 gt(std::max(B.ntiles, 1)), bt(kB);
  hipLaunchKernelGGL(k_b_zero_i32, dim3(1), dim3(64), 0, s, p->bad.ptr, 1);
  MPF_TRY(enqueue_assembly(p, true, lam, 0, B.lam_dev.ptr));
  hipLaunchKernelGGL(k_b_dot, gt, bt, 0, s, B.tlo.ptr, B.thi.ptr, p->rhs.ptr, p->rhs.ptr, B.part.ptr, 0);
  hipLaunchKernelGGL(k_b_newton, dim3(1), bt, 0, s, B.nmesh, B.mtile.ptr, B.part.ptr, B.ms.ptr, B.mf.ptr, 1, 0.0,
                     1.0, B.cnt.ptr);
  MPF_HIP(hipGetLastError());
  return MPF_OK;
}

static mpf_status reset_mesh_flags(mpf_problem* p, BatchWork& B) {
  std::vector<int32_t> init((size_t)B.nmesh * kMF, 0);
  for (int m = 0; m < B.nmesh; ++m) init[(size_t)m * kMF + kNA] = 1;
  MPF_HIP(hipMemcpyAsync(B.mf.ptr, init.data(), init.size() * sizeof(int32_t), hipMemcpyHostToDevice, p->stream));
  MPF_HIP(hipStreamSynchronize(p->stream));
  return MPF_OK;
}

// One load step of every mesh at load factor lam from the current converged state.  guess (device,
// [nNode*3], may be null): displacement to start Newton from; R0 is still taken at the plain start.
static mpf_status batch_step(mpf_problem* p, double lam, const mpf_batch_opts* o, const double* guess,
                             int32_t* iters_row, double* mesh_relres, mpf_batch_report* rep) {
  BatchWork& B = *p->batch;
  hipStream_t s = p->stream;
  const bool G = o->use_graph != 0;
  const bool band = o->solver == 1;  // exact corrections by banded LDL^T per mesh, else Jacobi-PCG
  const size_t n3 = (size_t)p->nNode * 3;
  auto t_all = std::chrono::steady_clock::now();
  // saveSolution, setBoundaryConditions
  MPF_HIP(hipMemcpyAsync(p->dispPrev.ptr, p->disp.ptr, n3 * sizeof(double), hipMemcpyDeviceToDevice, s));
  MPF_TRY(mpf_set_bc(p, lam));
  MPF_TRY(reset_mesh_flags(p, B));
  hipLaunchKernelGGL(k_b_set_scalar, dim3(1), dim3(1), 0, s, B.lam_dev.ptr, lam);
  MPF_HIP(hipGetLastError());
  // one Newton iteration: [update of the previous solve] + head + (K CG iterations | direct solve)
  auto iter_work = [&](int v, int kidx) -> mpf_status {
    if (v >= 2) MPF_TRY(do_update(p, v == 2, p->p0 ? B.emat.ptr : nullptr, B.mf.ptr + kNA, kMF));
    MPF_TRY(enqueue_newton_head(p, B, v == 0, lam, o->rtol, o->cg_rtol, !band));
    if (kidx > 0) {
      if (band) MPF_TRY(band_enqueue_solve(p, B.mf.ptr + kNA, kMF));
      else
        for (int k = 0; k < kKs[kidx]; ++k) MPF_TRY(enqueue_cg_iter(p, B, o->cg_maxit));
    }
    return MPF_OK;
  };
  auto chunk_work = [&]() -> mpf_status {
    for (int k = 0; k < B.chunk; ++k) MPF_TRY(enqueue_cg_iter(p, B, o->cg_maxit));
    return MPF_OK;
  };
  if (G && (B.g_solver != o->solver || B.g_rtol != o->rtol || B.g_cg_rtol != o->cg_rtol ||
            B.g_cg_maxit != o->cg_maxit)) {
    B.release_graphs();
    B.g_rtol = o->rtol;
    B.g_cg_rtol = o->cg_rtol;
    B.g_cg_maxit = o->cg_maxit;
    B.g_solver = o->solver;
  }
  // graph of (variant, K index), captured on first use
  auto iter_graph = [&](int v, int kidx) -> mpf_status {
    if (!G || B.g_it[v][kidx]) return MPF_OK;
    auto tc = std::chrono::steady_clock::now();
    MPF_TRY(capture(p, &B.g_it[v][kidx], [&] { return iter_work(v, kidx); }));
    rep->seconds_capture += secs(tc);
    return MPF_OK;
  };
  if (G && !B.g_chunk && !band) {
    auto tc = std::chrono::steady_clock::now();
    MPF_TRY(capture(p, &B.g_chunk, chunk_work));
    rep->seconds_capture += secs(tc);
  }
  if (G && guess && !B.g_r0) MPF_TRY(capture(p, &B.g_r0, [&] { return enqueue_r0(p, B, lam); }));
  if (guess) {
    auto ta = std::chrono::steady_clock::now();
    MPF_TRY(launch(p, B.g_r0, G, [&] { return enqueue_r0(p, B, lam); }));
    ++rep->graph_launches;
    MPF_HIP(hipMemcpyAsync(p->disp.ptr, guess, n3 * sizeof(double), hipMemcpyDeviceToDevice, s));
    const int32_t nbc = (int32_t)p->bc_dof_h.size();
    if (nbc > 0)
      hipLaunchKernelGGL(k_b_set_prescribed, dim3(grid_for(nbc, 256)), dim3(256), 0, s, nbc, p->bc_dof.ptr,
                         p->bc_val.ptr, lam, p->disp.ptr);
    MPF_TRY(reset_mesh_flags(p, B));
    rep->seconds_assembly += secs(ta);
  }
  mpf_status result = MPF_OK;
  bool step_conv = false;
  int kpred = band ? 1 : std::max(1, std::min(B.k_pred, BatchWork::kNK - 1));
  int prev_first = -1;  // first-iteration flag of the solve whose update is pending
  for (int iter = 1; iter <= o->itmax + 1; ++iter) {
    const int f = (iter == 1 && !guess) ? 1 : 0;
    const int v = iter == 1 ? (f ? 0 : 1) : (prev_first ? 2 : 3);
    const int kidx = iter > o->itmax ? 0 : kpred;
    auto ta = std::chrono::steady_clock::now();
    MPF_TRY(iter_graph(v, kidx));
    MPF_TRY(launch(p, B.g_it[v][kidx], G, [&] { return iter_work(v, kidx); }));
    p->assembled = true;  // the next update (Q1/P0: pressure recovery) reads this assembly
    ++rep->graph_launches;
    ++rep->iteration_graphs;
    int32_t bad = 0;
    MPF_HIP(hipMemcpyAsync(&bad, p->bad.ptr, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    MPF_TRY(read_counts(p, B));
    rep->seconds_assembly += secs(ta);
    if (bad) { result = MPF_ERR_NEG_JACOBIAN; break; }
    if (B.cnt_h[2] > 0) { result = MPF_ERR_DIVERGED; break; }
    if (iter == 1) {  // residual at the start point, relative to the plain-start R0
      std::vector<double> msh((size_t)B.nmesh * kMS);
      MPF_HIP(hipMemcpy(msh.data(), B.ms.ptr, msh.size() * sizeof(double), hipMemcpyDeviceToHost));
      double sum = 0.0, mx = 0.0;
      for (int m = 0; m < B.nmesh; ++m) {
        sum += msh[(size_t)m * kMS + kREL];
        mx = std::max(mx, msh[(size_t)m * kMS + kREL]);
      }
      rep->start_relres_mean = sum / B.nmesh;
      rep->start_relres_max = mx;
    }
    if (B.cnt_h[1] == 0) { step_conv = true; break; }
    if (iter > o->itmax) break;
    ++rep->newton_iters;
    auto ts = std::chrono::steady_clock::now();
    if (band) {
      int32_t pbad = 0;
      MPF_HIP(hipMemcpyAsync(&pbad, band_bad_flag(p), sizeof(int32_t), hipMemcpyDeviceToHost, s));
      MPF_HIP(hipStreamSynchronize(s));
      if (pbad) { result = MPF_ERR_DIVERGED; break; }
    } else {
      int its = kKs[kidx];
      while (B.cnt_h[0] > 0 && its < o->cg_maxit) {
        MPF_TRY(launch(p, B.g_chunk, G, chunk_work));
        ++rep->graph_launches;
        ++rep->chunk_graphs;
        its += B.chunk;
        MPF_TRY(read_counts(p, B));
      }
      const int most = B.cnt_h[3];
      rep->cg_iters_max += most;
      // next solve's K: the smallest listed count >= 1.25 x this solve's slowest mesh
      int k = 1;
      while (k < BatchWork::kNK - 1 && kKs[k] < most + most / 4) ++k;
      if (iter == 1) B.k_pred = k;
      kpred = k;
    }
    prev_first = f;
    rep->seconds_solve += secs(ts);
  }
  p->assembled = true;  // vals / rhs hold the step's last assembly
  std::vector<int32_t> mfh((size_t)B.nmesh * kMF);
  std::vector<double> msh((size_t)B.nmesh * kMS);
  MPF_HIP(hipMemcpyAsync(mfh.data(), B.mf.ptr, mfh.size() * sizeof(int32_t), hipMemcpyDeviceToHost, s));
  MPF_HIP(hipMemcpyAsync(msh.data(), B.ms.ptr, msh.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  MPF_HIP(hipStreamSynchronize(s));
  for (int m = 0; m < B.nmesh; ++m) {
    if (iters_row) iters_row[m] = mfh[(size_t)m * kMF + kNIT] + 1;  // assemblies to convergence
    if (mesh_relres) mesh_relres[m] = msh[(size_t)m * kMS + kREL];
    rep->cg_iters_total += mfh[(size_t)m * kMF + kCGT] + mfh[(size_t)m * kMF + kCGIT];
  }
  rep->seconds += secs(t_all);
  if (result != MPF_OK) return result;
  if (!step_conv) return MPF_ERR_NOT_CONVERGED;
  rep->steps_done += 1;
  return MPF_OK;
}

mpf_status batch_newton_step(mpf_problem* p, double lam, const mpf_batch_opts* o, const double* guess,
                             int32_t* mesh_iters, double* mesh_relres, mpf_batch_report* rep) {
  MPF_TRY(batch_configure(p, o, rep));
  return batch_step(p, lam, o, guess, mesh_iters, mesh_relres, rep);
}

__global__ void k_b_free_to_nodes(int64_t n3, const int32_t* __restrict__ nda, const double* __restrict__ du,
                                  double* __restrict__ out) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n3) return;
  const int32_t eq = nda[j];
  out[j] = eq >= 0 ? du[eq] : 0.0;
}

// Tangent increment for the next load step: out = K^-1 R(u; lam) on the free dofs (0 on prescribed
// dofs), R the residual at the current state with the load and prescribed values of lam (the lift
// of the first Newton iteration), K the factors each mesh's last banded direct solve left -- the
// tangent at that mesh's last Newton iterate.  At a converged u_k this is h u'(lam_k) to first order.
mpf_status batch_tangent_increment(mpf_problem* p, double lam, double* out) {
  if (!p || !out) return MPF_ERR_ARG;
  if (!p->batch || !p->prepared) return MPF_ERR_STATE;
  hipStream_t s = p->stream;
  MPF_TRY(mpf_set_bc(p, lam));
  hipLaunchKernelGGL(k_b_zero_i32, dim3(1), dim3(64), 0, s, p->bad.ptr, 1);
  MPF_TRY(enqueue_assembly(p, true, lam, 0));
  MPF_TRY(band_enqueue_resolve(p));
  const int64_t n3 = (int64_t)p->nNode * 3;
  hipLaunchKernelGGL(k_b_free_to_nodes, dim3(grid_for(n3, 256)), dim3(256), 0, s, n3, p->pat.nda.ptr, p->du.ptr, out);
  MPF_HIP(hipGetLastError());
  int32_t bad = 0;
  MPF_HIP(hipMemcpyAsync(&bad, p->bad.ptr, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  MPF_HIP(hipStreamSynchronize(s));
  return bad ? MPF_ERR_NEG_JACOBIAN : MPF_OK;
}

mpf_status batch_newton_solve(mpf_problem* p, const mpf_batch_opts* o, int32_t* mesh_iters, double* mesh_relres,
                              mpf_batch_report* rep) {
  if (!rep) return MPF_ERR_ARG;
  *rep = mpf_batch_report{};
  MPF_TRY(batch_configure(p, o, rep));
  if (o->nsteps <= 0 || o->dt <= 0.0) return MPF_ERR_ARG;
  const int nmesh = p->batch->nmesh;
  for (int step = 0; step < o->nsteps; ++step) {
    const double lam = (step + 1) * o->dt;  // lambda(t) = t
    MPF_TRY(batch_step(p, lam, o, nullptr, mesh_iters ? mesh_iters + (size_t)step * nmesh : nullptr, mesh_relres,
                       rep));
  }
  return MPF_OK;
}

}  // namespace mpf

extern "C" {

mpf_status mpf_set_batch(mpf_problem* p, int32_t nmesh, const int32_t* node_off, const int32_t* elem_off,
                         const double* mu, const double* Kinv) {
  if (!p || p->etype == 0) return MPF_ERR_ARG;
  return mpf::batch_set(p, nmesh, node_off, elem_off, mu, Kinv);
}

mpf_status mpf_batch_time_direct(mpf_problem* p, int32_t reps, double* ms_per_solve) {
  if (!p || !p->batch || reps < 1 || !ms_per_solve) return MPF_ERR_ARG;
  if (!p->band || !p->assembled) return MPF_ERR_STATE;
  hipStream_t s = p->stream;
  hipEvent_t e0, e1;
  MPF_HIP(hipEventCreate(&e0));
  MPF_HIP(hipEventCreate(&e1));
  MPF_HIP(hipEventRecord(e0, s));
  for (int r = 0; r < reps; ++r) MPF_TRY(mpf::band_enqueue_solve(p, nullptr, 0));
  MPF_HIP(hipEventRecord(e1, s));
  MPF_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  MPF_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *ms_per_solve = ms / reps;
  return MPF_OK;
}

mpf_status mpf_batch_assembly_choice(const mpf_problem* p, int32_t* mode, int32_t* stage_cap, double* ms_per_assembly) {
  if (!p || !p->batch) return MPF_ERR_ARG;
  if (p->batch->asm_mode < 0) return MPF_ERR_STATE;
  if (mode) *mode = p->batch->asm_mode;
  if (stage_cap) *stage_cap = p->batch->asm_cap;
  if (ms_per_assembly) *ms_per_assembly = p->batch->asm_ms;
  return MPF_OK;
}

mpf_status mpf_batch_autotune(mpf_problem* p, int32_t* tile, int32_t* lanes_per_row, double* ms_per_cg_iter) {
  if (!p || p->etype == 0) return MPF_ERR_ARG;
  return mpf::batch_autotune(p, tile, lanes_per_row, ms_per_cg_iter);
}

mpf_status mpf_batch_newton_solve(mpf_problem* p, const mpf_batch_opts* o, int32_t* mesh_iters, double* mesh_relres,
                                  mpf_batch_report* rep) {
  if (!p || p->etype == 0) return MPF_ERR_ARG;
  return mpf::batch_newton_solve(p, o, mesh_iters, mesh_relres, rep);
}

mpf_status mpf_batch_newton_step(mpf_problem* p, double lam, const mpf_batch_opts* o, const double* guess,
                                 int32_t* mesh_iters, double* mesh_relres, mpf_batch_report* rep) {
  if (!p || p->etype == 0) return MPF_ERR_ARG;
  return mpf::batch_newton_step(p, lam, o, guess, mesh_iters, mesh_relres, rep);
}

mpf_status mpf_batch_tangent_increment(mpf_problem* p, double lam, double* du_nodes) {
  return mpf::batch_tangent_increment(p, lam, du_nodes);
}

}  // extern "C"
