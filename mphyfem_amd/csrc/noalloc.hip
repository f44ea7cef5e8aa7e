// noalloc.hip -- mpf_solve_noalloc: the Jacobi-preconditioned CG of mpf_solve for callers that
// cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
// forbid hipMalloc).  No counterpart in the reference.
//
// Every work vector and partial sum lives in one file-scope device array, `scratch`, of fixed
// size: systems up to kNoAllocMaxDofs rows.  The array's device address is looked up once per
// device (hipGetSymbolAddress, with that device current) and only that address is used from the
// host.  All calls on one GPU share the array and leave their contents in it; every value a solve
// reads is written by that solve first.  Calls that overlap on one GPU race.
#include <cmath>
#include <cstring>
#include <mutex>

#include "mpf_internal.hpp"

namespace mpf {
mpf_status spmv(const int32_t* rowptr, const int32_t* colidx, const double* vals, int64_t n, int64_t nnz,
                const double* x, double* y, double* part, unsigned np, const double* sc, hipStream_t s);
}

namespace {

constexpr int64_t kNoAllocMaxDofs = MPF_NOALLOC_MAX_DOFS;  // 5 vectors of 4 MiB
constexpr int kParts = 1024;                  // workgroups of the vector kernels (partial sums)
constexpr int kB = 256;
enum { S_RZ = 0, S_PQ = 1, S_RR = 2, S_BB = 3, S_TOL2 = 4, S_DONE = 5, S_ITS = 6, S_MAXIT = 7, S_BETA = 8,
       S_N = 16 };  // S_DONE is slot 5, the same slot mpf::spmv's early exit reads
constexpr int64_t kScratchDoubles = 5 * kNoAllocMaxDofs + 2 * kParts + S_N;

__device__ __align__(256) double scratch[kScratchDoubles];

struct Layout {
  double *r, *z, *p, *q, *dinv, *part0, *part1, *sc;
};

Layout layout(double* base) {
  Layout L;
  L.r = base;
  L.z = L.r + kNoAllocMaxDofs;
  L.p = L.z + kNoAllocMaxDofs;
  L.q = L.p + kNoAllocMaxDofs;
  L.dinv = L.q + kNoAllocMaxDofs;
  L.part0 = L.dinv + kNoAllocMaxDofs;
  L.part1 = L.part0 + kParts;
  L.sc = L.part1 + kParts;
  return L;
}

// Device address of `scratch` on the current device, looked up once per device.
mpf_status scratch_ptr(double** out) {
  static std::mutex mu;
  static double* ptrs[64] = {};
  int dev = 0;
  MPF_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return MPF_ERR_ARG;
  std::lock_guard<std::mutex> lock(mu);
  if (!ptrs[dev]) {
    void* q = nullptr;
    MPF_HIP(hipGetSymbolAddress(&q, HIP_SYMBOL(scratch)));
    ptrs[dev] = static_cast<double*>(q);
  }
  *out = ptrs[dev];
  return MPF_OK;
}

__device__ __forceinline__ double na_wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ __forceinline__ double na_block_sum(double v, double* sh) {
  v = na_wave_sum(v);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += sh[i];
  __syncthreads();
  return s;
}

// x = 0, r = b, dinv = 1 / diag (Jacobi) or 1, z = dinv r, p = z; partials r.z and b.b
__global__ void __launch_bounds__(kB) k_na_init(int64_t n, const int32_t* __restrict__ rowptr,
                                                const int32_t* __restrict__ colidx, const double* __restrict__ vals,
                                                const double* __restrict__ b, double* __restrict__ x, Layout L,
                                                int jacobi) {
  __shared__ double sh[kB / 64];
  double rz = 0.0, bb = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kB) {
    double di = 1.0;
    if (jacobi) {
      double d = 0.0;
      for (int k = rowptr[i]; k < rowptr[i + 1]; ++k)
        if (colidx[k] == i) d = vals[k];
      di = d != 0.0 ? 1.0 / d : 1.0;
    }
    const double bi = b[i];
    x[i] = 0.0;
    L.r[i] = bi;
    L.dinv[i] = di;
    const double zi = di * bi;
    L.z[i] = zi;
    L.p[i] = zi;
    rz += bi * zi;
    bb += bi * bi;
  }
  rz = na_block_sum(rz, sh);
  bb = na_block_sum(bb, sh);
  if (threadIdx.x == 0) {
    L.part0[blockIdx.x] = rz;
    L.part1[blockIdx.x] = bb;
  }
}

__device__ double na_ordered_sum(const double* part, int np, double* sh) {
  double v = 0.0;
  for (int i = threadIdx.x; i < np; i += kB) v += part[i];
  return na_block_sum(v, sh);
}

__global__ void __launch_bounds__(kB) k_na_start(Layout L, int np, double rtol, int maxit) {
  __shared__ double sh[kB / 64];
  const double rz = na_ordered_sum(L.part0, np, sh);
  const double bb = na_ordered_sum(L.part1, np, sh);
  if (threadIdx.x == 0) {
    double* sc = L.sc;
    sc[S_RZ] = rz;
    sc[S_BB] = bb;
    sc[S_RR] = bb;
    sc[S_TOL2] = rtol * rtol * bb;
    sc[S_ITS] = 0.0;
    sc[S_MAXIT] = (double)maxit;
    sc[S_BETA] = 0.0;
    sc[S_PQ] = 0.0;
    sc[S_DONE] = (bb == 0.0 || maxit <= 0) ? 1.0 : 0.0;
  }
}

__global__ void __launch_bounds__(kB) k_na_pq(Layout L, int np) {
  __shared__ double sh[kB / 64];
  if (L.sc[S_DONE] != 0.0) return;
  const double pq = na_ordered_sum(L.part0, np, sh);
  if (threadIdx.x == 0) L.sc[S_PQ] = pq;
}

// x += alpha p, r -= alpha q, z = dinv r; partials r.z and r.r
__global__ void __launch_bounds__(kB) k_na_update(int64_t n, double* __restrict__ x, Layout L) {
  __shared__ double sh[kB / 64];
  const double* sc = L.sc;
  if (sc[S_DONE] != 0.0) return;
  const double alpha = sc[S_RZ] / sc[S_PQ];
  double rz = 0.0, rr = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kB) {
    x[i] += alpha * L.p[i];
    const double ri = L.r[i] - alpha * L.q[i];
    L.r[i] = ri;
    const double zi = L.dinv[i] * ri;
    L.z[i] = zi;
    rz += ri * zi;
    rr += ri * ri;
  }
  rz = na_block_sum(rz, sh);
  rr = na_block_sum(rr, sh);
  if (threadIdx.x == 0) {
    L.part0[blockIdx.x] = rz;
    L.part1[blockIdx.x] = rr;
  }
}

__global__ void __launch_bounds__(kB) k_na_check(Layout L, int np) {
  __shared__ double sh[kB / 64];
  double* sc = L.sc;
  if (sc[S_DONE] != 0.0) return;
  const double rz = na_ordered_sum(L.part0, np, sh);
  const double rr = na_ordered_sum(L.part1, np, sh);
  if (threadIdx.x == 0) {
    const double its = sc[S_ITS] + 1.0;
    sc[S_ITS] = its;
    sc[S_RR] = rr;
    sc[S_BETA] = rz / sc[S_RZ];
    sc[S_RZ] = rz;
    if (rr <= sc[S_TOL2] || its >= sc[S_MAXIT] || !(rr == rr)) sc[S_DONE] = 1.0;
  }
}

__global__ void __launch_bounds__(kB) k_na_p(int64_t n, Layout L) {
  if (L.sc[S_DONE] != 0.0) return;
  const double beta = L.sc[S_BETA];
  for (int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kB)
    L.p[i] = L.z[i] + beta * L.p[i];
}

}  // namespace

using namespace mpf;

extern "C" mpf_status mpf_solve_noalloc(mpf_problem* p, const mpf_krylov_opts* o, mpf_krylov_report* rep) {
  if (!p || !o || !rep) return MPF_ERR_ARG;
  if (o->ksp != MPF_KSP_CG || (o->pc != MPF_PC_NONE && o->pc != MPF_PC_JACOBI)) return MPF_ERR_ARG;
  const int64_t n = (int64_t)p->pat.dispDOF + p->pat.presDOF;
  if (n <= 0 || n > kNoAllocMaxDofs) return MPF_ERR_ARG;  // the scratch holds kNoAllocMaxDofs rows
  if (!p->assembled) return MPF_ERR_STATE;
  if (p->du.n < (size_t)n) return MPF_ERR_STATE;  // the solution buffer exists once the pattern does
  double* base = nullptr;
  MPF_TRY(scratch_ptr(&base));
  const Layout L = layout(base);
  hipStream_t s = p->stream;
  int64_t g = (n + kB - 1) / kB;
  const int np = (int)(g > kParts ? kParts : (g < 1 ? 1 : g));
  double* x = p->du.ptr;
  const int32_t *rp = p->pat.rowptr.ptr, *ci = p->pat.colidx.ptr;
  const double* vals = p->vals.ptr;
  hipLaunchKernelGGL(k_na_init, dim3(np), dim3(kB), 0, s, n, rp, ci, vals, (const double*)p->rhs.ptr, x, L,
                     o->pc == MPF_PC_JACOBI ? 1 : 0);
  hipLaunchKernelGGL(k_na_start, dim3(1), dim3(kB), 0, s, L, np, o->rtol, o->maxit);
  double hs[S_N];
  int chunk = 16;
  for (;;) {
    for (int c = 0; c < chunk; ++c) {
      MPF_TRY(spmv(rp, ci, vals, n, p->pat.nnz, L.p, L.q, L.part0, (unsigned)np, L.sc, s));
      hipLaunchKernelGGL(k_na_pq, dim3(1), dim3(kB), 0, s, L, np);
      hipLaunchKernelGGL(k_na_update, dim3(np), dim3(kB), 0, s, n, x, L);
      hipLaunchKernelGGL(k_na_check, dim3(1), dim3(kB), 0, s, L, np);
      hipLaunchKernelGGL(k_na_p, dim3(np), dim3(kB), 0, s, n, L);
    }
    MPF_HIP(hipGetLastError());
    MPF_HIP(hipMemcpyAsync(hs, L.sc, sizeof(hs), hipMemcpyDeviceToHost, s));
    MPF_HIP(hipStreamSynchronize(s));
    if (hs[S_DONE] != 0.0) break;
    if (chunk < 128) chunk *= 2;
  }
  std::memset(rep, 0, sizeof(*rep));
  rep->iterations = (int32_t)hs[S_ITS];
  rep->relres = hs[S_BB] > 0.0 ? std::sqrt(hs[S_RR] / hs[S_BB]) : 0.0;
  rep->converged = (hs[S_BB] == 0.0 || hs[S_RR] <= hs[S_TOL2]) ? 1 : 0;
  p->solved = true;
  return rep->converged ? MPF_OK : MPF_ERR_NOT_CONVERGED;
}
