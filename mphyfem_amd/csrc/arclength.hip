// arclength.hip -- arc-length continuation on the device: femSolidmechanics::solveWithArclength
// (femSolidmechanics2.cpp:855-1039) with SolverPetsc::solveArclengthSystem (SolverPetsc.cpp:671-757).
//
// Per attempt (MyTime::update adds dt to the time, which only counts attempts here):
//   saveSolution        dispPrev2 = dispPrev, dispPrev = disp (SolutionDataSolid.cpp:170-184)
//   predictor           steps > 1: disp = (1+f) dispPrev - f dispPrev2, lam = (1+f) lamPrev - f lamPrev2,
//                       f = ds / dsPrev (with saveSolution just before, the displacement predictor is
//                       the last state; the load factor is extrapolated)
//   DuFull = disp - dispPrev, Dl = lam - lamPrev
//   Newton (iter < iterationsMax): R = F_int-part + lam Fext; ||R|| < tol ends the step; else
//       du' = K^-1 R, du1 = K^-1 Fext (free dofs), with Du = DuFull on the free dofs,
//       A = Du.Du + Dl^2 F.F - ds^2, a = 2 Du, b = 2 Dl F.F (first step: A = 0, a = 0, b = 1),
//       dl = (a.(-du') - A) / (b + a.du1); disp, DuFull += du' + dl du1 on the free dofs;
//       lam += dl, Dl += dl
//   converged: first step sets ds = sqrt(DuFull.DuFull + lam^2 F.F) (max ds, min ds / 1024); the
//       previous states shift; ds doubles (within [min, max]) after two converged steps in a row
//   failed: ds *= 0.5 (after a converged step) or 0.25, not below the minimum
// F is the traction patches' force at load factor 1 (F.F over all nodal dofs, as Fext.dot(Fext)).
// Both solves of an iteration use the same matrix: with MPF_KSP_BAND the second is a re-solve with
// the factors of the first.  Reductions are two-stage with fixed partials (deterministic).
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "mpf_internal.hpp"

mpf_status assemble_sync(mpf_problem* p, bool first_iter, double lam, int scatterK);  // problem.hip
mpf_status rhs_norm(mpf_problem* p, double* out);                                    // problem.hip
mpf_status do_set_bc(mpf_problem* p, double lam);                                    // problem.hip

namespace mpf {
namespace {

constexpr int kB = 256;
constexpr int kParts = 1024;

__device__ __forceinline__ double al_block_sum(double v, double* sh) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += sh[i];
  __syncthreads();
  return s;
}

// part[b] = sum over the block's i of x[gx ? gx[i] : i] * y[i] (gx: gather the first operand)
__global__ void __launch_bounds__(kB) k_al_dot(int64_t n, const double* __restrict__ x, const int32_t* __restrict__ gx,
                                               const double* __restrict__ y, double* __restrict__ part) {
  __shared__ double sh[kB / 64];
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kB)
    s += x[gx ? gx[i] : i] * y[i];
  s = al_block_sum(s, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void k_al_finish(const double* __restrict__ part, int np, double* __restrict__ out) {
  __shared__ double sh[kB / 64];
  double v = 0.0;
  for (int i = threadIdx.x; i < np; i += kB) v += part[i];
  v = al_block_sum(v, sh);
  if (threadIdx.x == 0) *out = v;
}

// dispPrev2 = dispPrev, dispPrev = disp
__global__ void k_al_save(int64_t n, const double* __restrict__ disp, double* __restrict__ prev,
                          double* __restrict__ prev2) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  prev2[i] = prev[i];
  prev[i] = disp[i];
}

// disp = (1+f) prev - f prev2 (predictor, when f is given), DuFull = disp - prev
__global__ void k_al_predict(int64_t n, int predict, double f, double* __restrict__ disp, const double* __restrict__ prev,
                             const double* __restrict__ prev2, double* __restrict__ Du) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (predict) disp[i] = (1.0 + f) * prev[i] - f * prev2[i];
  Du[i] = disp[i] - prev[i];
}

// rhs = Fext on the free dofs (FextShort)
__global__ void k_al_fshort(int32_t nfree, const int32_t* __restrict__ assy, const double* __restrict__ F,
                            double* __restrict__ rhs) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nfree) rhs[i] = F[assy[i]];
}

// v = du' + dl du1 on the free dofs: disp and DuFull += v
__global__ void k_al_update(int32_t nfree, const int32_t* __restrict__ assy, const double* __restrict__ dup,
                            const double* __restrict__ du1, double dl, double* __restrict__ disp,
                            double* __restrict__ Du) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nfree) return;
  const double v = dup[i] + dl * du1[i];
  const int32_t d = assy[i];
  disp[d] += v;
  Du[d] += v;
}

struct AlWork {
  DevBuf<double> prev, prev2, Du, dup, part, sc;
};

unsigned nparts(int64_t n) {
  int64_t g = (n + kB - 1) / kB;
  return (unsigned)(g > kParts ? kParts : (g < 1 ? 1 : g));
}

// x . y (gx: gather x) into sc[slot], enqueued
void enqueue_dot(hipStream_t s, int64_t n, const double* x, const int32_t* gx, const double* y, AlWork& W, int slot) {
  const unsigned np = nparts(n);
  hipLaunchKernelGGL(k_al_dot, dim3(np), dim3(kB), 0, s, n, x, gx, y, W.part.ptr);
  hipLaunchKernelGGL(k_al_finish, dim3(1), dim3(kB), 0, s, W.part.ptr, (int)np, W.sc.ptr + slot);
}

}  // namespace
}  // namespace mpf

using namespace mpf;

extern "C" mpf_status mpf_arclength_solve(mpf_problem* p, const mpf_arclength_opts* o, double* lam_out, double* ds_out,
                                          mpf_arclength_report* rep) {
  if (!p || p->etype == 0 || !o || !rep || o->dt <= 0.0 || o->max_steps < 0 || o->max_iters < 1) return MPF_ERR_ARG;
  if (!p->prepared) return MPF_ERR_STATE;
  // displacement elements (the reference's loop leaves the pressure dofs unchanged), a traction
  // load, homogeneous prescribed values (the loop never applies prescribed increments)
  if (p->etype == MPF_TET10P1 || p->p0 || !p->has_fext || p->batch) return MPF_ERR_ARG;
  for (double v : p->bc_val_h)
    if (v != 0.0) return MPF_ERR_ARG;
  if (o->lin.ksp != MPF_KSP_CG && o->lin.ksp != MPF_KSP_GMRES && o->lin.ksp != MPF_KSP_BAND) return MPF_ERR_ARG;
  std::memset(rep, 0, sizeof(*rep));
  const auto t0 = std::chrono::steady_clock::now();
  hipStream_t s = p->stream;
  const int64_t n3 = (int64_t)p->nNode * 3;
  const int32_t nfree = p->pat.dispDOF;
  const int32_t* assy = p->pat.assy.ptr;
  AlWork W;
  MPF_TRY(W.prev.alloc(n3));
  MPF_TRY(W.prev2.alloc(n3));
  MPF_TRY(W.Du.alloc(n3));
  MPF_TRY(W.dup.alloc(nfree > 0 ? nfree : 1));
  MPF_TRY(W.part.alloc(kParts));
  MPF_TRY(W.sc.alloc(8));
  MPF_TRY(p->du.alloc(nfree > 0 ? nfree : 1));
  const unsigned g3 = (unsigned)((n3 + 255) / 256), gf = (unsigned)((nfree + 255) / 256 > 0 ? (nfree + 255) / 256 : 1);
  MPF_HIP(hipMemcpyAsync(W.prev.ptr, p->disp.ptr, n3 * sizeof(double), hipMemcpyDeviceToDevice, s));
  MPF_HIP(hipMemcpyAsync(W.prev2.ptr, p->disp.ptr, n3 * sizeof(double), hipMemcpyDeviceToDevice, s));
  // F.F over all nodal dofs
  enqueue_dot(s, n3, p->fext_unit.ptr, nullptr, p->fext_unit.ptr, W, 0);
  double FtF = 0.0;
  MPF_HIP(hipMemcpyAsync(&FtF, W.sc.ptr, sizeof(double), hipMemcpyDeviceToHost, s));
  MPF_HIP(hipStreamSynchronize(s));
  const double EPS = 1e-10;
  double loadFactor = o->dt, lamPrev = 0.0, lamPrev2 = 0.0;
  double ds = 0.0, dsPrev = 0.0, dsMax = 0.0, dsMin = 0.0;
  bool conv = false, convPrev = false;
  double cur = 0.0;
  int steps = 1;
  mpf_krylov_opts lin = o->lin;
  while (cur <= o->final_time - EPS && steps <= o->max_steps) {
    cur += o->dt;
    rep->attempts += 1;
    hipLaunchKernelGGL(k_al_save, dim3(g3), dim3(256), 0, s, n3, p->disp.ptr, W.prev.ptr, W.prev2.ptr);
    convPrev = conv;
    conv = false;
    const bool later = steps > 1;
    double f = 0.0;
    if (later) {
      f = ds / dsPrev;
      loadFactor = (1.0 + f) * lamPrev - f * lamPrev2;
    }
    hipLaunchKernelGGL(k_al_predict, dim3(g3), dim3(256), 0, s, n3, later ? 1 : 0, f, p->disp.ptr, W.prev.ptr,
                       W.prev2.ptr, W.Du.ptr);
    MPF_TRY(do_set_bc(p, loadFactor));  // prescribed values are zero: dispApplied = -disp = 0 there
    double Dl = loadFactor - lamPrev;
    for (int iter = 1; iter < o->max_iters; ++iter) {
      const mpf_status st = assemble_sync(p, iter == 1, loadFactor, 1);  // R includes lam Fext
      if (st == MPF_ERR_NEG_JACOBIAN) break;
      MPF_TRY(st);
      double rn = 0.0;
      MPF_TRY(rhs_norm(p, &rn));
      rep->newton_iters += 1;
      if (rn < o->tol) {
        conv = true;
        break;
      }
      // du' = K^-1 R
      mpf_krylov_report kr{};
      mpf_status ss = krylov_solve(p, &lin, &kr);
      rep->krylov_iters += kr.iterations;
      if (ss != MPF_OK && ss != MPF_ERR_NOT_CONVERGED) return ss;
      MPF_HIP(hipMemcpyAsync(W.dup.ptr, p->du.ptr, (size_t)nfree * sizeof(double), hipMemcpyDeviceToDevice, s));
      // du1 = K^-1 F (the same matrix: a re-solve with the band's factors)
      hipLaunchKernelGGL(k_al_fshort, dim3(gf), dim3(256), 0, s, nfree, assy, p->fext_unit.ptr, p->rhs.ptr);
      if (lin.ksp == MPF_KSP_BAND) {
        MPF_TRY(band_enqueue_resolve(p));
      } else {
        ss = krylov_solve(p, &lin, &kr);
        rep->krylov_iters += kr.iterations;
        if (ss != MPF_OK && ss != MPF_ERR_NOT_CONVERGED) return ss;
      }
      double h[3] = {0.0, 0.0, 0.0};
      if (later) {
        // Du.Du, Du.du', Du.du1 over the free dofs, Du gathered through assyForSoln into rhs
        hipLaunchKernelGGL(k_al_fshort, dim3(gf), dim3(256), 0, s, nfree, assy, W.Du.ptr, p->rhs.ptr);
        enqueue_dot(s, nfree, p->rhs.ptr, nullptr, p->rhs.ptr, W, 0);
        enqueue_dot(s, nfree, p->rhs.ptr, nullptr, W.dup.ptr, W, 1);
        enqueue_dot(s, nfree, p->rhs.ptr, nullptr, p->du.ptr, W, 2);
        MPF_HIP(hipMemcpyAsync(h, W.sc.ptr, 3 * sizeof(double), hipMemcpyDeviceToHost, s));
      }
      MPF_HIP(hipStreamSynchronize(s));
      double A = 0.0, b = 1.0, adu2 = 0.0, adu1 = 0.0;
      if (later) {
        A = h[0] + Dl * Dl * FtF - ds * ds;
        b = 2.0 * Dl * FtF;
        adu2 = -2.0 * h[1];  // du2 = -du'
        adu1 = 2.0 * h[2];
      }
      const double dl = (adu2 - A) / (b + adu1);
      hipLaunchKernelGGL(k_al_update, dim3(gf), dim3(256), 0, s, nfree, assy, W.dup.ptr, p->du.ptr, dl, p->disp.ptr,
                         W.Du.ptr);
      MPF_HIP(hipGetLastError());
      loadFactor += dl;
      Dl += dl;
    }
    if (conv) {
      if (steps == 1) {
        enqueue_dot(s, n3, W.Du.ptr, nullptr, W.Du.ptr, W, 0);
        double dd = 0.0;
        MPF_HIP(hipMemcpyAsync(&dd, W.sc.ptr, sizeof(double), hipMemcpyDeviceToHost, s));
        MPF_HIP(hipStreamSynchronize(s));
        ds = std::sqrt(dd + loadFactor * loadFactor * FtF);
        dsMax = ds;
        dsMin = ds / 1024.0;
      }
      if (lam_out) lam_out[steps - 1] = loadFactor;
      if (ds_out) ds_out[steps - 1] = steps > 1 ? ds : 0.0;
      lamPrev2 = lamPrev;
      lamPrev = loadFactor;
      hipLaunchKernelGGL(k_al_save, dim3(g3), dim3(256), 0, s, n3, p->disp.ptr, W.prev.ptr, W.prev2.ptr);
      dsPrev = ds;
      if (convPrev) ds = std::min(std::max(2.0 * ds, dsMin), dsMax);
      ++steps;
      rep->steps_done += 1;
      rep->final_lam = loadFactor;
    } else {
      ds = std::max(ds * (convPrev ? 0.5 : 0.25), dsMin);
    }
  }
  MPF_HIP(hipStreamSynchronize(s));
  rep->final_ds = ds;
  rep->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return MPF_OK;
}
