// problem.hip -- the femSolidmechanics-shaped driver on one device and the C ABI (include/mpf.h).
//
// Reference call sequence (chennachaos/mphyfem, src/femSolidmechanics/):
//   prepareMatrixPattern            femSolidmechanics2.cpp:121-478      -> mpf_prepare_pattern
//   solveWithNewtonRaphson          femSolidmechanics2.cpp:730-851      -> mpf_newton_solve / mpf_newton_step
//   timeUpdate / setBoundaryConditions femSolidmechanics1.cpp:1126-1164, 813-915 -> mpf_set_bc
//   calcStiffnessAndResidual        femSolidmechanics2.cpp:1047-1120    -> mpf_assemble
//   addExternalForces               femSolidmechanics1.cpp:922-1119     -> unit-load force vector, scaled
//   factoriseSolveAndUpdate         femSolidmechanics2.cpp:1124-1217    -> mpf_solve + mpf_update
//   converged / diverging           femSolidmechanics1.cpp:616-636

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>

#include "mpf_internal.hpp"

namespace mpf {
mpf_status spmv(const int32_t* rowptr, const int32_t* colidx, const double* vals, int64_t n, int64_t nnz,
                const double* x, double* y, double* part, unsigned np, const double* sc, hipStream_t s);
}

using namespace mpf;

namespace {

__global__ void k_fill_d(double* __restrict__ v, int64_t n, double val) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    v[i] = val;
}

__global__ void k_fill_int(int32_t* __restrict__ v, int64_t n, int32_t val) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    v[i] = val;
}

// dispApplied[dof] = value * lam - disp[dof] on the prescribed dofs (femSolidmechanics1.cpp:864-868, 890)
__global__ void k_set_bc(int32_t nbc, const int32_t* __restrict__ dof, const double* __restrict__ val, double lam,
                         const double* __restrict__ disp, double* __restrict__ applied) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nbc) applied[dof[i]] = val[i] * lam - disp[dof[i]];
}

// rhs[i] += lam * Fext[assyForSoln[i]]  (femSolidmechanics2.cpp:785-789)
// lam_dev (when given) holds the load factor in device memory, so that a captured graph reads the
// current one at replay
__global__ void k_add_fext(int32_t nfree, const int32_t* __restrict__ assy, const double* __restrict__ fext, double lam,
                           const double* __restrict__ lam_dev, double* __restrict__ rhs) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const double l = lam_dev ? *lam_dev : lam;
  if (i < nfree) rhs[i] += l * fext[assy[i]];
}

// factoriseSolveAndUpdate (femSolidmechanics2.cpp:1137-1196)
__global__ void k_update_free(int32_t nfree, const int32_t* __restrict__ assy, const double* __restrict__ du,
                              double* __restrict__ disp) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nfree) disp[assy[i]] += du[i];
}

__global__ void k_update_bc(int32_t nbc, const int32_t* __restrict__ dof, const double* __restrict__ applied,
                            double* __restrict__ disp) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nbc) disp[dof[i]] += applied[dof[i]];
}

__global__ void k_update_pres(int32_t npres, const int32_t* __restrict__ assyP, const double* __restrict__ du,
                              int32_t off, double* __restrict__ pres) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < npres) pres[assyP[i]] += du[off + i];
}

// prescribed dofs take value * lam (learned start point)
__global__ void k_set_prescribed(int32_t nbc, const int32_t* __restrict__ dof, const double* __restrict__ val,
                                 double lam, double* __restrict__ disp) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nbc) disp[dof[i]] = val[i] * lam;
}

// Material batch: F row-major -> sigma[9], C[81] in the reference's 9-index map.
__global__ void k_neohooke_batch(const double* __restrict__ F, const double* __restrict__ pr, mpf_matl m, int32_t n,
                                 double* __restrict__ sig, double* __restrict__ C) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  double f[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) f[i][j] = F[(size_t)t * 9 + 3 * i + j];
  const double J = f[0][0] * (f[1][1] * f[2][2] - f[1][2] * f[2][1]) - f[0][1] * (f[1][0] * f[2][2] - f[1][2] * f[2][0]) +
                   f[0][2] * (f[1][0] * f[2][1] - f[1][1] * f[2][0]);
  double b[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) b[i][j] = f[i][0] * f[j][0] + f[i][1] * f[j][1] + f[i][2] * f[j][2];
  const double K = 1.0 / m.Kinv;
  const double r1d3 = 1.0 / 3.0, r2d3 = 2.0 / 3.0;
  const double Jm2d3 = pow(J, -r2d3), Jm5d3 = Jm2d3 / J;
  const double Ib = b[0][0] + b[1][1] + b[2][2];
  double U1 = 0.0, U2 = 0.0;
  switch (m.Utype) {
    case 1: U1 = 0.1 * K * (J * J * J * J - 1.0 / (J * J * J * J * J * J)); U2 = 0.2 * K * (2.0 * J * J * J + 3.0 / (J * J * J * J * J * J * J)); break;
    case 2: U1 = 0.125 * K * (J * J - 1.0 / J / J) * (J + 1.0 / J / J / J); U2 = 0.125 * K * (3.0 * J * J + 5.0 / (J * J * J * J * J * J)); break;
    case 3: U1 = K * (J - 1.0); U2 = K; break;
    case 4: U1 = 0.5 * K * (J - 1.0 / J); U2 = 0.5 * K * (1.0 + 1.0 / J / J); break;
    case 5: U1 = 0.5 * K * (J - 1.0 + log(J) / J); U2 = 0.5 * K * (J * J + 1.0 - log(J)) / J / J; break;
    case 6: U1 = K * log(J); U2 = K / log(J); break;
    case 7: U1 = K * (1.0 - 1.0 / J); U2 = K / J / J; break;
    case 8: U1 = K * log(J) / J; U2 = K * (1.0 - log(J)) / J / J; break;
    default: break;
  }
  const double pres = m.mixed ? pr[t] : 0.0;
  double fact1 = m.mu * Jm5d3;
  const double fact2s = -fact1 * r1d3 * Ib + (m.mixed ? pres : U1);
  const int mp[9][2] = {{0, 0}, {1, 0}, {2, 0}, {0, 1}, {1, 1}, {2, 1}, {0, 2}, {1, 2}, {2, 2}};
  for (int i1 = 0; i1 < 9; ++i1) {
    const int i = mp[i1][0], j = mp[i1][1];
    sig[(size_t)t * 9 + i1] = fact1 * b[i][j] + (i == j ? fact2s : 0.0);
  }
  fact1 = r1d3 * m.mu * Jm5d3 * Ib;
  const double fact2 = -r2d3 * m.mu * Jm5d3;
  const double fact3 = r2d3 * r1d3 * m.mu * Jm5d3 * Ib;
  const double fact4 = m.mu * Jm5d3;
  const double fact7 = J * U2 + U1, fact8 = -U1, fact9 = m.tangent == MPF_TANGENT_CONSISTENT ? 0.0 : -U1;
  for (int i1 = 0; i1 < 9; ++i1) {
    const int i = mp[i1][0], j = mp[i1][1];
    const double dij = (i == j);
    for (int j1 = 0; j1 < 9; ++j1) {
      const int k = mp[j1][0], l = mp[j1][1];
      const double dik = (i == k), dil = (i == l), djk = (j == k), djl = (j == l), dkl = (k == l);
      double c = fact1 * (dil * djk) + fact2 * (b[i][j] * dkl + dij * b[k][l]) + fact3 * (dij * dkl) +
                 fact4 * (b[j][l] * dik);
      if (m.mixed) c += pres * (dij * dkl - djk * dil);
      else c += fact7 * (dij * dkl) + fact8 * (dil * djk) + fact9 * (djl * dik);
      C[(size_t)t * 81 + i1 * 9 + j1] = c;
    }
  }
}

inline unsigned g256(int64_t n) {
  int64_t g = (n + 255) / 256;
  if (g > 65535 * 16) g = 65535 * 16;
  return (unsigned)(g < 1 ? 1 : g);
}

// External force vector for lambda = 1 from the traction patches, on the host (one-time per
// problem; the per-step cost is the scaled add in k_add_fext).  femSolidmechanics1.cpp:1026-1082
// with LagrangeBasisFunsFace (BasisFunctionsLagrange.cpp:1688-1732): Quad4 faces with the 2x2 rule,
// Tria6 faces (LagrangeBasisFunsTria npElem 6, :364-418) with the 3-point rule (QuadratureUtil.cpp:
// 562-570); t = v0 n + v1 dx/dxi1 with n = dx/dxi2 x dx/dxi1 normalised.
void traction_unit(const std::vector<double>& X, const std::vector<mpf_problem::Trac>& tr, int nNode,
                   std::vector<double>& F) {
  F.assign((size_t)nNode * 3, 0.0);
  const double g = 0.577350269189626;
  const double gq[4][2] = {{-g, -g}, {g, -g}, {g, g}, {-g, g}};
  const double gt[3][2] = {{1.0 / 6.0, 1.0 / 6.0}, {4.0 / 6.0, 1.0 / 6.0}, {1.0 / 6.0, 4.0 / 6.0}};
  const double wt = 0.5 * (1.0 / 3.0);
  for (const auto& t : tr) {
    const int npf = t.npf;
    const int nf = (int)t.faces.size() / npf;
    const int ngp = npf == 4 ? 4 : 3;
    for (int f = 0; f < nf; ++f) {
      const int* nn = &t.faces[(size_t)npf * f];
      double Fl[18] = {0};
      for (int q = 0; q < ngp; ++q) {
        double N[6], d1[6], d2[6], w;
        if (npf == 4) {
          const double Nu[2] = {0.5 * (1 - gq[q][0]), 0.5 * (1 + gq[q][0])}, dNu[2] = {-0.5, 0.5};
          const double Nv[2] = {0.5 * (1 - gq[q][1]), 0.5 * (1 + gq[q][1])}, dNv[2] = {-0.5, 0.5};
          N[0] = Nv[0] * Nu[0]; N[1] = Nv[0] * Nu[1]; N[2] = Nv[1] * Nu[1]; N[3] = Nv[1] * Nu[0];
          d1[0] = Nv[0] * dNu[0]; d1[1] = Nv[0] * dNu[1]; d1[2] = Nv[1] * dNu[1]; d1[3] = Nv[1] * dNu[0];
          d2[0] = dNv[0] * Nu[0]; d2[1] = dNv[0] * Nu[1]; d2[2] = dNv[1] * Nu[1]; d2[3] = dNv[1] * Nu[0];
          w = 1.0;
        } else {
          const double x1 = gt[q][0], x2 = gt[q][1], x3 = 1.0 - x1 - x2;
          N[0] = x3 * (2.0 * x3 - 1.0); N[1] = x1 * (2.0 * x1 - 1.0); N[2] = x2 * (2.0 * x2 - 1.0);
          N[3] = 4.0 * x1 * x3; N[4] = 4.0 * x1 * x2; N[5] = 4.0 * x2 * x3;
          d1[0] = -4.0 * x3 + 1.0; d1[1] = 4.0 * x1 - 1.0; d1[2] = 0.0;
          d1[3] = 4.0 * (x3 - x1); d1[4] = 4.0 * x2; d1[5] = -4.0 * x2;
          d2[0] = -4.0 * x3 + 1.0; d2[1] = 0.0; d2[2] = 4.0 * x2 - 1.0;
          d2[3] = -4.0 * x1; d2[4] = 4.0 * x1; d2[5] = 4.0 * (x3 - x2);
          w = wt;
        }
        double du[3] = {0, 0, 0}, dv[3] = {0, 0, 0};
        for (int a = 0; a < npf; ++a)
          for (int d = 0; d < 3; ++d) { du[d] += X[3 * nn[a] + d] * d1[a]; dv[d] += X[3 * nn[a] + d] * d2[a]; }
        double n1[3] = {dv[1] * du[2] - dv[2] * du[1], dv[2] * du[0] - dv[0] * du[2], dv[0] * du[1] - dv[1] * du[0]};
        const double Jac = std::sqrt(n1[0] * n1[0] + n1[1] * n1[1] + n1[2] * n1[2]);
        for (int d = 0; d < 3; ++d) n1[d] /= Jac;
        const double dvol = w * Jac;
        for (int d = 0; d < 3; ++d) {
          const double tr_d = (t.v0 * n1[d] + t.v1 * du[d]) * dvol;
          for (int a = 0; a < npf; ++a) Fl[3 * a + d] += N[a] * tr_d;
        }
      }
      for (int a = 0; a < npf; ++a)
        for (int d = 0; d < 3; ++d) F[3 * nn[a] + d] += Fl[3 * a + d];
    }
  }
}

mpf_status check_device() {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return MPF_ERR_NO_DEVICE;
  return MPF_OK;
}

}  // namespace

uint64_t mpf::next_problem_uid() {
  static std::atomic<uint64_t> next_uid{1};
  return next_uid++;
}

// Per-problem host logic shared with the batched solver (batch.hip).
ElemArgs elem_args(mpf_problem* p) {
  ElemArgs a{};
  a.X = p->X.ptr;
  a.u = p->disp.ptr;
  a.uprev = p->dispPrev.ptr;
  a.pres = p->pres.ptr;
  a.conn = p->conn.ptr;
  a.mu = p->m.mu;
  a.Kinv = p->m.Kinv;
  a.Utype = p->m.Utype;
  a.consistent = p->m.tangent == MPF_TANGENT_CONSISTENT;
  a.bad = p->bad.ptr;
  return batch_elem_args(p, a);
}



// Enqueue the element kernels of one assembly (no zero fill, no external forces).
mpf_status enqueue_element_kernels(mpf_problem* p, bool first_iter, int scatterK) {
  hipStream_t s = p->stream;
  const PatternBufs& P = p->pat;
  ElemArgs A = elem_args(p);
  if (p->asm_mode == MPF_ASSEMBLY_GATHER || p->asm_mode == MPF_ASSEMBLY_GATHER_DIRECT) {
    return launch_gather_assembly(p, A, first_iter ? p->dispApplied.ptr : nullptr, scatterK, s);
  }
  ScatterArgs sc{};
  sc.slot = P.slot.ptr;
  sc.nodeinfo = P.nodeinfo.ptr;
  sc.nda = P.nda.ptr;
  sc.idp = P.idp.ptr;
  sc.dispDOF = P.dispDOF;
  sc.vals = p->vals.ptr;
  sc.rhs = p->rhs.ptr;
  sc.dispApplied = first_iter ? p->dispApplied.ptr : nullptr;
  sc.scatterK = scatterK;
  for (int c = 0; c < P.ncolors; ++c) {
    sc.elems = P.perm.ptr + P.color_off[c];
    sc.count = P.color_off[c + 1] - P.color_off[c];
    if (p->p0) MPF_TRY(launch_hex8p0_scatter(A, sc, p->q_Kup.ptr, p->q_Kpp.ptr, p->q_Fp.ptr, s));
    else if (p->etype == MPF_HEX8) MPF_TRY(launch_hex8_scatter(A, sc, s));
    else if (p->etype == MPF_TET10) MPF_TRY(launch_tet10d_scatter(A, sc, s));
    else MPF_TRY(launch_tet10_scatter(A, sc, s));
  }
  MPF_HIP(hipGetLastError());
  return MPF_OK;
}

// Enqueue one full assembly (no synchronisation).  scatterK = 0 assembles the residual only.
// zeroMtx (SolverPetsc.cpp:145-163) is only needed by the scatter, which accumulates.
mpf_status enqueue_assembly(mpf_problem* p, bool first_iter, double lam, int scatterK, const double* lam_dev = nullptr) {
  hipStream_t s = p->stream;
  const PatternBufs& P = p->pat;
  const int64_t n = (int64_t)P.dispDOF + P.presDOF;
  if (p->asm_mode == MPF_ASSEMBLY_COLORED_SCATTER) {
    if (scatterK) hipLaunchKernelGGL(k_fill_d, dim3(g256(P.nnz)), dim3(256), 0, s, p->vals.ptr, P.nnz, 0.0);
    hipLaunchKernelGGL(k_fill_d, dim3(g256(n)), dim3(256), 0, s, p->rhs.ptr, n, 0.0);
  }
  MPF_TRY(enqueue_element_kernels(p, first_iter, scatterK));
  if (p->has_fext && P.dispDOF > 0)
    hipLaunchKernelGGL(k_add_fext, dim3(grid_for(P.dispDOF, 256)), dim3(256), 0, s, P.dispDOF, P.assy.ptr,
                       p->fext_unit.ptr, lam, lam_dev, p->rhs.ptr);
  MPF_HIP(hipGetLastError());
  return MPF_OK;
}

// bad[0]: 1 + element with det F < 0 (smallest id).
mpf_status assemble_sync(mpf_problem* p, bool first_iter, double lam, int scatterK) {
  MPF_HIP(hipMemsetAsync(p->bad.ptr, 0, sizeof(int32_t), p->stream));
  MPF_TRY(enqueue_assembly(p, first_iter, lam, scatterK));
  int32_t bad = 0;
  MPF_HIP(hipMemcpyAsync(&bad, p->bad.ptr, sizeof(int32_t), hipMemcpyDeviceToHost, p->stream));
  MPF_HIP(hipStreamSynchronize(p->stream));
  p->assembled = true;
  if (bad) return MPF_ERR_NEG_JACOBIAN;
  return MPF_OK;
}

// emat / mflag / stride (batched meshes): Q1/P0 elements of meshes without a solve this iteration
// keep their pressures (their du is 0, but Fp / Kpp would still move them)
mpf_status do_update(mpf_problem* p, bool first_iter, const int32_t* emat = nullptr, const int32_t* mflag = nullptr,
                     int32_t stride = 0) {
  if (p->p0 && !p->assembled) return MPF_ERR_STATE;  // the pressure recovery reads the last assembly
  hipStream_t s = p->stream;
  const PatternBufs& P = p->pat;
  if (first_iter && !p->bc_dof_h.empty())
    hipLaunchKernelGGL(k_update_bc, dim3(grid_for(p->bc_dof_h.size(), 256)), dim3(256), 0, s,
                       (int32_t)p->bc_dof_h.size(), p->bc_dof.ptr, p->dispApplied.ptr, p->disp.ptr);
  if (P.dispDOF > 0)
    hipLaunchKernelGGL(k_update_free, dim3(grid_for(P.dispDOF, 256)), dim3(256), 0, s, P.dispDOF, P.assy.ptr, p->du.ptr,
                       p->disp.ptr);
  if (P.presDOF > 0)
    hipLaunchKernelGGL(k_update_pres, dim3(grid_for(P.presDOF, 256)), dim3(256), 0, s, P.presDOF, P.assyP.ptr,
                       p->du.ptr, P.dispDOF, p->pres.ptr);
  if (p->p0) {  // from the element matrices of the assembly this correction solved
    MPF_TRY(launch_hex8p0_pressure(p->nElem, p->conn.ptr, P.nda.ptr, p->du.ptr, first_iter ? p->dispApplied.ptr : nullptr,
                                   p->q_Kup.ptr, p->q_Kpp.ptr, p->q_Fp.ptr, p->pres.ptr, s, emat, mflag, stride));
  }
  MPF_HIP(hipGetLastError());
  return MPF_OK;
}

mpf_status do_set_bc(mpf_problem* p, double lam) {
  if (!p->bc_dof_h.empty())
    hipLaunchKernelGGL(k_set_bc, dim3(grid_for(p->bc_dof_h.size(), 256)), dim3(256), 0, p->stream,
                       (int32_t)p->bc_dof_h.size(), p->bc_dof.ptr, p->bc_val.ptr, lam, p->disp.ptr,
                       p->dispApplied.ptr);
  MPF_HIP(hipGetLastError());
  return MPF_OK;
}

mpf_status rhs_norm(mpf_problem* p, double* out) {
  if (p->scratch.n < (size_t)kMaxPartials + 1) MPF_TRY(p->scratch.alloc((size_t)kMaxPartials + 1));
  const int64_t n = (int64_t)p->pat.dispDOF + p->pat.presDOF;
  return dev_norm2(p->rhs.ptr, n, out, p->stream, p->scratch.ptr);
}

double seconds_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// One Newton load step at load factor lam from the current (converged) state.  Returns the
// number of iterations (assemblies) in *iters.  guess: host displacement to start from.
mpf_status newton_step(mpf_problem* p, double lam, const mpf_newton_opts* o, const double* guess, int32_t* iters,
                       mpf_newton_report* rep, double* relres_out) {
  hipStream_t s = p->stream;
  const size_t n3 = (size_t)p->nNode * 3;
  // saveSolution: dispPrev = disp, presPrev = pres (SolutionDataSolid.cpp:170-184)
  MPF_HIP(hipMemcpyAsync(p->dispPrev.ptr, p->disp.ptr, n3 * sizeof(double), hipMemcpyDeviceToDevice, s));
  MPF_HIP(hipMemcpyAsync(p->presPrev.ptr, p->pres.ptr, p->npres * sizeof(double), hipMemcpyDeviceToDevice, s));
  MPF_TRY(do_set_bc(p, lam));
  double R0 = -1.0;
  int iter = 1;
  bool conv = false;
  double lastrel = 1.0, rnPrev = -1.0;
  auto t_asm = std::chrono::steady_clock::now();
  if (guess) {
    // R0 from the plain start point (residual only, with the lift), then move to the guess.
    MPF_TRY(assemble_sync(p, true, lam, 0));
    MPF_TRY(rhs_norm(p, &R0));
    rep->seconds_assembly += seconds_since(t_asm);
    MPF_HIP(hipMemcpyAsync(p->disp.ptr, guess, n3 * sizeof(double), hipMemcpyHostToDevice, s));
    if (!p->bc_dof_h.empty())
      hipLaunchKernelGGL(k_set_prescribed, dim3(grid_for(p->bc_dof_h.size(), 256)), dim3(256), 0, s,
                         (int32_t)p->bc_dof_h.size(), p->bc_dof.ptr, p->bc_val.ptr, lam, p->disp.ptr);
  }
  for (; iter <= o->itmax; ++iter) {
    const bool first = (iter == 1) && !guess;
    t_asm = std::chrono::steady_clock::now();
    mpf_status st = assemble_sync(p, first, lam, 1);
    double rn = 0.0;
    if (st == MPF_OK) st = rhs_norm(p, &rn);
    rep->seconds_assembly += seconds_since(t_asm);
    rep->newton_iters += 1;
    if (st != MPF_OK) {
      if (iters) *iters = iter;
      return st;
    }
    if (std::isnan(rn) || std::isnan(R0)) {  // a NaN residual never passes the convergence test
      if (iters) *iters = iter;
      return MPF_ERR_DIVERGED;
    }
    if (R0 < 0) R0 = rn;
    lastrel = R0 > 0 ? rn / R0 : 0.0;
    if (R0 == 0.0 || lastrel < o->rtol) { conv = true; break; }
    if (iter > 3 && rnPrev > 0 && rn / rnPrev > 1.0e7) {
      if (iters) *iters = iter;
      return MPF_ERR_DIVERGED;
    }
    rnPrev = rn;
    mpf_krylov_opts lo = o->lin;
    if (o->forcing) lo.rtol = std::max(o->lin.rtol, std::min(0.1, 0.1 * lastrel));
    mpf_krylov_report kr{};
    auto t_sol = std::chrono::steady_clock::now();
    st = krylov_solve(p, &lo, &kr);
    rep->seconds_solve += seconds_since(t_sol);
    rep->krylov_iters += kr.iterations;
    if (st != MPF_OK) return st;
    MPF_TRY(do_update(p, first));
  }
  if (iters) *iters = conv ? iter : o->itmax;
  if (relres_out) *relres_out = lastrel;
  MPF_HIP(hipStreamSynchronize(s));
  return conv ? MPF_OK : MPF_ERR_NOT_CONVERGED;
}


// =============================================================================================
// C ABI
// =============================================================================================
extern "C" {

const char* mpf_status_string(int32_t st) {
  switch (st) {
    case MPF_OK: return "ok";
    case MPF_ERR_ARG: return "invalid argument";
    case MPF_ERR_HIP: return "HIP runtime error";
    case MPF_ERR_NEG_JACOBIAN: return "Negative Jacobian in the element";
    case MPF_ERR_STATE: return "call out of order";
    case MPF_ERR_NOT_CONVERGED: return "not converged";
    case MPF_ERR_DIVERGED: return "diverged";
    case MPF_ERR_NO_DEVICE: return "no HIP device";
    default: return "unknown status";
  }
}

int32_t mpf_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

mpf_status mpf_neohooke_stress_tangent_batch(const double* F, const double* p, const mpf_matl* m, int32_t n,
                                             double* sig, double* C, mpf_stream_t stream) {
  if (!F || !m || !sig || !C || n < 0 || (m->mixed && !p)) return MPF_ERR_ARG;
  if (m->Kinv <= 0.0) return MPF_ERR_ARG;
  if (n == 0) return MPF_OK;
  hipLaunchKernelGGL(k_neohooke_batch, dim3(grid_for(n, 128)), dim3(128), 0, (hipStream_t)stream, F, p, *m, n, sig, C);
  MPF_HIP(hipGetLastError());
  return MPF_OK;
}

mpf_status mpf_elem_hex8_nh_batch(const double* X, const double* u, const int32_t* conn, const mpf_matl* m,
                                  int32_t first, int32_t count, double* Ke, double* Re, int32_t* bad,
                                  mpf_stream_t stream) {
  if (!X || !u || !conn || !m || !Ke || !Re || first < 0 || count < 0 || m->Kinv <= 0.0) return MPF_ERR_ARG;
  ElemArgs a{};
  a.X = X; a.u = u; a.conn = conn; a.mu = m->mu; a.Kinv = m->Kinv; a.Utype = m->Utype; a.bad = bad;
  a.consistent = m->tangent == MPF_TANGENT_CONSISTENT;
  return launch_hex8_dense(a, first, count, Ke, Re, (hipStream_t)stream);
}

mpf_status mpf_elem_tet10_nh_batch(const double* X, const double* u, const int32_t* conn, const mpf_matl* m,
                                   int32_t first, int32_t count, double* Ke, double* Re, int32_t* bad,
                                   mpf_stream_t stream) {
  if (!X || !u || !conn || !m || !Ke || !Re || first < 0 || count < 0 || m->Kinv <= 0.0) return MPF_ERR_ARG;
  ElemArgs a{};
  a.X = X; a.u = u; a.conn = conn; a.mu = m->mu; a.Kinv = m->Kinv; a.Utype = m->Utype; a.bad = bad;
  a.consistent = m->tangent == MPF_TANGENT_CONSISTENT;
  return launch_tet10d_dense(a, first, count, Ke, Re, (hipStream_t)stream);
}

mpf_status mpf_elem_hex8p0_nh_batch(const double* X, const double* u, const double* u_prev, const double* pres_elem,
                                    const int32_t* conn, const mpf_matl* m, int32_t first, int32_t count, double* Kc,
                                    double* Fc, double* Kup, double* Kpp, double* Fp, int32_t* bad,
                                    mpf_stream_t stream) {
  if (!X || !u || !u_prev || !pres_elem || !conn || !m || !Kc || !Fc || !Kup || !Kpp || !Fp || first < 0 ||
      count < 0 || m->Kinv <= 0.0)
    return MPF_ERR_ARG;
  ElemArgs a{};
  a.X = X; a.u = u; a.uprev = u_prev; a.pres = pres_elem; a.conn = conn;
  a.mu = m->mu; a.Kinv = m->Kinv; a.Utype = m->Utype; a.bad = bad;
  return launch_hex8p0_dense(a, first, count, Kc, Fc, Kup, Kpp, Fp, (hipStream_t)stream);
}

mpf_status mpf_elem_tet10p1_nh_batch(const double* X, const double* u, const double* u_prev, const double* pres,
                                     const int32_t* conn, const mpf_matl* m, int32_t first, int32_t count, double* Kuu,
                                     double* Kup, double* Kpu, double* Kpp, double* Fu, double* Fp, int32_t* bad,
                                     mpf_stream_t stream) {
  if (!X || !u || !u_prev || !pres || !conn || !m || !Kuu || !Kup || !Kpu || !Kpp || !Fu || !Fp || first < 0 ||
      count < 0 || m->Kinv <= 0.0)
    return MPF_ERR_ARG;
  ElemArgs a{};
  a.X = X; a.u = u; a.uprev = u_prev; a.pres = pres; a.conn = conn;
  a.mu = m->mu; a.Kinv = m->Kinv; a.Utype = m->Utype; a.bad = bad;
  return launch_tet10_dense(a, first, count, Kuu, Kup, Kpu, Kpp, Fu, Fp, (hipStream_t)stream);
}

mpf_status mpf_problem_create(mpf_problem** out, int32_t elem_type, int32_t nNode, int32_t nElem, const double* X,
                              const int32_t* conn, const mpf_matl* m) {
  if (!out || !X || !conn || !m || nNode <= 0 || nElem <= 0) return MPF_ERR_ARG;
  if (elem_type != MPF_HEX8 && elem_type != MPF_HEX8P0 && elem_type != MPF_TET10P1 && elem_type != MPF_TET10)
    return MPF_ERR_ARG;
  if (m->Kinv <= 0.0) return MPF_ERR_ARG;
  const bool p0 = elem_type == MPF_HEX8P0;
  const int npe = (elem_type == MPF_TET10P1 || elem_type == MPF_TET10) ? 10 : 8;
  for (int64_t i = 0; i < (int64_t)nElem * npe; ++i)
    if (conn[i] < 0 || conn[i] >= nNode) return MPF_ERR_ARG;
  MPF_TRY(check_device());
  auto* p = new mpf_problem();
  p->uid = mpf::next_problem_uid();
  p->etype = p0 ? MPF_HEX8 : elem_type;  // Q1/P0 shares the Hex8 numbering and pattern (no pressure dofs)
  p->p0 = p0;
  p->npe = npe;
  p->nNode = nNode;
  p->nElem = nElem;
  p->npres = p0 ? nElem : nNode;
  p->m = *m;
  p->m.mixed = (elem_type == MPF_HEX8 || elem_type == MPF_TET10) ? 0 : 1;
  if (hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess) {
    delete p;
    return MPF_ERR_HIP;
  }
  const size_t n3 = (size_t)nNode * 3;
  mpf_status st = MPF_OK;
  if (st == MPF_OK) st = p->X.upload(X, n3, p->stream);
  if (st == MPF_OK) st = p->conn.upload(conn, (size_t)nElem * npe, p->stream);
  if (st == MPF_OK) st = p->disp.alloc(n3);
  if (st == MPF_OK) st = p->dispPrev.alloc(n3);
  if (st == MPF_OK) st = p->dispApplied.alloc(n3);
  if (st == MPF_OK) st = p->pres.alloc((size_t)p->npres);
  if (st == MPF_OK) st = p->presPrev.alloc((size_t)p->npres);
  if (st == MPF_OK) st = p->bad.alloc(2);
  if (st != MPF_OK) {
    mpf_problem_destroy(p);
    return st;
  }
  hipLaunchKernelGGL(k_fill_d, dim3(g256(n3)), dim3(256), 0, p->stream, p->disp.ptr, (int64_t)n3, 0.0);
  hipLaunchKernelGGL(k_fill_d, dim3(g256(n3)), dim3(256), 0, p->stream, p->dispPrev.ptr, (int64_t)n3, 0.0);
  hipLaunchKernelGGL(k_fill_d, dim3(g256(n3)), dim3(256), 0, p->stream, p->dispApplied.ptr, (int64_t)n3, 0.0);
  hipLaunchKernelGGL(k_fill_d, dim3(g256(p->npres)), dim3(256), 0, p->stream, p->pres.ptr, p->npres, 0.0);
  hipLaunchKernelGGL(k_fill_d, dim3(g256(p->npres)), dim3(256), 0, p->stream, p->presPrev.ptr, p->npres, 0.0);
  p->X_h.assign(X, X + n3);
  for (int c = 0; c < 3; ++c) {
    p->xmin[c] = nNode > 0 ? X[c] : 0.0;
    p->xmax[c] = p->xmin[c];
  }
  for (int64_t i = 0; i < (int64_t)nNode; ++i)
    for (int c = 0; c < 3; ++c) {
      p->xmin[c] = std::min(p->xmin[c], X[3 * i + c]);
      p->xmax[c] = std::max(p->xmax[c], X[3 * i + c]);
    }
  if (hipStreamSynchronize(p->stream) != hipSuccess) {
    mpf_problem_destroy(p);
    return MPF_ERR_HIP;
  }
  *out = p;
  return MPF_OK;
}

mpf_status mpf_problem_destroy(mpf_problem* p) {
  if (!p) return MPF_OK;
  if (p->stream) (void)hipStreamSynchronize(p->stream);
  krylov_free(p);
  pmg_free(p);
  batch_free(p);
  band_free(p);
  system_free(p);
  hipStream_t s = p->stream;
  delete p;
  if (s) (void)hipStreamDestroy(s);
  return MPF_OK;
}

mpf_status mpf_add_dirichlet(mpf_problem* p, int32_t n, const int32_t* nodes, int32_t dof, const double* value) {
  if (!p || p->etype == 0 || n < 0 || (n > 0 && !nodes) || dof > 2) return MPF_ERR_ARG;
  if (p->prepared) return MPF_ERR_STATE;
  // merge last-wins per dof (setBoundaryConditions loops the patches in order)
  std::map<int32_t, double> cur;
  for (size_t i = 0; i < p->bc_dof_h.size(); ++i) cur[p->bc_dof_h[i]] = p->bc_val_h[i];
  for (int i = 0; i < n; ++i) {
    if (nodes[i] < 0 || nodes[i] >= p->nNode) return MPF_ERR_ARG;
    if (dof >= 0) cur[3 * nodes[i] + dof] = value ? value[i] : 0.0;
    else for (int d = 0; d < 3; ++d) cur[3 * nodes[i] + d] = 0.0;
  }
  p->bc_dof_h.clear();
  p->bc_val_h.clear();
  for (auto& kv : cur) { p->bc_dof_h.push_back(kv.first); p->bc_val_h.push_back(kv.second); }
  return MPF_OK;
}

mpf_status mpf_add_traction(mpf_problem* p, int32_t nf, const int32_t* faces, double v0, double v1) {
  if (!p || p->etype == 0 || nf < 0 || (nf > 0 && !faces)) return MPF_ERR_ARG;
  if (p->prepared) return MPF_ERR_STATE;
  for (int64_t i = 0; i < (int64_t)nf * 4; ++i)
    if (faces[i] < 0 || faces[i] >= p->nNode) return MPF_ERR_ARG;
  mpf_problem::Trac t;
  t.faces.assign(faces, faces + (size_t)nf * 4);
  t.v0 = v0;
  t.v1 = v1;
  p->trac.push_back(t);
  return MPF_OK;
}

mpf_status mpf_add_traction_tri6(mpf_problem* p, int32_t nf, const int32_t* faces, double v0, double v1) {
  if (!p || p->etype == 0 || nf < 0 || (nf > 0 && !faces)) return MPF_ERR_ARG;
  if (p->prepared) return MPF_ERR_STATE;
  for (int64_t i = 0; i < (int64_t)nf * 6; ++i)
    if (faces[i] < 0 || faces[i] >= p->nNode) return MPF_ERR_ARG;
  mpf_problem::Trac t;
  t.faces.assign(faces, faces + (size_t)nf * 6);
  t.v0 = v0;
  t.v1 = v1;
  t.npf = 6;
  p->trac.push_back(t);
  return MPF_OK;
}

mpf_status mpf_prepare_pattern(mpf_problem* p) {
  if (!p) return MPF_ERR_ARG;
  if (p->prepared) return MPF_OK;
  hipStream_t s = p->stream;
  MPF_TRY(p->bc_dof.upload(p->bc_dof_h.data(), p->bc_dof_h.size(), s));
  MPF_TRY(p->bc_val.upload(p->bc_val_h.data(), p->bc_val_h.size(), s));
  MPF_TRY(build_pattern(p));
  const int64_t n = (int64_t)p->pat.dispDOF + p->pat.presDOF;
  MPF_TRY(p->vals.alloc(p->pat.nnz > 0 ? p->pat.nnz : 1));
  MPF_TRY(p->rhs.alloc(n > 0 ? n : 1));
  MPF_TRY(p->du.alloc(n > 0 ? n : 1));
  if (p->asm_mode != MPF_ASSEMBLY_COLORED_SCATTER) MPF_TRY(ensure_gather_tables(p));
  if (p->p0) {
    const size_t ne = (size_t)p->nElem;
    MPF_TRY(p->q_Kup.alloc(ne * 24));
    MPF_TRY(p->q_Kpp.alloc(ne));
    MPF_TRY(p->q_Fp.alloc(ne));
  }
  hipLaunchKernelGGL(k_fill_d, dim3(g256(n)), dim3(256), 0, s, p->du.ptr, n, 0.0);
  if (!p->trac.empty()) {
    std::vector<double> F;
    traction_unit(p->X_h, p->trac, p->nNode, F);
    MPF_TRY(p->fext_unit.upload(F.data(), F.size(), s));
    p->has_fext = true;
  }
  MPF_HIP(hipStreamSynchronize(s));
  p->prepared = true;
  return MPF_OK;
}

mpf_status mpf_get_sizes(const mpf_problem* p, int32_t* dispDOF, int32_t* presDOF, int64_t* nnz, int32_t* ncolors) {
  if (!p) return MPF_ERR_ARG;
  if (!p->prepared) return MPF_ERR_STATE;
  if (dispDOF) *dispDOF = p->pat.dispDOF;
  if (presDOF) *presDOF = p->pat.presDOF;
  if (nnz) *nnz = p->pat.nnz;
  if (ncolors) *ncolors = p->pat.ncolors;
  return MPF_OK;
}

mpf_status mpf_get_pattern(const mpf_problem* p, int32_t* rowptr, int32_t* colidx, int32_t* NodeDofArray,
                           int32_t* ID_pres, int32_t* color) {
  if (!p) return MPF_ERR_ARG;
  if (!p->prepared) return MPF_ERR_STATE;
  hipStream_t s = p->stream;
  const int64_t n = (int64_t)p->pat.dispDOF + p->pat.presDOF;
  if (rowptr) MPF_TRY(p->pat.rowptr.download(rowptr, n + 1, s));
  if (colidx) MPF_TRY(p->pat.colidx.download(colidx, p->pat.nnz, s));
  if (NodeDofArray) MPF_TRY(p->pat.nda.download(NodeDofArray, (size_t)p->nNode * 3, s));
  if (ID_pres) {
    if (p->etype == MPF_TET10P1) MPF_TRY(p->pat.idp.download(ID_pres, p->nNode, s));
    else std::fill(ID_pres, ID_pres + p->nNode, -1);
  }
  if (color) MPF_TRY(p->pat.color.download(color, p->nElem, s));
  MPF_HIP(hipStreamSynchronize(s));
  return MPF_OK;
}

mpf_status mpf_set_state(mpf_problem* p, const double* disp, const double* disp_prev, const double* pres,
                         const double* pres_prev) {
  if (!p || p->etype == 0) return MPF_ERR_ARG;
  hipStream_t s = p->stream;
  const size_t n3 = (size_t)p->nNode * 3;
  if (disp) MPF_HIP(hipMemcpyAsync(p->disp.ptr, disp, n3 * sizeof(double), hipMemcpyHostToDevice, s));
  if (disp_prev) MPF_HIP(hipMemcpyAsync(p->dispPrev.ptr, disp_prev, n3 * sizeof(double), hipMemcpyHostToDevice, s));
  if (pres) MPF_HIP(hipMemcpyAsync(p->pres.ptr, pres, p->npres * sizeof(double), hipMemcpyHostToDevice, s));
  if (pres_prev) MPF_HIP(hipMemcpyAsync(p->presPrev.ptr, pres_prev, p->npres * sizeof(double), hipMemcpyHostToDevice, s));
  MPF_HIP(hipStreamSynchronize(s));
  return MPF_OK;
}

mpf_status mpf_state_device_ptrs(mpf_problem* p, double** disp, double** pres) {
  if (!p || p->etype == 0) return MPF_ERR_ARG;
  if (disp) *disp = p->disp.ptr;
  if (pres) *pres = p->pres.ptr;
  return MPF_OK;
}

mpf_status mpf_copy_state_device(mpf_problem* p, double* disp_dst) {
  if (!p || p->etype == 0 || !disp_dst) return MPF_ERR_ARG;
  MPF_HIP(hipMemcpyAsync(disp_dst, p->disp.ptr, (size_t)p->nNode * 3 * sizeof(double), hipMemcpyDeviceToDevice,
                         p->stream));
  MPF_HIP(hipStreamSynchronize(p->stream));
  return MPF_OK;
}

mpf_status mpf_get_state(const mpf_problem* p, double* disp, double* pres) {
  if (!p || p->etype == 0) return MPF_ERR_ARG;
  hipStream_t s = p->stream;
  if (disp) MPF_TRY(p->disp.download(disp, (size_t)p->nNode * 3, s));
  if (pres) MPF_TRY(p->pres.download(pres, (size_t)p->npres, s));
  MPF_HIP(hipStreamSynchronize(s));
  return MPF_OK;
}

mpf_status mpf_set_assembly_mode(mpf_problem* p, int32_t mode) {
  if (!p || p->etype == 0 || (mode != MPF_ASSEMBLY_GATHER && mode != MPF_ASSEMBLY_COLORED_SCATTER &&
             mode != MPF_ASSEMBLY_GATHER_DIRECT))
    return MPF_ERR_ARG;
  p->asm_mode = mode;
  if (p->prepared && mode != MPF_ASSEMBLY_COLORED_SCATTER) MPF_TRY(ensure_gather_tables(p));
  return MPF_OK;
}

mpf_status mpf_set_chunk_stage_cap(mpf_problem* p, int32_t cap) {
  if (!p || p->etype == 0 || cap < 0 || cap > 32) return MPF_ERR_ARG;
  p->chunk_cap = cap;
  return MPF_OK;
}

mpf_status mpf_get_assembly_mode(const mpf_problem* p, int32_t* mode) {
  if (!p || !mode) return MPF_ERR_ARG;
  *mode = p->asm_mode;
  return MPF_OK;
}

mpf_status mpf_set_bc(mpf_problem* p, double lam) {
  if (!p || p->etype == 0) return MPF_ERR_ARG;
  if (!p->prepared) return MPF_ERR_STATE;
  MPF_TRY(do_set_bc(p, lam));
  MPF_HIP(hipStreamSynchronize(p->stream));
  return MPF_OK;
}

mpf_status mpf_assemble(mpf_problem* p, int32_t first_iter, double lam) {
  if (!p || p->etype == 0) return MPF_ERR_ARG;
  if (!p->prepared) return MPF_ERR_STATE;
  return assemble_sync(p, first_iter != 0, lam, 1);
}

mpf_status mpf_system_device_ptrs(mpf_problem* p, double** vals, double** rhs, const int32_t** rowptr,
                                  const int32_t** colidx) {
  if (!p) return MPF_ERR_ARG;
  if (!p->prepared) return MPF_ERR_STATE;
  if (vals) *vals = p->vals.ptr;
  if (rhs) *rhs = p->rhs.ptr;
  if (rowptr) *rowptr = p->pat.rowptr.ptr;
  if (colidx) *colidx = p->pat.colidx.ptr;
  return MPF_OK;
}

mpf_status mpf_get_system(const mpf_problem* p, double* vals, double* rhs) {
  if (!p) return MPF_ERR_ARG;
  if (!p->assembled) return MPF_ERR_STATE;
  hipStream_t s = p->stream;
  if (vals) MPF_TRY(p->vals.download(vals, p->pat.nnz, s));
  if (rhs) MPF_TRY(p->rhs.download(rhs, (size_t)p->pat.dispDOF + p->pat.presDOF, s));
  MPF_HIP(hipStreamSynchronize(s));
  return MPF_OK;
}

mpf_status mpf_rhs_norm(mpf_problem* p, double* norm) {
  if (!p || !norm) return MPF_ERR_ARG;
  if (!p->assembled) return MPF_ERR_STATE;
  return rhs_norm(p, norm);
}

mpf_status mpf_solve(mpf_problem* p, const mpf_krylov_opts* o, mpf_krylov_report* rep) {
  if (!p || !o || !rep) return MPF_ERR_ARG;
  if (!p->assembled) return MPF_ERR_STATE;  // SolverPetsc::factoriseAndSolve returns -1 (SolverPetsc.cpp:321-323)
  if (o->ksp != MPF_KSP_CG && o->ksp != MPF_KSP_GMRES && o->ksp != MPF_KSP_BAND) return MPF_ERR_ARG;
  if (o->ksp != MPF_KSP_BAND && (o->pc < MPF_PC_NONE || o->pc > MPF_PC_PMG)) return MPF_ERR_ARG;
  if (p->etype == 0 && o->ksp == MPF_KSP_BAND) return MPF_ERR_ARG;  // the band ordering needs the mesh
  if (p->etype == 0 && p->nNode == 0 && o->pc == MPF_PC_BJACOBI3 && p->pat.dispDOF > 0)
    return MPF_ERR_ARG;  // node blocks need the caller's NodeDofArray
  std::memset(rep, 0, sizeof(*rep));
  MPF_TRY(krylov_solve(p, o, rep));
  p->solved = true;
  return rep->converged ? MPF_OK : MPF_ERR_NOT_CONVERGED;
}

mpf_status mpf_get_solution(const mpf_problem* p, double* du) {
  if (!p || !du) return MPF_ERR_ARG;
  if (!p->solved) return MPF_ERR_STATE;
  MPF_TRY(p->du.download(du, (size_t)p->pat.dispDOF + p->pat.presDOF, p->stream));
  MPF_HIP(hipStreamSynchronize(p->stream));
  return MPF_OK;
}

mpf_status mpf_update(mpf_problem* p, int32_t first_iter) {
  if (!p || p->etype == 0) return MPF_ERR_ARG;
  if (!p->solved) return MPF_ERR_STATE;
  MPF_TRY(do_update(p, first_iter != 0));
  MPF_HIP(hipStreamSynchronize(p->stream));
  return MPF_OK;
}

mpf_status mpf_newton_step(mpf_problem* p, double lam, const mpf_newton_opts* o, const double* guess, int32_t* iters,
                           mpf_newton_report* rep) {
  if (!p || p->etype == 0 || !o || !rep) return MPF_ERR_ARG;
  if (!p->prepared) return MPF_ERR_STATE;
  std::memset(rep, 0, sizeof(*rep));
  const auto t0 = std::chrono::steady_clock::now();
  double rel = 0.0;
  const mpf_status st = newton_step(p, lam, o, guess, iters, rep, &rel);
  rep->final_relres = rel;
  rep->steps_done = st == MPF_OK ? 1 : 0;
  rep->seconds = seconds_since(t0);
  return st;
}

// MyTime (MyTime.cpp:26-106): update cur += dt; on failure cut (cur -= dt, dt *= 0.5, refuse below
// 1% of the initial dt) and reset the state; on success pop the step-size stack.
mpf_status mpf_newton_solve(mpf_problem* p, const mpf_newton_opts* o, int32_t* iters, mpf_newton_report* rep) {
  if (!p || p->etype == 0 || !o || !rep || o->nsteps < 0 || o->dt <= 0) return MPF_ERR_ARG;
  if (!p->prepared) return MPF_ERR_STATE;
  std::memset(rep, 0, sizeof(*rep));
  const auto t0 = std::chrono::steady_clock::now();
  const double EPS = 1e-10;  // MathBasic.h:17
  const double T = o->nsteps * o->dt;
  double cur = 0.0, dt = o->dt;
  std::vector<double> stack{dt};
  int done = 0;
  hipStream_t s = p->stream;
  const size_t n3 = (size_t)p->nNode * 3;
  while (cur <= T - EPS) {
    cur += dt;
    int32_t it = 0;
    double rel = 0.0;
    mpf_status st = newton_step(p, cur, o, nullptr, &it, rep, &rel);
    rep->final_relres = rel;
    if (st == MPF_OK) {
      if (iters && done < o->nsteps) iters[done] = it;
      ++done;
      const size_t m = stack.size();
      if (m == 1) dt = stack[0];
      else { dt = stack[m - 1]; stack.pop_back(); }
    } else if (st == MPF_ERR_NEG_JACOBIAN || st == MPF_ERR_NOT_CONVERGED || st == MPF_ERR_DIVERGED) {
      // reset (SolutionDataSolid.cpp:157-167) and cut
      MPF_HIP(hipMemcpyAsync(p->disp.ptr, p->dispPrev.ptr, n3 * sizeof(double), hipMemcpyDeviceToDevice, s));
      MPF_HIP(hipMemcpyAsync(p->pres.ptr, p->presPrev.ptr, p->npres * sizeof(double), hipMemcpyDeviceToDevice, s));
      cur -= dt;
      dt *= 0.5;
      if (dt < stack[0] * 0.01) {
        rep->steps_done = done;
        rep->seconds = seconds_since(t0);
        return st;
      }
      stack.push_back(dt);
    } else {
      return st;
    }
  }
  MPF_HIP(hipStreamSynchronize(s));
  rep->steps_done = done;
  rep->seconds = seconds_since(t0);
  return MPF_OK;
}

mpf_status mpf_time_assembly(mpf_problem* p, int32_t reps, int32_t first_iter, double lam, double* ms_per_assembly,
                             double* ms_scatter_kernels) {
  if (!p || p->etype == 0 || reps <= 0) return MPF_ERR_ARG;
  if (!p->prepared) return MPF_ERR_STATE;
  hipStream_t s = p->stream;
  hipEvent_t e0, e1;
  MPF_HIP(hipEventCreate(&e0));
  MPF_HIP(hipEventCreate(&e1));
  MPF_TRY(enqueue_assembly(p, first_iter != 0, lam, 1));  // warm-up
  MPF_HIP(hipEventRecord(e0, s));
  for (int r = 0; r < reps; ++r) MPF_TRY(enqueue_assembly(p, first_iter != 0, lam, 1));
  MPF_HIP(hipEventRecord(e1, s));
  MPF_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  MPF_HIP(hipEventElapsedTime(&ms, e0, e1));
  if (ms_per_assembly) *ms_per_assembly = ms / reps;
  if (ms_scatter_kernels) {
    // element kernels only (no zero fill / external forces)
    MPF_HIP(hipEventRecord(e0, s));
    for (int r = 0; r < reps; ++r) MPF_TRY(enqueue_element_kernels(p, first_iter != 0, 1));
    MPF_HIP(hipEventRecord(e1, s));
    MPF_HIP(hipEventSynchronize(e1));
    MPF_HIP(hipEventElapsedTime(&ms, e0, e1));
    *ms_scatter_kernels = ms / reps;
    // leave a complete system behind (the kernel-only sweeps accumulate into vals / rhs)
    MPF_TRY(enqueue_assembly(p, first_iter != 0, lam, 1));
    MPF_HIP(hipStreamSynchronize(s));
  }
  MPF_HIP(hipEventDestroy(e0));
  MPF_HIP(hipEventDestroy(e1));
  p->assembled = true;
  return MPF_OK;
}

}  // extern "C"
