// host_caller.cpp -- a compiled C++ caller of the GPU path through include/mpf_host.hpp, written the
// way femSolidmechanics::calcStiffnessAndResidual drives its elements and solver
// (femSolidmechanics2.cpp:1047-1120): zeroMtx, per element calcStiffnessAndResidual[Mixed], the
// first-iteration Dirichlet lift, assembleMatrixAndVector[2field], then factoriseAndSolve.
//
//   host_caller hex8 OUTDIR      10^3 Hex8 cube (x=0 fixed, x=1 UX = 0.1 lam), lam = 0.1, one
//                                Newton iteration from a perturbed state; per-element path and the
//                                device batch path; element, material and error-path checks
//   host_caller tet10p1 OUTDIR   2^3 x 6 Tet10/P1 cube (z=0 fixed, z=1 UZ = -0.05 lam), the 2field path
//   host_caller --link-check     loads and links only (no device call)
//
// Writes raw little-endian arrays (*.f64, *.i32) and report.txt into OUTDIR; the test compares
// them with the oracle.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "mpf_host.hpp"

using mpf_host::MatrixXd;
using mpf_host::VectorXd;

namespace {

std::string g_out;
std::FILE* g_rep = nullptr;

template <typename T>
void dump(const char* name, const T* a, size_t n) {
  std::ofstream f(g_out + "/" + name, std::ios::binary);
  f.write(reinterpret_cast<const char*>(a), (std::streamsize)(n * sizeof(T)));
}
template <typename T>
void dump(const char* name, const std::vector<T>& a) {
  dump(name, a.data(), a.size());
}
template <class Mat>
std::vector<double> rowmajor(const Mat& K) {
  std::vector<double> a((size_t)K.rows() * K.cols());
  for (int i = 0; i < K.rows(); ++i)
    for (int j = 0; j < K.cols(); ++j) a[(size_t)i * K.cols() + j] = K(i, j);
  return a;
}

// Pattern of a mesh from the device build (prepareMatrixPattern, femSolidmechanics2.cpp:121-478).
struct Pattern {
  int dispDOF = 0, presDOF = 0;
  int64_t nnz = 0;
  std::vector<int32_t> rowptr, colidx, nda, idp;
};

Pattern mesh_pattern(int etype, int nNode, int nElem, const std::vector<double>& X, const std::vector<int32_t>& conn,
                     const mpf_matl& m, const std::vector<int32_t>& fixed, const std::vector<int32_t>& loaded,
                     int ldof, double lval) {
  mpf_problem* p = nullptr;
  mpf_host::check(mpf_problem_create(&p, etype, nNode, nElem, X.data(), conn.data(), &m), "mpf_problem_create");
  std::vector<double> v(loaded.size(), lval);
  mpf_host::check(mpf_add_dirichlet(p, (int32_t)fixed.size(), fixed.data(), -1, nullptr), "mpf_add_dirichlet");
  mpf_host::check(mpf_add_dirichlet(p, (int32_t)loaded.size(), loaded.data(), ldof, v.data()), "mpf_add_dirichlet");
  mpf_host::check(mpf_prepare_pattern(p), "mpf_prepare_pattern");
  Pattern P;
  int32_t nc = 0;
  mpf_host::check(mpf_get_sizes(p, &P.dispDOF, &P.presDOF, &P.nnz, &nc), "mpf_get_sizes");
  const int n = P.dispDOF + P.presDOF;
  P.rowptr.resize(n + 1);
  P.colidx.resize((size_t)P.nnz);
  P.nda.resize(3 * (size_t)nNode);
  P.idp.resize((size_t)nNode);
  std::vector<int32_t> color((size_t)nElem);
  mpf_host::check(mpf_get_pattern(p, P.rowptr.data(), P.colidx.data(), P.nda.data(), P.idp.data(), color.data()),
                  "mpf_get_pattern");
  mpf_problem_destroy(p);
  return P;
}

std::vector<int32_t> face(int etype, int n, int f) {
  std::vector<int32_t> nodes((size_t)(2 * n + 1) * (2 * n + 1) + (size_t)(n + 1) * (n + 1));
  const int k = mpf_mesh_face_nodes(etype, n, n, n, f, nodes.data());
  nodes.resize(k);
  return nodes;
}

// dispApplied = target * lam - disp on the prescribed dofs (setBoundaryConditions,
// femSolidmechanics1.cpp:864-868, 890)
std::vector<double> disp_applied(int nNode, const std::vector<double>& u, const std::vector<int32_t>& fixed,
                                 const std::vector<int32_t>& loaded, int ldof, double target, double lam) {
  std::vector<double> a(3 * (size_t)nNode, 0.0);
  for (int32_t v : fixed)
    for (int d = 0; d < 3; ++d) a[3 * (size_t)v + d] = 0.0 * lam - u[3 * (size_t)v + d];
  for (int32_t v : loaded) a[3 * (size_t)v + ldof] = target * lam - u[3 * (size_t)v + ldof];
  return a;
}

int run_hex8() {
  const int n = 10;
  int nNode = 0, nElem = 0;
  mpf_mesh_hex_sizes(n, n, n, &nNode, &nElem);
  std::vector<double> X(3 * (size_t)nNode);
  std::vector<int32_t> conn(8 * (size_t)nElem);
  mpf_mesh_hex(n, n, n, 1.0, 1.0, 1.0, 0.05, 1, X.data(), conn.data());
  const double E = 1e3, nu = 0.3, mu = E / (2.0 * (1.0 + nu)), Kinv = 3.0 * (1.0 - 2.0 * nu) / E;
  mpf_host::MatlNeoHookean matl(mu, Kinv, 3);
  const std::vector<int32_t> fixed = face(MPF_HEX8, n, 0), loaded = face(MPF_HEX8, n, 1);
  const Pattern P = mesh_pattern(MPF_HEX8, nNode, nElem, X, conn, matl.data(), fixed, loaded, 0, 0.1);
  const double lam = 0.1;
  std::vector<double> u(3 * (size_t)nNode);
  for (size_t i = 0; i < u.size(); ++i) u[i] = 2e-3 * std::sin(0.37 * (double)i + 0.5);
  const std::vector<double> dispApplied = disp_applied(nNode, u, fixed, loaded, 0, 0.1, lam);
  dump("u.f64", u);
  dump("dispApplied.f64", dispApplied);
  dump("rowptr.i32", P.rowptr);
  dump("colidx.i32", P.colidx);

  mpf_host::MeshData geom(8, nNode, nElem, X.data(), conn.data());
  geom.setDisplacement(u);
  mpf_host::ElementWorkspace ws;
  std::vector<mpf_host::ElemSolid3DDisp> elems;
  elems.reserve(nElem);
  for (int e = 0; e < nElem; ++e) {
    elems.emplace_back(e, &geom, &matl, &ws);
    elems.back().prepareElemData(std::vector<int32_t>(conn.begin() + 8 * (size_t)e, conn.begin() + 8 * (size_t)e + 8),
                                 P.nda.data());
  }

  // solver on the caller's pattern; factoriseAndSolve before any assembly returns -1
  mpf_host::SolverGpu solver;
  solver.initialise(P.dispDOF, P.presDOF, P.rowptr, P.colidx, nNode, P.nda.data());
  std::fprintf(g_rep, "solve_before_assembly %d\n", solver.factoriseAndSolve());
  solver.opts.ksp = MPF_KSP_CG;
  solver.opts.pc = MPF_PC_BJACOBI3;
  solver.opts.rtol = 1e-13;
  solver.opts.maxit = 20000;

  // the reference's loop (femSolidmechanics2.cpp:1059-1113), first Newton iteration
  MatrixXd Kuu;
  VectorXd FlocalU;
  solver.zeroMtx();
  for (int e = 0; e < nElem; ++e) {
    elems[e].calcStiffnessAndResidual(Kuu, FlocalU);
    if (e == 0) {
      dump("Ke0.f64", rowmajor(Kuu));
      dump("Fe0.f64", FlocalU.data(), (size_t)FlocalU.size());
    }
    elems[e].applyDirichletBCs(Kuu, FlocalU, dispApplied);
    solver.assembleMatrixAndVector(0, 0, elems[e].forAssyVec, elems[e].forAssyVec, Kuu, FlocalU);
  }
  dump("vals.f64", solver.matrixValues());
  dump("rhs.f64", solver.rhsVec());
  const int st = solver.factoriseAndSolve();
  std::fprintf(g_rep, "solve_status %d iterations %d relres %.3e\n", st, solver.report.iterations,
               solver.report.relres);
  dump("soln.f64", solver.solnVec());

  // the same iteration through the device batch path: all elements in one launch, the lift on
  // the device, one scatter in element order
  {
    mpf_host::DeviceArray<double> Kb((size_t)nElem * 576), Fb((size_t)nElem * 24), dA;
    mpf_host::DeviceArray<int32_t> bad(1), fa, gd;
    std::vector<int32_t> hfa((size_t)nElem * 24), hgd((size_t)nElem * 24);
    for (int e = 0; e < nElem; ++e)
      for (int k = 0; k < 24; ++k) {
        hfa[(size_t)e * 24 + k] = elems[e].forAssyVec[k];
        hgd[(size_t)e * 24 + k] = elems[e].globalDOFnums[k];
      }
    fa.upload(hfa);
    gd.upload(hgd);
    dA.upload(dispApplied);
    int32_t z = 0;
    bad.upload(&z, 1);
    mpf_host::check(mpf_elem_hex8_nh_batch(geom.X(), geom.disp(), geom.conn(), &matl.data(), 0, nElem, Kb.get(),
                                           Fb.get(), bad.get(), nullptr),
                    "mpf_elem_hex8_nh_batch");
    mpf_host::check(mpf_elem_apply_dirichlet_batch(nElem, 24, fa.get(), gd.get(), dA.get(), Kb.get(), Fb.get(), nullptr),
                    "mpf_elem_apply_dirichlet_batch");
    mpf_host::SolverGpu s2;
    s2.initialise(P.dispDOF, P.presDOF, P.rowptr, P.colidx, nNode, P.nda.data());
    s2.zeroMtx();
    s2.assembleElementsDevice(nElem, 24, 24, fa.get(), fa.get(), Kb.get(), Fb.get());
    dump("vals_batch.f64", s2.matrixValues());
    dump("rhs_batch.f64", s2.rhsVec());
    // a second assembly on the same map (the next Newton iteration reuses it): zeroMtx, then add
    s2.zeroMtx();
    s2.assembleElementsDevice(nElem, 24, 24, fa.get(), fa.get(), Kb.get(), Fb.get());
    const std::vector<double> v2 = s2.matrixValues();
    std::fprintf(g_rep, "batch_reassembly_equal %d\n",
                 (int)(std::memcmp(v2.data(), solver.matrixValues().data(), v2.size() * sizeof(double)) == 0));
  }

  // material operator at one deformation gradient
  {
    MatrixXd F(3, 3), Fn(3, 3), C;
    VectorXd stre;
    const double a[9] = {1.05, 0.02, -0.01, 0.03, 0.97, 0.04, -0.02, 0.01, 1.1};
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) F(i, j) = Fn(i, j) = a[3 * i + j];
    double pres = 0.0;
    matl.computeStressAndTangent(true, 0, Fn, F, pres, stre, C);
    dump("matF.f64", a, 9);
    dump("stre.f64", stre.data(), 9);
    dump("Cmat.f64", rowmajor(C));
  }

  // a negative Jacobian throws the reference's runtime_error
  {
    std::vector<double> ub(u);
    const int32_t* c = &conn[0];
    for (int d = 0; d < 3; ++d) ub[3 * (size_t)c[6] + d] = -2.0 * (X[3 * (size_t)c[6] + d] - X[3 * (size_t)c[0] + d]);
    geom.setDisplacement(ub);
    std::string msg = "none";
    try {
      elems[0].calcStiffnessAndResidual(Kuu, FlocalU);
    } catch (const std::runtime_error& err) {
      msg = err.what();
    }
    std::fprintf(g_rep, "negative_jacobian %s\n", msg.c_str());
  }
  return 0;
}

int run_tet10p1() {
  const int n = 2;
  int nNode = 0, nElem = 0;
  mpf_mesh_tet10_sizes(n, &nNode, &nElem);
  std::vector<double> X(3 * (size_t)nNode);
  std::vector<int32_t> conn(10 * (size_t)nElem);
  mpf_mesh_tet10(n, 1.0, 0.1, 2, X.data(), conn.data());
  mpf_host::MatlNeoHookean matl(1.0, 1e-3, 3, true);
  const std::vector<int32_t> fixed = face(MPF_TET10P1, n, 4), loaded = face(MPF_TET10P1, n, 5);
  const Pattern P = mesh_pattern(MPF_TET10P1, nNode, nElem, X, conn, matl.data(), fixed, loaded, 2, -0.05);
  const double lam = 1.0;
  std::vector<double> u(3 * (size_t)nNode), up(3 * (size_t)nNode), pr((size_t)nNode), pa((size_t)nNode, 0.0);
  for (size_t i = 0; i < u.size(); ++i) {
    u[i] = 3e-3 * std::sin(0.41 * (double)i + 0.2);
    up[i] = 0.5 * u[i];
  }
  for (size_t i = 0; i < pr.size(); ++i) pr[i] = 0.01 * std::cos(0.73 * (double)i);
  const std::vector<double> dispApplied = disp_applied(nNode, u, fixed, loaded, 2, -0.05, lam);
  dump("u.f64", u);
  dump("uprev.f64", up);
  dump("pres.f64", pr);
  dump("rowptr.i32", P.rowptr);
  dump("colidx.i32", P.colidx);

  mpf_host::MeshData geom(10, nNode, nElem, X.data(), conn.data());
  geom.setDisplacement(u);
  geom.setPreviousDisplacement(up);
  geom.setPressure(pr);
  mpf_host::ElementWorkspace ws;
  mpf_host::SolverGpu solver;
  solver.initialise(P.dispDOF, P.presDOF, P.rowptr, P.colidx, nNode, P.nda.data());
  solver.opts.ksp = MPF_KSP_GMRES;
  solver.opts.pc = MPF_PC_BJACOBI3;
  solver.opts.restart = 100;
  solver.opts.rtol = 1e-12;
  solver.opts.maxit = 5000;
  MatrixXd Kuu, Kup, Kpu, Kpp;
  VectorXd Fu, Fp;
  solver.zeroMtx();
  for (int e = 0; e < nElem; ++e) {
    mpf_host::ElemSolid3DMixed1 el(e, &geom, &matl, &ws);
    el.prepareElemData(std::vector<int32_t>(conn.begin() + 10 * (size_t)e, conn.begin() + 10 * (size_t)e + 10),
                       P.nda.data(), P.idp.data());
    el.calcStiffnessAndResidualMixed(Kuu, Kup, Kpu, Kpp, Fu, Fp);
    el.applyDirichletBCs2field(Kuu, Kup, Kpu, Kpp, Fu, Fp, dispApplied, pa);
    solver.assembleMatrixAndVector2field(P.dispDOF, Kuu, Kup, Kpu, Kpp, Fu, Fp, el.forAssyVec, el.forAssyVecPres);
  }
  dump("vals.f64", solver.matrixValues());
  dump("rhs.f64", solver.rhsVec());
  const int st = solver.factoriseAndSolve();
  std::fprintf(g_rep, "solve_status %d iterations %d relres %.3e\n", st, solver.report.iterations,
               solver.report.relres);
  dump("soln.f64", solver.solnVec());
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc == 2 && std::strcmp(argv[1], "--link-check") == 0) {
    std::printf("linked: %s\n", mpf_status_string(MPF_ERR_NEG_JACOBIAN));
    return 0;
  }
  if (argc != 3) {
    std::fprintf(stderr, "usage: host_caller {hex8|tet10p1} OUTDIR | --link-check\n");
    return 2;
  }
  g_out = argv[2];
  g_rep = std::fopen((g_out + "/report.txt").c_str(), "w");
  if (!g_rep) return 2;
  int rc = 2;
  try {
    if (std::strcmp(argv[1], "hex8") == 0) rc = run_hex8();
    else if (std::strcmp(argv[1], "tet10p1") == 0) rc = run_tet10p1();
  } catch (const std::exception& e) {
    std::fprintf(g_rep, "exception %s\n", e.what());
    rc = 1;
  }
  std::fclose(g_rep);
  return rc;
}
