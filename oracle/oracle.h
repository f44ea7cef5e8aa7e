/*
 * oracle.h -- CPU FP64 restatement of the femSolidmechanics assembly and
 * Newton-Krylov path of chennachaos/mphyfem.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the test oracle: tests/, the smoke
 * entry point and the CPU-baseline leg of bench.py may load it; the product
 * library (mphyfem_amd/libmpf.so) never links, loads or calls it.
 *
 * Every function follows a reference file (paths relative to the reference
 * repository root, src/...).  Conventions kept from the reference:
 *   - nodal vectors interleaved u[node*3 + d]           (ElementBase.cpp:341-345)
 *   - element dof order 3*a + d                          (Elem_Solid_3D_Disp.cpp:322-324)
 *   - sigma[9] and Cmat[9x9] use the column-major index map
 *     (0,0),(1,0),(2,0),(0,1),(1,1),(2,1),(0,2),(1,2),(2,2)
 *                                                        (utilitiesmaterialHyperelastic.cpp:191)
 *   - all dense matrices crossing this C ABI are ROW-major.
 */
#ifndef MPF_ORACLE_H
#define MPF_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- quadrature and basis (QuadratureUtil.cpp, BasisFunctionsLagrange.cpp) ---- */
int  orc_gauss3d(int npElem, double* xi /*[ngp][3]*/, double* w /*[ngp]*/);   /* returns ngp */
int  orc_gauss_quad4(double* xi /*[4][2]*/, double* w /*[4]*/);
void orc_basis_hex8(const double xi[3], double* N /*8*/, double* dN /*[3][8]*/);
void orc_basis_tet(int npElem, const double xi[3], double* N, double* dN /*[3][npElem]*/);
void orc_basis_quad4(const double xi[2], double* N /*4*/, double* dN /*[2][4]*/);

/* 3x3 determinant and inverse by partial-pivot LU (the arithmetic Eigen's
 * MatrixXd::determinant()/inverse() use for a dynamic 3x3). */
void orc_det_inv3(const double A[9], double* det, double Ainv[9]);

/* ---- material (utilitiesmaterialHyperelastic.cpp:45-295) ---- */
double orc_dUdJ(int Utype, double K, double J);
double orc_d2UdJ2(int Utype, double K, double J);
void   orc_vol_constants(int finite, int Utype, double Jn, double K, double* Jhat, double* thetahat);
/* F row-major 3x3; stre[9] and Cmat[81] (Cmat row-major over the 9-index map) */
int  orc_neohooke(int mixed, int Utype, double Kinv, double mu, const double F[9], double pres,
                  double* stre, double* Cmat);

/* ---- elements ---- */
/* Elem_Solid_3D_Disp::calcStiffnessAndResidual (Elem_Solid_3D_Disp.cpp:196-374).
 * X,u: [npe][3]; Ke row-major [3npe][3npe]; Re [3npe].  Returns 0, or 1 on det F < 0. */
int  orc_elem_disp(int npe, const double* X, const double* u, double mu, double Kinv, int Utype,
                   double* Ke, double* Re);
/* Same residual; tangent with the volumetric stress term the reference omits (consistent
 * linearisation, K = -dR/du).  Used by Newton solves; not the reference's K_e. */
int  orc_elem_disp_consistent(int npe, const double* X, const double* u, double mu, double Kinv, int Utype,
                              double* Ke, double* Re);
/* ElementBase::calcStiffnessAndResidualMixed3D (ElementBase2.cpp:278-552), Tet10/P1.
 * X,u,uprev: [10][3]; pres: [4] (corner values).  Kuu [30][30], Kup [30][4],
 * Kpu [4][30], Kpp [4][4], Fu [30], Fp [4]. */
/* Hex8 Q1/P0 (Elem_Solid_3D_Mixed0): condensed Kc[24x24], Fc[24]; uncondensed Kuu, Fu, Kup[24],
   Kpu[24], Kpp[1], Fp[1] */
int  orc_elem_hex8p0(const double* X, const double* u, const double* uprev, double pe, double mu, double Kinv,
                     int Utype, double* Kc, double* Fc, double* Kuu, double* Fu, double* Kup, double* Kpu,
                     double* Kpp, double* Fp);
int  orc_elem_tet10p1(const double* X, const double* u, const double* uprev, const double* pres,
                      double mu, double Kinv, int Utype,
                      double* Kuu, double* Kup, double* Kpu, double* Kpp, double* Fu, double* Fp);

/* ---- problem: numbering, pattern, colouring, assembly, Newton ---- */
typedef struct orc_problem orc_problem;

/* elem_type 8 = Hex8 displacement, 9 = Hex8 Q1/P0, 10 = Tet10/P1 mixed, 11 = Tet10 displacement.
 * X [nNode][3], conn [nElem][npe]. */
orc_problem* orc_problem_create(int elem_type, int nNode, int nElem, const double* X, const int32_t* conn,
                                double mu, double Kinv, int Utype);
void orc_problem_destroy(orc_problem* p);
/* Tangent mode for the displacement element: 0 = as the reference codes it, 1 = consistent. */
void orc_set_tangent(orc_problem* p, int consistent);
/* Dirichlet patch: dof = -1 -> 'fixed' (all three dofs), else 'specified' for that dof with
 * value[i] * lambda(t) (femSolidmechanics1.cpp:1479-1543, :813-915). */
int  orc_add_dirichlet(orc_problem* p, int n, const int32_t* nodes, int dof, const double* value);
/* Quad4 dead-load traction patch (femSolidmechanics1.cpp:922-1119): faces [nf][4],
 * t = (v0*n + v1*tangent1) * lambda. */
int  orc_add_traction(orc_problem* p, int nf, const int32_t* faces, double v0, double v1);
/* Tria6 faces [nf][6] (corners, then (0,1), (1,2), (2,0)), 3-point rule: the Tet10 boundary faces */
int  orc_add_traction_tri6(orc_problem* p, int nf, const int32_t* faces, double v0, double v1);
/* The traction patches' external force at load factor lam, F [nNode * 3] (addExternalForces). */
void orc_external_force(const orc_problem* p, double lam, double* F);
/* DOF numbering + CSR pattern (femSolidmechanics2.cpp:121-478). */
int  orc_prepare_pattern(orc_problem* p);
void orc_sizes(const orc_problem* p, int32_t* dispDOF, int32_t* presDOF, int64_t* nnz);
void orc_get_pattern(const orc_problem* p, int32_t* rowptr, int32_t* colidx,
                     int32_t* NodeDofArray /*[nNode][3]*/, int32_t* ID_pres /*[nNode]*/);
/* Element colouring: sequential greedy first-fit in decreasing order of
 * mpf_color_priority(e).  Returns number of colours. */
int  orc_color(const orc_problem* p, int32_t* color /*[nElem]*/);

/* State access: disp [nNode*3], pres [nNode] (indexed by node id, as SolnData.pres). */
void orc_set_state(orc_problem* p, const double* disp, const double* dispPrev,
                   const double* pres, const double* presPrev);
void orc_get_state(const orc_problem* p, double* disp, double* pres);
/* setBoundaryConditions for load factor lam: dispApplied = value*lam - disp on prescribed dofs. */
void orc_set_bc(orc_problem* p, double lam);
/* calcStiffnessAndResidual + addExternalForces (femSolidmechanics2.cpp:1047-1120, :785-789).
 * first_iter != 0 applies the Dirichlet lift.  vals [nnz], rhs [n].  Returns 0 or 1 (det F<0). */
int  orc_assemble(orc_problem* p, int first_iter, double lam, double* vals, double* rhs);

/* Krylov solves on a CSR system (CPU).  precond: 0 none, 1 Jacobi, 2 3x3 node block-Jacobi.
 * Returns iterations; *relres = final ||b-Ax||/||b||. */
int  orc_pcg(const orc_problem* p, const double* vals, const double* b, double* x, int precond,
             double rtol, int maxit, double* relres);
int  orc_gmres(const orc_problem* p, const double* vals, const double* b, double* x, int precond,
               int restart, double rtol, int maxit, double* relres);
/* Banded LDL^T direct solve of the CSR system (symmetric positive definite). 0 on success. */
int  orc_direct_solve(const orc_problem* p, const double* vals, const double* b, double* x);
/* y = A x */
void orc_spmv(const orc_problem* p, const double* vals, const double* x, double* y);

/* Newton-Raphson over load steps (femSolidmechanics2.cpp:730-851, :1124-1217).
 * solver: 0 direct (banded LDL^T), 1 PCG (precond as above), 2 GMRES.
 * Convergence: ||R_free|| / R0 < rtol, R0 = ||R_free|| of iteration 1 of the step
 * (previous converged u with the step's Dirichlet increment lifted).
 * iters_out[nsteps] receives Newton iterations (assemblies) per step. Returns 0 on success. */
int  orc_newton(orc_problem* p, int nsteps, double dt, int solver, int precond, double lin_rtol,
                double rtol, int itmax, int32_t* iters_out, double* final_relres);

/* One element-assembly sweep only (timing helper for the CPU baseline): assembles elements
 * [first, first+count) into vals/rhs.  Returns 0 or 1. */
int  orc_assemble_range(orc_problem* p, int first, int count, double* vals, double* rhs);
/* all elements, colour by colour, OpenMP threads within a colour (CPU-baseline timing only) */
int  orc_assemble_colored(orc_problem* p, const int32_t* color, int nthreads, double* vals, double* rhs);

/* Colour priority shared by oracle and product: SplitMix64 finaliser of (e + golden). */
uint64_t orc_color_priority(int64_t e);
/* Post-processing (Elem_Solid_3D_Disp::projectStress :480-557; femSolidmechanics::projectStresses,
   femSolidmechanics3.cpp:918-960): out [nElem][nGP][11] = sigma (9, column-major map), von Mises,
   tr(sigma)/3; returns nGP.  elem [nElem][11] Gauss-point means, nodal [nNode][11] means over the
   elements around each node. */
int  orc_stress_gauss(const orc_problem* p, double* out);
/* Arc-length continuation (femSolidmechanics2.cpp:855-1039; SolverPetsc.cpp:671-757), direct solves:
   per converged step the load factor and the arc length used; returns the converged steps. */
int  orc_arclength(orc_problem* p, double dt, double timeFinal, int stepsMax, int iterationsMax, double conv_tol,
                   double* lam_out, double* ds_out, double* u_out, int32_t* attempts_out);
void orc_stress_project(const orc_problem* p, double* elem, double* nodal);

#ifdef __cplusplus
}
#endif
#endif
