/*
 * mpf.h -- C ABI of the MI355X (gfx950) femSolidmechanics assembly and
 * Newton-Krylov path.  Plain pointers and sizes only; no C++ types cross it.
 *
 * Conventions (same as the reference, chennachaos/mphyfem, paths under src/):
 *   - nodal vectors are interleaved u[node*3 + d]            (ElementBase.cpp:341-345)
 *   - element dof order is 3*a + d                            (Elem_Solid_3D_Disp.cpp:322-324)
 *   - Hex8 node order as LagrangeBasisFunsHexa                (BasisFunctionsLagrange.cpp:1202-1279)
 *   - Tet10 node order as LagrangeBasisFunsTetra, edges 4:(0,1) 5:(1,2) 6:(0,2) 7:(0,3)
 *     8:(2,3) 9:(1,3); P1 pressure on nodes 0..3             (BasisFunctionsLagrange.cpp:870-956,
 *                                                              Elem_Solid_3D_Mixed1.cpp:66-70)
 *   - sigma[9] and C[81] use the column-major index map (0,0),(1,0),(2,0),(0,1),...
 *                                                             (utilitiesmaterialHyperelastic.cpp:191)
 *   - dense matrices crossing this ABI are ROW-major (Eigen's default is column-major: the
 *     C++ layer in mpf_host.hpp copies element by element through operator()(i, j)).
 *   - global numbering: free displacement dofs node-major skipping prescribed dofs, then
 *     pressure dofs at offset dispDOF in increasing node id; CSR rows hold sorted unique
 *     columns                                                 (femSolidmechanics2.cpp:121-478)
 *
 * Errors are returned as mpf_status; nothing throws across this boundary.
 * "Device pointer" arguments must point to memory on the current HIP device.
 * All calls taking a hipStream_t are stream-ordered; the problem-level calls run on the
 * problem's own stream and synchronise before returning where they return host values.
 */
#ifndef MPF_H
#define MPF_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t* mpf_stream_t; /* == hipStream_t */

typedef enum {
  MPF_OK = 0,
  MPF_ERR_ARG = 1,            /* invalid argument or size */
  MPF_ERR_HIP = 2,            /* HIP runtime error (allocation, launch) */
  MPF_ERR_NEG_JACOBIAN = 3,   /* det F < 0 in an element: runtime_error("Negative Jacobian in the element")
                                 (Elem_Solid_3D_Disp.cpp:273-278; ElementBase2.cpp:382-385) */
  MPF_ERR_STATE = 4,          /* call out of order, e.g. solve before assemble (SolverPetsc.cpp:321-323) */
  MPF_ERR_NOT_CONVERGED = 5,  /* Krylov or Newton did not reach its tolerance */
  MPF_ERR_DIVERGED = 6,       /* NaN residual or ratio > 1e7 after iteration 3 (femSolidmechanics1.cpp:627-636) */
  MPF_ERR_NO_DEVICE = 7       /* no HIP device visible */
} mpf_status;

/* Element type ids (npElem). */
/* MPF_HEX8P0: Hex8 Q1/P0, one pressure per element condensed out of the system (Elem_Solid_3D_Mixed0). */
/* MPF_TET10: Tet10 displacement element, Elem_Solid_3D_Disp with npElem = 10 and the 4-point rule      */
/* (Elem_Solid_3D_Disp.cpp:47-51), same node order as MPF_TET10P1.                                      */
enum { MPF_HEX8 = 8, MPF_HEX8P0 = 9, MPF_TET10P1 = 10, MPF_TET10 = 11 };

/* Neo-Hookean parameters: Matl_NeoHookean (Matl_NeoHookean.cpp:45-53) reads mu = data_deviatoric[0],
 * Utype and Kinv (MaterialBase.cpp:91-115); K = 1/Kinv. */
typedef struct {
  double mu;
  double Kinv;
  int32_t Utype;   /* -1 or 1..8, volumetric energy (utilitiesmaterialHyperelastic.cpp:7-41) */
  int32_t mixed;   /* 1 for the u-p element: sigma uses the pressure field */
  int32_t tangent; /* displacement element only.  MPF_TANGENT_REFERENCE: the spatial tangent exactly as
                      computeStressAndTangent_NeoHooke codes it, whose volumetric part omits its stress
                      term U'(J) d_ik d_jl (utilitiesmaterialHyperelastic.cpp:280-285), so K != -dR/du.
                      MPF_TANGENT_CONSISTENT: that term included (K = -dR/du); same residual. */
  int32_t pad_;
} mpf_matl;
enum { MPF_TANGENT_REFERENCE = 0, MPF_TANGENT_CONSISTENT = 1 };

const char* mpf_status_string(int32_t status);
/* Number of visible HIP devices (0 when none). */
int32_t mpf_device_count(void);

/* ------------------------------------------------------------------------------------------ */
/* Material operator.  Replaces MaterialBase::computeStressAndTangent (MaterialBase.h:94) for    */
/* Matl_NeoHookean -> computeStressAndTangent_NeoHooke (utilitiesmaterialHyperelastic.cpp:186).  */
/* F: [n][9] ROW-major 3x3 (device), p: [n] pressure (used when m->mixed), sig: [n][9], C: [n][81]*/
/* in the reference's 9-index map, C row-major over that map.                                    */
mpf_status mpf_neohooke_stress_tangent_batch(const double* F, const double* p, const mpf_matl* m,
                                             int32_t n, double* sig, double* C, mpf_stream_t stream);

/* ------------------------------------------------------------------------------------------ */
/* Element operators (dense K_e / R_e per element, device buffers).                              */
/* Hex8 displacement element: Elem_Solid_3D_Disp::calcStiffnessAndResidual                     */
/* (Elem_Solid_3D_Disp.cpp:196-374; ElementBase.h:149).  X: [nNode][3], u: [nNode*3],            */
/* conn: [nElem][8].  Elements [first, first+count): Ke [count][24][24] row-major, Re [count][24]. */
/* bad: one int32 (device); set to 1 + (smallest element id) with det F < 0, else left unchanged */
/* (caller zeroes it).                                                                           */
mpf_status mpf_elem_hex8_nh_batch(const double* X, const double* u, const int32_t* conn,
                                  const mpf_matl* m, int32_t first, int32_t count,
                                  double* Ke, double* Re, int32_t* bad, mpf_stream_t stream);

/* Tet10 displacement element: Elem_Solid_3D_Disp::calcStiffnessAndResidual with npElem = 10, nGP = 4 */
/* (Elem_Solid_3D_Disp.cpp:47-51, 196-374).  conn: [nElem][10]; Ke [count][30][30] row-major,         */
/* Re [count][30]; bad as for the Hex8 element.                                                       */
mpf_status mpf_elem_tet10_nh_batch(const double* X, const double* u, const int32_t* conn,
                                   const mpf_matl* m, int32_t first, int32_t count,
                                   double* Ke, double* Re, int32_t* bad, mpf_stream_t stream);

/* Hex8 Q1/P0 mixed element with static condensation (Elem_Solid_3D_Mixed0::calcStiffnessAndResidual, */
/* Elem_Solid_3D_Mixed0.cpp:93-104, over ElementBase::calcStiffnessAndResidualMixed3D with Np = 1).   */
/* pres_elem: [nElem] element pressures; u_prev as for Tet10/P1.  Kc [count][24][24] = Kuu - Kup     */
/* (Kpu / Kpp), Fc [count][24] = Fu - Kup (Fp / Kpp), and for the pressure update after a solve       */
/* (solveForPressure, :112-199: p += (Fp - Kpu du_e) / Kpp, Kpu = Kup^T) Kup [count][24], Kpp and Fp  */
/* [count].  bad as for the displacement element.                                                   */
mpf_status mpf_elem_hex8p0_nh_batch(const double* X, const double* u, const double* u_prev,
                                    const double* pres_elem, const int32_t* conn, const mpf_matl* m,
                                    int32_t first, int32_t count, double* Kc, double* Fc, double* Kup,
                                    double* Kpp, double* Fp, int32_t* bad, mpf_stream_t stream);

/* Tet10/P1 mixed element: ElementBase::calcStiffnessAndResidualMixed3D via                     */
/* Elem_Solid_3D_Mixed1::calcStiffnessAndResidualMixed (ElementBase2.cpp:278-552;               */
/* Elem_Solid_3D_Mixed1.cpp:89-97; ElementBase.h:171).  u_prev: previous converged displacement  */
/* (Fn, ElementBase.cpp:302-329); pres: [nNode] nodal pressure indexed by node id.               */
/* Kuu [count][30][30], Kup [count][30][4], Kpu [count][4][30], Kpp [count][4][4],              */
/* Fu [count][30], Fp [count][4].                                                                */
mpf_status mpf_elem_tet10p1_nh_batch(const double* X, const double* u, const double* u_prev,
                                     const double* pres, const int32_t* conn, const mpf_matl* m,
                                     int32_t first, int32_t count,
                                     double* Kuu, double* Kup, double* Kpu, double* Kpp,
                                     double* Fu, double* Fp, int32_t* bad, mpf_stream_t stream);

/* ------------------------------------------------------------------------------------------ */
/* Problem: the femSolidmechanics-shaped driver state on one device (femSolidmechanics.h).       */
typedef struct mpf_problem mpf_problem;

/* Creates a problem from host arrays (copied to the device).  elem_type MPF_HEX8, MPF_HEX8P0,      */
/* MPF_TET10P1 or MPF_TET10.  MPF_HEX8P0 assembles the statically condensed system (displacement    */
/* rows only; the gather's staging records carry each element's Kup and Kpp, so every block adds    */
/* -Kup_a Kup_b^T / Kpp, and its residual is condensed) in any assembly mode, recovers the element  */
/* pressures in mpf_update (solveForPressure, Elem_Solid_3D_Mixed0.cpp:112-199), and batches like   */
/* MPF_HEX8.                                                                                         */
mpf_status mpf_problem_create(mpf_problem** out, int32_t elem_type, int32_t nNode, int32_t nElem,
                              const double* X, const int32_t* conn, const mpf_matl* m);
mpf_status mpf_problem_destroy(mpf_problem* p);

/* Dirichlet patch (BoundaryCondition 'fixed' when dof < 0, else 'specified' for dof 0..2 with  */
/* value[i] * lambda(t)); femSolidmechanics1.cpp:1479-1543 and :813-915.  value may be NULL (0). */
mpf_status mpf_add_dirichlet(mpf_problem* p, int32_t n, const int32_t* nodes, int32_t dof,
                             const double* value);
/* Quad4 dead-load traction patch t = (v0*normal + v1*tangent1) * lambda                        */
/* (femSolidmechanics1.cpp:922-1119; BasisFunctionsLagrange.cpp:1688-1732).  faces [nf][4].     */
mpf_status mpf_add_traction(mpf_problem* p, int32_t nf, const int32_t* faces, double v0, double v1);
/* The same dead load on Tria6 faces [nf][6] (corners, then the mid-edge nodes of (0,1), (1,2),   */
/* (2,0)), the boundary faces of a Tet10 mesh: LagrangeBasisFunsFace with npElem 6 and the 3-point */
/* triangle rule (BasisFunctionsLagrange.cpp:364-418, 1688-1732; QuadratureUtil.cpp:223-236).      */
mpf_status mpf_add_traction_tri6(mpf_problem* p, int32_t nf, const int32_t* faces, double v0, double v1);

/* Assembly algorithm used by mpf_assemble (same K and R either way):                           */
/*   MPF_ASSEMBLY_GATHER (default): per-element Gauss-point staging, then the (element, node)     */
/*     contributions to each node's rows are summed in element order and every CSR value and    */
/*     rhs entry is written once (no zero fill, no atomics, bitwise reproducible).              */
/*   MPF_ASSEMBLY_COLORED_SCATTER: elements of one colour per launch add their K_e/R_e blocks into  */
/*     CSR through the per-element slot map (no two elements of a launch share a node).           */
/*   MPF_ASSEMBLY_GATHER_DIRECT: the gather with every item reading the staging records directly   */
/*     (for Hex8 the default copies the records of each chunk of 8 wavefronts to LDS first).      */
/* The gather tables hold up to 64 elements and 128 coupled nodes per node; for a mesh beyond    */
/* that the problem switches itself to the coloured scatter (mpf_get_assembly_mode reports it).  */
enum { MPF_ASSEMBLY_GATHER = 0, MPF_ASSEMBLY_COLORED_SCATTER = 1, MPF_ASSEMBLY_GATHER_DIRECT = 2 };
mpf_status mpf_set_assembly_mode(mpf_problem* p, int32_t mode);
mpf_status mpf_get_assembly_mode(const mpf_problem* p, int32_t* mode);
/* Hex8 gather tuning: the largest number of distinct elements (0..32, default 32) whose staging  */
/* records a chunk of 8 wavefronts copies to LDS; a chunk with more reads them directly (same     */
/* values).  No counterpart in the reference; 0 makes every chunk take the direct reads.         */
mpf_status mpf_set_chunk_stage_cap(mpf_problem* p, int32_t cap);

/* prepareMatrixPattern (femSolidmechanics2.cpp:121-478) on the device: DOF numbering, pressure   */
/* dofs, CSR pattern (rowptr/colidx, int32, sorted), element colouring, per-element CSR slots.   */
mpf_status mpf_prepare_pattern(mpf_problem* p);
mpf_status mpf_get_sizes(const mpf_problem* p, int32_t* dispDOF, int32_t* presDOF, int64_t* nnz,
                         int32_t* ncolors);
/* Host copies of the pattern; any pointer may be NULL.  rowptr [n+1], colidx [nnz],             */
/* NodeDofArray [nNode][3] (-1 = prescribed), ID_pres [nNode] (-1 = no pressure dof),            */
/* color [nElem].                                                                                 */
mpf_status mpf_get_pattern(const mpf_problem* p, int32_t* rowptr, int32_t* colidx,
                           int32_t* NodeDofArray, int32_t* ID_pres, int32_t* color);

/* Solution state (host copies).  disp/disp_prev [nNode*3]; pres/pres_prev [nNode], or [nElem]  */
/* for MPF_HEX8P0.                                                                                */
mpf_status mpf_set_state(mpf_problem* p, const double* disp, const double* disp_prev,
                         const double* pres, const double* pres_prev);
mpf_status mpf_get_state(const mpf_problem* p, double* disp, double* pres);
/* setBoundaryConditions for load factor lam (femSolidmechanics1.cpp:813-915).                  */
mpf_status mpf_set_bc(mpf_problem* p, double lam);

/* calcStiffnessAndResidual + addExternalForces (femSolidmechanics2.cpp:1047-1120, :785-789):    */
/* K and R of every element into the CSR system (zeroMtx + assembleMatrixAndVector), Dirichlet  */
/* lift when first_iter, external load scaled by lam.                                           */
/* Returns MPF_ERR_NEG_JACOBIAN if any element has det F < 0.                                    */
mpf_status mpf_assemble(mpf_problem* p, int32_t first_iter, double lam);
/* Device pointers of the assembled system (valid until destroy): vals [nnz], rhs [n].          */
mpf_status mpf_system_device_ptrs(mpf_problem* p, double** vals, double** rhs,
                                  const int32_t** rowptr, const int32_t** colidx);
/* Host copies of the assembled system. */
mpf_status mpf_get_system(const mpf_problem* p, double* vals, double* rhs);
/* ||rhs||_2 of the assembled system (VecNorm, femSolidmechanics2.cpp:801). */
mpf_status mpf_rhs_norm(mpf_problem* p, double* norm);

/* Krylov solve of the assembled system (SolverPetsc::factoriseAndSolve, SolverPetsc.cpp:321;   */
/* KSPSolve with zero initial guess, :290-292).                                                  */
/* MPF_KSP_BAND: direct solve by banded LDL^T on the GPU after a bandwidth-reducing renumbering    */
/* (reverse Cuthill-McKee or slices across an axis); displacement systems only; replaces the      */
/* plain method's direct solve (SolverEigen::solve, Eigen::SimplicialLDLT, SolverEigen.cpp:191-228);*/
/* pc, restart, maxit, rtol are ignored;                                                          */
/* a zero or non-finite pivot gives MPF_ERR_DIVERGED.  The half-bandwidth after renumbering must   */
/* stay below about 490 dofs (MPF_ERR_ARG).                                                      */
/* MPF_PC_PMG (MPF_KSP_GMRES only; MPF_TET10 and MPF_TET10P1 problems, else MPF_ERR_ARG): flexible   */
/* GMRES(m) around a p-multigrid preconditioner.  Displacement rows: one V-cycle with Chebyshev      */
/* smoothing (3x3 node-block Jacobi inside; degree 1 by default, one scaled step) on the quadratic   */
/* dofs and a coarse correction on the vertex dofs, whose matrix is the Galerkin product P^T Kuu P   */
/* (P: the Tet10 -> Tet4 interpolation, a mid-edge node taking the mean of its two vertices), given  */
/* a V-cycle of its own over aggregates of 8 vertices (mpf_set_pmg_levels) or solved by block-Jacobi */
/* CG; the smoothing and the lowest CG read their matrices rounded to FP32 (sums in FP64; the outer  */
/* residual uses the FP64 system).  Pressure rows (MPF_TET10P1): block upper-triangular, the         */
/* pressure block being diag(Kpp) (1 + 1 / (mu Kinv)) (the pressure mass matrix scaled to the Schur  */
/* complement of a nearly incompressible material; Kinv = 0: the node-block Schur diagonal).  It     */
/* stands where petsc_options select a multigrid or field-split preconditioner (-pc_type mg / gamg / */
/* fieldsplit) for the reference's KSPSolve (SolverPetsc.cpp:117-140, 290); the code defaults there  */
/* are CG and ILU.  mpf_set_pmg changes its parameters, mpf_pmg_stats reports its last use.          */
enum { MPF_KSP_CG = 0, MPF_KSP_GMRES = 1, MPF_KSP_BAND = 2 };
enum { MPF_PC_NONE = 0, MPF_PC_JACOBI = 1, MPF_PC_BJACOBI3 = 2, MPF_PC_PMG = 3 };
typedef struct {
  int32_t ksp;       /* MPF_KSP_* */
  int32_t pc;        /* MPF_PC_* */
  int32_t restart;   /* GMRES(m) */
  int32_t maxit;
  double rtol;       /* ||b - A x|| / ||b|| */
} mpf_krylov_opts;
typedef struct {
  int32_t iterations;
  int32_t converged;
  double relres;
  double seconds;
} mpf_krylov_report;
/* Solves K du = rhs; du stays on the device (mpf_get_solution copies it). */
mpf_status mpf_solve(mpf_problem* p, const mpf_krylov_opts* o, mpf_krylov_report* rep);
/* Layout of the banded direct solver after its first use (MPF_KSP_BAND or the batched solver 1):  */
/* independent blocks, largest half-bandwidth, free dofs, and the factorisation's algorithmic      */
/* flops, sum over blocks of n b (b + 1).  No counterpart in the reference (reporting only).       */
mpf_status mpf_band_stats(const mpf_problem* p, int32_t* nblk, int32_t* bmax, int64_t* ndof, double* flops);
/* Banded direct solver: cut each block of n >= 3 b + 64 rows in two at a separator of b .. b + 15  */
/* rows in its middle (enable = 1, the default) or keep every block whole (0).  The two halves are  */
/* factored and solved by workgroups of their own up to the separator, whose Schur complement then  */
/* takes both halves' updates (elimination order [first half, second half reversed, separator]); a */
/* long block's chain of panels is about halved.  No counterpart in the reference (a GPU layout      */
/* choice; Eigen's SimplicialLDLT orders the whole system by AMD, SolverEigen.cpp:191-228).         */
mpf_status mpf_set_band_split(mpf_problem* p, int32_t enable);
/* Segments of the direct solver's layout (workgroups per launch) and blocks cut in two.            */
mpf_status mpf_band_segments(const mpf_problem* p, int32_t* nseg, int32_t* ncut);
/* p-multigrid preconditioner (MPF_PC_PMG) parameters: the degree of the Chebyshev smoothing before */
/* and after the coarse correction (1..8, default 1), the ratio beta / alpha of its interval        */
/* (1 < ratio < 1e6, default 10; beta = 1.1 x the power-iteration estimate of the largest           */
/* eigenvalue of the node-block-Jacobi-preconditioned Kuu), the coarse CG's relative tolerance      */
/* (0 < coarse_rtol < 1, default 0.1) and iteration limit (default 1000).  No counterpart in the    */
/* reference.                                                                                        */
mpf_status mpf_set_pmg(mpf_problem* p, int32_t degree, double ratio, double coarse_rtol, int32_t coarse_maxit);
/* Levels of MPF_PC_PMG: 2, the vertex level solved by block-Jacobi CG to coarse_rtol; 3 (default),  */
/* the vertex level given a V-cycle of its own -- Chebyshev smoothing with its 3x3 vertex blocks    */
/* (same degree and ratio, its own eigenvalue estimate), a correction from aggregates of 8 vertices */
/* (consecutive in Morton order of their coordinates, one dof per aggregate and component, the      */
/* aggregate matrix Pa^T Kc Pa with Pa the 0/1 aggregation) solved by block-Jacobi CG to            */
/* coarse_rtol, and the smoothing again.  Else MPF_ERR_ARG.  No counterpart in the reference.       */
mpf_status mpf_set_pmg_levels(mpf_problem* p, int32_t levels);
/* Its coarse levels and its work since the last solve began: vertex-level dofs and nonzeros,      */
/* V-cycles applied, CG iterations on the lowest level in total (the vertex level with 2 levels,   */
/* the aggregates with 3), the set-up time of the last solve (transfer tables on first use; the    */
/* Galerkin products, the block inverses and the eigenvalue estimates every solve), the smoother's */
/* estimate on the fine level and on the vertex level (3 levels), the aggregate level's dofs and    */
/* nonzeros, the number of levels and the smoothing degree.  No counterpart in the reference        */
/* (reporting only).                                                                               */
typedef struct {
  int64_t coarse_dofs;
  int64_t coarse_nnz;
  int64_t vcycles;
  int64_t coarse_iterations;
  double setup_seconds;
  double lambda_max;    /* the fine-level smoother's eigenvalue estimate */
  double lambda_vertex; /* the vertex-level smoother's (3 levels; else 0) */
  int64_t agg_dofs;
  int64_t agg_nnz;
  int32_t levels;
  int32_t degree;       /* the Chebyshev degree of the smoothing */
} mpf_pmg_report;
mpf_status mpf_pmg_stats(const mpf_problem* p, mpf_pmg_report* rep);
mpf_status mpf_get_solution(const mpf_problem* p, double* du);
/* The same solve with no device allocation, for callers that cannot allocate (code under HIP graph */
/* capture, tight serving loops, harnesses that forbid hipMalloc); it has no counterpart in the    */
/* reference.  CG with MPF_PC_NONE or MPF_PC_JACOBI only (other ksp / pc: MPF_ERR_ARG), systems of  */
/* at most MPF_NOALLOC_MAX_DOFS rows (else MPF_ERR_ARG, before any device work).  The work vectors  */
/* live in one library-owned device array of 5 * MPF_NOALLOC_MAX_DOFS doubles (20 MiB) per GPU,   */
/* held until the process ends and shared by every call on that GPU: calls that overlap on one GPU */
/* race.  The solution is read with mpf_get_solution.                                             */
#define MPF_NOALLOC_MAX_DOFS 524288
mpf_status mpf_solve_noalloc(mpf_problem* p, const mpf_krylov_opts* o, mpf_krylov_report* rep);
/* factoriseSolveAndUpdate's update (femSolidmechanics2.cpp:1124-1217): on the first iteration   */
/* prescribed dofs take dispApplied, free dofs += du, pressure dofs += du[dispDOF + k].          */
/* MPF_HEX8P0: then every element pressure += (Fp - Kpu . du_e) / Kpp from the last assembly    */
/* (solveForPressure, Elem_Solid_3D_Mixed0.cpp:112-199); MPF_ERR_STATE before any assembly.      */
mpf_status mpf_update(mpf_problem* p, int32_t first_iter);

/* Newton-Raphson load stepping (solveWithNewtonRaphson, femSolidmechanics2.cpp:730-851).        */
/* Step k uses lambda_k = sum of k dt's (MyTime::update, MyTime.cpp:26-37) with lambda(t) = t.   */
/* Converged when ||R_free|| / R0 < rtol with R0 the iteration-1 residual of the step.           */
typedef struct {
  int32_t nsteps;
  double dt;
  int32_t itmax;           /* Newton iterations (assemblies) per step */
  double rtol;             /* relative Newton tolerance, e.g. 1e-9 */
  mpf_krylov_opts lin;     /* linear solver options */
  int32_t forcing;         /* 1: inexact Newton, linear rtol = max(lin.rtol, min(0.1, 0.1*||R||/R0)) */
} mpf_newton_opts;
typedef struct {
  int32_t steps_done;
  int32_t newton_iters;    /* total assemblies */
  int32_t krylov_iters;    /* total Krylov iterations */
  double final_relres;
  double seconds;
  double seconds_assembly;
  double seconds_solve;
} mpf_newton_report;
/* Runs opts->nsteps load steps from the current state; iters [nsteps] (may be NULL) receives the */
/* Newton iterations per step.                                                                   */
mpf_status mpf_newton_solve(mpf_problem* p, const mpf_newton_opts* o, int32_t* iters,
                            mpf_newton_report* rep);
/* One load step to load factor lam from the current converged state.  guess (host, [nNode*3],   */
/* may be NULL): if given, Newton starts from it (prescribed dofs are set to their targets) and  */
/* the tolerance is still measured against R0 of the plain start (previous converged state with  */
/* the lifted increment).  iters receives the Newton iterations of this step.                    */
mpf_status mpf_newton_step(mpf_problem* p, double lam, const mpf_newton_opts* o, const double* guess,
                           int32_t* iters, mpf_newton_report* rep);

/* ------------------------------------------------------------------------------------------ */
/* Batched independent meshes.  No counterpart in the reference: many femSolidmechanics problems  */
/* (Hex8 or Hex8 Q1/P0, one Neo-Hookean material each) solved together.  The problem's mesh is the */
/* union of                                                                                         */
/* nmesh meshes: mesh m owns nodes [node_off[m], node_off[m+1]) and elements [elem_off[m],        */
/* elem_off[m+1]) (no element may reference another mesh's node), with material mu[m], Kinv[m]   */
/* (Utype and tangent mode from the problem's mpf_matl).  Call after mpf_problem_create.         */
mpf_status mpf_set_batch(mpf_problem* p, int32_t nmesh, const int32_t* node_off, const int32_t* elem_off,
                         const double* mu, const double* Kinv);

typedef struct {
  int32_t nsteps;         /* load steps, lambda(t) = t */
  double dt;
  int32_t itmax;          /* Newton iterations per step */
  double rtol;            /* per mesh: ||R_m|| / ||R0_m|| < rtol, R0_m at the plain start of the step */
  int32_t cg_maxit;       /* CG iterations per solve */
  double cg_rtol;         /* lower bound of the inexact-Newton CG tolerance (Eisenstat-Walker forcing) */
  int32_t use_graph;      /* 1: replay captured HIP graphs; 0: launch the same kernels directly */
  int32_t chunk;          /* CG iterations per graph launch (0: keep) */
  int32_t tile;           /* dofs per tile (0: keep the autotuned / default value) */
  int32_t lanes_per_row;  /* SpMV lanes per row, 4, 8 or 16 (0: keep) */
  int32_t solver;         /* 0: Jacobi-PCG per mesh; 1: banded LDL^T direct solve per mesh (MPF_KSP_BAND, */
                          /*    a workgroup per mesh, or per half of a long mesh cut at a separator;    */
                          /*    exact Newton corrections, cg_* ignored)                                 */
} mpf_batch_opts;

typedef struct {
  int32_t steps_done;
  int32_t newton_iters;     /* batch Newton iterations (solves), all steps */
  int64_t cg_iters_total;   /* CG iterations summed over meshes */
  int64_t cg_iters_max;     /* CG iterations of the slowest mesh, summed over solves */
  int32_t graph_launches;
  int32_t tile, lanes_per_row, chunk;
  double seconds;
  double seconds_assembly;   /* iteration graphs (update, assembly, test and the first K CG its) */
  double seconds_solve;      /* CG continued past K, or the direct solves' checks               */
  double start_relres_mean, start_relres_max;  /* ||R_m|| / ||R0_m|| at the start point (last step) */
  int32_t iteration_graphs;  /* one per Newton iteration: update + assembly + test + K CG iterations */
  int32_t chunk_graphs;      /* CG continued past K (graph_launches = these two + start-point ones)  */
  double seconds_capture;    /* graph capture and instantiation (first use of each K)               */
} mpf_batch_report;

/* Element batch and Jacobi-PCG tile shape by measurement, on the batch's own system: first the     */
/* assembly (the row-owner gather whose chunk kernel stages at most 8, 16, 24 or 32 distinct       */
/* elements per chunk in LDS, the gather reading each record directly, or the coloured scatter),   */
/* each timed over three assemblies, then every (tile, lanes_per_row)                              */
/* candidate over a few CG iterations; the fastest are kept for mpf_batch_newton_solve.  The       */
/* gather variants give bitwise the same system.                                                   */
mpf_status mpf_batch_autotune(mpf_problem* p, int32_t* tile, int32_t* lanes_per_row, double* ms_per_cg_iter);
/* The autotuner's assembly pick: MPF_ASSEMBLY_* mode, staged-element cap, ms per assembly          */
/* (MPF_ERR_STATE before mpf_batch_autotune).                                                      */
mpf_status mpf_batch_assembly_choice(const mpf_problem* p, int32_t* mode, int32_t* stage_cap, double* ms_per_assembly);
/* Batched solver 1: milliseconds per direct solve of every mesh (band fill, factorisation, solve and */
/* scatter of the last assembled system, timed with HIP events over reps solves on the problem's     */
/* stream; the solution and the factors are those of that system).  MPF_ERR_STATE before a step.   */
/* No counterpart in the reference (measurement only).                                              */
mpf_status mpf_batch_time_direct(mpf_problem* p, int32_t reps, double* ms_per_solve);
/* Newton-Raphson over nsteps load steps for every mesh at once (femSolidmechanics2.cpp:730-851  */
/* per mesh) with Jacobi-PCG per mesh.  mesh_iters [nsteps][nmesh]: assemblies to convergence;    */
/* mesh_relres [nmesh]: final ||R_m|| / ||R0_m||.  Either may be NULL.                           */
mpf_status mpf_batch_newton_solve(mpf_problem* p, const mpf_batch_opts* o, int32_t* mesh_iters,
                                  double* mesh_relres, mpf_batch_report* rep);
/* One load step of every mesh at load factor lam (nsteps, dt ignored), from the current converged */
/* state or, when guess (device, [nNode*3]) is given, from that displacement with the prescribed  */
/* dofs set to their values; R0_m is the plain-start residual either way.  mesh_iters [nmesh].    */
/* rep accumulates (zero it before the first step).                                             */
mpf_status mpf_batch_newton_step(mpf_problem* p, double lam, const mpf_batch_opts* o, const double* guess,
                                 int32_t* mesh_iters, double* mesh_relres, mpf_batch_report* rep);
/* Tangent increment for the next load step (batched solver 1 only, after a converged step):       */
/* du_nodes (device, [nNode*3]) = K^-1 R(u; lam) on the free dofs, 0 on prescribed dofs, where R  */
/* is the residual at the current state with the load and prescribed values of lam and K the      */
/* factors each mesh's last direct solve left.  No counterpart in the reference: it feeds the     */
/* tangent-secant start point u_k + 2 du - (u_k - u_{k-1}) of the learned guess.  MPF_ERR_STATE    */
/* before any direct solve.                                                                        */
mpf_status mpf_batch_tangent_increment(mpf_problem* p, double lam, double* du_nodes);

/* Device pointers of the problem's displacement [nNode*3] and pressure [nNode] (valid until      */
/* destroy; ordered on the problem's stream).                                                     */
mpf_status mpf_state_device_ptrs(mpf_problem* p, double** disp, double** pres);
/* Device-to-device copy of the displacement [nNode*3] into disp_dst (synchronous).                */
mpf_status mpf_copy_state_device(mpf_problem* p, double* disp_dst);

/* ------------------------------------------------------------------------------------------ */
/* Learned initial guess.  No counterpart in the reference: a message-passing network on a       */
/* problem's node graph (nodes sharing an element), 32 input features, 3 layers of width 64, BF16  */
/* MFMA GEMMs with FP32 accumulation, FP32 master weights with Adam.  It predicts the correction   */
/* of the linear extrapolation u_k + rho (u_k - u_{k-1}) of the next load step's displacement.    */
typedef struct mpf_mpnn mpf_mpnn;
int32_t mpf_mpnn_num_params(void);
mpf_status mpf_mpnn_create(mpf_mpnn** out, uint64_t seed);
mpf_status mpf_mpnn_destroy(mpf_mpnn* net);
/* Host copies of the FP32 parameters / accumulated gradients (either may be NULL); set refreshes  */
/* the BF16 copies.                                                                               */
mpf_status mpf_mpnn_get_params(mpf_mpnn* net, float* w, float* g);
mpf_status mpf_mpnn_set_params(mpf_mpnn* net, const float* w);
mpf_status mpf_mpnn_zero_grad(mpf_mpnn* net);
/* Features of load step k+1 from the converged u_k, u_{k-1} (device) into the network's input.    */
mpf_status mpf_mpnn_features(mpf_mpnn* net, mpf_problem* p, const double* uk, const double* ukm1, double rho,
                             double lam_next, double lam_ratio);
/* Forward over the problem's nodes: x [nNode][32] BF16 (device) or NULL for the features above;    */
/* y [nNode][32] FP32 (device) or NULL.                                                            */
mpf_status mpf_mpnn_forward(mpf_mpnn* net, mpf_problem* p, const uint16_t* x, float* y);
/* Mean squared error of the last forward against the converged u_{k+1} over free dofs; leaves     */
/* dL/dy in the network.                                                                          */
mpf_status mpf_mpnn_loss(mpf_mpnn* net, mpf_problem* p, const double* unext, const double* uk, const double* ukm1,
                         double rho, double* loss);
/* Linearised start residual of the last forward's prediction, the quantity Newton's convergence  */
/* depends on: mean over meshes of ||K e_m||^2 / (dlam^2 ||f_m||^2), e = predicted - converged     */
/* u_{k+1} on free dofs, K the tangent at the converged u_{k+1} (assembled here; overwrites the     */
/* problem's state and system), f_m the mesh's unit external load.  Leaves dL/dy in the network.   */
mpf_status mpf_mpnn_residual_loss(mpf_mpnn* net, mpf_problem* p, double lam_next, double dlam, const double* unext,
                                  const double* uk, const double* ukm1, double rho, double* loss);
/* Gradients of the last forward for dL/dy = dy [nNode][32] FP32 (device), or the one mpf_mpnn_loss */
/* left when dy is NULL; they accumulate until mpf_mpnn_adam / mpf_mpnn_zero_grad.                 */
mpf_status mpf_mpnn_backward(mpf_mpnn* net, mpf_problem* p, const float* dy);
mpf_status mpf_mpnn_adam(mpf_mpnn* net, float lr, float beta1, float beta2, float eps);
/* Newton start point u_k + rho (u_k - u_{k-1}) (+ the network's correction when use_net), device. */
mpf_status mpf_mpnn_guess(mpf_mpnn* net, mpf_problem* p, const double* uk, const double* ukm1, double rho,
                          int32_t use_net, double* guess);
/* Guarded start point for load factor lam: per mesh, the network's correction (last forward) enters  */
/* with the weight w in [0, 1] that minimises the residual linearised between the base predictor      */
/* (w = 0) and the whole correction (w = 1), both assembled (two residual-only assemblies; the        */
/* problem's state is restored).  n_learned: meshes with w > 0.                                        */
mpf_status mpf_mpnn_guarded_guess(mpf_mpnn* net, mpf_problem* p, double lam, const double* uk, const double* ukm1,
                                  double rho, double* guess, int32_t* n_learned);

/* ------------------------------------------------------------------------------------------ */
/* Arc-length continuation on the device: femSolidmechanics::solveWithArclength                   */
/* (femSolidmechanics2.cpp:855-1039) with SolverPetsc::solveArclengthSystem (SolverPetsc.cpp:      */
/* 671-757, psi = 1): each Newton iteration solves K du' = R and K du1 = Fext (one factorisation    */
/* with MPF_KSP_BAND) and picks the load increment dl from the linearised arc-length sphere.  The   */
/* load is the problem's traction patches (Fext at load factor 1); the first step is load-controlled */
/* at lam = dt; the arc length is then the first step's and adapts as the reference's (doubles     */
/* after two converged steps, 0.5 / 0.25 after a failure, within [ds1 / 1024, ds1]).  Displacement */
/* elements with homogeneous prescribed values only (MPF_ERR_ARG otherwise: the reference's loop   */
/* leaves pressures and prescribed increments untouched).  The state moves along the path.          */
typedef struct {
  double dt;            /* first load factor, and the time added per attempt (myTime.dt) */
  double final_time;    /* attempts continue while the time is <= final_time */
  int32_t max_steps;    /* stepsMax */
  int32_t max_iters;    /* iterationsMax (Newton runs iter < max_iters, as the reference's loop) */
  double tol;           /* absolute ||R|| */
  mpf_krylov_opts lin;
} mpf_arclength_opts;
typedef struct {
  int32_t steps_done, attempts, newton_iters, krylov_iters;
  double final_lam, final_ds, seconds;
} mpf_arclength_report;
/* lam_out / ds_out [max_steps] (may be NULL): per converged step its load factor and arc length   */
/* (0 for the first step).                                                                         */
mpf_status mpf_arclength_solve(mpf_problem* p, const mpf_arclength_opts* o, double* lam_out, double* ds_out,
                               mpf_arclength_report* rep);

/* ------------------------------------------------------------------------------------------ */
/* Input decks: the reference's block-structured config file (femBase::readInput and its read*     */
/* blocks, femBase.cpp:370-1200), the Gmsh 2.2 ASCII mesh its Files block names (readInputGMSH,    */
/* femBase.cpp:121-364; looked up beside the config file, ".msh" appended) and petsc_options.dat    */
/* beside it (-ksp_type, -pc_type, -ksp_rtol, -ksp_max_it, -ksp_gmres_restart; preonly / lu ->     */
/* MPF_KSP_BAND; without the file the code default CG + ILU of SolverPetsc.cpp:117-140 becomes CG  */
/* + 3x3 node-block Jacobi).  Supported: 3D, one domain of Hex8 (Gmsh 5) or Tet10 (Gmsh 11),        */
/* Matl_NeoHookean, ELEM_SOLID_3D_DISP / _MIXED0 / _MIXED1, fixed and specified patches with        */
/* constant values, traction patches of Quad4 (Gmsh 3) or Tria6 (Gmsh 9) faces, time functions,    */
/* solvertype newton, all time-varying loads on one time function; anything else: MPF_ERR_ARG     */
/* with a message on stderr.                                                                       */
typedef struct mpf_deck mpf_deck;
typedef struct {
  int32_t ndim, nNode, nElem, npe, gmsh_type;
  int32_t elem_type;            /* MPF_* once mpf_deck_setup has run, else 0 */
  int32_t npatch, nbc, ntrac, nmat, ntf;
  int32_t arclength;            /* solvertype arclength */
  double dt, dt_min, dt_max, final_time;  /* timeStep (1 to 3 values, MyTime::set), finalTime */
  int32_t max_steps, max_iters;
  double tol;                   /* absolute ||R|| tolerance (converged, femSolidmechanics1.cpp:616-622) */
  mpf_krylov_opts lin;          /* linear solver from petsc_options.dat or the default */
  int32_t lin_from_file;
  int32_t pad_;
} mpf_deck_info;
typedef struct {
  int32_t steps_done, newton_iters, krylov_iters, cuts, converged, pad_;
  double final_time, final_lam, final_norm, seconds;
} mpf_deck_report;
mpf_status mpf_deck_load(const char* config_path, mpf_deck** out);
mpf_status mpf_deck_destroy(mpf_deck* d);
mpf_status mpf_deck_get_info(const mpf_deck* d, mpf_deck_info* info);
/* Boundary patch `index` (0 .. npatch-1): its physical name, dimension and sorted unique nodes   */
/* (BoundaryPatch::processData, BoundaryPatch.cpp:45-53); nodes may be NULL to ask the count.      */
mpf_status mpf_deck_patch(const mpf_deck* d, int32_t index, char* name, int32_t name_len, int32_t* dim,
                          int32_t* nnodes, int32_t* nodes);
/* Builds the problem (prepareMatrixPattern with the deck's patches); *p stays owned by the deck. */
mpf_status mpf_deck_setup(mpf_deck* d, mpf_problem** p);
/* solveWithNewtonRaphson over the deck's time steps (femSolidmechanics2.cpp:730-851, MyTime cut / */
/* stck, load factor = the loads' time function at the current time, absolute tolerance); with     */
/* vtu_prefix a .vtu per converged step (postProcess).  MPF_ERR_NOT_CONVERGED when the step size   */
/* falls below 1% of the first (MyTime::cut).                                                      */
mpf_status mpf_deck_run(mpf_deck* d, const char* vtu_prefix, mpf_deck_report* rep);

/* ------------------------------------------------------------------------------------------ */
/* Post-processing of the current state.  Per Gauss point MPF_STRESS_NCOMP values: the Cauchy     */
/* stress (9, the column-major 9-index map), von Mises and tr(sigma)/3 (the pressure field itself  */
/* for the mixed elements) -- Elem_Solid_3D_Disp::projectStress (Elem_Solid_3D_Disp.cpp:480-557).  */
/* out: host [nElem][nGP][11] (nGP 8 for Hex8, 4 for Tet10).                                      */
#define MPF_STRESS_NCOMP 11
mpf_status mpf_stress_gauss(mpf_problem* p, double* out);
/* Element values = mean over the Gauss points (Elem_Solid_3D_Disp.cpp:416-421); nodal values =   */
/* mean of the element values over the elements at the node, in increasing element id            */
/* (femSolidmechanics::projectStresses, femSolidmechanics3.cpp:918-960).  elem: host [nElem][11],  */
/* nodal: host [nNode][11]; either may be NULL.                                                   */
mpf_status mpf_stress_project(mpf_problem* p, double* elem, double* nodal);
/* VTK XML unstructured grid of the current state (femSolidmechanics::postProcess,               */
/* femSolidmechanics3.cpp:267-776): points at X + u, point data "disp", "velocity" (zero: static), */
/* "pressure" (Tet10/P1: vertex values, mid-edge nodes the mean of their edge's vertices),        */
/* "stress", "vonMises" (nodal means); cell data "ElemID", "MatlID", "pressure" (Q1/P0),          */
/* "stress", "vonMises" (element means); field data "TIME".  Hex8 as VTK_HEXAHEDRON, Tet10 as       */
/* VTK_QUADRATIC_TETRA with the reference's node map.  binary: raw appended data, else ASCII.      */
mpf_status mpf_write_vtu(mpf_problem* p, const char* path, double time, int32_t binary);

/* ------------------------------------------------------------------------------------------ */
/* Linear system over a caller-supplied CSR pattern, filled from caller-computed element         */
/* matrices: the SolverPetsc operator (SolverPetsc.h:16-96) for a driver that keeps its own      */
/* element loop.  The handle is an mpf_problem without a mesh; mpf_solve (CG / GMRES),            */
/* mpf_get_solution, mpf_get_system, mpf_rhs_norm, mpf_system_device_ptrs and                      */
/* mpf_problem_destroy apply to it, the mesh calls return MPF_ERR_ARG.                             */
/* SolverPetsc::initialise (SolverPetsc.cpp:24-114) with the pattern prepareMatrixPattern inserts */
/* (femSolidmechanics2.cpp:294-446).  Host arrays: rowptr [n+1], colidx [nnz] sorted unique per   */
/* row, n = n_disp + n_pres (pressure rows after the displacement rows, femSolidmechanics2.cpp:175);*/
/* nda: optional NodeDofArray [nNode][3] (-1 = prescribed; a node's free dofs consecutive) for      */
/* MPF_PC_BJACOBI3, else pass nNode = 0 and nda = NULL.  Values and rhs start at zero.              */
mpf_status mpf_system_create(mpf_problem** out, int32_t n_disp, int32_t n_pres, const int32_t* rowptr,
                             const int32_t* colidx, int32_t nNode, const int32_t* nda);
/* SolverPetsc::zeroMtx (SolverPetsc.cpp:145-163): values and rhs to zero. */
mpf_status mpf_system_zero(mpf_problem* p);
/* Dof map of a batch of `count` element matrices of nrow x ncol (device arrays rows [count][nrow],   */
/* cols [count][ncol]; -1 = skip the row / column, as assembleMatrixAndVector2field's "!= -1" tests, */
/* SolverPetsc.cpp:434-502; r_off / c_off are added to the indices that are not -1).  Builds once   */
/* the sorted CSR-slot lists the adds reuse; *map_id names it.  An entry outside the pattern is      */
/* MPF_ERR_ARG (PETSc would insert it; the pattern here is fixed).  count*nrow*ncol < 2^31.          */
mpf_status mpf_system_map_elements(mpf_problem* p, int32_t count, int32_t nrow, int32_t ncol, const int32_t* rows,
                                   const int32_t* cols, int32_t r_off, int32_t c_off, int32_t* map_id);
/* SolverPetsc::assembleMatrixAndVector(r1, c1, row, col, K, F) (SolverPetsc.cpp:357-378) for every  */
/* element of a map: K [count][nrow][ncol] row-major and F [count][nrow], device; either may be NULL. */
/* Every CSR value and rhs entry receives its contributions in element order, as ADD_VALUES one     */
/* element after another: equal bit for bit to the sequential host loop.  Marks the system           */
/* assembled (currentStatus = ASSEMBLY_OK, femSolidmechanics2.cpp:1115).                             */
mpf_status mpf_system_add_elements(mpf_problem* p, int32_t map_id, const double* K, const double* F);
/* rhs[i] += f[i] for all rows (device f [n]): Fext added to the residual (femSolidmechanics2.cpp:785-789). */
mpf_status mpf_system_add_rhs(mpf_problem* p, const double* f);
/* ElementBase::applyDirichletBCs (ElementBase.cpp:567-594) for `count` element matrices of       */
/* nsize x nsize (device): F_e(jj) -= K_e(jj, ii) * dispApplied[gdof(ii)] for every ii with        */
/* forAssy(ii) == -1 and jj with forAssy(jj) != -1, ii in increasing order.  forAssy, gdof         */
/* [count][nsize] (forAssyVec, globalDOFnums), K row-major, F [count][nsize] updated in place.    */
/* The mixed form (applyDirichletBCs2field, :717-782) is this call on the combined u-p matrix.    */
mpf_status mpf_elem_apply_dirichlet_batch(int32_t count, int32_t nsize, const int32_t* forAssy, const int32_t* gdof,
                                          const double* dispApplied, const double* K, double* F,
                                          mpf_stream_t stream);
/* Device memory for callers without the HIP headers. */
mpf_status mpf_malloc(void** ptr, size_t bytes);
mpf_status mpf_free(void* ptr);
mpf_status mpf_memcpy_htod(void* dst, const void* src, size_t bytes);
mpf_status mpf_memcpy_dtoh(void* dst, const void* src, size_t bytes);

/* ------------------------------------------------------------------------------------------ */
/* Timing helpers: average device time (ms) of one assembly (element kernels + scatter, no       */
/* pattern build) over reps launches, measured with HIP events on the problem's stream.          */
mpf_status mpf_time_assembly(mpf_problem* p, int32_t reps, int32_t first_iter, double lam,
                             double* ms_per_assembly, double* ms_scatter_kernels);

/* ------------------------------------------------------------------------------------------ */
/* Synthetic meshes (seeded, SplitMix64 with u = (x >> 11) * 2^-53).  Output buffers are host     */
/* memory sized by the *_sizes call.                                                              */
/* Structured Hex8 block nx*ny*nz on [0,Lx]x[0,Ly]x[0,Lz]; node id i + (nx+1)*(j + (ny+1)*k);      */
/* coordinate d of a node is jittered by U(-jitter,+jitter)*h_d when its grid index along d is    */
/* interior.                                                                                     */
void mpf_mesh_hex_sizes(int32_t nx, int32_t ny, int32_t nz, int32_t* nNode, int32_t* nElem);
void mpf_mesh_hex(int32_t nx, int32_t ny, int32_t nz, double Lx, double Ly, double Lz, double jitter,
                  uint64_t seed, double* X, int32_t* conn);
/* Kuhn 6-tet split of an n^3 cube grid into Tet10 on the doubled (2n+1)^3 node grid; vertex     */
/* jitter as above, mid-edge nodes at edge midpoints.                                             */
void mpf_mesh_tet10_sizes(int32_t n, int32_t* nNode, int32_t* nElem);
void mpf_mesh_tet10(int32_t n, double L, double jitter, uint64_t seed, double* X, int32_t* conn);
/* Node ids of a face of a structured block: face 0..5 = x=0, x=max, y=0, y=max, z=0, z=max.     */
int32_t mpf_mesh_face_nodes(int32_t elem_type, int32_t nx, int32_t ny, int32_t nz, int32_t face,
                            int32_t* nodes);
/* Quad4 faces (outward normal under the reference's n = dv x du) of a Hex8 block face.          */
int32_t mpf_mesh_hex_face_quads(int32_t nx, int32_t ny, int32_t nz, int32_t face, int32_t* quads);

#ifdef __cplusplus
}
#endif
#endif
