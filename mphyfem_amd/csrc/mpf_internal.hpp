// mpf_internal.hpp -- shared internals of libmpf (gfx950).  Not part of the ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/mpf.h"

namespace mpf {

// Status carried through host code as an exception-free early return.
#define MPF_HIP(call)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess) {                                                             \
      std::fprintf(stderr, "libmpf: %s failed: %s (%s:%d)\n", #call, hipGetErrorString(e_), \
                   __FILE__, __LINE__);                                                 \
      return MPF_ERR_HIP;                                                               \
    }                                                                                   \
  } while (0)

#define MPF_TRY(call)                        \
  do {                                       \
    mpf_status s_ = (call);                  \
    if (s_ != MPF_OK) return s_;             \
  } while (0)

// Owned device allocation.
template <typename T>
struct DevBuf {
  T* ptr = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    n = 0;
  }
  mpf_status alloc(size_t count) {
    if (count == n && ptr) return MPF_OK;
    release();
    if (count == 0) return MPF_OK;
    if (hipMalloc(&ptr, count * sizeof(T)) != hipSuccess) {
      std::fprintf(stderr, "libmpf: hipMalloc of %zu bytes failed\n", count * sizeof(T));
      ptr = nullptr;
      return MPF_ERR_HIP;
    }
    n = count;
    return MPF_OK;
  }
  mpf_status upload(const T* h, size_t count, hipStream_t s) {
    MPF_TRY(alloc(count));
    if (count) MPF_HIP(hipMemcpyAsync(ptr, h, count * sizeof(T), hipMemcpyHostToDevice, s));
    return MPF_OK;
  }
  mpf_status download(T* h, size_t count, hipStream_t s) const {
    if (count > n) return MPF_ERR_ARG;
    if (count) MPF_HIP(hipMemcpyAsync(h, ptr, count * sizeof(T), hipMemcpyDeviceToHost, s));
    return MPF_OK;
  }
};

inline unsigned grid_for(size_t n, unsigned block) { return (unsigned)((n + block - 1) / block); }

// ---------------------------------------------------------------------------------------------
// Element-kernel launch description shared by the dense and the CSR-scatter variants.
// ---------------------------------------------------------------------------------------------
struct ElemArgs {
  const double* X;        // [nNode][3]
  const double* u;        // [nNode*3]
  const double* uprev;    // [nNode*3] (mixed)
  const double* pres;     // [nNode]   (mixed)
  const int32_t* conn;    // [nElem][npe]
  double mu, Kinv;
  const int32_t* emat;    // optional per-element material index (batched meshes), else mu / Kinv
  const double* mtab;     // [nmat][2]: mu, K = 1 / Kinv
  int32_t Utype;
  int32_t consistent;     // displacement element tangent mode (MPF_TANGENT_CONSISTENT)
  int32_t* bad;           // 1 + element id with det F < 0 (smallest id wins)
};

// Per-element CSR scatter description.
struct ScatterArgs {
  const int32_t* elems;     // element ids of this launch (colour slice of the permutation)
  int32_t count;
  const int32_t* slot;      // per-element slot bases (layout: see pattern.hip)
  const int32_t* nodeinfo;  // per node: (row length << 4) | (pressure flag << 3) | free mask (bits 0..2)
  const int32_t* nda;       // NodeDofArray [nNode][3]
  const int32_t* idp;       // ID_pres [nNode]
  int32_t dispDOF;
  double* vals;             // CSR values
  double* rhs;              // rhs [n]
  const double* dispApplied;  // [nNode*3]; non-null -> Dirichlet lift (first iteration)
  int32_t scatterK;           // 0: residual only (K kept in registers for the lift)
};

// Kernel launchers (elem.hip).
mpf_status launch_hex8_dense(const ElemArgs& a, int32_t first, int32_t count, double* Ke, double* Re,
                             hipStream_t s);
mpf_status launch_hex8p0_dense(const ElemArgs& a, int32_t first, int32_t count, double* Kc, double* Fc, double* Kup,
                               double* Kpp, double* Fp, hipStream_t s);
mpf_status launch_tet10_dense(const ElemArgs& a, int32_t first, int32_t count, double* Kuu, double* Kup,
                              double* Kpu, double* Kpp, double* Fu, double* Fp, hipStream_t s);
mpf_status launch_hex8_scatter(const ElemArgs& a, const ScatterArgs& sc, hipStream_t s);
mpf_status launch_hex8p0_scatter(const ElemArgs& a, const ScatterArgs& sc, double* Kup, double* Kpp, double* Fp,
                                 hipStream_t s);
mpf_status launch_hex8p0_pressure(int32_t nElem, const int32_t* conn, const int32_t* nda, const double* du,
                                  const double* dispApplied, const double* Kup, const double* Kpp, const double* Fp,
                                  double* pres, hipStream_t s, const int32_t* emat = nullptr,
                                  const int32_t* mflag = nullptr, int32_t stride = 0);
mpf_status launch_tet10_scatter(const ElemArgs& a, const ScatterArgs& sc, hipStream_t s);
// Tet10 displacement element (dense operator, coloured scatter with a 100-slot map per element)
mpf_status launch_tet10d_dense(const ElemArgs& a, int32_t first, int32_t count, double* Ke, double* Re,
                               hipStream_t s);
mpf_status launch_tet10d_scatter(const ElemArgs& a, const ScatterArgs& sc, hipStream_t s);

// Slot-map sizes per element.
constexpr int kHex8Slots = 64;                    // (a,b) node pairs
constexpr int kTet10Slots = 100 + 40 + 40 + 16;   // uu (a,b), up (a,q), pu (q,b), pp (q1,q2)
constexpr int kTet10DSlots = 100;                 // Tet10 displacement element: uu (a,b)

// ---------------------------------------------------------------------------------------------
// Problem state (the femSolidmechanics-shaped driver state on one device).
// ---------------------------------------------------------------------------------------------
struct PatternBufs {
  DevBuf<int32_t> nda;        // NodeDofArray [nNode][3]
  DevBuf<int32_t> assy;       // assyForSoln [dispDOF]
  DevBuf<int32_t> idp;        // ID_pres [nNode]
  DevBuf<int32_t> assyP;      // assyForSolnPres [presDOF]
  DevBuf<int32_t> nptr, nel;  // node -> element adjacency (CSR)
  DevBuf<int32_t> nbptr, nb;  // node neighbourhoods NB(n), sorted
  DevBuf<int32_t> rowptr, colidx;
  DevBuf<int32_t> nodeinfo;   // (row length << 4) | (pressure flag << 3) | free mask
  DevBuf<int32_t> color, perm;
  DevBuf<int32_t> slot;       // per-element slot bases
  // gather-assembly tables (gather.hip)
  DevBuf<int32_t> g_order;    // nodes in Morton (Z-curve) order of their coordinates
  DevBuf<int32_t> g_nelA;     // elements around each node (ascending), with the node's local index
  DevBuf<int2> g_fitem;       // laid-out (element, node) items of the owned blocks, 64 per wavefront
  DevBuf<int32_t> g_fnode;    // node whose rows each laid-out item belongs to
  DevBuf<uint8_t> g_fperm;    // per wavefront: the run-order position of the item each lane computes
  DevBuf<int32_t> g_btab;     // CSR location of every column block (n, j) and of its mirror (m, n)
  DevBuf<int32_t> g_near;     // node has a prescribed dof in an element around it
  DevBuf<int32_t> g_nearlist; // the flagged nodes
  int32_t g_nnear = 0;
  int64_t g_nwaves = 0;
  mpf::DevBuf<int32_t> g_cwave;  // chunked matrix kernel: first wavefront of each chunk
  mpf::DevBuf<int32_t> g_cnw;    // its number of wavefronts
  mpf::DevBuf<int32_t> g_ctab;   // distinct elements per chunk
  mpf::DevBuf<int32_t> g_ccnt;   // their number per chunk
  mpf::DevBuf<uint8_t> g_fslot;  // each laid-out item's element slot in its chunk
  int64_t g_nchunks = 0;
  bool gather_ready = false;
  std::vector<int32_t> color_off;  // host: [ncolors + 1] offsets into perm
  int32_t dispDOF = 0, presDOF = 0, ncolors = 0;
  int64_t nnz = 0;
};

struct KrylovWork;  // krylov.hip
struct PmgWork;     // pmg.hip
struct BatchWork;   // batch.hip
struct BandWork;    // band.hip
struct SystemWork;  // system.hip

}  // namespace mpf

struct mpf_problem {
  uint64_t uid = 0;  // unique per process (caches keyed by problem must not trust a reused address)
  int32_t etype = 0, npe = 0, nNode = 0, nElem = 0;
  mpf_matl m{};
  hipStream_t stream = nullptr;
  mpf::DevBuf<double> X;
  std::vector<double> X_h;  // host copy of X (traction integration at prepare)
  double xmin[3] = {0, 0, 0}, xmax[3] = {0, 0, 0};  // bounding box of the nodes
  mpf::DevBuf<int32_t> conn;
  // Dirichlet data (host lists, merged last-wins per dof like setBoundaryConditions)
  std::vector<int32_t> bc_dof_h;
  std::vector<double> bc_val_h;
  mpf::DevBuf<int32_t> bc_dof;
  mpf::DevBuf<double> bc_val;
  // traction patches (host) and the unit-load external force vector (device)
  struct Trac { std::vector<int32_t> faces; double v0, v1; int npf = 4; };  // Quad4 (4) or Tria6 (6) faces
  std::vector<Trac> trac;
  mpf::DevBuf<double> fext_unit;  // [nNode*3] for lambda = 1
  bool has_fext = false;
  mpf::PatternBufs pat;
  bool prepared = false;
  // Hex8 Q1/P0 (elem_type MPF_HEX8P0; etype is MPF_HEX8 for the pattern): one pressure per element,
  // and the last assembly's Kup [nElem][24], Kpp [nElem], Fp [nElem] for the pressure recovery
  bool p0 = false;
  mpf::DevBuf<double> q_Kup, q_Kpp, q_Fp;
  // state (pres, presPrev: [npres] = nNode, or nElem for Q1/P0)
  int64_t npres = 0;
  mpf::DevBuf<double> disp, dispPrev, pres, presPrev, dispApplied;
  // system
  mpf::DevBuf<double> vals, rhs, du;
  mpf::DevBuf<int32_t> bad;
  bool assembled = false;
  bool solved = false;
  mpf::KrylovWork* kw = nullptr;
  // assembly algorithm (MPF_ASSEMBLY_*) and the gather path's staging records
  int32_t asm_mode = MPF_ASSEMBLY_GATHER;
  int32_t band_split = 1;  // banded direct solve: cut long blocks at a separator (band.hip)
  int32_t chunk_cap = 32;  // Hex8 chunked matrix kernel: most distinct elements staged in LDS
  mpf::DevBuf<double> GS;
  mpf::DevBuf<double> scratch;  // reductions (norms)
  mpf::BatchWork* batch = nullptr;  // independent meshes solved together (batch.hip)
  mpf::BandWork* band = nullptr;    // banded LDL^T direct solver (band.hip)
  mpf::SystemWork* sys = nullptr;   // element maps of a caller-pattern system (etype 0, system.hip)
  mpf::PmgWork* pmg = nullptr;      // two-level p-multigrid preconditioner (pmg.hip)
};

namespace mpf {
uint64_t next_problem_uid();                        // problem.hip: unique per process
size_t stage_doubles_per_elem(int etype, bool p0);  // gather.hip
mpf_status ensure_gather_tables(mpf_problem* p);   // gather.hip (falls back to the scatter when unsupported)
mpf_status launch_gather_assembly(mpf_problem* p, const ElemArgs& A, const double* dispApplied, int writeK,
                                  hipStream_t s);
mpf_status build_pattern(mpf_problem* p);          // pattern.hip
// batch.hip
ElemArgs batch_elem_args(mpf_problem* p, ElemArgs a);
void batch_free(mpf_problem* p);
mpf_status batch_node_offsets(const mpf_problem* p, std::vector<int32_t>& off);
mpf_status batch_elem_offsets(const mpf_problem* p, std::vector<int32_t>& off);
mpf_status batch_set(mpf_problem* p, int32_t nmesh, const int32_t* node_off, const int32_t* elem_off,
                     const double* mu, const double* Kinv);
mpf_status batch_autotune(mpf_problem* p, int32_t* T, int32_t* LPR, double* ms_per_iter);
mpf_status batch_newton_solve(mpf_problem* p, const mpf_batch_opts* o, int32_t* mesh_iters, double* mesh_relres,
                              mpf_batch_report* rep);
mpf_status batch_newton_step(mpf_problem* p, double lam, const mpf_batch_opts* o, const double* guess,
                             int32_t* mesh_iters, double* mesh_relres, mpf_batch_report* rep);
mpf_status krylov_solve(mpf_problem* p, const mpf_krylov_opts* o, mpf_krylov_report* rep);  // krylov.hip
// band.hip
mpf_status band_prepare(mpf_problem* p, const std::vector<int32_t>& node_off);
mpf_status band_enqueue_solve(mpf_problem* p, const int32_t* active, int32_t astride);
mpf_status band_enqueue_resolve(mpf_problem* p);
mpf_status band_solve(mpf_problem* p, mpf_krylov_report* rep);
int32_t* band_bad_flag(mpf_problem* p);
void band_free(mpf_problem* p);
void krylov_free(mpf_problem* p);
void system_free(mpf_problem* p);                  // system.hip
// Most partial sums of one device reduction (one per workgroup of the vector kernels).
constexpr int kMaxPartials = 4096;
// krylov.hip: one CSR matrix and a block-diagonal preconditioner over runs of consecutive rows
// (blocks of 1..3 rows; binv holds each block's inverse as a 3x3 row-major array).
struct CsrView {
  int64_t n = 0, nnz = 0;
  const int32_t* rowptr = nullptr;
  const int32_t* colidx = nullptr;
  const double* vals = nullptr;
  const float* vals32 = nullptr;  // when set, the products read these FP32 copies of vals (sums in FP64)
};
struct BlockPcView {
  int64_t nblk = 0;
  const int32_t* bs = nullptr;    // first row of each block
  const int32_t* bm = nullptr;    // its number of rows
  const int32_t* rblk = nullptr;  // block of each row
  const double* binv = nullptr;   // [nblk][9]
  bool grp = false;               // every block's rows share one column list (row-group SpMV)
};
constexpr int kCgScalars = 16;    // device scalars of one CG run
struct CgScratch {                // n doubles each; part: 2 * kMaxPartials; sc: kCgScalars
  double *r, *r2, *z, *p, *q, *part, *sc;
};
// Block-preconditioned CG from x = 0 to ||b - A x|| <= rtol ||b|| or maxit iterations.
mpf_status pcg_csr(const CsrView& A, const BlockPcView& B, const CgScratch& w, const double* b, double* x,
                   double rtol, int maxit, hipStream_t s, mpf_krylov_report* rep);
// Block structure set-up: rblk from (bs, bm); *grp = 1 when every block's rows share one column list.
mpf_status block_pc_structure(const CsrView& A, int64_t nblk, const int32_t* bs, const int32_t* bm,
                              int32_t* rblk, bool* grp, hipStream_t s);
// Block inverses from the current values of A.
mpf_status block_pc_values(const CsrView& A, int64_t nblk, const int32_t* bs, const int32_t* bm, double* binv,
                           hipStream_t s);
// pmg.hip: the two-level preconditioner of Tet10 and Tet10/P1 systems.  pmg_setup builds the
// transfer operators once per pattern and the coarse matrix from the current values; fine: the
// 3x3 node blocks of the displacement rows (first nub blocks of the block-Jacobi list) and, for a
// mixed system, the pressure rows' Schur-diagonal blocks after them.  pmg_apply: z = M^-1 v.
mpf_status pmg_setup(mpf_problem* p, const BlockPcView& fine, int64_t nub);
mpf_status pmg_apply(mpf_problem* p, const BlockPcView& fine, int64_t nub, const double* v, double* z);
void pmg_free(mpf_problem* p);
// ||x||_2 to the host; scratch: kMaxPartials + 1 doubles of device memory.
mpf_status dev_norm2(const double* x, int64_t n, double* out_host, hipStream_t s, double* scratch);
}  // namespace mpf
