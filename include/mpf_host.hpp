// mpf_host.hpp -- header-only C++17 host layer over the C ABI in mpf.h, shaped like the reference's
// element, material and solver classes so that a femSolidmechanics-style driver can call the GPU
// path with its own loop (chennachaos/mphyfem, paths under src/):
//
//   mpf_host::ElemSolid3DDisp      Elem_Solid_3D_Disp::calcStiffnessAndResidual (Hex8, Tet10)
//                                  (ElementBase.h:149; Elem_Solid_3D_Disp.cpp:196-374) and
//                                  ElementBase::applyDirichletBCs (ElementBase.cpp:567-594)
//   mpf_host::ElemSolid3DMixed1    calcStiffnessAndResidualMixed for Tet10/P1 (ElementBase.h:171;
//                                  ElementBase2.cpp:278-552) and applyDirichletBCs2field (:717-782)
//   mpf_host::MatlNeoHookean       MaterialBase::computeStressAndTangent (MaterialBase.h:91-94;
//                                  Matl_NeoHookean.cpp:45-53)
//   mpf_host::SolverGpu            SolverPetsc: initialise / zeroMtx / assembleMatrixAndVector /
//                                  assembleMatrixAndVector2field / factoriseAndSolve / currentStatus
//                                  (SolverPetsc.h:16-96; SolverEigen.h:11 for the status values)
//
// Dense matrices: any type with resize(r, c), rows(), cols() and operator()(i, j), and vectors with
// resize(n), size() and operator()(i) -- Eigen::MatrixXd / VectorXd, or the small column-major
// mpf_host::MatrixXd / VectorXd below when Eigen is absent.  The C ABI passes row-major arrays; the
// copies here go through operator(), so the storage order of the caller's type does not matter.
//
// Errors: a negative Jacobian in an element throws std::runtime_error("Negative Jacobian in the
// element") as Elem_Solid_3D_Disp.cpp:277 does; other ABI failures throw mpf_host::Error.  The
// solver returns the reference's int codes and, unlike SolverPetsc::solve (SolverPetsc.cpp:301-302),
// a nonzero code when the Krylov method misses its tolerance.
//
// Link with libmpf.so (-lmpf); needs only mpf.h, no HIP headers.
#ifndef MPF_HOST_HPP
#define MPF_HOST_HPP

#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "mpf.h"

namespace mpf_host {

// ---------------------------------------------------------------------------------------------
// Errors
// ---------------------------------------------------------------------------------------------
class Error : public std::runtime_error {
 public:
  Error(int status, const std::string& what)
      : std::runtime_error(what + ": " + mpf_status_string(status)), status_(status) {}
  int status() const { return status_; }

 private:
  int status_;
};

inline void check(mpf_status st, const char* what) {
  if (st == MPF_OK) return;
  if (st == MPF_ERR_NEG_JACOBIAN) throw std::runtime_error("Negative Jacobian in the element");
  throw Error(st, what);
}

// ---------------------------------------------------------------------------------------------
// Minimal dense types (column-major, like Eigen's default) for callers without Eigen.
// ---------------------------------------------------------------------------------------------
class MatrixXd {
 public:
  MatrixXd() = default;
  MatrixXd(int r, int c) { resize(r, c); }
  void resize(int r, int c) {
    r_ = r;
    c_ = c;
    a_.assign((size_t)r * c, 0.0);
  }
  int rows() const { return r_; }
  int cols() const { return c_; }
  double& operator()(int i, int j) { return a_[(size_t)j * r_ + i]; }
  double operator()(int i, int j) const { return a_[(size_t)j * r_ + i]; }
  void setZero() { std::fill(a_.begin(), a_.end(), 0.0); }

 private:
  int r_ = 0, c_ = 0;
  std::vector<double> a_;
};

class VectorXd {
 public:
  VectorXd() = default;
  explicit VectorXd(int n) { resize(n); }
  void resize(int n) { a_.assign((size_t)n, 0.0); }
  int size() const { return (int)a_.size(); }
  double& operator()(int i) { return a_[(size_t)i]; }
  double operator()(int i) const { return a_[(size_t)i]; }
  double& operator[](int i) { return a_[(size_t)i]; }
  double operator[](int i) const { return a_[(size_t)i]; }
  void setZero() { std::fill(a_.begin(), a_.end(), 0.0); }
  double* data() { return a_.data(); }
  const double* data() const { return a_.data(); }

 private:
  std::vector<double> a_;
};

// ---------------------------------------------------------------------------------------------
// Device memory (through the ABI's allocation calls)
// ---------------------------------------------------------------------------------------------
template <typename T>
class DeviceArray {
 public:
  DeviceArray() = default;
  explicit DeviceArray(size_t n) { resize(n); }
  DeviceArray(const DeviceArray&) = delete;
  DeviceArray& operator=(const DeviceArray&) = delete;
  DeviceArray(DeviceArray&& o) noexcept : p_(o.p_), n_(o.n_) {
    o.p_ = nullptr;
    o.n_ = 0;
  }
  ~DeviceArray() { release(); }
  void release() {
    if (p_) (void)mpf_free(p_);
    p_ = nullptr;
    n_ = 0;
  }
  void resize(size_t n) {
    if (n == n_) return;
    release();
    void* q = nullptr;
    check(mpf_malloc(&q, n * sizeof(T)), "mpf_malloc");
    p_ = static_cast<T*>(q);
    n_ = n;
  }
  void upload(const T* h, size_t n) {
    resize(n);
    check(mpf_memcpy_htod(p_, h, n * sizeof(T)), "mpf_memcpy_htod");
  }
  void upload(const std::vector<T>& h) { upload(h.data(), h.size()); }
  void download(T* h, size_t n) const {
    if (n > n_) throw Error(MPF_ERR_ARG, "DeviceArray::download");
    check(mpf_memcpy_dtoh(h, p_, n * sizeof(T)), "mpf_memcpy_dtoh");
  }
  std::vector<T> to_host() const {
    std::vector<T> h(n_);
    download(h.data(), n_);
    return h;
  }
  T* get() const { return p_; }
  size_t size() const { return n_; }

 private:
  T* p_ = nullptr;
  size_t n_ = 0;
};

// ---------------------------------------------------------------------------------------------
// Mesh and solution data on the device: GeomData::NodePosOrig (GeomDataLagrange), the element
// connectivity, and SolnData's dispCur / dispPrev / pres / presPrev (SolutionDataSolid), shared by
// the elements through a non-owning pointer as in the reference (ElementBase.h:54-57).
// ---------------------------------------------------------------------------------------------
class MeshData {
 public:
  MeshData(int npe, int nNode, int nElem, const double* X, const int32_t* conn) : npe(npe), nNode(nNode), nElem(nElem) {
    X_.upload(X, (size_t)nNode * 3);
    conn_.upload(conn, (size_t)nElem * npe);
    std::vector<double> z((size_t)nNode * 3, 0.0), zp((size_t)nNode, 0.0);
    disp_.upload(z);
    dispPrev_.upload(z);
    pres_.upload(zp);
  }
  void setDisplacement(const std::vector<double>& u) { disp_.upload(check_len(u, 3 * (size_t)nNode)); }
  void setPreviousDisplacement(const std::vector<double>& u) { dispPrev_.upload(check_len(u, 3 * (size_t)nNode)); }
  void setPressure(const std::vector<double>& p) { pres_.upload(check_len(p, (size_t)nNode)); }
  const double* X() const { return X_.get(); }
  const int32_t* conn() const { return conn_.get(); }
  const double* disp() const { return disp_.get(); }
  const double* dispPrev() const { return dispPrev_.get(); }
  const double* pres() const { return pres_.get(); }
  const int npe, nNode, nElem;

 private:
  static const std::vector<double>& check_len(const std::vector<double>& v, size_t n) {
    if (v.size() != n) throw Error(MPF_ERR_ARG, "MeshData: vector length");
    return v;
  }
  DeviceArray<double> X_, disp_, dispPrev_, pres_;
  DeviceArray<int32_t> conn_;
};

// ---------------------------------------------------------------------------------------------
// Material: Matl_NeoHookean (mu = data_deviatoric[0], Kinv, Utype; MaterialBase.cpp:91-115)
// ---------------------------------------------------------------------------------------------
class MatlNeoHookean {
 public:
  MatlNeoHookean(double mu, double Kinv, int Utype, bool mixed = false, int tangent = MPF_TANGENT_REFERENCE) {
    m_.mu = mu;
    m_.Kinv = Kinv;
    m_.Utype = Utype;
    m_.mixed = mixed ? 1 : 0;
    m_.tangent = tangent;
    m_.pad_ = 0;
    F_.resize(9);
    p_.resize(1);
    sig_.resize(9);
    C_.resize(81);
  }
  const mpf_matl& data() const { return m_; }
  double getKinv() const { return m_.Kinv; }

  // computeStressAndTangent(tangFlag, sss, Fn, F, pres, stre, Cmat) (MaterialBase.h:91):
  // stre [9] and Cmat [9][9] in the reference's column-major 9-index map
  // (utilitiesmaterialHyperelastic.cpp:191); Fn is not used by the Neo-Hookean law.
  template <class Mat, class Vec>
  int computeStressAndTangent(bool tangFlag, int sss, const Mat& Fn, const Mat& F, double& pres, Vec& stre, Mat& Cmat) {
    (void)tangFlag;
    (void)sss;
    (void)Fn;
    double hF[9], hsig[9], hC[81];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) hF[3 * i + j] = F(i, j);
    F_.upload(hF, 9);
    p_.upload(&pres, 1);
    check(mpf_neohooke_stress_tangent_batch(F_.get(), p_.get(), &m_, 1, sig_.get(), C_.get(), nullptr),
          "mpf_neohooke_stress_tangent_batch");
    sig_.download(hsig, 9);
    C_.download(hC, 81);
    stre.resize(9);
    for (int i = 0; i < 9; ++i) stre(i) = hsig[i];
    Cmat.resize(9, 9);
    for (int i = 0; i < 9; ++i)
      for (int j = 0; j < 9; ++j) Cmat(i, j) = hC[9 * i + j];
    return 0;
  }
  // the form with internal variables (MaterialBase.h:94); the hyperelastic law has none
  template <class Mat, class Vec, class IVar>
  int computeStressAndTangent(bool tangFlag, int sss, const Mat& Fn, const Mat& F, double& pres, Vec& stre, Mat& Cmat,
                              IVar& ivar, int gp, double dt) {
    (void)ivar;
    (void)gp;
    (void)dt;
    return computeStressAndTangent(tangFlag, sss, Fn, F, pres, stre, Cmat);
  }

 private:
  mpf_matl m_{};
  DeviceArray<double> F_, p_, sig_, C_;
};

// Element workspace shared by the elements of one mesh (the reference reuses the caller's Klocal /
// Flocal the same way, Elem_Solid_3D_Disp.cpp:230-236).
struct ElementWorkspace {
  DeviceArray<double> K, K2, K3, K4, F, F2;
  DeviceArray<int32_t> bad;
  ElementWorkspace() {
    K.resize(30 * 30);
    K2.resize(30 * 4);
    K3.resize(4 * 30);
    K4.resize(16);
    F.resize(30);
    F2.resize(4);
    bad.resize(1);
  }
};

namespace detail {
inline void clear_bad(ElementWorkspace& ws) {
  int32_t z = 0;
  ws.bad.upload(&z, 1);
}
inline void raise_if_bad(const ElementWorkspace& ws) {
  int32_t b = 0;
  ws.bad.download(&b, 1);
  if (b) throw std::runtime_error("Negative Jacobian in the element");
}
template <class Mat>
void to_matrix(const double* rowmajor, int r, int c, Mat& K) {
  K.resize(r, c);
  for (int i = 0; i < r; ++i)
    for (int j = 0; j < c; ++j) K(i, j) = rowmajor[(size_t)i * c + j];
}
template <class Vec>
void to_vector(const double* a, int n, Vec& F) {
  F.resize(n);
  for (int i = 0; i < n; ++i) F(i) = a[i];
}
}  // namespace detail

// ---------------------------------------------------------------------------------------------
// Displacement element (Hex8, npElem 8, 2x2x2 Gauss points; Tet10, npElem 10, 4 points)
// ---------------------------------------------------------------------------------------------
class ElemSolid3DDisp {
 public:
  // elemId: this element's row in MeshData's connectivity; nodeNums as in the reference.
  ElemSolid3DDisp(int elemId, const MeshData* geom, const MatlNeoHookean* matl, ElementWorkspace* ws)
      : elemId_(elemId), geom_(geom), matl_(matl), ws_(ws) {
    if (geom->npe != 8 && geom->npe != 10) throw Error(MPF_ERR_ARG, "ElemSolid3DDisp: npElem");
    nsize = 3 * geom->npe;
  }
  // forAssyVec[3a+d] = NodeDofArray[nodeNums[a]][d]; globalDOFnums[3a+d] = 3 nodeNums[a] + d
  // (femSolidmechanics2.cpp:248-280; ElementBase.cpp:46-57)
  void prepareElemData(const std::vector<int32_t>& nodes, const int32_t* NodeDofArray) {
    nodeNums.assign(nodes.begin(), nodes.end());
    forAssyVec.resize(nsize);
    globalDOFnums.resize(nsize);
    for (int a = 0; a < geom_->npe; ++a)
      for (int d = 0; d < 3; ++d) {
        forAssyVec[3 * a + d] = NodeDofArray[3 * (size_t)nodeNums[a] + d];
        globalDOFnums[3 * a + d] = 3 * nodeNums[a] + d;
      }
  }
  template <class Mat, class Vec>
  int calcStiffnessAndResidual(Mat& Klocal, Vec& Flocal) {
    detail::clear_bad(*ws_);
    const mpf_matl* m = &matl_->data();
    if (geom_->npe == 8)
      check(mpf_elem_hex8_nh_batch(geom_->X(), geom_->disp(), geom_->conn(), m, elemId_, 1, ws_->K.get(), ws_->F.get(),
                                   ws_->bad.get(), nullptr),
            "mpf_elem_hex8_nh_batch");
    else
      check(mpf_elem_tet10_nh_batch(geom_->X(), geom_->disp(), geom_->conn(), m, elemId_, 1, ws_->K.get(), ws_->F.get(),
                                    ws_->bad.get(), nullptr),
            "mpf_elem_tet10_nh_batch");
    detail::raise_if_bad(*ws_);
    double K[30 * 30], F[30];
    ws_->K.download(K, (size_t)nsize * nsize);
    ws_->F.download(F, (size_t)nsize);
    detail::to_matrix(K, nsize, nsize, Klocal);
    detail::to_vector(F, nsize, Flocal);
    return 0;
  }
  // ElementBase::applyDirichletBCs (ElementBase.cpp:567-594): dispApplied over all 3 nNode dofs
  template <class Mat, class Vec>
  int applyDirichletBCs(const Mat& Klocal, Vec& Flocal, const std::vector<double>& dispApplied) const {
    for (int ii = 0; ii < nsize; ++ii) {
      if (forAssyVec[ii] != -1) continue;
      const double fact = dispApplied[globalDOFnums[ii]];
      for (int jj = 0; jj < nsize; ++jj)
        if (forAssyVec[jj] != -1) Flocal(jj) -= Klocal(jj, ii) * fact;
    }
    return 0;
  }
  int nsize = 0;
  std::vector<int32_t> nodeNums, forAssyVec, globalDOFnums;

 private:
  int elemId_;
  const MeshData* geom_;
  const MatlNeoHookean* matl_;
  ElementWorkspace* ws_;
};

// ---------------------------------------------------------------------------------------------
// Tet10/P1 mixed element (Elem_Solid_3D_Mixed1: P1 pressure on the 4 corner nodes)
// ---------------------------------------------------------------------------------------------
class ElemSolid3DMixed1 {
 public:
  ElemSolid3DMixed1(int elemId, const MeshData* geom, const MatlNeoHookean* matl, ElementWorkspace* ws)
      : elemId_(elemId), geom_(geom), matl_(matl), ws_(ws) {
    if (geom->npe != 10) throw Error(MPF_ERR_ARG, "ElemSolid3DMixed1: npElem");
  }
  // forAssyVecPres[q] = ID_pres[nodeNums[q]] (femSolidmechanics1.cpp:274-387; Elem_Solid_3D_Mixed1.cpp:66-70)
  void prepareElemData(const std::vector<int32_t>& nodes, const int32_t* NodeDofArray, const int32_t* ID_pres) {
    nodeNums.assign(nodes.begin(), nodes.end());
    nodeNumsPres.assign(nodes.begin(), nodes.begin() + 4);
    forAssyVec.resize(30);
    globalDOFnums.resize(30);
    forAssyVecPres.resize(4);
    for (int a = 0; a < 10; ++a)
      for (int d = 0; d < 3; ++d) {
        forAssyVec[3 * a + d] = NodeDofArray[3 * (size_t)nodeNums[a] + d];
        globalDOFnums[3 * a + d] = 3 * nodeNums[a] + d;
      }
    for (int q = 0; q < 4; ++q) forAssyVecPres[q] = ID_pres[nodeNums[q]];
  }
  template <class Mat, class Vec>
  int calcStiffnessAndResidualMixed(Mat& Kuu, Mat& Kup, Mat& Kpu, Mat& Kpp, Vec& FlocalU, Vec& FlocalP) {
    detail::clear_bad(*ws_);
    check(mpf_elem_tet10p1_nh_batch(geom_->X(), geom_->disp(), geom_->dispPrev(), geom_->pres(), geom_->conn(),
                                    &matl_->data(), elemId_, 1, ws_->K.get(), ws_->K2.get(), ws_->K3.get(),
                                    ws_->K4.get(), ws_->F.get(), ws_->F2.get(), ws_->bad.get(), nullptr),
          "mpf_elem_tet10p1_nh_batch");
    detail::raise_if_bad(*ws_);
    double a[900], b[120], c[120], d[16], f[30], g[4];
    ws_->K.download(a, 900);
    ws_->K2.download(b, 120);
    ws_->K3.download(c, 120);
    ws_->K4.download(d, 16);
    ws_->F.download(f, 30);
    ws_->F2.download(g, 4);
    detail::to_matrix(a, 30, 30, Kuu);
    detail::to_matrix(b, 30, 4, Kup);
    detail::to_matrix(c, 4, 30, Kpu);
    detail::to_matrix(d, 4, 4, Kpp);
    detail::to_vector(f, 30, FlocalU);
    detail::to_vector(g, 4, FlocalP);
    return 0;
  }
  // ElementBase::applyDirichletBCs2field (ElementBase.cpp:717-782)
  template <class Mat, class Vec>
  int applyDirichletBCs2field(const Mat& K11, const Mat& K12, const Mat& K21, const Mat& K22, Vec& F1, Vec& F2,
                              const std::vector<double>& dispApplied, const std::vector<double>& presApplied) const {
    for (int ii = 0; ii < 30; ++ii) {
      if (forAssyVec[ii] != -1) continue;
      const double fact = dispApplied[globalDOFnums[ii]];
      for (int jj = 0; jj < 30; ++jj)
        if (forAssyVec[jj] != -1) F1(jj) -= K11(jj, ii) * fact;
      for (int jj = 0; jj < 4; ++jj)
        if (forAssyVecPres[jj] != -1) F2(jj) -= K21(jj, ii) * fact;
    }
    for (int ii = 0; ii < 4; ++ii) {
      if (forAssyVecPres[ii] != -1) continue;
      const double fact = presApplied[nodeNums[ii]];
      for (int jj = 0; jj < 30; ++jj)
        if (forAssyVec[jj] != -1) F1(jj) -= K12(jj, ii) * fact;
      for (int jj = 0; jj < 4; ++jj)
        if (forAssyVecPres[jj] != -1) F2(jj) -= K22(jj, ii) * fact;
    }
    return 0;
  }
  std::vector<int32_t> nodeNums, nodeNumsPres, forAssyVec, forAssyVecPres, globalDOFnums;

 private:
  int elemId_;
  const MeshData* geom_;
  const MatlNeoHookean* matl_;
  ElementWorkspace* ws_;
};

// ---------------------------------------------------------------------------------------------
// Solver: SolverPetsc's assemble / solve calls over the device CSR system
// ---------------------------------------------------------------------------------------------
enum { SOLVER_EMPTY, PATTERN_OK, INIT_OK, ASSEMBLY_OK, FACTORISE_OK };  // SolverEigen.h:11

class SolverGpu {
 public:
  SolverGpu() {
    opts.ksp = MPF_KSP_CG;  // the reference's code default is CG + ILU (SolverPetsc.cpp:117-140)
    opts.pc = MPF_PC_JACOBI;
    opts.restart = 50;
    opts.maxit = 10000;
    opts.rtol = 1e-10;
  }
  SolverGpu(const SolverGpu&) = delete;
  SolverGpu& operator=(const SolverGpu&) = delete;
  ~SolverGpu() { free(); }

  // SolverPetsc::initialise with the pattern of prepareMatrixPattern: n = dispDOF + presDOF rows,
  // sorted unique columns per row; NodeDofArray (optional) enables the 3x3 block-Jacobi PC.
  int initialise(int dispDOF_, int presDOF_, const std::vector<int32_t>& rowptr, const std::vector<int32_t>& colidx,
                 int nNode = 0, const int32_t* NodeDofArray = nullptr) {
    free();
    dispDOF = dispDOF_;
    presDOF = presDOF_;
    nRow = nCol = dispDOF + presDOF;
    if ((int)rowptr.size() != nRow + 1) throw Error(MPF_ERR_ARG, "SolverGpu::initialise: rowptr size");
    nnz = rowptr.back();
    check(mpf_system_create(&sys_, dispDOF, presDOF, rowptr.data(), colidx.data(), NodeDofArray ? nNode : 0,
                            NodeDofArray),
          "mpf_system_create");
    currentStatus = INIT_OK;
    return 0;
  }
  int free() {
    groups_.clear();
    dmaps_.clear();
    if (sys_) (void)mpf_problem_destroy(sys_);
    sys_ = nullptr;
    currentStatus = SOLVER_EMPTY;
    return 0;
  }
  // SolverPetsc::zeroMtx
  int zeroMtx() {
    for (Group& g : groups_) g.clear_pending();
    check(mpf_system_zero(sys_), "mpf_system_zero");
    fext_pending_.clear();
    return 0;
  }
  // SolverPetsc::assembleMatrixAndVector(start1, start2, row, col, K, F) (SolverPetsc.cpp:357-378):
  // rows / columns equal to -1 are skipped (the solid driver calls it with start1 = start2 = 0 and
  // forAssyVec, whose -1 entries PETSc drops as negative indices).  Contributions are queued and
  // added on the device, in call order, before the next solve or read-back.
  template <class Mat, class Vec>
  int assembleMatrixAndVector(int start1, int start2, const std::vector<int32_t>& row, const std::vector<int32_t>& col,
                              const Mat& Klocal, const Vec& Flocal) {
    const int nr = (int)row.size(), nc = (int)col.size();
    Group& g = group(nr, nc);
    for (int i = 0; i < nr; ++i) g.rows.push_back(row[i] < 0 ? -1 : row[i] + start1);
    for (int j = 0; j < nc; ++j) g.cols.push_back(col[j] < 0 ? -1 : col[j] + start2);
    for (int i = 0; i < nr; ++i) {
      g.F.push_back(Flocal(i));
      for (int j = 0; j < nc; ++j) g.K.push_back(Klocal(i, j));
    }
    ++g.count;
    currentStatus = ASSEMBLY_OK;
    return 0;
  }
  // SolverPetsc::assembleMatrixAndVector2field (SolverPetsc.cpp:434-502): the u-p blocks as one
  // (size1 + size2)^2 element matrix whose pressure indices carry var2_offset.
  template <class Mat, class Vec>
  int assembleMatrixAndVector2field(int var2_offset, const Mat& K11, const Mat& K12, const Mat& K21, const Mat& K22,
                                    const Vec& F1, const Vec& F2, const std::vector<int32_t>& forAssyVar1,
                                    const std::vector<int32_t>& forAssyVar2) {
    const int s1 = (int)forAssyVar1.size(), s2 = (int)forAssyVar2.size(), s = s1 + s2;
    Group& g = group(s, s);
    for (int i = 0; i < s1; ++i) g.rows.push_back(forAssyVar1[i]);
    for (int i = 0; i < s2; ++i) g.rows.push_back(forAssyVar2[i] == -1 ? -1 : forAssyVar2[i] + var2_offset);
    g.cols.insert(g.cols.end(), g.rows.end() - s, g.rows.end());
    for (int i = 0; i < s; ++i) {
      g.F.push_back(i < s1 ? F1(i) : F2(i - s1));
      for (int j = 0; j < s; ++j) {
        double v;
        if (i < s1) v = j < s1 ? K11(i, j) : K12(i, j - s1);
        else v = j < s1 ? K21(i - s1, j) : K22(i - s1, j - s1);
        g.K.push_back(v);
      }
    }
    ++g.count;
    currentStatus = ASSEMBLY_OK;
    return 0;
  }
  // Batched form on device data: element matrices K [count][nrow][ncol] and F [count][nrow] already
  // on the device (e.g. from mpf_elem_*_batch), dof lists rows / cols on the device.  The map is built
  // once per distinct (rows, cols) pointer pair and size; pass rebuild = true when their contents change.
  int assembleElementsDevice(int count, int nrow, int ncol, const int32_t* rows_dev, const int32_t* cols_dev,
                             const double* K_dev, const double* F_dev, bool rebuild = false) {
    flush();
    DeviceMap* dm = nullptr;
    for (DeviceMap& m : dmaps_)
      if (m.rows == rows_dev && m.cols == cols_dev && m.count == count && m.nrow == nrow && m.ncol == ncol) dm = &m;
    if (!dm || rebuild) {
      int32_t id = -1;
      check(mpf_system_map_elements(sys_, count, nrow, ncol, rows_dev, cols_dev, 0, 0, &id), "mpf_system_map_elements");
      if (!dm) {
        dmaps_.push_back({rows_dev, cols_dev, count, nrow, ncol, id});
        dm = &dmaps_.back();
      } else {
        dm->id = id;
      }
    }
    check(mpf_system_add_elements(sys_, dm->id, K_dev, F_dev), "mpf_system_add_elements");
    currentStatus = ASSEMBLY_OK;
    return 0;
  }
  // Fext (SolverPetsc's public member) added to the residual before the solve (femSolidmechanics2.cpp:785-789)
  void addToRhs(const std::vector<double>& f) { fext_pending_ = f; }

  // SolverPetsc::factoriseAndSolve (SolverPetsc.cpp:318-330): -1 unless ASSEMBLY_OK; 0 when the
  // Krylov method reaches opts.rtol, else the mpf_status of the failure.
  int factoriseAndSolve() {
    if (currentStatus != ASSEMBLY_OK) return -1;
    flush();
    const mpf_status st = mpf_solve(sys_, &opts, &report);
    if (st != MPF_OK && st != MPF_ERR_NOT_CONVERGED) check(st, "mpf_solve");
    currentStatus = FACTORISE_OK;
    return st == MPF_OK ? 0 : (int)st;
  }
  // solnVec / rhsVec / matrix values on the host
  std::vector<double> solnVec() const {
    std::vector<double> x((size_t)nRow);
    check(mpf_get_solution(sys_, x.data()), "mpf_get_solution");
    return x;
  }
  std::vector<double> rhsVec() {
    flush();
    std::vector<double> r((size_t)nRow);
    check(mpf_get_system(sys_, nullptr, r.data()), "mpf_get_system");
    return r;
  }
  std::vector<double> matrixValues() {
    flush();
    std::vector<double> v((size_t)nnz);
    check(mpf_get_system(sys_, v.data(), nullptr), "mpf_get_system");
    return v;
  }
  // queued contributions to the device (called by the solve and read-backs)
  void flush() {
    for (Group& g : groups_) g.flush(sys_);
    if (!fext_pending_.empty()) {
      if ((int)fext_pending_.size() != nRow) throw Error(MPF_ERR_ARG, "SolverGpu: Fext length");
      DeviceArray<double> f;
      f.upload(fext_pending_);
      check(mpf_system_add_rhs(sys_, f.get()), "mpf_system_add_rhs");
      fext_pending_.clear();
    }
  }
  mpf_problem* handle() const { return sys_; }

  mpf_krylov_opts opts{};
  mpf_krylov_report report{};
  int currentStatus = SOLVER_EMPTY;
  int nRow = 0, nCol = 0, dispDOF = 0, presDOF = 0;
  int64_t nnz = 0;

 private:
  // queued element contributions of one size; the device map is kept while the dof lists repeat
  // (the same elements every Newton iteration), and rebuilt when they change
  struct Group {
    int nrow = 0, ncol = 0, count = 0;
    std::vector<int32_t> rows, cols, map_rows, map_cols;
    std::vector<double> K, F;
    int32_t map_id = -1;
    DeviceArray<int32_t> drows, dcols;
    DeviceArray<double> dK, dF;
    void clear_pending() {
      rows.clear();
      cols.clear();
      K.clear();
      F.clear();
      count = 0;
    }
    void flush(mpf_problem* sys) {
      if (count == 0) return;
      if (map_id < 0 || rows != map_rows || cols != map_cols) {
        drows.upload(rows);
        dcols.upload(cols);
        check(mpf_system_map_elements(sys, count, nrow, ncol, drows.get(), dcols.get(), 0, 0, &map_id),
              "mpf_system_map_elements");
        map_rows = rows;
        map_cols = cols;
      }
      dK.upload(K);
      dF.upload(F);
      check(mpf_system_add_elements(sys, map_id, dK.get(), dF.get()), "mpf_system_add_elements");
      clear_pending();
    }
  };
  struct DeviceMap {
    const int32_t* rows;
    const int32_t* cols;
    int count, nrow, ncol;
    int32_t id;
  };
  Group& group(int nr, int nc) {
    for (Group& g : groups_)
      if (g.nrow == nr && g.ncol == nc) return g;
    groups_.emplace_back();
    groups_.back().nrow = nr;
    groups_.back().ncol = nc;
    return groups_.back();
  }
  mpf_problem* sys_ = nullptr;
  std::vector<Group> groups_;
  std::vector<DeviceMap> dmaps_;
  std::vector<double> fext_pending_;
};

}  // namespace mpf_host

#endif  // MPF_HOST_HPP
