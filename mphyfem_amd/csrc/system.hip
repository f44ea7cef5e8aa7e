// system.hip -- a linear system over a caller-supplied CSR pattern, filled from caller-computed
// element matrices: the SolverPetsc-shaped operator (SolverPetsc.h:16-96) behind the C ABI.
//
//   mpf_system_create        SolverPetsc::initialise + the zero-valued pattern insertion
//                            (SolverPetsc.cpp:24-114; femSolidmechanics2.cpp:435-446)
//   mpf_system_zero          SolverPetsc::zeroMtx (SolverPetsc.cpp:145-163)
//   mpf_system_map_elements  the (row, col) -> CSR slot map of a batch of element dof lists
//   mpf_system_add_elements  SolverPetsc::assembleMatrixAndVector / ...2field for the batch
//                            (SolverPetsc.cpp:357-378, 434-502)
//   mpf_solve / mpf_get_solution / mpf_get_system work on the system as on a mesh problem.
//
// Summation order.  PETSc's ADD_VALUES adds an element's entries into an AIJ value one element at
// a time, so every CSR value is ((v + c_1) + c_2) + ... over the contributing elements in element
// order.  The device keeps exactly that order without atomics: the map sorts every contribution by
// its CSR slot with a stable radix sort (ties keep the contribution index, which grows with the
// element index), and the add kernel gives each slot one lane that walks its contributions in that
// order.  The result equals a sequential host loop over the elements bit for bit.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <climits>
#include <cstring>

#include "mpf_internal.hpp"

namespace mpf {

struct ElemMap {
  int32_t count = 0, nrow = 0, ncol = 0;
  int64_t nk = 0, nf = 0;            // matrix / vector contributions
  DevBuf<int32_t> kperm, koff;       // contributions sorted by CSR slot; slot s owns kperm[koff[s]..koff[s+1])
  DevBuf<int32_t> fperm, foff;       // the same for the right-hand side rows
};

struct SystemWork {
  std::vector<ElemMap*> maps;
  ~SystemWork() {
    for (ElemMap* m : maps) delete m;
  }
};

void system_free(mpf_problem* p) {
  delete p->sys;
  p->sys = nullptr;
}

namespace {

inline unsigned gsz(int64_t n, unsigned b = 256) {
  int64_t g = (n + b - 1) / b;
  return (unsigned)(g < 1 ? 1 : g);
}

__global__ void k_sys_fill(double* x, int64_t n, double v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = v;
}

// Key of every matrix contribution k = (e*nrow + i)*ncol + j: its CSR slot, or nnz when the row or
// the column is -1 (skipped, as assembleMatrixAndVector2field's "!= -1" tests).  A pair whose
// column is not in the row's pattern sets *missing.
__global__ void k_sys_kkeys(int64_t nk, int nrow, int ncol, const int32_t* __restrict__ rows,
                            const int32_t* __restrict__ cols, int32_t r_off, int32_t c_off,
                            const int32_t* __restrict__ rowptr, const int32_t* __restrict__ colidx, int32_t n,
                            int32_t nnz, int32_t* __restrict__ key, int32_t* __restrict__ idx,
                            int32_t* __restrict__ missing) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nk) return;
  const int64_t per = (int64_t)nrow * ncol;
  const int64_t e = k / per;
  const int rem = (int)(k - e * per);
  const int i = rem / ncol, j = rem - i * ncol;
  int32_t r = rows[e * nrow + i], c = cols[e * ncol + j];
  int32_t slot = nnz;
  if (r >= 0 && c >= 0) {
    r += r_off;
    c += c_off;
    if (r >= n || c >= n) {
      *missing = 1;
    } else {
      int lo = rowptr[r], hi = rowptr[r + 1];
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (colidx[mid] < c) lo = mid + 1;
        else hi = mid;
      }
      if (lo < rowptr[r + 1] && colidx[lo] == c) slot = lo;
      else *missing = 1;
    }
  }
  key[k] = slot;
  idx[k] = (int32_t)k;
}

// Key of every vector contribution k = e*nrow + i: its row, or n when skipped.
__global__ void k_sys_fkeys(int64_t nf, const int32_t* __restrict__ rows, int32_t r_off, int32_t n,
                            int32_t* __restrict__ key, int32_t* __restrict__ idx, int32_t* __restrict__ missing) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nf) return;
  int32_t r = rows[k];
  int32_t out = n;
  if (r >= 0) {
    r += r_off;
    if (r < n) out = r;
    else *missing = 1;
  }
  key[k] = out;
  idx[k] = (int32_t)k;
}

// off[s] = first position of key s in the sorted keys (lower bound), s = 0..m.
__global__ void k_sys_offsets(const int32_t* __restrict__ skey, int64_t total, int32_t m, int32_t* __restrict__ off) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s > m) return;
  int64_t lo = 0, hi = total;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (skey[mid] < (int32_t)s) lo = mid + 1;
    else hi = mid;
  }
  off[s] = (int32_t)lo;
}

// y[s] += sum of src[perm[k]] over the slot's contributions, in contribution order.
__global__ void k_sys_add(int32_t m, const int32_t* __restrict__ off, const int32_t* __restrict__ perm,
                          const double* __restrict__ src, double* __restrict__ y) {
  const int32_t s = (int32_t)((int64_t)blockIdx.x * blockDim.x + threadIdx.x);
  if (s >= m) return;
  const int32_t a = off[s], z = off[s + 1];
  if (a == z) return;
  double acc = y[s];
  for (int32_t k = a; k < z; ++k) acc += src[perm[k]];
  y[s] = acc;
}

__global__ void k_sys_axpy(int64_t n, const double* __restrict__ f, double* __restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] += f[i];
}

// ElementBase::applyDirichletBCs (ElementBase.cpp:567-594) for a batch: one lane per (element, row
// jj) with forAssy[jj] != -1 subtracts K(jj, ii) * dispApplied[gdof[ii]] over the prescribed ii in
// increasing order, the reference's order for that entry.
__global__ void k_elem_lift(int64_t total, int nsize, const int32_t* __restrict__ forAssy,
                            const int32_t* __restrict__ gdof, const double* __restrict__ dispApplied,
                            const double* __restrict__ K, double* __restrict__ F) {
#pragma clang fp contract(off)  // multiply, then subtract: the reference's host arithmetic (no FMA)
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  const int64_t e = t / nsize;
  const int jj = (int)(t - e * nsize);
  const int32_t* fa = forAssy + e * nsize;
  if (fa[jj] == -1) return;
  const int32_t* gd = gdof + e * nsize;
  const double* Kr = K + (e * nsize + jj) * nsize;
  double f = F[t];
  for (int ii = 0; ii < nsize; ++ii)
    if (fa[ii] == -1) f -= Kr[ii] * dispApplied[gd[ii]];
  F[t] = f;
}

int bits_for(int64_t v) {
  int b = 1;
  while (b < 31 && ((int64_t)1 << b) <= v) ++b;
  return b;
}

// Sort (key, idx) pairs by key, stably, into (skey, perm); offsets per key 0..m.
mpf_status sort_and_offset(int64_t total, int32_t m, DevBuf<int32_t>& key, DevBuf<int32_t>& idx, DevBuf<int32_t>& perm,
                           DevBuf<int32_t>& off, hipStream_t s) {
  DevBuf<int32_t> skey;
  MPF_TRY(skey.alloc(total > 0 ? total : 1));
  MPF_TRY(perm.alloc(total > 0 ? total : 1));
  MPF_TRY(off.alloc((size_t)m + 1));
  if (total > 0) {
    size_t tmp = 0;
    const int nb = bits_for(m);
    MPF_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, key.ptr, skey.ptr, idx.ptr, perm.ptr, (int)total, 0, nb, s));
    DevBuf<uint8_t> t;
    MPF_TRY(t.alloc(tmp > 0 ? tmp : 1));
    MPF_HIP(hipcub::DeviceRadixSort::SortPairs(t.ptr, tmp, key.ptr, skey.ptr, idx.ptr, perm.ptr, (int)total, 0, nb, s));
  }
  hipLaunchKernelGGL(k_sys_offsets, dim3(gsz((int64_t)m + 1)), dim3(256), 0, s, skey.ptr, total, m, off.ptr);
  MPF_HIP(hipGetLastError());
  MPF_HIP(hipStreamSynchronize(s));  // the temporaries are freed on return
  return MPF_OK;
}

bool is_system(const mpf_problem* p) { return p && p->etype == 0 && p->sys; }

}  // namespace
}  // namespace mpf

using namespace mpf;

extern "C" {

mpf_status mpf_system_create(mpf_problem** out, int32_t n_disp, int32_t n_pres, const int32_t* rowptr,
                             const int32_t* colidx, int32_t nNode, const int32_t* nda) {
  if (!out || !rowptr || n_disp < 0 || n_pres < 0 || (int64_t)n_disp + n_pres <= 0) return MPF_ERR_ARG;
  if ((int64_t)n_disp + n_pres > INT32_MAX - 1) return MPF_ERR_ARG;
  const int32_t n = n_disp + n_pres;
  if (rowptr[0] != 0) return MPF_ERR_ARG;
  for (int32_t r = 0; r < n; ++r) {
    if (rowptr[r + 1] < rowptr[r]) return MPF_ERR_ARG;
    for (int32_t k = rowptr[r]; k < rowptr[r + 1]; ++k) {
      if (!colidx || colidx[k] < 0 || colidx[k] >= n) return MPF_ERR_ARG;
      if (k > rowptr[r] && colidx[k] <= colidx[k - 1]) return MPF_ERR_ARG;  // sorted, unique
    }
  }
  const int64_t nnz = rowptr[n];
  // NodeDofArray (optional): a node's free displacement dofs are consecutive, in dof order, and
  // every displacement row belongs to exactly one node -- the node blocks of the 3x3 block-Jacobi
  // preconditioner
  if (nda) {
    if (nNode <= 0) return MPF_ERR_ARG;
    std::vector<int8_t> seen(n_disp, 0);
    for (int32_t v = 0; v < nNode; ++v) {
      int32_t prev = -1;
      for (int d = 0; d < 3; ++d) {
        const int32_t q = nda[3 * (int64_t)v + d];
        if (q < -1 || q >= n_disp) return MPF_ERR_ARG;
        if (q < 0) continue;
        if ((prev >= 0 && q != prev + 1) || seen[q]) return MPF_ERR_ARG;
        seen[q] = 1;
        prev = q;
      }
    }
    for (int32_t q = 0; q < n_disp; ++q)
      if (!seen[q]) return MPF_ERR_ARG;
  } else if (nNode != 0) {
    return MPF_ERR_ARG;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return MPF_ERR_NO_DEVICE;
  auto* p = new mpf_problem();
  p->uid = next_problem_uid();
  p->etype = 0;
  p->sys = new SystemWork();
  if (hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess) {
    delete p->sys;
    delete p;
    return MPF_ERR_HIP;
  }
  PatternBufs& P = p->pat;
  P.dispDOF = n_disp;
  P.presDOF = n_pres;
  P.nnz = nnz;
  p->nNode = nda ? nNode : 0;
  mpf_status st = P.rowptr.upload(rowptr, (size_t)n + 1, p->stream);
  if (st == MPF_OK) st = P.colidx.upload(colidx, (size_t)(nnz > 0 ? nnz : 0), p->stream);
  if (st == MPF_OK && nda) st = P.nda.upload(nda, 3 * (size_t)nNode, p->stream);
  if (st == MPF_OK) st = p->vals.alloc(nnz > 0 ? (size_t)nnz : 1);
  if (st == MPF_OK) st = p->rhs.alloc((size_t)n);
  if (st == MPF_OK) st = p->du.alloc((size_t)n);
  if (st == MPF_OK) st = p->scratch.alloc((size_t)kMaxPartials + 1);
  if (st == MPF_OK && hipStreamSynchronize(p->stream) != hipSuccess) st = MPF_ERR_HIP;
  if (st != MPF_OK) {
    mpf_problem_destroy(p);
    return st;
  }
  p->prepared = true;
  *out = p;
  return mpf_system_zero(p);
}

mpf_status mpf_system_zero(mpf_problem* p) {
  if (!is_system(p)) return MPF_ERR_ARG;
  const int64_t n = (int64_t)p->pat.dispDOF + p->pat.presDOF;
  hipLaunchKernelGGL(k_sys_fill, dim3(gsz(p->pat.nnz)), dim3(256), 0, p->stream, p->vals.ptr, p->pat.nnz, 0.0);
  hipLaunchKernelGGL(k_sys_fill, dim3(gsz(n)), dim3(256), 0, p->stream, p->rhs.ptr, n, 0.0);
  MPF_HIP(hipGetLastError());
  MPF_HIP(hipStreamSynchronize(p->stream));
  p->assembled = false;
  p->solved = false;
  return MPF_OK;
}

mpf_status mpf_system_map_elements(mpf_problem* p, int32_t count, int32_t nrow, int32_t ncol, const int32_t* rows,
                                   const int32_t* cols, int32_t r_off, int32_t c_off, int32_t* map_id) {
  if (!is_system(p) || !map_id || count < 0 || nrow <= 0 || ncol <= 0 || r_off < 0 || c_off < 0) return MPF_ERR_ARG;
  if (count > 0 && (!rows || !cols)) return MPF_ERR_ARG;
  const int64_t nk = (int64_t)count * nrow * ncol, nf = (int64_t)count * nrow;
  if (nk > INT32_MAX) return MPF_ERR_ARG;  // split the batch: contribution indices are int32
  const int32_t n = p->pat.dispDOF + p->pat.presDOF;
  const int32_t nnz = (int32_t)p->pat.nnz;
  hipStream_t s = p->stream;
  MPF_HIP(hipDeviceSynchronize());  // the caller's dof lists may come from work on any stream
  auto* M = new ElemMap();
  M->count = count;
  M->nrow = nrow;
  M->ncol = ncol;
  M->nk = nk;
  M->nf = nf;
  DevBuf<int32_t> key, idx, missing;
  int32_t hmiss = 0;
  mpf_status st = key.alloc(nk > 0 ? nk : 1);
  if (st == MPF_OK) st = idx.alloc(nk > 0 ? nk : 1);
  if (st == MPF_OK) st = missing.upload(&hmiss, 1, s);
  if (st == MPF_OK && nk > 0)
    hipLaunchKernelGGL(k_sys_kkeys, dim3(gsz(nk)), dim3(256), 0, s, nk, nrow, ncol, rows, cols, r_off, c_off,
                       p->pat.rowptr.ptr, p->pat.colidx.ptr, n, nnz, key.ptr, idx.ptr, missing.ptr);
  if (st == MPF_OK) st = sort_and_offset(nk, nnz, key, idx, M->kperm, M->koff, s);
  if (st == MPF_OK && nf > 0)
    hipLaunchKernelGGL(k_sys_fkeys, dim3(gsz(nf)), dim3(256), 0, s, nf, rows, r_off, n, key.ptr, idx.ptr, missing.ptr);
  if (st == MPF_OK) st = sort_and_offset(nf, n, key, idx, M->fperm, M->foff, s);
  if (st == MPF_OK && (hipMemcpy(&hmiss, missing.ptr, sizeof(int32_t), hipMemcpyDeviceToHost) != hipSuccess))
    st = MPF_ERR_HIP;
  if (st == MPF_OK && hmiss) st = MPF_ERR_ARG;  // an entry outside the pattern (SolverPetsc would insert it)
  if (st != MPF_OK) {
    delete M;
    return st;
  }
  p->sys->maps.push_back(M);
  *map_id = (int32_t)p->sys->maps.size() - 1;
  return MPF_OK;
}

mpf_status mpf_system_add_elements(mpf_problem* p, int32_t map_id, const double* K, const double* F) {
  if (!is_system(p) || map_id < 0 || map_id >= (int32_t)p->sys->maps.size()) return MPF_ERR_ARG;
  const ElemMap& M = *p->sys->maps[map_id];
  hipStream_t s = p->stream;
  MPF_HIP(hipDeviceSynchronize());  // element matrices from the caller's streams are complete
  const int32_t n = p->pat.dispDOF + p->pat.presDOF;
  const int32_t nnz = (int32_t)p->pat.nnz;
  if (K && M.nk > 0)
    hipLaunchKernelGGL(k_sys_add, dim3(gsz(nnz)), dim3(256), 0, s, nnz, M.koff.ptr, M.kperm.ptr, K, p->vals.ptr);
  if (F && M.nf > 0)
    hipLaunchKernelGGL(k_sys_add, dim3(gsz(n)), dim3(256), 0, s, n, M.foff.ptr, M.fperm.ptr, F, p->rhs.ptr);
  MPF_HIP(hipGetLastError());
  MPF_HIP(hipStreamSynchronize(s));
  p->assembled = true;  // currentStatus = ASSEMBLY_OK (femSolidmechanics2.cpp:1115)
  p->solved = false;
  return MPF_OK;
}

mpf_status mpf_system_add_rhs(mpf_problem* p, const double* f) {
  // Fext added to the residual (femSolidmechanics2.cpp:785-789): rhs[i] += f[i] over all rows
  if (!is_system(p) || !f) return MPF_ERR_ARG;
  const int64_t n = (int64_t)p->pat.dispDOF + p->pat.presDOF;
  MPF_HIP(hipDeviceSynchronize());
  hipLaunchKernelGGL(k_sys_axpy, dim3(gsz(n)), dim3(256), 0, p->stream, n, f, p->rhs.ptr);
  MPF_HIP(hipGetLastError());
  MPF_HIP(hipStreamSynchronize(p->stream));
  return MPF_OK;
}

mpf_status mpf_elem_apply_dirichlet_batch(int32_t count, int32_t nsize, const int32_t* forAssy, const int32_t* gdof,
                                          const double* dispApplied, const double* K, double* F, mpf_stream_t stream) {
  if (count < 0 || nsize <= 0 || nsize > 1024 || (count > 0 && (!forAssy || !gdof || !dispApplied || !K || !F)))
    return MPF_ERR_ARG;
  const int64_t total = (int64_t)count * nsize;
  if (total > 0)
    hipLaunchKernelGGL(k_elem_lift, dim3(gsz(total)), dim3(256), 0, (hipStream_t)stream, total, nsize, forAssy, gdof,
                       dispApplied, K, F);
  MPF_HIP(hipGetLastError());
  return MPF_OK;
}

// Device memory for callers without the HIP headers (the C++ host layer's staging buffers).
mpf_status mpf_malloc(void** ptr, size_t bytes) {
  if (!ptr) return MPF_ERR_ARG;
  *ptr = nullptr;
  if (bytes == 0) return MPF_OK;
  return hipMalloc(ptr, bytes) == hipSuccess ? MPF_OK : MPF_ERR_HIP;
}

mpf_status mpf_free(void* ptr) { return hipFree(ptr) == hipSuccess ? MPF_OK : MPF_ERR_HIP; }

mpf_status mpf_memcpy_htod(void* dst, const void* src, size_t bytes) {
  if (bytes && (!dst || !src)) return MPF_ERR_ARG;
  return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess ? MPF_OK : MPF_ERR_HIP;
}

mpf_status mpf_memcpy_dtoh(void* dst, const void* src, size_t bytes) {
  if (bytes && (!dst || !src)) return MPF_ERR_ARG;
  return hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost) == hipSuccess ? MPF_OK : MPF_ERR_HIP;
}

}  // extern "C"
