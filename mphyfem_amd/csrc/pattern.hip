// pattern.hip -- device construction of the DOF numbering, the CSR pattern, the element
// colouring and the per-element CSR slot map.
//
// Reference: femSolidmechanics::prepareMatrixPattern (femSolidmechanics2.cpp:121-478),
// setSpecifiedDOFs_Displacement (femSolidmechanics1.cpp:1479-1543) and
// prepareDataForMixedElements (femSolidmechanics1.cpp:274-387).  The reference pushes every
// (row, col) pair of every element into per-row vectors and sorts/uniques them (util.h:41-45).
// Here the same set is built from node neighbourhoods: all free u-rows of a node n (and its
// pressure row, for the mixed element) hold exactly the columns
//     [free dofs of m, m in NB(n) ascending]  ++  [dispDOF + ID_pres[m], m in NB(n), m a pressure node]
// where NB(n) is the sorted set of nodes sharing an element with n.  Because free dofs are
// numbered node-major and pressure dofs follow in node order, this list is already sorted,
// and equals the reference's sorted unique row (for Tet10/P1, every vertex node of an element
// containing n is a pressure node of that element).
//
// Colouring (no reference counterpart): Jones-Plassmann rounds with priority
// mpf_color_priority(e) and first-fit colours; the result equals sequential greedy first-fit
// in decreasing priority order, so it is deterministic and identical to the oracle's.

#include <hipcub/hipcub.hpp>

#include "mpf_internal.hpp"

namespace mpf {

namespace {

constexpr int kNBCap = 512;      // candidate capacity per node (adjacent elements x npe), first pass
constexpr int kNBCapBig = 8192;  // second pass, for the nodes above kNBCap (LDS 32 KiB)

__device__ __forceinline__ uint64_t color_priority(int64_t e) {
  uint64_t z = (uint64_t)e + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__global__ void k_mark_bc(const int32_t* __restrict__ bc_dof, int32_t nbc, int32_t* __restrict__ type) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nbc) type[bc_dof[i]] = 1;
}

__global__ void k_free_flags(const int32_t* __restrict__ type, int64_t n, int32_t* __restrict__ flag) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) flag[i] = type[i] ? 0 : 1;
}

// NodeDofArray from the exclusive scan of the free flags; assyForSoln scatter.
__global__ void k_number(const int32_t* __restrict__ flag, const int32_t* __restrict__ scan, int64_t n,
                         int32_t* __restrict__ nda, int32_t* __restrict__ assy) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (flag[i]) {
    nda[i] = scan[i];
    assy[scan[i]] = (int32_t)i;
  } else {
    nda[i] = -1;
  }
}

__global__ void k_mark_pres(const int32_t* __restrict__ conn, int32_t nElem, int32_t npe, int32_t* __restrict__ isP) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)nElem * 4) return;
  const int64_t e = t / 4, q = t % 4;
  isP[conn[e * npe + q]] = 1;
}

__global__ void k_number_pres(const int32_t* __restrict__ isP, const int32_t* __restrict__ scan, int32_t nNode,
                              int32_t* __restrict__ idp, int32_t* __restrict__ assyP) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nNode) return;
  if (isP[i]) {
    idp[i] = scan[i];
    assyP[scan[i]] = i;
  } else {
    idp[i] = -1;
  }
}

__global__ void k_node_count(const int32_t* __restrict__ conn, int64_t nent, int32_t* __restrict__ cnt) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < nent) atomicAdd(&cnt[conn[t]], 1);
}

__global__ void k_node_fill(const int32_t* __restrict__ conn, int64_t nent, int32_t npe, int32_t* __restrict__ cursor,
                            int32_t* __restrict__ nel) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < nent) {
    const int pos = atomicAdd(&cursor[conn[t]], 1);
    nel[pos] = (int32_t)(t / npe);
  }
}

// One wavefront per node: gather the nodes of the adjacent elements into LDS, bitonic-sort,
// unique.  COUNT pass writes |NB(n)|, the row length L(n) and the overflow flag; FILL pass
// writes NB(n) at nbptr[n].  Nodes with more than kNBCap candidates are left to a second launch
// with CAP = kNBCapBig (which skips the others); beyond that the overflow flag is set.
template <bool FILL, int CAP = kNBCap>
__global__ void __launch_bounds__(64) k_neighbours(const int32_t* __restrict__ conn, int32_t npe, int32_t nNode,
                                                     const int32_t* __restrict__ nptr, const int32_t* __restrict__ nel,
                                                     const int32_t* __restrict__ nda, const int32_t* __restrict__ isP,
                                                     int32_t* __restrict__ nbcount, int32_t* __restrict__ rowlen_node,
                                                     const int32_t* __restrict__ nbptr, int32_t* __restrict__ nb,
                                                     int32_t* __restrict__ overflow) {
  __shared__ int32_t B[CAP];  // one wavefront per block: the barriers below are uniform
  constexpr int LOWER = CAP == kNBCap ? 0 : kNBCap;  // this launch's nodes: LOWER < candidates <= CAP
  const int lane = threadIdx.x;
  const int n = blockIdx.x;
  if (n >= nNode) return;  // whole block
  int ncand = 0;
  {
    const int e0 = nptr[n], e1 = nptr[n + 1];
    ncand = (e1 - e0) * npe;
    if (LOWER > 0 && ncand <= LOWER) return;  // the first launch's node (whole block)
    if (ncand > CAP) {
      if (CAP == kNBCapBig && lane == 0) atomicOr(overflow, 1);
      return;                    // the second launch's node, or too many (whole block)
    }
    for (int t = lane; t < ncand; t += 64) {
      const int e = nel[e0 + t / npe];
      B[t] = conn[(size_t)e * npe + t % npe];
    }
  }
  int size = 1;
  while (size < ncand) size <<= 1;
  for (int t = ncand + lane; t < size; t += 64) B[t] = 0x7fffffff;
  __syncthreads();
  // bitonic sort of B[0..size)
  for (int k = 2; k <= size; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = lane; i < size; i += 64) {
        const int ixj = i ^ j;
        if (ixj > i) {
          const int32_t a = B[i], b = B[ixj];
          const bool up = (i & k) == 0;
          if ((a > b) == up) { B[i] = b; B[ixj] = a; }
        }
      }
      __syncthreads();
    }
  }
  // unique + compaction with a running wave prefix
  int total = 0, ulen = 0, plen = 0;
  const int base = FILL ? nbptr[n] : 0;
  for (int c = 0; c < ncand; c += 64) {
    const int i = c + lane;
    bool keep = false;
    int32_t v = 0;
    if (i < ncand) {
      v = B[i];
      keep = (i == 0) || (B[i - 1] != v);
    }
    const uint64_t mask = __ballot(keep);
    const int before = __popcll(mask & ((1ull << lane) - 1));
    if (keep) {
      if (FILL) nb[base + total + before] = v;
      else {
        ulen += (nda[3 * (size_t)v] >= 0 ? 1 : 0) + (nda[3 * (size_t)v + 1] >= 0 ? 1 : 0) +
                (nda[3 * (size_t)v + 2] >= 0 ? 1 : 0);
        if (isP) plen += isP[v] ? 1 : 0;
      }
    }
    total += __popcll(mask);
  }
  if (!FILL) {
    // wave sums of ulen/plen
    for (int off = 32; off > 0; off >>= 1) {
      ulen += __shfl_xor(ulen, off, 64);
      plen += __shfl_xor(plen, off, 64);
    }
    if (lane == 0) {
      nbcount[n] = total;
      rowlen_node[n] = ulen + plen;
    }
  }
}

// Row lengths in row order and per-node info.
__global__ void k_rowlen(int32_t nNode, const int32_t* __restrict__ nda, const int32_t* __restrict__ idp,
                         int32_t dispDOF, const int32_t* __restrict__ rowlen_node, int32_t* __restrict__ rowlen,
                         int32_t* __restrict__ nodeinfo) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nNode) return;
  const int L = rowlen_node[n];
  int mask = 0;
  for (int d = 0; d < 3; ++d) {
    const int r = nda[3 * (size_t)n + d];
    if (r >= 0) { rowlen[r] = L; mask |= 1 << d; }
  }
  int pflag = 0;
  if (idp) {
    const int q = idp[n];
    if (q >= 0) { rowlen[dispDOF + q] = L; pflag = 1; }
  }
  nodeinfo[n] = (L << 4) | (pflag << 3) | mask;
}

// Column fill: one wavefront per node writes the column list into each of the node's rows.
__global__ void __launch_bounds__(64) k_colfill(int32_t nNode, const int32_t* __restrict__ nbptr,
                                                  const int32_t* __restrict__ nb, const int32_t* __restrict__ nda,
                                                  const int32_t* __restrict__ idp, int32_t dispDOF,
                                                  const int32_t* __restrict__ rowptr, int32_t* __restrict__ colidx) {
  const int lane = threadIdx.x;
  const int n = blockIdx.x;
  if (n >= nNode) return;
  int rows[4], nr = 0;
  for (int d = 0; d < 3; ++d) {
    const int r = nda[3 * (size_t)n + d];
    if (r >= 0) rows[nr++] = r;
  }
  if (idp && idp[n] >= 0) rows[nr++] = dispDOF + idp[n];
  if (nr == 0) return;
  const int b0 = nbptr[n], b1 = nbptr[n + 1];
  int upos = 0;  // running position of u columns
  int ptotal = 0;
  // u columns
  for (int c = b0; c < b1; c += 64) {
    const int i = c + lane;
    int cnt = 0, f[3] = {-1, -1, -1};
    if (i < b1) {
      const int m = nb[i];
      for (int d = 0; d < 3; ++d) {
        const int r = nda[3 * (size_t)m + d];
        if (r >= 0) f[cnt++] = r;
      }
    }
    // exclusive wave scan of cnt
    int incl = cnt;
    for (int off = 1; off < 64; off <<= 1) {
      const int t = __shfl_up(incl, off, 64);
      if (lane >= off) incl += t;
    }
    const int excl = incl - cnt;
    for (int k = 0; k < nr; ++k) {
      int32_t* dst = colidx + rowptr[rows[k]] + upos + excl;
      for (int j = 0; j < cnt; ++j) dst[j] = f[j];
    }
    upos += __shfl(incl, 63, 64);
  }
  if (idp) {
    for (int c = b0; c < b1; c += 64) {
      const int i = c + lane;
      int has = 0, col = 0;
      if (i < b1) {
        const int m = nb[i];
        if (idp[m] >= 0) { has = 1; col = dispDOF + idp[m]; }
      }
      const uint64_t mask = __ballot(has);
      const int before = __popcll(mask & ((1ull << lane) - 1));
      if (has)
        for (int k = 0; k < nr; ++k) colidx[rowptr[rows[k]] + upos + ptotal + before] = col;
      ptotal += __popcll(mask);
    }
  }
}

__device__ __forceinline__ int find_col(const int32_t* __restrict__ colidx, int lo, int hi, int c) {
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (colidx[mid] < c) lo = mid + 1; else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ int first_free(const int32_t* __restrict__ nda, int n) {
  for (int d = 0; d < 3; ++d) {
    const int r = nda[3 * (size_t)n + d];
    if (r >= 0) return r;
  }
  return -1;
}

// Slot of (row, col) in CSR; -1 when either is absent.
__device__ __forceinline__ int slot_of(const int32_t* __restrict__ rowptr, const int32_t* __restrict__ colidx, int r,
                                       int c) {
  if (r < 0 || c < 0) return -1;
  const int pos = find_col(colidx, rowptr[r], rowptr[r + 1], c);
  return (pos < rowptr[r + 1] && colidx[pos] == c) ? pos : -1;
}

// Hex8 slot map: slot[e][a*8+b] = CSR position of (first free row of a, first free col of b).
// Displacement-element slot map (NPE^2 per element): [a*NPE+b] = (first free row of a, first free col of b)
template <int NPE>
__global__ void k_slots_disp(const int32_t* __restrict__ conn, int32_t nElem, const int32_t* __restrict__ nda,
                             const int32_t* __restrict__ rowptr, const int32_t* __restrict__ colidx,
                             int32_t* __restrict__ slot) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)nElem * NPE * NPE) return;
  const int64_t e = t / (NPE * NPE);
  const int a = (int)(t % (NPE * NPE)) / NPE, b = (int)(t % NPE);
  const int ra = first_free(nda, conn[e * NPE + a]);
  const int cb = first_free(nda, conn[e * NPE + b]);
  slot[t] = slot_of(rowptr, colidx, ra, cb);
}

// Tet10/P1 slot map (196 per element):
//   [a*10+b]        uu: (first free row of a, first free col of b)
//   100 + a*4+q     up: (first free row of a, pressure col of corner q)
//   140 + q*10+b    pu: (pressure row of q, first free col of b)
//   180 + q1*4+q2   pp: (pressure row of q1, pressure col of q2)
__global__ void k_slots_tet10(const int32_t* __restrict__ conn, int32_t nElem, const int32_t* __restrict__ nda,
                              const int32_t* __restrict__ idp, int32_t dispDOF, const int32_t* __restrict__ rowptr,
                              const int32_t* __restrict__ colidx, int32_t* __restrict__ slot) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)nElem * kTet10Slots) return;
  const int64_t e = t / kTet10Slots;
  const int s = (int)(t % kTet10Slots);
  const int32_t* cn = conn + e * 10;
  int r, c;
  if (s < 100) {
    r = first_free(nda, cn[s / 10]);
    c = first_free(nda, cn[s % 10]);
  } else if (s < 140) {
    const int a = (s - 100) / 4, q = (s - 100) % 4;
    r = first_free(nda, cn[a]);
    c = dispDOF + idp[cn[q]];
  } else if (s < 180) {
    const int q = (s - 140) / 10, b = (s - 140) % 10;
    r = dispDOF + idp[cn[q]];
    c = first_free(nda, cn[b]);
  } else {
    const int q1 = (s - 180) / 4, q2 = (s - 180) % 4;
    r = dispDOF + idp[cn[q1]];
    c = dispDOF + idp[cn[q2]];
  }
  slot[t] = slot_of(rowptr, colidx, r, c);
}

// One Jones-Plassmann round.  An uncoloured element whose uncoloured neighbours all have lower
// priority takes the smallest colour unused by its (already coloured, higher-priority) neighbours.
__global__ void k_color_round(const int32_t* __restrict__ conn, int32_t nElem, int32_t npe,
                              const int32_t* __restrict__ nptr, const int32_t* __restrict__ nel, int32_t* color,
                              int32_t* __restrict__ remaining, int32_t* __restrict__ overflow) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nElem) return;
  if (__hip_atomic_load(&color[e], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= 0) return;
  const uint64_t pe = color_priority(e);
  uint64_t used[4] = {0, 0, 0, 0};
  bool ready = true;
  for (int a = 0; a < npe && ready; ++a) {
    const int nd = conn[(size_t)e * npe + a];
    for (int k = nptr[nd]; k < nptr[nd + 1]; ++k) {
      const int f = nel[k];
      if (f == e) continue;
      const int cf = __hip_atomic_load(&color[f], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (cf < 0) {
        if (color_priority(f) > pe) { ready = false; break; }
      } else if (cf < 256) {
        used[cf >> 6] |= 1ull << (cf & 63);
      }
    }
  }
  if (!ready) {
    atomicAdd(remaining, 1);
    return;
  }
  int c = 0;
  for (int w = 0; w < 4; ++w) {
    if (~used[w]) { c = w * 64 + __ffsll((long long)~used[w]) - 1; break; }
    if (w == 3) { atomicOr(overflow, 2); c = 255; }
  }
  __hip_atomic_store(&color[e], c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ void k_iota(int32_t* __restrict__ v, int32_t n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = i;
}

__global__ void k_fill_i32(int32_t* __restrict__ v, int64_t n, int32_t val) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = val;
}

__global__ void k_color_hist(const int32_t* __restrict__ color, int32_t n, int32_t* __restrict__ hist) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) atomicAdd(&hist[color[i]], 1);
}

// exclusive scan of int32 with hipCUB into out (n entries); total returned via out_total (host)
mpf_status exscan(const int32_t* in, int32_t* out, int64_t n, int64_t* out_total, hipStream_t s) {
  size_t tmp = 0;
  MPF_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, in, out, (int)n, s));
  DevBuf<unsigned char> t;
  MPF_TRY(t.alloc(tmp + 1));
  MPF_HIP(hipcub::DeviceScan::ExclusiveSum(t.ptr, tmp, in, out, (int)n, s));
  if (out_total) {
    int32_t last_in = 0, last_out = 0;
    MPF_HIP(hipMemcpyAsync(&last_in, in + n - 1, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    MPF_HIP(hipMemcpyAsync(&last_out, out + n - 1, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    MPF_HIP(hipStreamSynchronize(s));
    *out_total = (int64_t)last_in + (int64_t)last_out;
  }
  return MPF_OK;
}

// Same for int64 totals (row pointers may only be int32 when nnz < 2^31).
mpf_status exscan_checked(const int32_t* in, int32_t* out, int64_t n, int64_t* total, hipStream_t s) {
  // compute the total in 64 bits first (sum of lengths) to refuse int32 overflow
  size_t tmp = 0;
  DevBuf<long long> sum;
  MPF_TRY(sum.alloc(1));
  auto conv = hipcub::TransformInputIterator<long long, hipcub::CastOp<long long>, const int32_t*>(
      in, hipcub::CastOp<long long>());
  MPF_HIP(hipcub::DeviceReduce::Sum(nullptr, tmp, conv, sum.ptr, (int)n, s));
  DevBuf<unsigned char> t;
  MPF_TRY(t.alloc(tmp + 1));
  MPF_HIP(hipcub::DeviceReduce::Sum(t.ptr, tmp, conv, sum.ptr, (int)n, s));
  long long h = 0;
  MPF_HIP(hipMemcpyAsync(&h, sum.ptr, sizeof(long long), hipMemcpyDeviceToHost, s));
  MPF_HIP(hipStreamSynchronize(s));
  if (h >= (1ll << 31) - 1) {
    std::fprintf(stderr, "libmpf: CSR nnz %lld exceeds int32 row pointers\n", h);
    return MPF_ERR_ARG;
  }
  *total = h;
  return exscan(in, out, n, nullptr, s);
}

}  // namespace

mpf_status build_pattern(mpf_problem* p) {
  hipStream_t s = p->stream;
  PatternBufs& P = p->pat;
  const int64_t ndof3 = (int64_t)p->nNode * 3;
  const int B = 256;

  // ---- a1: DOF numbering ----
  DevBuf<int32_t> type, flag, scan;
  MPF_TRY(type.alloc(ndof3 + 1));
  MPF_TRY(flag.alloc(ndof3 + 1));
  MPF_TRY(scan.alloc(ndof3 + 1));
  MPF_HIP(hipMemsetAsync(type.ptr, 0, (ndof3 + 1) * sizeof(int32_t), s));
  const int nbc = (int)p->bc_dof_h.size();
  if (nbc > 0) {
    hipLaunchKernelGGL(k_mark_bc, dim3(grid_for(nbc, B)), dim3(B), 0, s, p->bc_dof.ptr, nbc, type.ptr);
  }
  hipLaunchKernelGGL(k_free_flags, dim3(grid_for(ndof3, B)), dim3(B), 0, s, type.ptr, ndof3, flag.ptr);
  int64_t dispDOF = 0;
  MPF_TRY(exscan(flag.ptr, scan.ptr, ndof3, &dispDOF, s));
  P.dispDOF = (int32_t)dispDOF;
  MPF_TRY(P.nda.alloc(ndof3));
  MPF_TRY(P.assy.alloc(dispDOF > 0 ? dispDOF : 1));
  hipLaunchKernelGGL(k_number, dim3(grid_for(ndof3, B)), dim3(B), 0, s, flag.ptr, scan.ptr, ndof3, P.nda.ptr,
                     P.assy.ptr);

  // ---- a2: pressure dofs (mixed) ----
  DevBuf<int32_t> isP;
  const bool mixed = p->etype == MPF_TET10P1;
  P.presDOF = 0;
  if (mixed) {
    MPF_TRY(isP.alloc(p->nNode));
    MPF_HIP(hipMemsetAsync(isP.ptr, 0, p->nNode * sizeof(int32_t), s));
    hipLaunchKernelGGL(k_mark_pres, dim3(grid_for((int64_t)p->nElem * 4, B)), dim3(B), 0, s, p->conn.ptr, p->nElem,
                       p->npe, isP.ptr);
    int64_t presDOF = 0;
    MPF_TRY(exscan(isP.ptr, scan.ptr, p->nNode, &presDOF, s));
    P.presDOF = (int32_t)presDOF;
    MPF_TRY(P.idp.alloc(p->nNode));
    MPF_TRY(P.assyP.alloc(presDOF > 0 ? presDOF : 1));
    hipLaunchKernelGGL(k_number_pres, dim3(grid_for(p->nNode, B)), dim3(B), 0, s, isP.ptr, scan.ptr, p->nNode,
                       P.idp.ptr, P.assyP.ptr);
  }
  const int64_t ntot = (int64_t)P.dispDOF + P.presDOF;

  // ---- node -> element adjacency ----
  const int64_t nent = (int64_t)p->nElem * p->npe;
  DevBuf<int32_t> cnt;
  MPF_TRY(cnt.alloc(p->nNode + 1));
  MPF_TRY(P.nptr.alloc(p->nNode + 1));
  MPF_HIP(hipMemsetAsync(cnt.ptr, 0, (p->nNode + 1) * sizeof(int32_t), s));
  hipLaunchKernelGGL(k_node_count, dim3(grid_for(nent, B)), dim3(B), 0, s, p->conn.ptr, nent, cnt.ptr);
  int64_t nadj = 0;
  MPF_TRY(exscan(cnt.ptr, P.nptr.ptr, p->nNode + 1, &nadj, s));
  MPF_TRY(P.nel.alloc(nadj));
  MPF_HIP(hipMemcpyAsync(cnt.ptr, P.nptr.ptr, (p->nNode + 1) * sizeof(int32_t), hipMemcpyDeviceToDevice, s));
  hipLaunchKernelGGL(k_node_fill, dim3(grid_for(nent, B)), dim3(B), 0, s, p->conn.ptr, nent, p->npe, cnt.ptr,
                     P.nel.ptr);

  // ---- node neighbourhoods, row lengths ----
  DevBuf<int32_t> nbcount, rowlen_node, overflow;
  MPF_TRY(nbcount.alloc(p->nNode + 1));
  MPF_TRY(rowlen_node.alloc(p->nNode));
  MPF_TRY(overflow.alloc(1));
  MPF_HIP(hipMemsetAsync(overflow.ptr, 0, sizeof(int32_t), s));
  MPF_HIP(hipMemsetAsync(nbcount.ptr + p->nNode, 0, sizeof(int32_t), s));
  const unsigned gn = (unsigned)p->nNode;
  hipLaunchKernelGGL(k_neighbours<false>, dim3(gn), dim3(64), 0, s, p->conn.ptr, p->npe, p->nNode, P.nptr.ptr,
                     P.nel.ptr, P.nda.ptr, mixed ? isP.ptr : nullptr, nbcount.ptr, rowlen_node.ptr, nullptr, nullptr,
                     overflow.ptr);
  hipLaunchKernelGGL((k_neighbours<false, kNBCapBig>), dim3(gn), dim3(64), 0, s, p->conn.ptr, p->npe, p->nNode,
                     P.nptr.ptr, P.nel.ptr, P.nda.ptr, mixed ? isP.ptr : nullptr, nbcount.ptr, rowlen_node.ptr, nullptr,
                     nullptr, overflow.ptr);
  MPF_TRY(P.nbptr.alloc(p->nNode + 1));
  int64_t nbtot = 0;
  MPF_TRY(exscan(nbcount.ptr, P.nbptr.ptr, p->nNode + 1, &nbtot, s));
  MPF_TRY(P.nb.alloc(nbtot > 0 ? nbtot : 1));
  hipLaunchKernelGGL(k_neighbours<true>, dim3(gn), dim3(64), 0, s, p->conn.ptr, p->npe, p->nNode, P.nptr.ptr,
                     P.nel.ptr, P.nda.ptr, mixed ? isP.ptr : nullptr, nullptr, nullptr, P.nbptr.ptr, P.nb.ptr,
                     overflow.ptr);
  hipLaunchKernelGGL((k_neighbours<true, kNBCapBig>), dim3(gn), dim3(64), 0, s, p->conn.ptr, p->npe, p->nNode,
                     P.nptr.ptr, P.nel.ptr, P.nda.ptr, mixed ? isP.ptr : nullptr, nullptr, nullptr, P.nbptr.ptr,
                     P.nb.ptr, overflow.ptr);
  int32_t ovf = 0;
  MPF_HIP(hipMemcpyAsync(&ovf, overflow.ptr, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  MPF_HIP(hipStreamSynchronize(s));
  if (ovf) {
    std::fprintf(stderr, "libmpf: a node touches more than %d element-node entries\n", kNBCapBig);
    return MPF_ERR_ARG;
  }

  // ---- a3: CSR pattern ----
  DevBuf<int32_t> rowlen;
  MPF_TRY(rowlen.alloc(ntot + 1));
  MPF_HIP(hipMemsetAsync(rowlen.ptr, 0, (ntot + 1) * sizeof(int32_t), s));
  MPF_TRY(P.nodeinfo.alloc(p->nNode));
  hipLaunchKernelGGL(k_rowlen, dim3(grid_for(p->nNode, B)), dim3(B), 0, s, p->nNode, P.nda.ptr,
                     mixed ? P.idp.ptr : nullptr, P.dispDOF, rowlen_node.ptr, rowlen.ptr, P.nodeinfo.ptr);
  MPF_TRY(P.rowptr.alloc(ntot + 1));
  int64_t nnz = 0;
  MPF_TRY(exscan_checked(rowlen.ptr, P.rowptr.ptr, ntot + 1, &nnz, s));
  P.nnz = nnz;
  MPF_TRY(P.colidx.alloc(nnz > 0 ? nnz : 1));
  hipLaunchKernelGGL(k_colfill, dim3(gn), dim3(64), 0, s, p->nNode, P.nbptr.ptr, P.nb.ptr, P.nda.ptr,
                     mixed ? P.idp.ptr : nullptr, P.dispDOF, P.rowptr.ptr, P.colidx.ptr);

  // ---- per-element slot map ----
  if (p->etype == MPF_HEX8) {
    MPF_TRY(P.slot.alloc((size_t)p->nElem * kHex8Slots));
    hipLaunchKernelGGL(k_slots_disp<8>, dim3(grid_for((int64_t)p->nElem * kHex8Slots, B)), dim3(B), 0, s,
                       p->conn.ptr, p->nElem, P.nda.ptr, P.rowptr.ptr, P.colidx.ptr, P.slot.ptr);
  } else if (p->etype == MPF_TET10) {
    MPF_TRY(P.slot.alloc((size_t)p->nElem * kTet10DSlots));
    hipLaunchKernelGGL(k_slots_disp<10>, dim3(grid_for((int64_t)p->nElem * kTet10DSlots, B)), dim3(B), 0, s,
                       p->conn.ptr, p->nElem, P.nda.ptr, P.rowptr.ptr, P.colidx.ptr, P.slot.ptr);
  } else {
    MPF_TRY(P.slot.alloc((size_t)p->nElem * kTet10Slots));
    hipLaunchKernelGGL(k_slots_tet10, dim3(grid_for((int64_t)p->nElem * kTet10Slots, B)), dim3(B), 0, s,
                       p->conn.ptr, p->nElem, P.nda.ptr, P.idp.ptr, P.dispDOF, P.rowptr.ptr, P.colidx.ptr,
                       P.slot.ptr);
  }

  // ---- a4: colouring ----
  MPF_TRY(P.color.alloc(p->nElem));
  hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(p->nElem, B)), dim3(B), 0, s, P.color.ptr, (int64_t)p->nElem, -1);
  DevBuf<int32_t> remaining;
  MPF_TRY(remaining.alloc(1));
  MPF_HIP(hipMemsetAsync(overflow.ptr, 0, sizeof(int32_t), s));
  for (int round = 0;; ++round) {
    MPF_HIP(hipMemsetAsync(remaining.ptr, 0, sizeof(int32_t), s));
    hipLaunchKernelGGL(k_color_round, dim3(grid_for(p->nElem, B)), dim3(B), 0, s, p->conn.ptr, p->nElem, p->npe,
                       P.nptr.ptr, P.nel.ptr, P.color.ptr, remaining.ptr, overflow.ptr);
    int32_t rem = 0;
    MPF_HIP(hipMemcpyAsync(&rem, remaining.ptr, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    MPF_HIP(hipStreamSynchronize(s));
    if (rem == 0) break;
    if (round > 100000) return MPF_ERR_ARG;
  }
  MPF_HIP(hipMemcpyAsync(&ovf, overflow.ptr, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  MPF_HIP(hipStreamSynchronize(s));
  if (ovf) {
    std::fprintf(stderr, "libmpf: colouring needs more than 256 colours\n");
    return MPF_ERR_ARG;
  }
  // permutation sorted by (colour, element id)
  DevBuf<int32_t> ids, keys_out;
  MPF_TRY(ids.alloc(p->nElem));
  MPF_TRY(keys_out.alloc(p->nElem));
  MPF_TRY(P.perm.alloc(p->nElem));
  hipLaunchKernelGGL(k_iota, dim3(grid_for(p->nElem, B)), dim3(B), 0, s, ids.ptr, p->nElem);
  size_t tmp = 0;
  MPF_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, P.color.ptr, keys_out.ptr, ids.ptr, P.perm.ptr, p->nElem,
                                             0, 9, s));
  DevBuf<unsigned char> t;
  MPF_TRY(t.alloc(tmp + 1));
  MPF_HIP(hipcub::DeviceRadixSort::SortPairs(t.ptr, tmp, P.color.ptr, keys_out.ptr, ids.ptr, P.perm.ptr, p->nElem, 0,
                                             9, s));
  DevBuf<int32_t> hist;
  MPF_TRY(hist.alloc(256));
  MPF_HIP(hipMemsetAsync(hist.ptr, 0, 256 * sizeof(int32_t), s));
  hipLaunchKernelGGL(k_color_hist, dim3(grid_for(p->nElem, B)), dim3(B), 0, s, P.color.ptr, p->nElem, hist.ptr);
  std::vector<int32_t> h(256);
  MPF_HIP(hipMemcpyAsync(h.data(), hist.ptr, 256 * sizeof(int32_t), hipMemcpyDeviceToHost, s));
  MPF_HIP(hipStreamSynchronize(s));
  P.ncolors = 0;
  for (int c = 0; c < 256; ++c)
    if (h[c] > 0) P.ncolors = c + 1;
  P.color_off.assign(P.ncolors + 1, 0);
  for (int c = 0; c < P.ncolors; ++c) P.color_off[c + 1] = P.color_off[c] + h[c];
  MPF_HIP(hipGetLastError());
  return MPF_OK;
}

}  // namespace mpf
