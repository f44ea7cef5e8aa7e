// gather.hip -- row-owner ("gather") assembly of the tangent and residual into CSR.
//
// Same element arithmetic as the reference's per-element loop (Elem_Solid_3D_Disp.cpp:196-374,
// ElementBase2.cpp:278-552) and the same assembled K and R as SolverPetsc::assembleMatrixAndVector
// (SolverPetsc.cpp:357-378, 434-502) with the first-iteration Dirichlet lift (ElementBase.cpp:567-594,
// 717-782), organised for the GPU in two kernels:
//
//   stage kernel  (one wavefront per EPW elements, lane = (element, Gauss point)):
//       per element and Gauss point, grad_x N and b grad_x N of every node and the tangent
//       scalars, plus the element residual, into a staging record of TS doubles per element.
//   row kernel    (one wavefront per node n):
//       for every element e around n and every node b of e, the 3x3 block K_{n b} (and the u-p
//       couplings) from e's staging record, visited in (column block, element id) order from a table
//       built once with the pattern; each column block's run of contributions is summed in element
//       order, then every CSR value of n's rows and n's right-hand-side entries are written once.
//
// Nothing is read-modify-written, so there is no zero fill, no colouring and no atomics on the FP64
// path, and the result is bitwise reproducible.  Each column block's contributions are summed in
// element order, the reference's own order for a given row (femSolidmechanics2.cpp:1061).

#include "fem_device.hpp"

namespace mpf {

namespace {

using namespace dev;

constexpr int kMaxAdj = 64;   // elements around one node
constexpr int kMaxNB = 128;   // nodes coupled to one node (column blocks of its rows)

// Staging record layout (doubles) per element:
//   G  [NGP][NPE][6]   g_a (3), (b g)_a (3)
//   S  [NGP][SS]       Hex8: s1..s5 (SS = 5); Tet10/P1: s1..s4, dv, dv0, thetahat, resip (SS = 8)
//   R  [NPE][3]        element residual (no lift)
template <int NPE, int NGP, bool MIXED>
struct Stage {
  static constexpr int SS = MIXED ? 8 : 5;
  static constexpr int G0 = 0;
  static constexpr int S0 = NGP * NPE * 6;
  static constexpr int R0 = S0 + NGP * SS;
  static constexpr int TS = ((R0 + NPE * 3) + 7) / 8 * 8;
};

template <int NPE, int NGP, bool MIXED, int EPW>
__global__ void __launch_bounds__(64) k_stage(ElemArgs A, int32_t nElem, double* __restrict__ GS) {
  static_assert(EPW * NGP == 64, "one wavefront holds EPW elements x NGP Gauss points");
  using St = Stage<NPE, NGP, MIXED>;
  __shared__ double sX[EPW][NPE][3], sU[EPW][NPE][3], sUP[MIXED ? EPW : 1][NPE][3], sP[EPW][4];
  __shared__ double sdN[NGP][3][NPE];
  const int lane = threadIdx.x;
  const int base = blockIdx.x * EPW;
  if (lane < NGP) {
    double d[3][NPE];
    if (NPE == 8) {
      hex8_dN(lane, (double(*)[8])d);
    } else {
      double x1, x2, x3;
      tet_gp(lane, x1, x2, x3);
      tet10_dN(x1, x2, x3, (double*)d[0], (double*)d[1], (double*)d[2]);
    }
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int a = 0; a < NPE; ++a) sdN[lane][r][a] = d[r][a];
  }
  for (int t = lane; t < EPW * NPE; t += 64) {
    const int k = t / NPE, a = t % NPE;
    const int e = (base + k < nElem) ? base + k : base;  // idle slots repeat the first element
    const size_t nd = (size_t)A.conn[(size_t)e * NPE + a];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      sX[k][a][c] = A.X[3 * nd + c];
      sU[k][a][c] = A.u[3 * nd + c];
      if (MIXED) sUP[k][a][c] = A.uprev[3 * nd + c];
    }
    if (MIXED && a < 4) sP[k][a] = A.pres[nd];
  }
  __syncthreads();
  const int k = lane / NGP, gp = lane % NGP;
  const int e = base + k;
  const bool writer = e < nElem;
  double* rec = GS + (size_t)(writer ? e : base) * St::TS;
  const double wgp = (NPE == 8) ? 1.0 : 0.2500000000000000 * (1.0 / 6.0);  // QuadratureUtil.cpp:466-492, 706-709
  bool negJ = false;
  element_point<NPE, NGP, MIXED>(sX[k], sU[k], MIXED ? sUP[k] : nullptr, MIXED ? sP[k] : nullptr, sdN[gp], gp, wgp,
                                 A.mu, 1.0 / A.Kinv, A.Utype, A.consistent != 0, writer,
                                 rec + St::G0 + gp * NPE * 6, rec + St::S0 + gp * St::SS, rec + St::R0, negJ);
  if (writer && negJ) flag_bad(A.bad, e);
}

struct RowArgs {
  int32_t nNode;
  const int32_t* nptr;      // node -> elements (offsets)
  const int32_t* nelA;      // elements around each node, ascending: e * 16 + (local index of the node in e)
  const int32_t* items;     // per node, (column block j << 16) | (s << 4) | b sorted by (j, s)
  const int32_t* nbptr;     // node -> coupled nodes (sorted)
  const int32_t* nb;
  const int32_t* nodeinfo;  // (row length << 4) | (pressure << 3) | free mask
  const int32_t* nda;
  const int32_t* idp;
  int32_t dispDOF;
  const int32_t* rowptr;
  const double* GS;
  const double* dispApplied;  // non-null: first-iteration lift
  double* vals;
  double* rhs;
  int32_t writeK;
  int32_t* overflow;
};

// XCD-aware node order: consecutive workgroups land on different XCDs (round robin), so each XCD
// is given a contiguous range of nodes and the staging records of neighbouring nodes stay in its L2.
__device__ __forceinline__ int node_of_block(int b, int nNode) {
  const int per = (nNode + 7) / 8;
  return (b % 8) * per + b / 8;
}

// Ascending bitonic sort of n (a power of two, >= 64) keys in LDS by one wavefront.
__device__ __forceinline__ void wave_sort(int32_t* key, int n, int lane) {
  for (int k = 2; k <= n; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int t = lane; t < n; t += 64) {
        const int u = t ^ j;
        if (u > t) {
          const int32_t x = key[t], y = key[u];
          if ((x > y) == ((t & k) == 0)) { key[t] = y; key[u] = x; }
        }
      }
      __syncthreads();
    }
}

// Per-node gather tables, built once with the pattern (one wavefront per node):
//   nelA [nptr[n] + s]          = e * 16 + a: the elements around n in ascending id, a = local index of n in e
//   items[nptr[n] * NPE + t]    = (j << 16) | (s << 4) | b for every element s around n and node b of it,
//                                 sorted by (j, s), j = position of node conn[e][b] in the sorted coupled-node
//                                 list of n, i.e. the column block of n's rows that the (s, b) block lands in.
template <int NPE>
__global__ void __launch_bounds__(64) k_gather_tables(int32_t nNode, const int32_t* __restrict__ nptr,
                                                      const int32_t* __restrict__ nel,
                                                      const int32_t* __restrict__ nbptr,
                                                      const int32_t* __restrict__ nb,
                                                      const int32_t* __restrict__ conn, int32_t* __restrict__ nelA,
                                                      int32_t* __restrict__ items, int32_t* __restrict__ overflow) {
  constexpr int KCAP = NPE <= 8 ? 512 : 1024;  // kMaxAdj * NPE rounded up to a power of two
  static_assert(kMaxAdj * NPE <= KCAP, "item capacity");
  __shared__ int32_t sE[kMaxAdj], sNB[kMaxNB], sK[KCAP];
  const int lane = threadIdx.x;
  const int n = blockIdx.x;
  if (n >= nNode) return;
  const int e0 = nptr[n], m = nptr[n + 1] - e0;
  const int b0 = nbptr[n], nnb = nbptr[n + 1] - b0;
  if (m > kMaxAdj || nnb > kMaxNB) {
    if (lane == 0) atomicOr(overflow, 1);
    return;
  }
  sE[lane] = lane < m ? nel[e0 + lane] : 0x7fffffff;
  for (int j = lane; j < nnb; j += 64) sNB[j] = nb[b0 + j];
  __syncthreads();
  wave_sort(sE, 64, lane);
  if (lane < m) {
    const int32_t* c = conn + (size_t)sE[lane] * NPE;
    int a = 0;
    for (int t = 0; t < NPE; ++t)
      if (c[t] == n) a = t;
    nelA[e0 + lane] = sE[lane] * 16 + a;
  }
  const int nitems = m * NPE;
  int P2 = 64;
  while (P2 < nitems) P2 <<= 1;
  for (int t = lane; t < P2; t += 64) {
    int32_t key = 0x7fffffff;
    if (t < nitems) {
      const int s = t / NPE, b = t % NPE;
      const int vb = conn[(size_t)sE[s] * NPE + b];
      int lo = 0, hi = nnb;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (sNB[mid] < vb) lo = mid + 1; else hi = mid;
      }
      if (lo >= nnb || sNB[lo] != vb) atomicOr(overflow, 2);  // not a coupled node: inconsistent tables
      key = (min(lo, nnb - 1) << 16) | (s << 4) | b;
    }
    sK[t] = key;
  }
  __syncthreads();
  wave_sort(sK, P2, lane);
  for (int t = lane; t < nitems; t += 64) items[(size_t)e0 * NPE + t] = sK[t];
}

// Row kernel: one wavefront per node n, 64 (element, node) items per pass in (column, element) order.
// Item t computes the 3x3 block K_{n b} of element s (plus the u-p couplings) from s's staging record;
// the first item of each column run (the run's head) sums the run in element order, carrying a run
// that continues into the next pass, and writes the column block of n's rows once.
template <int NPE, int NGP, bool MIXED>
__global__ void __launch_bounds__(64) k_rows(RowArgs R) {
  using St = Stage<NPE, NGP, MIXED>;
  constexpr int NV = MIXED ? 16 : 9;  // uu 9 | up 3 | pu 3 | pp 1
  constexpr int NL = MIXED ? 4 : 3;   // right-hand-side rows of a node
  __shared__ int32_t sEA[kMaxAdj];
  __shared__ int32_t sNB[kMaxNB], sInfo[kMaxNB], sUpos[kMaxNB], sPpos[kMaxNB];
  __shared__ int32_t sJ[65];
  __shared__ double sPart[64][NV];
  __shared__ double sCarry[NV];
  __shared__ double sLift[64][NL];
  const int lane = threadIdx.x;
  const int n = node_of_block(blockIdx.x, R.nNode);
  if (n >= R.nNode) return;
  const int info_n = R.nodeinfo[n];
  const int mn = info_n & 7;
  const bool isPn = MIXED && ((info_n >> 3) & 1);
  if (!mn && !isPn) return;
  const int e0 = R.nptr[n], m = R.nptr[n + 1] - e0;
  const int b0 = R.nbptr[n], nnb = R.nbptr[n + 1] - b0;
  if (m > kMaxAdj || nnb > kMaxNB) {  // the table build rejects such meshes; never reached
    if (lane == 0) atomicOr(R.overflow, 1);
    return;
  }
  if (lane < m) sEA[lane] = R.nelA[e0 + lane];
  // coupled nodes, their info and the column offsets of their blocks in n's rows
  int cu0 = 0, cp0 = 0;
  for (int c = 0; c < nnb; c += 64) {
    const int j = c + lane;
    int inf = 0, nu = 0, np = 0;
    if (j < nnb) {
      const int v = R.nb[b0 + j];
      inf = R.nodeinfo[v];
      sNB[j] = v;
      sInfo[j] = inf;
      nu = __popc(inf & 7);
      np = MIXED ? ((inf >> 3) & 1) : 0;
    }
    int iu = nu, ip = np;
    for (int off = 1; off < 64; off <<= 1) {
      const int tu = __shfl_up(iu, off, 64), tp = __shfl_up(ip, off, 64);
      if (lane >= off) { iu += tu; ip += tp; }
    }
    if (j < nnb) { sUpos[j] = cu0 + iu - nu; sPpos[j] = cp0 + ip - np; }
    cu0 += __shfl(iu, 63, 64);
    cp0 += __shfl(ip, 63, 64);
  }
  const int Ucols = cu0;
  __syncthreads();

  double liftacc = 0.0;  // lanes < NL: sum over columns of K_{n j} dubar_j (first iteration)
  if (R.writeK || R.dispApplied) {
    const int nitems = m * NPE;
    const int32_t* items = R.items + (size_t)e0 * NPE;
    for (int p0 = 0; p0 < nitems; p0 += 64) {
      const int t = p0 + lane;
      const bool live = t < nitems;
      const int key = live ? items[t] : -1;
      const int j = live ? (key >> 16) : -1;
      double v[NV];
#pragma unroll
      for (int q = 0; q < NV; ++q) v[q] = 0.0;
      if (live) {
        const int s = (key >> 4) & 63, b = key & 15;
        const int ea = sEA[s];
        const int e = ea >> 4, a = ea & 15;
        const double* rec = R.GS + (size_t)e * St::TS;
        tangent_block<NGP, !MIXED>(rec + St::G0 + a * 6, rec + St::G0 + b * 6, NPE * 6, rec + St::S0, St::SS, v);
        if (MIXED) {
          const double* S = rec + St::S0;
#pragma unroll
          for (int gp = 0; gp < NGP; ++gp) {
            double Np[4];
            tet_Np(gp, Np);
            const double* ga = rec + St::G0 + gp * NPE * 6 + a * 6;
            const double* gb = rec + St::G0 + gp * NPE * 6 + b * 6;
            const double dv = S[gp * St::SS + 4], dv0 = S[gp * St::SS + 5], th = S[gp * St::SS + 6];
            const double Nb = b == 0 ? Np[0] : b == 1 ? Np[1] : b == 2 ? Np[2] : Np[3];
            const double Na = a == 0 ? Np[0] : a == 1 ? Np[1] : a == 2 ? Np[2] : Np[3];
            if (b < 4) {  // Kup(3a+i, b) += (dv g_a,i) Np_b   (ElementBase2.cpp:500-505)
              v[9] += (dv * ga[0]) * Nb; v[10] += (dv * ga[1]) * Nb; v[11] += (dv * ga[2]) * Nb;
            }
            if (a < 4) {  // Kpu(a, 3b+k) += (dv Np_a) g_b,k ; Kpp(a, b) -= (dv0 Np_a thetahat) Np_b (:524-539)
              const double bb4 = dv * Na;
              v[12] += bb4 * gb[0]; v[13] += bb4 * gb[1]; v[14] += bb4 * gb[2];
              if (b < 4) v[15] -= ((dv0 * Na) * th) * Nb;
            }
          }
        }
      }
      sJ[lane] = j;
      if (lane == 0) sJ[64] = (p0 + 64 < nitems) ? (items[p0 + 64] >> 16) : -1;
#pragma unroll
      for (int q = 0; q < NV; ++q) sPart[lane][q] = v[q];
      __syncthreads();
      bool hasLift = false;
      if (live && (lane == 0 || sJ[lane - 1] != j)) {  // head of a column run
        double acc[NV];
        const bool carried = lane == 0 && p0 > 0 && (items[p0 - 1] >> 16) == j;
#pragma unroll
        for (int q = 0; q < NV; ++q) acc[q] = carried ? sCarry[q] + v[q] : v[q];
        int t2 = lane + 1;
        for (; t2 < 64 && sJ[t2] == j; ++t2)
#pragma unroll
          for (int q = 0; q < NV; ++q) acc[q] += sPart[t2][q];
        if (t2 == 64 && sJ[64] == j) {
          // the run continues in the next pass (its lane 0 reads this after the pass-end barrier)
#pragma unroll
          for (int q = 0; q < NV; ++q) sCarry[q] = acc[q];
        } else {
          const int inf = sInfo[j];
          const int mb = inf & 7;
          const bool isPb = MIXED && ((inf >> 3) & 1);
          if (R.writeK) {
            const int up = sUpos[j], pp = Ucols + sPpos[j];
#pragma unroll
            for (int i = 0; i < 3; ++i) {
              if (!((mn >> i) & 1)) continue;
              double* row = R.vals + R.rowptr[R.nda[3 * (size_t)n + i]];
#pragma unroll
              for (int k = 0; k < 3; ++k)
                if ((mb >> k) & 1) row[up + rank3(mb, k)] = acc[3 * i + k];
              if (isPb) row[pp] = acc[9 + i];
            }
            if (isPn) {
              double* row = R.vals + R.rowptr[R.dispDOF + R.idp[n]];
#pragma unroll
              for (int k = 0; k < 3; ++k)
                if ((mb >> k) & 1) row[up + rank3(mb, k)] = acc[12 + k];
              if (isPb) row[pp] = acc[15];
            }
          }
          if (R.dispApplied && mb != 7) {  // F_n -= K_{n j} dubar_j over the prescribed dofs of j
            const int vb = sNB[j];
            double db[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) db[k] = ((mb >> k) & 1) ? 0.0 : R.dispApplied[3 * (size_t)vb + k];
#pragma unroll
            for (int i = 0; i < 3; ++i) sLift[lane][i] = acc[3 * i] * db[0] + acc[3 * i + 1] * db[1] + acc[3 * i + 2] * db[2];
            if (MIXED) sLift[lane][NL - 1] = acc[12] * db[0] + acc[13] * db[1] + acc[14] * db[2];
            hasLift = true;
          }
        }
      }
      const uint64_t lm = __ballot(hasLift);
      __syncthreads();
      if (lane < NL) {
        uint64_t mm = lm;
        while (mm) {
          const int h = __ffsll((unsigned long long)mm) - 1;
          liftacc += sLift[h][lane];
          mm &= mm - 1;
        }
      }
    }
  }

  // right-hand side of n: element residuals in element order, minus the lift
  if (lane < NL) {
    const int i = lane;
    double r = 0.0;
    bool own = false;
    if (i < 3) {
      own = (mn >> i) & 1;
      if (own)
        for (int s = 0; s < m; ++s) {
          const int ea = sEA[s];
          r += R.GS[(size_t)(ea >> 4) * St::TS + St::R0 + (ea & 15) * 3 + i];
        }
    } else if (isPn) {
      own = true;
      for (int s = 0; s < m; ++s) {
        const int ea = sEA[s];
        const int a = ea & 15;
        if (a >= 4) continue;
        const double* S = R.GS + (size_t)(ea >> 4) * St::TS + St::S0;
        double fp = 0.0;  // Fp_a -= (dv0 Np_a) resip   (ElementBase2.cpp:510, 522)
        for (int gp = 0; gp < NGP; ++gp) {
          double Np[4];
          tet_Np(gp, Np);
          const double Na = a == 0 ? Np[0] : a == 1 ? Np[1] : a == 2 ? Np[2] : Np[3];
          fp -= (S[gp * St::SS + 5] * Na) * S[gp * St::SS + 7];
        }
        r += fp;
      }
    }
    if (own) {
      const int row = i < 3 ? R.nda[3 * (size_t)n + i] : R.dispDOF + R.idp[n];
      R.rhs[row] = r - liftacc;
    }
  }
}

}  // namespace

size_t stage_doubles_per_elem(int etype) {
  return etype == MPF_HEX8 ? (size_t)Stage<8, 8, false>::TS : (size_t)Stage<10, 4, true>::TS;
}

mpf_status ensure_gather_tables(mpf_problem* p) {
  PatternBufs& P = p->pat;
  if (P.gather_ready) return MPF_OK;
  hipStream_t s = p->stream;
  const int64_t nadj = (int64_t)p->nElem * p->npe;
  if (p->nElem >= (1 << 27)) {  // nelA packs e * 16 + a into int32
    std::fprintf(stderr, "libmpf: more than 2^27 elements; using the coloured scatter\n");
    p->asm_mode = MPF_ASSEMBLY_COLORED_SCATTER;
    return MPF_OK;
  }
  MPF_TRY(P.g_nelA.alloc(nadj > 0 ? nadj : 1));
  MPF_TRY(P.g_items.alloc(nadj > 0 ? (size_t)nadj * p->npe : 1));
  MPF_TRY(p->GS.alloc((size_t)p->nElem * stage_doubles_per_elem(p->etype)));
  DevBuf<int32_t> ovf;
  MPF_TRY(ovf.alloc(1));
  MPF_HIP(hipMemsetAsync(ovf.ptr, 0, sizeof(int32_t), s));
  if (p->nNode > 0) {
    if (p->etype == MPF_HEX8)
      hipLaunchKernelGGL(k_gather_tables<8>, dim3(p->nNode), dim3(64), 0, s, p->nNode, P.nptr.ptr, P.nel.ptr,
                         P.nbptr.ptr, P.nb.ptr, p->conn.ptr, P.g_nelA.ptr, P.g_items.ptr, ovf.ptr);
    else
      hipLaunchKernelGGL(k_gather_tables<10>, dim3(p->nNode), dim3(64), 0, s, p->nNode, P.nptr.ptr, P.nel.ptr,
                         P.nbptr.ptr, P.nb.ptr, p->conn.ptr, P.g_nelA.ptr, P.g_items.ptr, ovf.ptr);
    MPF_HIP(hipGetLastError());
  }
  int32_t o = 0;
  MPF_HIP(hipMemcpyAsync(&o, ovf.ptr, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  MPF_HIP(hipStreamSynchronize(s));
  if (o) {
    std::fprintf(stderr, "libmpf: a node has more than %d elements or %d coupled nodes; using the coloured scatter\n",
                 kMaxAdj, kMaxNB);
    P.g_nelA.release();
    P.g_items.release();
    p->GS.release();
    p->asm_mode = MPF_ASSEMBLY_COLORED_SCATTER;
    return MPF_OK;
  }
  P.gather_ready = true;
  return MPF_OK;
}

mpf_status launch_gather_assembly(int etype, const ElemArgs& A, int32_t nElem, int32_t nNode, const PatternBufs& P,
                                  double* GS, double* vals, double* rhs, const double* dispApplied, int writeK,
                                  int32_t* overflow, hipStream_t s) {
  if (!P.gather_ready) return MPF_ERR_STATE;
  RowArgs R{};
  R.nNode = nNode;
  R.nptr = P.nptr.ptr; R.nelA = P.g_nelA.ptr; R.items = P.g_items.ptr; R.nbptr = P.nbptr.ptr; R.nb = P.nb.ptr;
  R.nodeinfo = P.nodeinfo.ptr; R.nda = P.nda.ptr; R.idp = P.idp.ptr; R.dispDOF = P.dispDOF;
  R.rowptr = P.rowptr.ptr; R.GS = GS; R.dispApplied = dispApplied; R.vals = vals; R.rhs = rhs;
  R.writeK = writeK; R.overflow = overflow;
  const unsigned nodes_grid = (unsigned)(((nNode + 7) / 8) * 8);
  if (nElem <= 0 || nNode <= 0) return MPF_OK;
  if (etype == MPF_HEX8) {
    hipLaunchKernelGGL((k_stage<8, 8, false, 8>), dim3(grid_for(nElem, 8)), dim3(64), 0, s, A, nElem, GS);
    hipLaunchKernelGGL((k_rows<8, 8, false>), dim3(nodes_grid), dim3(64), 0, s, R);
  } else {
    hipLaunchKernelGGL((k_stage<10, 4, true, 16>), dim3(grid_for(nElem, 16)), dim3(64), 0, s, A, nElem, GS);
    hipLaunchKernelGGL((k_rows<10, 4, true>), dim3(nodes_grid), dim3(64), 0, s, R);
  }
  MPF_HIP(hipGetLastError());
  return MPF_OK;
}

}  // namespace mpf
