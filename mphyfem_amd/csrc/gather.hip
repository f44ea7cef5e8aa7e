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
//   matrix kernel (one wavefront per 64 laid-out items of one node n):
//       for every element e around n and every node b of e, the 3x3 block K_{n b} (and the u-p
//       couplings) from e's staging record, visited in (column block, element id) order from a table
//       built once with the pattern; each column block's run of contributions is summed in element
//       order, then every CSR value of n's rows and n's right-hand-side entries are written once.
//       Hex8 runs it per chunk of 8 wavefronts whose elements' records are first copied to LDS.
//
// Nothing is read-modify-written, so there is no zero fill, no colouring and no atomics on the FP64
// path, and the result is bitwise reproducible.  Each column block's contributions are summed in
// element order, the reference's own order for a given row (femSolidmechanics2.cpp:1061).

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>

#include "fem_device.hpp"

namespace mpf {

namespace {

using namespace dev;

constexpr int kMaxAdj = 64;   // elements around one node
constexpr int kMaxNB = 128;   // nodes coupled to one node (column blocks of its rows)

// Staging record layout (doubles) per element:
//   Tet10, Tet10/P1:  G [NGP][NPE][4]  g_a (3) and a zero (two 16-byte groups per node)
//   Hex8:      J [NGP][9]       the inverse Jacobian (g_a = grad_jinv(J, dN_a/dxi))
//   B  [NGP][6]        b00 b11 b22 b01 b02 b12 (b g formed on use; Tet10/P1 records 248 doubles
//                      instead of 304 with b g stored, Hex8 184 instead of 448)
//   S  [NGP][SS]       displacement elements: s1..s5 (SS = 5); mixed elements: s1..s4, dv, dv0,
//                      thetahat, resip (SS = 8)
//   W  [26]            Hex8 Q1/P0 only: Kup [NPE][3], Kpp, 0 (the static condensation's coupling)
//   R  [NPE][3]        element residual (no lift; Q1/P0: condensed, Fu - Kup Fp / Kpp)
// The first SL doubles (G or J, B, S, W) are what the matrix kernel reads for a block.
// MIXED with NPE = 8 is the Hex8 Q1/P0 element (one pressure per element, condensed; the system has
// displacement rows only), with NPE = 10 the Tet10/P1 element (pressure rows on the corner nodes).
template <int NPE, int NGP, bool MIXED>
struct Stage {
  static constexpr bool P0 = MIXED && NPE == 8;
  static constexpr bool P1 = MIXED && NPE != 8;
  static constexpr bool COMPACT = NPE == 8;      // g formed from the inverse Jacobian
  static constexpr int GW = COMPACT ? 3 : 4;      // doubles per node and Gauss point of g (LDS copies, Tet10 records)
  static constexpr int SS = MIXED ? 8 : 5;
  static constexpr int G0 = 0;
  static constexpr int B0 = COMPACT ? NGP * 9 : NGP * NPE * GW;
  static constexpr int S0 = B0 + NGP * 6;
  static constexpr int W0 = S0 + NGP * SS;
  static constexpr int WN = P0 ? 26 : 0;
  static constexpr int R0 = W0 + WN;
  static constexpr int SL = R0;
  static constexpr int TS = ((R0 + NPE * 3) + 7) / 8 * 8;
  static_assert(B0 % 2 == 0 && S0 % 2 == 0 && W0 % 2 == 0 && SL % 2 == 0, "16-byte groups");
};

// Hex8 Q1/P0: the static condensation's part of block (a, b), -Kup_a Kup_b^T / Kpp
// (Elem_Solid_3D_Mixed0.cpp:93-104), from the record's W
__device__ __forceinline__ void add_condensed(const double* W, int a, int b, double Kb[9]) {
  const double kpp = W[24];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) Kb[3 * i + j] = Kb[3 * i + j] - W[3 * a + i] * (W[3 * b + j] / kpp);
}

// g, b g of node c at Gauss point gp from a staging record, either layout
// (dN: Hex8 dN/dxi table [NGP][NPE][3], e.g. in LDS; null: computed here)
template <int NPE, int NGP, bool MIXED>
__device__ __forceinline__ void record_g(const double* rec, int gp, int c, double g[3], double bg[3],
                                         const double* dN = nullptr) {
  using St = Stage<NPE, NGP, MIXED>;
  if (St::COMPACT) {
    double d[3];
    if (dN) {
      d[0] = dN[3 * (gp * NPE + c)]; d[1] = dN[3 * (gp * NPE + c) + 1]; d[2] = dN[3 * (gp * NPE + c) + 2];
    } else {
      hex8_dN_node(gp, c, d);
    }
    grad_jinv(rec + St::G0 + gp * 9, d, g);
    bsym_times(rec + St::B0 + gp * 6, g, bg);
  } else {
    const double* p = rec + St::G0 + (gp * NPE + c) * St::GW;
    g[0] = p[0]; g[1] = p[1]; g[2] = p[2];
    bsym_times(rec + St::B0 + gp * 6, g, bg);
  }
}

// 3x3 tangent block (a, b) of one element from its staging record, either layout
template <int NPE, int NGP, bool MIXED>
__device__ __forceinline__ void record_block(const double* rec, int a, int b, double Kb[9],
                                             const double* dN = nullptr) {
  using St = Stage<NPE, NGP, MIXED>;
#pragma unroll
  for (int t = 0; t < 9; ++t) Kb[t] = 0.0;
#pragma unroll 2
  for (int gp = 0; gp < NGP; ++gp) {
    double ga[3], ba[3], gb[3], bb[3];
    record_g<NPE, NGP, MIXED>(rec, gp, a, ga, ba, dN);
    record_g<NPE, NGP, MIXED>(rec, gp, b, gb, bb, dN);
    const double* S = rec + St::S0 + gp * St::SS;
    tangent_gp<!MIXED>(ga, ba, gb, bb, S[0], S[1], S[2], S[3], MIXED ? 0.0 : S[4], Kb);
  }
  if (St::P0) add_condensed(rec + St::W0, a, b, Kb);
}

// Stage kernel.  OUT_LDS: the wavefront's EPW records are written to LDS and then copied to the
// staging array, which holds them contiguously, with 16-byte stores that each cover 1 KiB (written
// directly, every lane's stores land 8 bytes apiece on 64 different cache lines).  Measured per
// assembly: Hex8 128^3 8.81 ms through LDS vs 9.53 ms direct; Tet10/P1 60^3x6 8.94 vs 9.08 ms.
#ifndef MPF_STAGE_LDS_HEX
#define MPF_STAGE_LDS_HEX 1
#endif
#ifndef MPF_STAGE_LDS_TET
#define MPF_STAGE_LDS_TET 1
#endif
// Q1/P0 (qKup, qKpp, qFp non-null): each element's Kup [24], Kpp and Fp also go to these arrays for
// the pressure recovery after the solve.
template <int NPE, int NGP, bool MIXED, int EPW, bool OUT_LDS>
__global__ void __launch_bounds__(64) k_stage(ElemArgs A, int32_t nElem, double* __restrict__ GS,
                                              double* __restrict__ qKup, double* __restrict__ qKpp,
                                              double* __restrict__ qFp) {
  static_assert(EPW * NGP == 64, "one wavefront holds EPW elements x NGP Gauss points");
  using St = Stage<NPE, NGP, MIXED>;
  static_assert(St::TS % 2 == 0, "16-byte records");
  static_assert(!St::P0 || NGP == NPE, "Q1/P0: one node per Gauss-point lane in the condensation");
  __shared__ double sX[EPW][NPE][3], sU[EPW][NPE][3], sUP[MIXED ? EPW : 1][NPE][3], sP[EPW][4];
  __shared__ double sdN[NGP][3][NPE];
  __shared__ double2 sOut[OUT_LDS ? EPW * St::TS / 2 : 1];
  __shared__ double sQ[St::P0 ? EPW * NGP * 26 : 1];  // per lane: dv g_a (24), -dv0 thetahat, -dv0 resip
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
  {
    // the wavefront's (element, node) pairs, NT per lane: every connectivity load issued, then every
    // node-data load, then the LDS stores (two round trips to memory instead of two per pair)
    constexpr int NT = (EPW * NPE + 63) / 64;
    size_t nd[NT];
    int ev[NT];
#pragma unroll
    for (int i = 0; i < NT; ++i) {
      const int t = lane + 64 * i, k = t / NPE, a = t % NPE;
      ev[i] = (base + k < nElem) ? base + k : base;  // idle slots repeat the first element
      nd[i] = t < EPW * NPE ? (size_t)A.conn[(size_t)ev[i] * NPE + a] : 0;
    }
    double xv[NT][3], uv[NT][3], upv[NT][3], pv[NT];
#pragma unroll
    for (int i = 0; i < NT; ++i) {
      const int t = lane + 64 * i, a = t % NPE;
      const bool on = t < EPW * NPE;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        xv[i][c] = on ? A.X[3 * nd[i] + c] : 0.0;
        uv[i][c] = on ? A.u[3 * nd[i] + c] : 0.0;
        upv[i][c] = (MIXED && on) ? A.uprev[3 * nd[i] + c] : 0.0;
      }
      pv[i] = 0.0;
      if (St::P1 && on && a < 4) pv[i] = A.pres[nd[i]];
      if (St::P0 && on && a == 0) pv[i] = A.pres[ev[i]];
    }
#pragma unroll
    for (int i = 0; i < NT; ++i) {
      const int t = lane + 64 * i, k = t / NPE, a = t % NPE;
      if (t >= EPW * NPE) continue;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        sX[k][a][c] = xv[i][c];
        sU[k][a][c] = uv[i][c];
        if (MIXED) sUP[k][a][c] = upv[i][c];
      }
      if (St::P1 && a < 4) sP[k][a] = pv[i];
      if (St::P0 && a == 0) sP[k][0] = pv[i];
    }
  }
  __syncthreads();
  const int k = lane / NGP, gp = lane % NGP;
  const int e = base + k;
  const bool writer = e < nElem;
  double* rec = OUT_LDS ? reinterpret_cast<double*>(sOut) + k * St::TS : GS + (size_t)(writer ? e : base) * St::TS;
  const double wgp = (NPE == 8) ? 1.0 : 0.2500000000000000 * (1.0 / 6.0);  // QuadratureUtil.cpp:466-492, 706-709
  bool negJ = false;
  double mu, K;
  elem_material(A, writer ? e : base, mu, K);
  element_point<NPE, NGP, MIXED, St::COMPACT, !St::COMPACT>(sX[k], sU[k], MIXED ? sUP[k] : nullptr, MIXED ? sP[k] : nullptr, sdN[gp],
                                              gp, wgp, mu, K, A.Utype, St::P0 ? false : A.consistent != 0, writer,
                                              rec + St::G0 + gp * (St::COMPACT ? 9 : NPE * St::GW),
                                              rec + St::S0 + gp * St::SS, rec + St::R0, negJ, rec + St::B0 + gp * 6);
  if (writer && negJ) flag_bad(A.bad, e);
  if constexpr (St::P0) {
    // Kup_a = sum_gp dv g_a (ElementBase2.cpp:500-505 with Np = 1), Kpp = -sum_gp dv0 thetahat,
    // Fp = -sum_gp dv0 resip (:510-539), each summed in Gauss-point order by the lane of node a = gp;
    // g_a from this lane's inverse Jacobian as element_point formed it
    const double* Sg = rec + St::S0 + gp * St::SS;
    double* q = sQ + (size_t)lane * 26;
    double Jr[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) Jr[i] = rec[St::G0 + gp * 9 + i];
    const double dv = Sg[4], dv0 = Sg[5];
#pragma unroll
    for (int a = 0; a < NPE; ++a) {
      double g[3];
      const double da[3] = {sdN[gp][0][a], sdN[gp][1][a], sdN[gp][2][a]};
      grad_jinv(Jr, da, g);
      q[3 * a] = dv * g[0]; q[3 * a + 1] = dv * g[1]; q[3 * a + 2] = dv * g[2];
    }
    q[24] = dv0 * Sg[6];
    q[25] = dv0 * Sg[7];
    __syncthreads();
    const int a = gp;
    const double* q0 = sQ + (size_t)k * NGP * 26;
    double u3[3] = {0.0, 0.0, 0.0}, kpp = 0.0, fp = 0.0;
    for (int g2 = 0; g2 < NGP; ++g2) {
#pragma unroll
      for (int i = 0; i < 3; ++i) u3[i] += q0[g2 * 26 + 3 * a + i];
      fp -= q0[g2 * 26 + 25];
      kpp -= q0[g2 * 26 + 24];
    }
    if (writer) {
      double* W = rec + St::W0;
      double* R = rec + St::R0;
      const double fps = fp / kpp;
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        W[3 * a + i] = u3[i];
        R[3 * a + i] = R[3 * a + i] - u3[i] * fps;  // the lane that wrote node a's residual (a = gp)
        qKup[(size_t)e * 24 + 3 * a + i] = u3[i];
      }
      if (a == 0) {
        W[24] = kpp;
        W[25] = 0.0;
        qKpp[e] = kpp;
        qFp[e] = fp;
      }
    }
  }
  if (OUT_LDS) {
    __syncthreads();
    const int nv = nElem - base < EPW ? nElem - base : EPW;
    double2* dst = reinterpret_cast<double2*>(GS + (size_t)base * St::TS);
    for (int t = lane; t < nv * (St::TS / 2); t += 64) dst[t] = sOut[t];
  }
}

// Morton (Z-curve) code of a node: 21 bits per axis of its position in the bounding box, interleaved.
__device__ __forceinline__ uint64_t spread3(uint64_t v) {
  v &= 0x1fffffull;
  v = (v | (v << 32)) & 0x1f00000000ffffull;
  v = (v | (v << 16)) & 0x1f0000ff0000ffull;
  v = (v | (v << 8)) & 0x100f00f00f00f00full;
  v = (v | (v << 4)) & 0x10c30c30c30c30c3ull;
  v = (v | (v << 2)) & 0x1249249249249249ull;
  return v;
}

__global__ void k_morton(int32_t nNode, const double* __restrict__ X, double x0, double y0, double z0, double sx,
                         double sy, double sz, uint64_t* __restrict__ code, int32_t* __restrict__ id) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nNode) return;
  const double q[3] = {(X[3 * (size_t)i] - x0) * sx, (X[3 * (size_t)i + 1] - y0) * sy, (X[3 * (size_t)i + 2] - z0) * sz};
  uint64_t c[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) c[d] = (uint64_t)fmin(fmax(q[d] + 0.5, 0.0), 2097151.0);
  code[i] = spread3(c[0]) | (spread3(c[1]) << 1) | (spread3(c[2]) << 2);
  id[i] = i;
}

// Batched meshes (node ranges [moff[m], moff[m + 1])): mesh index above a Morton code of the node on
// its own mesh's lattice (origin mq[4m..4m+2], inverse spacing mq[4m+3], 15 bits per axis), so the
// order runs mesh by mesh even where the meshes overlap in space.
__global__ void k_morton_mesh(int32_t nNode, int32_t nmesh, const int32_t* __restrict__ moff,
                              const double* __restrict__ mq, const double* __restrict__ X,
                              uint64_t* __restrict__ code, int32_t* __restrict__ id) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nNode) return;
  int lo = 0, hi = nmesh - 1;  // last mesh whose first node is <= i
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (moff[mid] <= i) lo = mid; else hi = mid - 1;
  }
  const double* q = mq + 4 * lo;
  uint64_t c[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) c[d] = (uint64_t)fmin(fmax((X[3 * (size_t)i + d] - q[d]) * q[3] + 0.5, 0.0), 32767.0);
  code[i] = ((uint64_t)lo << 45) | spread3(c[0]) | (spread3(c[1]) << 1) | (spread3(c[2]) << 2);
  id[i] = i;
}

// XCD-aware chunk order: consecutive workgroups land on different XCDs (round robin), so each XCD is
// given a contiguous 1/8 of the work.  Work is laid out in Morton order of the nodes, so that share is a
// compact region of the mesh whose elements' staging records stay in that XCD's L2.
__device__ __forceinline__ int64_t xcd_chunk(int64_t b, int64_t nb) {
  const int64_t per = (nb + 7) / 8;
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

constexpr int kItemCap = 1024;   // kMaxAdj * NPE rounded up to a power of two (Tet10: 640 -> 1024)
constexpr int kGroupNodes = 8;   // nodes per layout group (consecutive in Morton order)

// Block ownership.  K is symmetric (K_e(a, b) = K_e(b, a)^T for every element: the Neo-Hookean
// tangent and the u-p couplings alike), so the pair of blocks (n, m), (m, n) is summed once, by its
// owner, which writes K_nm into n's rows and its transpose into m's rows.  n owns (n, m) if m is n,
// or m has no rows, or n < m.  Only nodes with rows own anything.
template <bool MIXED>
__device__ __forceinline__ bool has_rows(int info) {
  return (info & 7) || (MIXED && ((info >> 3) & 1));
}

// Build step 1 (one wavefront per node n):
//   nelA[nptr[n] + s]       = e * 16 + a: the elements around n in ascending id, a = local index of n in e
//   keys[nptr[n] * NPE + t] = (s0 << 23) | (j << 16) | (s << 4) | b for every element s around n and
//                             node b of it such that n owns the block pair (n, conn[e][b]), sorted:
//                             j = position of node conn[e][b] in the sorted coupled-node list of n (the
//                             column block of n's rows the (s, b) block lands in), s0 = first element of
//                             column j's run
//   cnt[n]                  = number of those items (0 for a node without rows)
template <int NPE, bool MIXED>
__global__ void __launch_bounds__(64) k_node_items(int32_t nNode, const int32_t* __restrict__ nptr,
                                                   const int32_t* __restrict__ nel, const int32_t* __restrict__ nbptr,
                                                   const int32_t* __restrict__ nb, const int32_t* __restrict__ conn,
                                                   const int32_t* __restrict__ nodeinfo, int32_t* __restrict__ nelA,
                                                   int32_t* __restrict__ keys, int64_t* __restrict__ cnt,
                                                   int32_t* __restrict__ overflow) {
  static_assert(kMaxAdj * NPE <= kItemCap, "item capacity");
  __shared__ int32_t sE[kMaxAdj], sNB[kMaxNB], sK[kItemCap];
  const int lane = threadIdx.x;
  const int n = blockIdx.x;
  if (n >= nNode) return;
  const int e0 = nptr[n], m = nptr[n + 1] - e0;
  const int b0 = nbptr[n], nnb = nbptr[n + 1] - b0;
  if (m > kMaxAdj || nnb > kMaxNB) {
    if (lane == 0) { atomicOr(overflow, 1); cnt[n] = 0; }
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
  const bool rows = has_rows<MIXED>(nodeinfo[n]);
  const int nall = rows ? m * NPE : 0;
  int P2 = 64;
  while (P2 < nall) P2 <<= 1;
  int own = 0;
  for (int t = lane; t < P2; t += 64) {
    int32_t key = 0x7fffffff;
    if (t < nall) {
      const int s = t / NPE, b = t % NPE;
      const int vb = conn[(size_t)sE[s] * NPE + b];
      if (vb == n || !has_rows<MIXED>(nodeinfo[vb]) || n < vb) {
        int lo = 0, hi = nnb;
        while (lo < hi) {
          const int mid = (lo + hi) >> 1;
          if (sNB[mid] < vb) lo = mid + 1; else hi = mid;
        }
        if (lo >= nnb || sNB[lo] != vb) atomicOr(overflow, 2);  // not a coupled node: inconsistent tables
        key = (min(lo, nnb - 1) << 16) | (s << 4) | b;
        ++own;
      }
    }
    sK[t] = key;
  }
  for (int off = 32; off > 0; off >>= 1) own += __shfl_xor(own, off, 64);
  const int nitems = own;
  __syncthreads();
  wave_sort(sK, P2, lane);
  // Runs (columns) are laid out in order of their first element, so a wavefront's items come from
  // few elements; within a run, items stay in element order.  key |= (first element of the run) << 23.
  if (lane == 0) {
    int first = 0;
    for (int t = 0; t < nitems; ++t) {
      if (t == 0 || ((sK[t] >> 16) & 127) != ((sK[t - 1] >> 16) & 127)) first = (sK[t] >> 4) & 63;
      sK[t] |= first << 23;
    }
  }
  __syncthreads();
  wave_sort(sK, P2, lane);
  for (int t = lane; t < nitems; t += 64) keys[(size_t)e0 * NPE + t] = sK[t];
  if (lane == 0) cnt[n] = nitems;
}

// Layout walk over layout group g (nodes order[8g .. 8g+7]) and each node's sorted items: items are
// placed one after the other, except that a run (one column block of one node) never crosses a
// 64-item wavefront.  Calls emit(position, node, item index, key) for every item and returns the
// group's length padded to a multiple of 64.
template <int NPE, typename Emit>
__device__ __forceinline__ int64_t group_walk(int64_t g, int32_t nNode, const int32_t* __restrict__ order,
                                              const int32_t* __restrict__ nptr, const int32_t* __restrict__ keys,
                                              const int64_t* __restrict__ cnt, Emit emit) {
  int64_t pos = 0;
  const int64_t q1 = (g + 1) * kGroupNodes < nNode ? (g + 1) * kGroupNodes : nNode;
  for (int64_t q = g * kGroupNodes; q < q1; ++q) {
    const int n = order[q];
    const int c = (int)cnt[n];
    const int32_t* K = keys + (size_t)nptr[n] * NPE;
    int jprev = -1;
    for (int t = 0; t < c; ++t) {
      const int32_t key = K[t];
      const int j = (key >> 16) & 127;
      if (j != jprev) {
        int L = 1;
        while (t + L < c && ((K[t + L] >> 16) & 127) == j) ++L;
        if ((pos & 63) + L > 64) pos = (pos + 63) & ~(int64_t)63;
        jprev = j;
      }
      emit(pos, n, t, key);
      ++pos;
    }
  }
  return (pos + 63) & ~(int64_t)63;
}

// Build step 2a (one lane per layout group): padded length of each group.
template <int NPE>
__global__ void k_group_len(int64_t ngroups, int32_t nNode, const int32_t* __restrict__ order,
                            const int32_t* __restrict__ nptr, const int32_t* __restrict__ keys,
                            const int64_t* __restrict__ cnt, int64_t* __restrict__ glen) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ngroups) return;
  glen[g] = group_walk<NPE>(g, nNode, order, nptr, keys, cnt, [](int64_t, int, int, int32_t) {});
}

// Build step 2b (one lane per layout group): the laid-out items,
//   fitem[goff[g] + q] = {e * 16 + a, (blk << 4) | b}  (element e around node n, a = n's local index
//   in e, blk = nbptr[n] + j the column block of n's rows node b of e lands in),
//   fnode[goff[g] + q] = n; padding slots keep {-1, -1} and -1.
template <int NPE>
__global__ void k_group_fill(int64_t ngroups, int32_t nNode, const int32_t* __restrict__ order,
                             const int32_t* __restrict__ nptr, const int32_t* __restrict__ nbptr,
                             const int32_t* __restrict__ keys, const int64_t* __restrict__ cnt,
                             const int32_t* __restrict__ nelA, const int64_t* __restrict__ goff,
                             int2* __restrict__ fitem, int32_t* __restrict__ fnode) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ngroups) return;
  const int64_t base = goff[g];
  group_walk<NPE>(g, nNode, order, nptr, keys, cnt, [&](int64_t pos, int n, int, int32_t key) {
    const int blk = nbptr[n] + ((key >> 16) & 127);
    fitem[base + pos] = make_int2(nelA[nptr[n] + ((key >> 4) & 63)], (blk << 4) | (key & 15));
    fnode[base + pos] = n;
  });
}

// Build step 2c (one lane per layout group): the group's wavefronts split into chunks of at most CW
// (the chunked matrix kernel's workgroups); with count, only the number of chunks.
__global__ void k_group_chunks(int64_t ngroups, int32_t CW, const int64_t* __restrict__ glen,
                               const int64_t* __restrict__ goff, const int64_t* __restrict__ coff,
                               int64_t* __restrict__ ccount, int32_t* __restrict__ cwave, int32_t* __restrict__ cnw) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ngroups) return;
  const int64_t waves = glen[g] / 64;
  const int64_t nch = (waves + CW - 1) / CW;
  if (ccount) { ccount[g] = nch; return; }
  for (int64_t k = 0; k < nch; ++k) {
    cwave[coff[g] + k] = (int32_t)(goff[g] / 64 + k * CW);
    cnw[coff[g] + k] = (int32_t)(waves - k * CW < CW ? waves - k * CW : CW);
  }
}

// Build step 2d (one wavefront per laid-out wavefront): the order in which the lanes compute the
// wavefront's items.  Lane L computes the item at run-order position fperm[64 w + L]; positions are
// ranked by (element, local index of the row node, local index of the column node), padding last.
// The run sums stay in run order; only which lane forms which block changes.  A vector load's
// address processing takes one pass per distinct 128-byte line among each four consecutive lanes
// (measured on MI355X: 16-byte loads take 64 cycles per wavefront when the four lanes of every quad
// hit four lines, 16 when each quad stays in one line); in run order the lanes of a quad read four
// different elements' records, in this order mostly the same element's.
__global__ void __launch_bounds__(64) k_wave_perm(int64_t nwaves, const int2* __restrict__ fitem,
                                                  uint8_t* __restrict__ fperm) {
  const int64_t w = blockIdx.x;
  if (w >= nwaves) return;
  const int lane = threadIdx.x;
  const int2 it = fitem[w * 64 + lane];
  const long long key = it.x >= 0 ? ((long long)(it.x >> 4) << 8) | ((it.x & 15) << 4) | (it.y & 15)
                                   : 0x7fffffffffffffffLL;
  int rank = 0;
  for (int j = 0; j < 64; ++j) {
    const long long kj = __shfl(key, j, 64);
    rank += (kj < key || (kj == key && j < lane)) ? 1 : 0;
  }
  fperm[w * 64 + rank] = (uint8_t)lane;
}

// Build step 3b (one wavefront per node n): for each column block (n, j) with coupled node m, the
// entry of block (m, n) of m's rows, which receives its transpose when n owns the pair, is copied
// next to (n, j)'s own entry, so that a run's head reads both with one load; {-1, 0, ..} when the
// pair has no mirror (m is n, or m has no rows).  btab holds 2 BT ints per block.
template <bool MIXED>
__global__ void __launch_bounds__(64) k_mirror(int32_t nNode, const int32_t* __restrict__ nbptr,
                                               const int32_t* __restrict__ nb, const int32_t* __restrict__ nodeinfo,
                                               int32_t* __restrict__ btab) {
  constexpr int BT = MIXED ? 4 : 2;
  const int n = blockIdx.x;
  if (n >= nNode) return;
  const int b0 = nbptr[n], nnb = nbptr[n + 1] - b0;
  for (int j = threadIdx.x; j < nnb; j += 64) {
    const int mnode = nb[b0 + j];
    int out = -1;
    if (mnode != n && has_rows<MIXED>(nodeinfo[mnode])) {
      int lo = nbptr[mnode], hi = nbptr[mnode + 1];
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (nb[mid] < n) lo = mid + 1; else hi = mid;
      }
      out = (lo < nbptr[mnode + 1] && nb[lo] == n) ? lo : -1;
    }
    int32_t* dst = btab + 2 * BT * (size_t)(b0 + j) + BT;
#pragma unroll
    for (int q = 0; q < BT; ++q) dst[q] = out >= 0 ? btab[2 * BT * (size_t)out + q] : (q == 0 ? -1 : 0);
  }
}

// Build step 3 (one wavefront per node n): where each column block (n, j) of n's rows lives in the CSR,
// the first BT of the block's 2 BT ints (the mirror's entry follows, step 3b).
//   Hex8:      {A, (La << 9) | flags}                         (BT = 2)
//   Tet10/P1:  {A, Bp, (D << 16) | La, flags}                 (BT = 4)
//   flags = (self << 8) | (isPn << 7) | (isPb << 6) | (mn << 3) | mb
// A = start of n's first free u row + column offset of j's u dofs (-1 if n has no free u dof),
// Bp = the same in n's pressure row (-1 if n carries no pressure), D = offset of j's pressure column
// relative to A (same in every row of n), La = length of n's rows, mn / mb = free masks of n / j,
// self: j is n itself (its column run also sums n's right-hand side).
template <bool MIXED>
__global__ void __launch_bounds__(64) k_block_table(int32_t nNode, const int32_t* __restrict__ nbptr,
                                                    const int32_t* __restrict__ nb,
                                                    const int32_t* __restrict__ nodeinfo,
                                                    const int32_t* __restrict__ nda, const int32_t* __restrict__ idp,
                                                    int32_t dispDOF, const int32_t* __restrict__ rowptr,
                                                    int32_t* __restrict__ btab) {
  const int lane = threadIdx.x;
  const int n = blockIdx.x;
  if (n >= nNode) return;
  const int info_n = nodeinfo[n];
  const int mn = info_n & 7, La = info_n >> 4;
  const bool isPn = MIXED && ((info_n >> 3) & 1);
  const int b0 = nbptr[n], nnb = nbptr[n + 1] - b0;
  const int fr = (mn & 1) ? 0 : (mn & 2) ? 1 : (mn & 4) ? 2 : -1;
  const int RU = fr >= 0 ? rowptr[nda[3 * (size_t)n + fr]] : -1;
  const int RP = isPn ? rowptr[dispDOF + idp[n]] : -1;
  int Ucols = 0;
  for (int c = 0; c < nnb; c += 64) {
    const int j = c + lane;
    int nu = j < nnb ? __popc(nodeinfo[nb[b0 + j]] & 7) : 0;
    for (int off = 32; off > 0; off >>= 1) nu += __shfl_xor(nu, off, 64);
    Ucols += nu;
  }
  int cu0 = 0, cp0 = 0;
  for (int c = 0; c < nnb; c += 64) {
    const int j = c + lane;
    int inf = 0, nu = 0, np = 0;
    if (j < nnb) {
      inf = nodeinfo[nb[b0 + j]];
      nu = __popc(inf & 7);
      np = MIXED ? ((inf >> 3) & 1) : 0;
    }
    int iu = nu, ip = np;
    for (int off = 1; off < 64; off <<= 1) {
      const int tu = __shfl_up(iu, off, 64), tp = __shfl_up(ip, off, 64);
      if (lane >= off) { iu += tu; ip += tp; }
    }
    if (j < nnb) {
      const int upos = cu0 + iu - nu, ppos = cp0 + ip - np;
      const int mb = inf & 7;
      const size_t blk = (size_t)b0 + j;
      const int self = nb[b0 + j] == n;
      if (MIXED) {
        const int isPb = (inf >> 3) & 1;
        btab[8 * blk] = RU >= 0 ? RU + upos : -1;
        btab[8 * blk + 1] = RP >= 0 ? RP + upos : -1;
        btab[8 * blk + 2] = ((Ucols + ppos - upos) << 16) | La;
        btab[8 * blk + 3] = (self << 8) | ((int)isPn << 7) | (isPb << 6) | (mn << 3) | mb;
      } else {
        btab[4 * blk] = RU >= 0 ? RU + upos : -1;
        btab[4 * blk + 1] = (La << 9) | (self << 8) | (mn << 3) | mb;
      }
    }
    cu0 += __shfl(iu, 63, 64);
    cp0 += __shfl(ip, 63, 64);
  }
}

// Build step 4: near[n] = 1 if an element around n has a node with a prescribed dof (the first-
// iteration lift of n's rows is non-zero only there).
template <int NPE>
__global__ void k_near(int32_t nNode, const int32_t* __restrict__ nptr, const int32_t* __restrict__ nel,
                       const int32_t* __restrict__ conn, const int32_t* __restrict__ nodeinfo,
                       int32_t* __restrict__ near) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nNode) return;
  int f = 0;
  for (int s = nptr[n]; s < nptr[n + 1] && !f; ++s) {
    const int32_t* c = conn + (size_t)nel[s] * NPE;
    for (int b = 0; b < NPE; ++b)
      if ((nodeinfo[c[b]] & 7) != 7) f = 1;
  }
  near[n] = f;
}

// Gauss points gp0 .. gp0 + NG - 1 of one (element, node b) item of node n's rows, added to v: the
// 3x3 block K_{n b} (v[0..8]) and, for the mixed element, Kup(n, b) (v[9..11]), Kpu(n, b)
// (v[12..14]) and Kpp(n, b) (v[15]).  G, Bv, S point at the first of those Gauss points' g data
// (stride NPE * GW; FROM_J: the inverse Jacobians of a Hex8 record, stride 9, g formed here), b
// components (stride 6; compact records) and scalars (stride SS), in a staging record in global
// memory or in a copy in LDS.
template <int NPE, int NGP, bool MIXED, int NV, int NG, int UNROLL, bool FROM_J = false>
__device__ __forceinline__ void item_gps(const double* G, const double* Bv, const double* S0p, int gp0, int a, int b,
                                         double v[NV]) {
  using St = Stage<NPE, NGP, MIXED>;
#pragma unroll UNROLL
  for (int gp = 0; gp < NG; ++gp) {
    const double* Gp = G + gp * (FROM_J ? 9 : NPE * St::GW);
    const double* S = S0p + gp * St::SS;
    double ga[3], ba[3], gb[3], bb[3];
    if (St::COMPACT && FROM_J) {  // g from the record's inverse Jacobian
      double da[3], db[3];
      hex8_dN_node(gp0 + gp, a, da);
      hex8_dN_node(gp0 + gp, b, db);
      grad_jinv(Gp, da, ga);
      grad_jinv(Gp, db, gb);
    } else if (St::COMPACT) {  // g formed in LDS, 3 doubles per node
#pragma unroll
      for (int i = 0; i < 3; ++i) { ga[i] = Gp[a * 3 + i]; gb[i] = Gp[b * 3 + i]; }
    } else {  // Tet10 records: g padded to 4 doubles, two 16-byte loads per node
      const double2* ga2 = reinterpret_cast<const double2*>(Gp + a * 4);
      const double2* gb2 = reinterpret_cast<const double2*>(Gp + b * 4);
      const double2 p0 = ga2[0], p1 = ga2[1];
      const double2 q0 = gb2[0], q1 = gb2[1];
      ga[0] = p0.x; ga[1] = p0.y; ga[2] = p1.x;
      gb[0] = q0.x; gb[1] = q0.y; gb[2] = q1.x;
    }
    {
      const double2* B2 = reinterpret_cast<const double2*>(Bv + gp * 6);
      const double2 c0 = B2[0], c1 = B2[1], c2 = B2[2];
      const double B6[6] = {c0.x, c0.y, c1.x, c1.y, c2.x, c2.y};
      bsym_times(B6, ga, ba);
      bsym_times(B6, gb, bb);
    }
    tangent_gp<!MIXED>(ga, ba, gb, bb, S[0], S[1], S[2], S[3], MIXED ? 0.0 : S[4], v);
    if (St::P1) {
      double Np[4];
      tet_Np(gp0 + gp, Np);
      const double dv = S[4], dv0 = S[5], th = S[6];
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

// One (element, node b) item of node n's rows over all Gauss points, from the element's staging
// record at rec, UNROLL Gauss points at a time.
template <int NPE, int NGP, bool MIXED, int NV, int UNROLL = 2>
__device__ __forceinline__ void item_block(const double* rec, int a, int b, double v[NV]) {
  using St = Stage<NPE, NGP, MIXED>;
#pragma unroll
  for (int q = 0; q < NV; ++q) v[q] = 0.0;
  item_gps<NPE, NGP, MIXED, NV, NGP, UNROLL, St::COMPACT>(rec + St::G0, rec + St::B0, rec + St::S0, 0, a, b, v);
  if (St::P0) add_condensed(rec + St::W0, a, b, v);
}

// Node a's residual rows of one element (v[0..2]) and, for the mixed element, its pressure row (v[3])
template <int NPE, int NGP, bool MIXED>
__device__ __forceinline__ void item_residual(const double* rec, int a, double* v) {
  using St = Stage<NPE, NGP, MIXED>;
#pragma unroll
  for (int i = 0; i < 3; ++i) v[i] = rec[St::R0 + a * 3 + i];
  if (St::P1 && a < 4) {
    double fp = 0.0;  // Fp_a -= (dv0 Np_a) resip   (ElementBase2.cpp:510, 522)
#pragma unroll
    for (int gp = 0; gp < NGP; ++gp) {
      double Np[4];
      tet_Np(gp, Np);
      const double Na = a == 0 ? Np[0] : a == 1 ? Np[1] : a == 2 ? Np[2] : Np[3];
      fp -= (rec[St::S0 + gp * St::SS + 5] * Na) * rec[St::S0 + gp * St::SS + 7];
    }
    v[3] = fp;
  }
}

struct FlatArgs {
  int64_t nwaves;
  const int2* fitem;        // {e * 16 + a, (j << 10) | (s << 4) | b}, {-1, -1} in padding slots
  const int32_t* fnode;     // node whose rows each laid-out item belongs to (-1 in padding slots)
  const uint8_t* fperm;     // per wavefront, lane L computes the item at run-order position fperm[L]
  const int32_t* nbptr;
  const int32_t* btab;      // per column block: its CSR location, then its mirror's (2 BT ints)
  const int32_t* nda;
  const int32_t* idp;
  int32_t dispDOF;
  const double* GS;
  double* vals;
  double* rhs;
  // chunks of at most CW wavefronts of one layout group (chunked matrix kernel): first wavefront and
  // count, distinct elements of the chunk's items in ascending order (ECAP slots per chunk), their
  // number, and each item's slot
  int64_t nchunks;
  int32_t ecap;             // chunks with at most ecap (<= ECAP) distinct elements are staged
  const int32_t* cwave;
  const int32_t* cnw;
  const int32_t* ctab;
  const int32_t* ccnt;
  const uint8_t* fslot;
};

// Three consecutive CSR values x0 x1 x2 at row[0..2] (the three free dofs of one column node) as one
// 16-byte store on the 16-byte-aligned pair and one 8-byte store: two memory requests instead of
// three.  row is 8-byte aligned (an offset into the value array).
__device__ __forceinline__ void store3(double* row, double x0, double x1, double x2) {
  const bool odd = (reinterpret_cast<uintptr_t>(row) & 8) != 0;
  double2* pair = reinterpret_cast<double2*>(row + (odd ? 1 : 0));
  *pair = odd ? make_double2(x1, x2) : make_double2(x0, x1);
  row[odd ? 0 : 2] = odd ? x0 : x2;
}

// Writes a summed column block: K_nm (acc) into node n's rows at block blk and, through the mirror
// table, its transpose into m's rows.  For the mixed element, acc[9..11] = Kup(n, m) (n's u rows,
// m's pressure column), acc[12..14] = Kpu(n, m), acc[15] = Kpp(n, m); m's rows receive
// Kup(m, n) = Kpu(n, m)^T, Kpu(m, n) = Kup(n, m)^T, Kpp(m, n) = Kpp(n, m).
template <bool MIXED>
__device__ __forceinline__ void write_block(const FlatArgs& F, const int* bt, const double* acc) {
  constexpr int BT = MIXED ? 4 : 2;
  double* __restrict__ V = F.vals;
  {
    const int masks = MIXED ? bt[3] : (bt[1] & 0x1ff);
    const int La = MIXED ? (bt[2] & 0xffff) : (bt[1] >> 9);
    const int mn = (masks >> 3) & 7, mb = masks & 7;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      if (!((mn >> i) & 1)) continue;
      double* row = V + bt[0] + rank3(mn, i) * La;
      if (mb == 7) {
        store3(row, acc[3 * i], acc[3 * i + 1], acc[3 * i + 2]);
      } else {
#pragma unroll
        for (int k = 0; k < 3; ++k)
          if ((mb >> k) & 1) row[rank3(mb, k)] = acc[3 * i + k];
      }
      if (MIXED && ((masks >> 6) & 1)) row[bt[2] >> 16] = acc[9 + i];
    }
    if (MIXED && ((masks >> 7) & 1)) {
      double* row = V + bt[1];
      if (mb == 7) {
        store3(row, acc[12], acc[13], acc[14]);
      } else {
#pragma unroll
        for (int k = 0; k < 3; ++k)
          if ((mb >> k) & 1) row[rank3(mb, k)] = acc[12 + k];
      }
      if ((masks >> 6) & 1) row[bt[2] >> 16] = acc[15];
    }
  }
  const int* mt = bt + BT;  // the mirror block (m, n): rows of m, columns of n (all masks 0 if none)
  const int masks = MIXED ? mt[3] : (mt[1] & 0x1ff);
  const int La = MIXED ? (mt[2] & 0xffff) : (mt[1] >> 9);
  const int mm = (masks >> 3) & 7, mnn = masks & 7;  // free masks of m (rows) and n (columns)
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    if (!((mm >> k) & 1)) continue;
    double* row = V + mt[0] + rank3(mm, k) * La;
    if (mnn == 7) {
      store3(row, acc[k], acc[3 + k], acc[6 + k]);
    } else {
#pragma unroll
      for (int i = 0; i < 3; ++i)
        if ((mnn >> i) & 1) row[rank3(mnn, i)] = acc[3 * i + k];
    }
    if (MIXED && ((masks >> 6) & 1)) row[mt[2] >> 16] = acc[12 + k];
  }
  if (MIXED && ((masks >> 7) & 1)) {
    double* row = V + mt[1];
    if (mnn == 7) {
      store3(row, acc[9], acc[10], acc[11]);
    } else {
#pragma unroll
      for (int i = 0; i < 3; ++i)
        if ((mnn >> i) & 1) row[rank3(mnn, i)] = acc[9 + i];
    }
    if ((masks >> 6) & 1) row[mt[2] >> 16] = acc[15];
  }
}

// Column runs of one wavefront's items, after every item lane has stored its values at part[t * NT]
// and its block index (nbptr[n] + j, unique per node and column) at col[t]: the first lane of each
// run (its head) sums the run in lane order, which is element order, and writes the block and its
// mirror; the self column's head also writes n's right-hand-side rows.  Every CSR value is written
// exactly once.
// An item lane's column-block table entry (its block's CSR location and its mirror's), loaded
// before the lane's block is computed so that the run head has it at hand.
template <bool MIXED>
struct BlockEntry {
  int v[MIXED ? 8 : 4];
  __device__ __forceinline__ void load(const int32_t* btab, int blk) {
    const int4* p = reinterpret_cast<const int4*>(btab) + (MIXED ? 2 : 1) * (size_t)blk;
    const int4 x = p[0];
    v[0] = x.x; v[1] = x.y; v[2] = x.z; v[3] = x.w;
    if (MIXED) {
      const int4 y = p[MIXED ? 1 : 0];
      v[MIXED ? 4 : 0] = y.x; v[MIXED ? 5 : 1] = y.y; v[MIXED ? 6 : 2] = y.z; v[MIXED ? 7 : 3] = y.w;
    }
  }
  __device__ __forceinline__ int flags() const { return MIXED ? v[3] : (v[1] & 0x1ff); }
};

// Column runs of one wavefront's items, after every item lane has stored its values at part[t * NT]
// and its block index (nbptr[n] + j, unique per node and column) at col[t]: the first lane of each
// run (its head) sums the run in lane order, which is element order, and writes the block and its
// mirror (bt: the head's table entry); the self column's head also writes n's right-hand-side rows
// (n: the head's node).  Every CSR value is written exactly once.
template <bool MIXED, int NV, int NT>
__device__ __forceinline__ void run_heads(const FlatArgs& F, int n, int lane, int blk, const double v[NT],
                                          const int32_t* col, const double* part, const BlockEntry<MIXED>& bt) {
  if (lane > 0 && col[lane - 1] == blk) return;
  double acc[NT];
#pragma unroll
  for (int q = 0; q < NT; ++q) acc[q] = v[q];
  for (int t2 = lane + 1; t2 < 64 && col[t2] == blk; ++t2)
#pragma unroll
    for (int q = 0; q < NT; ++q) acc[q] += part[t2 * NT + q];
  const int flags = bt.flags();
  if ((flags >> 8) & 1) {
    // self block K_nn: its (i, k) and (k, i) entries (and n's u-p and p-u couplings) agree to
    // rounding; their mean makes the assembled matrix exactly symmetric
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int k = i + 1; k < 3; ++k) acc[3 * i + k] = acc[3 * k + i] = 0.5 * (acc[3 * i + k] + acc[3 * k + i]);
    if (MIXED)
#pragma unroll
      for (int i = 0; i < 3; ++i) acc[9 + i] = acc[12 + i] = 0.5 * (acc[9 + i] + acc[12 + i]);
    // n's right-hand side, its elements' residuals in element order
    const int mn = (flags >> 3) & 7;
#pragma unroll
    for (int i = 0; i < 3; ++i)
      if ((mn >> i) & 1) F.rhs[F.nda[3 * (size_t)n + i]] = acc[NV + i];
    if (MIXED && ((flags >> 7) & 1)) F.rhs[F.dispDOF + F.idp[n]] = acc[NT - 1];
  }
  write_block<MIXED>(F, bt.v, acc);
}

// Matrix kernel, one wavefront per 64 laid-out items, every item reading its element's staging record
// from global memory.  Lane L forms the block of the item at run-order position fperm[L] (items of
// one element on consecutive lanes) and stores it at that position in LDS.  The runs are then summed
// and written by lanes 0 .. nruns - 1, lane r taking the r-th run: the stores of a wavefront come
// from consecutive lanes rather than from run heads spread over the wavefront (a store's address
// processing, like a load's, takes one pass per distinct line among each four consecutive active
// lanes).  The sums are those of the run heads, in run order, so the values are the same bits.
template <int NPE, int NGP, bool MIXED>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) k_flat(FlatArgs F) {
  using St = Stage<NPE, NGP, MIXED>;
  constexpr bool P1 = St::P1;        // pressure rows (Tet10/P1)
  constexpr int NV = P1 ? 16 : 9;    // uu 9 | up 3 | pu 3 | pp 1
  constexpr int NR = P1 ? 4 : 3;     // right-hand-side rows of a node (self items)
  constexpr int NT = NV + NR;
  constexpr int BN = P1 ? 8 : 4;     // ints of a column-block table entry
  __shared__ int32_t sCol[64];
  __shared__ int32_t sStart[64];
  __shared__ double sPart[64 * NT];
  const int lane = threadIdx.x;
  const int64_t w = xcd_chunk(blockIdx.x, F.nwaves);
  if (w >= F.nwaves) return;
  const int2 it = F.fitem[w * 64 + lane];  // run-order position lane
  const int pc = F.fperm ? (int)F.fperm[w * 64 + lane] : lane;
  const int2 ic = make_int2(__shfl(it.x, pc, 64), __shfl(it.y, pc, 64));  // the item this lane forms
  const bool live = it.x >= 0;
  const int blk = live ? (it.y >> 4) : -1;
  const int pblk = __shfl_up(blk, 1, 64);
  const bool head = live && (lane == 0 || pblk != blk);  // first position of its run
  const bool self = head && (it.x & 15) == (it.y & 15);
  const bool clive = ic.x >= 0;
  const int e = ic.x >> 4, a = ic.x & 15, b = ic.y & 15;
  const bool cself = clive && a == b;  // the element's own item of n carries its residual rows
  double v[NT];
#pragma unroll
  for (int q = 0; q < NT; ++q) v[q] = 0.0;
  const double* rec = F.GS + (size_t)(clive ? e : 0) * St::TS;
  if (clive) item_block<NPE, NGP, MIXED, NV>(rec, a, b, v);
  if (cself) item_residual<NPE, NGP, MIXED>(rec, a, v + NV);
  // each run's table entry and node, loaded by its head before the barrier (registers are the
  // limit during the block computation)
  BlockEntry<P1> bt{};
  if (head) bt.load(F.btab, blk);
  const int n = self ? F.fnode[w * 64 + lane] : 0;
  sCol[lane] = blk;
#pragma unroll
  for (int q = 0; q < NT; ++q) sPart[pc * NT + q] = v[q];
  const unsigned long long hm = __ballot(head);
  if (head) sStart[__popcll(hm & ((1ull << lane) - 1))] = lane;
  __syncthreads();
  const int nruns = __popcll(hm);
  const int st = lane < nruns ? sStart[lane] : 0;
  BlockEntry<P1> hb;
#pragma unroll
  for (int k = 0; k < BN; ++k) hb.v[k] = __shfl(bt.v[k], st, 64);
  const int hn = __shfl(n, st, 64);
  if (lane < nruns) run_heads<P1, NV, NT>(F, hn, st, sCol[st], sPart + st * NT, sCol, sPart, hb);
}

// Chunked matrix kernel: one workgroup per chunk (the wavefronts of one layout group of 8 nodes in
// Morton order; for Hex8 on a structured grid an aligned 2x2x2 block of nodes and its 27 elements).
// The chunk's distinct elements' Gauss-point data (g, b, scalars) is put in LDS — the verbatim part
// of the records with coalesced 16-byte loads, g formed from the records' inverse Jacobians — and
// every item then reads LDS: each record is read from L2 once per chunk rather than once per item
// lane, and the per-lane gathers that bound the per-wavefront kernel on address processing become
// LDS reads.  A chunk with more than ECAP distinct elements (or more than the problem's cap) reads
// the records directly.  Workgroups of 8 wavefronts although a Hex8 group fills 5: measured per
// Hex8 128^3 assembly, the kernel takes 13.6, 13.7, 8.3 and 7.0 ms with workgroups of 5, 6, 7 and 8
// wavefronts for the same chunks.
#ifndef MPF_CHUNK_WAVES
#define MPF_CHUNK_WAVES 8
#endif
#ifndef MPF_CHUNK_ELEMS
#define MPF_CHUNK_ELEMS 32
#endif
// LDS slot stride padded by 6 doubles (Hex8: 286 doubles, 2288 bytes per slot) so that the slots'
// rows spread over the LDS banks: measured per Hex8 128^3 assembly, 10.24 ms unpadded, 9.97 ms
// with 2, 9.93 ms with 6, 9.98 ms with 10
#ifndef MPF_LDS_PAD
#define MPF_LDS_PAD 6
#endif
// ECAP: the chunk table's slots per chunk; SCAP (<= ECAP): the most distinct elements staged in LDS
// (Q1/P0 records are longer, so fewer fit where two workgroups share a CU's LDS).
template <int NPE, int NGP, bool MIXED, int CW, int ECAP, int SCAP = ECAP>
__global__ void __launch_bounds__(64 * CW) k_flat_chunk(FlatArgs F) {
  using St = Stage<NPE, NGP, MIXED>;
  constexpr bool P1 = St::P1;
  constexpr int NV = P1 ? 16 : 9;
  constexpr int NR = P1 ? 4 : 3;
  constexpr int NT = NV + NR;
  // LDS slot: g of every node and Gauss point (Hex8: formed here from the record's inverse
  // Jacobians), then the record's B and S copied verbatim (Tet10/P1: the record's G and S verbatim)
  constexpr int GX = St::COMPACT ? NGP * NPE * 3 : 0;   // g formed in LDS (doubles per slot)
  constexpr int CP0 = St::COMPACT ? St::B0 : 0;          // record part copied verbatim: [CP0, SL)
  constexpr int CP2 = (St::SL - CP0) / 2;                // its 16-byte groups
  constexpr int SLP = GX + (St::SL - CP0) + MPF_LDS_PAD; // LDS stride per element slot (doubles)
  static_assert(SLP % 2 == 0 && GX % 2 == 0 && CP0 % 2 == 0, "16-byte slots");
  static_assert(SCAP <= ECAP, "staged elements fit the chunk table");
  constexpr int STG = SCAP * SLP, PART = CW * 64 * NT;
  constexpr int LDSD = STG > PART ? STG : PART;  // record copies, then the run sums
  static_assert(ECAP <= 255, "slot tables");
  __shared__ double2 sMem[LDSD / 2];
  __shared__ int32_t sCol[CW * 64];
  __shared__ double sdN[St::COMPACT ? NGP * NPE * 3 : 1];
  const int tid = threadIdx.x, wid = tid >> 6, lane = tid & 63;
  const int64_t c = xcd_chunk(blockIdx.x, F.nchunks);
  if (c >= F.nchunks) return;  // whole workgroup, before any barrier
  const int cnt = F.ccnt[c];
  const bool staged = cnt <= SCAP && cnt <= F.ecap;
  // per-lane loads that need nothing from LDS, issued before the staging so that their latency
  // overlaps it: the item, its slot, its block's table entry, and (self items) the node and the
  // element's residual rows
  const int64_t w = (int64_t)F.cwave[c] + wid;
  const bool wv = wid < F.cnw[c];
  const int2 it = wv ? F.fitem[w * 64 + lane] : make_int2(-1, -1);
  const bool live = it.x >= 0;
  const int blk = live ? (it.y >> 4) : 0;
  const int e = it.x >> 4, a = it.x & 15, b = it.y & 15;
  const bool self = live && a == b;
  const int slot = live && staged ? F.fslot[w * 64 + lane] : 0;
  BlockEntry<P1> bt;
  bt.load(F.btab, blk);
  const int n = self ? F.fnode[w * 64 + lane] : 0;
  const double* rec = F.GS + (size_t)(live ? e : 0) * St::TS;
  double v[NT];
#pragma unroll
  for (int q = 0; q < NT; ++q) v[q] = 0.0;
  if (self) item_residual<NPE, NGP, MIXED>(rec, a, v + NV);
  if (St::COMPACT && tid < NGP * NPE) hex8_dN_node(tid / NPE, tid % NPE, &sdN[3 * tid]);
  __syncthreads();
  if (staged) {
    const int32_t* ct = F.ctab + c * ECAP;
    constexpr int BS = 64 * CW;
    // Hex8: one lane per (element slot, Gauss point) forms g of the 8 nodes from the record's inverse
    // Jacobian; its loads are issued before the verbatim copy, so both share one memory latency
    static_assert(!St::COMPACT || SCAP * NGP <= BS, "one (slot, Gauss point) per lane");
    const bool jl = St::COMPACT && tid < cnt * NGP;
    double Jr[9];
    if (St::COMPACT) {
      const double* J = F.GS + (size_t)ct[jl ? tid / NGP : 0] * St::TS + St::G0 + (tid % NGP) * 9;
#pragma unroll
      for (int i = 0; i < 9; ++i) Jr[i] = jl ? J[i] : 0.0;
    }
    // verbatim part: four 16-byte loads in flight per lane; indices past the end re-read the last group
    const int tot = cnt * CP2;
    auto ld = [&](int t) {
      t = t < tot ? t : tot - 1;
      const int k = t / CP2;
      return reinterpret_cast<const double2*>(F.GS + (size_t)ct[k] * St::TS + CP0)[t - k * CP2];
    };
    auto dst = [&](int t) {
      const int k = t / CP2;
      return k * (SLP / 2) + GX / 2 + (t - k * CP2);
    };
    for (int t0 = tid; t0 < tot; t0 += 4 * BS) {
      const double2 r0 = ld(t0), r1 = ld(t0 + BS), r2 = ld(t0 + 2 * BS), r3 = ld(t0 + 3 * BS);
      sMem[dst(t0)] = r0;
      if (t0 + BS < tot) sMem[dst(t0 + BS)] = r1;
      if (t0 + 2 * BS < tot) sMem[dst(t0 + 2 * BS)] = r2;
      if (t0 + 3 * BS < tot) sMem[dst(t0 + 3 * BS)] = r3;
    }
    if (jl) {
      const int k = tid / NGP, gp = tid % NGP;
      double* o = reinterpret_cast<double*>(sMem) + k * SLP + gp * NPE * 3;
#pragma unroll
      for (int a = 0; a < NPE; ++a) {
        double g[3];
        grad_jinv(Jr, &sdN[3 * (gp * NPE + a)], g);
        o[3 * a] = g[0]; o[3 * a + 1] = g[1]; o[3 * a + 2] = g[2];
      }
    }
  }
  __syncthreads();
  if (live) {
    // one Gauss point at a time: LDS latency needs no unrolling to hide, and registers are the limit
    if (staged) {
      const double* sl = reinterpret_cast<const double*>(sMem) + slot * SLP;
      item_gps<NPE, NGP, MIXED, NV, NGP, 1>(sl, sl + GX + (St::B0 - CP0), sl + GX + (St::S0 - CP0), 0, a, b, v);
      if (St::P0) add_condensed(sl + GX + (St::W0 - CP0), a, b, v);
    } else {
      item_gps<NPE, NGP, MIXED, NV, NGP, 1, St::COMPACT>(rec + St::G0, rec + St::B0, rec + St::S0, 0, a, b, v);
      if (St::P0) add_condensed(rec + St::W0, a, b, v);
    }
  }
  __syncthreads();  // the record copies become the run-sum area
  double* const part = reinterpret_cast<double*>(sMem) + wid * 64 * NT;
  sCol[tid] = live ? blk : -1;
#pragma unroll
  for (int q = 0; q < NT; ++q) part[lane * NT + q] = v[q];
  __syncthreads();
  if (live) run_heads<P1, NV, NT>(F, n, lane, blk, v, sCol + wid * 64, part, bt);
}

// Chunk tables (once per pattern), one workgroup of 512 lanes per chunk of at most 8 wavefronts: the
// chunk's distinct element ids in ascending order (bitonic sort of its items' elements in LDS, then a
// compaction), their number, and each item's slot among them.
template <int ECAP>
__global__ void __launch_bounds__(512) k_chunk_table(int64_t nchunks, const int32_t* __restrict__ cwave,
                                                     const int32_t* __restrict__ cnw, const int2* __restrict__ fitem,
                                                     int32_t* __restrict__ ctab, int32_t* __restrict__ ccnt,
                                                     uint8_t* __restrict__ fslot) {
  constexpr int N = 512, NW = N / 64;
  __shared__ int32_t key[N], uq[N], wsum[NW];
  const int tid = threadIdx.x, wid = tid >> 6, lane = tid & 63;
  const int64_t c = blockIdx.x;
  if (c >= nchunks) return;
  const bool valid = wid < cnw[c];
  const int64_t i = ((int64_t)cwave[c] + wid) * 64 + lane;
  const int2 it = valid ? fitem[i] : make_int2(-1, -1);
  const int32_t none = 0x7fffffff;
  const int32_t e = it.x >= 0 ? (it.x >> 4) : none;
  key[tid] = e;
  __syncthreads();
  for (int k = 2; k <= N; k <<= 1)
    for (int h = k >> 1; h > 0; h >>= 1) {
      const int p = tid ^ h;
      if (p > tid) {
        const int32_t x = key[tid], y = key[p];
        if ((x > y) == ((tid & k) == 0)) { key[tid] = y; key[p] = x; }
      }
      __syncthreads();
    }
  const int32_t x = key[tid];
  const bool u = x != none && (tid == 0 || key[tid - 1] != x);
  const uint64_t bal = __ballot(u);
  if (lane == 0) wsum[wid] = __popcll(bal);
  __syncthreads();
  int off = 0, tot = 0;
  for (int q = 0; q < NW; ++q) {
    if (q < wid) off += wsum[q];
    tot += wsum[q];
  }
  if (u) uq[off + __popcll(bal & ((1ull << lane) - 1ull))] = x;
  __syncthreads();
  int slot = 0;
  if (e != none && tot <= ECAP) {
    int lo = 0, hi = tot - 1;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (uq[mid] < e) lo = mid + 1; else hi = mid;
    }
    slot = lo;
  }
  if (valid) fslot[i] = (uint8_t)slot;
  if (tid < ECAP) ctab[c * ECAP + tid] = tid < tot ? uq[tid] : 0;
  if (tid == 0) ccnt[c] = tot;
}

struct RhsArgs {
  int32_t nNode;
  const int32_t* nptr;
  const int32_t* nelA;
  const int32_t* nodeinfo;
  const int32_t* nda;
  const int32_t* idp;
  int32_t dispDOF;
  const double* GS;
  double* rhs;
  int32_t lift_only;          // the residual sums are already in rhs (written by the matrix kernel)
};

// Right-hand-side kernel (one lane per node): n's rows of the element residuals, summed in element
// order.  With lift_only the matrix kernel has written them and only nodes in no element (whose rows
// no item wrote) get their zeros here.
template <int NPE, int NGP, bool MIXED>
__global__ void __launch_bounds__(256) k_rhs(RhsArgs R) {
  using St = Stage<NPE, NGP, MIXED>;
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= R.nNode) return;
  const int info = R.nodeinfo[n];
  const int mn = info & 7;
  const bool isPn = St::P1 && ((info >> 3) & 1);
  if (!mn && !isPn) return;
  const int e0 = R.nptr[n], m = R.nptr[n + 1] - e0;
  if (R.lift_only && m > 0) return;
  double r[4] = {0.0, 0.0, 0.0, 0.0};
  for (int s = 0; s < m; ++s) {
    const int ea = R.nelA[e0 + s];
    double t[4] = {0.0, 0.0, 0.0, 0.0};
    item_residual<NPE, NGP, MIXED>(R.GS + (size_t)(ea >> 4) * St::TS, ea & 15, t);
#pragma unroll
    for (int i = 0; i < 3; ++i) r[i] += t[i];
    if (isPn) r[3] += t[3];
  }
#pragma unroll
  for (int i = 0; i < 3; ++i)
    if ((mn >> i) & 1) R.rhs[R.nda[3 * (size_t)n + i]] = r[i];
  if (isPn) R.rhs[R.dispDOF + R.idp[n]] = r[3];
}

struct LiftArgs {
  int32_t nnear;
  const int32_t* list;        // nodes with a prescribed dof in an element around them
  const int32_t* nptr;
  const int32_t* nelA;
  const int32_t* conn;
  const int32_t* nodeinfo;
  const int32_t* nda;
  const int32_t* idp;
  int32_t dispDOF;
  const double* GS;
  const double* dispApplied;
  double* rhs;
};

// First-iteration lift (applyDirichletBCs / applyDirichletBCs2field, ElementBase.cpp:567-594,
// 717-782): n's right-hand side loses  sum_e K_e(n, b) dubar_b  over the prescribed dofs of the
// nodes b of n's elements.  One wavefront per listed node; its lanes take the (element, node b)
// pairs and the wavefront sums their products in a fixed butterfly order.
template <int NPE, int NGP, bool MIXED>
__global__ void __launch_bounds__(64) k_lift(LiftArgs L) {
  using St = Stage<NPE, NGP, MIXED>;
  __shared__ double sdN[St::COMPACT ? NGP * NPE * 3 : 1];
  const int q = blockIdx.x;
  if (q >= L.nnear) return;
  const int n = L.list[q], lane = threadIdx.x;
  static_assert(!St::COMPACT || NGP * NPE <= 64, "one table entry per lane");
  if (St::COMPACT && lane < NGP * NPE) hex8_dN_node(lane / NPE, lane % NPE, &sdN[3 * lane]);
  __syncthreads();
  const int info = L.nodeinfo[n];
  const int mn = info & 7;
  const bool isPn = St::P1 && ((info >> 3) & 1);
  if (!mn && !isPn) return;
  const int e0 = L.nptr[n], m = L.nptr[n + 1] - e0;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int t = lane; t < m * NPE; t += 64) {
    const int s = t / NPE, b = t - s * NPE;
    const int ea = L.nelA[e0 + s];
    const int e = ea >> 4, a = ea & 15;
    const int vb = L.conn[(size_t)e * NPE + b];
    const int mb = L.nodeinfo[vb] & 7;
    if (mb == 7) continue;
    double db[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) db[k] = ((mb >> k) & 1) ? 0.0 : L.dispApplied[3 * (size_t)vb + k];
    const double* rec = L.GS + (size_t)e * St::TS;
    double Kb[9];
    record_block<NPE, NGP, MIXED>(rec, a, b, Kb, St::COMPACT ? sdN : nullptr);
#pragma unroll
    for (int i = 0; i < 3; ++i) acc[i] += Kb[3 * i] * db[0] + Kb[3 * i + 1] * db[1] + Kb[3 * i + 2] * db[2];
    if (St::P1 && isPn && a < 4) {
      double pu[3] = {0.0, 0.0, 0.0};  // Kpu(a, 3b+k) = sum (dv Np_a) g_b,k
      for (int gp = 0; gp < NGP; ++gp) {
        double Np[4];
        tet_Np(gp, Np);
        const double Na = a == 0 ? Np[0] : a == 1 ? Np[1] : a == 2 ? Np[2] : Np[3];
        const double bb4 = rec[St::S0 + gp * St::SS + 4] * Na;
        const double* gb = rec + St::G0 + (gp * NPE + b) * St::GW;
        pu[0] += bb4 * gb[0]; pu[1] += bb4 * gb[1]; pu[2] += bb4 * gb[2];
      }
      acc[3] += pu[0] * db[0] + pu[1] * db[1] + pu[2] * db[2];
    }
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc[i] += shfl_xor_d(acc[i], off);
  if (lane != 0) return;
#pragma unroll
  for (int i = 0; i < 3; ++i)
    if ((mn >> i) & 1) L.rhs[L.nda[3 * (size_t)n + i]] -= acc[i];
  if (isPn) L.rhs[L.dispDOF + L.idp[n]] -= acc[3];
}

}  // namespace

size_t stage_doubles_per_elem(int etype, bool p0) {
  if (etype == MPF_HEX8) return p0 ? (size_t)Stage<8, 8, true>::TS : (size_t)Stage<8, 8, false>::TS;
  if (etype == MPF_TET10) return (size_t)Stage<10, 4, false>::TS;
  return (size_t)Stage<10, 4, true>::TS;
}

static mpf_status gather_unsupported(mpf_problem* p, const char* why) {
  std::fprintf(stderr, "libmpf: %s; using the coloured scatter\n", why);
  PatternBufs& P = p->pat;
  P.g_order.release();
  P.g_nelA.release();
  P.g_fitem.release();
  P.g_fnode.release();
  P.g_fperm.release();
  P.g_cwave.release();
  P.g_cnw.release();
  P.g_btab.release();
  P.g_near.release();
  P.g_nearlist.release();
  P.g_nnear = 0;
  P.g_ctab.release();
  P.g_ccnt.release();
  P.g_fslot.release();
  P.g_nchunks = 0;
  p->GS.release();
  p->asm_mode = MPF_ASSEMBLY_COLORED_SCATTER;
  return MPF_OK;
}

mpf_status ensure_gather_tables(mpf_problem* p) {
  PatternBufs& P = p->pat;
  if (P.gather_ready) return MPF_OK;
  hipStream_t s = p->stream;
  const int npe = p->npe, nNode = p->nNode;
  const bool mixed = p->etype == MPF_TET10P1;
  const int64_t nadj = (int64_t)p->nElem * npe;
  int64_t nbtot = 0;
  {
    int32_t last = 0;
    if (nNode > 0) MPF_HIP(hipMemcpyAsync(&last, P.nbptr.ptr + nNode, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    MPF_HIP(hipStreamSynchronize(s));
    nbtot = last;
  }
  if (p->nElem >= (1 << 27)) return gather_unsupported(p, "more than 2^27 elements");
  if (nbtot >= (1 << 27)) return gather_unsupported(p, "more than 2^27 node-pair blocks");
  if (nNode == 0 || p->nElem == 0) return gather_unsupported(p, "empty mesh");

  // processing order: Morton order of the node coordinates
  MPF_TRY(P.g_order.alloc(nNode));
  {
    DevBuf<uint64_t> code, code2;
    DevBuf<int32_t> ids;
    MPF_TRY(code.alloc(nNode));
    MPF_TRY(code2.alloc(nNode));
    MPF_TRY(ids.alloc(nNode));
    // positions quantised to a lattice of spacing h = (bounding-box volume / elements)^(1/3), about
    // the element size: on a structured grid each node gets its own lattice point, so 8 consecutive
    // nodes of the order form an aligned 2x2x2 block (the chunked matrix kernel's unit of reuse)
    double ext[3], h = 0.0;
    for (int c = 0; c < 3; ++c) ext[c] = p->xmax[c] - p->xmin[c];
    const double emax = std::max(ext[0], std::max(ext[1], ext[2]));
    if (ext[0] > 0.0 && ext[1] > 0.0 && ext[2] > 0.0) h = std::cbrt(ext[0] * ext[1] * ext[2] / (double)p->nElem);
    if (!(h > 0.0)) h = emax / std::cbrt((double)p->nElem);
    if (!(h > 0.0)) h = 1.0;
    h = std::max(h, emax / 2097151.0);
    const double sc[3] = {1.0 / h, 1.0 / h, 1.0 / h};
    std::vector<int32_t> moff, eoff;
    const bool multi = p->batch && batch_node_offsets(p, moff) == MPF_OK && batch_elem_offsets(p, eoff) == MPF_OK &&
                       moff.size() > 2 && moff.size() < (1u << 18);
    DevBuf<int32_t> moff_d;
    DevBuf<double> mq_d;
    if (multi) {  // each mesh on its own lattice, h_m = (box volume / elements)^(1/3) of that mesh
      const int nm = (int)moff.size() - 1;
      std::vector<double> Xh((size_t)nNode * 3), mq(4 * (size_t)nm);
      MPF_TRY(p->X.download(Xh.data(), Xh.size(), s));
      MPF_HIP(hipStreamSynchronize(s));
      for (int m = 0; m < nm; ++m) {
        double lo3[3] = {1e300, 1e300, 1e300}, hi3[3] = {-1e300, -1e300, -1e300};
        for (int32_t v = moff[m]; v < moff[m + 1]; ++v)
          for (int d = 0; d < 3; ++d) {
            lo3[d] = std::min(lo3[d], Xh[3 * (size_t)v + d]);
            hi3[d] = std::max(hi3[d], Xh[3 * (size_t)v + d]);
          }
        const int32_t ne = eoff[m + 1] - eoff[m];
        double e[3], hm = 0.0;
        for (int d = 0; d < 3; ++d) e[d] = moff[m + 1] > moff[m] ? hi3[d] - lo3[d] : 0.0;
        const double em = std::max(e[0], std::max(e[1], e[2]));
        if (e[0] > 0.0 && e[1] > 0.0 && e[2] > 0.0 && ne > 0) hm = std::cbrt(e[0] * e[1] * e[2] / (double)ne);
        if (!(hm > 0.0)) hm = ne > 0 ? em / std::cbrt((double)ne) : 1.0;
        if (!(hm > 0.0)) hm = 1.0;
        hm = std::max(hm, em / 32767.0);
        for (int d = 0; d < 3; ++d) mq[4 * (size_t)m + d] = moff[m + 1] > moff[m] ? lo3[d] : 0.0;
        mq[4 * (size_t)m + 3] = 1.0 / hm;
      }
      MPF_TRY(moff_d.upload(moff.data(), moff.size(), s));
      MPF_TRY(mq_d.upload(mq.data(), mq.size(), s));
      hipLaunchKernelGGL(k_morton_mesh, dim3(grid_for(nNode, 256)), dim3(256), 0, s, nNode, nm, moff_d.ptr, mq_d.ptr,
                         p->X.ptr, code.ptr, ids.ptr);
    } else {
      hipLaunchKernelGGL(k_morton, dim3(grid_for(nNode, 256)), dim3(256), 0, s, nNode, p->X.ptr, p->xmin[0],
                         p->xmin[1], p->xmin[2], sc[0], sc[1], sc[2], code.ptr, ids.ptr);
    }
    size_t tmp = 0;
    MPF_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, code.ptr, code2.ptr, ids.ptr, P.g_order.ptr, nNode, 0, 63,
                                               s));
    DevBuf<char> t;
    MPF_TRY(t.alloc(tmp > 0 ? tmp : 1));
    MPF_HIP(hipcub::DeviceRadixSort::SortPairs(t.ptr, tmp, code.ptr, code2.ptr, ids.ptr, P.g_order.ptr, nNode, 0, 63,
                                               s));
    MPF_HIP(hipStreamSynchronize(s));
  }

  // step 1: sorted owned items per node
  DevBuf<int32_t> keys, ovf;
  DevBuf<int64_t> cnt;
  MPF_TRY(P.g_nelA.alloc(nadj));
  MPF_TRY(keys.alloc((size_t)nadj * npe));
  MPF_TRY(cnt.alloc(nNode));
  MPF_TRY(ovf.alloc(1));
  MPF_HIP(hipMemsetAsync(ovf.ptr, 0, sizeof(int32_t), s));
  if (mixed)
    hipLaunchKernelGGL((k_node_items<10, true>), dim3(nNode), dim3(64), 0, s, nNode, P.nptr.ptr, P.nel.ptr,
                       P.nbptr.ptr, P.nb.ptr, p->conn.ptr, P.nodeinfo.ptr, P.g_nelA.ptr, keys.ptr, cnt.ptr, ovf.ptr);
  else if (npe == 10)
    hipLaunchKernelGGL((k_node_items<10, false>), dim3(nNode), dim3(64), 0, s, nNode, P.nptr.ptr, P.nel.ptr,
                       P.nbptr.ptr, P.nb.ptr, p->conn.ptr, P.nodeinfo.ptr, P.g_nelA.ptr, keys.ptr, cnt.ptr, ovf.ptr);
  else
    hipLaunchKernelGGL((k_node_items<8, false>), dim3(nNode), dim3(64), 0, s, nNode, P.nptr.ptr, P.nel.ptr,
                       P.nbptr.ptr, P.nb.ptr, p->conn.ptr, P.nodeinfo.ptr, P.g_nelA.ptr, keys.ptr, cnt.ptr, ovf.ptr);
  MPF_HIP(hipGetLastError());
  int32_t o = 0;
  MPF_HIP(hipMemcpyAsync(&o, ovf.ptr, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  MPF_HIP(hipStreamSynchronize(s));
  if (o & 1) return gather_unsupported(p, "a node has more than 64 elements or 128 coupled nodes");
  if (o & 2) return MPF_ERR_STATE;

  // step 2: layout groups of 8 nodes in Morton order, their padded lengths and offsets
  const int64_t ngroups = ((int64_t)nNode + kGroupNodes - 1) / kGroupNodes;
  DevBuf<int64_t> glen, goff;
  MPF_TRY(glen.alloc(ngroups + 1));
  MPF_TRY(goff.alloc(ngroups + 1));
  const unsigned ggrid = (unsigned)((ngroups + 255) / 256);
  if (npe == 10)
    hipLaunchKernelGGL(k_group_len<10>, dim3(ggrid), dim3(256), 0, s, ngroups, nNode, P.g_order.ptr, P.nptr.ptr,
                       keys.ptr, cnt.ptr, glen.ptr);
  else
    hipLaunchKernelGGL(k_group_len<8>, dim3(ggrid), dim3(256), 0, s, ngroups, nNode, P.g_order.ptr, P.nptr.ptr,
                       keys.ptr, cnt.ptr, glen.ptr);
  MPF_HIP(hipMemsetAsync(glen.ptr + ngroups, 0, sizeof(int64_t), s));
  auto exclusive_sum = [&](const int64_t* in, int64_t* out, int64_t n) -> mpf_status {
    size_t tmp = 0;
    MPF_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, in, out, n, s));
    DevBuf<char> t;
    MPF_TRY(t.alloc(tmp > 0 ? tmp : 1));
    MPF_HIP(hipcub::DeviceScan::ExclusiveSum(t.ptr, tmp, in, out, n, s));
    MPF_HIP(hipStreamSynchronize(s));
    return MPF_OK;
  };
  MPF_TRY(exclusive_sum(glen.ptr, goff.ptr, ngroups + 1));
  int64_t total = 0;
  MPF_HIP(hipMemcpyAsync(&total, goff.ptr + ngroups, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  MPF_HIP(hipStreamSynchronize(s));
  if (total >= ((int64_t)1 << 31) - 64) return gather_unsupported(p, "more than 2^31 laid-out items");

  // laid-out items
  MPF_TRY(P.g_fitem.alloc(total > 0 ? (size_t)total : 64));
  MPF_TRY(P.g_fnode.alloc(total > 0 ? (size_t)total : 64));
  MPF_HIP(hipMemsetAsync(P.g_fitem.ptr, 0xff, (total > 0 ? (size_t)total : 64) * sizeof(int2), s));
  MPF_HIP(hipMemsetAsync(P.g_fnode.ptr, 0xff, (total > 0 ? (size_t)total : 64) * sizeof(int32_t), s));
  if (npe == 10)
    hipLaunchKernelGGL(k_group_fill<10>, dim3(ggrid), dim3(256), 0, s, ngroups, nNode, P.g_order.ptr, P.nptr.ptr,
                       P.nbptr.ptr, keys.ptr, cnt.ptr, P.g_nelA.ptr, goff.ptr, P.g_fitem.ptr, P.g_fnode.ptr);
  else
    hipLaunchKernelGGL(k_group_fill<8>, dim3(ggrid), dim3(256), 0, s, ngroups, nNode, P.g_order.ptr, P.nptr.ptr,
                       P.nbptr.ptr, keys.ptr, cnt.ptr, P.g_nelA.ptr, goff.ptr, P.g_fitem.ptr, P.g_fnode.ptr);
  MPF_HIP(hipGetLastError());
  P.g_nwaves = total / 64;
  MPF_TRY(P.g_fperm.alloc(total > 0 ? (size_t)total : 64));
  if (P.g_nwaves > 0)
    hipLaunchKernelGGL(k_wave_perm, dim3((unsigned)P.g_nwaves), dim3(64), 0, s, P.g_nwaves, P.g_fitem.ptr,
                       P.g_fperm.ptr);
  MPF_HIP(hipGetLastError());

  // chunk tables of the chunked matrix kernel (Hex8): each group's wavefronts in chunks of <= CW
  if (p->etype == MPF_HEX8 && P.g_nwaves > 0) {
    constexpr int CW = MPF_CHUNK_WAVES, ECAP = MPF_CHUNK_ELEMS;
    static_assert(ECAP == 32, "mpf_set_chunk_stage_cap documents a cap of 32");
    static_assert(CW >= 1 && CW <= 8, "k_chunk_table covers 8 wavefronts");
    DevBuf<int64_t> ccount, coff;
    MPF_TRY(ccount.alloc(ngroups + 1));
    MPF_TRY(coff.alloc(ngroups + 1));
    hipLaunchKernelGGL(k_group_chunks, dim3(ggrid), dim3(256), 0, s, ngroups, CW, glen.ptr, goff.ptr, nullptr,
                       ccount.ptr, nullptr, nullptr);
    MPF_HIP(hipMemsetAsync(ccount.ptr + ngroups, 0, sizeof(int64_t), s));
    MPF_TRY(exclusive_sum(ccount.ptr, coff.ptr, ngroups + 1));
    int64_t nch = 0;
    MPF_HIP(hipMemcpyAsync(&nch, coff.ptr + ngroups, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    MPF_HIP(hipStreamSynchronize(s));
    P.g_nchunks = nch;
    MPF_TRY(P.g_cwave.alloc((size_t)nch));
    MPF_TRY(P.g_cnw.alloc((size_t)nch));
    MPF_TRY(P.g_ctab.alloc((size_t)nch * ECAP));
    MPF_TRY(P.g_ccnt.alloc((size_t)nch));
    MPF_TRY(P.g_fslot.alloc((size_t)P.g_nwaves * 64));
    hipLaunchKernelGGL(k_group_chunks, dim3(ggrid), dim3(256), 0, s, ngroups, CW, glen.ptr, goff.ptr, coff.ptr,
                       nullptr, P.g_cwave.ptr, P.g_cnw.ptr);
    hipLaunchKernelGGL((k_chunk_table<ECAP>), dim3((unsigned)nch), dim3(512), 0, s, nch, P.g_cwave.ptr, P.g_cnw.ptr,
                       P.g_fitem.ptr, P.g_ctab.ptr, P.g_ccnt.ptr, P.g_fslot.ptr);
    MPF_HIP(hipGetLastError());
    MPF_HIP(hipStreamSynchronize(s));
  }

  // step 3: column-block table; step 4: near-Dirichlet flags
  MPF_TRY(P.g_btab.alloc((size_t)(nbtot > 0 ? nbtot : 1) * (mixed ? 8 : 4)));
  if (mixed)
    hipLaunchKernelGGL(k_block_table<true>, dim3(nNode), dim3(64), 0, s, nNode, P.nbptr.ptr, P.nb.ptr, P.nodeinfo.ptr,
                       P.nda.ptr, P.idp.ptr, P.dispDOF, P.rowptr.ptr, P.g_btab.ptr);
  else
    hipLaunchKernelGGL(k_block_table<false>, dim3(nNode), dim3(64), 0, s, nNode, P.nbptr.ptr, P.nb.ptr, P.nodeinfo.ptr,
                       P.nda.ptr, nullptr, P.dispDOF, P.rowptr.ptr, P.g_btab.ptr);
  if (mixed)
    hipLaunchKernelGGL(k_mirror<true>, dim3(nNode), dim3(64), 0, s, nNode, P.nbptr.ptr, P.nb.ptr, P.nodeinfo.ptr,
                       P.g_btab.ptr);
  else
    hipLaunchKernelGGL(k_mirror<false>, dim3(nNode), dim3(64), 0, s, nNode, P.nbptr.ptr, P.nb.ptr, P.nodeinfo.ptr,
                       P.g_btab.ptr);
  MPF_HIP(hipGetLastError());
  MPF_TRY(P.g_near.alloc(nNode));
  if (npe == 10)
    hipLaunchKernelGGL(k_near<10>, dim3(grid_for(nNode, 256)), dim3(256), 0, s, nNode, P.nptr.ptr, P.nel.ptr,
                       p->conn.ptr, P.nodeinfo.ptr, P.g_near.ptr);
  else
    hipLaunchKernelGGL(k_near<8>, dim3(grid_for(nNode, 256)), dim3(256), 0, s, nNode, P.nptr.ptr, P.nel.ptr,
                       p->conn.ptr, P.nodeinfo.ptr, P.g_near.ptr);
  MPF_HIP(hipGetLastError());
  // list of the flagged nodes (the lift kernel's wavefronts)
  {
    MPF_TRY(P.g_nearlist.alloc(nNode));
    DevBuf<int32_t> nsel;
    MPF_TRY(nsel.alloc(1));
    hipcub::CountingInputIterator<int32_t> ids(0);
    size_t tmp = 0;
    MPF_HIP(hipcub::DeviceSelect::Flagged(nullptr, tmp, ids, P.g_near.ptr, P.g_nearlist.ptr, nsel.ptr, nNode, s));
    DevBuf<char> t;
    MPF_TRY(t.alloc(tmp > 0 ? tmp : 1));
    MPF_HIP(hipcub::DeviceSelect::Flagged(t.ptr, tmp, ids, P.g_near.ptr, P.g_nearlist.ptr, nsel.ptr, nNode, s));
    MPF_HIP(hipMemcpyAsync(&P.g_nnear, nsel.ptr, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    MPF_HIP(hipStreamSynchronize(s));
  }
  MPF_TRY(p->GS.alloc((size_t)p->nElem * stage_doubles_per_elem(p->etype, p->p0)));
  MPF_HIP(hipStreamSynchronize(s));
  P.gather_ready = true;
  return MPF_OK;
}

mpf_status launch_gather_assembly(mpf_problem* p, const ElemArgs& A, const double* dispApplied, int writeK,
                                  hipStream_t s) {
  const PatternBufs& P = p->pat;
  if (!P.gather_ready) return MPF_ERR_STATE;
  const int32_t nElem = p->nElem, nNode = p->nNode;
  if (nElem <= 0 || nNode <= 0) return MPF_OK;
  const bool mixed = p->etype == MPF_TET10P1;
  FlatArgs F{};
  F.nwaves = P.g_nwaves;
  F.fitem = P.g_fitem.ptr;
  F.fnode = P.g_fnode.ptr;
  F.fperm = P.g_fperm.ptr;
  F.nbptr = P.nbptr.ptr;
  F.btab = P.g_btab.ptr;
  F.nda = P.nda.ptr;
  F.idp = P.idp.ptr;
  F.dispDOF = P.dispDOF;
  F.GS = p->GS.ptr;
  F.vals = p->vals.ptr;
  F.rhs = p->rhs.ptr;
  F.nchunks = P.g_nchunks;
  F.ecap = p->chunk_cap;
  F.cwave = P.g_cwave.ptr;
  F.cnw = P.g_cnw.ptr;
  F.ctab = P.g_ctab.ptr;
  F.ccnt = P.g_ccnt.ptr;
  F.fslot = P.g_fslot.ptr;
  RhsArgs R{};
  R.nNode = nNode;
  R.nptr = P.nptr.ptr; R.nelA = P.g_nelA.ptr; R.nodeinfo = P.nodeinfo.ptr; R.nda = P.nda.ptr;
  R.idp = P.idp.ptr; R.dispDOF = P.dispDOF; R.GS = p->GS.ptr; R.rhs = p->rhs.ptr;
  LiftArgs L{};
  L.nnear = P.g_nnear; L.list = P.g_nearlist.ptr;
  L.nptr = P.nptr.ptr; L.nelA = P.g_nelA.ptr; L.conn = p->conn.ptr; L.nodeinfo = P.nodeinfo.ptr;
  L.nda = P.nda.ptr; L.idp = P.idp.ptr; L.dispDOF = P.dispDOF; L.GS = p->GS.ptr; L.dispApplied = dispApplied;
  L.rhs = p->rhs.ptr;
  const bool lift = dispApplied != nullptr && L.nnear > 0;
  // with the matrix the residual sums come from the matrix kernel; k_rhs then only lifts
  const bool flat = writeK && F.nwaves > 0;
  R.lift_only = flat ? 1 : 0;
  const unsigned flat_grid = (unsigned)(((F.nwaves + 7) / 8) * 8);
  const unsigned chunk_grid = (unsigned)(((F.nchunks + 7) / 8) * 8);
  const bool chunked = p->asm_mode != MPF_ASSEMBLY_GATHER_DIRECT && F.nchunks > 0;
  if (p->etype == MPF_TET10) {
    hipLaunchKernelGGL((k_stage<10, 4, false, 16, MPF_STAGE_LDS_TET>), dim3(grid_for(nElem, 16)), dim3(64), 0, s, A,
                       nElem, p->GS.ptr, nullptr, nullptr, nullptr);
    if (flat) hipLaunchKernelGGL((k_flat<10, 4, false>), dim3(flat_grid), dim3(64), 0, s, F);
    hipLaunchKernelGGL((k_rhs<10, 4, false>), dim3(grid_for(nNode, 256)), dim3(256), 0, s, R);
    if (lift) hipLaunchKernelGGL((k_lift<10, 4, false>), dim3(L.nnear), dim3(64), 0, s, L);
  } else if (p->p0) {  // Hex8 Q1/P0: condensed records, displacement rows only
    hipLaunchKernelGGL((k_stage<8, 8, true, 8, MPF_STAGE_LDS_HEX>), dim3(grid_for(nElem, 8)), dim3(64), 0, s, A,
                       nElem, p->GS.ptr, p->q_Kup.ptr, p->q_Kpp.ptr, p->q_Fp.ptr);
    if (flat && chunked)
      hipLaunchKernelGGL((k_flat_chunk<8, 8, true, MPF_CHUNK_WAVES, MPF_CHUNK_ELEMS, 28>),
                         dim3(chunk_grid), dim3(64 * MPF_CHUNK_WAVES), 0, s, F);
    else if (flat)
      hipLaunchKernelGGL((k_flat<8, 8, true>), dim3(flat_grid), dim3(64), 0, s, F);
    hipLaunchKernelGGL((k_rhs<8, 8, true>), dim3(grid_for(nNode, 256)), dim3(256), 0, s, R);
    if (lift) hipLaunchKernelGGL((k_lift<8, 8, true>), dim3(L.nnear), dim3(64), 0, s, L);
  } else if (!mixed) {
    hipLaunchKernelGGL((k_stage<8, 8, false, 8, MPF_STAGE_LDS_HEX>), dim3(grid_for(nElem, 8)), dim3(64), 0, s, A,
                       nElem, p->GS.ptr, nullptr, nullptr, nullptr);
    if (flat && chunked)
      hipLaunchKernelGGL((k_flat_chunk<8, 8, false, MPF_CHUNK_WAVES, MPF_CHUNK_ELEMS>),
                         dim3(chunk_grid), dim3(64 * MPF_CHUNK_WAVES), 0, s, F);
    else if (flat)
      hipLaunchKernelGGL((k_flat<8, 8, false>), dim3(flat_grid), dim3(64), 0, s, F);
    hipLaunchKernelGGL((k_rhs<8, 8, false>), dim3(grid_for(nNode, 256)), dim3(256), 0, s, R);
    if (lift) hipLaunchKernelGGL((k_lift<8, 8, false>), dim3(L.nnear), dim3(64), 0, s, L);
  } else {
    hipLaunchKernelGGL((k_stage<10, 4, true, 16, MPF_STAGE_LDS_TET>), dim3(grid_for(nElem, 16)), dim3(64), 0, s, A,
                       nElem, p->GS.ptr, nullptr, nullptr, nullptr);
    if (flat) hipLaunchKernelGGL((k_flat<10, 4, true>), dim3(flat_grid), dim3(64), 0, s, F);
    hipLaunchKernelGGL((k_rhs<10, 4, true>), dim3(grid_for(nNode, 256)), dim3(256), 0, s, R);
    if (lift) hipLaunchKernelGGL((k_lift<10, 4, true>), dim3(L.nnear), dim3(64), 0, s, L);
  }
  MPF_HIP(hipGetLastError());
  return MPF_OK;
}

}  // namespace mpf
