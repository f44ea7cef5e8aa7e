// elem.hip -- element operators (dense K_e / R_e per element) and the coloured-scatter assembly.
//
// Reference path (chennachaos/mphyfem, src/):
//   Elem_Solid_3D_Disp::calcStiffnessAndResidual      Elem_Solid_3D_Disp.cpp:196-374
//   ElementBase::calcStiffnessAndResidualMixed3D      ElementBase2.cpp:278-552
//   applyDirichletBCs / applyDirichletBCs2field        ElementBase.cpp:567-594, 717-782
//   SolverPetsc::assembleMatrixAndVector(2field)       SolverPetsc.cpp:357-378, 434-502
//
// One 64-lane wavefront (one workgroup) owns EPW elements (of one colour when scattering).
//   Phase 1: lane = (element slot k, Gauss point gp) runs dev::element_point (fem_device.hpp) and
//            stages grad_x N, b grad_x N and the Gauss-point scalars in LDS.
//   Phase 2: lane = node pair (a,b) (Tet10: 100 pairs as two blocks per lane, plus u-p lanes)
//            forms its 3x3 block with dev::tangent_block and either writes it densely (element
//            operator) or adds it into the CSR values through the per-element slot map.  One colour
//            per launch: no two elements of a launch share a node, so the FP64 read-modify-writes
//            never race.  Each lane issues all loads of its blocks before any store.

#include "fem_device.hpp"

namespace mpf {

namespace {

using namespace dev;

// CSR slots of a 3x3 node block: rows of node a (free mask ma, row length La) starting at slot s0,
// columns of node b (free mask mb).  -1 where a dof is prescribed.
__device__ __forceinline__ void block_slots(int s0, int ma, int mb, int La, int addr[9]) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int kk = 0; kk < 3; ++kk)
      addr[3 * i + kk] = (s0 >= 0 && ((ma >> i) & 1) && ((mb >> kk) & 1)) ? s0 + rank3(ma, i) * La + rank3(mb, kk) : -1;
}

// V[addr[t]] += v[t] for N distinct slots: every load is issued before the first store.
template <int N>
__device__ __forceinline__ void rmw_add(double* __restrict__ V, const int addr[N], const double v[N]) {
  double old[N];
#pragma unroll
  for (int t = 0; t < N; ++t) old[t] = addr[t] >= 0 ? V[addr[t]] : 0.0;
#pragma unroll
  for (int t = 0; t < N; ++t)
    if (addr[t] >= 0) V[addr[t]] = old[t] + v[t];
}

// ============================================================================================
// Hex8 displacement element
// ============================================================================================
template <int EPW>
struct Hex8Smem {
  double X[EPW][8][3];
  double U[EPW][8][3];
  double G[EPW][8][8][6];   // [k][gp][a]: g_a (3), (b g)_a (3)
  double S[EPW][8][5];      // [k][gp]: s1..s5
  double R[EPW][8][3];      // [k][a]: element residual
  double dN[8][3][8];       // [gp][r][a]: dN_a/dxi_r at the Gauss points
};

template <int EPW, bool DENSE>
__global__ void __launch_bounds__(64) hex8_kernel(ElemArgs A, int32_t first, int32_t count, double* __restrict__ Ke,
                                                  double* __restrict__ Re, ScatterArgs sc) {
  static_assert(EPW == 8, "one wave holds 8 elements x 8 Gauss points");
  __shared__ Hex8Smem<EPW> sm;
  const int lane = threadIdx.x;
  const int base_idx = blockIdx.x * EPW;
  if (lane < 8) {
    double d[3][8];
    hex8_dN(lane, d);
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int a = 0; a < 8; ++a) sm.dN[lane][r][a] = d[r][a];
  }
  // X, u of the EPW elements (idle slots of a partial block repeat the block's first element)
  for (int t = lane; t < EPW * 8; t += 64) {
    const int k = t >> 3, a = t & 7;
    const int idx = (base_idx + k < count) ? base_idx + k : base_idx;
    const int e = DENSE ? first + idx : sc.elems[idx];
    const size_t nd = (size_t)A.conn[(size_t)e * 8 + a];
#pragma unroll
    for (int c = 0; c < 3; ++c) { sm.X[k][a][c] = A.X[3 * nd + c]; sm.U[k][a][c] = A.u[3 * nd + c]; }
  }
  __syncthreads();
  {
    const int k = lane >> 3, gp = lane & 7;
    const bool act = base_idx + k < count;
    bool negJ = false;
    double mu, K;
    elem_material(A, DENSE ? first + (act ? base_idx + k : base_idx) : sc.elems[act ? base_idx + k : base_idx], mu, K);
    element_point<8, 8, false>(sm.X[k], sm.U[k], nullptr, nullptr, sm.dN[gp], gp, 1.0, mu, K, A.Utype,
                               A.consistent != 0, true, &sm.G[k][gp][0][0], sm.S[k][gp], &sm.R[k][0][0], negJ);
    if (act && negJ) flag_bad(A.bad, DENSE ? first + base_idx + k : sc.elems[base_idx + k]);
  }
  __syncthreads();

  // ---- phase 2: lane = (a, b) ----
  const int a = lane >> 3, b = lane & 7;
  if (DENSE) {
    for (int k = 0; k < EPW; ++k) {
      const int idx = base_idx + k;
      if (idx >= count) break;
      double Kb[9];
      tangent_block<8, true>(&sm.G[k][0][a][0], &sm.G[k][0][b][0], 48, &sm.S[k][0][0], 5, Kb);
      double* Ko = Ke + (size_t)idx * 576;
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int kk = 0; kk < 3; ++kk) Ko[(3 * a + i) * 24 + 3 * b + kk] = Kb[3 * i + kk];
      if (b == 0) {
        double* Ro = Re + (size_t)idx * 24;
        Ro[3 * a] = sm.R[k][a][0]; Ro[3 * a + 1] = sm.R[k][a][1]; Ro[3 * a + 2] = sm.R[k][a][2];
      }
    }
    return;
  }
  double* __restrict__ V = sc.vals;
  double* __restrict__ rhs = sc.rhs;
  // two elements per batch: 18 independent loads per lane
  for (int k0 = 0; k0 < EPW; k0 += 2) {
    if (base_idx + k0 >= count) break;  // uniform across the wave
    int addr[18];
    double val[18];
    double cl[2][3];
    int nar[2];
    int mar[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int k = k0 + h;
      const int idx = base_idx + k;
      const bool live = idx < count;
      double Kb[9];
      tangent_block<8, true>(&sm.G[k][0][a][0], &sm.G[k][0][b][0], 48, &sm.S[k][0][0], 5, Kb);
      const int e = live ? sc.elems[idx] : 0;
      const int na = A.conn[(size_t)e * 8 + a], nb = A.conn[(size_t)e * 8 + b];
      const int infa = sc.nodeinfo[na], infb = sc.nodeinfo[nb];
      const int ma = infa & 7, mb = infb & 7, La = infa >> 4;
      nar[h] = na;
      mar[h] = live ? ma : 0;
      // Dirichlet lift: F_a -= K_ab * dubar_b over prescribed components of b (first iteration)
      double c[3] = {0.0, 0.0, 0.0};
      if (sc.dispApplied) {
        double db[3];
#pragma unroll
        for (int kk = 0; kk < 3; ++kk) db[kk] = ((mb >> kk) & 1) ? 0.0 : sc.dispApplied[3 * (size_t)nb + kk];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          double v = Kb[3 * i] * db[0] + Kb[3 * i + 1] * db[1] + Kb[3 * i + 2] * db[2];
          v += shfl_xor_d(v, 1);
          v += shfl_xor_d(v, 2);
          v += shfl_xor_d(v, 4);
          c[i] = v;
        }
      }
      cl[h][0] = c[0]; cl[h][1] = c[1]; cl[h][2] = c[2];
      const int s0 = (live && sc.scatterK && ma && mb) ? sc.slot[(size_t)e * kHex8Slots + lane] : -1;
      block_slots(s0, ma, mb, La, addr + 9 * h);
#pragma unroll
      for (int t = 0; t < 9; ++t) val[9 * h + t] = Kb[t];
    }
    rmw_add<18>(V, addr, val);
    if (b == 0) {
      int ra[6];
      double rv[6];
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          const bool fr = (mar[h] >> i) & 1;
          ra[3 * h + i] = fr ? sc.nda[3 * (size_t)nar[h] + i] : -1;
          rv[3 * h + i] = sm.R[k0 + h][a][i] - cl[h][i];
        }
      rmw_add<6>(rhs, ra, rv);
    }
  }
}

// ============================================================================================
// Hex8 Q1/P0 mixed element, condensed (Elem_Solid_3D_Mixed0.cpp:93-104), dense outputs
// ============================================================================================
struct Hex8P0Smem {
  double X[8][8][3];
  double U[8][8][3];
  double UP[8][8][3];
  double Pe[8];              // element pressures
  double G[8][8][8][6];      // [k][gp][a]: g_a (3), (b g)_a (3)
  double S[8][8][8];         // [k][gp]: s1..s4, dv, dv0, thetahat, J - Jhat - thetahat p
  double R[8][8][3];         // [k][a]: Fu
  double Kup[8][24];         // [k]: int grad N dv
  double Kpp[8], Fp[8];
  double dN[8][3][8];
};

// One wavefront per 8 elements: phase 1 lane = (element, Gauss point); phase 2 lane = (a, b):
// Kc(a,b) = Kuu(a,b) - Kup_a (Kpu_b / Kpp) with Kpu = Kup^T, Fc_a = Fu_a - Kup_a (Fp / Kpp).
// SCATTER = false: dense Kc, Fc, Kup, Kpp, Fp of elements first .. first + count - 1 (output index idx).
// SCATTER = true: elements sc.elems[0 .. count) of one colour; each lane adds its condensed 3x3 block
// into CSR through the slot map, with the first-iteration lift Fc_a -= sum_b Kc_ab dubar_b
// (ElementBase.cpp:567-594), and Kup, Kpp, Fp are kept by element id for the pressure recovery.
template <bool SCATTER>
__global__ void __launch_bounds__(64) hex8p0_kernel(ElemArgs A, int32_t first, int32_t count,
                                                    double* __restrict__ Kc, double* __restrict__ Fc,
                                                    double* __restrict__ Kup, double* __restrict__ Kpp,
                                                    double* __restrict__ Fp, ScatterArgs sc) {
  __shared__ Hex8P0Smem sm;
  const int lane = threadIdx.x;
  const int base_idx = blockIdx.x * 8;
  auto elem_of = [&](int idx) { return SCATTER ? sc.elems[idx] : first + idx; };
  if (lane < 8) {
    double d[3][8];
    hex8_dN(lane, d);
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int a = 0; a < 8; ++a) sm.dN[lane][r][a] = d[r][a];
  }
  {  // lane = (k, a): node data; idle slots of a partial block repeat the block's first element
    const int k = lane >> 3, a = lane & 7;
    const int idx = (base_idx + k < count) ? base_idx + k : base_idx;
    const int e = elem_of(idx);
    const size_t nd = (size_t)A.conn[(size_t)e * 8 + a];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      sm.X[k][a][c] = A.X[3 * nd + c];
      sm.U[k][a][c] = A.u[3 * nd + c];
      sm.UP[k][a][c] = A.uprev[3 * nd + c];
    }
    if (a == 0) sm.Pe[k] = A.pres[e];
  }
  __syncthreads();
  {
    const int k = lane >> 3, gp = lane & 7;
    const bool act = base_idx + k < count;
    bool negJ = false;
    double mu, K;
    elem_material(A, elem_of(act ? base_idx + k : base_idx), mu, K);
    element_point<8, 8, true>(sm.X[k], sm.U[k], sm.UP[k], &sm.Pe[k], sm.dN[gp], gp, 1.0, mu, K, A.Utype, false, true,
                              &sm.G[k][gp][0][0], sm.S[k][gp], &sm.R[k][0][0], negJ);
    if (act && negJ) flag_bad(A.bad, elem_of(base_idx + k));
  }
  __syncthreads();
  {  // Kup (ElementBase2.cpp:500-505 with Np = 1), Kpp and Fp (:510-539): lane = (k, a)
    const int k = lane >> 3, a = lane & 7;
    double u3[3] = {0.0, 0.0, 0.0}, kpp = 0.0, fp = 0.0;
    for (int gp = 0; gp < 8; ++gp) {
      const double dv = sm.S[k][gp][4], dv0 = sm.S[k][gp][5];
#pragma unroll
      for (int i = 0; i < 3; ++i) u3[i] += dv * sm.G[k][gp][a][i];
      fp -= dv0 * sm.S[k][gp][7];
      kpp -= (dv0 * sm.S[k][gp][6]);
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) sm.Kup[k][3 * a + i] = u3[i];
    if (a == 0) { sm.Kpp[k] = kpp; sm.Fp[k] = fp; }
  }
  __syncthreads();
  const int a = lane >> 3, b = lane & 7;
  // condensed 3x3 block (a, b) and condensed residual of node a for element slot k
  auto condensed = [&](int k, double Kb[9], double fc[3]) {
    tangent_block<8, false>(&sm.G[k][0][a][0], &sm.G[k][0][b][0], 48, &sm.S[k][0][0], 8, Kb);
    const double kpp = sm.Kpp[k];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) Kb[3 * i + j] = Kb[3 * i + j] - sm.Kup[k][3 * a + i] * (sm.Kup[k][3 * b + j] / kpp);
    const double fps = sm.Fp[k] / kpp;
#pragma unroll
    for (int i = 0; i < 3; ++i) fc[i] = sm.R[k][a][i] - sm.Kup[k][3 * a + i] * fps;
  };
  auto keep_coupling = [&](int k, size_t o) {  // Kup, Kpp, Fp at output index o
    if (b == 0) {
#pragma unroll
      for (int i = 0; i < 3; ++i) Kup[o * 24 + 3 * a + i] = sm.Kup[k][3 * a + i];
      if (a == 0) { Kpp[o] = sm.Kpp[k]; Fp[o] = sm.Fp[k]; }
    }
  };
  if (!SCATTER) {
    for (int k = 0; k < 8; ++k) {
      const int idx = base_idx + k;
      if (idx >= count) break;  // uniform across the wave
      double Kb[9], fc[3];
      condensed(k, Kb, fc);
      double* Ko = Kc + (size_t)idx * 576;
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) Ko[(3 * a + i) * 24 + 3 * b + j] = Kb[3 * i + j];
      if (b == 0)
#pragma unroll
        for (int i = 0; i < 3; ++i) Fc[(size_t)idx * 24 + 3 * a + i] = fc[i];
      keep_coupling(k, (size_t)idx);
    }
    return;
  }
  // scatter: two elements per batch (18 independent CSR loads per lane); elements of one colour
  // share no node, so the 18 slots of a lane are distinct
  for (int k0 = 0; k0 < 8; k0 += 2) {
    if (base_idx + k0 >= count) break;  // uniform across the wave
    int addr[18];
    double val[18];
    int ra[6];
    double rv[6];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int k = k0 + h;
      const bool live = base_idx + k < count;
      double Kb[9], fc[3];
      condensed(k, Kb, fc);
      const int e = sc.elems[live ? base_idx + k : base_idx + k0];
      const int na = A.conn[(size_t)e * 8 + a], nb = A.conn[(size_t)e * 8 + b];
      const int infa = sc.nodeinfo[na], infb = sc.nodeinfo[nb];
      const int ma = live ? (infa & 7) : 0, mb = infb & 7, La = infa >> 4;
      double c[3] = {0.0, 0.0, 0.0};
      if (sc.dispApplied) {
        double db[3];
#pragma unroll
        for (int kk = 0; kk < 3; ++kk) db[kk] = ((mb >> kk) & 1) ? 0.0 : sc.dispApplied[3 * (size_t)nb + kk];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          double v = Kb[3 * i] * db[0] + Kb[3 * i + 1] * db[1] + Kb[3 * i + 2] * db[2];
          v += shfl_xor_d(v, 1);
          v += shfl_xor_d(v, 2);
          v += shfl_xor_d(v, 4);
          c[i] = v;
        }
      }
      const int s0 = (sc.scatterK && ma && mb) ? sc.slot[(size_t)e * kHex8Slots + lane] : -1;
      block_slots(s0, ma, mb, La, addr + 9 * h);
#pragma unroll
      for (int t = 0; t < 9; ++t) val[9 * h + t] = Kb[t];
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        ra[3 * h + i] = (b == 0 && ((ma >> i) & 1)) ? sc.nda[3 * (size_t)na + i] : -1;
        rv[3 * h + i] = fc[i] - c[i];
      }
      if (live) keep_coupling(k, (size_t)e);
    }
    rmw_add<18>(sc.vals, addr, val);
    rmw_add<6>(sc.rhs, ra, rv);
  }
}

// solveForPressure (Elem_Solid_3D_Mixed0.cpp:112-199): p_e += (Fp_e - Kpu_e . du_e) / Kpp_e, Kpu = Kup^T, where
// du_e is the element's displacement increment: the solved correction on free dofs and, in the first
// iteration, the applied Dirichlet increment on prescribed ones (femSolidmechanics2.cpp:1137-1186).
// emat (when given): element -> mesh of a batch; elements of meshes whose flag mflag[mesh * stride]
// is 0 (no solve this iteration) keep their pressure.
__global__ void k_hex8p0_pressure(int32_t nElem, const int32_t* __restrict__ conn, const int32_t* __restrict__ nda,
                                  const double* __restrict__ du, const double* __restrict__ dispApplied,
                                  const double* __restrict__ Kup, const double* __restrict__ Kpp,
                                  const double* __restrict__ Fp, double* __restrict__ pres,
                                  const int32_t* __restrict__ emat, const int32_t* __restrict__ mflag, int32_t stride) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nElem) return;
  if (emat && !mflag[(size_t)emat[e] * stride]) return;
  double s = 0.0;
  for (int a = 0; a < 8; ++a) {
    const size_t n = (size_t)conn[(size_t)e * 8 + a];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const int d = nda[3 * n + i];
      const double inc = d >= 0 ? du[d] : (dispApplied ? dispApplied[3 * n + i] : 0.0);
      s += Kup[(size_t)e * 24 + 3 * a + i] * inc;
    }
  }
  pres[e] += (Fp[e] - s) / Kpp[e];
}

// ============================================================================================
// Tet10 elements: the P1 mixed element (MIXED) and the displacement element (Elem_Solid_3D_Disp
// with npElem = 10, Elem_Solid_3D_Disp.cpp:47-51), which is the same kernel without the pressure
// field, the u-p couplings and the pressure rows
// ============================================================================================
template <int EPW>
struct Tet10Smem {
  double X[EPW][10][3];
  double U[EPW][10][3];
  double UP[EPW][10][3];
  double Pn[EPW][4];        // corner pressures
  double G[EPW][4][10][6];  // [k][gp][a]: g_a, (b g)_a
  double S[EPW][4][8];      // [k][gp]: s1..s4, dv, dv0, thetahat, resip
  double R[EPW][10][3];     // Fu
  double FpS[EPW][4];       // Fp
  double LP[EPW][10][4];    // first-iteration lift of Fp from the prescribed dofs of node a: Kpu(q, a) dubar_a
  double dN[4][3][10];      // [gp][r][a]: dN_a/dxi_r at the Gauss points
};

template <int EPW, bool HAS_S5>
__device__ __forceinline__ void tet10_uu_block(const Tet10Smem<EPW>& sm, int k, int a, int b, double Kb[9]) {
  tangent_block<4, HAS_S5>(&sm.G[k][0][a][0], &sm.G[k][0][b][0], 60, &sm.S[k][0][0], 8, Kb);
}

// Slots per element of the coloured scatter's slot map.
template <bool MIXED>
constexpr int tet10_slots() { return MIXED ? kTet10Slots : 100; }

template <int EPW, bool DENSE, bool MIXED = true>
__global__ void __launch_bounds__(64) tet10_kernel(ElemArgs A, int32_t first, int32_t count, double* __restrict__ Kuu,
                                                   double* __restrict__ Kup, double* __restrict__ Kpu,
                                                   double* __restrict__ Kpp, double* __restrict__ Fu,
                                                   double* __restrict__ Fp, ScatterArgs sc) {
  static_assert(EPW == 4 || EPW == 8 || EPW == 16, "phase 1 maps (k, gp) onto the 64 lanes");
  __shared__ Tet10Smem<EPW> sm;
  const int lane = threadIdx.x;
  const int base_idx = blockIdx.x * EPW;
  const double wq = 0.2500000000000000 * (1.0 / 6.0);  // QuadratureUtil.cpp:706-709
  if (lane < 4) {
    double x1, x2, x3, d[3][10];
    tet_gp(lane, x1, x2, x3);
    tet10_dN(x1, x2, x3, d[0], d[1], d[2]);
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int a = 0; a < 10; ++a) sm.dN[lane][r][a] = d[r][a];
  }
  for (int t = lane; t < EPW * 10; t += 64) {
    const int k = t / 10, a = t % 10;
    const int idx = (base_idx + k < count) ? base_idx + k : base_idx;
    const int e = DENSE ? first + idx : sc.elems[idx];
    const size_t nd = (size_t)A.conn[(size_t)e * 10 + a];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      sm.X[k][a][c] = A.X[3 * nd + c];
      sm.U[k][a][c] = A.u[3 * nd + c];
      if (MIXED) sm.UP[k][a][c] = A.uprev[3 * nd + c];
    }
    if (MIXED && a < 4) sm.Pn[k][a] = A.pres[nd];
  }
  __syncthreads();
  {
    // lane = (k, gp); lanes beyond 4*EPW mirror the lower ones and do not write
    const int k = (lane >> 2) & (EPW - 1), gp = lane & 3;
    const bool writer = (lane >> 2) < EPW;
    const bool act = writer && (base_idx + k < count);
    bool negJ = false;
    double mu, K;
    elem_material(A, DENSE ? first + (act ? base_idx + k : base_idx) : sc.elems[act ? base_idx + k : base_idx], mu, K);
    element_point<10, 4, MIXED>(sm.X[k], sm.U[k], MIXED ? sm.UP[k] : nullptr, MIXED ? sm.Pn[k] : nullptr, sm.dN[gp],
                                gp, wq, mu, K, A.Utype, !MIXED && A.consistent != 0, writer, &sm.G[k][gp][0][0],
                                sm.S[k][gp], &sm.R[k][0][0], negJ);
    if (act && negJ) flag_bad(A.bad, DENSE ? first + base_idx + k : sc.elems[base_idx + k]);
  }
  __syncthreads();

  // ---- phase 2: Kuu blocks (a,b) = items lane and lane + 64 (< 100); lanes 36..45 also take the
  //      Kup/Kpu coupling of node a = lane - 36; lanes 46..49 the Kpp row q = lane - 46 ----
  const bool lift = !DENSE && sc.dispApplied != nullptr;
  double* __restrict__ V = sc.vals;
  for (int k = 0; k < EPW; ++k) {
    const int idx = base_idx + k;
    if (idx >= count) break;  // uniform across the wave
    const int e = DENSE ? first + idx : sc.elems[idx];
    {
      const int a1 = lane / 10, b1 = lane % 10;
      const int item2 = lane + 64;
      const bool has2 = item2 < 100;
      const int a2 = has2 ? item2 / 10 : 0, b2 = has2 ? item2 % 10 : 0;
      double Kb[18];
      tet10_uu_block<EPW, !MIXED>(sm, k, a1, b1, Kb);
      tet10_uu_block<EPW, !MIXED>(sm, k, a2, b2, Kb + 9);
      if (DENSE) {
        double* Ko = Kuu + (size_t)idx * 900;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int kk = 0; kk < 3; ++kk) {
            Ko[(3 * a1 + i) * 30 + 3 * b1 + kk] = Kb[3 * i + kk];
            if (has2) Ko[(3 * a2 + i) * 30 + 3 * b2 + kk] = Kb[9 + 3 * i + kk];
          }
      } else {
        int addr[18];
        const int na1 = A.conn[(size_t)e * 10 + a1], nb1 = A.conn[(size_t)e * 10 + b1];
        const int i1a = sc.nodeinfo[na1], i1b = sc.nodeinfo[nb1];
        const int s1 =
            (sc.scatterK && (i1a & 7) && (i1b & 7)) ? sc.slot[(size_t)e * tet10_slots<MIXED>() + a1 * 10 + b1] : -1;
        block_slots(s1, i1a & 7, i1b & 7, i1a >> 4, addr);
        int s2 = -1, i2a = 0, i2b = 0;
        if (has2) {
          const int na2 = A.conn[(size_t)e * 10 + a2], nb2 = A.conn[(size_t)e * 10 + b2];
          i2a = sc.nodeinfo[na2];
          i2b = sc.nodeinfo[nb2];
          s2 = (sc.scatterK && (i2a & 7) && (i2b & 7)) ? sc.slot[(size_t)e * tet10_slots<MIXED>() + a2 * 10 + b2]
                                                       : -1;
        }
        block_slots(s2, i2a & 7, i2b & 7, i2a >> 4, addr + 9);
        rmw_add<18>(V, addr, Kb);
      }
    }
    if (!MIXED) {
      // displacement element: no pressure field
    } else if (lane >= 36 && lane < 46) {
      // Kup(3a+i, q) += (dv g_a,i) Np_q ;  Kpu(q, 3a+j) += (dv Np_q) g_a,j   (ElementBase2.cpp:500-505, 524-533)
      const int a = lane - 36;
      double up[3][4], pu[4][3];
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int q = 0; q < 4; ++q) { up[i][q] = 0.0; pu[q][i] = 0.0; }
#pragma unroll
      for (int gp = 0; gp < 4; ++gp) {
        const double* Ga = sm.G[k][gp][a];
        const double dv = sm.S[k][gp][4];
        double Np[4];
        tet_Np(gp, Np);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const double bb4 = dv * Np[q];
#pragma unroll
          for (int i = 0; i < 3; ++i) {
            up[i][q] += (dv * Ga[i]) * Np[q];
            pu[q][i] += bb4 * Ga[i];
          }
        }
      }
      if (DENSE) {
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            Kup[(size_t)idx * 120 + (3 * a + i) * 4 + q] = up[i][q];
            Kpu[(size_t)idx * 120 + q * 30 + 3 * a + i] = pu[q][i];
          }
      } else {
        const int na = A.conn[(size_t)e * 10 + a];
        const int infa = sc.nodeinfo[na];
        const int ma = infa & 7, La = infa >> 4;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          double c = 0.0;
          if (lift) {
#pragma unroll
            for (int j = 0; j < 3; ++j)
              if (!((ma >> j) & 1)) c += pu[q][j] * sc.dispApplied[3 * (size_t)na + j];
          }
          sm.LP[k][a][q] = c;
        }
        int addr[12];
        double v[12];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int sup = (sc.scatterK && ma) ? sc.slot[(size_t)e * kTet10Slots + 100 + a * 4 + q] : -1;
#pragma unroll
          for (int i = 0; i < 3; ++i) {
            addr[3 * q + i] = (sup >= 0 && ((ma >> i) & 1)) ? sup + rank3(ma, i) * La : -1;
            v[3 * q + i] = up[i][q];
          }
        }
        rmw_add<12>(V, addr, v);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int spu = (sc.scatterK && ma) ? sc.slot[(size_t)e * kTet10Slots + 140 + q * 10 + a] : -1;
#pragma unroll
          for (int i = 0; i < 3; ++i) {
            addr[3 * q + i] = (spu >= 0 && ((ma >> i) & 1)) ? spu + rank3(ma, i) : -1;
            v[3 * q + i] = pu[q][i];
          }
        }
        rmw_add<12>(V, addr, v);
      }
    } else if (lane >= 46 && lane < 50) {
      // Kpp(q, q2) -= (dv0 Np_q thetahat) Np_q2 ;  Fp_q -= (dv0 Np_q) resip   (ElementBase2.cpp:510-539)
      const int q = lane - 46;
      double pp[4] = {0, 0, 0, 0}, fp = 0.0;
#pragma unroll
      for (int gp = 0; gp < 4; ++gp) {
        double Np[4];
        tet_Np(gp, Np);
        const double Nq = q == 0 ? Np[0] : q == 1 ? Np[1] : q == 2 ? Np[2] : Np[3];
        const double* S = sm.S[k][gp];
        double bb5 = S[5] * Nq;
        fp -= bb5 * S[7];
        bb5 *= S[6];
#pragma unroll
        for (int q2 = 0; q2 < 4; ++q2) pp[q2] -= bb5 * Np[q2];
      }
      if (DENSE) {
#pragma unroll
        for (int q2 = 0; q2 < 4; ++q2) Kpp[(size_t)idx * 16 + q * 4 + q2] = pp[q2];
        Fp[(size_t)idx * 4 + q] = fp;
      } else {
        int addr[4];
#pragma unroll
        for (int q2 = 0; q2 < 4; ++q2) addr[q2] = sc.scatterK ? sc.slot[(size_t)e * kTet10Slots + 180 + q * 4 + q2] : -1;
        rmw_add<4>(V, addr, pp);
        sm.FpS[k][q] = fp;
      }
    }
    if (DENSE) {
      if (lane < 30) Fu[(size_t)idx * 30 + lane] = sm.R[k][lane / 3][lane % 3];
      continue;
    }
    __syncthreads();
    // residual rows with the first-iteration lift (applyDirichletBCs2field, ElementBase.cpp:717-782)
    if (lane < 10) {
      const int a = lane;
      const int na = A.conn[(size_t)e * 10 + a];
      const int ma = sc.nodeinfo[na] & 7;
      double c[3] = {0, 0, 0};
      if (lift && ma) {
        for (int b = 0; b < 10; ++b) {
          const int nb = A.conn[(size_t)e * 10 + b];
          const int mb = sc.nodeinfo[nb] & 7;
          if (mb == 7) continue;
          double db[3];
#pragma unroll
          for (int kk = 0; kk < 3; ++kk) db[kk] = ((mb >> kk) & 1) ? 0.0 : sc.dispApplied[3 * (size_t)nb + kk];
          double Kb[9];
          tet10_uu_block<EPW, !MIXED>(sm, k, a, b, Kb);
#pragma unroll
          for (int i = 0; i < 3; ++i) c[i] += Kb[3 * i] * db[0] + Kb[3 * i + 1] * db[1] + Kb[3 * i + 2] * db[2];
        }
      }
      int ra[3];
      double rv[3];
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        ra[i] = ((ma >> i) & 1) ? sc.nda[3 * (size_t)na + i] : -1;
        rv[i] = sm.R[k][a][i] - c[i];
      }
      rmw_add<3>(sc.rhs, ra, rv);
    } else if (MIXED && lane < 14) {
      const int q = lane - 10;
      const int nq = A.conn[(size_t)e * 10 + q];
      double c = 0.0;
      for (int a = 0; a < 10; ++a) c += sm.LP[k][a][q];
      sc.rhs[sc.dispDOF + sc.idp[nq]] += sm.FpS[k][q] - c;
    }
    __syncthreads();
  }
}

}  // namespace

// --------------------------------------------------------------------------------------------
// Launchers (elements per wavefront: Hex8 8, Tet10 MPF_TET_EPW, default 4).
// --------------------------------------------------------------------------------------------
#ifndef MPF_TET_EPW
#define MPF_TET_EPW 4
#endif
constexpr int kHexEPW = 8;
constexpr int kTetEPW = MPF_TET_EPW;

mpf_status launch_hex8_dense(const ElemArgs& a, int32_t first, int32_t count, double* Ke, double* Re,
                             hipStream_t s) {
  if (count <= 0) return MPF_OK;
  ScatterArgs sc{};
  hipLaunchKernelGGL((hex8_kernel<kHexEPW, true>), dim3(grid_for(count, kHexEPW)), dim3(64), 0, s, a, first, count,
                     Ke, Re, sc);
  MPF_HIP(hipGetLastError());
  return MPF_OK;
}

mpf_status launch_hex8_scatter(const ElemArgs& a, const ScatterArgs& sc, hipStream_t s) {
  if (sc.count <= 0) return MPF_OK;
  hipLaunchKernelGGL((hex8_kernel<kHexEPW, false>), dim3(grid_for(sc.count, kHexEPW)), dim3(64), 0, s, a, 0,
                     sc.count, nullptr, nullptr, sc);
  MPF_HIP(hipGetLastError());
  return MPF_OK;
}

mpf_status launch_hex8p0_dense(const ElemArgs& a, int32_t first, int32_t count, double* Kc, double* Fc, double* Kup,
                               double* Kpp, double* Fp, hipStream_t s) {
  if (count <= 0) return MPF_OK;
  ScatterArgs sc{};
  hipLaunchKernelGGL(hex8p0_kernel<false>, dim3(grid_for(count, 8)), dim3(64), 0, s, a, first, count, Kc, Fc, Kup,
                     Kpp, Fp, sc);
  MPF_HIP(hipGetLastError());
  return MPF_OK;
}

mpf_status launch_hex8p0_scatter(const ElemArgs& a, const ScatterArgs& sc, double* Kup, double* Kpp, double* Fp,
                                 hipStream_t s) {
  if (sc.count <= 0) return MPF_OK;
  hipLaunchKernelGGL(hex8p0_kernel<true>, dim3(grid_for(sc.count, 8)), dim3(64), 0, s, a, 0, sc.count, nullptr,
                     nullptr, Kup, Kpp, Fp, sc);
  MPF_HIP(hipGetLastError());
  return MPF_OK;
}

mpf_status launch_hex8p0_pressure(int32_t nElem, const int32_t* conn, const int32_t* nda, const double* du,
                                  const double* dispApplied, const double* Kup, const double* Kpp, const double* Fp,
                                  double* pres, hipStream_t s, const int32_t* emat, const int32_t* mflag,
                                  int32_t stride) {
  if (nElem <= 0) return MPF_OK;
  hipLaunchKernelGGL(k_hex8p0_pressure, dim3(grid_for(nElem, 256)), dim3(256), 0, s, nElem, conn, nda, du,
                     dispApplied, Kup, Kpp, Fp, pres, emat, mflag, stride);
  MPF_HIP(hipGetLastError());
  return MPF_OK;
}

mpf_status launch_tet10_dense(const ElemArgs& a, int32_t first, int32_t count, double* Kuu, double* Kup,
                              double* Kpu, double* Kpp, double* Fu, double* Fp, hipStream_t s) {
  if (count <= 0) return MPF_OK;
  ScatterArgs sc{};
  hipLaunchKernelGGL((tet10_kernel<kTetEPW, true>), dim3(grid_for(count, kTetEPW)), dim3(64), 0, s, a, first, count,
                     Kuu, Kup, Kpu, Kpp, Fu, Fp, sc);
  MPF_HIP(hipGetLastError());
  return MPF_OK;
}

mpf_status launch_tet10d_dense(const ElemArgs& a, int32_t first, int32_t count, double* Ke, double* Re,
                               hipStream_t s) {
  if (count <= 0) return MPF_OK;
  ScatterArgs sc{};
  hipLaunchKernelGGL((tet10_kernel<kTetEPW, true, false>), dim3(grid_for(count, kTetEPW)), dim3(64), 0, s, a, first,
                     count, Ke, nullptr, nullptr, nullptr, Re, nullptr, sc);
  MPF_HIP(hipGetLastError());
  return MPF_OK;
}

mpf_status launch_tet10d_scatter(const ElemArgs& a, const ScatterArgs& sc, hipStream_t s) {
  if (sc.count <= 0) return MPF_OK;
  hipLaunchKernelGGL((tet10_kernel<kTetEPW, false, false>), dim3(grid_for(sc.count, kTetEPW)), dim3(64), 0, s, a, 0,
                     sc.count, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, sc);
  MPF_HIP(hipGetLastError());
  return MPF_OK;
}

mpf_status launch_tet10_scatter(const ElemArgs& a, const ScatterArgs& sc, hipStream_t s) {
  if (sc.count <= 0) return MPF_OK;
  hipLaunchKernelGGL((tet10_kernel<kTetEPW, false>), dim3(grid_for(sc.count, kTetEPW)), dim3(64), 0, s, a, 0,
                     sc.count, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, sc);
  MPF_HIP(hipGetLastError());
  return MPF_OK;
}

}  // namespace mpf
