// band.hip -- direct solve of the assembled displacement system by banded LDL^T on the GPU.
//
// The reference's plain method solves every Newton correction directly (full Newton with a direct
// solve: SolverEigen::solve factors K by Eigen::SimplicialLDLT, SolverEigen.cpp:191-228).  Here the
// free dofs of each independent diagonal block of K (the whole problem, or each mesh of a batch) are
// renumbered node by node (reverse Cuthill-McKee or slices across an axis, whichever gives the
// smaller half-bandwidth b), which makes the block banded; the block is then factored as L D L^T
// without pivoting (K is the symmetric tangent of the consistent mode; a zero or non-finite pivot is
// reported) and solved, one workgroup per segment and kernel.
//
// Segments: a block long enough (n >= 3 b + 64 rows) is cut in two at a separator S of s rows
// (b <= s < b + 16) in its middle, and eliminated in the order [A1, A2 reversed, S]: segment 0
// holds A1 followed by S (band positions in order), segment 1 holds A2 from the block's far end
// followed by S (positions reversed, the separator's own entries left to segment 0).  A row of A1
// and a row of A2 are more than b apart, so the two segments factor independently up to S; S then
// takes the sum of both partial Schur complements and is factored at the end of segment 0.  The
// serial chain of panels of a long block is halved, and the segments (longest first) fill more of
// the GPU's compute units than one workgroup per block.  Shorter blocks are one segment.
//
//   k_band_factor: right-looking in panels of kW columns -- the panel factored column by column
//   (bands up to 176 wide: by one wavefront in registers, three rows per lane; wider: one thread
//   per panel row and one barrier per column), the rank-kW update of the b x b trailing triangle
//   on the matrix cores (FP64 MFMA);
//   k_band_solve: forward and backward substitution in blocks of kW rows, the unit triangle in one
//   wavefront's registers, the coupling to the next b rows by the whole workgroup.
//
// Every sum is taken in a fixed order, so the solution is bitwise reproducible.

#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>

#include "mpf_internal.hpp"

namespace mpf {

namespace {

constexpr int kW = 16;      // panel width (columns per panel, rows per solve block)
constexpr int kBS = 1024;   // threads per workgroup of the solve: 16 wavefronts, one per row of a backward block
constexpr int kFBS = 768;   // threads per workgroup of the factorisation: 12 wavefronts, so that one
                            // wavefront can hold three panel rows per lane in registers (170 VGPRs)
#ifndef MPF_BAND_TILES
#define MPF_BAND_TILES 2
#endif
constexpr int kNH = MPF_BAND_TILES;  // trailing-update tiles per wavefront and round (old values loaded first)
static_assert(kBS / 64 == kW, "one wavefront per panel row in the backward solve");

// One segment of a block (see the file comment).  Local row k of the segment is band position
// base + dir * k; its band is column-major lower, entry (r, c) at boff + c * (b + 1) + (r - c), and
// its right-hand side / solution y[yoff + k].
struct BandSeg {
  int64_t boff;     // band storage offset
  int32_t yoff;     // first entry in y (also the segment's first fill column)
  int32_t base;     // band position of local row 0
  int32_t dir;      // +1 or -1
  int32_t n;        // rows
  int32_t nf;       // columns factored before the separator's Schur complement is complete
  int32_t nend;     // columns factored in all (nf, or n for the segment that ends with the separator)
  int32_t zc;       // rows k < zc are this segment's own (system values, solution out); rows from zc
                    // (the separator's copy in segment 1) start at zero
  int32_t blk;      // block
  int32_t b;        // half-bandwidth of the block
  int32_t partner;  // the block's other segment, -1 for a block in one segment
};

struct BandArgs {
  const BandSeg* seg;
  double* band;
  double* y;     // permuted right-hand side, then the solution, per segment
  int32_t* bad;  // set on a zero or non-finite pivot
};

// Band of every active segment from the CSR rows, zeros included, and its right-hand side, in one
// pass: 16 lanes per band column, 16 columns per workgroup (segments of blocks whose active[blk *
// astride] is 0 keep their factors).  Column k of a segment holds L(i, k), i = k .. k + b, which
// K's symmetry puts in CSR row perm[base + dir k]: the row's entries whose local row falls in the
// band (and in the segment) are placed in an LDS copy of the column (colmax >= b + 1 doubles),
// zeros elsewhere, and the column goes out with consecutive lanes.  Columns from zc are zero.
// Also clears the pivot flag.  A column is a chain of dependent loads (segment, row, row pointer,
// column indices, positions), so many columns per compute unit are kept in flight.  (Measured on
// the 200 cantilevers: 2.1 ms as a zero pass of the band followed by a scatter of each CSR row
// into it; 1.6 ms in one pass with a wavefront per column.)
constexpr int kFillLanes = 16;                 // lanes per band column
constexpr int kFillCols = 256 / kFillLanes;    // columns per workgroup
__global__ void __launch_bounds__(256) k_band_fill(int32_t ncols, int32_t colmax, const int32_t* __restrict__ cseg,
                                                   const int32_t* __restrict__ rowptr,
                                                   const int32_t* __restrict__ colidx,
                                                   const double* __restrict__ vals, const double* __restrict__ rhs,
                                                   const int32_t* __restrict__ perm, const int32_t* __restrict__ iperm,
                                                   BandArgs A, const int32_t* __restrict__ active, int32_t astride) {
  extern __shared__ double sFill[];
  if (blockIdx.x == 0 && threadIdx.x == 0) *A.bad = 0;
  const int grp = threadIdx.x / kFillLanes, lane = threadIdx.x % kFillLanes;
  const int J = blockIdx.x * kFillCols + grp;
  // a column's lanes share its LDS copy within one wavefront: no workgroup barrier below
  if (J >= ncols) return;
  const BandSeg S = A.seg[cseg[J]];
  if (active && !active[(size_t)S.blk * astride]) return;
  const int k = J - S.yoff, b = S.b, B1 = b + 1;
  double* col = sFill + (size_t)grp * colmax;
  for (int t = lane; t < B1; t += kFillLanes) col[t] = 0.0;
  __builtin_amdgcn_wave_barrier();
  int r = -1;
  if (k < S.zc) {
    r = perm[S.base + S.dir * k];
    const int e1 = rowptr[r + 1];
    // in chunks of 8 entries per lane, each stage's loads issued together (a row of up to 128
    // entries is one round trip per stage, not one per entry)
    for (int e0 = rowptr[r] + lane; e0 < e1; e0 += 8 * kFillLanes) {
      int ci[8], pi[8];
      double vv[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int e = e0 + u * kFillLanes;
        ci[u] = e < e1 ? colidx[e] : -1;
        vv[u] = e < e1 ? vals[e] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) pi[u] = ci[u] >= 0 ? iperm[ci[u]] : 0;
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int i = S.dir * (pi[u] - S.base);
        if (ci[u] >= 0 && i >= k && i - k <= b && i < S.n) col[i - k] = vv[u];
      }
    }
  }
  __builtin_amdgcn_wave_barrier();
  double* L = A.band + S.boff + (size_t)k * B1;
  for (int t = lane; t < B1; t += kFillLanes) L[t] = col[t];
  if (lane == 0) A.y[J] = r >= 0 ? rhs[r] : 0.0;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// v from lane l of the wavefront (l uniform)
__device__ __forceinline__ double lane_value(double v, int l) {
  const uint64_t u = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, l);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), l);
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

typedef double double4_t __attribute__((ext_vector_type(4)));

// Panel of w columns from band column j0 into P (column-major, row stride PR: panel row r = band
// row j0 + r, rows past the block's end as zeros), by threads t0, t0 + ts, ...
// (in chunks of 4 entries per thread, each chunk's loads issued before its LDS stores: a band
// up to 176 wide is one chunk over 768 threads, one round trip to memory)
__device__ __forceinline__ void panel_load(double* P, const double* L, int j0, int w, int n, int B1, int PR, int t0,
                                           int ts) {
  for (int t1 = t0; t1 < w * B1; t1 += 4 * ts) {  // panel (r, c), r - c <= b
    double v[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int t = t1 + i * ts, c = t / B1, k = t - c * B1;
      v[i] = (t < w * B1 && j0 + c + k < n) ? L[(size_t)(j0 + c) * B1 + k] : 0.0;
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int t = t1 + i * ts, c = t / B1, k = t - c * B1;
      if (t < w * B1) P[c * PR + c + k] = v[i];
    }
  }
}

__device__ __forceinline__ void panel_store(const double* P, double* L, int j0, int w, int n, int B1, int PR, int t0,
                                            int ts) {
  for (int t = t0; t < w * B1; t += ts) {
    const int c = t / B1, k = t - c * B1, r = c + k;
    if (j0 + r < n) L[(size_t)(j0 + c) * B1 + k] = P[c * PR + r];
  }
}

// Panel factorisation by one wavefront (w + b <= 192 panel rows): lane l holds panel rows l, l + 64,
// l + 128 in registers; right-looking one column at a time, the pivot and the diagonal-block rows'
// entries of column c read across lanes, no barrier.  At column c every row r > c takes
// A(r, c') -= A(r, c) L(c', c) for c < c' <= min(r, w - 1), L(c', c) = A(c', c) / d_c, and
// L(r, c) = A(r, c) / d_c; P receives L below the diagonal and d on it, sd the pivots.
__device__ __forceinline__ void panel_factor_wave(double* P, double* sd, int j0, int w, int b, int n, int PR, int lane,
                                                  int32_t* bad) {
  double a[3][kW];
#pragma unroll
  for (int s = 0; s < 3; ++s) {
    const int r = lane + 64 * s;
    const bool rowv = r < w + b && j0 + r < n;
#pragma unroll
    for (int c = 0; c < kW; ++c) a[s][c] = (rowv && c < w && c <= r && r - c <= b) ? P[c * PR + r] : 0.0;
  }
#pragma unroll
  for (int c = 0; c < kW; ++c) {
    if (c < w) {  // uniform
      const double d = lane_value(a[0][c], c);
      if (lane == 0 && (d == 0.0 || !isfinite(d))) *bad = 1;
      const double inv = 1.0 / d;
      // every row takes the update, rows above cc included: the entries that changes (above the
      // diagonal) are never read or stored, and no test sits in the update
#pragma unroll
      for (int cc = c + 1; cc < kW; ++cc) {
        if (cc < w) {
          const double m = lane_value(a[0][c], cc) * inv;  // L(cc, c)
#pragma unroll
          for (int s = 0; s < 3; ++s) a[s][cc] -= a[s][c] * m;
        }
      }
      a[0][c] = lane > c ? a[0][c] * inv : a[0][c];  // L(r, c); the pivot row keeps d
      a[1][c] *= inv;
      a[2][c] *= inv;
    }
  }
#pragma unroll
  for (int s = 0; s < 3; ++s) {
    const int r = lane + 64 * s;
    const bool rowv = r < w + b && j0 + r < n;
#pragma unroll
    for (int c = 0; c < kW; ++c)
      if (rowv && c < w && c <= r && r - c <= b) P[c * PR + r] = a[s][c];  // L(r, c), d on the diagonal
  }
#pragma unroll
  for (int c = 0; c < kW; ++c)
    if (lane == c && c < w) sd[c] = a[0][c];
}

// Product tile (R0 .., C0 ..) of the rank-w update of factored panel P (pivots sd) by one wavefront
// on the matrix cores: lane l ends with sum_k L(R0 + 4 v + l / 16, k) d_k L(C0 + l % 16, k) in acc[v]
// (v_mfma_f64_16x16x4f64, k = 16 as 4 steps).
__device__ __forceinline__ double4_t update_tile(const double* P, const double* sd, int w, int b, int PR, int R0,
                                                 int C0, int lane) {
  const int i16 = lane & 15, k4 = lane >> 4;
  double4_t acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int s = 0; s < kW / 4; ++s) {
    const int k = 4 * s + k4;
    const int ra = R0 + i16, cb = C0 + i16;
    const double av = (k < w && ra < w + b && ra - k <= b) ? P[k * PR + ra] : 0.0;
    const double bv = (k < w && cb < w + b && cb - k <= b) ? P[k * PR + cb] * sd[k] : 0.0;
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc, 0, 0, 0);
  }
  return acc;
}

// Trailing update of panel (j0, w) in the band: tiles (ti, tj) of the lower triangle of the window
// of panel-relative rows and columns w .. w + b - 1, tj >= tj0, taken u0, u0 + us, ... in row-major
// order by one wavefront, NH tiles at a time with their old values loaded first.
template <int NH>
__device__ __forceinline__ void trailing_band(const double* P, const double* sd, double* L, int j0, int w, int b,
                                              int n, int B1, int PR, int T, int tj0, int u0, int us, int lane) {
  const int Tq = T - tj0, NTq = Tq * (Tq + 1) / 2;
  const int i16 = lane & 15, k4 = lane >> 4;
  for (int q0 = u0; q0 < NTq; q0 += NH * us) {
    int R0[NH], C0[NH];
    double old[NH][4];
#pragma unroll
    for (int h = 0; h < NH; ++h) {
      const int q = q0 + h * us;
      int t = 0;
      while ((t + 1) * (t + 2) / 2 <= q) ++t;
      R0[h] = w + 16 * (t + tj0);
      C0[h] = w + 16 * (tj0 + q - t * (t + 1) / 2);
      const int c = C0[h] + i16;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = R0[h] + 4 * v + k4;
        old[h][v] = (q < NTq && r >= c && r < w + b && j0 + r < n) ? L[(size_t)(j0 + c) * B1 + (r - c)] : 0.0;
      }
    }
#pragma unroll
    for (int h = 0; h < NH; ++h) {
      if (q0 + h * us >= NTq) break;
      const double4_t acc = update_tile(P, sd, w, b, PR, R0[h], C0[h], lane);
      const int c = C0[h] + i16;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = R0[h] + 4 * v + k4;
        if (r >= c && r < w + b && j0 + r < n) L[(size_t)(j0 + c) * B1 + (r - c)] = old[h][v] - acc[v];
      }
    }
  }
}

// Factor one segment per workgroup, columns [0, nf) (phase 0) or [nf, nend) (phase 1: the
// separator, after k_band_schur); segments of blocks whose active[blk * astride] is 0 are skipped
// (active may be null).  Stopping at nf leaves the trailing window below the last panel -- the
// separator's rows -- updated in the band.  Dynamic LDS: bands up to 176 wide, two panel buffers of (kW + bmax) x kW doubles
// (column-major, row stride PR = kW + bmax) each followed by its kW pivots; wider bands, one panel
// buffer, its pivots and 2 kW doubles of published column entries.
//
// Panel step at column j0 (w = min(kW, c1 - j0) columns, panel-relative rows 0 .. w - 1 + b):
//   (1) the panel is factored right-looking, one column at a time (panel_factor_wave); wider
//       panels give each row a thread and publish the diagonal-block rows' column c in LDS before
//       each step (one barrier per column);
//   (2) the trailing b x b window takes A(r, c') -= sum_k L(r, k) d_k L(c', k) for r >= c', in
//       16 x 16 tiles of the lower triangle on the matrix cores.
// Bands up to 176 wide look one panel ahead: the next panel is loaded into the other buffer and its
// columns (the window's first tile column) take (2) there; then wavefront 0 factors the next panel
// while the other wavefronts store this panel and apply (2) to the rest of the window in the band.
// Every stored value goes through the same operations in the same order either way.  Measured on
// the longest of 200 cantilevers (12 960 dofs, b = 167, 810 panels) before the look-ahead: 17.3 ms,
// of which 7.3 ms in (1), 7.3 ms in (2) and 2.7 ms loading and storing panels (timing-only builds
// that skip (1), (2) or both).
__global__ void __launch_bounds__(kFBS) k_band_factor(BandArgs A, int32_t bmax, int32_t phase,
                                                     const int32_t* __restrict__ active, int32_t astride,
                                                     int32_t* __restrict__ factored) {
  extern __shared__ double sP[];
  const BandSeg S = A.seg[blockIdx.x];
  if (active && !active[(size_t)S.blk * astride]) return;
  const int c0 = phase ? S.nf : 0, c1 = phase ? S.nend : S.nf;
  if (phase && c0 >= c1) return;
  const int PR = kW + bmax;
  const int tid = threadIdx.x, wid = tid >> 6, lane = tid & 63;
  const int n = S.n, b = S.b;
  double* L = A.band + S.boff;
  const int B1 = b + 1;
  const int T = (b + 15) >> 4;  // trailing tiles per side
  constexpr int NW = kFBS / 64;

  if (kW + bmax <= 3 * 64) {  // the same for every block of the launch
    double* const buf0 = sP;
    double* const buf1 = sP + (size_t)PR * kW + kW;
    if (c1 > c0) {
      panel_load(buf0, L, c0, min(kW, c1 - c0), n, B1, PR, tid, kFBS);
      __syncthreads();
      if (wid == 0) panel_factor_wave(buf0, buf0 + (size_t)PR * kW, c0, min(kW, c1 - c0), b, n, PR, lane, A.bad);
      __syncthreads();
    }
    for (int j0 = c0, q = 0; j0 < c1; j0 += kW, q ^= 1) {
      const int w = min(kW, c1 - j0);
      double* cur = q ? buf1 : buf0;
      double* nxt = q ? buf0 : buf1;
      const double* sdc = cur + (size_t)PR * kW;
      const int j1 = j0 + kW;
      const int w1 = j1 < c1 ? min(kW, c1 - j1) : 0;
      if (w1 > 0) {  // uniform
        panel_load(nxt, L, j1, w1, n, B1, PR, tid, kFBS);
        __syncthreads();
        // the window's first tile column is the next panel: its tiles are updated in LDS
        const int i16 = lane & 15, k4 = lane >> 4;
        for (int ti = wid; ti < T; ti += NW) {
          const double4_t acc = update_tile(cur, sdc, w, b, PR, w + 16 * ti, w, lane);
          const int c = w + i16;
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int r = w + 16 * ti + 4 * v + k4;
            if (r >= c && r < w + b && j0 + r < n) nxt[(c - w) * PR + (r - w)] -= acc[v];
          }
        }
        __syncthreads();
      }
      if (wid == 0) {
        if (w1 > 0) panel_factor_wave(nxt, nxt + (size_t)PR * kW, j1, w1, b, n, PR, lane, A.bad);
      } else {
        panel_store(cur, L, j0, w, n, B1, PR, tid - 64, kFBS - 64);
        // the last panel has no next panel in LDS: its window's first tile column goes to the band
        trailing_band<kNH>(cur, sdc, L, j0, w, b, n, B1, PR, T, w1 > 0 ? 1 : 0, wid - 1, NW - 1, lane);
      }
      __syncthreads();
    }
  } else {
    double* sd = sP + (size_t)PR * kW;  // kW pivots
    double* sC = sd + kW;               // 2 x kW: published column entries of the diagonal-block rows
    for (int j0 = c0; j0 < c1; j0 += kW) {
      const int w = min(kW, c1 - j0);
      panel_load(sP, L, j0, w, n, B1, PR, tid, kFBS);
      __syncthreads();
      {  // (1) thread r holds panel row r (r < w + b, the diagonal block's rows included); before
         // step c the diagonal-block rows' final entries of column c are published in LDS, so every
         // thread forms L(c', c) = A(c', c) / d_c itself and a step costs one barrier
        const int r = tid;
        const bool rowv = r < w + b && j0 + r < n;
        double a[kW];
#pragma unroll
        for (int c = 0; c < kW; ++c) a[c] = (rowv && c < w && c <= r && r - c <= b) ? sP[c * PR + r] : 0.0;
        if (rowv && r < w) sC[r] = a[0];
        __syncthreads();
#pragma unroll
        for (int c = 0; c < kW; ++c) {
          if (c < w) {  // uniform
            const double* col = sC + (c & 1) * kW;
            const double d = col[c];
            if (tid == 0 && (d == 0.0 || !isfinite(d))) *A.bad = 1;
            const double inv = 1.0 / d;
            if (rowv && r > c) {
              const double arc = a[c];
#pragma unroll
              for (int cc = c + 1; cc < kW; ++cc)
                if (cc < w && cc <= r) a[cc] -= arc * (col[cc] * inv);
              a[c] = arc * inv;  // L(r, c)
            }
            if (rowv && r > c && r < w && c + 1 < w) sC[((c + 1) & 1) * kW + r] = a[c + 1];
            __syncthreads();
          }
        }
#pragma unroll
        for (int c = 0; c < kW; ++c)
          if (rowv && c < w && c <= r && r - c <= b) sP[c * PR + r] = a[c];  // L(r, c), d on the diagonal
        if (rowv && r < w) sd[r] = a[r];
      }
      __syncthreads();
      panel_store(sP, L, j0, w, n, B1, PR, tid, kFBS);
      trailing_band<kNH>(sP, sd, L, j0, w, b, n, B1, PR, T, 0, wid, NW, lane);  // (2)
      __syncthreads();
    }
  }
  if (tid == 0 && phase == 0) factored[S.blk] = 1;  // the band holds this block's factors once the launches end
}

// Solve L D L^T x = y in place for one segment per workgroup (skipped like the factorisation), in
// blocks of kW rows: the kW x kW unit triangle by wavefront 0 with the segment's L in registers and
// the solved values read across lanes; the coupling to the next b rows by the workgroup (forward)
// or per row of the block by one wavefront with a butterfly sum (backward).  Over columns
// [c0, c1): forward L z = y, then z / d; backward L^T x = z, rows past c1 already solved.

// Band entries of forward block step j0 (w rows): thread gt of a group holds coupling row
// j0 + w + gt's entries in lc (zero outside the band), the group's wavefront 0 lane r row r of the
// unit triangle in l.
__device__ __forceinline__ void fwd_entries(const double* L, int B1, int b, int n, int j0, int w, int gt, int gw,
                                            int lane, double (&lc)[kW], double (&l)[kW]) {
  const int r = j0 + w + gt;
  const bool crow = r < min(n, j0 + w + b);
#pragma unroll
  for (int k = 0; k < kW; ++k)
    lc[k] = (crow && k < w && r - (j0 + k) <= b) ? L[(size_t)(j0 + k) * B1 + (r - (j0 + k))] : 0.0;
  if (gw == 0) {
#pragma unroll
    for (int c = 0; c < kW; ++c)
      l[c] = (lane < w && c < lane && lane - c <= b) ? L[(size_t)(j0 + c) * B1 + (lane - c)] : 0.0;
  }
}

// The rows a sweep is working on -- the b + kW rows of the window below (forward) or above
// (backward) the current block -- are kept in an LDS ring, row r at ring[r % kRing], so that a
// block step reads and updates them without a round trip through global memory.
constexpr int kRing = 1024;
static_assert(kRing >= 2 * (493 + kW), "the ring holds the window and the rows loaded ahead of it");

// forward substitution over columns [c0, c1) and the division by the pivots.  The two halves of
// the workgroup take the block steps in turn (b < kBS / 2 coupling rows each), and each half loads
// the band entries of its next step as soon as it has applied the current one, so that they
// arrive while the other half's step runs.  The ring starts with rows c0 .. c0 + kRing - 1; a row
// solved at a step frees its slot for the row kRing further on, loaded then and stored one step
// later; the rows past c1 that took updates are written back at the end.
__device__ __forceinline__ void band_forward(const double* L, double* y, double* sy, double* ring, int n, int b,
                                             int c0, int c1, int tid, int wid, int lane) {
  const int B1 = b + 1;
  const int g = tid / (kBS / 2), gt = tid % (kBS / 2), gw = gt >> 6;
  const int nsteps = (c1 - c0 + kW - 1) / kW;
  for (int i = tid; i < kRing; i += kBS) ring[(c0 + i) % kRing] = c0 + i < n ? y[c0 + i] : 0.0;
  double lc[kW], l[kW];
  if (g < nsteps) fwd_entries(L, B1, b, n, c0 + g * kW, min(kW, c1 - c0 - g * kW), gt, gw, lane, lc, l);
  double pend = 0.0;  // last wavefront, lanes 0 .. kW - 1: a row loaded ahead, stored into the ring one step later
  int prow = -1;
  __syncthreads();
  for (int st = 0; st < nsteps; ++st) {
    const int j0 = c0 + st * kW, w = min(kW, c1 - j0);
    const bool mine = (st & 1) == g;
    const int r = j0 + w + gt;
    const bool crow = mine && r < min(n, j0 + w + b);
    if (mine && gw == 0) {
      double z = lane < w ? ring[(j0 + lane) % kRing] : 0.0;
#pragma unroll
      for (int c = 0; c < kW; ++c) {
        if (c < w) {
          const double zc = lane_value(z, c);
          if (lane > c) z -= l[c] * zc;
        }
      }
      if (lane < w) {
        sy[lane] = z;
        y[j0 + lane] = z;
      }
    }
    __syncthreads();
    if (crow) {
      double acc = ring[r % kRing];
#pragma unroll
      for (int k = 0; k < kW; ++k)
        if (k < w) acc -= lc[k] * sy[k];  // lc is 0 beyond the band
      ring[r % kRing] = acc;
    }
    // the last wavefront (a coupling row only in bands wider than kBS / 2 - 64, so its wait for the row
    // loaded ahead is not a wait for band entries): the slots of the rows solved at the step before
    // take their rows kRing on
    if (wid == kBS / 64 - 1 && lane < kW) {
      if (prow >= 0) ring[prow % kRing] = pend;
      const int nr = j0 + lane + kRing;
      prow = (lane < w && nr < n) ? nr : -1;
      if (prow >= 0) pend = y[nr];
    }
    if (mine && st + 2 < nsteps)
      fwd_entries(L, B1, b, n, j0 + 2 * kW, min(kW, c1 - j0 - 2 * kW), gt, gw, lane, lc, l);
    __syncthreads();
  }
  for (int r = c1 + tid; r < min(n, c1 + b); r += kBS) y[r] = ring[r % kRing];
  for (int i0 = c0 + tid; i0 < c1; i0 += 4 * kBS) {  // 4 rows per thread per round trip
    double yv[4], dv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int i = i0 + u * kBS;
      yv[u] = i < c1 ? y[i] : 0.0;
      dv[u] = i < c1 ? L[(size_t)i * B1] : 1.0;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (i0 + u * kBS < c1) y[i0 + u * kBS] = yv[u] / dv[u];
  }
  __syncthreads();
}

constexpr int kBRow = 8;  // band entries per lane of a backward row sum: b < 64 kBRow
static_assert(64 * kBRow > 493, "the widest band the panel LDS allows");

// backward row k's entries L(r, k), r = j0 + w + lane + 64 t <= min(n - 1, k + b), k = j0 + wid,
// and its right-hand side
__device__ __forceinline__ void bwd_entries(const double* L, const double* y, int B1, int b, int n, int j0, int w,
                                            int wid, int lane, double (&v)[kBRow], double& yk) {
  const int k = j0 + wid;
#pragma unroll
  for (int t = 0; t < kBRow; ++t) {
    const int r = j0 + w + lane + 64 * t;
    v[t] = (wid < w && r <= min(n - 1, k + b)) ? L[(size_t)k * B1 + (r - k)] : 0.0;
  }
  yk = wid < w ? y[k] : 0.0;
}

// backward substitution over columns [c0, c1) (c0 a multiple of kW); each row's band entries and
// right-hand side for the next block step are loaded while this one is summed, and the solved
// rows the sums read are in the ring (rows c1 .. c1 + kRing - 1 from y at the start)
__device__ __forceinline__ void band_backward(const double* L, double* y, double* sy, double* ring, int n, int b,
                                              int c0, int c1, int tid, int wid, int lane) {
  const int B1 = b + 1;
  for (int i = tid; i < kRing; i += kBS) ring[(c1 + i) % kRing] = c1 + i < n ? y[c1 + i] : 0.0;
  double v[kBRow], vn[kBRow], yk = 0.0, ykn = 0.0;
  if (c1 > c0) {
    const int q = (c1 - 1) / kW;
    bwd_entries(L, y, B1, b, n, q * kW, min(kW, c1 - q * kW), wid, lane, v, yk);
  }
  __syncthreads();
  for (int q = (c1 - 1) / kW; c1 > c0 && q >= c0 / kW; --q) {
    const int j0 = q * kW, w = min(kW, c1 - j0);
    double lt[kW];  // wavefront 0, lane r: L(c, r), c > r, loaded before the row sums
    if (wid == 0) {
#pragma unroll
      for (int c = 0; c < kW; ++c)
        lt[c] = (lane < w && c < w && c > lane && c - lane <= b) ? L[(size_t)(j0 + lane) * B1 + (c - lane)] : 0.0;
    }
    if (q > c0 / kW) bwd_entries(L, y, B1, b, n, j0 - kW, kW, wid, lane, vn, ykn);
    if (wid < w) {  // row j0 + wid: minus the rows already solved beyond the block
      const int k = j0 + wid;
      double part = 0.0;
#pragma unroll
      for (int t = 0; t < kBRow; ++t) {
        const int r = j0 + w + lane + 64 * t;
        if (r <= min(n - 1, k + b)) part += v[t] * ring[r % kRing];
      }
      part = wave_sum(part);
      if (lane == 0) sy[wid] = yk - part;
    }
    __syncthreads();
    if (wid == 0) {
      double x = lane < w ? sy[lane] : 0.0;
#pragma unroll
      for (int c = kW - 1; c >= 0; --c) {
        if (c < w) {
          const double xc = lane_value(x, c);
          if (lane < c) x -= lt[c] * xc;
        }
      }
      if (lane < w) {
        ring[(j0 + lane) % kRing] = x;
        y[j0 + lane] = x;
      }
    }
#pragma unroll
    for (int t = 0; t < kBRow; ++t) v[t] = vn[t];
    yk = ykn;
    __syncthreads();
  }
}

// mode 0: forward and pivots over [0, nf); mode 1: the separator, [nf, nend) forward, pivots and
// backward (after k_band_sep has summed its right-hand side); mode 2: backward over [0, nf) (after
// k_band_sep has copied the separator's solution to segment 1).  A block in one segment: mode 0
// and mode 2 are the whole solve.
__global__ void __launch_bounds__(kBS) k_band_solve(BandArgs A, int32_t mode, const int32_t* __restrict__ active,
                                                   int32_t astride) {
  __shared__ double sy[kW];
  __shared__ double ring[kRing];
  const BandSeg S = A.seg[blockIdx.x];
  if (active && !active[(size_t)S.blk * astride]) return;
  const int tid = threadIdx.x, wid = tid >> 6, lane = tid & 63;
  const double* L = A.band + S.boff;
  double* y = A.y + S.yoff;
  if (mode == 0) {
    band_forward(L, y, sy, ring, S.n, S.b, 0, S.nf, tid, wid, lane);
  } else if (mode == 1) {
    if (S.nend <= S.nf) return;
    band_forward(L, y, sy, ring, S.n, S.b, S.nf, S.nend, tid, wid, lane);
    band_backward(L, y, sy, ring, S.n, S.b, S.nf, S.nend, tid, wid, lane);
  } else {
    band_backward(L, y, sy, ring, S.n, S.b, 0, S.nf, tid, wid, lane);
  }
}

// The separator of each cut block (one workgroup per segment; segment 0 of the pair acts).
// k_band_schur: segment 0's separator triangle takes segment 1's partial Schur complement, which
// holds the separator's entries in reversed order (separator row i is segment 0's row nf + i and
// segment 1's row n - 1 - i; the band offset i - j is the same in both).
__global__ void k_band_schur(BandArgs A, const int32_t* __restrict__ active, int32_t astride) {
  const BandSeg S = A.seg[blockIdx.x];
  if (S.partner < 0 || S.dir < 0) return;
  if (active && !active[(size_t)S.blk * astride]) return;
  const BandSeg P = A.seg[S.partner];
  const int B1 = S.b + 1, ns = S.n - S.nf;
  double* LA = A.band + S.boff;
  const double* LB = A.band + P.boff;
  for (int t0 = threadIdx.x; t0 < ns * B1; t0 += 8 * blockDim.x) {  // 8 entries per thread per round trip
    double va[8], vb[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int t = t0 + u * blockDim.x, i = t / B1, d = t - i * B1;  // entry (i, i - d) of the separator
      const bool on = t < ns * B1 && d <= i;
      va[u] = on ? LA[(size_t)(S.nf + i - d) * B1 + d] : 0.0;
      vb[u] = on ? LB[(size_t)(P.n - 1 - i) * B1 + d] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int t = t0 + u * blockDim.x, i = t / B1, d = t - i * B1;
      if (t < ns * B1 && d <= i) LA[(size_t)(S.nf + i - d) * B1 + d] = va[u] + vb[u];
    }
  }
}

// k_band_sep: copy 0 -- segment 0's separator right-hand side (after both forward sweeps) takes
// segment 1's; copy 1 -- segment 1's separator rows take the separator's solution.
__global__ void k_band_sep(BandArgs A, int32_t copy, const int32_t* __restrict__ active, int32_t astride) {
  const BandSeg S = A.seg[blockIdx.x];
  if (S.partner < 0 || S.dir < 0) return;
  if (active && !active[(size_t)S.blk * astride]) return;
  const BandSeg P = A.seg[S.partner];
  double* ya = A.y + S.yoff + S.nf;
  double* yb = A.y + P.yoff + (P.n - 1);
  for (int i = threadIdx.x; i < S.n - S.nf; i += blockDim.x) {
    if (copy) yb[-i] = ya[i];
    else ya[i] += yb[-i];
  }
}

// du[perm[position]] from each segment's own rows (k < zc: every band position once); 0 for the
// rows of blocks whose active[blk * astride] is 0
__global__ void k_band_scatter(int32_t ncols, const int32_t* __restrict__ cseg, const int32_t* __restrict__ perm,
                               BandArgs A, const int32_t* __restrict__ active, int32_t astride,
                               double* __restrict__ du) {
  const int J = blockIdx.x * blockDim.x + threadIdx.x;
  if (J >= ncols) return;
  const BandSeg S = A.seg[cseg[J]];
  const int k = J - S.yoff;
  if (k >= S.zc) return;
  const bool on = active == nullptr || active[(size_t)S.blk * astride] != 0;
  du[perm[S.base + S.dir * k]] = on ? A.y[J] : 0.0;
}

// each segment's right-hand side only (a solve with the factors already in the band)
__global__ void k_band_rhs(int32_t ncols, const int32_t* __restrict__ cseg, const int32_t* __restrict__ perm,
                           const double* __restrict__ rhs, BandArgs A) {
  const int J = blockIdx.x * blockDim.x + threadIdx.x;
  if (J >= ncols) return;
  const BandSeg S = A.seg[cseg[J]];
  const int k = J - S.yoff;
  A.y[J] = k < S.zc ? rhs[perm[S.base + S.dir * k]] : 0.0;
}

}  // namespace

struct BandWork {
  uint64_t key = 0;       // problem uid, block layout and cut setting it was built for
  int32_t nblk = 0, n = 0, bmax = 0;
  int32_t nseg = 0, ncut = 0, ncols = 0;  // segments, blocks cut in two, band columns of all segments
  int64_t nband = 0;
  std::vector<int32_t> bw_h, bstart_h;
  DevBuf<int32_t> perm, iperm, cseg;
  DevBuf<BandSeg> seg;
  DevBuf<double> band, y;
  DevBuf<int32_t> bad;
  DevBuf<int32_t> bok;    // per block: 1 once the band holds its factors
  bool factored = false;  // a factorisation has been enqueued since the layout was built
};

void band_free(mpf_problem* p) {
  delete p->band;
  p->band = nullptr;
}

// Node order of each node range [node_off[m], node_off[m + 1]) expanded to free dofs (a node's dofs
// in order), and the half-bandwidth of each block: of four candidate orders the one with the
// smallest half-bandwidth — reverse Cuthill-McKee on the coupled-node graph (best on elongated or
// irregular meshes) and the nodes sorted by position with each axis in turn slowest (slices
// across that axis; best on block-shaped structured meshes).
mpf_status band_prepare(mpf_problem* p, const std::vector<int32_t>& node_off) {
  const PatternBufs& P = p->pat;
  if (!p->prepared) return MPF_ERR_STATE;
  if (P.presDOF > 0) return MPF_ERR_ARG;  // displacement systems only (no pivoting)
  const int nblk = (int)node_off.size() - 1;
  const uint64_t key =
      p->uid * 1000003ull + (uint64_t)nblk * 7919ull + (uint64_t)node_off.back() * 2ull + (p->band_split ? 1 : 0);
  if (p->band && p->band->key == key) return MPF_OK;
  if (!p->band) p->band = new BandWork();
  BandWork& W = *p->band;
  W.factored = false;
  hipStream_t s = p->stream;
  const int nNode = p->nNode;
  std::vector<int32_t> nbptr(nNode + 1), nda(3 * (size_t)nNode);
  std::vector<double> X(3 * (size_t)nNode);
  MPF_HIP(hipMemcpyAsync(nbptr.data(), P.nbptr.ptr, nbptr.size() * sizeof(int32_t), hipMemcpyDeviceToHost, s));
  MPF_HIP(hipMemcpyAsync(nda.data(), P.nda.ptr, nda.size() * sizeof(int32_t), hipMemcpyDeviceToHost, s));
  MPF_HIP(hipMemcpyAsync(X.data(), p->X.ptr, X.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  MPF_HIP(hipStreamSynchronize(s));
  std::vector<int32_t> nb(nbptr[nNode] > 0 ? nbptr[nNode] : 1);
  MPF_HIP(hipMemcpy(nb.data(), P.nb.ptr, (size_t)nbptr[nNode] * sizeof(int32_t), hipMemcpyDeviceToHost));
  const int n = P.dispDOF;
  std::vector<int32_t> perm(n), iperm(n), bstart(nblk + 1, 0), bw(nblk, 0);
  std::vector<int32_t> first(nNode, -1), last(nNode, -1), mark(nNode, 0), level;
  auto deg = [&](int v) { return nbptr[v + 1] - nbptr[v]; };
  auto nfree = [&](int v) {
    return (nda[3 * (size_t)v] >= 0) + (nda[3 * (size_t)v + 1] >= 0) + (nda[3 * (size_t)v + 2] >= 0);
  };
  // half-bandwidth of a node order of [a, z)
  auto bandwidth = [&](const std::vector<int32_t>& order, int a, int z) {
    int q = 0;
    for (int v : order) {
      const int f = nfree(v);
      first[v] = f ? q : -1;
      last[v] = f ? q + f - 1 : -1;
      q += f;
    }
    int b = 0;
    for (int u = a; u < z; ++u) {
      if (last[u] < 0) continue;
      for (int k = nbptr[u]; k < nbptr[u + 1]; ++k) {
        const int v = nb[k];
        if (first[v] >= 0 && first[v] <= last[u]) b = std::max(b, last[u] - first[v]);
      }
    }
    return b;
  };
  int pos = 0;
  for (int m = 0; m < nblk; ++m) {
    const int a = node_off[m], z = node_off[m + 1];
    bstart[m] = pos;
    // reverse Cuthill-McKee: per connected component, breadth-first order from a pseudo-peripheral
    // start (the last node reached from the previous start, twice), neighbours in increasing
    // degree; then reversed
    std::vector<int32_t> best;
    {
      std::vector<int32_t> order;
      for (int v0 = a; v0 < z; ++v0) {
        if (mark[v0]) continue;
        int start = v0;
        for (int sweep = 0; sweep < 3; ++sweep) {
          level.assign(1, start);
          mark[start] = 2;
          for (size_t h = 0; h < level.size(); ++h) {
            const int u = level[h];
            const size_t from = level.size();
            for (int k = nbptr[u]; k < nbptr[u + 1]; ++k) {
              const int v = nb[k];
              if (v >= a && v < z && mark[v] == 0) { mark[v] = 2; level.push_back(v); }
            }
            std::stable_sort(level.begin() + from, level.end(), [&](int x, int y) { return deg(x) < deg(y); });
          }
          if (sweep < 2) {
            for (int v : level) mark[v] = 0;
            start = level.back();
          }
        }
        for (int v : level) mark[v] = 1;
        order.insert(order.end(), level.begin(), level.end());
      }
      std::reverse(order.begin(), order.end());
      best = order;
    }
    int bbest = bandwidth(best, a, z);
    // slice index of every node along each axis: the nodes sorted by that coordinate, a new slice
    // wherever the gap to the previous node exceeds half the largest gap (on a jittered structured
    // grid, exactly its grid lines)
    std::vector<int64_t> qk(3 * (size_t)(z - a), 0);
    for (int d = 0; d < 3; ++d) {
      std::vector<int32_t> srt(z - a);
      for (int v = a; v < z; ++v) srt[v - a] = v;
      std::stable_sort(srt.begin(), srt.end(),
                       [&](int x, int y) { return X[3 * (size_t)x + d] < X[3 * (size_t)y + d]; });
      double gmax = 0.0;
      for (size_t t = 1; t < srt.size(); ++t)
        gmax = std::max(gmax, X[3 * (size_t)srt[t] + d] - X[3 * (size_t)srt[t - 1] + d]);
      int64_t sl = 0;
      for (size_t t = 0; t < srt.size(); ++t) {
        if (t > 0 && X[3 * (size_t)srt[t] + d] - X[3 * (size_t)srt[t - 1] + d] > 0.5 * gmax) ++sl;
        qk[3 * (size_t)(srt[t] - a) + d] = sl;
      }
    }
    for (int ax = 0; ax < 3; ++ax) {  // slices across axis ax
      std::vector<int32_t> order(z - a);
      for (int v = a; v < z; ++v) order[v - a] = v;
      const int a1 = (ax + 1) % 3, a2 = (ax + 2) % 3;
      std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        const int64_t* px = &qk[3 * (size_t)(x - a)];
        const int64_t* py = &qk[3 * (size_t)(y - a)];
        if (px[ax] != py[ax]) return px[ax] < py[ax];
        if (px[a1] != py[a1]) return px[a1] < py[a1];
        return px[a2] < py[a2];
      });
      const int bb = bandwidth(order, a, z);
      if (bb < bbest) { bbest = bb; best = order; }
    }
    for (int v : best) {
      for (int d = 0; d < 3; ++d) {
        const int r = nda[3 * (size_t)v + d];
        if (r < 0) continue;
        perm[pos] = r;
        iperm[r] = pos;
        ++pos;
      }
    }
    bw[m] = bbest;
  }
  bstart[nblk] = pos;
  if (pos != n) return MPF_ERR_STATE;  // every free dof belongs to one block
  // segments: a block of nb >= 3 b + 64 rows is cut at a separator of b <= ns < b + kW rows in its
  // middle (A1 and A2 whole panels); the segments are then ordered longest first (by columns
  // factored before the separator), so that the longest start first
  std::vector<BandSeg> seg;
  int ncut = 0;
  for (int m = 0; m < nblk; ++m) {
    const int nb = bstart[m + 1] - bstart[m], b = bw[m];
    BandSeg g{};
    g.blk = m;
    g.b = b;
    g.partner = -1;
    if (p->band_split && nb >= 3 * b + 4 * kW) {
      const int a1 = kW * ((nb - b) / (2 * kW)), a2 = kW * ((nb - a1 - b) / kW), ns = nb - a1 - a2;
      BandSeg h = g;
      g.base = bstart[m];
      g.dir = 1;
      g.n = a1 + ns;
      g.nf = a1;
      g.nend = g.zc = a1 + ns;
      g.partner = (int)seg.size() + 1;
      h.base = bstart[m] + nb - 1;
      h.dir = -1;
      h.n = a2 + ns;
      h.nf = h.nend = h.zc = a2;
      h.partner = (int)seg.size();
      seg.push_back(g);
      seg.push_back(h);
      ++ncut;
    } else {
      g.base = bstart[m];
      g.dir = 1;
      g.n = g.nf = g.nend = g.zc = nb;
      seg.push_back(g);
    }
  }
  std::vector<int32_t> order(seg.size()), where(seg.size());
  for (size_t i = 0; i < seg.size(); ++i) order[i] = (int32_t)i;
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return seg[x].nf > seg[y].nf; });
  for (size_t i = 0; i < order.size(); ++i) where[order[i]] = (int32_t)i;
  std::vector<BandSeg> sorted(seg.size());
  int64_t nband = 0;
  int32_t ncols = 0;
  for (size_t i = 0; i < order.size(); ++i) {
    BandSeg g = seg[order[i]];
    if (g.partner >= 0) g.partner = where[g.partner];
    g.boff = nband;
    g.yoff = ncols;
    nband += (int64_t)g.n * (g.b + 1);
    ncols += g.n;
    sorted[i] = g;
  }
  std::vector<int32_t> cseg(std::max(ncols, 1), 0);
  for (size_t i = 0; i < sorted.size(); ++i)
    for (int k = 0; k < sorted[i].n; ++k) cseg[sorted[i].yoff + k] = (int32_t)i;
  W.nblk = nblk;
  W.n = n;
  W.nseg = (int32_t)sorted.size();
  W.ncut = ncut;
  W.ncols = ncols;
  W.bmax = nblk > 0 ? *std::max_element(bw.begin(), bw.end()) : 0;
  W.nband = nband;
  W.bw_h = bw;
  W.bstart_h = bstart;
  const int lds = (int)(((size_t)(kW + W.bmax) * kW + 3 * kW) * sizeof(double));
  if (lds > 64 * 1024) return MPF_ERR_ARG;  // half-bandwidth beyond the panel's LDS (about 490 dofs)
  if (W.bmax >= kBS / 2) return MPF_ERR_ARG;  // each half of the solve's workgroup gives a coupling row one thread
  static_assert(kW + 493 <= kFBS, "the wide-band panel path gives each panel row one thread");
  MPF_TRY(W.perm.upload(perm.data(), perm.size(), s));
  MPF_TRY(W.iperm.upload(iperm.data(), iperm.size(), s));
  MPF_TRY(W.cseg.upload(cseg.data(), cseg.size(), s));
  if (!sorted.empty()) MPF_TRY(W.seg.upload(sorted.data(), sorted.size(), s));
  MPF_TRY(W.band.alloc(W.nband > 0 ? (size_t)W.nband : 1));
  MPF_TRY(W.y.alloc(ncols > 0 ? (size_t)ncols : 1));
  MPF_TRY(W.bad.alloc(1));
  {
    const std::vector<int32_t> zero(std::max(nblk, 1), 0);
    MPF_TRY(W.bok.upload(zero.data(), zero.size(), s));
  }
  MPF_HIP(hipStreamSynchronize(s));
  W.key = key;
  return MPF_OK;
}

// Solve phases of every segment with the factors in the band (active as for the factorisation).
static void enqueue_band_solves(const BandWork* W, const BandArgs& A, const int32_t* active, int32_t astride,
                                hipStream_t s) {
  const dim3 g(std::max(W->nseg, 1));
  hipLaunchKernelGGL(k_band_solve, g, dim3(kBS), 0, s, A, 0, active, astride);
  if (W->ncut > 0) {
    hipLaunchKernelGGL(k_band_sep, g, dim3(256), 0, s, A, 0, active, astride);
    hipLaunchKernelGGL(k_band_solve, g, dim3(kBS), 0, s, A, 1, active, astride);
    hipLaunchKernelGGL(k_band_sep, g, dim3(256), 0, s, A, 1, active, astride);
  }
  hipLaunchKernelGGL(k_band_solve, g, dim3(kBS), 0, s, A, 2, active, astride);
}

// Enqueue: band from the assembled K, factor and solve every block, du = K^-1 rhs per block (0 for
// blocks whose active[m * astride] is 0; active may be null).  Allocates nothing (graph-capturable).
mpf_status band_enqueue_solve(mpf_problem* p, const int32_t* active, int32_t astride) {
  BandWork* W = p->band;
  if (!W) return MPF_ERR_STATE;
  if (W->nseg == 0) return MPF_OK;  // no free dofs
  hipStream_t s = p->stream;
  const PatternBufs& P = p->pat;
  const BandArgs A{W->seg.ptr, W->band.ptr, W->y.ptr, W->bad.ptr};
  static_assert(kFillCols * 494 * sizeof(double) <= 64 * 1024, "the fill's LDS holds columns of b <= 493");
  hipLaunchKernelGGL(k_band_fill, dim3(std::max((W->ncols + kFillCols - 1) / kFillCols, 1)), dim3(256),
                     (size_t)kFillCols * (W->bmax + 1) * sizeof(double), s, W->ncols, W->bmax + 1, W->cseg.ptr,
                     P.rowptr.ptr, P.colidx.ptr, p->vals.ptr, p->rhs.ptr, W->perm.ptr, W->iperm.ptr, A, active,
                     astride);
  const size_t lds = (kW + W->bmax <= 3 * 64 ? 2 * ((size_t)(kW + W->bmax) * kW + kW)
                                             : (size_t)(kW + W->bmax) * kW + 3 * kW) * sizeof(double);
  const dim3 g(std::max(W->nseg, 1));
  hipLaunchKernelGGL(k_band_factor, g, dim3(kFBS), lds, s, A, W->bmax, 0, active, astride, W->bok.ptr);
  if (W->ncut > 0) {
    hipLaunchKernelGGL(k_band_schur, g, dim3(1024), 0, s, A, active, astride);
    hipLaunchKernelGGL(k_band_factor, g, dim3(kFBS), lds, s, A, W->bmax, 1, active, astride, W->bok.ptr);
  }
  enqueue_band_solves(W, A, active, astride, s);
  hipLaunchKernelGGL(k_band_scatter, dim3(grid_for(std::max(W->ncols, 1), 256)), dim3(256), 0, s, W->ncols,
                     W->cseg.ptr, W->perm.ptr, A, active, astride, p->du.ptr);
  MPF_HIP(hipGetLastError());
  W->factored = true;
  return MPF_OK;
}

// Enqueue: du = K^-1 rhs for every block with the factors left by the last solve of that block
// (each block's own last factorisation; blocks skipped since keep theirs); du = 0 on a block that
// was never factored (a mesh converged at its start).  MPF_ERR_STATE before any factorisation.
mpf_status band_enqueue_resolve(mpf_problem* p) {
  BandWork* W = p->band;
  if (!W || !W->factored) return MPF_ERR_STATE;
  if (W->nseg == 0) return MPF_OK;  // no free dofs
  hipStream_t s = p->stream;
  const BandArgs A{W->seg.ptr, W->band.ptr, W->y.ptr, W->bad.ptr};
  hipLaunchKernelGGL(k_band_rhs, dim3(grid_for(std::max(W->ncols, 1), 256)), dim3(256), 0, s, W->ncols,
                     W->cseg.ptr, W->perm.ptr, p->rhs.ptr, A);
  enqueue_band_solves(W, A, W->bok.ptr, 1, s);
  hipLaunchKernelGGL(k_band_scatter, dim3(grid_for(std::max(W->ncols, 1), 256)), dim3(256), 0, s, W->ncols,
                     W->cseg.ptr, W->perm.ptr, A, W->bok.ptr, 1, p->du.ptr);
  MPF_HIP(hipGetLastError());
  return MPF_OK;
}

// Single problem (mpf_solve with MPF_KSP_BAND): one block holding every node.
mpf_status band_solve(mpf_problem* p, mpf_krylov_report* rep) {
  const auto t0 = std::chrono::steady_clock::now();
  MPF_TRY(p->du.alloc(p->pat.dispDOF > 0 ? p->pat.dispDOF : 1));
  const std::vector<int32_t> off{0, p->nNode};
  MPF_TRY(band_prepare(p, off));
  hipStream_t s = p->stream;
  MPF_TRY(band_enqueue_solve(p, nullptr, 0));
  int32_t bad = 0;
  MPF_HIP(hipMemcpyAsync(&bad, p->band->bad.ptr, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  MPF_HIP(hipStreamSynchronize(s));
  rep->iterations = 1;
  rep->converged = bad ? 0 : 1;
  rep->relres = 0.0;
  rep->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return bad ? MPF_ERR_DIVERGED : MPF_OK;
}

int32_t* band_bad_flag(mpf_problem* p) { return p->band ? p->band->bad.ptr : nullptr; }

}  // namespace mpf

extern "C" mpf_status mpf_band_stats(const mpf_problem* p, int32_t* nblk, int32_t* bmax, int64_t* ndof,
                                     double* flops) {
  if (!p || !p->band || p->band->key == 0) return MPF_ERR_STATE;
  const mpf::BandWork& W = *p->band;
  double f = 0.0;  // per column of each block: b (b + 1) / 2 multiply-adds in the trailing update
  for (int m = 0; m < W.nblk; ++m)
    f += (double)(W.bstart_h[m + 1] - W.bstart_h[m]) * (double)W.bw_h[m] * (double)(W.bw_h[m] + 1);
  if (nblk) *nblk = W.nblk;
  if (bmax) *bmax = W.bmax;
  if (ndof) *ndof = W.n;
  if (flops) *flops = f;
  return MPF_OK;
}

extern "C" mpf_status mpf_set_band_split(mpf_problem* p, int32_t enable) {
  if (!p || (enable != 0 && enable != 1)) return MPF_ERR_ARG;
  p->band_split = enable;  // the layout is rebuilt at the next direct solve (its key holds the setting)
  return MPF_OK;
}

extern "C" mpf_status mpf_band_segments(const mpf_problem* p, int32_t* nseg, int32_t* ncut) {
  if (!p || !p->band || p->band->key == 0) return MPF_ERR_STATE;
  if (nseg) *nseg = p->band->nseg;
  if (ncut) *ncut = p->band->ncut;
  return MPF_OK;
}
