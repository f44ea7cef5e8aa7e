// mpnn.hip -- the learned initial guess: a small message-passing network on the mesh graph, trained
// and run with hand-written BF16 MFMA kernels (FP32 accumulation) on gfx950.
//
// No counterpart in the reference.  Per load step k+1 of a problem, the network sees at every node
// the last converged displacement u_k, the last increment u_k - u_{k-1} (scaled per mesh), the load
// factors and the node position, and predicts the correction of the linear extrapolation
//   u_{k+1} ~ u_k + rho (u_k - u_{k-1}),   rho = (lam_{k+1} - lam_k) / (lam_k - lam_{k-1}),
// so the Newton start point is  u_k + rho (u_k - u_{k-1}) + s_m y  (s_m: per-mesh scale).
//
// Network (FIN = 32 input features, W = 64 hidden, 3 message-passing layers, FOUT = 32 outputs of
// which 3 are used):
//   h0 = relu(x W0 + b0)
//   a_l = mean_{j in NB(i)} h_{l-1,j}                        (NB: nodes sharing an element, self included)
//   h_l = h_{l-1} + relu([h_{l-1} | a_l] W_l + b_l),  l = 1..3
//   y  = h3 Wo + bo
// Weights live in FP32 (Adam) with BF16 copies for the MFMA operands; activations feeding a GEMM are
// BF16, every GEMM accumulates in FP32, the residual stream h is kept in FP32.
//
// GEMM kernels use __builtin_amdgcn_mfma_f32_16x16x32_bf16 (D 16x16 += A 16x32 * B 32x16): lane l
// holds A[l % 16][8 (l / 16) .. +7], B[8 (l / 16) .. +7][l % 16] and D[4 (l / 16) + i][l % 16], i < 4.

#include <algorithm>
#include <cmath>
#include <vector>

#include "mpf_internal.hpp"

mpf_status enqueue_assembly(mpf_problem* p, bool first_iter, double lam, int scatterK,
                            const double* lam_dev = nullptr);  // problem.hip

namespace mpf {
mpf_status spmv(const int32_t* rowptr, const int32_t* colidx, const double* vals, int64_t n, int64_t nnz,
                const double* x, double* y, double* part, unsigned np, const double* sc, hipStream_t s);  // krylov.hip
}

typedef short bf16x8_s __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace mpf {

constexpr int kFin = 32, kW = 64, kLayers = 3, kFout = 32;

struct MpnnParams {
  // offsets (floats) into the flat parameter vector
  static constexpr int W0 = 0, B0 = W0 + kFin * kW;
  static constexpr int WL = B0 + kW;                      // layer l: WL + l * (2W*W + W)
  static constexpr int LSZ = 2 * kW * kW + kW;
  static constexpr int WO = WL + kLayers * LSZ, BO = WO + kW * kFout;
  static constexpr int N = BO + kFout;
};

}  // namespace mpf

struct mpf_mpnn {
  mpf::DevBuf<float> w, g, m1, m2;        // parameters, gradients, Adam moments (flat, MpnnParams layout)
  mpf::DevBuf<uint16_t> wb, wtb;          // BF16 copies: every weight matrix as [K][C] and transposed [C][K]
  int64_t cap = 0;                         // node capacity of the activation buffers
  mpf::DevBuf<uint16_t> x;                 // [N][FIN] bf16 (input)
  mpf::DevBuf<float> h[mpf::kLayers + 1];  // residual stream [N][W] fp32
  mpf::DevBuf<uint16_t> ab[mpf::kLayers + 1];  // bf16 [h_l | a_{l+1}] (the next GEMM's input rows)
  mpf::DevBuf<uint16_t> r[mpf::kLayers + 1];  // relu outputs (bf16) -> relu masks
  mpf::DevBuf<float> y, dy, dh, dcat;      // [N][FOUT], [N][FOUT], [N][W], [N][2W]
  mpf::DevBuf<uint16_t> db16;              // bf16 copy of the current upstream gradient [N][2W max]
  mpf::DevBuf<float> part;                 // weight-gradient partials
  mpf::DevBuf<double> red;                 // loss partials
  hipStream_t stream = nullptr;
  int64_t adam_t = 0;
  // per-problem context of the learned-guess data path
  uint64_t ctx = 0;  // uid of the problem the context below was built for
  int32_t nmesh = 0;
  mpf::DevBuf<int32_t> node_off, nmesh_of, fixed;
  mpf::DevBuf<double> xref;
  mpf::DevBuf<float> scale;
  std::vector<int32_t> doff;          // per mesh: first free dof (free dofs of a mesh are contiguous)
  mpf::DevBuf<int32_t> doff_d;
  mpf::DevBuf<double> wres;           // per mesh: 1 / ||f_m||^2 over free dofs (unit external load)
  mpf::DevBuf<double> ebuf, rbuf, gbuf, msq;  // residual-loss work vectors
  mpf::DevBuf<double> usave, gbuf2;           // guarded guess: saved state, second candidate
  mpf::DevBuf<double> rnet, mdot;             // guarded guess: residual with the correction, per-mesh dots
  mpf::DevBuf<float> wmesh;                   // guarded guess: per-mesh weight of the correction
};

namespace mpf {
namespace {

__device__ __forceinline__ uint16_t f2bf(float f) {  // round to nearest even
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// ---- GEMM over rows: C[N][COUT] = A[N][KIN] (bf16, row stride lda) * B[KIN][COUT] (bf16, given as
// its transpose Bt[COUT][KIN] so each lane reads 8 consecutive k), one wavefront per 16 rows ----
enum { EPI_ENC = 0, EPI_MP = 1, EPI_OUT = 2, EPI_PLAIN = 3 };

struct MMArgs {
  int64_t n;
  const uint16_t* A;
  int lda;
  const uint16_t* Bt;
  const float* bias;
  const float* hprev;  // EPI_MP: residual input [N][W]
  float* out;          // fp32 output [N][COUT]
  uint16_t* outb;      // bf16 copy of out (row stride ldob), or null
  int ldob;
  uint16_t* rb;        // EPI_MP / EPI_ENC: relu output (bf16) [N][COUT]
};

template <int KIN, int COUT, int EPI>
__global__ void __launch_bounds__(256) k_mm(MMArgs a) {
  static_assert(KIN % 32 == 0 && COUT % 16 == 0, "MFMA tiles");
  __shared__ bf16x8_s sB[COUT * KIN / 8];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int t = tid; t < COUT * KIN / 8; t += 256) sB[t] = reinterpret_cast<const bf16x8_s*>(a.Bt)[t];
  __syncthreads();
  const int64_t r0 = ((int64_t)blockIdx.x * 4 + wave) * 16;
  if (r0 >= a.n) return;
  constexpr int NT = COUT / 16;
  f32x4 acc[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int64_t ra = r0 + (lane & 15);
  const int kg = lane >> 4;
#pragma unroll
  for (int kk = 0; kk < KIN / 32; ++kk) {
    bf16x8_s av = {0, 0, 0, 0, 0, 0, 0, 0};
    if (ra < a.n) av = *reinterpret_cast<const bf16x8_s*>(a.A + ra * a.lda + kk * 32 + kg * 8);
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const bf16x8_s bv = sB[((t * 16 + (lane & 15)) * KIN + kk * 32 + kg * 8) / 8];
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, av), __builtin_bit_cast(bf16x8, bv),
                                                      acc[t], 0, 0, 0);
    }
  }
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const int c = t * 16 + (lane & 15);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int64_t row = r0 + kg * 4 + i;
      if (row >= a.n) continue;
      float v = acc[t][i];
      if (EPI != EPI_PLAIN) v += a.bias[c];
      if (EPI == EPI_ENC || EPI == EPI_MP) {
        const float rr = v > 0.f ? v : 0.f;
        a.rb[row * COUT + c] = f2bf(rr);
        v = (EPI == EPI_MP) ? a.hprev[row * COUT + c] + rr : rr;
      }
      a.out[row * COUT + c] = v;
      if (a.outb) a.outb[row * a.ldob + c] = f2bf(v);
    }
  }
}

// ---- weight gradient: part[blk][K][C] = sum over the block's rows of A[r][k] * D[r][c] ----
// A bf16 [N][K] (row stride lda), D bf16 [N][C] (row stride ldd); rows in chunks of 32 staged in LDS
// transposed so both MFMA operands are 8 consecutive rows.
template <int K, int C>
__global__ void __launch_bounds__(256) k_wgrad(int64_t n, int64_t rows_per_block, const uint16_t* __restrict__ A,
                                               int lda, const uint16_t* __restrict__ D, int ldd,
                                               float* __restrict__ part) {
  constexpr int NT = (K / 16) * (C / 16);
  constexpr int TPW = (NT + 3) / 4;  // tiles per wave
  __shared__ uint16_t sAt[K][40];  // [k][r] (32 rows, padded)
  __shared__ uint16_t sDt[C][40];  // [c][r]
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t rb0 = (int64_t)blockIdx.x * rows_per_block;
  const int64_t rb1 = min(n, rb0 + rows_per_block);
  f32x4 acc[TPW];
#pragma unroll
  for (int t = 0; t < TPW; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int64_t c0 = rb0; c0 < rb1; c0 += 32) {
    __syncthreads();
    for (int t = tid; t < 32 * K; t += 256) {
      const int r = t / K, k = t % K;
      sAt[k][r] = (c0 + r < rb1) ? A[(c0 + r) * lda + k] : (uint16_t)0;
    }
    for (int t = tid; t < 32 * C; t += 256) {
      const int r = t / C, c = t % C;
      sDt[c][r] = (c0 + r < rb1) ? D[(c0 + r) * ldd + c] : (uint16_t)0;
    }
    __syncthreads();
    const int kg = lane >> 4;
#pragma unroll
    for (int q = 0; q < TPW; ++q) {
      const int t = wave * TPW + q;
      if (t >= NT) break;
      const int tk = t / (C / 16), tc = t % (C / 16);
      bf16x8_s av, bv;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        av[j] = (short)sAt[tk * 16 + (lane & 15)][kg * 8 + j];
        bv[j] = (short)sDt[tc * 16 + (lane & 15)][kg * 8 + j];
      }
      acc[q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, av), __builtin_bit_cast(bf16x8, bv),
                                                      acc[q], 0, 0, 0);
    }
  }
  float* P = part + (size_t)blockIdx.x * K * C;
#pragma unroll
  for (int q = 0; q < TPW; ++q) {
    const int t = wave * TPW + q;
    if (t >= NT) break;
    const int tk = t / (C / 16), tc = t % (C / 16);
#pragma unroll
    for (int i = 0; i < 4; ++i) P[(tk * 16 + (lane >> 4) * 4 + i) * C + tc * 16 + (lane & 15)] = acc[q][i];
  }
}

// g[K*C] += sum over blocks of part (block order); bias gradient g_b[c] += sum_r D32[r][c] (fp32 D)
__global__ void k_wgrad_reduce(int nblk, int KC, const float* __restrict__ part, float* __restrict__ g) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= KC) return;
  float s = 0.f;
  for (int b = 0; b < nblk; ++b) s += part[(size_t)b * KC + i];
  g[i] += s;
}

__global__ void __launch_bounds__(256) k_bgrad(int64_t n, int C, const float* __restrict__ D, int ldd,
                                               int64_t rows_per_block, float* __restrict__ part) {
  // part[blk][C]
  const int64_t r0 = (int64_t)blockIdx.x * rows_per_block, r1 = min(n, r0 + rows_per_block);
  for (int c = threadIdx.x; c < C; c += blockDim.x) {
    float s = 0.f;
    for (int64_t r = r0; r < r1; ++r) s += D[r * ldd + c];
    part[(size_t)blockIdx.x * C + c] = s;
  }
}

// ---- graph aggregation (mean over NB(i)), bf16 -> bf16 with fp32 sums; 8 channels per thread ----
__global__ void k_agg_fwd(int64_t n, const int32_t* __restrict__ nbptr, const int32_t* __restrict__ nb,
                          const uint16_t* __restrict__ hb, int ldh, uint16_t* __restrict__ out, int ldo) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t i = t >> 3;
  const int cg = (int)(t & 7);
  if (i >= n) return;
  float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  const int j0 = nbptr[i], j1 = nbptr[i + 1];
  for (int q = j0; q < j1; ++q) {
    const bf16x8_s v = *reinterpret_cast<const bf16x8_s*>(hb + (int64_t)nb[q] * ldh + cg * 8);
#pragma unroll
    for (int k = 0; k < 8; ++k) s[k] += bf2f((uint16_t)v[k]);
  }
  const float inv = j1 > j0 ? 1.f / (float)(j1 - j0) : 0.f;
  bf16x8_s o;
#pragma unroll
  for (int k = 0; k < 8; ++k) o[k] = (short)f2bf(s[k] * inv);
  *reinterpret_cast<bf16x8_s*>(out + i * ldo + cg * 8) = o;
}

// dh[j] += sum_{i in NB(j)} da[i] / |NB(i)|   (NB symmetric), fp32
__global__ void k_agg_bwd(int64_t n, const int32_t* __restrict__ nbptr, const int32_t* __restrict__ nb,
                          const float* __restrict__ da, int ldda, float* __restrict__ dh) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t j = t >> 6;
  const int c = (int)(t & 63);
  if (j >= n) return;
  float s = 0.f;
  for (int q = nbptr[j]; q < nbptr[j + 1]; ++q) {
    const int i = nb[q];
    s += da[(int64_t)i * ldda + c] / (float)(nbptr[i + 1] - nbptr[i]);
  }
  dh[j * kW + c] += s;
}

// dst[i][c] = mask(r[i][c] > 0) * src[i][c] (fp32) and a bf16 copy
__global__ void k_relu_bwd(int64_t n, int C, const float* __restrict__ src, int lds, const uint16_t* __restrict__ r,
                           float* __restrict__ dst, uint16_t* __restrict__ dstb) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * C) return;
  const int64_t i = t / C;
  const int c = (int)(t % C);
  const float v = r[t] != 0 ? src[i * lds + c] : 0.f;
  dst[t] = v;
  dstb[t] = f2bf(v);
}

__global__ void k_fill0(int64_t n, double* __restrict__ a) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) a[t] = 0.0;
}

__global__ void k_to_bf16(int64_t n, const float* __restrict__ a, uint16_t* __restrict__ b) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) b[t] = f2bf(a[t]);
}

__global__ void k_add(int64_t n, const float* __restrict__ a, int lda, int C, float* __restrict__ b) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * C) return;
  b[t] += a[(t / C) * lda + t % C];
}

// bf16 weight copies [K][C] and [C][K] from the fp32 parameter vector
__global__ void k_wcast(const float* __restrict__ w, uint16_t* __restrict__ wb, uint16_t* __restrict__ wtb) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  // matrices: W0 [FIN][W], layer l [2W][W], Wo [W][FOUT]; layout in wb / wtb follows the same order
  int off = 0, K = 0, C = 0, loc = -1, woff = 0;
  const int sizes[kLayers + 2][2] = {{kFin, kW}, {2 * kW, kW}, {2 * kW, kW}, {2 * kW, kW}, {kW, kFout}};
  const int woffs[kLayers + 2] = {MpnnParams::W0, MpnnParams::WL, MpnnParams::WL + MpnnParams::LSZ,
                                  MpnnParams::WL + 2 * MpnnParams::LSZ, MpnnParams::WO};
  for (int mtx = 0; mtx < kLayers + 2; ++mtx) {
    const int sz = sizes[mtx][0] * sizes[mtx][1];
    if (t >= off && t < off + sz) { loc = t - off; K = sizes[mtx][0]; C = sizes[mtx][1]; woff = woffs[mtx]; break; }
    off += sz;
  }
  if (loc < 0) return;
  const int k = loc / C, c = loc % C;
  const uint16_t b = f2bf(w[woff + loc]);
  wb[off + loc] = b;
  wtb[off + c * K + k] = b;
}

__global__ void k_adam(int n, float* __restrict__ w, float* __restrict__ g, float* __restrict__ m1,
                       float* __restrict__ m2, float lr, float b1, float b2, float eps, float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float a = b1 * m1[i] + (1.f - b1) * gi;
  const float b = b2 * m2[i] + (1.f - b2) * gi * gi;
  m1[i] = a;
  m2[i] = b;
  w[i] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
  g[i] = 0.f;
}

// ---- learned-guess data path ----
// per-mesh scale s_m = max(max_i |u_k - u_{k-1}|_inf over mesh m, tiny)
__global__ void __launch_bounds__(256) k_mesh_scale(const int32_t* __restrict__ node_off, const double* __restrict__ uk,
                                                    const double* __restrict__ ukm1, float* __restrict__ scale) {
  __shared__ double sh[256];
  const int m = blockIdx.x;
  double mx = 0.0;
  for (int64_t i = (int64_t)node_off[m] * 3 + threadIdx.x; i < (int64_t)node_off[m + 1] * 3; i += 256)
    mx = fmax(mx, fabs(uk[i] - ukm1[i]));
  sh[threadIdx.x] = mx;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) scale[m] = (float)fmax(sh[0], 1e-30);
}

struct FeatArgs {
  int64_t n;
  const int32_t* nmesh_of;  // node -> mesh
  const float* scale;       // per mesh
  const double* X;
  const double* xref;       // per mesh: [3] origin, [1] length scale  (4 doubles per mesh)
  const double* uk;
  const double* ukm1;
  const int32_t* fixed;     // node has a prescribed dof
  float rho, lam_next, lam_ratio;
  uint16_t* x;              // [N][FIN]
};

__global__ void k_features(FeatArgs f) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= f.n) return;
  const int m = f.nmesh_of[i];
  const float s = f.scale[m];
  const double* xr = f.xref + 4 * m;
  float v[kFin];
#pragma unroll
  for (int c = 0; c < kFin; ++c) v[c] = 0.f;
  for (int d = 0; d < 3; ++d) {
    const double uk = f.uk[3 * i + d], du = uk - f.ukm1[3 * i + d];
    v[d] = (float)(uk / s) * 0.1f;   // displacement in increments, damped
    v[3 + d] = (float)(du / s);      // last increment, in [-1, 1]
    v[6 + d] = (float)((f.X[3 * i + d] - xr[d]) / xr[3]);
  }
  v[9] = f.rho;
  v[10] = f.lam_next;
  v[11] = f.lam_ratio;
  v[12] = f.fixed[i] ? 1.f : 0.f;
  v[13] = 1.f;
#pragma unroll
  for (int c = 0; c < kFin; ++c) f.x[i * kFin + c] = f2bf(v[c]);
}

// loss = mean over free dofs of (y - t)^2, t = (u_next - u_k - rho (u_k - u_{k-1})) / s_m;  dy = dloss/dy
__global__ void __launch_bounds__(256) k_loss(int64_t n, const int32_t* __restrict__ nmesh_of,
                                              const float* __restrict__ scale, const int32_t* __restrict__ nda,
                                              const double* __restrict__ unext, const double* __restrict__ uk,
                                              const double* __restrict__ ukm1, float rho, const float* __restrict__ y,
                                              float inv_count, float* __restrict__ dy, double* __restrict__ part) {
  __shared__ double sh[256];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double l = 0.0;
  if (i < n) {
    const float s = scale[nmesh_of[i]];
    for (int d = 0; d < kFout; ++d) {
      float g = 0.f;
      if (d < 3 && nda[3 * i + d] >= 0) {
        const double uk_ = uk[3 * i + d];
        const float t = (float)((unext[3 * i + d] - uk_ - (double)rho * (uk_ - ukm1[3 * i + d])) / s);
        const float e = y[i * kFout + d] - t;
        l += (double)e * e;
        g = 2.f * e * inv_count;
      }
      dy[i * kFout + d] = g;
    }
  }
  sh[threadIdx.x] = l;
  __syncthreads();
  for (int s2 = 128; s2 > 0; s2 >>= 1) {
    if (threadIdx.x < s2) sh[threadIdx.x] += sh[threadIdx.x + s2];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// guess = u_k + rho (u_k - u_{k-1}) + s_m y   (prescribed dofs are reset by the solver)
// guess = u_k + rho (u_k - u_{k-1}) + w_m s_m y, w_m = wmesh[m] (or use_net when wmesh is null)
__global__ void k_guess(int64_t n, const int32_t* __restrict__ nmesh_of, const float* __restrict__ scale,
                        const double* __restrict__ uk, const double* __restrict__ ukm1, double rho, int use_net,
                        const float* __restrict__ wmesh, const float* __restrict__ y, double* __restrict__ guess) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int m = nmesh_of[i];
  const double s = scale[m];
  const double wgt = wmesh ? (double)wmesh[m] : (use_net ? 1.0 : 0.0);
  for (int d = 0; d < 3; ++d) {
    const double u = uk[3 * i + d];
    guess[3 * i + d] = u + rho * (u - ukm1[3 * i + d]) + (wgt != 0.0 ? wgt * s * (double)y[i * kFout + d] : 0.0);
  }
}

__global__ void k_set_bc_values(int32_t nbc, const int32_t* __restrict__ dof, const double* __restrict__ val,
                                double lam, double* __restrict__ u) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nbc) u[dof[i]] = val[i] * lam;
}

// e[dof] = s_m y - (u_next - u_k - rho (u_k - u_{k-1})) on the free dofs (the prediction's error)
__global__ void k_err_free(int64_t n, const int32_t* __restrict__ nmesh_of, const float* __restrict__ scale,
                           const int32_t* __restrict__ nda, const float* __restrict__ y,
                           const double* __restrict__ unext, const double* __restrict__ uk,
                           const double* __restrict__ ukm1, double rho, double* __restrict__ e) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double s = scale[nmesh_of[i]];
  for (int d = 0; d < 3; ++d) {
    const int dof = nda[3 * i + d];
    if (dof < 0) continue;
    const double u = uk[3 * i + d];
    e[dof] = s * (double)y[i * kFout + d] - (unext[3 * i + d] - u - rho * (u - ukm1[3 * i + d]));
  }
}

// per mesh: sum of r^2 over its free-dof range
__global__ void __launch_bounds__(256) k_mesh_sq(const int32_t* __restrict__ doff, const double* __restrict__ r,
                                                 double* __restrict__ out) {
  __shared__ double sh[256];
  const int m = blockIdx.x;
  double v = 0.0;
  for (int64_t i = doff[m] + threadIdx.x; i < doff[m + 1]; i += 256) v += r[i] * r[i];
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s2 = 128; s2 > 0; s2 >>= 1) {
    if (threadIdx.x < s2) sh[threadIdx.x] += sh[threadIdx.x + s2];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[m] = sh[0];
}

// per mesh: sum of a b over its free-dof range
__global__ void __launch_bounds__(256) k_mesh_dot(const int32_t* __restrict__ doff, const double* __restrict__ a,
                                                  const double* __restrict__ b, double* __restrict__ out) {
  __shared__ double sh[256];
  const int m = blockIdx.x;
  double v = 0.0;
  for (int64_t i = doff[m] + threadIdx.x; i < doff[m + 1]; i += 256) v += a[i] * b[i];
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s2 = 128; s2 > 0; s2 >>= 1) {
    if (threadIdx.x < s2) sh[threadIdx.x] += sh[threadIdx.x + s2];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[m] = sh[0];
}

// dy[i][d] = c w_m s_m g[dof] on free dofs, 0 elsewhere
__global__ void k_dy_from_g(int64_t n, const int32_t* __restrict__ nmesh_of, const float* __restrict__ scale,
                            const double* __restrict__ w, const int32_t* __restrict__ nda,
                            const double* __restrict__ g, double c, float* __restrict__ dy) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int m = nmesh_of[i];
  const double f = c * w[m] * (double)scale[m];
  for (int d = 0; d < kFout; ++d) {
    float v = 0.f;
    if (d < 3) {
      const int dof = nda[3 * i + d];
      if (dof >= 0) v = (float)(f * g[dof]);
    }
    dy[i * kFout + d] = v;
  }
}

inline unsigned g256(int64_t n) { return (unsigned)((n + 255) / 256); }

mpf_status ensure_cap(mpf_mpnn* net, int64_t n) {
  if (n <= net->cap) return MPF_OK;
  const int64_t c = ((n + 63) / 64) * 64;
  MPF_TRY(net->x.alloc((size_t)c * kFin));
  for (int l = 0; l <= kLayers; ++l) {
    MPF_TRY(net->h[l].alloc((size_t)c * kW));
    MPF_TRY(net->ab[l].alloc((size_t)c * 2 * kW));
    MPF_TRY(net->r[l].alloc((size_t)c * kW));
  }
  MPF_TRY(net->y.alloc((size_t)c * kFout));
  MPF_TRY(net->dy.alloc((size_t)c * kFout));
  MPF_TRY(net->dh.alloc((size_t)c * kW));
  MPF_TRY(net->dcat.alloc((size_t)c * 2 * kW));
  MPF_TRY(net->db16.alloc((size_t)c * 2 * kW));
  net->cap = c;
  return MPF_OK;
}

constexpr int64_t kRowsPerGradBlock = 4096;

// offsets of the bf16 weight matrices inside wb / wtb
constexpr int kWbOff[kLayers + 2] = {0, kFin * kW, kFin * kW + 2 * kW * kW, kFin * kW + 4 * kW * kW,
                                     kFin * kW + 6 * kW * kW};
constexpr int kWbTotal = kFin * kW + 6 * kW * kW + kW * kFout;

template <int K, int C>
mpf_status wgrad(mpf_mpnn* net, int64_t n, const uint16_t* A, int lda, const uint16_t* D, int ldd, const float* D32,
                 int ldd32, float* gW, float* gB) {
  hipStream_t s = net->stream;
  const int nblk = (int)((n + kRowsPerGradBlock - 1) / kRowsPerGradBlock);
  const size_t need = (size_t)nblk * K * C;
  if (net->part.n < need) MPF_TRY(net->part.alloc(need));
  hipLaunchKernelGGL((k_wgrad<K, C>), dim3(nblk), dim3(256), 0, s, n, kRowsPerGradBlock, A, lda, D, ldd,
                     net->part.ptr);
  hipLaunchKernelGGL(k_wgrad_reduce, dim3(g256(K * C)), dim3(256), 0, s, nblk, K * C, net->part.ptr, gW);
  hipLaunchKernelGGL(k_bgrad, dim3(nblk), dim3(64), 0, s, n, C, D32, ldd32, kRowsPerGradBlock, net->part.ptr);
  hipLaunchKernelGGL(k_wgrad_reduce, dim3(g256(C)), dim3(256), 0, s, nblk, C, net->part.ptr, gB);
  MPF_HIP(hipGetLastError());
  return MPF_OK;
}

}  // namespace

mpf_status mpnn_forward(mpf_mpnn* net, int64_t n, const int32_t* nbptr, const int32_t* nb) {
  hipStream_t s = net->stream;
  const unsigned gmm = (unsigned)((n + 63) / 64);
  const float* w = net->w.ptr;
  MMArgs a{};
  a.n = n;
  // h0 = relu(x W0 + b0); bf16 copy into the first half of ab[0] (concat [h0 | a1])
  a.A = net->x.ptr; a.lda = kFin; a.Bt = net->wtb.ptr + kWbOff[0]; a.bias = w + MpnnParams::B0;
  a.out = net->h[0].ptr; a.outb = net->ab[0].ptr; a.ldob = 2 * kW; a.rb = net->r[0].ptr;
  hipLaunchKernelGGL((k_mm<kFin, kW, EPI_ENC>), dim3(gmm), dim3(256), 0, s, a);
  for (int l = 1; l <= kLayers; ++l) {
    // a_l = mean_nb(h_{l-1}) into the second half of ab[l-1]
    hipLaunchKernelGGL(k_agg_fwd, dim3(g256(n * 8)), dim3(256), 0, s, n, nbptr, nb, net->ab[l - 1].ptr, 2 * kW,
                       net->ab[l - 1].ptr + kW, 2 * kW);
    MMArgs b{};
    b.n = n; b.A = net->ab[l - 1].ptr; b.lda = 2 * kW; b.Bt = net->wtb.ptr + kWbOff[l];
    b.bias = w + MpnnParams::WL + (l - 1) * MpnnParams::LSZ + 2 * kW * kW;
    b.hprev = net->h[l - 1].ptr; b.out = net->h[l].ptr; b.outb = net->ab[l].ptr; b.ldob = 2 * kW; b.rb = net->r[l].ptr;
    hipLaunchKernelGGL((k_mm<2 * kW, kW, EPI_MP>), dim3(gmm), dim3(256), 0, s, b);
  }
  MMArgs o{};
  o.n = n; o.A = net->ab[kLayers].ptr; o.lda = 2 * kW; o.Bt = net->wtb.ptr + kWbOff[kLayers + 1];
  o.bias = w + MpnnParams::BO; o.out = net->y.ptr;
  hipLaunchKernelGGL((k_mm<kW, kFout, EPI_OUT>), dim3(gmm), dim3(256), 0, s, o);
  MPF_HIP(hipGetLastError());
  return MPF_OK;
}

// gradients of the loss whose dL/dy is in net->dy, accumulated into net->g
mpf_status mpnn_backward(mpf_mpnn* net, int64_t n, const int32_t* nbptr, const int32_t* nb) {
  hipStream_t s = net->stream;
  const unsigned gmm = (unsigned)((n + 63) / 64);
  float* g = net->g.ptr;
  // output layer: dWo = h3^T dy, dbo = sum dy; dh = dy Wo^T
  hipLaunchKernelGGL(k_to_bf16, dim3(g256(n * kFout)), dim3(256), 0, s, n * kFout, net->dy.ptr, net->db16.ptr);
  MPF_TRY((wgrad<kW, kFout>(net, n, net->ab[kLayers].ptr, 2 * kW, net->db16.ptr, kFout, net->dy.ptr, kFout,
                            g + MpnnParams::WO, g + MpnnParams::BO)));
  MMArgs a{};
  a.n = n; a.A = net->db16.ptr; a.lda = kFout; a.Bt = net->wb.ptr + kWbOff[kLayers + 1]; a.out = net->dh.ptr;
  hipLaunchKernelGGL((k_mm<kFout, kW, EPI_PLAIN>), dim3(gmm), dim3(256), 0, s, a);
  for (int l = kLayers; l >= 1; --l) {
    // dr = dh * (r_l > 0)  -> dcat[:, :W] (fp32) and db16 (bf16)
    hipLaunchKernelGGL(k_relu_bwd, dim3(g256(n * kW)), dim3(256), 0, s, n, kW, net->dh.ptr, kW, net->r[l].ptr,
                       net->dcat.ptr, net->db16.ptr);
    const int wl = MpnnParams::WL + (l - 1) * MpnnParams::LSZ;
    MPF_TRY((wgrad<2 * kW, kW>(net, n, net->ab[l - 1].ptr, 2 * kW, net->db16.ptr, kW, net->dcat.ptr, kW, g + wl,
                               g + wl + 2 * kW * kW)));
    // d[h | a] = dr W_l^T  (N x 2W) into dcat
    MMArgs b{};
    b.n = n; b.A = net->db16.ptr; b.lda = kW; b.Bt = net->wb.ptr + kWbOff[l]; b.out = net->dcat.ptr;
    hipLaunchKernelGGL((k_mm<kW, 2 * kW, EPI_PLAIN>), dim3(gmm), dim3(256), 0, s, b);
    // dh_{l-1} = dh_l + d[:, :W] + agg_bwd(d[:, W:])
    hipLaunchKernelGGL(k_add, dim3(g256(n * kW)), dim3(256), 0, s, n, net->dcat.ptr, 2 * kW, kW, net->dh.ptr);
    hipLaunchKernelGGL(k_agg_bwd, dim3(g256(n * 64)), dim3(256), 0, s, n, nbptr, nb, net->dcat.ptr + kW, 2 * kW,
                       net->dh.ptr);
  }
  // encoder: dz0 = dh0 * (r0 > 0); dW0 = x^T dz0
  hipLaunchKernelGGL(k_relu_bwd, dim3(g256(n * kW)), dim3(256), 0, s, n, kW, net->dh.ptr, kW, net->r[0].ptr,
                     net->dcat.ptr, net->db16.ptr);
  MPF_TRY((wgrad<kFin, kW>(net, n, net->x.ptr, kFin, net->db16.ptr, kW, net->dcat.ptr, kW, g + MpnnParams::W0,
                           g + MpnnParams::B0)));
  MPF_HIP(hipGetLastError());
  return MPF_OK;
}

mpf_status mpnn_refresh(mpf_mpnn* net) {
  hipLaunchKernelGGL(k_wcast, dim3(g256(kWbTotal)), dim3(256), 0, net->stream, net->w.ptr, net->wb.ptr, net->wtb.ptr);
  MPF_HIP(hipGetLastError());
  return MPF_OK;
}

}  // namespace mpf

// =============================================================================================
// C ABI
// =============================================================================================
using namespace mpf;

namespace {
// node -> mesh, per-mesh reference frame, prescribed-node flags of a (batched) problem
mpf_status ensure_ctx(mpf_mpnn* net, mpf_problem* p) {
  if (net->ctx == p->uid) return MPF_OK;
  if (!p->prepared) return MPF_ERR_STATE;
  const int32_t nNode = p->nNode;
  std::vector<int32_t> off;
  if (!p->batch || batch_node_offsets(p, off) != MPF_OK) off = {0, nNode};  // a single mesh
  const int nm = (int)off.size() - 1;
  std::vector<double> X((size_t)nNode * 3);
  std::vector<int32_t> info(nNode), of(nNode), fx(nNode);
  MPF_TRY(p->X.download(X.data(), X.size(), p->stream));
  MPF_TRY(p->pat.nodeinfo.download(info.data(), info.size(), p->stream));
  MPF_HIP(hipStreamSynchronize(p->stream));
  std::vector<double> xr(4 * (size_t)nm);
  for (int m = 0; m < nm; ++m) {
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int32_t v = off[m]; v < off[m + 1]; ++v) {
      of[v] = m;
      fx[v] = (info[v] & 7) != 7;
      for (int d = 0; d < 3; ++d) { lo[d] = std::min(lo[d], X[3 * (size_t)v + d]); hi[d] = std::max(hi[d], X[3 * (size_t)v + d]); }
    }
    double L = 0.0;
    for (int d = 0; d < 3; ++d) { xr[4 * m + d] = off[m + 1] > off[m] ? lo[d] : 0.0; L = std::max(L, hi[d] - lo[d]); }
    xr[4 * m + 3] = L > 0.0 ? L : 1.0;
  }
  // free-dof ranges per mesh and the load weights 1 / ||f_m||^2
  std::vector<int32_t> nda((size_t)nNode * 3);
  MPF_TRY(p->pat.nda.download(nda.data(), nda.size(), p->stream));
  std::vector<double> fu;
  if (p->has_fext) {
    fu.resize((size_t)nNode * 3);
    MPF_TRY(p->fext_unit.download(fu.data(), fu.size(), p->stream));
  }
  MPF_HIP(hipStreamSynchronize(p->stream));
  std::vector<int32_t> doff(nm + 1, 0);
  std::vector<double> wr(nm, 0.0);
  int32_t cnt = 0;
  for (int m = 0; m < nm; ++m) {
    doff[m] = cnt;
    double f2 = 0.0;
    for (int32_t v = off[m]; v < off[m + 1]; ++v)
      for (int d = 0; d < 3; ++d) {
        const int32_t dof = nda[3 * (size_t)v + d];
        if (dof < 0) continue;
        if (dof != cnt) return MPF_ERR_STATE;  // free dofs numbered node-major
        ++cnt;
        if (!fu.empty()) f2 += fu[3 * (size_t)v + d] * fu[3 * (size_t)v + d];
      }
    wr[m] = f2 > 0.0 ? 1.0 / f2 : 0.0;
  }
  doff[nm] = cnt;
  net->doff = doff;
  MPF_TRY(net->doff_d.upload(doff.data(), doff.size(), net->stream));
  MPF_TRY(net->wres.upload(wr.data(), wr.size(), net->stream));
  const size_t nsys = (size_t)p->pat.dispDOF + p->pat.presDOF;
  MPF_TRY(net->ebuf.alloc(nsys > 0 ? nsys : 1));
  MPF_TRY(net->rbuf.alloc(nsys > 0 ? nsys : 1));
  MPF_TRY(net->gbuf.alloc(nsys > 0 ? nsys : 1));
  MPF_TRY(net->msq.alloc(nm));
  MPF_TRY(net->node_off.upload(off.data(), off.size(), net->stream));
  MPF_TRY(net->nmesh_of.upload(of.data(), of.size(), net->stream));
  MPF_TRY(net->fixed.upload(fx.data(), fx.size(), net->stream));
  MPF_TRY(net->xref.upload(xr.data(), xr.size(), net->stream));
  MPF_TRY(net->scale.alloc(nm));
  MPF_HIP(hipStreamSynchronize(net->stream));
  net->nmesh = nm;
  net->ctx = p->uid;
  return MPF_OK;
}

mpf_status graph_of(mpf_problem* p, const int32_t** nbptr, const int32_t** nb) {
  if (!p || !p->prepared) return MPF_ERR_STATE;
  *nbptr = p->pat.nbptr.ptr;
  *nb = p->pat.nb.ptr;
  return MPF_OK;
}
}  // namespace

extern "C" {

int32_t mpf_mpnn_num_params(void) { return MpnnParams::N; }

mpf_status mpf_mpnn_create(mpf_mpnn** out, uint64_t seed) {
  if (!out) return MPF_ERR_ARG;
  int nd = 0;
  if (hipGetDeviceCount(&nd) != hipSuccess || nd == 0) return MPF_ERR_NO_DEVICE;
  mpf_mpnn* net = new mpf_mpnn();
  mpf_status st = MPF_OK;
  if (hipStreamCreate(&net->stream) != hipSuccess) st = MPF_ERR_HIP;
  if (st == MPF_OK) st = net->w.alloc(MpnnParams::N);
  if (st == MPF_OK) st = net->g.alloc(MpnnParams::N);
  if (st == MPF_OK) st = net->m1.alloc(MpnnParams::N);
  if (st == MPF_OK) st = net->m2.alloc(MpnnParams::N);
  if (st == MPF_OK) st = net->wb.alloc(kWbTotal);
  if (st == MPF_OK) st = net->wtb.alloc(kWbTotal);
  if (st != MPF_OK) { delete net; return st; }
  // He-uniform init from SplitMix64 (biases 0; output layer scaled down: start near the extrapolation)
  std::vector<float> w(MpnnParams::N, 0.f);
  uint64_t x = seed;
  auto uni = [&]() {
    x += 0x9E3779B97F4A7C15ull;
    uint64_t z = x;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
  };
  auto fill = [&](int off, int K, int C, double gain) {
    const double lim = gain * std::sqrt(6.0 / K);
    for (int i = 0; i < K * C; ++i) w[off + i] = (float)((2.0 * uni() - 1.0) * lim);
  };
  fill(MpnnParams::W0, kFin, kW, 1.0);
  for (int l = 0; l < kLayers; ++l) fill(MpnnParams::WL + l * MpnnParams::LSZ, 2 * kW, kW, 0.5);
  fill(MpnnParams::WO, kW, kFout, 0.01);
  std::vector<float> z(MpnnParams::N, 0.f);
  if (hipMemcpy(net->w.ptr, w.data(), w.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(net->g.ptr, z.data(), z.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(net->m1.ptr, z.data(), z.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(net->m2.ptr, z.data(), z.size() * 4, hipMemcpyHostToDevice) != hipSuccess) {
    delete net;
    return MPF_ERR_HIP;
  }
  st = mpnn_refresh(net);
  if (st == MPF_OK && hipStreamSynchronize(net->stream) != hipSuccess) st = MPF_ERR_HIP;
  if (st != MPF_OK) { delete net; return st; }
  *out = net;
  return MPF_OK;
}

mpf_status mpf_mpnn_destroy(mpf_mpnn* net) {
  if (!net) return MPF_OK;
  if (net->stream) { (void)hipStreamSynchronize(net->stream); (void)hipStreamDestroy(net->stream); }
  delete net;
  return MPF_OK;
}

mpf_status mpf_mpnn_get_params(mpf_mpnn* net, float* w, float* g) {
  if (!net) return MPF_ERR_ARG;
  MPF_HIP(hipStreamSynchronize(net->stream));
  if (w) MPF_HIP(hipMemcpy(w, net->w.ptr, MpnnParams::N * 4, hipMemcpyDeviceToHost));
  if (g) MPF_HIP(hipMemcpy(g, net->g.ptr, MpnnParams::N * 4, hipMemcpyDeviceToHost));
  return MPF_OK;
}

mpf_status mpf_mpnn_set_params(mpf_mpnn* net, const float* w) {
  if (!net || !w) return MPF_ERR_ARG;
  MPF_HIP(hipStreamSynchronize(net->stream));
  MPF_HIP(hipMemcpy(net->w.ptr, w, MpnnParams::N * 4, hipMemcpyHostToDevice));
  MPF_TRY(mpnn_refresh(net));
  MPF_HIP(hipStreamSynchronize(net->stream));
  return MPF_OK;
}

mpf_status mpf_mpnn_zero_grad(mpf_mpnn* net) {
  if (!net) return MPF_ERR_ARG;
  std::vector<float> z(MpnnParams::N, 0.f);
  MPF_HIP(hipStreamSynchronize(net->stream));
  MPF_HIP(hipMemcpy(net->g.ptr, z.data(), z.size() * 4, hipMemcpyHostToDevice));
  return MPF_OK;
}

// Forward on a problem's node graph: x [nNode][32] bf16 (device; may be the network's own buffer
// filled by mpf_mpnn_features when null), y [nNode][32] fp32 (device, optional copy-out).
mpf_status mpf_mpnn_forward(mpf_mpnn* net, mpf_problem* p, const uint16_t* x, float* y) {
  if (!net || !p || p->etype == 0) return MPF_ERR_ARG;
  const int32_t *nbptr, *nb;
  MPF_TRY(graph_of(p, &nbptr, &nb));
  const int64_t n = p->nNode;
  MPF_TRY(ensure_cap(net, n));
  if (x) MPF_HIP(hipMemcpyAsync(net->x.ptr, x, (size_t)n * kFin * 2, hipMemcpyDeviceToDevice, net->stream));
  MPF_TRY(mpnn_forward(net, n, nbptr, nb));
  if (y) MPF_HIP(hipMemcpyAsync(y, net->y.ptr, (size_t)n * kFout * 4, hipMemcpyDeviceToDevice, net->stream));
  MPF_HIP(hipStreamSynchronize(net->stream));
  return MPF_OK;
}

// Backward of the last forward for dL/dy = dy [nNode][32] fp32 (device; NULL: the one left by
// mpf_mpnn_loss); gradients accumulate into the network.
mpf_status mpf_mpnn_backward(mpf_mpnn* net, mpf_problem* p, const float* dy) {
  if (!net || !p || p->etype == 0) return MPF_ERR_ARG;
  const int32_t *nbptr, *nb;
  MPF_TRY(graph_of(p, &nbptr, &nb));
  const int64_t n = p->nNode;
  if (n > net->cap) return MPF_ERR_STATE;
  if (dy) MPF_HIP(hipMemcpyAsync(net->dy.ptr, dy, (size_t)n * kFout * 4, hipMemcpyDeviceToDevice, net->stream));
  MPF_TRY(mpnn_backward(net, n, nbptr, nb));
  MPF_HIP(hipStreamSynchronize(net->stream));
  return MPF_OK;
}

// Input features of load step k+1 from the converged u_k, u_{k-1} (device, [nNode*3]); also sets the
// per-mesh scales used by mpf_mpnn_loss and mpf_mpnn_guess.  rho = (lam_{k+1} - lam_k)/(lam_k - lam_{k-1}).
mpf_status mpf_mpnn_features(mpf_mpnn* net, mpf_problem* p, const double* uk, const double* ukm1, double rho,
                             double lam_next, double lam_ratio) {
  if (!net || !p || p->etype == 0 || !uk || !ukm1) return MPF_ERR_ARG;
  MPF_TRY(ensure_ctx(net, p));
  const int64_t n = p->nNode;
  MPF_TRY(ensure_cap(net, n));
  hipLaunchKernelGGL(k_mesh_scale, dim3(net->nmesh), dim3(256), 0, net->stream, net->node_off.ptr, uk, ukm1,
                     net->scale.ptr);
  FeatArgs f{};
  f.n = n; f.nmesh_of = net->nmesh_of.ptr; f.scale = net->scale.ptr; f.X = p->X.ptr; f.xref = net->xref.ptr;
  f.uk = uk; f.ukm1 = ukm1; f.fixed = net->fixed.ptr; f.rho = (float)rho; f.lam_next = (float)lam_next;
  f.lam_ratio = (float)lam_ratio; f.x = net->x.ptr;
  hipLaunchKernelGGL(k_features, dim3(g256(n)), dim3(256), 0, net->stream, f);
  MPF_HIP(hipGetLastError());
  MPF_HIP(hipStreamSynchronize(net->stream));
  return MPF_OK;
}

// Mean squared error of the last forward against the converged u_{k+1} (free dofs only); dL/dy is
// left in the network for mpf_mpnn_backward (dy = NULL there).
mpf_status mpf_mpnn_loss(mpf_mpnn* net, mpf_problem* p, const double* unext, const double* uk, const double* ukm1,
                         double rho, double* loss) {
  if (!net || !p || !unext || !uk || !ukm1 || net->ctx != p->uid) return MPF_ERR_ARG;
  const int64_t n = p->nNode;
  const unsigned nb = g256(n);
  if (net->red.n < nb) MPF_TRY(net->red.alloc(nb));
  const float inv = p->pat.dispDOF > 0 ? 1.f / (float)p->pat.dispDOF : 0.f;
  hipLaunchKernelGGL(k_loss, dim3(nb), dim3(256), 0, net->stream, n, net->nmesh_of.ptr, net->scale.ptr, p->pat.nda.ptr,
                     unext, uk, ukm1, (float)rho, net->y.ptr, inv, net->dy.ptr, net->red.ptr);
  MPF_HIP(hipGetLastError());
  std::vector<double> h(nb);
  MPF_HIP(hipMemcpyAsync(h.data(), net->red.ptr, nb * sizeof(double), hipMemcpyDeviceToHost, net->stream));
  MPF_HIP(hipStreamSynchronize(net->stream));
  double sum = 0.0;
  for (double v : h) sum += v;
  if (loss) *loss = sum * inv;
  return MPF_OK;
}

// Linearised start residual of the prediction (what Newton's convergence depends on):
//   loss = mean over meshes of ||K e_m||^2 / (dlam^2 ||f_m||^2),
// e = predicted - converged u_{k+1} on the free dofs, K = tangent at the converged u_{k+1} (assembled
// here: the problem's state, system and solution buffers are overwritten).  dL/dy (K symmetric:
// 2 w_m s_m K K e / (dlam^2 nmesh)) is left in the network for mpf_mpnn_backward.
mpf_status mpf_mpnn_residual_loss(mpf_mpnn* net, mpf_problem* p, double lam_next, double dlam, const double* unext,
                                  const double* uk, const double* ukm1, double rho, double* loss) {
  if (!net || !p || !unext || !uk || !ukm1 || dlam <= 0.0 || net->ctx != p->uid) return MPF_ERR_ARG;
  const int64_t n = p->nNode;
  const int64_t nsys = (int64_t)p->pat.dispDOF + p->pat.presDOF;
  MPF_HIP(hipStreamSynchronize(net->stream));
  MPF_HIP(hipMemcpyAsync(p->disp.ptr, unext, (size_t)n * 3 * sizeof(double), hipMemcpyDeviceToDevice, p->stream));
  MPF_TRY(enqueue_assembly(p, false, lam_next, 1));
  hipLaunchKernelGGL(k_fill0, dim3(g256(nsys)), dim3(256), 0, p->stream, nsys, net->ebuf.ptr);
  hipLaunchKernelGGL(k_err_free, dim3(g256(n)), dim3(256), 0, p->stream, n, net->nmesh_of.ptr, net->scale.ptr,
                     p->pat.nda.ptr, net->y.ptr, unext, uk, ukm1, rho, net->ebuf.ptr);
  MPF_TRY(spmv(p->pat.rowptr.ptr, p->pat.colidx.ptr, p->vals.ptr, nsys, p->pat.nnz, net->ebuf.ptr, net->rbuf.ptr,
               nullptr, 0, nullptr, p->stream));
  hipLaunchKernelGGL(k_mesh_sq, dim3(net->nmesh), dim3(256), 0, p->stream, net->doff_d.ptr, net->rbuf.ptr,
                     net->msq.ptr);
  MPF_TRY(spmv(p->pat.rowptr.ptr, p->pat.colidx.ptr, p->vals.ptr, nsys, p->pat.nnz, net->rbuf.ptr, net->gbuf.ptr,
               nullptr, 0, nullptr, p->stream));
  const double c = 2.0 / (dlam * dlam * net->nmesh);
  hipLaunchKernelGGL(k_dy_from_g, dim3(g256(n)), dim3(256), 0, p->stream, n, net->nmesh_of.ptr, net->scale.ptr,
                     net->wres.ptr, p->pat.nda.ptr, net->gbuf.ptr, c, net->dy.ptr);
  MPF_HIP(hipGetLastError());
  std::vector<double> sq(net->nmesh), w(net->nmesh);
  MPF_HIP(hipMemcpyAsync(sq.data(), net->msq.ptr, sq.size() * sizeof(double), hipMemcpyDeviceToHost, p->stream));
  MPF_HIP(hipMemcpyAsync(w.data(), net->wres.ptr, w.size() * sizeof(double), hipMemcpyDeviceToHost, p->stream));
  MPF_HIP(hipStreamSynchronize(p->stream));
  double sum = 0.0;
  for (int m = 0; m < net->nmesh; ++m) sum += w[m] * sq[m];
  if (loss) *loss = sum / (dlam * dlam * net->nmesh);
  return MPF_OK;
}

// Newton start point for step k+1: u_k + rho (u_k - u_{k-1}) (+ s_m y of the last forward when use_net).
mpf_status mpf_mpnn_guess(mpf_mpnn* net, mpf_problem* p, const double* uk, const double* ukm1, double rho,
                          int32_t use_net, double* guess) {
  if (!net || !p || !uk || !ukm1 || !guess || net->ctx != p->uid) return MPF_ERR_ARG;
  const int64_t n = p->nNode;
  hipLaunchKernelGGL(k_guess, dim3(g256(n)), dim3(256), 0, net->stream, n, net->nmesh_of.ptr, net->scale.ptr, uk,
                     ukm1, rho, use_net, nullptr, net->y.ptr, guess);
  MPF_HIP(hipGetLastError());
  MPF_HIP(hipStreamSynchronize(net->stream));
  return MPF_OK;
}

// Guarded learned start: per mesh m the network's correction c_m enters with the weight w_m in [0, 1]
// that minimises the linearised residual at load factor lam (prescribed dofs at their values),
//   R(base + w c) ~ R_b + w (R_n - R_b),   w = (R_b.R_b - R_n.R_b) / |R_n - R_b|^2, clipped to [0, 1],
// from the residuals R_n with the whole correction and R_b at the base predictor: two residual-only
// assemblies, exact at w = 0 and w = 1, so the model never prefers a weight whose end points are
// both worse than the base.  The problem's displacement is restored.  n_learned: meshes with w > 0.
mpf_status mpf_mpnn_guarded_guess(mpf_mpnn* net, mpf_problem* p, double lam, const double* uk, const double* ukm1,
                                  double rho, double* guess, int32_t* n_learned) {
  if (!net || !p || !uk || !ukm1 || !guess || net->ctx != p->uid) return MPF_ERR_ARG;
  const int64_t n = p->nNode;
  const size_t n3 = (size_t)n * 3;
  if (net->usave.n < n3) MPF_TRY(net->usave.alloc(n3));
  if (net->gbuf2.n < n3) MPF_TRY(net->gbuf2.alloc(n3));
  const int64_t nfree = p->pat.dispDOF;
  if (net->rnet.n < (size_t)std::max<int64_t>(nfree, 1)) MPF_TRY(net->rnet.alloc(std::max<int64_t>(nfree, 1)));
  if (net->mdot.n < (size_t)net->nmesh) MPF_TRY(net->mdot.alloc(net->nmesh));
  MPF_HIP(hipStreamSynchronize(net->stream));
  hipStream_t s = p->stream;
  MPF_HIP(hipMemcpyAsync(net->usave.ptr, p->disp.ptr, n3 * sizeof(double), hipMemcpyDeviceToDevice, s));
  const int32_t nbc = (int32_t)p->bc_dof_h.size();
  std::vector<double> sq[2];
  for (int c = 0; c < 2; ++c) {  // 0: with the network, 1: extrapolation only
    double* cand = c == 0 ? guess : net->gbuf2.ptr;
    hipLaunchKernelGGL(k_guess, dim3(g256(n)), dim3(256), 0, s, n, net->nmesh_of.ptr, net->scale.ptr, uk, ukm1, rho,
                       c == 0 ? 1 : 0, nullptr, net->y.ptr, cand);
    MPF_HIP(hipMemcpyAsync(p->disp.ptr, cand, n3 * sizeof(double), hipMemcpyDeviceToDevice, s));
    if (nbc > 0)
      hipLaunchKernelGGL(k_set_bc_values, dim3(g256(nbc)), dim3(256), 0, s, nbc, p->bc_dof.ptr, p->bc_val.ptr, lam,
                         p->disp.ptr);
    MPF_TRY(enqueue_assembly(p, false, lam, 0));
    hipLaunchKernelGGL(k_mesh_sq, dim3(net->nmesh), dim3(256), 0, s, net->doff_d.ptr, p->rhs.ptr, net->msq.ptr);
    if (c == 0 && nfree > 0)
      MPF_HIP(hipMemcpyAsync(net->rnet.ptr, p->rhs.ptr, nfree * sizeof(double), hipMemcpyDeviceToDevice, s));
    if (c == 1)
      hipLaunchKernelGGL(k_mesh_dot, dim3(net->nmesh), dim3(256), 0, s, net->doff_d.ptr, net->rnet.ptr, p->rhs.ptr,
                         net->mdot.ptr);
    sq[c].resize(net->nmesh);
    MPF_HIP(hipMemcpyAsync(sq[c].data(), net->msq.ptr, net->nmesh * sizeof(double), hipMemcpyDeviceToHost, s));
    MPF_HIP(hipStreamSynchronize(s));
  }
  std::vector<double> dnb(net->nmesh);
  MPF_HIP(hipMemcpy(dnb.data(), net->mdot.ptr, net->nmesh * sizeof(double), hipMemcpyDeviceToHost));
  std::vector<float> w(net->nmesh);
  int32_t kept = 0;
  for (int m = 0; m < net->nmesh; ++m) {
    const double a = sq[0][m], b = sq[1][m], cr = dnb[m];  // |R_n|^2, |R_b|^2, R_n.R_b
    const double den = a + b - 2.0 * cr;                   // |R_n - R_b|^2
    double wm = den > 0.0 ? (b - cr) / den : 0.0;
    if (!(wm > 0.0)) wm = 0.0;
    if (wm > 1.0) wm = 1.0;
    w[m] = (float)wm;
    kept += w[m] != 0.f;
  }
  if (net->wmesh.n < (size_t)net->nmesh) MPF_TRY(net->wmesh.alloc(net->nmesh));
  MPF_HIP(hipMemcpyAsync(net->wmesh.ptr, w.data(), w.size() * sizeof(float), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_guess, dim3(g256(n)), dim3(256), 0, s, n, net->nmesh_of.ptr, net->scale.ptr, uk, ukm1, rho, 1,
                     net->wmesh.ptr, net->y.ptr, guess);
  MPF_HIP(hipMemcpyAsync(p->disp.ptr, net->usave.ptr, n3 * sizeof(double), hipMemcpyDeviceToDevice, s));
  MPF_HIP(hipGetLastError());
  MPF_HIP(hipStreamSynchronize(s));
  if (n_learned) *n_learned = kept;
  return MPF_OK;
}

mpf_status mpf_mpnn_adam(mpf_mpnn* net, float lr, float beta1, float beta2, float eps) {
  if (!net) return MPF_ERR_ARG;
  net->adam_t += 1;
  const float c1 = 1.f - std::pow(beta1, (float)net->adam_t), c2 = 1.f - std::pow(beta2, (float)net->adam_t);
  hipLaunchKernelGGL(k_adam, dim3(g256(MpnnParams::N)), dim3(256), 0, net->stream, MpnnParams::N, net->w.ptr,
                     net->g.ptr, net->m1.ptr, net->m2.ptr, lr, beta1, beta2, eps, c1, c2);
  MPF_TRY(mpnn_refresh(net));
  MPF_HIP(hipStreamSynchronize(net->stream));
  return MPF_OK;
}

}  // extern "C"
