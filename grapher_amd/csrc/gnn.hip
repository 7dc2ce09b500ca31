// gnn.hip -- message-passing network that proposes initial layouts for the
// force embedder (the learned warm start), with hand-written training and
// inference on gfx950.
//
// Model (no reference counterpart; it sits in front of
// gvizForceEmbedderBegin): input features x_v (16 wide: random start
// coordinates and their Chebyshev low-pass filters over the graph, a degree
// term, a constant, zero padding), three layers
//     h' = relu([h | mean_{u in N(v)} h_u] W + b)      (hidden width 64)
// and a linear head y = h W4 + b4 (2 outputs, padded to 16). Weights and
// activations are BF16 and every product accumulates in FP32 on the matrix
// cores (v_mfma_f32_16x16x32_bf16 where the depth is a multiple of 32,
// v_mfma_f32_16x16x16_bf16 otherwise); neighbour means, losses and the
// Adam update are FP32. Training fits a force-embedder layout T up to rotation,
// reflection and scale: with P and T centred and R, s the orthogonal
// Procrustes fit of P to T,
//     L = |s P R - T|^2 / |T|^2 = 1 - (sigma1 + sigma2)^2 / (|P|^2 |T|^2),
// dL/dP = 2 s (s P R - T) R^T / |T|^2 (R and s optimal: envelope theorem).
#include "../../include/gviz_hip.h"
#include "common.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

using namespace gvh;

namespace {

constexpr int kF = 16;    // input feature width (padded)
constexpr int kH = 64;    // hidden width
constexpr int kO = 16;    // head width (2 used)
constexpr int kB = 256;   // threads per block
constexpr int kRows = 256; // rows per weight-gradient block

// parameter layout (floats): W1 (2F x H), b1, W2 (2H x H), b2, W3, b3,
// W4 (H x O), b4
constexpr int64_t kOffW[4] = {0, 2 * kF * kH + kH,
                              2 * kF * kH + kH + 2 * kH * kH + kH,
                              2 * kF * kH + kH + 2 * (2 * kH * kH + kH)};
constexpr int kK[4] = {2 * kF, 2 * kH, 2 * kH, kH}; // rows of W_l
constexpr int kN[4] = {kH, kH, kH, kO};             // columns of W_l
constexpr int64_t kParams = kOffW[3] + kH * kO + kO;

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(uint16_t b) {
  return __uint_as_float((uint32_t)b << 16);
}
// round to nearest even (finite inputs)
__device__ __forceinline__ uint16_t f2bf(float f) {
  const uint32_t u = __float_as_uint(f);
  return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

// BF16 copies of the weights: Wt_l (N_l x K_l) for the forward product and
// Wr_l (K_l x N_l) for the backward-data product.
__global__ void k_weights_bf16(const float *__restrict__ p, uint16_t *__restrict__ wt,
                               uint16_t *__restrict__ wr, int64_t offW, int K, int N) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= K * N)
    return;
  const int r = i / N, c = i % N; // W[r][c], row-major K x N
  const uint16_t b = f2bf(p[offW + i]);
  wr[i] = b;
  wt[(int64_t)c * K + r] = b;
}

// Eight BF16 values of a 16-byte load, as floats (element 2m in the low half
// of word m).
__device__ __forceinline__ void bf8_to_f(const uint4 r, float *f) {
  const uint32_t w[4] = {r.x, r.y, r.z, r.w};
#pragma unroll
  for (int m = 0; m < 4; m++) {
    f[2 * m] = __uint_as_float(w[m] << 16);
    f[2 * m + 1] = __uint_as_float(w[m] & 0xFFFF0000u);
  }
}

// M[v] = mean over neighbours of X (width F in {16, 64}): F / 8 lanes per
// vertex, 16 bytes (8 features) per lane, neighbour rows fetched four at a
// time and added in adjacency order.
template <int F>
__global__ __launch_bounds__(kB) void k_aggregate(const int32_t *__restrict__ off,
                                                  const int32_t *__restrict__ adj,
                                                  const float *__restrict__ invDeg,
                                                  const uint16_t *__restrict__ X,
                                                  uint16_t *__restrict__ M, int64_t n) {
  constexpr int L = F / 8;
  const int64_t gt = (int64_t)blockIdx.x * kB + threadIdx.x;
  const int64_t v = gt / L;
  const int c = (int)(gt % L) * 8;
  if (v >= n)
    return;
  float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0}, f[8];
  const int32_t e1 = off[v + 1];
  int32_t e = off[v];
  for (; e + 4 <= e1; e += 4) {
    uint4 r[4];
#pragma unroll
    for (int k = 0; k < 4; k++)
      r[k] = *(const uint4 *)(X + (int64_t)adj[e + k] * F + c);
#pragma unroll
    for (int k = 0; k < 4; k++) {
      bf8_to_f(r[k], f);
#pragma unroll
      for (int j = 0; j < 8; j++)
        acc[j] += f[j];
    }
  }
  for (; e < e1; e++) {
    bf8_to_f(*(const uint4 *)(X + (int64_t)adj[e] * F + c), f);
#pragma unroll
    for (int j = 0; j < 8; j++)
      acc[j] += f[j];
  }
  const float s = invDeg[v];
  uint32_t w[4];
#pragma unroll
  for (int m = 0; m < 4; m++)
    w[m] = (uint32_t)f2bf(acc[2 * m] * s) | ((uint32_t)f2bf(acc[2 * m + 1] * s) << 16);
  *(uint4 *)(M + v * F + c) = make_uint4(w[0], w[1], w[2], w[3]);
}

// C = [A1 | A2] B with B given as Bt (Nc x K, BF16), one wavefront per 16
// rows, NT column tiles of 16. Epilogue: +bias, optional relu, to BF16 (outB,
// row stride Nc) and/or FP32 (outF, row stride Nc).
template <int NT>
__global__ __launch_bounds__(kB) void k_gemm(const uint16_t *__restrict__ A1,
                                             const uint16_t *__restrict__ A2, int Ka,
                                             int64_t n, const uint16_t *__restrict__ Bt,
                                             int K, const float *__restrict__ bias,
                                             int relu, uint16_t *__restrict__ outB,
                                             float *__restrict__ outF) {
  constexpr int Nc = 16 * NT;
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t r0 = ((int64_t)blockIdx.x * (kB / kWave) + threadIdx.x / kWave) * 16;
  if (r0 >= n)
    return;
  const int li = lane & 15, lg = lane >> 4;
  const int64_t row = r0 + li;
  v4f acc[NT];
#pragma unroll
  for (int t = 0; t < NT; t++)
    acc[t] = (v4f){0.0f, 0.0f, 0.0f, 0.0f};
  for (int kb = 0; kb < K; kb += 16) {
    const int kc = kb + 4 * lg;
    v4s a = (v4s){0, 0, 0, 0};
    if (row < n) {
      const uint16_t *src = kc < Ka ? A1 + row * Ka + kc : A2 + row * Ka + (kc - Ka);
      const uint2 raw = *(const uint2 *)src;
      a = __builtin_bit_cast(v4s, raw);
    }
#pragma unroll
    for (int t = 0; t < NT; t++) {
      const uint2 rb = *(const uint2 *)(Bt + (int64_t)(t * 16 + li) * K + kc);
      const v4s b = __builtin_bit_cast(v4s, rb);
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[t], 0, 0, 0);
    }
  }
#pragma unroll
  for (int t = 0; t < NT; t++) {
    const int col = t * 16 + li;
    const float bcol = bias ? bias[col] : 0.0f;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int64_t r = r0 + 4 * lg + i;
      if (r >= n)
        continue;
      float z = acc[t][i] + bcol;
      if (relu)
        z = z > 0.0f ? z : 0.0f;
      if (outB)
        outB[r * Nc + col] = f2bf(z);
      if (outF)
        outF[r * Nc + col] = z;
    }
  }
}

// k_gemm on the 32-deep BF16 product (v_mfma_f32_16x16x32_bf16) for K a
// multiple of 32: each lane loads 8 consecutive k of its row of A and its
// column of B (16 bytes each), so a 128-deep row takes four loads.
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
template <int NT>
__global__ __launch_bounds__(kB) void k_gemm32(const uint16_t *__restrict__ A1,
                                               const uint16_t *__restrict__ A2, int Ka,
                                               int64_t n, const uint16_t *__restrict__ Bt,
                                               int K, const float *__restrict__ bias,
                                               int relu, uint16_t *__restrict__ outB,
                                               float *__restrict__ outF) {
  constexpr int Nc = 16 * NT;
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t r0 = ((int64_t)blockIdx.x * (kB / kWave) + threadIdx.x / kWave) * 16;
  if (r0 >= n)
    return;
  const int li = lane & 15, lg = lane >> 4;
  const int64_t row = r0 + li;
  v4f acc[NT];
#pragma unroll
  for (int t = 0; t < NT; t++)
    acc[t] = (v4f){0.0f, 0.0f, 0.0f, 0.0f};
  for (int kb = 0; kb < K; kb += 32) {
    const int kc = kb + 8 * lg;
    uint4 raw = make_uint4(0, 0, 0, 0);
    if (row < n)
      raw = *(const uint4 *)(kc < Ka ? A1 + row * Ka + kc : A2 + row * Ka + (kc - Ka));
    const v8bf a = __builtin_bit_cast(v8bf, raw);
#pragma unroll
    for (int t = 0; t < NT; t++) {
      const uint4 rb = *(const uint4 *)(Bt + (int64_t)(t * 16 + li) * K + kc);
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, __builtin_bit_cast(v8bf, rb),
                                                       acc[t], 0, 0, 0);
    }
  }
#pragma unroll
  for (int t = 0; t < NT; t++) {
    const int col = t * 16 + li;
    const float bcol = bias ? bias[col] : 0.0f;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int64_t r = r0 + 4 * lg + i;
      if (r >= n)
        continue;
      float z = acc[t][i] + bcol;
      if (relu)
        z = z > 0.0f ? z : 0.0f;
      if (outB)
        outB[r * Nc + col] = f2bf(z);
      if (outF)
        outF[r * Nc + col] = z;
    }
  }
}

// Per-block partial of dW = [A1 | A2]^T D over rows [b*kRows, (b+1)*kRows):
// both operands staged in LDS, 16 x 16 output tiles spread over the
// wavefronts, partial[b][K][Nc] in FP32.
template <int NT>
__global__ __launch_bounds__(kB) void k_wgrad(const uint16_t *__restrict__ A1,
                                              const uint16_t *__restrict__ A2, int Ka,
                                              int64_t n, const uint16_t *__restrict__ D,
                                              int K, float *__restrict__ partial) {
  constexpr int Nc = 16 * NT;
  __shared__ uint16_t sA[kRows * 2 * kH];
  __shared__ uint16_t sD[kRows * kH];
  const int64_t rBase = (int64_t)blockIdx.x * kRows;
  for (int i = threadIdx.x; i < kRows * K; i += kB) {
    const int rr = i / K, c = i % K;
    const int64_t r = rBase + rr;
    uint16_t v = 0;
    if (r < n)
      v = c < Ka ? A1[r * Ka + c] : A2[r * Ka + (c - Ka)];
    sA[i] = v;
  }
  for (int i = threadIdx.x; i < kRows * Nc; i += kB) {
    const int64_t r = rBase + i / Nc;
    sD[i] = r < n ? D[r * Nc + (i % Nc)] : (uint16_t)0;
  }
  __syncthreads();
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  const int li = lane & 15, lg = lane >> 4;
  const int tilesM = K / 16;
  for (int tile = wave; tile < tilesM * NT; tile += kB / kWave) {
    const int mt = tile / NT, nt = tile % NT;
    v4f acc = (v4f){0.0f, 0.0f, 0.0f, 0.0f};
    for (int rs = 0; rs < kRows; rs += 16) {
      v4s a, b;
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int rr = rs + 4 * lg + j;
        a[j] = (short)sA[rr * K + mt * 16 + li];
        b[j] = (short)sD[rr * Nc + nt * 16 + li];
      }
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
    float *out = partial + (int64_t)blockIdx.x * K * Nc;
#pragma unroll
    for (int i = 0; i < 4; i++)
      out[(mt * 16 + 4 * lg + i) * Nc + nt * 16 + li] = acc[i];
  }
}

// Per-block column sums of D (bias gradient), rows as k_wgrad.
__global__ __launch_bounds__(kB) void k_colsum(const uint16_t *__restrict__ D, int Nc,
                                               int64_t n, float *__restrict__ partial) {
  const int64_t rBase = (int64_t)blockIdx.x * kRows;
  for (int c = threadIdx.x; c < Nc; c += kB) {
    float s = 0.0f;
    for (int rr = 0; rr < kRows; rr++) {
      const int64_t r = rBase + rr;
      if (r < n)
        s += bf2f(D[r * Nc + c]);
    }
    partial[(int64_t)blockIdx.x * Nc + c] = s;
  }
}

// out[i] = sum over blocks of partial[b][i], in block order.
__global__ void k_reduce_blocks(const float *__restrict__ partial, int blocks,
                                int64_t count, float *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count)
    return;
  float s = 0.0f;
  for (int b = 0; b < blocks; b++)
    s += partial[(int64_t)b * count + i];
  out[i] = s;
}

// dZ = relu'(H) * dA for the head's input (widths 64).
__global__ void k_mask(const float *__restrict__ dA, const uint16_t *__restrict__ H,
                       int64_t count, uint16_t *__restrict__ dZ) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count)
    return;
  dZ[i] = bf2f(H[i]) > 0.0f ? f2bf(dA[i]) : (uint16_t)0;
}

// Backward through [h | mean h]: dH[u] = dA[u][0:64] + sum over neighbours v
// of dA[v][64:128] / deg(v) (the graph is symmetric), then the relu mask of
// the layer that produced h. One wavefront per vertex.
__global__ __launch_bounds__(kB) void k_bwd_combine(
    const int32_t *__restrict__ off, const int32_t *__restrict__ adj,
    const float *__restrict__ invDeg, const float *__restrict__ dA,
    const uint16_t *__restrict__ H, int64_t n, uint16_t *__restrict__ dZ) {
  const int64_t u = ((int64_t)blockIdx.x * kB + threadIdx.x) / kWave;
  const int f = threadIdx.x & (kWave - 1);
  if (u >= n)
    return;
  float s = dA[u * 2 * kH + f];
  for (int32_t e = off[u]; e < off[u + 1]; e++) {
    const int32_t v = adj[e];
    s += dA[(int64_t)v * 2 * kH + kH + f] * invDeg[v];
  }
  const uint16_t h = H[u * kH + f];
  dZ[u * kH + f] = bf2f(h) > 0.0f ? f2bf(s) : (uint16_t)0;
}

// Raw moments of the head output P = Y[:, 0:2] and the target T, FP64 per
// block: sum p (2), sum t (2), sum p t^T (4), sum |p|^2, sum |t|^2.
__global__ __launch_bounds__(kB) void k_head_moments(const float *__restrict__ Y,
                                                     const float *__restrict__ T,
                                                     int64_t n, double *__restrict__ part) {
  __shared__ double sm[kB / kWave][10];
  double m[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int64_t v = (int64_t)blockIdx.x * kB + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * kB) {
    const double px = Y[v * kO], py = Y[v * kO + 1];
    const double tx = T ? T[2 * v] : 0.0, ty = T ? T[2 * v + 1] : 0.0;
    m[0] += px;
    m[1] += py;
    m[2] += tx;
    m[3] += ty;
    m[4] += px * tx;
    m[5] += px * ty;
    m[6] += py * tx;
    m[7] += py * ty;
    m[8] += px * px + py * py;
    m[9] += tx * tx + ty * ty;
  }
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
#pragma unroll
  for (int j = 0; j < 10; j++) {
    const double w = wave_sum(m[j]);
    if (lane == 0)
      sm[wave][j] = w;
  }
  __syncthreads();
  if (threadIdx.x < 10) {
    double s = 0.0;
    for (int w = 0; w < kB / kWave; w++)
      s += sm[w][threadIdx.x];
    part[blockIdx.x * 10 + threadIdx.x] = s;
  }
}

struct LossFit {
  float mpx, mpy, mtx, mty; // means
  float r00, r01, r10, r11; // R
  float s, scale;           // s, 2 s / |T|^2
};

// dY[v] = 2 s (s (p - mp) R - (t - mt)) R^T / |T|^2 in columns 0..1; zeros
// elsewhere. (The row mean of this gradient is zero, so the centring of P
// needs no correction.)
__global__ void k_loss_grad(const float *__restrict__ Y, const float *__restrict__ T,
                            int64_t n, LossFit fit, uint16_t *__restrict__ dY) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n)
    return;
  const float px = Y[v * kO] - fit.mpx, py = Y[v * kO + 1] - fit.mpy;
  const float tx = T[2 * v] - fit.mtx, ty = T[2 * v + 1] - fit.mty;
  // e = s p R - t  (row vector times R)
  const float ex = fit.s * (px * fit.r00 + py * fit.r10) - tx;
  const float ey = fit.s * (px * fit.r01 + py * fit.r11) - ty;
  // g = scale * e R^T
  const float gx = fit.scale * (ex * fit.r00 + ey * fit.r01);
  const float gy = fit.scale * (ex * fit.r10 + ey * fit.r11);
  dY[v * kO] = f2bf(gx);
  dY[v * kO + 1] = f2bf(gy);
  for (int c = 2; c < kO; c++)
    dY[v * kO + c] = 0;
}

__global__ void k_adam(float *__restrict__ p, const float *__restrict__ g,
                       float *__restrict__ m, float *__restrict__ v, int64_t count,
                       float lr, float b1, float b2, float eps, float c1, float c2) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count)
    return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}


// Chebyshev low-pass filter of the lazy random walk P = (I + D^-1 A) / 2
// (isolated vertices: P x = x). mu(P) = (2 P - c I) / c maps the eigenvalues
// in [0, c] onto [-1, 1]; y_0 = x, y_1 = mu(P) y_0,
// y_{k+1} = 2 mu(P) y_k - y_{k-1}, so y_d = T_d(mu(P)) x: every
// eigen-component of P below c is damped by at least 1 / T_d(mu(1))
// relative to the component at eigenvalue 1 (d = 16, c = 0.9: 5.5e-5), in
// d neighbour sweeps where the plain walk P^t needs t ~ d^2.
constexpr float kChebC = 0.9f;
constexpr int kChebDeg = 16;
constexpr int kChebMarks[4] = {2, 4, 8, 16};

// out = mu(P) y (first) or 2 mu(P) y - yPrev. kChebLanes lanes per vertex:
// lane j adds neighbours j, j + 8, ... in adjacency order, then the lanes'
// sums are combined pairwise (j + 4, j + 2, j + 1).
constexpr int kChebLanes = 8;
__global__ __launch_bounds__(kB) void k_cheb(const int32_t *__restrict__ off,
                                             const int32_t *__restrict__ adj,
                                             const float *__restrict__ invDeg,
                                             const float2 *__restrict__ y,
                                             const float2 *__restrict__ yPrev, int first,
                                             int64_t n, float2 *__restrict__ out) {
  const int64_t gt = (int64_t)blockIdx.x * kB + threadIdx.x;
  const int64_t v = gt / kChebLanes;
  const int j = (int)(gt % kChebLanes);
  // whole vertex groups share a wavefront, so a group exits together
  if (v >= n)
    return;
  const int32_t e0 = off[v], e1 = off[v + 1];
  float sx = 0.0f, sy = 0.0f;
  for (int32_t e = e0 + j; e < e1; e += kChebLanes) {
    const float2 u = y[adj[e]];
    sx += u.x;
    sy += u.y;
  }
#pragma unroll
  for (int o = kChebLanes / 2; o > 0; o >>= 1) {
    sx += __shfl_xor(sx, o, kChebLanes);
    sy += __shfl_xor(sy, o, kChebLanes);
  }
  if (j != 0)
    return;
  const float2 yv = y[v];
  float px = yv.x, py = yv.y;
  if (e1 > e0) {
    px = 0.5f * (yv.x + sx * invDeg[v]);
    py = 0.5f * (yv.y + sy * invDeg[v]);
  }
  // mu(P) y = (2 / c) P y - y
  float mx = (2.0f / kChebC) * px - yv.x, my = (2.0f / kChebC) * py - yv.y;
  if (!first) {
    const float2 q = yPrev[v];
    mx = 2.0f * mx - q.x;
    my = 2.0f * my - q.y;
  }
  out[v] = make_float2(mx, my);
}

// SplitMix64 (Steele, Lea & Flood 2014) of seed + (v + 1) * golden gamma:
// bits 40-63 and 16-39 as two uniform coordinates in [-1, 1).
__device__ __host__ inline uint64_t splitmix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__global__ void k_uniform_start(uint64_t seed, int64_t n, float2 *__restrict__ out) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n)
    return;
  const uint64_t z = splitmix64(seed + (uint64_t)(v + 1) * 0x9E3779B97F4A7C15ull);
  const float sc = 1.0f / 16777216.0f;
  out[v] = make_float2((float)(z >> 40) * sc * 2.0f - 1.0f,
                       (float)((z >> 16) & 0xFFFFFFull) * sc * 2.0f - 1.0f);
}

// Per-block sums of x, y, x^2, y^2 (FP64).
__global__ __launch_bounds__(kB) void k_moments2(const float2 *__restrict__ x, int64_t n,
                                                 double *__restrict__ part) {
  __shared__ double sm[kB / kWave][4];
  double m[4] = {0, 0, 0, 0};
  for (int64_t v = (int64_t)blockIdx.x * kB + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * kB) {
    const double a = x[v].x, b = x[v].y;
    m[0] += a;
    m[1] += b;
    m[2] += a * a;
    m[3] += b * b;
  }
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const double w = wave_sum(m[j]);
    if (lane == 0)
      sm[wave][j] = w;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    double t = 0.0;
    for (int w = 0; w < kB / kWave; w++)
      t += sm[w][threadIdx.x];
    part[blockIdx.x * 4 + threadIdx.x] = t;
  }
}

// Totals of k_moments2's per-block partials: stats = {sum x, sum y,
// sum x^2, sum y^2} (one block; partials summed in block order).
__global__ void k_stats_finish(const double *__restrict__ part, int blocks,
                               double *__restrict__ stats) {
  __shared__ double sm[kB / kWave][4];
  double m[4] = {0, 0, 0, 0};
  for (int b = threadIdx.x; b < blocks; b += kB)
#pragma unroll
    for (int j = 0; j < 4; j++)
      m[j] += part[4 * b + j];
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const double w = wave_sum(m[j]);
    if (lane == 0)
      sm[wave][j] = w;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    double t = 0.0;
    for (int w = 0; w < kB / kWave; w++)
      t += sm[w][threadIdx.x];
    stats[threadIdx.x] = t;
  }
}

// Feature columns c, c+1 = (x - mean) * invStd per coordinate, as BF16, from
// the totals of k_stats_finish.
__global__ void k_put_feature(const float2 *__restrict__ x, int64_t n, int c,
                              const double *__restrict__ stats, uint16_t *__restrict__ X0) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n)
    return;
  const double nn = (double)n;
  const double mx = stats[0] / nn, my = stats[1] / nn;
  const double vx = stats[2] / nn - mx * mx, vy = stats[3] / nn - my * my;
  const float sx = vx > 1e-30 ? (float)(1.0 / sqrt(vx)) : 0.0f;
  const float sy = vy > 1e-30 ? (float)(1.0 / sqrt(vy)) : 0.0f;
  X0[v * kF + c] = f2bf((x[v].x - (float)mx) * sx);
  X0[v * kF + c + 1] = f2bf((x[v].y - (float)my) * sy);
}

// Proposal placement: out = (y - mean) * radius / RMS radius of y, in FP64,
// from the totals of k_stats_finish (a zero spread leaves y centred).
__global__ void k_place(const float2 *__restrict__ y, int64_t n,
                        const double *__restrict__ stats, double radius,
                        double2 *__restrict__ out) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n)
    return;
  const double nn = (double)n;
  const double mx = stats[0] / nn, my = stats[1] / nn;
  const double var = stats[2] / nn - mx * mx + stats[3] / nn - my * my;
  const double s = var > 1e-300 ? radius / sqrt(var) : 1.0;
  out[v] = make_double2(((double)y[v].x - mx) * s, ((double)y[v].y - my) * s);
}

// Degree term, constant and zero padding (columns 10..15).
__global__ void k_put_tail(const int32_t *__restrict__ off, int64_t n,
                           uint16_t *__restrict__ X0) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n)
    return;
  X0[v * kF + 10] = f2bf(logf(1.0f + (float)(off[v + 1] - off[v])) * 0.25f);
  X0[v * kF + 11] = f2bf(1.0f);
  for (int c = 12; c < kF; c++)
    X0[v * kF + c] = 0;
}

// ---- pivot MDS (Brandes & Pich): classical MDS restricted to the columns of
// k <= 64 pivot vertices. D = BFS hop distances (n x k), C = -1/2 of D^2
// double-centred, axes = C v_i / mu_i^(1/4) for the top two eigenpairs
// (v_i, mu_i) of C^T C (so each axis carries sqrt of the corresponding
// classical-MDS eigenvalue, as in full MDS).
constexpr int kPiv = 64;
constexpr uint16_t kFar = 0xFFFF;

// dist = kFar everywhere; visited/frontier = the pivots' bits.
__global__ void k_piv_init(uint16_t *__restrict__ dist, uint64_t *__restrict__ vis,
                           uint64_t *__restrict__ front, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n * kPiv)
    dist[i] = kFar;
  if (i < n) {
    vis[i] = 0;
    front[i] = 0;
  }
}

__global__ void k_piv_seed(const int32_t *__restrict__ piv, int k, uint16_t *__restrict__ dist,
                           uint64_t *__restrict__ vis, uint64_t *__restrict__ front) {
  // one thread: pivots may repeat, so their bits are merged serially
  if (blockIdx.x != 0 || threadIdx.x != 0)
    return;
  for (int j = 0; j < k; j++) {
    const int32_t v = piv[j];
    dist[(int64_t)v * kPiv + j] = 0;
    vis[v] |= 1ull << j;
    front[v] |= 1ull << j;
  }
}

// One BFS level for all pivots at once: a vertex joins pivot j's next
// frontier when a neighbour is on j's frontier and v has not been reached
// from j. Reads only `front` and writes only `next`, `vis[v]` and v's row.
__global__ __launch_bounds__(kB) void k_piv_level(const int32_t *__restrict__ off,
                                                  const int32_t *__restrict__ adj,
                                                  const uint64_t *__restrict__ front,
                                                  uint64_t *__restrict__ vis,
                                                  uint64_t *__restrict__ next,
                                                  uint16_t *__restrict__ dist, int64_t n,
                                                  int level, int32_t *__restrict__ changed) {
  const int64_t v = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (v >= n)
    return;
  uint64_t m = 0;
  for (int32_t e = off[v]; e < off[v + 1]; e++)
    m |= front[adj[e]];
  m &= ~vis[v];
  next[v] = m;
  if (m) {
    vis[v] |= m;
    const uint16_t d = (uint16_t)(level + 1);
    for (uint64_t b = m; b; b &= b - 1)
      dist[v * kPiv + __builtin_ctzll(b)] = d;
    *changed = 1;
  }
}

__device__ __forceinline__ double piv_d2(uint16_t d, uint16_t far) {
  const double x = (double)(d == kFar ? far : d);
  return x * x;
}

// Row means of D^2 per vertex and per-block column sums (fixed order).
__global__ __launch_bounds__(kB) void k_piv_moments(const uint16_t *__restrict__ dist,
                                                    int64_t n, int k, uint16_t far,
                                                    double *__restrict__ rowMean,
                                                    double *__restrict__ colPart) {
  const int64_t chunk = (n + gridDim.x - 1) / gridDim.x;
  const int64_t b0 = blockIdx.x * chunk, b1 = b0 + chunk < n ? b0 + chunk : n;
  for (int64_t v = b0 + threadIdx.x; v < b1; v += kB) {
    double r = 0.0;
    for (int j = 0; j < k; j++)
      r += piv_d2(dist[v * kPiv + j], far);
    rowMean[v] = r / k;
  }
  if (threadIdx.x < k) {
    double c = 0.0;
    for (int64_t v = b0; v < b1; v++)
      c += piv_d2(dist[v * kPiv + threadIdx.x], far);
    colPart[blockIdx.x * kPiv + threadIdx.x] = c;
  }
}

__device__ __forceinline__ double piv_c(const uint16_t *__restrict__ dist, int64_t v, int j,
                                        uint16_t far, const double *__restrict__ rowMean,
                                        const double *__restrict__ colMean, double all) {
  return -0.5 * (piv_d2(dist[v * kPiv + j], far) - rowMean[v] - colMean[j] + all);
}

// Per-block partial Gram matrix C^T C (k x k, FP64), rows staged in LDS.
constexpr int kGramRows = 32;
__global__ __launch_bounds__(kB) void k_piv_gram(const uint16_t *__restrict__ dist, int64_t n,
                                                 int k, uint16_t far,
                                                 const double *__restrict__ rowMean,
                                                 const double *__restrict__ colMean,
                                                 double all, double *__restrict__ gramPart) {
  __shared__ double cs[kGramRows][kPiv];
  double acc[kPiv * kPiv / kB];
#pragma unroll
  for (int q = 0; q < kPiv * kPiv / kB; q++)
    acc[q] = 0.0;
  const int64_t chunk = (n + gridDim.x - 1) / gridDim.x;
  const int64_t b0 = blockIdx.x * chunk, b1 = b0 + chunk < n ? b0 + chunk : n;
  for (int64_t base = b0; base < b1; base += kGramRows) {
    __syncthreads();
    for (int i = threadIdx.x; i < kGramRows * kPiv; i += kB) {
      const int r = i / kPiv, j = i % kPiv;
      const int64_t v = base + r;
      cs[r][j] = (v < b1 && j < k) ? piv_c(dist, v, j, far, rowMean, colMean, all) : 0.0;
    }
    __syncthreads();
    const int rows = b1 - base < kGramRows ? (int)(b1 - base) : kGramRows;
    for (int r = 0; r < rows; r++) {
#pragma unroll
      for (int q = 0; q < kPiv * kPiv / kB; q++) {
        const int e = q * kB + threadIdx.x;
        acc[q] += cs[r][e / kPiv] * cs[r][e % kPiv];
      }
    }
  }
#pragma unroll
  for (int q = 0; q < kPiv * kPiv / kB; q++)
    gramPart[(int64_t)blockIdx.x * kPiv * kPiv + q * kB + threadIdx.x] = acc[q];
}

// Sums `blocks` partial k x k (padded 64 x 64) matrices in block order.
__global__ void k_piv_gram_sum(const double *__restrict__ part, int blocks,
                               double *__restrict__ out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= kPiv * kPiv)
    return;
  double s = 0.0;
  for (int b = 0; b < blocks; b++)
    s += part[(int64_t)b * kPiv * kPiv + e];
  out[e] = s;
}

// Coordinates = C [e1 e2] (axes pre-scaled on the host), FP64 accumulation.
__global__ __launch_bounds__(kB) void k_piv_project(const uint16_t *__restrict__ dist,
                                                    int64_t n, int k, uint16_t far,
                                                    const double *__restrict__ rowMean,
                                                    const double *__restrict__ colMean,
                                                    double all, const double *__restrict__ axes,
                                                    double2 *__restrict__ out) {
  const int64_t v = (int64_t)blockIdx.x * kB + threadIdx.x;
  if (v >= n)
    return;
  double x = 0.0, y = 0.0;
  for (int j = 0; j < k; j++) {
    const double c = piv_c(dist, v, j, far, rowMean, colMean, all);
    x += c * axes[j];
    y += c * axes[kPiv + j];
  }
  out[v] = make_double2(x, y);
}

// Per-block sums of x, y, x^2 + y^2 of the raw coordinates (fixed order).
__global__ __launch_bounds__(kB) void k_piv_stats(const double2 *__restrict__ raw, int64_t n,
                                                  double *__restrict__ part) {
  __shared__ double sm[kB / kWave][3];
  double a = 0.0, b = 0.0, c = 0.0;
  for (int64_t v = (int64_t)blockIdx.x * kB + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * kB) {
    const double2 p = raw[v];
    a += p.x;
    b += p.y;
    c += p.x * p.x + p.y * p.y;
  }
  a = wave_sum(a);
  b = wave_sum(b);
  c = wave_sum(c);
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  if (lane == 0) {
    sm[wave][0] = a;
    sm[wave][1] = b;
    sm[wave][2] = c;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    double t = 0.0;
    for (int w = 0; w < kB / kWave; w++)
      t += sm[w][threadIdx.x];
    part[blockIdx.x * 3 + threadIdx.x] = t;
  }
}

// Skip term: centred, scaled to RMS radius 1 (one scale for both axes);
// FP32 copy for the head and BF16 feature columns 12-13.
__global__ void k_piv_finish(const double2 *__restrict__ raw, int64_t n, double mx, double my,
                             double inv, float2 *__restrict__ skip, uint16_t *__restrict__ X0) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n)
    return;
  const float x = (float)((raw[v].x - mx) * inv), y = (float)((raw[v].y - my) * inv);
  skip[v] = make_float2(x, y);
  X0[v * kF + 12] = f2bf(x);
  X0[v * kF + 13] = f2bf(y);
}

// Head outputs 0-1 of every row, packed.
__global__ void k_head_xy(const float *__restrict__ Y, int64_t n, float2 *__restrict__ out) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v < n)
    out[v] = make_float2(Y[v * kO], Y[v * kO + 1]);
}

// Head outputs 0-1 += skip (after the head product).
__global__ void k_add_skip(const float2 *__restrict__ skip, int64_t n, float *__restrict__ Y) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n)
    return;
  Y[v * kO] += skip[v].x;
  Y[v * kO + 1] += skip[v].y;
}

// Eigenpairs of a symmetric m x m matrix (row-major, m <= 64) by cyclic
// Jacobi rotations; returns eigenvalues in w and eigenvectors as columns of V.
void jacobi_eigen(std::vector<double> A, int m, std::vector<double> &w,
                  std::vector<double> &V) {
  V.assign((size_t)m * m, 0.0);
  for (int i = 0; i < m; i++)
    V[(size_t)i * m + i] = 1.0;
  for (int sweep = 0; sweep < 100; sweep++) {
    double off = 0.0, diag = 0.0;
    for (int p = 0; p < m; p++)
      for (int q = 0; q < m; q++)
        (p == q ? diag : off) += A[(size_t)p * m + q] * A[(size_t)p * m + q];
    if (off <= 1e-30 * (diag > 0.0 ? diag : 1.0))
      break;
    for (int p = 0; p < m - 1; p++)
      for (int q = p + 1; q < m; q++) {
        const double apq = A[(size_t)p * m + q];
        if (apq == 0.0)
          continue;
        const double app = A[(size_t)p * m + p], aqq = A[(size_t)q * m + q];
        const double theta = (aqq - app) / (2.0 * apq);
        const double t = (theta >= 0.0 ? 1.0 : -1.0) /
                         (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), sn = t * c;
        for (int r = 0; r < m; r++) {
          const double arp = A[(size_t)r * m + p], arq = A[(size_t)r * m + q];
          A[(size_t)r * m + p] = c * arp - sn * arq;
          A[(size_t)r * m + q] = sn * arp + c * arq;
        }
        for (int r = 0; r < m; r++) {
          const double apr = A[(size_t)p * m + r], aqr = A[(size_t)q * m + r];
          A[(size_t)p * m + r] = c * apr - sn * aqr;
          A[(size_t)q * m + r] = sn * apr + c * aqr;
        }
        for (int r = 0; r < m; r++) {
          const double vrp = V[(size_t)r * m + p], vrq = V[(size_t)r * m + q];
          V[(size_t)r * m + p] = c * vrp - sn * vrq;
          V[(size_t)r * m + q] = sn * vrp + c * vrq;
        }
      }
  }
  w.resize(m);
  for (int i = 0; i < m; i++)
    w[i] = A[(size_t)i * m + i];
}

} // namespace

struct gvizhip_gnn {
  int64_t cap = 0, capNnz = 0, n = 0;
  hipStream_t stream = nullptr;
  DevBuf<int32_t> off, adj;
  DevBuf<float> invDeg, feat32;
  DevBuf<uint16_t> X0, M0, H[3], M[2], dY, dZ;
  DevBuf<float> Y, dA, target;
  DevBuf<float> params, grads, adamM, adamV;
  DevBuf<uint16_t> wt[4], wr[4];
  DevBuf<float> partial;
  DevBuf<double> moments;
  DevBuf<float> diff[3];
  DevBuf<double> stats;
  DevBuf<double> placed;
  double *pinned = nullptr; // host staging of proposals (2 cap doubles)
  // pivot MDS
  DevBuf<int32_t> piv, changed;
  DevBuf<uint16_t> pdist;
  DevBuf<uint64_t> pvis, pfront[2];
  DevBuf<double> prow, pcol, pgram, paxes, praw;
  DevBuf<float> skip;
  bool hasSkip = false;
  int64_t step = 0;
  bool weightsDirty = true;
};

namespace {

int rows_blocks(int64_t n) { return (int)((n + kRows - 1) / kRows); }

int refresh_weights(gvizhip_gnn *g) {
  if (!g->weightsDirty)
    return 0;
  for (int l = 0; l < 4; l++) {
    const int cnt = kK[l] * kN[l];
    hipLaunchKernelGGL(k_weights_bf16, dim3(grid_for(cnt, kB)), dim3(kB), 0, g->stream,
                       g->params.p, g->wt[l].p, g->wr[l].p, kOffW[l], kK[l], kN[l]);
  }
  GVH_CHECK_LAUNCH();
  g->weightsDirty = false;
  return 0;
}

template <int NT>
void launch_gemm(gvizhip_gnn *g, const uint16_t *A1, const uint16_t *A2, int Ka,
                 const uint16_t *Bt, int K, const float *bias, int relu, uint16_t *outB,
                 float *outF) {
  const int64_t waves = (g->n + 15) / 16;
  const unsigned blocks = (unsigned)((waves + kB / kWave - 1) / (kB / kWave));
  if (K % 32 == 0)
    hipLaunchKernelGGL(k_gemm32<NT>, dim3(blocks), dim3(kB), 0, g->stream, A1, A2, Ka,
                       g->n, Bt, K, bias, relu, outB, outF);
  else
    hipLaunchKernelGGL(k_gemm<NT>, dim3(blocks), dim3(kB), 0, g->stream, A1, A2, Ka, g->n,
                       Bt, K, bias, relu, outB, outF);
}

int forward(gvizhip_gnn *g) {
  if (refresh_weights(g))
    return -1;
  const int64_t n = g->n;
  hipStream_t st = g->stream;
  const float *b[4];
  for (int l = 0; l < 4; l++)
    b[l] = g->params.p + kOffW[l] + kK[l] * kN[l];
  // layer 1
  hipLaunchKernelGGL(k_aggregate<kF>, dim3(grid_for(n * (kF / 8), kB)), dim3(kB), 0, st,
                     g->off.p, g->adj.p, g->invDeg.p, g->X0.p, g->M0.p, n);
  launch_gemm<4>(g, g->X0.p, g->M0.p, kF, g->wt[0].p, 2 * kF, b[0], 1, g->H[0].p, nullptr);
  // layers 2 and 3
  for (int l = 1; l < 3; l++) {
    hipLaunchKernelGGL(k_aggregate<kH>, dim3(grid_for(n * (kH / 8), kB)), dim3(kB), 0, st,
                       g->off.p, g->adj.p, g->invDeg.p, g->H[l - 1].p, g->M[l - 1].p, n);
    launch_gemm<4>(g, g->H[l - 1].p, g->M[l - 1].p, kH, g->wt[l].p, 2 * kH, b[l], 1,
                   g->H[l].p, nullptr);
  }
  // head (+ the pivot-MDS skip term when one is set)
  launch_gemm<1>(g, g->H[2].p, g->H[2].p, kH, g->wt[3].p, kH, b[3], 0, nullptr, g->Y.p);
  if (g->hasSkip)
    hipLaunchKernelGGL(k_add_skip, dim3(grid_for(n, kB)), dim3(kB), 0, st,
                       (const float2 *)g->skip.p, n, g->Y.p);
  GVH_CHECK_LAUNCH();
  return 0;
}

// Procrustes fit of the head output to the target: loss and the epilogue
// constants of k_loss_grad.
int fit_loss(gvizhip_gnn *g, double *lossOut, LossFit *fitOut) {
  const int blocks = (int)std::min<int64_t>(grid_for(g->n, kB), 256);
  hipLaunchKernelGGL(k_head_moments, dim3(blocks), dim3(kB), 0, g->stream, g->Y.p,
                     g->target.p, g->n, g->moments.p);
  GVH_CHECK_LAUNCH();
  std::vector<double> part(10 * blocks);
  GVH_CHECK(hipMemcpyAsync(part.data(), g->moments.p, sizeof(double) * 10 * blocks,
                           hipMemcpyDeviceToHost, g->stream));
  GVH_CHECK(hipStreamSynchronize(g->stream));
  double m[10] = {0};
  for (int b = 0; b < blocks; b++)
    for (int j = 0; j < 10; j++)
      m[j] += part[b * 10 + j];
  const double nn = (double)g->n;
  const double mpx = m[0] / nn, mpy = m[1] / nn, mtx = m[2] / nn, mty = m[3] / nn;
  // centred cross moments M = P^T T and squared norms
  const double a = m[4] - nn * mpx * mtx, bq = m[5] - nn * mpx * mty;
  const double c = m[6] - nn * mpy * mtx, d = m[7] - nn * mpy * mty;
  const double pp = m[8] - nn * (mpx * mpx + mpy * mpy);
  const double tt = m[9] - nn * (mtx * mtx + mty * mty);
  // best rotation or reflection maximising tr(R^T M)
  const double rotV = std::hypot(a + d, c - bq), refV = std::hypot(a - d, bq + c);
  double r00, r01, r10, r11, tr;
  if (rotV >= refV) {
    const double th = std::atan2(c - bq, a + d);
    r00 = std::cos(th), r01 = -std::sin(th), r10 = std::sin(th), r11 = std::cos(th);
    tr = rotV;
  } else {
    const double th = std::atan2(bq + c, a - d);
    r00 = std::cos(th), r01 = std::sin(th), r10 = std::sin(th), r11 = -std::cos(th);
    tr = refV;
  }
  const double s = pp > 0.0 ? tr / pp : 0.0;
  const double loss = (tt > 0.0 && pp > 0.0) ? 1.0 - tr * tr / (pp * tt) : 1.0;
  if (lossOut)
    *lossOut = loss;
  if (fitOut) {
    fitOut->mpx = (float)mpx, fitOut->mpy = (float)mpy;
    fitOut->mtx = (float)mtx, fitOut->mty = (float)mty;
    fitOut->r00 = (float)r00, fitOut->r01 = (float)r01;
    fitOut->r10 = (float)r10, fitOut->r11 = (float)r11;
    fitOut->s = (float)s;
    fitOut->scale = tt > 0.0 ? (float)(2.0 * s / tt) : 0.0f;
  }
  return 0;
}

template <int NT>
int weight_grad(gvizhip_gnn *g, const uint16_t *A1, const uint16_t *A2, int Ka, int K,
                const uint16_t *D, int layer) {
  constexpr int Nc = 16 * NT;
  const int blocks = rows_blocks(g->n);
  hipLaunchKernelGGL(k_wgrad<NT>, dim3(blocks), dim3(kB), 0, g->stream, A1, A2, Ka, g->n,
                     D, K, g->partial.p);
  hipLaunchKernelGGL(k_reduce_blocks, dim3(grid_for((int64_t)K * Nc, kB)), dim3(kB), 0,
                     g->stream, g->partial.p, blocks, (int64_t)K * Nc,
                     g->grads.p + kOffW[layer]);
  hipLaunchKernelGGL(k_colsum, dim3(blocks), dim3(kB), 0, g->stream, D, Nc, g->n,
                     g->partial.p);
  hipLaunchKernelGGL(k_reduce_blocks, dim3(1), dim3(kB), 0, g->stream, g->partial.p,
                     blocks, (int64_t)Nc, g->grads.p + kOffW[layer] + (int64_t)K * Nc);
  GVH_CHECK_LAUNCH();
  return 0;
}

// Gradients of the loss with respect to every parameter (after forward and
// fit_loss).
int backward(gvizhip_gnn *g, const LossFit &fit) {
  const int64_t n = g->n;
  hipStream_t st = g->stream;
  hipLaunchKernelGGL(k_loss_grad, dim3(grid_for(n, kB)), dim3(kB), 0, st, g->Y.p,
                     g->target.p, n, fit, g->dY.p);
  // head: dW4 = H3^T dY, dH3 = dY W4^T
  if (weight_grad<1>(g, g->H[2].p, g->H[2].p, kH, kH, g->dY.p, 3))
    return -1;
  launch_gemm<4>(g, g->dY.p, g->dY.p, kO, g->wr[3].p, kO, nullptr, 0, nullptr, g->dA.p);
  hipLaunchKernelGGL(k_mask, dim3(grid_for(n * kH, kB)), dim3(kB), 0, st, g->dA.p,
                     g->H[2].p, n * kH, g->dZ.p);
  // layers 3 and 2: dW = [H | M]^T dZ, then dZ of the layer below
  for (int l = 2; l >= 1; l--) {
    if (weight_grad<4>(g, g->H[l - 1].p, g->M[l - 1].p, kH, 2 * kH, g->dZ.p, l))
      return -1;
    launch_gemm<8>(g, g->dZ.p, g->dZ.p, kH, g->wr[l].p, kH, nullptr, 0, nullptr, g->dA.p);
    hipLaunchKernelGGL(k_bwd_combine, dim3(grid_for(n * kWave, kB)), dim3(kB), 0, st,
                       g->off.p, g->adj.p, g->invDeg.p, g->dA.p, g->H[l - 1].p, n,
                       g->dZ.p);
  }
  // layer 1
  if (weight_grad<4>(g, g->X0.p, g->M0.p, kF, 2 * kF, g->dZ.p, 0))
    return -1;
  GVH_CHECK_LAUNCH();
  return 0;
}

// Totals {sum x, sum y, sum x^2, sum y^2} of a 2-column signal into
// stats[0..3] on the device (no host synchronisation).
int moments_of(gvizhip_gnn *g, const float2 *x, double *stats) {
  const int blocks = (int)std::min<int64_t>(grid_for(g->n, kB), 256);
  hipLaunchKernelGGL(k_moments2, dim3(blocks), dim3(kB), 0, g->stream, x, g->n,
                     g->moments.p);
  hipLaunchKernelGGL(k_stats_finish, dim3(1), dim3(kB), 0, g->stream,
                     (const double *)g->moments.p, blocks, stats);
  GVH_CHECK_LAUNCH();
  return 0;
}

// Input features from the start in diff[0]: columns 0-1 the start, 2-3 /
// 4-5 / 6-7 / 8-9 its Chebyshev filters of degree 2 / 4 / 8 / 16 (each pair
// centred and scaled to unit variance per coordinate), column 10
// log(1 + degree) / 4, column 11 one, the rest zero. Enqueued only.
int build_features(gvizhip_gnn *g) {
  const int64_t n = g->n;
  g->hasSkip = false;
  const unsigned grid = grid_for(n, kB);
  float2 *buf[3] = {(float2 *)g->diff[0].p, (float2 *)g->diff[1].p,
                    (float2 *)g->diff[2].p};
  // y_k lives in buf[k % 3]; the start stays in buf[0] until y_3 replaces it
  int mark = 0;
  for (int k = 0; k <= kChebDeg; k++) {
    if (k > 0)
      hipLaunchKernelGGL(k_cheb, dim3(grid_for(n * kChebLanes, kB)), dim3(kB), 0, g->stream,
                         g->off.p, g->adj.p,
                         g->invDeg.p, (const float2 *)buf[(k - 1) % 3],
                         (const float2 *)buf[(k + 1) % 3], k == 1 ? 1 : 0, n,
                         buf[k % 3]);
    if (k == 0 || (mark < 4 && k == kChebMarks[mark])) {
      const int f = k == 0 ? 0 : ++mark;
      double *st = g->stats.p + 4 * f;
      if (moments_of(g, buf[k % 3], st))
        return -1;
      hipLaunchKernelGGL(k_put_feature, dim3(grid), dim3(kB), 0, g->stream,
                         (const float2 *)buf[k % 3], n, 2 * f, (const double *)st, g->X0.p);
    }
  }
  hipLaunchKernelGGL(k_put_tail, dim3(grid), dim3(kB), 0, g->stream, g->off.p, n, g->X0.p);
  GVH_CHECK_LAUNCH();
  return 0;
}

int check_graph(gvizhip_gnn *g) {
  if (!g || g->n <= 0) {
    set_error("gvizhip_gnn: no graph set");
    return -1;
  }
  return 0;
}

} // namespace

extern "C" {

int gvizhip_gnn_create(gvizhip_gnn **out, int64_t maxVertices, int64_t maxEntries) {
  if (!out || maxVertices <= 0 || maxEntries < 0 || maxVertices > INT32_MAX ||
      maxEntries > INT32_MAX) {
    set_error("gvizhip_gnn_create: invalid arguments");
    return -1;
  }
  *out = nullptr;
  int dev = 0;
  if (hipGetDeviceCount(&dev) != hipSuccess || dev <= 0) {
    set_error("gvizhip_gnn_create: no GPU");
    return -1;
  }
  auto *g = new gvizhip_gnn;
  const int64_t n = maxVertices;
  g->cap = n;
  g->capNnz = maxEntries;
  const int64_t rb = (n + kRows - 1) / kRows;
  bool bad = hipStreamCreate(&g->stream) != hipSuccess || g->off.alloc(n + 1) ||
             g->adj.alloc(maxEntries) || g->invDeg.alloc(n) || g->X0.alloc(n * kF) ||
             g->M0.alloc(n * kF) || g->dY.alloc(n * kO) || g->dZ.alloc(n * kH) ||
             g->Y.alloc(n * kO) || g->dA.alloc(n * 2 * kH) || g->target.alloc(2 * n) ||
             g->feat32.alloc(1) || g->diff[0].alloc(2 * n) || g->diff[1].alloc(2 * n) ||
             g->diff[2].alloc(2 * n) || g->stats.alloc(4 * 8) || g->placed.alloc(2 * n) ||
             g->params.alloc(kParams) || g->grads.alloc(kParams) ||
             g->adamM.alloc(kParams) || g->adamV.alloc(kParams) ||
             g->partial.alloc(rb * 2 * kH * kH) || g->moments.alloc(10 * 256);
  for (int l = 0; l < 3 && !bad; l++)
    bad = g->H[l].alloc(n * kH) != 0;
  for (int l = 0; l < 2 && !bad; l++)
    bad = g->M[l].alloc(n * kH) != 0;
  for (int l = 0; l < 4 && !bad; l++)
    bad = g->wt[l].alloc(kK[l] * kN[l]) || g->wr[l].alloc(kK[l] * kN[l]);
  if (bad) {
    if (!gvizhip_last_error()[0])
      set_error("gvizhip_gnn_create: allocation failed");
    gvizhip_gnn_destroy(g);
    return -1;
  }
  GVH_CHECK(hipMemsetAsync(g->params.p, 0, sizeof(float) * kParams, g->stream));
  GVH_CHECK(hipMemsetAsync(g->adamM.p, 0, sizeof(float) * kParams, g->stream));
  GVH_CHECK(hipMemsetAsync(g->adamV.p, 0, sizeof(float) * kParams, g->stream));
  GVH_CHECK(hipStreamSynchronize(g->stream));
  *out = g;
  return 0;
}

int64_t gvizhip_gnn_param_count(void) { return kParams; }

int gvizhip_gnn_set_params(gvizhip_gnn *g, const float *params) {
  if (!g || !params) {
    set_error("gvizhip_gnn_set_params: invalid arguments");
    return -1;
  }
  GVH_CHECK(hipMemcpyAsync(g->params.p, params, sizeof(float) * kParams,
                           hipMemcpyHostToDevice, g->stream));
  GVH_CHECK(hipMemsetAsync(g->adamM.p, 0, sizeof(float) * kParams, g->stream));
  GVH_CHECK(hipMemsetAsync(g->adamV.p, 0, sizeof(float) * kParams, g->stream));
  GVH_CHECK(hipStreamSynchronize(g->stream));
  g->step = 0;
  g->weightsDirty = true;
  return 0;
}

int gvizhip_gnn_get_params(gvizhip_gnn *g, float *params) {
  if (!g || !params) {
    set_error("gvizhip_gnn_get_params: invalid arguments");
    return -1;
  }
  GVH_CHECK(hipMemcpyAsync(params, g->params.p, sizeof(float) * kParams,
                           hipMemcpyDeviceToHost, g->stream));
  GVH_CHECK(hipStreamSynchronize(g->stream));
  return 0;
}

int gvizhip_gnn_set_graph(gvizhip_gnn *g, int64_t n, const int64_t *offsets,
                          const int64_t *adj, const float *features) {
  if (!g || n <= 0 || n > g->cap || !offsets || !features ||
      offsets[n] > g->capNnz || offsets[0] != 0 || (offsets[n] > 0 && !adj)) {
    set_error("gvizhip_gnn_set_graph: invalid arguments");
    return -1;
  }
  const int64_t m = offsets[n];
  std::vector<int32_t> o(n + 1), a(m > 0 ? m : 1);
  std::vector<float> inv(n);
  std::vector<uint16_t> x(n * kF);
  for (int64_t v = 0; v <= n; v++)
    o[v] = (int32_t)offsets[v];
  for (int64_t v = 0; v < n; v++) {
    if (offsets[v + 1] < offsets[v]) {
      set_error("gvizhip_gnn_set_graph: offsets are not monotone");
      return -1;
    }
    const int64_t d = offsets[v + 1] - offsets[v];
    inv[v] = d > 0 ? 1.0f / (float)d : 0.0f;
  }
  for (int64_t e = 0; e < m; e++) {
    if (adj[e] < 0 || adj[e] >= n) {
      set_error("gvizhip_gnn_set_graph: neighbour id out of range");
      return -1;
    }
    a[e] = (int32_t)adj[e];
  }
  for (int64_t i = 0; i < n * kF; i++) {
    const uint32_t u = [&] {
      uint32_t b;
      std::memcpy(&b, &features[i], 4);
      return b;
    }();
    x[i] = (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
  }
  g->n = n;
  g->hasSkip = false;
  GVH_CHECK(hipMemcpyAsync(g->off.p, o.data(), sizeof(int32_t) * (n + 1),
                           hipMemcpyHostToDevice, g->stream));
  if (m > 0)
    GVH_CHECK(hipMemcpyAsync(g->adj.p, a.data(), sizeof(int32_t) * m,
                             hipMemcpyHostToDevice, g->stream));
  GVH_CHECK(hipMemcpyAsync(g->invDeg.p, inv.data(), sizeof(float) * n,
                           hipMemcpyHostToDevice, g->stream));
  GVH_CHECK(hipMemcpyAsync(g->X0.p, x.data(), sizeof(uint16_t) * n * kF,
                           hipMemcpyHostToDevice, g->stream));
  GVH_CHECK(hipStreamSynchronize(g->stream));
  return 0;
}


// Uploads the graph like gvizhip_gnn_set_graph and builds the features on
// the device from start (n x 2): columns 0-1 the start, 2-3 / 4-5 / 6-7 / 8-9
// the start after 4 / 16 / 64 / 256 lazy random-walk steps (each pair
// centred and scaled to unit variance per coordinate), 10 log(1 + degree) /
// 4, 11 one, the rest zero.
int gvizhip_gnn_set_graph_diffused(gvizhip_gnn *g, int64_t n, const int64_t *offsets,
                                   const int64_t *adj, const double *start) {
  if (!g || !start || n <= 0 || n > g->cap) {
    set_error("gvizhip_gnn_set_graph_diffused: invalid arguments");
    return -1;
  }
  std::vector<float> zeros((size_t)n * kF, 0.0f);
  if (gvizhip_gnn_set_graph(g, n, offsets, adj, zeros.data()))
    return -1;
  return gvizhip_gnn_set_start(g, start);
}

// Rebuilds the diffused features of the current graph from a new start
// (columns 0-11; 12-15 zero, the skip term removed).
int gvizhip_gnn_set_start(gvizhip_gnn *g, const double *start) {
  if (check_graph(g) || !start) {
    if (g && !start)
      set_error("gvizhip_gnn_set_start: invalid arguments");
    return -1;
  }
  const int64_t n = g->n;
  std::vector<float> x(2 * n);
  for (int64_t i = 0; i < 2 * n; i++)
    x[i] = (float)start[i];
  GVH_CHECK(hipMemcpyAsync(g->diff[0].p, x.data(), sizeof(float) * 2 * n,
                           hipMemcpyHostToDevice, g->stream));
  if (build_features(g))
    return -1;
  GVH_CHECK(hipStreamSynchronize(g->stream));
  return 0;
}

int gvizhip_gnn_set_start_seeded(gvizhip_gnn *g, uint64_t seed) {
  if (check_graph(g))
    return -1;
  hipLaunchKernelGGL(k_uniform_start, dim3(grid_for(g->n, kB)), dim3(kB), 0, g->stream,
                     seed, g->n, (float2 *)g->diff[0].p);
  GVH_CHECK_LAUNCH();
  if (build_features(g))
    return -1;
  GVH_CHECK(hipStreamSynchronize(g->stream));
  return 0;
}

int gvizhip_gnn_propose(gvizhip_gnn *g, uint64_t seed, double radius, double *positions) {
  if (check_graph(g) || !positions || !(radius > 0.0)) {
    if (g && g->n > 0)
      set_error("gvizhip_gnn_propose: invalid arguments");
    return -1;
  }
  const int64_t n = g->n;
  if (!g->pinned &&
      hipHostMalloc((void **)&g->pinned, sizeof(double) * 2 * g->cap) != hipSuccess) {
    g->pinned = nullptr;
    set_error("gvizhip_gnn_propose: pinned staging allocation failed");
    return -1;
  }
  hipLaunchKernelGGL(k_uniform_start, dim3(grid_for(n, kB)), dim3(kB), 0, g->stream, seed,
                     n, (float2 *)g->diff[0].p);
  GVH_CHECK_LAUNCH();
  if (build_features(g) || forward(g))
    return -1;
  hipLaunchKernelGGL(k_head_xy, dim3(grid_for(n, kB)), dim3(kB), 0, g->stream, g->Y.p, n,
                     (float2 *)g->diff[0].p);
  if (moments_of(g, (const float2 *)g->diff[0].p, g->stats.p + 20))
    return -1;
  hipLaunchKernelGGL(k_place, dim3(grid_for(n, kB)), dim3(kB), 0, g->stream,
                     (const float2 *)g->diff[0].p, n, (const double *)(g->stats.p + 20),
                     radius, (double2 *)g->placed.p);
  GVH_CHECK_LAUNCH();
  GVH_CHECK(hipMemcpyAsync(g->pinned, g->placed.p, sizeof(double) * 2 * n,
                           hipMemcpyDeviceToHost, g->stream));
  GVH_CHECK(hipStreamSynchronize(g->stream));
  std::memcpy(positions, g->pinned, sizeof(double) * 2 * n);
  return 0;
}

// Pivot MDS of the current graph from BFS distances to `count` pivots (see
// k_piv_* above): feature columns 12-13 and the skip term added to the head.
int gvizhip_gnn_set_pivots(gvizhip_gnn *g, const int64_t *pivots, int64_t count,
                           double *coords) {
  if (check_graph(g))
    return -1;
  const int64_t n = g->n;
  hipStream_t st = g->stream;
  if (count == 0) {
    g->hasSkip = false;
    return 0;
  }
  if (!pivots || count < 0 || count > kPiv) {
    set_error("gvizhip_gnn_set_pivots: 0 to %d pivots", kPiv);
    return -1;
  }
  std::vector<int32_t> pv(count);
  for (int64_t j = 0; j < count; j++) {
    if (pivots[j] < 0 || pivots[j] >= n) {
      set_error("gvizhip_gnn_set_pivots: pivot out of range");
      return -1;
    }
    pv[j] = (int32_t)pivots[j];
  }
  const int64_t cap = g->cap;
  constexpr int kParts = 256;
  if (g->piv.alloc(kPiv) || g->changed.alloc(1) || g->pdist.alloc(cap * kPiv) ||
      g->pvis.alloc(cap) || g->pfront[0].alloc(cap) || g->pfront[1].alloc(cap) ||
      g->prow.alloc(cap) || g->pcol.alloc(kParts * kPiv + kPiv) ||
      g->pgram.alloc((int64_t)kParts * kPiv * kPiv + kPiv * kPiv) ||
      g->paxes.alloc(2 * kPiv) || g->praw.alloc(2 * cap) || g->skip.alloc(2 * cap))
    return -1;
  const int k = (int)count;
  GVH_CHECK(hipMemcpyAsync(g->piv.p, pv.data(), sizeof(int32_t) * k, hipMemcpyHostToDevice,
                           st));
  hipLaunchKernelGGL(k_piv_init, dim3(grid_for(n * kPiv, kB)), dim3(kB), 0, st, g->pdist.p,
                     g->pvis.p, g->pfront[0].p, n);
  hipLaunchKernelGGL(k_piv_seed, dim3(1), dim3(1), 0, st, g->piv.p, k, g->pdist.p, g->pvis.p,
                     g->pfront[0].p);
  GVH_CHECK_LAUNCH();
  int cur = 0, level = 0;
  for (; level < (int)kFar - 2; level++) {
    int32_t zero = 0, changed = 0;
    GVH_CHECK(hipMemcpyAsync(g->changed.p, &zero, sizeof(int32_t), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_piv_level, dim3(grid_for(n, kB)), dim3(kB), 0, st, g->off.p,
                       g->adj.p, g->pfront[cur].p, g->pvis.p, g->pfront[1 - cur].p,
                       g->pdist.p, n, level, g->changed.p);
    GVH_CHECK_LAUNCH();
    GVH_CHECK(hipMemcpyAsync(&changed, g->changed.p, sizeof(int32_t), hipMemcpyDeviceToHost,
                             st));
    GVH_CHECK(hipStreamSynchronize(st));
    cur = 1 - cur;
    if (!changed)
      break;
  }
  // vertices a pivot never reaches count as one hop beyond the farthest
  const uint16_t far = (uint16_t)(level + 1);
  const int parts = (int)std::min<int64_t>(kParts, (n + kB - 1) / kB);
  hipLaunchKernelGGL(k_piv_moments, dim3(parts), dim3(kB), 0, st, g->pdist.p, n, k, far,
                     g->prow.p, g->pcol.p);
  GVH_CHECK_LAUNCH();
  std::vector<double> colPart((size_t)parts * kPiv);
  GVH_CHECK(hipMemcpyAsync(colPart.data(), g->pcol.p, sizeof(double) * parts * kPiv,
                           hipMemcpyDeviceToHost, st));
  GVH_CHECK(hipStreamSynchronize(st));
  std::vector<double> colMean(kPiv, 0.0);
  double all = 0.0;
  for (int j = 0; j < k; j++) {
    double c = 0.0;
    for (int b = 0; b < parts; b++)
      c += colPart[(size_t)b * kPiv + j];
    colMean[j] = c / (double)n;
    all += colMean[j];
  }
  all /= k;
  double *colMeanDev = g->pcol.p + kParts * kPiv;
  GVH_CHECK(hipMemcpyAsync(colMeanDev, colMean.data(), sizeof(double) * kPiv,
                           hipMemcpyHostToDevice, st));
  const int gparts = (int)std::min<int64_t>(kParts, (n + kGramRows - 1) / kGramRows);
  hipLaunchKernelGGL(k_piv_gram, dim3(gparts), dim3(kB), 0, st, g->pdist.p, n, k, far,
                     g->prow.p, colMeanDev, all, g->pgram.p);
  double *gramDev = g->pgram.p + (int64_t)kParts * kPiv * kPiv;
  hipLaunchKernelGGL(k_piv_gram_sum, dim3(kPiv * kPiv / kB), dim3(kB), 0, st, g->pgram.p,
                     gparts, gramDev);
  GVH_CHECK_LAUNCH();
  std::vector<double> gram(kPiv * kPiv);
  GVH_CHECK(hipMemcpyAsync(gram.data(), gramDev, sizeof(double) * kPiv * kPiv,
                           hipMemcpyDeviceToHost, st));
  GVH_CHECK(hipStreamSynchronize(st));
  std::vector<double> A((size_t)k * k), w, V;
  for (int a = 0; a < k; a++)
    for (int b = 0; b < k; b++)
      A[(size_t)a * k + b] = gram[(size_t)a * kPiv + b];
  jacobi_eigen(A, k, w, V);
  int top[2] = {-1, -1};
  for (int i = 0; i < k; i++) {
    if (top[0] < 0 || w[i] > w[top[0]]) {
      top[1] = top[0];
      top[0] = i;
    } else if (top[1] < 0 || w[i] > w[top[1]]) {
      top[1] = i;
    }
  }
  std::vector<double> axes(2 * kPiv, 0.0);
  for (int a = 0; a < 2; a++) {
    const int c = top[a];
    if (c < 0 || !(w[c] > 0.0))
      continue;
    // sign: the largest-magnitude component is positive
    int big = 0;
    for (int j = 1; j < k; j++)
      if (std::fabs(V[(size_t)j * k + c]) > std::fabs(V[(size_t)big * k + c]))
        big = j;
    const double sgn = V[(size_t)big * k + c] < 0.0 ? -1.0 : 1.0;
    const double sc = sgn / std::pow(w[c], 0.25);
    for (int j = 0; j < k; j++)
      axes[a * kPiv + j] = V[(size_t)j * k + c] * sc;
  }
  GVH_CHECK(hipMemcpyAsync(g->paxes.p, axes.data(), sizeof(double) * 2 * kPiv,
                           hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_piv_project, dim3(grid_for(n, kB)), dim3(kB), 0, st, g->pdist.p, n, k,
                     far, g->prow.p, colMeanDev, all, g->paxes.p, (double2 *)g->praw.p);
  GVH_CHECK_LAUNCH();
  const int sparts = (int)std::min<int64_t>(kParts, grid_for(n, kB));
  hipLaunchKernelGGL(k_piv_stats, dim3(sparts), dim3(kB), 0, st, (const double2 *)g->praw.p,
                     n, g->pcol.p);
  GVH_CHECK_LAUNCH();
  std::vector<double> sp(3 * sparts);
  GVH_CHECK(hipMemcpyAsync(sp.data(), g->pcol.p, sizeof(double) * 3 * sparts,
                           hipMemcpyDeviceToHost, st));
  GVH_CHECK(hipStreamSynchronize(st));
  double sx = 0.0, sy = 0.0, s2 = 0.0;
  for (int b = 0; b < sparts; b++) {
    sx += sp[3 * b];
    sy += sp[3 * b + 1];
    s2 += sp[3 * b + 2];
  }
  const double mx = sx / n, my = sy / n;
  const double var = s2 / n - (mx * mx + my * my);
  const double rms = var > 0.0 ? std::sqrt(var) : 0.0;
  const double inv = rms > 0.0 ? 1.0 / rms : 0.0;
  hipLaunchKernelGGL(k_piv_finish, dim3(grid_for(n, kB)), dim3(kB), 0, st,
                     (const double2 *)g->praw.p, n, mx, my, inv, (float2 *)g->skip.p, g->X0.p);
  GVH_CHECK_LAUNCH();
  if (coords) {
    std::vector<float> c(2 * n);
    GVH_CHECK(hipMemcpyAsync(c.data(), g->skip.p, sizeof(float) * 2 * n,
                             hipMemcpyDeviceToHost, st));
    GVH_CHECK(hipStreamSynchronize(st));
    for (int64_t i = 0; i < 2 * n; i++)
      coords[i] = c[i];
  }
  GVH_CHECK(hipStreamSynchronize(st));
  g->hasSkip = true;
  return 0;
}

int gvizhip_gnn_forward(gvizhip_gnn *g, double *positions) {
  if (check_graph(g) || forward(g))
    return -1;
  if (positions) {
    // outputs 0-1 of every row, packed on the device before the copy
    hipLaunchKernelGGL(k_head_xy, dim3(grid_for(g->n, kB)), dim3(kB), 0, g->stream, g->Y.p,
                       g->n, (float2 *)g->diff[0].p);
    GVH_CHECK_LAUNCH();
    std::vector<float> y(2 * g->n);
    GVH_CHECK(hipMemcpyAsync(y.data(), g->diff[0].p, sizeof(float) * 2 * g->n,
                             hipMemcpyDeviceToHost, g->stream));
    GVH_CHECK(hipStreamSynchronize(g->stream));
    for (int64_t i = 0; i < 2 * g->n; i++)
      positions[i] = y[i];
  }
  GVH_CHECK(hipStreamSynchronize(g->stream));
  return 0;
}

static int upload_target(gvizhip_gnn *g, const double *target) {
  std::vector<float> t(2 * g->n);
  for (int64_t i = 0; i < 2 * g->n; i++)
    t[i] = (float)target[i];
  GVH_CHECK(hipMemcpyAsync(g->target.p, t.data(), sizeof(float) * 2 * g->n,
                           hipMemcpyHostToDevice, g->stream));
  return 0;
}

int gvizhip_gnn_gradients(gvizhip_gnn *g, const double *target, double *loss,
                          float *grads) {
  if (check_graph(g) || !target) {
    set_error("gvizhip_gnn_gradients: invalid arguments");
    return -1;
  }
  LossFit fit;
  if (upload_target(g, target) || forward(g) || fit_loss(g, loss, &fit) ||
      backward(g, fit))
    return -1;
  if (grads) {
    GVH_CHECK(hipMemcpyAsync(grads, g->grads.p, sizeof(float) * kParams,
                             hipMemcpyDeviceToHost, g->stream));
  }
  GVH_CHECK(hipStreamSynchronize(g->stream));
  return 0;
}

int gvizhip_gnn_train_step(gvizhip_gnn *g, const double *target, double lr,
                           double *loss) {
  if (check_graph(g) || !target || !(lr > 0.0)) {
    set_error("gvizhip_gnn_train_step: invalid arguments");
    return -1;
  }
  LossFit fit;
  if (upload_target(g, target) || forward(g) || fit_loss(g, loss, &fit) ||
      backward(g, fit))
    return -1;
  g->step++;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.0f - std::pow(b1, (float)g->step);
  const float c2 = 1.0f - std::pow(b2, (float)g->step);
  hipLaunchKernelGGL(k_adam, dim3(grid_for(kParams, kB)), dim3(kB), 0, g->stream,
                     g->params.p, g->grads.p, g->adamM.p, g->adamV.p, kParams, (float)lr,
                     b1, b2, 1e-8f, c1, c2);
  GVH_CHECK_LAUNCH();
  g->weightsDirty = true;
  GVH_CHECK(hipStreamSynchronize(g->stream));
  return 0;
}

void gvizhip_gnn_destroy(gvizhip_gnn *g) {
  if (!g)
    return;
  if (g->stream) {
    (void)hipStreamSynchronize(g->stream);
    (void)hipStreamDestroy(g->stream);
  }
  if (g->pinned)
    (void)hipHostFree(g->pinned);
  delete g;
}

} // extern "C"
