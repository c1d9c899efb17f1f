// Termination-depth proposal network: a 63-dim positional encoding of the ray -> 3 x 128 ReLU -> 1 output
// (log of the proposed depth), BF16 weights and activations on the MFMA units with FP32 accumulation.
// No counterpart in the reference; its output bounds the first traversal of a ray (ever_rays.proposal).
//
// Model (the CPU check in the tests restates it): x = bf16(pe(ray)), h_l = bf16(relu(h_{l-1} . bf16(W_l) + b_l))
// for l = 1..3 (FP32 accumulation), y = h_3 . w_4 + b_4 in FP32. Training minimises the pinball loss of y against
// log(target depth) with Adam on FP32 master weights; the gradients that feed an MFMA product are rounded to
// bf16 at the points the CPU check rounds them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <utility>

#include "ever_common.hpp"
#include "ever_internal.hpp"

namespace ever {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// gfx950's v_mfma_f32_16x16x32_bf16: D(16x16) += A(16x32) B(32x16), FP32 accumulation. Lane l supplies row l % 16 of
// A and column l % 16 of B, each with the eight k of its lane group (k = 8 * (l / 16) + 0..7 here: the product pairs
// A's and B's elements by lane group and slot, so any k order both operands share gives the same sum), and receives
// D[4 * (l / 16) + q][l % 16], q = 0..3.
__device__ __forceinline__ f32x4 mfma16x16x32(const uint16_t* a, const uint16_t* b, f32x4 c) {
  const uint4 av = *reinterpret_cast<const uint4*>(a), bv = *reinterpret_cast<const uint4*>(b);
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, av), __builtin_bit_cast(bf16x8, bv), c,
                                                 0, 0, 0);
}

__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// ---------------------------------------------------------------------------------------------------------------
// C[m][n] = sum_k A(m,k) B(k,n), A(m,k) = A[m*sam + k*sak], B(k,n) = B[k*sbk + n*sbn] (bf16), FP32 accumulation on
// v_mfma_f32_16x16x32_bf16. Block tile 64 x 64 (4 waves, 32 x 32 each = 2 x 2 MFMA tiles), K in steps of 32 (a chunk's
// last step zero-padded past its end). blockIdx.z selects a K chunk [z*kc, min(K,(z+1)*kc)); with more than one chunk
// the block writes FP32 partials part[z][m][n], else the epilogue: + bias[n], ReLU, FP32 (c32 with strides) and/or
// bf16 (c16 with strides). Requires M % 64 == 0, N % 64 == 0, kc % 16 == 0, K % 16 == 0 (checked by the host
// launcher).
struct GemmArgs {
  const uint16_t* A;
  int64_t sam, sak;
  const uint16_t* B;
  int64_t sbk, sbn;
  int64_t M, N, K, kc;
  const float* bias;
  int relu;
  float* c32;
  int64_t scm, scn;
  uint16_t* c16;
  int64_t s16m, s16n;
  float* part;
};

__global__ __launch_bounds__(256) void k_gemm_bf16(GemmArgs g) {
  __shared__ __align__(16) uint16_t As[64][32 + 8];  // [m][k], padded rows (16-byte aligned)
  __shared__ __align__(16) uint16_t Bs[64][32 + 8];  // [n][k]
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t m0 = (int64_t)blockIdx.x * 64, n0 = (int64_t)blockIdx.y * 64;
  const int64_t kb = (int64_t)blockIdx.z * g.kc, ke = min(g.K, kb + g.kc);
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  f32x4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int64_t k0 = kb; k0 < ke; k0 += 32) {
    // 64 x 32 elements of A and of B: 8 each per thread (zero past the chunk's end)
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int idx = tid + 256 * e, r = idx >> 5, c = idx & 31;
      const bool in = k0 + c < ke;
      As[r][c] = in ? g.A[(m0 + r) * g.sam + (k0 + c) * g.sak] : (uint16_t)0;
      Bs[r][c] = in ? g.B[(k0 + c) * g.sbk + (n0 + r) * g.sbn] : (uint16_t)0;
    }
    __syncthreads();
    const int row = lane & 15, k8 = (lane >> 4) * 8;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[i][j] = mfma16x16x32(&As[wm + 16 * i + row][k8], &Bs[wn + 16 * j + row][k8], acc[i][j]);
    __syncthreads();
  }
  // lane holds D[4*(lane/16) + q][lane % 16] of each 16 x 16 tile
  const int col = lane & 15, rq = (lane >> 4) * 4;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int64_t m = m0 + wm + 16 * i + rq + q, n = n0 + wn + 16 * j + col;
        float v = acc[i][j][q];
        if (gridDim.z > 1) {
          g.part[((int64_t)blockIdx.z * g.M + m) * g.N + n] = v;
          continue;
        }
        if (g.bias) v += g.bias[n];
        if (g.relu) v = fmaxf(v, 0.f);
        if (g.c32) g.c32[m * g.scm + n * g.scn] = v;
        if (g.c16) g.c16[m * g.s16m + n * g.s16n] = f2bf(v);
      }
}

// sum of the split-K partials + the same epilogue
__global__ void k_gemm_reduce(GemmArgs g, int nz) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.M * g.N) return;
  const int64_t m = i / g.N, n = i % g.N;
  float v = 0.f;
  for (int z = 0; z < nz; ++z) v += g.part[(int64_t)z * g.M * g.N + i];
  if (g.bias) v += g.bias[n];
  if (g.relu) v = fmaxf(v, 0.f);
  if (g.c32) g.c32[m * g.scm + n * g.scn] = v;
  if (g.c16) g.c16[m * g.s16m + n * g.s16n] = f2bf(v);
}

hipError_t launch_gemm(GemmArgs g, int nz, hipStream_t st) {
  if (g.M % 64 || g.N % 64 || g.K % 16 || nz < 1) return hipErrorInvalidValue;
  g.kc = ((g.K / 16 + nz - 1) / nz) * 16;
  nz = (int)((g.K + g.kc - 1) / g.kc);
  dim3 grid((unsigned)(g.M / 64), (unsigned)(g.N / 64), (unsigned)nz);
  hipLaunchKernelGGL(k_gemm_bf16, grid, dim3(256), 0, st, g);
  if (nz > 1) hipLaunchKernelGGL(k_gemm_reduce, dim3(grid_for(g.M * g.N, 256)), dim3(256), 0, st, g, nz);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------------------------------------------
// Ray encoding (63 features + one zero column): o/4 (3); p/4 and sin/cos(2^j pi p/4), j = 0..4 (33), p the point of
// the ray closest to the scene origin; d and sin/cos(2^j pi d), j = 0..3 (27). o = origin + tmin * direction,
// d = unit direction.
__device__ __forceinline__ void encode_ray(f3 o, f3 d, float* x) {
  const float pd = dot3(o, d);
  const f3 p = o - pd * d;
  int c = 0;
  x[c++] = 0.25f * o.x; x[c++] = 0.25f * o.y; x[c++] = 0.25f * o.z;
  const float pv[3] = {0.25f * p.x, 0.25f * p.y, 0.25f * p.z};
  const float dv[3] = {d.x, d.y, d.z};
  for (int a = 0; a < 3; ++a) x[c++] = pv[a];
  for (int j = 0; j < 5; ++j)
    for (int a = 0; a < 3; ++a) {
      float s, cs;
      sincospif(ldexpf(pv[a], j), &s, &cs);
      x[c++] = s;
      x[c++] = cs;
    }
  for (int a = 0; a < 3; ++a) x[c++] = dv[a];
  for (int j = 0; j < 4; ++j)
    for (int a = 0; a < 3; ++a) {
      float s, cs;
      sincospif(ldexpf(dv[a], j), &s, &cs);
      x[c++] = s;
      x[c++] = cs;
    }
  x[c] = 0.f;  // c == 63
}

// The features of encode_ray at the same indices, split over four lanes per ray by `part` (0..3): part 0 the plain
// features (o, p, d) and sin/cos of p at frequencies 1, 2; part 1 p at 4, 8; part 2 p at 16 and d at 1; part 3 d at
// 2, 4, 8 (nine sin/cos pairs at most per lane) and the zero column. Written as bf16 into row x.
__device__ __forceinline__ void encode_ray_part(f3 o, f3 d, int part, uint16_t* x) {
  const float pd = dot3(o, d);
  const f3 p = o - pd * d;
  const float pv[3] = {0.25f * p.x, 0.25f * p.y, 0.25f * p.z};
  const float dv[3] = {d.x, d.y, d.z};
  auto pair = [&](const float* v, int base, int j) {  // sin/cos of 2^j pi v[a] at base + 2 * (3 j + a) + 0/1
    for (int a = 0; a < 3; ++a) {
      float sn, cs;
      sincospif(ldexpf(v[a], j), &sn, &cs);
      x[base + 2 * (3 * j + a)] = f2bf(sn);
      x[base + 2 * (3 * j + a) + 1] = f2bf(cs);
    }
  };
  if (part == 0) {
    x[0] = f2bf(0.25f * o.x); x[1] = f2bf(0.25f * o.y); x[2] = f2bf(0.25f * o.z);
    for (int a = 0; a < 3; ++a) x[3 + a] = f2bf(pv[a]);
    for (int a = 0; a < 3; ++a) x[36 + a] = f2bf(dv[a]);
    pair(pv, 6, 0);
    pair(pv, 6, 1);
  } else if (part == 1) {
    pair(pv, 6, 2);
    pair(pv, 6, 3);
  } else if (part == 2) {
    pair(pv, 6, 4);
    pair(dv, 39, 0);
  } else {
    pair(dv, 39, 1);
    pair(dv, 39, 2);
    pair(dv, 39, 3);
    x[63] = 0;
  }
}

__global__ void k_encode(const float* __restrict__ rayo, const float* __restrict__ rayd, int64_t nr, int64_t mpad,
                         float tmin, uint16_t* __restrict__ X) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= mpad) return;
  float x[kMlpIn];
  if (r < nr) {
    const f3 rd = ld3(rayd + 3 * r);
    encode_ray(ld3(rayo + 3 * r) + tmin * rd, l2n(rd), x);
  } else {
    for (int i = 0; i < kMlpIn; ++i) x[i] = 0.f;
  }
  for (int i = 0; i < kMlpIn; ++i) X[r * kMlpIn + i] = f2bf(x[i]);
}

// y[m] = h3[m] . w4 + b4; optionally the proposal exp(y) * margin
__global__ void k_head(const uint16_t* __restrict__ H3, const float* __restrict__ w4, const float* __restrict__ b4,
                       int64_t n, float* __restrict__ y, float* __restrict__ proposal, float margin) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n) return;
  float s = 0.f;
  for (int k = 0; k < kMlpHid; ++k) s = fmaf(bf2f(H3[m * kMlpHid + k]), w4[k], s);
  s += b4[0];
  if (y) y[m] = s;
  if (proposal) proposal[m] = expf(s) * margin;
}

// pinball loss of y against log(target): dL/dy = -(tau - [u < 0]) / count, u = log t - y; rows with target <= 0
// are ignored. Also dh3 = bf16(dy * w4 masked by h3 > 0) = dz3 and the loss sum.
__global__ void k_loss_head_bwd(const float* __restrict__ y, const float* __restrict__ target, int64_t n,
                                int64_t mpad, float tau, float inv_count, const uint16_t* __restrict__ H3,
                                const float* __restrict__ w4, float* __restrict__ dy, uint16_t* __restrict__ dZ3,
                                float* __restrict__ loss_sum) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= mpad) return;
  float g = 0.f, l = 0.f;
  if (m < n && target[m] > 0.f) {
    const float u = logf(target[m]) - y[m];
    g = -(tau - (u < 0.f ? 1.f : 0.f)) * inv_count;
    l = u * (tau - (u < 0.f ? 1.f : 0.f)) * inv_count;
  }
  dy[m] = g;
  for (int k = 0; k < kMlpHid; ++k) {
    const bool on = bf2f(H3[m * kMlpHid + k]) > 0.f;
    dZ3[m * kMlpHid + k] = f2bf(on ? g * w4[k] : 0.f);
  }
  if (l != 0.f) atomicAdd(loss_sum, l);
}

// dz_l = bf16(dh_l masked by h_l > 0), in place over the FP32 product dh (written as bf16 into dz)
__global__ void k_relu_bwd(const float* __restrict__ dH, const uint16_t* __restrict__ H, int64_t count,
                           uint16_t* __restrict__ dZ) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  dZ[i] = f2bf(bf2f(H[i]) > 0.f ? dH[i] : 0.f);
}

// column sums over rows: out[n] = sum_m in[m*cols + n] (bf16 or FP32 input), one block per column chunk
__global__ void k_colsum_bf16(const uint16_t* __restrict__ in, int64_t rows, int cols, float* __restrict__ out) {
  const int n = blockIdx.x;
  float s = 0.f;
  for (int64_t m = threadIdx.x; m < rows; m += blockDim.x) s += bf2f(in[m * cols + n]);
  __shared__ float red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[n] = red[0];
}

// dw4[k] = sum_m dy[m] h3[m][k]; db4 = sum_m dy[m]
__global__ void k_head_wgrad(const float* __restrict__ dy, const uint16_t* __restrict__ H3, int64_t rows,
                             float* __restrict__ dw4, float* __restrict__ db4) {
  const int k = blockIdx.x;  // 0..kMlpHid (the last block sums dy)
  float s = 0.f;
  for (int64_t m = threadIdx.x; m < rows; m += blockDim.x)
    s += k < kMlpHid ? dy[m] * bf2f(H3[m * kMlpHid + k]) : dy[m];
  __shared__ float red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    if (k < kMlpHid) dw4[k] = red[0];
    else db4[0] = red[0];
  }
}

// Adam on the FP32 master parameters; refreshes the bf16 copies of the three weight matrices
__global__ void k_adam(float* __restrict__ p, float* __restrict__ mom, float* __restrict__ vel,
                       const float* __restrict__ grad, int64_t count, float lr, float b1, float b2, float eps,
                       float c1, float c2, uint16_t* __restrict__ pb) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const float gi = grad[i];
  const float m = b1 * mom[i] + (1.f - b1) * gi;
  const float v = b2 * vel[i] + (1.f - b2) * gi * gi;
  mom[i] = m;
  vel[i] = v;
  const float w = p[i] - lr * (m / c1) / (sqrtf(v / c2) + eps);
  p[i] = w;
  pb[i] = f2bf(w);
}

__global__ void k_to_bf16(const float* __restrict__ p, int64_t count, uint16_t* __restrict__ pb) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) pb[i] = f2bf(p[i]);
}

__global__ void k_fill_f32(float* p, int64_t n, float v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------------------------------------------------
// Parameter block (FP32): W1 [64][128] (row 63 unused), b1 [128], W2 [128][128], b2, W3 [128][128], b3, w4 [128],
// b4 [1]; the bf16 block mirrors the same offsets.
MlpLayout mlp_layout() {
  MlpLayout L;
  int64_t o = 0;
  L.w1 = o; o += (int64_t)kMlpIn * kMlpHid;
  L.b1 = o; o += kMlpHid;
  L.w2 = o; o += (int64_t)kMlpHid * kMlpHid;
  L.b2 = o; o += kMlpHid;
  L.w3 = o; o += (int64_t)kMlpHid * kMlpHid;
  L.b3 = o; o += kMlpHid;
  L.w4 = o; o += kMlpHid;
  L.b4 = o; o += 1;
  L.count = o;
  return L;
}

// Activation workspace for m (padded) rows: X [m][64], H1..H3 [m][128] bf16, dH FP32 [m][128], dZ bf16 [m][128],
// dy FP32 [m], split-K partials FP32, gradient block FP32 [count], loss FP32.
MlpWork mlp_work(int64_t m) {
  MlpWork W;
  const MlpLayout L = mlp_layout();
  int64_t o = 0;
  auto take = [&](int64_t bytes) {
    const int64_t at = o;
    o = (o + bytes + 255) & ~(int64_t)255;
    return at;
  };
  W.x = take(m * kMlpIn * 2);
  W.h1 = take(m * kMlpHid * 2);
  W.h2 = take(m * kMlpHid * 2);
  W.h3 = take(m * kMlpHid * 2);
  W.dh = take(m * kMlpHid * 4);
  W.dz = take(m * kMlpHid * 2);
  W.dz2 = take(m * kMlpHid * 2);
  W.y = take(m * 4);
  W.dy = take(m * 4);
  W.part = take((int64_t)kMlpSplit * kMlpHid * kMlpHid * 4);
  W.grad = take(L.count * 4);
  W.loss = take(4);
  W.total = o;
  return W;
}

static GemmArgs gemm(const uint16_t* A, int64_t sam, int64_t sak, const uint16_t* B, int64_t sbk, int64_t sbn,
                     int64_t M, int64_t N, int64_t K) {
  GemmArgs g{};
  g.A = A; g.sam = sam; g.sak = sak;
  g.B = B; g.sbk = sbk; g.sbn = sbn;
  g.M = M; g.N = N; g.K = K;
  return g;
}

// forward through the three hidden layers for m rows (X already encoded); H1..H3 written
static hipError_t mlp_hidden(const float* p, const uint16_t* pb, char* ws, const MlpWork& W, int64_t m,
                             hipStream_t st) {
  const MlpLayout L = mlp_layout();
  const uint16_t* X = reinterpret_cast<const uint16_t*>(ws + W.x);
  uint16_t* H[3] = {reinterpret_cast<uint16_t*>(ws + W.h1), reinterpret_cast<uint16_t*>(ws + W.h2),
                    reinterpret_cast<uint16_t*>(ws + W.h3)};
  const int64_t wo[3] = {L.w1, L.w2, L.w3}, bo[3] = {L.b1, L.b2, L.b3};
  const uint16_t* in = X;
  int64_t kin = kMlpIn;
  for (int l = 0; l < 3; ++l) {
    GemmArgs g = gemm(in, kin, 1, pb + wo[l], kMlpHid, 1, m, kMlpHid, kin);
    g.bias = p + bo[l];
    g.relu = 1;
    g.c16 = H[l];
    g.s16m = kMlpHid;
    g.s16n = 1;
    hipError_t e = launch_gemm(g, 1, st);
    if (e != hipSuccess) return e;
    in = H[l];
    kin = kMlpHid;
  }
  return hipSuccess;
}

// ---------------------------------------------------------------------------------------------------------------
// Fused inference: a persistent kernel whose blocks copy the three weight matrices into LDS once (transposed,
// [n][k] with padded rows) and then run 16-row tiles through the three layers and the head, one tile per wave at a
// time; activations stay in LDS. Same arithmetic as the layer-by-layer path (bf16 operands, FP32 accumulation,
// bias + ReLU, bf16 activations, FP32 head).
constexpr int kFuseWaves = 8;
// row pitches in bf16 elements: multiples of 8 (16-byte operand loads), offset by 8 from the row length (bank spread)
constexpr int kPadIn = kMlpIn + 8, kPadHid = kMlpHid + 8;
struct __align__(16) FusedLds {
  uint16_t w1[kMlpHid][kPadIn];
  uint16_t w2[kMlpHid][kPadHid];
  uint16_t w3[kMlpHid][kPadHid];
  float bias[3][kMlpHid];
  float w4[kMlpHid];
  uint16_t act[kFuseWaves][16][kPadHid];
};

// one layer for the wave's 16 rows: act[r][0..K) -> act[r][0..128), in place (read fully before written)
template <int KIN, int PITCH>
__device__ __forceinline__ void fused_layer(const uint16_t (*W)[PITCH], const float* bias, uint16_t (*act)[kPadHid],
                                            int lane) {
  f32x4 acc[8];
#pragma unroll
  for (int t = 0; t < 8; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int row = lane & 15, k8 = (lane >> 4) * 8;
#pragma unroll
  for (int k0 = 0; k0 < KIN; k0 += 32) {
#pragma unroll
    for (int t = 0; t < 8; ++t) acc[t] = mfma16x16x32(&act[row][k0 + k8], &W[t * 16 + row][k0 + k8], acc[t]);
  }
  __builtin_amdgcn_wave_barrier();  // every lane has read its operands before any lane overwrites act
  const int col = lane & 15, rq = (lane >> 4) * 4;
#pragma unroll
  for (int t = 0; t < 8; ++t)
#pragma unroll
    for (int q = 0; q < 4; ++q) act[rq + q][t * 16 + col] = f2bf(fmaxf(acc[t][q] + bias[t * 16 + col], 0.f));
  __builtin_amdgcn_wave_barrier();
}

__global__ __launch_bounds__(64 * kFuseWaves) void k_mlp_fused_infer(const float* __restrict__ p,
                                                                      const uint16_t* __restrict__ pb,
                                                                      const float* __restrict__ rayo,
                                                                      const float* __restrict__ rayd, int64_t nr,
                                                                      float tmin, float margin, float* __restrict__ y,
                                                                      float* __restrict__ proposal, MlpLayout L) {
  __shared__ FusedLds S;  // 118 KB of the 160 KB LDS of a gfx950 CU
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int i = tid; i < kMlpHid * kMlpIn; i += blockDim.x) {  // W1 [k][n] -> w1[n][k]
    const int k = i / kMlpHid, n = i % kMlpHid;
    S.w1[n][k] = pb[L.w1 + i];
  }
  for (int i = tid; i < kMlpHid * kMlpHid; i += blockDim.x) {
    const int k = i / kMlpHid, n = i % kMlpHid;
    S.w2[n][k] = pb[L.w2 + i];
    S.w3[n][k] = pb[L.w3 + i];
  }
  for (int i = tid; i < kMlpHid; i += blockDim.x) {
    S.bias[0][i] = p[L.b1 + i];
    S.bias[1][i] = p[L.b2 + i];
    S.bias[2][i] = p[L.b3 + i];
    S.w4[i] = p[L.w4 + i];
  }
  const float b4 = p[L.b4];
  __syncthreads();
  uint16_t (*act)[kPadHid] = S.act[wave];
  const int64_t tiles = (nr + 15) / 16;
  const int64_t stride = (int64_t)gridDim.x * kFuseWaves;
  // the ray of this lane in a tile (lane % 16), loaded one tile ahead so that the loads overlap the previous tile
  auto load_ray = [&](int64_t tile, f3& o, f3& d) {
    const int64_t r = tile * 16 + (lane & 15);
    if (tile < tiles && r < nr) {
      o = ld3(rayo + 3 * r);
      d = ld3(rayd + 3 * r);
    } else {
      o = mk3(0.f, 0.f, 0.f);
      d = mk3(0.f, 0.f, 0.f);
    }
  };
  f3 o_next, d_next;
  load_ray((int64_t)blockIdx.x * kFuseWaves + wave, o_next, d_next);
  for (int64_t tile = (int64_t)blockIdx.x * kFuseWaves + wave; tile < tiles; tile += stride) {
    const f3 ro = o_next, rd = d_next;
    load_ray(tile + stride, o_next, d_next);
    // encoding: four lanes per ray (lane % 16 the ray, lane / 16 a quarter of the features)
    {
      const int64_t r = tile * 16 + (lane & 15);
      if (r < nr) {
        encode_ray_part(ro + tmin * rd, l2n(rd), lane >> 4, act[lane & 15]);
      } else {
        for (int i = lane >> 4; i < kMlpIn; i += 4) act[lane & 15][i] = 0;
      }
    }
    __builtin_amdgcn_wave_barrier();
    fused_layer<kMlpIn, kPadIn>(S.w1, S.bias[0], act, lane);
    fused_layer<kMlpHid, kPadHid>(S.w2, S.bias[1], act, lane);
    fused_layer<kMlpHid, kPadHid>(S.w3, S.bias[2], act, lane);
    // head: lane -> row lane % 16, quarter lane / 16 of the 128 inputs; sum the quarters across the wave
    const int row = lane & 15, qtr = lane >> 4;
    float s = 0.f;
    const uint4* hv = reinterpret_cast<const uint4*>(&act[row][qtr * 32]);
    const float4* wv = reinterpret_cast<const float4*>(&S.w4[qtr * 32]);
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const uint4 h = hv[c];
      const uint32_t hw[4] = {h.x, h.y, h.z, h.w};
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const float4 w = wv[2 * c + j / 2];
        const float w0 = (j & 1) ? w.z : w.x, w1 = (j & 1) ? w.w : w.y;
        s = fmaf(__uint_as_float(hw[j] << 16), w0, s);
        s = fmaf(__uint_as_float(hw[j] & 0xffff0000u), w1, s);
      }
    }
    s += __shfl_xor(s, 16);
    s += __shfl_xor(s, 32);
    const int64_t r = tile * 16 + row;
    if (qtr == 0 && r < nr) {
      const float v = s + b4;
      if (y) y[r] = v;
      if (proposal) proposal[r] = expf(v) * margin;
    }
    __builtin_amdgcn_wave_barrier();
  }
}

hipError_t launch_mlp_infer(const float* p, const uint16_t* pb, const float* rayo, const float* rayd, int64_t nr,
                            float tmin, float margin, float* y, float* proposal, void* ws, hipStream_t st) {
  (void)ws;
  int dev = 0, cus = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e != hipSuccess) return e;
  const int64_t tiles = (nr + 15) / 16;
  const int64_t blocks = std::min<int64_t>((int64_t)cus, (tiles + kFuseWaves - 1) / kFuseWaves);
  hipLaunchKernelGGL(k_mlp_fused_infer, dim3((unsigned)std::max<int64_t>(blocks, 1)), dim3(64 * kFuseWaves),
                     0, st, p, pb, rayo, rayd, nr, tmin, margin, y, proposal, mlp_layout());
  return hipGetLastError();
}

hipError_t launch_mlp_train_step(float* p, uint16_t* pb, float* mom, float* vel, const float* rayo,
                                 const float* rayd, const float* target, int64_t nr, float tmin, float tau, float lr,
                                 int step, void* ws, float* loss_out, float* grad_out, hipStream_t st) {
  const int64_t m = (nr + 63) / 64 * 64;
  const MlpWork W = mlp_work(m);
  const MlpLayout L = mlp_layout();
  char* w = reinterpret_cast<char*>(ws);
  uint16_t* X = reinterpret_cast<uint16_t*>(w + W.x);
  uint16_t* H[3] = {reinterpret_cast<uint16_t*>(w + W.h1), reinterpret_cast<uint16_t*>(w + W.h2),
                    reinterpret_cast<uint16_t*>(w + W.h3)};
  float* dH = reinterpret_cast<float*>(w + W.dh);
  uint16_t* dZ = reinterpret_cast<uint16_t*>(w + W.dz);
  uint16_t* dZ2 = reinterpret_cast<uint16_t*>(w + W.dz2);
  float* y = reinterpret_cast<float*>(w + W.y);
  float* dy = reinterpret_cast<float*>(w + W.dy);
  float* grad = grad_out ? grad_out : reinterpret_cast<float*>(w + W.grad);
  float* loss = reinterpret_cast<float*>(w + W.loss);
  float* part = reinterpret_cast<float*>(w + W.part);

  hipLaunchKernelGGL(k_encode, dim3(grid_for(m, 256)), dim3(256), 0, st, rayo, rayd, nr, m, tmin, X);
  hipError_t e = mlp_hidden(p, pb, w, W, m, st);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_head, dim3(grid_for(m, 256)), dim3(256), 0, st, H[2], p + L.w4, p + L.b4, m, y,
                     (float*)nullptr, 1.f);
  int64_t valid = nr;
  hipLaunchKernelGGL(k_fill_f32, dim3(1), dim3(1), 0, st, loss, (int64_t)1, 0.f);
  hipLaunchKernelGGL(k_loss_head_bwd, dim3(grid_for(m, 256)), dim3(256), 0, st, y, target, nr, m, tau,
                     1.f / (float)(valid > 0 ? valid : 1), H[2], p + L.w4, dy, dZ, loss);
  hipLaunchKernelGGL(k_head_wgrad, dim3(kMlpHid + 1), dim3(256), 0, st, dy, H[2], m, grad + L.w4, grad + L.b4);
  // layers 3, 2, 1: dW_l = h_{l-1}^T dz_l (split over rows), db_l = colsum(dz_l), dh_{l-1} = dz_l W_l^T
  const int64_t wo[3] = {L.w1, L.w2, L.w3}, bo[3] = {L.b1, L.b2, L.b3};
  const uint16_t* ins[3] = {X, H[0], H[1]};
  const int64_t kin[3] = {kMlpIn, kMlpHid, kMlpHid};
  uint16_t* dz_cur = dZ;
  uint16_t* dz_next = dZ2;
  for (int l = 2; l >= 0; --l) {
    {
      GemmArgs g = gemm(ins[l], 1, kin[l], dz_cur, kMlpHid, 1, kin[l], kMlpHid, m);
      g.part = part;
      g.c32 = grad + wo[l];
      g.scm = kMlpHid;
      g.scn = 1;
      const int nz = (int)std::min<int64_t>(kMlpSplit, m / 16);
      e = launch_gemm(g, nz, st);
      if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_colsum_bf16, dim3(kMlpHid), dim3(256), 0, st, dz_cur, m, kMlpHid, grad + bo[l]);
    if (l == 0) break;
    {
      GemmArgs g = gemm(dz_cur, kMlpHid, 1, pb + wo[l], 1, kMlpHid, m, kMlpHid, kMlpHid);
      g.c32 = dH;
      g.scm = kMlpHid;
      g.scn = 1;
      e = launch_gemm(g, 1, st);
      if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_relu_bwd, dim3(grid_for(m * kMlpHid, 256)), dim3(256), 0, st, dH, H[l - 1],
                       m * kMlpHid, dz_next);
    std::swap(dz_cur, dz_next);
  }
  if (lr > 0.f) {
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
    hipLaunchKernelGGL(k_adam, dim3(grid_for(L.count, 256)), dim3(256), 0, st, p, mom, vel, grad, L.count, lr, b1,
                       b2, 1e-8f, c1, c2, pb);
  }
  if (loss_out) {
    e = hipMemcpyAsync(loss_out, loss, sizeof(float), hipMemcpyDeviceToDevice, st);
    if (e != hipSuccess) return e;
  }
  return hipGetLastError();
}

hipError_t launch_mlp_sync_bf16(const float* p, uint16_t* pb, hipStream_t st) {
  const MlpLayout L = mlp_layout();
  hipLaunchKernelGGL(k_to_bf16, dim3(grid_for(L.count, 256)), dim3(256), 0, st, p, L.count, pb);
  return hipGetLastError();
}

hipError_t launch_gemm_public(const uint16_t* A, int64_t sam, int64_t sak, const uint16_t* B, int64_t sbk,
                              int64_t sbn, int64_t M, int64_t N, int64_t K, const float* bias, int relu, float* C,
                              int splits, float* part, hipStream_t st) {
  GemmArgs g = gemm(A, sam, sak, B, sbk, sbn, M, N, K);
  g.bias = bias;
  g.relu = relu;
  g.c32 = C;
  g.scm = N;
  g.scn = 1;
  g.part = part;
  return launch_gemm(g, splits, st);
}

}  // namespace ever
