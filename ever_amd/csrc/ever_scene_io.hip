// Scene I/O on the GPU: primitive arrays from the records of a PLY file (one row of float32 columns per
// primitive), with the viewer's activations for 3D Gaussian Splatting files (ply_file_loader.cc:150-193), and a
// Morton order of the primitives (ply_file_loader.cc:72-148) computed with the library's radix sort.
#include <cfloat>

#include <mutex>

#include "ever_internal.hpp"

namespace ever {
namespace {

// ply_file_loader.cc:150-162, evaluated with the same float/double roundings
__device__ __forceinline__ float sigmoid_f(float x) { return 1.0f / (1.0f + expf(-x)); }
__device__ __forceinline__ float softplus_f(float x) { return x > 5.0f ? x : logf(1.0f + expf(x)); }
__device__ __forceinline__ float inv_opacity_f(float y) {
  return (float)fmax(0.0, -log(fmax((double)(1.0f - y), 1e-10)));
}

__global__ __launch_bounds__(256) void k_scene_from_records(const float* __restrict__ rec, int64_t n,
                                                            ever_ply_layout Lc, const int32_t* __restrict__ order,
                                                            float* __restrict__ means, float* __restrict__ scales,
                                                            float* __restrict__ quats, float* __restrict__ dens,
                                                            float* __restrict__ feats) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int64_t i = order ? (int64_t)order[j] : j;
  const float* r = rec + i * Lc.num_columns;
  float m[3], s[3], q[4];
  for (int a = 0; a < 3; ++a) m[a] = r[Lc.mean[a]];
  for (int a = 0; a < 3; ++a) s[a] = r[Lc.scale[a]];
  for (int a = 0; a < 4; ++a) q[a] = r[Lc.rot[a]];
  float d;
  if (Lc.activation == EVER_PLY_ACTIVATE_3DGS) {
    // rotations divided by their squared length, as the viewer does; scales softplus + 3e-4 capped at 25;
    // density = -log(1 - 0.99 sigmoid(opacity)) / min(2 s)
    float len = 0.f;
    for (int a = 0; a < 4; ++a) len += q[a] * q[a];
    for (int a = 0; a < 4; ++a) q[a] /= len;
    float nrm = 9999.f;
    for (int a = 0; a < 3; ++a) {
      s[a] = (float)fmin((double)softplus_f(s[a]) + 3e-4, 25.0);
      nrm = fminf(2.f * s[a], nrm);
    }
    const float alpha = (float)(0.99 * (double)sigmoid_f(r[Lc.opacity]));
    d = inv_opacity_f(alpha) / nrm;
  } else {
    d = r[Lc.density];
  }
  for (int a = 0; a < 3; ++a) {
    means[3 * j + a] = m[a];
    scales[3 * j + a] = s[a];
  }
  for (int a = 0; a < 4; ++a) quats[4 * j + a] = q[a];
  dens[j] = d;
  // features (d, 3) interleaved: [0][c] = f_dc_c, [1 + k][c] = f_rest[c * R + k] (channel-major in the file)
  const int R = Lc.num_rest;
  float* f = feats + j * (int64_t)(1 + R) * 3;
  for (int c = 0; c < 3; ++c) f[c] = r[Lc.f_dc[c]];
  for (int k = 0; k < R; ++k)
    for (int c = 0; c < 3; ++c) f[3 * (1 + k) + c] = r[Lc.f_rest0 + c * R + k];
}

// bounds of the means: per-block partials, then one block. Minimum starts at FLT_MAX and maximum at FLT_MIN (the
// smallest positive float), as the viewer's loader starts them.
__global__ __launch_bounds__(256) void k_mean_bounds_partial(const float* __restrict__ means, int64_t n,
                                                             float* __restrict__ part) {
  __shared__ float red[6][256];
  float v[6] = {FLT_MAX, FLT_MAX, FLT_MAX, FLT_MIN, FLT_MIN, FLT_MIN};
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    for (int a = 0; a < 3; ++a) {
      const float x = means[3 * i + a];
      v[a] = fminf(v[a], x);
      v[3 + a] = fmaxf(v[3 + a], x);
    }
  }
  for (int a = 0; a < 6; ++a) red[a][threadIdx.x] = v[a];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int a = 0; a < 6; ++a)
        red[a][threadIdx.x] = a < 3 ? fminf(red[a][threadIdx.x], red[a][threadIdx.x + w])
                                    : fmaxf(red[a][threadIdx.x], red[a][threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x < 6) part[6 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

__global__ __launch_bounds__(256) void k_mean_bounds_final(const float* __restrict__ part, int nparts,
                                                           float* __restrict__ box) {
  __shared__ float red[6][256];
  float v[6] = {FLT_MAX, FLT_MAX, FLT_MAX, FLT_MIN, FLT_MIN, FLT_MIN};
  for (int p = threadIdx.x; p < nparts; p += blockDim.x)
    for (int a = 0; a < 6; ++a) v[a] = a < 3 ? fminf(v[a], part[6 * p + a]) : fmaxf(v[a], part[6 * p + a]);
  for (int a = 0; a < 6; ++a) red[a][threadIdx.x] = v[a];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int a = 0; a < 6; ++a)
        red[a][threadIdx.x] = a < 3 ? fminf(red[a][threadIdx.x], red[a][threadIdx.x + w])
                                    : fmaxf(red[a][threadIdx.x], red[a][threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x < 6) box[threadIdx.x] = red[threadIdx.x][0];
}

// 63-bit Morton code: 21 bits per axis of (m - lo) / (hi - lo) * (2^21 - 1), truncated; bit b of x, y, z at
// 3b, 3b + 1, 3b + 2. Split into the low 32 bits (keys) and the high 31 bits (hi).
__global__ __launch_bounds__(256) void k_morton63(const float* __restrict__ means, int64_t n,
                                                  const float* __restrict__ box, uint32_t* __restrict__ lo32,
                                                  uint32_t* __restrict__ hi31, uint32_t* __restrict__ idx) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint64_t code = 0;
  for (int a = 0; a < 3; ++a) {
    const float rel = (means[3 * i + a] - box[a]) / (box[3 + a] - box[a]);
    const float sc = 2097151.0f * rel;
    // a zero extent gives NaN: no bits (the truncation of NaN to int has none in its low 21 bits either)
    const uint32_t u = (sc == sc) ? (uint32_t)(int32_t)sc : 0u;
    for (int b = 0; b < 21; ++b) code |= (uint64_t)((u >> b) & 1u) << (3 * b + a);
  }
  lo32[i] = (uint32_t)code;
  hi31[i] = (uint32_t)(code >> 32);
  idx[i] = (uint32_t)i;
}

__global__ __launch_bounds__(256) void k_gather_u32(const uint32_t* __restrict__ src, const uint32_t* __restrict__ idx,
                                                    int64_t n, uint32_t* __restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[idx[i]];
}


}  // namespace

hipError_t launch_scene_from_records(const float* records, int64_t n, const ever_ply_layout& Lc, const int32_t* order,
                                     float* means, float* scales, float* quats, float* densities, float* features,
                                     hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_scene_from_records, dim3(grid_for(n, 256)), dim3(256), 0, st, records, n, Lc, order, means,
                     scales, quats, densities, features);
  return hipGetLastError();
}

hipError_t launch_morton_order(const float* means, int64_t n, int32_t* order, void* ws, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  const MortonLayout L = morton_layout(n);
  float* box = at<float>(ws, L.box);
  float* part = at<float>(ws, L.part);
  uint32_t* k0 = at<uint32_t>(ws, L.keys0);
  uint32_t* v0 = at<uint32_t>(ws, L.vals0);
  uint32_t* k1 = at<uint32_t>(ws, L.keys1);
  uint32_t* v1 = at<uint32_t>(ws, L.vals1);
  uint32_t* hi = at<uint32_t>(ws, L.hi);
  uint32_t* hist = at<uint32_t>(ws, L.hist);
  const int nb = (int)std::min<int64_t>(kBoundsBlocks, grid_for(n, 256));
  hipLaunchKernelGGL(k_mean_bounds_partial, dim3(nb), dim3(256), 0, st, means, n, part);
  hipLaunchKernelGGL(k_mean_bounds_final, dim3(1), dim3(256), 0, st, part, nb, box);
  hipLaunchKernelGGL(k_morton63, dim3(grid_for(n, 256)), dim3(256), 0, st, means, n, box, k0, hi, v0);
  // LSD over the 63 bits: stable sort by the low word, then stable sort by the high word
  hipError_t e = launch_sort_pairs(k0, v0, k1, v1, hist, n, 32, st);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_gather_u32, dim3(grid_for(n, 256)), dim3(256), 0, st, hi, v0, n, k0);
  e = launch_sort_pairs(k0, v0, k1, v1, hist, n, 31, st);
  if (e != hipSuccess) return e;
  return hipMemcpyAsync(order, v0, 4 * (size_t)n, hipMemcpyDeviceToDevice, st);
}

// Scratch of ever_morton_order_noalloc: one array per GPU the library runs on, held until the process ends. Every
// region is written before it is read (bounds, keys, values, histograms), so nothing needs clearing between calls;
// calls that overlap on one GPU share it and race.
constexpr size_t kMortonScratchBytes = morton_layout(kMortonNoallocMaxPrims).total;
static __device__ __align__(256) unsigned char scratch[kMortonScratchBytes];

// The array's device address on each GPU, looked up (hipGetSymbolAddress) on the first call with that GPU current.
// In host code `scratch` names a host-side stand-in, never the device memory: kernels get this address instead.
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[64];

hipError_t launch_morton_order_noalloc(const float* means, int64_t n, int32_t* order, hipStream_t st) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= 64) return hipErrorInvalidDevice;
  void* ptr = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
      e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
      if (e != hipSuccess) return e;
      g_scratch_ptr[dev] = ptr;
    }
    ptr = g_scratch_ptr[dev];
  }
  return launch_morton_order(means, n, order, ptr, st);
}

}  // namespace ever
