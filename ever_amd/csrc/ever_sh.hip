// Standalone per-primitive SH colour and its backward, one lane per primitive, for callers that keep the
// reference's two-step form eval_sh -> RGB2SH -> trace_rays (eval_sh.py:28-82; splinetracers/slang/sh_kernel.slang).
#include "ever_common.hpp"
#include "ever_internal.hpp"

namespace ever {

__global__ __launch_bounds__(256) void k_eval_sh(const float* __restrict__ means, const float* __restrict__ feat,
                                                 const float* __restrict__ origin, int64_t n, int d, int deg,
                                                 float* __restrict__ out) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const f3 dir = sh_dir(ld3(means + 3 * p), ld3(origin));
  float b[16];
  sh_basis(deg, dir.x, dir.y, dir.z, b);
  const int nc = (deg + 1) * (deg + 1);
  const float* f = feat + (int64_t)3 * d * p;
  float c0 = 0.5f, c1 = 0.5f, c2 = 0.5f;
  for (int k = 0; k < nc; ++k) {
    c0 += b[k] * f[3 * k + 0];
    c1 += b[k] * f[3 * k + 1];
    c2 += b[k] * f[3 * k + 2];
  }
  out[3 * p + 0] = c0;
  out[3 * p + 1] = c1;
  out[3 * p + 2] = c2;
}

// bw_sh_kernel (sh_kernel.slang:45-98): dL/df_k = Y_k(dir) * dL/drgb; primitives whose colour gradient is
// below 1e-8 in every channel get zero. No gradient flows to the means through the direction.
__global__ __launch_bounds__(256) void k_eval_sh_bwd(const float* __restrict__ means, const float* __restrict__ origin,
                                                     const float* __restrict__ drgb, int64_t n, int d, int deg,
                                                     float* __restrict__ dfeat) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const f3 g = ld3(drgb + 3 * p);
  float* df = dfeat + (int64_t)3 * d * p;
  const bool skip = fabsf(g.x) < 1e-8f && fabsf(g.y) < 1e-8f && fabsf(g.z) < 1e-8f;
  const int nc = (deg + 1) * (deg + 1);
  float b[16];
  if (!skip) {
    const f3 dir = sh_dir(ld3(means + 3 * p), ld3(origin));
    sh_basis(deg, dir.x, dir.y, dir.z, b);
  }
  for (int k = 0; k < d; ++k) {
    const bool live = !skip && k < nc;
    df[3 * k + 0] = live ? b[k] * g.x : 0.f;
    df[3 * k + 1] = live ? b[k] * g.y : 0.f;
    df[3 * k + 2] = live ? b[k] * g.z : 0.f;
  }
}

hipError_t launch_eval_sh(const float* means, const float* feat, const float* origin, int64_t n, int d, int deg,
                          float* out, hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL(k_eval_sh, dim3(grid_for(n, 256)), dim3(256), 0, st, means, feat, origin, n, d, deg, out);
  return hipGetLastError();
}

hipError_t launch_eval_sh_bwd(const float* means, const float* feat, const float* origin, const float* drgb,
                              int64_t n, int d, int deg, float* dfeat, hipStream_t st) {
  (void)feat;
  if (n > 0)
    hipLaunchKernelGGL(k_eval_sh_bwd, dim3(grid_for(n, 256)), dim3(256), 0, st, means, origin, drgb, n, d, deg, dfeat);
  return hipGetLastError();
}

}  // namespace ever
