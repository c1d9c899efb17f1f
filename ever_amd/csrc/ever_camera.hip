// Camera ray generation: one ray per pixel of a perspective, fisheye or panoramic camera with optional radial and
// tangential lens distortion. Replaces the caller-side tensor code of integration_utils/gaussian_renderer/ever.py:
// 29-121 (get_ray_directions + get_rays for the pinhole camera, camera2rays_full -> camera_utils_zipnerf.py:909-1040
// pixels_to_rays for the others), which issues one launch per tensor operation and, for distorted cameras, ten
// Newton iterations of them. Here it is one pass: each lane computes its pixel's ray and writes 24 bytes.
#include "ever_internal.hpp"

namespace ever {
namespace {

// Brown-Conrady model: distorted = p * (1 + k1 r + k2 r^2 + k3 r^3 + k4 r^4) + tangential(p1, p2), r = |p|^2.
// Inverted by Newton's method from the distorted point, ten steps, a step skipped where the Jacobian's
// determinant is below 1e-9 in magnitude (camera_utils_zipnerf.py:816-906).
__device__ __forceinline__ void undistort(float& x, float& y, const float (&k)[6]) {
  const float xd = x, yd = y;
  const float k1 = k[0], k2 = k[1], k3 = k[2], k4 = k[3], p1 = k[4], p2 = k[5];
#pragma unroll 1
  for (int it = 0; it < 10; ++it) {
    const float r = x * x + y * y;
    const float d = 1.f + r * (k1 + r * (k2 + r * (k3 + r * k4)));
    // residuals of the forward model at (x, y)
    const float ex = d * x + 2.f * p1 * x * y + p2 * (r + 2.f * x * x) - xd;
    const float ey = d * y + 2.f * p2 * x * y + p1 * (r + 2.f * y * y) - yd;
    // Jacobian of the forward model
    const float dd = k1 + r * (2.f * k2 + r * (3.f * k3 + r * 4.f * k4));  // d'(r)
    const float ddx = 2.f * x * dd, ddy = 2.f * y * dd;
    const float jxx = d + ddx * x + 2.f * p1 * y + 6.f * p2 * x;
    const float jxy = ddy * x + 2.f * p1 * x + 2.f * p2 * y;
    const float jyx = ddx * y + 2.f * p2 * y + 2.f * p1 * x;
    const float jyy = d + ddy * y + 2.f * p2 * x + 6.f * p1 * y;
    const float det = jxx * jyy - jxy * jyx;
    if (fabsf(det) > 1e-9f) {
      x -= (ex * jyy - ey * jxy) / det;
      y -= (ey * jxx - ex * jyx) / det;
    }
  }
}

__global__ __launch_bounds__(256) void k_camera_rays(ever_camera C, const float2* __restrict__ jitter,
                                                     float* __restrict__ origins, float* __restrict__ directions) {
  const int64_t n = (int64_t)C.width * C.height;
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int y = (int)(r / C.width);
  const int x = (int)(r - (int64_t)y * C.width);
  const float2 j = jitter ? jitter[r] : make_float2(0.5f, 0.5f);
  float u = ((float)x + j.x - C.cx) / C.fx;
  float v = ((float)y + j.y - C.cy) / C.fy;
  if (C.distorted) {
    const float k[6] = {C.dist[0], C.dist[1], C.dist[2], C.dist[3], C.dist[4], C.dist[5]};
    undistort(u, v, k);
  }
  float dc[3];
  if (C.model == EVER_CAMERA_FISHEYE) {
    const float th = fminf(sqrtf(u * u + v * v), 3.14159265358979f);
    // sin(theta)/theta -> 1 at the optical axis
    const float s = th > 0.f ? sinf(th) / th : 1.f;
    dc[0] = u * s;
    dc[1] = v * s;
    dc[2] = cosf(th);
  } else if (C.model == EVER_CAMERA_PANORAMIC) {
    const float su = sinf(u), cu = cosf(u), sv = sinf(v), cv = cosf(v);
    dc[0] = -sv * su;
    dc[1] = -cv;
    dc[2] = -sv * cu;
  } else {
    dc[0] = u;
    dc[1] = v;
    dc[2] = 1.f;
  }
  const float* M = C.c2w;  // uniform: scalar loads
  float w[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) w[a] = M[4 * a] * dc[0] + M[4 * a + 1] * dc[1] + M[4 * a + 2] * dc[2];
  const float inv = 1.f / sqrtf(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    directions[3 * r + a] = w[a] * inv;
    origins[3 * r + a] = M[4 * a + 3];
  }
}

}  // namespace

hipError_t launch_camera_rays(const ever_camera& C, const float* jitter, float* origins, float* directions,
                              hipStream_t st) {
  const int64_t n = (int64_t)C.width * C.height;
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_camera_rays, dim3(grid_for(n, 256)), dim3(256), 0, st, C,
                     reinterpret_cast<const float2*>(jitter), origins, directions);
  return hipGetLastError();
}

}  // namespace ever
