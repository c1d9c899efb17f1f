// Device-side math shared by the build, forward and backward kernels (gfx950, wave64).
//
// Semantics restated from the reference (paths relative to the reference repository root):
//   safe_div / bw_safe_div / safe_sqrt / l2_normalize / tukey_power_ladder / rotate_vector
//                                              splinetracers/slang/safe-math.slang:68-139, 237-253
//   ellipsoid roots (fast + safe variants)     splinetracers/fast_ellipsoid_splinetracer/slang/tri-intersect.slang:268-344
//   SH constants and basis                     splinetracers/slang/sh.slang:18-103
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ever {

constexpr float kTiny = 1.0754944e-20f;  // safe-math.slang:16
constexpr float kMinVal = -1e20f;
constexpr float kMaxVal = 1e20f;
constexpr float kLogCutoff = 5.54f;       // shaders.slang:17
constexpr float kPreMulti = 1000.f;       // spline-machine.slang:17
constexpr float kShC0 = 0.28209479177387814f;
constexpr float kShC1 = 0.4886025119029199f;
constexpr float kEmptyT = 1e10f;          // empty k-buffer slot (shaders.slang:199-201)
constexpr float kEndT = 1e9f;             // end-of-events threshold (shaders.slang:214)

struct f3 {
  float x, y, z;
};
__device__ __forceinline__ f3 mk3(float x, float y, float z) { return {x, y, z}; }
__device__ __forceinline__ f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 operator*(float s, f3 a) { return {s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ f3 cross3(f3 a, f3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ f3 ld3(const float* p) { return {p[0], p[1], p[2]}; }

__device__ __forceinline__ float clipf(float v, float lo, float hi) { return fmaxf(fminf(v, hi), lo); }
__device__ __forceinline__ float sgnf(float v) { return v > 0.f ? 1.f : (v < 0.f ? -1.f : 0.f); }

__device__ __forceinline__ float safe_div(float a, float b) {
  return clipf(a / (fabsf(b) < kTiny ? kTiny : b), kMinVal, kMaxVal);
}
__device__ __forceinline__ void bw_safe_div(float a, float b, float g, float& da, float& db) {
  da = clipf(g / (fabsf(b) < kTiny ? kTiny : b), kMinVal, kMaxVal);
  float b2 = b * b;
  db = clipf(-a / (b2 < kTiny ? kTiny : b2) * g, kMinVal, kMaxVal);
}
__device__ __forceinline__ float safe_sqrt(float a) { return a < kTiny ? 0.f : sqrtf(a); }

// rotation by the conjugate quaternion of q = (w, x, y, z)  (safe-math.slang:237-244)
__device__ __forceinline__ f3 qrot(f3 v, float4 q) {
  f3 u = mk3(-q.y, -q.z, -q.w);
  f3 t = 2.f * cross3(u, v);
  return v + q.x * t + cross3(u, t);
}
__device__ __forceinline__ void qrot_bwd(f3 v, float4 q, f3 ob, f3& vb, float4& qb) {
  f3 u = mk3(-q.y, -q.z, -q.w);
  f3 t = 2.f * cross3(u, v);
  f3 ub = cross3(t, ob);
  f3 tb = cross3(ob, u) + q.x * ob;
  float wb = dot3(ob, t);
  ub = ub + 2.f * cross3(v, tb);
  vb = vb + ob + 2.f * cross3(tb, u);
  qb.x += wb;
  qb.y -= ub.x;
  qb.z -= ub.y;
  qb.w -= ub.z;
}

__device__ __forceinline__ f3 l2n(f3 x) {
  float n = sqrtf(fmaxf(dot3(x, x), kTiny));
  return mk3(safe_div(x.x, n), safe_div(x.y, n), safe_div(x.z, n));
}
__device__ __forceinline__ f3 l2n_bwd(f3 x, f3 yb) {
  float m = dot3(x, x);
  float n = sqrtf(fmaxf(m, kTiny));
  float ax, ay, az, b0, b1, b2;
  bw_safe_div(x.x, n, yb.x, ax, b0);
  bw_safe_div(x.y, n, yb.y, ay, b1);
  bw_safe_div(x.z, n, yb.z, az, b2);
  f3 xb = mk3(ax, ay, az);
  if (m > kTiny) xb = xb + ((b0 + b1 + b2) / n) * x;
  return xb;
}

// tukey_power_ladder(x, -0.1) = 11 * sign(x) * (1 - (1 + |x|/1.1)^-0.1)   (safe-math.slang:246-253)
__device__ __forceinline__ float ladder(float x) {
  float xs = fabsf(x) * (1.f / 1.1f);
  return sgnf(x) * (-11.f) * (exp2f(-0.1f * log2f(xs + 1.f)) - 1.f);
}
__device__ __forceinline__ float ladder_grad(float x) {
  if (x == 0.f) return 0.f;
  float xs = fabsf(x) * (1.f / 1.1f);
  return exp2f(-1.1f * log2f(xs + 1.f));
}

// Real SH basis up to degree 3 (sh.slang:48-88), PlenOctree signs.
__device__ __forceinline__ void sh_basis(int deg, float x, float y, float z, float* b) {
  b[0] = kShC0;
  if (deg > 0) {
    b[1] = -kShC1 * y;
    b[2] = kShC1 * z;
    b[3] = -kShC1 * x;
    if (deg > 1) {
      float xx = x * x, yy = y * y, zz = z * z, xy = x * y, yz = y * z, xz = x * z;
      b[4] = 1.0925484305920792f * xy;
      b[5] = -1.0925484305920792f * yz;
      b[6] = 0.31539156525252005f * (2.f * zz - xx - yy);
      b[7] = -1.0925484305920792f * xz;
      b[8] = 0.5462742152960396f * (xx - yy);
      if (deg > 2) {
        b[9] = -0.5900435899266435f * y * (3.f * xx - yy);
        b[10] = 2.890611442640554f * xy * z;
        b[11] = -0.4570457994644658f * y * (4.f * zz - xx - yy);
        b[12] = 0.3731763325901154f * z * (2.f * zz - 3.f * xx - 3.f * yy);
        b[13] = -0.4570457994644658f * x * (4.f * zz - xx - yy);
        b[14] = 1.445305721320277f * z * (xx - yy);
        b[15] = -0.5900435899266435f * x * (xx - 3.f * yy);
      }
    }
  }
}

// Direction of the per-primitive SH prepass: normalize(mean - origin) (sh_kernel.slang:37-39).
__device__ __forceinline__ f3 sh_dir(f3 mean, f3 origin) {
  f3 v = mean - origin;
  float inv = 1.f / sqrtf(dot3(v, v));
  return inv * v;
}

// softplus(x, beta=10) as torch.nn.functional.softplus (threshold 20), used by the reference's caller
// (integration_utils/gaussian_renderer/ever.py:179).
__device__ __forceinline__ float softplus10(float x) {
  float bx = 10.f * x;
  return bx > 20.f ? x : log1pf(expf(bx)) * 0.1f;
}
__device__ __forceinline__ float softplus10_grad(float x) {
  float bx = 10.f * x;
  return bx > 20.f ? 1.f : 1.f / (1.f + expf(-bx));
}

// Safe ellipsoid roots with the backward's conventions (tri-intersect.slang:286-344): the rotated raw
// direction is l2-normalized, divisions use safe_div. Returns false on a miss.
struct SafeRoots {
  f3 Tro, Trd, rd_rot, ocn, rdn;
  float a, bp, ba, ll, h, c, sh, sb, qq, r0, r1;
  f3 l;
};
__device__ __forceinline__ bool safe_roots(f3 ro, f3 rd, f3 s, float4 q, SafeRoots& S) {
  S.rd_rot = qrot(rd, q);
  S.Trd = l2n(S.rd_rot);
  S.Tro = qrot(ro, q);
  S.ocn = mk3(safe_div(S.Tro.x, s.x), safe_div(S.Tro.y, s.y), safe_div(S.Tro.z, s.z));
  S.rdn = mk3(safe_div(S.Trd.x, s.x), safe_div(S.Trd.y, s.y), safe_div(S.Trd.z, s.z));
  S.a = dot3(S.rdn, S.rdn);
  S.bp = -dot3(S.ocn, S.rdn);
  S.ba = safe_div(S.bp, S.a);
  S.l = S.ocn + S.ba * S.rdn;
  S.ll = dot3(S.l, S.l);
  S.h = S.a * (1.f - S.ll);
  S.c = dot3(S.ocn, S.ocn) - 1.f;
  if (S.h < 0.f) {
    S.r0 = S.r1 = -1.f;
    return false;
  }
  S.sh = safe_sqrt(S.h);
  S.sb = sgnf(S.bp);
  S.qq = S.bp + S.sb * S.sh;
  S.r0 = safe_div(S.c, S.qq);
  S.r1 = safe_div(S.qq, S.a);
  return true;
}
__device__ __forceinline__ float safe_root_select(const SafeRoots& S, int face) {
  return face == 1 ? fminf(S.r0, S.r1) : fmaxf(S.r0, S.r1);
}
// Adjoint of the selected root w.r.t. ro = origin - mean, rd = raw direction, scale and raw quaternion.
__device__ __forceinline__ void safe_roots_bwd(const SafeRoots& S, f3 ro, f3 rd, f3 s, float4 q, int face,
                                               float tb, f3& rob, f3& rdb, f3& sbar, float4& qb) {
  float r0b = 0.f, r1b = 0.f;
  if (face == 1) {
    if (S.r0 < S.r1) r0b = tb; else r1b = tb;
  } else {
    if (S.r0 > S.r1) r0b = tb; else r1b = tb;
  }
  float cb, qqb, ab = 0.f, bpb, hb = 0.f, da, db;
  bw_safe_div(S.c, S.qq, r0b, da, db);
  cb = da;
  qqb = db;
  bw_safe_div(S.qq, S.a, r1b, da, db);
  qqb += da;
  ab += db;
  bpb = qqb;
  if (!(S.h < kTiny)) hb = qqb * S.sb * 0.5f / S.sh;
  ab += hb * (1.f - S.ll);
  f3 lb = (2.f * (-hb * S.a)) * S.l;
  f3 ocnb = (2.f * cb) * S.ocn + lb;
  float bab = dot3(lb, S.rdn);
  f3 rdnb = S.ba * lb;
  bw_safe_div(S.bp, S.a, bab, da, db);
  bpb += da;
  ab += db;
  ocnb = ocnb + (-bpb) * S.rdn;
  rdnb = rdnb + (-bpb) * S.ocn + (2.f * ab) * S.rdn;
  f3 Trob, Trdb;
  bw_safe_div(S.Tro.x, s.x, ocnb.x, da, db); Trob.x = da; sbar.x += db;
  bw_safe_div(S.Tro.y, s.y, ocnb.y, da, db); Trob.y = da; sbar.y += db;
  bw_safe_div(S.Tro.z, s.z, ocnb.z, da, db); Trob.z = da; sbar.z += db;
  bw_safe_div(S.Trd.x, s.x, rdnb.x, da, db); Trdb.x = da; sbar.x += db;
  bw_safe_div(S.Trd.y, s.y, rdnb.y, da, db); Trdb.y = da; sbar.y += db;
  bw_safe_div(S.Trd.z, s.z, rdnb.z, da, db); Trdb.z = da; sbar.z += db;
  qrot_bwd(ro, q, Trob, rob, qb);
  f3 rrb = l2n_bwd(S.rd_rot, Trdb);
  qrot_bwd(rd, q, rrb, rdb, qb);
}

// Relaxed device-scope float atomic add (global_atomic_add_f32 on gfx950; result unused).
__device__ __forceinline__ void atomic_addf(float* p, float v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace ever
