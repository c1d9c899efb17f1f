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

// One safe_div denominator used several times in an adjoint: the reciprocals of b' (b, or kTiny where |b| <
// kTiny) and of b2' (b*b, or kTiny below it), from the hardware reciprocal (within 1 ulp). With them
// safe_div(a, b) ~ clip(a * r1) and bw_safe_div(a, b, g) ~ (clip(g * r1), clip(-a * r2 * g)), a few ulp from
// the divisions; used in gradients only, never in the forward's colours.
struct SafeRcp {
  float r1, r2;
};
__device__ __forceinline__ SafeRcp safe_rcp(float b) {
  const float b2 = b * b;
  return {__builtin_amdgcn_rcpf(fabsf(b) < kTiny ? kTiny : b), __builtin_amdgcn_rcpf(b2 < kTiny ? kTiny : b2)};
}
__device__ __forceinline__ float safe_div_r(float a, const SafeRcp& r) { return clipf(a * r.r1, kMinVal, kMaxVal); }
__device__ __forceinline__ void bw_safe_div_r(float a, const SafeRcp& r, float g, float& da, float& db) {
  da = clipf(g * r.r1, kMinVal, kMaxVal);
  db = clipf(-a * r.r2 * g, kMinVal, kMaxVal);
}

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
  const SafeRcp rn = safe_rcp(n);
  float ax, ay, az, b0, b1, b2;
  bw_safe_div_r(x.x, rn, yb.x, ax, b0);
  bw_safe_div_r(x.y, rn, yb.y, ay, b1);
  bw_safe_div_r(x.z, rn, yb.z, az, b2);
  f3 xb = mk3(ax, ay, az);
  if (m > kTiny) xb = xb + ((b0 + b1 + b2) * __builtin_amdgcn_rcpf(n)) * x;
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
// The same derivative from the activated value y = softplus10(x): sigmoid(10 x) = 1 - exp(-10 y), evaluated as
// -expm1(-10 y) (a few ulp; above the threshold, y = x > 2 and it rounds to 1 as the derivative is there)
__device__ __forceinline__ float softplus10_grad_from_value(float y) { return -expm1f(-10.f * y); }

// FP64 vector helpers for the root evaluation of the backward.
struct d3 {
  double x, y, z;
};
__device__ __forceinline__ d3 mkd(double x, double y, double z) { return {x, y, z}; }
__device__ __forceinline__ d3 operator+(d3 a, d3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ d3 operator-(d3 a, d3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ d3 operator*(double s, d3 a) { return {s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ double dotd(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ d3 crossd(d3 a, d3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ d3 tod(f3 a) { return {(double)a.x, (double)a.y, (double)a.z}; }
__device__ __forceinline__ f3 tof(d3 a) { return {(float)a.x, (float)a.y, (float)a.z}; }
// safe_div(a, b) in FP64 as a * (1 / b'), b' the guarded denominator, clamped as safe_div clamps
__device__ __forceinline__ double safe_mul_d(double a, double rb) {
  return fmax(fmin(a * rb, (double)kMaxVal), (double)kMinVal);
}
__device__ __forceinline__ d3 qrot_d(d3 v, float4 q) {
  d3 u = mkd(-q.y, -q.z, -q.w);
  d3 t = 2.0 * crossd(u, v);
  return v + (double)q.x * t + crossd(u, t);
}

// Start-inside containment in FP64 (the point of a surface through the start is decided at the oracle's precision):
// p - mean rotated by the quaternion normalised by its length, divided by the scales, squared length <= 1.
// `adjoint`: the form of backwards_kernel.slang:395-405 (guarded divisions, scales clipped at 1e-8, length <= 1);
// else initialize_density.cu:107-150.
__device__ __forceinline__ double safe_div_d(double a, double b) {
  return fmax(fmin(a / (fabs(b) < (double)kTiny ? (double)kTiny : b), (double)kMaxVal), (double)kMinVal);
}
__device__ __forceinline__ bool inside_d(d3 p, f3 mean, f3 s, float4 q, bool adjoint) {
  const double qn = sqrt((double)q.x * q.x + (double)q.y * q.y + (double)q.z * q.z + (double)q.w * q.w);
  const double w = adjoint ? safe_div_d(q.x, qn) : q.x / qn, x = adjoint ? safe_div_d(q.y, qn) : q.y / qn,
               y = adjoint ? safe_div_d(q.z, qn) : q.z / qn, z = adjoint ? safe_div_d(q.w, qn) : q.w / qn;
  const d3 v = p - tod(mean);
  const d3 u = mkd(-x, -y, -z);
  const d3 t = 2.0 * crossd(u, v);
  const d3 r = v + w * t + crossd(u, t);
  if (adjoint) {
    const d3 tr = mkd(safe_div_d(r.x, fmax((double)s.x, 1e-8)), safe_div_d(r.y, fmax((double)s.y, 1e-8)),
                      safe_div_d(r.z, fmax((double)s.z, 1e-8)));
    return sqrt(dotd(tr, tr)) <= 1.0;
  }
  const d3 tr = mkd(r.x / s.x, r.y / s.y, r.z / s.z);
  return dotd(tr, tr) <= 1.0;
}

// FP64 reciprocal and reciprocal square root from the hardware estimates (v_rcp_f64, v_rsq_f64) refined by two
// Newton steps: within a few ulp of the IEEE operations.
__device__ __forceinline__ double rcp_nr_d(double b) {
  double r = __builtin_amdgcn_rcp(b);
  r = fma(fma(-b, r, 1.0), r, r);
  return fma(fma(-b, r, 1.0), r, r);
}
__device__ __forceinline__ double rsq_nr_d(double x) {
  double r = __builtin_amdgcn_rsq(x);
  r = r * fma(-0.5 * x * r, r, 1.5);
  return r * fma(-0.5 * x * r, r, 1.5);
}

// The ray of the integrator in FP64: unit direction (l2 normalization of the raw direction); the origin is
// rayo + tmin * rayd evaluated in FP64 (shaders.slang:178-179).
__device__ __forceinline__ d3 ray_dir_fp64(f3 rdraw) {
  d3 d = tod(rdraw);
  const double nn = 1.0 / sqrt(dotd(d, d));
  return mkd(d.x * nn, d.y * nn, d.z * nn);
}
__device__ __forceinline__ d3 ray_origin_fp64(f3 ro, f3 rdraw, float tmin) {
  return mkd((double)ro.x + (double)tmin * rdraw.x, (double)ro.y + (double)tmin * rdraw.y,
             (double)ro.z + (double)tmin * rdraw.z);
}

// Roots of one primitive in FP64 from its raw parameters, as the forward integrates them: eliIntersect
// (tri-intersect.slang:268-283) on the ray rotated by the raw quaternion (rotate_vector, safe-math.slang:237-244)
// and divided by the scales. Returns whether 1 - |l|^2 >= 0 (the hit test of :275); the roots use it clamped at 0.
// Written without contraction, so that every kernel that inlines it (both tree walks) computes identical depths.
__device__ __forceinline__ bool raw_roots_d(d3 o, d3 d, float4 q, f3 sc, f3 mu, double& t0, double& t1) {
#pragma clang fp contract(off)
  const d3 Tro = qrot_d(o - tod(mu), q), Trd = qrot_d(d, q);
  // one reciprocal per shared denominator (within an ulp of the quotients)
  const double rx = 1.0 / (double)sc.x, ry = 1.0 / (double)sc.y, rz = 1.0 / (double)sc.z;
  const d3 ocn = mkd(Tro.x * rx, Tro.y * ry, Tro.z * rz);
  const d3 rdn = mkd(Trd.x * rx, Trd.y * ry, Trd.z * rz);
  const double a = dotd(rdn, rdn);
  const double ia = 1.0 / a;
  const double bp = -dotd(ocn, rdn);
  const d3 l = ocn + (bp * ia) * rdn;
  const double h = a * (1.0 - dotd(l, l));
  const double c = dotd(ocn, ocn) - 1.0;
  const double qq = bp + (bp > 0.0 ? 1.0 : (bp < 0.0 ? -1.0 : 0.0)) * sqrt(fmax(h, 0.0));
  const double x0 = c / qq, x1 = qq * ia;
  t0 = fmin(x0, x1);
  t1 = fmax(x0, x1);
  return !(h < 0.0);
}

// FP64 forms of the safe-math adjoint helpers (safe-math.slang:68-106), with the refined hardware reciprocals.
__device__ __forceinline__ double clipd(double v) { return fmax(fmin(v, (double)kMaxVal), (double)kMinVal); }
// 1/b' and 1/b2' of a guarded denominator: b' = b, or kTiny where |b| < kTiny; b2' = b*b, or kTiny below it
struct SafeRcpD {
  double r1, r2;
};
__device__ __forceinline__ SafeRcpD safe_rcp_fd(double b) {
  const double r1 = rcp_nr_d(fabs(b) < (double)kTiny ? (double)kTiny : b);
  return {r1, b * b < (double)kTiny ? 1.0 / (double)kTiny : r1 * r1};
}

// Safe ellipsoid roots with the backward's conventions (tri-intersect.slang:286-344: the rotated raw direction is
// l2-normalized, divisions use safe_div) and, when the ray meets the ellipsoid, the gradient of the root the face
// selects (entry: the smaller, safe_intersect, tri-intersect.slang:443-456) w.r.t. rod = origin - mean, rd = raw
// direction, the scale and the raw quaternion: the adjoint of the root with a unit upstream gradient. Evaluated in
// FP64 (reciprocals and reciprocal square roots from the hardware estimates refined by Newton steps) in the
// primitive's normalized frame (o = rotated origin / scale, d = rotated unit direction / scale), so that 1 - |l|^2
// and the depths keep their precision at grazing incidence and far from the ray origin, and so that the scale
// gradient -(o * ob + d * db) / s, which cancels by |o| along a thin axis (o and t * d are large and nearly opposite
// there, their sum being the hit point on the unit sphere), loses no digits that matter. Returns the selected root
// (-1 for a miss); hit: 1 - |l|^2 >= 0.
__device__ __forceinline__ double safe_roots_jac_d(d3 rod, f3 rd, f3 s, float4 q, bool entry, f3& jro, f3& jrd,
                                                   f3& js, float4& jq, bool& hit) {
  // the roots
  const d3 rd_rot = qrot_d(tod(rd), q);
  const double rnrm = rsq_nr_d(fmax(dotd(rd_rot, rd_rot), (double)kTiny));
  const d3 Trd = mkd(safe_mul_d(rd_rot.x, rnrm), safe_mul_d(rd_rot.y, rnrm), safe_mul_d(rd_rot.z, rnrm));
  const d3 Tro = qrot_d(rod, q);
  const SafeRcpD rx = safe_rcp_fd(s.x), ry = safe_rcp_fd(s.y), rz = safe_rcp_fd(s.z);
  const d3 o = mkd(safe_mul_d(Tro.x, rx.r1), safe_mul_d(Tro.y, ry.r1), safe_mul_d(Tro.z, rz.r1));
  const d3 d = mkd(safe_mul_d(Trd.x, rx.r1), safe_mul_d(Trd.y, ry.r1), safe_mul_d(Trd.z, rz.r1));
  const double a = dotd(d, d);
  const double bp = -dotd(o, d);
  const SafeRcpD ra = safe_rcp_fd(a);
  const double ba = safe_mul_d(bp, ra.r1);
  const d3 l = o + ba * d;
  const double omll = 1.0 - dotd(l, l);
  const double h = a * omll;
  const double c = dotd(o, o) - 1.0;
  hit = !(h < 0.0);
  double rs = 0.0, sh = 0.0;
  if (!(h < (double)kTiny)) {
    rs = rsq_nr_d(h);
    sh = h * rs;
  }
  const double sb = bp > 0.0 ? 1.0 : (bp < 0.0 ? -1.0 : 0.0);
  const double qq = bp + sb * sh;
  const SafeRcpD rq = safe_rcp_fd(qq);
  const double r0 = safe_mul_d(c, rq.r1), r1 = safe_mul_d(qq, ra.r1);
  jro = jrd = js = mk3(0.f, 0.f, 0.f);
  jq = make_float4(0.f, 0.f, 0.f, 0.f);
  if (!hit) return -1.0;
  const bool sel_r0 = entry ? (r0 < r1) : (r0 > r1);
  // the adjoint of the selected root with a unit upstream gradient
  const double r0b = sel_r0 ? 1.0 : 0.0, r1b = sel_r0 ? 0.0 : 1.0;
  const double cb = clipd(r0b * rq.r1);
  double qqb = clipd(-c * rq.r2 * r0b) + clipd(r1b * ra.r1);
  double ab = clipd(-qq * ra.r2 * r1b);
  const double hb = qqb * sb * 0.5 * rs;  // 0 where h < kTiny
  ab += hb * omll;
  const d3 lb = (-2.0 * hb * a) * l;
  d3 ocnb = (2.0 * cb) * o + lb;
  const double bab = dotd(lb, d);
  d3 rdnb = ba * lb;
  const double bpb = qqb + clipd(bab * ra.r1);
  ab += clipd(-bp * ra.r2 * bab);
  ocnb = ocnb + (-bpb) * d;
  rdnb = rdnb + (-bpb) * o + (2.0 * ab) * d;
  const f3 Trob = mk3((float)clipd(ocnb.x * rx.r1), (float)clipd(ocnb.y * ry.r1), (float)clipd(ocnb.z * rz.r1));
  const f3 Trdb = mk3((float)clipd(rdnb.x * rx.r1), (float)clipd(rdnb.y * ry.r1), (float)clipd(rdnb.z * rz.r1));
  const double sx = (double)s.x, sy = (double)s.y, sz = (double)s.z;
  js.x = (float)(clipd(-(o.x * sx) * rx.r2 * ocnb.x) + clipd(-(d.x * sx) * rx.r2 * rdnb.x));
  js.y = (float)(clipd(-(o.y * sy) * ry.r2 * ocnb.y) + clipd(-(d.y * sy) * ry.r2 * rdnb.y));
  js.z = (float)(clipd(-(o.z * sz) * rz.r2 * ocnb.z) + clipd(-(d.z * sz) * rz.r2 * rdnb.z));
  qrot_bwd(tof(rod), q, Trob, jro, jq);
  const f3 rrb = l2n_bwd(qrot(rd, q), Trdb);
  qrot_bwd(rd, q, rrb, jrd, jq);
  return entry ? fmin(r0, r1) : fmax(r0, r1);
}

// Relaxed device-scope float atomic add (global_atomic_add_f32 on gfx950; result unused).
__device__ __forceinline__ void atomic_addf(float* p, float v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace ever
