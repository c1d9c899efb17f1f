// Forward tracer: one lane per ray, LBVH traversal with a short LDS stack (global spill beyond 32 entries),
// a per-ray register k-buffer of the K nearest enter/exit events beyond the last processed depth, and the
// exact piecewise-constant integrator applied to those events in order; re-traversal until the ray's
// log-transmittance reaches 5.54, max_iters events were processed or no event is left.
//
// Semantics follow the reference's OptiX programs (splinetracers/fast_ellipsoid_splinetracer/slang/shaders.slang):
//   intersection + report rules :84-107, any-hit k-buffer insertion (strict <) :117-164, raygen loop :171-235,
//   integrator spline-machine.slang:103-153, start-inside density initialize_density.cu:69-150.
// Differences, all toward the FP64 oracle: event depths fed to the integrator, the density/colour accumulator
// and the log-transmittance are FP64 (the accumulators exactly invertible by the adjoint); 1-exp(-x) is
// evaluated as -expm1(-x); touch counts use atomics; the re-query continues from a (t, event key) cursor
// instead of t > t_cur, so events of equal depth straddling two k-buffer fills are processed once.
#include "ever_common.hpp"
#include "ever_internal.hpp"

namespace ever {

constexpr int kFwdBlock = 64;
constexpr int kLdsDepth = 32;     // LDS stack entries per lane
constexpr int kSpillDepth = kSpillPerRay;  // global spill entries per ray

struct FwdArgs {
  const float4* rec;
  const float4* nodes4;
  const float4* colr;
  const int32_t* cand;
  const uint32_t* header;
  const float* means;
  const float* scales;
  const float* quats;
  const float* dens;
  const float* feat;
  int color_mode;
  int64_t n;
  const float* rayo;
  const float* rayd;
  int64_t nr;
  float tmin, tmax;
  int max_iters;
  float* color;
  float* states;
  int32_t* iters;
  int32_t* tri;
  int32_t* touch;
  float* init_drgb;
  double* st64;
  int32_t* spill;
  const uint32_t* order;  // processing order: thread slot -> ray index
};

// kern_prefilter (initialize_density.cu:69-105): primitives whose AABB (create_aabbs.cu:77-114) lies within tmin
// of ray 0's origin.
__global__ __launch_bounds__(256) void k_prefilter(ever_prims P, const float* __restrict__ rayo, float tmin,
                                                   int32_t* __restrict__ cand, uint32_t* __restrict__ header) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.num_prims) return;
  const float4 q = reinterpret_cast<const float4*>(P.quats)[i];
  const float qn = rsqrtf(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
  const float r = q.x * qn, x = q.y * qn, y = q.z * qn, z = q.w * qn;
  const float Rm[3][3] = {{1.f - 2.f * (y * y + z * z), 2.f * (x * y - r * z), 2.f * (x * z + r * y)},
                          {2.f * (x * y + r * z), 1.f - 2.f * (x * x + z * z), 2.f * (y * z - r * x)},
                          {2.f * (x * z - r * y), 2.f * (y * z + r * x), 1.f - 2.f * (x * x + y * y)}};
  const float sv[3] = {P.scales[3 * i], P.scales[3 * i + 1], P.scales[3 * i + 2]};
  const float o[3] = {rayo[0], rayo[1], rayo[2]};
  float dmin = 0.f;
  for (int k = 0; k < 3; ++k) {
    const float e = sqrtf(Rm[k][0] * Rm[k][0] * sv[0] * sv[0] + Rm[k][1] * Rm[k][1] * sv[1] * sv[1] +
                          Rm[k][2] * Rm[k][2] * sv[2] * sv[2]) * (1.f + 1e-6f);
    const float m = P.means[3 * i + k];
    if (o[k] < m - e) dmin += (o[k] - (m - e)) * (o[k] - (m - e));
    else if (o[k] > m + e) dmin += (o[k] - (m + e)) * (o[k] - (m + e));
  }
  if (dmin <= tmin * tmin) {
    uint32_t pos = atomicAdd(header + 8, 1u);
    cand[pos] = (int32_t)i;
  }
}

__global__ void k_reset_count(uint32_t* header) { header[8] = 0u; }

__device__ __forceinline__ uint32_t expand10(uint32_t v) {
  v = (v * 0x00010001u) & 0xFF0000FFu;
  v = (v * 0x00000101u) & 0x0F00F00Fu;
  v = (v * 0x00000011u) & 0xC30C30C3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}

// Sort key of a ray: 30-bit Morton code of the point o + d * |c - o| (c = centre of the primitive centroids),
// quantized over a cube of four times the scene extent around c. Rays of neighbouring pixels of one camera map
// to neighbouring points, so a wave traces a compact screen tile; rays from different origins separate.
__global__ __launch_bounds__(256) void k_ray_keys(const float* __restrict__ rayo, const float* __restrict__ rayd,
                                                  int64_t nr, float tmin, const uint32_t* __restrict__ header,
                                                  uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nr) return;
  auto o2f = [](uint32_t u) { return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u); };
  const f3 lo = mk3(o2f(header[0]), o2f(header[1]), o2f(header[2]));
  const f3 hi = mk3(o2f(header[3]), o2f(header[4]), o2f(header[5]));
  const f3 c = 0.5f * (lo + hi);
  const float ext = fmaxf(fmaxf(fmaxf(hi.x - lo.x, hi.y - lo.y), hi.z - lo.z), 1e-6f);
  const f3 rd = ld3(rayd + 3 * r);
  const f3 o = ld3(rayo + 3 * r) + tmin * rd;
  const f3 d = l2n(rd);
  const f3 co = c - o;
  const f3 p = o + sqrtf(dot3(co, co)) * d;
  const float inv = 1.f / (4.f * ext);
  const uint32_t qx = (uint32_t)fminf(fmaxf(((p.x - c.x) * inv + 0.5f) * 1024.f, 0.f), 1023.f);
  const uint32_t qy = (uint32_t)fminf(fmaxf(((p.y - c.y) * inv + 0.5f) * 1024.f, 0.f), 1023.f);
  const uint32_t qz = (uint32_t)fminf(fmaxf(((p.z - c.z) * inv + 0.5f) * 1024.f, 0.f), 1023.f);
  keys[r] = (expand10(qx) << 2) | (expand10(qy) << 1) | expand10(qz);
  vals[r] = (uint32_t)r;
}

__global__ void k_iota(uint32_t* v, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = (uint32_t)i;
}

__global__ void k_copy_u32(const uint32_t* __restrict__ a, int32_t* __restrict__ b, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (int32_t)a[i];
}

__global__ void k_zero_i32(int32_t* p, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// k-buffer entries are 64-bit keys (depth bits << 32 | event key), event key = 2*prim + (exit ? 1 : 0): for
// depths >= 0 the float bit pattern is monotonic, so one unsigned compare orders by (t, key) with the entry of a
// primitive before its exit on equal t. Insertion keeps the reference's stable bubble (shaders.slang:146-157).
template <int K>
__device__ __forceinline__ void kbuf_insert(uint64_t (&bk)[K], uint64_t v) {
  if (!(v < bk[K - 1])) return;
  // the fragments of one primitive report identical events: keep one
  bool dup = false;
#pragma unroll
  for (int k = 0; k < K; ++k) dup |= (bk[k] == v);
  if (dup) return;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const bool sw = v < bk[k];
    const uint64_t old = bk[k];
    bk[k] = sw ? v : old;
    v = sw ? old : v;
  }
}

__device__ __forceinline__ uint64_t pack_event(float t, uint32_t key) {
  return ((uint64_t)__float_as_uint(t + 0.f) << 32) | key;
}

__device__ __forceinline__ bool slab(float lx, float ly, float lz, float hx, float hy, float hz, f3 o, f3 inv,
                                     float t_lo, float t_hi, float& t_near) {
  float tx0 = (lx - o.x) * inv.x, tx1 = (hx - o.x) * inv.x;
  float ty0 = (ly - o.y) * inv.y, ty1 = (hy - o.y) * inv.y;
  float tz0 = (lz - o.z) * inv.z, tz1 = (hz - o.z) * inv.z;
  float tn = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), t_lo));
  float tf = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fminf(fmaxf(tz0, tz1), t_hi));
  t_near = tn;
  return tn <= tf;
}

// FP64 depth of one event for the integrator: roots of |M (o + t d - mean)| = 1 with the record's M rows and the
// primitive's mean (the same formula as tri-intersect.slang:268-283, evaluated in double so that depth
// differences keep their precision far from the origin).
__device__ __forceinline__ double event_depth_fp64(float4 RA, float4 RB, float4 RC, f3 mu, double ox, double oy,
                                                   double oz, double dx, double dy, double dz, bool entry) {
  const double vx = ox - (double)mu.x, vy = oy - (double)mu.y, vz = oz - (double)mu.z;
  const double o0 = RA.x * vx + RA.y * vy + RA.z * vz, o1 = RB.x * vx + RB.y * vy + RB.z * vz,
               o2 = RC.x * vx + RC.y * vy + RC.z * vz;
  const double r0 = RA.x * dx + RA.y * dy + RA.z * dz, r1 = RB.x * dx + RB.y * dy + RB.z * dz,
               r2 = RC.x * dx + RC.y * dy + RC.z * dz;
  const double a = r0 * r0 + r1 * r1 + r2 * r2;
  const double bp = -(o0 * r0 + o1 * r1 + o2 * r2);
  const double ba = bp / a;
  const double l0 = o0 + ba * r0, l1 = o1 + ba * r1, l2 = o2 + ba * r2;
  const double h = fmax(a * (1.0 - (l0 * l0 + l1 * l1 + l2 * l2)), 0.0);
  const double c = o0 * o0 + o1 * o1 + o2 * o2 - 1.0;
  const double q = bp + (bp > 0.0 ? 1.0 : (bp < 0.0 ? -1.0 : 0.0)) * sqrt(h);
  const double x0 = c / q, x1 = q / a;
  return entry ? fmin(x0, x1) : fmax(x0, x1);
}

template <int K>
__global__ __launch_bounds__(kFwdBlock) void k_trace_forward(FwdArgs A) {
  __shared__ int32_t stk[kLdsDepth * kFwdBlock];
  const int lane = threadIdx.x;
  const int64_t slot = (int64_t)blockIdx.x * kFwdBlock + lane;
  if (slot >= A.nr) return;
  const int64_t r = A.order[slot];

  const f3 ro = ld3(A.rayo + 3 * r), rdraw = ld3(A.rayd + 3 * r);
  const f3 d = l2n(rdraw);
  const f3 o = ro + A.tmin * rdraw;  // shaders.slang:178-179: origin shifted along the raw direction
  const f3 inv = mk3(fabsf(d.x) > 1e-20f ? 1.f / d.x : copysignf(1e20f, d.x),
                     fabsf(d.y) > 1e-20f ? 1.f / d.y : copysignf(1e20f, d.y),
                     fabsf(d.z) > 1e-20f ? 1.f / d.z : copysignf(1e20f, d.z));
  const double odx = (double)ro.x + (double)A.tmin * rdraw.x, ody = (double)ro.y + (double)A.tmin * rdraw.y,
               odz = (double)ro.z + (double)A.tmin * rdraw.z;
  double ddx = rdraw.x, ddy = rdraw.y, ddz = rdraw.z;
  {
    const double nn = 1.0 / sqrt(ddx * ddx + ddy * ddy + ddz * ddz);
    ddx *= nn;
    ddy *= nn;
    ddz *= nn;
  }

  // initial density / colour of ellipsoids containing rayo + tmin * d (initialize_density.cu:107-150)
  double D0 = 0.0, D1 = 0.0, D2 = 0.0, D3 = 0.0;
  {
    const int32_t ncand = (int32_t)A.header[8];
    const f3 p = ro + A.tmin * d;
    for (int32_t j = 0; j < ncand; ++j) {
      const int64_t i = A.cand[j];
      float4 q = reinterpret_cast<const float4*>(A.quats)[i];
      float qn = 1.f / sqrtf(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
      q = make_float4(q.x * qn, q.y * qn, q.z * qn, q.w * qn);
      f3 v = qrot(p - ld3(A.means + 3 * i), q);
      f3 s = ld3(A.scales + 3 * i);
      f3 tr = mk3(v.x / s.x, v.y / s.y, v.z / s.z);
      if (dot3(tr, tr) <= 1.f) {
        float den = A.dens[i];
        float c0, c1, c2;
        if (A.color_mode == EVER_COLOR_SH) {
          float4 cr = A.colr[i];
          c0 = cr.y; c1 = cr.z; c2 = cr.w;
        } else {  // the reference reads the features at flat index prim*3 + c (initialize_density.cu:140-142)
          c0 = A.feat[3 * i + 0] * kShC0 + 0.5f;
          c1 = A.feat[3 * i + 1] * kShC0 + 0.5f;
          c2 = A.feat[3 * i + 2] * kShC0 + 0.5f;
        }
        D0 += den;
        D1 += den * c0;
        D2 += den * c1;
        D3 += den * c2;
      }
    }
  }
  reinterpret_cast<float4*>(A.init_drgb)[r] = make_float4((float)D0, (float)D1, (float)D2, (float)D3);

  float C0 = 0.f, C1 = 0.f, C2 = 0.f, csx = 0.f, csy = 0.f, dpx = 0.f, dpy = 0.f;
  double t_state = 0.0, logT = 0.0;
  // cursor over (t, key): initially nothing at t == 0 (the reference requires t > t_cur = 0)
  float cur_t = 0.f;
  uint64_t cursor = pack_event(0.f, 0xffffffffu);
  int iter = 0;
  const int64_t nleaves = A.header[9];
  const int32_t root = nleaves == 1 ? ~0 : 0;
  int32_t* spill = A.spill + slot * kSpillDepth;

  while (logT < (double)kLogCutoff && iter < A.max_iters) {
    uint64_t bk[K];
#pragma unroll
    for (int k = 0; k < K; ++k) bk[k] = pack_event(kEmptyT, 0xffffffffu);
    // ---- traversal: gather the K nearest events after the cursor ----
    int sp = 0;
    int32_t node = root;
    int64_t steps = 0;
    const int64_t max_steps = 2 * nleaves + 2;  // a tree visits each node at most once per traversal
    while (++steps <= max_steps) {
      if (node >= 0) {
        // 4-wide record: slab test of up to four children, nearest first
        const float4* nd = A.nodes4 + 8 * (int64_t)node;
        const float4 lx = nd[0], hx = nd[1], ly = nd[2], hy = nd[3], lz = nd[4], hz = nd[5], cf = nd[6];
        const float t_hi = fminf(A.tmax, __uint_as_float((uint32_t)(bk[K - 1] >> 32)));
        float d[4];
        int32_t c[4] = {__float_as_int(cf.x), __float_as_int(cf.y), __float_as_int(cf.z), __float_as_int(cf.w)};
        const float lxs[4] = {lx.x, lx.y, lx.z, lx.w}, hxs[4] = {hx.x, hx.y, hx.z, hx.w};
        const float lys[4] = {ly.x, ly.y, ly.z, ly.w}, hys[4] = {hy.x, hy.y, hy.z, hy.w};
        const float lzs[4] = {lz.x, lz.y, lz.z, lz.w}, hzs[4] = {hz.x, hz.y, hz.z, hz.w};
        int nhit = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          float tn;
          // unused slots (kEmptyChild) never hit, whatever their stored box
          const bool h = slab(lxs[k], lys[k], lzs[k], hxs[k], hys[k], hzs[k], o, inv, cur_t, t_hi, tn) &&
                         c[k] != kEmptyChild;
          d[k] = h ? tn : INFINITY;
          nhit += h ? 1 : 0;
        }
        if (nhit > 0) {
          // sorting network on (d, c), ascending
#define EVER_CSWAP(a, b)                                      \
  if (d[b] < d[a]) {                                          \
    const float td = d[a]; d[a] = d[b]; d[b] = td;            \
    const int32_t tc = c[a]; c[a] = c[b]; c[b] = tc;          \
  }
          EVER_CSWAP(0, 1) EVER_CSWAP(2, 3) EVER_CSWAP(0, 2) EVER_CSWAP(1, 3) EVER_CSWAP(1, 2)
#undef EVER_CSWAP
#pragma unroll
          for (int k = 3; k >= 1; --k) {
            if (k < nhit && sp < kLdsDepth + kSpillDepth) {  // bounded by 3 x (LBVH depth / 2)
              if (sp < kLdsDepth) stk[sp * kFwdBlock + lane] = c[k];
              else spill[sp - kLdsDepth] = c[k];
              ++sp;
            }
          }
          node = c[0];
          continue;
        }
      } else {
        // leaf: ellipsoid roots (tri-intersect.slang:268-283) in the record's frame
        const uint32_t s = (uint32_t)(~node);
        const float4 RA = A.rec[4 * (int64_t)s + 0], RB = A.rec[4 * (int64_t)s + 1], RC = A.rec[4 * (int64_t)s + 2];
        const f3 ocn = mk3(RA.x * o.x + RA.y * o.y + RA.z * o.z + RA.w, RB.x * o.x + RB.y * o.y + RB.z * o.z + RB.w,
                           RC.x * o.x + RC.y * o.y + RC.z * o.z + RC.w);
        const f3 rdn = mk3(RA.x * d.x + RA.y * d.y + RA.z * d.z, RB.x * d.x + RB.y * d.y + RB.z * d.z,
                           RC.x * d.x + RC.y * d.y + RC.z * d.z);
        const float aa = dot3(rdn, rdn);
        const float bp = -dot3(ocn, rdn);
        const f3 l = ocn + (bp / aa) * rdn;
        const float h = aa * (1.f - dot3(l, l));
        if (h >= 0.f) {
          const float cc = dot3(ocn, ocn) - 1.f;
          const float qq = bp + sgnf(bp) * sqrtf(h);
          const float r0 = cc / qq, r1 = qq / aa;
          const float t0 = fminf(r0, r1), t1 = fmaxf(r0, r1);
          // intersection program (shaders.slang:96-106) + any-hit insertion of entry and exit (:138-159), with
          // the (t, key) cursor in place of t > t_cur
          if (!(t1 < cur_t)) {
            const uint64_t ve = pack_event(t0, 2u * s), vx = pack_event(t1, 2u * s + 1u);
            if (t0 >= cur_t && ve > cursor) {
              if (t0 <= A.tmax) {
                kbuf_insert<K>(bk, ve);
                kbuf_insert<K>(bk, vx);
              }
            } else if (vx > cursor && t1 <= A.tmax) {
              kbuf_insert<K>(bk, vx);
            }
          }
        }
      }
      if (sp == 0) break;
      --sp;
      node = sp < kLdsDepth ? stk[sp * kFwdBlock + lane] : spill[sp - kLdsDepth];
    }
    // ---- integrate the gathered events in order (shaders.slang:210-226) ----
    // The head of the buffer is consumed and the rest shifted down, so the loop body exists once.
    bool end = false;
    for (int k = 0; k < K; ++k) {
      const uint64_t head = bk[0];
#pragma unroll
      for (int j = 0; j < K - 1; ++j) bk[j] = bk[j + 1];
      const float tk = __uint_as_float((uint32_t)(head >> 32));
      if (tk > kEndT) {
        end = true;
        break;
      }
      const uint32_t key = (uint32_t)head;
      const uint32_t s = key >> 1;
      const bool entry = (key & 1u) == 0u;
      const float4 RA = A.rec[4 * (int64_t)s + 0], RB = A.rec[4 * (int64_t)s + 1], RC = A.rec[4 * (int64_t)s + 2];
      const float4 Dr = A.rec[4 * (int64_t)s + 3];
      const uint32_t orig = s;  // records are in original primitive order
      const f3 mu = ld3(A.means + 3 * (int64_t)orig);
      const double t = event_depth_fp64(RA, RB, RC, mu, odx, ody, odz, ddx, ddy, ddz, entry);
      const float hgt = entry ? Dr.x : -Dr.x;
      // update (spline-machine.slang:103-138)
      const float dt = fmaxf((float)(t - t_state), 0.f);
      const float Df = (float)D0;
      const float area = fmaxf(Df * dt, 0.f);
      const float rn0 = safe_div((float)D1, Df), rn1 = safe_div((float)D2, Df), rn2 = safe_div((float)D3, Df);
      const float e2 = expf(-(float)logT);
      const float w = clipf(-expm1f(-area) * e2, 0.f, 1.f);
      logT = fmax((double)area + logT, 0.0);
      C0 += w * rn0;
      C1 += w * rn1;
      C2 += w * rn2;
      const float m = ladder((float)((t + t_state) * (0.5 * kPreMulti)));
      dpx += 2.f * w * m * csx;
      dpy += 2.f * w * csy;
      csx += w;
      csy += w * m;
      D0 += (double)hgt;
      D1 += (double)(hgt * Dr.y);
      D2 += (double)(hgt * Dr.z);
      D3 += (double)(hgt * Dr.w);
      t_state = t;
      cur_t = tk;
      cursor = head;
      if (A.tri) A.tri[r + (int64_t)iter * A.nr] = (int32_t)(2u * orig + (entry ? 1u : 0u));
      if (A.touch) atomicAdd(A.touch + orig, 1);
      ++iter;
      if (!(logT < (double)kLogCutoff && iter < A.max_iters)) break;
    }
    if (end) break;
  }

  // extract_color (spline-machine.slang:146-153) + distortion assembly (fast_ellipsoid_splinetracer.py:81-85)
  float* oc = A.color + 5 * r;
  oc[0] = C0;
  oc[1] = C1;
  oc[2] = C2;
  oc[3] = (float)(-expm1(-logT));
  oc[4] = dpx - dpy;
  float4* os = reinterpret_cast<float4*>(A.states + 16 * r);
  os[0] = make_float4(dpx, dpy, csx, csy);
  os[1] = make_float4(0.f, 0.f, 0.f, (float)t_state);
  os[2] = make_float4((float)D0, (float)D1, (float)D2, (float)D3);
  os[3] = make_float4((float)logT, C0, C1, C2);
  A.iters[r] = iter;
  double* s64 = A.st64 + 6 * r;
  s64[0] = D0;
  s64[1] = D1;
  s64[2] = D2;
  s64[3] = D3;
  s64[4] = logT;
  s64[5] = t_state;
}

hipError_t launch_forward(const ever_prims& P, void* ws, const ever_rays& R, const ever_forward_out& O,
                          hipStream_t st) {
  SceneLayout L = scene_layout(P.num_prims);
  uint32_t* header = at<uint32_t>(ws, L.header);
  int32_t* cand = at<int32_t>(ws, L.cand);
  hipLaunchKernelGGL(k_reset_count, dim3(1), dim3(1), 0, st, header);
  if (O.touch_count && P.num_prims > 0)
    hipLaunchKernelGGL(k_zero_i32, dim3(grid_for(P.num_prims, 256)), dim3(256), 0, st, O.touch_count, P.num_prims);
  if (R.num_rays <= 0) return hipGetLastError();
  if (P.num_prims > 0)
    hipLaunchKernelGGL(k_prefilter, dim3(grid_for(P.num_prims, 256)), dim3(256), 0, st, P, R.origins, R.tmin, cand,
                       header);
  // processing order
  ForwardLayout FL = forward_layout(R.num_rays);
  uint32_t* k0 = at<uint32_t>(O.workspace, FL.keys0);
  uint32_t* v0 = at<uint32_t>(O.workspace, FL.vals0);
  if (R.num_rays >= kMinRaysToSort && P.num_prims > 0) {
    hipLaunchKernelGGL(k_ray_keys, dim3(grid_for(R.num_rays, 256)), dim3(256), 0, st, R.origins, R.directions,
                       R.num_rays, R.tmin, header, k0, v0);
    hipError_t e = launch_sort_pairs(k0, v0, at<uint32_t>(O.workspace, FL.keys1), at<uint32_t>(O.workspace, FL.vals1),
                                     at<uint32_t>(O.workspace, FL.hist), R.num_rays, 30, st);
    if (e != hipSuccess) return e;
  } else {
    hipLaunchKernelGGL(k_iota, dim3(grid_for(R.num_rays, 256)), dim3(256), 0, st, v0, R.num_rays);
  }
  hipLaunchKernelGGL(k_copy_u32, dim3(grid_for(R.num_rays, 256)), dim3(256), 0, st, v0, O.ray_order, R.num_rays);
  FwdArgs A;
  A.rec = at<float4>(ws, L.rec);
  A.nodes4 = at<float4>(ws, L.nodes4);
  A.colr = at<float4>(ws, L.colr);
  A.cand = cand;
  A.header = header;
  A.means = P.means;
  A.scales = P.scales;
  A.quats = P.quats;
  A.dens = P.densities;
  A.feat = P.features;
  A.color_mode = P.color_mode;
  A.n = P.num_prims;
  A.rayo = R.origins;
  A.rayd = R.directions;
  A.nr = R.num_rays;
  A.tmin = R.tmin;
  A.tmax = R.tmax;
  A.max_iters = R.max_iters;
  A.color = O.color;
  A.states = O.states;
  A.iters = O.iters;
  A.tri = O.tri_collection;
  A.touch = O.touch_count;
  A.init_drgb = O.initial_drgb;
  A.st64 = O.ray_state64;
  A.spill = at<int32_t>(O.workspace, FL.spill);
  A.order = v0;
  dim3 grid(grid_for(R.num_rays, kFwdBlock));
  if (P.num_prims <= 0) {
    // no primitive: every ray sees no event; run the kernel with an empty loop by setting max_iters = 0
    A.max_iters = 0;
  }
  switch (R.kbuf) {
    case 8: hipLaunchKernelGGL(k_trace_forward<8>, grid, dim3(kFwdBlock), 0, st, A); break;
    case 32: hipLaunchKernelGGL(k_trace_forward<32>, grid, dim3(kFwdBlock), 0, st, A); break;
    default: hipLaunchKernelGGL(k_trace_forward<16>, grid, dim3(kFwdBlock), 0, st, A); break;
  }
  return hipGetLastError();
}

}  // namespace ever
