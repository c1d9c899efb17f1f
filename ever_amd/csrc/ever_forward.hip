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
#include <algorithm>

#include "ever_common.hpp"
#include "ever_internal.hpp"

namespace ever {

constexpr int kFwdBlock = 64;
constexpr int kLdsDepth = kLdsStackPerRay;  // LDS stack entries per lane
constexpr int kSpillDepth = kSpillPerRay;    // global spill entries per ray
constexpr int kWaveBlock = 64;   // one wave per block in the per-wave traversal
constexpr int64_t kXcdChunk = 16;  // consecutive 64-ray tiles given to one XCD (per-wave traversal)

// Waves per SIMD that the per-wave walk is compiled for (the VGPR budget: 128 at 4 waves). The block's LDS (K events
// of 8 B and a 4-B stack row per lane) admits 18 blocks per CU at K = 16 and 16 at K = 19.
template <int K>
constexpr int wave_walk_waves() { return 4; }

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
  float tmax_lo, tmax_hi;  // tmax widened by the gathering band (node culling, leaf tests)
  int max_iters;
  float* color;
  float* states;
  int32_t* iters;
  int32_t* tri;
  int32_t* touch;
  float* init_drgb;
  double* st64;
  float* last_dirac;  // (R,4), nullable
  int32_t* spill;          // per-ray walk: kSpillPerRay stack entries per ray
  char* ist;               // integrator state per processing slot (IntState, 112 B per slot, 16-B aligned)
  const uint32_t* order;  // processing order: thread slot -> ray index
  const float* proposal;  // per-ray proposed termination depth, nullable
  const float4* vcol;         // (V,N) per-view (density, colour) when several views, else null
  const int32_t* view_index;  // (R) view of each ray when several views
  int32_t tile;               // persistent launch: 64-ray tiles fetched at a time (0: one wave per block)
  int32_t events_by_slot;     // tri column = processing slot instead of ray index
  uint32_t* tile_counter;     // persistent launch: next tile
};

// kern_prefilter (initialize_density.cu:69-105): primitives whose AABB (create_aabbs.cu:77-114) lies within tmin
// of ray 0's origin. With several views (each one camera of the reference), within tmin of any view's origin, with
// the mask of those views stored after the list (cand + N, 64-bit): a ray's start-inside density then comes from
// the candidates of its own view, as in a call of its camera alone.
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
  float lo[3], hi[3];
  for (int k = 0; k < 3; ++k) {
    const float e = sqrtf(Rm[k][0] * Rm[k][0] * sv[0] * sv[0] + Rm[k][1] * Rm[k][1] * sv[1] * sv[1] +
                          Rm[k][2] * Rm[k][2] * sv[2] * sv[2]) * (1.f + 1e-6f);
    const float m = P.means[3 * i + k];
    lo[k] = m - e;
    hi[k] = m + e;
  }
  auto near = [&](const float* o) {
    float dmin = 0.f;
    for (int k = 0; k < 3; ++k) {
      if (o[k] < lo[k]) dmin += (o[k] - lo[k]) * (o[k] - lo[k]);
      else if (o[k] > hi[k]) dmin += (o[k] - hi[k]) * (o[k] - hi[k]);
    }
    return dmin <= tmin * tmin;
  };
  if (P.num_views > 1) {
    uint64_t mask = 0;
    for (int v = 0; v < P.num_views; ++v)
      if (near(P.view_origins + 3 * v)) mask |= uint64_t(1) << v;
    if (mask != 0) {
      const uint32_t pos = atomicAdd(header + 8, 1u);
      cand[pos] = (int32_t)i;
      reinterpret_cast<uint64_t*>(cand + cand_mask_offset(P.num_prims))[pos] = mask;
    }
    return;
  }
  if (near(rayo)) {
    const uint32_t pos = atomicAdd(header + 8, 1u);
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

__device__ __forceinline__ uint32_t expand14_2d(uint32_t v) {  // 14 bits -> every other bit of 28
  v &= 0x3fffu;
  v = (v | (v << 8)) & 0x00ff00ffu;
  v = (v | (v << 4)) & 0x0f0f0f0fu;
  v = (v | (v << 2)) & 0x33333333u;
  v = (v | (v << 1)) & 0x55555555u;
  return v;
}

// The reference camera of a ray batch for the sort keys: ray 0's origin o0, looking at the scene centre c (centre of
// the primitive centroids), an orthonormal frame (u, v, m) with m along c - o0, and a ray's gnomonic coordinates
// (dot(d, u), dot(d, v)) / dot(d, m) on the plane normal to m: for a pinhole camera at o0 an image of its pixel grid.
struct KeyFrame {
  f3 o0, u, v, m;
  bool ok;  // the frame exists (o0 is not at the centre)
};
__device__ __forceinline__ KeyFrame key_frame(f3 o0, f3 c, float ext) {
  KeyFrame F;
  F.o0 = o0;
  const f3 m = c - F.o0;
  const float mn = sqrtf(dot3(m, m));
  F.ok = mn > 1e-6f * ext;
  F.m = F.ok ? (1.f / mn) * m : mk3(0.f, 0.f, 1.f);
  // u from the coordinate axis least aligned with m
  const f3 ax = fabsf(F.m.x) <= fabsf(F.m.y) && fabsf(F.m.x) <= fabsf(F.m.z) ? mk3(1.f, 0.f, 0.f)
                : (fabsf(F.m.y) <= fabsf(F.m.z) ? mk3(0.f, 1.f, 0.f) : mk3(0.f, 0.f, 1.f));
  F.u = l2n(cross3(ax, F.m));
  F.v = cross3(F.m, F.u);
  return F;
}
// gnomonic coordinates of a ray of the reference camera; false when the ray starts elsewhere or points more than
// acos(0.25) away from m
__device__ __forceinline__ bool key_plane(const KeyFrame& F, f3 ro, f3 d, float& a, float& b) {
  const float dm = dot3(d, F.m);
  if (!(F.ok && ro.x == F.o0.x && ro.y == F.o0.y && ro.z == F.o0.z && dm > 0.25f)) return false;
  a = dot3(d, F.u) / dm;
  b = dot3(d, F.v) / dm;
  return true;
}

// Sort key of a ray (30 bits). A ray of the reference camera (above) is keyed by the 2D Morton code of its gnomonic
// coordinates on a grid whose cell is the spacing between the camera's first two rays (neighbouring pixels of an
// image in scanline order; at least 1/2048 of the plane's +-4 range): 64 consecutive keys then form a square screen
// tile of 8 x 8 pixels (measured at 1920x1080: forward 8.60 -> 7.11 ms against the 3D key below). With one camera the
// reference is ray 0's origin and the code has 14 bits per axis; with several views (view_index) each view is its own
// camera (its origin, its first two rays), the view's index leads the key and the code has (29 - view bits) / 2 bits
// per axis. Other rays (other origins, wide angles) follow, keyed by the 3D Morton code (29 bits) of the point
// o + d * |c - o|, quantized over a cube of four times the scene extent around c, so that rays from different origins
// separate. Any order gives the same results.
// first ray of each view's run of rays (the last run's, if a view's rays are not contiguous)
__global__ __launch_bounds__(256) void k_view_first(const int32_t* __restrict__ view_index, int64_t nr, int32_t nv,
                                                    int32_t* __restrict__ first) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nr) return;
  const int32_t v = view_index[r];
  if (v >= 0 && v < nv && (r == 0 || view_index[r - 1] != v)) first[v] = (int32_t)r;
}

__global__ __launch_bounds__(256) void k_ray_keys(const float* __restrict__ rayo, const float* __restrict__ rayd,
                                                  int64_t nr, float tmin, const uint32_t* __restrict__ header,
                                                  const int32_t* __restrict__ view_index,
                                                  const float* __restrict__ view_origins, int32_t nv,
                                                  const int32_t* __restrict__ view_first,
                                                  uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nr) return;
  auto o2f = [](uint32_t u) { return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u); };
  const f3 lo = mk3(o2f(header[0]), o2f(header[1]), o2f(header[2]));
  const f3 hi = mk3(o2f(header[3]), o2f(header[4]), o2f(header[5]));
  const f3 c = 0.5f * (lo + hi);
  const float ext = fmaxf(fmaxf(fmaxf(hi.x - lo.x, hi.y - lo.y), hi.z - lo.z), 1e-6f);
  const f3 rd = ld3(rayd + 3 * r);
  const f3 ro = ld3(rayo + 3 * r);
  const f3 d = l2n(rd);
  // the ray's camera: its view's origin and first ray, or ray 0
  const int32_t v = nv > 1 ? view_index[r] : 0;
  const bool views = nv > 1 && v >= 0 && v < nv;
  const int64_t r0 = views ? (int64_t)view_first[v] : 0;
  const KeyFrame F = key_frame(views ? ld3(view_origins + 3 * (int64_t)v) : ld3(rayo), c, ext);
  // view bits leading the key, and bits per axis of the 2D code
  int vb = 0;
  while (views && vb < 8 && (1 << vb) < nv) ++vb;
  const int axb = (29 - vb) / 2 > 14 ? 14 : (29 - vb) / 2;
  float a, b;
  if (key_plane(F, ro, d, a, b)) {
    // grid cell: the spacing of the camera's first two rays on the plane, within [8 / 16384, 8 / 256]
    float cell = 8.f / 16384.f;
    float a0, b0, a1, b1;
    if (r0 + 1 < nr && (!views || view_index[r0 + 1] == v) &&
        key_plane(F, ld3(rayo + 3 * r0), l2n(ld3(rayd + 3 * r0)), a0, b0) &&
        key_plane(F, ld3(rayo + 3 * (r0 + 1)), l2n(ld3(rayd + 3 * (r0 + 1))), a1, b1))
      cell = fminf(fmaxf(sqrtf((a1 - a0) * (a1 - a0) + (b1 - b0) * (b1 - b0)), 8.f / 16384.f), 8.f / 256.f);
    const float inv = 1.f / cell;
    const float half = (float)(1 << (axb - 1)), top = (float)((1 << axb) - 1);
    const uint32_t qa = (uint32_t)fminf(fmaxf(a * inv + half, 0.f), top);
    const uint32_t qb = (uint32_t)fminf(fmaxf(b * inv + half, 0.f), top);
    const uint32_t code = (expand14_2d(qb) << 1) | expand14_2d(qa);  // 2 * axb bits
    keys[r] = views ? ((uint32_t)(v & ((1 << vb) - 1)) << (2 * axb)) | code : code;
  } else {
    const f3 o = ro + tmin * rd;
    const f3 co = c - o;
    const f3 p = o + sqrtf(dot3(co, co)) * d;
    const float inv = 1.f / (4.f * ext);
    const uint32_t qx = (uint32_t)fminf(fmaxf(((p.x - c.x) * inv + 0.5f) * 1024.f, 0.f), 1023.f);
    const uint32_t qy = (uint32_t)fminf(fmaxf(((p.y - c.y) * inv + 0.5f) * 1024.f, 0.f), 1023.f);
    const uint32_t qz = (uint32_t)fminf(fmaxf(((p.z - c.z) * inv + 0.5f) * 1024.f, 0.f), 1023.f);
    keys[r] = (1u << 29) | (((expand10(qx) << 2) | (expand10(qy) << 1) | expand10(qz)) >> 1);
  }
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

constexpr uint64_t kNoEvent = ((uint64_t)0x7f800000u << 32) | 0xffffffffu;  // pack_event(kEmptyT, ~0)

// Inserts up to two events without branches: an event that must not enter is passed as kNoEvent, which equals
// the empty slots and moves nothing. Same result as kbuf_insert of each event in turn.
template <int K>
__device__ __forceinline__ void kbuf_insert_pred(uint64_t (&bk)[K], uint64_t v) {
  bool dup = false;
#pragma unroll
  for (int k = 0; k < K; ++k) dup |= (bk[k] == v);
  v = dup ? kNoEvent : v;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const bool sw = v < bk[k];
    const uint64_t old = bk[k];
    bk[k] = sw ? v : old;
    v = sw ? old : v;
  }
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

// Grazing incidence. The record rows M (fp32 products of the rotation and 1/scale) and b = -M mean carry a relative
// error of a few float32 ulps, so 1 - |l|^2 evaluated from them is off by up to about eps32 * |M o + b| terms, i.e.
// eps32 * (|ocn| + |b|). Where 1 - |l|^2 lies within a band of kGrazeUlps such units of zero, whether the ray meets
// the ellipsoid and the length of its chord are taken from the raw parameters in FP64, as the FP64 oracle does; the
// chord of a grazing hit is short but its derivatives grow as 1 / sqrt(1 - |l|^2), so both must match.
constexpr float kGrazeUlps = 64.f * 5.9604645e-8f;
constexpr float kGraze2 = kGrazeUlps * kGrazeUlps;

// The ray in FP64 (ray_origin_fp64, ray_dir_fp64).
__device__ __forceinline__ void ray_fp64(const FwdArgs& A, int64_t r, d3& o, d3& d) {
  const f3 ro = ld3(A.rayo + 3 * r), rdraw = ld3(A.rayd + 3 * r);
  o = ray_origin_fp64(ro, rdraw, A.tmin);
  d = ray_dir_fp64(rdraw);
}

// Roots of primitive s in FP64 from its raw parameters (raw_roots_d).
__device__ __forceinline__ bool raw_roots_fp64(const FwdArgs& A, uint32_t s, d3 o, d3 d, double& t0, double& t1) {
  return raw_roots_d(o, d, reinterpret_cast<const float4*>(A.quats)[s], ld3(A.scales + 3 * (int64_t)s),
                     ld3(A.means + 3 * (int64_t)s), t0, t1);
}

// Per-ray state of the traversal: the ray in fp32 (box and leaf tests), the (t, key) cursor and what decides which
// events a traversal gathers. The integrator's state (IntState) is kept in the forward workspace between batches, so
// that the traversal loops carry none of it in registers.
struct RayState {
  int32_t r;     // ray index
  int32_t e;     // column of the ray's events in tri: the ray index, or its processing slot (events by slot)
  int32_t slot;  // processing slot (the ray's IntState)
  uint64_t cursor;  // (t, key) of the last event taken from a buffer (integrated, skipped or held); t = cur_t(S)
  float bound;      // events beyond this depth are left for a later (plain) traversal; +inf once plain
  // float32 depth of the last event of the last completed 16-event round of the reference (the start of its current
  // trace), lowered by the gathering band: exits beyond tmax are gathered when their entry lies after it
  float tref;
  bool live;  // the ray has not terminated
};

// The ray in float32 for the box and leaf tests, rebuilt from the inputs at the start of each traversal so that it
// occupies no registers while the integrator runs: origin shifted by tmin along the raw direction
// (shaders.slang:178-179), unit direction, its reciprocal and origin * reciprocal (slab distances as
// fma(plane, inv, -o * inv)).
struct RayGeo {
  f3 o, d, inv, oi;
};
__device__ __forceinline__ RayGeo ray_geo(const FwdArgs& A, int64_t r) {
  RayGeo G;
  const f3 ro = ld3(A.rayo + 3 * r), rdraw = ld3(A.rayd + 3 * r);
  G.d = l2n(rdraw);
  G.o = ro + A.tmin * rdraw;
  G.inv = mk3(fabsf(G.d.x) > 1e-20f ? 1.f / G.d.x : copysignf(1e20f, G.d.x),
              fabsf(G.d.y) > 1e-20f ? 1.f / G.d.y : copysignf(1e20f, G.d.y),
              fabsf(G.d.z) > 1e-20f ? 1.f / G.d.z : copysignf(1e20f, G.d.z));
  G.oi = mk3(G.o.x * G.inv.x, G.o.y * G.inv.y, G.o.z * G.inv.z);
  return G;
}

// Integrator state (the reference's SplineState, structs.h:41-54, with FP64 density/colour accumulators and
// log-transmittance), one 112-byte record per processing slot in the forward workspace, moved as seven 16-byte
// words from one base address: (D0, D1) (D2, D3) (t_state, logT) (ts_state, logT_s) (tref64, hold)
// (C0, C1, C2, csx) (csy, dpx, dpy, iter).
struct IntState {
  double D0, D1, D2, D3, t_state, logT;
  // the same accumulation with the backward's safe roots (rotated direction l2-normalized, safe_roots_jac_d): the
  // log-transmittance the adjoint starts from and unwinds event by event with exactly these depths
  double ts_state, logT_s;
  double tref64;  // FP64 depth of the last event of the last completed 16-event round
  float C0, C1, C2, csx, csy, dpx, dpy;
  int32_t iter;
  uint64_t hold;  // an event taken from the previous buffer but not yet integrated (0: none)
};
constexpr int64_t kIstBytes = 112;

__device__ __forceinline__ void ist_load(const FwdArgs& A, int64_t slot, IntState& I) {
  const double2* d = reinterpret_cast<const double2*>(A.ist + kIstBytes * slot);
  const double2 a = d[0], b = d[1], c = d[2], e = d[3], g = d[4];
  const float4 f = reinterpret_cast<const float4*>(d)[5], h = reinterpret_cast<const float4*>(d)[6];
  I.D0 = a.x; I.D1 = a.y; I.D2 = b.x; I.D3 = b.y; I.t_state = c.x; I.logT = c.y; I.ts_state = e.x; I.logT_s = e.y;
  I.tref64 = g.x; I.hold = (uint64_t)__double_as_longlong(g.y);
  I.C0 = f.x; I.C1 = f.y; I.C2 = f.z; I.csx = f.w; I.csy = h.x; I.dpx = h.y; I.dpy = h.z; I.iter = __float_as_int(h.w);
}
__device__ __forceinline__ void ist_store(const FwdArgs& A, int64_t slot, const IntState& I) {
  double2* d = reinterpret_cast<double2*>(A.ist + kIstBytes * slot);
  d[0] = make_double2(I.D0, I.D1);
  d[1] = make_double2(I.D2, I.D3);
  d[2] = make_double2(I.t_state, I.logT);
  d[3] = make_double2(I.ts_state, I.logT_s);
  d[4] = make_double2(I.tref64, __longlong_as_double((long long)I.hold));
  reinterpret_cast<float4*>(d)[5] = make_float4(I.C0, I.C1, I.C2, I.csx);
  reinterpret_cast<float4*>(d)[6] = make_float4(I.csy, I.dpx, I.dpy, __int_as_float(I.iter));
}

__device__ __forceinline__ bool int_alive(const FwdArgs& A, const IntState& I) {
  return I.logT < (double)kLogCutoff && I.iter < A.max_iters;
}

// Gathering band around tmax and the round boundary: the float32 leaf depths gather every event whose eligibility
// could hold in FP64 (their error is far below this band), and the integrator decides it from the FP64 depths.
constexpr float kBand = 1e-4f;
__device__ __forceinline__ float band_lo(float t) { return t - kBand * (fabsf(t) + 1.f); }
__device__ __forceinline__ float band_hi(float t) { return t + kBand * (fabsf(t) + 1.f); }

// depth of the cursor (the float32 depth of the last event taken)
__device__ __forceinline__ float cur_t(const RayState& S) { return __uint_as_float((uint32_t)(S.cursor >> 32)); }

// A lane of the per-wave walk with no ray: it stays in the wave (the wave's stack keeps entry i in lane i, and a
// lane that has left the wave would drop the entries written to it) and is never live, so it never enters a box.
__device__ __forceinline__ void ray_idle(const FwdArgs& A, RayState& S) {
  S.r = 0;
  S.e = 0;
  S.slot = 0;
  S.cursor = 0;
  S.bound = INFINITY;
  S.tref = 0.f;
  S.live = false;
}

template <bool kViews>
__device__ __forceinline__ void ray_setup(const FwdArgs& A, int64_t r, int64_t slot, RayState& S) {
  S.r = (int32_t)r;
  S.e = (int32_t)(A.events_by_slot ? slot : r);
  S.slot = (int32_t)slot;
  const f3 ro = ld3(A.rayo + 3 * r);
  // initial density / colour of ellipsoids containing rayo + tmin * d (initialize_density.cu:107-150), decided in FP64
  // (a start point on a surface is placed as the oracle places it)
  double D0 = 0.0, D1 = 0.0, D2 = 0.0, D3 = 0.0;
  const int32_t ncand = (int32_t)A.header[8];
  const d3 rdd = tod(ld3(A.rayd + 3 * r));
  const d3 p = tod(ro) + ((double)A.tmin / sqrt(dotd(rdd, rdd))) * rdd;  // (FP64, as the oracle)
  const int32_t view = kViews ? A.view_index[r] : 0;
  for (int32_t j = 0; j < ncand; ++j) {
    const int64_t i = A.cand[j];
    // several views: only the candidates of this ray's view
    if (kViews && ((reinterpret_cast<const uint64_t*>(A.cand + cand_mask_offset(A.n))[j] >> view) & 1u) == 0) continue;
    if (inside_d(p, ld3(A.means + 3 * i), ld3(A.scales + 3 * i), reinterpret_cast<const float4*>(A.quats)[i], false)) {
      float den = A.dens[i];
      float c0, c1, c2;
      if (A.color_mode == EVER_COLOR_SH) {
        const float4 cr = kViews ? A.vcol[(int64_t)view * A.n + i] : A.colr[i];
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
  reinterpret_cast<float4*>(A.init_drgb)[r] = make_float4((float)D0, (float)D1, (float)D2, (float)D3);
  IntState I;
  I.D0 = D0;
  I.D1 = D1;
  I.D2 = D2;
  I.D3 = D3;
  I.t_state = I.logT = I.ts_state = I.logT_s = I.tref64 = 0.0;
  I.C0 = I.C1 = I.C2 = I.csx = I.csy = I.dpx = I.dpy = 0.f;
  I.iter = 0;
  I.hold = 0;
  ist_store(A, slot, I);
  // cursor over (t, key): below every event (events at float32 depths down to the start band are packed at t = 0
  // and decided from their FP64 depths, which the reference requires to be > t_cur = 0; see leaf_events)
  S.cursor = 0;
  S.bound = INFINITY;
  S.tref = band_lo(0.f);
  S.live = A.max_iters > 0;
}

// Depth at which the ray leaves the union of the root record's child boxes (+inf when the root is a leaf): a
// proposal beyond it leaves no event for a later traversal.
__device__ __forceinline__ float scene_exit(const FwdArgs& A, const RayGeo& S) {
  if (A.header[9] <= 1) return INFINITY;
  const float4* nd = A.nodes4;
  const float4 cf = nd[6];
  const bool used[4] = {__float_as_int(cf.x) != kEmptyChild, __float_as_int(cf.y) != kEmptyChild,
                        __float_as_int(cf.z) != kEmptyChild, __float_as_int(cf.w) != kEmptyChild};
  const float iv[3] = {S.inv.x, S.inv.y, S.inv.z}, oiv[3] = {S.oi.x, S.oi.y, S.oi.z};
  float t = INFINITY;
  for (int a = 0; a < 3; ++a) {
    const float4 l = nd[2 * a], h = nd[2 * a + 1];
    float lo = INFINITY, hi = -INFINITY;
    lo = used[0] ? fminf(lo, l.x) : lo; hi = used[0] ? fmaxf(hi, h.x) : hi;
    lo = used[1] ? fminf(lo, l.y) : lo; hi = used[1] ? fmaxf(hi, h.y) : hi;
    lo = used[2] ? fminf(lo, l.z) : lo; hi = used[2] ? fmaxf(hi, h.z) : hi;
    lo = used[3] ? fminf(lo, l.w) : lo; hi = used[3] ? fmaxf(hi, h.w) : hi;
    t = fminf(t, fmaxf(fmaf(lo, iv[a], -oiv[a]), fmaf(hi, iv[a], -oiv[a])));
  }
  return t;
}

// the ray's proposal, if any (NaN or negative proposals are ignored)
__device__ __forceinline__ void ray_bound(const FwdArgs& A, RayState& S) {
  const float b = A.proposal[S.r];
  S.bound = b >= 0.f ? b : INFINITY;
}

// The roots are written with explicit fused multiply-adds and no other contraction, so both tree walks (records
// in vector or in scalar registers) produce bit-identical depths and therefore the same event order.
__device__ __forceinline__ float dot3_fma(f3 a, f3 b) { return fmaf(a.z, b.z, fmaf(a.y, b.y, a.x * b.x)); }

// Leaf test of one primitive: ellipsoid roots (tri-intersect.slang:268-283) in the record's frame, then the
// intersection program (shaders.slang:96-106) and the any-hit choice of which events enter the k-buffer (:138-159),
// with the (t, key) cursor in place of t > t_cur. Sets e (entry) and x (exit) to the events to insert, kNoEvent
// where none. Entry and exit go in together when the entry lies after the cursor (and within tmax); otherwise the
// exit alone when it lies after the cursor and within tmax, or beyond tmax with its entry in the current round.
// The float32 depths decide within a band around tmax and the round's start: the integrator takes the final
// decision from the FP64 depths (integrate_seq).
template <bool kBounded>
__device__ __forceinline__ void leaf_events(const FwdArgs& A, const RayState& S, const RayGeo& G, float4 RA,
                                            float4 RB, float4 RC, uint32_t s, uint64_t& e, uint64_t& x) {
#pragma clang fp contract(off)
  const f3 o = G.o, d = G.d;
  const f3 ocn = mk3(fmaf(RA.z, o.z, fmaf(RA.y, o.y, fmaf(RA.x, o.x, RA.w))),
                     fmaf(RB.z, o.z, fmaf(RB.y, o.y, fmaf(RB.x, o.x, RB.w))),
                     fmaf(RC.z, o.z, fmaf(RC.y, o.y, fmaf(RC.x, o.x, RC.w))));
  const f3 rdn = mk3(fmaf(RA.z, d.z, fmaf(RA.y, d.y, RA.x * d.x)), fmaf(RB.z, d.z, fmaf(RB.y, d.y, RB.x * d.x)),
                     fmaf(RC.z, d.z, fmaf(RC.y, d.y, RC.x * d.x)));
  const float aa = dot3_fma(rdn, rdn);
  const float bp = -dot3_fma(ocn, rdn);
  // hardware reciprocal and square root (each within 1 ulp) instead of the IEEE sequences: these float roots
  // only order the events and place the cursor (the integrator recomputes every depth in FP64), and both tree
  // walks share this function, so their event lists stay identical
  const float ia = __builtin_amdgcn_rcpf(aa);
  const float ba = bp * ia;
  const f3 l = mk3(fmaf(ba, rdn.x, ocn.x), fmaf(ba, rdn.y, ocn.y), fmaf(ba, rdn.z, ocn.z));
  const float omll = 1.f - dot3_fma(l, l);
  const float h = aa * omll;
  const float cc = dot3_fma(ocn, ocn) - 1.f;
  const float qq = fmaf(sgnf(bp), __builtin_amdgcn_sqrtf(fmaxf(h, 0.f)), bp);
  const float r0 = cc * __builtin_amdgcn_rcpf(qq), r1 = qq * ia;
  const float t0 = fminf(r0, r1), t1 = fmaxf(r0, r1);
  // at the ray's start (cursor depth 0) events down to the gathering band below it are gathered too, packed at depth
  // 0: the float32 root of a surface through the start point can fall on either side of it, and the integrator
  // decides from the FP64 depth, which the reference requires to be > 0
  const float lo_t = cur_t(S) > 0.f ? cur_t(S) : band_lo(0.f);
  const uint64_t ve = pack_event(fmaxf(t0, 0.f), 2u * s), vx = pack_event(fmaxf(t1, 0.f), 2u * s + 1u);
  // a miss within the grazing band still enters the buffer (with the roots of 1 - |l|^2 clamped at 0): the
  // integrator decides such events from the raw parameters in FP64 and skips the pair if the ray misses
  const float bb = fmaf(RC.w, RC.w, fmaf(RB.w, RB.w, RA.w * RA.w));
  const bool meets = h >= 0.f || omll * omll < kGraze2 * (cc + 1.f + 4.f * bb);
  const bool hit = meets && !(t1 < lo_t);
  const bool with_entry = t0 >= lo_t && ve > S.cursor;
  const float tmax = A.tmax_hi;
  const bool ins_e = hit && with_entry && t0 <= tmax;
  // an exit beyond tmax enters only with its entry: in the reference's trace of the current round (from tref) the
  // entry is reported with the exit as its attribute; an entry before tref was reported in an earlier round
  const bool ins_x =
      hit && (with_entry ? t0 <= tmax : (vx > S.cursor && (t1 <= tmax || (t0 > S.tref && t0 <= tmax))));
  if (kBounded) {
    // under a proposal only events up to it are gathered, so that every gathered event precedes every culled one
    e = ins_e && t0 <= S.bound ? ve : kNoEvent;
    x = ins_x && t1 <= S.bound ? vx : kNoEvent;
  } else {
    e = ins_e ? ve : kNoEvent;
    x = ins_x ? vx : kNoEvent;
  }
}

// The control point of an integrated event (get_ctrl_pt, shaders.slang:65-82): +-density * (1, colour).
template <bool kViews>
__device__ __forceinline__ void write_last_dirac(const FwdArgs& A, const RayState& S, uint32_t key) {
  const uint32_t s = key >> 1;
  const float4 Dr = kViews ? A.vcol[(int64_t)A.view_index[S.r] * A.n + s] : A.rec[4 * (int64_t)s + 3];
  const float h = (key & 1u) == 0u ? Dr.x : -Dr.x;
  reinterpret_cast<float4*>(A.last_dirac)[S.r] = make_float4(h, h * Dr.y, h * Dr.z, h * Dr.w);
}

// Integrates the gathered events head_at(0), head_at(1), ... (each fetched once, in buffer order) as the reference's
// raygen loop does (shaders.slang:210-226); returns true when the buffer held fewer than K events, i.e. no event
// is left up to the bound of this traversal.
// Depths, order and eligibility are those of FP64, from the raw parameters, as in the oracle:
// - the buffer is ordered by the float32 depths of the leaf test; one event is held back, across buffers too
//   (IntState.hold), so that where two neighbours tie within float32 resolution and their FP64 depths order them
//   the other way, the later-fetched one is integrated first (the clipped step max(t - t_prev, 0) makes the
//   gradients depend on that order);
// - a grazing miss in FP64 is skipped; an entry beyond tmax is skipped; an exit beyond tmax counts only while its
//   entry lies after the last event of the reference's last completed 16-event round (BUFFER_SIZE,
//   shaders.slang:18), i.e. when the reference's trace of the current round reports it with its entry, which
//   makes results independent of K.
template <int K, bool kBounded, bool kViews, typename Head>
__device__ __forceinline__ bool integrate_seq(const FwdArgs& A, RayState& S, Head head_at) {
  IntState I;
  ist_load(A, S.slot, I);
  d3 od, dd;  // the ray in FP64
  ray_fp64(A, S.r, od, dd);
  uint32_t last = 0xffffffffu;  // key of the last event integrated in this batch
  uint64_t ph = 0;              // the held-back event, its FP64 depth and its round condition (below)
  double pt = 0.0, pt0 = 0.0;
  bool held = false;
  uint64_t carried = I.hold;  // the previous buffer's held event is fetched first
  I.hold = 0;
  bool end = false;  // the buffer's end marker was fetched
  int k = 0;         // next buffer entry to fetch
  bool ret = false;
  while (true) {
    // fetch the next eligible event: its FP64 depth, and for an exit beyond tmax the depth of its entry, which
    // must lie in the current round when the exit is integrated (-inf: no condition)
    uint64_t nh = 0;
    double nt = 0.0, nt0 = 0.0;
    bool got = false;
    while (carried != 0 || (!end && k < K)) {
      uint64_t head = carried;
      carried = 0;
      if (head == 0) {
        head = head_at(k++);
        if (__uint_as_float((uint32_t)(head >> 32)) > kEndT) {
          end = true;
          break;
        }
      }
      S.cursor = S.cursor > head ? S.cursor : head;
      double t0, t1;
      if (!raw_roots_fp64(A, (uint32_t)head >> 1, od, dd, t0, t1)) continue;  // a grazing miss
      if (t0 > (double)A.tmax) continue;  // an entry beyond tmax, or an exit whose entry is beyond it
      const bool entry = (head & 1u) == 0u;
      nt = entry ? t0 : t1;
      // the reference takes only events after the last one it processed (t > t_cur, shaders.slang:146; t_cur = 0 at
      // the start): one gathered from the start band whose FP64 depth is not > 0 is not an event of its trace
      if (!(nt > 0.0)) continue;
      nt0 = (entry || t1 <= (double)A.tmax) ? -INFINITY : t0;
      nh = head;
      got = true;
      break;
    }
    if (!got) {
      if (!held) {
        ret = end;
        break;
      }
      // a full buffer, or the end of the events up to the bound of a ray that continues beyond its proposal: the
      // held event waits for the next traversal, whose first event may precede it in FP64
      if (!end || (kBounded && S.bound != INFINITY && S.bound < scene_exit(A, ray_geo(A, S.r)))) {
        I.hold = ph;
        ret = end;
        break;
      }
    } else if (!held) {
      ph = nh;
      pt = nt;
      pt0 = nt0;
      held = true;
      continue;
    }
    // the earlier of the held-back and the fetched event in (FP64 depth, key) order; the other stays held
    const bool take_new = got && (nt < pt || (nt == pt && (uint32_t)nh < (uint32_t)ph));
    const uint64_t head = take_new ? nh : ph;
    const double t = take_new ? nt : pt;
    const double t0r = take_new ? nt0 : pt0;
    if (!take_new) {
      ph = nh;
      pt = nt;
      pt0 = nt0;
      held = got;
    }
    if (t0r != -INFINITY && !(t0r > I.tref64)) continue;  // its entry lies in an earlier round
    const uint32_t key = (uint32_t)head;
    const uint32_t s = key >> 1;
    const bool entry = (key & 1u) == 0u;
    // (density, colour): the record's, or this ray's view's when several views share the batch
    const float4 Dr = kViews ? A.vcol[(int64_t)A.view_index[S.r] * A.n + s] : A.rec[4 * (int64_t)s + 3];
    const uint32_t orig = s;  // records are in original primitive order
    const float hgt = entry ? Dr.x : -Dr.x;
    // update (spline-machine.slang:103-138)
    const float dt = fmaxf((float)(t - I.t_state), 0.f);
    const float Df = (float)I.D0;
    const float area = fmaxf(Df * dt, 0.f);
    {
      // the adjoint's depth of this event and its area, as ever_backward evaluates them: its safe roots measure
      // depth along the rotated direction normalized after the rotation, which the raw quaternion scales, so the
      // safe depth is the raw one times |qrot(d, q)| (the same point of the same line; FP64, within a few ulp)
      const d3 rd = qrot_d(dd, reinterpret_cast<const float4*>(A.quats)[s]);
      const double ts = t * sqrt(dotd(rd, rd));
      const float dts = fmaxf((float)(ts - I.ts_state), 0.f);
      I.logT_s = fmax((double)fmaxf(Df * dts, 0.f) + I.logT_s, 0.0);
      I.ts_state = ts;
    }
    const float rn0 = safe_div((float)I.D1, Df), rn1 = safe_div((float)I.D2, Df), rn2 = safe_div((float)I.D3, Df);
    const float e2 = expf(-(float)I.logT);
    const float w = clipf(-expm1f(-area) * e2, 0.f, 1.f);
    I.logT = fmax((double)area + I.logT, 0.0);
    I.C0 += w * rn0;
    I.C1 += w * rn1;
    I.C2 += w * rn2;
    const float m = ladder((float)((t + I.t_state) * (0.5 * kPreMulti)));
    I.dpx += 2.f * w * m * I.csx;
    I.dpy += 2.f * w * I.csy;
    I.csx += w;
    I.csy += w * m;
    I.D0 += (double)hgt;
    I.D1 += (double)(hgt * Dr.y);
    I.D2 += (double)(hgt * Dr.z);
    I.D3 += (double)(hgt * Dr.w);
    I.t_state = t;
    last = key;
    if (A.tri) A.tri[S.e + (int64_t)I.iter * A.nr] = (int32_t)(2u * orig + (entry ? 1u : 0u));
    if (A.touch) atomicAdd(A.touch + orig, 1);
    ++I.iter;
    if ((I.iter & 15) == 0) {
      I.tref64 = t;
      S.tref = band_lo((float)t);
    }
    if (!int_alive(A, I)) {
      ret = false;
      break;
    }
  }
  ist_store(A, S.slot, I);
  S.live = int_alive(A, I);
  // the caller's last integrated control point (shaders.slang:229-232)
  if (A.last_dirac && last != 0xffffffffu) write_last_dirac<kViews>(A, S, last);
  return ret;
}

template <int K, bool kBounded, bool kViews>
__device__ __forceinline__ bool integrate_events(const FwdArgs& A, RayState& S, uint64_t (&bk)[K]) {
  // the head of the buffer is consumed and the rest shifted down, so the loop body exists once
  return integrate_seq<K, kBounded, kViews>(A, S, [&](int) {
    const uint64_t head = bk[0];
#pragma unroll
    for (int j = 0; j < K - 1; ++j) bk[j] = bk[j + 1];
    return head;
  });
}

// extract_color (spline-machine.slang:146-153) + distortion assembly (fast_ellipsoid_splinetracer.py:81-85)
__device__ __forceinline__ void ray_write(const FwdArgs& A, const RayState& S) {
  const int64_t r = S.r;
  IntState I;
  ist_load(A, S.slot, I);
  float* oc = A.color + 5 * r;
  oc[0] = I.C0;
  oc[1] = I.C1;
  oc[2] = I.C2;
  oc[3] = (float)(-expm1(-I.logT));
  oc[4] = I.dpx - I.dpy;
  float4* os = reinterpret_cast<float4*>(A.states + 16 * r);
  os[0] = make_float4(I.dpx, I.dpy, I.csx, I.csy);
  os[1] = make_float4(0.f, 0.f, 0.f, (float)I.t_state);
  os[2] = make_float4((float)I.D0, (float)I.D1, (float)I.D2, (float)I.D3);
  os[3] = make_float4((float)I.logT, I.C0, I.C1, I.C2);
  A.iters[r] = I.iter;
  double* s64 = A.st64 + 8 * r;
  s64[0] = I.D0;
  s64[1] = I.D1;
  s64[2] = I.D2;
  s64[3] = I.D3;
  s64[4] = I.logT;
  s64[5] = I.t_state;
  s64[6] = I.logT_s;
  s64[7] = I.ts_state;
  // the last integrated control point is written by the integrator; a ray without events has none
  if (A.last_dirac && I.iter == 0) reinterpret_cast<float4*>(A.last_dirac)[r] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// Per-ray traversal: one lane walks the tree for its own ray with a short LDS stack (global spill beyond
// kLdsDepth entries). Used for rays in no particular order, where the rays of a wave share few nodes.
template <int K, int kFlags>  // kFlags: 1 proposal-bounded, 2 several views
__global__ __launch_bounds__(kFwdBlock) void k_trace_forward(FwdArgs A) {
  constexpr bool kBounded = (kFlags & 1) != 0, kViews = (kFlags & 2) != 0;
  __shared__ int32_t stk[kLdsDepth * kFwdBlock];
  const int lane = threadIdx.x;
  const int64_t slot = (int64_t)blockIdx.x * kFwdBlock + lane;
  if (slot >= A.nr) return;
  RayState S;
  ray_setup<kViews>(A, A.order[slot], slot, S);
  if (kBounded) ray_bound(A, S);
  const int64_t nleaves = A.header[9];
  const int32_t root = nleaves == 1 ? ~0 : 0;
  int32_t* spill = A.spill + slot * kSpillDepth;

  while (S.live) {
    const RayGeo G = ray_geo(A, S.r);
    const f3 inv = G.inv, oi = G.oi;
    uint64_t bk[K];
#pragma unroll
    for (int k = 0; k < K; ++k) bk[k] = pack_event(kEmptyT, 0xffffffffu);
    // the lower culling depth is the cursor, held at tmax once the cursor has passed it: exits beyond tmax whose
    // entries lie before it may remain, and their boxes straddle tmax
    const float t_lo = fminf(cur_t(S) > 0.f ? cur_t(S) : band_lo(0.f), A.tmax_lo);  // (the start band: leaf_events)
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
        const float t_hi =
            fminf(kBounded ? fminf(A.tmax_hi, S.bound) : A.tmax_hi, __uint_as_float((uint32_t)(bk[K - 1] >> 32)));
        const int32_t c0 = __float_as_int(cf.x), c1 = __float_as_int(cf.y), c2 = __float_as_int(cf.z),
                      c3 = __float_as_int(cf.w);
        const bool used[4] = {c0 != kEmptyChild, c1 != kEmptyChild, c2 != kEmptyChild, c3 != kEmptyChild};
        const float lxs[4] = {lx.x, lx.y, lx.z, lx.w}, hxs[4] = {hx.x, hx.y, hx.z, hx.w};
        const float lys[4] = {ly.x, ly.y, ly.z, ly.w}, hys[4] = {hy.x, hy.y, hy.z, hy.w};
        const float lzs[4] = {lz.x, lz.y, lz.z, lz.w}, hzs[4] = {hz.x, hz.y, hz.z, hz.w};
        // sort keys: entry distance bits (>= 0, so ordered as unsigned) with the child slot in the low 2 bits;
        // a miss sorts last. Child refs stay in four scalars: an array indexed by the sorted slot would be
        // placed in scratch memory.
        uint32_t key[4];
        int nhit = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const float x0 = fmaf(lxs[k], inv.x, -oi.x), x1 = fmaf(hxs[k], inv.x, -oi.x);
          const float y0 = fmaf(lys[k], inv.y, -oi.y), y1 = fmaf(hys[k], inv.y, -oi.y);
          const float z0 = fmaf(lzs[k], inv.z, -oi.z), z1 = fmaf(hzs[k], inv.z, -oi.z);
          const float tn = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fmaxf(fminf(z0, z1), t_lo));
          const float tf = fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fminf(fmaxf(z0, z1), t_hi));
          // unused slots (kEmptyChild) never hit, whatever their stored box
          const bool h = tn <= tf && used[k];
          key[k] = h ? ((__float_as_uint(tn) & ~3u) | (uint32_t)k) : 0xffffffffu;
          nhit += h ? 1 : 0;
        }
        if (nhit > 0) {
#define EVER_CSWAP(a, b)                           \
  {                                                \
    const uint32_t lo_ = min(key[a], key[b]);      \
    key[b] = max(key[a], key[b]);                  \
    key[a] = lo_;                                  \
  }
          EVER_CSWAP(0, 1) EVER_CSWAP(2, 3) EVER_CSWAP(0, 2) EVER_CSWAP(1, 3) EVER_CSWAP(1, 2)
#undef EVER_CSWAP
          auto child = [&](uint32_t kk) {
            const bool odd = (kk & 1u) != 0u;
            const int32_t lo = odd ? c1 : c0, hi = odd ? c3 : c2;
            return (kk & 2u) ? hi : lo;
          };
#pragma unroll
          for (int k = 3; k >= 1; --k) {
            if (k < nhit) {  // sp < kStackBound <= kLdsDepth + kSpillDepth (tree bound, ever_internal.hpp)
              const int32_t ck = child(key[k]);
              if (sp < kLdsDepth) stk[sp * kFwdBlock + lane] = ck;
              else spill[sp - kLdsDepth] = ck;
              ++sp;
            }
          }
          node = child(key[0]);
          continue;
        }
      } else {
        const uint32_t s = (uint32_t)(~node);
        const float4 RA = A.rec[4 * (int64_t)s + 0], RB = A.rec[4 * (int64_t)s + 1], RC = A.rec[4 * (int64_t)s + 2];
        uint64_t e, x;
        leaf_events<kBounded>(A, S, G, RA, RB, RC, s, e, x);
        kbuf_insert<K>(bk, e);
        kbuf_insert<K>(bk, x);
      }
      if (sp == 0) break;
      --sp;
      node = sp < kLdsDepth ? stk[sp * kFwdBlock + lane] : spill[sp - kLdsDepth];
    }
    if (integrate_events<K, kBounded, kViews>(A, S, bk)) {
      // fewer than K events up to the bound: done under the plain walk or when the bound lies beyond the scene,
      // else continue without the proposal
      if (!kBounded || S.bound == INFINITY || S.bound >= scene_exit(A, ray_geo(A, S.r))) break;
      S.bound = INFINITY;
    }
  }
  ray_write(A, S);
}

// Loads through the constant address space: with a wave-uniform address the compiler emits scalar loads.
typedef __attribute__((address_space(4))) const float const_float;
__device__ __forceinline__ float4 ld_uniform4(const float4* p, int64_t i) {
  const const_float* q = (const const_float*)(p + i);
  return make_float4(q[0], q[1], q[2], q[3]);
}

// Per-wave traversal: the 64 rays of a wave (neighbours in the Morton order of the ray sort) walk the tree
// together. A node is entered when the box test of any live lane passes, so every lane sees a superset of the
// nodes its own walk would visit; which events reach a lane's k-buffer does not depend on the visiting order.
// Node and leaf records are read once per wave through the scalar cache, the stack is wave-uniform (one entry
// per lane of two registers), and children are ordered by the entry distances of the first live lane.
// kSigns: every lane of the wave has the direction signs `sg` (bit a set: component a negative), so the entry
// and exit planes of each axis are known per wave and the slab test needs no per-lane min/max.
// The k-buffer of a wave lives in LDS (entry k of lane l at ev[k][l]) as a max-heap per lane; only its top, the
// culling bound, is kept in a register. An insertion runs only when some lane of the wave has an event below its
// bound.

// Comparator list of Batcher's odd-even merge sort for K entries: the network for the next power of two with the
// comparators that reach past K left out (those entries would be +inf, so such comparators move nothing).
struct CmpNet {
  int n;
  int8_t a[192], b[192];
};
constexpr CmpNet batcher_net(int K) {
  CmpNet c{};
  int P = 1;
  while (P < K) P <<= 1;
  for (int p = 1; p < P; p <<= 1)
    for (int k = p; k >= 1; k >>= 1)
      for (int j = k % p; j + k < P; j += 2 * k)
        for (int i = 0; i < k && i + j + k < P; ++i)
          if ((i + j) / (2 * p) == (i + j + k) / (2 * p) && i + j + k < K) {
            c.a[c.n] = (int8_t)(i + j);
            c.b[c.n] = (int8_t)(i + j + k);
            ++c.n;
          }
  return c;
}
constexpr int heap_levels(int K) { return K < 2 ? 0 : 1 + heap_levels(K / 2); }  // floor(log2 K)

// The k-buffer of a lane during the walk is a binary max-heap over its K entries (children of entry k at 2k + 1
// and 2k + 2; when K is even a row of zeros follows the last entry, so the last parent's missing child never
// wins), started as K copies of the empty event. An event below the lane's K-th (the top, kept in a register as
// the culling bound) replaces the top and sinks: at most log2(K) moves where a sorted insertion shifts up to K.
// The fragments of one primitive report identical events (a quarter of the insertions at 1M/1080p): an event
// whose key the lane's entries already hold is dropped (a key is one primitive's entry or exit, which a ray meets
// once, so within a traversal equal keys are equal events). The scan reads the keys only.

// v replaces the top (v below it) and sinks to its place.
template <int K>
__device__ __forceinline__ void kheap_replace_sift(uint64_t (*ev)[kWaveBlock], int lane, uint64_t v, uint64_t& kmax) {
  int i = 0;
  uint64_t top = v;
#pragma unroll
  for (int level = 0; level < heap_levels(K); ++level) {
    const int c = 2 * i + 1;
    if (c >= K) break;
    const uint64_t l = ev[c][lane], r = ev[c + 1][lane];
    const bool right = r > l;
    const uint64_t mv = right ? r : l;
    if (!(mv > v)) break;
    ev[i][lane] = mv;
    if (level == 0) top = mv;
    i = c + (right ? 1 : 0);
  }
  ev[i][lane] = v;
  kmax = top;
}
// A leaf's entry e and exit x (x > e, kNoEvent where none) inserted in turn, with one key scan for both (their keys
// differ, so the scan before the entry's insertion answers for the exit too; one scan per leaf instead of one per
// event: forward 6.69 -> 6.50 ms at 1M/1080p).
template <int K>
__device__ __forceinline__ void kheap_insert2(uint64_t (*ev)[kWaveBlock], int lane, uint64_t e, uint64_t x,
                                              uint64_t& kmax) {
  const bool te = e < kmax, tx = x < kmax;
  if (!te && !tx) return;
  bool de = false, dx = false;
  const uint32_t* evk = reinterpret_cast<const uint32_t*>(&ev[0][lane]);  // (low word of entry k: its key)
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const uint32_t kk = evk[k * 2 * kWaveBlock];
    de |= kk == (uint32_t)e;
    dx |= kk == (uint32_t)x;
  }
  if (te && !de) kheap_replace_sift<K>(ev, lane, e, kmax);
  if (x < kmax && !dx) kheap_replace_sift<K>(ev, lane, x, kmax);
}

// The heap sorted ascending for the integrator, in registers by Batcher's network.
template <int K>
__device__ __forceinline__ void kheap_sort(uint64_t (*ev)[kWaveBlock], int lane) {
  constexpr CmpNet net = batcher_net(K);
  uint64_t r[K];
#pragma unroll
  for (int k = 0; k < K; ++k) r[k] = ev[k][lane];
#pragma unroll
  for (int c = 0; c < net.n; ++c) {
    const uint64_t x = r[net.a[c]], y = r[net.b[c]];
    r[net.a[c]] = x < y ? x : y;
    r[net.b[c]] = x < y ? y : x;
  }
#pragma unroll
  for (int k = 0; k < K; ++k) ev[k][lane] = r[k];
}

template <int K, bool kSigns, bool kBounded, bool kViews>
__device__ __forceinline__ void wave_trace(const FwdArgs& A, RayState& S, int lane, uint32_t sg,
                                           uint64_t (*ev)[kWaveBlock], int32_t* row2) {
  const int64_t nleaves = A.header[9];
  const int32_t root = nleaves == 1 ? ~0 : 0;
  bool done = false;  // (a lane without a ray is never alive: ray_idle)

  while (true) {
    const bool alive = !done && S.live;
    const uint64_t live = __ballot(alive);
    if (live == 0) break;
    const int leader = __builtin_ctzll(live);
    const RayGeo G = ray_geo(A, S.r);
    const f3 inv = G.inv, oi = G.oi;
#pragma unroll
    for (int k = 0; k < K; ++k) ev[k][lane] = kNoEvent;
    uint64_t kmax = kNoEvent;
    // a finished lane never enters a box; the cursor is held at tmax once past it (exits beyond tmax whose entries
    // lie before it may remain, in boxes that straddle tmax)
    float t_lo = alive ? fminf(cur_t(S) > 0.f ? cur_t(S) : band_lo(0.f), A.tmax_lo) : INFINITY;  // (leaf_events)
    // kept opaque: otherwise the select is folded into every box test below (two extra selects per child)
    asm volatile("" : "+v"(t_lo));
    // wave stack: entry i in lane i % 64 of row i / 64; rows 0 and 1 in VGPRs, row 2 in LDS (entries beyond 128,
    // needed only by trees deeper than 42 records; 192 entries cover the bound of 180). Entry 128 + i is written
    // and read back by lane i only.
    int32_t stk0 = 0, stk1 = 0;
    int sp = 0;
    int32_t node = root;
    // the last 64 primitives whose leaves the wave tested in this traversal (lane i holds entry i, replaced in turn):
    // another fragment of one of them is skipped, since its test would give every lane the same events, which the
    // lane then already holds or has rejected against a bound that has only come down since (measured 1.5-2% of the
    // forward at 1M/1080p; 128 entries in two registers, no better)
    uint32_t seen = 0xffffffffu;
    uint32_t nseen = 0;
    // the upper culling depth: the lane's k-th event (recomputed when an insertion changes it); -inf on a finished
    // lane, which then enters no box (its lower depth is +inf). Kept opaque like t_lo.
    auto upper = [&]() {
      float h =
          alive ? fminf(kBounded ? fminf(A.tmax_hi, S.bound) : A.tmax_hi, __uint_as_float((uint32_t)(kmax >> 32)))
                : -INFINITY;
      asm volatile("" : "+v"(h));
      return h;
    };
    float t_hi = upper();
    uint32_t steps = 0;
    const uint32_t max_steps = 2u * (uint32_t)nleaves + 2u;  // a tree visits each node at most once per traversal
    while (++steps <= max_steps) {
      if (node >= 0) {
        const float4* nd = A.nodes4 + 8 * (int64_t)node;
        // with known signs the rows are read as entry / exit planes (row 2a + 1 holds the upper planes of axis a)
        const int ex = kSigns ? (int)(sg & 1u) : 0, ey = kSigns ? (int)((sg >> 1) & 1u) : 0,
                  ez = kSigns ? (int)((sg >> 2) & 1u) : 0;
        const float4 lx = ld_uniform4(nd, ex), hx = ld_uniform4(nd, 1 - ex), ly = ld_uniform4(nd, 2 + ey),
                     hy = ld_uniform4(nd, 3 - ey), lz = ld_uniform4(nd, 4 + ez), hz = ld_uniform4(nd, 5 - ez),
                     cf = ld_uniform4(nd, 6);
        const int32_t c0 = __float_as_int(cf.x), c1 = __float_as_int(cf.y), c2 = __float_as_int(cf.z),
                      c3 = __float_as_int(cf.w);
        // the child references are needed only after the box tests: held here so that their load is issued with the
        // planes' (one wait for the whole record) rather than after the tests
        asm volatile("" ::"s"(c0), "s"(c1), "s"(c2), "s"(c3));
        // an unused slot holds the empty box (+inf lower, -inf upper planes); with known signs and finite inv (clamped
        // at 1e20) and oi (checked before the signed walk), its entry distance is +inf and its exit -inf, so the box
        // test rejects it without the slot check
        const bool used[4] = {kSigns || c0 != kEmptyChild, kSigns || c1 != kEmptyChild, kSigns || c2 != kEmptyChild,
                              kSigns || c3 != kEmptyChild};
        const float lxs[4] = {lx.x, lx.y, lx.z, lx.w}, hxs[4] = {hx.x, hx.y, hx.z, hx.w};
        const float lys[4] = {ly.x, ly.y, ly.z, ly.w}, hys[4] = {hy.x, hy.y, hy.z, hy.w};
        const float lzs[4] = {lz.x, lz.y, lz.z, lz.w}, hzs[4] = {hz.x, hz.y, hz.z, hz.w};
        // the children some lane enters, as a wave-uniform bit mask, and each child's sort key per lane: the entry
        // distance where the lane enters it (+inf where not), the slot in the low 2 bits
        uint32_t tks[4];
        uint32_t hmask = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          float tn, tf;
          if (kSigns) {  // lxs.. hold the entry planes, hxs.. the exit planes
            tn = fmaxf(fmaxf(fmaf(lxs[k], inv.x, -oi.x), fmaf(lys[k], inv.y, -oi.y)),
                       fmaxf(fmaf(lzs[k], inv.z, -oi.z), t_lo));
            tf = fminf(fminf(fmaf(hxs[k], inv.x, -oi.x), fmaf(hys[k], inv.y, -oi.y)),
                       fminf(fmaf(hzs[k], inv.z, -oi.z), t_hi));
          } else {
            const float x0 = fmaf(lxs[k], inv.x, -oi.x), x1 = fmaf(hxs[k], inv.x, -oi.x);
            const float y0 = fmaf(lys[k], inv.y, -oi.y), y1 = fmaf(hys[k], inv.y, -oi.y);
            const float z0 = fmaf(lzs[k], inv.z, -oi.z), z1 = fmaf(hzs[k], inv.z, -oi.z);
            tn = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fmaxf(fminf(z0, z1), t_lo));
            tf = fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fminf(fmaxf(z0, z1), t_hi));
          }
          // the lanes entering the box (ordered compare: a NaN bound enters nothing)
          const bool enter = tn <= tf;
          const uint64_t hm = __builtin_amdgcn_ballot_w64(enter);
          tks[k] = enter ? ((__float_as_uint(tn) & ~3u) | (uint32_t)k) : (0x7f800000u | (uint32_t)k);
          // (hm != 0) as 0/1 in a scalar register: written as a compare and select because the compiler otherwise
          // keeps it as a lane mask, and turns it into an integer through a vector select and a readfirstlane
          uint32_t any;
          asm("s_cmp_lg_u64 %1, 0\n\ts_cselect_b32 %0, 1, 0" : "=s"(any) : "s"(hm) : "scc");
          hmask |= (used[k] ? any : 0u) << k;
        }
        auto child = [&](uint32_t kk) {
          const bool odd = (kk & 1u) != 0u;
          const int32_t lo = odd ? c1 : c0, hi = odd ? c3 : c2;
          return (kk & 2u) ? hi : lo;
        };
        if (hmask != 0) {
          if ((hmask & (hmask - 1u)) == 0u) {  // one child entered: nothing to order or push
            node = child((uint32_t)__builtin_ctz(hmask));
            continue;
          }
          // wave-uniform sort keys: the first live lane's entry distance (+inf where it does not enter) with the
          // slot in the low 2 bits; children no lane enters sort last
          uint32_t key[4];
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            const uint32_t tl = (uint32_t)__builtin_amdgcn_readlane((int)tks[k], leader);
            key[k] = ((hmask >> k) & 1u) ? tl : 0xffffffffu;
          }
          const int nhit = __builtin_popcount(hmask);
#define EVER_CSWAP(a, b)                           \
  {                                                \
    const uint32_t lo_ = min(key[a], key[b]);      \
    key[b] = max(key[a], key[b]);                  \
    key[a] = lo_;                                  \
  }
          EVER_CSWAP(0, 1) EVER_CSWAP(2, 3) EVER_CSWAP(0, 2) EVER_CSWAP(1, 3) EVER_CSWAP(1, 2)
#undef EVER_CSWAP
#pragma unroll
          for (int k = 3; k >= 1; --k) {
            if (k < nhit) {  // sp < kStackBound <= 3 * 64 (tree bound, ever_internal.hpp)
              const int32_t ck = child(key[k]);
              if (sp < 64) stk0 = (lane == sp) ? ck : stk0;
              else if (sp < 128) stk1 = (lane == sp - 64) ? ck : stk1;
              else if (lane == sp - 128) row2[lane] = ck;
              ++sp;
            }
          }
          node = child(key[0]);
          continue;
        }
      } else {
        const uint32_t s = (uint32_t)(~node);
        if (__builtin_amdgcn_ballot_w64(seen == s) == 0) {
          seen = (lane == (int)(nseen & 63u)) ? s : seen;
          ++nseen;
          const float4* rc = A.rec + 4 * (int64_t)s;
          const float4 RA = ld_uniform4(rc, 0), RB = ld_uniform4(rc, 1), RC = ld_uniform4(rc, 2);
          uint64_t e = kNoEvent, x = kNoEvent;
          if (alive) leaf_events<kBounded>(A, S, G, RA, RB, RC, s, e, x);
          if (__builtin_amdgcn_ballot_w64(e < kmax || x < kmax) != 0) {
            kheap_insert2<K>(ev, lane, e, x, kmax);
            t_hi = upper();
          }
        }
      }
      if (sp == 0) break;
      --sp;
      if (sp < 128) node = __builtin_amdgcn_readlane(sp < 64 ? stk0 : stk1, sp & 63);
      else node = __builtin_amdgcn_readlane(row2[lane], sp - 128);  // lane sp - 128 reads back its own store
    }
    kheap_sort<K>(ev, lane);
    if (alive && integrate_seq<K, kBounded, kViews>(A, S, [&](int k) { return ev[k][lane]; })) {
      // fewer than K events up to the bound: done under the plain walk or when the bound lies beyond the scene,
      // else continue without the proposal
      if (!kBounded || S.bound == INFINITY || S.bound >= scene_exit(A, ray_geo(A, S.r))) done = true;
      else S.bound = INFINITY;
    }
  }
}

template <int K, int kFlags>  // kFlags: 1 proposal-bounded, 2 several views, 4 persistent
__global__ __launch_bounds__(kWaveBlock, (kFlags & 4) ? 3 : wave_walk_waves<K>()) void k_trace_forward_wave(FwdArgs A) {
  constexpr bool kBounded = (kFlags & 1) != 0, kViews = (kFlags & 2) != 0, kPersistent = (kFlags & 4) != 0;
  const int lane = threadIdx.x;
  __shared__ uint64_t ev[K + (K % 2 == 0 ? 1 : 0)][kWaveBlock];  // (the zero row after a heap of even size)
  if constexpr (K % 2 == 0) ev[K][lane] = 0;
  __shared__ int32_t row2[kWaveBlock];  // third row of the wave's traversal stack
  // the 64 rays of processing slots [base, base + 64); lanes past the batch stay in the wave without a ray
  // (ray_idle), so that all 64 lanes hold the wave's stack
  auto trace_tile = [&](int64_t base) {
    if (base >= A.nr) return;  // the whole wave: uniform
    const int64_t slot = base + lane;
    const bool valid = slot < A.nr;
    RayState S;
    if (valid) {
      ray_setup<kViews>(A, A.order[slot], slot, S);
      if (kBounded) ray_bound(A, S);
    } else {
      ray_idle(A, S);
    }
    // lane 0 always has a ray; rayless lanes take its signs so they do not break the wave's sign uniformity.
    // The signed walk also needs finite origin * reciprocal on every lane (it rejects unused child slots by their
    // empty boxes, which a NaN slab distance would not do); a wave with a non-finite one takes the general walk.
    const RayGeo G0 = ray_geo(A, S.r);
    const f3 inv0 = G0.inv;
    uint32_t sg = (signbit(inv0.x) ? 1u : 0u) | (signbit(inv0.y) ? 2u : 0u) | (signbit(inv0.z) ? 4u : 0u);
    const uint32_t sg0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)sg);
    if (!valid) sg = sg0;
    const bool fin = isfinite(G0.oi.x) && isfinite(G0.oi.y) && isfinite(G0.oi.z);
    if (__ballot(sg != sg0 || !fin) == 0) wave_trace<K, true, kBounded, kViews>(A, S, lane, sg0, ev, row2);
    else wave_trace<K, false, kBounded, kViews>(A, S, lane, sg0, ev, row2);
    if (base + lane < A.nr) ray_write(A, S);
  };
  if (!kPersistent) {
    // blocks are dispatched to the 8 XCDs in turn (block b on XCD b % 8): give each XCD runs of kXcdChunk consecutive
    // tiles (a 32x32 pixel block of one camera) so that its L2 serves neighbouring tiles (measured 1-2% faster than
    // one tile per XCD at a time; runs of 256 tiles, 8% slower: the XCDs then stop sharing the last-level cache)
    const int64_t nb = gridDim.x, b = blockIdx.x;
    const int64_t full = nb / (8 * kXcdChunk) * (8 * kXcdChunk);
    const int64_t t = b < full ? ((b / 8) / kXcdChunk) * (8 * kXcdChunk) + (b % 8) * kXcdChunk + (b / 8) % kXcdChunk : b;
    trace_tile(t * kWaveBlock);
    return;
  }
  // persistent: fetch A.tile consecutive tiles at a time until the batch is exhausted (every wave reaches the
  // exit once the counter passes the last tile)
  while (true) {
    uint32_t t = 0;
    if (lane == 0) t = atomicAdd(A.tile_counter, 1u);
    t = (uint32_t)__builtin_amdgcn_readfirstlane((int)t);
    const int64_t first = (int64_t)t * A.tile * kWaveBlock;
    if (first >= A.nr) break;
    for (int i = 0; i < A.tile; ++i) trace_tile(first + (int64_t)i * kWaveBlock);
  }
}

// the instantiation for (walk, k-buffer size, flags)
template <int K, bool kWave>
static void (*pick_flags(int flags))(FwdArgs) {
  if (kWave) {
    switch (flags) {  // persistent launches (4) exist for the plain and the several-views forms
      case 1: return k_trace_forward_wave<K, 1>;
      case 2: return k_trace_forward_wave<K, 2>;
      case 3: return k_trace_forward_wave<K, 3>;
      case 4: return k_trace_forward_wave<K, 4>;
      case 6: return k_trace_forward_wave<K, 6>;
      default: return k_trace_forward_wave<K, 0>;
    }
  }
  switch (flags) {
    case 1: return k_trace_forward<K, 1>;
    case 2: return k_trace_forward<K, 2>;
    case 3: return k_trace_forward<K, 3>;
    default: return k_trace_forward<K, 0>;
  }
}

template <bool kWave>
static void (*pick_kernel(int kbuf, int flags))(FwdArgs) {
  return kbuf == 8    ? pick_flags<8, kWave>(flags)
         : kbuf == 19 ? pick_flags<19, kWave>(flags)
         : kbuf == 32 ? pick_flags<32, kWave>(flags)
                      : pick_flags<16, kWave>(flags);
}

hipError_t launch_forward(const ever_prims& P, void* ws, const ever_rays& R, const ever_forward_out& O,
                          hipStream_t st, bool keep_candidates) {
  SceneLayout L = scene_layout(P.num_prims);
  uint32_t* header = at<uint32_t>(ws, L.header);
  int32_t* cand = at<int32_t>(ws, L.cand);
  if (!keep_candidates) hipLaunchKernelGGL(k_reset_count, dim3(1), dim3(1), 0, st, header);
  if (O.touch_count && P.num_prims > 0)
    hipLaunchKernelGGL(k_zero_i32, dim3(grid_for(P.num_prims, 256)), dim3(256), 0, st, O.touch_count, P.num_prims);
  if (R.num_rays <= 0) return hipGetLastError();
  if (P.num_prims > 0 && !keep_candidates)
    hipLaunchKernelGGL(k_prefilter, dim3(grid_for(P.num_prims, 256)), dim3(256), 0, st, P, R.origins, R.tmin, cand,
                       header);
  // processing order
  ForwardLayout FL = forward_layout(R.num_rays, per_ray_walk(R.num_rays, R.traversal));
  uint32_t* k0 = at<uint32_t>(O.workspace, FL.keys0);
  uint32_t* v0 = at<uint32_t>(O.workspace, FL.vals0);
  if (R.num_rays >= kMinRaysToSort && P.num_prims > 0) {
    // several views: each view keys its rays as its own camera; the first ray of each view goes to the sort's second
    // value buffer, unused until the sort starts
    const bool views = P.num_views > 1 && R.view_index && P.num_views <= R.num_rays;
    int32_t* vfirst = at<int32_t>(O.workspace, FL.vals1);
    if (views)
      hipLaunchKernelGGL(k_view_first, dim3(grid_for(R.num_rays, 256)), dim3(256), 0, st, R.view_index, R.num_rays,
                         P.num_views, vfirst);
    hipLaunchKernelGGL(k_ray_keys, dim3(grid_for(R.num_rays, 256)), dim3(256), 0, st, R.origins, R.directions,
                       R.num_rays, R.tmin, header, R.view_index, P.view_origins, views ? P.num_views : 0, vfirst, k0,
                       v0);
    // the low 6 bits order rays within an 8 x 8 tile (or a cell of 4 x 4 x 4) and are left unsorted (three passes
    // instead of four): a wave takes 64 consecutive rays, and a tile keeps its input order, which moves rays between
    // waves only where a tile straddles two (at 1920 x 1080 a wave's pixels then span 64.9 pixels of bounding box on
    // average, against 64.8 with the full sort)
    hipError_t e = launch_sort_pairs(k0, v0, at<uint32_t>(O.workspace, FL.keys1), at<uint32_t>(O.workspace, FL.vals1),
                                     at<uint32_t>(O.workspace, FL.hist), R.num_rays, 30, st, nullptr, 6, &v0);
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
  A.tmax_lo = R.tmax - 1e-4f * (fabsf(R.tmax) + 1.f);
  A.tmax_hi = R.tmax + 1e-4f * (fabsf(R.tmax) + 1.f);
  A.max_iters = R.max_iters;
  A.color = O.color;
  A.states = O.states;
  A.iters = O.iters;
  A.tri = O.tri_collection;
  A.events_by_slot = R.event_layout == EVER_EVENTS_BY_SLOT ? 1 : 0;
  A.touch = O.touch_count;
  A.init_drgb = O.initial_drgb;
  A.st64 = O.ray_state64;
  A.last_dirac = O.last_dirac;
  A.spill = at<int32_t>(O.workspace, FL.spill);
  A.ist = at<char>(O.workspace, FL.ist);
  A.order = v0;
  A.proposal = R.proposal;
  A.vcol = P.num_views > 1 ? reinterpret_cast<const float4*>(P.view_colors) : nullptr;
  A.view_index = R.view_index;
  A.tile = 0;
  A.tile_counter = nullptr;
  dim3 grid(grid_for(R.num_rays, kFwdBlock));
  if (P.num_prims <= 0) {
    // no primitive: every ray sees no event; run the kernel with an empty loop by setting max_iters = 0
    A.max_iters = 0;
  }
  const bool wave = !per_ray_walk(R.num_rays, R.traversal);  // (the forward workspace is sized for this walk)
  profile_begin(EVER_KERNEL_FORWARD, st);
  int flags = (R.proposal ? 1 : 0) | (A.vcol ? 2 : 0);
  if (wave) {
    dim3 wgrid(grid_for(R.num_rays, kWaveBlock));
    if (R.tile > 0 && !R.proposal) {
      // one block per wave slot of the GPU (16 per CU), no more than there are fetches
      int dev = 0, cus = 0;
      hipError_t e = hipGetDevice(&dev);
      if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
      if (e != hipSuccess) return e;
      const int64_t fetches = (R.num_rays + (int64_t)R.tile * kWaveBlock - 1) / ((int64_t)R.tile * kWaveBlock);
      wgrid = dim3((unsigned)std::max<int64_t>(1, std::min<int64_t>(fetches, 16 * (int64_t)cus)));
      A.tile = R.tile;
      A.tile_counter = at<uint32_t>(O.workspace, FL.counter);
      if (!R.proposal) flags |= 4;  // the proposal-bounded form has no persistent variant
      hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, st, reinterpret_cast<int32_t*>(A.tile_counter),
                         (int64_t)1);
    }
    hipLaunchKernelGGL(pick_kernel<true>(R.kbuf, flags), wgrid, dim3(kWaveBlock), 0, st, A);
  } else {
    hipLaunchKernelGGL(pick_kernel<false>(R.kbuf, flags), grid, dim3(kFwdBlock), 0, st, A);
  }
  profile_end(EVER_KERNEL_FORWARD, st);
  return hipGetLastError();
}

}  // namespace ever
