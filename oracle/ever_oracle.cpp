// EVER tracer oracle — TEST INFRASTRUCTURE ONLY.
//
// A CPU restatement (templated on the scalar type, instantiated in FP64 for checking and FP32 for the CPU
// timing baseline) of the reference's exact volumetric ellipsoid tracer and of its backward pass. Only the
// tests, __graft_entry__.smoke() and bench.py's cpu_baseline leg load this library, and only as the checker
// or the timed CPU baseline. The product path (ever_amd/) never loads it.
//
// Reference files followed (paths relative to the reference repository root):
//   AABB                 splinetracers/fast_ellipsoid_splinetracer/create_aabbs.cu:69-115
//   start-inside density splinetracers/fast_ellipsoid_splinetracer/initialize_density.cu:69-150
//   event loop           splinetracers/fast_ellipsoid_splinetracer/slang/shaders.slang:84-164 (intersection,
//                        any-hit k-buffer), :171-235 (raygen loop), constants :15-18
//   roots                splinetracers/fast_ellipsoid_splinetracer/slang/tri-intersect.slang:268-302, 317-344
//   integrator           splinetracers/fast_ellipsoid_splinetracer/slang/spline-machine.slang:103-153
//   guarded math         splinetracers/slang/safe-math.slang:16-18, 44-144, 237-253
//   colour (DC)          splinetracers/slang/sh.slang:18-51 ; shaders.slang:65-82
//   backward             splinetracers/fast_ellipsoid_splinetracer/slang/backwards_kernel.slang:132-367
//                        (adjoint of update + safe_intersect + eval_color, means2D heuristic),
//                        :374-428 (start-inside adjoint); clamps splinetracers/fast_ellipsoid_splinetracer.py:201-219
//
// Where the reference reconstructs earlier states by subtraction (spline-machine.slang:65-100), this oracle
// recomputes every state forward in the working precision and stores them per ray, then runs the exact
// reverse-mode adjoint of `update` (derived by hand; pinned by central finite differences in the tests).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <climits>
#include <vector>
#include <omp.h>

namespace {

constexpr double kTiny = 1.0754944e-20;   // safe-math.slang:16
constexpr double kMinVal = -1e20;         // safe-math.slang:17
constexpr double kMaxVal = 1e20;          // safe-math.slang:18
constexpr float kLogCutoff = 5.54f;       // shaders.slang:17 (an fp32 literal in the reference)
constexpr double kPreMulti = 1000.0;      // spline-machine.slang:17
constexpr double kLadderP = -0.1;         // spline-machine.slang:18
constexpr double kShC0 = 0.28209479177387814;  // sh.slang:18

template <class R> struct V3 {
  R x, y, z;
};
template <class R> inline V3<R> mk(R x, R y, R z) { return {x, y, z}; }
template <class R> inline V3<R> operator+(V3<R> a, V3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class R> inline V3<R> operator-(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class R> inline V3<R> operator*(R s, V3<R> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class R> inline V3<R> operator-(V3<R> a) { return {-a.x, -a.y, -a.z}; }
template <class R> inline R dot(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class R> inline V3<R> cross(V3<R> a, V3<R> b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class R> inline V3<R> load3(const R* p) { return {p[0], p[1], p[2]}; }

template <class R> inline R clipv(R v, R lo, R hi) { return std::max(std::min(v, hi), lo); }
template <class R> inline R sgn(R v) { return v > 0 ? R(1) : (v < 0 ? R(-1) : R(0)); }

// safe_div (safe-math.slang:84-91): divides by +TINY when |b| < TINY, result clipped to +-1e20.
template <class R> inline R safe_div(R a, R b) {
  if (std::fabs(b) < R(kTiny)) return clipv<R>(a / R(kTiny), R(kMinVal), R(kMaxVal));
  return clipv<R>(a / b, R(kMinVal), R(kMaxVal));
}
// bw_safe_div (safe-math.slang:68-83). Returns (da, db) for upstream g.
template <class R> inline void bw_safe_div(R a, R b, R g, R& da, R& db) {
  if (std::fabs(b) < R(kTiny)) da = clipv<R>(g / R(kTiny), R(kMinVal), R(kMaxVal));
  else da = clipv<R>(g / b, R(kMinVal), R(kMaxVal));
  R b2 = b * b;
  if (b2 < R(kTiny)) db = clipv<R>(-a / R(kTiny) * g, R(kMinVal), R(kMaxVal));
  else db = clipv<R>(-a / b2 * g, R(kMinVal), R(kMaxVal));
}
// safe_sqrt (safe-math.slang:122-129)
template <class R> inline R safe_sqrt(R a) { return a < R(kTiny) ? R(0) : std::sqrt(a); }
// safe_exp (safe-math.slang:141-144)
template <class R> inline R safe_exp(R v) { return std::exp(clipv<R>(v, R(kMinVal), R(std::log(kMaxVal)))); }

// rotate_vector (safe-math.slang:237-244): rotation by the conjugate of q = (w, x, y, z).
template <class R> inline V3<R> rotate(V3<R> v, const R* q) {
  V3<R> u = mk(-q[1], -q[2], -q[3]);
  V3<R> t = R(2) * cross(u, v);
  return v + q[0] * t + cross(u, t);
}
// Adjoint of rotate: accumulates into vbar (3) and qbar (4).
template <class R> inline void rotate_bwd(V3<R> v, const R* q, V3<R> obar, V3<R>& vbar, R* qbar) {
  V3<R> u = mk(-q[1], -q[2], -q[3]);
  V3<R> t = R(2) * cross(u, v);
  // out = v + w*t + cross(u, t)
  V3<R> ubar = cross(t, obar);           // d cross(u,t) / du
  V3<R> tbar = cross(obar, u) + q[0] * obar;
  R wbar = dot(obar, t);
  // t = 2 cross(u, v)
  ubar = ubar + R(2) * cross(v, tbar);
  vbar = vbar + obar + R(2) * cross(tbar, u);
  qbar[0] += wbar;
  qbar[1] -= ubar.x;
  qbar[2] -= ubar.y;
  qbar[3] -= ubar.z;
}

// tukey_power_ladder (safe-math.slang:246-253) with p = LADDER_P, and its derivative.
template <class R> inline R ladder(R x) {
  const R p = R(kLadderP);
  R xp = std::fabs(x);
  R xs = xp / std::max(R(kTiny), std::fabs(p - R(1)));
  return sgn(x) * std::fabs(p - R(1)) / p * (std::pow(xs + R(1), p) - R(1));
}
template <class R> inline R ladder_grad(R x) {
  // d/dx = sign(x)^2 * (|x|/|p-1| + 1)^(p-1); zero at x == 0 like the automatic derivative of sign().
  const R p = R(kLadderP);
  if (x == R(0)) return R(0);
  R xs = std::fabs(x) / std::fabs(p - R(1));
  return std::pow(xs + R(1), p - R(1));
}

// l2_normalize (safe-math.slang:136-139)
template <class R> inline V3<R> l2n(V3<R> x) {
  R n = std::sqrt(std::max(dot(x, x), R(kTiny)));
  return mk(safe_div(x.x, n), safe_div(x.y, n), safe_div(x.z, n));
}
template <class R> inline V3<R> l2n_bwd(V3<R> x, V3<R> ybar) {
  R m = dot(x, x);
  R n = std::sqrt(std::max(m, R(kTiny)));
  R dax, day, daz, db0, db1, db2;
  bw_safe_div(x.x, n, ybar.x, dax, db0);
  bw_safe_div(x.y, n, ybar.y, day, db1);
  bw_safe_div(x.z, n, ybar.z, daz, db2);
  R nbar = db0 + db1 + db2;
  V3<R> xbar = mk(dax, day, daz);
  if (m > R(kTiny)) xbar = xbar + (nbar / n) * x;  // d sqrt(m)/dx = x / sqrt(m)
  return xbar;
}

// Ellipsoid roots. Unsafe variant: eliIntersect (tri-intersect.slang:268-283) used by the forward.
// Safe variant: safe_eliIntersect (tri-intersect.slang:286-302) used by the backward. Both return
// (min, max) of the two roots (ray_intersect_ellipsoid :317-329) or (-1, -1) on a miss.
template <class R> inline bool eli_roots(V3<R> ro, V3<R> rd, V3<R> s, bool safe, R& t0, R& t1) {
  V3<R> ocn, rdn;
  if (safe) {
    ocn = mk(safe_div(ro.x, s.x), safe_div(ro.y, s.y), safe_div(ro.z, s.z));
    rdn = mk(safe_div(rd.x, s.x), safe_div(rd.y, s.y), safe_div(rd.z, s.z));
  } else {
    ocn = mk(ro.x / s.x, ro.y / s.y, ro.z / s.z);
    rdn = mk(rd.x / s.x, rd.y / s.y, rd.z / s.z);
  }
  R a = dot(rdn, rdn);
  R bp = -dot(ocn, rdn);
  V3<R> l = ocn + (safe ? safe_div(bp, a) : bp / a) * rdn;
  R h = a * (R(1) - dot(l, l));
  R c = dot(ocn, ocn) - R(1);
  if (h < R(0)) {
    t0 = t1 = R(-1);
    return false;
  }
  R q = bp + sgn(bp) * (safe ? safe_sqrt(h) : std::sqrt(h));
  R r0 = safe ? safe_div(c, q) : c / q;
  R r1 = safe ? safe_div(q, a) : q / a;
  t0 = std::min(r0, r1);
  t1 = std::max(r0, r1);
  return true;
}

// Adjoint of safe_ray_intersect_ellipsoid (tri-intersect.slang:332-344) composed with the root selection of
// safe_intersect (:443-456). Given tbar for the selected root, accumulates into robar/rdbar/sbar/qbar where
// ro = origin - mean, rd = raw direction.
template <class R>
inline void safe_root_bwd(V3<R> ro_w, V3<R> rd_w, V3<R> s, const R* q, int face, R tbar, V3<R>& ro_bar,
                          V3<R>& rd_bar, V3<R>& s_bar, R* q_bar) {
  V3<R> rd_rot = rotate(rd_w, q);
  V3<R> Trd = l2n(rd_rot);
  V3<R> Tro = rotate(ro_w, q);
  // forward of safe_eliIntersect
  V3<R> ocn = mk(safe_div(Tro.x, s.x), safe_div(Tro.y, s.y), safe_div(Tro.z, s.z));
  V3<R> rdn = mk(safe_div(Trd.x, s.x), safe_div(Trd.y, s.y), safe_div(Trd.z, s.z));
  R a = dot(rdn, rdn);
  R bp = -dot(ocn, rdn);
  R ba = safe_div(bp, a);
  V3<R> l = ocn + ba * rdn;
  R ll = dot(l, l);
  R h = a * (R(1) - ll);
  R c = dot(ocn, ocn) - R(1);
  if (h < R(0)) return;  // miss: constant (-1, -1), no derivative
  R sh = safe_sqrt(h);
  R sb = sgn(bp);
  R qq = bp + sb * sh;
  R r0 = safe_div(c, qq);
  R r1 = safe_div(qq, a);
  // min/max selection: min(r0,r1) picks r0 iff r0 < r1; max picks r0 iff r0 > r1.
  R r0bar = 0, r1bar = 0;
  if (face == 1) {
    if (r0 < r1) r0bar = tbar; else r1bar = tbar;
  } else {
    if (r0 > r1) r0bar = tbar; else r1bar = tbar;
  }
  R cbar = 0, qqbar = 0, abar = 0, bpbar = 0, hbar = 0;
  {
    R da, db;
    bw_safe_div(c, qq, r0bar, da, db);
    cbar += da;
    qqbar += db;
    bw_safe_div(qq, a, r1bar, da, db);
    qqbar += da;
    abar += db;
  }
  bpbar += qqbar;
  R shbar = qqbar * sb;
  if (!(h < R(kTiny))) hbar += shbar * R(0.5) / std::sqrt(h);
  // h = a * (1 - ll)
  abar += hbar * (R(1) - ll);
  R llbar = -hbar * a;
  V3<R> lbar = (R(2) * llbar) * l;
  // c = ocn.ocn - 1
  V3<R> ocnbar = (R(2) * cbar) * ocn;
  // l = ocn + ba * rdn
  ocnbar = ocnbar + lbar;
  R babar = dot(lbar, rdn);
  V3<R> rdnbar = ba * lbar;
  {
    R da, db;
    bw_safe_div(bp, a, babar, da, db);
    bpbar += da;
    abar += db;
  }
  // bp = -ocn.rdn
  ocnbar = ocnbar + (-bpbar) * rdn;
  rdnbar = rdnbar + (-bpbar) * ocn;
  // a = rdn.rdn
  rdnbar = rdnbar + (R(2) * abar) * rdn;
  // ocn = Tro / s ; rdn = Trd / s
  V3<R> Trobar, Trdbar;
  {
    R da, db;
    bw_safe_div(Tro.x, s.x, ocnbar.x, da, db); Trobar.x = da; s_bar.x += db;
    bw_safe_div(Tro.y, s.y, ocnbar.y, da, db); Trobar.y = da; s_bar.y += db;
    bw_safe_div(Tro.z, s.z, ocnbar.z, da, db); Trobar.z = da; s_bar.z += db;
    bw_safe_div(Trd.x, s.x, rdnbar.x, da, db); Trdbar.x = da; s_bar.x += db;
    bw_safe_div(Trd.y, s.y, rdnbar.y, da, db); Trdbar.y = da; s_bar.y += db;
    bw_safe_div(Trd.z, s.z, rdnbar.z, da, db); Trdbar.z = da; s_bar.z += db;
  }
  rotate_bwd(ro_w, q, Trobar, ro_bar, q_bar);
  V3<R> rdrotbar = l2n_bwd(rd_rot, Trdbar);
  rotate_bwd(rd_w, q, rdrotbar, rd_bar, q_bar);
}

template <class R> struct State {
  R dpx, dpy, csx, csy, t;
  R drgb[4];
  R logT;
  R C[3];
};

template <class R> inline void state_to16(const State<R>& s, R* o) {
  o[0] = s.dpx; o[1] = s.dpy; o[2] = s.csx; o[3] = s.csy;
  o[4] = 0; o[5] = 0; o[6] = 0; o[7] = s.t;
  o[8] = s.drgb[0]; o[9] = s.drgb[1]; o[10] = s.drgb[2]; o[11] = s.drgb[3];
  o[12] = s.logT; o[13] = s.C[0]; o[14] = s.C[1]; o[15] = s.C[2];
}

// update (spline-machine.slang:103-138)
template <class R> inline State<R> update(const State<R>& s, R t, const R* dirac) {
  State<R> n;
  R dt = std::max(t - s.t, R(0));
  for (int i = 0; i < 4; ++i) n.drgb[i] = s.drgb[i] + dirac[i];
  n.t = t;
  R area = std::max(s.drgb[0] * dt, R(0));
  R rgbn[3] = {safe_div(s.drgb[1], s.drgb[0]), safe_div(s.drgb[2], s.drgb[0]), safe_div(s.drgb[3], s.drgb[0])};
  n.logT = std::max(area + s.logT, R(0));
  R w = clipv<R>((R(1) - safe_exp(-area)) * safe_exp(-s.logT), R(0), R(1));
  for (int i = 0; i < 3; ++i) n.C[i] = s.C[i] + w * rgbn[i];
  R m = ladder((n.t + s.t) / R(2) * R(kPreMulti));
  n.dpx = s.dpx + R(2) * w * m * s.csx;
  n.dpy = s.dpy + R(2) * w * s.csy;
  n.csx = s.csx + w;
  n.csy = s.csy + w * m;
  return n;
}

// Reverse-mode adjoint of update. nb = adjoint of the new state; returns adjoint of the old state `sb` and of
// the control point (tbar, diracbar). Derivative conventions follow the automatic derivative of max/min
// (gradient passes only on the strict side) and bw_safe_div.
template <class R>
inline void update_bwd(const State<R>& s, R t, const State<R>& nb, State<R>& sb, R& tbar, R* diracbar) {
  R dt_raw = t - s.t;
  R dt = std::max(dt_raw, R(0));
  R D = s.drgb[0];
  R a_raw = D * dt;
  R area = std::max(a_raw, R(0));
  R rgbn[3] = {safe_div(s.drgb[1], D), safe_div(s.drgb[2], D), safe_div(s.drgb[3], D)};
  R L_raw = area + s.logT;
  R e1 = safe_exp(-area), e2 = safe_exp(-s.logT);
  R w_raw = (R(1) - e1) * e2;
  R w = clipv<R>(w_raw, R(0), R(1));
  R mid = (t + s.t) / R(2) * R(kPreMulti);
  R m = ladder(mid);

  R wbar = nb.C[0] * rgbn[0] + nb.C[1] * rgbn[1] + nb.C[2] * rgbn[2] + nb.dpx * R(2) * m * s.csx +
           nb.dpy * R(2) * s.csy + nb.csx + nb.csy * m;
  R mbar = nb.dpx * R(2) * w * s.csx + nb.csy * w;
  sb = State<R>{};
  for (int i = 0; i < 3; ++i) sb.C[i] = nb.C[i];
  sb.dpx = nb.dpx;
  sb.dpy = nb.dpy;
  sb.csx = nb.csx + nb.dpx * R(2) * w * m;
  sb.csy = nb.csy + nb.dpy * R(2) * w;
  for (int i = 0; i < 4; ++i) {
    sb.drgb[i] = nb.drgb[i];
    diracbar[i] = nb.drgb[i];
  }
  tbar = nb.t;
  R midbar = mbar * ladder_grad(mid);
  tbar += midbar * R(kPreMulti) / R(2);
  sb.t += midbar * R(kPreMulti) / R(2);
  // w = clip(w_raw, 0, 1) = max(min(w_raw, 1), 0)
  R wrawbar = (w_raw < R(1) && std::min(w_raw, R(1)) > R(0)) ? wbar : R(0);
  R e1bar = -wrawbar * e2;
  R e2bar = wrawbar * (R(1) - e1);
  R areabar = -e1bar * e1;
  sb.logT += -e2bar * e2;
  R Lrawbar = (L_raw > R(0)) ? nb.logT : R(0);
  areabar += Lrawbar;
  sb.logT += Lrawbar;
  // rgbn = safe_div(R, D)
  R Dbar = 0;
  for (int i = 0; i < 3; ++i) {
    R da, db;
    bw_safe_div(s.drgb[1 + i], D, nb.C[i] * w, da, db);
    sb.drgb[1 + i] += da;
    Dbar += db;
  }
  R arawbar = (a_raw > R(0)) ? areabar : R(0);
  Dbar += arawbar * dt;
  R dtbar = arawbar * D;
  sb.drgb[0] += Dbar;
  R dtrawbar = (dt_raw > R(0)) ? dtbar : R(0);
  tbar += dtrawbar;
  sb.t -= dtrawbar;
}

// Simple binary BVH over AABBs (median split) used only to accelerate candidate enumeration at large N.
struct Bvh {
  struct Node {
    float lo[3], hi[3];
    int left, right, first, count;
  };
  std::vector<Node> nodes;
  std::vector<int> order;
};

inline void bvh_build(Bvh& bvh, const float* aabb, int64_t n) {
  bvh.order.resize(n);
  for (int64_t i = 0; i < n; ++i) bvh.order[i] = (int)i;
  bvh.nodes.clear();
  bvh.nodes.reserve(2 * n + 1);
  struct Item { int node, first, count; };
  std::vector<Item> stack;
  bvh.nodes.push_back({});
  stack.push_back({0, 0, (int)n});
  while (!stack.empty()) {
    Item it = stack.back();
    stack.pop_back();
    Bvh::Node nd{};
    for (int k = 0; k < 3; ++k) { nd.lo[k] = 1e30f; nd.hi[k] = -1e30f; }
    float clo[3] = {1e30f, 1e30f, 1e30f}, chi[3] = {-1e30f, -1e30f, -1e30f};
    for (int i = it.first; i < it.first + it.count; ++i) {
      const float* b = aabb + 6 * (int64_t)bvh.order[i];
      for (int k = 0; k < 3; ++k) {
        nd.lo[k] = std::min(nd.lo[k], b[k]);
        nd.hi[k] = std::max(nd.hi[k], b[3 + k]);
        float c = 0.5f * (b[k] + b[3 + k]);
        clo[k] = std::min(clo[k], c);
        chi[k] = std::max(chi[k], c);
      }
    }
    if (it.count <= 4) {
      nd.left = nd.right = -1;
      nd.first = it.first;
      nd.count = it.count;
      bvh.nodes[it.node] = nd;
      continue;
    }
    int axis = 0;
    for (int k = 1; k < 3; ++k)
      if (chi[k] - clo[k] > chi[axis] - clo[axis]) axis = k;
    int mid = it.first + it.count / 2;
    std::nth_element(bvh.order.begin() + it.first, bvh.order.begin() + mid, bvh.order.begin() + it.first + it.count,
                     [&](int a, int b) {
                       const float* ba = aabb + 6 * (int64_t)a;
                       const float* bb = aabb + 6 * (int64_t)b;
                       return ba[axis] + ba[3 + axis] < bb[axis] + bb[3 + axis];
                     });
    nd.left = (int)bvh.nodes.size();
    bvh.nodes.push_back({});
    nd.right = (int)bvh.nodes.size();
    bvh.nodes.push_back({});
    nd.first = 0;
    nd.count = 0;
    bvh.nodes[it.node] = nd;
    stack.push_back({nd.left, it.first, mid - it.first});
    stack.push_back({nd.right, mid, it.first + it.count - mid});
  }
}

template <class R> struct Scene {
  const R *means, *scales, *quats, *dens, *feat;
  int64_t n;
  int d;  // SH coefficients per primitive
};

// AABB of one primitive (create_aabbs.cu:69-115): half-extent_i = sqrt(sum_j (R_ij s_j)^2), q normalized.
template <class R> void prim_aabb(const Scene<R>& sc, int64_t i, float* out) {
  const R* q0 = sc.quats + 4 * i;
  double qn = std::sqrt((double)q0[0] * q0[0] + (double)q0[1] * q0[1] + (double)q0[2] * q0[2] + (double)q0[3] * q0[3]);
  double r = q0[0] / qn, x = q0[1] / qn, y = q0[2] / qn, z = q0[3] / qn;
  double Rm[3][3] = {{1 - 2 * (y * y + z * z), 2 * (x * y - r * z), 2 * (x * z + r * y)},
                     {2 * (x * y + r * z), 1 - 2 * (x * x + z * z), 2 * (y * z - r * x)},
                     {2 * (x * z - r * y), 2 * (y * z + r * x), 1 - 2 * (x * x + y * y)}};
  const R* s = sc.scales + 3 * i;
  const R* m = sc.means + 3 * i;
  for (int k = 0; k < 3; ++k) {
    double e = std::sqrt(Rm[k][0] * Rm[k][0] * s[0] * s[0] + Rm[k][1] * Rm[k][1] * s[1] * s[1] +
                         Rm[k][2] * Rm[k][2] * s[2] * s[2]);
    e = e * (1.0 + 1e-5) + 1e-7;  // conservative for the candidate search only
    out[k] = (float)(m[k] - e);
    out[3 + k] = (float)(m[k] + e);
  }
}

// Containment test of initialize_density.cu:126-134 (normalized q, conjugate rotation, divide by scale).
template <class R> bool inside_fwd(const Scene<R>& sc, int64_t i, V3<R> p) {
  const R* q0 = sc.quats + 4 * i;
  R qn = std::sqrt(q0[0] * q0[0] + q0[1] * q0[1] + q0[2] * q0[2] + q0[3] * q0[3]);
  R q[4] = {q0[0] / qn, q0[1] / qn, q0[2] / qn, q0[3] / qn};
  V3<R> v = rotate(p - load3(sc.means + 3 * i), q);
  const R* s = sc.scales + 3 * i;
  V3<R> tr = mk(v.x / s[0], v.y / s[1], v.z / s[2]);
  return dot(tr, tr) <= R(1);
}
// Containment test of backwards_kernel.slang:395-405 (safe_div by |q|, scales clipped at 1e-8, length <= 1).
template <class R> bool inside_bwd(const Scene<R>& sc, int64_t i, V3<R> p) {
  const R* q0 = sc.quats + 4 * i;
  R qn = std::sqrt(q0[0] * q0[0] + q0[1] * q0[1] + q0[2] * q0[2] + q0[3] * q0[3]);
  R q[4] = {safe_div(q0[0], qn), safe_div(q0[1], qn), safe_div(q0[2], qn), safe_div(q0[3], qn)};
  V3<R> v = rotate(p - load3(sc.means + 3 * i), q);
  const R* s = sc.scales + 3 * i;
  V3<R> tr = mk(safe_div(v.x, std::max(s[0], R(1e-8))), safe_div(v.y, std::max(s[1], R(1e-8))),
                safe_div(v.z, std::max(s[2], R(1e-8))));
  return std::sqrt(dot(tr, tr)) <= R(1);
}

template <class R> struct Tracer {
  Scene<R> sc;
  const Bvh* bvh;  // nullable: brute force
  std::vector<float> aabb;
  R tmin, tmax;
  int max_iters, kbuf;
  bool safe;

  // Candidate enumeration restating the any-hit k-buffer (shaders.slang:84-164): collect the kbuf smallest
  // events after the cursor, ordered by (t, key) with key = 2*prim + (exit ? 1 : 0) (entry before exit of the
  // same primitive on equal t). The reference re-queries with a strict t > t_cur (shaders.slang:146), which
  // drops an event whose t equals the last processed one when the two straddle a k-buffer boundary; the
  // (t, key) cursor processes tied events exactly once and agrees with the reference everywhere else.
  int query(V3<R> o, V3<R> d, R cur_t, int64_t cur_key, std::vector<std::pair<R, int64_t>>& buf,
            std::vector<int>& stack) const {
    buf.clear();
    auto valid = [&](R t, int64_t key) { return t > cur_t || (t == cur_t && key > cur_key); };
    auto consider = [&](int64_t i) {
      V3<R> ro = o - load3(sc.means + 3 * i);
      const R* q = sc.quats + 4 * i;
      V3<R> Tro = rotate(ro, q);
      V3<R> Trd = safe ? l2n(rotate(d, q)) : rotate(d, q);
      R t0, t1;
      eli_roots(Tro, Trd, load3(sc.scales + 3 * i), safe, t0, t1);
      if (t1 < cur_t) return;
      const int64_t ke = 2 * i, kx = 2 * i + 1;
      if (valid(t0, ke)) {
        if (t0 > tmax) return;  // optixReportHit outside [tmin, tmax] is rejected
        buf.push_back({t0, ke});
        buf.push_back({t1, kx});
      } else if (valid(t1, kx)) {
        if (t1 > tmax) return;
        buf.push_back({t1, kx});
      }
    };
    if (!bvh) {
      for (int64_t i = 0; i < sc.n; ++i) consider(i);
    } else {
      // AABB slab test against [cur_t, tmax]; conservative superset of the primitives the roots accept.
      R inv[3] = {R(1) / d.x, R(1) / d.y, R(1) / d.z};
      R oo[3] = {o.x, o.y, o.z};
      stack.clear();
      stack.push_back(0);
      while (!stack.empty()) {
        const Bvh::Node& nd = bvh->nodes[stack.back()];
        stack.pop_back();
        R lo = cur_t, hi = tmax;
        bool hit = true;
        for (int k = 0; k < 3 && hit; ++k) {
          R ta = (R(nd.lo[k]) - oo[k]) * inv[k], tb = (R(nd.hi[k]) - oo[k]) * inv[k];
          if (std::isnan(ta) || std::isnan(tb)) { ta = -INFINITY; tb = INFINITY; }
          if (ta > tb) std::swap(ta, tb);
          lo = std::max(lo, ta);
          hi = std::min(hi, tb);
          if (lo > hi) hit = false;
        }
        if (!hit) continue;
        if (nd.left < 0) {
          for (int j = nd.first; j < nd.first + nd.count; ++j) consider(bvh->order[j]);
        } else {
          stack.push_back(nd.left);
          stack.push_back(nd.right);
        }
      }
    }
    // Exit events may lie beyond tmax when their entry was accepted in the same query (shaders.slang:138-158).
    std::sort(buf.begin(), buf.end());
    if ((int)buf.size() > kbuf) buf.resize(kbuf);
    return (int)buf.size();
  }

  R dc_color(int64_t prim, int c) const { return R(kShC0) * sc.feat[prim * sc.d * 3 + c] + R(0.5); }

  // raygen loop (shaders.slang:171-235). Returns the processed event ids (2*prim + entry) and per-event t.
  void trace(V3<R> rayo, V3<R> rayd, const R* init_drgb, State<R>& st, std::vector<int>& ids,
             std::vector<R>* ts) const {
    V3<R> d = l2n(rayd);
    V3<R> o = rayo + tmin * rayd;
    st = State<R>{};
    for (int i = 0; i < 4; ++i) st.drgb[i] = init_drgb[i];
    ids.clear();
    if (ts) ts->clear();
    std::vector<std::pair<R, int64_t>> buf;
    std::vector<int> stack;
    int iter = 0;
    R cur_t = 0;
    int64_t cur_key = INT64_MAX;  // nothing at t == 0 is an event (t > t_cur initially)
    while (st.logT < R(kLogCutoff) && iter < max_iters) {
      query(o, d, cur_t, cur_key, buf, stack);
      bool end = false;
      for (int k = 0; k < kbuf; ++k) {
        if (k >= (int)buf.size() || buf[k].first > R(1e9)) {
          end = true;
          break;
        }
        const int64_t key = buf[k].second;
        const int64_t prim = key / 2;
        const bool entry = (key & 1) == 0;
        R h = entry ? sc.dens[prim] : -sc.dens[prim];
        R dirac[4] = {h, h * dc_color(prim, 0), h * dc_color(prim, 1), h * dc_color(prim, 2)};
        st = update(st, buf[k].first, dirac);
        ids.push_back((int)(2 * prim + (entry ? 1 : 0)));
        if (ts) ts->push_back(buf[k].first);
        cur_t = buf[k].first;
        cur_key = key;
        ++iter;
        if (!(st.logT < R(kLogCutoff) && iter < max_iters)) break;
      }
      if (end) break;
    }
  }
};

template <class R>
int forward_impl(const R* means, const R* scales, const R* quats, const R* dens, const R* feat, int64_t n, int d,
                 const R* rayo, const R* rayd, int64_t nr, R tmin, R tmax, int max_iters, int kbuf, int use_bvh,
                 int safe, R* out_color, R* out_state, int32_t* out_iters, int32_t* tri, int32_t* touch,
                 R* init_drgb, int32_t* init_inds, int32_t* init_count, int64_t ray_stride) {
  Tracer<R> tr;
  tr.sc = {means, scales, quats, dens, feat, n, d};
  tr.tmin = tmin;
  tr.tmax = tmax;
  tr.max_iters = max_iters;
  tr.kbuf = kbuf;
  tr.safe = safe != 0;
  tr.aabb.resize(6 * n);
  for (int64_t i = 0; i < n; ++i) prim_aabb(tr.sc, i, &tr.aabb[6 * i]);
  Bvh bvh;
  tr.bvh = nullptr;
  if (use_bvh && n > 0) {
    bvh_build(bvh, tr.aabb.data(), n);
    tr.bvh = &bvh;
  }
  // kern_prefilter (initialize_density.cu:69-105): AABBs within tmin of ray 0's origin.
  std::vector<int64_t> cand;
  if (nr > 0) {
    V3<R> r0 = load3(rayo);
    for (int64_t i = 0; i < n; ++i) {
      const float* b = &tr.aabb[6 * i];
      // exact (unpadded) AABB for the prefilter
      float eb[6];
      {
        const R* q0 = quats + 4 * i;
        double qn = std::sqrt((double)q0[0] * q0[0] + (double)q0[1] * q0[1] + (double)q0[2] * q0[2] + (double)q0[3] * q0[3]);
        double r = q0[0] / qn, x = q0[1] / qn, y = q0[2] / qn, z = q0[3] / qn;
        double Rm[3][3] = {{1 - 2 * (y * y + z * z), 2 * (x * y - r * z), 2 * (x * z + r * y)},
                           {2 * (x * y + r * z), 1 - 2 * (x * x + z * z), 2 * (y * z - r * x)},
                           {2 * (x * z - r * y), 2 * (y * z + r * x), 1 - 2 * (x * x + y * y)}};
        for (int k = 0; k < 3; ++k) {
          double e = std::sqrt(Rm[k][0] * Rm[k][0] * scales[3 * i] * scales[3 * i] +
                               Rm[k][1] * Rm[k][1] * scales[3 * i + 1] * scales[3 * i + 1] +
                               Rm[k][2] * Rm[k][2] * scales[3 * i + 2] * scales[3 * i + 2]);
          eb[k] = (float)(means[3 * i + k] - e);
          eb[3 + k] = (float)(means[3 * i + k] + e);
        }
      }
      (void)b;
      double dmin = 0, ro[3] = {r0.x, r0.y, r0.z};
      for (int k = 0; k < 3; ++k) {
        if (ro[k] < eb[k]) dmin += (ro[k] - eb[k]) * (ro[k] - eb[k]);
        else if (ro[k] > eb[3 + k]) dmin += (ro[k] - eb[3 + k]) * (ro[k] - eb[3 + k]);
      }
      if (dmin <= (double)tmin * tmin) cand.push_back(i);
    }
  }
  if (init_count) *init_count = (int32_t)cand.size();
  if (init_inds)
    for (size_t j = 0; j < cand.size(); ++j) init_inds[j] = (int32_t)cand[j];
  if (touch) std::memset(touch, 0, sizeof(int32_t) * n);
  std::vector<int32_t> touch_acc(touch ? n : 0, 0);

#pragma omp parallel
  {
    std::vector<int> ids;
    std::vector<int32_t> local_touch;
#pragma omp for schedule(dynamic, 16)
    for (int64_t r = 0; r < nr; r += ray_stride) {
      V3<R> ro = load3(rayo + 3 * r), rd = load3(rayd + 3 * r);
      // kern_initialize_density (initialize_density.cu:107-150): point rayo + tmin * normalize(rayd); colour
      // read at flat index prim*3 + c as the reference does.
      R idr[4] = {0, 0, 0, 0};
      V3<R> p = ro + tmin * l2n(rd);
      for (int64_t c : cand) {
        if (inside_fwd(tr.sc, c, p)) {
          R s = dens[c];
          idr[0] += s;
          for (int k = 0; k < 3; ++k) idr[1 + k] += s * (feat[c * 3 + k] * R(kShC0) + R(0.5));
        }
      }
      if (init_drgb)
        for (int k = 0; k < 4; ++k) init_drgb[4 * r + k] = idr[k];
      State<R> st;
      tr.trace(ro, rd, idr, st, ids, nullptr);
      int E = (int)ids.size();
      out_color[5 * r + 0] = st.C[0];
      out_color[5 * r + 1] = st.C[1];
      out_color[5 * r + 2] = st.C[2];
      out_color[5 * r + 3] = R(1) - std::exp(-st.logT);
      out_color[5 * r + 4] = st.dpx - st.dpy;
      if (out_state) state_to16(st, out_state + 16 * r);
      if (out_iters) out_iters[r] = E;
      if (tri)
        for (int k = 0; k < E; ++k) tri[r + (int64_t)k * nr] = ids[k];
      if (touch)
        for (int k = 0; k < E; ++k) {
#pragma omp atomic
          touch[ids[k] / 2] += 1;
        }
    }
  }
  return 0;
}

template <class R>
int backward_impl(const R* means, const R* scales, const R* quats, const R* dens, const R* feat, int64_t n, int d,
                  const R* rayo, const R* rayd, int64_t nr, R tmin, int max_iters, const R* grad_out,
                  const int32_t* iters, const int32_t* tri, const R* init_drgb, const int32_t* init_inds,
                  int32_t init_count, const R* wct, int64_t n_wct, R* dmeans, R* dscales, R* dquats, R* ddens,
                  R* dfeat, R* drayo, R* drayd, R* dmeans2D, R* dinit_out, int apply_clamp) {
  Scene<R> sc{means, scales, quats, dens, feat, n, d};
  std::memset(dmeans, 0, sizeof(R) * 3 * n);
  std::memset(dscales, 0, sizeof(R) * 3 * n);
  std::memset(dquats, 0, sizeof(R) * 4 * n);
  std::memset(ddens, 0, sizeof(R) * n);
  std::memset(dfeat, 0, sizeof(R) * 3 * d * n);
  std::memset(drayo, 0, sizeof(R) * 3 * nr);
  std::memset(drayd, 0, sizeof(R) * 3 * nr);
  if (dmeans2D) std::memset(dmeans2D, 0, sizeof(R) * 2 * n);
  std::vector<R> dinit(4 * nr, R(0));
  // inverse of each world-to-camera matrix (safe-math.slang:187-225 computes the plain 4x4 inverse).
  std::vector<double> inv(16 * std::max<int64_t>(n_wct, 1), 0.0);
  for (int64_t w = 0; wct && w < n_wct; ++w) {
    double m[4][4], a[4][8];
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j) m[i][j] = wct[16 * w + 4 * i + j];
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 8; ++j) a[i][j] = j < 4 ? m[i][j] : (j - 4 == i ? 1.0 : 0.0);
    for (int c = 0; c < 4; ++c) {
      int piv = c;
      for (int r2 = c + 1; r2 < 4; ++r2)
        if (std::fabs(a[r2][c]) > std::fabs(a[piv][c])) piv = r2;
      for (int j = 0; j < 8; ++j) std::swap(a[c][j], a[piv][j]);
      double p = a[c][c];
      for (int j = 0; j < 8; ++j) a[c][j] /= p;
      for (int r2 = 0; r2 < 4; ++r2)
        if (r2 != c) {
          double f = a[r2][c];
          for (int j = 0; j < 8; ++j) a[r2][j] -= f * a[c][j];
        }
    }
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j) inv[16 * w + 4 * i + j] = a[i][j + 4];
  }

  const int nthreads = omp_get_max_threads();
  std::vector<std::vector<R>> acc(nthreads);
  const int64_t stride_prim = 3 + 3 + 4 + 1 + 3 + 2;  // mean, scale, quat, density, colour, means2D
#pragma omp parallel
  {
    int tid = omp_get_thread_num();
    std::vector<R>& A = acc[tid];
    A.assign(stride_prim * n, R(0));
    std::vector<State<R>> states;
    std::vector<R> ts;
    std::vector<R> dirs;
#pragma omp for schedule(dynamic, 16)
    for (int64_t r = 0; r < nr; ++r) {
      int E = iters[r];
      if (E >= max_iters - 1 || E <= 0) continue;  // backwards_kernel.slang:304
      V3<R> dir = load3(rayd + 3 * r);              // raw direction (backwards_kernel.slang:288-289)
      V3<R> org = load3(rayo + 3 * r) + tmin * dir;
      // Recompute control points with safe_intersect and the forward states from the initial density.
      states.resize(E + 1);
      ts.resize(E);
      dirs.resize(4 * E);
      State<R> st{};
      for (int k = 0; k < 4; ++k) st.drgb[k] = init_drgb[4 * r + k];
      states[0] = st;
      for (int k = 0; k < E; ++k) {
        int id = tri[r + (int64_t)k * nr];
        int64_t p = id / 2;
        int face = id % 2;
        V3<R> ro = org - load3(means + 3 * p);
        const R* q = quats + 4 * p;
        R t0, t1;
        eli_roots(rotate(ro, q), l2n(rotate(dir, q)), load3(scales + 3 * p), true, t0, t1);
        ts[k] = face == 1 ? t0 : t1;
        R h = face == 1 ? dens[p] : -dens[p];
        R col[3];
        for (int c = 0; c < 3; ++c) col[c] = R(kShC0) * feat[p * d * 3 + c] + R(0.5);
        dirs[4 * k] = h;
        for (int c = 0; c < 3; ++c) dirs[4 * k + 1 + c] = h * col[c];
        states[k + 1] = update(states[k], ts[k], &dirs[4 * k]);
      }
      // bwd of extract_color (spline-machine.slang:146-153)
      const State<R>& fin = states[E];
      State<R> nb{};
      for (int c = 0; c < 3; ++c) nb.C[c] = grad_out[5 * r + c];
      nb.logT = grad_out[5 * r + 3] * std::exp(-fin.logT);
      nb.dpx = grad_out[5 * r + 4];
      nb.dpy = -grad_out[5 * r + 4];
      V3<R> orgbar = mk<R>(0, 0, 0), dirbar = mk<R>(0, 0, 0);
      const double* iv = nullptr;
      if (wct && n_wct > 0) iv = &inv[16 * (r < n_wct ? r : 0)];
      for (int k = E - 1; k >= 0; --k) {
        State<R> sb;
        R tbar, dbar[4];
        update_bwd(states[k], ts[k], nb, sb, tbar, dbar);
        int id = tri[r + (int64_t)k * nr];
        int64_t p = id / 2;
        int face = id % 2;
        R sgn_face = face == 1 ? R(1) : R(-1);
        R* a = &A[stride_prim * p];
        // dirac = +-density * (1, colour)
        R col[3];
        for (int c = 0; c < 3; ++c) col[c] = R(kShC0) * feat[p * d * 3 + c] + R(0.5);
        a[10] += sgn_face * (dbar[0] + dbar[1] * col[0] + dbar[2] * col[1] + dbar[3] * col[2]);
        for (int c = 0; c < 3; ++c) a[11 + c] += sgn_face * dens[p] * dbar[1 + c];
        // t = selected root of safe_intersect
        V3<R> ro_bar = mk<R>(0, 0, 0), rd_bar = mk<R>(0, 0, 0), s_bar = mk<R>(0, 0, 0);
        R q_bar[4] = {0, 0, 0, 0};
        V3<R> ro = org - load3(means + 3 * p);
        safe_root_bwd(ro, dir, load3(scales + 3 * p), quats + 4 * p, face, tbar, ro_bar, rd_bar, s_bar, q_bar);
        a[0] -= ro_bar.x; a[1] -= ro_bar.y; a[2] -= ro_bar.z;
        a[3] += s_bar.x; a[4] += s_bar.y; a[5] += s_bar.z;
        for (int c = 0; c < 4; ++c) a[6 + c] += q_bar[c];
        orgbar = orgbar + ro_bar;
        dirbar = dirbar + rd_bar;
        if (iv) {
          // means2D heuristic (backwards_kernel.slang:213-220): bwd of inv_project w.r.t. xy, upstream = dL/dmean
          // of this event; dist = z of (mean,1) * wct.
          const R* W = wct + 16 * (r < n_wct ? r : 0);
          R dist = means[3 * p] * W[2] + means[3 * p + 1] * W[6] + means[3 * p + 2] * W[10] + W[14];
          R mb[3] = {-ro_bar.x, -ro_bar.y, -ro_bar.z};
          for (int k2 = 0; k2 < 2; ++k2) {
            double pb = iv[4 * k2 + 0] * mb[0] + iv[4 * k2 + 1] * mb[1] + iv[4 * k2 + 2] * mb[2];
            a[14 + k2] += R(pb) * dist;
          }
        }
        nb = sb;
      }
      for (int c = 0; c < 4; ++c) dinit[4 * r + c] = nb.drgb[c];
      drayo[3 * r + 0] = orgbar.x; drayo[3 * r + 1] = orgbar.y; drayo[3 * r + 2] = orgbar.z;
      drayd[3 * r + 0] = dirbar.x; drayd[3 * r + 1] = dirbar.y; drayd[3 * r + 2] = dirbar.z;
    }
    // start-inside adjoint (backwards_kernel.slang:374-428): ray 0's origin + tmin * raw direction of ray r.
if (init_count > 0) {
#pragma omp for schedule(static)
    for (int64_t r = 0; r < nr; ++r) {
      V3<R> rd = load3(rayd + 3 * r);
      V3<R> p0 = load3(rayo) + tmin * rd;
      for (int j = 0; j < init_count; ++j) {
        int64_t p = init_inds[j];
        if (!inside_bwd(sc, p, p0)) continue;
        R* a = &A[stride_prim * p];
        R col[3];
        for (int c = 0; c < 3; ++c) col[c] = R(kShC0) * feat[p * d * 3 + c] + R(0.5);
        const R* g = &dinit[4 * r];
        a[10] += g[0] + g[1] * col[0] + g[2] * col[1] + g[3] * col[2];
        for (int c = 0; c < 3; ++c) a[11 + c] += dens[p] * g[1 + c];
      }
    }
    }
  }
  for (int t = 0; t < nthreads; ++t) {
    if (acc[t].empty()) continue;
    for (int64_t p = 0; p < n; ++p) {
      const R* a = &acc[t][stride_prim * p];
      for (int c = 0; c < 3; ++c) dmeans[3 * p + c] += a[c];
      for (int c = 0; c < 3; ++c) dscales[3 * p + c] += a[3 + c];
      for (int c = 0; c < 4; ++c) dquats[4 * p + c] += a[6 + c];
      ddens[p] += a[10];
      for (int c = 0; c < 3; ++c) dfeat[p * d * 3 + c] += R(kShC0) * a[11 + c];  // d colour / d f0 = SH_C0
      if (dmeans2D) {
        dmeans2D[2 * p] += a[14];
        dmeans2D[2 * p + 1] += a[15];
      }
    }
  }
  if (dinit_out) std::memcpy(dinit_out, dinit.data(), sizeof(R) * 4 * nr);
  if (apply_clamp) {
    // fast_ellipsoid_splinetracer.py:201-211
    for (int64_t i = 0; i < 3 * n; ++i) {
      dmeans[i] = clipv<R>(dmeans[i], -1e3, 1e3);
      dscales[i] = clipv<R>(dscales[i], -1e3, 1e3);
    }
    for (int64_t i = 0; i < 4 * n; ++i) dquats[i] = clipv<R>(dquats[i], -1e3, 1e3);
    for (int64_t i = 0; i < n; ++i) ddens[i] = clipv<R>(ddens[i], -50, 50);
    for (int64_t i = 0; i < 3 * d * n; ++i) dfeat[i] = clipv<R>(dfeat[i], -1e3, 1e3);
    for (int64_t i = 0; i < 3 * nr; ++i) {
      drayo[i] = clipv<R>(drayo[i], -1e3, 1e3);
      drayd[i] = clipv<R>(drayd[i], -1e3, 1e3);
    }
  }
  return 0;
}

// Real spherical-harmonic colour (sh.slang:48-103; PlenOctree constants sh.slang:18-27), unclamped.
template <class R> void eval_sh_one(const R* f, int deg, R x, R y, R z, R* out) {
  static const double C1 = 0.4886025119029199;
  static const double C2[5] = {1.0925484305920792, -1.0925484305920792, 0.31539156525252005, -1.0925484305920792,
                               0.5462742152960396};
  static const double C3[7] = {-0.5900435899266435, 2.890611442640554, -0.4570457994644658, 0.3731763325901154,
                               -0.4570457994644658, 1.445305721320277, -0.5900435899266435};
  R b[16];
  b[0] = R(kShC0);
  if (deg > 0) {
    b[1] = R(-C1) * y; b[2] = R(C1) * z; b[3] = R(-C1) * x;
  }
  if (deg > 1) {
    R xx = x * x, yy = y * y, zz = z * z, xy = x * y, yz = y * z, xz = x * z;
    b[4] = R(C2[0]) * xy; b[5] = R(C2[1]) * yz; b[6] = R(C2[2]) * (R(2) * zz - xx - yy);
    b[7] = R(C2[3]) * xz; b[8] = R(C2[4]) * (xx - yy);
    if (deg > 2) {
      b[9] = R(C3[0]) * y * (R(3) * xx - yy); b[10] = R(C3[1]) * xy * z;
      b[11] = R(C3[2]) * y * (R(4) * zz - xx - yy); b[12] = R(C3[3]) * z * (R(2) * zz - R(3) * xx - R(3) * yy);
      b[13] = R(C3[4]) * x * (R(4) * zz - xx - yy); b[14] = R(C3[5]) * z * (xx - yy);
      b[15] = R(C3[6]) * x * (xx - R(3) * yy);
    }
  }
  int nc = (deg + 1) * (deg + 1);
  for (int c = 0; c < 3; ++c) {
    R acc = R(0.5);
    for (int k = 0; k < nc; ++k) acc += b[k] * f[3 * k + c];
    out[c] = acc;
  }
}

}  // namespace

extern "C" {

int oracle_forward_f64(const double* means, const double* scales, const double* quats, const double* dens,
                       const double* feat, int64_t n, int d, const double* rayo, const double* rayd, int64_t nr,
                       double tmin, double tmax, int max_iters, int kbuf, int use_bvh, int safe, double* out_color,
                       double* out_state, int32_t* out_iters, int32_t* tri, int32_t* touch, double* init_drgb,
                       int32_t* init_inds, int32_t* init_count) {
  return forward_impl<double>(means, scales, quats, dens, feat, n, d, rayo, rayd, nr, tmin, tmax, max_iters, kbuf,
                              use_bvh, safe, out_color, out_state, out_iters, tri, touch, init_drgb, init_inds,
                              init_count, 1);
}

// FP32 instantiation, used as the timed CPU baseline. ray_stride > 1 traces a deterministic subsample.
int oracle_forward_f32(const float* means, const float* scales, const float* quats, const float* dens,
                       const float* feat, int64_t n, int d, const float* rayo, const float* rayd, int64_t nr,
                       float tmin, float tmax, int max_iters, int kbuf, int use_bvh, int safe, float* out_color,
                       float* out_state, int32_t* out_iters, int32_t* tri, int32_t* touch, float* init_drgb,
                       int32_t* init_inds, int32_t* init_count, int64_t ray_stride) {
  return forward_impl<float>(means, scales, quats, dens, feat, n, d, rayo, rayd, nr, tmin, tmax, max_iters, kbuf,
                             use_bvh, safe, out_color, out_state, out_iters, tri, touch, init_drgb, init_inds,
                             init_count, ray_stride);
}

int oracle_backward_f64(const double* means, const double* scales, const double* quats, const double* dens,
                        const double* feat, int64_t n, int d, const double* rayo, const double* rayd, int64_t nr,
                        double tmin, int max_iters, const double* grad_out, const int32_t* iters, const int32_t* tri,
                        const double* init_drgb, const int32_t* init_inds, int32_t init_count, const double* wct,
                        int64_t n_wct, double* dmeans, double* dscales, double* dquats, double* ddens, double* dfeat,
                        double* drayo, double* drayd, double* dmeans2D, double* dinit, int apply_clamp) {
  return backward_impl<double>(means, scales, quats, dens, feat, n, d, rayo, rayd, nr, tmin, max_iters, grad_out,
                               iters, tri, init_drgb, init_inds, init_count, wct, n_wct, dmeans, dscales, dquats,
                               ddens, dfeat, drayo, drayd, dmeans2D, dinit, apply_clamp);
}

int oracle_backward_f32(const float* means, const float* scales, const float* quats, const float* dens,
                        const float* feat, int64_t n, int d, const float* rayo, const float* rayd, int64_t nr,
                        float tmin, int max_iters, const float* grad_out, const int32_t* iters, const int32_t* tri,
                        const float* init_drgb, const int32_t* init_inds, int32_t init_count, const float* wct,
                        int64_t n_wct, float* dmeans, float* dscales, float* dquats, float* ddens, float* dfeat,
                        float* drayo, float* drayd, float* dmeans2D, float* dinit, int apply_clamp) {
  return backward_impl<float>(means, scales, quats, dens, feat, n, d, rayo, rayd, nr, tmin, max_iters, grad_out,
                              iters, tri, init_drgb, init_inds, init_count, wct, n_wct, dmeans, dscales, dquats,
                              ddens, dfeat, drayo, drayd, dmeans2D, dinit, apply_clamp);
}

// Per-primitive SH colour at direction normalize(mean - origin) (sh_kernel.slang:19-42), FP64.
int oracle_eval_sh_f64(const double* means, const double* feat, int64_t n, int deg, int d, const double* origin,
                       double* out_rgb) {
  for (int64_t i = 0; i < n; ++i) {
    double v[3] = {means[3 * i] - origin[0], means[3 * i + 1] - origin[1], means[3 * i + 2] - origin[2]};
    double nn = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    eval_sh_one<double>(feat + 3 * d * i, deg, v[0] / nn, v[1] / nn, v[2] / nn, out_rgb + 3 * i);
  }
  return 0;
}

int oracle_num_threads(void) { return omp_get_max_threads(); }

}  // extern "C"
