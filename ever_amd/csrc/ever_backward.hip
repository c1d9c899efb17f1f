// Backward tracer: one lane per ray replays the ray's stored events in reverse, reconstructing each earlier
// state from the FP64 accumulators written by the forward (exact inverse of the forward's FP64 additions),
// applies the hand-derived adjoint of `update` and of the safe ellipsoid roots, and scatters per-primitive
// gradients (mean, scale, quaternion, density, colour) into a 64-byte per-primitive accumulator with float
// atomics. A per-primitive epilogue turns the colour gradient into SH-feature gradients, applies the
// reference's clamps and the means2D statistic.
//
// Semantics restated from the reference (splinetracers/fast_ellipsoid_splinetracer/slang/backwards_kernel.slang):
//   per-ray skip rule :303-304, bwd of extract_color :307, reverse loop :327-361 (bwd_diff(update) :154,
//   bwd_diff(safe_intersect) :172, colour :184, means2D :213-220), dL/d initial drgb :363-366,
//   start-inside adjoint :374-428; clamps splinetracers/fast_ellipsoid_splinetracer.py:201-211.
#include "ever_common.hpp"
#include "ever_internal.hpp"

namespace ever {

constexpr int kBwdBlock = 256;
constexpr int64_t kXcdChunk = 16;  // consecutive blocks given to one XCD
// per-block LDS table of the primitives the block's rays touch (a 256-ray screen tile touches about 50 at 1M/1080p);
// 256 slots of 14 FP64 fields keep the block's LDS at 30 KB, so that three blocks share a CU
constexpr int kHashSlots = 256;
constexpr int kAccFields = 14;    // mean 3, scale 3, quat 4, density 1, colour 3
constexpr uint32_t kEmptySlot = 0xffffffffu;

struct BwdArgs {
  const float* means;
  const float* scales;
  const float* quats;
  const float4* colr;
  const float* rayo;
  const float* rayd;
  int64_t nr;
  float tmin;
  int max_iters;
  const float* grad_out;
  const int32_t* iters;
  const int32_t* tri;
  const float* states;
  const double* st64;
  const float* wcts;
  int64_t nw;
  const float* inv;  // (nw,16) inverses
  float* acc;        // (N,16)
  float* drayo;
  float* drayd;
  float* dinit;      // (R,4) scratch or caller buffer
  const int32_t* order;  // processing order (nullable: identity)
  int32_t events_by_slot;  // tri column = processing slot instead of ray index
  int64_t ray0;            // the slots cover rays ray0 .. ray0 + nr - 1 (order entries are relative to ray0)
  int64_t tri_stride;      // row stride of the event list (one row per event index)
  // several views: per-view (density, colour) records, each ray's view, per-view colour gradients (V,N,4)
  const float4* vcol;
  const int32_t* view_index;
  float* vgrad;
  int64_t n;
};

// Accumulator key of a primitive seen from a view: the primitive itself, or view * N + primitive with several
// views (whose colour gradients are kept per view). Fields 0..10 (geometry, density) always go to the
// primitive's accumulator, fields 11..13 (colour) to the view's when there are several.
__device__ __forceinline__ void add_global(const BwdArgs& A, uint32_t key, int j, float v) {
  if (A.vgrad) {
    if (j >= 11) atomic_addf(A.vgrad + 4 * (int64_t)key + (j - 11), v);
    else atomic_addf(A.acc + 16 * (int64_t)(key % (uint32_t)A.n) + j, v);
  } else {
    atomic_addf(A.acc + 16 * (int64_t)key + j, v);
  }
}

// Accumulate one event's per-primitive gradient into the block's LDS table (open addressing, linear probing);
// a primitive that finds no slot within 16 probes goes straight to the global accumulator. The table is
// flushed to the global accumulator with one atomic per field and primitive at the end of the block, which
// turns the per-event atomics (contended: neighbouring rays hit the same primitives) into per-block ones.
// The table's fields are FP64: on gfx950 an LDS FP64 atomic add (ds_add_f64) costs far less LDS time than an
// FP32 one (measured: adjoint 3.1 against 4.5 ms at 1M/1080p), and the per-block sums gain precision.
__device__ __forceinline__ void scatter_prim(const BwdArgs& A, uint32_t* hkeys, double (*hvals)[kHashSlots], uint32_t p,
                                             const float (&v)[kAccFields]) {
  // find the slot first (the lanes of a wave take different numbers of probes), then add with all lanes together:
  // one copy of the 14 atomics, issued once per flush by every flushing lane of the wave
  uint32_t h = (p * 2654435761u) >> 24;  // 8 bits: kHashSlots = 256
  bool found = false;
#pragma unroll 1
  for (int probe = 0; probe < 16; ++probe) {
    uint32_t k = __hip_atomic_load(hkeys + h, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (k == kEmptySlot) {
      const uint32_t old = atomicCAS(hkeys + h, kEmptySlot, p);
      k = (old == kEmptySlot) ? p : old;
    }
    if (k == p) {
      found = true;
      break;
    }
    h = (h + 1) & (kHashSlots - 1);
  }
  if (found) {
#pragma unroll
    for (int j = 0; j < kAccFields; ++j) atomicAdd(&hvals[j][h], (double)v[j]);
  } else {
#pragma unroll
    for (int j = 0; j < kAccFields; ++j) add_global(A, p, j, v[j]);
  }
}

__global__ void k_zero_f32(float* p, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.f;
}

// Inverse of each 4x4 world-to-camera matrix (the reference inverts per ray, safe-math.slang:187-225), FP64
// Gauss-Jordan with partial pivoting.
__global__ void k_invert4(const float* __restrict__ w, int64_t nw, float* __restrict__ out) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nw) return;
  double a[4][8];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 8; ++j) a[i][j] = j < 4 ? (double)w[16 * k + 4 * i + j] : (j - 4 == i ? 1.0 : 0.0);
  for (int c = 0; c < 4; ++c) {
    int piv = c;
    for (int r = c + 1; r < 4; ++r)
      if (fabs(a[r][c]) > fabs(a[piv][c])) piv = r;
    for (int j = 0; j < 8; ++j) {
      double t = a[c][j];
      a[c][j] = a[piv][j];
      a[piv][j] = t;
    }
    double p = a[c][c];
    for (int j = 0; j < 8; ++j) a[c][j] /= p;
    for (int r = 0; r < 4; ++r)
      if (r != c) {
        double f = a[r][c];
        for (int j = 0; j < 8; ++j) a[r][j] -= f * a[c][j];
      }
  }
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) out[16 * k + 4 * i + j] = (float)a[i][j + 4];
}

// The primitive data one event needs, loaded an event ahead of its use (the loop below walks each ray's events
// back to front, and a load issued one iteration early has that iteration's arithmetic to arrive).
struct EventRaw {
  float4 q, cr;  // quaternion; (density, colour) of the event's view
  f3 mean, scl;
};
__device__ __forceinline__ void event_load(const BwdArgs& A, int32_t id, int32_t view, EventRaw& R) {
  const int64_t p = id >> 1;
  R.q = reinterpret_cast<const float4*>(A.quats)[p];
  R.cr = A.vcol ? A.vcol[(int64_t)view * A.n + p] : A.colr[p];
  R.mean = ld3(A.means + 3 * p);
  R.scl = ld3(A.scales + 3 * p);
}

// One event's primitive and its selected safe root: the FP64 depth and the root's gradient w.r.t. the ray origin
// relative to the mean, the raw direction, the scale and the raw quaternion (safe_roots_jac_d; the chain from the
// root's adjoint tb to these four is linear in tb, so an event's contribution is tb times it: the adjoint's clamps at
// +-1e20 bind only where such a product exceeds them). The forward accumulated its log-transmittance for this kernel
// with the same depths, so the states unwind exactly. The loop below needs them twice,
// as the depth of the previous event and then as the event whose adjoint is taken; they are computed once and
// carried over.
struct EventGeo {
  float4 cr;
  f3 mean;
  f3 jro, jrd, js;  // d root / d(origin - mean), d root / d raw direction, d root / d scale
  float4 jq;        // d root / d raw quaternion
  bool ok;          // the ray meets the ellipsoid
};
__device__ __forceinline__ double event_geo(d3 org, f3 dir, int32_t id, const EventRaw& R, EventGeo& G) {
  G.cr = R.cr;
  G.mean = R.mean;
  // entry (face 1): the smaller root, exit: the larger (safe_intersect, tri-intersect.slang:443-456)
  return safe_roots_jac_d(org - tod(G.mean), dir, R.scl, R.q, (id & 1) != 0, G.jro, G.jrd, G.js, G.jq, G.ok);
}

__device__ __forceinline__ void backward_ray(const BwdArgs& A, int64_t r, int64_t col, uint32_t* hkeys,
                                             double (*hvals)[kHashSlots]) {
  const int E = A.iters[r];
  float4* dinit = reinterpret_cast<float4*>(A.dinit) + r;
  if (E >= A.max_iters - 1 || E <= 0) {  // backwards_kernel.slang:304
    *dinit = make_float4(0.f, 0.f, 0.f, 0.f);
    float* o = A.drayo + 3 * r;
    float* dd = A.drayd + 3 * r;
    o[0] = o[1] = o[2] = 0.f;
    dd[0] = dd[1] = dd[2] = 0.f;
    return;
  }
  const f3 dir = ld3(A.rayd + 3 * r);  // raw direction (backwards_kernel.slang:288-289)
  const f3 rof = ld3(A.rayo + 3 * r);
  const d3 org = ray_origin_fp64(rof, dir, A.tmin);
  const float* g = A.grad_out + 5 * r;
  const float gC0 = g[0], gC1 = g[1], gC2 = g[2], gO = g[3], gD = g[4];
  const float4* st = reinterpret_cast<const float4*>(A.states + 16 * r);
  const float4 st0 = st[0];
  const double* s64 = A.st64 + 8 * r;
  // state after the current event (starts as the final state; the log-transmittance the forward accumulated with
  // this kernel's depths, so that unwinding it event by event returns exactly to each earlier value)
  double Da0 = s64[0], Da1 = s64[1], Da2 = s64[2], Da3 = s64[3], La = s64[6];
  float csxa = st0.z, csya = st0.w;
  // adjoints of the state after the current event
  const float dpxb = gD, dpyb = -gD;
  float Lb = gO * (float)exp(-La);
  float csxb = 0.f, csyb = 0.f, tcarry = 0.f;
  float Db0 = 0.f, Db1 = 0.f, Db2 = 0.f, Db3 = 0.f;
  f3 orgb = mk3(0.f, 0.f, 0.f), dirb = mk3(0.f, 0.f, 0.f);
  const float* inv = nullptr;
  const float* W = nullptr;
  if (A.nw > 1) {
    const int64_t wi = r < A.nw ? r : 0;  // backwards_kernel.slang:313
    inv = A.inv + 16 * wi;
    W = A.wcts + 16 * wi;
  }

  const int32_t view = A.vcol ? A.view_index[r] : 0;
  // software pipeline over the events k = E-1 .. 0: event ids are read three events ahead; the primitive data of
  // event k - 1, which iteration k needs first (its depth), is loaded during iteration k + 1
  const int32_t* tcol = A.tri + col;
  int32_t id = tcol[(int64_t)(E - 1) * A.tri_stride];
  int32_t id_next = E >= 2 ? tcol[(int64_t)(E - 2) * A.tri_stride] : 0;   // event k - 1
  int32_t id_next2 = E >= 3 ? tcol[(int64_t)(E - 3) * A.tri_stride] : 0;  // event k - 2
  EventGeo G;
  double tk;
  {
    EventRaw R0;
    event_load(A, id, view, R0);
    tk = event_geo(org, dir, id, R0, G);
  }
  EventRaw Rn;  // raw data of event k - 1
  if (E >= 2) event_load(A, id_next, view, Rn);
  for (int k = E - 1; k >= 0; --k) {
    int32_t id_next3 = 0;
    if (k >= 3) id_next3 = tcol[(int64_t)(k - 3) * A.tri_stride];
    int32_t id_prev = 0;
    double tprev = 0.0;
    EventGeo Gp;
    if (k > 0) {
      id_prev = id_next;
      tprev = event_geo(org, dir, id_prev, Rn, Gp);
    }
    // raw data of event k - 2 into the registers event k - 1's just left, for the next iteration
    if (k >= 2) event_load(A, id_next2, view, Rn);
    const int64_t p = id >> 1;
    const int face = id & 1;
    const float4 cr = G.cr;
    const float hgt = face ? cr.x : -cr.x;
    // reconstruct the state before event k (inverse of the forward's FP64 accumulation)
    const double Dbf0 = Da0 - (double)hgt, Dbf1 = Da1 - (double)(hgt * cr.y), Dbf2 = Da2 - (double)(hgt * cr.z),
                 Dbf3 = Da3 - (double)(hgt * cr.w);
    const float dt_raw = (float)(tk - tprev);
    const float dt = fmaxf(dt_raw, 0.f);
    const float Df = (float)Dbf0;
    const float a_raw = Df * dt;
    const float area = fmaxf(a_raw, 0.f);
    const double Lbf = fmax(La - (double)area, 0.0);
    const float e1 = expf(-area), om = -expm1f(-area);
    const float e2 = expf(-(float)Lbf);
    const float w_raw = om * e2;
    const float w = clipf(w_raw, 0.f, 1.f);
    const float mid = (float)((tk + tprev) * (0.5 * kPreMulti));
    const float m = ladder(mid);
    const float csxbf = csxa - w, csybf = csya - w * m;
    const float R1 = (float)Dbf1, R2 = (float)Dbf2, R3 = (float)Dbf3;
    // IEEE divisions here: gradients at high optical depth are sensitive to this step's rounding
    const float rn0 = safe_div(R1, Df), rn1 = safe_div(R2, Df), rn2 = safe_div(R3, Df);
    // adjoint of update (spline-machine.slang:103-138)
    const float wb = gC0 * rn0 + gC1 * rn1 + gC2 * rn2 + dpxb * 2.f * m * csxbf + dpyb * 2.f * csybf + csxb + csyb * m;
    const float mb = dpxb * 2.f * w * csxbf + csyb * w;
    const float dirac0 = Db0, dirac1 = Db1, dirac2 = Db2, dirac3 = Db3;  // adjoint of the new state's drgb
    const float csxb_n = csxb + dpxb * 2.f * w * m;
    const float csyb_n = csyb + dpyb * 2.f * w;
    const float midb = mb * ladder_grad(mid) * (0.5f * kPreMulti);
    float tb = tcarry + midb;
    float tprevb = midb;
    const float wrawb = (w_raw < 1.f && w_raw > 0.f) ? wb : 0.f;
    float areab = wrawb * e2 * e1;
    const float L_raw = area + (float)Lbf;
    float Lb_n = -(wrawb * om) * e2;
    if (L_raw > 0.f) {
      areab += Lb;
      Lb_n += Lb;
    }
    float Dbar = 0.f, da, db;
    bw_safe_div(R1, Df, gC0 * w, da, db);
    const float Db1_n = Db1 + da;
    Dbar += db;
    bw_safe_div(R2, Df, gC1 * w, da, db);
    const float Db2_n = Db2 + da;
    Dbar += db;
    bw_safe_div(R3, Df, gC2 * w, da, db);
    const float Db3_n = Db3 + da;
    Dbar += db;
    const float arawb = a_raw > 0.f ? areab : 0.f;
    Dbar += arawb * dt;
    const float dtb = arawb * Df;
    const float Db0_n = Db0 + Dbar;
    if (dt_raw > 0.f) {
      tb += dtb;
      tprevb -= dtb;
    }
    // control point of event k: dirac = +-density * (1, colour), t = selected safe root
    const float sg = face ? 1.f : -1.f;
    const float densb = sg * (dirac0 + dirac1 * cr.y + dirac2 * cr.z + dirac3 * cr.w);
    const float colb0 = sg * cr.x * dirac1, colb1 = sg * cr.x * dirac2, colb2 = sg * cr.x * dirac3;
    const f3 mean = G.mean;
    f3 rob = mk3(0.f, 0.f, 0.f), rdb = mk3(0.f, 0.f, 0.f), sbar = mk3(0.f, 0.f, 0.f);
    float4 qb = make_float4(0.f, 0.f, 0.f, 0.f);
    if (G.ok) {
      rob = tb * G.jro;
      rdb = tb * G.jrd;
      sbar = tb * G.js;
      qb = make_float4(tb * G.jq.x, tb * G.jq.y, tb * G.jq.z, tb * G.jq.w);
    }
    orgb = orgb + rob;
    dirb = dirb + rdb;
    {
      const float v[kAccFields] = {-rob.x, -rob.y, -rob.z, sbar.x, sbar.y, sbar.z, qb.x, qb.y, qb.z, qb.w,
                                   densb, colb0, colb1, colb2};
      const uint32_t key = A.vcol ? (uint32_t)(view * A.n + p) : (uint32_t)p;
      scatter_prim(A, hkeys, hvals, key, v);
    }
    float* ac = A.acc + 16 * p;
    if (inv) {
      // means2D statistic (backwards_kernel.slang:213-220) with this ray's matrix
      const float dist = mean.x * W[2] + mean.y * W[6] + mean.z * W[10] + W[14];
      atomic_addf(ac + 14, dist * (inv[0] * -rob.x + inv[1] * -rob.y + inv[2] * -rob.z));
      atomic_addf(ac + 15, dist * (inv[4] * -rob.x + inv[5] * -rob.y + inv[6] * -rob.z));
    }
    // step to the previous event
    Da0 = Dbf0; Da1 = Dbf1; Da2 = Dbf2; Da3 = Dbf3;
    La = Lbf;
    csxa = csxbf;
    csya = csybf;
    Lb = Lb_n;
    csxb = csxb_n;
    csyb = csyb_n;
    Db0 = Db0_n; Db1 = Db1_n; Db2 = Db2_n; Db3 = Db3_n;
    tcarry = tprevb;
    id = id_prev;
    tk = tprev;
    G = Gp;
    id_next = id_next2;
    id_next2 = id_next3;
  }
  *dinit = make_float4(Db0, Db1, Db2, Db3);
  float* o = A.drayo + 3 * r;
  float* dd = A.drayd + 3 * r;
  o[0] = clipf(orgb.x, -1e3f, 1e3f);
  o[1] = clipf(orgb.y, -1e3f, 1e3f);
  o[2] = clipf(orgb.z, -1e3f, 1e3f);
  dd[0] = clipf(dirb.x, -1e3f, 1e3f);
  dd[1] = clipf(dirb.y, -1e3f, 1e3f);
  dd[2] = clipf(dirb.z, -1e3f, 1e3f);
}

// Compiled for three waves per SIMD (168 VGPRs, a few spilled): measured faster than two without spills.
__global__ __launch_bounds__(kBwdBlock, 3) void k_trace_backward(BwdArgs A) {
  __shared__ uint32_t hkeys[kHashSlots];
  __shared__ double hvals[kAccFields][kHashSlots];
  for (int i = threadIdx.x; i < kHashSlots; i += kBwdBlock) {
    hkeys[i] = kEmptySlot;
#pragma unroll
    for (int j = 0; j < kAccFields; ++j) hvals[j][i] = 0.0;
  }
  __syncthreads();
  // blocks are dispatched to the 8 XCDs in turn: each XCD takes runs of kXcdChunk consecutive blocks (64x64 pixels
  // of one camera) so that its L2 serves neighbouring blocks' primitive data (measured 3% faster than one block per
  // XCD at a time)
  const int64_t nb = gridDim.x, bb = blockIdx.x;
  const int64_t full = nb / (8 * kXcdChunk) * (8 * kXcdChunk);
  const int64_t tb = bb < full ? ((bb / 8) / kXcdChunk) * (8 * kXcdChunk) + (bb % 8) * kXcdChunk + (bb / 8) % kXcdChunk : bb;
  const int64_t slot = tb * kBwdBlock + threadIdx.x;
  if (slot < A.nr) {
    const int64_t loc = A.order ? (int64_t)A.order[slot] : slot;
    backward_ray(A, A.ray0 + loc, A.events_by_slot ? slot : loc, hkeys, hvals);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < kHashSlots; i += kBwdBlock) {
    const uint32_t key = hkeys[i];
    if (key == kEmptySlot) continue;
#pragma unroll
    for (int j = 0; j < kAccFields; ++j) {
      const float v = (float)hvals[j][i];
      if (v != 0.f) add_global(A, key, j, v);
    }
  }
}

// backwards_initial_drgb_kernel (backwards_kernel.slang:374-428): ray 0's origin + tmin * raw direction. With
// several views (each one camera of the reference), the origin of the ray's view and the candidates of that view.
__global__ __launch_bounds__(256) void k_init_adjoint(const float* __restrict__ means, const float* __restrict__ scales,
                                                      const float* __restrict__ quats, const float4* __restrict__ colr,
                                                      const int32_t* __restrict__ cand,
                                                      const uint32_t* __restrict__ header,
                                                      const float* __restrict__ rayo, const float* __restrict__ rayd,
                                                      int64_t nr, float tmin, const float* __restrict__ dinit,
                                                      float* __restrict__ acc, const float4* __restrict__ vcol,
                                                      const int32_t* __restrict__ view_index,
                                                      const float* __restrict__ view_origins,
                                                      float* __restrict__ vgrad, int64_t n) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nr) return;
  const int32_t ncand = (int32_t)header[8];
  if (ncand == 0) return;
  const float4 gv = reinterpret_cast<const float4*>(dinit)[r];
  if (gv.x == 0.f && gv.y == 0.f && gv.z == 0.f && gv.w == 0.f) return;
  const int32_t view = vcol ? view_index[r] : 0;
  const d3 p0 = tod(ld3(vcol ? view_origins + 3 * view : rayo)) + (double)tmin * tod(ld3(rayd + 3 * r));  // (FP64)
  for (int32_t j = 0; j < ncand; ++j) {
    const int64_t p = cand[j];
    if (vcol && ((reinterpret_cast<const uint64_t*>(cand + cand_mask_offset(n))[j] >> view) & 1u) == 0) continue;
    if (inside_d(p0, ld3(means + 3 * p), ld3(scales + 3 * p), reinterpret_cast<const float4*>(quats)[p], true)) {
      const int64_t kv = vcol ? (int64_t)view * n + p : p;
      const float4 cr = vcol ? vcol[kv] : colr[p];
      float* ac = acc + 16 * p;
      atomic_addf(ac + 10, gv.x + gv.y * cr.y + gv.z * cr.z + gv.w * cr.w);
      float* cg = vcol ? vgrad + 4 * kv : ac + 11;
      atomic_addf(cg + 0, cr.x * gv.y);
      atomic_addf(cg + 1, cr.x * gv.z);
      atomic_addf(cg + 2, cr.x * gv.w);
    }
  }
}

struct EpiArgs {
  const float* acc;
  const float* means;
  const float4* colr;  // (density, colour) of each primitive as the scene build computed it (view 0)
  int64_t n;
  int d;
  int color_mode, sh_degree, activation;
  float ox, oy, oz;
  const float* wcts;
  const float* inv;
  int64_t nw;
  float* dmeans;
  float* dscales;
  float* dquats;
  float* ddens;
  float* dfeat;
  float* dmeans2D;
  float* grgb;  // (N,4) colour-space gradient before the SH basis (colour mode EVER_COLOR_SH)
  // several views: view origins (V,3) and the per-view colour gradients (V,N,4), turned in place into the
  // colour-space gradients before the SH basis
  const float* vorig;
  const float4* vcol;  // (V,N) per-view (density, colour) of the build
  int nv;
  float* vgrad;
};

// dL/df[p,k,c] = Y_k(dir_p) * g[p,c], one lane per primitive: the basis once, then its d x 3 gradients
// (coefficients beyond the degree are zero). With d = 16 the block's 256 x 192 B of gradients are staged in LDS
// and stored as one contiguous run.
__global__ __launch_bounds__(256) void k_sh_feature_grads(const float* __restrict__ grgb,
                                                          const float* __restrict__ means, float ox, float oy,
                                                          float oz, int64_t n, int d, int deg,
                                                          float* __restrict__ dfeat) {
  const int64_t p0 = (int64_t)blockIdx.x * blockDim.x, p = p0 + threadIdx.x;
  const bool valid = p < n;
  const int nc = (deg + 1) * (deg + 1);
  const float4 g = valid ? reinterpret_cast<const float4*>(grgb)[p] : make_float4(0.f, 0.f, 0.f, 0.f);
  float bas[16];
  if (g.x != 0.f || g.y != 0.f || g.z != 0.f) {
    const f3 dir = sh_dir(ld3(means + 3 * p), mk3(ox, oy, oz));
    sh_basis(deg, dir.x, dir.y, dir.z, bas);
  } else {
#pragma unroll
    for (int t = 0; t < 16; ++t) bas[t] = 0.f;
  }
  if (d == 16 && (reinterpret_cast<uintptr_t>(dfeat) & 15u) == 0) {
    __shared__ float4 st[256 * 13];  // 12 float4 per primitive, rows padded to 13 against bank conflicts
    if (valid) {
      float v[48];
#pragma unroll
      for (int t = 0; t < 16; ++t) {
        const float b = t < nc ? bas[t] : 0.f;
        v[3 * t] = b * g.x;
        v[3 * t + 1] = b * g.y;
        v[3 * t + 2] = b * g.z;
      }
#pragma unroll
      for (int q = 0; q < 12; ++q)
        st[threadIdx.x * 13 + q] = make_float4(v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3]);
    }
    __syncthreads();
    const int nblk = n - p0 < 256 ? (int)(n - p0) : 256;
    float4* out4 = reinterpret_cast<float4*>(dfeat + 48 * p0);
#pragma unroll
    for (int it = 0; it < 12; ++it) {
      const int q = it * 256 + (int)threadIdx.x;
      if (q < 12 * nblk) out4[q] = st[(q / 12) * 13 + q % 12];
    }
    return;
  }
  if (!valid) return;
  float* out = dfeat + 3 * (int64_t)d * p;
  for (int t = 0; t < d; ++t) {
    float b = 0.f;
#pragma unroll
    for (int u = 0; u < 16; ++u) b = (u == t && t < nc) ? bas[u] : b;
    out[3 * t] = b * g.x;
    out[3 * t + 1] = b * g.y;
    out[3 * t + 2] = b * g.z;
  }
}

// Several views: dL/df[p,k,c] = sum_v Y_k(dir_{p,v}) * g_v[p,c], one lane per primitive.
__global__ __launch_bounds__(256) void k_sh_feature_grads_views(const float* __restrict__ vgrad,
                                                                const float* __restrict__ means,
                                                                const float* __restrict__ vorig, int nv, int64_t n,
                                                                int d, int deg, float* __restrict__ dfeat) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int nc = (deg + 1) * (deg + 1);
  float acc[16][3];
#pragma unroll
  for (int k = 0; k < 16; ++k) acc[k][0] = acc[k][1] = acc[k][2] = 0.f;
  const f3 mean = ld3(means + 3 * p);
  for (int v = 0; v < nv; ++v) {
    const float4 g = reinterpret_cast<const float4*>(vgrad)[(int64_t)v * n + p];
    if (g.x == 0.f && g.y == 0.f && g.z == 0.f) continue;
    const f3 dir = sh_dir(mean, ld3(vorig + 3 * v));
    float bas[16];
    sh_basis(deg, dir.x, dir.y, dir.z, bas);
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      acc[k][0] += bas[k] * g.x;
      acc[k][1] += bas[k] * g.y;
      acc[k][2] += bas[k] * g.z;
    }
  }
  float* out = dfeat + (int64_t)3 * d * p;
#pragma unroll
  for (int k = 0; k < 16; ++k)
    if (k < d)
#pragma unroll
      for (int ch = 0; ch < 3; ++ch) out[3 * k + ch] = k < nc ? acc[k][ch] : 0.f;
  for (int k = 16; k < d; ++k)
    for (int ch = 0; ch < 3; ++ch) out[3 * k + ch] = 0.f;
}

// Per primitive: clamps (fast_ellipsoid_splinetracer.py:201-211), colour -> SH features, means2D.
__global__ __launch_bounds__(256) void k_epilogue(EpiArgs E) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= E.n) return;
  const float4* a4 = reinterpret_cast<const float4*>(E.acc + 16 * p);
  const float4 a0 = a4[0], a1 = a4[1], a2 = a4[2], a3 = a4[3];
  // a0 = (mx,my,mz,sx) a1 = (sy,sz,qw,qx) a2 = (qy,qz,dens,c0) a3 = (c1,c2,m2x,m2y)
  E.dmeans[3 * p + 0] = clipf(a0.x, -1e3f, 1e3f);
  E.dmeans[3 * p + 1] = clipf(a0.y, -1e3f, 1e3f);
  E.dmeans[3 * p + 2] = clipf(a0.z, -1e3f, 1e3f);
  E.dscales[3 * p + 0] = clipf(a0.w, -1e3f, 1e3f);
  E.dscales[3 * p + 1] = clipf(a1.x, -1e3f, 1e3f);
  E.dscales[3 * p + 2] = clipf(a1.y, -1e3f, 1e3f);
  E.dquats[4 * p + 0] = clipf(a1.z, -1e3f, 1e3f);
  E.dquats[4 * p + 1] = clipf(a1.w, -1e3f, 1e3f);
  E.dquats[4 * p + 2] = clipf(a2.x, -1e3f, 1e3f);
  E.dquats[4 * p + 3] = clipf(a2.y, -1e3f, 1e3f);
  E.ddens[p] = clipf(a2.z, -50.f, 50.f);
  // dL/d(deg-0 feature) = SH_C0 * dL/dcolour, clamped like the reference's feature gradient
  const float g0 = clipf(kShC0 * a2.w, -1e3f, 1e3f), g1 = clipf(kShC0 * a3.x, -1e3f, 1e3f),
              g2 = clipf(kShC0 * a3.y, -1e3f, 1e3f);
  float* df = E.dfeat + (int64_t)3 * E.d * p;
  if (E.color_mode == EVER_COLOR_SH) {
    // chain through RGB2SH (1/SH_C0) and the optional softplus (its derivative from the colour the build
    // stored, so the SH features are not read again); the SH basis is applied by k_sh_feature_grads
    float gr[3] = {g0 / kShC0, g1 / kShC0, g2 / kShC0};
    if (E.activation == 1) {
      const float4 c = E.colr[p];
      gr[0] *= softplus10_grad_from_value(c.y);
      gr[1] *= softplus10_grad_from_value(c.z);
      gr[2] *= softplus10_grad_from_value(c.w);
    }
    // bw_sh_kernel skips primitives whose colour gradient is below 1e-8 in every channel (sh_kernel.slang:167-170)
    const bool skip = fabsf(gr[0]) < 1e-8f && fabsf(gr[1]) < 1e-8f && fabsf(gr[2]) < 1e-8f;
    reinterpret_cast<float4*>(E.grgb)[p] = skip ? make_float4(0.f, 0.f, 0.f, 0.f) : make_float4(gr[0], gr[1], gr[2], 0.f);
    // several views: the same chain per view (each view is one camera's call in the reference), in place
    for (int v = 0; E.nv > 1 && v < E.nv; ++v) {
      float4* gv4 = reinterpret_cast<float4*>(E.vgrad) + (int64_t)v * E.n + p;
      const float4 cg = *gv4;
      float g[3] = {clipf(kShC0 * cg.x, -1e3f, 1e3f) / kShC0, clipf(kShC0 * cg.y, -1e3f, 1e3f) / kShC0,
                    clipf(kShC0 * cg.z, -1e3f, 1e3f) / kShC0};
      if (E.activation == 1) {
        const float4 c = E.vcol[(int64_t)v * E.n + p];
        g[0] *= softplus10_grad_from_value(c.y);
        g[1] *= softplus10_grad_from_value(c.z);
        g[2] *= softplus10_grad_from_value(c.w);
      }
      const bool sk = fabsf(g[0]) < 1e-8f && fabsf(g[1]) < 1e-8f && fabsf(g[2]) < 1e-8f;
      *gv4 = sk ? make_float4(0.f, 0.f, 0.f, 0.f) : make_float4(g[0], g[1], g[2], 0.f);
    }
  } else {
    df[0] = g0;
    df[1] = g1;
    df[2] = g2;
    for (int k = 3; k < 3 * E.d; ++k) df[k] = 0.f;
  }
  if (E.dmeans2D) {
    float m2x, m2y;
    if (E.nw == 1) {
      // the statistic is linear in the per-event mean gradient: apply it to the unclamped total
      const f3 mean = ld3(E.means + 3 * p);
      const float* W = E.wcts;
      const float* iv = E.inv;
      const float dist = mean.x * W[2] + mean.y * W[6] + mean.z * W[10] + W[14];
      m2x = dist * (iv[0] * a0.x + iv[1] * a0.y + iv[2] * a0.z);
      m2y = dist * (iv[4] * a0.x + iv[5] * a0.y + iv[6] * a0.z);
    } else {
      m2x = a3.z;
      m2y = a3.w;
    }
    E.dmeans2D[2 * p + 0] = m2x;
    E.dmeans2D[2 * p + 1] = m2y;
  }
}

namespace {

// Regions and arguments shared by the stored-event and the recomputing adjoint.
struct BackwardCall {
  int64_t n, nr, nw;
  bool views;
  float* acc;
  float* inv;
  float* dinit;
};

BackwardCall backward_begin(const ever_prims& P, const ever_rays& R, const ever_backward_io& B, hipStream_t st) {
  BackwardCall C;
  C.n = P.num_prims;
  C.nr = R.num_rays;
  C.nw = B.wcts ? B.num_wcts : 0;
  BackwardLayout BL = backward_layout(C.n, C.nr, C.nw);
  C.acc = at<float>(B.workspace, BL.acc);
  C.inv = at<float>(B.workspace, BL.inv);
  C.dinit = B.dinitial_drgb ? B.dinitial_drgb : at<float>(B.workspace, BL.dinit);
  C.views = P.num_views > 1 && C.n > 0;
  if (C.n > 0) hipLaunchKernelGGL(k_zero_f32, dim3(grid_for(16 * C.n, 256)), dim3(256), 0, st, C.acc, 16 * C.n);
  if (C.views)
    hipLaunchKernelGGL(k_zero_f32, dim3(grid_for(4 * C.n * P.num_views, 256)), dim3(256), 0, st, B.view_color_grads,
                       4 * C.n * (int64_t)P.num_views);
  if (C.nw > 0) hipLaunchKernelGGL(k_invert4, dim3(grid_for(C.nw, 64)), dim3(64), 0, st, B.wcts, C.nw, C.inv);
  return C;
}

// Arguments of k_trace_backward over all rays of the call, events read from the caller's stored list.
BwdArgs backward_args(const ever_prims& P, void* ws, const ever_rays& R, const ever_backward_io& B,
                      const BackwardCall& C) {
  SceneLayout L = scene_layout(C.n);
  BwdArgs A;
  A.means = P.means;
  A.scales = P.scales;
  A.quats = P.quats;
  A.colr = at<float4>(ws, L.colr);
  A.rayo = R.origins;
  A.rayd = R.directions;
  A.nr = C.nr;
  A.tmin = R.tmin;
  A.max_iters = R.max_iters;
  A.grad_out = B.grad_out;
  A.iters = B.iters;
  A.tri = B.tri_collection;
  A.states = B.states;
  A.st64 = B.ray_state64;
  A.wcts = B.wcts;
  A.nw = C.nw;
  A.inv = C.inv;
  A.acc = C.acc;
  A.drayo = B.drayo;
  A.drayd = B.drayd;
  A.dinit = C.dinit;
  A.order = B.ray_order;
  A.events_by_slot = R.event_layout == EVER_EVENTS_BY_SLOT ? 1 : 0;
  A.ray0 = 0;
  A.tri_stride = C.nr;
  A.vcol = C.views ? reinterpret_cast<const float4*>(P.view_colors) : nullptr;
  A.view_index = C.views ? R.view_index : nullptr;
  A.vgrad = C.views ? B.view_color_grads : nullptr;
  A.n = C.n;
  return A;
}

void backward_trace(const BwdArgs& A, hipStream_t st) {
  profile_begin(EVER_KERNEL_BACKWARD, st);
  hipLaunchKernelGGL(k_trace_backward, dim3(grid_for(A.nr, kBwdBlock)), dim3(kBwdBlock), 0, st, A);
  profile_end(EVER_KERNEL_BACKWARD, st);
}

// Start-inside adjoint over all rays, then the per-primitive epilogue (clamps, means2D, SH feature gradients).
void backward_end(const ever_prims& P, void* ws, const ever_rays& R, const ever_backward_io& B,
                  const BackwardCall& C, hipStream_t st) {
  const int64_t n = C.n;
  SceneLayout L = scene_layout(n);
  BackwardLayout BL = backward_layout(n, C.nr, C.nw);
  const float4* vcol = C.views ? reinterpret_cast<const float4*>(P.view_colors) : nullptr;
  if (C.nr > 0 && n > 0)
    hipLaunchKernelGGL(k_init_adjoint, dim3(grid_for(C.nr, 256)), dim3(256), 0, st, P.means, P.scales, P.quats,
                       at<float4>(ws, L.colr), at<int32_t>(ws, L.cand), at<uint32_t>(ws, L.header), R.origins,
                       R.directions, C.nr, R.tmin, C.dinit, C.acc, vcol, C.views ? R.view_index : nullptr,
                       P.view_origins, C.views ? B.view_color_grads : nullptr, n);
  if (n <= 0) return;
  EpiArgs E;
  E.acc = C.acc;
  E.means = P.means;
  E.colr = at<float4>(ws, L.colr);
  E.n = n;
  E.d = P.feature_size;
  E.color_mode = P.color_mode;
  E.sh_degree = P.sh_degree;
  E.activation = P.activation;
  E.ox = P.sh_origin[0];
  E.oy = P.sh_origin[1];
  E.oz = P.sh_origin[2];
  E.wcts = B.wcts;
  E.inv = C.inv;
  E.nw = C.nw;
  E.dmeans = B.dmeans;
  E.dscales = B.dscales;
  E.dquats = B.dquats;
  E.ddens = B.ddensities;
  E.dfeat = B.dfeatures;
  E.dmeans2D = C.nw > 0 ? B.dmeans2D : nullptr;
  E.grgb = at<float>(B.workspace, BL.grgb);
  E.vorig = C.views ? P.view_origins : nullptr;
  E.vcol = vcol;
  E.nv = C.views ? P.num_views : 0;
  E.vgrad = C.views ? B.view_color_grads : nullptr;
  hipLaunchKernelGGL(k_epilogue, dim3(grid_for(n, 256)), dim3(256), 0, st, E);
  if (P.color_mode == EVER_COLOR_SH) {
    if (C.views)
      hipLaunchKernelGGL(k_sh_feature_grads_views, dim3(grid_for(n, 256)), dim3(256), 0, st, B.view_color_grads,
                         P.means, P.view_origins, P.num_views, n, P.feature_size, P.sh_degree, B.dfeatures);
    else
      hipLaunchKernelGGL(k_sh_feature_grads, dim3(grid_for(n, 256)), dim3(256), 0, st, E.grgb, P.means,
                         P.sh_origin[0], P.sh_origin[1], P.sh_origin[2], n, P.feature_size, P.sh_degree,
                         B.dfeatures);
  }
}

}  // namespace

hipError_t launch_backward(const ever_prims& P, void* ws, const ever_rays& R, const ever_backward_io& B,
                           hipStream_t st) {
  const BackwardCall C = backward_begin(P, R, B, st);
  if (C.nr > 0) backward_trace(backward_args(P, ws, R, B, C), st);
  backward_end(P, ws, R, B, C, st);
  return hipGetLastError();
}

// The adjoint with its event lists traced again, one chunk of rays at a time: the chunk's forward (with the
// start-inside candidates of the call's own forward, still in the scene workspace) writes the chunk's events by
// processing slot into a list of chunk x max_iters entries, and k_trace_backward consumes it before the next chunk
// overwrites it. A ray's events do not depend on the other rays traced with it, so every ray sees the events of
// the stored-list adjoint; the accumulators, the start-inside adjoint and the epilogue are shared by all chunks.
hipError_t launch_backward_recompute(const ever_prims& P, void* ws, const ever_rays& R, const ever_backward_io& B,
                                     int64_t chunk, hipStream_t st) {
  const BackwardCall C = backward_begin(P, R, B, st);
  if (C.nr > 0) {
    const RecomputeLayout RL = recompute_layout(C.n, C.nr, C.nw, R.max_iters, R.traversal, chunk);
    BwdArgs A = backward_args(P, ws, R, B, C);
    int32_t* tri = at<int32_t>(B.workspace, RL.tri);
    A.tri = tri;
    A.events_by_slot = 1;
    for (int64_t a = 0; a < C.nr; a += RL.chunk) {
      const int64_t c = std::min<int64_t>(RL.chunk, C.nr - a);
      ever_rays Rc = R;
      Rc.origins = R.origins + 3 * a;
      Rc.directions = R.directions + 3 * a;
      Rc.num_rays = c;
      Rc.proposal = R.proposal ? R.proposal + a : nullptr;
      Rc.view_index = R.view_index ? R.view_index + a : nullptr;
      Rc.tile = 0;
      Rc.event_layout = EVER_EVENTS_BY_SLOT;
      // the chunk's forward workspace is sized for the walk of a full chunk; a shorter last chunk keeps that walk
      Rc.traversal = per_ray_walk(RL.chunk, R.traversal) ? EVER_TRAVERSAL_RAY : EVER_TRAVERSAL_WAVE;
      ever_forward_out O{};
      O.color = at<float>(B.workspace, RL.color);
      O.states = at<float>(B.workspace, RL.states);
      O.iters = at<int32_t>(B.workspace, RL.iters);
      O.tri_collection = tri;
      O.touch_count = nullptr;
      O.initial_drgb = at<float>(B.workspace, RL.init);
      O.ray_state64 = at<double>(B.workspace, RL.st64);
      O.ray_order = at<int32_t>(B.workspace, RL.order);
      O.workspace = at<char>(B.workspace, RL.fwd);
      O.workspace_bytes = forward_layout(RL.chunk, per_ray_walk(RL.chunk, R.traversal)).total;
      O.last_dirac = nullptr;
      hipError_t e = launch_forward(P, ws, Rc, O, st, true);
      if (e != hipSuccess) return e;
      A.ray0 = a;
      A.nr = c;
      A.tri_stride = c;  // the chunk's forward wrote its list with one row of c slots per event index
      A.order = O.ray_order;
      backward_trace(A, st);
    }
  }
  backward_end(P, ws, R, B, C, st);
  return hipGetLastError();
}

}  // namespace ever
