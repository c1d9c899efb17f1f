// Internal host-side declarations: workspace layouts and kernel launchers.
#pragma once

#include <hip/hip_runtime.h>
#include <algorithm>
#include <stddef.h>
#include <stdint.h>
#include <string>

#include "../../include/ever_amd.h"

namespace ever {

constexpr int kSortThreads = 256;
constexpr int kSortItems = 16;
constexpr int kSortTile = kSortThreads * kSortItems;  // keys per radix-sort tile
constexpr int kRadix = 256;

inline size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

constexpr int kLeafBlock = 64;    // leaves per block of the range-union refit
constexpr int kMaxFragments = 4;  // leaf fragments per primitive (elongated ellipsoids are split along the major axis)
constexpr int32_t kEmptyChild = (int32_t)0x80000000;  // unused child slot of a 4-wide node

// int32 offset of the view masks in the start-inside candidate region (8-byte aligned, after N int32 entries)
__host__ __device__ constexpr int64_t cand_mask_offset(int64_t n) { return (n + 1) & ~int64_t(1); }

// Scene workspace: everything the LBVH build writes and the tracers read. All offsets 256-byte aligned.
// Leaves are fragments: each primitive owns kMaxFragments slots (S = kMaxFragments * N); unused slots carry the
// sort key 0xffffffff and end behind the L used leaves (L is counted on the device, header[9]).
//   header   : uint32[16]  (0..5 centroid bounds as ordered ints, 8 = start-inside candidate count, 9 = L)
//   aabb     : float[S*6]  fragment boxes (slot order); once the leaf blocks are built, the binary-node boxes
//   fprim    : uint32[S]   slot -> primitive
//   keys0/1, vals0/1 : uint32[S] radix-sort ping-pong (vals0 ends as sorted leaf -> slot)
//   hist     : uint32[256 * tiles + 256] (digit-major tile counts, then digit totals)
//   nodes    : float4[S-1] binary hierarchy: child indices (leaf child = ~sorted leaf)
//   nodes4   : float4[(S-1) * 8] 4-wide records (128 B): children's boxes per axis + child refs (leaf = ~primitive)
//   range    : int2[S-1] leaf range of each internal node
//   sbox/pre/suf : float[S*6] leaf boxes in sorted order, in-block prefix and suffix unions
//   tab      : float[levels * blocks * 6] sparse table of block unions
//   rec      : float4[N * 4] per-primitive traversal record, original order (64 B)
//   colr     : float4[N] (density, R, G, B) in original order
//   cand     : int32[N] start-inside candidates, then (at cand_mask_offset) uint64[N] their view masks (several
//              views: bit v set when the candidate's box lies within tmin of view v's origin)
//   bpart    : float[6 * ceil(N/256)] per-block centroid bounds
struct SceneLayout {
  int64_t n = 0, slots = 0, tiles = 0, leaf_blocks = 0;
  int tab_levels = 1;
  size_t header, aabb, fprim, keys0, vals0, keys1, vals1, hist, nodes, nodes4, range, sbox, pre, suf, tab, rec, colr, cand,
      bpart, total;
};

inline SceneLayout scene_layout(int64_t n) {
  SceneLayout L;
  L.n = n;
  L.slots = (int64_t)kMaxFragments * n;
  const int64_t S = L.slots;
  L.tiles = (S + kSortTile - 1) / kSortTile;
  L.leaf_blocks = (S + kLeafBlock - 1) / kLeafBlock;
  L.tab_levels = 1;
  while ((int64_t(1) << L.tab_levels) <= L.leaf_blocks) ++L.tab_levels;
  const int64_t ni = S > 1 ? S - 1 : 1;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    size_t o = off;
    off = align_up(off + bytes, 256);
    return o;
  };
  L.header = take(64);
  L.aabb = take(sizeof(float) * 6 * S);
  L.fprim = take(4 * S);
  L.keys0 = take(4 * S);
  L.vals0 = take(4 * S);
  L.keys1 = take(4 * S);
  L.vals1 = take(4 * S);
  L.hist = take(4 * (size_t)kRadix * ((L.tiles > 0 ? L.tiles : 1) + 1));  // + digit totals
  L.nodes = take(16 * ni);
  L.nodes4 = take(128 * ni);
  L.range = take(8 * ni);
  L.sbox = take(24 * S);
  L.pre = take(24 * S);
  L.suf = take(24 * S);
  L.tab = take(24 * (size_t)L.tab_levels * (L.leaf_blocks > 0 ? L.leaf_blocks : 1));
  L.rec = take(64 * n);
  L.colr = take(16 * n);
  L.cand = take(4 * (size_t)cand_mask_offset(n) + 8 * (size_t)n);
  L.bpart = take(24 * (size_t)((n + 255) / 256 > 0 ? (n + 255) / 256 : 1));
  L.total = off;
  return L;
}

// Backward workspace: per-primitive gradient accumulators (16 floats: mean 3, scale 3, quat 4, density 1,
// colour 3, means2D 2) and inverted world-to-camera matrices.
// ... and the adjoint of the start-inside density per ray (used when the caller passes no buffer for it).
struct BackwardLayout {
  size_t acc, inv, dinit, grgb, total;
};
inline BackwardLayout backward_layout(int64_t n, int64_t nr, int64_t nw) {
  BackwardLayout L;
  size_t off = 0;
  L.acc = off;
  off = align_up(off + 64 * (size_t)(n > 0 ? n : 1), 256);
  L.inv = off;
  off = align_up(off + 64 * (size_t)(nw > 0 ? nw : 1), 256);
  L.dinit = off;
  off = align_up(off + 16 * (size_t)(nr > 0 ? nr : 1), 256);
  L.grgb = off;
  off = align_up(off + 16 * (size_t)(n > 0 ? n : 1), 256);
  L.total = off;
  return L;
}

template <class T> inline T* at(void* base, size_t off) { return reinterpret_cast<T*>(static_cast<char*>(base) + off); }
template <class T> inline const T* at(const void* base, size_t off) {
  return reinterpret_cast<const T*>(static_cast<const char*>(base) + off);
}

// Workgroup barrier for LDS traffic only: waits for this wave's LDS operations and meets the other waves, without
// draining its global stores as __syncthreads does (a loop that scatters to global memory between barriers would
// otherwise wait out every store's latency).
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// launchers (return hipSuccess or the first error)
// Stable LSD radix sort of (key, value) pairs on key bits [lo_bit, bits) (8 per pass); the result ends in (k0, v0),
// or, when vres is given, in whichever pair the last pass wrote, with *vres set to its value buffer (no copy back).
// With dcount, the pairs sorted are the first min(*dcount, n) (a count known only on the device); n sizes the grid.
hipError_t launch_sort_pairs(uint32_t* k0, uint32_t* v0, uint32_t* k1, uint32_t* v1, uint32_t* hist, int64_t n,
                             int bits, hipStream_t st, const uint32_t* dcount = nullptr, int lo_bit = 0,
                             uint32_t** vres = nullptr);
hipError_t launch_build(const ever_prims& P, void* ws, hipStream_t s);
// Rewrites the per-primitive colours of a built scene from new features (geometry and tree untouched).
hipError_t launch_update_colors(const ever_prims& P, void* ws, hipStream_t s);
// keep_candidates: trace with the start-inside candidate list an earlier forward call left in the scene workspace
// (the rays are a part of that call's rays, whose first ray chose the candidates) instead of building it again
hipError_t launch_forward(const ever_prims& P, void* ws, const ever_rays& R, const ever_forward_out& O,
                          hipStream_t s, bool keep_candidates = false);
hipError_t launch_backward(const ever_prims& P, void* ws, const ever_rays& R, const ever_backward_io& B,
                           hipStream_t s);
// The adjoint without stored event lists: the rays are traced again chunk by chunk into one chunk-sized list
hipError_t launch_backward_recompute(const ever_prims& P, void* ws, const ever_rays& R, const ever_backward_io& B,
                                     int64_t chunk, hipStream_t s);
struct MortonLayout {
  size_t box, part, keys0, vals0, keys1, vals1, hi, hist, total;
};
constexpr int kBoundsBlocks = 1024;  // blocks of the partial bound reduction of the Morton order
// Morton-order workspace: mean bounds (partial and final), 64-bit keys as two radix-sort passes over 32 + 31 bits
// (ping-pong keys/values, the high words) and the digit histograms.
constexpr MortonLayout morton_layout(int64_t n) {
  MortonLayout L{};
  size_t off = 0;
  auto take = [&off](size_t bytes) {
    const size_t o = off;
    off += (bytes + 255) & ~(size_t)255;
    return o;
  };
  const int64_t tiles = (n + kSortTile - 1) / kSortTile;
  L.box = take(6 * sizeof(float));
  L.part = take(6 * sizeof(float) * kBoundsBlocks);
  L.keys0 = take(4 * (size_t)n);
  L.vals0 = take(4 * (size_t)n);
  L.keys1 = take(4 * (size_t)n);
  L.vals1 = take(4 * (size_t)n);
  L.hi = take(4 * (size_t)n);
  L.hist = take(4 * (size_t)kRadix * ((tiles > 0 ? tiles : 1) + 1));
  L.total = off;
  return L;
}
// Morton order without a caller workspace: the library's own scratch array, for at most this many primitives
constexpr int64_t kMortonNoallocMaxPrims = EVER_MORTON_NOALLOC_MAX_PRIMS;
hipError_t launch_morton_order_noalloc(const float* means, int64_t n, int32_t* order, hipStream_t s);
hipError_t launch_scene_from_records(const float* records, int64_t n, const ever_ply_layout& L, const int32_t* order,
                                     float* means, float* scales, float* quats, float* densities, float* features,
                                     hipStream_t s);
hipError_t launch_morton_order(const float* means, int64_t n, int32_t* order, void* ws, hipStream_t s);
hipError_t launch_camera_rays(const ever_camera& C, const float* jitter, float* origins, float* directions,
                              hipStream_t s);
hipError_t launch_eval_sh(const float* means, const float* feat, const float* origin, int64_t n, int d, int deg,
                          float* out, hipStream_t s);
hipError_t launch_eval_sh_bwd(const float* means, const float* feat, const float* origin, const float* drgb,
                              int64_t n, int d, int deg, float* dfeat, hipStream_t s);

// Traversal stack bound. Along a root-to-leaf path of the Karras hierarchy the split level strictly increases.
// Used leaf keys are below 2^30, so distinct keys split at levels 2..31; equal keys split on the sorted leaf index
// (level 32 + clz(a ^ b), a, b < 2^30 leaves: levels 34..63). A path therefore holds at most 60 internal nodes,
// and so at most 60 4-wide records (each record descends at least one binary level). A record pushes at most three
// children, so a traversal stack never holds more than 180 entries, whatever the scene.
constexpr int kMaxPathNodes = 60;
constexpr int kStackBound = 3 * kMaxPathNodes;
constexpr int kLdsStackPerRay = 32;   // per-ray walk: stack entries per lane in LDS
constexpr int kSpillPerRay = 160;     // ... and in the per-ray spill area of the forward workspace
static_assert(kLdsStackPerRay + kSpillPerRay >= kStackBound, "per-ray traversal stack below the tree bound");
constexpr int kWaveStackRows = 3;     // per-wave walk: 64-entry rows, two in VGPRs and one in LDS
static_assert(64 * kWaveStackRows >= kStackBound, "per-wave traversal stack below the tree bound");
constexpr int64_t kMinRaysToSort = 65536;  // below this the rays are traced in the caller's order

// Whether a call traces with the per-ray walk (EVER_TRAVERSAL_RAY, or AUTO on a batch too small to sort).
inline bool per_ray_walk(int64_t nr, int traversal) {
  return traversal == EVER_TRAVERSAL_RAY || (traversal == EVER_TRAVERSAL_AUTO && nr < kMinRaysToSort);
}

// Forward workspace: traversal-stack spill of the per-ray walk (the per-wave walk keeps its whole stack on chip), the
// ray sort (30-bit Morton key of a point on each ray) and the integrator's state per processing slot (112 B).
struct ForwardLayout {
  int64_t tiles = 0;
  size_t spill, keys0, vals0, keys1, vals1, hist, counter, ist, total;
};
inline ForwardLayout forward_layout(int64_t nr, bool per_ray) {
  ForwardLayout L;
  const int64_t r = nr > 0 ? nr : 1;
  L.tiles = (r + kSortTile - 1) / kSortTile;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    size_t o = off;
    off = align_up(off + bytes, 256);
    return o;
  };
  L.spill = take(per_ray ? sizeof(int32_t) * kSpillPerRay * (size_t)r : 0);
  L.keys0 = take(4 * (size_t)r);
  L.vals0 = take(4 * (size_t)r);
  L.keys1 = take(4 * (size_t)r);
  L.vals1 = take(4 * (size_t)r);
  L.hist = take(4 * (size_t)kRadix * (L.tiles + 1));  // + digit totals
  L.counter = take(4);                                 // next tile of a persistent launch
  L.ist = take(112 * (size_t)r);
  L.total = off;
  return L;
}

// Backward workspace of the recomputing adjoint: the backward's own regions, then one chunk's forward workspace,
// event list (slot-major, chunk x max_iters) and forward outputs, reused by every chunk.
struct RecomputeLayout {
  int64_t chunk = 0;
  size_t fwd, tri, color, states, iters, init, st64, order, total;
};
inline RecomputeLayout recompute_layout(int64_t n, int64_t nr, int64_t nw, int max_iters, int traversal,
                                        int64_t chunk) {
  RecomputeLayout L;
  L.chunk = std::max<int64_t>(1, std::min<int64_t>(chunk, nr));
  const size_t c = (size_t)L.chunk;
  size_t off = align_up(backward_layout(n, nr, nw).total, 256);
  auto take = [&](size_t bytes) {
    size_t o = off;
    off = align_up(off + bytes, 256);
    return o;
  };
  L.fwd = take(forward_layout(L.chunk, per_ray_walk(L.chunk, traversal)).total);
  L.tri = take(4 * c * (size_t)std::max(max_iters, 0));
  L.color = take(20 * c);
  L.states = take(64 * c);
  L.iters = take(4 * c);
  L.init = take(16 * c);
  L.st64 = take(64 * c);
  L.order = take(4 * c);
  L.total = off;
  return L;
}

inline unsigned grid_for(int64_t n, int block) { return (unsigned)((n + block - 1) / block); }

// Proposal network (ever_mlp.hip)
constexpr int kMlpIn = 64;     // 63 encoded features + one zero column
constexpr int kMlpHid = 128;   // hidden width
constexpr int kMlpSplit = 32;  // row chunks of the weight-gradient products
struct MlpLayout {
  int64_t w1, b1, w2, b2, w3, b3, w4, b4, count;
};
struct MlpWork {
  int64_t x, h1, h2, h3, dh, dz, dz2, y, dy, part, grad, loss, total;
};
MlpLayout mlp_layout();
MlpWork mlp_work(int64_t rows_padded);
hipError_t launch_mlp_infer(const float* p, const uint16_t* pb, const float* rayo, const float* rayd, int64_t nr,
                            float tmin, float margin, float* y, float* proposal, void* ws, hipStream_t st);
hipError_t launch_mlp_train_step(float* p, uint16_t* pb, float* mom, float* vel, const float* rayo,
                                 const float* rayd, const float* target, int64_t nr, float tmin, float tau, float lr,
                                 int step, void* ws, float* loss_out, float* grad_out, hipStream_t st);
hipError_t launch_mlp_sync_bf16(const float* p, uint16_t* pb, hipStream_t st);
hipError_t launch_gemm_public(const uint16_t* A, int64_t sam, int64_t sak, const uint16_t* B, int64_t sbk,
                              int64_t sbn, int64_t M, int64_t N, int64_t K, const float* bias, int relu, float* C,
                              int splits, float* part, hipStream_t st);

// Event pairs around the trace kernels while ever_profile_enable(1) is in effect (ever_capi.hip).
void profile_begin(int kernel, hipStream_t st);
void profile_end(int kernel, hipStream_t st);

}  // namespace ever
