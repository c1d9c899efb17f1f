// Internal host-side declarations: workspace layouts and kernel launchers.
#pragma once

#include <hip/hip_runtime.h>
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

// Scene workspace: everything the LBVH build writes and the tracers read. All offsets 256-byte aligned.
//   header   : uint32[16]  (0..5 centroid bounds as ordered ints, 8 = start-inside candidate count)
//   aabb     : float[N*6]  per-primitive box (original order)
//   keys0/1, vals0/1 : uint32[N] radix-sort ping-pong (vals end as the sorted->original permutation)
//   hist     : uint32[256 * tiles]
//   nodes    : float4[(N-1) * 4] BVH2 nodes, children boxes stored in the parent (64 B per node)
//   parent   : int32[N-1], leaf_parent : int32[N], flags : uint32[N-1]
//   rec      : float4[N * 4] per-primitive traversal record in LBVH order (64 B)
//   colr     : float4[N] (density, R, G, B) in original order
//   cand     : int32[N] start-inside candidates
struct SceneLayout {
  int64_t n = 0, tiles = 0;
  size_t header, aabb, keys0, vals0, keys1, vals1, hist, nodes, parent, leaf_parent, flags, rec, colr, cand, total;
};

inline SceneLayout scene_layout(int64_t n) {
  SceneLayout L;
  L.n = n;
  L.tiles = (n + kSortTile - 1) / kSortTile;
  int64_t ni = n > 1 ? n - 1 : 1;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    size_t o = off;
    off = align_up(off + bytes, 256);
    return o;
  };
  L.header = take(64);
  L.aabb = take(sizeof(float) * 6 * n);
  L.keys0 = take(4 * n);
  L.vals0 = take(4 * n);
  L.keys1 = take(4 * n);
  L.vals1 = take(4 * n);
  L.hist = take(4 * (size_t)kRadix * (L.tiles > 0 ? L.tiles : 1));
  L.nodes = take(64 * ni);
  L.parent = take(4 * ni);
  L.leaf_parent = take(4 * n);
  L.flags = take(4 * ni);
  L.rec = take(64 * n);
  L.colr = take(16 * n);
  L.cand = take(4 * n);
  L.total = off;
  return L;
}

// Backward workspace: per-primitive gradient accumulators (16 floats: mean 3, scale 3, quat 4, density 1,
// colour 3, means2D 2) and inverted world-to-camera matrices.
// ... and the adjoint of the start-inside density per ray (used when the caller passes no buffer for it).
struct BackwardLayout {
  size_t acc, inv, dinit, total;
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
  L.total = off;
  return L;
}

template <class T> inline T* at(void* base, size_t off) { return reinterpret_cast<T*>(static_cast<char*>(base) + off); }
template <class T> inline const T* at(const void* base, size_t off) {
  return reinterpret_cast<const T*>(static_cast<const char*>(base) + off);
}

// launchers (return hipSuccess or the first error)
hipError_t launch_build(const ever_prims& P, void* ws, hipStream_t s);
hipError_t launch_forward(const ever_prims& P, void* ws, const ever_rays& R, const ever_forward_out& O,
                          int32_t* spill, hipStream_t s);
hipError_t launch_backward(const ever_prims& P, void* ws, const ever_rays& R, const ever_backward_io& B,
                           hipStream_t s);
hipError_t launch_eval_sh(const float* means, const float* feat, const float* origin, int64_t n, int d, int deg,
                          float* out, hipStream_t s);
hipError_t launch_eval_sh_bwd(const float* means, const float* feat, const float* origin, const float* drgb,
                              int64_t n, int d, int deg, float* dfeat, hipStream_t s);

constexpr int kSpillPerRay = 32;  // int32 traversal-stack spill entries per ray

inline unsigned grid_for(int64_t n, int block) { return (unsigned)((n + block - 1) / block); }

}  // namespace ever
