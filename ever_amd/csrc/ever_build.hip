// LBVH build over ellipsoid AABBs, rebuilt on every call: per-primitive prepass (colour, traversal record,
// centroid bounds), leaf fragments (elongated ellipsoids split along their major axis), Morton codes with a box size class, an LDS radix sort written for wave64, Karras' parallel hierarchy
// construction, and node boxes computed as unions over each node's contiguous leaf range (in-block
// prefix/suffix unions + a sparse table over 64-leaf blocks), which needs no atomics or cross-XCD coherence.
//
// Replaces the reference's OptiX acceleration structure (splinetracers/fast_ellipsoid_splinetracer/GAS.cpp:56-153)
// and its AABB kernel (create_aabbs.cu:69-115), plus the per-primitive SH prepass (splinetracers/slang/sh_kernel.slang:19-42)
// when colour mode EVER_COLOR_SH is selected.
#include "ever_common.hpp"
#include "ever_internal.hpp"

namespace ever {

__device__ __forceinline__ uint32_t f2o(float f) {
  uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float o2f(uint32_t u) {
  return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

__global__ void k_fill_u32(uint32_t* p, int64_t n, uint32_t v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_build_init(uint32_t* header) {
  int t = threadIdx.x;
  if (t < 3) header[t] = 0xffffffffu;   // min (ordered)
  else if (t < 6) header[t] = 0u;       // max (ordered)
  else if (t < 16) header[t] = 0u;
}

// Leaf fragments of a primitive with scales sv: ceil(s_max / s_mid) clamped to [1, kMaxFragments]; ax is the axis
// of s_max (the split axis). The prepass counts with it and the fragment emission emits with it.
__device__ __forceinline__ int fragment_count(const float sv[3], int& ax) {
  ax = 0;
  if (sv[1] > sv[ax]) ax = 1;
  if (sv[2] > sv[ax]) ax = 2;
  float mid = 0.f;
  for (int k = 0; k < 3; ++k)
    if (k != ax) mid = fmaxf(mid, sv[k]);
  const int F = mid > 0.f ? (int)ceilf(sv[ax] / mid) : kMaxFragments;
  return F < 1 ? 1 : (F > kMaxFragments ? kMaxFragments : F);
}

// Per primitive (original order): colour record, traversal record, centroid bounds and the block's fragment count.
//   record (64 B): rows (M_k, b_k), k = 0..2, with M = diag(1/s) L_q, b = -M mean, so that the ellipsoid-frame
//   origin is ocn = M o + b and direction rdn = M d; L_q is the linear map of rotate_vector (safe-math.slang:237-244)
//   for the raw quaternion, as the reference's intersection uses it (tri-intersect.slang:317-329).
//   row 3 = (density, R, G, B).
// kColorsOnly: only the colour part of the records (rec row 3, colr, per-view colours) is rewritten, for new
// features on an unchanged geometry (ever_update_colors)
template <bool kColorsOnly>
__global__ __launch_bounds__(256) void k_prim_prepass(ever_prims P, float4* __restrict__ rec,
                                                      float4* __restrict__ colr, float* __restrict__ bpart,
                                                      uint32_t* __restrict__ fcount) {
  // the block's SH features staged through LDS with contiguous loads (a lane reading its own primitive's
  // coefficients would touch a different line per lane and load), 49 words per primitive so the lanes' reads
  // fall in distinct banks
  constexpr int kStride = 49;
  __shared__ float fs[256 * kStride];
  const int64_t i0 = (int64_t)blockIdx.x * blockDim.x;
  const int per = 3 * P.feature_size;
  const bool staged = P.color_mode == EVER_COLOR_SH && per <= 48;
  if (staged) {
    // block-local indices in 32 bits; twelve loads in flight per thread (a whole SH3 block in one pass) before
    // their LDS stores, and the primitive of a word from a float reciprocal (idx < 12288 and per <= 48 keep
    // (idx + 0.5) / per at least 0.01 from an integer, far above the rounding error)
    const int nb = P.num_prims - i0 < 256 ? (int)(P.num_prims - i0) : 256;
    const int total = nb * per;
    const float* src = P.features + per * i0;
    const float rper = 1.f / (float)per;
    for (int b0 = 0; b0 < total; b0 += 12 * 256) {
      float w[12];
#pragma unroll
      for (int u = 0; u < 12; ++u) {
        const int idx = b0 + u * 256 + (int)threadIdx.x;
        w[u] = idx < total ? src[idx] : 0.f;
      }
#pragma unroll
      for (int u = 0; u < 12; ++u) {
        const int idx = b0 + u * 256 + (int)threadIdx.x;
        const int pl = (int)(((float)idx + 0.5f) * rper);
        if (idx < total) fs[pl * kStride + (idx - pl * per)] = w[u];
      }
    }
    __syncthreads();
  }
  int64_t i = i0 + threadIdx.x;
  bool valid = i < P.num_prims;
  float cx = 0.f, cy = 0.f, cz = 0.f;
  int nfrag = 0;
  if (valid) {
    const float4 q = reinterpret_cast<const float4*>(P.quats)[i];
    const f3 s = ld3(P.scales + 3 * i);
    if (!kColorsOnly) {
      const float sv[3] = {s.x, s.y, s.z};
      int ax;
      nfrag = fragment_count(sv, ax);
    }
    const f3 m = ld3(P.means + 3 * i);
    cx = m.x; cy = m.y; cz = m.z;
    float dens = P.densities[i];
    float3 c;
    const float* f = staged ? fs + threadIdx.x * kStride : P.features + (int64_t)3 * P.feature_size * i;
    if (P.color_mode == EVER_COLOR_SH) {
      auto colour_at = [&](f3 origin) {
        const f3 dir = sh_dir(m, origin);
        float bas[16];
        sh_basis(P.sh_degree, dir.x, dir.y, dir.z, bas);
        const int nc = (P.sh_degree + 1) * (P.sh_degree + 1);
        float acc[3] = {0.5f, 0.5f, 0.5f};
        for (int k = 0; k < nc; ++k)
          for (int ch = 0; ch < 3; ++ch) acc[ch] += bas[k] * f[3 * k + ch];
        if (P.activation == 1)
          for (int ch = 0; ch < 3; ++ch) acc[ch] = softplus10(acc[ch]);
        // colour of the deg-0 feature RGB2SH(rgb) seen through SH_C0 * f + 0.5 (ever.py:178-180,
        // shaders.slang:65-82)
        return make_float3(acc[0], acc[1], acc[2]);
      };
      if (P.num_views > 1) {
        float4* vc = reinterpret_cast<float4*>(P.view_colors);
        for (int v = 0; v < P.num_views; ++v) {
          const float3 cv = colour_at(ld3(P.view_origins + 3 * v));
          vc[(int64_t)v * P.num_prims + i] = make_float4(dens, cv.x, cv.y, cv.z);
          if (v == 0) c = cv;
        }
      } else {
        c = colour_at(mk3(P.sh_origin[0], P.sh_origin[1], P.sh_origin[2]));
      }
    } else {
      c = make_float3(kShC0 * f[0] + 0.5f, kShC0 * f[1] + 0.5f, kShC0 * f[2] + 0.5f);
    }
    const float4 cr = make_float4(dens, c.x, c.y, c.z);
    colr[i] = cr;
    rec[4 * i + 3] = cr;
    if (kColorsOnly) return;
    const f3 c0 = qrot(mk3(1.f, 0.f, 0.f), q), c1 = qrot(mk3(0.f, 1.f, 0.f), q), c2 = qrot(mk3(0.f, 0.f, 1.f), q);
    const float M[3][3] = {{c0.x / s.x, c1.x / s.x, c2.x / s.x},
                           {c0.y / s.y, c1.y / s.y, c2.y / s.y},
                           {c0.z / s.z, c1.z / s.z, c2.z / s.z}};
    for (int k = 0; k < 3; ++k) {
      const float bb = -(M[k][0] * m.x + M[k][1] * m.y + M[k][2] * m.z);
      rec[4 * i + k] = make_float4(M[k][0], M[k][1], M[k][2], bb);
    }
  }
  if (kColorsOnly) return;
  // block reduction of centroid bounds into a per-block partial (reduced by k_bounds_final), and of the fragment
  // count (scanned into fragment offsets by k_bounds_final)
  __shared__ float red[6][4];
  __shared__ uint32_t nred[4];
  float v[6] = {valid ? cx : INFINITY, valid ? cy : INFINITY, valid ? cz : INFINITY,
                valid ? -cx : INFINITY, valid ? -cy : INFINITY, valid ? -cz : INFINITY};
  for (int o = 32; o > 0; o >>= 1) {
    for (int k = 0; k < 6; ++k) v[k] = fminf(v[k], __shfl_xor(v[k], o));
    nfrag += __shfl_xor(nfrag, o);
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    for (int k = 0; k < 6; ++k) red[k][w] = v[k];
    nred[w] = (uint32_t)nfrag;
  }
  __syncthreads();
  if (threadIdx.x < 6) {
    const int k = threadIdx.x;
    bpart[6 * (int64_t)blockIdx.x + k] = fminf(fminf(red[k][0], red[k][1]), fminf(red[k][2], red[k][3]));
  }
  if (threadIdx.x == 0) fcount[blockIdx.x] = nred[0] + nred[1] + nred[2] + nred[3];
}

// Final reduction of the per-block bounds (one block); writes the ordered-int bounds into the header. Also turns
// the per-block fragment counts into exclusive offsets (in place) and writes their total, the leaf count L, into
// header[9].
__global__ __launch_bounds__(256) void k_bounds_final(const float* __restrict__ bpart, int64_t nb,
                                                      uint32_t* __restrict__ header, uint32_t* __restrict__ fcount) {
  __shared__ float red[6][256];
  float v[6] = {INFINITY, INFINITY, INFINITY, INFINITY, INFINITY, INFINITY};
  for (int64_t b = threadIdx.x; b < nb; b += 256)
    for (int k = 0; k < 6; ++k) v[k] = fminf(v[k], bpart[6 * b + k]);
  for (int k = 0; k < 6; ++k) red[k][threadIdx.x] = v[k];
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o)
      for (int k = 0; k < 6; ++k) red[k][threadIdx.x] = fminf(red[k][threadIdx.x], red[k][threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x < 3) header[threadIdx.x] = f2o(red[threadIdx.x][0]);
  else if (threadIdx.x < 6) header[threadIdx.x] = f2o(-red[threadIdx.x][0]);
  // fragment offsets: each thread scans a contiguous chunk of blocks
  __shared__ uint32_t part[256];
  const int t = threadIdx.x;
  const int64_t chunk = (nb + 255) / 256;
  const int64_t b0 = t * chunk, b1 = b0 + chunk < nb ? b0 + chunk : nb;
  uint32_t sum = 0;
  for (int64_t b = b0; b < b1; ++b) sum += fcount[b];
  part[t] = sum;
  __syncthreads();
  for (int o = 1; o < 256; o <<= 1) {
    const uint32_t u = t >= o ? part[t - o] : 0u;
    __syncthreads();
    part[t] += u;
    __syncthreads();
  }
  uint32_t run = part[t] - sum;
  for (int64_t b = b0; b < b1; ++b) {
    const uint32_t c = fcount[b];
    fcount[b] = run;
    run += c;
  }
  if (t == 255) header[9] = part[255];
}

__device__ __forceinline__ uint32_t expand_bits10(uint32_t v) {
  v = (v * 0x00010001u) & 0xFF0000FFu;
  v = (v * 0x00000101u) & 0x0F00F00Fu;
  v = (v * 0x00000011u) & 0xC30C30C3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}

// Leaf fragments of one primitive: the ellipsoid is cut along its longest local axis into
// F = clamp(ceil(s_max / s_mid), 1, kMaxFragments) slabs; fragment j's box is the exact AABB of the slab
// a_j <= u <= b_j of the ellipsoid (u its coordinate along that axis in the unit-ball frame), which is the
// ellipsoid's tight AABB (create_aabbs.cu:77-114: half-extent_k = sqrt(sum_l (R_kl s_l)^2), q normalized) on the
// sides whose extreme point lies inside the slab. (Measured at 1M/1080p against the AABB of the slab's local box
// [a_j, b_j] x [-rho_j, rho_j]^2 clipped to the tight AABB: forward 6.44 -> 6.08 ms, build +0.09 ms with the
// ceil(s_max / s_mid) rule; ceil(0.75 s_max / s_mid), ceil(1.25 ...), ceil(1.5 ...) and at most 3 or 6 fragments
// all cost more in forward + build.)
// Boxes are padded by a relative 4e-6 so they stay conservative under rounding. Each fragment gets the 27-bit
// Morton code (9 bits per axis) of its box centre with a 3-bit box size class inserted after its first 15 bits.
// Fragments are written compactly in primitive order (fragment j of primitive i at foff[block] + the fragment count
// of the block's earlier primitives + j), so the sort and the tree see only the L fragments in use.
__global__ __launch_bounds__(256) void k_emit_fragments(ever_prims P, const uint32_t* __restrict__ header,
                                                        const uint32_t* __restrict__ foff, float* __restrict__ aabb,
                                                        uint32_t* __restrict__ fprim, uint32_t* __restrict__ keys,
                                                        uint32_t* __restrict__ vals) {
  // the block's fragments (at most 4 per primitive) are staged in LDS and stored as one contiguous run
  __shared__ uint32_t wsum[4];
  __shared__ float sbox[6 * kMaxFragments * 256];
  __shared__ uint32_t skey[kMaxFragments * 256], sprim[kMaxFragments * 256];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int F = 0, ax = 0;
  float sv[3] = {0.f, 0.f, 0.f};
  if (i < P.num_prims) {
    sv[0] = P.scales[3 * i];
    sv[1] = P.scales[3 * i + 1];
    sv[2] = P.scales[3 * i + 2];
    F = fragment_count(sv, ax);
  }
  // the primitive's first fragment: block offset + exclusive prefix of F over the block
  int inc = F;
  for (int o = 1; o < 64; o <<= 1) {
    const int u = __shfl_up(inc, o);
    inc += lane >= o ? u : 0;
  }
  if (lane == 63) wsum[wave] = (uint32_t)inc;
  __syncthreads();
  const uint32_t fb = foff[blockIdx.x];
  uint32_t first = (uint32_t)(inc - F);  // within the block
  for (int w = 0; w < wave; ++w) first += wsum[w];
  const int total = (int)(wsum[0] + wsum[1] + wsum[2] + wsum[3]);
  if (i < P.num_prims) {
    float lo[3] = {o2f(header[0]), o2f(header[1]), o2f(header[2])};
    float hi[3] = {o2f(header[3]), o2f(header[4]), o2f(header[5])};
    const float4 q = reinterpret_cast<const float4*>(P.quats)[i];
    const float qn = rsqrtf(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
    const float r = q.x * qn, x = q.y * qn, y = q.z * qn, z = q.w * qn;
    const float Rm[3][3] = {{1.f - 2.f * (y * y + z * z), 2.f * (x * y - r * z), 2.f * (x * z + r * y)},
                            {2.f * (x * y + r * z), 1.f - 2.f * (x * x + z * z), 2.f * (y * z - r * x)},
                            {2.f * (x * z - r * y), 2.f * (y * z + r * x), 1.f - 2.f * (x * x + y * y)}};
    const float mv[3] = {P.means[3 * i], P.means[3 * i + 1], P.means[3 * i + 2]};
    float tight[3];
    for (int k = 0; k < 3; ++k)
      tight[k] = sqrtf(Rm[k][0] * Rm[k][0] * sv[0] * sv[0] + Rm[k][1] * Rm[k][1] * sv[1] * sv[1] +
                       Rm[k][2] * Rm[k][2] * sv[2] * sv[2]);
    for (int j = 0; j < F; ++j) {
      const int slot = (int)first + j;
      const float a = -1.f + 2.f * j / F, b = -1.f + 2.f * (j + 1) / F;
      float* bx = sbox + 6 * slot;
      uint32_t qk[3];
      float maxext = 0.f, sext = 0.f;
      for (int k = 0; k < 3; ++k) {
        // world coordinate k over the slab a <= u <= b of the unit ball (u the split axis, x = mean + R diag(s) y):
        // x_k - mean_k = m_u u + m_p . y_p with |y_p| <= sqrt(1 - u^2), so its largest value is
        // f(u) = m_u u + |m_p| sqrt(1 - u^2), which peaks (at |m| = tight_k) where u = m_u / |m|: inside the slab
        // that is the ellipsoid's own extent, else the larger of f(a) and f(b); the smallest value likewise with -m_u
        float mu = 0.f, mp2 = 0.f;
        for (int l = 0; l < 3; ++l) {
          const float m = Rm[k][l] * sv[l];
          if (l == ax) mu = m;
          else mp2 += m * m;
        }
        const float mp = sqrtf(mp2), mm = tight[k];
        const float ca = sqrtf(fmaxf(1.f - a * a, 0.f)), cb = sqrtf(fmaxf(1.f - b * b, 0.f));
        const float us = mm > 0.f ? mu / mm : 0.f;
        const float up = (us >= a && us <= b) ? mm : fmaxf(mu * a + mp * ca, mu * b + mp * cb);
        const float dn = (-us >= a && -us <= b) ? mm : fmaxf(-mu * a + mp * ca, -mu * b + mp * cb);
        const float lo_k = mv[k] - fminf(dn, mm), hi_k = mv[k] + fminf(up, mm);
        const float pad = 4e-6f * fmaxf(hi_k - lo_k, tight[k]) + 1e-7f * fabsf(mv[k]);
        bx[k] = lo_k - pad;
        bx[3 + k] = hi_k + pad;
        const float ext = hi[k] - lo[k];
        const float u = ext > 0.f ? (0.5f * (lo_k + hi_k) - lo[k]) / ext : 0.5f;
        qk[k] = (uint32_t)fminf(fmaxf(u * 512.f, 0.f), 511.f);
        maxext = fmaxf(maxext, hi_k - lo_k);
        sext = fmaxf(sext, ext);
      }
      // size class of the box (0: the largest, an eighth of the scene or more; one class per halving, 7: the
      // smallest), inserted into the Morton code after its first 15 bits: boxes of one region are grouped by size
      // before they are split further, so large boxes end up high in the tree instead of widening the small-box
      // nodes they would otherwise share (measured: forward 9.3 -> 8.1 ms at 1M/1080p; at the top of the code or
      // after 9 bits it is worse, after 12, 18 or 21 bits or with 2 or 4 size bits no better)
      const float rel = sext > 0.f && maxext > 0.f ? sext / maxext : 256.f;
      const uint32_t sc = (uint32_t)fminf(fmaxf(floorf(log2f(rel)) - 1.f, 0.f), 7.f);
      const uint32_t m27 = (expand_bits10(qk[0]) << 2) | (expand_bits10(qk[1]) << 1) | expand_bits10(qk[2]);
      skey[slot] = ((m27 >> 12) << 15) | (sc << 12) | (m27 & 0xfffu);
      sprim[slot] = (uint32_t)i;
    }
  }
  __syncthreads();
  for (int l = threadIdx.x; l < 6 * total; l += 256) aabb[6 * (int64_t)fb + l] = sbox[l];
  for (int l = threadIdx.x; l < total; l += 256) {
    keys[fb + l] = skey[l];
    fprim[fb + l] = sprim[l];
    vals[fb + l] = fb + (uint32_t)l;
  }
}

// ---- LSD radix sort, 8-bit digits, tiles of kSortTile keys --------------------------------------------------
// Histogram: hist is digit-major, hist[digit * tiles + tile].
__global__ __launch_bounds__(kSortThreads) void k_radix_hist(const uint32_t* __restrict__ keys, int64_t n, int shift,
                                                             uint32_t* __restrict__ hist, int64_t tiles,
                                                             const uint32_t* __restrict__ dcount) {
  __shared__ uint32_t h[kRadix];
  if (dcount) n = (int64_t)*dcount < n ? (int64_t)*dcount : n;
  h[threadIdx.x] = 0;
  __syncthreads();
  int64_t base = (int64_t)blockIdx.x * kSortTile;
#pragma unroll
  for (int it = 0; it < kSortItems; ++it) {
    int64_t idx = base + it * kSortThreads + threadIdx.x;
    if (idx < n) atomicAdd(&h[(keys[idx] >> shift) & 0xffu], 1u);
  }
  __syncthreads();
  hist[(int64_t)threadIdx.x * tiles + blockIdx.x] = h[threadIdx.x];
}

// Exclusive scan of each digit's row of the histogram (one block per digit), and the digit totals.
__global__ __launch_bounds__(256) void k_scan_digits(uint32_t* __restrict__ hist, int64_t tiles,
                                                     uint32_t* __restrict__ totals) {
  __shared__ uint32_t part[256];
  const int t = threadIdx.x;
  uint32_t* row = hist + (int64_t)blockIdx.x * tiles;
  const int64_t chunk = (tiles + 255) / 256;
  const int64_t b = t * chunk, e = b + chunk < tiles ? b + chunk : tiles;
  uint32_t s = 0;
  for (int64_t i = b; i < e; ++i) s += row[i];
  part[t] = s;
  __syncthreads();
  for (int o = 1; o < 256; o <<= 1) {
    const uint32_t v = t >= o ? part[t - o] : 0u;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  uint32_t run = part[t] - s;
  for (int64_t i = b; i < e; ++i) {
    const uint32_t v = row[i];
    row[i] = run;
    run += v;
  }
  if (t == 255) totals[blockIdx.x] = part[255];
}

// Stable scatter of one tile of kSortTile keys. The tile is first ranked in LDS: kSortItems rounds of 256
// consecutive keys, ranks within a wave from 8 ballots (multi-split), across the 4 waves from an LDS prefix, placing
// each key at its tile-local sorted position (the tile's digit start + the keys of that digit before it). The
// per-wave digit counts and the running digit offsets are double-buffered by round parity, so a round needs two
// barriers. Then the sorted tile is written out in order: consecutive threads write consecutive addresses of a
// digit's run (a few hundred bytes on average) instead of one isolated word per digit and round.
__global__ __launch_bounds__(kSortThreads) void k_radix_scatter(const uint32_t* __restrict__ kin,
                                                                const uint32_t* __restrict__ vin,
                                                                uint32_t* __restrict__ kout,
                                                                uint32_t* __restrict__ vout, int64_t n, int shift,
                                                                const uint32_t* __restrict__ hist, int64_t tiles,
                                                                const uint32_t* __restrict__ totals,
                                                                const uint32_t* __restrict__ dcount) {
  __shared__ uint32_t skey[kSortTile], sval[kSortTile];
  __shared__ uint32_t lstart[kRadix], gbase[kRadix], th[kRadix];
  __shared__ uint32_t offs[2][kRadix];
  __shared__ uint32_t cnt[2][4][kRadix];
  __shared__ uint32_t wsum[2][4];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (dcount) n = (int64_t)*dcount < n ? (int64_t)*dcount : n;
  const int64_t base = (int64_t)blockIdx.x * kSortTile;
  if (base >= n) return;
  const int nv = n - base < kSortTile ? (int)(n - base) : kSortTile;
  // the tile's keys and values are loaded at once, so their latency is paid once
  uint32_t key[kSortItems], val[kSortItems];
#pragma unroll
  for (int it = 0; it < kSortItems; ++it) {
    const int l = it * kSortThreads + tid;
    key[it] = l < nv ? kin[base + l] : 0u;
    val[it] = l < nv ? vin[base + l] : 0u;
  }
  th[tid] = 0;
#pragma unroll
  for (int w = 0; w < 4; ++w) cnt[0][w][tid] = 0;
  lds_barrier();
#pragma unroll
  for (int it = 0; it < kSortItems; ++it)
    if (it * kSortThreads + tid < nv) atomicAdd(&th[(key[it] >> shift) & 0xffu], 1u);
  lds_barrier();
  // exclusive scans over the 256 digits (one per thread): the tile's digit counts give the tile-local digit starts,
  // the digit totals the digits' bases in the output
  const uint32_t tc = th[tid], dt = totals[tid];
  uint32_t inc = tc, inc2 = dt;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t u = __shfl_up(inc, o), u2 = __shfl_up(inc2, o);
    inc += lane >= o ? u : 0u;
    inc2 += lane >= o ? u2 : 0u;
  }
  if (lane == 63) {
    wsum[0][wave] = inc;
    wsum[1][wave] = inc2;
  }
  lds_barrier();
  uint32_t before = 0, before2 = 0;
  for (int w = 0; w < wave; ++w) {
    before += wsum[0][w];
    before2 += wsum[1][w];
  }
  const uint32_t ls = before + inc - tc;
  lstart[tid] = ls;
  offs[0][tid] = ls;
  gbase[tid] = hist[(int64_t)tid * tiles + blockIdx.x] + before2 + inc2 - dt;
  const uint64_t lt_mask = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
#pragma unroll
  for (int it = 0; it < kSortItems; ++it) {
    const int pb = it & 1;
    const bool valid = it * kSortThreads + tid < nv;
    const uint32_t digit = (key[it] >> shift) & 0xffu;
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      bool bit = (digit >> b) & 1u;
      uint64_t m = __ballot(bit);
      peers &= bit ? m : ~m;
    }
    const uint32_t rank = (uint32_t)__popcll(peers & lt_mask);
    const uint32_t count = (uint32_t)__popcll(peers);
    if (valid && rank == 0) cnt[pb][wave][digit] = count;
    lds_barrier();
    {
      const uint32_t c0 = cnt[pb][0][tid], c1 = cnt[pb][1][tid], c2 = cnt[pb][2][tid], c3 = cnt[pb][3][tid];
      cnt[pb][0][tid] = 0;
      cnt[pb][1][tid] = c0;
      cnt[pb][2][tid] = c0 + c1;
      cnt[pb][3][tid] = c0 + c1 + c2;
      offs[pb ^ 1][tid] = offs[pb][tid] + c0 + c1 + c2 + c3;
#pragma unroll
      for (int w = 0; w < 4; ++w) cnt[pb ^ 1][w][tid] = 0;
    }
    lds_barrier();
    if (valid) {
      const uint32_t pos = offs[pb][digit] + cnt[pb][wave][digit] + rank;
      skey[pos] = key[it];
      sval[pos] = val[it];
    }
  }
  lds_barrier();
#pragma unroll
  for (int it = 0; it < kSortItems; ++it) {
    const int l = it * kSortThreads + tid;
    if (l < nv) {
      const uint32_t k = skey[l];
      const uint32_t d = (k >> shift) & 0xffu;
      const uint32_t gp = gbase[d] + ((uint32_t)l - lstart[d]);
      kout[gp] = k;
      vout[gp] = sval[l];
    }
  }
}

// ---- Karras (2012) hierarchy over the sorted codes ---------------------------------------------------------
__device__ __forceinline__ int kdelta(const uint32_t* keys, int64_t n, int64_t a, int64_t b) {
  if (b < 0 || b >= n) return -1;
  uint32_t ka = keys[a], kb = keys[b];
  if (ka != kb) return __clz(ka ^ kb);
  return 32 + __clz((uint32_t)a ^ (uint32_t)b);
}

__global__ __launch_bounds__(256) void k_karras(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ header,
                                                float4* __restrict__ nodes, int2* __restrict__ range) {
  const int64_t n = header[9];  // leaves in use
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n - 1) return;
  int d = (kdelta(keys, n, i, i + 1) - kdelta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
  int dmin = kdelta(keys, n, i, i - d);
  int64_t lmax = 2;
  while (kdelta(keys, n, i, i + lmax * d) > dmin) lmax *= 2;
  int64_t l = 0;
  for (int64_t t = lmax / 2; t >= 1; t /= 2)
    if (kdelta(keys, n, i, i + (l + t) * d) > dmin) l += t;
  int64_t j = i + l * d;
  int dnode = kdelta(keys, n, i, j);
  int64_t s = 0;
  int64_t t = l;
  do {
    t = (t + 1) / 2;
    if (kdelta(keys, n, i, i + (s + t) * d) > dnode) s += t;
  } while (t > 1);
  int64_t gamma = i + s * d + (d < 0 ? d : 0);
  int64_t lo = i < j ? i : j, hi = i < j ? j : i;
  const int32_t left = (lo == gamma) ? ~(int32_t)gamma : (int32_t)gamma;
  const int32_t right = (hi == gamma + 1) ? ~(int32_t)(gamma + 1) : (int32_t)(gamma + 1);
  nodes[i] = make_float4(__int_as_float(left), __int_as_float(right), 0.f, 0.f);
  range[i] = make_int2((int32_t)lo, (int32_t)hi);
}

struct Box {
  float lx, ly, lz, hx, hy, hz;
};
__device__ __forceinline__ Box box_empty() { return {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY}; }
__device__ __forceinline__ Box box_union(Box a, Box b) {
  return {fminf(a.lx, b.lx), fminf(a.ly, b.ly), fminf(a.lz, b.lz),
          fmaxf(a.hx, b.hx), fmaxf(a.hy, b.hy), fmaxf(a.hz, b.hz)};
}
__device__ __forceinline__ Box box_load(const float* p) { return {p[0], p[1], p[2], p[3], p[4], p[5]}; }
__device__ __forceinline__ void box_store(float* p, Box b) {
  p[0] = b.lx; p[1] = b.ly; p[2] = b.lz; p[3] = b.hx; p[4] = b.hy; p[5] = b.hz;
}

// Per sorted leaf, blocks of kLeafBlock = 64 leaves (one wave): the leaf box in sorted order, in-block prefix and
// suffix unions (lane shuffles), and the block union (level 0 of the sparse table). Blocks beyond the L used
// leaves get an empty block box.
static_assert(kLeafBlock == 64, "k_leaf_blocks scans with the lanes of one wave");
__global__ __launch_bounds__(kLeafBlock) void k_leaf_blocks(const uint32_t* __restrict__ header,
                                                            const float* __restrict__ aabb,
                                                            const uint32_t* __restrict__ order,
                                                            float* __restrict__ sbox, float* __restrict__ pre,
                                                            float* __restrict__ suf, float* __restrict__ tab0) {
  const int64_t n = header[9];
  const int lane = threadIdx.x;
  if ((int64_t)blockIdx.x * kLeafBlock >= n) {
    if (lane == 0) box_store(tab0 + 6 * (int64_t)blockIdx.x, box_empty());
    return;
  }
  const int64_t s0 = (int64_t)blockIdx.x * kLeafBlock, sidx = s0 + lane;
  // the block's leaf, prefix and suffix boxes are staged in LDS and stored as contiguous runs
  __shared__ float st[3][6 * kLeafBlock];
  Box b = box_empty();
  if (sidx < n) b = box_load(aabb + 6 * (int64_t)order[sidx]);
  box_store(&st[0][6 * lane], b);
  // all six bounds as minima
  const float v[6] = {b.lx, b.ly, b.lz, -b.hx, -b.hy, -b.hz};
  float p[6], q[6];
  for (int k = 0; k < 6; ++k) p[k] = q[k] = v[k];
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      const float up = __shfl_up(p[k], o), dn = __shfl_down(q[k], o);
      if (lane >= o) p[k] = fminf(p[k], up);
      if (lane + o < 64) q[k] = fminf(q[k], dn);
    }
  }
  box_store(&st[1][6 * lane], {p[0], p[1], p[2], -p[3], -p[4], -p[5]});
  box_store(&st[2][6 * lane], {q[0], q[1], q[2], -q[3], -q[4], -q[5]});
  if (lane == 63) box_store(tab0 + 6 * (int64_t)blockIdx.x, {p[0], p[1], p[2], -p[3], -p[4], -p[5]});
  __syncthreads();
  const int nv = n - s0 < kLeafBlock ? (int)(n - s0) : kLeafBlock;
  for (int l = lane; l < 6 * nv; l += kLeafBlock) {
    sbox[6 * s0 + l] = st[0][l];
    pre[6 * s0 + l] = st[1][l];
    suf[6 * s0 + l] = st[2][l];
  }
}

// Sparse table levels 1..min(levels - 1, kTabFused) in one launch: a block loads its 256 entries of level 0 and the
// 2^kTabFused - 1 after them into LDS (entries beyond nb as empty boxes) and builds the levels there, each one
// shorter by its half-width, storing its own 256 entries of every level.
constexpr int kTabFused = 8;
__global__ __launch_bounds__(256) void k_table_levels_low(float* __restrict__ tab, int64_t nb, int levels) {
  constexpr int W = 256 + (1 << kTabFused) - 1;
  __shared__ float lds[2][6][W];
  const int tid = threadIdx.x;
  const int64_t b0 = (int64_t)blockIdx.x * 256;
  for (int j = tid; j < W; j += 256) {
    const Box x = b0 + j < nb ? box_load(tab + 6 * (b0 + j)) : box_empty();
    lds[0][0][j] = x.lx; lds[0][1][j] = x.ly; lds[0][2][j] = x.lz;
    lds[0][3][j] = -x.hx; lds[0][4][j] = -x.hy; lds[0][5][j] = -x.hz;
  }
  __syncthreads();
  const int top = levels - 1 < kTabFused ? levels - 1 : kTabFused;
  int valid = W;
  for (int k = 1; k <= top; ++k) {
    const int half = 1 << (k - 1), sb = (k - 1) & 1, db = k & 1;
    valid -= half;
    for (int j = tid; j < valid; j += 256)
      for (int f = 0; f < 6; ++f) lds[db][f][j] = fminf(lds[sb][f][j], lds[sb][f][j + half]);
    __syncthreads();
    if (b0 + tid < nb)
      box_store(tab + (int64_t)k * 6 * nb + 6 * (b0 + tid),
                {lds[db][0][tid], lds[db][1][tid], lds[db][2][tid], -lds[db][3][tid], -lds[db][4][tid],
                 -lds[db][5][tid]});
  }
}

// Sparse table level k over block unions: tab_k[b] = tab_{k-1}[b] U tab_{k-1}[b + 2^(k-1)].
__global__ __launch_bounds__(256) void k_table_level(const float* __restrict__ prev, float* __restrict__ cur,
                                                     int64_t nb, int64_t half) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  Box x = box_load(prev + 6 * b);
  if (b + half < nb) x = box_union(x, box_load(prev + 6 * (b + half)));
  box_store(cur + 6 * b, x);
}

__device__ __forceinline__ Box range_union(int32_t lo, int32_t hi, const float* sbox, const float* pre,
                                           const float* suf, const float* tab, int64_t nb) {
  const int32_t bl = lo / kLeafBlock, bh = hi / kLeafBlock;
  if (bl == bh) {
    Box x = box_empty();
    for (int32_t s = lo; s <= hi; ++s) x = box_union(x, box_load(sbox + 6 * (int64_t)s));
    return x;
  }
  Box x = box_union(box_load(suf + 6 * (int64_t)lo), box_load(pre + 6 * (int64_t)hi));
  const int32_t len = bh - bl - 1;
  if (len > 0) {
    const int k = 31 - __clz(len);
    const float* t = tab + (int64_t)k * 6 * nb;
    x = box_union(x, box_union(box_load(t + 6 * (int64_t)(bl + 1)), box_load(t + 6 * (int64_t)(bh - (1 << k)))));
  }
  return x;
}

// 4-wide node records: the record of binary node i describes up to four nodes of its binary subtree (its
// children, then the internal slot with the largest box surface opened twice), boxes stored per axis (SoA) so one
// record is 128 B:
//   f4[0] = lo.x of children 0..3, f4[1] = hi.x, f4[2] = lo.y, f4[3] = hi.y, f4[4] = lo.z, f4[5] = hi.z,
//   f4[6] = child refs (internal: binary node index of the child's record; leaf: ~primitive; empty: kEmptyChild)
// Only records reachable from the root (binary node 0) are visited.
// Box of every binary node (the union over its leaf range), once per node, into ibox: the records below gather
// their four child boxes from it instead of evaluating four range unions each.
__global__ __launch_bounds__(256) void k_binbox(const int2* __restrict__ range, const float* __restrict__ sbox,
                                                const float* __restrict__ pre, const float* __restrict__ suf,
                                                const float* __restrict__ tab, const uint32_t* __restrict__ header,
                                                float* __restrict__ ibox, int64_t nb) {
  const int64_t n = header[9];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n - 1) return;
  const int2 r = range[i];
  const Box b = range_union(r.x, r.y, sbox, pre, suf, tab, nb);
  float* o = ibox + 6 * i;
  o[0] = b.lx; o[1] = b.ly; o[2] = b.lz; o[3] = b.hx; o[4] = b.hy; o[5] = b.hz;
}

__device__ __forceinline__ Box child_box(int32_t c, const float* sbox, const float* ibox) {
  return c < 0 ? box_load(sbox + 6 * (int64_t)(~c)) : box_load(ibox + 6 * (int64_t)c);
}

__device__ __forceinline__ void node4_record(const float4* __restrict__ nodes2, const float* __restrict__ sbox,
                                             const float* __restrict__ ibox, const uint32_t* __restrict__ order,
                                             const uint32_t* __restrict__ fprim, int64_t i, float4 (*stage)[257],
                                             int t) {
  const float4 e = nodes2[i];
  // Children of the 4-wide record: start from the node's two children and open, twice, the internal slot with the
  // largest box surface (it is replaced by its two children), so a record has four children whenever its binary
  // subtree has four nodes to offer. Slots are scalars filled with selects (dynamically indexed arrays would live
  // in scratch); a slot holds the binary child encoding (internal >= 0, leaf ~index) or kSlotNone.
  constexpr int32_t kSlotNone = INT32_MIN;
  auto area = [](const Box& b) {
    const float dx = b.hx - b.lx, dy = b.hy - b.ly, dz = b.hz - b.lz;
    return dx * dy + dy * dz + dz * dx;
  };
  int32_t r0 = __float_as_int(e.x), r1 = __float_as_int(e.y), r2 = kSlotNone, r3 = kSlotNone;
  Box b0 = child_box(r0, sbox, ibox), b1 = child_box(r1, sbox, ibox), b2 = box_empty(), b3 = box_empty();
  {
    const float a0 = r0 >= 0 ? area(b0) : -1.f, a1 = r1 >= 0 ? area(b1) : -1.f;
    if (a0 >= 0.f || a1 >= 0.f) {
      const bool o0 = a0 >= a1;
      const float4 ge = nodes2[o0 ? r0 : r1];
      const int32_t y = __float_as_int(ge.x), z = __float_as_int(ge.y);
      const Box by = child_box(y, sbox, ibox), bz = child_box(z, sbox, ibox);
      if (o0) {
        r0 = y;
        b0 = by;
      } else {
        r1 = y;
        b1 = by;
      }
      r2 = z;
      b2 = bz;
      const float e0 = r0 >= 0 ? area(b0) : -1.f, e1 = r1 >= 0 ? area(b1) : -1.f, e2 = r2 >= 0 ? area(b2) : -1.f;
      if (e0 >= 0.f || e1 >= 0.f || e2 >= 0.f) {
        const int k = (e0 >= e1 && e0 >= e2) ? 0 : (e1 >= e2 ? 1 : 2);
        const float4 ge2 = nodes2[k == 0 ? r0 : (k == 1 ? r1 : r2)];
        const int32_t u = __float_as_int(ge2.x), v = __float_as_int(ge2.y);
        const Box bu = child_box(u, sbox, ibox), bv = child_box(v, sbox, ibox);
        if (k == 0) {
          r0 = u;
          b0 = bu;
        } else if (k == 1) {
          r1 = u;
          b1 = bu;
        } else {
          r2 = u;
          b2 = bu;
        }
        r3 = v;
        b3 = bv;
      }
    }
  }
  auto ref_of = [&](int32_t r) {
    return r == kSlotNone ? kEmptyChild : (r < 0 ? ~(int32_t)fprim[order[~r]] : r);
  };
  const int32_t ref[4] = {ref_of(r0), ref_of(r1), ref_of(r2), ref_of(r3)};
  const Box bx[4] = {b0, b1, b2, b3};
  stage[0][t] = make_float4(bx[0].lx, bx[1].lx, bx[2].lx, bx[3].lx);
  stage[1][t] = make_float4(bx[0].hx, bx[1].hx, bx[2].hx, bx[3].hx);
  stage[2][t] = make_float4(bx[0].ly, bx[1].ly, bx[2].ly, bx[3].ly);
  stage[3][t] = make_float4(bx[0].hy, bx[1].hy, bx[2].hy, bx[3].hy);
  stage[4][t] = make_float4(bx[0].lz, bx[1].lz, bx[2].lz, bx[3].lz);
  stage[5][t] = make_float4(bx[0].hz, bx[1].hz, bx[2].hz, bx[3].hz);
  stage[6][t] = make_float4(__int_as_float(ref[0]), __int_as_float(ref[1]), __int_as_float(ref[2]),
                            __int_as_float(ref[3]));
  stage[7][t] = make_float4(0.f, 0.f, 0.f, 0.f);
}

__global__ __launch_bounds__(256) void k_node4(const float4* __restrict__ nodes2, float4* __restrict__ nodes4,
                                               const float* __restrict__ sbox, const float* __restrict__ ibox,
                                               const uint32_t* __restrict__ header,
                                               const uint32_t* __restrict__ order, const uint32_t* __restrict__ fprim) {
  // the block's records are staged in LDS (one row per record float4, padded against bank conflicts) and stored
  // as one contiguous run, so each store instruction of a wave writes 1 KiB in sequence
  __shared__ float4 stage[8][257];
  const int64_t n = header[9];
  const int64_t i0 = (int64_t)blockIdx.x * blockDim.x;
  if (i0 >= n - 1) return;
  const int64_t i = i0 + threadIdx.x;
  if (i < n - 1) node4_record(nodes2, sbox, ibox, order, fprim, i, stage, threadIdx.x);
  __syncthreads();
  const int64_t nrec = (n - 1 - i0) < (int64_t)blockDim.x ? (n - 1 - i0) : (int64_t)blockDim.x;
  float4* nd = nodes4 + 8 * i0;
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    const int q = m * (int)blockDim.x + (int)threadIdx.x;
    if (q < 8 * nrec) nd[q] = stage[q & 7][q >> 3];
  }
}

hipError_t launch_sort_pairs(uint32_t* k0, uint32_t* v0, uint32_t* k1, uint32_t* v1, uint32_t* hist, int64_t n,
                             int bits, hipStream_t st, const uint32_t* dcount, int lo_bit, uint32_t** vres) {
  if (n <= 0) return hipSuccess;
  const int64_t tiles = (n + kSortTile - 1) / kSortTile;
  uint32_t *kin = k0, *vin = v0, *kout = k1, *vout = v1;
  int passes = 0;
  for (int shift = lo_bit; shift < bits; shift += 8, ++passes) {
    hipLaunchKernelGGL(k_radix_hist, dim3((unsigned)tiles), dim3(kSortThreads), 0, st, kin, n, shift, hist, tiles,
                       dcount);
    hipLaunchKernelGGL(k_scan_digits, dim3(kRadix), dim3(256), 0, st, hist, tiles, hist + (int64_t)kRadix * tiles);
    hipLaunchKernelGGL(k_radix_scatter, dim3((unsigned)tiles), dim3(kSortThreads), 0, st, kin, vin, kout, vout, n,
                       shift, hist, tiles, hist + (int64_t)kRadix * tiles, dcount);
    uint32_t* t = kin; kin = kout; kout = t;
    t = vin; vin = vout; vout = t;
  }
  if (vres) {
    *vres = vin;
    return hipGetLastError();
  }
  if (passes & 1) {  // bring the result back to (k0, v0)
    hipError_t e = hipMemcpyAsync(k0, kin, 4 * (size_t)n, hipMemcpyDeviceToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(v0, vin, 4 * (size_t)n, hipMemcpyDeviceToDevice, st);
    if (e != hipSuccess) return e;
  }
  return hipGetLastError();
}

hipError_t launch_update_colors(const ever_prims& P, void* ws, hipStream_t st) {
  const int64_t n = P.num_prims;
  if (n <= 0) return hipSuccess;
  SceneLayout L = scene_layout(n);
  hipLaunchKernelGGL(k_prim_prepass<true>, dim3(grid_for(n, 256)), dim3(256), 0, st, P, at<float4>(ws, L.rec),
                     at<float4>(ws, L.colr), at<float>(ws, L.bpart), nullptr);
  return hipGetLastError();
}

hipError_t launch_build(const ever_prims& P, void* ws, hipStream_t st) {
  const int64_t n = P.num_prims;
  SceneLayout L = scene_layout(n);
  uint32_t* header = at<uint32_t>(ws, L.header);
  float* aabb = at<float>(ws, L.aabb);
  uint32_t* fprim = at<uint32_t>(ws, L.fprim);
  uint32_t* k0 = at<uint32_t>(ws, L.keys0);
  uint32_t* v0 = at<uint32_t>(ws, L.vals0);
  uint32_t* k1 = at<uint32_t>(ws, L.keys1);
  uint32_t* v1 = at<uint32_t>(ws, L.vals1);
  uint32_t* hist = at<uint32_t>(ws, L.hist);
  float4* nodes = at<float4>(ws, L.nodes);
  float4* nodes4 = at<float4>(ws, L.nodes4);
  int2* range = at<int2>(ws, L.range);
  float* sbox = at<float>(ws, L.sbox);
  float* pre = at<float>(ws, L.pre);
  float* suf = at<float>(ws, L.suf);
  float* tab = at<float>(ws, L.tab);
  float4* rec = at<float4>(ws, L.rec);
  float4* colr = at<float4>(ws, L.colr);
  float* bpart = at<float>(ws, L.bpart);

  hipLaunchKernelGGL(k_build_init, dim3(1), dim3(64), 0, st, header);
  if (n <= 0) return hipGetLastError();
  const unsigned nb256 = grid_for(n, 256);
  // per-block fragment counts, then offsets, live in vals1 (free until the sort, nb256 <= S words)
  uint32_t* foff = v1;
  hipLaunchKernelGGL(k_prim_prepass<false>, dim3(nb256), dim3(256), 0, st, P, rec, colr, bpart, foff);
  hipLaunchKernelGGL(k_bounds_final, dim3(1), dim3(256), 0, st, bpart, (int64_t)nb256, header, foff);
  hipLaunchKernelGGL(k_emit_fragments, dim3(nb256), dim3(256), 0, st, P, header, foff, aabb, fprim, k0, v0);
  const int64_t S = L.slots;
  // the L fragments in use (L <= S, read by the sort kernels from header[9]); keys have 30 bits
  hipError_t e = launch_sort_pairs(k0, v0, k1, v1, hist, S, 30, st, header + 9);
  if (e != hipSuccess) return e;
  // sorted codes in keys0 (used leaves first), sorted leaf -> slot in vals0
  hipLaunchKernelGGL(k_karras, dim3(grid_for(S, 256)), dim3(256), 0, st, k0, header, nodes, range);
  const int64_t nb = L.leaf_blocks;
  hipLaunchKernelGGL(k_leaf_blocks, dim3((unsigned)nb), dim3(kLeafBlock), 0, st, header, aabb, v0, sbox, pre, suf,
                     tab);
  if (L.tab_levels > 1)
    hipLaunchKernelGGL(k_table_levels_low, dim3(grid_for(nb, 256)), dim3(256), 0, st, tab, nb, L.tab_levels);
  for (int k = kTabFused + 1; k < L.tab_levels; ++k)
    hipLaunchKernelGGL(k_table_level, dim3(grid_for(nb, 256)), dim3(256), 0, st, tab + (int64_t)(k - 1) * 6 * nb,
                       tab + (int64_t)k * 6 * nb, nb, (int64_t)1 << (k - 1));
  // the fragment boxes (aabb) are dead once the leaf blocks are built: their region holds the binary-node boxes
  float* ibox = aabb;
  hipLaunchKernelGGL(k_binbox, dim3(grid_for(S, 256)), dim3(256), 0, st, range, sbox, pre, suf, tab, header, ibox, nb);
  hipLaunchKernelGGL(k_node4, dim3(grid_for(S, 256)), dim3(256), 0, st, nodes, nodes4, sbox, ibox, header, v0, fprim);
  return hipGetLastError();
}

}  // namespace ever
