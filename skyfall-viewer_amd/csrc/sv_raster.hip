// sv_raster.hip -- tiled forward rasterizer and adjoint backward for gfx950.
//
// One wave64 owns one 16x16 pixel tile of one view. The tile is four 8x8
// quadrants; lane l shades pixel (l & 7, l >> 3) of every quadrant j =
// 0..3 (slot j of its 4 pixels), so one wave-wide reduction in the
// backward covers all 256 pixels, and a splat that misses a quadrant costs
// nothing there: the quadrant test is uniform over the wave. The tile's
// sorted splat list is staged through LDS in batches of B 64-byte records.
// Work items are (view, global tile); with a work counter the grid is
// persistent and pulls items from a device queue.
#include "sv_internal.hpp"

namespace sv {

constexpr int WT = 64;  // threads per tile (one wave)
constexpr int PPL = 4;  // pixels per lane

// Sorted id buffer and pair count of the step from its metadata; false
// when the step overflowed (nothing is rendered or written).
__device__ __forceinline__ bool resolve_meta(RasterArgs &a) {
  const StepMeta *m = a.meta;
  if (m->overflow) return false;
  a.vals = m->sel ? a.vals1 : a.vals0;
  a.n_isect = m->n_isect;
  return true;
}

// Workgroups are dispatched to the 8 XCDs round-robin (block b runs on XCD
// b mod 8). With xcd_remap the blocks of one XCD take one contiguous range
// of work items (a bijection on [0, n)), so neighbouring tiles, which read
// the same splats' records and write adjacent adjoint records, run on the
// same XCD and share its L2.
constexpr uint32_t NUM_XCD = 8;
__device__ __forceinline__ uint32_t xcd_remap(uint32_t b, uint32_t n) {
  const uint32_t q = n / NUM_XCD, r = n % NUM_XCD, x = b % NUM_XCD;
  return x * q + min(x, r) + b / NUM_XCD;
}

// Persistent launch: one work queue per XCD, each a contiguous part of the
// work items (the same parts as xcd_remap), its counter on a line of its
// own (WORK_QUEUE_STRIDE words apart: no two XCDs' atomics share a cache
// line). A workgroup pulls from its own XCD's queue and, once that is
// drained, from the others in turn; when every queue is drained it gets
// n_tiles_total (no item). A launch covers tiles [tile_begin,
// n_tiles_total).
static_assert(NUM_WORK_QUEUES == NUM_XCD, "one work queue per XCD");

// The tile of work item i of a launch: its list order, or with a tile
// order the i-th entry of it (a permutation of the launch's items).
__device__ __forceinline__ uint32_t item_tile(const RasterArgs &a, uint32_t i) {
  return a.tile_begin + (a.tile_order ? a.tile_order[i] : i);
}

// Called by the whole wave (the raster workgroups are one wave): lane 0 does
// the atomics, the item is read from it, every value stays scalar.
__device__ __forceinline__ uint32_t pull_item(const RasterArgs &a) {
  const uint32_t n = a.n_tiles_total - a.tile_begin, q = n / NUM_XCD, r = n % NUM_XCD;
  const uint32_t home = blockIdx.x % NUM_XCD;
#pragma unroll 1
  for (uint32_t d = 0; d < NUM_XCD; ++d) {
    const uint32_t x = (home + d) % NUM_XCD;
    const uint32_t len = q + (x < r ? 1u : 0u);
    uint32_t t = 0;
    if (threadIdx.x == 0) t = atomicAdd(a.work_counter + x * WORK_QUEUE_STRIDE, 1u);
    t = __builtin_amdgcn_readfirstlane(t);
    if (t < len) return item_tile(a, x * q + min(x, r) + t);
  }
  return a.n_tiles_total;
}

__device__ __forceinline__ uint32_t first_tile(const RasterArgs &a) {
  if (a.work_counter == nullptr)
    return item_tile(a, a.cfg.xcd_remap ? xcd_remap(blockIdx.x, gridDim.x) : blockIdx.x);
  return pull_item(a);
}

// ---------------------------------------------------------------------
// Longest-first work order. A tile's raster time grows with its list, and
// the lists are very uneven (most tiles of a view hold no splat, a few
// central ones thousands), so taking the work items in tile order leaves
// the longest tiles to start late and run alone at the end of the launch.
// The order lists the items each XCD takes, longest first, in the
// places that XCD takes them: the eight contiguous item ranges when
// workgroups of one XCD take one range (xcd_remap, the persistent queues),
// else the eight classes of items i = x mod 8 (workgroups reach the XCDs
// round-robin in item order). A tile's work is its list length for the
// forward and, for the backward, the entries some pixel visited in the
// forward. One workgroup per XCD buckets its items by work (four buckets
// per octave) with LDS counters and writes them
// bucket by bucket, longest first; the order inside a bucket is whatever the
// atomics give, which changes no result (every tile is computed on its own).
// ---------------------------------------------------------------------
constexpr int ORDER_BUCKETS = 124;
constexpr int ORDER_THREADS = 1024;

__device__ __forceinline__ uint32_t order_bucket(uint32_t len) {
  if (len < 4u) return len;
  const uint32_t lg = 31u - (uint32_t)__clz(len);  // >= 2
  return 4u + 4u * (lg - 2u) + ((len >> (lg - 2u)) & 3u);
}

// A tile's work: its forward's visited entries when given, else its list.
__device__ __forceinline__ uint32_t tile_work(const uint32_t *ranges, const uint32_t *cost, uint32_t t) {
  return cost ? cost[t] : ranges[t + 1] - ranges[t];
}

__global__ __launch_bounds__(ORDER_THREADS) void k_tile_order(const uint32_t *__restrict__ ranges,
                                                             const uint32_t *__restrict__ cost,
                                                             const StepMeta *__restrict__ meta, uint32_t tile_begin,
                                                             uint32_t n, uint32_t contiguous,
                                                             uint32_t *__restrict__ order) {
  if (meta->overflow) return;
  __shared__ uint32_t cnt[ORDER_BUCKETS];
  // this XCD's items: the j-th is base + j * step, j < m; its k-th place in
  // the order is base + k * step
  const uint32_t x = blockIdx.x, q = n / NUM_XCD, r = n % NUM_XCD;
  const uint32_t base = contiguous ? x * q + min(x, r) : x;
  const uint32_t step = contiguous ? 1u : NUM_XCD;
  const uint32_t m = contiguous ? q + (x < r ? 1u : 0u) : (x < n ? (n - x + NUM_XCD - 1u) / NUM_XCD : 0u);
  for (uint32_t b = threadIdx.x; b < (uint32_t)ORDER_BUCKETS; b += ORDER_THREADS) cnt[b] = 0u;
  __syncthreads();
  for (uint32_t j = threadIdx.x; j < m; j += ORDER_THREADS) {
    const uint32_t t = tile_begin + base + j * step;
    atomicAdd(&cnt[order_bucket(tile_work(ranges, cost, t))], 1u);
  }
  __syncthreads();
  // first place of each bucket, the longest bucket first
  if (threadIdx.x == 0) {
    uint32_t run = 0;
    for (int b = ORDER_BUCKETS - 1; b >= 0; --b) {
      const uint32_t c = cnt[b];
      cnt[b] = run;
      run += c;
    }
  }
  __syncthreads();
  for (uint32_t j = threadIdx.x; j < m; j += ORDER_THREADS) {
    const uint32_t i = base + j * step, t = tile_begin + i;
    const uint32_t k = atomicAdd(&cnt[order_bucket(tile_work(ranges, cost, t))], 1u);
    order[base + k * step] = i;
  }
}

void launch_tile_order(hipStream_t st, const uint32_t *ranges, const uint32_t *cost, const StepMeta *meta,
                       uint32_t tile_begin, uint32_t n, bool contiguous, uint32_t *order) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_tile_order, dim3(NUM_XCD), dim3(ORDER_THREADS), 0, st, ranges, cost, meta, tile_begin, n,
                     contiguous ? 1u : 0u, order);
}

// The next work item of a persistent launch (false: none left, or one
// workgroup per tile).
__device__ __forceinline__ bool next_tile(const RasterArgs &a, uint32_t &tg) {
  if (a.work_counter == nullptr) return false;
  tg = pull_item(a);
  return tg < a.n_tiles_total;
}

// DPP move of a float (0 in lanes disabled by the row mask).
template <int CTRL, int ROW_MASK = 0xF>
__device__ __forceinline__ float dpp(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROW_MASK, 0xF, false));
}

// Sum over each row of 16 lanes (fixed association; every lane of the row
// receives the row sum).
__device__ __forceinline__ float row_total(float v) {
  v = v + dpp<0xB1>(v);   // quad_perm [1,0,3,2]
  v = v + dpp<0x4E>(v);   // quad_perm [2,3,0,1]
  v = v + dpp<0x124>(v);  // row_ror:4
  v = v + dpp<0x128>(v);  // row_ror:8
  return v;
}

// Sum over the 64 lanes; the total is exact in lane 63 (fixed association
// (R3 + R2) + (R1 + R0) of the row sums; deterministic).
__device__ __forceinline__ float wave_total63(float v) {
  v = v + dpp<0xB1>(v);        // quad_perm [1,0,3,2]
  v = v + dpp<0x4E>(v);        // quad_perm [2,3,0,1]
  v = v + dpp<0x124>(v);       // row_ror:4
  v = v + dpp<0x128>(v);       // row_ror:8 -> every lane holds its row sum
  v = v + dpp<0x142, 0xA>(v);  // row_bcast:15 into rows 1 and 3
  v = v + dpp<0x143, 0xC>(v);  // row_bcast:31 into rows 2 and 3
  return v;
}

// Butterfly levels of a reduce-scatter over the wave. Each takes two
// per-lane values x, y and returns, in the half of the lanes whose level
// bit is 0, x summed with its partner lane's x, and in the other half y
// summed with its partner's y.
// Level 32: v_permlane32_swap exchanges the upper half of x with the lower
// half of y.
__device__ __forceinline__ float pair_sum32(float x, float y) {
  const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(x), __float_as_uint(y), false, false);
  return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}
// Level 16: v_permlane16_swap exchanges the odd rows of x with the even
// rows of y.
__device__ __forceinline__ float pair_sum16(float x, float y) {
  const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(x), __float_as_uint(y), false, false);
  return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}
// Level 8: row_ror:8 is lane ^ 8 within a row.
__device__ __forceinline__ float pair_sum8(float x, float y) {
  const bool hi = (threadIdx.x & 8u) != 0;
  const float mine = hi ? y : x, other = hi ? x : y;
  return mine + dpp<0x128>(other);
}

// Sums of eight per-lane values over the 64 lanes: lane l receives the sum
// of value l >> 3 (the eight lanes of a group hold the same bits). Three
// butterfly levels, then a sum over the group (quad_perm, quad_perm,
// row_half_mirror); fixed association, deterministic.
__device__ __forceinline__ float wave_scatter8(float v0, float v1, float v2, float v3, float v4, float v5,
                                               float v6, float v7) {
  const float a0 = pair_sum32(v0, v4), a1 = pair_sum32(v1, v5);
  const float a2 = pair_sum32(v2, v6), a3 = pair_sum32(v3, v7);
  const float b0 = pair_sum16(a0, a2), b1 = pair_sum16(a1, a3);
  float c = pair_sum8(b0, b1);
  c = c + dpp<0xB1>(c);   // quad_perm [1,0,3,2]
  c = c + dpp<0x4E>(c);   // quad_perm [2,3,0,1]
  c = c + dpp<0x141>(c);  // row_half_mirror
  return c;
}

__device__ __forceinline__ uint32_t wave_max_u32(uint32_t m) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = max(m, (uint32_t)__shfl_xor((int)m, o, 64));
  return m;
}

// 0.0f from a register move the optimiser does not fold.
__device__ __forceinline__ float opaque_zero() {
  float z;
  asm volatile("v_mov_b32 %0, 0" : "=v"(z));
  return z;
}

__device__ __forceinline__ float sign_ref(float x) {
  return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f);
}

struct TileGeom {
  uint32_t s, tx, ty;
  const DevView *V;
};

__device__ __forceinline__ TileGeom tile_geom(const RasterArgs &a, uint32_t tg) {
  TileGeom t;
  t.s = a.tile_view[tg];
  t.V = &a.views[t.s];
  const uint32_t lt = tg - t.V->tile_base;
  t.tx = lt % t.V->tiles_x;
  t.ty = lt / t.V->tiles_x;
  return t;
}

// Pixel of slot j (quadrant j) of this lane.
__device__ __forceinline__ uint32_t pix_x(const TileGeom &t, int j) {
  return t.tx * 16u + 8u * (uint32_t)(j & 1) + (threadIdx.x & 7u);
}
__device__ __forceinline__ uint32_t pix_y(const TileGeom &t, int j) {
  return t.ty * 16u + 8u * (uint32_t)(j >> 1) + ((threadIdx.x & 63u) >> 3);
}

// Position of an intersection in emit order: the Gaussian's first
// intersection (its projection block's base + q3.x, the prefix within the
// block) plus its row-major index inside the Gaussian's tile rectangle
// (training-tiles.wgsl:238-246).
__device__ __forceinline__ uint32_t emit_pos(const float4 &q3, uint32_t base, uint32_t tx, uint32_t ty) {
  const uint32_t u = __float_as_uint(q3.y), w = __float_as_uint(q3.z);
  const uint32_t x0 = u & 0xffffu, y0 = u >> 16, x1 = w & 0xffffu;
  return base + __float_as_uint(q3.x) + (ty - y0) * (x1 - x0) + (tx - x0);
}

// Pixel-centre bounds of the quadrants of a tile: x of quadrant columns
// 0/1, y of quadrant rows 0/1 (an empty range has lo > hi).
struct QuadBounds {
  float x0[2], x1[2], y0[2], y1[2];
};

__device__ __forceinline__ QuadBounds quad_bounds(const TileGeom &t, const DevView &V) {
  QuadBounds b;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const uint32_t xs = t.tx * 16u + 8u * h, ys = t.ty * 16u + 8u * h;
    b.x0[h] = (float)xs + 0.5f;
    b.x1[h] = xs < V.W ? (float)min(xs + 7u, V.W - 1u) + 0.5f : -1.0f;
    b.y0[h] = (float)ys + 0.5f;
    b.y1[h] = ys < V.H ? (float)min(ys + 7u, V.H - 1u) + 0.5f : -1.0f;
  }
  return b;
}

// Minimum over the quadrant's pixel-centre rectangle of the power p(d) =
// 0.5 (a dx^2 + c dy^2) + b dx dy (a convex quadratic with its minimum 0 at
// the mean). With the mean outside the rectangle the minimum lies on an edge
// that faces the mean (on the segment from any point of the rectangle to the
// mean, p decreases, and the segment leaves the rectangle through such an
// edge); along the edge x = xe the minimum is at dy = -b dx / c, clamped to
// the edge, and likewise for y = ye. Also returns the sum of the absolute
// terms at the evaluated points (the scale of their rounding error).
__device__ __forceinline__ float rect_min_power(float mx, float my, float a, float b, float c, float inv_a,
                                                float inv_c, float x0, float x1, float y0, float y1,
                                                float *scale) {
  const float dx = fminf(fmaxf(mx, x0), x1) - mx;  // 0 when mx is inside [x0, x1]
  const float dy = fminf(fmaxf(my, y0), y1) - my;
  float best = 0.0f, s = 0.0f;
  if (dx != 0.0f || dy != 0.0f) {
    best = INFINITY;
    if (dx != 0.0f) {
      const float t = fminf(fmaxf(-b * dx * inv_c, y0 - my), y1 - my);
      best = fminf(best, 0.5f * (a * dx * dx + c * t * t) + b * dx * t);
      s = 0.5f * (a * dx * dx + c * t * t) + fabsf(b * dx * t);
    }
    if (dy != 0.0f) {
      const float t = fminf(fmaxf(-b * dy * inv_a, x0 - mx), x1 - mx);
      best = fminf(best, 0.5f * (a * t * t + c * dy * dy) + b * t * dy);
      s = fmaxf(s, 0.5f * (a * t * t + c * dy * dy) + fabsf(b * t * dy));
    }
  }
  *scale = s;
  return best;
}

// Quadrants of the tile that an entry can colour: bit h + 2 v for the
// quadrant of column h, row v. First the box |dx| <= ext_x, |dy| <= ext_y
// around the mean; then the power over the quadrant's pixel centres must
// come within power_max. That bound is kept conservative: power_max already
// sits 0.02 above ln(255 opacity), and the minimum must exceed it by more
// than 2e-6 times the size of the terms (the rounding of this estimate and
// of the per-pixel power, a few float ulps of the largest term, on any pixel
// inside the box). A quadrant dropped here holds no pixel that takes the
// entry.
__device__ __forceinline__ uint32_t quad_mask(const float4 &q0, float c, const float4 &q2, const QuadBounds &qb) {
  bool cx[2], cy[2];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    cx[h] = qb.x1[h] >= qb.x0[h] && fabsf(q0.x - fminf(fmaxf(q0.x, qb.x0[h]), qb.x1[h])) <= q2.z;
    cy[h] = qb.y1[h] >= qb.y0[h] && fabsf(q0.y - fminf(fmaxf(q0.y, qb.y0[h]), qb.y1[h])) <= q2.w;
  }
  uint32_t qm = (uint32_t)(cx[0] && cy[0]) | ((uint32_t)(cx[1] && cy[0]) << 1) |
                ((uint32_t)(cx[0] && cy[1]) << 2) | ((uint32_t)(cx[1] && cy[1]) << 3);
  if (qm == 0u) return 0u;
  const float a = q0.z, b = q0.w;
  const float inv_a = __builtin_amdgcn_rcpf(a), inv_c = __builtin_amdgcn_rcpf(c);
  // the largest sum of absolute terms at a pixel inside the box
  const float sbox = 0.5f * (a * q2.z * q2.z + c * q2.w * q2.w) + fabsf(b) * q2.z * q2.w;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    if (!((qm >> j) & 1u)) continue;
    float s;
    const float mn = rect_min_power(q0.x, q0.y, a, b, c, inv_a, inv_c, qb.x0[j & 1], qb.x1[j & 1], qb.y0[j >> 1],
                                    qb.y1[j >> 1], &s);
    if (mn > q2.y + 2e-6f * (sbox + s)) qm &= ~(1u << j);
  }
  return qm;
}

// Load up to B entries of a tile list from `first` (count cnt) into LDS,
// keeping only those whose power <= power_max region (axis-aligned bounds
// ext_x, ext_y around the mean) overlaps the pixel centres of some quadrant
// of `wq`; kept entries stay in list order. s_k holds the entry's index
// in the batch (bits 0..15) and the mask of quadrants it overlaps (bits
// 16..19). With zmask, zmask[first + k] = 0 for every entry not kept.
// Returns how many were kept.
template <int B>
__device__ __forceinline__ uint32_t load_batch(const RasterArgs &a, size_t vbase, uint32_t first,
                                               uint32_t cnt, const QuadBounds &qb, uint32_t wq,
                                               float4 *s_q0, float4 *s_q1, float4 *s_q2, uint32_t *s_k,
                                               uint8_t *zmask) {
  const int lane = threadIdx.x & 63;
  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  uint32_t kept = 0;
  for (uint32_t r = 0; r < cnt; r += WT) {
    const uint32_t k = r + lane;
    bool touch = false;
    uint32_t qm = 0;
    float4 q0, q1, q2;
    if (k < cnt) {
      const uint32_t id = min(a.vals[first + k], a.N - 1u);
      const float4 *q = a.rec + (vbase + id) * 4;
      q0 = q[0];
      q1 = q[1];
      q2 = q[2];
      qm = quad_mask(q0, q1.x, q2, qb);
      touch = (qm & wq) != 0;
      if (zmask && !touch) zmask[first + k] = 0;
    }
    const uint64_t m = __ballot(touch);
    if (touch) {
      const uint32_t pos = kept + (uint32_t)__popcll(m & lt_mask);
      s_q0[pos] = q0;
      s_q1[pos] = q1;
      s_q2[pos] = q2;
      s_k[pos] = k | (qm << 16);
    }
    kept += (uint32_t)__popcll(m);
  }
  return kept;
}

// The backward's batch: up to B entries of a tile list from `first` (count
// cnt), in list order. With the forward's taken masks (a.tmask) the entries
// some pixel took, with the quadrants where one did (only their records are
// loaded); without them every record is loaded and the entries and
// quadrants quad_mask keeps are kept. s_k as in load_batch; s_p holds the
// kept entry's emit position. Returns how many were kept.
template <int B>
__device__ __forceinline__ uint32_t load_batch_bwd(const RasterArgs &a, size_t vbase, uint32_t first,
                                                   uint32_t cnt, const QuadBounds &qb, float4 *s_q0,
                                                   float4 *s_q1, float4 *s_q2, uint32_t *s_k, uint32_t *s_p,
                                                   uint32_t slot, uint32_t tx, uint32_t ty) {
  const int lane = threadIdx.x & 63;
  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  uint32_t kept = 0;
  for (uint32_t r = 0; r < cnt; r += WT) {
    const uint32_t k = r + lane;
    uint32_t qm = 0, id = 0;
    float4 q0, q1, q2, q3;
    if (k < cnt) {
      if (a.tmask) qm = a.tmask[first + k];
      if (!a.tmask || qm) {
        id = min(a.vals[first + k], a.N - 1u);
        const float4 *q = a.rec + (vbase + id) * 4;
        q0 = q[0];
        q1 = q[1];
        q2 = q[2];
        q3 = q[3];
        if (!a.tmask) qm = quad_mask(q0, q1.x, q2, qb);
      }
    }
    const bool touch = qm != 0u;
    const uint64_t m = __ballot(touch);
    if (touch) {
      const uint32_t pos = kept + (uint32_t)__popcll(m & lt_mask);
      s_q0[pos] = q0;
      s_q1[pos] = q1;
      s_q2[pos] = q2;
      s_k[pos] = k | (qm << 16);
      s_p[pos] = emit_pos(q3, a.bbase[(size_t)slot * a.nb + (id >> 8)], tx, ty);
    }
    kept += (uint32_t)__popcll(m);
  }
  return kept;
}

// Alpha of one splat record (q0, q1) at pixel centre (fx, fy), and
// whether the pixel takes it (training-forward-tiles.wgsl:183-197): power
// = 0.5 (a dx^2 + c dy^2) + b dx dy, skipped if < 0; alpha = min(opacity
// e^-power, 0.999), skipped if < 1/255. The power goes into the exp
// clamped to 80 (the exp's argument range): e^-80 opacity is far below
// 1/255, so a clamped pair is skipped as the unclamped one is.
__device__ __forceinline__ bool frag_alpha(float fx, float fy, const float4 &x0, const float4 &x1, float *as) {
  const float dx = fx - x0.x;
  const float dy = fy - x0.y;
  const float sg = sv_power(dx, dy, x0.z, x0.w, x1.x);
  const float gauss = sv_exp_alpha(-fminf(sg, 80.0f));
  *as = fminf(x1.y * gauss, SV_MAX_ALPHA);
  return !(sg < 0.0f) && !(*as < SV_MIN_ALPHA);
}

// Front-to-back compositing of one taken entry into one pixel; the pixel
// is done, with the entry as its last index, once 1 - A < 1e-4.
__device__ __forceinline__ void composite(float as, float r, float g, float b, uint32_t idx, float &c0, float &c1,
                                          float &c2, float &A, bool &done, uint32_t &last) {
  const float contrib = as * (1.0f - A);
  // the colour sums as fused multiply-adds (the alpha and transmittance
  // chain, which decides every skip and the end, stays unfused)
  c0 = fmaf(r, contrib, c0);
  c1 = fmaf(g, contrib, c1);
  c2 = fmaf(b, contrib, c2);
  A = A + contrib;
  const bool fin = A > SV_A_DONE;  // 1 - A < 1e-4; selects, not a branch
  done = done | fin;
  last = fin ? idx : last;
}

// The LDS batch of one wave and its synchronisation: a wave's LDS accesses
// are processed in issue order, so one wave needs only a fence that keeps
// the compiler from moving them across (and waits for them); workgroups of
// several independent waves use it instead of a workgroup barrier.
template <int B>
struct FwdLds {
  float4 q0[B], q1[B], q2[B];
  uint32_t k[B];
  // byte j of word e: some pixel of quadrant j composited kept entry e
  uint32_t t[B];
};

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// ---------------------------------------------------------------------
// training_tiled_forward (training-forward-tiles.wgsl:107-215). Per pixel
// (centre +0.5) over the tile list front to back: power = 0.5(a dx^2 +
// c dy^2) + b dx dy, skipped if < 0; alpha = min(sigmoid(logit) e^-power,
// 0.999), skipped if < 1/255; C += c alpha (1 - A); A += alpha (1 - A);
// stop when 1 - A < 1e-4 after the add. last = last visited index (the
// saturating entry, else the end of the list). Output C + (1 - A) bg.
// Empty tile: bg, A = 0, last = NO_LAST.
//
// One wave renders the quadrants `wq` (bit h + 2 v) of the 16x16 tile tg
// (the binning tile: its list is what every pixel composites, in order).
// The raster tile is a launch choice only: 16 px = one wave, all four
// quadrants; 8 px = one wave per quadrant; 32 px = a workgroup of four
// waves, one 16x16 tile each. Per pixel the operations and their order are
// the same in all three, so the images are identical bit for bit.
// ---------------------------------------------------------------------
template <int B>
__device__ __forceinline__ void forward_tile(const RasterArgs &a, uint32_t tg, uint32_t wq, FwdLds<B> &L) {
  const TileGeom t = tile_geom(a, tg);
  const DevView &V = *t.V;
  const size_t P = (size_t)V.W * V.H;
  float *col = a.color + 3 * V.pix_off;
  const uint32_t start = a.ranges[tg], end = a.ranges[tg + 1];
  // per tile, for the backward's work order: how many entries of its list
  // some pixel visits (a wave that renders the whole tile writes it)
  uint32_t *const cost = wq == 0xFu ? a.tile_cost : nullptr;
  if (start >= end) {
    if (cost && (threadIdx.x & 63u) == 0u) cost[tg] = 0u;
#pragma unroll
    for (int j = 0; j < PPL; ++j) {
      const uint32_t px = pix_x(t, j), py = pix_y(t, j);
      if (((wq >> j) & 1u) && px < V.W && py < V.H) {
        const size_t lp = (size_t)py * V.W + px;
        col[lp] = a.cfg.bg[0];
        col[P + lp] = a.cfg.bg[1];
        col[2 * P + lp] = a.cfg.bg[2];
        a.alpha[V.pix_off + lp] = 0.0f;
        a.last[V.pix_off + lp] = NO_LAST;
      }
    }
    return;
  }
  const size_t vbase = (size_t)t.s * a.N;
  const QuadBounds qb = quad_bounds(t, V);
  float c0[PPL], c1[PPL], c2[PPL], A[PPL], fxh[2], fyh[2];
  uint32_t last[PPL];
  bool done[PPL];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    fxh[h] = (float)pix_x(t, h) + 0.5f;
    fyh[h] = (float)pix_y(t, 2 * h) + 0.5f;
  }
#pragma unroll
  for (int j = 0; j < PPL; ++j) {
    c0[j] = c1[j] = c2[j] = A[j] = 0.0f;
    last[j] = end - 1u;
    done[j] = !(((wq >> j) & 1u) && pix_x(t, j) < V.W && pix_y(t, j) < V.H);
  }
  // the quadrant bits of an entry's word that this wave renders
  const uint32_t wmask = wq << 16;
  // Taken masks, for the backward: per entry, the quadrants where some pixel
  // composited it. Every lane that composites kept entry e in quadrant j
  // stores 1 into byte j of L.t[e] (all of them the same value; always, as
  // a branch would cost more than the store), and after the batch every
  // loaded entry's mask goes to its list position (0 for the entries culled
  // here or left when every pixel was done). Only a wave that renders the
  // whole tile writes them: it loads every entry any pixel of the tile
  // reaches.
  uint8_t *const tm = wq == 0xFu ? a.tmask : nullptr;
  uint8_t *const lt = reinterpret_cast<uint8_t *>(L.t);
  const uint32_t lane = threadIdx.x & 63u;
  for (uint32_t b0 = start; b0 < end; b0 += B) {
    const uint32_t cnt = min((uint32_t)B, end - b0);
    const uint32_t kept = load_batch<B>(a, vbase, b0, cnt, qb, wq, L.q0, L.q1, L.q2, L.k, tm);
    for (uint32_t x = lane; x < kept; x += WT) L.t[x] = 0u;
    wave_sync();
    // Two entries per iteration: their alphas are independent (only the
    // compositing that follows is ordered), so the two exp chains
    // interleave. Per pixel the same operations run in the same order.
    uint32_t e = 0;
    for (; e + 1u < kept; e += 2u) {
      // same words on every lane: the quadrant tests below are uniform
      const uint32_t ka = __builtin_amdgcn_readfirstlane(L.k[e]) & (wmask | 0xFFFFu);
      const uint32_t kb = __builtin_amdgcn_readfirstlane(L.k[e + 1u]) & (wmask | 0xFFFFu);
      const float4 a0 = L.q0[e], a1 = L.q1[e], a2 = L.q2[e];
      const float4 b0r = L.q0[e + 1u], b1 = L.q1[e + 1u], b2 = L.q2[e + 1u];
#pragma unroll
      for (int j = 0; j < PPL; ++j) {
        const bool ua = (ka >> (16 + j)) & 1u, ub = (kb >> (16 + j)) & 1u;
        if (!(ua || ub)) continue;
        if (done[j]) continue;
        float asa, asb;
        const bool va = ua && frag_alpha(fxh[j & 1], fyh[j >> 1], a0, a1, &asa);
        const bool vb = ub && frag_alpha(fxh[j & 1], fyh[j >> 1], b0r, b1, &asb);
        if (va) {
          composite(asa, a1.z, a1.w, a2.x, b0 + (ka & 0xFFFFu), c0[j], c1[j], c2[j], A[j], done[j], last[j]);
          lt[4u * e + j] = 1u;
        }
        if (vb && !done[j]) {
          composite(asb, b1.z, b1.w, b2.x, b0 + (kb & 0xFFFFu), c0[j], c1[j], c2[j], A[j], done[j], last[j]);
          lt[4u * e + 4u + j] = 1u;
        }
      }
      if (__all((done[0] && done[1]) && (done[2] && done[3]))) break;
    }
    if (e + 1u == kept) {
      const uint32_t ka = __builtin_amdgcn_readfirstlane(L.k[e]) & (wmask | 0xFFFFu);
      const float4 a0 = L.q0[e], a1 = L.q1[e], a2 = L.q2[e];
#pragma unroll
      for (int j = 0; j < PPL; ++j) {
        if (!((ka >> (16 + j)) & 1u)) continue;
        if (done[j]) continue;
        float asa;
        if (!frag_alpha(fxh[j & 1], fyh[j >> 1], a0, a1, &asa)) continue;
        composite(asa, a1.z, a1.w, a2.x, b0 + (ka & 0xFFFFu), c0[j], c1[j], c2[j], A[j], done[j], last[j]);
        lt[4u * e + j] = 1u;
      }
    }
    // a wave vote, then the sync that protects the LDS batch
    const bool all_done = __all((done[0] && done[1]) && (done[2] && done[3]));
    if (tm) {
      wave_sync();
      for (uint32_t x = lane; x < kept; x += WT) {
        const uint32_t w = L.t[x];
        tm[b0 + (L.k[x] & 0xFFFFu)] = (uint8_t)((w & 1u) | ((w >> 7) & 2u) | ((w >> 14) & 4u) | ((w >> 21) & 8u));
      }
    }
    wave_sync();
    if (all_done) break;
  }
#pragma unroll
  for (int j = 0; j < PPL; ++j) {
    const uint32_t px = pix_x(t, j), py = pix_y(t, j);
    if (((wq >> j) & 1u) && px < V.W && py < V.H) {
      const size_t lp = (size_t)py * V.W + px;
      const float T = 1.0f - A[j];
      col[lp] = c0[j] + T * a.cfg.bg[0];
      col[P + lp] = c1[j] + T * a.cfg.bg[1];
      col[2 * P + lp] = c2[j] + T * a.cfg.bg[2];
      a.alpha[V.pix_off + lp] = A[j];
      a.last[V.pix_off + lp] = last[j];
    }
  }
  if (cost) {
    uint32_t m = start;
#pragma unroll
    for (int j = 0; j < PPL; ++j)
      if (pix_x(t, j) < V.W && pix_y(t, j) < V.H) m = max(m, last[j] + 1u);
    m = wave_max_u32(m);
    if ((threadIdx.x & 63u) == 0u) cost[tg] = m - start;
  }
}

// 16-px raster tiles: one wave per tile (work item = global tile; the
// persistent launch pulls them from the work queues).
template <int B>
__global__ __launch_bounds__(WT, 7) void k_forward(RasterArgs ain) {
  RasterArgs a = ain;
  if (!resolve_meta(a)) return;
  __shared__ FwdLds<B> L;
  uint32_t tg = first_tile(a);
  while (tg < a.n_tiles_total) {
    forward_tile<B>(a, tg, 0xFu, L);
    if (!next_tile(a, tg)) break;
  }
}

// 8-px raster tiles: one wave per quadrant (work item = 4 global tile + q).
template <int B>
__global__ __launch_bounds__(WT, 7) void k_forward8(RasterArgs ain) {
  RasterArgs a = ain;
  if (!resolve_meta(a)) return;
  __shared__ FwdLds<B> L;
  const uint32_t n = 4u * a.n_tiles_total;
  const uint32_t item = a.cfg.xcd_remap ? xcd_remap(blockIdx.x, gridDim.x) : blockIdx.x;
  if (item >= n) return;
  forward_tile<B>(a, item >> 2, 1u << (item & 3u), L);
}

// 32-px raster tiles: a workgroup of four waves, wave w the 16-px tile
// (2 bx + (w & 1), 2 by + (w >> 1)) of block (bx, by) of its view.
template <int B>
__global__ __launch_bounds__(4 * WT, 7) void k_forward32(RasterArgs ain) {
  RasterArgs a = ain;
  if (!resolve_meta(a)) return;
  __shared__ FwdLds<B> L[4];
  const uint32_t blk = a.cfg.xcd_remap ? xcd_remap(blockIdx.x, gridDim.x) : blockIdx.x;
  if (blk >= a.n_blocks_total) return;
  const uint32_t w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const DevView &V = a.views[a.block_view[blk]];
  const uint32_t lb = blk - V.blk_base;
  const uint32_t tx = 2u * (lb % V.blk_x) + (w & 1u), ty = 2u * (lb / V.blk_x) + (w >> 1);
  if (tx >= V.tiles_x || ty >= V.tiles_y) return;
  forward_tile<B>(a, V.tile_base + ty * V.tiles_x + tx, 0xFu, L[w]);
}

// ---------------------------------------------------------------------
// Adjoint backward, order-correct. Math per (pixel, splat) is that of
// training_tiled_backward (training-backward-tiles.wgsl:318-396): reverse
// traversal from A_final and the last index, T_before = T / max(1 -
// alpha, 1e-4), no alpha > 0.98 skip. Instead of per-pixel float CAS
// atomics into 54 per-Gaussian slots, the wave sums the 9 screen-space
// quantities of one splat (sum of res * alpha * T_before per channel,
// dL/dlogit, and the moments of q = -dL/dpower over the pixel offsets
// (dx, dy): sums of q dx^2, q dx dy, q dy^2, q dx, q dy, from which the
// record gather forms dL/dconic abc and dL/dmean xy) over the tile's 256
// pixels (in registers over each lane's 4 pixels, then a fixed butterfly
// reduce-scatter over the wave for eight of them and a DPP tree for the
// ninth) and, for an entry some pixel took, writes one 36-byte record at
// the intersection's emit position and sets its flag. With the forward's
// taken masks it visits only the taken entries and quadrants. The SH chain
// is applied per Gaussian afterwards (exact: the SH basis depends only on
// the Gaussian and the view). Deterministic.
// The residual is fused (w_l2 (p - t) + w_l1 sign(p - t)) or read from
// a.residual.
// ---------------------------------------------------------------------
template <int B>
__global__ __launch_bounds__(WT, 5) void k_backward(RasterArgs ain) {
  RasterArgs a = ain;
  if (!resolve_meta(a)) return;
  __shared__ float4 s_q0[B];
  __shared__ float4 s_q1[B];
  __shared__ float4 s_q2[B];
  __shared__ uint32_t s_k[B];
  __shared__ uint32_t s_p[B];
  const int lane = threadIdx.x;
  uint32_t tg = first_tile(a);
  while (tg < a.n_tiles_total) {
    const TileGeom t = tile_geom(a, tg);
    const DevView &V = *t.V;
    const size_t P = (size_t)V.W * V.H;
    const uint32_t start = a.ranges[tg], end = a.ranges[tg + 1];
    const size_t vbase = (size_t)t.s * a.N;
    if (start < end) {
      // sacc: the residual's dot product with the colour accumulated behind
      // the current entry (<r, acc>, background first); the update acc' =
      // alpha c + (1 - alpha) acc is linear, so <r, acc'> = alpha <r, c> +
      // (1 - alpha) <r, acc> keeps one sum per pixel instead of three
      float r0[PPL], r1[PPL], r2[PPL], T[PPL], sacc[PPL], fxh[2], fyh[2];
      uint32_t lastp[PPL];
      bool valid[PPL];
      float l1 = 0.0f;
      uint32_t m = 0;
      const float *colv = a.color + 3 * V.pix_off;
      const float *tgt = a.target + 3 * V.pix_off;
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        fxh[h] = (float)pix_x(t, h) + 0.5f;
        fyh[h] = (float)pix_y(t, 2 * h) + 0.5f;
      }
#pragma unroll
      for (int j = 0; j < PPL; ++j) {
        const uint32_t px = pix_x(t, j), py = pix_y(t, j);
        r0[j] = r1[j] = r2[j] = 0.0f;
        T[j] = 1.0f;
        lastp[j] = 0;
        valid[j] = false;
        if (px < V.W && py < V.H) {
          const size_t lp = (size_t)py * V.W + px;
          lastp[j] = a.last[V.pix_off + lp];
          T[j] = 1.0f - a.alpha[V.pix_off + lp];
          const float d0 = colv[lp] - tgt[lp];
          const float d1 = colv[P + lp] - tgt[P + lp];
          const float d2 = colv[2 * P + lp] - tgt[2 * P + lp];
          l1 = l1 + ((fabsf(d0) + fabsf(d1)) + fabsf(d2));
          if (a.residual) {
            const float *rr = a.residual + 3 * V.pix_off;
            r0[j] = rr[lp];
            r1[j] = rr[P + lp];
            r2[j] = rr[2 * P + lp];
          } else {
            r0[j] = a.cfg.w_l2 * d0 + a.cfg.w_l1 * sign_ref(d0);
            r1[j] = a.cfg.w_l2 * d1 + a.cfg.w_l1 * sign_ref(d1);
            r2[j] = a.cfg.w_l2 * d2 + a.cfg.w_l1 * sign_ref(d2);
          }
          valid[j] = (lastp[j] != NO_LAST) && lastp[j] >= start;
          if (valid[j]) m = max(m, lastp[j] + 1u);
        }
        sacc[j] = fmaf(r2[j], a.cfg.bg[2], fmaf(r1[j], a.cfg.bg[1], r0[j] * a.cfg.bg[0]));
      }
      const float l1t = wave_total63(l1);
      if (lane == 63 && a.l1_partial) a.l1_partial[tg] = l1t;
      // entries from tend on are reached by no pixel
      uint32_t tend = wave_max_u32(m);
      if (tend < start) tend = start;
      const QuadBounds qb = quad_bounds(t, V);
      for (uint32_t hi = tend; hi > start;) {
        const uint32_t lo = (hi - start > (uint32_t)B) ? hi - B : start;
        const uint32_t cnt = hi - lo;
        const uint32_t kept = load_batch_bwd<B>(a, vbase, lo, cnt, qb, s_q0, s_q1, s_q2, s_k, s_p, t.s, t.tx, t.ty);
        __syncthreads();
        for (int e = (int)kept - 1; e >= 0; --e) {
          // same word on every lane: the quadrant tests below are uniform
          const uint32_t ke = __builtin_amdgcn_readfirstlane(s_k[e]);
          const uint32_t k = ke & 0xFFFFu;
          const uint32_t i = lo + k;
          const float4 x0 = s_q0[e];
          const float4 x1 = s_q1[e];
          const float4 x2 = s_q2[e];
          // the nine sums start at a zero the compiler cannot see through:
          // with a literal 0 it rematerialises the zeros on every skip path
          float v0 = opaque_zero(), v1 = v0, v2 = v0, v3 = v0, v4 = v0, v5 = v0, v6 = v0, v7 = v0, v8 = v0;
          bool hit = false;
          {
#pragma unroll
            for (int j = 0; j < PPL; ++j) {
              if (!((ke >> (16 + j)) & 1u)) continue;
              // one divergent branch per slot: the alpha is computed on every
              // lane (out-of-range lanes' values are discarded)
              const float dx = fxh[j & 1] - x0.x;
              const float dy = fyh[j >> 1] - x0.y;
              const float sg = sv_power(dx, dy, x0.z, x0.w, x1.x);
              const float gauss = sv_exp_alpha(-fminf(sg, 80.0f));  // as frag_alpha
              const float op = x1.y;
              const float as = fminf(op * gauss, SV_MAX_ALPHA);
              const bool take = (valid[j] & (i <= lastp[j])) & !(sg < 0.0f) & !(as < SV_MIN_ALPHA);
              if (!take) continue;
              hit = true;
              const float oma = fmaxf(1.0f - as, 0.0001f);
              const float Tb = T[j] * __builtin_amdgcn_rcpf(oma);
              // the gradient terms in fused multiply-adds (they feed only
              // sums; the alpha and T chain above stays the forward's)
              const float contrib = as * Tb;
              v0 = fmaf(r0[j], contrib, v0);
              v1 = fmaf(r1[j], contrib, v1);
              v2 = fmaf(r2[j], contrib, v2);
              const float fg = fmaf(r2[j], x2.x, fmaf(r1[j], x1.w, r0[j] * x1.z));
              const float dL_dalpha = (fg - sacc[j]) * Tb;
              v3 = fmaf(dL_dalpha * as, 1.0f - op, v3);
              // second and first moments of -dL/dpower over the pixel
              // offsets; the conic gradient is -(1/2 m_xx, m_xy, 1/2 m_yy) and
              // the mean gradient (a m_x + b m_y, b m_x + c m_y), applied per
              // (view, Gaussian) when the records are gathered
              const float nds = dL_dalpha * (op * gauss);  // -dL/dpower
              const float ux = nds * dx, uy = nds * dy;
              v4 = fmaf(ux, dx, v4);
              v5 = fmaf(ux, dy, v5);
              v6 = fmaf(uy, dy, v6);
              v7 = v7 + ux;
              v8 = v8 + uy;
              sacc[j] = fmaf(as, fg, (1.0f - as) * sacc[j]);
              T[j] = Tb;
            }
          }
          // an entry some pixel took: its record at its emit position, the
          // sums over the tile's 256 pixels (components 0..7 end in lanes
          // 8c, component 8 in lane 63; one store of nine lanes), and its
          // flag. No other entry writes either: the record gather reads a
          // record only where the flag is set.
          if (__any(hit)) {
            const uint32_t p = s_p[e];
            const float s07 = wave_scatter8(v0, v1, v2, v3, v4, v5, v6, v7);
            const float s8 = wave_total63(v8);
            if (((lane & 7) == 0 || lane == 63) && (uint64_t)p < a.n_isect) {
              a.isect_rec[(size_t)p * REC9 + (lane == 63 ? 8u : (uint32_t)lane >> 3)] = lane == 63 ? s8 : s07;
              if (lane == 0) a.rflag[p] = 1u;
            }
          }
        }
        __syncthreads();
        hi = lo;
      }
    } else if (lane == 0 && a.l1_partial) {
      a.l1_partial[tg] = 0.0f;
    }
    if (!next_tile(a, tg)) break;
  }
}

// ---------------------------------------------------------------------
// backward_tiles_local (backward-tiles-local.wgsl:185-382): the reference
// wavefront backward, one 256-thread workgroup per 16x16 tile as in the
// shader. 64-entry batches; per (pixel, splat) the 54 values (logit, conic
// abc, mean xy, 48 SH) are converted with i32(v * 65536) (truncating,
// saturating) and added with LDS integer atomics; the i32 sums are
// written per intersection. flags: WF_REFERENCE_ORDER walks batches
// front-to-back (:262), else back-to-front (order-correct); WF_SKIP_098
// keeps the alpha > 0.98 skip (:329-331).
// ---------------------------------------------------------------------
constexpr int RT = 256;
constexpr int WF_B = 64;
constexpr int WF_S = 54;

__global__ __launch_bounds__(RT) void k_backward_fixed(RasterArgs ain, const float *__restrict__ P,
                                                       size_t cap, uint32_t flags,
                                                       int32_t *__restrict__ fixed,
                                                       uint32_t *__restrict__ inv) {
  RasterArgs a = ain;
  if (!resolve_meta(a)) return;
  __shared__ int32_t s_acc[WF_B * WF_S];
  __shared__ float4 s_x0[WF_B];
  __shared__ float4 s_x1[WF_B];
  __shared__ float s_x2[WF_B];
  __shared__ float s_basis[WF_B][16];
  const uint32_t tg = blockIdx.x;
  if (tg >= a.n_tiles_total) return;
  const uint32_t s = a.tile_view[tg];
  const DevView &V = a.views[s];
  const uint32_t lt = tg - V.tile_base;
  const uint32_t tx = lt % V.tiles_x, ty = lt / V.tiles_x;
  const uint32_t px = tx * 16u + (threadIdx.x & 15u), py = ty * 16u + (threadIdx.x >> 4);
  const bool inside = px < V.W && py < V.H;
  const size_t Pn = (size_t)V.W * V.H;
  const size_t lp = (size_t)py * V.W + px;
  const uint32_t start = a.ranges[tg], end = a.ranges[tg + 1];
  if (start >= end) return;
  const size_t vbase = (size_t)s * a.N;
  float res0 = 0.f, res1 = 0.f, res2 = 0.f, T = 1.0f;
  uint32_t bin_final = 0;
  bool valid = false;
  if (inside) {
    bin_final = a.last[V.pix_off + lp];
    T = 1.0f - a.alpha[V.pix_off + lp];
    const float *r = a.residual + 3 * V.pix_off;
    res0 = r[lp];
    res1 = r[Pn + lp];
    res2 = r[2 * Pn + lp];
    valid = bin_final != NO_LAST && bin_final >= start;
  }
  float acc0 = a.cfg.bg[0], acc1 = a.cfg.bg[1], acc2 = a.cfg.bg[2];
  const float fx = (float)px + 0.5f, fy = (float)py + 0.5f;
  const uint32_t count = end - start;
  const uint32_t nb = (count + WF_B - 1) / WF_B;
  const bool ref_order = (flags & WF_REFERENCE_ORDER) != 0;
  const bool skip098 = (flags & WF_SKIP_098) != 0;
  const uint32_t deg = a.cfg.sh_deg;
  for (uint32_t bi = 0; bi < nb; ++bi) {
    const uint32_t batch = ref_order ? bi : nb - 1u - bi;
    const uint32_t bs = start + batch * WF_B;
    const uint32_t be = min(bs + WF_B, end);
    const uint32_t bsz = be - bs;
    for (int z = threadIdx.x; z < WF_B * WF_S; z += RT) s_acc[z] = 0;
    if (threadIdx.x < bsz) {
      const uint32_t g = min(a.vals[bs + threadIdx.x], a.N - 1u);
      const float4 *q = a.rec + (vbase + g) * 4;
      s_x0[threadIdx.x] = q[0];
      s_x1[threadIdx.x] = q[1];
      s_x2[threadIdx.x] = q[2].x;
      // sorted position of this (Gaussian, tile) pair, filed under its emit
      // position: the reduce walks each Gaussian's pairs in tile order
      inv[emit_pos(q[3], a.bbase[(size_t)s * a.nb + (g >> 8)], tx, ty)] = bs + threadIdx.x;
      float wp[3];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const float v = P[(size_t)(P_POS + k) * cap + g];
        wp[k] = a.cfg.render_f16 ? f16_round(v) : v;
      }
      float pc[4];
      m4_mul_pos(V.view, wp[0], wp[1], wp[2], pc);
      const float nx = -pc[0], ny = -pc[1], nz = -pc[2];
      const float len = sqrtf(dot3(nx, ny, nz, nx, ny, nz));
      float b[16];
      sh_basis(nx / len, ny / len, nz / len, deg, b);
#pragma unroll
      for (int k = 0; k < 16; ++k) s_basis[threadIdx.x][k] = b[k];
    }
    __syncthreads();
    if (valid) {
      const uint32_t pe = min(be, bin_final + 1u);
      for (uint32_t i1 = pe; i1 > bs; --i1) {
        const uint32_t i = i1 - 1u;
        const uint32_t k = i - bs;
        const float4 x0 = s_x0[k];
        const float4 x1 = s_x1[k];
        const float dx = fx - x0.x, dy = fy - x0.y;
        const float sg = sv_power(dx, dy, x0.z, x0.w, x1.x);
        if (sg < 0.0f) continue;
        const float gauss = sv_exp_alpha(-fminf(sg, 80.0f));
        const float op = x1.y;
        const float as = fminf(op * gauss, SV_MAX_ALPHA);
        if (as < SV_MIN_ALPHA) continue;
        if (skip098 && as > 0.98f) continue;
        const float oma = fmaxf(1.0f - as, 0.0001f);
        const float Tb = T / oma;
        const float contrib = as * Tb;
        const float rs0 = res0 * contrib, rs1 = res1 * contrib, rs2 = res2 * contrib;
        int32_t *acc = &s_acc[k * WF_S];
#pragma unroll
        for (int q = 0; q < 16; ++q) {
          const float bq = s_basis[k][q];
          atomicAdd(&acc[6 + 3 * q + 0], f2i_sat((bq * rs0) * 65536.0f));
          atomicAdd(&acc[6 + 3 * q + 1], f2i_sat((bq * rs1) * 65536.0f));
          atomicAdd(&acc[6 + 3 * q + 2], f2i_sat((bq * rs2) * 65536.0f));
        }
        const float cr = x1.z, cg = x1.w, cb = s_x2[k];
        const float fg = (res0 * cr + res1 * cg) + res2 * cb;
        const float bgt = (res0 * acc0 + res1 * acc1) + res2 * acc2;
        const float dL_dalpha = (fg - bgt) * Tb;
        const float dlogit = (dL_dalpha * as) * (1.0f - op);
        const float dL_ds = (dL_dalpha * op) * (-gauss);
        atomicAdd(&acc[0], f2i_sat(dlogit * 65536.0f));
        atomicAdd(&acc[1], f2i_sat((((dL_ds * 0.5f) * dx) * dx) * 65536.0f));
        atomicAdd(&acc[2], f2i_sat(((dL_ds * dx) * dy) * 65536.0f));
        atomicAdd(&acc[3], f2i_sat((((dL_ds * 0.5f) * dy) * dy) * 65536.0f));
        atomicAdd(&acc[4], f2i_sat((-(dL_ds * (x0.z * dx + x0.w * dy))) * 65536.0f));
        atomicAdd(&acc[5], f2i_sat((-(dL_ds * (x0.w * dx + x1.x * dy))) * 65536.0f));
        acc0 = as * cr + (1.0f - as) * acc0;
        acc1 = as * cg + (1.0f - as) * acc1;
        acc2 = as * cb + (1.0f - as) * acc2;
        T = Tb;
      }
    }
    __syncthreads();
    for (uint32_t z = threadIdx.x; z < bsz * WF_S; z += RT) fixed[(size_t)bs * WF_S + z] = s_acc[z];
    __syncthreads();
  }
}

uint32_t raster_resident_blocks(bool backward, uint32_t batch) {
  int per_cu = 0, cus = 0, dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  const void *k = backward ? (batch <= 64 ? (const void *)k_backward<64>
                                          : batch <= 128 ? (const void *)k_backward<128> : (const void *)k_backward<256>)
                           : (batch <= 64 ? (const void *)k_forward<64>
                                          : batch <= 128 ? (const void *)k_forward<128> : (const void *)k_forward<256>);
  if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, WT, 0);
  if (e != hipSuccess || per_cu <= 0 || cus <= 0) {
    (void)hipGetLastError();
    return 0;
  }
  return (uint32_t)per_cu * (uint32_t)cus;
}

void launch_forward(hipStream_t st, const RasterArgs &a, uint32_t batch, uint32_t grid, uint32_t raster_px) {
  if (a.n_tiles_total == 0 || grid == 0) return;
  if (raster_px == 8) {
    const uint32_t g8 = 4u * a.n_tiles_total;
    if (batch <= 64) hipLaunchKernelGGL(k_forward8<64>, dim3(g8), dim3(WT), 0, st, a);
    else if (batch <= 128) hipLaunchKernelGGL(k_forward8<128>, dim3(g8), dim3(WT), 0, st, a);
    else hipLaunchKernelGGL(k_forward8<256>, dim3(g8), dim3(WT), 0, st, a);
  } else if (raster_px == 32) {
    if (a.n_blocks_total == 0) return;
    if (batch <= 64) hipLaunchKernelGGL(k_forward32<64>, dim3(a.n_blocks_total), dim3(4 * WT), 0, st, a);
    else if (batch <= 128) hipLaunchKernelGGL(k_forward32<128>, dim3(a.n_blocks_total), dim3(4 * WT), 0, st, a);
    else hipLaunchKernelGGL(k_forward32<256>, dim3(a.n_blocks_total), dim3(4 * WT), 0, st, a);
  } else {
    if (batch <= 64) hipLaunchKernelGGL(k_forward<64>, dim3(grid), dim3(WT), 0, st, a);
    else if (batch <= 128) hipLaunchKernelGGL(k_forward<128>, dim3(grid), dim3(WT), 0, st, a);
    else hipLaunchKernelGGL(k_forward<256>, dim3(grid), dim3(WT), 0, st, a);
  }
}

void launch_backward(hipStream_t st, const RasterArgs &a, uint32_t batch, uint32_t grid) {
  if (a.n_tiles_total == 0 || grid == 0) return;
  if (batch <= 64) hipLaunchKernelGGL(k_backward<64>, dim3(grid), dim3(WT), 0, st, a);
  else if (batch <= 128) hipLaunchKernelGGL(k_backward<128>, dim3(grid), dim3(WT), 0, st, a);
  else hipLaunchKernelGGL(k_backward<256>, dim3(grid), dim3(WT), 0, st, a);
}

void launch_backward_fixed(hipStream_t st, const RasterArgs &a, const float *P, size_t cap,
                           uint32_t flags, int32_t *fixed, uint32_t *inv) {
  if (a.n_tiles_total == 0) return;
  hipLaunchKernelGGL(k_backward_fixed, dim3(a.n_tiles_total), dim3(RT), 0, st, a, P, cap, flags,
                     fixed, inv);
}

}  // namespace sv
