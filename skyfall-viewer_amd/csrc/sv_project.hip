// sv_project.hip -- per-Gaussian projection, device prefix sum, key emit,
// tile ranges and the loss residual (gfx950).
#include "sv_internal.hpp"

#include <algorithm>

namespace sv {

// ---------------------------------------------------------------------
// Projection: project_gaussians_for_tiling (training-tiles.wgsl:39-180)
// fused with the per-view SH colour (training-forward-tiles.wgsl:65-105,
// evaluated once per (Gaussian, view) instead of once per pixel; the
// inputs of that evaluation depend only on the Gaussian and the view, so
// the value is the same) and the opacity sigmoid (:188).
//
// One thread per Gaussian, looping over the views of the step, so the 59
// parameters are read once per step. Each (view, Gaussian) record is 64
// bytes (one memory burst) of four float4:
//   q0 = (mean.x, mean.y, conic.a, conic.b)
//   q1 = (conic.c, opacity, r, g)
//   q2 = (b, power_max, ext_x, ext_y)
//   q3 = (emit base within the projection block, rect x0|y0<<16,
//         rect x1|y1<<16, tight-rectangle insets)
// and aux = (depth, radius). power_max = ln(255 opacity) + 0.02: a pair with
// a larger power has opacity * exp(-power) < 1/255 even with the ulp error
// of exp, so it would be skipped. ext_x, ext_y bound the region power <=
// power_max: max |dx| = sqrt(2 power_max (Q^-1)_xx) for the conic Q, with a
// 1% + 0.01 px margin for rounding. The rasterizers' batch loads use both
// to skip whole quadrants of a tile without evaluating their pixels; every
// skipped pair is one the reference skips, so no result changes.
// Exclusive prefix of two u32 values over the 256 threads of the block,
// and their block totals (wave scans, then the four wave totals).
__device__ __forceinline__ void block_scan2(uint32_t a, uint32_t b, uint32_t &pa, uint32_t &pb, uint32_t &ta,
                                            uint32_t &tb, uint32_t *s_w) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t x = a, y = b;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t xo = (uint32_t)__shfl_up((int)x, o, 64), yo = (uint32_t)__shfl_up((int)y, o, 64);
    if (lane >= o) {
      x += xo;
      y += yo;
    }
  }
  if (lane == 63) {
    s_w[wave] = x;
    s_w[4 + wave] = y;
  }
  __syncthreads();
  uint32_t ba = 0, bb = 0;
  ta = 0;
  tb = 0;
#pragma unroll
  for (int w = 0; w < 4; ++w) {
    if (w < wave) {
      ba += s_w[w];
      bb += s_w[4 + w];
    }
    ta += s_w[w];
    tb += s_w[4 + w];
  }
  pa = ba + x - a;
  pb = bb + y - b;
  __syncthreads();
}

// ---------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_project(const float *__restrict__ P, size_t cap, uint32_t N,
                                                 const DevView *__restrict__ views, int nv,
                                                 DevConfig cfg, float4 *__restrict__ rec,
                                                 float2 *__restrict__ aux,
                                                 uint32_t *__restrict__ counts,
                                                 uint32_t *__restrict__ max_dq,
                                                 uint32_t *__restrict__ scounts,
                                                 uint2 *__restrict__ rects,
                                                 uint32_t *__restrict__ tins,
                                                 uint32_t *__restrict__ offs_loc,
                                                 uint32_t *__restrict__ soffs_loc,
                                                 uint32_t *__restrict__ btot,
                                                 uint32_t *__restrict__ sbtot) {
  // Records of one view for the block's 256 Gaussians are staged in LDS
  // and stored as whole contiguous lines (each store instruction of a wave
  // writes 1 KB), instead of 64-byte pieces 64 bytes apart per lane.
  __shared__ float4 s_rec[256 * 4];
  __shared__ uint32_t s_scan[8];
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = g < N;
  const uint32_t gl = live ? g : N - 1u;  // lanes past N compute on a valid Gaussian, store nothing
  uint32_t local_max_dq = 0;
  {
    float prm[11];
#pragma unroll
    for (int k = 0; k < 11; ++k) {
      const float v = P[(size_t)k * cap + gl];
      prm[k] = cfg.render_f16 ? f16_round(v) : v;
    }
    const float wx = prm[P_POS], wy = prm[P_POS + 1], wz = prm[P_POS + 2];
    float sigma[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) sigma[k] = sv_expf(clampf(prm[P_SCALE + k], -10.0f, 10.0f));
    const float qd = ((prm[P_QUAT] * prm[P_QUAT] + prm[P_QUAT + 1] * prm[P_QUAT + 1]) +
                      prm[P_QUAT + 2] * prm[P_QUAT + 2]) +
                     prm[P_QUAT + 3] * prm[P_QUAT + 3];
    const float ql = sqrtf(qd);
    const M3 R = quat_to_mat3(prm[P_QUAT] / ql, prm[P_QUAT + 1] / ql, prm[P_QUAT + 2] / ql,
                              prm[P_QUAT + 3] / ql);
    M3 S;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) S.m[i][j] = (i == j) ? sigma[i] : 0.0f;
    // C3D = transpose(S * R) * S * R  (training-tiles.wgsl:101)
    const M3 C3D = mul(mul(transpose(mul(S, R)), S), R);
    const float logit = prm[P_LOGIT];
    const float opac = sv_sigmoid(logit);
    // the power bound ln(255 opacity) + 0.02 depends on the Gaussian only
    const float pmax = opac > 0.0f ? logf(255.0f * opac) + 0.02f : -1.0f;
    // SH coefficients once per Gaussian (re-reading them per view made the
    // kernel fetch 12x its parameters from memory)
    float shc[48];
#pragma unroll
    for (int k = 0; k < 48; ++k) {
      const float v = P[(size_t)(P_SH + k) * cap + gl];
      shc[k] = cfg.render_f16 ? f16_round(v) : v;
    }

    for (int s = 0; s < nv; ++s) {
      const DevView &V = views[s];
      const size_t vg = (size_t)s * N + g;
      float pc[4], clip[4];
      m4_mul_pos(V.view, wx, wy, wz, pc);
      m4_mul_v4(V.proj, pc, clip);
      const float w = clip[3];
      const float cx = clip[0] / w, cy = clip[1] / w;
      bool alive = !(fabsf(cx) > 1.2f || fabsf(cy) > 1.2f || pc[2] <= 0.0f);
      const float depth = pc[2];
      if (depth < cfg.near_plane || depth > cfg.far_plane) alive = false;
      float4 r0 = make_float4(0.f, 0.f, 0.f, 0.f), r1 = r0, r2 = r0, r3 = r0;
      float2 ax = make_float2(0.f, 0.f);
      uint32_t count = 0, scount = 0;
      uint2 rect = make_uint2(0u, 0u);
      uint32_t tin = TIGHT_EMPTY;
      if (alive) {
        const float tx = pc[0], ty = pc[1], tz = pc[2];
        const float fx = V.focal[0], fy = V.focal[1];
        M3 J;
        J.m[0][0] = fx / tz;
        J.m[0][1] = 0.0f;
        J.m[0][2] = -(fx * tx) / (tz * tz);
        J.m[1][0] = 0.0f;
        J.m[1][1] = fy / tz;
        J.m[1][2] = -(fy * ty) / (tz * tz);
        J.m[2][0] = 0.0f;
        J.m[2][1] = 0.0f;
        J.m[2][2] = 0.0f;
        M3 Rwc;
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
          for (int r = 0; r < 3; ++r) Rwc.m[c][r] = V.view[c * 4 + r];
        const M3 T = mul(J, Rwc);
        const M3 C2D = mul(mul(T, C3D), transpose(T));
        const float c00 = C2D.m[0][0] + cfg.eps2d;
        const float c11 = C2D.m[1][1] + cfg.eps2d;
        const float c01 = C2D.m[0][1];
        const float det = c00 * c11 - c01 * c01;
        if (det <= 0.0f) alive = false;
        const float mid = 0.5f * (c00 + c11);
        const float disc = sqrtf(fmaxf(mid * mid - det, 0.0f));
        const float l1 = mid + disc, l2 = mid - disc;
        const float radius = ceilf(3.0f * sqrtf(fmaxf(l1, l2)));
        if (radius > cfg.radius_clip || radius < 1.0f) alive = false;
        if (alive) {
          const float det_inv = 1.0f / det;
          const float aa = c11 * det_inv;
          const float bb = -c01 * det_inv;
          const float cc = c00 * det_inv;
          const float mx = (0.5f * (cx + 1.0f)) * (float)V.W;
          const float my = (0.5f * (cy + 1.0f)) * (float)V.H;
          // tile rect, training-tiles.wgsl:161-171; the tile size is 16, so
          // the divisions are multiplications by 2^-4 (exact)
          const float its = 1.0f / (float)cfg.tile_size;
          const float tlx = mx * its, tly = my * its, trad = radius * its;
          uint32_t x0 = f2u_sat(floorf(tlx - trad)), y0 = f2u_sat(floorf(tly - trad));
          uint32_t x1 = f2u_sat(ceilf(tlx + trad)), y1 = f2u_sat(ceilf(tly + trad));
          x0 = min(x0, V.tiles_x);
          y0 = min(y0, V.tiles_y);
          x1 = min(x1, V.tiles_x);
          y1 = min(y1, V.tiles_y);
          const int32_t nx = (int32_t)x1 - (int32_t)x0, ny = (int32_t)y1 - (int32_t)y0;
          count = (nx <= 0 || ny <= 0) ? 0u : (uint32_t)(nx * ny);
          // supertiles of BIN_S x BIN_S tiles the rectangle touches (binned build)
          if (count) scount = ((x1 - 1u) / BIN_S - x0 / BIN_S + 1u) * ((y1 - 1u) / BIN_S - y0 / BIN_S + 1u);
          rect = make_uint2(x0 | (y0 << 16), x1 | (y1 << 16));
          // viewDir = normalize(-(view * pos).xyz), training-forward-tiles.wgsl:185
          const float dxv = -pc[0], dyv = -pc[1], dzv = -pc[2];
          const float len = sqrtf(dot3(dxv, dyv, dzv, dxv, dyv, dzv));
          float b[16];
          sh_basis(dxv / len, dyv / len, dzv / len, cfg.sh_deg, b);
          float rgb[3];
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            float coef[16];
#pragma unroll
            for (int k = 0; k < 16; ++k) coef[k] = shc[3 * k + c];
            rgb[c] = sh_channel(b, coef, 1, cfg.sh_deg);
          }
          const uint32_t dq = f2u_sat(clampf(depth, 0.0f, 65535.0f)) & 0xffffu;
          if (count > 0u && live) local_max_dq = max(local_max_dq, dq);
          const float qdet = aa * cc - bb * bb;
          float ex = -1.0f, ey = -1.0f;
          if (pmax > 0.0f && qdet > 0.0f) {
            ex = sqrtf(2.0f * pmax * cc / qdet) * 1.01f + 0.01f;
            ey = sqrtf(2.0f * pmax * aa / qdet) * 1.01f + 0.01f;
          }
          r0 = make_float4(mx, my, aa, bb);
          r1 = make_float4(cc, opac, rgb[0], rgb[1]);
          r2 = make_float4(rgb[2], pmax, ex, ey);
          tin = count ? tight_insets(mx, ex, my, ey, x0, y0, x1, y1) : TIGHT_EMPTY;
          r3 = make_float4(0.0f, __uint_as_float(x0 | (y0 << 16)), __uint_as_float(x1 | (y1 << 16)),
                           __uint_as_float(tin));
          ax = make_float2(depth, radius);
        }
      }
      // first level of the scans over (view, Gaussian): exclusive prefix of
      // the tile and supertile counts within the block; the block totals
      // are scanned afterwards, and a position is block base + prefix
      uint32_t pa, pb, ta, tb;
      block_scan2(live ? count : 0u, live ? scount : 0u, pa, pb, ta, tb, s_scan);
      r3.x = __uint_as_float(pa);  // emit base within the block
      s_rec[threadIdx.x * 4 + 0] = r0;
      s_rec[threadIdx.x * 4 + 1] = r1;
      s_rec[threadIdx.x * 4 + 2] = r2;
      s_rec[threadIdx.x * 4 + 3] = r3;
      if (live) {
        aux[vg] = ax;
        counts[vg] = count;
        if (scounts) scounts[vg] = scount;
        rects[vg] = rect;
        tins[vg] = tin;
        offs_loc[vg] = pa;
        if (soffs_loc) soffs_loc[vg] = pb;
      }
      if (threadIdx.x == 0) {
        btot[(size_t)s * gridDim.x + blockIdx.x] = ta;
        if (sbtot) sbtot[(size_t)s * gridDim.x + blockIdx.x] = tb;
      }
      __syncthreads();
      // piece j of the block's 1024 float4 belongs to Gaussian
      // blockIdx.x * 256 + j / 4 of this view
      const uint32_t g0 = blockIdx.x * blockDim.x;
      float4 *dst = rec + ((size_t)s * N + g0) * 4;
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const uint32_t j = (uint32_t)u * 256u + threadIdx.x;
        if (g0 + j / 4u < N) dst[j] = s_rec[j];
      }
      __syncthreads();
    }
  }
  // block max of the depth quantum, then one atomic per block
  uint32_t m = local_max_dq;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = max(m, (uint32_t)__shfl_xor((int)m, o, 64));
  __shared__ uint32_t s_m[4];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0) s_m[wave] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint32_t mm = max(max(s_m[0], s_m[1]), max(s_m[2], s_m[3]));
    if (mm) atomicMax(max_dq, mm);
  }
}

void launch_project(hipStream_t st, const float *P, size_t cap, uint32_t N, const DevView *views,
                    int nv, const DevConfig &cfg, float4 *rec, float2 *aux, uint32_t *counts,
                    uint32_t *max_dq, uint32_t *scounts, uint2 *rects, uint32_t *tins, uint32_t *offs_loc,
                    uint32_t *soffs_loc, uint32_t *btot, uint32_t *sbtot) {
  if (N == 0) return;
  const uint32_t blocks = project_blocks(N);
  hipLaunchKernelGGL(k_project, dim3(blocks), dim3(256), 0, st, P, cap, N, views, nv, cfg, rec, aux,
                     counts, max_dq, scounts, rects, tins, offs_loc, soffs_loc, btot, sbtot);
}

// ---------------------------------------------------------------------
// Exclusive prefix sum of u32 counts (replaces the host cumsum of
// gaussian-renderer.ts:1814-1828). Reduce-then-scan, 4096 per block.
// ---------------------------------------------------------------------
constexpr int SCAN_T = 256;
constexpr int SCAN_ITEMS = 16;
constexpr int SCAN_TILE = SCAN_T * SCAN_ITEMS;

__device__ __forceinline__ uint64_t block_exclusive_scan_u64(uint64_t v, uint64_t *s_w,
                                                             uint64_t *block_total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint64_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint64_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) s_w[wave] = x;
  __syncthreads();
  uint64_t wbase = 0, tot = 0;
  for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
    if (w < wave) wbase += s_w[w];
    tot += s_w[w];
  }
  if (block_total) *block_total = tot;
  return wbase + x - v;
}

__global__ __launch_bounds__(SCAN_T) void k_scan_reduce(const uint32_t *__restrict__ in, size_t M,
                                                        uint64_t *__restrict__ bsum) {
  const size_t base = (size_t)blockIdx.x * SCAN_TILE;
  uint64_t s = 0;
#pragma unroll
  for (int j = 0; j < SCAN_ITEMS; ++j) {
    const size_t i = base + (size_t)j * SCAN_T + threadIdx.x;
    if (i < M) s += in[i];
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  __shared__ uint64_t s_w[SCAN_T / 64];
  if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t t = 0;
    for (int w = 0; w < SCAN_T / 64; ++w) t += s_w[w];
    bsum[blockIdx.x] = t;
  }
}

// Single workgroup: exclusive scan of the block sums in place; total out.
__global__ __launch_bounds__(1024) void k_scan_top(uint64_t *__restrict__ bsum, size_t nb,
                                                   uint64_t *__restrict__ total) {
  __shared__ uint64_t s_w[16];
  uint64_t carry = 0;
  for (size_t base = 0; base < nb; base += 1024) {
    const size_t i = base + threadIdx.x;
    const uint64_t v = (i < nb) ? bsum[i] : 0;
    uint64_t tot;
    const uint64_t ex = block_exclusive_scan_u64(v, s_w, &tot);
    if (i < nb) bsum[i] = carry + ex;
    carry += tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = carry;
}

__global__ __launch_bounds__(SCAN_T) void k_scan_apply(const uint32_t *__restrict__ in, size_t M,
                                                       const uint64_t *__restrict__ bsum,
                                                       uint32_t *__restrict__ out) {
  const size_t base = (size_t)blockIdx.x * SCAN_TILE + (size_t)threadIdx.x * SCAN_ITEMS;
  uint32_t v[SCAN_ITEMS];
  uint64_t s = 0;
#pragma unroll
  for (int j = 0; j < SCAN_ITEMS; ++j) {
    const size_t i = base + j;
    v[j] = (i < M) ? in[i] : 0u;
    s += v[j];
  }
  __shared__ uint64_t s_w[SCAN_T / 64];
  uint64_t run = block_exclusive_scan_u64(s, s_w, nullptr) + bsum[blockIdx.x];
#pragma unroll
  for (int j = 0; j < SCAN_ITEMS; ++j) {
    const size_t i = base + j;
    if (i < M) out[i] = (uint32_t)run;
    run += v[j];
  }
}

size_t scan_tmp_words(size_t M) { return (M + SCAN_TILE - 1) / SCAN_TILE + 1; }

void launch_scan(hipStream_t st, const uint32_t *counts, uint32_t *offs, size_t M, uint64_t *tmp,
                 uint64_t *total) {
  const size_t nb = (M + SCAN_TILE - 1) / SCAN_TILE;
  if (nb == 0) {
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, st, tmp, (size_t)0, total);
    return;
  }
  hipLaunchKernelGGL(k_scan_reduce, dim3((uint32_t)nb), dim3(SCAN_T), 0, st, counts, M, tmp);
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, st, tmp, nb, total);
  hipLaunchKernelGGL(k_scan_apply, dim3((uint32_t)nb), dim3(SCAN_T), 0, st, counts, M, tmp, offs);
}

// ---------------------------------------------------------------------
// emit_intersections (training-tiles.wgsl:201-247); also stores the
// Gaussian's first intersection in its record. Key = (global tile <<
// dbits) | depth_q with depth_q = u32(clamp(depth, 0, 65535)) & 0xffff
// (encode_isect_key, :182-199). dbits is the bit width of the largest
// depth_q of the step, so the key order equals the reference key order
// (camera, tile, depth_q) restricted to the keys present. Value =
// Gaussian index. Order within a Gaussian: ty-major, then tx (:238-246).
// ---------------------------------------------------------------------
// Key layout of the step from the scan total and the largest depth
// quantum: dbits = bit length of max depth_q, key bits = tile bits + dbits,
// 8-bit sort passes; overflow 1 = more pairs than the buffers hold (the
// host grows them and runs the step again), 2 = key wider than the key
// type. With an overflow nothing after this point writes any state.
__global__ void k_keyinfo(StepMeta *m, uint32_t tile_bits, uint32_t key_type_bits, uint64_t cap) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  const uint64_t n = m->n_isect;
  const uint32_t dbits = 32u - (uint32_t)__clz((int)m->max_dq);
  const uint32_t kb = tile_bits + dbits;
  uint32_t of = 0;
  if (n > cap || n >= 0xffffffffull) of = 1;
  else if (kb > key_type_bits) of = 2;
  const uint32_t passes = (of || n <= 1) ? 0u : (kb + 7u) / 8u;
  m->dbits = dbits;
  m->key_bits = kb;
  m->passes = passes;
  m->sel = passes & 1u;
  m->overflow = of;
}

void launch_keyinfo(hipStream_t st, StepMeta *m, uint32_t tile_bits, uint32_t key_type_bits, uint64_t cap) {
  hipLaunchKernelGGL(k_keyinfo, dim3(1), dim3(64), 0, st, m, tile_bits, key_type_bits, cap);
}

// One workgroup emits the pairs of 256 consecutive (view, Gaussian) items.
// Their output ranges are adjacent (exclusive scan), so the workgroup
// writes one contiguous span: thread t takes span positions t, t + 256, ...
// and finds the owning item by binary search over the items' first
// positions in LDS. Stores are coalesced whole lines instead of one
// sequential run per thread.
template <typename K>
__global__ __launch_bounds__(256) void k_emit(float4 *__restrict__ rec, const uint2 *__restrict__ rects,
                                              const float2 *__restrict__ aux,
                                              const uint32_t *__restrict__ counts,
                                              const uint32_t *__restrict__ offs,
                                              const uint32_t *__restrict__ bbase, uint32_t nb, uint32_t N, int nv,
                                              const DevView *__restrict__ views, const StepMeta *m,
                                              K *__restrict__ keys, uint32_t *__restrict__ vals) {
  __shared__ uint32_t s_lo[256];   // first output position (0xffffffff past the items)
  __shared__ uint32_t s_rect[256]; // x0 | width << 16
  __shared__ uint32_t s_y0[256];
  __shared__ K s_kb[256];          // key of the item's tile (tx = x0, ty = y0) without depth bits
  __shared__ uint32_t s_dq[256];
  __shared__ uint32_t s_g[256];
  __shared__ uint32_t s_tx[256];   // tiles_x of the item's view
  if (m->overflow) return;
  const uint32_t dbits = m->dbits;
  const size_t M = (size_t)nv * N;
  const uint32_t t = threadIdx.x;
  const size_t vg = (size_t)blockIdx.x * 256u + t;
  uint32_t lo = 0xffffffffu, hi = 0;
  if (vg < M) {
    const uint32_t cnt = counts[vg];
    const uint32_t s = (uint32_t)(vg / N), g = (uint32_t)(vg - (size_t)s * N);
    lo = bbase[(size_t)s * nb + (g >> 8)] + offs[vg];
    hi = lo + cnt;
    if (cnt) {
      const DevView &V = views[s];
      const uint2 rc = rects[vg];
      const uint32_t a = rc.x, b = rc.y;
      const uint32_t x0 = a & 0xffffu, y0 = a >> 16, x1 = b & 0xffffu;
      s_rect[t] = x0 | ((x1 - x0) << 16);
      s_y0[t] = y0;
      s_kb[t] = (K)(V.tile_base + y0 * V.tiles_x + x0);
      s_dq[t] = f2u_sat(clampf(aux[vg].x, 0.0f, 65535.0f)) & 0xffffu;
      s_g[t] = g;
      s_tx[t] = V.tiles_x;
    }
  }
  s_lo[t] = lo;
  // span of the block: from the first item's position to the end of the last
  __shared__ uint32_t s_span[2];
  if (t == 0) s_span[0] = lo;
  const size_t last = min((size_t)blockIdx.x * 256u + 255u, M - 1);
  if (vg == last) s_span[1] = hi;
  __syncthreads();
  const uint32_t span_lo = s_span[0], span_hi = s_span[1];
  for (uint32_t k = span_lo + t; k < span_hi; k += 256u) {
    // largest j with s_lo[j] <= k: that item's range contains k
    uint32_t j = 0;
#pragma unroll
    for (uint32_t step = 128; step > 0; step >>= 1)
      if (s_lo[j + step] <= k) j += step;
    const uint32_t local = k - s_lo[j];
    const uint32_t rect = s_rect[j], w = rect >> 16;
    const uint32_t dy = local / w, dx = local - dy * w;
    const K tile = s_kb[j] + (K)(dy * s_tx[j] + dx);
    keys[k] = (tile << dbits) | (K)s_dq[j];
    vals[k] = s_g[j];
  }
}

template <typename K>
void launch_emit(hipStream_t st, float4 *rec, const uint2 *rects, const float2 *aux, const uint32_t *counts,
                 const uint32_t *offs, const uint32_t *bbase, uint32_t N, int nv, const DevView *views,
                 const StepMeta *m, K *keys, uint32_t *vals) {
  const size_t M = (size_t)nv * N;
  if (M == 0) return;
  hipLaunchKernelGGL(k_emit<K>, dim3((uint32_t)((M + 255) / 256)), dim3(256), 0, st, rec, rects, aux, counts,
                     offs, bbase, project_blocks(N), N, nv, views, m, keys, vals);
}
template void launch_emit<uint32_t>(hipStream_t, float4 *, const uint2 *, const float2 *, const uint32_t *,
                                    const uint32_t *, const uint32_t *, uint32_t, int, const DevView *,
                                    const StepMeta *, uint32_t *, uint32_t *);
template void launch_emit<uint64_t>(hipStream_t, float4 *, const uint2 *, const float2 *, const uint32_t *,
                                    const uint32_t *, const uint32_t *, uint32_t, int, const DevView *,
                                    const StepMeta *, uint64_t *, uint32_t *);

// ---------------------------------------------------------------------
// Tile ranges: ranges[t] = first sorted index whose tile is >= t, which is
// what build_tile_offsets' atomicMin plus the host reverse-min fix-up
// computes (training-tiles.wgsl:249-271, gaussian-renderer.ts:2014-2020).
// Every entry is written by exactly one thread; no atomics.
// ---------------------------------------------------------------------
// Launched for max(cap, n_tiles + 1) threads; the live count, key layout
// and sorted buffer come from the step metadata.
template <typename K>
__global__ __launch_bounds__(256) void k_ranges(const K *__restrict__ keys0, const K *__restrict__ keys1,
                                                const StepMeta *m, uint32_t n_tiles,
                                                uint32_t *__restrict__ ranges) {
  if (m->overflow) return;
  const uint64_t I = m->n_isect;
  // a fixed grid strides over the live pairs (a thread per pair of the
  // capacity would launch mostly idle blocks)
  const uint64_t i0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  if (I == 0) {
    for (uint64_t i = i0; i <= n_tiles; i += stride) ranges[i] = 0u;
    return;
  }
  const K *__restrict__ keys = m->sel ? keys1 : keys0;
  const uint32_t dbits = m->dbits;
  // bits above the key bits carry a payload (binned build)
  const K kmask = m->key_bits >= 8 * sizeof(K) ? ~(K)0 : (((K)1 << m->key_bits) - 1);
  for (uint64_t i = i0; i < I; i += stride) {
    const uint32_t t = (uint32_t)((keys[i] & kmask) >> dbits);
    const uint32_t lo = (i == 0) ? 0u : (uint32_t)((keys[i - 1] & kmask) >> dbits) + 1u;
    for (uint32_t u = lo; u <= t && u <= n_tiles; ++u) ranges[u] = (uint32_t)i;
    if (i == I - 1)
      for (uint32_t u = t + 1u; u <= n_tiles; ++u) ranges[u] = (uint32_t)I;
  }
}

__global__ void k_fill_u32(uint32_t *p, uint32_t v, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_fill_f32(float *p, float v, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
void launch_fill_u32(hipStream_t st, uint32_t *p, uint32_t v, size_t n) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_fill_u32, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, st, p, v, n);
}
void launch_fill_f32(hipStream_t st, float *p, float v, size_t n) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_fill_f32, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, st, p, v, n);
}

// A kernel, not a memset: the step is captured into a graph.
__global__ __launch_bounds__(256) void k_clear_flags(uint4 *__restrict__ f, const StepMeta *m) {
  if (m->overflow) return;
  const uint64_t words = (m->n_isect + 15) / 16;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += stride)
    f[i] = make_uint4(0u, 0u, 0u, 0u);
}
void launch_clear_flags(hipStream_t st, uint8_t *flags, uint64_t cap, const StepMeta *m) {
  const uint64_t words = flag_array_bytes(cap) / 16;
  if (words == 0) return;
  const uint32_t blocks = (uint32_t)std::min<uint64_t>((words + 255) / 256, 2048);
  hipLaunchKernelGGL(k_clear_flags, dim3(blocks), dim3(256), 0, st, (uint4 *)flags, m);
}

template <typename K>
void launch_ranges(hipStream_t st, const K *keys0, const K *keys1, uint64_t cap, uint32_t n_tiles,
                   const StepMeta *m, uint32_t *ranges) {
  const uint64_t threads = std::max<uint64_t>(cap, (uint64_t)n_tiles + 1);
  const uint32_t blocks = (uint32_t)std::min<uint64_t>((threads + 255) / 256, 4096);
  hipLaunchKernelGGL(k_ranges<K>, dim3(blocks), dim3(256), 0, st, keys0, keys1, m, n_tiles, ranges);
}
template void launch_ranges<uint32_t>(hipStream_t, const uint32_t *, const uint32_t *, uint64_t, uint32_t,
                                      const StepMeta *, uint32_t *);
template void launch_ranges<uint64_t>(hipStream_t, const uint64_t *, const uint64_t *, uint64_t, uint32_t,
                                      const StepMeta *, uint32_t *);

// ---------------------------------------------------------------------
// training_loss (training-loss.wgsl:86-126) for one view: residual =
// w_l2 (p - t) + w_l1 sign(p - t) [+ w_ssim * dssim (p - t), :33-84].
// Colour planes: c * P + y * W + x within the view's block of 3P floats.
// ---------------------------------------------------------------------
__device__ __forceinline__ float signf_ref(float x) {
  return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f);
}

// One workgroup per 16x16 pixel tile. With the SSIM term, the tile's
// 20x20 patch of both images (every pixel the tile's 5x5 windows reach,
// coordinates clamped at the image border as the shader clamps them) is
// staged in LDS once; each pixel then sums its window from there in the
// shader's order (rows, then columns, channels apart), so every sum is the
// same float as from the image itself.
constexpr int LOSS_T = 16;               // output tile side
constexpr int LOSS_R = 2;                // window radius (5 x 5)
constexpr int LOSS_P = LOSS_T + 2 * LOSS_R;

__global__ __launch_bounds__(256) void k_loss(uint32_t W, uint32_t H, DevConfig cfg,
                                              const float *__restrict__ color,
                                              const float *__restrict__ target,
                                              float *__restrict__ residual) {
  const uint32_t tx = threadIdx.x & 15u, ty = threadIdx.x >> 4;
  const uint32_t x = blockIdx.x * LOSS_T + tx;
  const uint32_t y = blockIdx.y * LOSS_T + ty;
  const bool in = x < W && y < H;
  const size_t P = (size_t)W * H;
  const size_t pix = (size_t)y * W + x;
  float gs[3] = {0.f, 0.f, 0.f};
  if (cfg.w_ssim > 0.0f) {  // the same for the whole workgroup
    __shared__ float sx[3][LOSS_P][LOSS_P + 1], sy[3][LOSS_P][LOSS_P + 1];
    for (uint32_t i = threadIdx.x; i < (uint32_t)(LOSS_P * LOSS_P); i += 256) {
      const int pr = (int)(i / LOSS_P), pc = (int)(i % LOSS_P);
      const int gy = min(max((int)(blockIdx.y * LOSS_T) - LOSS_R + pr, 0), (int)H - 1);
      const int gx = min(max((int)(blockIdx.x * LOSS_T) - LOSS_R + pc, 0), (int)W - 1);
      const size_t q = (size_t)gy * W + gx;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        sx[c][pr][pc] = color[c * P + q];
        sy[c][pr][pc] = target[c * P + q];
      }
    }
    __syncthreads();
    // pixel (x + dx, y + dy) of the window is patch entry (ty + 2 + dy, tx + 2 + dx)
    float mux[3] = {0, 0, 0}, muy[3] = {0, 0, 0};
    const float n = 25.0f;
    for (int r = 0; r < 2 * LOSS_R + 1; ++r)
      for (int k = 0; k < 2 * LOSS_R + 1; ++k) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          mux[c] = mux[c] + sx[c][ty + r][tx + k];
          muy[c] = muy[c] + sy[c][ty + r][tx + k];
        }
      }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      mux[c] = mux[c] / n;
      muy[c] = muy[c] / n;
    }
    float sx2[3] = {0, 0, 0}, sy2[3] = {0, 0, 0}, sxy[3] = {0, 0, 0};
    for (int r = 0; r < 2 * LOSS_R + 1; ++r)
      for (int k = 0; k < 2 * LOSS_R + 1; ++k) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const float a = sx[c][ty + r][tx + k] - mux[c];
          const float b = sy[c][ty + r][tx + k] - muy[c];
          sx2[c] = sx2[c] + a * a;
          sy2[c] = sy2[c] + b * b;
          sxy[c] = sxy[c] + a * b;
        }
      }
    const float C1 = 0.01f * 0.01f, C2 = 0.03f * 0.03f;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      sx2[c] = sx2[c] / n;
      sy2[c] = sy2[c] / n;
      sxy[c] = sxy[c] / n;
      const float num1 = (2.0f * mux[c]) * muy[c] + C1;
      const float num2 = 2.0f * sxy[c] + C2;
      const float den1 = (mux[c] * mux[c] + muy[c] * muy[c]) + C1;
      const float den2 = (sx2[c] + sy2[c]) + C2;
      const float ssim = (num1 * num2) / (den1 * den2);
      const float dssim = (1.0f - ssim) * 0.5f;
      gs[c] = dssim * (sx[c][ty + LOSS_R][tx + LOSS_R] - sy[c][ty + LOSS_R][tx + LOSS_R]);
    }
  }
  if (!in) return;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const float d = color[c * P + pix] - target[c * P + pix];
    float g = cfg.w_l2 * d + cfg.w_l1 * signf_ref(d);
    if (cfg.w_ssim > 0.0f) g = g + cfg.w_ssim * gs[c];
    residual[c * P + pix] = g;
  }
}

void launch_loss(hipStream_t st, const DevView &hv, const DevConfig &cfg, const float *color,
                 const float *target, float *residual) {
  const dim3 grid((hv.W + 15) / 16, (hv.H + 15) / 16);
  hipLaunchKernelGGL(k_loss, grid, dim3(256), 0, st, hv.W, hv.H, cfg, color, target, residual);
}

}  // namespace sv
