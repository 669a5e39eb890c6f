// sv_raster.hip -- tiled forward rasterizer and adjoint backward for gfx950.
//
// One 256-thread workgroup (four wave64) owns one 16x16 pixel tile of one
// view; its sorted splat list is staged through LDS in batches of B
// records. Work items are (view, global tile); with a work counter the
// grid is persistent and pulls items from a device queue.
#include "sv_internal.hpp"

namespace sv {

constexpr int RT = 256;  // threads per raster workgroup (16 x 16 pixels)

__device__ __forceinline__ bool next_tile(const RasterArgs &a, uint32_t &tg, uint32_t *s_tg) {
  if (a.work_counter == nullptr) return false;
  __syncthreads();
  if (threadIdx.x == 0) *s_tg = atomicAdd(a.work_counter, 1u);
  __syncthreads();
  tg = *s_tg;
  return tg < a.n_tiles_total;
}

__device__ __forceinline__ uint32_t first_tile(const RasterArgs &a, uint32_t *s_tg) {
  if (a.work_counter == nullptr) return blockIdx.x;
  if (threadIdx.x == 0) *s_tg = atomicAdd(a.work_counter, 1u);
  __syncthreads();
  return *s_tg;
}

// ---------------------------------------------------------------------
// training_tiled_forward (training-forward-tiles.wgsl:107-215). Per pixel
// (centre +0.5) over the tile list front to back: power = 0.5(a dx^2 +
// c dy^2) + b dx dy, skipped if < 0; alpha = min(sigmoid(logit) e^-power,
// 0.999), skipped if < 1/255; C += c alpha (1 - A); A += alpha (1 - A);
// stop when 1 - A < 1e-4 after the add. last = last visited index. Output
// C + (1 - A) bg. Empty tile: bg, A = 0, last = NO_LAST.
// ---------------------------------------------------------------------
template <int B>
__global__ __launch_bounds__(RT) void k_forward(RasterArgs a) {
  __shared__ float4 s_x0[B];
  __shared__ float4 s_x1[B];
  __shared__ float s_x2[B];
  __shared__ uint32_t s_tg;
  uint32_t tg = first_tile(a, &s_tg);
  while (tg < a.n_tiles_total) {
    const uint32_t s = a.tile_view[tg];
    const DevView &V = a.views[s];
    const uint32_t lt = tg - V.tile_base;
    const uint32_t tx = lt % V.tiles_x, ty = lt / V.tiles_x;
    const uint32_t px = tx * 16u + (threadIdx.x & 15u), py = ty * 16u + (threadIdx.x >> 4);
    const bool inside = px < V.W && py < V.H;
    const size_t P = (size_t)V.W * V.H;
    const size_t lp = (size_t)py * V.W + px;
    const size_t pix = V.pix_off + lp;
    float *col = a.color + 3 * V.pix_off;
    const uint32_t start = a.ranges[tg], end = a.ranges[tg + 1];
    if (start >= end) {
      if (inside) {
        col[lp] = a.cfg.bg[0];
        col[P + lp] = a.cfg.bg[1];
        col[2 * P + lp] = a.cfg.bg[2];
        a.alpha[pix] = 0.0f;
        a.last[pix] = NO_LAST;
      }
    } else {
      const float fx = (float)px + 0.5f, fy = (float)py + 0.5f;
      const size_t vbase = (size_t)s * a.N;
      float c0 = 0.f, c1 = 0.f, c2 = 0.f, A = 0.f;
      uint32_t last = start;
      bool done = !inside;
      for (uint32_t b0 = start; b0 < end; b0 += B) {
        const uint32_t cnt = min((uint32_t)B, end - b0);
        for (uint32_t k = threadIdx.x; k < cnt; k += RT) {
          const size_t vg = vbase + min(a.vals[b0 + k], a.N - 1u);
          s_x0[k] = a.rec[vg];
          s_x1[k] = a.rec[a.rs + vg];
          s_x2[k] = a.rec[2 * a.rs + vg].x;
        }
        __syncthreads();
        if (!done) {
          for (uint32_t k = 0; k < cnt; ++k) {
            last = b0 + k;
            const float4 x0 = s_x0[k];
            const float4 x1 = s_x1[k];
            const float dx = fx - x0.x, dy = fy - x0.y;
            const float sg = 0.5f * (x0.z * dx * dx + x1.x * dy * dy) + x0.w * dx * dy;
            if (sg < 0.0f) continue;
            const float gauss = sv_expf(-sg);
            const float as = fminf(x1.y * gauss, SV_MAX_ALPHA);
            if (as < SV_MIN_ALPHA) continue;
            const float contrib = as * (1.0f - A);
            c0 = c0 + x1.z * contrib;
            c1 = c1 + x1.w * contrib;
            c2 = c2 + s_x2[k] * contrib;
            A = A + contrib;
            if (1.0f - A < SV_T_EPS) {
              done = true;
              break;
            }
          }
        }
        if (__syncthreads_and(done)) break;
      }
      if (inside) {
        const float T = 1.0f - A;
        col[lp] = c0 + T * a.cfg.bg[0];
        col[P + lp] = c1 + T * a.cfg.bg[1];
        col[2 * P + lp] = c2 + T * a.cfg.bg[2];
        a.alpha[pix] = A;
        a.last[pix] = last;
      }
    }
    if (!next_tile(a, tg, &s_tg)) break;
  }
}

// Position of an intersection in emit order: the Gaussian's first slot
// plus its row-major index inside the Gaussian's tile rectangle
// (training-tiles.wgsl:238-246).
__device__ __forceinline__ uint32_t emit_pos(const RasterArgs &a, size_t vg, uint32_t tx, uint32_t ty) {
  const float4 r3 = a.rec[3 * a.rs + vg];
  const uint32_t u = __float_as_uint(r3.x), w = __float_as_uint(r3.y);
  const uint32_t x0 = u & 0xffffu, y0 = u >> 16, x1 = w & 0xffffu;
  return a.offs[vg] + (ty - y0) * (x1 - x0) + (tx - x0);
}

__device__ __forceinline__ float sign_ref(float x) {
  return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f);
}

// ---------------------------------------------------------------------
// Adjoint backward, order-correct. Math per (pixel, splat) is that of
// training_tiled_backward (training-backward-tiles.wgsl:318-396): reverse
// traversal from A_final and the last index, T_before = T / max(1 -
// alpha, 1e-4), no alpha > 0.98 skip. Instead of per-pixel float CAS
// atomics into 54 per-Gaussian slots, each wave sums the 9 screen-space
// quantities (sum of res * alpha * T_before per channel, dL/dlogit,
// dL/dconic abc, dL/dmean xy) of one splat over its 64 pixels with a
// fixed butterfly, the four waves are summed in order, and the tile's
// record is written at the intersection's emit position. The SH chain is
// applied per Gaussian afterwards (exact: the SH basis depends only on the
// Gaussian and the view). Deterministic.
// The residual is either fused (w_l2 (p - t) + w_l1 sign(p - t)) or read
// from a.residual.
// ---------------------------------------------------------------------
template <int B>
__global__ __launch_bounds__(RT) void k_backward(RasterArgs a) {
  __shared__ float4 s_x0[B];
  __shared__ float4 s_x1[B];
  __shared__ float s_x2[B];
  __shared__ uint32_t s_p[B];
  __shared__ float s_part[RT / 64][B][REC9];
  __shared__ uint32_t s_wmax[RT / 64];
  __shared__ float s_l1[RT / 64];
  __shared__ uint32_t s_tg;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  uint32_t tg = first_tile(a, &s_tg);
  while (tg < a.n_tiles_total) {
    const uint32_t s = a.tile_view[tg];
    const DevView &V = a.views[s];
    const uint32_t lt = tg - V.tile_base;
    const uint32_t tx = lt % V.tiles_x, ty = lt / V.tiles_x;
    const uint32_t px = tx * 16u + (threadIdx.x & 15u), py = ty * 16u + (threadIdx.x >> 4);
    const bool inside = px < V.W && py < V.H;
    const size_t P = (size_t)V.W * V.H;
    const size_t lp = (size_t)py * V.W + px;
    const size_t pix = V.pix_off + lp;
    const uint32_t start = a.ranges[tg], end = a.ranges[tg + 1];
    const size_t vbase = (size_t)s * a.N;
    if (start < end) {
      float res0 = 0.f, res1 = 0.f, res2 = 0.f, T = 1.0f, l1 = 0.f;
      uint32_t lastp = 0;
      bool valid = false;
      if (inside) {
        lastp = a.last[pix];
        T = 1.0f - a.alpha[pix];
        const float *col = a.color + 3 * V.pix_off;
        const float *tgt = a.target + 3 * V.pix_off;
        const float d0 = col[lp] - tgt[lp];
        const float d1 = col[P + lp] - tgt[P + lp];
        const float d2 = col[2 * P + lp] - tgt[2 * P + lp];
        l1 = (fabsf(d0) + fabsf(d1)) + fabsf(d2);
        if (a.residual) {
          const float *r = a.residual + 3 * V.pix_off;
          res0 = r[lp];
          res1 = r[P + lp];
          res2 = r[2 * P + lp];
        } else {
          res0 = a.cfg.w_l2 * d0 + a.cfg.w_l1 * sign_ref(d0);
          res1 = a.cfg.w_l2 * d1 + a.cfg.w_l1 * sign_ref(d1);
          res2 = a.cfg.w_l2 * d2 + a.cfg.w_l1 * sign_ref(d2);
        }
        valid = (lastp != NO_LAST) && lastp >= start;
      }
      float acc0 = a.cfg.bg[0], acc1 = a.cfg.bg[1], acc2 = a.cfg.bg[2];
      const float fx = (float)px + 0.5f, fy = (float)py + 0.5f;
      // wave / tile maximum of the last index (+1), and the tile L1 sum
      uint32_t m = valid ? lastp + 1u : 0u;
      float l1w = l1;
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        m = max(m, (uint32_t)__shfl_xor((int)m, o, 64));
        l1w = l1w + __shfl_xor(l1w, o, 64);
      }
      if (lane == 0) {
        s_wmax[wave] = m;
        s_l1[wave] = l1w;
      }
      __syncthreads();
      const uint32_t wend = s_wmax[wave];  // this wave's entries lie below wend
      uint32_t tend = max(max(s_wmax[0], s_wmax[1]), max(s_wmax[2], s_wmax[3]));
      if (tend < start) tend = start;
      if (threadIdx.x == 0 && a.l1_partial)
        a.l1_partial[tg] = ((s_l1[0] + s_l1[1]) + s_l1[2]) + s_l1[3];
      // entries no pixel reaches get zero records
      for (uint32_t i = tend + threadIdx.x; i < end; i += RT) {
        const size_t vg = vbase + min(a.vals[i], a.N - 1u);
        const uint32_t p = emit_pos(a, vg, tx, ty);
        if ((uint64_t)p < a.n_isect) {
          float *o = a.isect_rec + (size_t)p * REC9;
#pragma unroll
          for (int q = 0; q < REC9; ++q) o[q] = 0.0f;
        }
      }
      for (uint32_t hi = tend; hi > start;) {
        const uint32_t lo = (hi - start > (uint32_t)B) ? hi - B : start;
        const uint32_t cnt = hi - lo;
        __syncthreads();
        for (uint32_t k = threadIdx.x; k < cnt; k += RT) {
          const size_t vg = vbase + min(a.vals[lo + k], a.N - 1u);
          s_x0[k] = a.rec[vg];
          s_x1[k] = a.rec[a.rs + vg];
          s_x2[k] = a.rec[2 * a.rs + vg].x;
          s_p[k] = emit_pos(a, vg, tx, ty);
        }
        for (uint32_t z = threadIdx.x; z < (RT / 64) * (uint32_t)B * REC9; z += RT)
          (&s_part[0][0][0])[z] = 0.0f;
        __syncthreads();
        // entries this wave must visit: indices < wend
        int kmax = (int)cnt - 1;
        if (wend <= lo) kmax = -1;
        else if (wend - lo < cnt) kmax = (int)(wend - lo) - 1;
        for (int k = kmax; k >= 0; --k) {
          const uint32_t i = lo + (uint32_t)k;
          float v0 = 0.f, v1 = 0.f, v2 = 0.f, v3 = 0.f, v4 = 0.f, v5 = 0.f, v6 = 0.f, v7 = 0.f,
                v8 = 0.f;
          bool hit = false;
          if (valid && i <= lastp) {
            const float4 x0 = s_x0[k];
            const float4 x1 = s_x1[k];
            const float dx = fx - x0.x, dy = fy - x0.y;
            const float sg = 0.5f * (x0.z * dx * dx + x1.x * dy * dy) + x0.w * dx * dy;
            if (sg >= 0.0f) {
              const float gauss = sv_expf(-sg);
              const float op = x1.y;
              const float as = fminf(op * gauss, SV_MAX_ALPHA);
              if (as >= SV_MIN_ALPHA) {
                hit = true;
                const float oma = fmaxf(1.0f - as, 0.0001f);
                const float Tb = T / oma;
                const float contrib = as * Tb;
                v0 = res0 * contrib;
                v1 = res1 * contrib;
                v2 = res2 * contrib;
                const float cr = x1.z, cg = x1.w, cb = s_x2[k];
                const float fg = (res0 * cr + res1 * cg) + res2 * cb;
                const float bgt = (res0 * acc0 + res1 * acc1) + res2 * acc2;
                const float dL_dalpha = (fg - bgt) * Tb;
                v3 = (dL_dalpha * as) * (1.0f - op);
                const float dL_dG = dL_dalpha * op;
                const float dL_ds = dL_dG * (-gauss);
                v4 = ((dL_ds * 0.5f) * dx) * dx;
                v5 = (dL_ds * dx) * dy;
                v6 = ((dL_ds * 0.5f) * dy) * dy;
                v7 = -(dL_ds * (x0.z * dx + x0.w * dy));
                v8 = -(dL_ds * (x0.w * dx + x1.x * dy));
                acc0 = as * cr + (1.0f - as) * acc0;
                acc1 = as * cg + (1.0f - as) * acc1;
                acc2 = as * cb + (1.0f - as) * acc2;
                T = Tb;
              }
            }
          }
          if (__ballot(hit) == 0ull) continue;
          v0 = wave_sum(v0);
          v1 = wave_sum(v1);
          v2 = wave_sum(v2);
          v3 = wave_sum(v3);
          v4 = wave_sum(v4);
          v5 = wave_sum(v5);
          v6 = wave_sum(v6);
          v7 = wave_sum(v7);
          v8 = wave_sum(v8);
          if (lane == 0) {
            float *o = &s_part[wave][k][0];
            o[0] = v0;
            o[1] = v1;
            o[2] = v2;
            o[3] = v3;
            o[4] = v4;
            o[5] = v5;
            o[6] = v6;
            o[7] = v7;
            o[8] = v8;
          }
        }
        __syncthreads();
        for (uint32_t k = threadIdx.x; k < cnt; k += RT) {
          if ((uint64_t)s_p[k] >= a.n_isect) continue;
          float *o = a.isect_rec + (size_t)s_p[k] * REC9;
#pragma unroll
          for (int q = 0; q < REC9; ++q)
            o[q] = ((s_part[0][k][q] + s_part[1][k][q]) + s_part[2][k][q]) + s_part[3][k][q];
        }
        hi = lo;
      }
    } else if (threadIdx.x == 0 && a.l1_partial) {
      // empty tile: the loss still sees bg vs target
      a.l1_partial[tg] = 0.0f;
    }
    if (!next_tile(a, tg, &s_tg)) break;
  }
}

// ---------------------------------------------------------------------
// backward_tiles_local (backward-tiles-local.wgsl:185-382): reference
// wavefront backward. 64-entry batches; per (pixel, splat) the 54 values
// (logit, conic abc, mean xy, 48 SH) are converted with i32(v * 65536)
// (truncating, saturating) and added with LDS integer atomics; the i32
// sums are written per intersection. flags: SV_WF_REFERENCE_ORDER walks
// batches front-to-back (:262), else back-to-front (order-correct);
// SV_WF_SKIP_098 keeps the alpha > 0.98 skip (:329-331). Tile size 16.
// ---------------------------------------------------------------------
constexpr int WF_B = 64;
constexpr int WF_S = 54;

__global__ __launch_bounds__(RT) void k_backward_fixed(RasterArgs a, const float *__restrict__ P,
                                                       size_t cap, uint32_t flags,
                                                       int32_t *__restrict__ fixed) {
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
      const size_t vg = vbase + g;
      s_x0[threadIdx.x] = a.rec[vg];
      s_x1[threadIdx.x] = a.rec[a.rs + vg];
      s_x2[threadIdx.x] = a.rec[2 * a.rs + vg].x;
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
        const float sg = 0.5f * (x0.z * dx * dx + x1.x * dy * dy) + x0.w * dx * dy;
        if (sg < 0.0f) continue;
        const float gauss = sv_expf(-sg);
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
    for (uint32_t z = threadIdx.x; z < bsz * WF_S; z += RT)
      fixed[(size_t)bs * WF_S + z] = s_acc[z];
    __syncthreads();
  }
}

template <int B>
static void launch_fwd_b(hipStream_t st, const RasterArgs &a, uint32_t grid) {
  hipLaunchKernelGGL(k_forward<B>, dim3(grid), dim3(RT), 0, st, a);
}
template <int B>
static void launch_bwd_b(hipStream_t st, const RasterArgs &a, uint32_t grid) {
  hipLaunchKernelGGL(k_backward<B>, dim3(grid), dim3(RT), 0, st, a);
}

void launch_forward(hipStream_t st, const RasterArgs &a, uint32_t batch, uint32_t raster_px,
                    uint32_t grid) {
  (void)raster_px;
  if (a.n_tiles_total == 0 || grid == 0) return;
  if (batch <= 64) launch_fwd_b<64>(st, a, grid);
  else if (batch <= 128) launch_fwd_b<128>(st, a, grid);
  else launch_fwd_b<256>(st, a, grid);
}

void launch_backward(hipStream_t st, const RasterArgs &a, uint32_t batch, uint32_t grid) {
  if (a.n_tiles_total == 0 || grid == 0) return;
  if (batch <= 64) launch_bwd_b<64>(st, a, grid);
  else if (batch <= 128) launch_bwd_b<128>(st, a, grid);
  else launch_bwd_b<256>(st, a, grid);
}

void launch_backward_fixed(hipStream_t st, const RasterArgs &a, const float *P, size_t cap,
                           uint32_t flags, int32_t *fixed) {
  if (a.n_tiles_total == 0) return;
  hipLaunchKernelGGL(k_backward_fixed, dim3(a.n_tiles_total), dim3(RT), 0, st, a, P, cap, flags,
                     fixed);
}

}  // namespace sv
