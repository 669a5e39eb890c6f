// sv_score.hip -- multi-view scoring for prune / densify on the device
// (generateMultiViewScores, gaussian-renderer.ts:2640-2781, with the
// shaders of metric-map.wgsl:32-147).
//
// Per rendered view of the batch: the per-pixel l1 metric, its min / max /
// mean (the reference reads the metric map back and reduces it on the
// host, computePixelLossStats :2709-2745), the threshold flag, and, for
// every flagged pixel, one count per splat of its tile list (up to the
// pixel's last index) whose power lies in [0, 6]. The per-view counts are
// then folded into importance (sum) and pruning (+= count * mean, a double
// sum stored to float, as the host's Float32Array accumulation) in view
// order. Counts are integer atomics (order-free); everything else is one
// thread per output, so the results are deterministic.
#include "sv_internal.hpp"

namespace sv {

namespace {

constexpr uint32_t SC_T = 256;
constexpr float METRIC_THRESH = 0.1f;  // MULTI_VIEW_LOSS_THRESH, gaussian-renderer.ts:70

// compute_metric_map (metric-map.wgsl:32-50): l1 = (|dr| + |dg| + |db|) / 3.
__global__ __launch_bounds__(SC_T) void k_metric_map(const DevView *__restrict__ views,
                                                     const float *__restrict__ color,
                                                     const float *__restrict__ target,
                                                     float *__restrict__ loss) {
  const DevView &V = views[blockIdx.y];
  const size_t P = (size_t)V.W * V.H;
  const size_t lp = (size_t)blockIdx.x * SC_T + threadIdx.x;
  if (lp >= P) return;
  const float *c = color + 3 * V.pix_off, *t = target + 3 * V.pix_off;
  const float d0 = fabsf(c[lp] - t[lp]), d1 = fabsf(c[P + lp] - t[P + lp]), d2 = fabsf(c[2 * P + lp] - t[2 * P + lp]);
  loss[V.pix_off + lp] = ((d0 + d1) + d2) / 3.0f;
}

// One workgroup per view: min and max over the finite values, and their
// sum in double (thread t adds pixels t, t + 1024, ... in order, then a
// fixed pairwise tree); mean = sum / pixels. stats[s]: min, max (as
// floats), inv_range = f32(1 / (max - min)) in double, valid flag, mean.
__global__ __launch_bounds__(1024) void k_metric_stats(const DevView *__restrict__ views,
                                                       const float *__restrict__ loss,
                                                       MetricStats *__restrict__ stats) {
  __shared__ double s_sum[1024];
  __shared__ float s_mn[1024], s_mx[1024];
  const uint32_t s = blockIdx.x;
  const DevView &V = views[s];
  const size_t P = (size_t)V.W * V.H;
  const float *l = loss + V.pix_off;
  double sum = 0.0;
  float mn = INFINITY, mx = -INFINITY;
  for (size_t i = threadIdx.x; i < P; i += 1024) {
    const float x = l[i];
    if (!isfinite(x)) continue;
    mn = fminf(mn, x);
    mx = fmaxf(mx, x);
    sum += (double)x;
  }
  s_sum[threadIdx.x] = sum;
  s_mn[threadIdx.x] = mn;
  s_mx[threadIdx.x] = mx;
  __syncthreads();
  for (uint32_t st = 512; st > 0; st >>= 1) {
    if (threadIdx.x < st) {
      s_sum[threadIdx.x] = s_sum[threadIdx.x] + s_sum[threadIdx.x + st];
      s_mn[threadIdx.x] = fminf(s_mn[threadIdx.x], s_mn[threadIdx.x + st]);
      s_mx[threadIdx.x] = fmaxf(s_mx[threadIdx.x], s_mx[threadIdx.x + st]);
    }
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  MetricStats o;
  o.min = s_mn[0];
  o.max = s_mx[0];
  o.valid = (isfinite(o.min) && isfinite(o.max)) ? 1u : 0u;
  const double lo = o.min, hi = o.max;
  o.inv_range = (o.valid && hi > lo) ? (float)(1.0 / (hi - lo)) : 0.0f;
  o.mean = P ? s_sum[0] / (double)P : 0.0;
  stats[s] = o;
}

// threshold_metric_map + accumulate_metric_counts (metric-map.wgsl:52-147):
// one thread per pixel of view blockIdx.y; counts are per (slot, Gaussian).
__global__ __launch_bounds__(SC_T) void k_metric_counts(const DevView *__restrict__ views,
                                                        const MetricStats *__restrict__ stats,
                                                        const float *__restrict__ loss,
                                                        const uint32_t *__restrict__ ranges,
                                                        const uint32_t *__restrict__ vals,
                                                        const uint32_t *__restrict__ last,
                                                        const float4 *__restrict__ rec, uint32_t N,
                                                        uint32_t *__restrict__ counts) {
  const uint32_t s = blockIdx.y;
  const DevView &V = views[s];
  const size_t lp = (size_t)blockIdx.x * SC_T + threadIdx.x;
  if (lp >= (size_t)V.W * V.H) return;
  const MetricStats st = stats[s];
  if (!st.valid) return;
  const float value = (loss[V.pix_off + lp] - st.min) * st.inv_range;
  if (!(value > METRIC_THRESH)) return;
  const uint32_t x = (uint32_t)(lp % V.W), y = (uint32_t)(lp / V.W);
  const uint32_t tx = x / 16u, ty = y / 16u;
  if (tx >= V.tiles_x || ty >= V.tiles_y) return;
  const uint32_t tg = V.tile_base + ty * V.tiles_x + tx;
  const uint32_t start = ranges[tg], end = ranges[tg + 1];
  if (start >= end) return;
  const uint32_t li = last[V.pix_off + lp];
  if (li < start) return;
  const uint32_t fin = min(li, end - 1u);
  const float px = (float)x + 0.5f, py = (float)y + 0.5f;
  const float4 *vrec = rec + (size_t)s * N * 4;
  uint32_t *cnt = counts + (size_t)s * N;
  for (uint32_t i = start; i <= fin; ++i) {
    const uint32_t g = vals[i];
    if (g >= N) continue;
    const float4 q0 = vrec[(size_t)g * 4], q1 = vrec[(size_t)g * 4 + 1];
    const float dx = px - q0.x, dy = py - q0.y;
    const float p = 0.5f * (q0.z * dx * dx + q1.x * dy * dy) + q0.w * dx * dy;
    if (p < 0.0f || p > 6.0f) continue;
    atomicAdd(&cnt[g], 1u);
  }
}

// Fold the per-view counts in view order: importance += c, pruning =
// f32(pruning + c * mean) (gaussian-renderer.ts:2770-2775).
__global__ __launch_bounds__(SC_T) void k_metric_fold(const MetricStats *__restrict__ stats, int nv, uint32_t N,
                                                      const uint32_t *__restrict__ counts,
                                                      uint32_t *__restrict__ importance,
                                                      float *__restrict__ pruning) {
  const uint32_t g = blockIdx.x * SC_T + threadIdx.x;
  if (g >= N) return;
  uint32_t imp = 0;
  float pr = 0.0f;
  for (int s = 0; s < nv; ++s) {
    if (!stats[s].valid) continue;
    const uint32_t c = counts[(size_t)s * N + g];
    imp += c;
    pr = (float)((double)pr + (double)c * stats[s].mean);
  }
  importance[g] = imp;
  pruning[g] = pr;
}

}  // namespace

void launch_multiview_scores(hipStream_t st, const DevView *views, int nv, uint32_t max_pixels,
                             const float *color, const float *target, const uint32_t *ranges, const uint32_t *vals,
                             const uint32_t *last, const float4 *rec, uint32_t N, float *loss, MetricStats *stats,
                             uint32_t *counts, uint32_t *importance, float *pruning) {
  if (nv <= 0 || N == 0) return;
  const dim3 pg((max_pixels + SC_T - 1) / SC_T, (uint32_t)nv);
  hipLaunchKernelGGL(k_metric_map, pg, dim3(SC_T), 0, st, views, color, target, loss);
  hipLaunchKernelGGL(k_metric_stats, dim3((uint32_t)nv), dim3(1024), 0, st, views, loss, stats);
  launch_fill_u32(st, counts, 0u, (size_t)nv * N);
  hipLaunchKernelGGL(k_metric_counts, pg, dim3(SC_T), 0, st, views, stats, loss, ranges, vals, last, rec, N,
                     counts);
  hipLaunchKernelGGL(k_metric_fold, dim3((N + SC_T - 1) / SC_T), dim3(SC_T), 0, st, stats, nv, N, counts,
                     importance, pruning);
}

}  // namespace sv
