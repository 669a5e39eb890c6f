/*
 * sv_oracle.h -- CPU restatement of Skyfall-Viewer's splat training step.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker for the HIP
 * implementation under skyfall-viewer_amd/. Only tests/, the smoke check
 * in __graft_entry__.py and the cpu_baseline leg of bench.py may load it.
 * The product library never links or calls it.
 *
 * Every function restates one WGSL entry point (or one host routine) of the
 * reference, with the file:line it follows. Floating-point operations are
 * written in an explicit order and the library is compiled with
 * -ffp-contract=off, so each float is the IEEE result of that order.
 * exp() is orc_expf(): Cody-Waite range reduction plus the Cephes expf
 * minimax polynomial (S. L. Moshier, Cephes Math Library), so the value is
 * reproducible on any IEEE machine (WGSL leaves exp() implementation
 * defined, see src/shaders/training-forward-tiles.wgsl:188,192).
 *
 * Parity status of each function (how it is pinned) is listed in the
 * header comment of sv_oracle.c.
 */
#ifndef SV_ORACLE_H
#define SV_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* CameraUniforms, src/shaders/common.wgsl:21-28 (272 bytes, column-major). */
typedef struct orc_camera {
  float view[16];
  float view_inv[16];
  float proj[16];
  float proj_inv[16];
  float viewport[2];
  float focal[2];
} orc_camera;

/* TilingParams + ProjectionParams + BackgroundParams + RenderSettings.sh_deg
 * (training-tiles.wgsl:1-15, training-forward-tiles.wgsl:10-13,
 * common.wgsl:30-37). */
typedef struct orc_view {
  uint32_t width, height, tile_size, tiles_x, tiles_y;
  uint32_t sh_deg;
  float near_plane, far_plane, radius_clip, eps2d;
  float bg[3];
  uint32_t geom_flags; /* 1: the geometry backward's mean path takes the
                          forward's y convention (the library's
                          SV_GEOM_MEAN_Y_FORWARD); 0: the shader as written */
} orc_view;

/* GaussProjection, common.wgsl:57-65 (48 bytes). */
typedef struct orc_proj {
  float mean2d[2];
  float radius;
  float depth;
  float conic[3];
  uint32_t tiles_count;
  float position_world[3];
  float opacity_logit;
} orc_proj;

/* Training parameters in the reference's f32 training layout
 * (gaussian-renderer.ts:130-137): pos[3N], log-scale[3N], quaternion
 * (x,y,z,w)[4N], opacity logit[N], SH[48N] (coefficient k, channel c at
 * 3k+c). */
typedef struct orc_params {
  const float *pos, *logscale, *quat, *logit, *sh;
} orc_params;

/* Per-Gaussian gradient buffers (gaussian-renderer.ts:216-231). */
typedef struct orc_grads {
  float *pos, *logscale, *quat, *logit, *sh, *geom; /* geom: 5N conic a,b,c, mean x,y */
} orc_grads;

#define ORC_TILE_GRAD_STRIDE 54u

/* Wavefront backward options (backward-tiles-local.wgsl). */
#define ORC_WF_REFERENCE_ORDER 0x1u /* batches front-to-back, :262 */
#define ORC_WF_SKIP_098        0x2u /* skip alpha > 0.98, :329-331 */

/* exp of the per-pixel splat weight (2^(x log2 e), degree-5 polynomial),
 * for x in [-80, 1/2]; within 3 + 2|x| ulp of e^x. */
float orc_exp_alpha(float x);
void orc_exp_alpha_v(size_t n, const float *x, float *y);
float orc_expf(float x);
float orc_f16_round(float x);

void orc_project(size_t n, const orc_params *p, const orc_camera *cam,
                 const orc_view *v, orc_proj *out);
void orc_sh_color(const float dir[3], const float *sh48, uint32_t deg,
                  float rgb[3]);
void orc_view_dir(const orc_camera *cam, const float pos[3], float dir[3]);
uint32_t orc_encode_key(uint32_t cam, uint32_t tile, float depth);
uint64_t orc_count_isects(size_t n, const orc_proj *proj);
void orc_emit(size_t n, const orc_proj *proj, const orc_view *v,
              uint32_t *keys, uint32_t *vals);
void orc_sort_pairs(size_t m, uint32_t *keys, uint32_t *vals);
/* The same stable sort as a parallel LSD radix sort (OpenMP threads). */
void orc_sort_pairs_radix(size_t m, uint32_t *keys, uint32_t *vals);
void orc_tile_offsets(size_t m, const uint32_t *keys, uint32_t n_tiles,
                      uint32_t *offsets);
void orc_forward(const orc_params *p, const orc_proj *proj,
                 const orc_camera *cam, const orc_view *v,
                 const uint32_t *offsets, const uint32_t *vals, size_t m,
                 float *color, float *alpha, uint32_t *last_ids);
void orc_loss(const float *color, const float *target, uint32_t w, uint32_t h,
              float w_l2, float w_l1, float w_ssim, float *residual);
void orc_backward_legacy(size_t n, const orc_params *p, const orc_proj *proj,
                         const orc_camera *cam, const orc_view *v,
                         const uint32_t *offsets, const uint32_t *vals,
                         size_t m, const float *residual, const float *alpha,
                         const uint32_t *last_ids, orc_grads *g);
void orc_backward_wavefront(const orc_params *p, const orc_proj *proj,
                            const orc_camera *cam, const orc_view *v,
                            const uint32_t *offsets, const uint32_t *vals,
                            size_t m, const float *residual,
                            const float *alpha, const uint32_t *last_ids,
                            uint32_t flags, int32_t *tile_grads_fixed);
void orc_backward_reduce(size_t n, const int32_t *tile_grads_fixed,
                         const uint32_t *vals, size_t m, orc_grads *g);
void orc_backward_geom(size_t n, const orc_params *p, const orc_camera *cam,
                       uint32_t geom_flags, orc_grads *g);
void orc_adam(size_t count, float *param, float *m1, float *m2,
              const float *grad, float lr, float beta1, float beta2,
              float eps, float bc1_rcp, float bc2_sqrt_rcp);
void orc_adam_bias(uint32_t t, float beta1, float beta2, float *bc1_rcp,
                   float *bc2_sqrt_rcp);
void orc_accumulate_importance(size_t n, const orc_proj *proj, float decay,
                               float *stats4);
void orc_accumulate_grads(size_t n, const float *grad_geom, float *accum4);

/* Whole-step driver: project, emit, sort, offsets, forward, loss, backward
 * (legacy order-correct float when fixed == NULL, else the wavefront
 * fixed-point kernel + reduce with wf_flags), geometry backward.
 * Returns the intersection count. Buffers the caller does not need may be
 * NULL, except the gradient planes in g. */
typedef struct orc_step_out {
  orc_proj *proj;           /* n */
  uint32_t *keys, *vals;    /* capacity cap_isect each, sorted on return */
  size_t cap_isect;
  uint32_t *offsets;        /* tiles + 1 */
  float *color, *alpha;     /* 3P, P */
  uint32_t *last_ids;       /* P */
  float *residual;          /* 3P */
  int32_t *tile_grads_fixed;/* 54 * cap_isect, wavefront path only */
} orc_step_out;

int64_t orc_render_and_grad(size_t n, const orc_params *p,
                            const orc_camera *cam, const orc_view *v,
                            const float *target, float w_l2, float w_l1,
                            float w_ssim, int use_wavefront, uint32_t wf_flags,
                            orc_step_out *o, orc_grads *g);

/* Float64 forward + loss (for central finite differences). The loss is
 * sum over pixels and channels of w_l2/2 (p-t)^2 + w_l1 |p-t|, whose
 * gradient is the reference residual (training-loss.wgsl:100-107). */
double orc_forward_loss_f64(size_t n, const double *pos, const double *logscale,
                            const double *quat, const double *logit,
                            const double *sh, const orc_camera *cam,
                            const orc_view *v, const float *target,
                            double w_l2, double w_l1, double *color_out);

/* The same loss summed over the pixels of Gaussian `only`'s tile rectangle
 * alone (0 when it is culled). A change of that Gaussian's SH coefficients
 * leaves its projection, every tile list and every other pixel unchanged,
 * so the central difference of this sum is that of the whole loss. */
double orc_forward_loss_f64_only(size_t n, const double *pos, const double *logscale,
                                 const double *quat, const double *logit,
                                 const double *sh, const orc_camera *cam,
                                 const orc_view *v, const float *target,
                                 double w_l2, double w_l1, size_t only);

/* Stable compaction (gaussian-renderer.ts:2917-3000): rows with keep[i] != 0
 * are copied in order. Returns the kept count. `stride` floats per row. */
size_t orc_compact_rows(size_t n, size_t stride, const float *src,
                        const uint8_t *keep, float *dst);

/* Prune rule: keep[i] = !(sigmoid(logit_i) < min_opacity), sigmoid as
 * training-forward-tiles.wgsl:188. */
void orc_prune_keep(size_t n, const float *logit, float min_opacity, uint8_t *keep);

/* performDensification (gaussian-renderer.ts:3211-3466) on rows of 59
 * floats (position 3, log-scale 3, quaternion xyzw 4, logit 1, SH 48) with
 * their two moment rows. actions: 0 keep, 1 clone, 2 split. Output: every
 * row that is not split, in order, then the children in parent order (clone:
 * 1, split: 2) with zero moments. The reference's Math.random draws come
 * from the counter stream u(seed, 16 g + k). Returns the output count. */
size_t orc_densify(size_t n, const float *rows, const float *m1, const float *m2,
                   const uint8_t *actions, uint64_t seed, float *out, float *out_m1,
                   float *out_m2);

/* Per-tile L1 error of one view: e_p = (|dr| + |dg|) + |db| per pixel, the
 * 256 pixel slots of the tile in row-major order (0 outside the image),
 * summed by a fixed pairwise tree (slot i += slot i + s, s = 128 .. 1).
 * color / target: H x W x 3. */
void orc_tile_l1(const orc_view *v, const float *color, const float *target, float *tile_l1);

/* Densify rule (the 300-step training run): a Gaussian is a candidate when a tile of its
 * rectangle in a view where it is listed has L1 error above that view's
 * threshold; candidates split when their largest projected radius over
 * those views exceeds split_px, else clone. proj[v] / tile_l1[v] per view.
 * score (may be NULL): per candidate the largest tile error over its tiles
 * divided by the view's threshold (the error where the threshold <= 1e-8),
 * 0 for the others (the normalised score of gaussian-renderer.ts:3280-3286). */
void orc_densify_decide(size_t n, size_t nviews, const orc_proj *const *proj,
                        const orc_view *views, const float *const *tile_l1,
                        const float *thr, float split_px, uint8_t *actions, float *score);

/* The reference trainer's densify candidates (performDensification,
 * gaussian-renderer.ts:3249-3290), in double as the JS evaluates them.
 * rule[6] = grad_threshold, grad_abs_threshold, size_threshold,
 * min_visibility, importance_threshold, pixel_scale; accum4 = per Gaussian
 * (sum |dL/dmean2d|, unused, sum |dL/dconic|, visible steps); importance =
 * multi-view counts over n_samples views. score (may be NULL): the
 * normalised score of each candidate, 0 for the others. */
void orc_densify_decide_grad(size_t n, const float *accum4, const float *logscale,
                             const uint32_t *importance, uint32_t n_samples, const double *rule,
                             uint8_t *actions, double *score);

/* knn_init (knn-init.wgsl:40-90; constants gaussian-renderer.ts:77-79):
 * for each of the n points (positions xyz interleaved, rounded to f16 as
 * read_position does) the three smallest distances to the other points
 * (self excluded by index, duplicates count; 1e10 when fewer than three),
 * density = (d0 + d1 + d2) / 3, sigma = min(clamp(max(density / 3, 0.0015
 * e), 0.0007 e, 0.007 e), 0.1) with e = extent, logscale = f16(log(max(sigma,
 * 1e-6))). O(n^2), as the shader. density / logscale: n floats each. */
void orc_knn_init(size_t n, const float *pos, float extent, float *density, float *logscale);

/* Scoring of one rendered view for generateMultiViewScores
 * (gaussian-renderer.ts:2640-2781, metric-map.wgsl:32-147):
 * l1 = ((|dr| + |dg|) + |db|) / 3 per pixel; min / max over finite values
 * and mean = (sequential double sum) / pixels (computePixelLossStats); a
 * pixel is flagged when (l1 - min) * f32(1 / (max - min)) > 0.1; for every
 * flagged pixel, counts[g] += 1 for each splat g of its tile list from the
 * start to min(last id, end - 1) with 0 <= power <= 6. Returns 0 when the
 * view has no finite value (the reference skips it), else 1; *mean set.
 * color / target: H x W x 3; counts: n, accumulated (not cleared). */
int orc_metric_view(const orc_view *v, size_t n, const orc_proj *proj, const uint32_t *offsets,
                    const uint32_t *vals, const uint32_t *last_ids, const float *color,
                    const float *target, uint32_t *counts, double *mean);

/* Vector helpers for the tests. */
void orc_expf_v(size_t n, const float *x, float *y);
void orc_f16_round_v(size_t n, const float *x, float *y);
void orc_sh_color_v(size_t n, const float *dirs, const float *sh48,
                    uint32_t deg, float *rgb);

#ifdef __cplusplus
}
#endif
#endif
