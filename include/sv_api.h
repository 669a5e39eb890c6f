/*
 * sv_api.h -- C ABI of the MI355X splat training step (libsv.so).
 *
 * Drop-in boundary for Skyfall-Viewer's training path. The reference runs
 * the step as WGSL passes recorded by the GaussianRenderer host
 * (src/renderers/gaussian-renderer.ts); each entry point below replaces one
 * host routine of that file, cited per function. A Node N-API addon
 * (skyfall-viewer_amd/node/) binds these symbols for the TypeScript host,
 * and Python binds them with ctypes.
 *
 * Conventions
 *  - Every function returns an sv_status (0 = SV_OK) unless stated.
 *  - Pointers are caller-owned HOST memory, copied in or out before the call
 *    returns. No device pointer crosses the ABI.
 *  - One sv_ctx owns one HIP stream on one device (and a side stream that
 *    the batched step forks work onto and joins). A context is not
 *    thread-safe; calls block until the streams are idle.
 *  - Gaussian parameters use the reference's f32 training layout
 *    (gaussian-renderer.ts:130-137): pos[3N], log-scale[3N], quaternion
 *    (x,y,z,w)[4N], opacity logit[N], SH[48N] (coefficient k, channel c at
 *    index 3k+c). Images are row-major H x W x 3 float.
 */
#ifndef SV_API_H
#define SV_API_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum sv_status {
  SV_OK = 0,
  SV_E_ARG = 1,      /* bad argument (null pointer, index out of range) */
  SV_E_CAPACITY = 2, /* more Gaussians/intersections than allocated */
  SV_E_HIP = 3,      /* HIP runtime error; see sv_last_error */
  SV_E_STATE = 4,    /* call out of order (no views, no target, ...) */
  SV_E_NOMEM = 5     /* device allocation failed */
} sv_status;

typedef struct sv_ctx sv_ctx;

/* CameraUniforms, src/shaders/common.wgsl:21-28 and
 * src/camera/camera-uniform.ts:8 (272 bytes, matrices column-major).
 * viewport = (W, H) in pixels, focal = (fx, fy) in pixels. */
typedef struct sv_camera {
  float view[16];
  float view_inv[16];
  float proj[16];
  float proj_inv[16];
  float viewport[2];
  float focal[2];
} sv_camera;

/* Training configuration. Defaults (sv_default_config) are the reference's:
 * tile 16 (gaussian-renderer.ts:366), projection params 0.2/1e4/512/0.3
 * (:474-488), background black (:490-504), loss weights (0, 0.8, 0.2)
 * (:506-522), learning rates (:88-96), Adam 0.9/0.999/1e-8 (:2421-2423),
 * SH update interval 16 (:84), importance decay 0.98 (importance.ts:8-13). */
typedef struct sv_config {
  uint32_t tile_size;  /* binning tile in pixels: 16 */
  uint32_t batch;      /* splats staged per LDS batch: 64, 128 or 256 */
  uint32_t raster_px;  /* raster tile of the forward in pixels: 16 (one wave
                          per binning tile), 8 (one wave per 8x8 quadrant) or
                          32 (four waves per 2x2 tiles); the lists stay the
                          16-px tiles', so the images are the same */
  uint32_t sh_deg;     /* RenderSettings.sh_deg, 0..3 */
  float bg[3];
  float near_plane, far_plane, radius_clip, eps2d;
  float w_l2, w_l1, w_ssim;
  float lr_pos, lr_sh, lr_opacity, lr_scale, lr_rot;
  float beta1, beta2, eps;
  uint32_t sh_update_interval;
  uint32_t pos_lr_schedule; /* 1: computePositionLearningRate (:284-298) */
  uint32_t render_f16;      /* 1: render from f16-rounded params (A6) */
  float importance_decay;
  uint32_t persistent;      /* 1: raster kernels pull (view, tile) work items
                               from a device queue in one persistent launch */
  uint32_t use_graph;       /* 1: capture the batched step in a hipGraph */
  uint32_t keep_grads;      /* 1: batched steps also write the gradient
                               buffers (sv_read_grads) before Adam */
  uint32_t plain_wf_flags;  /* SV_WF_* flags of the SV_MODE_PLAIN backward */
  uint32_t xcd_remap;       /* 1: raster workgroups on one XCD take a
                               contiguous range of tiles (L2 locality of
                               neighbouring tiles' records) */
  uint32_t binned_build;    /* 1: build the tile lists by binning splats into
                               8x8-tile supertiles (sort of one pair per
                               splat and supertile) instead of sorting every
                               (tile, splat) pair; same lists and ranges */
  uint32_t geom_flags;      /* SV_GEOM_* flags of the geometry backward; 0 =
                               the reference's training_backward_geom as
                               written */
  uint32_t split_backward;  /* 1: a batched step of two or more views runs
                               its backward as two launches over halves of
                               the views, the first half's record gather
                               overlapping the second half's backward on a
                               second stream (same results; default 0, the
                               autotuner tries both) */
  uint32_t tile_order;      /* 1 (default): the 16-px forward and backward
                               take each XCD's tiles longest list first (a
                               per-step bucket sort of the tiles by list
                               length), so the longest tiles do not start
                               last; same results */
} sv_config;

/* Step modes for sv_step. */
#define SV_MODE_PLAIN 0      /* the reference's method, one view at a time:
                                wavefront fixed-point backward (54 i32 slots per
                                intersection), atomic reduce, separate
                                geometry backward, then Adam */
#define SV_MODE_BATCHED 1    /* adjoint backward, all views in one pipeline
                                with (view, tile, depth) keys */
#define SV_MODE_SEQUENTIAL 2 /* adjoint backward, one view at a time (same
                                kernels and results as SV_MODE_BATCHED) */

/* Wavefront fixed-point backward flags (backward-tiles-local.wgsl). */
#define SV_WF_REFERENCE_ORDER 0x1u /* 64-entry batches front-to-back (:262) */
#define SV_WF_SKIP_098 0x2u        /* skip alpha > 0.98 (:329-331) */

/* Geometry backward flags (training-backward-geom.wgsl). */
#define SV_GEOM_MEAN_Y_FORWARD 0x1u /* the mean path's y terms take the
                                       forward's pixel convention, y = H - c_y
                                       - f_y y / z (training-tiles.wgsl:149-151
                                       with the projection of
                                       training-camera.ts:39-51), instead of
                                       the +f_y of :264-266, whose position-y
                                       gradient has the opposite sign to the
                                       forward's dependence */

typedef struct sv_step_stats {
  uint64_t n_isect;     /* intersections sorted this step (all views) */
  uint32_t n_views;
  uint32_t sort_passes; /* 8-bit digit passes of the last sort (of the
                           supertile pairs when binned) */
  uint32_t key_bits;    /* significant bits of the (tile, depth) key */
  uint32_t binned;      /* 1: the tile lists were built by binning */
  double l1_sum;        /* sum |pred - target| over all pixels and channels */
} sv_step_stats;

/* Per-kernel timing record (sv_timings). */
typedef struct sv_kernel_time {
  char name[32];
  double total_ms;
  uint64_t launches;
} sv_kernel_time;

/* One candidate of sv_autotune. */
typedef struct sv_tune_entry {
  uint32_t batch;       /* LDS batch of the rasterizers */
  uint32_t persistent;  /* persistent raster launch */
  uint32_t xcd_remap;   /* XCD-contiguous tile ranges */
  uint32_t tile_size;   /* raster tile of the forward: 8, 16 or 32 */
  uint32_t split_backward; /* the split backward of the batched step */
  double ms_per_step;   /* mean device time of one batched step */
} sv_tune_entry;

/* --- library / context ------------------------------------------------ */
/* Version string. */
const char *sv_version(void);
/* Fill *out with the reference defaults listed above. */
sv_status sv_default_config(sv_config *out);
/* get_renderer(...) buffer creation (gaussian-renderer.ts:154-161,612-768):
 * create a context on HIP device `device` with room for `capacity`
 * Gaussians. */
sv_status sv_create(int device, size_t capacity, sv_ctx **out);
sv_status sv_destroy(sv_ctx *ctx);
/* Message of the last failing call on ctx (or of sv_create when ctx is
 * NULL). */
const char *sv_last_error(const sv_ctx *ctx);
/* setTrainingParams / setLossWeights (gaussian-renderer.ts:3922-3930,
 * 4058-4074) plus the uniform buffers of :464-522. */
sv_status sv_set_config(sv_ctx *ctx, const sv_config *cfg);
sv_status sv_get_config(const sv_ctx *ctx, sv_config *cfg);

/* --- parameters ------------------------------------------------------- */
/* reloadPointCloud + training_init_params (gaussian-renderer.ts:2092-2126,
 * training-params.wgsl:35-79): upload n Gaussians; clears moments, grads,
 * statistics and the step counter. Any pointer may not be NULL. A NaN or
 * infinite value returns SV_E_ARG and changes nothing (so do non-finite
 * moments, targets and residuals below). */
sv_status sv_set_gaussians(sv_ctx *ctx, size_t n, const float *pos,
                           const float *logscale, const float *quat,
                           const float *logit, const float *sh);
/* readBufferToCPU of the train_* buffers (gaussian-renderer.ts:3129-3205).
 * NULL pointers are skipped. */
sv_status sv_get_gaussians(sv_ctx *ctx, float *pos, float *logscale,
                           float *quat, float *logit, float *sh);
/* getGaussianCount (gaussian-renderer.ts:4110-4153). */
size_t sv_get_count(const sv_ctx *ctx);
/* readAdamMoments / writeAdamMoments (gaussian-renderer.ts:3129-3205);
 * which = 1 (first moment) or 2 (second moment). */
sv_status sv_get_moments(sv_ctx *ctx, int which, float *pos, float *logscale,
                         float *quat, float *logit, float *sh);
sv_status sv_set_moments(sv_ctx *ctx, int which, const float *pos,
                         const float *logscale, const float *quat,
                         const float *logit, const float *sh);
/* Adam step counter (adam_step_counter, gaussian-renderer.ts:2544). */
sv_status sv_get_step(const sv_ctx *ctx, uint32_t *t);
sv_status sv_set_step(sv_ctx *ctx, uint32_t t);

/* --- views and targets ------------------------------------------------ */
/* setTrainingCameras (gaussian-renderer.ts:4084-4094): nv cameras; the
 * render size of view v is round(viewport). */
sv_status sv_set_views(sv_ctx *ctx, int nv, const sv_camera *cams);
/* setTrainingImages (gaussian-renderer.ts:4096-4105): target of view v as
 * H x W x 3 float (the reference samples rgba8 textures at texel centres,
 * training-loss.wgsl:94-102). */
sv_status sv_set_target(sv_ctx *ctx, int view, const float *rgb);

/* --- Function-style pieces of the plain path ------------------------- */
/* buildTileIntersections (gaussian-renderer.ts:1780-2023): project, count,
 * device scan, emit, radix sort, tile ranges for view v. */
sv_status sv_build_tiles(sv_ctx *ctx, int view, uint64_t *n_isect);
/* run_training_tiled_forward (gaussian-renderer.ts:2261-2290). */
sv_status sv_forward(sv_ctx *ctx, int view);
/* run_training_loss (gaussian-renderer.ts:2292-2310): residual of view v. */
sv_status sv_loss(sv_ctx *ctx, int view);
/* Supply dL/d(image) of view v directly (H x W x 3), replacing sv_loss. */
sv_status sv_set_residual(sv_ctx *ctx, int view, const float *rgb);
/* run_training_tiled_backward + run_training_backward_geom
 * (gaussian-renderer.ts:2312-2407), order-correct adjoint: reverse
 * traversal from A_final and last index, then the per-Gaussian SH chain
 * and geometry VJP. Adds into the gradient buffers. */
sv_status sv_backward(sv_ctx *ctx, int view);
/* run_wavefront_backward (gaussian-renderer.ts:2327-2385): the reference's
 * i32 fixed-point per-intersection sums (backward-tiles-local.wgsl) and
 * the per-Gaussian reduce (backward-reduce.wgsl) + geometry backward.
 * flags: SV_WF_*. Adds into the gradient buffers. */
sv_status sv_backward_fixed(sv_ctx *ctx, int view, uint32_t flags);
/* zeroGradBuffers (gaussian-renderer.ts:3596-3623). */
sv_status sv_zero_grad(sv_ctx *ctx);
/* run_adam_updates (gaussian-renderer.ts:2455-2545) with per-tensor
 * learning rates; increments the step counter first (t = step index). */
sv_status sv_adam(sv_ctx *ctx);

/* --- whole step -------------------------------------------------------- */
/* trainStep (gaussian-renderer.ts:3638-3660) over nv views: zero grads,
 * render + loss + backward for every view (gradients summed), one Adam
 * step. mode SV_MODE_PLAIN runs the views one after another;
 * SV_MODE_BATCHED runs them through one pipeline. stats may be NULL. */
sv_status sv_step(sv_ctx *ctx, const int *views, int nv, int mode,
                  sv_step_stats *stats);
/* Forward only for nv views in one batched pipeline. */
sv_status sv_render_batched(sv_ctx *ctx, const int *views, int nv);
/* The viewer path on the training pipeline (in place of the instanced-quad
 * render of preprocess.wgsl:116-222 / gaussian.wgsl:15-67,
 * gaussian-renderer.ts:2217-2259: the tile lists and the tiled forward of
 * training-forward-tiles.wgsl:107-215): renders the current Gaussians from
 * any camera, not one of the training views, with the training forward's
 * rules (alpha <= 0.999, 1/255 cut, early stop, background). rgb: H x W x 3
 * floats; alpha: H x W accumulated alpha, or NULL. The training views, their
 * images and targets are untouched. */
sv_status sv_render_camera(sv_ctx *ctx, const sv_camera *cam, float *rgb, float *alpha);
/* No counterpart in the reference. sv_render_camera for callers that cannot
 * allocate device memory (serving loops that render a frame per request,
 * harnesses that forbid hipMalloc): the same image and alpha, but every
 * intermediate (projection records, pairs, sort buffers, tile ranges, pixel
 * planes) lives in one fixed 64 MiB scratch array that each device holds
 * from library load until the process ends; the call allocates nothing.
 * Limits: at most 65,536 Gaussians, 1,048,576 pixels, 8,192 tiles and
 * 1,048,576 (tile, splat) pairs; larger inputs return SV_E_CAPACITY (the
 * pair count is known only after projection: then the call fails after
 * running, with no output written). Every call on a device shares that
 * device's scratch, so calls that overlap on one device (contexts on
 * several host threads) race; calls in sequence are independent. Uses the
 * context's Gaussians, configuration and stream; rgb / alpha as in
 * sv_render_camera (host memory). */
sv_status sv_render_camera_noalloc(sv_ctx *ctx, const sv_camera *cam, float *rgb, float *alpha);

/* --- readbacks --------------------------------------------------------- */
sv_status sv_read_image(sv_ctx *ctx, int view, float *rgb);
sv_status sv_read_alpha(sv_ctx *ctx, int view, float *alpha);
/* last_ids (training-forward-tiles.wgsl:214); 0xffffffff in empty tiles. */
sv_status sv_read_last_ids(sv_ctx *ctx, int view, uint32_t *ids);
sv_status sv_read_residual(sv_ctx *ctx, int view, float *rgb);
/* Sorted (isect_ids, flatten_ids) of the last sv_build_tiles of view v,
 * with keys in the reference encoding (training-tiles.wgsl:182-199). */
sv_status sv_read_keys(sv_ctx *ctx, int view, uint32_t *keys, uint32_t *vals);
/* tile_offsets after the host fix-up (gaussian-renderer.ts:2014-2020):
 * tiles+1 entries. */
sv_status sv_read_tile_offsets(sv_ctx *ctx, int view, uint32_t *offsets);
/* GaussProjection records of view v (common.wgsl:57-65): 12 floats per
 * Gaussian (mean2d.xy, radius, depth, conic.abc, tiles_count as f32 bits,
 * position_world.xyz, opacity_logit). */
sv_status sv_read_projections(sv_ctx *ctx, int view, float *rec);
/* Gradient buffers; geom is grad_geom (5N: conic a,b,c, mean x,y) of the
 * last backward. NULL pointers are skipped. */
sv_status sv_read_grads(sv_ctx *ctx, float *pos, float *logscale, float *quat,
                        float *logit, float *sh, float *geom);
/* i32 fixed-point tile_grads of the last sv_backward_fixed (54 per
 * intersection, backward-reduce.wgsl:2-5). */
sv_status sv_read_tile_grads_fixed(sv_ctx *ctx, int view, int32_t *out);
/* ImportanceStats (importance-accum.wgsl:1-6) and grad_accum
 * (densify-accum.wgsl:19-46), 4 floats per Gaussian each. */
sv_status sv_read_stats(sv_ctx *ctx, float *importance4, float *grad_accum4);

/* --- prune / densify ------------------------------------------------- */
/* applyPruneMask (gaussian-renderer.ts:2917-3000): remove[i] != 0 drops
 * Gaussian i; params and moments are compacted in order on the device. */
sv_status sv_prune_mask(sv_ctx *ctx, const uint8_t *remove, size_t *n_removed);
/* Prune rule: drop Gaussians with sigmoid(logit) < min_opacity
 * (prune-decide.wgsl:178-180), decided and compacted on the device. */
sv_status sv_prune_opacity(sv_ctx *ctx, float min_opacity, size_t *n_removed);
/* resetOpacity (gaussian-renderer.ts:3867-3920): logits above
 * logit(max_opacity) are set to it, on the device; moments are kept. */
sv_status sv_reset_opacity(sv_ctx *ctx, float max_opacity, size_t *n_clamped);
/* initializeKnn (gaussian-renderer.ts:2131-2146, the knn_init shader
 * knn-init.wgsl:40-90, constants :77-79): for every Gaussian the mean of the
 * distances to its three nearest other Gaussians (positions rounded to f16;
 * 1e10 stands in for a missing neighbour), sigma = min(clamp(max(mean / 3,
 * 0.0015 e), 0.0007 e, 0.007 e), 0.1) with e = scene_extent, and all three
 * log-scales set to f16(log(max(sigma, 1e-6))); on the device, by an exact
 * grid search instead of the shader's all-pairs scan (same results). Moments
 * are kept (initializeKnnAndSync, :2159-2172, runs before training).
 * density: count floats (the shader's density_buffer) or NULL. */
sv_status sv_knn_init(sv_ctx *ctx, float scene_extent, float *density);
/* generateMultiViewScores (gaussian-renderer.ts:2754-2781) over the given
 * views, in order, on the device (metric-map.wgsl:32-147; the per-view
 * min / max / mean of computePixelLossStats, :2709-2745, reduced on the
 * device instead of read back): each view is rendered (one batched build),
 * pixels with (l1 - min) / (max - min) > 0.1 are flagged, and each splat
 * of a flagged pixel's tile list up to its last index with power in
 * [0, 6] counts once. importance[i] = sum of counts, pruning[i] = sum of
 * counts x the view's mean l1 (count floats each). The caller picks the
 * views (the reference samples up to 10 with Math.random). Views without
 * targets are skipped, as views without a finite metric. n_scored: views
 * scored (may be NULL). */
sv_status sv_multiview_scores(sv_ctx *ctx, const int *views, int nv, uint32_t *importance, float *pruning,
                              int *n_scored);
/* performDensification (gaussian-renderer.ts:3211-3466) with a seeded RNG:
 * actions[i] = 0 none, 1 clone, 2 split. Output order: non-split originals
 * in order, then children in parent order (clone: 1, split: 2); children
 * get zero moments. */
sv_status sv_densify(sv_ctx *ctx, const uint8_t *actions, uint64_t seed,
                     size_t *n_added);
/* Densify decision on the device, replacing the candidate selection of
 * performDensification (gaussian-renderer.ts:3249-3326) and the per-view
 * scoring of getMultiViewScores (:2640-2798): the views are rendered, each
 * tile's L1 error against its target is scored, each view's threshold is
 * its tile errors' `percentile` (value at floor(p (T - 1)) of the sorted
 * errors); a Gaussian listed on a tile above its view's threshold is
 * split (largest projected radius > split_px) or cloned. max_new > 0
 * applies the reference's growth cap (:3297-3306, which uses
 * max(1, floor(0.1 N))): only the max_new candidates of highest score keep
 * their action, ties in index order; a candidate's score is its largest
 * tile error over the views where it is listed divided by that view's
 * threshold (the error itself for a threshold <= 1e-8), as the reference
 * normalises its scores (:3280-3286). max_new = 0 keeps every candidate.
 * actions: N bytes (0 none, 1 clone, 2 split), ready for sv_densify. */
sv_status sv_densify_decide(sv_ctx *ctx, const int *views, int nviews, float percentile,
                            float split_px, size_t max_new, uint8_t *actions, size_t *n_clone,
                            size_t *n_split);
/* Thresholds of the reference trainer's densify rule
 * (DEFAULT_DENSIFICATION_CONFIG, densification.ts:26-33, and
 * MULTI_VIEW_IMPORTANCE_THRESHOLD, gaussian-renderer.ts:71), in double as
 * the JS computes them. */
typedef struct sv_densify_rule {
  double grad_threshold;       /* clone: mean |dL/dmean2d| >= this x pixel_scale (0.003) */
  double grad_abs_threshold;   /* split: mean |dL/dconic| >= this x pixel_scale (0.015) */
  double size_threshold;       /* world units: largest scale <= it takes the clone
                                  test, > it the split test (scene extent x 0.001,
                                  densification.ts:104-106) */
  double min_visibility;       /* visible steps required, max(1, .) (25) */
  double importance_threshold; /* importance / max(1, samples) must exceed it (5) */
  double pixel_scale;          /* training image width / 2 (:3251) */
} sv_densify_rule;
/* The candidate selection of performDensification as the reference trainer
 * runs it (gaussian-renderer.ts:3249-3326), decided on the device from the
 * step statistics accumulated since the last prune / densify (grad_accum,
 * densify-accum.wgsl: sum of |dL/dmean2d|, sum of |dL/dconic|, visible
 * steps) and a multi-view importance count (sv_multiview_scores over
 * `n_samples` views): a Gaussian with at least min_visibility visible
 * steps and importance / max(1, n_samples) above importance_threshold is a
 * clone candidate when its largest scale exp(log-scale) is at most
 * size_threshold and its mean 2-D gradient reaches grad_threshold x
 * pixel_scale, a split candidate when the scale is larger and its mean
 * conic gradient reaches grad_abs_threshold x pixel_scale; its score is
 * that mean over the threshold (the mean itself for a threshold <= 1e-8).
 * max_new > 0 keeps the max_new candidates of highest score (ties in
 * index order; the reference uses max(1, floor(0.1 N))); 0 keeps all.
 * importance: N counts (host). actions: N bytes (0 none, 1 clone, 2
 * split), ready for sv_densify. */
sv_status sv_densify_decide_grad(sv_ctx *ctx, const sv_densify_rule *rule, const uint32_t *importance,
                                 uint32_t n_samples, size_t max_new, uint8_t *actions, size_t *n_clone,
                                 size_t *n_split);
/* Per-tile L1 error of view `view` from the last sv_densify_decide
 * (tiles_x * tiles_y floats, row-major tiles). */
sv_status sv_read_tile_l1(sv_ctx *ctx, int view, float *out);

/* --- on-disk inputs (host only; errors via sv_ingest_error) ----------- */
/* Message of the last failed sv_ply_* / sv_colmap_* / sv_points_* call on
 * this thread. */
const char *sv_ingest_error(void);
/* decodeHeader (plyreader.ts:1-55): vertex count and SH degree (from the
 * number of f_rest_* properties, load.ts:44-62) of a binary little-endian
 * PLY. */
sv_status sv_ply_info(const char *path, size_t *n, uint32_t *sh_deg);
/* load (load.ts:14-206): 3DGS PLY (x, y, z, f_dc_*, f_rest_* channel-major,
 * opacity, scale_*, rot_0..3 = (w, x, y, z)) or a plain point cloud (red,
 * green, blue) into the library's arrays (quaternion (x, y, z, w), SH
 * coefficient-major with rgb inner, unused bands 0), f16-rounded as the
 * reference stores them; the reference's defaults fill what the file lacks.
 * scene_extent (may be NULL) gets the largest bounding-box side. */
sv_status sv_ply_load(const char *path, size_t n, float *pos, float *logscale, float *quat,
                      float *logit, float *sh, float *scene_extent);
/* Writes a 3DGS PLY of sh_deg (the format sv_ply_load reads). No
 * counterpart in the reference's training path. */
sv_status sv_ply_save(const char *path, size_t n, const float *pos, const float *logscale,
                      const float *quat, const float *logit, const float *sh, uint32_t sh_deg);
/* Image and point counts of a COLMAP sparse model directory (images.bin,
 * points3D.bin). */
sv_status sv_colmap_info(const char *sparse_dir, size_t *n_images, size_t *n_points);
/* loadColmapDatasetScene (colmap-dataset.ts:129-326): training cameras of
 * every image (view = [R | t] from qvec / tvec, view_inv its inverse, the
 * pinhole projection of training-camera.ts:29-52, viewport, focal) for
 * SIMPLE_PINHOLE / PINHOLE cameras, image names (name_len bytes each, may
 * be NULL) and the 3D points with colours. */
sv_status sv_colmap_load(const char *sparse_dir, size_t n_images, sv_camera *cams, char *names,
                         size_t name_len, size_t n_points, float *xyz, uint8_t *rgb);
/* createPointCloudFromXYZRGB (pointcloud-loader.ts:29-100): one Gaussian
 * per point, colour as the SH DC term, the reference's default opacity,
 * scale and rotation, f16-rounded. */
sv_status sv_points_to_gaussians(size_t n, const float *xyz, const uint8_t *rgb, float *pos,
                                 float *logscale, float *quat, float *logit, float *sh,
                                 float *scene_extent);

/* --- checkpoint / resume ------------------------------------------- */
/* No counterpart in the reference (its nearest is the CPU sync of
 * gaussian-renderer.ts:3129-3205). Writes the training state to `path`:
 * header ("SVCKPT01", version, N, step, sv_config), then parameters, the
 * two Adam moments (59 planes of N floats each), the importance and
 * densify statistics (4 floats per Gaussian), then an FNV-1a 64 checksum
 * of everything before it. The random draws of densify come from caller
 * seeds, so there is no generator state to save. */
sv_status sv_save_checkpoint(sv_ctx *ctx, const char *path);
/* Restores a checkpoint written by sv_save_checkpoint; a file holding more
 * Gaussians than the context's capacity grows it to the next power of two
 * (as sv_densify does). Views and targets are not part of it. */
sv_status sv_load_checkpoint(sv_ctx *ctx, const char *path);

/* --- autotuner ------------------------------------------------------ */
/* No counterpart in the reference (its tile and batch sizes are
 * constants: 16x16 tiles, 64-entry batches, backward-tiles-local.wgsl:49-53,
 * 185). Times `steps` batched steps of `views` (after one untimed step) for
 * every candidate raster tile 8/16/32 x LDS batch 64/128/256 x XCD tile
 * mapping 0/1 (and persistent launch 0/1 with 16-px raster tiles), then,
 * with two views or more, the fastest of those with split_backward
 * switched, restoring parameters, moments, statistics and the step counter
 * before each candidate and at the end, and leaves the fastest candidate in
 * the configuration. The binning tile stays 16 (gaussian-renderer.ts:366): it
 * decides which fringe pixels (power between the 3-sigma radius and
 * ln(255 a)) a splat is composited into; the raster tile only regroups the
 * same per-pixel work, so every candidate renders the same image.
 * out: up to *count entries (NULL to query); *count receives the number of
 * candidates. */
sv_status sv_autotune(sv_ctx *ctx, const int *views, int nviews, int steps,
                      sv_tune_entry *out, int *count);

/* --- device memory ------------------------------------------------- */
/* No counterpart in the reference (WebGPU buffer sizes are not reported).
 * Device memory the context holds now (its parameter, moment, gradient,
 * pixel and tile-build buffers) and the most it has held since sv_create
 * or the last sv_reset_memory_peak. Either pointer may be NULL. */
sv_status sv_memory_stats(const sv_ctx *ctx, uint64_t *device_bytes, uint64_t *peak_bytes);
sv_status sv_reset_memory_peak(sv_ctx *ctx);

/* --- timing --------------------------------------------------------- */
/* Enable per-kernel hipEvent timing on the context stream. */
sv_status sv_enable_timing(sv_ctx *ctx, int enable);
/* Copy up to *count records; *count receives the number available. */
sv_status sv_timings(sv_ctx *ctx, sv_kernel_time *out, int *count);
sv_status sv_reset_timings(sv_ctx *ctx);
/* Block until the context stream is idle. */
sv_status sv_sync(sv_ctx *ctx);

/* --- seeded synthetic scenes (harness; host only, no device) ---------- */
/* n Gaussians: means U[-h,h]^3, log-scales N(-4.5,0.3), quaternion
 * normalize(N(0,1)^4) stored (x,y,z,w), logit N(0,1), SH DC N(0.5,0.2),
 * other coefficients N(0,0.05); every value rounded to f16 (RNE). */
sv_status sv_scene_gaussians(uint64_t seed, size_t n, float half_extent,
                             float *pos, float *logscale, float *quat,
                             float *logit, float *sh);
/* pos += N(0, sigma) per component (then f16-rounded). */
sv_status sv_scene_jitter(uint64_t seed, size_t n, float sigma, float *pos);
/* Pinhole camera at eye looking at target, world up +y (+x if parallel),
 * camera +z forward, fx = fy = (H/2)/tan(fovy/2), principal point at the
 * image centre; proj = get_pinhole_projection_matrix(0.01, 4e6, ...)
 * (training-camera.ts:29-52). */
sv_status sv_scene_camera(const float eye[3], const float target[3], uint32_t w,
                          uint32_t h, float fovy_deg, sv_camera *out);
/* Seeded uniform and normal draws (SplitMix64 + Box-Muller). */
double sv_scene_uniform(uint64_t seed, uint64_t counter);
double sv_scene_normal(uint64_t seed, uint64_t counter);

#ifdef __cplusplus
}
#endif
#endif
