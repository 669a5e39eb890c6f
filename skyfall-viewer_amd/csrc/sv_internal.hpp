// sv_internal.hpp -- launch wrappers exported by the .hip translation units
// to the context code in sv_api.cpp. All launches go to the given stream.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sv_device.hpp"

namespace sv {

// sv_project.hip --------------------------------------------------------
// Projection + SH colour + tile count for nv views per Gaussian
// (project_gaussians_for_tiling, training-tiles.wgsl:39-180, with the SH
// colour of training-forward-tiles.wgsl:65-105 evaluated once per view).
// rec: 64-byte record (4 float4) per (slot, Gaussian); aux: (depth,
// radius); counts: u32 per (slot, Gaussian).
// Also the first level of the scans over (slot, Gaussian): offs_loc /
// soffs_loc = exclusive prefix of the tile / supertile counts within the
// projection block (256 Gaussians of one slot; q3.x holds the tile one) and
// btot / sbtot = the block totals, slot-major (project_blocks(N) per slot).
// A (slot, Gaussian)'s first position = scan of the totals at its block +
// its local prefix (block_pos).
inline uint32_t project_blocks(uint32_t N) { return (N + 255u) / 256u; }
void launch_project(hipStream_t st, const float *P, size_t cap, uint32_t N, const DevView *views,
                    int nv, const DevConfig &cfg, float4 *rec, float2 *aux, uint32_t *counts,
                    uint32_t *max_dq, uint32_t *scounts, uint2 *rects, uint32_t *tins, uint32_t *offs_loc,
                    uint32_t *soffs_loc, uint32_t *btot, uint32_t *sbtot);

// Tight tile rectangle of a (view, Gaussian): the tiles of its rectangle
// whose pixel centres come within one pixel plus its extent (ext_x, ext_y)
// of the mean. Every tile where a quadrant test can pass lies inside, so
// the backward writes, and the record gather reads, adjoint records only
// there. Stored as four 8-bit insets from the rectangle's sides (left | top
// << 8 | right << 16 | bottom << 24); TIGHT_EMPTY (all insets 255) leaves
// nothing, as rectangles are at most 66 tiles wide.
constexpr uint32_t TIGHT_EMPTY = 0xffffffffu;
__device__ __forceinline__ uint32_t tight_insets(float m, float e, float mm, float ee, uint32_t x0, uint32_t y0,
                                                 uint32_t x1, uint32_t y1) {
  if (!(e >= 0.0f && ee >= 0.0f)) return TIGHT_EMPTY;
  // tile t is kept when 16 t + 0.5 - 1 <= m + e and 16 t + 15.5 + 1 >= m - e
  const auto lo = [](float c, float r) {
    return (int32_t)ceilf(fminf(fmaxf((c - r - 16.5f) * 0.0625f, -1e6f), 1e6f));
  };
  const auto hi = [](float c, float r) {
    return (int32_t)floorf(fminf(fmaxf((c + r + 0.5f) * 0.0625f, -1e6f), 1e6f)) + 1;
  };
  const int32_t tx0 = max(lo(m, e), (int32_t)x0), tx1 = min(hi(m, e), (int32_t)x1);
  const int32_t ty0 = max(lo(mm, ee), (int32_t)y0), ty1 = min(hi(mm, ee), (int32_t)y1);
  if (tx0 >= tx1 || ty0 >= ty1) return TIGHT_EMPTY;
  const auto c8 = [](int32_t v) { return (uint32_t)min(v, 254); };
  return c8(tx0 - (int32_t)x0) | (c8(ty0 - (int32_t)y0) << 8) | (c8((int32_t)x1 - tx1) << 16) |
         (c8((int32_t)y1 - ty1) << 24);
}
// Bounds [tx0, tx1) x [ty0, ty1) of the tight rectangle (empty: tx0 >= tx1).
struct TightRect {
  int32_t tx0, ty0, tx1, ty1;
};
__device__ __forceinline__ TightRect tight_rect(uint32_t ins, uint32_t x0, uint32_t y0, uint32_t x1, uint32_t y1) {
  TightRect t;
  t.tx0 = (int32_t)x0 + (int32_t)(ins & 255u);
  t.ty0 = (int32_t)y0 + (int32_t)((ins >> 8) & 255u);
  t.tx1 = (int32_t)x1 - (int32_t)((ins >> 16) & 255u);
  t.ty1 = (int32_t)y1 - (int32_t)(ins >> 24);
  return t;
}
__device__ __forceinline__ uint32_t block_pos(const uint32_t *bbase, uint32_t nb, uint32_t s, uint32_t g,
                                              uint32_t local) {
  return bbase[(size_t)s * nb + (g >> 8)] + local;
}
// Exclusive scan of u32 counts (M entries) into u32 offsets; total (u64)
// written to *total. tmp must hold scan_tmp_words(M) u64 words.
size_t scan_tmp_words(size_t M);
void launch_scan(hipStream_t st, const uint32_t *counts, uint32_t *offs, size_t M, uint64_t *tmp,
                 uint64_t *total);
// Derive dbits, key bits, sort passes, result buffer and overflow from
// the scan total and the largest depth quantum.
void launch_keyinfo(hipStream_t st, StepMeta *m, uint32_t tile_bits, uint32_t key_type_bits, uint64_t cap);
// emit_intersections (training-tiles.wgsl:201-247) for nv slots.
// rects: per (slot, Gaussian) tile rectangle (x0 | y0 << 16, x1 | y1 << 16);
// offs / bbase: local prefixes and scanned block totals of the tile counts.
template <typename K>
void launch_emit(hipStream_t st, float4 *rec, const uint2 *rects, const float2 *aux, const uint32_t *counts,
                 const uint32_t *offs, const uint32_t *bbase, uint32_t N, int nv, const DevView *views,
                 const StepMeta *m, K *keys, uint32_t *vals);
// Tile ranges = lower_bound per global tile (build_tile_offsets +
// host fix-up, training-tiles.wgsl:249-271, gaussian-renderer.ts:2014-2020).
// Grid sized for the capacity; n, dbits and the result buffer from m.
template <typename K>
void launch_ranges(hipStream_t st, const K *keys0, const K *keys1, uint64_t cap, uint32_t n_tiles,
                   const StepMeta *m, uint32_t *ranges);
// Binned build (sv_bin.hip): pass info of the supertile pairs from m[0]
// and m[1].n_isect; supertile pairs of each item; tile counts, ranges and
// tile lists from the sorted supertile pairs (m[2], m[3]: scan totals).
void launch_bin_info(hipStream_t st, StepMeta *m, uint32_t super_bits, uint64_t cap_super);
void launch_emit_super(hipStream_t st, const uint2 *rects, const float2 *aux, const uint32_t *scounts,
                       const uint32_t *soffs, const uint32_t *sbbase, uint32_t N, int nv, const DevView *views,
                       const StepMeta *m, uint32_t *keys, uint32_t *vals);
struct BinArgs {
  const uint32_t *keys0, *vals0, *keys1, *vals1;  // supertile pairs, sorted into buffer m[1].sel
  const uint32_t *sranges;      // per global supertile + 1: lower bound in the sorted pairs
  const uint32_t *super_view;   // global supertile -> slot
  const DevView *views;
  StepMeta *m;                  // m[0..3]
  uint32_t n_super_tiles, n_tiles;
  size_t chunk_cap;             // bin_chunk_cap(super pair capacity, n_super_tiles)
  uint32_t *nch, *cbase;        // per supertile + 1: chunks, first chunk
  uint32_t *chunk_super;        // per chunk: supertile
  uint32_t *ccount;             // per chunk x 64 tiles
  uint32_t *tcount;             // per global tile + 1
  uint64_t *scan_tmp;           // scan_tmp_words(max(n_tiles, n_super_tiles) + 1)
  uint32_t *ranges, *out;       // tile ranges (n_tiles + 1) and lists (pair capacity)
};
size_t bin_chunk_cap(uint64_t super_cap, uint32_t n_super_tiles);
void launch_fill_tiles(hipStream_t st, const BinArgs &b);
// Residual of training_loss (training-loss.wgsl:86-126) for one slot.
void launch_loss(hipStream_t st, const DevView &hv, const DevConfig &cfg, const float *color,
                 const float *target, float *residual);

// sv_sort.hip -----------------------------------------------------------
// Stable LSD radix sort of (key, value) pairs, 8-bit digits. Launches
// max_passes passes sized for `cap` pairs; each reads n, the key width and
// the number of passes from m and exits when not needed. Pass p reads
// buffer p % 2 and writes buffer (p + 1) % 2; the result is in buffer
// m->sel.
template <typename K>
size_t sort_tmp_words(uint64_t n);
template <typename K>
void launch_sort(hipStream_t st, K *keys0, uint32_t *vals0, K *keys1, uint32_t *vals1, uint64_t cap,
                 int max_passes, uint32_t *tmp, const StepMeta *m);

// sv_raster.hip ---------------------------------------------------------
struct RasterArgs {
  const DevView *views;
  const uint32_t *tile_view;  // global tile -> slot
  const uint32_t *ranges;     // n_tiles_total + 1
  const uint32_t *vals0;      // sorted Gaussian ids: buffer 0 ...
  const uint32_t *vals1;      // ... or buffer 1, per meta->sel
  const StepMeta *meta;       // intersections, result buffer, overflow
  // set on the device from meta at kernel entry (resolve_meta)
  const uint32_t *vals;
  uint64_t n_isect;
  const float4 *rec;          // 64-byte record per (slot, Gaussian)
  uint32_t N;                 // Gaussians (record stride per slot)
  uint32_t n_tiles_total;     // end of this launch's tiles (global tile ids)
  uint32_t tile_begin;        // first tile of this launch (0: all tiles)
  DevConfig cfg;
  // forward outputs / backward inputs (pixel arrays of the batch)
  float *color;     // 3 planes per view
  float *alpha;
  uint32_t *last;
  const float *target;    // 3 planes per view
  const float *residual;  // 3 planes per view, or null for the fused loss
  // backward outputs
  const uint32_t *bbase;  // scanned block totals of the tile counts (emit positions)
  uint32_t nb;            // projection blocks per slot
  float *isect_rec;       // 9-float adjoint record per intersection (emit order)
  // per intersection (emit order): 1 where the backward wrote a record,
  // 0 (cleared before the backward) elsewhere
  uint8_t *rflag;
  // per sorted entry: the quadrants of its tile (bit h + 2 v) where some
  // pixel composited the entry. The forward writes it for every entry it
  // loads (null: not written); the backward, when not null, visits only
  // the entries and quadrants it names instead of culling.
  uint8_t *tmask;
  const float2 *aux;      // per (slot, Gaussian) depth and radius
  float *l1_partial;      // per-tile sum |pred - target| (may be null)
  uint32_t *work_counter; // persistent launch: NUM_WORK_QUEUES counters,
                          // WORK_QUEUE_STRIDE words apart (null: one block per tile)
  // 32-px raster tiles: blocks of 2 x 2 tiles, global block -> slot
  const uint32_t *block_view;
  uint32_t n_blocks_total;
  // 16-px raster tiles: work item i of the launch renders tile tile_begin +
  // tile_order[i] (a permutation of 0 .. n - 1, longest list first; null:
  // tile tile_begin + i)
  const uint32_t *tile_order;
  // forward output (may be null): per tile, the entries of its list some
  // pixel visits (the backward's work), written by a wave that renders the
  // whole tile
  uint32_t *tile_cost;
};
constexpr uint32_t NUM_WORK_QUEUES = 8;     // one per XCD
constexpr uint32_t WORK_QUEUE_STRIDE = 32;  // words between counters (a 128-byte line each)
// Workgroups of the forward / backward raster kernel (LDS batch `batch`)
// resident on the whole device at once: the persistent grid. 0 if unknown.
uint32_t raster_resident_blocks(bool backward, uint32_t batch);
// raster_px: 16 = one wave per 16x16 tile (grid_blocks: tiles or the
// persistent grid); 8 = one wave per 8x8 quadrant; 32 = four waves per 2x2
// block of tiles. The lists are the 16-px tiles' in every case (same images).
void launch_forward(hipStream_t st, const RasterArgs &a, uint32_t batch, uint32_t grid_blocks, uint32_t raster_px);
void launch_backward(hipStream_t st, const RasterArgs &a, uint32_t batch, uint32_t grid_blocks);
// The longest-first work order of a 16-px raster launch over tiles
// [tile_begin, tile_begin + n), within the items each XCD takes: contiguous
// eighths of the launch (xcd_remap, the persistent queues) or, with
// `contiguous` false, the items i = x mod 8 of XCD x. A tile's work is
// cost[t] when cost is given (the forward's visited entries), else its list
// length.
void launch_tile_order(hipStream_t st, const uint32_t *ranges, const uint32_t *cost, const StepMeta *meta,
                       uint32_t tile_begin, uint32_t n, bool contiguous, uint32_t *order);
// backward_tiles_local (backward-tiles-local.wgsl:185-382) for one slot:
// i32 sums per (intersection, slot of 54) in sorted order; inv[emit
// position] = sorted position of every intersection.
void launch_backward_fixed(hipStream_t st, const RasterArgs &a, const float *P, size_t cap,
                           uint32_t flags, int32_t *fixed, uint32_t *inv);

// sv_optim.hip ----------------------------------------------------------
struct VjpArgs {
  const float *P;  // parameter planes
  size_t cap;
  uint32_t N;
  const DevView *views;
  int nv;
  int slot_begin, slot_count;  // the slots the record gather covers
  DevConfig cfg;
  const uint32_t *counts;  // per (slot, Gaussian)
  const uint32_t *offs;    // local prefix of the tile counts in the projection block
  const uint32_t *bbase;   // scanned block totals: first emit position = block_pos(...)
  uint32_t nb;
  const float4 *rec;       // projection records
  const uint2 *rects;      // tile rectangle per record
  const uint32_t *tins;    // tight-rectangle insets per record (tight_insets)
  const float2 *aux;       // (depth, radius) per record (radius for the importance stats)
  const float *isect_rec;  // 9-float adjoint record per intersection
  const uint8_t *rflag;    // per intersection: 1 where the backward wrote a record
  float *vrec;             // 9 planes of nv * N: per (slot, Gaussian) sum of its reached records
  float4 *importance;      // importance statistics (may be null)
  float *G;                // gradient planes (accumulate), or null when fused
  float *geom;             // 5 planes: grad_geom of the last slot (may be null)
  float4 *grad_accum;      // densify statistics (may be null)
  const StepMeta *meta;    // overflow: no update
  // fused Adam (when G == null)
  float *Pw;               // parameters to update (== P)
  float *M1, *M2;
  const AdamParams *adam;
};
// Per (slot, Gaussian): sum of the adjoint records of its reached pairs
// into vrec (one thread per item, so the dependent record walks of all
// items overlap); then the per-Gaussian VJP reads 9 coalesced planes.
void launch_gather_records(hipStream_t st, const VjpArgs &a);
void launch_vjp(hipStream_t st, const VjpArgs &a);
void launch_adam(hipStream_t st, float *P, float *M1, float *M2, const float *G, size_t cap,
                 uint32_t N, const AdamParams *adam);
// Copy the step's Adam parameters to the device: from mapped host memory
// (src, re-read on every graph replay) or, with src null, from `value`.
void launch_load_adam(hipStream_t st, const AdamParams &value, const AdamParams *src, AdamParams *dst);
// backward_reduce (backward-reduce.wgsl:35-71) + geometry backward input
// for one slot: one thread per Gaussian adds its intersections' values in
// sorted order (the Gaussian's tiles in ascending order, found through inv
// from its emit positions), so the sums do not depend on scheduling.
void launch_reduce_fixed(hipStream_t st, const int32_t *fixed, const uint32_t *inv, const uint32_t *counts,
                         const uint32_t *offs, const uint32_t *bbase, uint32_t N, const StepMeta *m, size_t cap,
                         float *G, float *geom);
// training_backward_geom (training-backward-geom.wgsl:36-299) from the
// per-Gaussian grad_geom planes of one view.
// Also accumulate_importance (importance-accum.wgsl:19-40) for that view
// from its records and counts.
void launch_geom(hipStream_t st, const float *P, size_t cap, uint32_t N, const DevView *view,
                 const DevConfig &cfg, const float *geom, float *G, float4 *grad_accum,
                 const float2 *aux, const uint32_t *counts, float4 *importance);
// Stable compaction of rows by keep flags (exclusive scan of flags).
void launch_compact(hipStream_t st, const float *src, float *dst, size_t cap_src, size_t cap_dst,
                    int planes, const uint32_t *keep_offs, const uint8_t *keep, uint32_t N);
void launch_prune_flags(hipStream_t st, const float *P, size_t cap, uint32_t N, float min_opacity,
                        uint8_t *keep, uint32_t *keep_u32);
// Clamp logits above max_logit; adds the number clamped to *n_clamped.
void launch_clamp_logit(hipStream_t st, float *P, size_t cap, uint32_t N, float max_logit, uint32_t *n_clamped);
// Per-tile L1 error of every tile of the batch (scoring for densify).
void launch_tile_l1(hipStream_t st, const DevView *views, const uint32_t *tile_view, uint32_t n_tiles,
                    const float *color, const float *target, float *tile_l1);
// Densify rule from the batch's records, tile L1 and per-view thresholds;
// score (may be null) = the candidates' normalised tile error.
void launch_densify_decide(hipStream_t st, const float4 *rec, const float2 *aux, const uint32_t *counts, uint32_t N,
                           const DevView *views, int nv, const float *tile_l1, const float *thr,
                           float split_px, uint8_t *actions, float *score);
// The reference trainer's densify rule from the step statistics; the
// thresholds already multiplied out in double (grad_px = grad_threshold x
// pixel_scale, and so on) as the JS forms them.
struct DensifyRuleDev {
  double grad_px, abs_px, size, min_vis, imp_threshold, samples;
};
void launch_densify_decide_grad(hipStream_t st, const float4 *grad_accum, const float *P, size_t cap, uint32_t N,
                                const uint32_t *importance, const DensifyRuleDev &r, uint8_t *actions,
                                double *score);
void launch_densify(hipStream_t st, const float *P, const float *M1, const float *M2, size_t cap,
                    uint32_t N, const uint8_t *actions, const uint32_t *keep_offs,
                    const uint32_t *child_offs, uint32_t n_keep, uint64_t seed, float *Pn,
                    float *M1n, float *M2n, size_t cap_n);
void launch_u8_to_u32(hipStream_t st, const uint8_t *a, uint32_t *o, uint32_t N, int mode);
void launch_fill_u32(hipStream_t st, uint32_t *p, uint32_t v, size_t n);
void launch_fill_f32(hipStream_t st, float *p, float v, size_t n);
// Zero the first m->n_isect bytes of a per-intersection flag array of cap
// bytes, rounded up to whole 16-byte words (the array holds
// flag_array_bytes(cap)).
inline size_t flag_array_bytes(uint64_t cap) { return (size_t)((cap + 15) / 16 * 16); }
void launch_clear_flags(hipStream_t st, uint8_t *flags, uint64_t cap, const StepMeta *m);

// sv_score.hip ----------------------------------------------------------
// Multi-view scoring (generateMultiViewScores) of the rendered batch.
struct MetricStats {
  float min, max, inv_range;
  uint32_t valid;  // the view has a finite metric value
  double mean;
};
void launch_multiview_scores(hipStream_t st, const DevView *views, int nv, uint32_t max_pixels,
                             const float *color, const float *target, const uint32_t *ranges, const uint32_t *vals,
                             const uint32_t *last, const float4 *rec, uint32_t N, float *loss, MetricStats *stats,
                             uint32_t *counts, uint32_t *importance, float *pruning);

// sv_knn.hip ------------------------------------------------------------
// knn_init (knn-init.wgsl:40-90) by an exact grid search: cells per axis
// for N points, the cell count, and the device scratch of one call.
uint32_t knn_grid_dim(uint32_t N);
size_t knn_grid_cells(uint32_t N);
struct KnnScratch {
  void *grid;           // 64 bytes: grid box, cell size, cells per axis
  uint32_t *cnt;        // per cell
  uint32_t *start;      // per cell
  uint32_t *cell;       // per point
  uint32_t *rank;       // per point
  float4 *spos;         // per point, in cell order: xyz (f16-rounded), id
  uint64_t *scan_tmp;   // scan_tmp_words(cells)
  uint64_t *total;
  float *density;       // per point
};
// Writes the three log-scale planes of P and w.density.
hipError_t launch_knn_init(hipStream_t st, float *P, size_t cap, uint32_t N, float extent, const KnnScratch &w);

// sv_noalloc.hip --------------------------------------------------------
// Fixed device scratch of sv_render_camera_noalloc (one copy per device, held
// until the process ends) and the largest render it serves.
constexpr size_t NOALLOC_SCRATCH_BYTES = 64ull << 20;
constexpr uint32_t NOALLOC_MAX_SPLATS = 1u << 16;
constexpr uint32_t NOALLOC_MAX_PAIRS = 1u << 20;
constexpr uint32_t NOALLOC_MAX_PIXELS = 1u << 20;
constexpr uint32_t NOALLOC_MAX_TILES = 8192;
// Device address of the scratch on `device`, which must be current;
// resolved once per device (hipGetSymbolAddress) and kept.
hipError_t noalloc_scratch(int device, void **out);
// Colour planes (c * P + i) -> interleaved rgb (3 i + c).
void launch_interleave3(hipStream_t st, const float *planes, size_t P, float *out);

}  // namespace sv
