// sv_api.cpp -- context, device memory and step orchestration behind the
// C ABI of include/sv_api.h. One context = one HIP stream on one device.
//
// Device memory layout (all planar, one plane per scalar):
//   P, M1, M2, G      59 planes x cap floats (parameters, Adam moments, grads)
//   geom              5 planes x cap (grad_geom of the last backward)
//   importance, grad_accum   float4 x cap (prune/densify statistics)
//   pixel arena       every view's pixels back to back (view v at pix_off):
//                     colour, residual, target as 3 planes of P_v floats,
//                     alpha and last index as 1 plane
//   batch arrays      per (slot, Gaussian): 4 record planes, count, offset;
//                     per intersection: key, value (ping-pong), 9-float
//                     adjoint record; per global tile: range start, slot
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <unordered_map>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "../../include/sv_api.h"
#include "sv_internal.hpp"

using namespace sv;

namespace {

struct HostView {
  sv_camera cam;
  uint32_t W, H, tiles_x, tiles_y, n_tiles;
  uint64_t pix_off;
  bool has_target;
};

struct TimedPhase {
  int id;
  hipEvent_t a, b;
};

const char *kPhaseNames[] = {"project", "scan", "emit", "sort", "ranges", "forward", "backward",
                             "vjp_adam", "adam", "loss", "backward_fixed", "compact", "readback_sync",
                             "fill"};
constexpr int kNumPhases = sizeof(kPhaseNames) / sizeof(kPhaseNames[0]);
enum Phase {
  PH_PROJECT,
  PH_SCAN,
  PH_EMIT,
  PH_SORT,
  PH_RANGES,
  PH_FORWARD,
  PH_BACKWARD,
  PH_VJP,
  PH_ADAM,
  PH_LOSS,
  PH_FIXED,
  PH_COMPACT,
  PH_SYNC,
  PH_FILL
};

uint32_t bitlen(uint64_t v) {
  uint32_t b = 0;
  while (v) {
    ++b;
    v >>= 1;
  }
  return b;
}

}  // namespace

struct sv_ctx {
  int device = 0;
  hipStream_t st = nullptr;
  // the split backward: the first views' record gather runs on st2 (forked
  // and joined with the two events) while the other views' backward runs
  hipStream_t st2 = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  size_t cap = 0;
  uint32_t N = 0;
  sv_config cfg{};
  std::string err;
  uint32_t step_t = 0;

  float *P = nullptr, *M1 = nullptr, *M2 = nullptr, *G = nullptr, *geom = nullptr;
  float4 *importance = nullptr, *grad_accum = nullptr;

  std::vector<HostView> views;
  uint64_t pix_total = 0;
  float *color = nullptr, *alpha = nullptr, *residual = nullptr, *target = nullptr;
  uint32_t *last = nullptr;
  std::vector<uint8_t> residual_ready;

  // batch state
  std::vector<int> slot_views;  // views of the current batch
  std::vector<DevView> h_slots;
  DevView *d_views = nullptr;
  size_t d_views_cap = 0;
  uint32_t tiles_total = 0;
  uint32_t *tile_view = nullptr, *ranges = nullptr;
  // longest-first work orders of the forward and of the backward launches
  uint32_t *order_f = nullptr, *order_b = nullptr;
  // per tile: the entries some pixel visited in the last forward
  uint32_t *tile_cost = nullptr;
  // 2x2-tile blocks of the batch (32-px raster tiles): count, block -> slot
  uint32_t blocks_total = 0;
  uint32_t *block_view = nullptr;
  size_t blocks_cap = 0;
  float *l1_partial = nullptr;
  float *tile_l1 = nullptr;  // densify scoring, per global tile
  size_t tiles_cap = 0;
  std::vector<int> scored_views;  // batch of the last sv_densify_decide
  std::vector<uint32_t> scored_base;
  float4 *rec = nullptr;
  float2 *aux = nullptr;
  uint32_t *counts = nullptr, *offs = nullptr;
  uint32_t *scounts = nullptr, *soffs = nullptr;  // supertile pairs per (slot, Gaussian)
  uint2 *rects = nullptr;  // tile rectangle per (slot, Gaussian)
  uint32_t *tins = nullptr;  // tight-rectangle insets per (slot, Gaussian)
  // projection-block totals of the tile / supertile counts and their scans
  uint32_t *btot = nullptr, *sbtot = nullptr, *bbase = nullptr, *sbbase = nullptr;
  size_t btab_cap = 0;
  float *vrec = nullptr;  // 9 planes of slots x N: reached-record sums for the VJP
  size_t vg_cap = 0;
  uint64_t *scan_tmp = nullptr;
  size_t scan_tmp_cap = 0;
  uint64_t *d_small = nullptr;  // scan totals of prune / densify
  StepMeta *d_meta = nullptr;   // device-side sizes of the current tile build
  StepMeta *h_meta = nullptr;   // pinned copy read after a build
  AdamParams *d_adam = nullptr;
  AdamParams *h_adam = nullptr;      // pinned, mapped: written before a graph replay
  AdamParams *h_adam_dev = nullptr;  // its device address
  void *keys = nullptr, *keys2 = nullptr;
  uint32_t *vals = nullptr, *vals2 = nullptr;
  uint64_t isect_cap = 0;
  uint32_t *sort_tmp = nullptr;
  size_t sort_tmp_cap = 0;
  float *isect_rec = nullptr;
  uint8_t *rflag = nullptr;  // per intersection (emit order): the backward wrote its record
  uint8_t *tmask = nullptr;  // per sorted entry: the quadrants that took it in the forward
  bool tmask_ready = false;  // the last forward wrote tmask
  bool cost_ready = false;  // the last forward wrote tile_cost
  int32_t *fixed = nullptr;
  uint32_t *fixed_inv = nullptr;  // emit position -> sorted position (plain mode)
  uint64_t fixed_cap = 0;
  uint64_t *h_small = nullptr;  // pinned

  // binned build: supertiles of the batch, pair buffers, difference array
  uint32_t super_total = 0, super_bits = 0, max_tiles_x = 0, max_tiles_y = 0;
  uint32_t *super_view = nullptr, *sranges = nullptr;
  size_t super_tab_cap = 0;
  uint32_t *diff = nullptr;  // per global tile + 1: tile counts
  uint32_t *nch = nullptr, *cbase = nullptr;  // per supertile + 1
  uint32_t *skeys = nullptr, *skeys2 = nullptr, *svals = nullptr, *svals2 = nullptr, *ssort_tmp = nullptr;
  uint64_t super_cap = 0;
  uint32_t *chunk_super = nullptr, *ccount = nullptr;
  size_t chunk_cap = 0;
  int max_super_passes = 0;
  bool binned = false;     // the next build bins
  bool bin_unfit = false;  // a build found supertile key + rectangle wider than 32 bits
  bool last_binned = false;  // the last completed build binned
  uint32_t spasses = 0;

  // key plan of the batch (host bounds) and the last tile build (read back)
  uint32_t tile_bits = 0, max_key_bits = 0;
  bool key64 = false;
  int max_passes = 0;
  bool built = false;
  bool forward_done = false;
  uint64_t I = 0;
  uint32_t dbits = 0, key_bits = 0;
  int passes = 0;
  int sel = 0;  // sorted pairs in (keys, vals) = 0 or (keys2, vals2) = 1

  // viewer render target (sv_render_camera)
  float *vw_color = nullptr, *vw_alpha = nullptr;
  uint32_t *vw_last = nullptr;
  size_t vw_cap = 0;

  // device address of the fixed scratch of sv_render_camera_noalloc
  unsigned char *noalloc = nullptr;
  // persistent raster launches: per-XCD work counters and the resident
  // workgroups of [forward, backward] x LDS batch [64, 128, 256]
  uint32_t *d_work = nullptr;
  uint32_t resident[2][3] = {{0, 0, 0}, {0, 0, 0}};

  // captured batched step
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  std::vector<uint64_t> gsig;
  bool capturing = false;

  // device memory held by the context (dalloc / dfree) and its peak
  uint64_t dev_bytes = 0, dev_peak = 0;

  // timing
  bool timing = false;
  std::vector<TimedPhase> pending;
  std::vector<hipEvent_t> event_pool;
  double phase_ms[kNumPhases] = {0};
  uint64_t phase_n[kNumPhases] = {0};
};

namespace {

thread_local std::string g_create_err;

sv_status fail(sv_ctx *c, sv_status s, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if (c) c->err = buf;
  else g_create_err = buf;
  return s;
}

#define HIP_OK(c, expr)                                                                   \
  do {                                                                                    \
    hipError_t _e = (expr);                                                               \
    if (_e != hipSuccess)                                                                 \
      return fail((c), SV_E_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), \
                  __FILE__, __LINE__);                                                    \
  } while (0)

sv_status check_launch(sv_ctx *c, const char *what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(c, SV_E_HIP, "%s: %s", what, hipGetErrorString(e));
  return SV_OK;
}

// Every device allocation of a context, so that sv_memory_stats can report
// what the context holds (allocation -> owning context, bytes).
std::mutex g_alloc_mu;
std::unordered_map<const void *, std::pair<sv_ctx *, size_t>> g_allocs;

template <typename T>
sv_status dalloc(sv_ctx *c, T **p, size_t n) {
  *p = nullptr;
  if (n == 0) n = 1;
  hipError_t e = hipMalloc((void **)p, n * sizeof(T));
  if (e != hipSuccess) {
    *p = nullptr;
    (void)hipGetLastError();
    return fail(c, SV_E_NOMEM, "hipMalloc of %zu bytes failed: %s", n * sizeof(T), hipGetErrorString(e));
  }
  if (c) {
    std::lock_guard<std::mutex> lock(g_alloc_mu);
    g_allocs[(const void *)*p] = {c, n * sizeof(T)};
    c->dev_bytes += n * sizeof(T);
    c->dev_peak = std::max(c->dev_peak, c->dev_bytes);
  }
  return SV_OK;
}

template <typename T>
void dfree(T *&p) {
  if (p) {
    {
      std::lock_guard<std::mutex> lock(g_alloc_mu);
      auto it = g_allocs.find((const void *)p);
      if (it != g_allocs.end()) {
        it->second.first->dev_bytes -= it->second.second;
        g_allocs.erase(it);
      }
    }
    (void)hipFree((void *)p);
  }
  p = nullptr;
}

// Grow a device array to at least n elements (contents are not kept).
template <typename T>
sv_status grow(sv_ctx *c, T **p, size_t &capn, size_t n) {
  if (n <= capn && *p) return SV_OK;
  dfree(*p);
  size_t want = std::max(n, capn + capn / 4);
  sv_status s = dalloc(c, p, want);
  if (s) {
    capn = 0;
    return s;
  }
  capn = want;
  return SV_OK;
}

void ph_begin(sv_ctx *c, int id, TimedPhase *tp) {
  tp->id = -1;
  if (!c->timing || c->capturing) return;
  if (c->event_pool.size() < 2) {
    for (int i = 0; i < 16; ++i) {
      hipEvent_t e;
      if (hipEventCreate(&e) == hipSuccess) c->event_pool.push_back(e);
    }
  }
  if (c->event_pool.size() < 2) return;
  tp->a = c->event_pool.back();
  c->event_pool.pop_back();
  tp->b = c->event_pool.back();
  c->event_pool.pop_back();
  tp->id = id;
  (void)hipEventRecord(tp->a, c->st);
}
void ph_end(sv_ctx *c, TimedPhase *tp) {
  if (tp->id < 0) return;
  (void)hipEventRecord(tp->b, c->st);
  c->pending.push_back(*tp);
}
void resolve_timings(sv_ctx *c) {
  for (auto &tp : c->pending) {
    float ms = 0.f;
    if (hipEventSynchronize(tp.b) == hipSuccess && hipEventElapsedTime(&ms, tp.a, tp.b) == hipSuccess) {
      c->phase_ms[tp.id] += ms;
      c->phase_n[tp.id] += 1;
    }
    c->event_pool.push_back(tp.a);
    c->event_pool.push_back(tp.b);
  }
  c->pending.clear();
}

struct Scoped {
  sv_ctx *c;
  TimedPhase tp;
  Scoped(sv_ctx *cc, int id) : c(cc) { ph_begin(c, id, &tp); }
  ~Scoped() { ph_end(c, &tp); }
};

sv_status sync(sv_ctx *c) {
  HIP_OK(c, hipStreamSynchronize(c->st));
  resolve_timings(c);
  return SV_OK;
}

DevConfig dev_config(const sv_config &k) {
  DevConfig d;
  d.bg[0] = k.bg[0];
  d.bg[1] = k.bg[1];
  d.bg[2] = k.bg[2];
  d.near_plane = k.near_plane;
  d.far_plane = k.far_plane;
  d.radius_clip = k.radius_clip;
  d.eps2d = k.eps2d;
  d.w_l2 = k.w_l2;
  d.w_l1 = k.w_l1;
  d.w_ssim = k.w_ssim;
  d.tile_size = k.tile_size;
  d.sh_deg = k.sh_deg;
  d.render_f16 = k.render_f16;
  d.importance_decay = k.importance_decay;
  d.xcd_remap = k.xcd_remap;
  d.geom_flags = k.geom_flags;
  return d;
}

sv_status check_ctx(sv_ctx *c) {
  if (!c) return SV_E_ARG;
  HIP_OK(c, hipSetDevice(c->device));
  return SV_OK;
}

sv_status check_view(sv_ctx *c, int v) {
  if (v < 0 || v >= (int)c->views.size())
    return fail(c, SV_E_ARG, "view %d out of range (%zu views)", v, c->views.size());
  return SV_OK;
}

// ---------------------------------------------------------------------
// Batch: a list of views rendered through one pipeline.
// ---------------------------------------------------------------------
void key_plan(sv_ctx *c);
sv_status ensure_isect_capacity(sv_ctx *c, uint64_t n);
sv_status ensure_super_capacity(sv_ctx *c, uint64_t n);

sv_status prepare_batch(sv_ctx *c, const int *vs, int nv) {
  if (nv <= 0) return fail(c, SV_E_ARG, "no views");
  if (c->N == 0) return fail(c, SV_E_STATE, "no Gaussians");
  for (int i = 0; i < nv; ++i)
    if (sv_status s = check_view(c, vs[i])) return s;
  const bool same = c->slot_views.size() == (size_t)nv &&
                    std::equal(c->slot_views.begin(), c->slot_views.end(), vs);
  c->built = false;
  c->forward_done = false;
  if (!same) {
    c->slot_views.assign(vs, vs + nv);
    c->h_slots.resize(nv);
    uint32_t tb = 0, sb = 0, bb = 0;
    c->max_tiles_x = c->max_tiles_y = 0;
    for (int s = 0; s < nv; ++s) {
      const HostView &hv = c->views[vs[s]];
      DevView &d = c->h_slots[s];
      std::memcpy(d.view, hv.cam.view, sizeof(d.view));
      std::memcpy(d.proj, hv.cam.proj, sizeof(d.proj));
      d.focal[0] = hv.cam.focal[0];
      d.focal[1] = hv.cam.focal[1];
      d.viewport[0] = hv.cam.viewport[0];
      d.viewport[1] = hv.cam.viewport[1];
      d.W = hv.W;
      d.H = hv.H;
      d.tiles_x = hv.tiles_x;
      d.tiles_y = hv.tiles_y;
      d.n_tiles = hv.n_tiles;
      d.tile_base = tb;
      d.pix_off = hv.pix_off;
      d.super_x = (hv.tiles_x + BIN_S - 1) / BIN_S;
      d.super_base = sb;
      d.blk_x = (hv.tiles_x + 1) / 2;
      d.blk_base = bb;
      bb += d.blk_x * ((hv.tiles_y + 1) / 2);
      tb += hv.n_tiles;
      sb += d.super_x * ((hv.tiles_y + BIN_S - 1) / BIN_S);
      c->max_tiles_x = std::max(c->max_tiles_x, hv.tiles_x);
      c->max_tiles_y = std::max(c->max_tiles_y, hv.tiles_y);
    }
    c->tiles_total = tb;
    c->super_total = sb;
    c->blocks_total = bb;
    if (c->blocks_cap < (size_t)bb + 1) {
      dfree(c->block_view);
      c->blocks_cap = 0;
      if (sv_status s = dalloc(c, &c->block_view, (size_t)bb + 1)) return s;
      c->blocks_cap = (size_t)bb + 1;
    }
    if (sv_status s = grow(c, &c->d_views, c->d_views_cap, (size_t)nv)) return s;
    HIP_OK(c, hipMemcpyAsync(c->d_views, c->h_slots.data(), nv * sizeof(DevView), hipMemcpyHostToDevice, c->st));
    if (c->tiles_cap < (size_t)tb + 1) {
      dfree(c->tile_view);
      dfree(c->ranges);
      dfree(c->l1_partial);
      dfree(c->tile_l1);
      dfree(c->diff);
      dfree(c->order_f);
      dfree(c->order_b);
      dfree(c->tile_cost);
      c->tiles_cap = 0;
      c->scored_views.clear();
      size_t want = (size_t)tb + 1;
      if (sv_status s = dalloc(c, &c->tile_view, want)) return s;
      if (sv_status s = dalloc(c, &c->ranges, want)) return s;
      if (sv_status s = dalloc(c, &c->l1_partial, want)) return s;
      if (sv_status s = dalloc(c, &c->tile_l1, want)) return s;
      if (sv_status s = dalloc(c, &c->diff, want)) return s;
      if (sv_status s = dalloc(c, &c->order_f, want)) return s;
      if (sv_status s = dalloc(c, &c->order_b, want)) return s;
      if (sv_status s = dalloc(c, &c->tile_cost, want)) return s;
      c->tiles_cap = want;
    }
    if (c->super_tab_cap < (size_t)sb + 1) {
      dfree(c->super_view);
      dfree(c->sranges);
      dfree(c->nch);
      dfree(c->cbase);
      c->super_tab_cap = 0;
      if (sv_status s = dalloc(c, &c->super_view, (size_t)sb + 1)) return s;
      if (sv_status s = dalloc(c, &c->sranges, (size_t)sb + 1)) return s;
      if (sv_status s = dalloc(c, &c->nch, (size_t)sb + 1)) return s;
      if (sv_status s = dalloc(c, &c->cbase, (size_t)sb + 1)) return s;
      c->super_tab_cap = (size_t)sb + 1;
    }
    std::vector<uint32_t> stv(sb);
    for (int s = 0; s < nv; ++s) {
      const DevView &d = c->h_slots[s];
      const uint32_t ns = d.super_x * ((d.tiles_y + BIN_S - 1) / BIN_S);
      for (uint32_t t = 0; t < ns; ++t) stv[d.super_base + t] = (uint32_t)s;
    }
    if (sb) HIP_OK(c, hipMemcpyAsync(c->super_view, stv.data(), sb * sizeof(uint32_t), hipMemcpyHostToDevice, c->st));
    std::vector<uint32_t> tv(tb);
    for (int s = 0; s < nv; ++s)
      for (uint32_t t = 0; t < c->h_slots[s].n_tiles; ++t) tv[c->h_slots[s].tile_base + t] = (uint32_t)s;
    if (tb) HIP_OK(c, hipMemcpyAsync(c->tile_view, tv.data(), tb * sizeof(uint32_t), hipMemcpyHostToDevice, c->st));
    std::vector<uint32_t> bv(bb);
    for (int s = 0; s < nv; ++s) {
      const DevView &d = c->h_slots[s];
      const uint32_t nbk = d.blk_x * ((d.tiles_y + 1) / 2);
      for (uint32_t t = 0; t < nbk; ++t) bv[d.blk_base + t] = (uint32_t)s;
    }
    if (bb) HIP_OK(c, hipMemcpyAsync(c->block_view, bv.data(), bb * sizeof(uint32_t), hipMemcpyHostToDevice, c->st));
    // the host vectors must outlive the async copies
    HIP_OK(c, hipStreamSynchronize(c->st));
  }
  const size_t M = (size_t)nv * c->N;
  if (M > c->vg_cap || !c->rec) {
    dfree(c->rec);
    dfree(c->aux);
    dfree(c->counts);
    dfree(c->offs);
    dfree(c->scounts);
    dfree(c->soffs);
    dfree(c->rects);
    dfree(c->tins);
    dfree(c->vrec);
    c->vg_cap = 0;
    if (sv_status s = dalloc(c, &c->vrec, (size_t)REC9 * M)) return s;
    if (sv_status s = dalloc(c, &c->rec, 4 * M)) return s;
    if (sv_status s = dalloc(c, &c->aux, M)) return s;
    if (sv_status s = dalloc(c, &c->counts, M)) return s;
    if (sv_status s = dalloc(c, &c->offs, M)) return s;
    if (sv_status s = dalloc(c, &c->scounts, M)) return s;
    if (sv_status s = dalloc(c, &c->soffs, M)) return s;
    if (sv_status s = dalloc(c, &c->rects, M)) return s;
    if (sv_status s = dalloc(c, &c->tins, M)) return s;
    c->vg_cap = M;
  }
  // projection-block tables: slots x blocks per slot (+1 for the scan total)
  const size_t nbk = (size_t)nv * project_blocks(c->N) + 1;
  if (nbk > c->btab_cap || !c->btot) {
    dfree(c->btot);
    dfree(c->sbtot);
    dfree(c->bbase);
    dfree(c->sbbase);
    c->btab_cap = 0;
    if (sv_status s = dalloc(c, &c->btot, nbk)) return s;
    if (sv_status s = dalloc(c, &c->sbtot, nbk)) return s;
    if (sv_status s = dalloc(c, &c->bbase, nbk)) return s;
    if (sv_status s = dalloc(c, &c->sbbase, nbk)) return s;
    c->btab_cap = nbk;
  }
  const size_t scan_n = std::max(M, (size_t)c->tiles_total + 1);
  if (sv_status s = grow(c, &c->scan_tmp, c->scan_tmp_cap, scan_tmp_words(scan_n))) return s;
  key_plan(c);
  if (sv_status s = ensure_isect_capacity(c, 1)) return s;
  return c->binned ? ensure_super_capacity(c, 1) : SV_OK;
}

size_t rec_stride(const sv_ctx *c) { return (size_t)c->slot_views.size() * c->N; }

// Key type and the most sort passes the batch can need, from bounds the
// host knows: tile bits of the batch and the depth quantum bound. A listed
// splat has depth <= far_plane, so depth_q <= min(65535, far_plane).
void key_plan(sv_ctx *c) {
  c->tile_bits = bitlen(c->tiles_total > 0 ? c->tiles_total - 1 : 0);
  const double far = (double)c->cfg.far_plane;
  const uint64_t dq_max = (far >= 0.0) ? (uint64_t)std::floor(std::min(far, 65535.0)) : 65535u;
  c->max_key_bits = c->tile_bits + bitlen(dq_max);
  c->key64 = c->max_key_bits > 32;
  c->max_passes = (int)((c->max_key_bits + 7) / 8);
  // binned build: supertile key + depth bits (the device checks the live
  // width against the 32-bit key with the rectangle bits)
  c->super_bits = bitlen(c->super_total > 0 ? c->super_total - 1 : 0);
  // a binned step's key has at most 32 - BIN_RECT_BITS bits (wider keys
  // fall back to the pair sort), so three 8-bit passes cover it
  c->max_super_passes =
      (int)std::min<uint32_t>((32u - BIN_RECT_BITS + 7u) / 8u, (c->super_bits + bitlen(dq_max) + 7) / 8);
  c->binned = c->cfg.binned_build && !c->bin_unfit && c->super_bits + BIN_RECT_BITS < 32;
}

// Supertile pair buffers for at least n pairs (binned build).
sv_status ensure_super_capacity(sv_ctx *c, uint64_t n) {
  if (n <= c->super_cap && c->skeys) {
    // chunks of the supertile lists: bounded by pairs / BIN_CH + supertiles
    const size_t cc = bin_chunk_cap(c->super_cap, c->super_total);
    if (cc <= c->chunk_cap && c->chunk_super) return SV_OK;
    dfree(c->chunk_super);
    dfree(c->ccount);
    c->chunk_cap = 0;
    if (sv_status s = dalloc(c, &c->chunk_super, cc)) return s;
    if (sv_status s = dalloc(c, &c->ccount, cc * 64)) return s;
    c->chunk_cap = cc;
    return SV_OK;
  }
  const uint64_t want = std::max<uint64_t>(n + n / 4, 1 << 16);
  dfree(c->skeys);
  dfree(c->skeys2);
  dfree(c->svals);
  dfree(c->svals2);
  dfree(c->ssort_tmp);
  c->super_cap = 0;
  if (sv_status s = dalloc(c, &c->skeys, want)) return s;
  if (sv_status s = dalloc(c, &c->skeys2, want)) return s;
  if (sv_status s = dalloc(c, &c->svals, want)) return s;
  if (sv_status s = dalloc(c, &c->svals2, want)) return s;
  if (sv_status s = dalloc(c, &c->ssort_tmp, sort_tmp_words<uint32_t>(want))) return s;
  c->super_cap = want;
  return ensure_super_capacity(c, n);
}

// Pair buffers for at least n intersections (contents are not kept).
sv_status ensure_isect_capacity(sv_ctx *c, uint64_t n) {
  if (n <= c->isect_cap && c->keys) return SV_OK;
  const uint64_t want = std::max<uint64_t>(n + n / 4, 1 << 16);
  dfree(c->keys);
  dfree(c->keys2);
  dfree(c->vals);
  dfree(c->vals2);
  dfree(c->isect_rec);
  dfree(c->rflag);
  dfree(c->tmask);
  dfree(c->sort_tmp);
  c->isect_cap = 0;
  c->tmask_ready = false;
  c->sort_tmp_cap = 0;
  if (sv_status s = dalloc(c, (uint64_t **)&c->keys, want)) return s;
  if (sv_status s = dalloc(c, (uint64_t **)&c->keys2, want)) return s;
  if (sv_status s = dalloc(c, &c->vals, want)) return s;
  if (sv_status s = dalloc(c, &c->vals2, want)) return s;
  if (sv_status s = dalloc(c, &c->isect_rec, (size_t)want * REC9)) return s;
  if (sv_status s = dalloc(c, &c->rflag, flag_array_bytes(want))) return s;
  if (sv_status s = dalloc(c, &c->tmask, (size_t)want)) return s;
  const size_t stw = sort_tmp_words<uint64_t>(want);
  if (sv_status s = dalloc(c, &c->sort_tmp, stw)) return s;
  c->sort_tmp_cap = stw;
  c->isect_cap = want;
  return SV_OK;
}

// buildTileIntersections (gaussian-renderer.ts:1780-2023) for the batch,
// enqueued without any host readback: the pair count, key layout, sort
// passes and result buffer stay on the device (StepMeta), and every launch
// is sized by capacities, so the same launches serve any step.
sv_status enqueue_build(sv_ctx *c) {
  const int nv = (int)c->slot_views.size();
  const DevConfig dc = dev_config(c->cfg);
  const size_t M = (size_t)nv * c->N;
  const bool bin = c->binned;
  StepMeta *m = c->d_meta;
  launch_fill_u32(c->st, (uint32_t *)m, 0u, 4 * sizeof(StepMeta) / 4);
  {
    Scoped t(c, PH_PROJECT);
    launch_project(c->st, c->P, c->cap, c->N, c->d_views, nv, dc, c->rec, c->aux, c->counts, &m->max_dq,
                   bin ? c->scounts : nullptr, c->rects, c->tins, c->offs, bin ? c->soffs : nullptr, c->btot,
                   bin ? c->sbtot : nullptr);
  }
  {
    Scoped t(c, PH_SCAN);
    // second level of the scans: over the projection blocks' totals
    const size_t nbk = (size_t)nv * project_blocks(c->N);
    launch_scan(c->st, c->btot, c->bbase, nbk, c->scan_tmp, &m->n_isect);
    if (bin) {
      launch_scan(c->st, c->sbtot, c->sbbase, nbk, c->scan_tmp, &m[1].n_isect);
      launch_keyinfo(c->st, m, c->tile_bits, 64u, c->isect_cap);
      launch_bin_info(c->st, m, c->super_bits, c->super_cap);
    } else {
      launch_keyinfo(c->st, m, c->tile_bits, c->max_key_bits, c->isect_cap);
    }
  }
  if (bin) {
    {
      Scoped t(c, PH_EMIT);
      launch_emit_super(c->st, c->rects, c->aux, c->scounts, c->soffs, c->sbbase, c->N, nv, c->d_views, m,
                        c->skeys, c->svals);
    }
    {
      Scoped t(c, PH_SORT);
      launch_sort<uint32_t>(c->st, c->skeys, c->svals, c->skeys2, c->svals2, c->super_cap, c->max_super_passes,
                            c->ssort_tmp, &m[1]);
    }
    {
      Scoped t(c, PH_RANGES);
      launch_ranges<uint32_t>(c->st, c->skeys, c->skeys2, c->super_cap, c->super_total, &m[1], c->sranges);
    }
    {
      Scoped t(c, PH_FILL);
      BinArgs b{};
      b.keys0 = c->skeys;
      b.vals0 = c->svals;
      b.keys1 = c->skeys2;
      b.vals1 = c->svals2;
      b.sranges = c->sranges;
      b.super_view = c->super_view;
      b.views = c->d_views;
      b.m = m;
      b.n_super_tiles = c->super_total;
      b.n_tiles = c->tiles_total;
      b.chunk_cap = c->chunk_cap;
      b.nch = c->nch;
      b.cbase = c->cbase;
      b.chunk_super = c->chunk_super;
      b.ccount = c->ccount;
      b.tcount = c->diff;
      b.scan_tmp = c->scan_tmp;
      b.ranges = c->ranges;
      b.out = c->vals;
      launch_fill_tiles(c->st, b);
    }
    return check_launch(c, "binned tile build");
  }
  {
    Scoped t(c, PH_EMIT);
    if (c->key64)
      launch_emit<uint64_t>(c->st, c->rec, c->rects, c->aux, c->counts, c->offs, c->bbase, c->N, nv, c->d_views, m,
                            (uint64_t *)c->keys, c->vals);
    else
      launch_emit<uint32_t>(c->st, c->rec, c->rects, c->aux, c->counts, c->offs, c->bbase, c->N, nv, c->d_views, m,
                            (uint32_t *)c->keys, c->vals);
  }
  {
    Scoped t(c, PH_SORT);
    if (c->key64)
      launch_sort<uint64_t>(c->st, (uint64_t *)c->keys, c->vals, (uint64_t *)c->keys2, c->vals2,
                            c->isect_cap, c->max_passes, c->sort_tmp, m);
    else
      launch_sort<uint32_t>(c->st, (uint32_t *)c->keys, c->vals, (uint32_t *)c->keys2, c->vals2,
                            c->isect_cap, c->max_passes, c->sort_tmp, m);
  }
  {
    Scoped t(c, PH_RANGES);
    if (c->key64)
      launch_ranges<uint64_t>(c->st, (const uint64_t *)c->keys, (const uint64_t *)c->keys2, c->isect_cap,
                              c->tiles_total, m, c->ranges);
    else
      launch_ranges<uint32_t>(c->st, (const uint32_t *)c->keys, (const uint32_t *)c->keys2, c->isect_cap,
                              c->tiles_total, m, c->ranges);
  }
  return check_launch(c, "tile build");
}

// Wait for the stream and read the build's metadata. *retry: nothing past
// the scans wrote any state and the build must run again, after the pair
// buffers grew (overflow 1) or with the tile pairs sorted instead of
// binned (overflow 3: supertile key and rectangle wider than 32 bits).
sv_status read_meta(sv_ctx *c, bool *retry) {
  *retry = false;
  HIP_OK(c, hipMemcpyAsync(c->h_meta, c->d_meta, 2 * sizeof(StepMeta), hipMemcpyDeviceToHost, c->st));
  if (sv_status s = sync(c)) return s;
  const StepMeta &m = c->h_meta[0];
  const StepMeta &m1 = c->h_meta[1];
  if (m.overflow == 2)
    return fail(c, SV_E_CAPACITY, "sort key of %u bits exceeds the planned %u bits", m.key_bits, c->max_key_bits);
  if (m.overflow == 3) {
    c->bin_unfit = true;
    key_plan(c);
    *retry = true;
    return SV_OK;
  }
  if (m.overflow == 1) {
    if (m.n_isect >= 0xffffffffull || (c->binned && m1.n_isect >= 0xffffffffull))
      return fail(c, SV_E_CAPACITY, "%llu intersections exceed the 32-bit index range",
                  (unsigned long long)std::max(m.n_isect, m1.n_isect));
    if (sv_status s = ensure_isect_capacity(c, m.n_isect)) return s;
    if (c->binned)
      if (sv_status s = ensure_super_capacity(c, m1.n_isect)) return s;
    *retry = true;
    return SV_OK;
  }
  c->I = m.n_isect;
  c->dbits = m.dbits;
  c->key_bits = m.key_bits;
  c->passes = (int)(c->binned ? m1.passes : m.passes);
  c->spasses = m1.passes;
  c->sel = (int)m.sel;
  c->last_binned = c->binned;
  return SV_OK;
}

// Build and wait (the per-view entry points and the plain / sequential
// steps read the pair count on the host).
sv_status build_batch(sv_ctx *c) {
  for (int attempt = 0; attempt < 3; ++attempt) {
    if (sv_status s = enqueue_build(c)) return s;
    bool retry = false;
    if (sv_status s = read_meta(c, &retry)) return s;
    if (!retry) {
      c->built = true;
      return SV_OK;
    }
  }
  return fail(c, SV_E_STATE, "tile build did not fit after growing the buffers");
}

const uint32_t *cur_vals(const sv_ctx *c) { return c->sel ? c->vals2 : c->vals; }
const void *cur_keys(const sv_ctx *c) { return c->sel ? c->keys2 : c->keys; }

RasterArgs raster_args(sv_ctx *c) {
  RasterArgs a{};
  a.views = c->d_views;
  a.tile_view = c->tile_view;
  a.ranges = c->ranges;
  a.vals0 = c->vals;
  a.vals1 = c->vals2;
  a.meta = c->d_meta;
  a.vals = nullptr;
  a.n_isect = 0;
  a.rec = c->rec;
  a.N = c->N;
  a.n_tiles_total = c->tiles_total;
  a.cfg = dev_config(c->cfg);
  a.color = c->color;
  a.alpha = c->alpha;
  a.last = c->last;
  a.target = c->target;
  a.residual = nullptr;
  a.bbase = c->bbase;
  a.nb = project_blocks(c->N);
  a.isect_rec = c->isect_rec;
  a.rflag = c->rflag;
  a.tmask = nullptr;
  a.aux = c->aux;
  a.l1_partial = c->l1_partial;
  a.work_counter = nullptr;
  a.block_view = c->block_view;
  a.n_blocks_total = c->blocks_total;
  return a;
}

// Grid of a raster launch: one workgroup per tile, or with `persistent` as
// many workgroups as the device holds at once (occupancy x CUs), pulling
// tiles from the per-XCD work queues.
uint32_t raster_grid(sv_ctx *c, RasterArgs &a, bool backward) {
  const uint32_t n = a.n_tiles_total - a.tile_begin;
  if (c->cfg.persistent) {
    const int bi = c->cfg.batch <= 64 ? 0 : (c->cfg.batch <= 128 ? 1 : 2);
    const uint32_t resident = c->resident[backward ? 1 : 0][bi];
    a.work_counter = c->d_work;
    launch_fill_u32(c->st, a.work_counter, 0u, (size_t)NUM_WORK_QUEUES * WORK_QUEUE_STRIDE);
    return std::min<uint32_t>(n, resident ? resident : 2048u);
  }
  return n;
}

// Work items of a 16-px raster launch go to the XCDs in contiguous ranges
// (xcd_remap, the persistent queues) or round-robin in item order.
bool order_contiguous(const sv_ctx *c) { return c->cfg.persistent || c->cfg.xcd_remap; }

sv_status enqueue_forward(sv_ctx *c) {
  RasterArgs a = raster_args(c);
  // taken masks for the backward; an 8-px raster wave renders one quadrant
  // of a tile, stops when that quadrant is done and writes none (the
  // backward then culls by itself)
  a.tmask = c->cfg.raster_px != 8 ? c->tmask : nullptr;
  c->tmask_ready = a.tmask != nullptr;
  // 8- and 32-px raster tiles launch a workgroup per quadrant / block
  const uint32_t grid = c->cfg.raster_px == 16 ? raster_grid(c, a, false) : c->tiles_total;
  // the work order (16-px tiles) and, for the backward's, each tile's
  // visited entries (written by waves that render whole tiles)
  if (c->cfg.tile_order) a.tile_cost = c->tile_cost;
  c->cost_ready = c->cfg.tile_order && c->cfg.raster_px != 8;
  {
    Scoped t(c, PH_FORWARD);
    if (c->cfg.tile_order && c->cfg.raster_px == 16) {
      launch_tile_order(c->st, c->ranges, nullptr, c->d_meta, 0u, c->tiles_total, order_contiguous(c), c->order_f);
      a.tile_order = c->order_f;
    }
    launch_forward(c->st, a, c->cfg.batch, grid, c->cfg.raster_px);
  }
  return check_launch(c, "forward");
}

sv_status forward_batch(sv_ctx *c) {
  if (!c->built) return fail(c, SV_E_STATE, "tiles not built");
  if (sv_status s = enqueue_forward(c)) return s;
  c->forward_done = true;
  return SV_OK;
}

// use_residual: read the residual arena instead of the fused loss.
// The backward's raster arguments: the forward's taken masks when it wrote
// them.
RasterArgs backward_args(sv_ctx *c, bool use_residual) {
  RasterArgs a = raster_args(c);
  if (use_residual) a.residual = c->residual;
  a.tmask = c->tmask_ready ? c->tmask : nullptr;
  return a;
}

sv_status enqueue_backward(sv_ctx *c, bool use_residual) {
  RasterArgs a = backward_args(c, use_residual);
  const uint32_t grid = raster_grid(c, a, true);
  {
    Scoped t(c, PH_BACKWARD);
    launch_clear_flags(c->st, c->rflag, c->isect_cap, c->d_meta);
    // longest first by the entries the forward visited (by list length
    // when it wrote no costs)
    if (c->cfg.tile_order) {
      launch_tile_order(c->st, c->ranges, c->cost_ready ? c->tile_cost : nullptr, c->d_meta, 0u, c->tiles_total,
                        order_contiguous(c), c->order_b);
      a.tile_order = c->order_b;
    }
    launch_backward(c->st, a, c->cfg.batch, grid);
  }
  return check_launch(c, "backward");
}

sv_status backward_batch(sv_ctx *c, bool use_residual) {
  if (!c->forward_done) return fail(c, SV_E_STATE, "forward not run for this batch");
  for (int v : c->slot_views)
    if (!use_residual && !c->views[v].has_target) return fail(c, SV_E_STATE, "view %d has no target", v);
  return enqueue_backward(c, use_residual);
}

// Adam hyper-parameters of step t (run_adam_on_buffer,
// gaussian-renderer.ts:2409-2452).
AdamParams adam_params(const sv_ctx *c, uint32_t t) {
  const sv_config &k = c->cfg;
  AdamParams h{};
  double lr_pos = k.lr_pos;
  if (k.pos_lr_schedule && k.lr_pos > 0.0f) {
    // computePositionLearningRate (gaussian-renderer.ts:284-298)
    const double lr_final = 0.0000016, delay_mult = 0.01, max_steps = 30000.0;
    const double step = std::max<double>(t, 0.0);
    const double tt = std::min(1.0, step / max_steps);
    const double loglerp = std::exp(std::log((double)k.lr_pos) * (1 - tt) + std::log(lr_final) * tt);
    const double phase = std::min(1.0, step / std::max(1.0, std::round(0.1 * max_steps)));
    const double delay = delay_mult + (1.0 - delay_mult) * std::sin(0.5 * M_PI * phase);
    lr_pos = loglerp * delay;
  }
  h.lr[0] = (float)lr_pos;
  h.lr[1] = k.lr_scale;
  h.lr[2] = k.lr_rot;
  h.lr[3] = k.lr_opacity;
  h.lr[4] = k.lr_sh;
  h.beta1 = k.beta1;
  h.beta2 = k.beta2;
  h.eps = k.eps;
  // bias terms in double, stored as f32 (gaussian-renderer.ts:2425-2440)
  const double tt = t < 1 ? 1.0 : (double)t;
  h.bc1_rcp = (float)(1.0 / (1.0 - std::pow((double)k.beta1, tt)));
  h.bc2_sqrt_rcp = (float)(1.0 / std::sqrt(1.0 - std::pow((double)k.beta2, tt)));
  h.update_sh = (k.sh_update_interval == 0) ? 1u : ((t % k.sh_update_interval) == 0 ? 1u : 0u);
  return h;
}

VjpArgs vjp_args(sv_ctx *c) {
  VjpArgs v{};
  v.P = c->P;
  v.cap = c->cap;
  v.N = c->N;
  v.views = c->d_views;
  v.nv = (int)c->slot_views.size();
  v.slot_begin = 0;
  v.slot_count = v.nv;
  v.cfg = dev_config(c->cfg);
  v.counts = c->counts;
  v.offs = c->offs;
  v.bbase = c->bbase;
  v.nb = project_blocks(c->N);
  v.rec = c->rec;
  v.rects = c->rects;
  v.tins = c->tins;
  v.aux = c->aux;
  v.isect_rec = c->isect_rec;
  v.rflag = c->rflag;
  v.vrec = c->vrec;
  v.importance = c->importance;
  v.G = c->G;
  v.geom = c->geom;
  v.grad_accum = c->grad_accum;
  v.meta = c->d_meta;
  v.Pw = c->P;
  v.M1 = c->M1;
  v.M2 = c->M2;
  v.adam = c->d_adam;
  return v;
}

sv_status zero_grads(sv_ctx *c) {
  launch_fill_f32(c->st, c->G, 0.0f, (size_t)P_NUM * c->cap);
  launch_fill_f32(c->st, c->geom, 0.0f, (size_t)5 * c->cap);
  return check_launch(c, "zero grads");
}

// Adam over the gradient planes with parameters already in d_adam.
sv_status enqueue_adam_planes(sv_ctx *c) {
  {
    Scoped t(c, PH_ADAM);
    launch_adam(c->st, c->P, c->M1, c->M2, c->G, c->cap, c->N, c->d_adam);
  }
  return check_launch(c, "adam");
}

sv_status adam_planes(sv_ctx *c) {
  c->step_t += 1;
  launch_load_adam(c->st, adam_params(c, c->step_t), nullptr, c->d_adam);
  return enqueue_adam_planes(c);
}

sv_status upload_planar(sv_ctx *c, float *dst, size_t cap, const float *src, size_t n, int comps,
                        int plane0) {
  if (!src) return SV_OK;
  std::vector<float> tmp(n);
  for (int k = 0; k < comps; ++k) {
    for (size_t i = 0; i < n; ++i) tmp[i] = src[i * comps + k];
    HIP_OK(c, hipMemcpy(dst + (size_t)(plane0 + k) * cap, tmp.data(), n * sizeof(float), hipMemcpyHostToDevice));
  }
  return SV_OK;
}
sv_status download_planar(sv_ctx *c, const float *src, size_t cap, float *dst, size_t n, int comps,
                          int plane0) {
  if (!dst) return SV_OK;
  std::vector<float> tmp(n);
  for (int k = 0; k < comps; ++k) {
    HIP_OK(c, hipMemcpy(tmp.data(), src + (size_t)(plane0 + k) * cap, n * sizeof(float), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; ++i) dst[i * comps + k] = tmp[i];
  }
  return SV_OK;
}

// The first non-finite value of an input array (n * comps floats), or
// SV_OK: NaN and infinities are refused at the boundary, before any kernel
// sees them (the reference has no such check; a NaN would spread through
// every later step).
sv_status check_finite(sv_ctx *c, const char *what, const float *v, size_t count) {
  if (!v) return SV_OK;
  for (size_t i = 0; i < count; ++i)
    if (!std::isfinite(v[i])) return fail(c, SV_E_ARG, "%s[%zu] is not finite", what, i);
  return SV_OK;
}

sv_status check_params_finite(sv_ctx *c, size_t n, const float *pos, const float *ls, const float *q,
                              const float *lg, const float *sh) {
  if (sv_status s = check_finite(c, "pos", pos, 3 * n)) return s;
  if (sv_status s = check_finite(c, "logscale", ls, 3 * n)) return s;
  if (sv_status s = check_finite(c, "quat", q, 4 * n)) return s;
  if (sv_status s = check_finite(c, "logit", lg, n)) return s;
  return check_finite(c, "sh", sh, 48 * n);
}

sv_status set_params(sv_ctx *c, float *base, const float *pos, const float *ls, const float *q,
                     const float *lg, const float *sh) {
  if (sv_status s = upload_planar(c, base, c->cap, pos, c->N, 3, P_POS)) return s;
  if (sv_status s = upload_planar(c, base, c->cap, ls, c->N, 3, P_SCALE)) return s;
  if (sv_status s = upload_planar(c, base, c->cap, q, c->N, 4, P_QUAT)) return s;
  if (sv_status s = upload_planar(c, base, c->cap, lg, c->N, 1, P_LOGIT)) return s;
  return upload_planar(c, base, c->cap, sh, c->N, 48, P_SH);
}
sv_status get_params(sv_ctx *c, const float *base, float *pos, float *ls, float *q, float *lg, float *sh) {
  HIP_OK(c, hipStreamSynchronize(c->st));
  if (sv_status s = download_planar(c, base, c->cap, pos, c->N, 3, P_POS)) return s;
  if (sv_status s = download_planar(c, base, c->cap, ls, c->N, 3, P_SCALE)) return s;
  if (sv_status s = download_planar(c, base, c->cap, q, c->N, 4, P_QUAT)) return s;
  if (sv_status s = download_planar(c, base, c->cap, lg, c->N, 1, P_LOGIT)) return s;
  return download_planar(c, base, c->cap, sh, c->N, 48, P_SH);
}

sv_status alloc_gaussian_arrays(sv_ctx *c, size_t cap) {
  if (sv_status s = dalloc(c, &c->P, (size_t)P_NUM * cap)) return s;
  if (sv_status s = dalloc(c, &c->M1, (size_t)P_NUM * cap)) return s;
  if (sv_status s = dalloc(c, &c->M2, (size_t)P_NUM * cap)) return s;
  if (sv_status s = dalloc(c, &c->G, (size_t)P_NUM * cap)) return s;
  if (sv_status s = dalloc(c, &c->geom, (size_t)5 * cap)) return s;
  if (sv_status s = dalloc(c, &c->importance, cap)) return s;
  if (sv_status s = dalloc(c, &c->grad_accum, cap)) return s;
  return SV_OK;
}

void clear_state(sv_ctx *c) {
  launch_fill_f32(c->st, c->M1, 0.f, (size_t)P_NUM * c->cap);
  launch_fill_f32(c->st, c->M2, 0.f, (size_t)P_NUM * c->cap);
  launch_fill_f32(c->st, c->G, 0.f, (size_t)P_NUM * c->cap);
  launch_fill_f32(c->st, c->geom, 0.f, (size_t)5 * c->cap);
  launch_fill_f32(c->st, (float *)c->importance, 0.f, (size_t)4 * c->cap);
  launch_fill_f32(c->st, (float *)c->grad_accum, 0.f, (size_t)4 * c->cap);
}

sv_status ensure_view_built(sv_ctx *c, int v) {
  if (c->built && c->slot_views.size() == 1 && c->slot_views[0] == v) return SV_OK;
  if (sv_status s = prepare_batch(c, &v, 1)) return s;
  return build_batch(c);
}

sv_status read_pixels3(sv_ctx *c, const float *arena, int v, float *rgb) {
  if (sv_status s = check_view(c, v)) return s;
  if (!rgb) return fail(c, SV_E_ARG, "null output");
  const HostView &hv = c->views[v];
  const size_t P = (size_t)hv.W * hv.H;
  std::vector<float> tmp(3 * P);
  HIP_OK(c, hipStreamSynchronize(c->st));
  HIP_OK(c, hipMemcpy(tmp.data(), arena + 3 * hv.pix_off, 3 * P * sizeof(float), hipMemcpyDeviceToHost));
  for (size_t i = 0; i < P; ++i)
    for (int ch = 0; ch < 3; ++ch) rgb[3 * i + ch] = tmp[ch * P + i];
  return SV_OK;
}

sv_status write_pixels3(sv_ctx *c, float *arena, int v, const float *rgb) {
  const HostView &hv = c->views[v];
  const size_t P = (size_t)hv.W * hv.H;
  std::vector<float> tmp(3 * P);
  for (size_t i = 0; i < P; ++i)
    for (int ch = 0; ch < 3; ++ch) tmp[ch * P + i] = rgb[3 * i + ch];
  HIP_OK(c, hipStreamSynchronize(c->st));
  HIP_OK(c, hipMemcpy(arena + 3 * hv.pix_off, tmp.data(), 3 * P * sizeof(float), hipMemcpyHostToDevice));
  return SV_OK;
}

// The batched step needs every view's residual either fused or read.
bool batch_needs_residual_buffer(sv_ctx *c) { return c->cfg.w_ssim > 0.0f; }

sv_status loss_views(sv_ctx *c) {
  const DevConfig dc = dev_config(c->cfg);
  Scoped t(c, PH_LOSS);
  for (size_t s = 0; s < c->slot_views.size(); ++s) {
    const HostView &hv = c->views[c->slot_views[s]];
    launch_loss(c->st, c->h_slots[s], dc, c->color + 3 * hv.pix_off, c->target + 3 * hv.pix_off,
                c->residual + 3 * hv.pix_off);
  }
  return check_launch(c, "loss");
}

// run_wavefront_backward + run_training_backward_geom
// (gaussian-renderer.ts:2312-2407) for the single-view batch, reading the
// residual arena. Adds into G; grad_geom of the view is left in geom.
sv_status backward_fixed_batch(sv_ctx *c, uint32_t flags, bool stats) {
  if (c->I > c->fixed_cap || !c->fixed) {
    dfree(c->fixed);
    dfree(c->fixed_inv);
    c->fixed_cap = 0;
    const uint64_t want = std::max<uint64_t>(c->I + c->I / 4, 1024);
    if (sv_status s = dalloc(c, &c->fixed, (size_t)want * 54)) return s;
    if (sv_status s = dalloc(c, &c->fixed_inv, (size_t)want)) return s;
    c->fixed_cap = want;
  }
  RasterArgs a = raster_args(c);
  a.residual = c->residual;
  launch_fill_f32(c->st, c->geom, 0.f, (size_t)5 * c->cap);
  {
    Scoped t(c, PH_FIXED);
    launch_backward_fixed(c->st, a, c->P, c->cap, flags, c->fixed, c->fixed_inv);
    launch_reduce_fixed(c->st, c->fixed, c->fixed_inv, c->counts, c->offs, c->bbase, c->N, c->d_meta, c->cap, c->G,
                        c->geom);
    launch_geom(c->st, c->P, c->cap, c->N, c->d_views, dev_config(c->cfg), c->geom, c->G,
                stats ? c->grad_accum : nullptr, c->aux, c->counts, stats ? c->importance : nullptr);
  }
  return check_launch(c, "backward_fixed");
}

// The batched step after prepare_batch, enqueued with no host readback
// and no allocation, so that it can be captured: build, forward,
// [loss], adjoint backward, per-Gaussian VJP with fused Adam (or VJP into
// the gradient planes + Adam when keep_grads). Adam parameters come from
// `hp` or, when hp is null, from the mapped host block h_adam (re-read on
// every replay).
// The batched step's backward as two launches over two view ranges of
// about half the tiles each (split at a view boundary); the first range's
// record gather runs on the side stream while the second range's backward
// runs (the gather of a range reads only the records and flags that range's
// backward wrote; the flags of every range are cleared before the first
// launch). Every slot's records are gathered when it returns.
sv_status enqueue_split_backward(sv_ctx *c, bool use_res) {
  const int nv = (int)c->slot_views.size();
  int k = 1;
  while (k < nv - 1 && c->h_slots[k].tile_base < c->tiles_total / 2) ++k;
  const uint32_t tk = c->h_slots[k].tile_base;
  {
    Scoped t(c, PH_BACKWARD);
    RasterArgs a = backward_args(c, use_res);
    a.n_tiles_total = tk;
    launch_clear_flags(c->st, c->rflag, c->isect_cap, c->d_meta);
    // each range's work order over its own items (the XCD ranges of a
    // launch follow its size)
    if (c->cfg.tile_order) {
      const uint32_t *cost = c->cost_ready ? c->tile_cost : nullptr;
      launch_tile_order(c->st, c->ranges, cost, c->d_meta, 0u, tk, order_contiguous(c), c->order_b);
      launch_tile_order(c->st, c->ranges, cost, c->d_meta, tk, c->tiles_total - tk, order_contiguous(c),
                        c->order_b + tk);
      a.tile_order = c->order_b;
    }
    launch_backward(c->st, a, c->cfg.batch, raster_grid(c, a, true));
    HIP_OK(c, hipEventRecord(c->ev_fork, c->st));
    HIP_OK(c, hipStreamWaitEvent(c->st2, c->ev_fork, 0));
    VjpArgs ga = vjp_args(c);
    ga.slot_begin = 0;
    ga.slot_count = k;
    launch_gather_records(c->st2, ga);
    RasterArgs b = backward_args(c, use_res);
    b.tile_begin = tk;
    if (c->cfg.tile_order) b.tile_order = c->order_b + tk;
    launch_backward(c->st, b, c->cfg.batch, raster_grid(c, b, true));
    VjpArgs gb = vjp_args(c);
    gb.slot_begin = k;
    gb.slot_count = nv - k;
    launch_gather_records(c->st, gb);
    HIP_OK(c, hipEventRecord(c->ev_join, c->st2));
    HIP_OK(c, hipStreamWaitEvent(c->st, c->ev_join, 0));
  }
  return check_launch(c, "split backward");
}

sv_status enqueue_batched_step(sv_ctx *c, const AdamParams *hp) {
  if (sv_status s = enqueue_build(c)) return s;
  if (sv_status s = enqueue_forward(c)) return s;
  const bool use_res = batch_needs_residual_buffer(c);
  if (use_res)
    if (sv_status s = loss_views(c)) return s;
  // (with the phase timers on, the backward stays one launch, so that its
  // time is the kernel's)
  const bool split = c->cfg.split_backward && c->slot_views.size() >= 2 && !c->cfg.keep_grads && !c->timing;
  if (split) {
    if (sv_status s = enqueue_split_backward(c, use_res)) return s;
  } else if (sv_status s = enqueue_backward(c, use_res)) {
    return s;
  }
  launch_load_adam(c->st, hp ? *hp : AdamParams{}, hp ? nullptr : c->h_adam_dev, c->d_adam);
  VjpArgs va = vjp_args(c);
  if (c->cfg.keep_grads) {
    if (sv_status s = zero_grads(c)) return s;
    {
      Scoped t(c, PH_VJP);
      launch_gather_records(c->st, va);
      launch_vjp(c->st, va);
    }
    if (sv_status s = check_launch(c, "vjp")) return s;
    return enqueue_adam_planes(c);
  }
  va.G = nullptr;
  va.geom = nullptr;
  {
    Scoped t(c, PH_VJP);
    if (!split) launch_gather_records(c->st, va);
    launch_vjp(c->st, va);
  }
  return check_launch(c, "vjp+adam");
}

// Everything a captured step depends on: buffers, sizes, the batch's
// views and the configuration (passed by value into kernels).
std::vector<uint64_t> graph_signature(const sv_ctx *c) {
  std::vector<uint64_t> g;
  const void *ptrs[] = {c->P, c->M1, c->M2, c->G, c->geom, c->importance, c->grad_accum, c->color,
                        c->alpha, c->residual, c->target, c->last, c->d_views, c->tile_view, c->ranges,
                        c->l1_partial, c->rflag, c->tmask, c->rec, c->aux, c->counts, c->offs, c->vrec, c->scan_tmp, c->keys,
                        c->keys2, c->vals, c->vals2, c->sort_tmp, c->isect_rec, c->d_meta, c->d_adam,
                        c->h_adam_dev, c->scounts, c->soffs, c->rects, c->tins, c->btot, c->sbtot, c->bbase, c->sbbase, c->super_view, c->sranges, c->diff, c->order_f, c->order_b, c->tile_cost,
                        c->skeys, c->skeys2, c->svals, c->svals2, c->ssort_tmp, c->nch, c->cbase,
                        c->chunk_super, c->ccount, c->d_work, c->block_view};
  for (const void *p : ptrs) g.push_back((uint64_t)(uintptr_t)p);
  g.push_back(c->N);
  g.push_back(c->cap);
  g.push_back(c->isect_cap);
  g.push_back(c->super_cap);
  g.push_back(c->chunk_cap);
  g.push_back(c->binned ? 1u : 0u);
  g.push_back(c->tiles_total);
  g.push_back(c->blocks_total);
  g.push_back(c->super_total);
  g.push_back(c->super_bits);
  g.push_back((uint64_t)c->max_super_passes);
  g.push_back((uint64_t)c->max_passes);
  g.push_back(c->key64 ? 1u : 0u);
  g.push_back(c->max_key_bits);
  g.push_back(c->tile_bits);
  // per slot: the view, its size and pixel offset (the loss launches take
  // them by value)
  for (size_t s = 0; s < c->slot_views.size(); ++s) {
    const int v = c->slot_views[s];
    g.push_back((uint64_t)(uint32_t)v);
    g.push_back(((uint64_t)c->views[v].W << 32) | c->views[v].H);
    g.push_back(c->views[v].pix_off);
  }
  uint64_t w[(sizeof(sv_config) + 7) / 8] = {0};
  std::memcpy(w, &c->cfg, sizeof(sv_config));
  for (uint64_t x : w) g.push_back(x);
  return g;
}

void drop_graph(sv_ctx *c) {
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  c->gexec = nullptr;
  c->graph = nullptr;
  c->gsig.clear();
}

// Capture the batched step into a graph unless the current one matches.
sv_status ensure_graph(sv_ctx *c) {
  std::vector<uint64_t> sig = graph_signature(c);
  if (c->gexec && sig == c->gsig) return SV_OK;
  drop_graph(c);
  HIP_OK(c, hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
  c->capturing = true;
  sv_status s = enqueue_batched_step(c, nullptr);
  c->capturing = false;
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(c->st, &g);
  if (s) {
    if (g) (void)hipGraphDestroy(g);
    return s;
  }
  if (e != hipSuccess || !g) return fail(c, SV_E_HIP, "graph capture failed: %s", hipGetErrorString(e));
  hipGraphExec_t x = nullptr;
  e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(g);
    return fail(c, SV_E_HIP, "graph instantiate failed: %s", hipGetErrorString(e));
  }
  c->graph = g;
  c->gexec = x;
  c->gsig.swap(sig);
  return SV_OK;
}

// One batched step. With use_graph (and timing off) the step is one graph
// launch; the Adam step counter only advances when the step ran. If the
// pair buffers were too small, nothing past the scan wrote any state: the
// buffers grow and the step runs again.
sv_status step_batched(sv_ctx *c, const int *vs, int nv) {
  if (sv_status s = prepare_batch(c, vs, nv)) return s;
  const uint32_t t = c->step_t + 1;
  const AdamParams hp = adam_params(c, t);
  for (int attempt = 0; attempt < 3; ++attempt) {
    if (c->cfg.use_graph && !c->timing) {
      if (sv_status s = ensure_graph(c)) return s;
      *c->h_adam = hp;
      HIP_OK(c, hipGraphLaunch(c->gexec, c->st));
    } else {
      if (sv_status s = enqueue_batched_step(c, &hp)) return s;
    }
    bool retry = false;
    if (sv_status s = read_meta(c, &retry)) return s;
    if (!retry) {
      c->step_t = t;
      c->built = true;
      c->forward_done = true;
      return SV_OK;
    }
  }
  return fail(c, SV_E_STATE, "batched step did not fit after growing the buffers");
}

double sum_l1(sv_ctx *c) {
  std::vector<float> t(c->tiles_total);
  if (c->tiles_total == 0) return 0.0;
  if (hipMemcpy(t.data(), c->l1_partial, c->tiles_total * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
    return -1.0;
  double s = 0.0;
  for (float v : t) s += v;
  return s;
}

// Replace the Gaussian arrays by a compacted / grown copy of n_new rows.
sv_status swap_in(sv_ctx *c, float *Pn, float *M1n, float *M2n, size_t cap_n, uint32_t n_new) {
  dfree(c->P);
  dfree(c->M1);
  dfree(c->M2);
  c->P = Pn;
  c->M1 = M1n;
  c->M2 = M2n;
  if (cap_n != c->cap) {
    dfree(c->G);
    dfree(c->geom);
    dfree(c->importance);
    dfree(c->grad_accum);
    if (sv_status s = dalloc(c, &c->G, (size_t)P_NUM * cap_n)) return s;
    if (sv_status s = dalloc(c, &c->geom, (size_t)5 * cap_n)) return s;
    if (sv_status s = dalloc(c, &c->importance, cap_n)) return s;
    if (sv_status s = dalloc(c, &c->grad_accum, cap_n)) return s;
    c->cap = cap_n;
  }
  c->N = n_new;
  // statistics restart after prune/densify (gaussian-renderer.ts:2990-2992)
  launch_fill_f32(c->st, c->G, 0.f, (size_t)P_NUM * c->cap);
  launch_fill_f32(c->st, c->geom, 0.f, (size_t)5 * c->cap);
  launch_fill_f32(c->st, (float *)c->importance, 0.f, (size_t)4 * c->cap);
  launch_fill_f32(c->st, (float *)c->grad_accum, 0.f, (size_t)4 * c->cap);
  c->slot_views.clear();
  c->built = false;
  c->forward_done = false;
  return check_launch(c, "swap_in");
}

// Compact by keep flags already on the device (keep8 bytes, keep32 words).
sv_status compact_by_keep(sv_ctx *c, uint8_t *keep8, uint32_t *keep32, size_t *n_removed) {
  uint32_t *offs = nullptr;
  uint64_t *tmp = nullptr;
  if (sv_status s = dalloc(c, &offs, c->N)) return s;
  if (sv_status s = dalloc(c, &tmp, scan_tmp_words(c->N))) {
    dfree(offs);
    return s;
  }
  launch_scan(c->st, keep32, offs, c->N, tmp, c->d_small);
  HIP_OK(c, hipMemcpyAsync(c->h_small, c->d_small, sizeof(uint64_t), hipMemcpyDeviceToHost, c->st));
  HIP_OK(c, hipStreamSynchronize(c->st));
  const uint32_t n_keep = (uint32_t)c->h_small[0];
  const size_t removed = c->N - n_keep;
  if (n_removed) *n_removed = removed;
  // applyPruneMask: nothing to do when nothing or everything is removed
  // (gaussian-renderer.ts:2922-2925)
  if (removed == 0 || n_keep == 0) {
    if (n_removed) *n_removed = 0;
    dfree(offs);
    dfree(tmp);
    return SV_OK;
  }
  float *Pn = nullptr, *M1n = nullptr, *M2n = nullptr;
  if (sv_status s = dalloc(c, &Pn, (size_t)P_NUM * c->cap)) return s;
  if (sv_status s = dalloc(c, &M1n, (size_t)P_NUM * c->cap)) return s;
  if (sv_status s = dalloc(c, &M2n, (size_t)P_NUM * c->cap)) return s;
  {
    Scoped t(c, PH_COMPACT);
    launch_compact(c->st, c->P, Pn, c->cap, c->cap, P_NUM, offs, keep8, c->N);
    launch_compact(c->st, c->M1, M1n, c->cap, c->cap, P_NUM, offs, keep8, c->N);
    launch_compact(c->st, c->M2, M2n, c->cap, c->cap, P_NUM, offs, keep8, c->N);
  }
  if (sv_status s = check_launch(c, "compact")) return s;
  HIP_OK(c, hipStreamSynchronize(c->st));
  dfree(offs);
  dfree(tmp);
  return swap_in(c, Pn, M1n, M2n, c->cap, n_keep);
}

}  // namespace

// =====================================================================
// C ABI
// =====================================================================
extern "C" {

const char *sv_version(void) { return "skyfall-mi355x 0.1 (gfx950)"; }

sv_status sv_default_config(sv_config *o) {
  if (!o) return SV_E_ARG;
  std::memset(o, 0, sizeof(*o));
  o->tile_size = 16;
  o->batch = 64;
  o->raster_px = 16;
  o->sh_deg = 3;
  o->near_plane = 0.2f;
  o->far_plane = 10000.0f;
  o->radius_clip = 512.0f;
  o->eps2d = 0.3f;
  o->w_l2 = 0.0f;
  o->w_l1 = 0.8f;
  o->w_ssim = 0.2f;
  o->lr_pos = 0.00016f;
  o->lr_sh = 0.0025f;
  o->lr_opacity = 0.025f;
  o->lr_scale = 0.005f;
  o->lr_rot = 0.001f;
  o->beta1 = 0.9f;
  o->beta2 = 0.999f;
  o->eps = 1e-8f;
  o->sh_update_interval = 16;
  o->pos_lr_schedule = 1;
  o->render_f16 = 0;
  o->importance_decay = 0.98f;
  o->persistent = 0;
  o->use_graph = 0;
  o->keep_grads = 0;
  o->plain_wf_flags = 0;
  o->xcd_remap = 1;
  o->binned_build = 1;
  o->geom_flags = 0;
  o->split_backward = 0;
  o->tile_order = 1;
  return SV_OK;
}

const char *sv_last_error(const sv_ctx *c) { return c ? c->err.c_str() : g_create_err.c_str(); }

sv_status sv_create(int device, size_t capacity, sv_ctx **out) {
  if (!out) return fail(nullptr, SV_E_ARG, "null out");
  *out = nullptr;
  if (capacity == 0) return fail(nullptr, SV_E_ARG, "capacity must be > 0");
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev == 0)
    return fail(nullptr, SV_E_HIP, "no HIP device available (%s)", hipGetErrorString(e));
  if (device < 0 || device >= ndev) return fail(nullptr, SV_E_ARG, "device %d out of range", device);
  sv_ctx *c = new sv_ctx();
  c->device = device;
  sv_default_config(&c->cfg);
  if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&c->st) != hipSuccess ||
      hipStreamCreate(&c->st2) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming) != hipSuccess) {
    fail(nullptr, SV_E_HIP, "cannot create streams and events on device %d", device);
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    if (c->ev_join) (void)hipEventDestroy(c->ev_join);
    if (c->st2) (void)hipStreamDestroy(c->st2);
    if (c->st) (void)hipStreamDestroy(c->st);
    delete c;
    return SV_E_HIP;
  }
  c->cap = capacity;
  {
    void *p = nullptr;
    const hipError_t e = noalloc_scratch(device, &p);
    if (e != hipSuccess) {
      fail(nullptr, SV_E_HIP, "scratch address on device %d: %s", device, hipGetErrorString(e));
      (void)hipEventDestroy(c->ev_fork);
      (void)hipEventDestroy(c->ev_join);
      (void)hipStreamDestroy(c->st2);
      (void)hipStreamDestroy(c->st);
      delete c;
      return SV_E_HIP;
    }
    c->noalloc = (unsigned char *)p;
  }
  sv_status s = alloc_gaussian_arrays(c, capacity);
  if (!s) s = dalloc(c, &c->d_small, 8);
  if (!s) s = dalloc(c, &c->d_meta, 4);
  if (!s) s = dalloc(c, &c->d_work, (size_t)NUM_WORK_QUEUES * WORK_QUEUE_STRIDE);
  if (!s) {
    const uint32_t bs[3] = {64, 128, 256};
    for (int k = 0; k < 2; ++k)
      for (int b = 0; b < 3; ++b) c->resident[k][b] = raster_resident_blocks(k == 1, bs[b]);
  }
  if (!s) s = dalloc(c, &c->d_adam, 1);
  if (!s && hipHostMalloc((void **)&c->h_small, 8 * sizeof(uint64_t), hipHostMallocDefault) != hipSuccess)
    s = fail(c, SV_E_NOMEM, "hipHostMalloc failed");
  if (!s && hipHostMalloc((void **)&c->h_meta, 3 * sizeof(StepMeta), hipHostMallocDefault) != hipSuccess)
    s = fail(c, SV_E_NOMEM, "hipHostMalloc failed");
  if (!s && (hipHostMalloc((void **)&c->h_adam, sizeof(AdamParams), hipHostMallocMapped) != hipSuccess ||
             hipHostGetDevicePointer((void **)&c->h_adam_dev, c->h_adam, 0) != hipSuccess))
    s = fail(c, SV_E_NOMEM, "mapped hipHostMalloc failed");
  if (s) {
    g_create_err = c->err;
    sv_destroy(c);
    return s;
  }
  clear_state(c);
  launch_fill_f32(c->st, c->P, 0.f, (size_t)P_NUM * c->cap);
  if (hipStreamSynchronize(c->st) != hipSuccess) {
    g_create_err = "initial clear failed";
    sv_destroy(c);
    return SV_E_HIP;
  }
  *out = c;
  return SV_OK;
}

sv_status sv_destroy(sv_ctx *c) {
  if (!c) return SV_E_ARG;
  (void)hipSetDevice(c->device);
  if (c->st) (void)hipStreamSynchronize(c->st);
  for (auto &tp : c->pending) {
    c->event_pool.push_back(tp.a);
    c->event_pool.push_back(tp.b);
  }
  for (auto e : c->event_pool) (void)hipEventDestroy(e);
  dfree(c->P);
  dfree(c->M1);
  dfree(c->M2);
  dfree(c->G);
  dfree(c->geom);
  dfree(c->importance);
  dfree(c->grad_accum);
  dfree(c->color);
  dfree(c->alpha);
  dfree(c->residual);
  dfree(c->target);
  dfree(c->last);
  dfree(c->d_views);
  dfree(c->tile_view);
  dfree(c->block_view);
  dfree(c->ranges);
  dfree(c->l1_partial);
  dfree(c->tile_l1);
  dfree(c->order_f);
  dfree(c->order_b);
  dfree(c->tile_cost);
  dfree(c->rec);
  dfree(c->aux);
  dfree(c->counts);
  dfree(c->offs);
  dfree(c->vrec);
  dfree(c->scan_tmp);
  drop_graph(c);
  dfree(c->d_small);
  dfree(c->d_meta);
  dfree(c->d_work);
  dfree(c->d_adam);
  if (c->h_meta) (void)hipHostFree(c->h_meta);
  if (c->h_adam) (void)hipHostFree(c->h_adam);
  dfree(c->keys);
  dfree(c->keys2);
  dfree(c->vals);
  dfree(c->vals2);
  dfree(c->sort_tmp);
  dfree(c->isect_rec);
  dfree(c->rflag);
  dfree(c->tmask);
  dfree(c->scounts);
  dfree(c->soffs);
  dfree(c->rects);
  dfree(c->tins);
  dfree(c->btot);
  dfree(c->sbtot);
  dfree(c->bbase);
  dfree(c->sbbase);
  dfree(c->super_view);
  dfree(c->sranges);
  dfree(c->diff);
  dfree(c->skeys);
  dfree(c->skeys2);
  dfree(c->svals);
  dfree(c->svals2);
  dfree(c->ssort_tmp);
  dfree(c->vw_color);
  dfree(c->vw_alpha);
  dfree(c->vw_last);
  dfree(c->nch);
  dfree(c->cbase);
  dfree(c->chunk_super);
  dfree(c->ccount);
  dfree(c->fixed);
  dfree(c->fixed_inv);
  if (c->h_small) (void)hipHostFree(c->h_small);
  if (c->st2) (void)hipStreamSynchronize(c->st2);
  if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
  if (c->ev_join) (void)hipEventDestroy(c->ev_join);
  if (c->st2) (void)hipStreamDestroy(c->st2);
  if (c->st) (void)hipStreamDestroy(c->st);
  delete c;
  return SV_OK;
}

sv_status sv_set_config(sv_ctx *c, const sv_config *k) {
  if (!c || !k) return SV_E_ARG;
  if (k->tile_size != 16) return fail(c, SV_E_ARG, "tile_size must be 16 (got %u)", k->tile_size);
  if (k->batch != 64 && k->batch != 128 && k->batch != 256)
    return fail(c, SV_E_ARG, "batch must be 64, 128 or 256 (got %u)", k->batch);
  if (k->raster_px != 8 && k->raster_px != 16 && k->raster_px != 32)
    return fail(c, SV_E_ARG, "raster_px must be 8, 16 or 32 (got %u)", k->raster_px);
  if (k->sh_deg > 3) return fail(c, SV_E_ARG, "sh_deg must be <= 3");
  if (k->plain_wf_flags > 3) return fail(c, SV_E_ARG, "plain_wf_flags must be a combination of SV_WF_*");
  if (k->xcd_remap > 1) return fail(c, SV_E_ARG, "xcd_remap must be 0 or 1");
  if (k->binned_build > 1) return fail(c, SV_E_ARG, "binned_build must be 0 or 1");
  static_assert(SV_GEOM_MEAN_Y_FORWARD == GEOM_MEAN_Y_FORWARD, "geometry flag values");
  if (k->split_backward > 1) return fail(c, SV_E_ARG, "split_backward must be 0 or 1");
  if (k->tile_order > 1) return fail(c, SV_E_ARG, "tile_order must be 0 or 1");
  if (k->geom_flags & ~SV_GEOM_MEAN_Y_FORWARD)
    return fail(c, SV_E_ARG, "geom_flags must be a combination of SV_GEOM_*");
  c->cfg = *k;
  return SV_OK;
}

sv_status sv_get_config(const sv_ctx *c, sv_config *k) {
  if (!c || !k) return SV_E_ARG;
  *k = c->cfg;
  return SV_OK;
}

sv_status sv_set_gaussians(sv_ctx *c, size_t n, const float *pos, const float *ls, const float *q,
                           const float *lg, const float *sh) {
  if (sv_status s = check_ctx(c)) return s;
  c->bin_unfit = false;
  if (!pos || !ls || !q || !lg || !sh) return fail(c, SV_E_ARG, "null parameter pointer");
  if (n > c->cap) return fail(c, SV_E_CAPACITY, "%zu Gaussians exceed capacity %zu", n, c->cap);
  if (n > 0xffffffffull) return fail(c, SV_E_CAPACITY, "too many Gaussians");
  if (sv_status s = check_params_finite(c, n, pos, ls, q, lg, sh)) return s;
  HIP_OK(c, hipStreamSynchronize(c->st));
  drop_graph(c);
  c->N = (uint32_t)n;
  if (sv_status s = set_params(c, c->P, pos, ls, q, lg, sh)) return s;
  clear_state(c);
  c->step_t = 0;
  c->slot_views.clear();
  c->built = false;
  c->forward_done = false;
  return sync(c);
}

sv_status sv_get_gaussians(sv_ctx *c, float *pos, float *ls, float *q, float *lg, float *sh) {
  if (sv_status s = check_ctx(c)) return s;
  return get_params(c, c->P, pos, ls, q, lg, sh);
}

size_t sv_get_count(const sv_ctx *c) { return c ? c->N : 0; }

sv_status sv_get_moments(sv_ctx *c, int which, float *pos, float *ls, float *q, float *lg, float *sh) {
  if (sv_status s = check_ctx(c)) return s;
  if (which != 1 && which != 2) return fail(c, SV_E_ARG, "which must be 1 or 2");
  return get_params(c, which == 1 ? c->M1 : c->M2, pos, ls, q, lg, sh);
}

sv_status sv_set_moments(sv_ctx *c, int which, const float *pos, const float *ls, const float *q,
                         const float *lg, const float *sh) {
  if (sv_status s = check_ctx(c)) return s;
  if (which != 1 && which != 2) return fail(c, SV_E_ARG, "which must be 1 or 2");
  if (sv_status s = check_params_finite(c, c->N, pos, ls, q, lg, sh)) return s;
  HIP_OK(c, hipStreamSynchronize(c->st));
  return set_params(c, which == 1 ? c->M1 : c->M2, pos, ls, q, lg, sh);
}

sv_status sv_get_step(const sv_ctx *c, uint32_t *t) {
  if (!c || !t) return SV_E_ARG;
  *t = c->step_t;
  return SV_OK;
}
sv_status sv_set_step(sv_ctx *c, uint32_t t) {
  if (!c) return SV_E_ARG;
  c->step_t = t;
  return SV_OK;
}

sv_status sv_set_views(sv_ctx *c, int nv, const sv_camera *cams) {
  if (sv_status s = check_ctx(c)) return s;
  c->bin_unfit = false;
  if (nv <= 0 || !cams) return fail(c, SV_E_ARG, "need at least one camera");
  HIP_OK(c, hipStreamSynchronize(c->st));
  drop_graph(c);
  std::vector<HostView> vs(nv);
  uint64_t off = 0;
  const uint32_t ts = c->cfg.tile_size;
  for (int i = 0; i < nv; ++i) {
    HostView &h = vs[i];
    h.cam = cams[i];
    const double w = std::round((double)cams[i].viewport[0]), hh = std::round((double)cams[i].viewport[1]);
    if (!(w >= 1.0) || !(hh >= 1.0) || w > 65535.0 * ts || hh > 65535.0 * ts)
      return fail(c, SV_E_ARG, "camera %d has an invalid viewport", i);
    h.W = (uint32_t)w;
    h.H = (uint32_t)hh;
    h.tiles_x = (h.W + ts - 1) / ts;
    h.tiles_y = (h.H + ts - 1) / ts;
    h.n_tiles = h.tiles_x * h.tiles_y;
    h.pix_off = off;
    h.has_target = false;
    off += (uint64_t)h.W * h.H;
  }
  dfree(c->color);
  dfree(c->alpha);
  dfree(c->residual);
  dfree(c->target);
  dfree(c->last);
  if (sv_status s = dalloc(c, &c->color, 3 * off)) return s;
  if (sv_status s = dalloc(c, &c->residual, 3 * off)) return s;
  if (sv_status s = dalloc(c, &c->target, 3 * off)) return s;
  if (sv_status s = dalloc(c, &c->alpha, off)) return s;
  if (sv_status s = dalloc(c, &c->last, off)) return s;
  launch_fill_f32(c->st, c->target, 0.f, 3 * off);
  launch_fill_f32(c->st, c->residual, 0.f, 3 * off);
  launch_fill_f32(c->st, c->color, 0.f, 3 * off);
  launch_fill_f32(c->st, c->alpha, 0.f, off);
  launch_fill_u32(c->st, c->last, NO_LAST, off);
  c->views.swap(vs);
  c->pix_total = off;
  c->residual_ready.assign(nv, 0);
  c->slot_views.clear();
  c->built = false;
  c->forward_done = false;
  return sync(c);
}

sv_status sv_set_target(sv_ctx *c, int v, const float *rgb) {
  if (sv_status s = check_ctx(c)) return s;
  if (sv_status s = check_view(c, v)) return s;
  if (!rgb) return fail(c, SV_E_ARG, "null target");
  if (sv_status s = check_finite(c, "target", rgb, 3 * (size_t)c->views[v].W * c->views[v].H)) return s;
  if (sv_status s = write_pixels3(c, c->target, v, rgb)) return s;
  c->views[v].has_target = true;
  return SV_OK;
}

sv_status sv_build_tiles(sv_ctx *c, int v, uint64_t *n_isect) {
  if (sv_status s = check_ctx(c)) return s;
  if (sv_status s = check_view(c, v)) return s;
  if (sv_status s = prepare_batch(c, &v, 1)) return s;
  if (sv_status s = build_batch(c)) return s;
  if (n_isect) *n_isect = c->I;
  return sync(c);
}

sv_status sv_forward(sv_ctx *c, int v) {
  if (sv_status s = check_ctx(c)) return s;
  if (sv_status s = check_view(c, v)) return s;
  if (sv_status s = ensure_view_built(c, v)) return s;
  if (sv_status s = forward_batch(c)) return s;
  c->residual_ready[v] = 0;
  return sync(c);
}

sv_status sv_loss(sv_ctx *c, int v) {
  if (sv_status s = check_ctx(c)) return s;
  if (sv_status s = check_view(c, v)) return s;
  if (!c->views[v].has_target) return fail(c, SV_E_STATE, "view %d has no target", v);
  if (!(c->forward_done && c->slot_views.size() == 1 && c->slot_views[0] == v))
    return fail(c, SV_E_STATE, "sv_forward(view %d) must run before sv_loss", v);
  if (sv_status s = loss_views(c)) return s;
  c->residual_ready[v] = 1;
  return sync(c);
}

sv_status sv_set_residual(sv_ctx *c, int v, const float *rgb) {
  if (sv_status s = check_ctx(c)) return s;
  if (sv_status s = check_view(c, v)) return s;
  if (!rgb) return fail(c, SV_E_ARG, "null residual");
  if (sv_status s = check_finite(c, "residual", rgb, 3 * (size_t)c->views[v].W * c->views[v].H)) return s;
  if (sv_status s = write_pixels3(c, c->residual, v, rgb)) return s;
  c->residual_ready[v] = 1;
  return SV_OK;
}

sv_status sv_backward(sv_ctx *c, int v) {
  if (sv_status s = check_ctx(c)) return s;
  if (sv_status s = check_view(c, v)) return s;
  if (!(c->forward_done && c->slot_views.size() == 1 && c->slot_views[0] == v))
    return fail(c, SV_E_STATE, "sv_forward(view %d) must run before sv_backward", v);
  const bool use_res = c->residual_ready[v] != 0;
  if (!use_res && batch_needs_residual_buffer(c))
    if (sv_status s = loss_views(c)) return s;
  if (sv_status s = backward_batch(c, use_res || batch_needs_residual_buffer(c))) return s;
  VjpArgs va = vjp_args(c);
  {
    Scoped t(c, PH_VJP);
    launch_gather_records(c->st, va);
    launch_vjp(c->st, va);
  }
  if (sv_status s = check_launch(c, "vjp")) return s;
  c->residual_ready[v] = 0;
  return sync(c);
}

sv_status sv_backward_fixed(sv_ctx *c, int v, uint32_t flags) {
  if (sv_status s = check_ctx(c)) return s;
  if (sv_status s = check_view(c, v)) return s;
  if (!(c->forward_done && c->slot_views.size() == 1 && c->slot_views[0] == v))
    return fail(c, SV_E_STATE, "sv_forward(view %d) must run before sv_backward_fixed", v);
  if (!c->residual_ready[v]) {
    if (!c->views[v].has_target) return fail(c, SV_E_STATE, "view %d has no target", v);
    if (sv_status s = loss_views(c)) return s;
  }
  if (sv_status s = backward_fixed_batch(c, flags, true)) return s;
  c->residual_ready[v] = 0;
  return sync(c);
}

sv_status sv_zero_grad(sv_ctx *c) {
  if (sv_status s = check_ctx(c)) return s;
  if (sv_status s = zero_grads(c)) return s;
  return sync(c);
}

sv_status sv_adam(sv_ctx *c) {
  if (sv_status s = check_ctx(c)) return s;
  if (sv_status s = adam_planes(c)) return s;
  return sync(c);
}

sv_status sv_step(sv_ctx *c, const int *vs, int nv, int mode, sv_step_stats *stats) {
  if (sv_status s = check_ctx(c)) return s;
  if (!vs || nv <= 0) return fail(c, SV_E_ARG, "no views");
  for (int i = 0; i < nv; ++i) {
    if (sv_status s = check_view(c, vs[i])) return s;
    if (!c->views[vs[i]].has_target) return fail(c, SV_E_STATE, "view %d has no target", vs[i]);
  }
  uint64_t total_isect = 0;
  double l1 = 0.0;
  sv_status s;
  if (mode == SV_MODE_PLAIN) {
    if ((s = zero_grads(c))) return s;
    for (int i = 0; i < nv; ++i) {
      const int v = vs[i];
      if ((s = prepare_batch(c, &v, 1))) return s;
      if ((s = build_batch(c))) return s;
      total_isect += c->I;
      if ((s = forward_batch(c))) return s;
      if ((s = loss_views(c))) return s;
      if ((s = backward_fixed_batch(c, c->cfg.plain_wf_flags, true))) return s;
    }
    if ((s = adam_planes(c))) return s;
  } else if (mode == SV_MODE_SEQUENTIAL) {
    if ((s = zero_grads(c))) return s;
    for (int i = 0; i < nv; ++i) {
      const int v = vs[i];
      if ((s = prepare_batch(c, &v, 1))) return s;
      if ((s = build_batch(c))) return s;
      total_isect += c->I;
      if ((s = forward_batch(c))) return s;
      const bool use_res = batch_needs_residual_buffer(c);
      if (use_res && (s = loss_views(c))) return s;
      if ((s = backward_batch(c, use_res))) return s;
      VjpArgs va = vjp_args(c);
      {
        Scoped t(c, PH_VJP);
        launch_gather_records(c->st, va);
        launch_vjp(c->st, va);
      }
      if ((s = check_launch(c, "vjp"))) return s;
      if (stats) l1 += sum_l1(c);
    }
    if ((s = adam_planes(c))) return s;
  } else if (mode == SV_MODE_BATCHED) {
    if ((s = step_batched(c, vs, nv))) return s;
    total_isect = c->I;
    if (stats) l1 = sum_l1(c);
  } else {
    return fail(c, SV_E_ARG, "unknown mode %d", mode);
  }
  if ((s = sync(c))) return s;
  if (stats) {
    stats->n_isect = total_isect;
    stats->n_views = (uint32_t)nv;
    stats->sort_passes = (uint32_t)c->passes;
    stats->key_bits = c->key_bits;
    stats->binned = c->last_binned ? 1u : 0u;
    stats->l1_sum = l1;
  }
  return SV_OK;
}

sv_status sv_render_batched(sv_ctx *c, const int *vs, int nv) {
  if (sv_status s = check_ctx(c)) return s;
  if (sv_status s = prepare_batch(c, vs, nv)) return s;
  if (sv_status s = build_batch(c)) return s;
  if (sv_status s = forward_batch(c)) return s;
  return sync(c);
}

sv_status sv_render_camera(sv_ctx *c, const sv_camera *cam, float *rgb, float *alpha) {
  if (sv_status s = check_ctx(c)) return s;
  if (!cam || !rgb) return fail(c, SV_E_ARG, "camera and output are required");
  if (c->N == 0) return fail(c, SV_E_STATE, "no Gaussians");
  const uint32_t ts = c->cfg.tile_size;
  const double w = std::round((double)cam->viewport[0]), hh = std::round((double)cam->viewport[1]);
  if (!(w >= 1.0) || !(hh >= 1.0) || w > 65535.0 * ts || hh > 65535.0 * ts)
    return fail(c, SV_E_ARG, "camera has an invalid viewport");
  HostView h{};
  h.cam = *cam;
  h.W = (uint32_t)w;
  h.H = (uint32_t)hh;
  h.tiles_x = (h.W + ts - 1) / ts;
  h.tiles_y = (h.H + ts - 1) / ts;
  h.n_tiles = h.tiles_x * h.tiles_y;
  h.pix_off = 0;
  h.has_target = false;
  const size_t P = (size_t)h.W * h.H;
  if (P > c->vw_cap || !c->vw_color) {
    dfree(c->vw_color);
    dfree(c->vw_alpha);
    dfree(c->vw_last);
    c->vw_cap = 0;
    if (sv_status s = dalloc(c, &c->vw_color, 3 * P)) return s;
    if (sv_status s = dalloc(c, &c->vw_alpha, P)) return s;
    if (sv_status s = dalloc(c, &c->vw_last, P)) return s;
    c->vw_cap = P;
  }
  // The camera renders as one more view whose pixels live in the viewer
  // buffers; the training views, their pixels and the last build are left
  // as they were (the next step rebuilds its batch).
  const int vid = (int)c->views.size();
  c->views.push_back(h);
  c->residual_ready.push_back(0);
  std::swap(c->color, c->vw_color);
  std::swap(c->alpha, c->vw_alpha);
  std::swap(c->last, c->vw_last);
  sv_status s = prepare_batch(c, &vid, 1);
  if (!s) s = build_batch(c);
  if (!s) s = forward_batch(c);
  if (!s) s = read_pixels3(c, c->color, vid, rgb);
  if (!s && alpha) {
    hipError_t e = hipMemcpy(alpha, c->alpha, P * sizeof(float), hipMemcpyDeviceToHost);
    if (e != hipSuccess) s = fail(c, SV_E_HIP, "viewer alpha readback: %s", hipGetErrorString(e));
  }
  std::swap(c->color, c->vw_color);
  std::swap(c->alpha, c->vw_alpha);
  std::swap(c->last, c->vw_last);
  c->views.pop_back();
  c->residual_ready.pop_back();
  c->slot_views.clear();
  c->built = false;
  c->forward_done = false;
  return s;
}

sv_status sv_render_camera_noalloc(sv_ctx *c, const sv_camera *cam, float *rgb, float *alpha) {
  if (sv_status s = check_ctx(c)) return s;
  if (!cam || !rgb) return fail(c, SV_E_ARG, "camera and output are required");
  if (c->N == 0) return fail(c, SV_E_STATE, "no Gaussians");
  if (!c->noalloc) return fail(c, SV_E_STATE, "no scratch on device %d", c->device);
  const uint32_t ts = c->cfg.tile_size;
  const double w = std::round((double)cam->viewport[0]), hh = std::round((double)cam->viewport[1]);
  if (!(w >= 1.0) || !(hh >= 1.0) || w > 65535.0 * ts || hh > 65535.0 * ts)
    return fail(c, SV_E_ARG, "camera has an invalid viewport");
  const uint32_t W = (uint32_t)w, H = (uint32_t)hh, N = c->N;
  const uint32_t tx = (W + ts - 1) / ts, ty = (H + ts - 1) / ts;
  const uint64_t P = (uint64_t)W * H, T = (uint64_t)tx * ty;
  if (N > NOALLOC_MAX_SPLATS)
    return fail(c, SV_E_CAPACITY, "%u Gaussians exceed the scratch render's maximum of %u", N, NOALLOC_MAX_SPLATS);
  if (P > NOALLOC_MAX_PIXELS || T > NOALLOC_MAX_TILES)
    return fail(c, SV_E_CAPACITY, "%ux%u exceeds the scratch render's maximum of %u pixels and %u tiles", W, H,
                NOALLOC_MAX_PIXELS, NOALLOC_MAX_TILES);
  // carve the scratch for the maxima (the same offsets on every call)
  uintptr_t off = 0;
  auto take = [&](size_t bytes) -> void * {
    void *p = c->noalloc + off;
    off += (bytes + 255) & ~(size_t)255;
    return p;
  };
  const size_t NS = NOALLOC_MAX_SPLATS, NP = NOALLOC_MAX_PAIRS, NX = NOALLOC_MAX_PIXELS, NT = NOALLOC_MAX_TILES;
  const size_t nbmax = project_blocks(NOALLOC_MAX_SPLATS) + 1;
  DevView *d_view = (DevView *)take(sizeof(DevView));
  StepMeta *m = (StepMeta *)take(4 * sizeof(StepMeta));
  float4 *rec = (float4 *)take(4 * NS * sizeof(float4));
  float2 *aux = (float2 *)take(NS * sizeof(float2));
  uint32_t *counts = (uint32_t *)take(NS * 4), *offs = (uint32_t *)take(NS * 4), *tins = (uint32_t *)take(NS * 4);
  uint2 *rects = (uint2 *)take(NS * sizeof(uint2));
  uint32_t *btot = (uint32_t *)take(nbmax * 4), *bbase = (uint32_t *)take(nbmax * 4);
  uint64_t *scan_tmp = (uint64_t *)take(scan_tmp_words(std::max(NS, NT + 1)) * 8);
  uint32_t *keys = (uint32_t *)take(NP * 4), *keys2 = (uint32_t *)take(NP * 4);
  uint32_t *vals = (uint32_t *)take(NP * 4), *vals2 = (uint32_t *)take(NP * 4);
  uint32_t *sort_tmp = (uint32_t *)take(sort_tmp_words<uint32_t>(NP) * 4);
  uint32_t *tile_view = (uint32_t *)take((NT + 1) * 4), *ranges = (uint32_t *)take((NT + 1) * 4);
  float *color = (float *)take(3 * NX * 4), *alph = (float *)take(NX * 4), *out = (float *)take(3 * NX * 4);
  uint32_t *last = (uint32_t *)take(NX * 4);
  if (off > NOALLOC_SCRATCH_BYTES)
    return fail(c, SV_E_STATE, "scratch layout needs %zu bytes of %zu", (size_t)off, (size_t)NOALLOC_SCRATCH_BYTES);
  // the one view, as prepare_batch sets up a batch of one
  DevView d{};
  std::memcpy(d.view, cam->view, sizeof(d.view));
  std::memcpy(d.proj, cam->proj, sizeof(d.proj));
  d.focal[0] = cam->focal[0];
  d.focal[1] = cam->focal[1];
  d.viewport[0] = cam->viewport[0];
  d.viewport[1] = cam->viewport[1];
  d.W = W;
  d.H = H;
  d.tiles_x = tx;
  d.tiles_y = ty;
  d.n_tiles = (uint32_t)T;
  d.tile_base = 0;
  d.pix_off = 0;
  d.super_x = (tx + BIN_S - 1) / BIN_S;
  d.super_base = 0;
  const DevConfig dc = dev_config(c->cfg);
  const uint32_t tile_bits = bitlen(T > 0 ? T - 1 : 0);
  const double far = (double)c->cfg.far_plane;
  const uint64_t dq_max = (far >= 0.0) ? (uint64_t)std::floor(std::min(far, 65535.0)) : 65535u;
  const uint32_t max_key_bits = tile_bits + bitlen(dq_max);  // <= 13 + 16 bits: a 32-bit key
  hipStream_t st = c->st;
  HIP_OK(c, hipMemcpyAsync(d_view, &d, sizeof(DevView), hipMemcpyHostToDevice, st));
  launch_fill_u32(st, (uint32_t *)m, 0u, 4 * sizeof(StepMeta) / 4);
  launch_fill_u32(st, tile_view, 0u, T);
  launch_project(st, c->P, c->cap, N, d_view, 1, dc, rec, aux, counts, &m->max_dq, nullptr, rects, tins, offs,
                 nullptr, btot, nullptr);
  launch_scan(st, btot, bbase, project_blocks(N), scan_tmp, &m->n_isect);
  launch_keyinfo(st, m, tile_bits, max_key_bits, NP);
  launch_emit<uint32_t>(st, rec, rects, aux, counts, offs, bbase, N, 1, d_view, m, keys, vals);
  launch_sort<uint32_t>(st, keys, vals, keys2, vals2, NP, (int)((max_key_bits + 7) / 8), sort_tmp, m);
  launch_ranges<uint32_t>(st, keys, keys2, NP, (uint32_t)T, m, ranges);
  RasterArgs a{};
  a.views = d_view;
  a.tile_view = tile_view;
  a.ranges = ranges;
  a.vals0 = vals;
  a.vals1 = vals2;
  a.meta = m;
  a.rec = rec;
  a.N = N;
  a.n_tiles_total = (uint32_t)T;
  a.cfg = dc;
  a.color = color;
  a.alpha = alph;
  a.last = last;
  a.bbase = bbase;
  a.nb = project_blocks(N);
  a.aux = aux;
  launch_forward(st, a, c->cfg.batch, (uint32_t)T, 16u);
  launch_interleave3(st, color, P, out);
  if (sv_status s = check_launch(c, "scratch render")) return s;
  StepMeta hm{};
  HIP_OK(c, hipMemcpyAsync(&hm, m, sizeof(StepMeta), hipMemcpyDeviceToHost, st));
  HIP_OK(c, hipStreamSynchronize(st));
  if (hm.overflow == 1)
    return fail(c, SV_E_CAPACITY, "%llu intersections exceed the scratch render's maximum of %u",
                (unsigned long long)hm.n_isect, NOALLOC_MAX_PAIRS);
  if (hm.overflow)
    return fail(c, SV_E_STATE, "scratch render: sort key of %u bits", hm.key_bits);
  HIP_OK(c, hipMemcpy(rgb, out, 3 * P * sizeof(float), hipMemcpyDeviceToHost));
  if (alpha) HIP_OK(c, hipMemcpy(alpha, alph, P * sizeof(float), hipMemcpyDeviceToHost));
  return SV_OK;
}

sv_status sv_read_image(sv_ctx *c, int v, float *rgb) {
  if (sv_status s = check_ctx(c)) return s;
  return read_pixels3(c, c->color, v, rgb);
}

sv_status sv_read_residual(sv_ctx *c, int v, float *rgb) {
  if (sv_status s = check_ctx(c)) return s;
  return read_pixels3(c, c->residual, v, rgb);
}

sv_status sv_read_alpha(sv_ctx *c, int v, float *out) {
  if (sv_status s = check_ctx(c)) return s;
  if (sv_status s = check_view(c, v)) return s;
  if (!out) return fail(c, SV_E_ARG, "null output");
  const HostView &hv = c->views[v];
  HIP_OK(c, hipStreamSynchronize(c->st));
  HIP_OK(c, hipMemcpy(out, c->alpha + hv.pix_off, (size_t)hv.W * hv.H * sizeof(float), hipMemcpyDeviceToHost));
  return SV_OK;
}

sv_status sv_read_last_ids(sv_ctx *c, int v, uint32_t *out) {
  if (sv_status s = check_ctx(c)) return s;
  if (sv_status s = check_view(c, v)) return s;
  if (!out) return fail(c, SV_E_ARG, "null output");
  const HostView &hv = c->views[v];
  HIP_OK(c, hipStreamSynchronize(c->st));
  HIP_OK(c, hipMemcpy(out, c->last + hv.pix_off, (size_t)hv.W * hv.H * sizeof(uint32_t), hipMemcpyDeviceToHost));
  return SV_OK;
}

sv_status sv_read_keys(sv_ctx *c, int v, uint32_t *keys, uint32_t *vals) {
  if (sv_status s = check_ctx(c)) return s;
  if (sv_status s = check_view(c, v)) return s;
  if (!(c->built && c->slot_views.size() == 1 && c->slot_views[0] == v))
    return fail(c, SV_E_STATE, "tiles of view %d are not the last build", v);
  HIP_OK(c, hipStreamSynchronize(c->st));
  const uint64_t I = c->I;
  if (vals && I) HIP_OK(c, hipMemcpy(vals, cur_vals(c), I * sizeof(uint32_t), hipMemcpyDeviceToHost));
  if (keys && I && c->last_binned) {
    // binned build: no key array; the key of entry i of tile t is (t,
    // depth_q of its splat), depth_q formed as emit_intersections does
    std::vector<uint32_t> rg((size_t)c->tiles_total + 1), vv(I);
    std::vector<float2> ax(c->N);
    HIP_OK(c, hipMemcpy(rg.data(), c->ranges, rg.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
    HIP_OK(c, hipMemcpy(vv.data(), cur_vals(c), I * sizeof(uint32_t), hipMemcpyDeviceToHost));
    HIP_OK(c, hipMemcpy(ax.data(), c->aux, (size_t)c->N * sizeof(float2), hipMemcpyDeviceToHost));
    for (uint32_t t = 0; t < c->tiles_total; ++t)
      for (uint32_t i = rg[t]; i < rg[t + 1] && i < I; ++i) {
        const float d = std::fmin(std::fmax(ax[std::min(vv[i], c->N - 1)].x, 0.0f), 65535.0f);
        keys[i] = ((t & 0xfffu) << 16) | ((uint32_t)d & 0xffffu);
      }
  } else if (keys && I) {
    const uint64_t dmask = c->dbits ? ((1ull << c->dbits) - 1ull) : 0ull;
    if (c->key64) {
      std::vector<uint64_t> k(I);
      HIP_OK(c, hipMemcpy(k.data(), cur_keys(c), I * sizeof(uint64_t), hipMemcpyDeviceToHost));
      for (uint64_t i = 0; i < I; ++i)
        keys[i] = (uint32_t)((((k[i] >> c->dbits) & 0xfffu) << 16) | (k[i] & dmask));
    } else {
      std::vector<uint32_t> k(I);
      HIP_OK(c, hipMemcpy(k.data(), cur_keys(c), I * sizeof(uint32_t), hipMemcpyDeviceToHost));
      for (uint64_t i = 0; i < I; ++i)
        keys[i] = (uint32_t)(((((uint64_t)k[i] >> c->dbits) & 0xfffu) << 16) | (k[i] & dmask));
    }
  }
  return SV_OK;
}

sv_status sv_read_tile_offsets(sv_ctx *c, int v, uint32_t *out) {
  if (sv_status s = check_ctx(c)) return s;
  if (sv_status s = check_view(c, v)) return s;
  if (!out) return fail(c, SV_E_ARG, "null output");
  if (!(c->built && c->slot_views.size() == 1 && c->slot_views[0] == v))
    return fail(c, SV_E_STATE, "tiles of view %d are not the last build", v);
  HIP_OK(c, hipStreamSynchronize(c->st));
  HIP_OK(c, hipMemcpy(out, c->ranges, ((size_t)c->tiles_total + 1) * sizeof(uint32_t), hipMemcpyDeviceToHost));
  return SV_OK;
}

sv_status sv_read_projections(sv_ctx *c, int v, float *out) {
  if (sv_status s = check_ctx(c)) return s;
  if (sv_status s = check_view(c, v)) return s;
  if (!out) return fail(c, SV_E_ARG, "null output");
  if (!(c->built && c->slot_views.size() == 1 && c->slot_views[0] == v))
    return fail(c, SV_E_STATE, "tiles of view %d are not the last build", v);
  HIP_OK(c, hipStreamSynchronize(c->st));
  const size_t N = c->N;
  std::vector<float4> r(4 * N);
  std::vector<float2> ax(N);
  std::vector<uint32_t> cnt(N);
  HIP_OK(c, hipMemcpy(r.data(), c->rec, 4 * N * sizeof(float4), hipMemcpyDeviceToHost));
  HIP_OK(c, hipMemcpy(ax.data(), c->aux, N * sizeof(float2), hipMemcpyDeviceToHost));
  HIP_OK(c, hipMemcpy(cnt.data(), c->counts, N * sizeof(uint32_t), hipMemcpyDeviceToHost));
  std::vector<float> pos(3 * N), lg(N);
  if (sv_status s = get_params(c, c->P, pos.data(), nullptr, nullptr, lg.data(), nullptr)) return s;
  for (size_t g = 0; g < N; ++g) {
    const float4 r0 = r[4 * g], r1 = r[4 * g + 1];
    float *o = out + 12 * g;
    if (ax[g].y == 0.0f) {  // culled (radius 0): the record is all zero
      for (int k = 0; k < 12; ++k) o[k] = 0.0f;
      continue;
    }
    o[0] = r0.x;
    o[1] = r0.y;
    o[2] = ax[g].y;
    o[3] = ax[g].x;
    o[4] = r0.z;
    o[5] = r0.w;
    o[6] = r1.x;
    std::memcpy(&o[7], &cnt[g], 4);
    o[8] = pos[3 * g];
    o[9] = pos[3 * g + 1];
    o[10] = pos[3 * g + 2];
    o[11] = lg[g];
  }
  return SV_OK;
}

sv_status sv_read_grads(sv_ctx *c, float *pos, float *ls, float *q, float *lg, float *sh, float *geom) {
  if (sv_status s = check_ctx(c)) return s;
  if (sv_status s = get_params(c, c->G, pos, ls, q, lg, sh)) return s;
  if (geom) {
    std::vector<float> t(c->N);
    for (int k = 0; k < 5; ++k) {
      HIP_OK(c, hipMemcpy(t.data(), c->geom + (size_t)k * c->cap, c->N * sizeof(float), hipMemcpyDeviceToHost));
      for (size_t g = 0; g < c->N; ++g) geom[5 * g + k] = t[g];
    }
  }
  return SV_OK;
}

sv_status sv_read_tile_grads_fixed(sv_ctx *c, int v, int32_t *out) {
  if (sv_status s = check_ctx(c)) return s;
  if (sv_status s = check_view(c, v)) return s;
  if (!out) return fail(c, SV_E_ARG, "null output");
  if (!c->fixed || !(c->slot_views.size() == 1 && c->slot_views[0] == v))
    return fail(c, SV_E_STATE, "no fixed-point backward for view %d", v);
  HIP_OK(c, hipStreamSynchronize(c->st));
  if (c->I) HIP_OK(c, hipMemcpy(out, c->fixed, c->I * 54 * sizeof(int32_t), hipMemcpyDeviceToHost));
  return SV_OK;
}

sv_status sv_read_stats(sv_ctx *c, float *imp, float *acc) {
  if (sv_status s = check_ctx(c)) return s;
  HIP_OK(c, hipStreamSynchronize(c->st));
  if (imp) HIP_OK(c, hipMemcpy(imp, c->importance, c->N * sizeof(float4), hipMemcpyDeviceToHost));
  if (acc) HIP_OK(c, hipMemcpy(acc, c->grad_accum, c->N * sizeof(float4), hipMemcpyDeviceToHost));
  return SV_OK;
}

sv_status sv_prune_mask(sv_ctx *c, const uint8_t *remove, size_t *n_removed) {
  if (sv_status s = check_ctx(c)) return s;
  if (!remove) return fail(c, SV_E_ARG, "null mask");
  if (c->N == 0) {
    if (n_removed) *n_removed = 0;
    return SV_OK;
  }
  uint8_t *rm = nullptr, *keep8 = nullptr;
  uint32_t *keep32 = nullptr;
  if (sv_status s = dalloc(c, &rm, c->N)) return s;
  if (sv_status s = dalloc(c, &keep8, c->N)) return s;
  if (sv_status s = dalloc(c, &keep32, c->N)) return s;
  HIP_OK(c, hipMemcpy(rm, remove, c->N, hipMemcpyHostToDevice));
  launch_u8_to_u32(c->st, rm, keep32, c->N, 1);
  std::vector<uint8_t> keep(c->N);
  for (size_t i = 0; i < c->N; ++i) keep[i] = remove[i] ? 0 : 1;
  HIP_OK(c, hipMemcpy(keep8, keep.data(), c->N, hipMemcpyHostToDevice));
  sv_status s = compact_by_keep(c, keep8, keep32, n_removed);
  dfree(rm);
  dfree(keep8);
  dfree(keep32);
  return s;
}

sv_status sv_prune_opacity(sv_ctx *c, float min_opacity, size_t *n_removed) {
  if (sv_status s = check_ctx(c)) return s;
  if (c->N == 0) {
    if (n_removed) *n_removed = 0;
    return SV_OK;
  }
  uint8_t *keep8 = nullptr;
  uint32_t *keep32 = nullptr;
  if (sv_status s = dalloc(c, &keep8, c->N)) return s;
  if (sv_status s = dalloc(c, &keep32, c->N)) return s;
  launch_prune_flags(c->st, c->P, c->cap, c->N, min_opacity, keep8, keep32);
  sv_status s = compact_by_keep(c, keep8, keep32, n_removed);
  dfree(keep8);
  dfree(keep32);
  return s;
}

sv_status sv_densify(sv_ctx *c, const uint8_t *actions, uint64_t seed, size_t *n_added) {
  if (sv_status s = check_ctx(c)) return s;
  if (!actions) return fail(c, SV_E_ARG, "null actions");
  const uint32_t N = c->N;
  if (N == 0) {
    if (n_added) *n_added = 0;
    return SV_OK;
  }
  size_t n_clone = 0, n_split = 0;
  for (uint32_t i = 0; i < N; ++i) {
    if (actions[i] > 2) return fail(c, SV_E_ARG, "action %u at %u is not 0, 1 or 2", actions[i], i);
    n_clone += actions[i] == 1;
    n_split += actions[i] == 2;
  }
  const size_t n_new = (size_t)N - n_split + n_clone + 2 * n_split;
  if (n_added) *n_added = n_new - N;
  if (n_clone == 0 && n_split == 0) return SV_OK;
  size_t cap_n = c->cap;
  if (n_new > cap_n) {
    // ensureCapacity (gaussian-renderer.ts:2044-2064): next power of two
    cap_n = 1;
    while (cap_n < n_new) cap_n <<= 1;
  }
  uint8_t *act = nullptr;
  uint32_t *flags = nullptr, *keep_offs = nullptr, *child_offs = nullptr;
  uint64_t *tmp = nullptr;
  if (sv_status s = dalloc(c, &act, N)) return s;
  if (sv_status s = dalloc(c, &flags, N)) return s;
  if (sv_status s = dalloc(c, &keep_offs, N)) return s;
  if (sv_status s = dalloc(c, &child_offs, N)) return s;
  if (sv_status s = dalloc(c, &tmp, scan_tmp_words(N))) return s;
  HIP_OK(c, hipMemcpy(act, actions, N, hipMemcpyHostToDevice));
  launch_u8_to_u32(c->st, act, flags, N, 2);
  launch_scan(c->st, flags, keep_offs, N, tmp, c->d_small);
  launch_u8_to_u32(c->st, act, flags, N, 3);
  launch_scan(c->st, flags, child_offs, N, tmp, c->d_small + 1);
  float *Pn = nullptr, *M1n = nullptr, *M2n = nullptr;
  if (sv_status s = dalloc(c, &Pn, (size_t)P_NUM * cap_n)) return s;
  if (sv_status s = dalloc(c, &M1n, (size_t)P_NUM * cap_n)) return s;
  if (sv_status s = dalloc(c, &M2n, (size_t)P_NUM * cap_n)) return s;
  {
    Scoped t(c, PH_COMPACT);
    launch_densify(c->st, c->P, c->M1, c->M2, c->cap, N, act, keep_offs, child_offs,
                   (uint32_t)(N - n_split), seed, Pn, M1n, M2n, cap_n);
  }
  if (sv_status s = check_launch(c, "densify")) return s;
  HIP_OK(c, hipStreamSynchronize(c->st));
  dfree(act);
  dfree(flags);
  dfree(keep_offs);
  dfree(child_offs);
  dfree(tmp);
  return swap_in(c, Pn, M1n, M2n, cap_n, (uint32_t)n_new);
}

sv_status sv_reset_opacity(sv_ctx *c, float max_opacity, size_t *n_clamped) {
  if (sv_status s = check_ctx(c)) return s;
  if (!(max_opacity > 0.0f && max_opacity < 1.0f)) return fail(c, SV_E_ARG, "max_opacity must be in (0, 1)");
  const float max_logit = (float)std::log((double)max_opacity / (1.0 - (double)max_opacity));
  launch_fill_u32(c->st, (uint32_t *)c->d_small, 0u, 2);
  launch_clamp_logit(c->st, c->P, c->cap, c->N, max_logit, (uint32_t *)c->d_small);
  if (sv_status s = check_launch(c, "reset_opacity")) return s;
  HIP_OK(c, hipMemcpyAsync(c->h_small, c->d_small, sizeof(uint64_t), hipMemcpyDeviceToHost, c->st));
  if (sv_status s = sync(c)) return s;
  if (n_clamped) *n_clamped = (size_t)(c->h_small[0] & 0xffffffffu);
  c->built = false;
  c->forward_done = false;
  return SV_OK;
}

sv_status sv_knn_init(sv_ctx *c, float scene_extent, float *density) {
  if (sv_status s = check_ctx(c)) return s;
  if (!(scene_extent >= 0.0f) || !std::isfinite(scene_extent))
    return fail(c, SV_E_ARG, "scene_extent must be finite and >= 0");
  const uint32_t N = c->N;
  if (N == 0) return SV_OK;
  const size_t cells = knn_grid_cells(N);
  KnnScratch w{};
  uint8_t *grid = nullptr;
  auto release = [&]() {
    dfree(grid);
    dfree(w.cnt);
    dfree(w.start);
    dfree(w.cell);
    dfree(w.rank);
    dfree(w.spos);
    dfree(w.scan_tmp);
    dfree(w.total);
    dfree(w.density);
  };
  sv_status s = dalloc(c, &grid, 64);
  if (!s) s = dalloc(c, &w.cnt, cells);
  if (!s) s = dalloc(c, &w.start, cells);
  if (!s) s = dalloc(c, &w.cell, N);
  if (!s) s = dalloc(c, &w.rank, N);
  if (!s) s = dalloc(c, &w.spos, N);
  if (!s) s = dalloc(c, &w.scan_tmp, scan_tmp_words(cells));
  if (!s) s = dalloc(c, &w.total, 1);
  if (!s) s = dalloc(c, &w.density, N);
  if (s) {
    release();
    return s;
  }
  w.grid = grid;
  hipError_t e = launch_knn_init(c->st, c->P, c->cap, N, scene_extent, w);
  if (e == hipSuccess && density)
    e = hipMemcpyAsync(density, w.density, (size_t)N * sizeof(float), hipMemcpyDeviceToHost, c->st);
  if (e == hipSuccess) e = hipStreamSynchronize(c->st);
  release();
  if (e != hipSuccess) return fail(c, SV_E_HIP, "knn_init: %s", hipGetErrorString(e));
  c->built = false;
  c->forward_done = false;
  return SV_OK;
}

sv_status sv_multiview_scores(sv_ctx *c, const int *vs, int nv, uint32_t *importance, float *pruning,
                              int *n_scored) {
  if (sv_status s = check_ctx(c)) return s;
  if (!vs || nv <= 0 || !importance || !pruning) return fail(c, SV_E_ARG, "views and outputs are required");
  if (n_scored) *n_scored = 0;
  std::vector<int> use;
  for (int i = 0; i < nv; ++i) {
    if (sv_status s = check_view(c, vs[i])) return s;
    if (c->views[vs[i]].has_target) use.push_back(vs[i]);
  }
  const uint32_t N = c->N;
  if (use.empty() || N == 0) {
    std::fill(importance, importance + N, 0u);
    std::fill(pruning, pruning + N, 0.0f);
    return SV_OK;
  }
  const int k = (int)use.size();
  if (sv_status s = prepare_batch(c, use.data(), k)) return s;
  if (sv_status s = build_batch(c)) return s;
  if (sv_status s = forward_batch(c)) return s;
  uint32_t max_pixels = 0;
  for (int i = 0; i < k; ++i) max_pixels = std::max(max_pixels, c->h_slots[i].W * c->h_slots[i].H);
  float *loss = nullptr, *pr = nullptr;
  MetricStats *stats = nullptr;
  uint32_t *counts = nullptr, *imp = nullptr;
  auto release = [&]() {
    dfree(loss);
    dfree(pr);
    dfree(stats);
    dfree(counts);
    dfree(imp);
  };
  sv_status s = dalloc(c, &loss, c->pix_total);
  if (!s) s = dalloc(c, &stats, (size_t)k);
  if (!s) s = dalloc(c, &counts, (size_t)k * N);
  if (!s) s = dalloc(c, &imp, N);
  if (!s) s = dalloc(c, &pr, N);
  if (s) {
    release();
    return s;
  }
  launch_multiview_scores(c->st, c->d_views, k, max_pixels, c->color, c->target, c->ranges, cur_vals(c), c->last,
                          c->rec, N, loss, stats, counts, imp, pr);
  hipError_t e = hipGetLastError();
  std::vector<MetricStats> hs(k);
  if (e == hipSuccess) e = hipMemcpyAsync(importance, imp, (size_t)N * sizeof(uint32_t), hipMemcpyDeviceToHost, c->st);
  if (e == hipSuccess) e = hipMemcpyAsync(pruning, pr, (size_t)N * sizeof(float), hipMemcpyDeviceToHost, c->st);
  if (e == hipSuccess) e = hipMemcpyAsync(hs.data(), stats, k * sizeof(MetricStats), hipMemcpyDeviceToHost, c->st);
  if (e == hipSuccess) e = hipStreamSynchronize(c->st);
  release();
  if (e != hipSuccess) return fail(c, SV_E_HIP, "multiview scores: %s", hipGetErrorString(e));
  if (n_scored) {
    int ns = 0;
    for (const MetricStats &m : hs) ns += m.valid ? 1 : 0;
    *n_scored = ns;
  }
  return SV_OK;
}

sv_status sv_densify_decide(sv_ctx *c, const int *vs, int nv, float percentile, float split_px,
                            size_t max_new, uint8_t *actions, size_t *n_clone, size_t *n_split) {
  if (sv_status s = check_ctx(c)) return s;
  if (!vs || nv <= 0 || !actions) return fail(c, SV_E_ARG, "views and actions are required");
  if (!(percentile >= 0.0f && percentile <= 1.0f)) return fail(c, SV_E_ARG, "percentile must be in [0, 1]");
  for (int i = 0; i < nv; ++i) {
    if (sv_status s = check_view(c, vs[i])) return s;
    if (!c->views[vs[i]].has_target) return fail(c, SV_E_STATE, "view %d has no target", vs[i]);
  }
  if (sv_status s = prepare_batch(c, vs, nv)) return s;
  if (sv_status s = build_batch(c)) return s;
  if (sv_status s = forward_batch(c)) return s;
  launch_tile_l1(c->st, c->d_views, c->tile_view, c->tiles_total, c->color, c->target, c->tile_l1);
  if (sv_status s = check_launch(c, "tile_l1")) return s;
  std::vector<float> l1(c->tiles_total);
  HIP_OK(c, hipMemcpyAsync(l1.data(), c->tile_l1, l1.size() * sizeof(float), hipMemcpyDeviceToHost, c->st));
  HIP_OK(c, hipStreamSynchronize(c->st));
  std::vector<float> thr(nv);
  c->scored_views.assign(vs, vs + nv);
  c->scored_base.resize(nv);
  for (int s = 0; s < nv; ++s) {
    const DevView &d = c->h_slots[s];
    c->scored_base[s] = d.tile_base;
    std::vector<float> v(l1.begin() + d.tile_base, l1.begin() + d.tile_base + d.n_tiles);
    std::sort(v.begin(), v.end());
    thr[s] = v.empty() ? 0.0f : v[(size_t)std::floor((double)percentile * (double)(v.size() - 1))];
  }
  float *d_thr = nullptr, *d_score = nullptr;
  uint8_t *d_act = nullptr;
  auto release = [&]() {
    dfree(d_thr);
    dfree(d_act);
    dfree(d_score);
  };
  sv_status st = dalloc(c, &d_thr, (size_t)nv);
  if (!st) st = dalloc(c, &d_act, (size_t)c->N);
  if (!st && max_new) st = dalloc(c, &d_score, (size_t)c->N);
  if (!st && hipMemcpy(d_thr, thr.data(), nv * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
    st = fail(c, SV_E_HIP, "threshold upload failed");
  if (!st) {
    launch_densify_decide(c->st, c->rec, c->aux, c->counts, c->N, c->d_views, nv, c->tile_l1, d_thr, split_px, d_act,
                          d_score);
    st = check_launch(c, "densify_decide");
  }
  if (!st && hipMemcpy(actions, d_act, c->N, hipMemcpyDeviceToHost) != hipSuccess)
    st = fail(c, SV_E_HIP, "actions readback failed");
  std::vector<uint32_t> cand;
  for (uint32_t i = 0; !st && i < c->N; ++i)
    if (actions[i]) cand.push_back(i);
  if (!st && max_new && cand.size() > max_new) {
    // performDensification's growth cap (gaussian-renderer.ts:3297-3306):
    // candidates by score, highest first, ties in index order (a stable
    // sort of the index-ordered list); the first max_new keep their action
    std::vector<float> score(c->N);
    if (hipMemcpy(score.data(), d_score, c->N * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
      st = fail(c, SV_E_HIP, "score readback failed");
    if (!st) {
      std::stable_sort(cand.begin(), cand.end(), [&](uint32_t a, uint32_t b) { return score[a] > score[b]; });
      for (size_t k = max_new; k < cand.size(); ++k) actions[cand[k]] = 0;
    }
  }
  release();
  if (st) return st;
  size_t nc = 0, ns = 0;
  for (uint32_t i = 0; i < c->N; ++i) {
    nc += actions[i] == 1;
    ns += actions[i] == 2;
  }
  if (n_clone) *n_clone = nc;
  if (n_split) *n_split = ns;
  return SV_OK;
}

sv_status sv_densify_decide_grad(sv_ctx *c, const sv_densify_rule *rule, const uint32_t *importance,
                                 uint32_t n_samples, size_t max_new, uint8_t *actions, size_t *n_clone,
                                 size_t *n_split) {
  if (sv_status s = check_ctx(c)) return s;
  if (!rule || !actions || (!importance && c->N)) return fail(c, SV_E_ARG, "rule, importance and actions are required");
  const double vals[6] = {rule->grad_threshold, rule->grad_abs_threshold, rule->size_threshold,
                          rule->min_visibility, rule->importance_threshold, rule->pixel_scale};
  for (double v : vals)
    if (!std::isfinite(v)) return fail(c, SV_E_ARG, "densify rule thresholds must be finite");
  // performDensification (gaussian-renderer.ts:3249-3255): thresholds in
  // pixels, visibility at least 1, importance over max(1, samples)
  DensifyRuleDev r;
  r.grad_px = rule->grad_threshold * rule->pixel_scale;
  r.abs_px = rule->grad_abs_threshold * rule->pixel_scale;
  r.size = rule->size_threshold;
  r.min_vis = std::max(1.0, rule->min_visibility);
  r.imp_threshold = rule->importance_threshold;
  r.samples = (double)std::max<uint32_t>(1u, n_samples);
  if (n_clone) *n_clone = 0;
  if (n_split) *n_split = 0;
  if (c->N == 0) return SV_OK;
  uint32_t *d_imp = nullptr;
  uint8_t *d_act = nullptr;
  double *d_score = nullptr;
  auto release = [&]() {
    dfree(d_imp);
    dfree(d_act);
    dfree(d_score);
  };
  sv_status st = dalloc(c, &d_imp, (size_t)c->N);
  if (!st) st = dalloc(c, &d_act, (size_t)c->N);
  if (!st) st = dalloc(c, &d_score, (size_t)c->N);
  if (!st && hipMemcpy(d_imp, importance, c->N * sizeof(uint32_t), hipMemcpyHostToDevice) != hipSuccess)
    st = fail(c, SV_E_HIP, "importance upload failed");
  if (!st) {
    launch_densify_decide_grad(c->st, c->grad_accum, c->P, c->cap, c->N, d_imp, r, d_act, d_score);
    st = check_launch(c, "densify_decide_grad");
  }
  if (!st && hipMemcpy(actions, d_act, c->N, hipMemcpyDeviceToHost) != hipSuccess)
    st = fail(c, SV_E_HIP, "actions readback failed");
  std::vector<uint32_t> cand;
  for (uint32_t i = 0; !st && i < c->N; ++i)
    if (actions[i]) cand.push_back(i);
  if (!st && max_new && cand.size() > max_new) {
    // the growth cap (:3297-3306): candidates by score, highest first, ties
    // in index order (the JS sort is stable); the first max_new are kept
    std::vector<double> score(c->N);
    if (hipMemcpy(score.data(), d_score, c->N * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
      st = fail(c, SV_E_HIP, "score readback failed");
    if (!st) {
      std::stable_sort(cand.begin(), cand.end(), [&](uint32_t a, uint32_t b) { return score[a] > score[b]; });
      for (size_t k = max_new; k < cand.size(); ++k) actions[cand[k]] = 0;
    }
  }
  release();
  if (st) return st;
  size_t nc = 0, ns = 0;
  for (uint32_t i = 0; i < c->N; ++i) {
    nc += actions[i] == 1;
    ns += actions[i] == 2;
  }
  if (n_clone) *n_clone = nc;
  if (n_split) *n_split = ns;
  return SV_OK;
}

sv_status sv_read_tile_l1(sv_ctx *c, int v, float *out) {
  if (sv_status s = check_ctx(c)) return s;
  if (sv_status s = check_view(c, v)) return s;
  if (!out) return fail(c, SV_E_ARG, "null output");
  for (size_t s = 0; s < c->scored_views.size(); ++s)
    if (c->scored_views[s] == v) {
      HIP_OK(c, hipStreamSynchronize(c->st));
      HIP_OK(c, hipMemcpy(out, c->tile_l1 + c->scored_base[s], c->views[v].n_tiles * sizeof(float),
                          hipMemcpyDeviceToHost));
      return SV_OK;
    }
  return fail(c, SV_E_STATE, "view %d was not scored by the last sv_densify_decide", v);
}

namespace {

struct Fnv {
  uint64_t h = 1469598103934665603ull;
  void add(const void *p, size_t n) {
    const unsigned char *b = (const unsigned char *)p;
    for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 1099511628211ull;
  }
};

const char kCkptMagic[8] = {'S', 'V', 'C', 'K', 'P', 'T', '0', '1'};

struct CkptHeader {
  char magic[8];
  uint32_t version, planes;
  uint64_t n;
  uint32_t step, config_bytes;
};

}  // namespace

sv_status sv_save_checkpoint(sv_ctx *c, const char *path) {
  if (sv_status s = check_ctx(c)) return s;
  if (!path) return fail(c, SV_E_ARG, "null path");
  HIP_OK(c, hipStreamSynchronize(c->st));
  FILE *f = std::fopen(path, "wb");
  if (!f) return fail(c, SV_E_ARG, "cannot open %s for writing", path);
  Fnv h;
  bool ok = true;
  auto put = [&](const void *p, size_t n) {
    if (ok && std::fwrite(p, 1, n, f) != n) ok = false;
    h.add(p, n);
  };
  CkptHeader hd{};
  std::memcpy(hd.magic, kCkptMagic, 8);
  hd.version = 1;
  hd.planes = P_NUM;
  hd.n = c->N;
  hd.step = c->step_t;
  hd.config_bytes = (uint32_t)sizeof(sv_config);
  put(&hd, sizeof(hd));
  put(&c->cfg, sizeof(sv_config));
  std::vector<float> plane(std::max<size_t>(c->N, 1) * 4);
  sv_status st = SV_OK;
  const float *arrays[3] = {c->P, c->M1, c->M2};
  for (int a = 0; a < 3 && !st; ++a)
    for (int k = 0; k < P_NUM && !st; ++k) {
      if (hipMemcpy(plane.data(), arrays[a] + (size_t)k * c->cap, c->N * sizeof(float), hipMemcpyDeviceToHost) !=
          hipSuccess)
        st = fail(c, SV_E_HIP, "checkpoint readback failed");
      else
        put(plane.data(), c->N * sizeof(float));
    }
  const float4 *stats[2] = {c->importance, c->grad_accum};
  for (int a = 0; a < 2 && !st; ++a) {
    if (hipMemcpy(plane.data(), stats[a], c->N * sizeof(float4), hipMemcpyDeviceToHost) != hipSuccess)
      st = fail(c, SV_E_HIP, "checkpoint readback failed");
    else
      put(plane.data(), c->N * sizeof(float4));
  }
  const uint64_t sum = h.h;
  if (ok && std::fwrite(&sum, 1, sizeof(sum), f) != sizeof(sum)) ok = false;
  if (std::fclose(f) != 0) ok = false;
  if (st) return st;
  if (!ok) return fail(c, SV_E_ARG, "writing %s failed", path);
  return SV_OK;
}

sv_status sv_load_checkpoint(sv_ctx *c, const char *path) {
  if (sv_status s = check_ctx(c)) return s;
  if (!path) return fail(c, SV_E_ARG, "null path");
  FILE *f = std::fopen(path, "rb");
  if (!f) return fail(c, SV_E_ARG, "cannot open %s", path);
  Fnv h;
  bool ok = true;
  auto get = [&](void *p, size_t n) {
    if (ok && std::fread(p, 1, n, f) != n) ok = false;
    if (ok) h.add(p, n);
  };
  CkptHeader hd{};
  get(&hd, sizeof(hd));
  if (!ok || std::memcmp(hd.magic, kCkptMagic, 8) != 0 || hd.version != 1 || hd.planes != (uint32_t)P_NUM ||
      hd.config_bytes != sizeof(sv_config)) {
    std::fclose(f);
    return fail(c, SV_E_ARG, "%s is not a version-1 checkpoint of this library", path);
  }
  // the file's size must match its count before anything is allocated
  const uint64_t per = (uint64_t)(3 * P_NUM + 8) * sizeof(float);
  long here = std::ftell(f);
  long size = -1;
  if (here >= 0 && std::fseek(f, 0, SEEK_END) == 0) size = std::ftell(f);
  if (size < 0 || std::fseek(f, here, SEEK_SET) != 0 || hd.n > 0xffffffffull ||
      (uint64_t)size != (uint64_t)here + sizeof(sv_config) + hd.n * per + sizeof(uint64_t)) {
    std::fclose(f);
    return fail(c, SV_E_ARG, "%s: size does not match its %llu Gaussians", path, (unsigned long long)hd.n);
  }
  sv_config cfg;
  get(&cfg, sizeof(cfg));
  const size_t n = (size_t)hd.n;
  std::vector<float> data((size_t)3 * P_NUM * n + 8 * n);
  get(data.data(), data.size() * sizeof(float));
  uint64_t sum = 0;
  const uint64_t want = h.h;
  if (ok && std::fread(&sum, 1, sizeof(sum), f) != sizeof(sum)) ok = false;
  std::fclose(f);
  if (!ok) return fail(c, SV_E_ARG, "%s is truncated", path);
  if (sum != want) return fail(c, SV_E_ARG, "%s: checksum mismatch", path);
  sv_config old = c->cfg;
  if (sv_status s = sv_set_config(c, &cfg)) {
    c->cfg = old;
    return s;
  }
  HIP_OK(c, hipStreamSynchronize(c->st));
  drop_graph(c);
  if (n > c->cap) {
    // a checkpoint saved after densify grew the scene: grow to the next
    // power of two, as densify does (ensureCapacity, gaussian-renderer.ts:2044-2064)
    size_t cap_n = 1;
    while (cap_n < n) cap_n <<= 1;
    float *Pn = nullptr, *M1n = nullptr, *M2n = nullptr;
    if (sv_status s = dalloc(c, &Pn, (size_t)P_NUM * cap_n)) return s;
    if (sv_status s = dalloc(c, &M1n, (size_t)P_NUM * cap_n)) {
      dfree(Pn);
      return s;
    }
    if (sv_status s = dalloc(c, &M2n, (size_t)P_NUM * cap_n)) {
      dfree(Pn);
      dfree(M1n);
      return s;
    }
    if (sv_status s = swap_in(c, Pn, M1n, M2n, cap_n, (uint32_t)n)) return s;
  }
  clear_state(c);
  float *arrays[3] = {c->P, c->M1, c->M2};
  const float *src = data.data();
  for (int a = 0; a < 3; ++a)
    for (int k = 0; k < P_NUM; ++k, src += n)
      HIP_OK(c, hipMemcpy(arrays[a] + (size_t)k * c->cap, src, n * sizeof(float), hipMemcpyHostToDevice));
  HIP_OK(c, hipMemcpy(c->importance, src, n * sizeof(float4), hipMemcpyHostToDevice));
  src += 4 * n;
  HIP_OK(c, hipMemcpy(c->grad_accum, src, n * sizeof(float4), hipMemcpyHostToDevice));
  c->N = (uint32_t)n;
  c->step_t = hd.step;
  c->slot_views.clear();
  c->built = false;
  c->forward_done = false;
  return sync(c);
}

sv_status sv_autotune(sv_ctx *c, const int *vs, int nv, int steps, sv_tune_entry *out, int *count) {
  if (sv_status s = check_ctx(c)) return s;
  if (!count) return fail(c, SV_E_ARG, "count is required");
  if (!vs || nv <= 0 || steps <= 0) return fail(c, SV_E_ARG, "views and a positive step count are required");
  for (int i = 0; i < nv; ++i) {
    if (sv_status s = check_view(c, vs[i])) return s;
    if (!c->views[vs[i]].has_target) return fail(c, SV_E_STATE, "view %d has no target", vs[i]);
  }
  const uint32_t batches[3] = {64, 128, 256};
  std::vector<sv_tune_entry> cand;
  // raster tile 16 x batch x persistent x XCD mapping; the 8- and 32-px
  // raster tiles (forward only) launch one workgroup per quadrant / block
  const uint32_t rasters[3] = {8, 16, 32};
  const uint32_t split0 = c->cfg.split_backward;
  for (uint32_t r : rasters)
    for (uint32_t b : batches)
      for (uint32_t p = 0; p < (r == 16 ? 2u : 1u); ++p)
        for (uint32_t x = 0; x < 2; ++x) cand.push_back(sv_tune_entry{b, p, x, r, split0, 0.0});
  // then the fastest of those with the split backward switched (it decides
  // only with two views or more): one candidate filled in after the sweep
  const bool try_split = nv >= 2;
  const size_t n_sweep = cand.size();
  if (try_split) cand.push_back(sv_tune_entry{0, 0, 0, 0, split0 ? 0u : 1u, 0.0});
  if (!out) {
    *count = (int)cand.size();
    return SV_OK;
  }
  // saved state: parameters, moments, statistics, step counter
  const size_t plane = (size_t)P_NUM * c->cap;
  float *sP = nullptr, *sM1 = nullptr, *sM2 = nullptr;
  float4 *sImp = nullptr, *sAcc = nullptr;
  sv_status st = dalloc(c, &sP, plane);
  if (!st) st = dalloc(c, &sM1, plane);
  if (!st) st = dalloc(c, &sM2, plane);
  if (!st) st = dalloc(c, &sImp, c->cap);
  if (!st) st = dalloc(c, &sAcc, c->cap);
  auto copy_all = [&](bool save) -> sv_status {
    float *dst[3] = {sP, sM1, sM2}, *src[3] = {c->P, c->M1, c->M2};
    for (int k = 0; k < 3; ++k)
      HIP_OK(c, hipMemcpyAsync(save ? dst[k] : src[k], save ? src[k] : dst[k], plane * sizeof(float),
                               hipMemcpyDeviceToDevice, c->st));
    HIP_OK(c, hipMemcpyAsync(save ? sImp : c->importance, save ? c->importance : sImp, c->cap * sizeof(float4),
                             hipMemcpyDeviceToDevice, c->st));
    HIP_OK(c, hipMemcpyAsync(save ? sAcc : c->grad_accum, save ? c->grad_accum : sAcc, c->cap * sizeof(float4),
                             hipMemcpyDeviceToDevice, c->st));
    return sync(c);
  };
  const sv_config saved_cfg = c->cfg;
  const uint32_t t0 = c->step_t;
  const bool timing0 = c->timing;
  c->timing = false;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (!st && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
    st = fail(c, SV_E_HIP, "hipEventCreate failed");
  if (!st) st = copy_all(true);
  for (size_t i = 0; i < cand.size() && !st; ++i) {
    if (i == n_sweep) {
      // the switched split backward on the sweep's fastest candidate
      size_t best = 0;
      for (size_t j = 1; j < n_sweep; ++j)
        if (cand[j].ms_per_step < cand[best].ms_per_step) best = j;
      const uint32_t sp = cand[i].split_backward;
      cand[i] = cand[best];
      cand[i].split_backward = sp;
      cand[i].ms_per_step = 0.0;
    }
    c->cfg = saved_cfg;
    c->cfg.batch = cand[i].batch;
    c->cfg.persistent = cand[i].persistent;
    c->cfg.xcd_remap = cand[i].xcd_remap;
    c->cfg.raster_px = cand[i].tile_size;
    c->cfg.split_backward = cand[i].split_backward;
    c->step_t = t0;
    if ((st = step_batched(c, vs, nv))) break;  // untimed: compiles, grows buffers, captures
    if ((st = copy_all(false))) break;
    c->step_t = t0;
    if (hipEventRecord(e0, c->st) != hipSuccess) st = fail(c, SV_E_HIP, "hipEventRecord failed");
    for (int k = 0; k < steps && !st; ++k) st = step_batched(c, vs, nv);
    if (st) break;
    float ms = 0.0f;
    if (hipEventRecord(e1, c->st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
        hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
      st = fail(c, SV_E_HIP, "tuner timing failed");
    cand[i].ms_per_step = (double)ms / steps;
    if (!st) st = copy_all(false);
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  dfree(sP);
  dfree(sM1);
  dfree(sM2);
  dfree(sImp);
  dfree(sAcc);
  c->step_t = t0;
  c->timing = timing0;
  c->cfg = saved_cfg;
  if (st) return st;
  size_t best = 0;
  for (size_t i = 1; i < cand.size(); ++i)
    if (cand[i].ms_per_step < cand[best].ms_per_step) best = i;
  c->cfg.batch = cand[best].batch;
  c->cfg.persistent = cand[best].persistent;
  c->cfg.xcd_remap = cand[best].xcd_remap;
  c->cfg.raster_px = cand[best].tile_size;
  c->cfg.split_backward = cand[best].split_backward;
  const int n = std::min(*count, (int)cand.size());
  for (int i = 0; i < n; ++i) out[i] = cand[i];
  *count = (int)cand.size();
  return SV_OK;
}

sv_status sv_memory_stats(const sv_ctx *c, uint64_t *device_bytes, uint64_t *peak_bytes) {
  if (!c) return SV_E_ARG;
  std::lock_guard<std::mutex> lock(g_alloc_mu);
  if (device_bytes) *device_bytes = c->dev_bytes;
  if (peak_bytes) *peak_bytes = c->dev_peak;
  return SV_OK;
}

sv_status sv_reset_memory_peak(sv_ctx *c) {
  if (!c) return SV_E_ARG;
  std::lock_guard<std::mutex> lock(g_alloc_mu);
  c->dev_peak = c->dev_bytes;
  return SV_OK;
}

sv_status sv_enable_timing(sv_ctx *c, int enable) {
  if (!c) return SV_E_ARG;
  c->timing = enable != 0;
  return SV_OK;
}

sv_status sv_timings(sv_ctx *c, sv_kernel_time *out, int *count) {
  if (!c || !count) return SV_E_ARG;
  if (sv_status s = sync(c)) return s;
  const int n = std::min(*count, kNumPhases);
  for (int i = 0; i < n && out; ++i) {
    std::memset(out[i].name, 0, sizeof(out[i].name));
    std::snprintf(out[i].name, sizeof(out[i].name), "%s", kPhaseNames[i]);
    out[i].total_ms = c->phase_ms[i];
    out[i].launches = c->phase_n[i];
  }
  *count = kNumPhases;
  return SV_OK;
}

sv_status sv_reset_timings(sv_ctx *c) {
  if (!c) return SV_E_ARG;
  if (sv_status s = sync(c)) return s;
  for (int i = 0; i < kNumPhases; ++i) {
    c->phase_ms[i] = 0.0;
    c->phase_n[i] = 0;
  }
  return SV_OK;
}

sv_status sv_sync(sv_ctx *c) {
  if (!c) return SV_E_ARG;
  return sync(c);
}

}  // extern "C"
