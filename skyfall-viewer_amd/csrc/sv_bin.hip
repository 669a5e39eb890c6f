// sv_bin.hip -- tile lists of a batch by two-level binning.
//
// Produces exactly what emit_intersections + the stable radix sort by
// (tile, depth_q) + build_tile_offsets produce (training-tiles.wgsl:182-271,
// sort.ts:260-370): for every tile the ids of the splats that list it,
// ordered by (depth_q, emission order), and the tile ranges. Emission order
// within one view is Gaussian order, and a Gaussian lists a tile at most
// once, so the order within a tile is (depth_q, Gaussian id).
//
// Instead of sorting all I (tile, splat) pairs, splats are binned into
// supertiles of 8 x 8 tiles: one pair per (splat, supertile) with key
// (global supertile, depth_q) and, above the key bits, the splat's tile
// rectangle clipped to the supertile (3 bits per bound); those pairs are
// radix-sorted (stable). Each supertile's sorted list is cut into chunks
// of BIN_CH pairs. A workgroup per chunk counts, per tile of the supertile,
// the pairs whose rectangle covers it (wave ballots); a prefix over the
// supertile's chunks gives each chunk its offset within every tile list
// and each tile its count; a scan of the counts gives the tile ranges; a
// second pass per chunk writes each pair into the lists of the tiles it
// covers, in list order. No global atomics, no serial walk of a long list.
#include "sv_internal.hpp"

namespace sv {

namespace {

constexpr uint32_t BIN_T = 256;

__device__ __forceinline__ uint32_t rect_lo(uint32_t v) { return v & 0xffffu; }
__device__ __forceinline__ uint32_t rect_hi(uint32_t v) { return v >> 16; }

// Pass info of the supertile pairs: supertile bits, dbits, key bits,
// passes, overflow. Runs after the scans of both counts and k_keyinfo.
// m[0]: tile pairs (n_isect from the count scan), m[1]: supertile pairs.
// Overflow of the tile build: 1 = grow (either buffer), 3 = the key and
// the rectangle bits do not fit 32 bits (the host falls back to sorting
// the tile pairs).
__global__ void k_bin_info(StepMeta *m, uint32_t super_bits, uint64_t cap_super) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  StepMeta &m0 = m[0];
  StepMeta &m1 = m[1];
  const uint32_t dbits = m0.dbits;
  const uint32_t kb = super_bits + dbits;
  const uint64_t n = m1.n_isect;
  uint32_t of = 0;
  if (n > cap_super || n >= 0xffffffffull) of = 1;
  else if (kb + BIN_RECT_BITS > 32u) of = 3;
  if (m0.overflow == 1 || of == 1) of = 1;
  m1.dbits = dbits;
  m1.key_bits = kb;
  m1.passes = (of || n <= 1) ? 0u : (kb + 7u) / 8u;
  m1.sel = m1.passes & 1u;
  m1.overflow = of;
  // the tile lists are written, not sorted, into buffer 0
  m0.passes = 0;
  m0.sel = 0;
  m0.overflow = of ? of : m0.overflow;
}

// Per (slot, Gaussian) item: its supertile pairs (supertiles row-major)
// from its first supertile-pair position on. One workgroup writes the pairs of
// 256 consecutive items as one contiguous span (binary search of the
// owning item per position), so the stores are whole lines.
__global__ __launch_bounds__(BIN_T) void k_emit_super(const uint2 *__restrict__ rects,
                                                      const float2 *__restrict__ aux,
                                                      const uint32_t *__restrict__ scounts,
                                                      const uint32_t *__restrict__ soffs,
                                                      const uint32_t *__restrict__ sbbase, uint32_t nb,
                                                      uint32_t N, int nv, const DevView *__restrict__ views,
                                                      const StepMeta *m, uint32_t *__restrict__ keys,
                                                      uint32_t *__restrict__ vals) {
  __shared__ uint32_t s_lo[BIN_T];    // first pair position (0xffffffff past the items)
  __shared__ uint32_t s_srect[BIN_T]; // supertile rect: sx0 | sw << 8 | sy0 << 16
  __shared__ uint32_t s_trect0[BIN_T];  // tile rect x0 | y0 << 16
  __shared__ uint32_t s_trect1[BIN_T];  // tile rect x1 | y1 << 16
  __shared__ uint32_t s_kbase[BIN_T];  // key of the item's first supertile, with depth_q
  __shared__ uint32_t s_g[BIN_T];
  __shared__ uint32_t s_sx[BIN_T];     // supertiles per row of the item's view
  __shared__ uint32_t s_span[2];
  if (m[0].overflow) return;
  const uint32_t dbits = m[1].dbits, kb = m[1].key_bits;
  const size_t M = (size_t)nv * N;
  const uint32_t t = threadIdx.x;
  const size_t vg = (size_t)blockIdx.x * BIN_T + t;
  uint32_t lo = 0xffffffffu, hi = 0;
  if (vg < M) {
    const uint32_t cnt = scounts[vg];
    const uint32_t s = (uint32_t)(vg / N), g = (uint32_t)(vg - (size_t)s * N);
    lo = block_pos(sbbase, nb, s, g, soffs[vg]);
    hi = lo + cnt;
    if (cnt) {
      const DevView &V = views[s];
      const uint2 rc = rects[vg];
      const uint32_t a = rc.x, b = rc.y;
      const uint32_t x0 = rect_lo(a), y0 = rect_hi(a), x1 = rect_lo(b), y1 = rect_hi(b);
      const uint32_t sx0 = x0 / BIN_S, sy0 = y0 / BIN_S;
      const uint32_t sx1 = (x1 - 1u) / BIN_S + 1u;
      s_srect[t] = sx0 | ((sx1 - sx0) << 8) | (sy0 << 16);
      s_trect0[t] = a;
      s_trect1[t] = b;
      const uint32_t dq = f2u_sat(clampf(aux[vg].x, 0.0f, 65535.0f)) & 0xffffu;
      s_kbase[t] = ((V.super_base + sy0 * V.super_x + sx0) << dbits) | dq;
      s_g[t] = g;
      s_sx[t] = V.super_x;
    }
  }
  s_lo[t] = lo;
  if (t == 0) s_span[0] = lo;
  const size_t last = min((size_t)blockIdx.x * BIN_T + (BIN_T - 1), M - 1);
  if (vg == last) s_span[1] = hi;
  __syncthreads();
  const uint32_t span_lo = s_span[0], span_hi = s_span[1];
  for (uint32_t k = span_lo + t; k < span_hi; k += BIN_T) {
    uint32_t j = 0;
#pragma unroll
    for (uint32_t step = BIN_T / 2; step > 0; step >>= 1)
      if (s_lo[j + step] <= k) j += step;
    const uint32_t local = k - s_lo[j];
    const uint32_t sr = s_srect[j], sw = (sr >> 8) & 0xffu;
    // local < 2^16 and sw < 2^8: (local + 0.5) / sw in float is at least
    // 0.5 / sw from an integer, far beyond its rounding error, so the
    // truncation is the integer quotient
    const uint32_t dy = (uint32_t)(((float)local + 0.5f) * __builtin_amdgcn_rcpf((float)sw));
    const uint32_t dx = local - dy * sw;
    const uint32_t sx = (sr & 0xffu) + dx, sy = (sr >> 16) + dy;
    // the tile rectangle clipped to supertile (sx, sy), local to it
    const uint32_t a = s_trect0[j], b = s_trect1[j];
    const uint32_t bx = sx * BIN_S, by = sy * BIN_S;
    const uint32_t lx0 = max(rect_lo(a), bx) - bx, lx1 = min(rect_lo(b), bx + BIN_S) - 1u - bx;
    const uint32_t ly0 = max(rect_hi(a), by) - by, ly1 = min(rect_hi(b), by + BIN_S) - 1u - by;
    const uint32_t r = lx0 | (lx1 << 3) | (ly0 << 6) | (ly1 << 9);
    keys[k] = (s_kbase[j] + ((dy * s_sx[j] + dx) << dbits)) | (r << kb);
    vals[k] = s_g[j];
  }
}

// Chunks of the supertile lists: nch[s] = ceil(len / BIN_CH) per
// supertile (sranges from the sorted pairs); after the scan of nch,
// chunk_super maps each chunk to its supertile.
__global__ __launch_bounds__(BIN_T) void k_chunk_counts(const uint32_t *__restrict__ sranges, uint32_t n_super_tiles,
                                                        const StepMeta *m, uint32_t *__restrict__ nch) {
  const uint32_t s = blockIdx.x * BIN_T + threadIdx.x;
  if (s > n_super_tiles) return;
  uint32_t c = 0;
  if (s < n_super_tiles && !m[0].overflow) {
    const uint32_t n1 = (uint32_t)m[1].n_isect;
    const uint32_t lo = min(sranges[s], n1), hi = max(lo, min(sranges[s + 1], n1));
    c = (hi - lo + BIN_CH - 1) / BIN_CH;
  }
  nch[s] = c;  // nch[n_super_tiles] = 0: the scan's total lands past the end
}

__global__ __launch_bounds__(BIN_T) void k_chunk_map(const uint32_t *__restrict__ nch,
                                                     const uint32_t *__restrict__ cbase, uint32_t n_super_tiles,
                                                     uint32_t *__restrict__ chunk_super) {
  const uint32_t s = blockIdx.x * BIN_T + threadIdx.x;
  if (s >= n_super_tiles) return;
  const uint32_t b = cbase[s], c = nch[s];
  for (uint32_t k = 0; k < c; ++k) chunk_super[b + k] = s;
}

// The pairs of chunk c: supertile s, list positions [lo, hi).
struct Chunk {
  uint32_t s, k, lo, hi;
  bool live;
};
__device__ __forceinline__ Chunk chunk_of(uint32_t c, const StepMeta *m, const uint32_t *chunk_super,
                                          const uint32_t *cbase, const uint32_t *sranges) {
  Chunk ch{};
  ch.live = !m[0].overflow && (uint64_t)c < m[3].n_isect;
  if (!ch.live) return ch;
  ch.s = chunk_super[c];
  ch.k = c - cbase[ch.s];
  const uint32_t n1 = (uint32_t)m[1].n_isect;
  const uint32_t s0 = min(sranges[ch.s], n1), s1 = max(s0, min(sranges[ch.s + 1], n1));
  ch.lo = min(s0 + ch.k * BIN_CH, s1);
  ch.hi = min(ch.lo + BIN_CH, s1);
  return ch;
}

// Thread i of a chunk workgroup takes pair lo + i (wave w pairs lo + 64 w
// .. lo + 64 w + 63). Loads the lane's key and Gaussian id (the key's bits
// from kb on hold the pair's rectangle: x0, x1, y0, y1 in bits 0-2, 3-5,
// 6-8, 9-11) and returns whether the lane has a pair. The loads are
// unconditional (lanes past the chunk's end read its last pair, then
// unused), so no branch makes the wave wait for them before their use.
__device__ __forceinline__ bool chunk_pair(const Chunk &ch, const uint32_t *keys, const uint32_t *vals, uint32_t *key,
                                           uint32_t *g) {
  const uint32_t k = ch.lo + threadIdx.x;
  if (ch.hi <= ch.lo) return false;  // uniform
  const uint32_t kc = min(k, ch.hi - 1u);
  *key = keys[kc];
  *g = vals[kc];
  return k < ch.hi;
}
// The 64 tiles of the supertile a pair's rectangle covers, as a bit mask
// (bit lx + 8 ly).
__device__ __forceinline__ uint64_t cover_mask(uint32_t r) {
  const uint32_t x0 = r & 7u, x1 = (r >> 3) & 7u, y0 = (r >> 6) & 7u, y1 = (r >> 9) & 7u;
  const uint64_t cols = (uint64_t)((0xffu >> (7u - (x1 - x0))) << x0) * 0x0101010101010101ull;
  const uint64_t below = y1 == 7u ? ~0ull : ((1ull << (8u * (y1 + 1u))) - 1ull);
  const uint64_t above = (1ull << (8u * y0)) - 1ull;
  return cols & below & ~above;
}

// One block-swap stage of the transpose on a 32-bit half: lanes with bit
// S set take the upper S-bit blocks of their own word and the lower
// blocks of the partner's (lane ^ S), shifted down; the others the
// reverse.
template <int S, uint32_t MASK>
__device__ __forceinline__ uint32_t tstage(uint32_t x, uint32_t y) {
  return (threadIdx.x & S) ? ((x & MASK) | ((y >> S) & ~MASK)) : ((x & ~MASK) | ((y << S) & MASK));
}

// 64 x 64 bit transpose across the wave: lane l holds row l (bit t =
// entry (l, t)); returns column l (bit p = entry (p, l)), i.e. lane t gets
// the ballot of bit t. Six block-swap stages: v_permlane32_swap (halves),
// v_permlane16_swap (rows), row_ror:8, ds_swizzle xor 4, quad_perm.
__device__ __forceinline__ uint64_t wave_transpose64(uint64_t row) {
  uint32_t lo = (uint32_t)row, hi = (uint32_t)(row >> 32);
  {
    const auto r = __builtin_amdgcn_permlane32_swap(lo, hi, false, false);
    lo = r[0];
    hi = r[1];
  }
  uint32_t w[2] = {lo, hi};
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    uint32_t x = w[h];
    const auto r = __builtin_amdgcn_permlane16_swap(x, x, false, false);
    x = tstage<16, 0xFFFF0000u>(x, (threadIdx.x & 16) ? r[0] : r[1]);
    x = tstage<8, 0xFF00FF00u>(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x128, 0xF, 0xF, false));
    x = tstage<4, 0xF0F0F0F0u>(x, (uint32_t)__builtin_amdgcn_ds_swizzle((int)x, 0x101F));
    x = tstage<2, 0xCCCCCCCCu>(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x4E, 0xF, 0xF, false));
    x = tstage<1, 0xAAAAAAAAu>(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0xB1, 0xF, 0xF, false));
    w[h] = x;
  }
  return (uint64_t)w[0] | ((uint64_t)w[1] << 32);
}

__global__ __launch_bounds__(BIN_CH) void k_fill_count(const uint32_t *__restrict__ keys0,
                                                       const uint32_t *__restrict__ vals0,
                                                       const uint32_t *__restrict__ keys1,
                                                       const uint32_t *__restrict__ vals1,
                                                       const uint32_t *__restrict__ sranges,
                                                       const uint32_t *__restrict__ cbase,
                                                       const uint32_t *__restrict__ chunk_super, const StepMeta *m,
                                                       uint32_t *__restrict__ ccount) {
  __shared__ uint32_t s_cnt[BIN_CH / 64][64];
  const uint32_t sel = m[1].sel, kb = m[1].key_bits;
  const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  // a fixed grid strides over the chunks (one short workgroup per chunk
  // made the launch the cost)
  for (uint32_t c = blockIdx.x;; c += gridDim.x) {
    const Chunk ch = chunk_of(c, m, chunk_super, cbase, sranges);
    if (!ch.live) return;
    uint32_t key = 0, g = 0;
    const bool ok = chunk_pair(ch, sel ? keys1 : keys0, sel ? vals1 : vals0, &key, &g);
    // lane t: the wave's pairs covering tile t
    const uint64_t col = wave_transpose64(ok ? cover_mask(key >> kb) : 0ull);
    s_cnt[wave][lane] = (uint32_t)__popcll(col);
    __syncthreads();
    if (threadIdx.x < 64) {
      uint32_t tot = 0;
#pragma unroll
      for (int w = 0; w < BIN_CH / 64; ++w) tot += s_cnt[w][threadIdx.x];
      ccount[(size_t)c * 64 + threadIdx.x] = tot;
    }
    __syncthreads();
  }
}

// Per (supertile, tile of it): prefix of the chunk counts in chunk order
// (ccount becomes each chunk's offset within the tile's list) and the
// tile's count (tiles outside the view have none).
__global__ __launch_bounds__(64) void k_tile_totals(const uint32_t *__restrict__ nch,
                                                    const uint32_t *__restrict__ cbase,
                                                    const uint32_t *__restrict__ super_view,
                                                    const DevView *__restrict__ views,
                                                    uint32_t *__restrict__ ccount,
                                                    uint32_t *__restrict__ tcount) {
  const uint32_t s = blockIdx.x, t = threadIdx.x;
  const DevView &V = views[super_view[s]];
  const uint32_t ls = s - V.super_base;
  const uint32_t tx = (ls % V.super_x) * BIN_S + (t & 7u), ty = (ls / V.super_x) * BIN_S + (t >> 3);
  const uint32_t b = cbase[s], c = nch[s];
  uint32_t run = 0;
  // eight chunks per round: their counts load together, then the prefix
  constexpr uint32_t R = 8;
  for (uint32_t k0 = 0; k0 < c; k0 += R) {
    uint32_t v[R];
#pragma unroll
    for (uint32_t u = 0; u < R; ++u) v[u] = k0 + u < c ? ccount[(size_t)(b + k0 + u) * 64 + t] : 0u;
#pragma unroll
    for (uint32_t u = 0; u < R; ++u) {
      if (k0 + u < c) ccount[(size_t)(b + k0 + u) * 64 + t] = run;
      run += v[u];
    }
  }
  if (tx < V.tiles_x && ty < V.tiles_y) tcount[V.tile_base + ty * V.tiles_x + tx] = run;
}

// Second pass per chunk: each pair goes to position ranges[tile] + the
// chunk's offset + the count of covering pairs before it in the chunk.
// The chunk's entries are first staged in LDS grouped by tile (groups of
// tiles whose entries fit BIN_STAGE), then stored as contiguous runs, so
// every store instruction writes consecutive addresses.
constexpr uint32_t BIN_STAGE = 8192;

__global__ __launch_bounds__(BIN_CH) void k_fill_write(const uint32_t *__restrict__ keys0,
                                                       const uint32_t *__restrict__ vals0,
                                                       const uint32_t *__restrict__ keys1,
                                                       const uint32_t *__restrict__ vals1,
                                                       const uint32_t *__restrict__ sranges,
                                                       const uint32_t *__restrict__ cbase,
                                                       const uint32_t *__restrict__ chunk_super,
                                                       const uint32_t *__restrict__ super_view,
                                                       const DevView *__restrict__ views, const StepMeta *m,
                                                       const uint32_t *__restrict__ ccount,
                                                       const uint32_t *__restrict__ ranges,
                                                       uint32_t *__restrict__ out) {
  __shared__ uint32_t s_pre[BIN_CH / 64][64];  // per (wave, tile): covering pairs of earlier waves
  __shared__ uint32_t s_gid[BIN_CH / 64][64];  // per (wave, lane): the pair's Gaussian
  __shared__ uint32_t s_tot[64], s_base[64], s_end[64], s_soff[64];
  __shared__ uint32_t s_stage[BIN_STAGE];
  const Chunk ch = chunk_of(blockIdx.x, m, chunk_super, cbase, sranges);
  if (!ch.live) return;
  const uint32_t sel = m[1].sel, kb = m[1].key_bits;
  const uint32_t tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  uint32_t key = 0, g = 0;
  const bool ok = chunk_pair(ch, sel ? keys1 : keys0, sel ? vals1 : vals0, &key, &g);
  // while the pairs load, the first wave looks up its tile's list range and
  // the chunk's offset in it (a chain of dependent loads: view, ranges),
  // summed only after the transpose so that nothing waits for them here
  uint32_t rb = 0, co = 0, e = 0;
  if (tid < 64) {
    const DevView &V = views[super_view[ch.s]];
    const uint32_t ls = ch.s - V.super_base;
    const uint32_t tx = (ls % V.super_x) * BIN_S + (tid & 7u), ty = (ls / V.super_x) * BIN_S + (tid >> 3);
    if (tx < V.tiles_x && ty < V.tiles_y) {
      const uint32_t tg = V.tile_base + ty * V.tiles_x + tx;
      rb = ranges[tg];
      co = ccount[(size_t)blockIdx.x * 64 + tid];
      e = ranges[tg + 1];
    }
  }
  const uint64_t mine = ok ? cover_mask(key >> kb) : 0ull;
  const uint64_t col = wave_transpose64(mine);
  s_pre[wave][lane] = (uint32_t)__popcll(col);
  s_gid[wave][lane] = g;
  __syncthreads();
  if (tid < 64) {
    uint32_t run = 0;
#pragma unroll
    for (uint32_t w = 0; w < BIN_CH / 64; ++w) {
      const uint32_t c = s_pre[w][tid];
      s_pre[w][tid] = run;
      run += c;
    }
    s_tot[tid] = run;
    s_base[tid] = rb + co;
    s_end[tid] = e;
  }
  __syncthreads();
  // groups of tiles [t0, t1) whose entries fit the stage (one tile has at
  // most BIN_CH <= BIN_STAGE entries, so every group holds a tile). The
  // first wave scans the tile totals; with all of them within the stage
  // (the usual case) there is one group.
  __shared__ uint32_t s_incl[64];
  if (tid < 64) {
    uint32_t x = s_tot[tid];
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t y = (uint32_t)__shfl_up((int)x, o, 64);
      if (lane >= (uint32_t)o) x += y;
    }
    s_incl[tid] = x;
  }
  __syncthreads();
  for (uint32_t t0 = 0; t0 < 64;) {
    // the group is the longest run of tiles from t0 whose entries fit
    const uint32_t before = t0 ? s_incl[t0 - 1] : 0u;
    uint32_t t1 = t0 + 1;
    if (s_incl[63] - before <= BIN_STAGE) {
      t1 = 64;
    } else {
      while (t1 < 64 && s_incl[t1] - before <= BIN_STAGE) ++t1;
    }
    if (tid < 64) s_soff[tid] = (tid >= t0 && tid < t1) ? (tid ? s_incl[tid - 1] : 0u) - before : 0xffffffffu;
    __syncthreads();
    // lane t stages tile t's entries from this wave's pairs: the pairs
    // covering it (bits of its transposed column), in pair order, at the
    // tile's stage offset plus the covering pairs of earlier waves
    if (lane >= t0 && lane < t1) {
      uint32_t idx = s_soff[lane] + s_pre[wave][lane];
      for (uint64_t mm = col; mm; mm &= mm - 1ull, ++idx) {
        const uint32_t p = (uint32_t)__ffsll((unsigned long long)mm) - 1u;
        if (idx < BIN_STAGE) s_stage[idx] = s_gid[wave][p];
      }
    }
    __syncthreads();
    // each wave stores whole tiles: a tile's entries are one contiguous run
    // of its list
    for (uint32_t t = t0 + wave; t < t1; t += BIN_CH / 64) {
      const uint32_t n = s_tot[t], so = s_soff[t], b = s_base[t], e = s_end[t];
      for (uint32_t i = lane; i < n; i += 64) {
        const uint32_t pos = b + i;
        if (pos < e) out[pos] = s_stage[so + i];  // the counts bound every list
      }
    }
    __syncthreads();
    t0 = t1;
  }
}

}  // namespace

void launch_bin_info(hipStream_t st, StepMeta *m, uint32_t super_bits, uint64_t cap_super) {
  hipLaunchKernelGGL(k_bin_info, dim3(1), dim3(64), 0, st, m, super_bits, cap_super);
}

void launch_emit_super(hipStream_t st, const uint2 *rects, const float2 *aux, const uint32_t *scounts,
                       const uint32_t *soffs, const uint32_t *sbbase, uint32_t N, int nv, const DevView *views,
                       const StepMeta *m, uint32_t *keys, uint32_t *vals) {
  const size_t M = (size_t)nv * N;
  if (M == 0) return;
  hipLaunchKernelGGL(k_emit_super, dim3((uint32_t)((M + BIN_T - 1) / BIN_T)), dim3(BIN_T), 0, st, rects, aux,
                     scounts, soffs, sbbase, project_blocks(N), N, nv, views, m, keys, vals);
}

size_t bin_chunk_cap(uint64_t super_cap, uint32_t n_super_tiles) {
  return (size_t)(super_cap / BIN_CH) + n_super_tiles + 1;
}

void launch_fill_tiles(hipStream_t st, const BinArgs &b) {
  const uint32_t S = b.n_super_tiles;
  if (S == 0) return;
  hipLaunchKernelGGL(k_chunk_counts, dim3((S + 1 + BIN_T - 1) / BIN_T), dim3(BIN_T), 0, st, b.sranges, S, b.m,
                     b.nch);
  launch_scan(st, b.nch, b.cbase, (size_t)S + 1, b.scan_tmp, &b.m[3].n_isect);
  hipLaunchKernelGGL(k_chunk_map, dim3((S + BIN_T - 1) / BIN_T), dim3(BIN_T), 0, st, b.nch, b.cbase, S,
                     b.chunk_super);
  const uint32_t grid = (uint32_t)b.chunk_cap;
  hipLaunchKernelGGL(k_fill_count, dim3(std::min<uint32_t>(grid, 2048u)), dim3(BIN_CH), 0, st, b.keys0, b.vals0,
                     b.keys1, b.vals1, b.sranges, b.cbase, b.chunk_super, b.m, b.ccount);
  launch_fill_u32(st, b.tcount, 0u, (size_t)b.n_tiles + 1);
  hipLaunchKernelGGL(k_tile_totals, dim3(S), dim3(64), 0, st, b.nch, b.cbase, b.super_view, b.views, b.ccount,
                     b.tcount);
  launch_scan(st, b.tcount, b.ranges, (size_t)b.n_tiles + 1, b.scan_tmp, &b.m[2].n_isect);
  hipLaunchKernelGGL(k_fill_write, dim3(grid), dim3(BIN_CH), 0, st, b.keys0, b.vals0, b.keys1, b.vals1, b.sranges,
                     b.cbase, b.chunk_super, b.super_view, b.views, b.m, b.ccount, b.ranges, b.out);
}

}  // namespace sv
