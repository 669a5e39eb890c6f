// sv_sort.hip -- stable LSD radix sort of (key, u32 value) pairs for gfx950.
//
// Replaces the reference's 4-pass WGSL sort (src/sort/sort.ts:260-370,
// src/sort/radix_sort.wgsl), whose ranking emulates 32-lane subgroup
// matches through LDS. Here each pass is reduce-then-scan with no
// inter-workgroup waiting:
//   count    per 4096-key tile digit counts -> hist[digit][tile]
//   rowsum   per digit total over tiles
//   rowscan  hist -> global output offset of (digit, tile)
//   scatter  wave64 ballot match ranking (8 ballots per key row), per-wave
//            digit counters, block scan, LDS staging, coalesced scatter
// Every pass is stable, so the whole sort equals a stable sort by key.
#include "sv_internal.hpp"

namespace sv {

constexpr int SORT_T = 256;
constexpr int SORT_ITEMS = 16;
constexpr int SORT_TILE = SORT_T * SORT_ITEMS;  // 4096 keys per workgroup
constexpr int NWAVE = SORT_T / 64;

// Pairs, tiles and whether pass `pass` runs, from the step metadata.
// dmask keeps the digit's bits below key_bits: bits above the key carry a
// payload (the binned build's tile rectangle) and must not order the pairs.
struct PassInfo {
  uint64_t n;
  uint32_t nb, shift, dmask;
  bool run;
};
__device__ __forceinline__ PassInfo pass_info(const StepMeta *m, uint32_t pass) {
  PassInfo p;
  p.n = m->n_isect;
  p.run = pass < m->passes;
  p.nb = (uint32_t)((p.n + SORT_TILE - 1) / SORT_TILE);
  p.shift = 8u * pass;
  const uint32_t live = m->key_bits > p.shift ? m->key_bits - p.shift : 0u;
  p.dmask = live >= 8u ? 0xffu : ((1u << live) - 1u);
  return p;
}

template <typename K, int RB>
__global__ __launch_bounds__(SORT_T) void k_sort_count(const K *__restrict__ keys, const StepMeta *m,
                                                      uint32_t pass, uint32_t *__restrict__ hist) {
  constexpr int RADIX = 1 << RB;
  const PassInfo pi = pass_info(m, pass);
  if (!pi.run || blockIdx.x >= pi.nb) return;
  const uint64_t n = pi.n;
  const uint32_t shift = pi.shift, nb = pi.nb, dmask = pi.dmask & (RADIX - 1);
  __shared__ uint32_t s_h[NWAVE][RADIX];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int i = threadIdx.x; i < NWAVE * RADIX; i += SORT_T) (&s_h[0][0])[i] = 0u;
  __syncthreads();
  const uint64_t base = (uint64_t)blockIdx.x * SORT_TILE + (uint64_t)wave * (64 * SORT_ITEMS);
  // A pass over at most 4 live key bits has at most 16 distinct digits:
  // there, lanes with equal digits are matched by ballots and counted with
  // one LDS add by the lowest of them (per-lane atomics would serialise on
  // the few addresses); otherwise one LDS atomic per key.
  // the tile's keys load together first
  K kk[SORT_ITEMS];
#pragma unroll
  for (int j = 0; j < SORT_ITEMS; ++j) {
    const uint64_t i = base + (uint64_t)j * 64 + lane;
    kk[j] = i < n ? keys[i] : (K)0;
  }
  if (m->key_bits <= shift + 4u) {
#pragma unroll
    for (int j = 0; j < SORT_ITEMS; ++j) {
      const uint64_t i = base + (uint64_t)j * 64 + lane;
      const bool valid = i < n;
      const uint32_t d = valid ? (uint32_t)(kk[j] >> shift) & dmask : 0u;
      uint64_t mk = __ballot(valid);
#pragma unroll
      for (int bit = 0; bit < 4; ++bit) {
        const bool on = (d >> bit) & 1u;
        const uint64_t bb = __ballot(on);
        mk &= on ? bb : ~bb;
      }
      if (valid && lane == __ffsll((unsigned long long)mk) - 1) atomicAdd(&s_h[wave][d], (uint32_t)__popcll(mk));
    }
  } else {
#pragma unroll
    for (int j = 0; j < SORT_ITEMS; ++j) {
      const uint64_t i = base + (uint64_t)j * 64 + lane;
      if (i < n) atomicAdd(&s_h[wave][(uint32_t)(kk[j] >> shift) & dmask], 1u);
    }
  }
  __syncthreads();
  for (int d = threadIdx.x; d < RADIX; d += SORT_T) {
    uint32_t c = 0;
#pragma unroll
    for (int w = 0; w < NWAVE; ++w) c += s_h[w][d];
    hist[(size_t)d * nb + blockIdx.x] = c;
  }
}

__device__ __forceinline__ uint32_t block_reduce_u32(uint32_t v, uint32_t *s_w) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += (uint32_t)__shfl_xor((int)v, o, 64);
  if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = v;
  __syncthreads();
  uint32_t t = 0;
  for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += s_w[w];
  __syncthreads();
  return t;
}

// One workgroup per digit: total of the digit over all tiles.
__global__ __launch_bounds__(SORT_T) void k_sort_rowsum(const uint32_t *__restrict__ hist,
                                                       const StepMeta *m, uint32_t pass,
                                                       uint32_t *__restrict__ dtot) {
  const PassInfo pi = pass_info(m, pass);
  if (!pi.run) return;
  const uint32_t nb = pi.nb;
  __shared__ uint32_t s_w[NWAVE];
  const uint32_t d = blockIdx.x;
  uint32_t s = 0;
  for (uint32_t b = threadIdx.x; b < nb; b += SORT_T) s += hist[(size_t)d * nb + b];
  s = block_reduce_u32(s, s_w);
  if (threadIdx.x == 0) dtot[d] = s;
}

// One workgroup per digit: hist row -> exclusive scan + digit base.
__global__ __launch_bounds__(SORT_T) void k_sort_rowscan(uint32_t *__restrict__ hist,
                                                        const StepMeta *m, uint32_t pass,
                                                        const uint32_t *__restrict__ dtot) {
  const PassInfo pi = pass_info(m, pass);
  if (!pi.run) return;
  const uint32_t nb = pi.nb;
  __shared__ uint32_t s_w[NWAVE];
  const uint32_t d = blockIdx.x;
  uint32_t part = 0;
  for (uint32_t e = threadIdx.x; e < d; e += SORT_T) part += dtot[e];
  uint32_t carry = block_reduce_u32(part, s_w);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (uint32_t base = 0; base < nb; base += SORT_T) {
    const uint32_t b = base + threadIdx.x;
    const uint32_t v = (b < nb) ? hist[(size_t)d * nb + b] : 0u;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t y = (uint32_t)__shfl_up((int)x, o, 64);
      if (lane >= o) x += y;
    }
    if (lane == 63) s_w[wave] = x;
    __syncthreads();
    uint32_t wbase = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < NWAVE; ++w) {
      if (w < wave) wbase += s_w[w];
      tot += s_w[w];
    }
    if (b < nb) hist[(size_t)d * nb + b] = carry + wbase + x - v;
    carry += tot;
    __syncthreads();
  }
}

template <typename K, int RB>
__global__ __launch_bounds__(SORT_T) void k_sort_scatter(const K *__restrict__ kin,
                                                        const uint32_t *__restrict__ vin,
                                                        K *__restrict__ kout,
                                                        uint32_t *__restrict__ vout,
                                                        const StepMeta *m, uint32_t pass,
                                                        const uint32_t *__restrict__ hist) {
  constexpr int RADIX = 1 << RB;
  constexpr int DPT = RADIX / SORT_T;  // digits per thread in the block scan
  const PassInfo pi = pass_info(m, pass);
  if (!pi.run || blockIdx.x >= pi.nb) return;
  const uint64_t n = pi.n;
  const uint32_t shift = pi.shift, nb = pi.nb, dmask = pi.dmask & (RADIX - 1);
  __shared__ uint32_t s_wcnt[NWAVE][RADIX];
  __shared__ uint32_t s_lbase[RADIX];
  __shared__ uint32_t s_gbase[RADIX];
  __shared__ K s_keys[SORT_TILE];
  __shared__ uint32_t s_vals[SORT_TILE];
  __shared__ uint32_t s_w[NWAVE];

  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int i = threadIdx.x; i < NWAVE * RADIX; i += SORT_T) (&s_wcnt[0][0])[i] = 0u;
  for (int d = threadIdx.x; d < RADIX; d += SORT_T) s_gbase[d] = hist[(size_t)d * nb + blockIdx.x];
  __syncthreads();

  const uint64_t tile0 = (uint64_t)blockIdx.x * SORT_TILE;
  const uint64_t base = tile0 + (uint64_t)wave * (64 * SORT_ITEMS);
  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  K key[SORT_ITEMS];
  uint32_t val[SORT_ITEMS];
  uint32_t rank[SORT_ITEMS];
  // all the tile's loads first, so they are in flight together (the
  // ranking below is ordered through LDS, item by item)
#pragma unroll
  for (int j = 0; j < SORT_ITEMS; ++j) {
    const uint64_t i = base + (uint64_t)j * 64 + lane;
    const bool valid = i < n;
    key[j] = valid ? kin[i] : (K)0;
    val[j] = valid ? vin[i] : 0u;
  }
#pragma unroll
  for (int j = 0; j < SORT_ITEMS; ++j) {
    const uint64_t i = base + (uint64_t)j * 64 + lane;
    const bool valid = i < n;
    const uint32_t d = (uint32_t)(key[j] >> shift) & dmask;
    // lanes holding the same digit: AND of ballots over the 8 digit bits
    uint64_t m = __ballot(valid);
#pragma unroll
    for (int bit = 0; bit < RB; ++bit) {
      const bool on = (d >> bit) & 1u;
      const uint64_t bb = __ballot(on);
      m &= on ? bb : ~bb;
    }
    uint32_t before = 0;
    if (valid) before = s_wcnt[wave][d];
    __builtin_amdgcn_wave_barrier();
    if (valid) {
      const int leader = __ffsll((unsigned long long)m) - 1;
      if (lane == leader) s_wcnt[wave][d] = before + (uint32_t)__popcll(m);
    }
    __builtin_amdgcn_wave_barrier();
    rank[j] = before + (uint32_t)__popcll(m & lt_mask);
  }
  __syncthreads();
  // per digit: prefix over waves, block total, block exclusive scan; each
  // thread owns DPT consecutive digits
  {
    uint32_t dtot[DPT];
    uint32_t tsum = 0;
#pragma unroll
    for (int u = 0; u < DPT; ++u) {
      const int d = threadIdx.x * DPT + u;
      uint32_t tot = 0;
#pragma unroll
      for (int w = 0; w < NWAVE; ++w) {
        const uint32_t c = s_wcnt[w][d];
        s_wcnt[w][d] = tot;
        tot += c;
      }
      dtot[u] = tot;
      tsum += tot;
    }
    uint32_t x = tsum;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t y = (uint32_t)__shfl_up((int)x, o, 64);
      if (lane >= o) x += y;
    }
    if (lane == 63) s_w[wave] = x;
    __syncthreads();
    uint32_t run = x - tsum;
#pragma unroll
    for (int w = 0; w < NWAVE; ++w)
      if (w < wave) run += s_w[w];
#pragma unroll
    for (int u = 0; u < DPT; ++u) {
      s_lbase[threadIdx.x * DPT + u] = run;
      run += dtot[u];
    }
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < SORT_ITEMS; ++j) {
    const uint64_t i = base + (uint64_t)j * 64 + lane;
    if (i < n) {
      const uint32_t d = (uint32_t)(key[j] >> shift) & dmask;
      const uint32_t lpos = s_lbase[d] + s_wcnt[wave][d] + rank[j];
      if (lpos < (uint32_t)SORT_TILE) {
        s_keys[lpos] = key[j];
        s_vals[lpos] = val[j];
      }
    }
  }
  __syncthreads();
  const uint32_t cnt = (uint32_t)((n - tile0) < (uint64_t)SORT_TILE ? (n - tile0) : SORT_TILE);
  for (uint32_t k = threadIdx.x; k < cnt; k += SORT_T) {
    const K kk = s_keys[k];
    const uint32_t d = (uint32_t)(kk >> shift) & dmask;
    const uint32_t g = s_gbase[d] + (k - s_lbase[d]);
    if ((uint64_t)g < n) {
      kout[g] = kk;
      vout[g] = s_vals[k];
    }
  }
}

// 8-bit digits: a 10-bit digit saves a pass for 17-20 bit keys but its
// per-tile histogram (1024 x tiles) and 56 KB of LDS staging made each pass
// about six times slower on gfx950, so every pass uses 8 bits.
constexpr int DIGIT_BITS = 8;

template <typename K>
size_t sort_tmp_words(uint64_t n) {
  const uint64_t nb = (n + SORT_TILE - 1) / SORT_TILE;
  return (size_t)(nb * (1u << DIGIT_BITS) + (1u << DIGIT_BITS));
}

// Passes are launched for `cap` pairs; blocks past the live pair count and
// passes past m->passes return at once, so the same launches serve any
// pair count up to cap (and can be captured in a graph).
template <typename K>
void launch_sort(hipStream_t st, K *keys0, uint32_t *vals0, K *keys1, uint32_t *vals1, uint64_t cap,
                 int max_passes, uint32_t *tmp, const StepMeta *m) {
  const uint32_t nb = (uint32_t)((cap + SORT_TILE - 1) / SORT_TILE);
  if (nb == 0) return;
  constexpr int RADIX = 1 << DIGIT_BITS;
  uint32_t *hist = tmp;
  uint32_t *dtot = tmp + (size_t)nb * RADIX;
  for (int p = 0; p < max_passes; ++p) {
    K *kin = (p & 1) ? keys1 : keys0, *kout = (p & 1) ? keys0 : keys1;
    uint32_t *vin = (p & 1) ? vals1 : vals0, *vout = (p & 1) ? vals0 : vals1;
    hipLaunchKernelGGL((k_sort_count<K, DIGIT_BITS>), dim3(nb), dim3(SORT_T), 0, st, kin, m, (uint32_t)p,
                       hist);
    hipLaunchKernelGGL(k_sort_rowsum, dim3(RADIX), dim3(SORT_T), 0, st, hist, m, (uint32_t)p, dtot);
    hipLaunchKernelGGL(k_sort_rowscan, dim3(RADIX), dim3(SORT_T), 0, st, hist, m, (uint32_t)p, dtot);
    hipLaunchKernelGGL((k_sort_scatter<K, DIGIT_BITS>), dim3(nb), dim3(SORT_T), 0, st, kin, vin, kout,
                       vout, m, (uint32_t)p, hist);
  }
}

template size_t sort_tmp_words<uint32_t>(uint64_t);
template size_t sort_tmp_words<uint64_t>(uint64_t);
template void launch_sort<uint32_t>(hipStream_t, uint32_t *, uint32_t *, uint32_t *, uint32_t *,
                                    uint64_t, int, uint32_t *, const StepMeta *);
template void launch_sort<uint64_t>(hipStream_t, uint64_t *, uint32_t *, uint64_t *, uint32_t *,
                                    uint64_t, int, uint32_t *, const StepMeta *);

}  // namespace sv
