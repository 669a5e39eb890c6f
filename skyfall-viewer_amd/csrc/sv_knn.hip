// sv_knn.hip -- density-aware initial scales: knn_init (knn-init.wgsl:40-90)
// on a uniform grid.
//
// The shader compares every point with every other (O(N^2)). Here the
// points are bucketed into cubic cells (counting sort by cell), and each
// query walks shells of cells of growing Chebyshev radius around its own
// cell until no unvisited cell can hold a point closer than its third
// nearest distance. The three smallest distances are the same numbers the
// full scan finds (the search is exact; ties only exchange equal values), so
// density and log-scale are the shader's, without its N^2 cost.
//
// Distances are compared squared: sqrt is monotone, so the k-th smallest
// length is sqrt of the k-th smallest squared length. The shader's slots
// start at 1e10 and only take strictly smaller lengths, so slot k ends at
// min(k-th smallest length, 1e10).
#include <math.h>

#include <algorithm>
#include <cmath>

#include "sv_internal.hpp"

namespace sv {

namespace {

constexpr int KT = 256;

struct KnnGrid {
  float lo[3];
  float h, inv_h;
  uint32_t g[3];
};

__device__ __forceinline__ float3 knn_pos(const float *__restrict__ P, size_t cap, uint32_t i) {
  // read_position (knn-init.wgsl:20-25): positions as stored in f16
  return make_float3(f16_round(P[(size_t)(P_POS + 0) * cap + i]), f16_round(P[(size_t)(P_POS + 1) * cap + i]),
                     f16_round(P[(size_t)(P_POS + 2) * cap + i]));
}

// Cell coordinate along one axis, clamped into the grid (points outside the
// grid box fall into its border cells; NaN goes to cell 0).
__device__ __forceinline__ int knn_cell(float x, float lo, float inv_h, uint32_t g) {
  const float t = (x - lo) * inv_h;
  if (!(t >= 0.0f)) return 0;
  if (t >= (float)g) return (int)g - 1;
  return min((int)t, (int)g - 1);
}

// Grid box: per axis [max(min, mean - 3 sd), min(max, mean + 3 sd)], so a
// few far outliers do not stretch the cells; cubic cells with at most
// gmax along the longest axis. One workgroup of 1024 threads.
__global__ __launch_bounds__(1024) void k_knn_bounds(const float *__restrict__ P, size_t cap, uint32_t N,
                                                     uint32_t gmax, KnnGrid *__restrict__ grid) {
  __shared__ float s_mn[3][16], s_mx[3][16];
  __shared__ double s_s[3][16], s_q[3][16];
  float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  double s[3] = {0, 0, 0}, q[3] = {0, 0, 0};
  for (uint32_t i = threadIdx.x; i < N; i += blockDim.x) {
    const float3 p = knn_pos(P, cap, i);
    const float v[3] = {p.x, p.y, p.z};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      if (!isfinite(v[a])) continue;
      mn[a] = fminf(mn[a], v[a]);
      mx[a] = fmaxf(mx[a], v[a]);
      s[a] += v[a];
      q[a] += (double)v[a] * v[a];
    }
  }
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    for (int o = 32; o > 0; o >>= 1) {
      mn[a] = fminf(mn[a], __shfl_xor(mn[a], o, 64));
      mx[a] = fmaxf(mx[a], __shfl_xor(mx[a], o, 64));
      s[a] += __shfl_xor(s[a], o, 64);
      q[a] += __shfl_xor(q[a], o, 64);
    }
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      s_mn[a][w] = mn[a];
      s_mx[a][w] = mx[a];
      s_s[a][w] = s[a];
      s_q[a][w] = q[a];
    }
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  const int nw = (int)(blockDim.x >> 6);
  float lo[3], hi[3], ext = 0.0f;
  for (int a = 0; a < 3; ++a) {
    float m0 = INFINITY, m1 = -INFINITY;
    double ss = 0, qq = 0;
    for (int k = 0; k < nw; ++k) {
      m0 = fminf(m0, s_mn[a][k]);
      m1 = fmaxf(m1, s_mx[a][k]);
      ss += s_s[a][k];
      qq += s_q[a][k];
    }
    if (!(m0 <= m1)) m0 = m1 = 0.0f;  // no finite coordinate
    const double mean = N ? ss / N : 0.0;
    const double sd = N ? sqrt(fmax(qq / N - mean * mean, 0.0)) : 0.0;
    lo[a] = fmaxf(m0, (float)(mean - 3.0 * sd));
    hi[a] = fminf(m1, (float)(mean + 3.0 * sd));
    if (!(lo[a] <= hi[a])) lo[a] = hi[a] = m0;
    ext = fmaxf(ext, hi[a] - lo[a]);
  }
  float h = ext / (float)gmax;
  if (!(h > 0.0f) || !isfinite(h)) h = 1.0f;
  for (int a = 0; a < 3; ++a) {
    grid->lo[a] = lo[a];
    const float cells = ceilf((hi[a] - lo[a]) / h);
    grid->g[a] = (uint32_t)fminf(fmaxf(cells, 1.0f), (float)gmax);
  }
  grid->h = h;
  grid->inv_h = 1.0f / h;
}

__device__ __forceinline__ uint32_t knn_cell_id(const KnnGrid &G, float3 p, int *c) {
  c[0] = knn_cell(p.x, G.lo[0], G.inv_h, G.g[0]);
  c[1] = knn_cell(p.y, G.lo[1], G.inv_h, G.g[1]);
  c[2] = knn_cell(p.z, G.lo[2], G.inv_h, G.g[2]);
  return ((uint32_t)c[2] * G.g[1] + (uint32_t)c[1]) * G.g[0] + (uint32_t)c[0];
}

__global__ __launch_bounds__(KT) void k_knn_count(const float *__restrict__ P, size_t cap, uint32_t N,
                                                  const KnnGrid *__restrict__ grid, uint32_t *__restrict__ cnt,
                                                  uint32_t *__restrict__ cell, uint32_t *__restrict__ rank) {
  const uint32_t i = blockIdx.x * KT + threadIdx.x;
  if (i >= N) return;
  const KnnGrid G = *grid;
  int c[3];
  const uint32_t id = knn_cell_id(G, knn_pos(P, cap, i), c);
  cell[i] = id;
  rank[i] = atomicAdd(&cnt[id], 1u);
}

__global__ __launch_bounds__(KT) void k_knn_scatter(const float *__restrict__ P, size_t cap, uint32_t N,
                                                    const uint32_t *__restrict__ start,
                                                    const uint32_t *__restrict__ cell,
                                                    const uint32_t *__restrict__ rank, float4 *__restrict__ spos) {
  const uint32_t i = blockIdx.x * KT + threadIdx.x;
  if (i >= N) return;
  const float3 p = knn_pos(P, cap, i);
  spos[start[cell[i]] + rank[i]] = make_float4(p.x, p.y, p.z, __uint_as_float(i));
}

__device__ __forceinline__ void knn_insert(float d2, float &b0, float &b1, float &b2) {
  if (d2 < b2) {
    if (d2 < b1) {
      b2 = b1;
      if (d2 < b0) {
        b1 = b0;
        b0 = d2;
      } else {
        b1 = d2;
      }
    } else {
      b2 = d2;
    }
  }
}

// One thread per point in cell order (neighbouring lanes search
// neighbouring cells). After the shell of radius r every unvisited point
// lies outside the (2r+1)^3 box of cells, at least as far as the nearest
// box face that is not the grid's border; the search stops once the third
// smallest squared distance is below that bound (less a margin of 1e-3
// cells and 1e-5 relative, which covers the rounding of the cell
// assignment) or the box covers the grid.
__global__ __launch_bounds__(KT) void k_knn_query(const float4 *__restrict__ spos, uint32_t N,
                                                  const KnnGrid *__restrict__ grid,
                                                  const uint32_t *__restrict__ start,
                                                  const uint32_t *__restrict__ cnt, float extent,
                                                  float *__restrict__ P, size_t cap, float *__restrict__ density) {
  const uint32_t s = blockIdx.x * KT + threadIdx.x;
  if (s >= N) return;
  const KnnGrid G = *grid;
  const float4 me = spos[s];
  const uint32_t self = __float_as_uint(me.w);
  const float3 p = make_float3(me.x, me.y, me.z);
  const float pv[3] = {p.x, p.y, p.z};
  int c[3];
  (void)knn_cell_id(G, p, c);
  const int gx = (int)G.g[0], gy = (int)G.g[1], gz = (int)G.g[2];
  float b0 = INFINITY, b1 = INFINITY, b2 = INFINITY;
  for (int r = 0;; ++r) {
    for (int z = c[2] - r; z <= c[2] + r; ++z) {
      if (z < 0 || z >= gz) continue;
      const bool zf = (z == c[2] - r) || (z == c[2] + r);
      for (int y = c[1] - r; y <= c[1] + r; ++y) {
        if (y < 0 || y >= gy) continue;
        const bool face = zf || (y == c[1] - r) || (y == c[1] + r);
        const int step = face ? 1 : 2 * r;  // inside the shell only the two x faces
        for (int x = c[0] - r; x <= c[0] + r; x += step) {
          if (x < 0 || x >= gx) continue;
          const uint32_t id = ((uint32_t)z * G.g[1] + (uint32_t)y) * G.g[0] + (uint32_t)x;
          const uint32_t j0 = start[id], j1 = j0 + cnt[id];
          for (uint32_t j = j0; j < j1; ++j) {
            const float4 o = spos[j];
            if (__float_as_uint(o.w) == self) continue;
            const float dx = p.x - o.x, dy = p.y - o.y, dz = p.z - o.z;
            knn_insert((dx * dx + dy * dy) + dz * dz, b0, b1, b2);
          }
        }
      }
    }
    bool covered = true;
    float lb = INFINITY;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const int ga = (int)G.g[a];
      if (c[a] - r > 0) {
        covered = false;
        lb = fminf(lb, pv[a] - (G.lo[a] + (float)(c[a] - r) * G.h));
      }
      if (c[a] + r < ga - 1) {
        covered = false;
        lb = fminf(lb, (G.lo[a] + (float)(c[a] + r + 1) * G.h) - pv[a]);
      }
    }
    if (covered) break;
    lb = lb - 1e-3f * G.h;
    if (b2 < INFINITY && lb > 0.0f && b2 <= (lb * lb) * 0.99999f) break;
  }
  // knn-init.wgsl:64-89
  const float n0 = fminf(sqrtf(b0), 1e10f), n1 = fminf(sqrtf(b1), 1e10f), n2 = fminf(sqrtf(b2), 1e10f);
  const float mean_dist = (n0 + n1 + n2) / 3.0f;
  density[self] = mean_dist;
  const float base_sigma = mean_dist / 3.0f;
  const float min_sigma = extent * 0.0007f, max_sigma = extent * 0.007f, bias_sigma = extent * 0.0015f;
  float sigma = fmaxf(base_sigma, bias_sigma);
  sigma = fminf(fmaxf(sigma, min_sigma), max_sigma);
  sigma = fminf(sigma, 0.1f);
  const float ls = f16_round((float)log((double)fmaxf(sigma, 1e-6f)));
  P[(size_t)(P_SCALE + 0) * cap + self] = ls;
  P[(size_t)(P_SCALE + 1) * cap + self] = ls;
  P[(size_t)(P_SCALE + 2) * cap + self] = ls;
}

}  // namespace

uint32_t knn_grid_dim(uint32_t N) {
  // about two points per cell for a volume-filling cloud; 160^3 cells at most
  const double g = std::cbrt((double)N / 2.0);
  return (uint32_t)std::min(160.0, std::max(1.0, std::floor(g + 0.5)));
}

size_t knn_grid_cells(uint32_t N) {
  const uint32_t g = knn_grid_dim(N);
  return (size_t)g * g * g;
}

hipError_t launch_knn_init(hipStream_t st, float *P, size_t cap, uint32_t N, float extent, const KnnScratch &w) {
  if (N == 0) return hipSuccess;
  const uint32_t gmax = knn_grid_dim(N);
  const size_t cells = (size_t)gmax * gmax * gmax;
  KnnGrid *grid = reinterpret_cast<KnnGrid *>(w.grid);
  hipLaunchKernelGGL(k_knn_bounds, dim3(1), dim3(1024), 0, st, P, cap, N, gmax, grid);
  launch_fill_u32(st, w.cnt, 0u, cells);
  const uint32_t nb = (N + KT - 1) / KT;
  hipLaunchKernelGGL(k_knn_count, dim3(nb), dim3(KT), 0, st, P, cap, N, grid, w.cnt, w.cell, w.rank);
  launch_scan(st, w.cnt, w.start, cells, w.scan_tmp, w.total);
  hipLaunchKernelGGL(k_knn_scatter, dim3(nb), dim3(KT), 0, st, P, cap, N, w.start, w.cell, w.rank, w.spos);
  hipLaunchKernelGGL(k_knn_query, dim3(nb), dim3(KT), 0, st, w.spos, N, grid, w.start, w.cnt, extent, P, cap,
                     w.density);
  return hipGetLastError();
}

}  // namespace sv
