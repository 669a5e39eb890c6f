// sv_scene.cpp -- seeded synthetic scenes and cameras (host only).
//
// Draw j of stream `seed` is SplitMix64 evaluated at counter j (a
// counter-based generator, so any subset can be drawn in parallel and in
// any order). Normals use Box-Muller on two consecutive uniforms.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/sv_api.h"

namespace {

uint64_t mix64(uint64_t seed, uint64_t ctr) {
  uint64_t z = seed + (ctr + 1ull) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Round an f32 to the nearest f16 (ties to even) and back.
float f16_round(float x) {
  uint32_t u;
  std::memcpy(&u, &x, 4);
  const uint32_t sign = u & 0x80000000u, ax = u & 0x7fffffffu;
  uint32_t h;
  if (ax >= 0x7f800000u) {
    h = ax > 0x7f800000u ? 0x7e00u : 0x7c00u;
  } else if (ax >= 0x477ff000u) {
    h = 0x7c00u;
  } else if (ax < 0x38800000u) {
    if (ax < 0x33000000u) {
      h = 0;
    } else {
      const uint32_t e = ax >> 23, m = (ax & 0x7fffffu) | 0x800000u, sh = 126u - e;
      uint32_t q = m >> sh;
      const uint32_t rem = m & ((1u << sh) - 1u), half = 1u << (sh - 1u);
      if (rem > half || (rem == half && (q & 1u))) ++q;
      h = q;
    }
  } else {
    uint32_t q = (((ax >> 23) - 112u) << 10) | ((ax & 0x7fffffu) >> 13);
    const uint32_t rem = ax & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (q & 1u))) ++q;
    h = q;
  }
  const uint32_t he = (h >> 10) & 0x1fu, hm = h & 0x3ffu;
  uint32_t out;
  if (he == 0) {
    const float f = (float)hm * 5.9604644775390625e-8f;
    std::memcpy(&out, &f, 4);
  } else if (he == 31) {
    out = 0x7f800000u | (hm << 13);
  } else {
    out = ((he + 112u) << 23) | (hm << 13);
  }
  out |= sign;
  float r;
  std::memcpy(&r, &out, 4);
  return r;
}

void invert4(const double m[16], double out[16]) {
  // Gauss-Jordan on a column-major 4x4 (row r, col c at c*4+r).
  double a[4][8];
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) {
      a[r][c] = m[c * 4 + r];
      a[r][c + 4] = (r == c) ? 1.0 : 0.0;
    }
  for (int c = 0; c < 4; ++c) {
    int piv = c;
    for (int r = c + 1; r < 4; ++r)
      if (std::fabs(a[r][c]) > std::fabs(a[piv][c])) piv = r;
    if (piv != c)
      for (int k = 0; k < 8; ++k) std::swap(a[c][k], a[piv][k]);
    const double d = a[c][c];
    if (d == 0.0) {
      for (int k = 0; k < 16; ++k) out[k] = 0.0;
      return;
    }
    for (int k = 0; k < 8; ++k) a[c][k] /= d;
    for (int r = 0; r < 4; ++r) {
      if (r == c) continue;
      const double f = a[r][c];
      for (int k = 0; k < 8; ++k) a[r][k] -= f * a[c][k];
    }
  }
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) out[c * 4 + r] = a[r][c + 4];
}

// Run fn(begin, end) over [0, n) on up to 16 threads.
template <typename F>
void parallel_for(size_t n, F fn) {
  const size_t nt = std::max<size_t>(1, std::min<size_t>(16, std::thread::hardware_concurrency()));
  if (n < 4096 || nt == 1) {
    fn(0, n);
    return;
  }
  std::vector<std::thread> th;
  const size_t chunk = (n + nt - 1) / nt;
  for (size_t t = 0; t < nt; ++t) {
    const size_t b = t * chunk, e = std::min(n, b + chunk);
    if (b < e) th.emplace_back(fn, b, e);
  }
  for (auto &x : th) x.join();
}

}  // namespace

extern "C" {

double sv_scene_uniform(uint64_t seed, uint64_t counter) {
  return (double)(mix64(seed, counter) >> 11) * (1.0 / 9007199254740992.0);
}

double sv_scene_normal(uint64_t seed, uint64_t counter) {
  const double u1 = sv_scene_uniform(seed, 2 * counter);
  const double u2 = sv_scene_uniform(seed, 2 * counter + 1);
  return std::sqrt(-2.0 * std::log(1.0 - u1)) * std::cos(2.0 * M_PI * u2);
}

sv_status sv_scene_gaussians(uint64_t seed, size_t n, float half_extent, float *pos, float *logscale,
                             float *quat, float *logit, float *sh) {
  if (!pos || !logscale || !quat || !logit || !sh) return SV_E_ARG;
  parallel_for(n, [=](size_t i0, size_t i1) {
  for (size_t i = i0; i < i1; ++i) {
    const uint64_t b = (uint64_t)i * 64ull;
    for (int k = 0; k < 3; ++k)
      pos[3 * i + k] = f16_round((float)((2.0 * sv_scene_uniform(seed, b + k) - 1.0) * half_extent));
    for (int k = 0; k < 3; ++k)
      logscale[3 * i + k] = f16_round((float)(-4.5 + 0.3 * sv_scene_normal(seed, b + 3 + k)));
    double q[4], qn = 0.0;
    for (int k = 0; k < 4; ++k) {
      q[k] = sv_scene_normal(seed, b + 6 + k);
      qn += q[k] * q[k];
    }
    qn = std::sqrt(qn);
    if (qn == 0.0) {
      q[0] = q[1] = q[2] = 0.0;
      q[3] = qn = 1.0;
    }
    for (int k = 0; k < 4; ++k) quat[4 * i + k] = f16_round((float)(q[k] / qn));
    logit[i] = f16_round((float)sv_scene_normal(seed, b + 10));
    for (int j = 0; j < 48; ++j) {
      const double z = sv_scene_normal(seed, b + 11 + j);
      sh[48 * i + j] = f16_round((float)(j < 3 ? 0.5 + 0.2 * z : 0.05 * z));
    }
  }
  });
  return SV_OK;
}

sv_status sv_scene_jitter(uint64_t seed, size_t n, float sigma, float *pos) {
  if (!pos) return SV_E_ARG;
  parallel_for(n, [=](size_t i0, size_t i1) {
    for (size_t i = i0; i < i1; ++i)
      for (int k = 0; k < 3; ++k)
        pos[3 * i + k] =
            f16_round((float)((double)pos[3 * i + k] + sigma * sv_scene_normal(seed, 3ull * i + k)));
  });
  return SV_OK;
}

sv_status sv_scene_camera(const float eye[3], const float target[3], uint32_t w, uint32_t h,
                          float fovy_deg, sv_camera *out) {
  if (!eye || !target || !out || w == 0 || h == 0 || !(fovy_deg > 0.0f)) return SV_E_ARG;
  double f[3] = {(double)target[0] - eye[0], (double)target[1] - eye[1], (double)target[2] - eye[2]};
  double fl = std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
  if (fl == 0.0) return SV_E_ARG;
  for (double &v : f) v /= fl;
  double up[3] = {0.0, 1.0, 0.0};
  if (std::fabs(f[1]) > 0.999) {
    up[0] = 1.0;
    up[1] = 0.0;
  }
  double r[3] = {f[1] * up[2] - f[2] * up[1], f[2] * up[0] - f[0] * up[2], f[0] * up[1] - f[1] * up[0]};
  const double rl = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
  for (double &v : r) v /= rl;
  const double u[3] = {r[1] * f[2] - r[2] * f[1], r[2] * f[0] - r[0] * f[2], r[0] * f[1] - r[1] * f[0]};
  const double e[3] = {eye[0], eye[1], eye[2]};
  double V[16] = {0};
  const double *rows[3] = {r, u, f};
  for (int row = 0; row < 3; ++row) {
    for (int c = 0; c < 3; ++c) V[c * 4 + row] = rows[row][c];
    V[12 + row] = -(rows[row][0] * e[0] + rows[row][1] * e[1] + rows[row][2] * e[2]);
  }
  V[15] = 1.0;
  const double fy = (0.5 * h) / std::tan(0.5 * fovy_deg * M_PI / 180.0);
  const double fx = fy;
  const double cx = 0.5 * w, cy = 0.5 * h;
  // get_pinhole_projection_matrix (training-camera.ts:29-52) after its
  // transpose: rows (sx,0,ox,0), (0,sy,oy,0), (0,0,1,0), (0,0,1,0).
  const double sx = 2.0 * fx / w, sy = -2.0 * fy / h, ox = 2.0 * cx / w - 1.0, oy = 1.0 - 2.0 * cy / h;
  double Pm[16] = {0};
  Pm[0] = sx;
  Pm[5] = sy;
  Pm[8] = ox;
  Pm[9] = oy;
  Pm[10] = 1.0;
  Pm[11] = 1.0;
  double Vi[16], Pi[16];
  invert4(V, Vi);
  invert4(Pm, Pi);  // singular (two equal rows): zeros, as wgpu-matrix's inverse degrades
  for (int k = 0; k < 16; ++k) {
    out->view[k] = (float)V[k];
    out->view_inv[k] = (float)Vi[k];
    out->proj[k] = (float)Pm[k];
    out->proj_inv[k] = (float)Pi[k];
  }
  out->viewport[0] = (float)w;
  out->viewport[1] = (float)h;
  out->focal[0] = (float)fx;
  out->focal[1] = (float)fy;
  return SV_OK;
}

}  // extern "C"
