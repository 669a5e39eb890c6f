/*
 * sv_oracle.c -- CPU restatement of Skyfall-Viewer's splat training step.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product). Built by
 * oracle/Makefile into oracle/liboracle.so with -ffp-contract=off.
 *
 * Matrices follow WGSL: float m[col][row], and a WGSL product A * B is
 * restated as (A*B)[i][j] = (A[0][j]*B[i][0] + A[1][j]*B[i][1]) + A[2][j]*B[i][2].
 * Left-associative WGSL expressions keep their parse order.
 *
 * How each function is pinned (tests/test_oracle_*.py):
 *   orc_expf            libm expf over 2^20 samples, <= 2 ulp
 *   orc_exp_alpha       float64 exp, within 3 + 2|x| ulp (WGSL's exp accuracy)
 *                       over [-80, 1/2]
 *   orc_f16_round       numpy float16 round-to-nearest-even
 *   orc_project         closed form for an isotropic on-axis splat (conic,
 *                       radius); pixel-mean closed form (fx x/z + cx,
 *                       H - cy - fy y/z); culling invariants; a float64
 *                       numpy restatement to 1e-5 rel
 *   orc_sh_color        real spherical harmonics from scipy.special
 *                       (Condon-Shortley convention) at random directions
 *   orc_emit/_encode    integer brute force in Python
 *   orc_sort_pairs      numpy stable argsort
 *   orc_sort_pairs_radix  = orc_sort_pairs and numpy stable argsort (ties,
 *                       unordered values, every digit width)
 *   orc_tile_offsets    numpy searchsorted (lower_bound)
 *   orc_forward         pure-Python compositing loop on a tiny scene;
 *                       empty tile -> background; single opaque splat
 *   orc_loss            numpy closed form
 *   orc_backward_legacy torch float64 autograd of the compositing model
 *                       (conic, mean2d, logit, colour), and central finite
 *                       differences of orc_forward_loss_f64 for the logit
 *                       and every SH coefficient of 256 Gaussians
 *   orc_forward_loss_f64_only  its central differences for an SH change
 *                       equal orc_forward_loss_f64's
 *   orc_backward_wavefront  equals orc_backward_legacy (to fixed-point
 *                       truncation) where every pixel list fits one batch
 *                       and alpha <= 0.98; integer sums brute force
 *   orc_backward_reduce integer brute force
 *   orc_backward_geom   torch float64 autograd of the covariance model the
 *                       shader differentiates (its conic-b factor kept)
 *   orc_adam            textbook Adam; first step from zero moments gives
 *                       -lr*g/(|g|+eps)
 *   orc_accumulate_*    closed forms
 *   orc_compact_rows    numpy boolean indexing
 *   orc_metric_view     numpy restatement of the per-pixel metric and
 *                       brute-force splat walks on small scenes
 *   orc_knn_init        scipy cKDTree 3-NN distances in float64 and the
 *                       sigma rule in numpy; special cases n <= 3,
 *                       duplicate points
 *   orc_prune_keep, orc_densify, orc_tile_l1, orc_densify_decide
 *                       parity unpinned against the reference (its rule is
 *                       weighted random, its draws Math.random); pinned by
 *                       numpy restatements and invariants (counts, order,
 *                       zero child moments, split offsets ~ N(0, sigma))
 */
#include "sv_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ */
/* Constants: src/shaders/common.wgsl:1-18                              */
/* ------------------------------------------------------------------ */
static const float SH_C0 = 0.28209479177387814f;
static const float SH_C1 = 0.4886025119029199f;
static const float SH_C2[5] = {1.0925484305920792f, -1.0925484305920792f,
                               0.31539156525252005f, -1.0925484305920792f,
                               0.5462742152960396f};
static const float SH_C3[7] = {-0.5900435899266435f, 2.890611442640554f,
                               -0.4570457994644658f, 0.3731763325901154f,
                               -0.4570457994644658f, 1.445305721320277f,
                               -0.5900435899266435f};

/* training-forward-tiles.wgsl:38-41 */
#define MIN_ALPHA ((float)(1.0 / 255.0))
#define MAX_FRAGMENT_ALPHA 0.999f
#define TRANSMITTANCE_THRESHOLD 1e-4f
/* backward-tiles-local.wgsl:49-52 */
#define FIXED_SCALE 65536.0f
#define WF_BATCH 64u

/* ------------------------------------------------------------------ */
/* Deterministic exp: Cody-Waite reduction + Cephes expf polynomial.    */
/* ------------------------------------------------------------------ */
static inline float bits_to_f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static inline uint32_t f_to_bits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

/* exp of the per-pixel splat weight e^-power (training-forward-tiles.wgsl:
 * 190 and the backward passes), for x in [-80, 1/2]: 2^(x log2 e) with the
 * product rounded, k = rint(y), f = y - k, 2^f by a degree-5 polynomial with
 * constant term 1 fitted for relative error on [-1/2, 1/2], times 2^k built
 * in the exponent bits. Within 3 + 2|x| ulp of e^x (WGSL's accuracy
 * requirement for exp). Callers clamp the power to 80 first, as the device
 * does (a power above 80 gives an alpha far below 1/255 either way). */
float orc_exp_alpha(float x) {
  const float y = x * 0x1.715476p+0f;
  const float k = rintf(y);
  const float f = y - k;
  float p = 0x1.5bba14p-10f;
  p = fmaf(p, f, 0x1.3cea88p-7f);
  p = fmaf(p, f, 0x1.c6b752p-5f);
  p = fmaf(p, f, 0x1.ebf9bcp-3f);
  p = fmaf(p, f, 0x1.62e42ap-1f);
  p = fmaf(p, f, 1.0f);
  return p * bits_to_f((uint32_t)((int)k + 127) << 23);
}

void orc_exp_alpha_v(size_t n, const float *x, float *y) {
  for (size_t i = 0; i < n; ++i) y[i] = orc_exp_alpha(x[i]);
}

float orc_expf(float x) {
  if (x != x) return x;
  if (x > 88.72283935546875f) return INFINITY;
  if (x < -103.972084f) return 0.0f;
  const float n = rintf(x * 1.44269504088896341f);
  float r = fmaf(-n, 0.693359375f, x);
  r = fmaf(-n, -2.12194440e-4f, r);
  float p = 1.9875691500e-4f;
  p = fmaf(p, r, 1.3981999507e-3f);
  p = fmaf(p, r, 8.3334519073e-3f);
  p = fmaf(p, r, 4.1665795894e-2f);
  p = fmaf(p, r, 1.6666665459e-1f);
  p = fmaf(p, r, 5.0000001201e-1f);
  const float r2 = r * r;
  p = fmaf(p, r2, r);
  p = p + 1.0f;
  int ni = (int)n;
  if (ni < -126) {
    /* two exact-power scalings; the final multiply rounds once */
    p = p * bits_to_f((uint32_t)(ni + 64 + 127) << 23);
    return p * bits_to_f((uint32_t)(127 - 64) << 23);
  }
  if (ni > 127) {
    p = p * 2.0f;
    ni -= 1;
  }
  return p * bits_to_f((uint32_t)(ni + 127) << 23);
}

/* f32 -> f16 -> f32 with round-to-nearest-even (pack2x16float,
 * training-params.wgsl:81-128; @petamoriken/float16 on the host). */
float orc_f16_round(float x) {
  uint32_t u = f_to_bits(x);
  uint32_t sign = u & 0x80000000u;
  uint32_t ax = u & 0x7fffffffu;
  uint16_t h;
  if (ax >= 0x7f800000u) {
    h = (uint16_t)((ax > 0x7f800000u) ? 0x7e00u : 0x7c00u);
  } else if (ax >= 0x477ff000u) {
    h = 0x7c00u; /* rounds to inf */
  } else if (ax < 0x38800000u) {
    /* subnormal half */
    if (ax < 0x33000000u) {
      h = 0;
    } else {
      uint32_t e = ax >> 23;
      uint32_t m = (ax & 0x7fffffu) | 0x800000u;
      uint32_t shift = 126u - e; /* 14 + (113 - e) - 1 + 1 */
      uint32_t q = m >> (shift);
      uint32_t rem = m & ((1u << shift) - 1u);
      uint32_t half = 1u << (shift - 1u);
      if (rem > half || (rem == half && (q & 1u))) q++;
      h = (uint16_t)q;
    }
  } else {
    uint32_t e = (ax >> 23) - 112u;
    uint32_t m = ax & 0x7fffffu;
    uint32_t q = (e << 10) | (m >> 13);
    uint32_t rem = m & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (q & 1u))) q++;
    h = (uint16_t)q;
  }
  /* back to f32 */
  uint32_t he = (h >> 10) & 0x1fu, hm = h & 0x3ffu, out;
  if (he == 0) {
    if (hm == 0) {
      out = 0;
    } else {
      float f = (float)hm * 5.9604644775390625e-8f; /* 2^-24 exact */
      out = f_to_bits(f);
    }
  } else if (he == 31) {
    out = 0x7f800000u | (hm << 13);
  } else {
    out = ((he + 112u) << 23) | (hm << 13);
  }
  return bits_to_f(out | sign);
}

/* ------------------------------------------------------------------ */
/* WGSL matrix helpers (m[col][row])                                    */
/* ------------------------------------------------------------------ */
typedef float m3[3][3];

static void m3_mul(const m3 A, const m3 B, m3 C) {
  m3 t;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      t[i][j] = (A[0][j] * B[i][0] + A[1][j] * B[i][1]) + A[2][j] * B[i][2];
  memcpy(C, t, sizeof(m3));
}
static void m3_transpose(const m3 A, m3 T) {
  m3 t;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) t[i][j] = A[j][i];
  memcpy(T, t, sizeof(m3));
}
static void m3_mul_v(const m3 A, const float v[3], float r[3]) {
  float t[3];
  for (int j = 0; j < 3; ++j)
    t[j] = (A[0][j] * v[0] + A[1][j] * v[1]) + A[2][j] * v[2];
  r[0] = t[0]; r[1] = t[1]; r[2] = t[2];
}
/* mat4x4 (column-major float[16]) * vec4 */
static void m4_mul_v(const float *M, const float v[4], float r[4]) {
  float t[4];
  for (int j = 0; j < 4; ++j)
    t[j] = ((M[0 * 4 + j] * v[0] + M[1 * 4 + j] * v[1]) + M[2 * 4 + j] * v[2]) +
           M[3 * 4 + j] * v[3];
  r[0] = t[0]; r[1] = t[1]; r[2] = t[2]; r[3] = t[3];
}
static inline float dot3(const float a[3], const float b[3]) {
  return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}
static inline void normalize3(const float v[3], float o[3]) {
  float len = sqrtf(dot3(v, v));
  o[0] = v[0] / len; o[1] = v[1] / len; o[2] = v[2] / len;
}
static inline float clampf(float x, float lo, float hi) {
  return fminf(fmaxf(x, lo), hi);
}
/* WGSL float -> u32 conversion: truncation, saturating, NaN -> 0. */
static inline uint32_t f2u_sat(float x) {
  if (!(x > 0.0f)) return 0u;
  if (x >= 4294967296.0f) return 0xffffffffu;
  return (uint32_t)x;
}
/* WGSL float -> i32 conversion: truncation, saturating, NaN -> 0. */
static inline int32_t f2i_sat(float x) {
  if (x != x) return 0;
  if (x >= 2147483648.0f) return INT32_MAX;
  if (x <= -2147483648.0f) return INT32_MIN;
  return (int32_t)x;
}
static inline float sigmoid_ref(float logit) {
  /* training-forward-tiles.wgsl:188 */
  return 1.0f / (1.0f + orc_expf(-logit));
}

/* quat_to_mat3_tiling, training-tiles.wgsl:30-37 (q = (x,y,z,w)). The
 * geometry backward's quat_to_mat3 (training-backward-geom.wgsl:21-31)
 * builds the same matrix. */
static void quat_to_mat3(const float q[4], m3 R) {
  const float rx = q[0], ry = q[1], rz = q[2], rw = q[3];
  R[0][0] = 1.0f - 2.0f * (ry * ry + rz * rz);
  R[0][1] = 2.0f * (rx * ry - rw * rz);
  R[0][2] = 2.0f * (rx * rz + rw * ry);
  R[1][0] = 2.0f * (rx * ry + rw * rz);
  R[1][1] = 1.0f - 2.0f * (rx * rx + rz * rz);
  R[1][2] = 2.0f * (ry * rz - rw * rx);
  R[2][0] = 2.0f * (rx * rz - rw * ry);
  R[2][1] = 2.0f * (ry * rz + rw * rx);
  R[2][2] = 1.0f - 2.0f * (rx * rx + ry * ry);
}

static void normalize4(const float *q, float o[4]) {
  float d = ((q[0] * q[0] + q[1] * q[1]) + q[2] * q[2]) + q[3] * q[3];
  float len = sqrtf(d);
  o[0] = q[0] / len; o[1] = q[1] / len; o[2] = q[2] / len; o[3] = q[3] / len;
}

/* ------------------------------------------------------------------ */
/* a4: project_gaussians_for_tiling, training-tiles.wgsl:39-180         */
/* ------------------------------------------------------------------ */
static void tile_rect(const orc_view *v, float mx, float my, float radius,
                      uint32_t r[4]) {
  /* training-tiles.wgsl:161-171 (and :225-234 in emit) */
  const float ts = (float)v->tile_size;
  const float tile_x = mx / ts, tile_y = my / ts;
  const float trx = radius / ts, try_ = radius / ts;
  uint32_t x0 = f2u_sat(floorf(tile_x - trx));
  uint32_t y0 = f2u_sat(floorf(tile_y - try_));
  uint32_t x1 = f2u_sat(ceilf(tile_x + trx));
  uint32_t y1 = f2u_sat(ceilf(tile_y + try_));
  r[0] = x0 < v->tiles_x ? x0 : v->tiles_x;
  r[1] = y0 < v->tiles_y ? y0 : v->tiles_y;
  r[2] = x1 < v->tiles_x ? x1 : v->tiles_x;
  r[3] = y1 < v->tiles_y ? y1 : v->tiles_y;
}

void orc_project(size_t n, const orc_params *p, const orc_camera *cam,
                 const orc_view *v, orc_proj *out) {
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < n; ++i) {
    orc_proj *o = &out[i];
    /* The reference writes only tiles_count/radius/depth on a cull and
     * leaves the rest stale; the restatement zeroes the record. */
    memset(o, 0, sizeof(*o));
    const float pw[4] = {p->pos[3 * i], p->pos[3 * i + 1], p->pos[3 * i + 2], 1.0f};
    float pc[4], clip[4];
    m4_mul_v(cam->view, pw, pc);
    m4_mul_v(cam->proj, pc, clip);
    const float w = clip[3];
    clip[0] = clip[0] / w; clip[1] = clip[1] / w;
    clip[2] = clip[2] / w; clip[3] = clip[3] / w;
    if (fabsf(clip[0]) > 1.2f || fabsf(clip[1]) > 1.2f || pc[2] <= 0.0f) continue;
    const float depth = pc[2];
    if (depth < v->near_plane || depth > v->far_plane) continue;

    float sigma[3];
    for (int k = 0; k < 3; ++k)
      sigma[k] = orc_expf(clampf(p->logscale[3 * i + k], -10.0f, 10.0f));
    m3 S = {{sigma[0], 0, 0}, {0, sigma[1], 0}, {0, 0, sigma[2]}};
    float q[4];
    normalize4(&p->quat[4 * i], q);
    m3 R;
    quat_to_mat3(q, R);
    /* C3D = transpose(S * R) * S * R  (:101) */
    m3 SR, SRt, C3D;
    m3_mul(S, R, SR);
    m3_transpose(SR, SRt);
    m3_mul(SRt, S, C3D);
    m3_mul(C3D, R, C3D);
    const float tx = pc[0], ty = pc[1], tz = pc[2];
    const float fx = cam->focal[0], fy = cam->focal[1];
    m3 J = {{fx / tz, 0.0f, -(fx * tx) / (tz * tz)},
            {0.0f, fy / tz, -(fy * ty) / (tz * tz)},
            {0.0f, 0.0f, 0.0f}};
    m3 Rwc;
    for (int c = 0; c < 3; ++c)
      for (int r = 0; r < 3; ++r) Rwc[c][r] = cam->view[c * 4 + r];
    m3 T, TC, Tt, C2D;
    m3_mul(J, Rwc, T);
    m3_mul(T, C3D, TC);
    m3_transpose(T, Tt);
    m3_mul(TC, Tt, C2D);
    C2D[0][0] = C2D[0][0] + v->eps2d;
    C2D[1][1] = C2D[1][1] + v->eps2d;
    const float det = C2D[0][0] * C2D[1][1] - C2D[0][1] * C2D[0][1];
    if (det <= 0.0f) continue;
    const float mid = 0.5f * (C2D[0][0] + C2D[1][1]);
    const float disc = sqrtf(fmaxf(mid * mid - det, 0.0f));
    const float l1 = mid + disc, l2 = mid - disc;
    const float radius = ceilf(3.0f * sqrtf(fmaxf(l1, l2)));
    if (radius > v->radius_clip || radius < 1.0f) continue;
    const float det_inv = 1.0f / det;
    const float aa = C2D[1][1] * det_inv;
    const float bb = -C2D[0][1] * det_inv;
    const float cc = C2D[0][0] * det_inv;
    const float mx = (0.5f * (clip[0] + 1.0f)) * (float)v->width;
    const float my = (0.5f * (clip[1] + 1.0f)) * (float)v->height;
    o->mean2d[0] = mx; o->mean2d[1] = my;
    o->radius = radius;
    o->depth = depth;
    o->conic[0] = aa; o->conic[1] = bb; o->conic[2] = cc;
    o->position_world[0] = pw[0]; o->position_world[1] = pw[1]; o->position_world[2] = pw[2];
    o->opacity_logit = p->logit[i];
    uint32_t rect[4];
    tile_rect(v, mx, my, radius, rect);
    const int32_t nx = (int32_t)rect[2] - (int32_t)rect[0];
    const int32_t ny = (int32_t)rect[3] - (int32_t)rect[1];
    o->tiles_count = (nx <= 0 || ny <= 0) ? 0u : (uint32_t)(nx * ny);
  }
}

/* ------------------------------------------------------------------ */
/* a5: computeColorFromSH_tf, training-forward-tiles.wgsl:65-105        */
/* ------------------------------------------------------------------ */
void orc_sh_color(const float dir[3], const float *sh, uint32_t deg, float rgb[3]) {
  float res[3];
  for (int c = 0; c < 3; ++c) res[c] = SH_C0 * sh[c];
  if (deg > 0u) {
    const float x = dir[0], y = dir[1], z = dir[2];
    const float k1 = (-SH_C1) * y, k2 = SH_C1 * z, k3 = (-SH_C1) * x;
    for (int c = 0; c < 3; ++c)
      res[c] = res[c] + ((k1 * sh[3 + c] + k2 * sh[6 + c]) - (SH_C1 * x) * sh[9 + c]);
    (void)k3;
    if (deg > 1u) {
      const float xx = x * x, yy = y * y, zz = z * z;
      const float xy = x * y, yz = y * z, xz = x * z;
      const float b4 = SH_C2[0] * xy, b5 = SH_C2[1] * yz;
      const float b6 = SH_C2[2] * ((2.0f * zz - xx) - yy);
      const float b7 = SH_C2[3] * xz, b8 = SH_C2[4] * (xx - yy);
      for (int c = 0; c < 3; ++c)
        res[c] = res[c] + ((((b4 * sh[12 + c] + b5 * sh[15 + c]) + b6 * sh[18 + c]) +
                            b7 * sh[21 + c]) + b8 * sh[24 + c]);
      if (deg > 2u) {
        const float b9 = (SH_C3[0] * y) * (3.0f * xx - yy);
        const float b10 = (SH_C3[1] * xy) * z;
        const float b11 = (SH_C3[2] * y) * ((4.0f * zz - xx) - yy);
        const float b12 = (SH_C3[3] * z) * ((2.0f * zz - 3.0f * xx) - 3.0f * yy);
        const float b13 = (SH_C3[4] * x) * ((4.0f * zz - xx) - yy);
        const float b14 = (SH_C3[5] * z) * (xx - yy);
        const float b15 = (SH_C3[6] * x) * (xx - 3.0f * yy);
        for (int c = 0; c < 3; ++c)
          res[c] = res[c] + ((((((b9 * sh[27 + c] + b10 * sh[30 + c]) + b11 * sh[33 + c]) +
                                 b12 * sh[36 + c]) + b13 * sh[39 + c]) + b14 * sh[42 + c]) +
                             b15 * sh[45 + c]);
      }
    }
  }
  for (int c = 0; c < 3; ++c) {
    const float t = res[c] + 0.5f;
    rgb[c] = fmaxf(0.0f, t);
  }
}

/* Real SH basis values in the reference's evaluation order, used by the
 * backward's per-coefficient gradients (backward-tiles-local.wgsl:104-183).
 * basis[k] multiplies res_scaled for coefficient k. */
static void sh_basis(const float dir[3], uint32_t deg, float basis[16]) {
  const float x = dir[0], y = dir[1], z = dir[2];
  for (int k = 0; k < 16; ++k) basis[k] = 0.0f;
  basis[0] = SH_C0;
  if (deg > 0u) {
    basis[1] = (-SH_C1) * y;
    basis[2] = SH_C1 * z;
    basis[3] = (-SH_C1) * x;
    if (deg > 1u) {
      const float xx = x * x, yy = y * y, zz = z * z;
      const float xy = x * y, yz = y * z, xz = x * z;
      basis[4] = SH_C2[0] * xy;
      basis[5] = SH_C2[1] * yz;
      basis[6] = SH_C2[2] * ((2.0f * zz - xx) - yy);
      basis[7] = SH_C2[3] * xz;
      basis[8] = SH_C2[4] * (xx - yy);
      if (deg > 2u) {
        basis[9] = (SH_C3[0] * y) * (3.0f * xx - yy);
        basis[10] = (SH_C3[1] * xy) * z;
        basis[11] = (SH_C3[2] * y) * ((4.0f * zz - xx) - yy);
        basis[12] = (SH_C3[3] * z) * ((2.0f * zz - 3.0f * xx) - 3.0f * yy);
        basis[13] = (SH_C3[4] * x) * ((4.0f * zz - xx) - yy);
        basis[14] = (SH_C3[5] * z) * (xx - yy);
        basis[15] = (SH_C3[6] * x) * (xx - 3.0f * yy);
      }
    }
  }
}

/* viewDir = normalize(-(camera.view * position_world).xyz),
 * training-forward-tiles.wgsl:184-185 */
void orc_view_dir(const orc_camera *cam, const float pos[3], float dir[3]) {
  const float pw[4] = {pos[0], pos[1], pos[2], 1.0f};
  float pc[4];
  m4_mul_v(cam->view, pw, pc);
  const float neg[3] = {-pc[0], -pc[1], -pc[2]};
  normalize3(neg, dir);
}

/* ------------------------------------------------------------------ */
/* a6/a7: host cumsum (gaussian-renderer.ts:1814-1828) + emit           */
/* ------------------------------------------------------------------ */
uint32_t orc_encode_key(uint32_t cam, uint32_t tile, float depth) {
  /* training-tiles.wgsl:182-199 */
  const float dc = clampf(depth, 0.0f, 65535.0f);
  const uint32_t dq = f2u_sat(dc) & 0xffffu;
  return ((cam & 0xfu) << 28) | ((tile & 0xfffu) << 16) | dq;
}

uint64_t orc_count_isects(size_t n, const orc_proj *proj) {
  uint64_t s = 0;
  for (size_t i = 0; i < n; ++i) s += proj[i].tiles_count;
  return s;
}

void orc_emit(size_t n, const orc_proj *proj, const orc_view *v, uint32_t *keys,
              uint32_t *vals) {
  /* training-tiles.wgsl:201-247; start = inclusive cumsum of the previous
   * Gaussian (the host prefix sum, gaussian-renderer.ts:1814-1828). */
  uint64_t *start = (uint64_t *)malloc((n + 1) * sizeof(uint64_t));
  start[0] = 0;
  for (size_t i = 0; i < n; ++i) start[i + 1] = start[i] + proj[i].tiles_count;
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < n; ++i) {
    const orc_proj *gp = &proj[i];
    if (gp->tiles_count == 0u) continue;
    uint32_t r[4];
    tile_rect(v, gp->mean2d[0], gp->mean2d[1], gp->radius, r);
    uint64_t cur = start[i];
    for (uint32_t ty = r[1]; ty < r[3]; ++ty)
      for (uint32_t tx = r[0]; tx < r[2]; ++tx) {
        keys[cur] = orc_encode_key(0u, ty * v->tiles_x + tx, gp->depth);
        vals[cur] = (uint32_t)i;
        ++cur;
      }
  }
  free(start);
}

/* ------------------------------------------------------------------ */
/* a8: stable LSD radix sort (sort.ts:260-370, radix_sort.wgsl) ==      */
/* stable sort by key. Values are emitted in increasing order per key   */
/* (each Gaussian emits a tile at most once), so sorting (key, value)   */
/* pairs lexicographically is the same permutation.                     */
/* ------------------------------------------------------------------ */
static int cmp_u64(const void *a, const void *b) {
  const uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
  return (x > y) - (x < y);
}

void orc_sort_pairs(size_t m, uint32_t *keys, uint32_t *vals) {
  uint64_t *kv = (uint64_t *)malloc((m ? m : 1) * sizeof(uint64_t));
  /* Stability: ties on key keep their input order. When the values are
   * non-decreasing in input order (true for emit output), ordering ties by
   * value is the same thing; otherwise ties are ordered by input index. */
  uint32_t *pos_of = NULL;
  int need_pos = 0;
  for (size_t i = 1; i < m; ++i)
    if (vals[i] < vals[i - 1]) { need_pos = 1; break; }
  if (!need_pos) {
    for (size_t i = 0; i < m; ++i) kv[i] = ((uint64_t)keys[i] << 32) | vals[i];
    qsort(kv, m, sizeof(uint64_t), cmp_u64);
    for (size_t i = 0; i < m; ++i) {
      keys[i] = (uint32_t)(kv[i] >> 32);
      vals[i] = (uint32_t)kv[i];
    }
  } else {
    /* general stable path: sort (key, original index) */
    pos_of = (uint32_t *)malloc(m * sizeof(uint32_t));
    memcpy(pos_of, vals, m * sizeof(uint32_t));
    for (size_t i = 0; i < m; ++i) kv[i] = ((uint64_t)keys[i] << 32) | (uint64_t)i;
    qsort(kv, m, sizeof(uint64_t), cmp_u64);
    for (size_t i = 0; i < m; ++i) {
      keys[i] = (uint32_t)(kv[i] >> 32);
      vals[i] = pos_of[(uint32_t)kv[i]];
    }
    free(pos_of);
  }
  free(kv);
}

/* Parallel stable LSD radix sort of the same pairs (8-bit digits over the
 * key's significant bytes): each thread counts its contiguous slice, the
 * per-(digit, thread) counts are scanned digit-major, and each thread
 * scatters its slice in order, so equal keys keep their input order. The
 * same output as orc_sort_pairs (a stable sort by key); it is the CPU
 * baseline's sort. */
void orc_sort_pairs_radix(size_t m, uint32_t *keys, uint32_t *vals) {
  if (m < 2) return;
  uint32_t kmax = 0;
  for (size_t i = 0; i < m; ++i) kmax |= keys[i];
  int passes = 0;
  while (passes < 4 && (kmax >> (8 * passes))) ++passes;
  uint32_t *k2 = (uint32_t *)malloc(m * sizeof(uint32_t));
  uint32_t *v2 = (uint32_t *)malloc(m * sizeof(uint32_t));
  int nt = 1;
#ifdef _OPENMP
  nt = omp_get_max_threads();
#endif
  size_t *cnt = (size_t *)calloc((size_t)nt * 256, sizeof(size_t));
  uint32_t *ks = keys, *vs = vals, *kd = k2, *vd = v2;
  for (int p = 0; p < passes; ++p) {
    const int sh = 8 * p;
    memset(cnt, 0, (size_t)nt * 256 * sizeof(size_t));
#pragma omp parallel num_threads(nt)
    {
      int t = 0;
#ifdef _OPENMP
      t = omp_get_thread_num();
#endif
      const size_t lo = m * (size_t)t / (size_t)nt, hi = m * (size_t)(t + 1) / (size_t)nt;
      size_t *c = cnt + (size_t)t * 256;
      for (size_t i = lo; i < hi; ++i) c[(ks[i] >> sh) & 255u]++;
#pragma omp barrier
#pragma omp single
      {
        size_t run = 0;
        for (int d = 0; d < 256; ++d)
          for (int u = 0; u < nt; ++u) {
            const size_t x = cnt[(size_t)u * 256 + d];
            cnt[(size_t)u * 256 + d] = run;
            run += x;
          }
      }
      for (size_t i = lo; i < hi; ++i) {
        const size_t o = c[(ks[i] >> sh) & 255u]++;
        kd[o] = ks[i];
        vd[o] = vs[i];
      }
    }
    uint32_t *tk = ks, *tv = vs;
    ks = kd;
    vs = vd;
    kd = tk;
    vd = tv;
  }
  if (ks != keys) {
    memcpy(keys, ks, m * sizeof(uint32_t));
    memcpy(vals, vs, m * sizeof(uint32_t));
  }
  free(cnt);
  free(k2);
  free(v2);
}

/* ------------------------------------------------------------------ */
/* a9: build_tile_offsets (training-tiles.wgsl:249-271) with the host    */
/* init to I and reverse-min fix-up (gaussian-renderer.ts:1966-2020).   */
/* ------------------------------------------------------------------ */
void orc_tile_offsets(size_t m, const uint32_t *keys, uint32_t n_tiles,
                      uint32_t *offsets) {
  for (uint32_t t = 0; t <= n_tiles; ++t) offsets[t] = (uint32_t)m;
  for (size_t i = 0; i < m; ++i) {
    const uint32_t tile_and_cam = keys[i] >> 16;
    const uint32_t tile_id = tile_and_cam & 0xfffu;
    const uint32_t cam = tile_and_cam >> 12;
    const uint32_t idx = cam * n_tiles + tile_id;
    if (idx < n_tiles && (uint32_t)i < offsets[idx]) offsets[idx] = (uint32_t)i;
  }
  for (int64_t t = (int64_t)n_tiles - 1; t >= 0; --t)
    if (offsets[t] > offsets[t + 1]) offsets[t] = offsets[t + 1];
}

/* ------------------------------------------------------------------ */
/* a10: training_tiled_forward, training-forward-tiles.wgsl:107-215     */
/* ------------------------------------------------------------------ */
#define ORC_NO_LAST 0xffffffffu

void orc_forward(const orc_params *p, const orc_proj *proj, const orc_camera *cam,
                 const orc_view *v, const uint32_t *offsets, const uint32_t *vals,
                 size_t m, float *color, float *alpha, uint32_t *last_ids) {
  const uint32_t W = v->width, H = v->height;
  const uint32_t n_tiles = v->tiles_x * v->tiles_y;
  const float ts = (float)v->tile_size;
#pragma omp parallel for schedule(dynamic, 4)
  for (int64_t gy = 0; gy < (int64_t)H; ++gy) {
    for (uint32_t gx = 0; gx < W; ++gx) {
      const size_t pix = (size_t)gy * W + gx;
      const float px = (float)gx + 0.5f, py = (float)gy + 0.5f;
      const uint32_t tile_x = f2u_sat(px / ts), tile_y = f2u_sat(py / ts);
      if (tile_x >= v->tiles_x || tile_y >= v->tiles_y) continue;
      const uint32_t tile_id = tile_y * v->tiles_x + tile_x;
      const uint32_t start = offsets[tile_id];
      const uint32_t end = (tile_id + 1u < n_tiles + 1u) ? offsets[tile_id + 1] : (uint32_t)m;
      if (start >= end || start >= (uint32_t)m) {
        color[3 * pix] = v->bg[0]; color[3 * pix + 1] = v->bg[1]; color[3 * pix + 2] = v->bg[2];
        alpha[pix] = 0.0f;
        last_ids[pix] = ORC_NO_LAST; /* the reference leaves it stale */
        continue;
      }
      float col[3] = {0.0f, 0.0f, 0.0f};
      float a = 0.0f;
      uint32_t last = start;
      for (uint32_t i = start; i < end; ++i) {
        last = i;
        const uint32_t g = vals[i];
        const orc_proj *gp = &proj[g];
        const float ca = gp->conic[0], cb = gp->conic[1], cc = gp->conic[2];
        const float dx = px - gp->mean2d[0], dy = py - gp->mean2d[1];
        const float s = 0.5f * (ca * dx * dx + cc * dy * dy) + cb * dx * dy;
        if (s < 0.0f) continue;
        float dir[3], vc[3];
        orc_view_dir(cam, gp->position_world, dir);
        const float opacity = sigmoid_ref(gp->opacity_logit);
        orc_sh_color(dir, &p->sh[48 * (size_t)g], v->sh_deg, vc);
        const float gauss = orc_exp_alpha(-fminf(s, 80.0f));
        const float al = opacity * gauss;
        const float as = fminf(al, MAX_FRAGMENT_ALPHA);
        if (as < MIN_ALPHA) continue;
        const float oma = 1.0f - a;
        const float contrib = as * oma;
        /* the colour sums fused (WGSL leaves the contraction of
         * col + c * contrib to the implementation; the library fuses it);
         * the alpha chain, which decides the skips and the end, unfused */
        col[0] = fmaf(vc[0], contrib, col[0]);
        col[1] = fmaf(vc[1], contrib, col[1]);
        col[2] = fmaf(vc[2], contrib, col[2]);
        a = a + contrib;
        if (1.0f - a < TRANSMITTANCE_THRESHOLD) break;
      }
      const float T = 1.0f - a;
      color[3 * pix] = col[0] + T * v->bg[0];
      color[3 * pix + 1] = col[1] + T * v->bg[1];
      color[3 * pix + 2] = col[2] + T * v->bg[2];
      alpha[pix] = a;
      last_ids[pix] = last;
    }
  }
}

/* ------------------------------------------------------------------ */
/* a11: training_loss, training-loss.wgsl:86-126 (SSIM term :33-84).     */
/* Targets are f32 images at the render resolution, so the bilinear     */
/* sample at texel centres is the texel itself.                         */
/* ------------------------------------------------------------------ */
static inline float signf_ref(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }

static void ssim_grad(const float *color, const float *target, uint32_t W,
                      uint32_t H, int cx, int cy, float out[3]) {
  float mux[3] = {0, 0, 0}, muy[3] = {0, 0, 0};
  const float n = 25.0f;
  for (int dy = -2; dy <= 2; ++dy)
    for (int dx = -2; dx <= 2; ++dx) {
      int x = cx + dx, y = cy + dy;
      int xp = x < 0 ? 0 : (x > (int)W - 1 ? (int)W - 1 : x);
      int yp = y < 0 ? 0 : (y > (int)H - 1 ? (int)H - 1 : y);
      /* samplePred clamps; sampleTarg uses the sampler (clamp-to-edge) */
      for (int c = 0; c < 3; ++c) {
        mux[c] = mux[c] + color[3 * ((size_t)yp * W + xp) + c];
        muy[c] = muy[c] + target[3 * ((size_t)yp * W + xp) + c];
      }
    }
  for (int c = 0; c < 3; ++c) { mux[c] = mux[c] / n; muy[c] = muy[c] / n; }
  float sx2[3] = {0, 0, 0}, sy2[3] = {0, 0, 0}, sxy[3] = {0, 0, 0};
  for (int dy = -2; dy <= 2; ++dy)
    for (int dx = -2; dx <= 2; ++dx) {
      int x = cx + dx, y = cy + dy;
      int xp = x < 0 ? 0 : (x > (int)W - 1 ? (int)W - 1 : x);
      int yp = y < 0 ? 0 : (y > (int)H - 1 ? (int)H - 1 : y);
      for (int c = 0; c < 3; ++c) {
        const float a = color[3 * ((size_t)yp * W + xp) + c] - mux[c];
        const float b = target[3 * ((size_t)yp * W + xp) + c] - muy[c];
        sx2[c] = sx2[c] + a * a;
        sy2[c] = sy2[c] + b * b;
        sxy[c] = sxy[c] + a * b;
      }
    }
  const float C1 = 0.01f * 0.01f, C2 = 0.03f * 0.03f;
  for (int c = 0; c < 3; ++c) {
    sx2[c] = sx2[c] / n; sy2[c] = sy2[c] / n; sxy[c] = sxy[c] / n;
    const float num1 = (2.0f * mux[c]) * muy[c] + C1;
    const float num2 = 2.0f * sxy[c] + C2;
    const float den1 = (mux[c] * mux[c] + muy[c] * muy[c]) + C1;
    const float den2 = (sx2[c] + sy2[c]) + C2;
    const float ssim = (num1 * num2) / (den1 * den2);
    const size_t pc = 3 * ((size_t)cy * W + cx) + c;
    const float dssim = (1.0f - ssim) * 0.5f;
    out[c] = dssim * (color[pc] - target[pc]);
  }
}

void orc_loss(const float *color, const float *target, uint32_t W, uint32_t H,
              float w_l2, float w_l1, float w_ssim, float *residual) {
#pragma omp parallel for schedule(static)
  for (int64_t y = 0; y < (int64_t)H; ++y)
    for (uint32_t x = 0; x < W; ++x) {
      const size_t pix = (size_t)y * W + x;
      float gs[3] = {0, 0, 0};
      if (w_ssim > 0.0f) ssim_grad(color, target, W, H, (int)x, (int)y, gs);
      for (int c = 0; c < 3; ++c) {
        const float d = color[3 * pix + c] - target[3 * pix + c];
        float g = w_l2 * d + w_l1 * signf_ref(d);
        if (w_ssim > 0.0f) g = g + w_ssim * gs[c];
        residual[3 * pix + c] = g;
      }
    }
}

/* ------------------------------------------------------------------ */
/* Per-(pixel, splat) adjoint shared by a12 and a13                      */
/* ------------------------------------------------------------------ */
typedef struct pair_grad {
  float sh[48];
  float logit;
  float conic[3];
  float mean[2];
} pair_grad;

/* One reverse-traversal step; returns 0 if the entry is skipped.
 * Body of training-backward-tiles.wgsl:325-395 (identical to
 * backward-tiles-local.wgsl:296-362 without the 0.98 skip). */
static int adjoint_pair(const orc_params *p, const orc_proj *proj,
                        const orc_camera *cam, const orc_view *v, uint32_t g,
                        float px, float py, const float res[3], float *T_current,
                        float accum_rec[3], int skip_098, pair_grad *out) {
  const orc_proj *gp = &proj[g];
  if (gp->tiles_count == 0u) return 0;
  const float aa = gp->conic[0], bb = gp->conic[1], cc = gp->conic[2];
  const float dx = px - gp->mean2d[0], dy = py - gp->mean2d[1];
  const float s = 0.5f * (aa * dx * dx + cc * dy * dy) + bb * dx * dy;
  if (s < 0.0f) return 0;
  const float gauss = orc_exp_alpha(-fminf(s, 80.0f));
  float dir[3];
  orc_view_dir(cam, gp->position_world, dir);
  const float opacity = sigmoid_ref(gp->opacity_logit);
  const float al = opacity * gauss;
  const float as = fminf(al, MAX_FRAGMENT_ALPHA);
  if (as < MIN_ALPHA) return 0;
  if (skip_098 && as > 0.98f) return 0;
  const float oma = fmaxf(1.0f - as, 0.0001f);
  const float T_before = *T_current / oma;
  const float contrib = as * T_before;
  const float rs[3] = {res[0] * contrib, res[1] * contrib, res[2] * contrib};
  float basis[16];
  sh_basis(dir, v->sh_deg, basis);
  for (int k = 0; k < 16; ++k)
    for (int c = 0; c < 3; ++c) out->sh[3 * k + c] = basis[k] * rs[c];
  float shc[3];
  orc_sh_color(dir, &p->sh[48 * (size_t)g], v->sh_deg, shc);
  const float fg = dot3(res, shc);
  const float bgt = dot3(res, accum_rec);
  const float dL_dalpha = (fg - bgt) * T_before;
  out->logit = (dL_dalpha * as) * (1.0f - opacity);
  const float dL_dG = dL_dalpha * opacity;
  const float dL_ds = dL_dG * (-gauss);
  out->conic[0] = ((dL_ds * 0.5f) * dx) * dx;
  out->conic[1] = (dL_ds * dx) * dy;
  out->conic[2] = ((dL_ds * 0.5f) * dy) * dy;
  const float ddx = dL_ds * (aa * dx + bb * dy);
  const float ddy = dL_ds * (bb * dx + cc * dy);
  out->mean[0] = -ddx;
  out->mean[1] = -ddy;
  for (int c = 0; c < 3; ++c) accum_rec[c] = as * shc[c] + (1.0f - as) * accum_rec[c];
  *T_current = T_before;
  return 1;
}

/* Slot order of tile_grads, backward-reduce.wgsl:2-5. */
static inline void pair_to_slots(const pair_grad *pg, float s[54]) {
  s[0] = pg->logit;
  s[1] = pg->conic[0]; s[2] = pg->conic[1]; s[3] = pg->conic[2];
  s[4] = pg->mean[0]; s[5] = pg->mean[1];
  for (int k = 0; k < 48; ++k) s[6 + k] = pg->sh[k];
}

/* ------------------------------------------------------------------ */
/* a13: training_tiled_backward (legacy, order-correct),                */
/* training-backward-tiles.wgsl:250-397. The float CAS atomics are      */
/* restated as: per intersection, pixels of its tile in row-major order */
/* add into a float partial; then partials are added per Gaussian in    */
/* sorted-intersection order.                                           */
/* ------------------------------------------------------------------ */
void orc_backward_legacy(size_t n, const orc_params *p, const orc_proj *proj,
                         const orc_camera *cam, const orc_view *v,
                         const uint32_t *offsets, const uint32_t *vals, size_t m,
                         const float *residual, const float *alpha,
                         const uint32_t *last_ids, orc_grads *g) {
  const uint32_t W = v->width, H = v->height, ts = v->tile_size;
  const uint32_t n_tiles = v->tiles_x * v->tiles_y;
  float *part = (float *)calloc((m ? m : 1) * 54, sizeof(float));
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t t = 0; t < (int64_t)n_tiles; ++t) {
    const uint32_t start = offsets[t];
    const uint32_t end = offsets[t + 1];
    if (start >= end || start >= (uint32_t)m) continue;
    const uint32_t tx = (uint32_t)t % v->tiles_x, ty = (uint32_t)t / v->tiles_x;
    for (uint32_t ly = 0; ly < ts; ++ly)
      for (uint32_t lx = 0; lx < ts; ++lx) {
        const uint32_t gx = tx * ts + lx, gy = ty * ts + ly;
        if (gx >= W || gy >= H) continue;
        const size_t pix = (size_t)gy * W + gx;
        const float px = (float)gx + 0.5f, py = (float)gy + 0.5f;
        const uint32_t bin_final = last_ids[pix];
        if (bin_final == ORC_NO_LAST || bin_final < start) continue;
        const float res[3] = {residual[3 * pix], residual[3 * pix + 1], residual[3 * pix + 2]};
        float T = 1.0f - alpha[pix];
        float acc[3] = {v->bg[0], v->bg[1], v->bg[2]};
        const uint32_t num = bin_final - start + 1u;
        for (uint32_t j = 0; j < num; ++j) {
          const uint32_t i = bin_final - j;
          if (i < start || i >= end) continue;
          pair_grad pg;
          if (!adjoint_pair(p, proj, cam, v, vals[i], px, py, res, &T, acc, 0, &pg)) continue;
          float s[54];
          pair_to_slots(&pg, s);
          float *dst = &part[(size_t)i * 54];
          for (int k = 0; k < 54; ++k) dst[k] = dst[k] + s[k];
        }
      }
  }
  for (size_t i = 0; i < m; ++i) {
    const uint32_t gi = vals[i];
    const float *s = &part[i * 54];
    g->logit[gi] = g->logit[gi] + s[0];
    for (int k = 0; k < 5; ++k) g->geom[5 * (size_t)gi + k] = g->geom[5 * (size_t)gi + k] + s[1 + k];
    for (int k = 0; k < 48; ++k) g->sh[48 * (size_t)gi + k] = g->sh[48 * (size_t)gi + k] + s[6 + k];
  }
  free(part);
  (void)n;
}

/* ------------------------------------------------------------------ */
/* a12: backward_tiles_local (wavefront, i32 fixed point),               */
/* backward-tiles-local.wgsl:185-382. Output is the i32 LDS sum per      */
/* (intersection, slot) before the /65536 flush (:369-378).              */
/* flags: ORC_WF_REFERENCE_ORDER walks 64-entry batches front-to-back as  */
/* the shader does (:262); without it batches run back-to-front, which   */
/* is order-correct. ORC_WF_SKIP_098 keeps the alpha > 0.98 skip.        */
/* ------------------------------------------------------------------ */
void orc_backward_wavefront(const orc_params *p, const orc_proj *proj,
                            const orc_camera *cam, const orc_view *v,
                            const uint32_t *offsets, const uint32_t *vals, size_t m,
                            const float *residual, const float *alpha,
                            const uint32_t *last_ids, uint32_t flags,
                            int32_t *tile_grads_fixed) {
  const uint32_t W = v->width, H = v->height, ts = v->tile_size;
  const uint32_t n_tiles = v->tiles_x * v->tiles_y;
  memset(tile_grads_fixed, 0, (m ? m : 1) * 54 * sizeof(int32_t));
  if (m == 0) return;
  const int ref_order = (flags & ORC_WF_REFERENCE_ORDER) != 0;
  const int skip098 = (flags & ORC_WF_SKIP_098) != 0;
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t t = 0; t < (int64_t)n_tiles; ++t) {
    const uint32_t start = offsets[t];
    const uint32_t end = offsets[t + 1];
    if (start >= end || start >= (uint32_t)m) continue;
    const uint32_t count = end - start;
    const uint32_t nb = (count + WF_BATCH - 1u) / WF_BATCH;
    const uint32_t tx = (uint32_t)t % v->tiles_x, ty = (uint32_t)t / v->tiles_x;
    uint32_t *acc = (uint32_t *)&tile_grads_fixed[(size_t)start * 54];
    for (uint32_t ly = 0; ly < ts; ++ly)
      for (uint32_t lx = 0; lx < ts; ++lx) {
        const uint32_t gx = tx * ts + lx, gy = ty * ts + ly;
        if (gx >= W || gy >= H) continue;
        const size_t pix = (size_t)gy * W + gx;
        const float px = (float)gx + 0.5f, py = (float)gy + 0.5f;
        const uint32_t bin_final = last_ids[pix];
        if (bin_final == ORC_NO_LAST || bin_final < start) continue;
        const float res[3] = {residual[3 * pix], residual[3 * pix + 1], residual[3 * pix + 2]};
        float T = 1.0f - alpha[pix];
        float accr[3] = {v->bg[0], v->bg[1], v->bg[2]};
        for (uint32_t bi = 0; bi < nb; ++bi) {
          const uint32_t batch = ref_order ? bi : (nb - 1u - bi);
          const uint32_t bs = start + batch * WF_BATCH;
          const uint32_t be = (bs + WF_BATCH < end) ? bs + WF_BATCH : end;
          const uint32_t pe = (be < bin_final + 1u) ? be : bin_final + 1u;
          if (bs >= pe) continue;
          for (uint32_t j = 0; j < pe - bs; ++j) {
            const uint32_t i = pe - 1u - j;
            if (i < start || i > bin_final) continue;
            pair_grad pg;
            if (!adjoint_pair(p, proj, cam, v, vals[i], px, py, res, &T, accr, skip098, &pg))
              continue;
            float s[54];
            pair_to_slots(&pg, s);
            uint32_t *dst = &acc[(size_t)(i - start) * 54];
            for (int k = 0; k < 54; ++k)
              dst[k] = dst[k] + (uint32_t)f2i_sat(s[k] * FIXED_SCALE);
          }
        }
      }
  }
}

/* a14: backward_reduce, backward-reduce.wgsl:35-71. The CAS float adds
 * are restated in sorted-intersection order. */
void orc_backward_reduce(size_t n, const int32_t *fixed, const uint32_t *vals,
                         size_t m, orc_grads *g) {
  const float inv = 1.0f / 65536.0f;
  for (size_t i = 0; i < m; ++i) {
    const uint32_t gi = vals[i];
    const int32_t *s = &fixed[i * 54];
    float v0 = (float)s[0] * inv;
    if (fabsf(v0) > 1e-10f) g->logit[gi] = g->logit[gi] + v0;
    for (int k = 0; k < 5; ++k) {
      float vk = (float)s[1 + k] * inv;
      if (fabsf(vk) > 1e-10f) g->geom[5 * (size_t)gi + k] = g->geom[5 * (size_t)gi + k] + vk;
    }
    for (int k = 0; k < 48; ++k) {
      float vk = (float)s[6 + k] * inv;
      if (fabsf(vk) > 1e-10f) g->sh[48 * (size_t)gi + k] = g->sh[48 * (size_t)gi + k] + vk;
    }
  }
  (void)n;
}

/* ------------------------------------------------------------------ */
/* a15: training_backward_geom, training-backward-geom.wgsl:36-299      */
/* ------------------------------------------------------------------ */
void orc_backward_geom(size_t n, const orc_params *p, const orc_camera *cam,
                       uint32_t geom_flags, orc_grads *g) {
  const float EPS_2D = 0.3f, LIM = 1.3f;
  /* geom_flags bit 0: the mean path's y terms with the forward's pixel
   * convention y = H - c_y - f_y y / z (training-tiles.wgsl:149-151), i.e.
   * the :264-266 terms times -1 (exact) */
  const float ysign = (geom_flags & 1u) ? -1.0f : 1.0f;
#pragma omp parallel for schedule(static)
  for (size_t idx = 0; idx < n; ++idx) {
    const float *gg = &g->geom[5 * idx];
    const float gA = gg[0], gB = gg[1], gC = gg[2], gmx = gg[3], gmy = gg[4];
    const float conic_mag = (fabsf(gA) + fabsf(gB)) + fabsf(gC);
    const float mean_mag = fabsf(gmx) + fabsf(gmy);
    if (conic_mag < 1e-10f && mean_mag < 1e-10f) continue;
    const float pw[4] = {p->pos[3 * idx], p->pos[3 * idx + 1], p->pos[3 * idx + 2], 1.0f};
    float t4[4];
    m4_mul_v(cam->view, pw, t4);
    const float t[3] = {t4[0], t4[1], t4[2]};
    if (t[2] <= 0.0f) continue;
    const float fx = cam->focal[0], fy = cam->focal[1];
    const float tan_fovx = cam->viewport[0] / (2.0f * fx);
    const float tan_fovy = cam->viewport[1] / (2.0f * fy);
    const float limx = LIM * tan_fovx, limy = LIM * tan_fovy;
    const float txtz = t[0] / t[2], tytz = t[1] / t[2];
    const float tc[3] = {clampf(txtz, -limx, limx) * t[2], clampf(tytz, -limy, limy) * t[2], t[2]};
    const float xgm = (txtz >= -limx && txtz <= limx) ? 1.0f : 0.0f;
    const float ygm = (tytz >= -limy && tytz <= limy) ? 1.0f : 0.0f;
    const float tz_inv = 1.0f / tc[2];
    const float tz2_inv = tz_inv * tz_inv;
    m3 J = {{fx * tz_inv, 0.0f, ((-fx) * tc[0]) * tz2_inv},
            {0.0f, fy * tz_inv, ((-fy) * tc[1]) * tz2_inv},
            {0.0f, 0.0f, 0.0f}};
    m3 W;
    /* mat3x3(view[0].x, view[1].x, view[2].x, view[0].y, ...) (:116-120) */
    for (int c = 0; c < 3; ++c)
      for (int r = 0; r < 3; ++r) W[c][r] = cam->view[r * 4 + c];
    m3 Tm;
    m3_mul(W, J, Tm);
    float sigma[3];
    for (int k = 0; k < 3; ++k)
      sigma[k] = orc_expf(clampf(p->logscale[3 * idx + k], -10.0f, 10.0f));
    float q[4];
    normalize4(&p->quat[4 * idx], q);
    m3 R;
    quat_to_mat3(q, R);
    m3 S = {{sigma[0], 0, 0}, {0, sigma[1], 0}, {0, 0, sigma[2]}};
    m3 M, Mt, cov3D;
    m3_mul(S, R, M);
    m3_transpose(M, Mt);
    m3_mul(Mt, M, cov3D);
    m3 Tt, c3t, tmp, cov2D;
    m3_transpose(Tm, Tt);
    m3_transpose(cov3D, c3t);
    m3_mul(Tt, c3t, tmp);
    m3_mul(tmp, Tm, cov2D);
    const float a = cov2D[0][0] + EPS_2D;
    const float b = cov2D[0][1];
    const float c = cov2D[1][1] + EPS_2D;
    const float det = a * c - b * b;
    if (det <= 0.0f) continue;
    /* STEP 1 (:164-168) */
    const float d2i = 1.0f / (det * det + 1e-7f);
    const float dL_da = d2i * ((((-c) * c) * gA + ((2.0f * b) * c) * gB) + (det - a * c) * gC);
    const float dL_dc = d2i * ((((-a) * a) * gC + ((2.0f * a) * b) * gB) + (det - a * c) * gA);
    const float dL_db = (d2i * 2.0f) * ((((b * c) * gA) - (det + (2.0f * b) * b) * gB) + (a * b) * gC);
    /* STEP 2 (:173-185) */
#define TM(i, j) Tm[i][j]
    const float d00 = ((TM(0,0) * TM(0,0)) * dL_da + (TM(0,0) * TM(1,0)) * dL_db) + (TM(1,0) * TM(1,0)) * dL_dc;
    const float d11 = ((TM(0,1) * TM(0,1)) * dL_da + (TM(0,1) * TM(1,1)) * dL_db) + (TM(1,1) * TM(1,1)) * dL_dc;
    const float d22 = ((TM(0,2) * TM(0,2)) * dL_da + (TM(0,2) * TM(1,2)) * dL_db) + (TM(1,2) * TM(1,2)) * dL_dc;
    const float d01 = (((2.0f * TM(0,0)) * TM(0,1)) * dL_da + (TM(0,0) * TM(1,1) + TM(0,1) * TM(1,0)) * dL_db) + ((2.0f * TM(1,0)) * TM(1,1)) * dL_dc;
    const float d02 = (((2.0f * TM(0,0)) * TM(0,2)) * dL_da + (TM(0,0) * TM(1,2) + TM(0,2) * TM(1,0)) * dL_db) + ((2.0f * TM(1,0)) * TM(1,2)) * dL_dc;
    const float d12 = (((2.0f * TM(0,2)) * TM(0,1)) * dL_da + (TM(0,1) * TM(1,2) + TM(0,2) * TM(1,1)) * dL_db) + ((2.0f * TM(1,1)) * TM(1,2)) * dL_dc;
    m3 dS = {{d00, 0.5f * d01, 0.5f * d02}, {0.5f * d01, d11, 0.5f * d12}, {0.5f * d02, 0.5f * d12, d22}};
    /* STEP 3 (:190-192) */
    m3 M2, dM, dMt, Rt;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) M2[i][j] = 2.0f * M[i][j];
    m3_mul(M2, dS, dM);
    m3_transpose(dM, dMt);
    m3_transpose(R, Rt);
    /* STEP 4 (:197-201) */
    const float ds_raw[3] = {dot3(Rt[0], dMt[0]), dot3(Rt[1], dMt[1]), dot3(Rt[2], dMt[2])};
    /* STEP 5 (:206-231) */
    m3 Ms;
    memcpy(Ms, dMt, sizeof(m3));
    for (int j = 0; j < 3; ++j) {
      Ms[0][j] = Ms[0][j] * sigma[0];
      Ms[1][j] = Ms[1][j] * sigma[1];
      Ms[2][j] = Ms[2][j] * sigma[2];
    }
    const float r = q[3], x = q[0], y = q[1], z = q[2];
    float dq[4];
    dq[3] = ((2.0f * z) * (Ms[0][1] - Ms[1][0]) + (2.0f * y) * (Ms[2][0] - Ms[0][2])) +
            (2.0f * x) * (Ms[1][2] - Ms[2][1]);
    dq[0] = (((2.0f * y) * (Ms[1][0] + Ms[0][1]) + (2.0f * z) * (Ms[2][0] + Ms[0][2])) +
             (2.0f * r) * (Ms[1][2] - Ms[2][1])) - (4.0f * x) * (Ms[2][2] + Ms[1][1]);
    dq[1] = (((2.0f * x) * (Ms[1][0] + Ms[0][1]) + (2.0f * r) * (Ms[2][0] - Ms[0][2])) +
             (2.0f * z) * (Ms[1][2] + Ms[2][1])) - (4.0f * y) * (Ms[2][2] + Ms[0][0]);
    dq[2] = (((2.0f * r) * (Ms[0][1] - Ms[1][0]) + (2.0f * x) * (Ms[2][0] + Ms[0][2])) +
             (2.0f * y) * (Ms[1][2] + Ms[2][1])) - (4.0f * z) * (Ms[1][1] + Ms[0][0]);
    /* STEP 6 (:236-259) */
#define C3(i, j) cov3D[i][j]
    const float dT00 = ((2.0f * ((TM(0,0) * C3(0,0) + TM(0,1) * C3(0,1)) + TM(0,2) * C3(0,2))) * dL_da) +
                       (((TM(1,0) * C3(0,0) + TM(1,1) * C3(0,1)) + TM(1,2) * C3(0,2)) * dL_db);
    const float dT01 = ((2.0f * ((TM(0,0) * C3(1,0) + TM(0,1) * C3(1,1)) + TM(0,2) * C3(1,2))) * dL_da) +
                       (((TM(1,0) * C3(1,0) + TM(1,1) * C3(1,1)) + TM(1,2) * C3(1,2)) * dL_db);
    const float dT02 = ((2.0f * ((TM(0,0) * C3(2,0) + TM(0,1) * C3(2,1)) + TM(0,2) * C3(2,2))) * dL_da) +
                       (((TM(1,0) * C3(2,0) + TM(1,1) * C3(2,1)) + TM(1,2) * C3(2,2)) * dL_db);
    const float dT10 = ((2.0f * ((TM(1,0) * C3(0,0) + TM(1,1) * C3(0,1)) + TM(1,2) * C3(0,2))) * dL_dc) +
                       (((TM(0,0) * C3(0,0) + TM(0,1) * C3(0,1)) + TM(0,2) * C3(0,2)) * dL_db);
    const float dT11 = ((2.0f * ((TM(1,0) * C3(1,0) + TM(1,1) * C3(1,1)) + TM(1,2) * C3(1,2))) * dL_dc) +
                       (((TM(0,0) * C3(1,0) + TM(0,1) * C3(1,1)) + TM(0,2) * C3(1,2)) * dL_db);
    const float dT12 = ((2.0f * ((TM(1,0) * C3(2,0) + TM(1,1) * C3(2,1)) + TM(1,2) * C3(2,2))) * dL_dc) +
                       (((TM(0,0) * C3(2,0) + TM(0,1) * C3(2,1)) + TM(0,2) * C3(2,2)) * dL_db);
#undef C3
#undef TM
    const float dJ00 = (W[0][0] * dT00 + W[0][1] * dT01) + W[0][2] * dT02;
    const float dJ02 = (W[2][0] * dT00 + W[2][1] * dT01) + W[2][2] * dT02;
    const float dJ11 = (W[1][0] * dT10 + W[1][1] * dT11) + W[1][2] * dT12;
    const float dJ12 = (W[2][0] * dT10 + W[2][1] * dT11) + W[2][2] * dT12;
    const float tz3_inv = tz2_inv * tz_inv;
    const float dtx_cov = xgm * (((-fx) * tz2_inv) * dJ02);
    const float dty_cov = ygm * (((-fy) * tz2_inv) * dJ12);
    const float dtz_cov = ((((-fx) * tz2_inv) * dJ00 - (fy * tz2_inv) * dJ11) +
                           (((2.0f * fx) * tc[0]) * tz3_inv) * dJ02) +
                          (((2.0f * fy) * tc[1]) * tz3_inv) * dJ12;
    /* STEP 7 (:264-274) */
    const float dtx_mean = (gmx * fx) * tz_inv;
    const float dty_mean = ysign * ((gmy * fy) * tz_inv);
    const float dtz_mean = ((((-gmx) * fx) * t[0]) * tz2_inv) - ysign * (((gmy * fy) * t[1]) * tz2_inv);
    const float dLdt[3] = {dtx_cov + dtx_mean, dty_cov + dty_mean, dtz_cov + dtz_mean};
    m3 Wt;
    m3_transpose(W, Wt);
    float dpos[3];
    m3_mul_v(Wt, dLdt, dpos);
    /* STEP 8-9 (:279-298) */
    for (int k = 0; k < 3; ++k) {
      g->pos[3 * idx + k] = g->pos[3 * idx + k] + dpos[k];
      g->logscale[3 * idx + k] = g->logscale[3 * idx + k] + ds_raw[k] * sigma[k];
    }
    g->quat[4 * idx + 0] = g->quat[4 * idx + 0] + dq[0];
    g->quat[4 * idx + 1] = g->quat[4 * idx + 1] + dq[1];
    g->quat[4 * idx + 2] = g->quat[4 * idx + 2] + dq[2];
    g->quat[4 * idx + 3] = g->quat[4 * idx + 3] + dq[3];
  }
}

/* ------------------------------------------------------------------ */
/* a16: adam_step, adam.wgsl:18-38; bias terms as run_adam_on_buffer   */
/* computes them in double and stores f32 (gaussian-renderer.ts:2425-   */
/* 2440).                                                              */
/* ------------------------------------------------------------------ */
void orc_adam_bias(uint32_t t, float beta1, float beta2, float *bc1_rcp,
                   float *bc2_sqrt_rcp) {
  const double tt = t < 1u ? 1.0 : (double)t;
  const double bc1 = 1.0 - pow((double)beta1, tt);
  const double bc2 = 1.0 - pow((double)beta2, tt);
  *bc1_rcp = (float)(1.0 / bc1);
  *bc2_sqrt_rcp = (float)(1.0 / sqrt(bc2));
}

void orc_adam(size_t count, float *param, float *m1, float *m2, const float *grad,
              float lr, float beta1, float beta2, float eps, float bc1_rcp,
              float bc2_sqrt_rcp) {
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < count; ++i) {
    const float gr = grad[i];
    const float a = beta1 * m1[i] + (1.0f - beta1) * gr;
    const float b = beta2 * m2[i] + ((1.0f - beta2) * gr) * gr;
    const float denom = sqrtf(b) * bc2_sqrt_rcp + eps;
    const float step = lr * bc1_rcp;
    param[i] = param[i] - step * (a / denom);
    m1[i] = a;
    m2[i] = b;
  }
}

/* a17: accumulate_importance (importance-accum.wgsl:19-40), decay 0.98
 * (importance.ts:8-13); stats4 = (importance_sum, visibility, max_r, pad). */
void orc_accumulate_importance(size_t n, const orc_proj *proj, float decay,
                               float *s) {
  for (size_t i = 0; i < n; ++i) {
    float *st = &s[4 * i];
    st[0] = st[0] * decay;
    st[1] = st[1] * decay;
    st[2] = st[2] * decay;
    if (proj[i].tiles_count > 0u) {
      st[0] = st[0] + (float)proj[i].tiles_count;
      st[1] = st[1] + 1.0f;
      st[2] = fmaxf(st[2], proj[i].radius);
    }
  }
}

/* a17: accumulate_grads (densify-accum.wgsl:19-46); accum4 = (|g2d|,
 * 0, |gconic|, count). */
void orc_accumulate_grads(size_t n, const float *gg, float *acc) {
  for (size_t i = 0; i < n; ++i) {
    const float *g = &gg[5 * i];
    const float n2 = sqrtf(g[3] * g[3] + g[4] * g[4]);
    const float na = sqrtf((g[0] * g[0] + g[1] * g[1]) + g[2] * g[2]);
    if (n2 > 1e-10f || na > 1e-10f) {
      acc[4 * i + 0] = acc[4 * i + 0] + n2;
      acc[4 * i + 1] = 0.0f;
      acc[4 * i + 2] = acc[4 * i + 2] + na;
      acc[4 * i + 3] = acc[4 * i + 3] + 1.0f;
    }
  }
}

/* ------------------------------------------------------------------ */
/* Whole render + gradient (trainStep order, gaussian-renderer.ts:     */
/* 3638-3660, with buildTileIntersections :1780-2023 first).           */
/* ------------------------------------------------------------------ */
int64_t orc_render_and_grad(size_t n, const orc_params *p, const orc_camera *cam,
                            const orc_view *v, const float *target, float w_l2,
                            float w_l1, float w_ssim, int use_wavefront,
                            uint32_t wf_flags, orc_step_out *o, orc_grads *g) {
  orc_project(n, p, cam, v, o->proj);
  const uint64_t m = orc_count_isects(n, o->proj);
  if (m > o->cap_isect) return -1;
  orc_emit(n, o->proj, v, o->keys, o->vals);
  orc_sort_pairs_radix((size_t)m, o->keys, o->vals);
  const uint32_t n_tiles = v->tiles_x * v->tiles_y;
  orc_tile_offsets((size_t)m, o->keys, n_tiles, o->offsets);
  orc_forward(p, o->proj, cam, v, o->offsets, o->vals, (size_t)m, o->color, o->alpha,
              o->last_ids);
  if (!target || !g) return (int64_t)m;
  orc_loss(o->color, target, v->width, v->height, w_l2, w_l1, w_ssim, o->residual);
  if (use_wavefront) {
    orc_backward_wavefront(p, o->proj, cam, v, o->offsets, o->vals, (size_t)m, o->residual,
                           o->alpha, o->last_ids, wf_flags, o->tile_grads_fixed);
    orc_backward_reduce(n, o->tile_grads_fixed, o->vals, (size_t)m, g);
  } else {
    orc_backward_legacy(n, p, o->proj, cam, v, o->offsets, o->vals, (size_t)m, o->residual,
                        o->alpha, o->last_ids, g);
  }
  orc_backward_geom(n, p, cam, v->geom_flags, g);
  return (int64_t)m;
}

/* ------------------------------------------------------------------ */
/* Float64 forward (for finite differences). Same algorithm as          */
/* orc_project + emit + stable sort + orc_forward, in double with libm. */
/* ------------------------------------------------------------------ */
typedef double d3[3][3];
static void d3_mul(const d3 A, const d3 B, d3 C) {
  d3 t;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      t[i][j] = A[0][j] * B[i][0] + A[1][j] * B[i][1] + A[2][j] * B[i][2];
  memcpy(C, t, sizeof(d3));
}
static void d3_tr(const d3 A, d3 T) {
  d3 t;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) t[i][j] = A[j][i];
  memcpy(T, t, sizeof(d3));
}
static void d4_mul_v(const float *M, const double v[4], double r[4]) {
  for (int j = 0; j < 4; ++j)
    r[j] = M[j] * v[0] + M[4 + j] * v[1] + M[8 + j] * v[2] + M[12 + j] * v[3];
}

typedef struct dproj {
  double mx, my, radius, depth, a, b, c, opac, rgb[3];
  uint32_t rect[4];
  int alive;
} dproj;

static void sh_color_d(const double dir[3], const double *sh, uint32_t deg, double rgb[3]) {
  const double C0 = 0.28209479177387814, C1 = 0.4886025119029199;
  const double C2[5] = {1.0925484305920792, -1.0925484305920792, 0.31539156525252005,
                        -1.0925484305920792, 0.5462742152960396};
  const double C3[7] = {-0.5900435899266435, 2.890611442640554, -0.4570457994644658,
                        0.3731763325901154, -0.4570457994644658, 1.445305721320277,
                        -0.5900435899266435};
  const double x = dir[0], y = dir[1], z = dir[2];
  double b[16] = {0};
  b[0] = C0;
  if (deg > 0) { b[1] = -C1 * y; b[2] = C1 * z; b[3] = -C1 * x; }
  if (deg > 1) {
    const double xx = x * x, yy = y * y, zz = z * z;
    b[4] = C2[0] * x * y; b[5] = C2[1] * y * z; b[6] = C2[2] * (2 * zz - xx - yy);
    b[7] = C2[3] * x * z; b[8] = C2[4] * (xx - yy);
    if (deg > 2) {
      b[9] = C3[0] * y * (3 * xx - yy); b[10] = C3[1] * x * y * z;
      b[11] = C3[2] * y * (4 * zz - xx - yy); b[12] = C3[3] * z * (2 * zz - 3 * xx - 3 * yy);
      b[13] = C3[4] * x * (4 * zz - xx - yy); b[14] = C3[5] * z * (xx - yy);
      b[15] = C3[6] * x * (xx - 3 * yy);
    }
  }
  for (int c = 0; c < 3; ++c) {
    double s = 0;
    for (int k = 0; k < 16; ++k) s += b[k] * sh[3 * k + c];
    s += 0.5;
    rgb[c] = s > 0 ? s : 0;
  }
}

static int cmp_isect_d(const void *a, const void *b) {
  const uint64_t *x = (const uint64_t *)a, *y = (const uint64_t *)b;
  if (x[0] != y[0]) return x[0] < y[0] ? -1 : 1;
  return (x[1] > y[1]) - (x[1] < y[1]);
}

/* The FP64 forward and loss; with only < n the loss sums the pixels of
 * Gaussian `only`'s tile rectangle alone (0 when it is culled). */
static double forward_loss_f64(size_t n, const double *pos, const double *logscale,
                               const double *quat, const double *logit, const double *sh,
                               const orc_camera *cam, const orc_view *v, const float *target,
                               double w_l2, double w_l1, double *color_out, size_t only) {
  dproj *pr = (dproj *)calloc(n ? n : 1, sizeof(dproj));
  const double W = v->width, H = v->height, ts = v->tile_size;
  uint64_t total = 0;
  for (size_t i = 0; i < n; ++i) {
    dproj *o = &pr[i];
    const double pw[4] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], 1.0};
    double pc[4], cl[4];
    d4_mul_v(cam->view, pw, pc);
    d4_mul_v(cam->proj, pc, cl);
    for (int k = 0; k < 4; ++k) cl[k] /= (k == 3 ? 1.0 : cl[3]);
    if (fabs(cl[0]) > 1.2 || fabs(cl[1]) > 1.2 || pc[2] <= 0) continue;
    if (pc[2] < v->near_plane || pc[2] > v->far_plane) continue;
    double sg[3];
    for (int k = 0; k < 3; ++k) {
      double l = logscale[3 * i + k];
      l = l < -10 ? -10 : (l > 10 ? 10 : l);
      sg[k] = exp(l);
    }
    const double *qq = &quat[4 * i];
    const double ql = sqrt(qq[0] * qq[0] + qq[1] * qq[1] + qq[2] * qq[2] + qq[3] * qq[3]);
    const double x = qq[0] / ql, y = qq[1] / ql, z = qq[2] / ql, w = qq[3] / ql;
    d3 R = {{1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)},
            {2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)},
            {2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)}};
    d3 S = {{sg[0], 0, 0}, {0, sg[1], 0}, {0, 0, sg[2]}};
    d3 SR, SRt, C3D;
    d3_mul(S, R, SR); d3_tr(SR, SRt); d3_mul(SRt, S, C3D); d3_mul(C3D, R, C3D);
    const double fx = cam->focal[0], fy = cam->focal[1];
    const double tx = pc[0], ty = pc[1], tz = pc[2];
    d3 J = {{fx / tz, 0, -(fx * tx) / (tz * tz)}, {0, fy / tz, -(fy * ty) / (tz * tz)}, {0, 0, 0}};
    d3 Rwc;
    for (int c = 0; c < 3; ++c)
      for (int r = 0; r < 3; ++r) Rwc[c][r] = cam->view[c * 4 + r];
    d3 T, TC, Tt, C2;
    d3_mul(J, Rwc, T); d3_mul(T, C3D, TC); d3_tr(T, Tt); d3_mul(TC, Tt, C2);
    C2[0][0] += v->eps2d; C2[1][1] += v->eps2d;
    const double det = C2[0][0] * C2[1][1] - C2[0][1] * C2[0][1];
    if (det <= 0) continue;
    const double mid = 0.5 * (C2[0][0] + C2[1][1]);
    const double disc = sqrt(fmax(mid * mid - det, 0));
    const double rad = ceil(3 * sqrt(fmax(mid + disc, mid - disc)));
    if (rad > v->radius_clip || rad < 1) continue;
    o->a = C2[1][1] / det; o->b = -C2[0][1] / det; o->c = C2[0][0] / det;
    o->mx = 0.5 * (cl[0] + 1) * W; o->my = 0.5 * (cl[1] + 1) * H;
    o->radius = rad; o->depth = pc[2];
    o->opac = 1.0 / (1.0 + exp(-logit[i]));
    double nd[3] = {-pc[0], -pc[1], -pc[2]};
    const double nl = sqrt(nd[0] * nd[0] + nd[1] * nd[1] + nd[2] * nd[2]);
    for (int k = 0; k < 3; ++k) nd[k] /= nl;
    sh_color_d(nd, &sh[48 * i], v->sh_deg, o->rgb);
    const double txf = o->mx / ts, tyf = o->my / ts, tr = rad / ts;
    double f0 = floor(txf - tr), f1 = floor(tyf - tr), f2 = ceil(txf + tr), f3 = ceil(tyf + tr);
    uint32_t r0 = f0 > 0 ? (uint32_t)f0 : 0, r1 = f1 > 0 ? (uint32_t)f1 : 0;
    uint32_t r2 = f2 > 0 ? (uint32_t)f2 : 0, r3 = f3 > 0 ? (uint32_t)f3 : 0;
    o->rect[0] = r0 < v->tiles_x ? r0 : v->tiles_x; o->rect[1] = r1 < v->tiles_y ? r1 : v->tiles_y;
    o->rect[2] = r2 < v->tiles_x ? r2 : v->tiles_x; o->rect[3] = r3 < v->tiles_y ? r3 : v->tiles_y;
    if (o->rect[2] > o->rect[0] && o->rect[3] > o->rect[1]) {
      o->alive = 1;
      total += (uint64_t)(o->rect[2] - o->rect[0]) * (o->rect[3] - o->rect[1]);
    }
  }
  uint64_t *isect = (uint64_t *)malloc((total ? total : 1) * 2 * sizeof(uint64_t));
  size_t cur = 0;
  for (size_t i = 0; i < n; ++i) {
    if (!pr[i].alive) continue;
    double dc = pr[i].depth < 0 ? 0 : (pr[i].depth > 65535 ? 65535 : pr[i].depth);
    const uint64_t dq = (uint64_t)dc & 0xffffu;
    for (uint32_t ty = pr[i].rect[1]; ty < pr[i].rect[3]; ++ty)
      for (uint32_t tx = pr[i].rect[0]; tx < pr[i].rect[2]; ++tx) {
        isect[2 * cur] = ((uint64_t)(ty * v->tiles_x + tx) << 16) | dq;
        isect[2 * cur + 1] = i;
        ++cur;
      }
  }
  qsort(isect, cur, 2 * sizeof(uint64_t), cmp_isect_d);
  const uint32_t nt = v->tiles_x * v->tiles_y;
  size_t *off = (size_t *)malloc((nt + 1) * sizeof(size_t));
  {
    size_t k = 0;
    for (uint32_t t = 0; t <= nt; ++t) {
      while (k < cur && (isect[2 * k] >> 16) < t) ++k;
      off[t] = k;
    }
  }
  double loss = 0;
  const double bg[3] = {v->bg[0], v->bg[1], v->bg[2]};
  // pixel window: the whole image, or Gaussian `only`'s tile rectangle
  uint32_t wx0 = 0, wy0 = 0, wx1 = v->width, wy1 = v->height;
  if (only < n) {
    const dproj *o = &pr[only];
    const uint32_t t = v->tile_size;
    wx0 = wy0 = wx1 = wy1 = 0;
    if (o->alive) {
      wx0 = o->rect[0] * t;
      wy0 = o->rect[1] * t;
      wx1 = o->rect[2] * t < v->width ? o->rect[2] * t : v->width;
      wy1 = o->rect[3] * t < v->height ? o->rect[3] * t : v->height;
    }
  }
#pragma omp parallel for reduction(+ : loss) schedule(dynamic, 4)
  for (int64_t gy = wy0; gy < (int64_t)wy1; ++gy)
    for (uint32_t gx = wx0; gx < wx1; ++gx) {
      const double px = gx + 0.5, py = gy + 0.5;
      const uint32_t tid = (uint32_t)(py / ts) * v->tiles_x + (uint32_t)(px / ts);
      double col[3] = {0, 0, 0}, a = 0;
      for (size_t k = off[tid]; k < off[tid + 1]; ++k) {
        const dproj *o = &pr[isect[2 * k + 1]];
        const double dx = px - o->mx, dy = py - o->my;
        const double s = 0.5 * (o->a * dx * dx + o->c * dy * dy) + o->b * dx * dy;
        if (s < 0) continue;
        double al = o->opac * exp(-s);
        if (al > 0.999) al = 0.999;
        if (al < 1.0 / 255.0) continue;
        const double contrib = al * (1 - a);
        for (int c = 0; c < 3; ++c) col[c] += o->rgb[c] * contrib;
        a += contrib;
        if (1 - a < 1e-4) break;
      }
      const size_t pix = (size_t)gy * v->width + gx;
      for (int c = 0; c < 3; ++c) {
        const double pcol = col[c] + (1 - a) * bg[c];
        if (color_out) color_out[3 * pix + c] = pcol;
        const double d = pcol - (double)target[3 * pix + c];
        loss += 0.5 * w_l2 * d * d + w_l1 * fabs(d);
      }
    }
  free(off);
  free(isect);
  free(pr);
  return loss;
}

double orc_forward_loss_f64(size_t n, const double *pos, const double *logscale,
                            const double *quat, const double *logit, const double *sh,
                            const orc_camera *cam, const orc_view *v, const float *target,
                            double w_l2, double w_l1, double *color_out) {
  return forward_loss_f64(n, pos, logscale, quat, logit, sh, cam, v, target, w_l2, w_l1, color_out, (size_t)-1);
}

double orc_forward_loss_f64_only(size_t n, const double *pos, const double *logscale,
                                 const double *quat, const double *logit, const double *sh,
                                 const orc_camera *cam, const orc_view *v, const float *target,
                                 double w_l2, double w_l1, size_t only) {
  return forward_loss_f64(n, pos, logscale, quat, logit, sh, cam, v, target, w_l2, w_l1, NULL, only);
}

size_t orc_compact_rows(size_t n, size_t stride, const float *src, const uint8_t *keep,
                        float *dst) {
  size_t w = 0;
  for (size_t i = 0; i < n; ++i) {
    if (!keep[i]) continue;
    memcpy(&dst[w * stride], &src[i * stride], stride * sizeof(float));
    ++w;
  }
  return w;
}

/* Vector helpers for the tests. */
void orc_expf_v(size_t n, const float *x, float *y) {
  for (size_t i = 0; i < n; ++i) y[i] = orc_expf(x[i]);
}
void orc_f16_round_v(size_t n, const float *x, float *y) {
  for (size_t i = 0; i < n; ++i) y[i] = orc_f16_round(x[i]);
}
void orc_sh_color_v(size_t n, const float *dirs, const float *sh48, uint32_t deg, float *rgb) {
  for (size_t i = 0; i < n; ++i) orc_sh_color(&dirs[3 * i], &sh48[48 * i], deg, &rgb[3 * i]);
}

/* ------------------------------------------------------------------ */
/* a18: prune / densify compaction                                      */
/* ------------------------------------------------------------------ */
void orc_prune_keep(size_t n, const float *logit, float min_opacity, uint8_t *keep) {
  for (size_t i = 0; i < n; ++i) keep[i] = sigmoid_ref(logit[i]) < min_opacity ? 0u : 1u;
}

static uint64_t splitmix_at(uint64_t seed, uint64_t ctr) {
  uint64_t z = seed + (ctr + 1ull) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static const double ORC_PI = 3.14159265358979323846;
static double unif_at(uint64_t seed, uint64_t ctr) {
  return (double)(splitmix_at(seed, ctr) >> 11) * (1.0 / 9007199254740992.0);
}

size_t orc_densify(size_t n, const float *rows, const float *m1, const float *m2,
                   const uint8_t *actions, uint64_t seed, float *out, float *out_m1,
                   float *out_m2) {
  enum { S = 59 };
  size_t w = 0;
  /* survivors (gaussian-renderer.ts:3344-3370) */
  for (size_t i = 0; i < n; ++i) {
    if (actions[i] == 2) continue;
    memcpy(&out[w * S], &rows[i * S], S * sizeof(float));
    memcpy(&out_m1[w * S], &m1[i * S], S * sizeof(float));
    memcpy(&out_m2[w * S], &m2[i * S], S * sizeof(float));
    ++w;
  }
  /* children (:3372-3433) */
  for (size_t i = 0; i < n; ++i) {
    const uint8_t a = actions[i];
    if (a == 0) continue;
    const float *r = &rows[i * S];
    const double ls[3] = {r[3], r[4], r[5]};
    const double p[3] = {r[0], r[1], r[2]};
    const uint64_t base = 16ull * (uint64_t)i;
    const int nc = a == 1 ? 1 : 2;
    for (int c = 0; c < nc; ++c) {
      float *o = &out[w * S];
      memcpy(o, r, S * sizeof(float));
      memset(&out_m1[w * S], 0, S * sizeof(float));
      memset(&out_m2[w * S], 0, S * sizeof(float));
      double off[3];
      if (a == 1) {
        /* clone: max sigma * 0.25 * (0.5 + 0.5 u) along a random direction */
        const double js = fmax(fmax(exp(ls[0]), exp(ls[1])), exp(ls[2]));
        const double mag = js * 0.25 * (0.5 + unif_at(seed, base + 0) * 0.5);
        double d[3] = {unif_at(seed, base + 1) * 2.0 - 1.0, unif_at(seed, base + 2) * 2.0 - 1.0,
                       unif_at(seed, base + 3) * 2.0 - 1.0};
        double nrm = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        if (nrm == 0.0) nrm = 1.0;
        for (int k = 0; k < 3; ++k) off[k] = d[k] / nrm * mag;
      } else {
        /* split: R * N(0, sigma) (sampleGaussian3D, Box-Muller on four
         * draws, :3103-3120; quatToMat3 of the stored quaternion,
         * :3086-3098), log-scale - ln(1.6) */
        const double sg[3] = {exp(ls[0]), exp(ls[1]), exp(ls[2])};
        const uint64_t cb = base + 4ull * (uint64_t)c;
        const double u1 = unif_at(seed, cb + 0), u2 = unif_at(seed, cb + 1);
        const double u3 = unif_at(seed, cb + 2), u4 = unif_at(seed, cb + 3);
        const double r1 = sqrt(-2.0 * log(1.0 - u1)), r3 = sqrt(-2.0 * log(1.0 - u3));
        const double z[3] = {r1 * cos(2.0 * ORC_PI * u2) * sg[0], r1 * sin(2.0 * ORC_PI * u2) * sg[1],
                             r3 * cos(2.0 * ORC_PI * u4) * sg[2]};
        const double x = r[6], y = r[7], zq = r[8], wq = r[9];
        const double x2 = x + x, y2 = y + y, z2 = zq + zq;
        const double xx = x * x2, xy = x * y2, xz = x * z2, yy = y * y2, yz = y * z2, zz = zq * z2;
        const double wx = wq * x2, wy = wq * y2, wz = wq * z2;
        const double R[3][3] = {{1 - (yy + zz), xy - wz, xz + wy},
                                {xy + wz, 1 - (xx + zz), yz - wx},
                                {xz - wy, yz + wx, 1 - (xx + yy)}};
        for (int k = 0; k < 3; ++k) off[k] = R[k][0] * z[0] + R[k][1] * z[1] + R[k][2] * z[2];
        const double red = log(0.8 * 2.0);
        for (int k = 0; k < 3; ++k) o[3 + k] = (float)(ls[k] - red);
      }
      for (int k = 0; k < 3; ++k) o[k] = (float)(p[k] + off[k]);
      ++w;
    }
  }
  return w;
}

void orc_tile_l1(const orc_view *v, const float *color, const float *target, float *tile_l1) {
  const uint32_t W = v->width, H = v->height, ts = v->tile_size;
  const uint32_t nt = v->tiles_x * v->tiles_y;
  for (uint32_t t = 0; t < nt; ++t) {
    const uint32_t tx = t % v->tiles_x, ty = t / v->tiles_x;
    float slot[256];
    for (uint32_t k = 0; k < 256; ++k) {
      const uint32_t x = tx * ts + (k & 15u), y = ty * ts + (k >> 4);
      slot[k] = 0.0f;
      if (x < W && y < H) {
        const size_t q = ((size_t)y * W + x) * 3;
        slot[k] = (fabsf(color[q] - target[q]) + fabsf(color[q + 1] - target[q + 1])) +
                  fabsf(color[q + 2] - target[q + 2]);
      }
    }
    for (uint32_t s = 128; s > 0; s >>= 1)
      for (uint32_t k = 0; k < s; ++k) slot[k] = slot[k] + slot[k + s];
    tile_l1[t] = slot[0];
  }
}

void orc_densify_decide(size_t n, size_t nviews, const orc_proj *const *proj,
                        const orc_view *views, const float *const *tile_l1,
                        const float *thr, float split_px, uint8_t *actions, float *score) {
  for (size_t i = 0; i < n; ++i) {
    int cand = 0;
    float rmax = 0.0f, best = 0.0f;
    for (size_t s = 0; s < nviews; ++s) {
      const orc_proj *gp = &proj[s][i];
      if (gp->tiles_count == 0u) continue;
      rmax = fmaxf(rmax, gp->radius);
      uint32_t r[4];
      tile_rect(&views[s], gp->mean2d[0], gp->mean2d[1], gp->radius, r);
      for (uint32_t ty = r[1]; ty < r[3]; ++ty)
        for (uint32_t tx = r[0]; tx < r[2]; ++tx) {
          const float e = tile_l1[s][ty * views[s].tiles_x + tx];
          if (e > thr[s]) cand = 1;
          /* normalised score (gaussian-renderer.ts:3280-3286) */
          best = fmaxf(best, thr[s] > 1e-8f ? e / thr[s] : e);
        }
    }
    actions[i] = !cand ? 0u : (rmax > split_px ? 2u : 1u);
    if (score) score[i] = cand ? best : 0.0f;
  }
}

void orc_densify_decide_grad(size_t n, const float *accum4, const float *logscale,
                             const uint32_t *importance, uint32_t n_samples, const double *rule,
                             uint8_t *actions, double *score) {
  /* :3251-3255 */
  const double grad_px = rule[0] * rule[5], abs_px = rule[1] * rule[5];
  const double size_thr = rule[2];
  const double min_vis = rule[3] > 1.0 ? rule[3] : 1.0;
  const double samples = n_samples > 1u ? (double)n_samples : 1.0;
  for (size_t i = 0; i < n; ++i) {
    uint8_t act = 0;
    double sc = 0.0;
    const double denom = accum4[4 * i + 3];
    if (!(denom < min_vis)) { /* :3263-3266 */
      const double avg = (double)accum4[4 * i] / denom;
      const double avg_abs = (double)accum4[4 * i + 2] / denom;
      const double s0 = exp((double)logscale[3 * i]), s1 = exp((double)logscale[3 * i + 1]),
                   s2 = exp((double)logscale[3 * i + 2]);
      const double smax = fmax(fmax(s0, s1), s2); /* :3270-3275 */
      const double imp = (double)importance[i] / samples;
      if (!(imp <= rule[4])) { /* :3276-3280 */
        if (smax <= size_thr && avg >= grad_px) {
          act = 1;
          sc = grad_px > 1e-8 ? avg / grad_px : avg;
        } else if (smax > size_thr && avg_abs >= abs_px) {
          act = 2;
          sc = abs_px > 1e-8 ? avg_abs / abs_px : avg_abs;
        }
      }
    }
    actions[i] = act;
    if (score) score[i] = sc;
  }
}

/* ------------------------------------------------------------------ */
/* knn_init: knn-init.wgsl:40-90                                        */
/* ------------------------------------------------------------------ */
static void knn_insert_sorted(float d[3], float v) { /* insert_sorted, :27-39 */
  if (v < d[0]) {
    d[2] = d[1];
    d[1] = d[0];
    d[0] = v;
  } else if (v < d[1]) {
    d[2] = d[1];
    d[1] = v;
  } else if (v < d[2]) {
    d[2] = v;
  }
}

void orc_knn_init(size_t n, const float *pos, float extent, float *density, float *logscale) {
  const float min_ratio = 0.0007f, max_ratio = 0.007f, bias_ratio = 0.0015f;
  float *p = (float *)malloc((n ? n : 1) * 3 * sizeof(float));
  for (size_t i = 0; i < 3 * n; ++i) p[i] = orc_f16_round(pos[i]); /* read_position, :20-25 */
#pragma omp parallel for schedule(dynamic, 64)
  for (long long ii = 0; ii < (long long)n; ++ii) {
    const size_t idx = (size_t)ii;
    float d[3] = {1e10f, 1e10f, 1e10f};
    const float *me = p + 3 * idx;
    for (size_t i = 0; i < n; ++i) {
      if (i == idx) continue;
      const float *o = p + 3 * i;
      const float dx = me[0] - o[0], dy = me[1] - o[1], dz = me[2] - o[2];
      knn_insert_sorted(d, sqrtf((dx * dx + dy * dy) + dz * dz)); /* length(my_pos - other_pos) */
    }
    const float mean_dist = (d[0] + d[1] + d[2]) / 3.0f;
    density[idx] = mean_dist;
    const float base_sigma = mean_dist / 3.0f;
    const float min_sigma = extent * min_ratio, max_sigma = extent * max_ratio;
    const float bias_sigma = extent * bias_ratio;
    float sigma = fmaxf(base_sigma, bias_sigma);
    sigma = fminf(fmaxf(sigma, min_sigma), max_sigma); /* clamp */
    sigma = fminf(sigma, 0.1f);                        /* MAX_INITIAL_SIGMA */
    logscale[idx] = orc_f16_round((float)log((double)fmaxf(sigma, 1e-6f)));
  }
  free(p);
}

/* ------------------------------------------------------------------ */
/* Multi-view scoring of one view: metric-map.wgsl:32-147 and          */
/* gaussian-renderer.ts:2709-2745 (stats, threshold uniform)           */
/* ------------------------------------------------------------------ */
int orc_metric_view(const orc_view *v, size_t n, const orc_proj *proj, const uint32_t *offsets,
                    const uint32_t *vals, const uint32_t *last_ids, const float *color,
                    const float *target, uint32_t *counts, double *mean) {
  const uint32_t W = v->width, H = v->height, ts = v->tile_size;
  const size_t P = (size_t)W * H;
  float *loss = (float *)malloc((P ? P : 1) * sizeof(float));
  for (size_t i = 0; i < P; ++i) { /* compute_metric_map, :32-50 */
    const float d0 = fabsf(color[3 * i] - target[3 * i]);
    const float d1 = fabsf(color[3 * i + 1] - target[3 * i + 1]);
    const float d2 = fabsf(color[3 * i + 2] - target[3 * i + 2]);
    loss[i] = (d0 + d1 + d2) / 3.0f;
  }
  /* computePixelLossStats (gaussian-renderer.ts:2709-2745) */
  double mn = INFINITY, mx = -INFINITY, sum = 0.0;
  for (size_t i = 0; i < P; ++i) {
    const double x = loss[i];
    if (!isfinite(x)) continue;
    if (x < mn) mn = x;
    if (x > mx) mx = x;
    sum += x;
  }
  if (!isfinite(mn) || !isfinite(mx)) {
    free(loss);
    return 0;
  }
  *mean = P > 0 ? sum / (double)P : 0.0;
  const float min_f = (float)mn;
  const float inv_range = (float)(mx > mn ? 1.0 / (mx - mn) : 0.0);
  const float thresh = 0.1f; /* MULTI_VIEW_LOSS_THRESH, :70 */
  for (size_t idx = 0; idx < P; ++idx) {
    const float value = (loss[idx] - min_f) * inv_range; /* threshold_metric_map, :60-62 */
    if (!(value > thresh)) continue;
    const uint32_t x = (uint32_t)(idx % W), y = (uint32_t)(idx / W); /* accumulate_metric_counts */
    const uint32_t tx = x / ts, ty = y / ts;
    if (tx >= v->tiles_x || ty >= v->tiles_y) continue;
    const uint32_t tile = ty * v->tiles_x + tx;
    const uint32_t start = offsets[tile], end = offsets[tile + 1];
    if (start >= end) continue;
    const uint32_t last = last_ids[idx];
    if (last < start) continue;
    const uint32_t fin = last < end - 1u ? last : end - 1u;
    const float px = (float)x + 0.5f, py = (float)y + 0.5f;
    for (uint32_t i = start; i <= fin; ++i) {
      const uint32_t g = vals[i];
      if (g >= n) continue;
      const orc_proj *gp = &proj[g];
      const float dx = px - gp->mean2d[0], dy = py - gp->mean2d[1];
      const float s2 = 0.5f * (gp->conic[0] * dx * dx + gp->conic[2] * dy * dy) + gp->conic[1] * dx * dy;
      if (s2 < 0.0f || s2 > 6.0f) continue;
      counts[g] += 1u;
    }
  }
  free(loss);
  return 1;
}
