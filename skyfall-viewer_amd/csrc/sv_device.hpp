// sv_device.hpp -- device math and data layout shared by the gfx950 kernels.
//
// All kernels are compiled with -ffp-contract=off and correctly rounded
// f32 division and square root, so every float below is the IEEE result of
// the operation order written here. That order follows the WGSL source of
// the reference (cited per function) so that sort keys, tile ranges,
// images, last indices and fixed-point gradient integers are reproducible
// bit for bit on any IEEE machine.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sv {

// ---------------------------------------------------------------------
// Parameter planes: one f32 plane of `cap` floats per scalar parameter.
// ---------------------------------------------------------------------
enum : int {
  P_POS = 0,    // 3 planes
  P_SCALE = 3,  // 3 planes (log sigma)
  P_QUAT = 6,   // 4 planes (x, y, z, w)
  P_LOGIT = 10, // 1 plane
  P_SH = 11,    // 48 planes, coefficient k channel c at P_SH + 3k + c
  P_NUM = 59
};

constexpr uint32_t NO_LAST = 0xffffffffu;
// Wavefront fixed-point backward flags (same values as SV_WF_* in sv_api.h).
constexpr uint32_t WF_REFERENCE_ORDER = 0x1u;
constexpr uint32_t WF_SKIP_098 = 0x2u;
constexpr int REC9 = 9;  // per-(splat, tile) adjoint record: sum res alpha T (rgb), dL/dlogit, and the
                         // moments of -dL/dpower (m_xx, m_xy, m_yy, m_x, m_y); the record gather turns the
                         // moments into dL/dconic (3) and dL/dmean (2)

// Per-view constants resident in device memory.
struct DevView {
  float view[16];      // world -> camera, column-major (CameraUniforms.view)
  float proj[16];      // CameraUniforms.proj
  float focal[2];
  float viewport[2];
  uint32_t W, H, tiles_x, tiles_y;
  uint32_t tile_base;  // first global tile id of this view
  uint32_t n_tiles;
  uint32_t super_base; // first global supertile id of this view (binned build)
  uint32_t super_x;    // supertiles per row
  uint64_t pix_off;    // first pixel of this view in the pixel arrays
  uint32_t blk_base;   // first global 2x2-tile block of this view (32-px raster tiles)
  uint32_t blk_x;      // blocks per row
};

// Binned tile build: supertiles of BIN_S x BIN_S tiles; a supertile pair
// carries its tile rectangle (3 bits per bound) above the key bits.
constexpr uint32_t BIN_S = 8;
constexpr uint32_t BIN_RECT_BITS = 12;
constexpr uint32_t BIN_CH = 512;  // pairs per chunk of a supertile list

// Values of the current tile build that only the device knows; kernels
// read them so that no launch depends on a host readback (the batched step
// is captured in a hipGraph).
struct StepMeta {
  uint64_t n_isect;            // intersections (written by the scan)
  uint32_t max_dq;             // largest depth quantum of a listed splat
  uint32_t work;               // persistent raster work counter
  uint32_t dbits;              // key bits of the depth quantum
  uint32_t key_bits;           // tile bits + dbits
  uint32_t passes;             // 8-bit sort passes needed
  uint32_t sel;                // ping-pong buffer holding the sorted pairs
  uint32_t overflow;           // n_isect > capacity: nothing past emit runs
  uint32_t pad[7];
};

// Adam hyper-parameters of the current step (run_adam_on_buffer,
// gaussian-renderer.ts:2409-2452), loaded on the device each step.
struct AdamParams {
  float lr[5];  // position, log-scale, quaternion, logit, SH
  float beta1, beta2, eps, bc1_rcp, bc2_sqrt_rcp;
  uint32_t update_sh;
  uint32_t pad;
};

struct DevConfig {
  float bg[3];
  float near_plane, far_plane, radius_clip, eps2d;
  float w_l2, w_l1, w_ssim;
  uint32_t tile_size;
  uint32_t sh_deg;
  uint32_t render_f16;
  float importance_decay;
  uint32_t xcd_remap;
  uint32_t geom_flags;  // SV_GEOM_* of the geometry backward
};
constexpr uint32_t GEOM_MEAN_Y_FORWARD = 0x1u;  // SV_GEOM_MEAN_Y_FORWARD

// ---------------------------------------------------------------------
// Constants of src/shaders/common.wgsl:1-18
// ---------------------------------------------------------------------
#define SV_SH_C0 0.28209479177387814f
#define SV_SH_C1 0.4886025119029199f
#define SV_SH_C2_0 1.0925484305920792f
#define SV_SH_C2_1 -1.0925484305920792f
#define SV_SH_C2_2 0.31539156525252005f
#define SV_SH_C2_3 -1.0925484305920792f
#define SV_SH_C2_4 0.5462742152960396f
#define SV_SH_C3_0 -0.5900435899266435f
#define SV_SH_C3_1 2.890611442640554f
#define SV_SH_C3_2 -0.4570457994644658f
#define SV_SH_C3_3 0.3731763325901154f
#define SV_SH_C3_4 -0.4570457994644658f
#define SV_SH_C3_5 1.445305721320277f
#define SV_SH_C3_6 -0.5900435899266435f

// training-forward-tiles.wgsl:38-41
#define SV_MIN_ALPHA ((float)(1.0 / 255.0))
#define SV_MAX_ALPHA 0.999f
#define SV_T_EPS 1e-4f
// 1 - A < 1e-4 exactly when A > SV_A_DONE, the largest float below the real
// number 1 - 1e-4f (for A >= 1/2 the subtraction is exact, below it both
// tests are false; checked over every float in [0, 2])
#define SV_A_DONE 0x1.fff2e4p-1f

// ---------------------------------------------------------------------
// exp: Cody-Waite reduction by ln2 (split 0.693359375 - 2.12194440e-4) and
// the Cephes expf degree-5 minimax polynomial (S. L. Moshier, Cephes Math
// Library). Only IEEE add/mul/fma/rint are used, so the result is the same
// on the GPU and on any IEEE CPU; accuracy is within 2 ulp of expf.
// ---------------------------------------------------------------------
__device__ __forceinline__ float sv_expf(float x) {
  if (x != x) return x;
  if (x > 88.72283935546875f) return __builtin_inff();
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
    p = p * __uint_as_float((uint32_t)(ni + 64 + 127) << 23);
    return p * __uint_as_float((uint32_t)(127 - 64) << 23);
  }
  if (ni > 127) {
    p = p * 2.0f;
    ni -= 1;
  }
  return p * __uint_as_float((uint32_t)(ni + 127) << 23);
}

// exp of the per-pixel splat weight e^-power (training-forward-tiles.wgsl:
// 190, and the same weight in every backward): 2^(x log2 e) with the product
// rounded to float, k = rint(y), f = y - k in [-1/2, 1/2] (exact), 2^f by a
// degree-5 polynomial with constant term 1, fitted for relative error on
// [-1/2, 1/2] (9.2e-8; Lawson's iteratively reweighted least squares),
// scaled by 2^k through the exponent bits. For x in [-80, 1/2] (the
// rasterizers clamp the power to 80) it is within 3 + 2|x| ulp of e^x, the
// accuracy WGSL requires of exp (at most 2.5 ulp for |x| < 1, 7.1 ulp for
// |x| < 6; e^0 = 1). IEEE mul / fma / rint only, so the result is the same
// float on any IEEE machine. Ten instructions against seventeen for sv_expf.
__device__ __forceinline__ float sv_exp_alpha(float x) {
  const float y = x * 0x1.715476p+0f;
  const float k = rintf(y);
  const float f = y - k;
  float p = 0x1.5bba14p-10f;
  p = fmaf(p, f, 0x1.3cea88p-7f);
  p = fmaf(p, f, 0x1.c6b752p-5f);
  p = fmaf(p, f, 0x1.ebf9bcp-3f);
  p = fmaf(p, f, 0x1.62e42ap-1f);
  p = fmaf(p, f, 1.0f);
  // p 2^k by ldexp: exact for every k of the argument range [-80, 1/2]
  // (2^k normal), and it issues alongside the FP32 multiply-adds
  return __builtin_ldexpf(p, (int)k);
}

// The splat's power 0.5 (a dx^2 + c dy^2) + b dx dy in the shader's order
// (training-forward-tiles.wgsl:183-186). The halving is an ldexp by -1: the
// same correctly rounded value as the multiply by 0.5, subnormals included.
__device__ __forceinline__ float sv_power(float dx, float dy, float a, float b, float c) {
  return __builtin_amdgcn_ldexpf(a * dx * dx + c * dy * dy, -1) + b * dx * dy;
}

// Logistic function as written at training-forward-tiles.wgsl:188.
__device__ __forceinline__ float sv_sigmoid(float logit) {
  return 1.0f / (1.0f + sv_expf(-logit));
}

// WGSL float -> u32: truncation toward zero, saturating, NaN -> 0.
__device__ __forceinline__ uint32_t f2u_sat(float x) {
  if (!(x > 0.0f)) return 0u;
  if (x >= 4294967296.0f) return 0xffffffffu;
  return (uint32_t)x;
}
// WGSL float -> i32: truncation toward zero, saturating, NaN -> 0.
__device__ __forceinline__ int32_t f2i_sat(float x) {
  if (x != x) return 0;
  if (x >= 2147483648.0f) return 2147483647;
  if (x <= -2147483648.0f) return (int32_t)0x80000000u;
  return (int32_t)x;
}
__device__ __forceinline__ float clampf(float x, float lo, float hi) {
  return fminf(fmaxf(x, lo), hi);
}
// f32 -> f16 -> f32, round to nearest even (pack2x16float).
__device__ __forceinline__ float f16_round(float x) {
  return static_cast<float>(static_cast<_Float16>(x));
}

// ---------------------------------------------------------------------
// WGSL matrices: m[col][row]. A * B -> (A*B)[i][j] =
// (A[0][j]*B[i][0] + A[1][j]*B[i][1]) + A[2][j]*B[i][2].
// ---------------------------------------------------------------------
struct M3 {
  float m[3][3];
};

__device__ __forceinline__ M3 mul(const M3 &A, const M3 &B) {
  M3 C;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      C.m[i][j] = (A.m[0][j] * B.m[i][0] + A.m[1][j] * B.m[i][1]) + A.m[2][j] * B.m[i][2];
  return C;
}
__device__ __forceinline__ M3 transpose(const M3 &A) {
  M3 T;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) T.m[i][j] = A.m[j][i];
  return T;
}
__device__ __forceinline__ float dot3(float ax, float ay, float az, float bx, float by, float bz) {
  return (ax * bx + ay * by) + az * bz;
}
// mat4x4 (column-major) * vec4: sum over columns 0..3 in order.
__device__ __forceinline__ void m4_mul_v4(const float *M, const float v[4], float out[4]) {
  float t[4];
#pragma unroll
  for (int j = 0; j < 4; ++j)
    t[j] = ((M[0 * 4 + j] * v[0] + M[1 * 4 + j] * v[1]) + M[2 * 4 + j] * v[2]) + M[3 * 4 + j] * v[3];
#pragma unroll
  for (int j = 0; j < 4; ++j) out[j] = t[j];
}
__device__ __forceinline__ void m4_mul_pos(const float *M, float x, float y, float z,
                                           float out[4]) {
  const float v[4] = {x, y, z, 1.0f};
  m4_mul_v4(M, v, out);
}

// quat_to_mat3_tiling, training-tiles.wgsl:30-37, q = (x, y, z, w); the
// geometry backward's quat_to_mat3 (training-backward-geom.wgsl:21-31) is
// the same matrix.
__device__ __forceinline__ M3 quat_to_mat3(float rx, float ry, float rz, float rw) {
  M3 R;
  R.m[0][0] = 1.0f - 2.0f * (ry * ry + rz * rz);
  R.m[0][1] = 2.0f * (rx * ry - rw * rz);
  R.m[0][2] = 2.0f * (rx * rz + rw * ry);
  R.m[1][0] = 2.0f * (rx * ry + rw * rz);
  R.m[1][1] = 1.0f - 2.0f * (rx * rx + rz * rz);
  R.m[1][2] = 2.0f * (ry * rz - rw * rx);
  R.m[2][0] = 2.0f * (rx * rz - rw * ry);
  R.m[2][1] = 2.0f * (ry * rz + rw * rx);
  R.m[2][2] = 1.0f - 2.0f * (rx * rx + ry * ry);
  return R;
}

// Real SH basis in the reference's evaluation order (the scalar factor of
// each term of computeColorFromSH_tf, training-forward-tiles.wgsl:65-105,
// and of wg_accumulate_sh_grads, backward-tiles-local.wgsl:104-183).
__device__ __forceinline__ void sh_basis(float x, float y, float z, uint32_t deg, float b[16]) {
#pragma unroll
  for (int k = 0; k < 16; ++k) b[k] = 0.0f;
  b[0] = SV_SH_C0;
  if (deg > 0u) {
    b[1] = (-SV_SH_C1) * y;
    b[2] = SV_SH_C1 * z;
    b[3] = (-SV_SH_C1) * x;
    if (deg > 1u) {
      const float xx = x * x, yy = y * y, zz = z * z;
      const float xy = x * y, yz = y * z, xz = x * z;
      b[4] = SV_SH_C2_0 * xy;
      b[5] = SV_SH_C2_1 * yz;
      b[6] = SV_SH_C2_2 * ((2.0f * zz - xx) - yy);
      b[7] = SV_SH_C2_3 * xz;
      b[8] = SV_SH_C2_4 * (xx - yy);
      if (deg > 2u) {
        b[9] = (SV_SH_C3_0 * y) * (3.0f * xx - yy);
        b[10] = (SV_SH_C3_1 * xy) * z;
        b[11] = (SV_SH_C3_2 * y) * ((4.0f * zz - xx) - yy);
        b[12] = (SV_SH_C3_3 * z) * ((2.0f * zz - 3.0f * xx) - 3.0f * yy);
        b[13] = (SV_SH_C3_4 * x) * ((4.0f * zz - xx) - yy);
        b[14] = (SV_SH_C3_5 * z) * (xx - yy);
        b[15] = (SV_SH_C3_6 * x) * (xx - 3.0f * yy);
      }
    }
  }
}

// computeColorFromSH_tf (training-forward-tiles.wgsl:65-105) for one
// channel given the basis b and the 16 coefficients c[k*stride]. The
// reference sums per band: result = C0*c0; result += (t1 + t2) - t3;
// result += (((t4+t5)+t6)+t7)+t8; result += (((((t9+..)+..)+..)+t15);
// max(0, result + 0.5). The degree-1 term is subtracted as (C1*x)*c3; with
// b[3] = (-C1)*x the sum (t1+t2) + b[3]*c3 is the same float because IEEE
// a - b equals a + (-b) and negation is exact.
__device__ __forceinline__ float sh_channel(const float b[16], const float *c, int stride,
                                            uint32_t deg) {
  float res = b[0] * c[0];
  if (deg > 0u) {
    res = res + ((b[1] * c[1 * stride] + b[2] * c[2 * stride]) + b[3] * c[3 * stride]);
    if (deg > 1u) {
      res = res + ((((b[4] * c[4 * stride] + b[5] * c[5 * stride]) + b[6] * c[6 * stride]) +
                    b[7] * c[7 * stride]) + b[8] * c[8 * stride]);
      if (deg > 2u) {
        res = res + ((((((b[9] * c[9 * stride] + b[10] * c[10 * stride]) + b[11] * c[11 * stride]) +
                        b[12] * c[12 * stride]) + b[13] * c[13 * stride]) + b[14] * c[14 * stride]) +
                     b[15] * c[15 * stride]);
      }
    }
  }
  const float t = res + 0.5f;
  return fmaxf(0.0f, t);
}

// Sum of a float over the 64 lanes of a wavefront in a fixed butterfly
// order (deterministic). Every lane receives the total.
__device__ __forceinline__ float wave_sum(float v) {
  v = v + __shfl_xor(v, 1, 64);
  v = v + __shfl_xor(v, 2, 64);
  v = v + __shfl_xor(v, 4, 64);
  v = v + __shfl_xor(v, 8, 64);
  v = v + __shfl_xor(v, 16, 64);
  v = v + __shfl_xor(v, 32, 64);
  return v;
}

}  // namespace sv
