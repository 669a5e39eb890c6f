// sv_optim.hip -- per-Gaussian backward (SH chain + geometry VJP), Adam,
// the fixed-point reduce, and prune/densify compaction (gfx950).
//
// All per-Gaussian kernels use one thread per Gaussian over the planar
// parameter layout (plane k of parameter k, stride cap), so every load and
// store of a wave is one contiguous 256-byte segment.
#include "sv_internal.hpp"

namespace sv {

// View-independent part of training_backward_geom
// (training-backward-geom.wgsl:127-145).
struct GeomPre {
  float wx, wy, wz;
  float sigma[3];
  float q[4];  // normalised (x, y, z, w)
  M3 R, M, cov3D;
};

__device__ __forceinline__ GeomPre geom_pre(const float prm[11]) {
  GeomPre g;
  g.wx = prm[P_POS];
  g.wy = prm[P_POS + 1];
  g.wz = prm[P_POS + 2];
#pragma unroll
  for (int k = 0; k < 3; ++k) g.sigma[k] = sv_expf(clampf(prm[P_SCALE + k], -10.0f, 10.0f));
  const float d = ((prm[P_QUAT] * prm[P_QUAT] + prm[P_QUAT + 1] * prm[P_QUAT + 1]) +
                   prm[P_QUAT + 2] * prm[P_QUAT + 2]) +
                  prm[P_QUAT + 3] * prm[P_QUAT + 3];
  const float l = sqrtf(d);
#pragma unroll
  for (int k = 0; k < 4; ++k) g.q[k] = prm[P_QUAT + k] / l;
  g.R = quat_to_mat3(g.q[0], g.q[1], g.q[2], g.q[3]);
  M3 S;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) S.m[i][j] = (i == j) ? g.sigma[i] : 0.0f;
  g.M = mul(S, g.R);
  g.cov3D = mul(transpose(g.M), g.M);
  return g;
}

// training_backward_geom (training-backward-geom.wgsl:36-299) for one
// view: dL/d(conic a,b,c; mean x,y) -> dL/d(position, log-scale,
// quaternion). Operation order is that of the WGSL source (frustum clamp
// 1.3 tan_fov, EWA covariance with EPS_2D = 0.3, 3DGS conic VJP, M = S R,
// cov3D = M^T M, the quaternion formula on the normalised q, the mean path
// with +fy, dL/dlog(sigma) = dL/dsigma * sigma). Returns false where the
// shader returns early. ysign = -1 gives the mean path's y terms the
// forward's convention (pixel y = H - c_y - f_y y / z); +1 is the shader's
// (the sign flips are exact, so +1 is the shader bit for bit).
__device__ __forceinline__ bool geom_vjp(const DevView &V, const GeomPre &G, float gA, float gB,
                                         float gC, float gmx, float gmy, float ysign, float dpos[3],
                                         float dls[3], float dq[4]) {
  const float conic_mag = (fabsf(gA) + fabsf(gB)) + fabsf(gC);
  const float mean_mag = fabsf(gmx) + fabsf(gmy);
  if (conic_mag < 1e-10f && mean_mag < 1e-10f) return false;
  float t4[4];
  m4_mul_pos(V.view, G.wx, G.wy, G.wz, t4);
  const float tx = t4[0], ty = t4[1], tz = t4[2];
  if (tz <= 0.0f) return false;
  const float fx = V.focal[0], fy = V.focal[1];
  const float tan_fovx = V.viewport[0] / (2.0f * fx);
  const float tan_fovy = V.viewport[1] / (2.0f * fy);
  const float limx = 1.3f * tan_fovx, limy = 1.3f * tan_fovy;
  const float txtz = tx / tz, tytz = ty / tz;
  const float tcx = clampf(txtz, -limx, limx) * tz;
  const float tcy = clampf(tytz, -limy, limy) * tz;
  const float tcz = tz;
  const float xgm = (txtz >= -limx && txtz <= limx) ? 1.0f : 0.0f;
  const float ygm = (tytz >= -limy && tytz <= limy) ? 1.0f : 0.0f;
  const float tz_inv = 1.0f / tcz;
  const float tz2_inv = tz_inv * tz_inv;
  M3 J;
  J.m[0][0] = fx * tz_inv;
  J.m[0][1] = 0.0f;
  J.m[0][2] = ((-fx) * tcx) * tz2_inv;
  J.m[1][0] = 0.0f;
  J.m[1][1] = fy * tz_inv;
  J.m[1][2] = ((-fy) * tcy) * tz2_inv;
  J.m[2][0] = 0.0f;
  J.m[2][1] = 0.0f;
  J.m[2][2] = 0.0f;
  M3 W;  // mat3x3(view[0].x, view[1].x, view[2].x, view[0].y, ...)
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int r = 0; r < 3; ++r) W.m[c][r] = V.view[r * 4 + c];
  const M3 Tm = mul(W, J);
  const M3 cov2D = mul(mul(transpose(Tm), transpose(G.cov3D)), Tm);
  const float a = cov2D.m[0][0] + 0.3f;
  const float b = cov2D.m[0][1];
  const float c = cov2D.m[1][1] + 0.3f;
  const float det = a * c - b * b;
  if (det <= 0.0f) return false;
  // STEP 1
  const float d2i = 1.0f / (det * det + 1e-7f);
  const float dL_da = d2i * ((((-c) * c) * gA + ((2.0f * b) * c) * gB) + (det - a * c) * gC);
  const float dL_dc = d2i * ((((-a) * a) * gC + ((2.0f * a) * b) * gB) + (det - a * c) * gA);
  const float dL_db = (d2i * 2.0f) * ((((b * c) * gA) - (det + (2.0f * b) * b) * gB) + (a * b) * gC);
  // STEP 2
#define TM(i, j) Tm.m[i][j]
  const float d00 = ((TM(0, 0) * TM(0, 0)) * dL_da + (TM(0, 0) * TM(1, 0)) * dL_db) + (TM(1, 0) * TM(1, 0)) * dL_dc;
  const float d11 = ((TM(0, 1) * TM(0, 1)) * dL_da + (TM(0, 1) * TM(1, 1)) * dL_db) + (TM(1, 1) * TM(1, 1)) * dL_dc;
  const float d22 = ((TM(0, 2) * TM(0, 2)) * dL_da + (TM(0, 2) * TM(1, 2)) * dL_db) + (TM(1, 2) * TM(1, 2)) * dL_dc;
  const float d01 = (((2.0f * TM(0, 0)) * TM(0, 1)) * dL_da + (TM(0, 0) * TM(1, 1) + TM(0, 1) * TM(1, 0)) * dL_db) +
                    ((2.0f * TM(1, 0)) * TM(1, 1)) * dL_dc;
  const float d02 = (((2.0f * TM(0, 0)) * TM(0, 2)) * dL_da + (TM(0, 0) * TM(1, 2) + TM(0, 2) * TM(1, 0)) * dL_db) +
                    ((2.0f * TM(1, 0)) * TM(1, 2)) * dL_dc;
  const float d12 = (((2.0f * TM(0, 2)) * TM(0, 1)) * dL_da + (TM(0, 1) * TM(1, 2) + TM(0, 2) * TM(1, 1)) * dL_db) +
                    ((2.0f * TM(1, 1)) * TM(1, 2)) * dL_dc;
  M3 dS;
  dS.m[0][0] = d00;
  dS.m[0][1] = 0.5f * d01;
  dS.m[0][2] = 0.5f * d02;
  dS.m[1][0] = 0.5f * d01;
  dS.m[1][1] = d11;
  dS.m[1][2] = 0.5f * d12;
  dS.m[2][0] = 0.5f * d02;
  dS.m[2][1] = 0.5f * d12;
  dS.m[2][2] = d22;
  // STEP 3
  M3 M2;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) M2.m[i][j] = 2.0f * G.M.m[i][j];
  const M3 dM = mul(M2, dS);
  const M3 dMt = transpose(dM);
  const M3 Rt = transpose(G.R);
  // STEP 4
  const float ds0 = dot3(Rt.m[0][0], Rt.m[0][1], Rt.m[0][2], dMt.m[0][0], dMt.m[0][1], dMt.m[0][2]);
  const float ds1 = dot3(Rt.m[1][0], Rt.m[1][1], Rt.m[1][2], dMt.m[1][0], dMt.m[1][1], dMt.m[1][2]);
  const float ds2 = dot3(Rt.m[2][0], Rt.m[2][1], Rt.m[2][2], dMt.m[2][0], dMt.m[2][1], dMt.m[2][2]);
  // STEP 5
  M3 Ms = dMt;
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    Ms.m[0][j] = Ms.m[0][j] * G.sigma[0];
    Ms.m[1][j] = Ms.m[1][j] * G.sigma[1];
    Ms.m[2][j] = Ms.m[2][j] * G.sigma[2];
  }
  const float r = G.q[3], x = G.q[0], y = G.q[1], z = G.q[2];
  dq[3] = ((2.0f * z) * (Ms.m[0][1] - Ms.m[1][0]) + (2.0f * y) * (Ms.m[2][0] - Ms.m[0][2])) +
          (2.0f * x) * (Ms.m[1][2] - Ms.m[2][1]);
  dq[0] = (((2.0f * y) * (Ms.m[1][0] + Ms.m[0][1]) + (2.0f * z) * (Ms.m[2][0] + Ms.m[0][2])) +
           (2.0f * r) * (Ms.m[1][2] - Ms.m[2][1])) -
          (4.0f * x) * (Ms.m[2][2] + Ms.m[1][1]);
  dq[1] = (((2.0f * x) * (Ms.m[1][0] + Ms.m[0][1]) + (2.0f * r) * (Ms.m[2][0] - Ms.m[0][2])) +
           (2.0f * z) * (Ms.m[1][2] + Ms.m[2][1])) -
          (4.0f * y) * (Ms.m[2][2] + Ms.m[0][0]);
  dq[2] = (((2.0f * r) * (Ms.m[0][1] - Ms.m[1][0]) + (2.0f * x) * (Ms.m[2][0] + Ms.m[0][2])) +
           (2.0f * y) * (Ms.m[1][2] + Ms.m[2][1])) -
          (4.0f * z) * (Ms.m[1][1] + Ms.m[0][0]);
  // STEP 6
#define C3(i, j) G.cov3D.m[i][j]
  const float dT00 = ((2.0f * ((TM(0, 0) * C3(0, 0) + TM(0, 1) * C3(0, 1)) + TM(0, 2) * C3(0, 2))) * dL_da) +
                     (((TM(1, 0) * C3(0, 0) + TM(1, 1) * C3(0, 1)) + TM(1, 2) * C3(0, 2)) * dL_db);
  const float dT01 = ((2.0f * ((TM(0, 0) * C3(1, 0) + TM(0, 1) * C3(1, 1)) + TM(0, 2) * C3(1, 2))) * dL_da) +
                     (((TM(1, 0) * C3(1, 0) + TM(1, 1) * C3(1, 1)) + TM(1, 2) * C3(1, 2)) * dL_db);
  const float dT02 = ((2.0f * ((TM(0, 0) * C3(2, 0) + TM(0, 1) * C3(2, 1)) + TM(0, 2) * C3(2, 2))) * dL_da) +
                     (((TM(1, 0) * C3(2, 0) + TM(1, 1) * C3(2, 1)) + TM(1, 2) * C3(2, 2)) * dL_db);
  const float dT10 = ((2.0f * ((TM(1, 0) * C3(0, 0) + TM(1, 1) * C3(0, 1)) + TM(1, 2) * C3(0, 2))) * dL_dc) +
                     (((TM(0, 0) * C3(0, 0) + TM(0, 1) * C3(0, 1)) + TM(0, 2) * C3(0, 2)) * dL_db);
  const float dT11 = ((2.0f * ((TM(1, 0) * C3(1, 0) + TM(1, 1) * C3(1, 1)) + TM(1, 2) * C3(1, 2))) * dL_dc) +
                     (((TM(0, 0) * C3(1, 0) + TM(0, 1) * C3(1, 1)) + TM(0, 2) * C3(1, 2)) * dL_db);
  const float dT12 = ((2.0f * ((TM(1, 0) * C3(2, 0) + TM(1, 1) * C3(2, 1)) + TM(1, 2) * C3(2, 2))) * dL_dc) +
                     (((TM(0, 0) * C3(2, 0) + TM(0, 1) * C3(2, 1)) + TM(0, 2) * C3(2, 2)) * dL_db);
#undef C3
#undef TM
  const float dJ00 = (W.m[0][0] * dT00 + W.m[0][1] * dT01) + W.m[0][2] * dT02;
  const float dJ02 = (W.m[2][0] * dT00 + W.m[2][1] * dT01) + W.m[2][2] * dT02;
  const float dJ11 = (W.m[1][0] * dT10 + W.m[1][1] * dT11) + W.m[1][2] * dT12;
  const float dJ12 = (W.m[2][0] * dT10 + W.m[2][1] * dT11) + W.m[2][2] * dT12;
  const float tz3_inv = tz2_inv * tz_inv;
  const float dtx_cov = xgm * (((-fx) * tz2_inv) * dJ02);
  const float dty_cov = ygm * (((-fy) * tz2_inv) * dJ12);
  const float dtz_cov = ((((-fx) * tz2_inv) * dJ00 - (fy * tz2_inv) * dJ11) + (((2.0f * fx) * tcx) * tz3_inv) * dJ02) +
                        (((2.0f * fy) * tcy) * tz3_inv) * dJ12;
  // STEP 7
  const float dtx_mean = (gmx * fx) * tz_inv;
  const float dty_mean = ysign * ((gmy * fy) * tz_inv);
  const float dtz_mean = ((((-gmx) * fx) * tx) * tz2_inv) - ysign * (((gmy * fy) * ty) * tz2_inv);
  const float l0 = dtx_cov + dtx_mean, l1 = dty_cov + dty_mean, l2 = dtz_cov + dtz_mean;
  const M3 Wt = transpose(W);
#pragma unroll
  for (int j = 0; j < 3; ++j) dpos[j] = (Wt.m[0][j] * l0 + Wt.m[1][j] * l1) + Wt.m[2][j] * l2;
  // STEP 8
  dls[0] = ds0 * G.sigma[0];
  dls[1] = ds1 * G.sigma[1];
  dls[2] = ds2 * G.sigma[2];
  return true;
}

__device__ __forceinline__ float geom_ysign(const DevConfig &cfg) {
  return (cfg.geom_flags & GEOM_MEAN_Y_FORWARD) ? -1.0f : 1.0f;
}

// accumulate_grads (densify-accum.wgsl:19-46).
__device__ __forceinline__ void densify_accum(float4 &acc, float gA, float gB, float gC, float gmx,
                                              float gmy) {
  const float n2 = sqrtf(gmx * gmx + gmy * gmy);
  const float na = sqrtf((gA * gA + gB * gB) + gC * gC);
  if (n2 > 1e-10f || na > 1e-10f) {
    acc.x = acc.x + n2;
    acc.y = 0.0f;
    acc.z = acc.z + na;
    acc.w = acc.w + 1.0f;
  }
}

// adam_step (adam.wgsl:18-38) on one element.
__device__ __forceinline__ void adam_elem(float &p, float &m1, float &m2, float g, float lr, float b1,
                                          float b2, float eps, float bc1, float bc2s) {
  const float a = b1 * m1 + (1.0f - b1) * g;
  const float b = b2 * m2 + ((1.0f - b2) * g) * g;
  const float denom = sqrtf(b) * bc2s + eps;
  const float step = lr * bc1;
  p = p - step * (a / denom);
  m1 = a;
  m2 = b;
}

__device__ __forceinline__ int tensor_of_plane(int k) {
  return k < P_SCALE ? 0 : (k < P_QUAT ? 1 : (k < P_LOGIT ? 2 : (k < P_SH ? 3 : 4)));
}

// ---------------------------------------------------------------------
// Per-Gaussian backward over the views of a step: sum the Gaussian's
// per-tile adjoint records of each view (emit order, fixed), apply the SH
// chain (grad_sh[k][c] = Y_k(dir_view) * sum res*alpha*T), the opacity
// logit gradient and the geometry VJP, then either add into the gradient
// planes (plain path) or run Adam on all 59 parameters in place (batched
// path; every Gaussian is updated, as the reference's Adam dispatches
// cover all active elements).
// ---------------------------------------------------------------------
// Sum of the adjoint records of the taken pairs of one (slot, Gaussian)
// (slots slot_begin .. slot_begin + slot_count - 1, one grid row each),
// in emit order (ty-major over its tile rectangle): only a pair some pixel
// took has a record, and its flag says so; no pair outside the Gaussian's
// tight rectangle can be taken, so the walk covers that rectangle only.
__global__ __launch_bounds__(256) void k_gather_records(VjpArgs a) {
  if (a.meta && a.meta->overflow) return;
  const size_t M = (size_t)a.nv * a.N;
  // grid: x over Gaussians, y over view slots
  const uint32_t s = (uint32_t)a.slot_begin + blockIdx.y, g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= a.N) return;
  const size_t vg = (size_t)s * a.N + g;
  float r[REC9];
#pragma unroll
  for (int q = 0; q < REC9; ++q) r[q] = 0.0f;
  const uint32_t cnt = a.counts[vg];
  if (cnt) {
    const uint2 rc = a.rects[vg];
    const uint32_t ra = rc.x, rb = rc.y;
    const uint32_t x0 = ra & 0xffffu, y0 = ra >> 16, x1 = rb & 0xffffu, y1 = rb >> 16;
    const TightRect tr = tight_rect(a.tins[vg], x0, y0, x1, y1);
    const uint32_t w = x1 - x0;
    const uint32_t tw = tr.tx1 > tr.tx0 ? (uint32_t)(tr.tx1 - tr.tx0) : 0u;
    const uint32_t n = tr.ty1 > tr.ty0 ? tw * (uint32_t)(tr.ty1 - tr.ty0) : 0u;
    // entry (dx, dy) of the tight rectangle is record (oy + dy) w + ox + dx
    // of the Gaussian's rectangle
    const uint32_t ox = n ? (uint32_t)tr.tx0 - x0 : 0u, oy = n ? (uint32_t)tr.ty0 - y0 : 0u;
    const size_t e0 = (size_t)block_pos(a.bbase, a.nb, s, g, a.offs[vg]) + oy * w + ox;
    const float *rp = a.isect_rec + e0 * REC9;
    const uint8_t *fp = a.rflag + e0;
    bool first = true;
    // four tight-rectangle entries per round: their flag and record loads
    // are issued together, then folded in emit order
    constexpr uint32_t U = 4;
    // (cx, cy): column and row of entry k0, stepped without divisions
    uint32_t cx = 0, cy = 0;
    for (uint32_t k0 = 0; k0 < n; k0 += U) {
      bool ok[U];
      uint32_t ex[U], ey[U];
      float rr[U][REC9];
#pragma unroll
      for (uint32_t u = 0; u < U; ++u) {
        ex[u] = cx;
        ey[u] = cy;
        if (++cx == tw) {
          cx = 0;
          ++cy;
        }
        ok[u] = k0 + u < n && fp[ey[u] * w + ex[u]] != 0u;
      }
#pragma unroll
      for (uint32_t u = 0; u < U; ++u) {
        if (!ok[u]) continue;
        const float *src = rp + (size_t)(ey[u] * w + ex[u]) * REC9;
#pragma unroll
        for (int q = 0; q < REC9; ++q) rr[u][q] = src[q];
      }
#pragma unroll
      for (uint32_t u = 0; u < U; ++u) {
        if (!ok[u]) continue;
        if (first) {
#pragma unroll
          for (int q = 0; q < REC9; ++q) r[q] = rr[u][q];
          first = false;
        } else {
#pragma unroll
          for (int q = 0; q < REC9; ++q) r[q] = r[q] + rr[u][q];
        }
      }
    }
    // the records hold moments of -dL/dpower over pixel offsets (m_xx,
    // m_xy, m_yy, m_x, m_y); with this (view, Gaussian)'s conic (a, b, c):
    // dL/dconic = -(m_xx / 2, m_xy, m_yy / 2), dL/dmean = (a m_x + b m_y,
    // b m_x + c m_y). A pair no pixel reached (most of them: occluded
    // splats) has no record and keeps its zeros, without loading the conic.
    if (!first) {
      const float4 q0 = a.rec[vg * 4];
      const float cc = a.rec[vg * 4 + 1].x;
      const float mx = r[7], my = r[8];
      r[4] = -0.5f * r[4];
      r[5] = -r[5];
      r[6] = -0.5f * r[6];
      r[7] = fmaf(q0.z, mx, q0.w * my);
      r[8] = fmaf(q0.w, mx, cc * my);
    }
  }
#pragma unroll
  for (int q = 0; q < REC9; ++q) a.vrec[(size_t)q * M + vg] = r[q];
}

void launch_gather_records(hipStream_t st, const VjpArgs &a) {
  const size_t M = (size_t)a.nv * a.N;
  if (M == 0) return;
  if (a.slot_count <= 0) return;
  hipLaunchKernelGGL(k_gather_records, dim3((a.N + 255) / 256, (uint32_t)a.slot_count), dim3(256), 0, st, a);
}

// The per-Gaussian backward runs as two kernels with the same arithmetic
// as one: k_vjp_sh (the 48 SH gradients and their Adam) and then
// k_vjp_geom (logit, geometry VJP, statistics and the 11 non-SH
// parameters, whose fused Adam moves the positions k_vjp_sh reads).
// Split, each holds about half the registers, so twice as many waves hide
// the memory latency of the plane reads.
__global__ __launch_bounds__(256) void k_vjp_geom(VjpArgs a) {
  if (a.meta && a.meta->overflow) return;
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= a.N) return;
  const size_t M = (size_t)a.nv * a.N;
  const uint32_t gl = g;
  float prm[11];
#pragma unroll
  for (int k = 0; k < 11; ++k) {
    const float v = a.P[(size_t)k * a.cap + gl];
    prm[k] = a.cfg.render_f16 ? f16_round(v) : v;
  }
  const GeomPre pre = geom_pre(prm);
  float glogit = 0.0f;
  float gp[3] = {0.f, 0.f, 0.f}, gs[3] = {0.f, 0.f, 0.f}, gq[4] = {0.f, 0.f, 0.f, 0.f};
  bool any = false;
  float4 dacc = a.grad_accum ? a.grad_accum[gl] : make_float4(0.f, 0.f, 0.f, 0.f);
  float geo[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
  float4 imp = a.importance ? a.importance[gl] : make_float4(0.f, 0.f, 0.f, 0.f);
  for (int s = 0; s < a.nv; ++s) {
    const size_t vg = (size_t)s * a.N + gl;
    const uint32_t cnt = a.counts[vg];
    // accumulate_importance (importance-accum.wgsl:19-40), once per view
    imp.x = imp.x * a.cfg.importance_decay;
    imp.y = imp.y * a.cfg.importance_decay;
    imp.z = imp.z * a.cfg.importance_decay;
#pragma unroll
    for (int q = 0; q < 5; ++q) geo[q] = 0.0f;
    if (cnt == 0u) continue;
    imp.x = imp.x + (float)cnt;
    imp.y = imp.y + 1.0f;
    const float2 ax = a.aux[vg];
    imp.z = fmaxf(imp.z, ax.y);
    any = true;
    // sum of the Gaussian's reached records of this view (k_gather_records);
    // all zero: no pixel reached it, and every term below is zero
    float r[REC9];
    bool reached = false;
#pragma unroll
    for (int q = 0; q < REC9; ++q) {
      r[q] = a.vrec[(size_t)q * M + vg];
      reached = reached || r[q] != 0.0f;
    }
    if (!reached) continue;
    const DevView &V = a.views[s];
    glogit = glogit + r[3];
    float dp[3], dl[3], dqq[4];
    if (geom_vjp(V, pre, r[4], r[5], r[6], r[7], r[8], geom_ysign(a.cfg), dp, dl, dqq)) {
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        gp[k] = gp[k] + dp[k];
        gs[k] = gs[k] + dl[k];
      }
#pragma unroll
      for (int k = 0; k < 4; ++k) gq[k] = gq[k] + dqq[k];
    }
    densify_accum(dacc, r[4], r[5], r[6], r[7], r[8]);
#pragma unroll
    for (int q = 0; q < 5; ++q) geo[q] = r[4 + q];
  }
  if (a.grad_accum) a.grad_accum[g] = dacc;
  if (a.importance) a.importance[g] = imp;
  if (a.geom) {
#pragma unroll
    for (int q = 0; q < 5; ++q) a.geom[(size_t)q * a.cap + g] = geo[q];
  }
  float grad[P_SH];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    grad[P_POS + k] = gp[k];
    grad[P_SCALE + k] = gs[k];
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) grad[P_QUAT + k] = gq[k];
  grad[P_LOGIT] = glogit;
  if (a.G) {
    if (!any) return;
#pragma unroll
    for (int k = 0; k < P_SH; ++k) {
      float *dst = a.G + (size_t)k * a.cap + g;
      *dst = *dst + grad[k];
    }
    return;
  }
  const AdamParams &H = *a.adam;
#pragma unroll
  for (int k = 0; k < P_SH; ++k) {
    const int t = tensor_of_plane(k);
    const float lr = H.lr[t];
    if (t == 0 && !(lr > 0.0f)) continue;
    const size_t o = (size_t)k * a.cap + g;
    float p = a.Pw[o], m1 = a.M1[o], m2 = a.M2[o];
    adam_elem(p, m1, m2, grad[k], lr, H.beta1, H.beta2, H.eps, H.bc1_rcp, H.bc2_sqrt_rcp);
    a.Pw[o] = p;
    a.M1[o] = m1;
    a.M2[o] = m2;
  }
}

// SH chain: grad_sh[k][c] = sum over views of Y_k(dir_view) * sum res * alpha
// * T (channel c), in view order. A view whose colour sums are all zero adds
// nothing (its products are zeros), as in the skipped views of k_vjp_geom.
__global__ __launch_bounds__(256) void k_vjp_sh(VjpArgs a) {
  if (a.meta && a.meta->overflow) return;
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= a.N) return;
  const size_t M = (size_t)a.nv * a.N;
  float w[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float v = a.P[(size_t)(P_POS + k) * a.cap + g];
    w[k] = a.cfg.render_f16 ? f16_round(v) : v;
  }
  float gsh[48];
#pragma unroll
  for (int k = 0; k < 48; ++k) gsh[k] = 0.0f;
  bool any = false;
  for (int s = 0; s < a.nv; ++s) {
    const size_t vg = (size_t)s * a.N + g;
    if (a.counts[vg] == 0u) continue;
    any = true;
    const float r0 = a.vrec[vg], r1 = a.vrec[M + vg], r2 = a.vrec[2 * M + vg];
    if (r0 == 0.0f && r1 == 0.0f && r2 == 0.0f) continue;
    const DevView &V = a.views[s];
    float pc[4];
    m4_mul_pos(V.view, w[0], w[1], w[2], pc);
    const float nx = -pc[0], ny = -pc[1], nz = -pc[2];
    const float len = sqrtf(dot3(nx, ny, nz, nx, ny, nz));
    float b[16];
    sh_basis(nx / len, ny / len, nz / len, a.cfg.sh_deg, b);
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      gsh[3 * k + 0] = gsh[3 * k + 0] + b[k] * r0;
      gsh[3 * k + 1] = gsh[3 * k + 1] + b[k] * r1;
      gsh[3 * k + 2] = gsh[3 * k + 2] + b[k] * r2;
    }
  }
  if (a.G) {
    if (!any) return;
#pragma unroll
    for (int k = 0; k < 48; ++k) {
      float *dst = a.G + (size_t)(P_SH + k) * a.cap + g;
      *dst = *dst + gsh[k];
    }
    return;
  }
  const AdamParams &H = *a.adam;
  if (!H.update_sh) return;
  const float lr = H.lr[4];
#pragma unroll
  for (int k = 0; k < 48; ++k) {
    const size_t o = (size_t)(P_SH + k) * a.cap + g;
    float p = a.Pw[o], m1 = a.M1[o], m2 = a.M2[o];
    adam_elem(p, m1, m2, gsh[k], lr, H.beta1, H.beta2, H.eps, H.bc1_rcp, H.bc2_sqrt_rcp);
    a.Pw[o] = p;
    a.M1[o] = m1;
    a.M2[o] = m2;
  }
}

void launch_vjp(hipStream_t st, const VjpArgs &a) {
  if (a.N == 0) return;
  // the SH kernel reads the positions (view directions) that the fused Adam
  // of k_vjp_geom updates: it runs first
  hipLaunchKernelGGL(k_vjp_sh, dim3((a.N + 255) / 256), dim3(256), 0, st, a);
  hipLaunchKernelGGL(k_vjp_geom, dim3((a.N + 255) / 256), dim3(256), 0, st, a);
}

// run_adam_updates (gaussian-renderer.ts:2455-2545) over the gradient
// planes: one thread per (plane, Gaussian), learning rate per tensor.
__global__ __launch_bounds__(256) void k_adam(float *__restrict__ P, float *__restrict__ M1,
                                              float *__restrict__ M2, const float *__restrict__ G,
                                              size_t cap, uint32_t N, const AdamParams *__restrict__ hp) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = blockIdx.y;
  if (g >= N) return;
  const AdamParams &H = *hp;
  const int t = tensor_of_plane(k);
  if (t == 4 && !H.update_sh) return;
  const float lr = H.lr[t];
  if (t == 0 && !(lr > 0.0f)) return;
  const size_t o = (size_t)k * cap + g;
  float p = P[o], m1 = M1[o], m2 = M2[o];
  adam_elem(p, m1, m2, G[o], lr, H.beta1, H.beta2, H.eps, H.bc1_rcp, H.bc2_sqrt_rcp);
  P[o] = p;
  M1[o] = m1;
  M2[o] = m2;
}

void launch_adam(hipStream_t st, float *P, float *M1, float *M2, const float *G, size_t cap,
                 uint32_t N, const AdamParams *adam) {
  if (N == 0) return;
  hipLaunchKernelGGL(k_adam, dim3((N + 255) / 256, P_NUM), dim3(256), 0, st, P, M1, M2, G, cap, N, adam);
}

// One wave copies the step's Adam parameters into device memory: from
// `src` (host memory mapped into the device address space, rewritten by
// the host before each step) or, when src is null, from the value passed.
__global__ void k_load_adam(AdamParams v, const AdamParams *src, AdamParams *dst) {
  constexpr int W = sizeof(AdamParams) / 4;
  const int i = threadIdx.x;
  if (i >= W) return;
  const uint32_t *s = src ? (const uint32_t *)src : (const uint32_t *)&v;
  ((uint32_t *)dst)[i] = s[i];
}

void launch_load_adam(hipStream_t st, const AdamParams &value, const AdamParams *src, AdamParams *dst) {
  hipLaunchKernelGGL(k_load_adam, dim3(1), dim3(64), 0, st, value, src, dst);
}

// backward_reduce (backward-reduce.wgsl:35-71): per intersection, the 54
// fixed-point sums / 65536 with |v| > 1e-10 are added with float atomics
// to grad_opacity, grad_geom and grad_sh (order not deterministic, as in
// the reference).
__global__ __launch_bounds__(256) void k_reduce_fixed(const int32_t *__restrict__ fixed,
                                                      const uint32_t *__restrict__ inv,
                                                      const uint32_t *__restrict__ counts,
                                                      const uint32_t *__restrict__ offs,
                                                      const uint32_t *__restrict__ bbase, uint32_t N,
                                                      const StepMeta *m, size_t cap,
                                                      float *__restrict__ G, float *__restrict__ geom) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (m->overflow || g >= N) return;
  const uint32_t cnt = counts[g];
  if (cnt == 0u) return;
  const uint32_t e0 = bbase[g >> 8] + offs[g];
  const uint64_t nI = m->n_isect;  // every sorted position is below it
  const float inv65536 = 1.0f / 65536.0f;
  // the Gaussian's values, added in sorted order as the shader's atomics
  // would in a sequential schedule; skip |v| <= 1e-10 (backward-reduce.wgsl:44-68)
  float acc[6];
  acc[0] = G[(size_t)P_LOGIT * cap + g];
#pragma unroll
  for (int k = 0; k < 5; ++k) acc[1 + k] = geom[(size_t)k * cap + g];
  for (uint32_t e = e0; e < e0 + cnt; ++e) {
    if (inv[e] >= nI) continue;
    const int32_t *s = fixed + (size_t)inv[e] * 54;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      const float v = (float)s[k] * inv65536;
      if (fabsf(v) > 1e-10f) acc[k] = acc[k] + v;
    }
  }
  G[(size_t)P_LOGIT * cap + g] = acc[0];
#pragma unroll
  for (int k = 0; k < 5; ++k) geom[(size_t)k * cap + g] = acc[1 + k];
  // SH: 48 sums, 12 at a time to bound registers
  for (int k0 = 0; k0 < 48; k0 += 12) {
    float sh[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) sh[k] = G[(size_t)(P_SH + k0 + k) * cap + g];
    for (uint32_t e = e0; e < e0 + cnt; ++e) {
      if (inv[e] >= nI) continue;
      const int32_t *s = fixed + (size_t)inv[e] * 54 + 6 + k0;
#pragma unroll
      for (int k = 0; k < 12; ++k) {
        const float v = (float)s[k] * inv65536;
        if (fabsf(v) > 1e-10f) sh[k] = sh[k] + v;
      }
    }
#pragma unroll
    for (int k = 0; k < 12; ++k) G[(size_t)(P_SH + k0 + k) * cap + g] = sh[k];
  }
}

void launch_reduce_fixed(hipStream_t st, const int32_t *fixed, const uint32_t *inv, const uint32_t *counts,
                         const uint32_t *offs, const uint32_t *bbase, uint32_t N, const StepMeta *m, size_t cap,
                         float *G, float *geom) {
  if (N == 0) return;
  hipLaunchKernelGGL(k_reduce_fixed, dim3((N + 255) / 256), dim3(256), 0, st, fixed, inv, counts, offs, bbase, N,
                     m, cap, G, geom);
}

// training_backward_geom from the grad_geom planes (wavefront path) +
// accumulate_grads.
__global__ __launch_bounds__(256) void k_geom(const float *__restrict__ P, size_t cap, uint32_t N,
                                              const DevView *__restrict__ view, DevConfig cfg,
                                              const float *__restrict__ geom, float *__restrict__ G,
                                              float4 *__restrict__ grad_accum,
                                              const float2 *__restrict__ aux,
                                              const uint32_t *__restrict__ counts,
                                              float4 *__restrict__ importance) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= N) return;
  if (importance) {
    float4 imp = importance[g];
    imp.x = imp.x * cfg.importance_decay;
    imp.y = imp.y * cfg.importance_decay;
    imp.z = imp.z * cfg.importance_decay;
    const uint32_t cnt = counts[g];
    if (cnt > 0u) {
      imp.x = imp.x + (float)cnt;
      imp.y = imp.y + 1.0f;
      imp.z = fmaxf(imp.z, aux[g].y);
    }
    importance[g] = imp;
  }
  float prm[11];
#pragma unroll
  for (int k = 0; k < 11; ++k) {
    const float v = P[(size_t)k * cap + g];
    prm[k] = cfg.render_f16 ? f16_round(v) : v;
  }
  float r[5];
#pragma unroll
  for (int q = 0; q < 5; ++q) r[q] = geom[(size_t)q * cap + g];
  if (grad_accum) {
    float4 acc = grad_accum[g];
    densify_accum(acc, r[0], r[1], r[2], r[3], r[4]);
    grad_accum[g] = acc;
  }
  const GeomPre pre = geom_pre(prm);
  float dp[3], dl[3], dq[4];
  if (!geom_vjp(*view, pre, r[0], r[1], r[2], r[3], r[4], geom_ysign(cfg), dp, dl, dq)) return;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    G[(size_t)(P_POS + k) * cap + g] = G[(size_t)(P_POS + k) * cap + g] + dp[k];
    G[(size_t)(P_SCALE + k) * cap + g] = G[(size_t)(P_SCALE + k) * cap + g] + dl[k];
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) G[(size_t)(P_QUAT + k) * cap + g] = G[(size_t)(P_QUAT + k) * cap + g] + dq[k];
}

void launch_geom(hipStream_t st, const float *P, size_t cap, uint32_t N, const DevView *view,
                 const DevConfig &cfg, const float *geom, float *G, float4 *grad_accum,
                 const float2 *aux, const uint32_t *counts, float4 *importance) {
  if (N == 0) return;
  hipLaunchKernelGGL(k_geom, dim3((N + 255) / 256), dim3(256), 0, st, P, cap, N, view, cfg, geom, G,
                     grad_accum, aux, counts, importance);
}

// ---------------------------------------------------------------------
// Prune / densify on the device.
// ---------------------------------------------------------------------
// mode 0: o = a != 0 (keep flags given directly); mode 1: o = (a == 0)
// (a = remove flags); mode 2: o = (a != 2) (densify: originals kept unless
// split); mode 3: o = child count (clone 1, split 2).
__global__ void k_u8_to_u32(const uint8_t *__restrict__ a, uint32_t *__restrict__ o, uint32_t N,
                            int mode) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const uint32_t v = a[i];
  uint32_t r;
  if (mode == 0) r = v != 0u;
  else if (mode == 1) r = v == 0u;
  else if (mode == 2) r = v != 2u;
  else r = v == 1u ? 1u : (v == 2u ? 2u : 0u);
  o[i] = r;
}
void launch_u8_to_u32(hipStream_t st, const uint8_t *a, uint32_t *o, uint32_t N, int mode) {
  if (N == 0) return;
  hipLaunchKernelGGL(k_u8_to_u32, dim3((N + 255) / 256), dim3(256), 0, st, a, o, N, mode);
}

// prune_decide's opacity rule (prune-decide.wgsl:168-180).
__global__ void k_prune_flags(const float *__restrict__ P, size_t cap, uint32_t N, float min_opacity,
                              uint8_t *__restrict__ keep, uint32_t *__restrict__ keep_u32) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= N) return;
  const float op = sv_sigmoid(P[(size_t)P_LOGIT * cap + g]);
  const uint32_t k = (op < min_opacity) ? 0u : 1u;
  keep[g] = (uint8_t)k;
  keep_u32[g] = k;
}
void launch_prune_flags(hipStream_t st, const float *P, size_t cap, uint32_t N, float min_opacity,
                        uint8_t *keep, uint32_t *keep_u32) {
  if (N == 0) return;
  hipLaunchKernelGGL(k_prune_flags, dim3((N + 255) / 256), dim3(256), 0, st, P, cap, N, min_opacity,
                     keep, keep_u32);
}

// resetOpacity (gaussian-renderer.ts:3867-3920): clamp logits to
// max_logit; one counter add per wave of clamped entries.
__global__ void k_clamp_logit(float *__restrict__ P, size_t cap, uint32_t N, float max_logit,
                              uint32_t *__restrict__ n_clamped) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  bool hit = false;
  if (g < N) {
    float *p = P + (size_t)P_LOGIT * cap + g;
    if (*p > max_logit) {
      *p = max_logit;
      hit = true;
    }
  }
  const uint64_t m = __ballot(hit);
  if ((threadIdx.x & 63) == 0 && m) atomicAdd(n_clamped, (uint32_t)__popcll(m));
}
void launch_clamp_logit(hipStream_t st, float *P, size_t cap, uint32_t N, float max_logit, uint32_t *n_clamped) {
  if (N == 0) return;
  hipLaunchKernelGGL(k_clamp_logit, dim3((N + 255) / 256), dim3(256), 0, st, P, cap, N, max_logit, n_clamped);
}

// Order-preserving compaction of `planes` planes (applyPruneMask,
// gaussian-renderer.ts:2917-3000): row g goes to keep_offs[g].
__global__ __launch_bounds__(256) void k_compact(const float *__restrict__ src, float *__restrict__ dst,
                                                 size_t cap_src, size_t cap_dst, int planes,
                                                 const uint32_t *__restrict__ keep_offs,
                                                 const uint8_t *__restrict__ keep, uint32_t N) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= N || !keep[g]) return;
  const uint32_t o = keep_offs[g];
  for (int k = 0; k < planes; ++k) dst[(size_t)k * cap_dst + o] = src[(size_t)k * cap_src + g];
}
void launch_compact(hipStream_t st, const float *src, float *dst, size_t cap_src, size_t cap_dst,
                    int planes, const uint32_t *keep_offs, const uint8_t *keep, uint32_t N) {
  if (N == 0) return;
  hipLaunchKernelGGL(k_compact, dim3((N + 255) / 256), dim3(256), 0, st, src, dst, cap_src, cap_dst,
                     planes, keep_offs, keep, N);
}

// Per-tile L1 error of the batch (scoring for the densify rule): one
// 256-thread workgroup per global tile; slot k = pixel (k & 15, k >> 4) of
// the tile holds (|dr| + |dg|) + |db| (0 outside the view), summed by the
// fixed pairwise tree slot[k] += slot[k + s], s = 128 .. 1.
__global__ __launch_bounds__(256) void k_tile_l1(const DevView *__restrict__ views,
                                                 const uint32_t *__restrict__ tile_view,
                                                 uint32_t n_tiles, const float *__restrict__ color,
                                                 const float *__restrict__ target,
                                                 float *__restrict__ tile_l1) {
  __shared__ float s_v[256];
  const uint32_t tg = blockIdx.x;
  if (tg >= n_tiles) return;
  const DevView &V = views[tile_view[tg]];
  const uint32_t t = tg - V.tile_base;
  const uint32_t x = (t % V.tiles_x) * 16u + (threadIdx.x & 15u);
  const uint32_t y = (t / V.tiles_x) * 16u + (threadIdx.x >> 4);
  float e = 0.0f;
  if (x < V.W && y < V.H) {
    const size_t P = (size_t)V.W * V.H, lp = (size_t)y * V.W + x;
    const float *c = color + 3 * V.pix_off, *g = target + 3 * V.pix_off;
    e = (fabsf(c[lp] - g[lp]) + fabsf(c[P + lp] - g[P + lp])) + fabsf(c[2 * P + lp] - g[2 * P + lp]);
  }
  s_v[threadIdx.x] = e;
  __syncthreads();
  for (uint32_t st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) s_v[threadIdx.x] = s_v[threadIdx.x] + s_v[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) tile_l1[tg] = s_v[0];
}

void launch_tile_l1(hipStream_t st, const DevView *views, const uint32_t *tile_view, uint32_t n_tiles,
                    const float *color, const float *target, float *tile_l1) {
  if (n_tiles == 0) return;
  hipLaunchKernelGGL(k_tile_l1, dim3(n_tiles), dim3(256), 0, st, views, tile_view, n_tiles, color, target,
                     tile_l1);
}

// Densify rule: candidate when a tile of the Gaussian's rectangle in a view
// where it is listed has L1 error above that view's threshold; a candidate
// splits when its largest projected radius over those views exceeds
// split_px, else it is cloned. Replaces the decision block of
// performDensification (gaussian-renderer.ts:3249-3326). score (may be
// null): the largest tile error over the Gaussian's tiles in those views
// divided by the view's threshold (the error itself where the threshold is
// <= 1e-8), the normalised score the reference ranks candidates by
// (:3280-3286); > 1 exactly for candidates with a positive threshold.
__global__ __launch_bounds__(256) void k_densify_decide(const float4 *__restrict__ rec,
                                                        const float2 *__restrict__ aux,
                                                        const uint32_t *__restrict__ counts, uint32_t N,
                                                        const DevView *__restrict__ views, int nv,
                                                        const float *__restrict__ tile_l1,
                                                        const float *__restrict__ thr, float split_px,
                                                        uint8_t *__restrict__ actions, float *__restrict__ score) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= N) return;
  bool cand = false;
  float rmax = 0.0f, best = 0.0f;
  for (int s = 0; s < nv; ++s) {
    const size_t vg = (size_t)s * N + g;
    if (counts[vg] == 0u) continue;
    const float4 q3 = rec[vg * 4 + 3];
    rmax = fmaxf(rmax, aux[vg].y);
    if (cand && !score) continue;
    const DevView &V = views[s];
    const uint32_t a = __float_as_uint(q3.y), b = __float_as_uint(q3.z);
    const uint32_t x0 = a & 0xffffu, y0 = a >> 16, x1 = b & 0xffffu, y1 = b >> 16;
    const float th = thr[s];
    for (uint32_t ty = y0; ty < y1; ++ty)
      for (uint32_t tx = x0; tx < x1; ++tx) {
        const float e = tile_l1[V.tile_base + ty * V.tiles_x + tx];
        if (e > th) cand = true;
        best = fmaxf(best, th > 1e-8f ? e / th : e);
      }
  }
  actions[g] = !cand ? 0u : (rmax > split_px ? 2u : 1u);
  if (score) score[g] = cand ? best : 0.0f;
}

// performDensification's candidate loop (gaussian-renderer.ts:3262-3290),
// in double as the JS evaluates it: visible steps (grad_accum.w) at least
// min_vis; mean 2-D gradient (x / w) and mean conic gradient (z / w);
// largest exp(log-scale); importance / samples above the threshold; clone
// if the largest scale <= size and the mean gradient >= grad_px, split if
// larger and the mean conic gradient >= abs_px; score = mean / threshold
// (the mean itself for a threshold <= 1e-8).
__global__ __launch_bounds__(256) void k_densify_decide_grad(const float4 *__restrict__ grad_accum,
                                                             const float *__restrict__ P, size_t cap, uint32_t N,
                                                             const uint32_t *__restrict__ importance,
                                                             DensifyRuleDev r, uint8_t *__restrict__ actions,
                                                             double *__restrict__ score) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= N) return;
  const float4 acc = grad_accum[g];
  const double denom = (double)acc.w;
  uint8_t act = 0;
  double sc = 0.0;
  if (!(denom < r.min_vis)) {
    const double avg = (double)acc.x / denom, avg_abs = (double)acc.z / denom;
    const double s0 = exp((double)P[(size_t)(P_SCALE + 0) * cap + g]);
    const double s1 = exp((double)P[(size_t)(P_SCALE + 1) * cap + g]);
    const double s2 = exp((double)P[(size_t)(P_SCALE + 2) * cap + g]);
    const double smax = fmax(fmax(s0, s1), s2);
    const double imp = (double)importance[g] / r.samples;
    if (!(imp <= r.imp_threshold)) {
      if (smax <= r.size && avg >= r.grad_px) {
        act = 1;
        sc = r.grad_px > 1e-8 ? avg / r.grad_px : avg;
      } else if (smax > r.size && avg_abs >= r.abs_px) {
        act = 2;
        sc = r.abs_px > 1e-8 ? avg_abs / r.abs_px : avg_abs;
      }
    }
  }
  actions[g] = act;
  score[g] = sc;
}

void launch_densify_decide_grad(hipStream_t st, const float4 *grad_accum, const float *P, size_t cap, uint32_t N,
                                const uint32_t *importance, const DensifyRuleDev &r, uint8_t *actions,
                                double *score) {
  if (N == 0) return;
  hipLaunchKernelGGL(k_densify_decide_grad, dim3((N + 255) / 256), dim3(256), 0, st, grad_accum, P, cap, N,
                     importance, r, actions, score);
}

void launch_densify_decide(hipStream_t st, const float4 *rec, const float2 *aux, const uint32_t *counts, uint32_t N,
                           const DevView *views, int nv, const float *tile_l1, const float *thr,
                           float split_px, uint8_t *actions, float *score) {
  if (N == 0) return;
  hipLaunchKernelGGL(k_densify_decide, dim3((N + 255) / 256), dim3(256), 0, st, rec, aux, counts, N, views, nv,
                     tile_l1, thr, split_px, actions, score);
}

// Counter-based SplitMix64 stream: value `ctr` of stream `seed`.
__device__ __forceinline__ uint64_t rng_u64(uint64_t seed, uint64_t ctr) {
  uint64_t z = seed + (ctr + 1ull) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__device__ __forceinline__ double rng_uniform(uint64_t seed, uint64_t ctr) {
  return (double)(rng_u64(seed, ctr) >> 11) * (1.0 / 9007199254740992.0);
}

// performDensification (gaussian-renderer.ts:3211-3466) with the seeded
// stream above in place of Math.random: originals that are not split go
// to keep_offs[g]; children go to n_keep + child_offs[g]. Clone: one copy
// moved by max(sigma) * 0.25 * (0.5 + 0.5 u) along a random direction
// ((:3390-3398). Split: two copies at R * N(0, sigma) with R =
// quatToMat3(rot) (:3086-3098, un-normalised quaternion as there) and
// log-scale - ln(1.6) (:3421-3424). Children keep SH, opacity and
// rotation and get zero moments (:3334-3340). Stream counters per parent:
// 16 g + {0..3} clone, 16 g + {0..7} split.
__global__ __launch_bounds__(256) void k_densify(const float *__restrict__ P, const float *__restrict__ M1,
                                                 const float *__restrict__ M2, size_t cap, uint32_t N,
                                                 const uint8_t *__restrict__ actions,
                                                 const uint32_t *__restrict__ keep_offs,
                                                 const uint32_t *__restrict__ child_offs, uint32_t n_keep,
                                                 uint64_t seed, float *__restrict__ Pn,
                                                 float *__restrict__ M1n, float *__restrict__ M2n,
                                                 size_t cap_n) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= N) return;
  const uint32_t act = actions[g];
  if (act != 2u) {
    const uint32_t o = keep_offs[g];
    for (int k = 0; k < P_NUM; ++k) {
      Pn[(size_t)k * cap_n + o] = P[(size_t)k * cap + g];
      M1n[(size_t)k * cap_n + o] = M1[(size_t)k * cap + g];
      M2n[(size_t)k * cap_n + o] = M2[(size_t)k * cap + g];
    }
  }
  if (act == 0u) return;
  const uint32_t c0 = n_keep + child_offs[g];
  const double ls0 = P[(size_t)(P_SCALE + 0) * cap + g], ls1 = P[(size_t)(P_SCALE + 1) * cap + g],
               ls2 = P[(size_t)(P_SCALE + 2) * cap + g];
  const double px = P[(size_t)P_POS * cap + g], py = P[(size_t)(P_POS + 1) * cap + g],
               pz = P[(size_t)(P_POS + 2) * cap + g];
  const uint64_t base = 16ull * g;
  const int nchild = act == 1u ? 1 : 2;
  for (int c = 0; c < nchild; ++c) {
    const uint32_t o = c0 + (uint32_t)c;
    for (int k = 0; k < P_NUM; ++k) {
      Pn[(size_t)k * cap_n + o] = P[(size_t)k * cap + g];
      M1n[(size_t)k * cap_n + o] = 0.0f;
      M2n[(size_t)k * cap_n + o] = 0.0f;
    }
    double off[3];
    if (act == 1u) {
      const double js = fmax(fmax(exp(ls0), exp(ls1)), exp(ls2));
      const double mag = js * 0.25 * (0.5 + rng_uniform(seed, base + 0) * 0.5);
      double d[3] = {rng_uniform(seed, base + 1) * 2.0 - 1.0, rng_uniform(seed, base + 2) * 2.0 - 1.0,
                     rng_uniform(seed, base + 3) * 2.0 - 1.0};
      double nrm = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
      if (nrm == 0.0) nrm = 1.0;
      for (int k = 0; k < 3; ++k) off[k] = d[k] / nrm * mag;
    } else {
      const double sg[3] = {exp(ls0), exp(ls1), exp(ls2)};
      const uint64_t cb = base + 4ull * (uint64_t)c;
      const double u1 = rng_uniform(seed, cb + 0), u2 = rng_uniform(seed, cb + 1);
      const double u3 = rng_uniform(seed, cb + 2), u4 = rng_uniform(seed, cb + 3);
      const double r1 = sqrt(-2.0 * log(1.0 - u1)), r3 = sqrt(-2.0 * log(1.0 - u3));
      const double n0 = r1 * cos(2.0 * M_PI * u2), n1 = r1 * sin(2.0 * M_PI * u2);
      const double n2 = r3 * cos(2.0 * M_PI * u4);
      const double ls[3] = {n0 * sg[0], n1 * sg[1], n2 * sg[2]};
      const double x = P[(size_t)(P_QUAT + 0) * cap + g], y = P[(size_t)(P_QUAT + 1) * cap + g],
                   z = P[(size_t)(P_QUAT + 2) * cap + g], w = P[(size_t)(P_QUAT + 3) * cap + g];
      const double x2 = x + x, y2 = y + y, z2 = z + z;
      const double xx = x * x2, xy = x * y2, xz = x * z2, yy = y * y2, yz = y * z2, zz = z * z2;
      const double wx = w * x2, wy = w * y2, wz = w * z2;
      const double Rm[3][3] = {{1 - (yy + zz), xy - wz, xz + wy},
                               {xy + wz, 1 - (xx + zz), yz - wx},
                               {xz - wy, yz + wx, 1 - (xx + yy)}};
      for (int r = 0; r < 3; ++r) off[r] = Rm[r][0] * ls[0] + Rm[r][1] * ls[1] + Rm[r][2] * ls[2];
      const double red = log(0.8 * 2.0);
      Pn[(size_t)(P_SCALE + 0) * cap_n + o] = (float)(ls0 - red);
      Pn[(size_t)(P_SCALE + 1) * cap_n + o] = (float)(ls1 - red);
      Pn[(size_t)(P_SCALE + 2) * cap_n + o] = (float)(ls2 - red);
    }
    Pn[(size_t)(P_POS + 0) * cap_n + o] = (float)(px + off[0]);
    Pn[(size_t)(P_POS + 1) * cap_n + o] = (float)(py + off[1]);
    Pn[(size_t)(P_POS + 2) * cap_n + o] = (float)(pz + off[2]);
  }
}

void launch_densify(hipStream_t st, const float *P, const float *M1, const float *M2, size_t cap,
                    uint32_t N, const uint8_t *actions, const uint32_t *keep_offs,
                    const uint32_t *child_offs, uint32_t n_keep, uint64_t seed, float *Pn, float *M1n,
                    float *M2n, size_t cap_n) {
  if (N == 0) return;
  hipLaunchKernelGGL(k_densify, dim3((N + 255) / 256), dim3(256), 0, st, P, M1, M2, cap, N, actions,
                     keep_offs, child_offs, n_keep, seed, Pn, M1n, M2n, cap_n);
}

}  // namespace sv
