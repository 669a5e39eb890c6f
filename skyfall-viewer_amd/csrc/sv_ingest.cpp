// sv_ingest.cpp -- on-disk inputs of the training step (host only):
// 3DGS / point-cloud PLY files and COLMAP sparse models, decoded into the
// library's f32 arrays (position, log-scale, quaternion (x, y, z, w), logit,
// 48 SH coefficients) and cameras.
//
// Values are rounded to f16 (round to nearest even) as the reference stores
// them in Float16Array buffers (load.ts:90-172, pointcloud-loader.ts:42-100).
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/sv_api.h"

namespace {

thread_local std::string g_ingest_err;

sv_status ingest_fail(sv_status s, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_ingest_err = buf;
  return s;
}

float to_f16(float x) {
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

bool read_file(const char *path, std::vector<unsigned char> &out) {
  FILE *f = std::fopen(path, "rb");
  if (!f) return false;
  std::fseek(f, 0, SEEK_END);
  const long n = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  out.resize(n > 0 ? (size_t)n : 0);
  const bool ok = n >= 0 && std::fread(out.data(), 1, out.size(), f) == out.size();
  std::fclose(f);
  return ok;
}

// ---------------------------------------------------------------- PLY
// decodeHeader / readRawVertex (plyreader.ts:1-75): binary little-endian
// vertices whose properties are `float` (4 bytes) or `uchar` (1 byte, read
// as value / 255), in header order.
struct Ply {
  size_t n = 0;
  std::vector<std::pair<std::string, bool>> props;  // name, is_float
  size_t stride = 0;
  size_t data = 0;  // byte offset of the first vertex
};

sv_status parse_ply(const std::vector<unsigned char> &buf, Ply &p) {
  const std::string head(buf.begin(), buf.begin() + std::min<size_t>(buf.size(), 1 << 16));
  const size_t end = head.find("end_header");
  if (head.compare(0, 3, "ply") != 0 || end == std::string::npos)
    return ingest_fail(SV_E_ARG, "not a PLY file (no ply / end_header)");
  size_t pos = 0;
  bool binary_le = false;
  while (pos < end) {
    size_t nl = head.find('\n', pos);
    if (nl == std::string::npos || nl > end) nl = end;
    std::string line = head.substr(pos, nl - pos);
    while (!line.empty() && (line.back() == '\r' || line.back() == ' ')) line.pop_back();
    pos = nl + 1;
    char a[64], b[64], c[64];
    if (line.rfind("format", 0) == 0) {
      binary_le = line.find("binary_little_endian") != std::string::npos;
    } else if (line.rfind("element vertex", 0) == 0) {
      unsigned long long v = 0;
      if (std::sscanf(line.c_str(), "element vertex %llu", &v) != 1) return ingest_fail(SV_E_ARG, "bad vertex count");
      p.n = (size_t)v;
    } else if (line.rfind("property", 0) == 0 && std::sscanf(line.c_str(), "%63s %63s %63s", a, b, c) == 3) {
      const std::string t = b;
      if (t != "float" && t != "uchar")
        return ingest_fail(SV_E_ARG, "property %s has type %s (float and uchar are read)", c, b);
      p.props.emplace_back(c, t == "float");
      p.stride += t == "float" ? 4 : 1;
    }
  }
  if (!binary_le) return ingest_fail(SV_E_ARG, "only binary_little_endian PLY is read");
  p.data = end + std::strlen("end_header") + 1;
  // n * stride may not fit in size_t for a crafted count: compare by division
  if (p.data > buf.size() || (p.stride != 0 && p.n > (buf.size() - p.data) / p.stride))
    return ingest_fail(SV_E_ARG, "PLY vertex data is truncated");
  return SV_OK;
}

int prop_index(const Ply &p, const std::string &name) {
  for (size_t i = 0; i < p.props.size(); ++i)
    if (p.props[i].first == name) return (int)i;
  return -1;
}

int ply_sh_deg(const Ply &p) {
  int rest = 0;
  for (auto &x : p.props)
    if (x.first.rfind("f_rest_", 0) == 0) ++rest;
  const int per = rest / 3;
  return (int)std::lround(std::sqrt((double)per + 1.0)) - 1;
}

void write_u8(FILE *f, const std::string &s) { std::fwrite(s.data(), 1, s.size(), f); }

// ------------------------------------------------------------- COLMAP
struct Reader {
  const std::vector<unsigned char> &b;
  size_t off = 0;
  bool ok = true;
  // off never passes b.size(), so b.size() - off does not wrap
  template <typename T>
  T get() {
    T v{};
    if (!ok || sizeof(T) > b.size() - off) {
      ok = false;
      return v;
    }
    std::memcpy(&v, b.data() + off, sizeof(T));
    off += sizeof(T);
    return v;
  }
  // skip `count` records of `each` bytes (count comes from the file)
  void skip(uint64_t count, size_t each) {
    if (!ok || (each != 0 && count > (uint64_t)((b.size() - off) / each))) {
      ok = false;
      off = b.size();
      return;
    }
    off += (size_t)count * each;
  }
};

// Parameters per COLMAP camera model id (SIMPLE_PINHOLE, PINHOLE,
// SIMPLE_RADIAL, RADIAL, OPENCV, OPENCV_FISHEYE, FULL_OPENCV, FOV,
// SIMPLE_RADIAL_FISHEYE, RADIAL_FISHEYE, THIN_PRISM_FISHEYE).
int model_params(int id) {
  static const int n[] = {3, 4, 4, 5, 8, 8, 12, 5, 4, 5, 12};
  return (id >= 0 && id < 11) ? n[id] : -1;
}

struct ColmapCam {
  uint32_t id;
  int model;
  uint64_t w, h;
  double fx, fy, cx, cy;
};

// get_pinhole_projection_matrix (training-camera.ts:29-52) after its
// transpose, column-major.
void pinhole_proj(double fx, double fy, double cx, double cy, double w, double h, float out[16]) {
  std::memset(out, 0, 16 * sizeof(float));
  out[0] = (float)(2.0 * fx / w);
  out[5] = (float)(-2.0 * fy / h);
  out[8] = (float)(2.0 * cx / w - 1.0);
  out[9] = (float)(1.0 - 2.0 * cy / h);
  out[10] = 1.0f;
  out[11] = 1.0f;
}

}  // namespace

extern "C" {

const char *sv_ingest_error(void) { return g_ingest_err.c_str(); }

sv_status sv_ply_info(const char *path, size_t *n, uint32_t *sh_deg) {
  std::vector<unsigned char> buf;
  if (!path || !read_file(path, buf)) return ingest_fail(SV_E_ARG, "cannot read %s", path ? path : "(null)");
  Ply p;
  if (sv_status s = parse_ply(buf, p)) return s;
  if (n) *n = p.n;
  if (sh_deg) *sh_deg = prop_index(p, "f_dc_0") >= 0 ? (uint32_t)std::max(0, ply_sh_deg(p)) : 0u;
  return SV_OK;
}

sv_status sv_ply_load(const char *path, size_t n, float *pos, float *logscale, float *quat, float *logit,
                      float *sh, float *scene_extent) {
  if (!pos || !logscale || !quat || !logit || !sh) return ingest_fail(SV_E_ARG, "null output array");
  std::vector<unsigned char> buf;
  if (!path || !read_file(path, buf)) return ingest_fail(SV_E_ARG, "cannot read %s", path ? path : "(null)");
  Ply p;
  if (sv_status s = parse_ply(buf, p)) return s;
  if (n != p.n) return ingest_fail(SV_E_ARG, "expected %zu vertices, the file has %zu", n, p.n);
  // load.ts:30-70: Gaussian format when f_dc_0, opacity, rot_0, scale_0
  // exist; a plain point cloud when red, green, blue exist.
  const int i_dc = prop_index(p, "f_dc_0"), i_op = prop_index(p, "opacity"), i_rot = prop_index(p, "rot_0"),
            i_sc = prop_index(p, "scale_0");
  const bool gauss = i_dc >= 0 && i_op >= 0 && i_rot >= 0 && i_sc >= 0;
  const bool rgb = prop_index(p, "red") >= 0 && prop_index(p, "green") >= 0 && prop_index(p, "blue") >= 0 && !gauss;
  const int deg = gauss ? std::max(0, ply_sh_deg(p)) : 0;
  const int per = (deg + 1) * (deg + 1) - 1;  // f_rest coefficients per colour
  if (deg > 3) return ingest_fail(SV_E_ARG, "SH degree %d > 3", deg);
  std::vector<int> sh_src(48, -1);
  if (gauss) {
    // shFeatureOrder (load.ts:52-62): f_dc_c, then f_rest_{c * per + i}
    for (int c = 0; c < 3; ++c) sh_src[c] = prop_index(p, "f_dc_" + std::to_string(c));
    for (int i = 0; i < per; ++i)
      for (int c = 0; c < 3; ++c) sh_src[3 * (i + 1) + c] = prop_index(p, "f_rest_" + std::to_string(c * per + i));
  }
  std::vector<size_t> off(p.props.size());
  for (size_t k = 0, o = 0; k < p.props.size(); ++k) {
    off[k] = o;
    o += p.props[k].second ? 4 : 1;
  }
  auto val = [&](size_t v, int k) -> float {
    const unsigned char *d = buf.data() + p.data + v * p.stride + off[k];
    if (p.props[k].second) {
      float f;
      std::memcpy(&f, d, 4);
      return f;
    }
    return (float)(*d) / 255.0f;
  };
  const int ix = prop_index(p, "x"), iy = prop_index(p, "y"), iz = prop_index(p, "z");
  if (ix < 0 || iy < 0 || iz < 0) return ingest_fail(SV_E_ARG, "PLY has no x, y, z");
  const int i_r1 = prop_index(p, "rot_1"), i_r2 = prop_index(p, "rot_2"), i_r3 = prop_index(p, "rot_3");
  const int i_s1 = prop_index(p, "scale_1"), i_s2 = prop_index(p, "scale_2");
  if (gauss && (i_r1 < 0 || i_r2 < 0 || i_r3 < 0 || i_s1 < 0 || i_s2 < 0))
    return ingest_fail(SV_E_ARG, "PLY Gaussian format without rot_1..3 / scale_1..2");
  const double SH_C0 = 0.28209479177387814;
  const float default_logit = -2.197f, default_ls = (float)std::log(0.01);
  double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (size_t v = 0; v < p.n; ++v) {
    const float xyz[3] = {val(v, ix), val(v, iy), val(v, iz)};
    for (int k = 0; k < 3; ++k) {
      pos[3 * v + k] = to_f16(xyz[k]);
      mn[k] = std::min(mn[k], (double)xyz[k]);
      mx[k] = std::max(mx[k], (double)xyz[k]);
    }
    for (int j = 0; j < 48; ++j) sh[48 * v + j] = 0.0f;
    if (gauss) {
      for (int j = 0; j < 3 * (per + 1); ++j) sh[48 * v + j] = sh_src[j] >= 0 ? to_f16(val(v, sh_src[j])) : 0.0f;
      logit[v] = to_f16(val(v, i_op));
      // rot_0..3 = (w, x, y, z) in the file and in the reference's half[4..7]
      quat[4 * v + 0] = to_f16(val(v, i_r1));
      quat[4 * v + 1] = to_f16(val(v, i_r2));
      quat[4 * v + 2] = to_f16(val(v, i_r3));
      quat[4 * v + 3] = to_f16(val(v, i_rot));
      logscale[3 * v + 0] = to_f16(val(v, i_sc));
      logscale[3 * v + 1] = to_f16(val(v, i_s1));
      logscale[3 * v + 2] = to_f16(val(v, i_s2));
    } else {
      if (rgb) {
        // load.ts:117-126: the reader already divided uchar colours by 255
        // and the loader divides again (reproduced as written)
        const int ic[3] = {prop_index(p, "red"), prop_index(p, "green"), prop_index(p, "blue")};
        for (int c = 0; c < 3; ++c)
          sh[48 * v + c] = to_f16((float)(((double)val(v, ic[c]) / 255.0 - 0.5) / SH_C0));
      }
      logit[v] = to_f16(default_logit);
      // defaultRot [0, 0, 0, 1] written to the (w, x, y, z) slots (load.ts:157-162)
      quat[4 * v + 0] = 0.0f;
      quat[4 * v + 1] = 0.0f;
      quat[4 * v + 2] = 1.0f;
      quat[4 * v + 3] = 0.0f;
      for (int k = 0; k < 3; ++k) logscale[3 * v + k] = to_f16(default_ls);
    }
  }
  if (scene_extent)
    *scene_extent = p.n ? (float)std::max(mx[0] - mn[0], std::max(mx[1] - mn[1], mx[2] - mn[2])) : 0.0f;
  return SV_OK;
}

sv_status sv_ply_save(const char *path, size_t n, const float *pos, const float *logscale, const float *quat,
                      const float *logit, const float *sh, uint32_t sh_deg) {
  if (!path || !pos || !logscale || !quat || !logit || !sh) return ingest_fail(SV_E_ARG, "null argument");
  if (sh_deg > 3) return ingest_fail(SV_E_ARG, "sh_deg must be <= 3");
  FILE *f = std::fopen(path, "wb");
  if (!f) return ingest_fail(SV_E_ARG, "cannot open %s", path);
  const int per = (int)((sh_deg + 1) * (sh_deg + 1)) - 1;
  std::string h = "ply\nformat binary_little_endian 1.0\nelement vertex " + std::to_string(n) + "\n";
  for (const char *nm : {"x", "y", "z", "nx", "ny", "nz", "f_dc_0", "f_dc_1", "f_dc_2"})
    h += std::string("property float ") + nm + "\n";
  for (int i = 0; i < 3 * per; ++i) h += "property float f_rest_" + std::to_string(i) + "\n";
  for (const char *nm : {"opacity", "scale_0", "scale_1", "scale_2", "rot_0", "rot_1", "rot_2", "rot_3"})
    h += std::string("property float ") + nm + "\n";
  h += "end_header\n";
  write_u8(f, h);
  std::vector<float> row;
  bool ok = true;
  for (size_t v = 0; v < n && ok; ++v) {
    row.clear();
    for (int k = 0; k < 3; ++k) row.push_back(pos[3 * v + k]);
    row.insert(row.end(), {0.0f, 0.0f, 0.0f});
    for (int c = 0; c < 3; ++c) row.push_back(sh[48 * v + c]);
    for (int c = 0; c < 3; ++c)
      for (int i = 0; i < per; ++i) row.push_back(sh[48 * v + 3 * (i + 1) + c]);
    row.push_back(logit[v]);
    for (int k = 0; k < 3; ++k) row.push_back(logscale[3 * v + k]);
    row.push_back(quat[4 * v + 3]);
    for (int k = 0; k < 3; ++k) row.push_back(quat[4 * v + k]);
    ok = std::fwrite(row.data(), sizeof(float), row.size(), f) == row.size();
  }
  if (std::fclose(f) != 0) ok = false;
  return ok ? SV_OK : ingest_fail(SV_E_ARG, "writing %s failed", path);
}

sv_status sv_colmap_info(const char *sparse_dir, size_t *n_images, size_t *n_points) {
  if (!sparse_dir) return ingest_fail(SV_E_ARG, "null path");
  const std::string d = sparse_dir;
  std::vector<unsigned char> ib, pb;
  if (!read_file((d + "/images.bin").c_str(), ib) || !read_file((d + "/points3D.bin").c_str(), pb))
    return ingest_fail(SV_E_ARG, "cannot read %s/images.bin or points3D.bin", sparse_dir);
  Reader ri{ib}, rp{pb};
  const uint64_t ni = ri.get<uint64_t>(), np = rp.get<uint64_t>();
  if (!ri.ok || !rp.ok) return ingest_fail(SV_E_ARG, "COLMAP files are truncated");
  if (n_images) *n_images = (size_t)ni;
  if (n_points) *n_points = (size_t)np;
  return SV_OK;
}

sv_status sv_colmap_load(const char *sparse_dir, size_t n_images, sv_camera *cams, char *names,
                         size_t name_len, size_t n_points, float *xyz, uint8_t *rgb) {
  if (!sparse_dir || !cams || !xyz || !rgb) return ingest_fail(SV_E_ARG, "null argument");
  const std::string d = sparse_dir;
  std::vector<unsigned char> cb, ib, pb;
  if (!read_file((d + "/cameras.bin").c_str(), cb) || !read_file((d + "/images.bin").c_str(), ib) ||
      !read_file((d + "/points3D.bin").c_str(), pb))
    return ingest_fail(SV_E_ARG, "cannot read the COLMAP files in %s", sparse_dir);
  // parseColmapCamerasBin (colmap-dataset.ts:129-152)
  Reader rc{cb};
  const uint64_t ncam = rc.get<uint64_t>();
  std::vector<ColmapCam> cmap;
  for (uint64_t i = 0; i < ncam && rc.ok; ++i) {
    ColmapCam c{};
    c.id = rc.get<uint32_t>();
    c.model = rc.get<int32_t>();
    c.w = rc.get<uint64_t>();
    c.h = rc.get<uint64_t>();
    const int np = model_params(c.model);
    if (np < 0) return ingest_fail(SV_E_ARG, "unknown COLMAP camera model %d", c.model);
    std::vector<double> prm(np);
    for (int k = 0; k < np; ++k) prm[k] = rc.get<double>();
    if (c.model == 0) {  // SIMPLE_PINHOLE f, cx, cy
      c.fx = c.fy = prm[0];
      c.cx = prm[1];
      c.cy = prm[2];
    } else if (c.model == 1) {  // PINHOLE fx, fy, cx, cy
      c.fx = prm[0];
      c.fy = prm[1];
      c.cx = prm[2];
      c.cy = prm[3];
    } else {
      return ingest_fail(SV_E_ARG, "COLMAP camera model %d is not a pinhole model", c.model);
    }
    cmap.push_back(c);
  }
  if (!rc.ok) return ingest_fail(SV_E_ARG, "cameras.bin is truncated");
  // parseColmapImagesBin (:154-184) and the camera construction (:258-300)
  Reader ri{ib};
  const uint64_t ni = ri.get<uint64_t>();
  if (ni != n_images) return ingest_fail(SV_E_ARG, "expected %zu images, the model has %llu", n_images,
                                         (unsigned long long)ni);
  for (uint64_t i = 0; i < ni && ri.ok; ++i) {
    ri.get<uint32_t>();  // image id
    double q[4], t[3];
    for (double &x : q) x = ri.get<double>();
    for (double &x : t) x = ri.get<double>();
    const uint32_t cam_id = ri.get<uint32_t>();
    std::string name;
    while (ri.ok) {
      const char ch = (char)ri.get<uint8_t>();
      if (ch == 0) break;
      name.push_back(ch);
    }
    const uint64_t n2 = ri.get<uint64_t>();
    ri.skip(n2, 24);
    const ColmapCam *cam = nullptr;
    for (auto &c : cmap)
      if (c.id == cam_id) cam = &c;
    if (!cam) return ingest_fail(SV_E_ARG, "image %llu refers to unknown camera %u", (unsigned long long)i, cam_id);
    // qvecToRotMat (:206-221): R = world-to-camera rotation (row-major)
    const double qw = q[0], qx = q[1], qy = q[2], qz = q[3];
    const double R[9] = {1 - 2 * (qy * qy + qz * qz), 2 * (qx * qy - qw * qz), 2 * (qx * qz + qw * qy),
                         2 * (qx * qy + qw * qz), 1 - 2 * (qx * qx + qz * qz), 2 * (qy * qz - qw * qx),
                         2 * (qx * qz - qw * qy), 2 * (qy * qz + qw * qx), 1 - 2 * (qx * qx + qy * qy)};
    sv_camera &o = cams[i];
    std::memset(&o, 0, sizeof(o));
    // view = [R | t] (the inverse of the reference's c2w), column-major
    for (int r = 0; r < 3; ++r) {
      for (int c = 0; c < 3; ++c) o.view[c * 4 + r] = (float)R[3 * r + c];
      o.view[12 + r] = (float)t[r];
    }
    o.view[15] = 1.0f;
    // view_inv = c2w: rotation R^T, centre -R^T t
    for (int r = 0; r < 3; ++r) {
      for (int c = 0; c < 3; ++c) o.view_inv[c * 4 + r] = (float)R[3 * c + r];
      o.view_inv[12 + r] = (float)(-(R[r] * t[0] + R[3 + r] * t[1] + R[6 + r] * t[2]));
    }
    o.view_inv[15] = 1.0f;
    pinhole_proj(cam->fx, cam->fy, cam->cx, cam->cy, (double)cam->w, (double)cam->h, o.proj);
    o.viewport[0] = (float)cam->w;
    o.viewport[1] = (float)cam->h;
    o.focal[0] = (float)cam->fx;
    o.focal[1] = (float)cam->fy;
    if (names && name_len) {
      std::memset(names + i * name_len, 0, name_len);
      std::strncpy(names + i * name_len, name.c_str(), name_len - 1);
    }
  }
  if (!ri.ok) return ingest_fail(SV_E_ARG, "images.bin is truncated");
  // parseColmapPoints3DBin (:186-204)
  Reader rp{pb};
  const uint64_t np = rp.get<uint64_t>();
  if (np != n_points) return ingest_fail(SV_E_ARG, "expected %zu points, the model has %llu", n_points,
                                         (unsigned long long)np);
  for (uint64_t i = 0; i < np && rp.ok; ++i) {
    rp.get<uint64_t>();
    for (int k = 0; k < 3; ++k) xyz[3 * i + k] = (float)rp.get<double>();
    for (int k = 0; k < 3; ++k) rgb[3 * i + k] = rp.get<uint8_t>();
    rp.get<double>();
    const uint64_t track = rp.get<uint64_t>();
    rp.skip(track, 8);
  }
  if (!rp.ok) return ingest_fail(SV_E_ARG, "points3D.bin is truncated");
  return SV_OK;
}

// createPointCloudFromXYZRGB (pointcloud-loader.ts:29-100): one Gaussian
// per point with its colour as the SH DC term and the reference defaults.
sv_status sv_points_to_gaussians(size_t n, const float *xyz, const uint8_t *rgb, float *pos, float *logscale,
                                 float *quat, float *logit, float *sh, float *scene_extent) {
  if (!xyz || !rgb || !pos || !logscale || !quat || !logit || !sh) return ingest_fail(SV_E_ARG, "null argument");
  const double SH_C0 = 0.28209479177387814;
  double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (size_t v = 0; v < n; ++v) {
    for (int k = 0; k < 3; ++k) {
      pos[3 * v + k] = to_f16(xyz[3 * v + k]);
      mn[k] = std::min(mn[k], (double)xyz[3 * v + k]);
      mx[k] = std::max(mx[k], (double)xyz[3 * v + k]);
    }
    for (int j = 0; j < 48; ++j) sh[48 * v + j] = 0.0f;
    for (int c = 0; c < 3; ++c) sh[48 * v + c] = to_f16((float)(((double)rgb[3 * v + c] / 255.0 - 0.5) / SH_C0));
    logit[v] = to_f16(-2.197f);
    quat[4 * v + 0] = 0.0f;
    quat[4 * v + 1] = 0.0f;
    quat[4 * v + 2] = 1.0f;
    quat[4 * v + 3] = 0.0f;
    for (int k = 0; k < 3; ++k) logscale[3 * v + k] = to_f16((float)std::log(0.01));
  }
  if (scene_extent)
    *scene_extent = n ? (float)std::max(mx[0] - mn[0], std::max(mx[1] - mn[1], mx[2] - mn[2])) : 0.0f;
  return SV_OK;
}

}  // extern "C"
