// Sanitizer check (AddressSanitizer + UndefinedBehaviorSanitizer) of the
// CPU oracle and the library's host-only sources: runs every oracle routine
// on small seeded scenes (both backward methods with both flag sets, the
// SSIM loss, the float64 forward, Adam, KNN, prune / densify, scoring, both
// sorts, the exp / f16 helpers over edge values) and the PLY / COLMAP
// readers on valid, truncated and oversized-count files. Exits 0 and prints
// "sanitize check ok" when no sanitizer fires and every result is sane.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include <unistd.h>

#include "../../include/sv_api.h"
#include "../../oracle/sv_oracle.h"

namespace {

int g_fail = 0;
#define CHECK(cond)                                                        \
  do {                                                                     \
    if (!(cond)) {                                                         \
      std::fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #cond); \
      ++g_fail;                                                            \
    }                                                                      \
  } while (0)

struct Scene {
  size_t n;
  std::vector<float> pos, ls, q, lg, sh;
  orc_params p() const { return orc_params{pos.data(), ls.data(), q.data(), lg.data(), sh.data()}; }
};

Scene scene(uint64_t seed, size_t n) {
  Scene s;
  s.n = n;
  s.pos.resize(3 * n);
  s.ls.resize(3 * n);
  s.q.resize(4 * n);
  s.lg.resize(n);
  s.sh.resize(48 * n);
  CHECK(sv_scene_gaussians(seed, n, 1.0f, s.pos.data(), s.ls.data(), s.q.data(), s.lg.data(), s.sh.data()) == SV_OK);
  return s;
}

orc_camera camera(float x, float y, float z, uint32_t w, uint32_t h) {
  const float eye[3] = {x, y, z}, tgt[3] = {0, 0, 0};
  sv_camera c;
  CHECK(sv_scene_camera(eye, tgt, w, h, 60.0f, &c) == SV_OK);
  orc_camera o;
  static_assert(sizeof(o) == sizeof(c), "camera layouts");
  std::memcpy(&o, &c, sizeof(o));
  return o;
}

orc_view view(uint32_t w, uint32_t h) {
  orc_view v{};
  v.width = w;
  v.height = h;
  v.tile_size = 16;
  v.tiles_x = (w + 15) / 16;
  v.tiles_y = (h + 15) / 16;
  v.sh_deg = 3;
  v.near_plane = 0.2f;
  v.far_plane = 1e4f;
  v.radius_clip = 512.0f;
  v.eps2d = 0.3f;
  return v;
}

struct Step {
  std::vector<orc_proj> proj;
  std::vector<uint32_t> keys, vals, offs, last;
  std::vector<float> color, alpha, residual;
  std::vector<int32_t> fixed;
  orc_step_out out;
};

void alloc_step(Step &s, size_t n, const orc_view &v, size_t cap) {
  const size_t P = (size_t)v.width * v.height;
  s.proj.resize(n);
  s.keys.resize(cap);
  s.vals.resize(cap);
  s.offs.resize((size_t)v.tiles_x * v.tiles_y + 1);
  s.last.resize(P);
  s.color.resize(3 * P);
  s.alpha.resize(P);
  s.residual.resize(3 * P);
  s.fixed.resize(cap * 54);
  s.out = orc_step_out{s.proj.data(), s.keys.data(), s.vals.data(), cap, s.offs.data(), s.color.data(),
                       s.alpha.data(), s.last.data(), s.residual.data(), s.fixed.data()};
}

struct Grads {
  std::vector<float> pos, ls, q, lg, sh, geom;
  explicit Grads(size_t n) : pos(3 * n), ls(3 * n), q(4 * n), lg(n), sh(48 * n), geom(5 * n) {}
  orc_grads g() { return orc_grads{pos.data(), ls.data(), q.data(), lg.data(), sh.data(), geom.data()}; }
};

bool all_finite(const std::vector<float> &v) {
  for (float x : v)
    if (!std::isfinite(x)) return false;
  return true;
}

void check_step() {
  const size_t n = 3000;
  Scene s = scene(0, n), t = scene(0, n);
  CHECK(sv_scene_jitter(1, n, 0.02f, t.pos.data()) == SV_OK);
  const orc_camera cam = camera(0.3f, 0.2f, 3.0f, 96, 72);
  const orc_view v = view(96, 72);
  const orc_params tp = t.p(), sp = s.p();
  std::vector<orc_proj> pr(n);
  orc_project(n, &tp, &cam, &v, pr.data());
  const uint64_t m = orc_count_isects(n, pr.data());
  CHECK(m > 0);
  Step tgt;
  alloc_step(tgt, n, v, m);
  CHECK(orc_render_and_grad(n, &tp, &cam, &v, nullptr, 0, 0, 0, 0, 0, &tgt.out, nullptr) == (int64_t)m);
  // both methods, both flag sets, with and without the SSIM term
  for (int wf = 0; wf < 2; ++wf)
    for (uint32_t flags : {0u, 3u})
      for (float ssim : {0.0f, 0.2f}) {
        if (!wf && flags) continue;
        Step st;
        alloc_step(st, n, v, 4 * m);
        Grads g(n);
        orc_grads gg = g.g();
        const int64_t r = orc_render_and_grad(n, &sp, &cam, &v, tgt.color.data(), 0.8f, 0.2f, ssim, wf, flags,
                                              &st.out, &gg);
        CHECK(r > 0);
        CHECK(all_finite(g.sh) && all_finite(g.pos) && all_finite(g.lg));
        if (wf && flags == 0 && ssim == 0.0f) {
          // the fixed-point sums reduce to the same gradients again
          Grads g2(n);
          orc_grads gg2 = g2.g();
          orc_backward_reduce(n, st.fixed.data(), st.vals.data(), (size_t)r, &gg2);
          CHECK(std::memcmp(g2.lg.data(), g.lg.data(), n * sizeof(float)) == 0);
        }
      }
  // a large loss weight: fixed-point sums wrap (defined unsigned arithmetic)
  {
    Step st;
    alloc_step(st, n, v, 4 * m);
    Grads g(n);
    orc_grads gg = g.g();
    CHECK(orc_render_and_grad(n, &sp, &cam, &v, tgt.color.data(), 3e4f, 0.0f, 0.0f, 1, 0, &st.out, &gg) > 0);
  }
  // float64 forward
  std::vector<double> dp(s.pos.begin(), s.pos.end()), dl(s.ls.begin(), s.ls.end()), dq(s.q.begin(), s.q.end()),
      dg(s.lg.begin(), s.lg.end()), dsh(s.sh.begin(), s.sh.end());
  std::vector<double> col(3 * 96 * 72);
  const double loss = orc_forward_loss_f64(n, dp.data(), dl.data(), dq.data(), dg.data(), dsh.data(), &cam, &v,
                                           tgt.color.data(), 0.8, 0.2, col.data());
  CHECK(std::isfinite(loss) && loss > 0.0);
  // Adam
  std::vector<float> p(s.sh), m1(s.sh.size()), m2(s.sh.size()), gr(s.sh.size(), 0.01f);
  float bc1, bc2;
  orc_adam_bias(3, 0.9f, 0.999f, &bc1, &bc2);
  orc_adam(p.size(), p.data(), m1.data(), m2.data(), gr.data(), 1e-3f, 0.9f, 0.999f, 1e-15f, bc1, bc2);
  CHECK(all_finite(p));
  // statistics
  std::vector<float> st4(4 * n), acc4(4 * n), geom(5 * n, 0.5f);
  orc_accumulate_importance(n, tgt.proj.data(), 0.98f, st4.data());
  orc_accumulate_grads(n, geom.data(), acc4.data());
  // scoring, tile L1, densify decision with scores
  std::vector<float> l1((size_t)v.tiles_x * v.tiles_y);
  Step cur;
  alloc_step(cur, n, v, 4 * m);
  const int64_t r = orc_render_and_grad(n, &sp, &cam, &v, nullptr, 0, 0, 0, 0, 0, &cur.out, nullptr);
  orc_tile_l1(&v, cur.color.data(), tgt.color.data(), l1.data());
  std::vector<uint32_t> counts(n);
  double mean = 0.0;
  CHECK(orc_metric_view(&v, n, cur.proj.data(), cur.offs.data(), cur.vals.data(), cur.last.data(), cur.color.data(),
                        tgt.color.data(), counts.data(), &mean) == 1);
  const orc_proj *projs[1] = {cur.proj.data()};
  const float *l1s[1] = {l1.data()};
  const float thr = 0.1f;
  std::vector<uint8_t> act(n);
  std::vector<float> score(n);
  orc_densify_decide(n, 1, projs, &v, l1s, &thr, 8.0f, act.data(), score.data());
  // sorts agree
  std::vector<uint32_t> k1(cur.keys.begin(), cur.keys.begin() + r), v1(cur.vals.begin(), cur.vals.begin() + r);
  std::vector<uint32_t> k2 = k1, v2 = v1;
  for (size_t i = 0; i + 1 < k1.size(); i += 2) std::swap(k1[i], k1[i + 1]), std::swap(k2[i], k2[i + 1]);
  orc_sort_pairs(k1.size(), k1.data(), v1.data());
  orc_sort_pairs_radix(k2.size(), k2.data(), v2.data());
  CHECK(k1 == k2 && v1 == v2);
  // prune / densify / compaction on rows of 59
  std::vector<float> rows(59 * n), mm1(59 * n, 0.1f), mm2(59 * n, 0.2f);
  for (size_t i = 0; i < n; ++i) {
    std::memcpy(&rows[59 * i], &s.pos[3 * i], 12);
    std::memcpy(&rows[59 * i + 3], &s.ls[3 * i], 12);
    std::memcpy(&rows[59 * i + 6], &s.q[4 * i], 16);
    rows[59 * i + 10] = s.lg[i];
    std::memcpy(&rows[59 * i + 11], &s.sh[48 * i], 192);
  }
  for (size_t i = 0; i < n; ++i) act[i] = (uint8_t)(i % 3);
  size_t n_out = n;
  for (size_t i = 0; i < n; ++i) n_out += act[i] != 0;
  std::vector<float> o(59 * n_out), o1(59 * n_out), o2(59 * n_out);
  CHECK(orc_densify(n, rows.data(), mm1.data(), mm2.data(), act.data(), 7, o.data(), o1.data(), o2.data()) == n_out);
  CHECK(all_finite(o));
  std::vector<uint8_t> keep(n);
  orc_prune_keep(n, s.lg.data(), 0.3f, keep.data());
  std::vector<float> kept(59 * n);
  size_t nk = 0;
  for (uint8_t x : keep) nk += x != 0;
  CHECK(orc_compact_rows(n, 59, rows.data(), keep.data(), kept.data()) == nk);
  // KNN on points with duplicates
  std::vector<float> pts(3 * 500), dens(500), kls(500);
  for (size_t i = 0; i < 500; ++i)
    for (int k = 0; k < 3; ++k) pts[3 * i + k] = s.pos[3 * (i % 250) + k];
  orc_knn_init(500, pts.data(), 2.0f, dens.data(), kls.data());
  CHECK(all_finite(kls));
}

void check_helpers() {
  std::vector<float> x, y;
  for (int i = 0; i <= 200000; ++i) x.push_back(-80.0f + 80.5f * (float)i / 200000.0f);
  y.resize(x.size());
  orc_exp_alpha_v(x.size(), x.data(), y.data());
  CHECK(all_finite(y));
  const float sp[] = {0.0f, -0.0f, 100.0f, -200.0f, INFINITY, -INFINITY, NAN, 88.7f, -103.0f};
  std::vector<float> e(9);
  orc_expf_v(9, sp, e.data());
  std::vector<float> bits(1 << 16), r16(1 << 16);
  for (uint32_t i = 0; i < (1u << 16); ++i) {
    const uint32_t u = i * 65537u + 0x3000u;
    std::memcpy(&bits[i], &u, 4);
  }
  orc_f16_round_v(bits.size(), bits.data(), r16.data());
  const float dirs[6] = {1, 0, 0, 0, 0.6f, 0.8f};
  std::vector<float> sh48(96, 0.1f), rgb(6);
  orc_sh_color_v(2, dirs, sh48.data(), 3, rgb.data());
  CHECK(all_finite(rgb));
}

std::string tmpdir() {
  char tmpl[] = "/tmp/sv_sanitize_XXXXXX";
  const char *d = mkdtemp(tmpl);
  CHECK(d != nullptr);
  return d ? d : "/tmp";
}

void write_bytes(const std::string &path, const std::vector<unsigned char> &b) {
  FILE *f = std::fopen(path.c_str(), "wb");
  CHECK(f != nullptr);
  if (!f) return;
  std::fwrite(b.data(), 1, b.size(), f);
  std::fclose(f);
}

template <typename T>
void put(std::vector<unsigned char> &b, T v) {
  const unsigned char *p = (const unsigned char *)&v;
  b.insert(b.end(), p, p + sizeof(T));
}

void check_ingest() {
  const std::string d = tmpdir();
  const size_t n = 200;
  Scene s = scene(3, n);
  const std::string ply = d + "/g.ply";
  CHECK(sv_ply_save(ply.c_str(), n, s.pos.data(), s.ls.data(), s.q.data(), s.lg.data(), s.sh.data(), 3) == SV_OK);
  size_t nn = 0;
  uint32_t deg = 0;
  CHECK(sv_ply_info(ply.c_str(), &nn, &deg) == SV_OK && nn == n && deg == 3);
  Scene l = scene(4, n);
  float ext = 0.0f;
  CHECK(sv_ply_load(ply.c_str(), n, l.pos.data(), l.ls.data(), l.q.data(), l.lg.data(), l.sh.data(), &ext) == SV_OK);
  // truncated file and a vertex count whose byte size overflows size_t
  std::vector<unsigned char> full;
  {
    FILE *f = std::fopen(ply.c_str(), "rb");
    CHECK(f != nullptr);
    if (f) {
      int ch;
      while ((ch = std::fgetc(f)) != EOF) full.push_back((unsigned char)ch);
      std::fclose(f);
    }
  }
  std::vector<unsigned char> cut(full.begin(), full.begin() + full.size() / 2);
  write_bytes(d + "/cut.ply", cut);
  CHECK(sv_ply_info((d + "/cut.ply").c_str(), &nn, &deg) != SV_OK);
  std::string head(full.begin(), full.end());
  const size_t at = head.find("element vertex ");
  CHECK(at != std::string::npos);
  if (at != std::string::npos) {
    const size_t eol = head.find('\n', at);
    std::string big = head.substr(0, at) + "element vertex 18446744073709551000" + head.substr(eol);
    write_bytes(d + "/big.ply", std::vector<unsigned char>(big.begin(), big.end()));
    CHECK(sv_ply_info((d + "/big.ply").c_str(), &nn, &deg) != SV_OK);
  }
  // COLMAP: one PINHOLE camera, one image, three points
  std::vector<unsigned char> cb, ib, pb;
  put<uint64_t>(cb, 1);
  put<uint32_t>(cb, 1);
  put<int32_t>(cb, 1);
  put<uint64_t>(cb, 64);
  put<uint64_t>(cb, 48);
  for (double v : {50.0, 50.0, 32.0, 24.0}) put<double>(cb, v);
  put<uint64_t>(ib, 1);
  put<uint32_t>(ib, 1);
  for (double v : {1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 3.0}) put<double>(ib, v);
  put<uint32_t>(ib, 1);
  for (char ch : std::string("img0.png")) put<char>(ib, ch);
  put<char>(ib, 0);
  put<uint64_t>(ib, 2);
  for (int k = 0; k < 2; ++k) {
    put<double>(ib, 1.0);
    put<double>(ib, 2.0);
    put<int64_t>(ib, -1);
  }
  put<uint64_t>(pb, 3);
  for (int i = 0; i < 3; ++i) {
    put<uint64_t>(pb, i);
    for (double v : {0.1 * i, 0.2, 0.3}) put<double>(pb, v);
    for (int k = 0; k < 3; ++k) put<uint8_t>(pb, (uint8_t)(40 * i + k));
    put<double>(pb, 0.5);
    put<uint64_t>(pb, 1);
    put<uint32_t>(pb, 1);
    put<uint32_t>(pb, 0);
  }
  write_bytes(d + "/cameras.bin", cb);
  write_bytes(d + "/images.bin", ib);
  write_bytes(d + "/points3D.bin", pb);
  size_t ni = 0, np = 0;
  CHECK(sv_colmap_info(d.c_str(), &ni, &np) == SV_OK && ni == 1 && np == 3);
  sv_camera cams[1];
  char names[64];
  float xyz[9];
  uint8_t rgb[9];
  CHECK(sv_colmap_load(d.c_str(), 1, cams, names, 64, 3, xyz, rgb) == SV_OK);
  // a track length past the end of points3D.bin and an observation count
  // whose byte size overflows: errors, not reads past the buffers
  std::vector<unsigned char> pb2(pb);
  const uint64_t huge = 0x2000000000000001ull;
  std::memcpy(&pb2[8 + 8 + 24 + 3 + 8], &huge, 8);
  write_bytes(d + "/points3D.bin", pb2);
  CHECK(sv_colmap_load(d.c_str(), 1, cams, names, 64, 3, xyz, rgb) != SV_OK);
  write_bytes(d + "/points3D.bin", pb);
  std::vector<unsigned char> ib2(ib);
  const size_t obs = 8 + 4 + 7 * 8 + 4 + 9;
  std::memcpy(&ib2[obs], &huge, 8);
  write_bytes(d + "/images.bin", ib2);
  CHECK(sv_colmap_load(d.c_str(), 1, cams, names, 64, 3, xyz, rgb) != SV_OK);
  // points to Gaussians
  float extent = 0.0f;
  Scene pg = scene(5, 3);
  CHECK(sv_points_to_gaussians(3, xyz, rgb, pg.pos.data(), pg.ls.data(), pg.q.data(), pg.lg.data(), pg.sh.data(),
                               &extent) == SV_OK);
  for (const char *f : {"g.ply", "cut.ply", "big.ply", "cameras.bin", "images.bin", "points3D.bin"})
    std::remove((d + "/" + f).c_str());
  rmdir(d.c_str());
}

}  // namespace

int main() {
  check_helpers();
  check_step();
  check_ingest();
  if (g_fail) {
    std::fprintf(stderr, "%d checks failed\n", g_fail);
    return 1;
  }
  std::printf("sanitize check ok\n");
  return 0;
}
