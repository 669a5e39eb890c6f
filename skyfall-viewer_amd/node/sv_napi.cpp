// sv_napi.cpp -- Node N-API addon over libsv.so (include/sv_api.h).
//
// The JS side (index.js) keeps the reference renderer's training interface
// (GaussianRenderer, src/renderers/gaussian-renderer.ts:30-58); this file is
// the thin binding: typed arrays in and out (copied), a context handle as an
// external, status codes turned into thrown Errors carrying sv_last_error.
// stepAsync runs training steps on a libuv worker and never touches JS from
// there (one context per addon call chain; contexts are not thread-safe).
#define NAPI_VERSION 4
#include <node_api.h>

#include <cmath>

#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/sv_api.h"

namespace {

#define CK(call)                                                   \
  do {                                                             \
    if ((call) != napi_ok) {                                       \
      napi_throw_error(env, nullptr, "N-API call failed: " #call); \
      return nullptr;                                              \
    }                                                              \
  } while (0)

napi_value throw_sv(napi_env env, sv_ctx *c, sv_status s, const char *what) {
  std::string m = std::string(what) + " failed (status " + std::to_string((int)s) + "): " + sv_last_error(c);
  napi_throw_error(env, nullptr, m.c_str());
  return nullptr;
}

#define SV(ctx, call, what)                                   \
  do {                                                        \
    sv_status _s = (call);                                    \
    if (_s != SV_OK) return throw_sv(env, (ctx), _s, (what)); \
  } while (0)

struct Args {
  size_t n = 8;
  napi_value v[8];
  napi_value self;
};

bool get_args(napi_env env, napi_callback_info info, Args &a, size_t need) {
  a.n = 8;
  if (napi_get_cb_info(env, info, &a.n, a.v, &a.self, nullptr) != napi_ok) return false;
  if (a.n < need) {
    napi_throw_type_error(env, nullptr, "missing arguments");
    return false;
  }
  return true;
}

// Contexts are externals; destroy() clears the pointer held in a box so a
// second destroy or a later call fails cleanly.
struct Box {
  sv_ctx *c;
};
void finalize_box(napi_env, void *data, void *) {
  Box *b = (Box *)data;
  if (b) {
    if (b->c) sv_destroy(b->c);
    delete b;
  }
}

sv_ctx *ctx_of(napi_env env, napi_value v) {
  void *p = nullptr;
  if (napi_get_value_external(env, v, &p) != napi_ok || !p || !((Box *)p)->c) {
    napi_throw_type_error(env, nullptr, "not a live sv context");
    return nullptr;
  }
  return ((Box *)p)->c;
}

bool get_u32(napi_env env, napi_value v, uint32_t *o) { return napi_get_value_uint32(env, v, o) == napi_ok; }
bool get_i32(napi_env env, napi_value v, int32_t *o) { return napi_get_value_int32(env, v, o) == napi_ok; }
bool get_f64(napi_env env, napi_value v, double *o) { return napi_get_value_double(env, v, o) == napi_ok; }

// Typed array data pointer of the expected element type and minimum length.
template <typename T>
T *typed(napi_env env, napi_value v, napi_typedarray_type want, size_t min_len, size_t *len_out = nullptr) {
  bool is = false;
  if (napi_is_typedarray(env, v, &is) != napi_ok || !is) {
    napi_throw_type_error(env, nullptr, "expected a typed array");
    return nullptr;
  }
  napi_typedarray_type t;
  size_t len = 0, off = 0;
  void *data = nullptr;
  napi_value ab;
  if (napi_get_typedarray_info(env, v, &t, &len, &data, &ab, &off) != napi_ok || t != want) {
    napi_throw_type_error(env, nullptr, "typed array of the wrong element type");
    return nullptr;
  }
  if (len < min_len) {
    napi_throw_range_error(env, nullptr, "typed array too short");
    return nullptr;
  }
  if (len_out) *len_out = len;
  return (T *)data;
}

template <typename T>
napi_value new_typed(napi_env env, napi_typedarray_type t, size_t n, T **data) {
  napi_value ab, arr;
  void *p = nullptr;
  if (napi_create_arraybuffer(env, n * sizeof(T), &p, &ab) != napi_ok) return nullptr;
  if (napi_create_typedarray(env, t, n, ab, 0, &arr) != napi_ok) return nullptr;
  *data = (T *)p;
  return arr;
}

napi_value prop(napi_env env, napi_value obj, const char *name) {
  napi_value v = nullptr;
  napi_get_named_property(env, obj, name, &v);
  return v;
}

// ------------------------------------------------------------------ config
enum FieldType { F_U32, F_F32 };
struct Field {
  const char *name;
  size_t off;
  FieldType t;
};
#define FU(n) {#n, offsetof(sv_config, n), F_U32}
#define FF(n) {#n, offsetof(sv_config, n), F_F32}
const Field kFields[] = {FU(tile_size), FU(batch), FU(raster_px), FU(sh_deg), FF(near_plane), FF(far_plane),
                         FF(radius_clip), FF(eps2d), FF(w_l2), FF(w_l1), FF(w_ssim), FF(lr_pos), FF(lr_sh),
                         FF(lr_opacity), FF(lr_scale), FF(lr_rot), FF(beta1), FF(beta2), FF(eps),
                         FU(sh_update_interval), FU(pos_lr_schedule), FU(render_f16), FF(importance_decay),
                         FU(persistent), FU(use_graph), FU(keep_grads), FU(plain_wf_flags), FU(xcd_remap),
                         FU(binned_build), FU(geom_flags), FU(split_backward), FU(tile_order)};

napi_value config_to_js(napi_env env, const sv_config &k) {
  napi_value o;
  CK(napi_create_object(env, &o));
  for (const Field &f : kFields) {
    napi_value v;
    const char *p = (const char *)&k + f.off;
    if (f.t == F_U32) CK(napi_create_uint32(env, *(const uint32_t *)p, &v));
    else CK(napi_create_double(env, *(const float *)p, &v));
    CK(napi_set_named_property(env, o, f.name, v));
  }
  napi_value bg;
  CK(napi_create_array_with_length(env, 3, &bg));
  for (uint32_t i = 0; i < 3; ++i) {
    napi_value v;
    CK(napi_create_double(env, k.bg[i], &v));
    CK(napi_set_element(env, bg, i, v));
  }
  CK(napi_set_named_property(env, o, "bg", bg));
  return o;
}

// Unknown keys are an error, as a misspelt hyper-parameter would otherwise
// be ignored silently.
bool config_from_js(napi_env env, napi_value o, sv_config &k) {
  napi_value names;
  if (napi_get_property_names(env, o, &names) != napi_ok) return false;
  uint32_t n = 0;
  napi_get_array_length(env, names, &n);
  for (uint32_t i = 0; i < n; ++i) {
    napi_value key;
    napi_get_element(env, names, i, &key);
    char buf[64];
    size_t len = 0;
    napi_get_value_string_utf8(env, key, buf, sizeof(buf), &len);
    napi_value v = prop(env, o, buf);
    if (!strcmp(buf, "bg")) {
      for (uint32_t j = 0; j < 3; ++j) {
        napi_value e;
        double d;
        if (napi_get_element(env, v, j, &e) != napi_ok || !get_f64(env, e, &d)) {
          napi_throw_type_error(env, nullptr, "bg must be three numbers");
          return false;
        }
        k.bg[j] = (float)d;
      }
      continue;
    }
    const Field *f = nullptr;
    for (const Field &x : kFields)
      if (!strcmp(x.name, buf)) f = &x;
    if (!f) {
      napi_throw_type_error(env, nullptr, (std::string("unknown config field ") + buf).c_str());
      return false;
    }
    double d;
    if (!get_f64(env, v, &d)) {
      napi_throw_type_error(env, nullptr, (std::string("config field ") + buf + " must be a number").c_str());
      return false;
    }
    char *p = (char *)&k + f->off;
    if (f->t == F_U32) *(uint32_t *)p = (uint32_t)d;
    else *(float *)p = (float)d;
  }
  return true;
}

// --------------------------------------------------------------- functions
napi_value Version(napi_env env, napi_callback_info) {
  napi_value s;
  CK(napi_create_string_utf8(env, sv_version(), NAPI_AUTO_LENGTH, &s));
  return s;
}

napi_value Create(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 1)) return nullptr;
  double cap = 0;
  int32_t dev = 0;
  if (!get_f64(env, a.v[0], &cap) || cap < 1) {
    napi_throw_range_error(env, nullptr, "capacity must be a positive number");
    return nullptr;
  }
  if (a.n > 1) get_i32(env, a.v[1], &dev);
  sv_ctx *c = nullptr;
  sv_status s = sv_create(dev, (size_t)cap, &c);
  if (s != SV_OK) return throw_sv(env, nullptr, s, "sv_create");
  Box *b = new Box{c};
  napi_value ext;
  if (napi_create_external(env, b, finalize_box, nullptr, &ext) != napi_ok) {
    sv_destroy(c);
    delete b;
    napi_throw_error(env, nullptr, "cannot wrap the context");
    return nullptr;
  }
  return ext;
}

napi_value Destroy(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 1)) return nullptr;
  void *p = nullptr;
  if (napi_get_value_external(env, a.v[0], &p) == napi_ok && p) {
    Box *b = (Box *)p;
    if (b->c) sv_destroy(b->c);
    b->c = nullptr;
  }
  return nullptr;
}

napi_value GetConfig(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 1)) return nullptr;
  sv_ctx *c = ctx_of(env, a.v[0]);
  if (!c) return nullptr;
  sv_config k;
  SV(c, sv_get_config(c, &k), "sv_get_config");
  return config_to_js(env, k);
}

napi_value SetConfig(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 2)) return nullptr;
  sv_ctx *c = ctx_of(env, a.v[0]);
  if (!c) return nullptr;
  sv_config k;
  SV(c, sv_get_config(c, &k), "sv_get_config");
  if (!config_from_js(env, a.v[1], k)) return nullptr;
  SV(c, sv_set_config(c, &k), "sv_set_config");
  return nullptr;
}

const char *kParamNames[5] = {"pos", "logscale", "quat", "logit", "sh"};
const size_t kParamWidth[5] = {3, 3, 4, 1, 48};

bool params_from_js(napi_env env, napi_value o, size_t n, float *ptr[5]) {
  for (int k = 0; k < 5; ++k) {
    napi_value v = prop(env, o, kParamNames[k]);
    ptr[k] = typed<float>(env, v, napi_float32_array, n * kParamWidth[k]);
    if (!ptr[k]) return false;
  }
  return true;
}

napi_value params_to_js(napi_env env, size_t n, float *ptr[5]) {
  napi_value o;
  CK(napi_create_object(env, &o));
  for (int k = 0; k < 5; ++k) {
    napi_value arr = new_typed<float>(env, napi_float32_array, n * kParamWidth[k], &ptr[k]);
    if (!arr) return nullptr;
    CK(napi_set_named_property(env, o, kParamNames[k], arr));
  }
  return o;
}

napi_value SetGaussians(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 2)) return nullptr;
  sv_ctx *c = ctx_of(env, a.v[0]);
  if (!c) return nullptr;
  size_t n = 0;
  if (!typed<float>(env, prop(env, a.v[1], "logit"), napi_float32_array, 0, &n)) return nullptr;
  float *p[5];
  if (!params_from_js(env, a.v[1], n, p)) return nullptr;
  SV(c, sv_set_gaussians(c, n, p[0], p[1], p[2], p[3], p[4]), "sv_set_gaussians");
  return nullptr;
}

napi_value GetGaussians(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 1)) return nullptr;
  sv_ctx *c = ctx_of(env, a.v[0]);
  if (!c) return nullptr;
  float *p[5];
  napi_value o = params_to_js(env, sv_get_count(c), p);
  if (!o) return nullptr;
  SV(c, sv_get_gaussians(c, p[0], p[1], p[2], p[3], p[4]), "sv_get_gaussians");
  return o;
}

napi_value SetMoments(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 3)) return nullptr;
  sv_ctx *c = ctx_of(env, a.v[0]);
  if (!c) return nullptr;
  int32_t which;
  if (!get_i32(env, a.v[1], &which)) return nullptr;
  float *p[5];
  if (!params_from_js(env, a.v[2], sv_get_count(c), p)) return nullptr;
  SV(c, sv_set_moments(c, which, p[0], p[1], p[2], p[3], p[4]), "sv_set_moments");
  return nullptr;
}

napi_value GetMoments(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 2)) return nullptr;
  sv_ctx *c = ctx_of(env, a.v[0]);
  if (!c) return nullptr;
  int32_t which;
  if (!get_i32(env, a.v[1], &which)) return nullptr;
  float *p[5];
  napi_value o = params_to_js(env, sv_get_count(c), p);
  if (!o) return nullptr;
  SV(c, sv_get_moments(c, which, p[0], p[1], p[2], p[3], p[4]), "sv_get_moments");
  return o;
}

napi_value GetCount(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 1)) return nullptr;
  sv_ctx *c = ctx_of(env, a.v[0]);
  if (!c) return nullptr;
  napi_value v;
  CK(napi_create_double(env, (double)sv_get_count(c), &v));
  return v;
}

bool camera_from_js(napi_env env, napi_value o, sv_camera *cam) {
  std::memset(cam, 0, sizeof(*cam));
  struct {
    const char *name;
    float *dst;
    size_t n;
    bool required;
  } f[] = {{"view", cam->view, 16, true},          {"view_inv", cam->view_inv, 16, false},
           {"proj", cam->proj, 16, true},          {"proj_inv", cam->proj_inv, 16, false},
           {"viewport", cam->viewport, 2, true},   {"focal", cam->focal, 2, true}};
  for (auto &x : f) {
    napi_value v = prop(env, o, x.name);
    napi_valuetype t;
    napi_typeof(env, v, &t);
    if (t == napi_undefined) {
      if (x.required) {
        napi_throw_type_error(env, nullptr, (std::string("camera needs ") + x.name).c_str());
        return false;
      }
      continue;
    }
    for (uint32_t i = 0; i < x.n; ++i) {
      napi_value e;
      double d;
      if (napi_get_element(env, v, i, &e) != napi_ok || !get_f64(env, e, &d)) {
        napi_throw_type_error(env, nullptr, (std::string("camera field ") + x.name + " is not numeric").c_str());
        return false;
      }
      x.dst[i] = (float)d;
    }
  }
  return true;
}

napi_value SetViews(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 2)) return nullptr;
  sv_ctx *c = ctx_of(env, a.v[0]);
  if (!c) return nullptr;
  uint32_t n = 0;
  CK(napi_get_array_length(env, a.v[1], &n));
  std::vector<sv_camera> cams(n);
  for (uint32_t i = 0; i < n; ++i) {
    napi_value e;
    CK(napi_get_element(env, a.v[1], i, &e));
    if (!camera_from_js(env, e, &cams[i])) return nullptr;
  }
  SV(c, sv_set_views(c, (int)n, cams.data()), "sv_set_views");
  return nullptr;
}

// f(ctx, view, Float32Array rgb) for target / residual
napi_value set_pixels(napi_env env, napi_callback_info info, sv_status (*fn)(sv_ctx *, int, const float *),
                      const char *what) {
  Args a;
  if (!get_args(env, info, a, 3)) return nullptr;
  sv_ctx *c = ctx_of(env, a.v[0]);
  if (!c) return nullptr;
  int32_t v;
  if (!get_i32(env, a.v[1], &v)) return nullptr;
  const float *p = typed<float>(env, a.v[2], napi_float32_array, 0);
  if (!p) return nullptr;
  SV(c, fn(c, v, p), what);
  return nullptr;
}
napi_value SetTarget(napi_env env, napi_callback_info info) { return set_pixels(env, info, sv_set_target, "sv_set_target"); }
napi_value SetResidual(napi_env env, napi_callback_info info) {
  return set_pixels(env, info, sv_set_residual, "sv_set_residual");
}

// f(ctx, view)
napi_value view_call(napi_env env, napi_callback_info info, sv_status (*fn)(sv_ctx *, int), const char *what) {
  Args a;
  if (!get_args(env, info, a, 2)) return nullptr;
  sv_ctx *c = ctx_of(env, a.v[0]);
  if (!c) return nullptr;
  int32_t v;
  if (!get_i32(env, a.v[1], &v)) return nullptr;
  SV(c, fn(c, v), what);
  return nullptr;
}
napi_value Forward(napi_env env, napi_callback_info info) { return view_call(env, info, sv_forward, "sv_forward"); }
napi_value Loss(napi_env env, napi_callback_info info) { return view_call(env, info, sv_loss, "sv_loss"); }
napi_value Backward(napi_env env, napi_callback_info info) { return view_call(env, info, sv_backward, "sv_backward"); }

napi_value BuildTiles(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 2)) return nullptr;
  sv_ctx *c = ctx_of(env, a.v[0]);
  if (!c) return nullptr;
  int32_t v;
  if (!get_i32(env, a.v[1], &v)) return nullptr;
  uint64_t n = 0;
  SV(c, sv_build_tiles(c, v, &n), "sv_build_tiles");
  napi_value r;
  CK(napi_create_double(env, (double)n, &r));
  return r;
}

napi_value BackwardFixed(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 3)) return nullptr;
  sv_ctx *c = ctx_of(env, a.v[0]);
  if (!c) return nullptr;
  int32_t v;
  uint32_t flags;
  if (!get_i32(env, a.v[1], &v) || !get_u32(env, a.v[2], &flags)) return nullptr;
  SV(c, sv_backward_fixed(c, v, flags), "sv_backward_fixed");
  return nullptr;
}

napi_value ZeroGrad(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 1)) return nullptr;
  sv_ctx *c = ctx_of(env, a.v[0]);
  if (!c) return nullptr;
  SV(c, sv_zero_grad(c), "sv_zero_grad");
  return nullptr;
}

napi_value Adam(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 1)) return nullptr;
  sv_ctx *c = ctx_of(env, a.v[0]);
  if (!c) return nullptr;
  SV(c, sv_adam(c), "sv_adam");
  return nullptr;
}

bool views_from_js(napi_env env, napi_value v, std::vector<int> &out) {
  uint32_t n = 0;
  if (napi_get_array_length(env, v, &n) != napi_ok) {
    bool is = false;
    napi_is_typedarray(env, v, &is);
    if (!is) {
      napi_throw_type_error(env, nullptr, "views must be an array");
      return false;
    }
    size_t len = 0;
    const int32_t *p = typed<int32_t>(env, v, napi_int32_array, 0, &len);
    if (!p) return false;
    out.assign(p, p + len);
    return true;
  }
  out.resize(n);
  for (uint32_t i = 0; i < n; ++i) {
    napi_value e;
    if (napi_get_element(env, v, i, &e) != napi_ok || !get_i32(env, e, &out[i])) {
      napi_throw_type_error(env, nullptr, "views must be integers");
      return false;
    }
  }
  return true;
}

napi_value stats_to_js(napi_env env, const sv_step_stats &s) {
  napi_value o, v;
  CK(napi_create_object(env, &o));
  CK(napi_create_double(env, (double)s.n_isect, &v));
  CK(napi_set_named_property(env, o, "nIsect", v));
  CK(napi_create_uint32(env, s.n_views, &v));
  CK(napi_set_named_property(env, o, "nViews", v));
  CK(napi_create_uint32(env, s.sort_passes, &v));
  CK(napi_set_named_property(env, o, "sortPasses", v));
  CK(napi_create_uint32(env, s.key_bits, &v));
  CK(napi_set_named_property(env, o, "keyBits", v));
  CK(napi_create_double(env, s.l1_sum, &v));
  CK(napi_set_named_property(env, o, "l1Sum", v));
  return o;
}

napi_value Step(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 3)) return nullptr;
  sv_ctx *c = ctx_of(env, a.v[0]);
  if (!c) return nullptr;
  std::vector<int> vs;
  int32_t mode;
  if (!views_from_js(env, a.v[1], vs) || !get_i32(env, a.v[2], &mode)) return nullptr;
  sv_step_stats st{};
  SV(c, sv_step(c, vs.data(), (int)vs.size(), mode, &st), "sv_step");
  return stats_to_js(env, st);
}

// stepAsync(ctx, views, mode, count) -> Promise<stats of the last step>:
// `count` steps on a worker thread.
struct StepWork {
  napi_async_work work;
  napi_deferred deferred;
  sv_ctx *c;
  std::vector<int> vs;
  int mode;
  uint32_t count;
  sv_status status;
  std::string err;
  sv_step_stats st;
};

void step_execute(napi_env, void *data) {
  StepWork *w = (StepWork *)data;
  w->status = SV_OK;
  for (uint32_t i = 0; i < w->count && w->status == SV_OK; ++i)
    w->status = sv_step(w->c, w->vs.data(), (int)w->vs.size(), w->mode, &w->st);
  if (w->status != SV_OK) w->err = sv_last_error(w->c);
}

void step_complete(napi_env env, napi_status, void *data) {
  StepWork *w = (StepWork *)data;
  if (w->status == SV_OK) {
    napi_resolve_deferred(env, w->deferred, stats_to_js(env, w->st));
  } else {
    napi_value msg, err;
    std::string m = "sv_step failed (status " + std::to_string((int)w->status) + "): " + w->err;
    napi_create_string_utf8(env, m.c_str(), NAPI_AUTO_LENGTH, &msg);
    napi_create_error(env, nullptr, msg, &err);
    napi_reject_deferred(env, w->deferred, err);
  }
  napi_delete_async_work(env, w->work);
  delete w;
}

napi_value StepAsync(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 4)) return nullptr;
  sv_ctx *c = ctx_of(env, a.v[0]);
  if (!c) return nullptr;
  StepWork *w = new StepWork();
  w->c = c;
  int32_t mode;
  if (!views_from_js(env, a.v[1], w->vs) || !get_i32(env, a.v[2], &mode) || !get_u32(env, a.v[3], &w->count)) {
    delete w;
    return nullptr;
  }
  w->mode = mode;
  napi_value promise, name;
  CK(napi_create_promise(env, &w->deferred, &promise));
  CK(napi_create_string_utf8(env, "sv_step", NAPI_AUTO_LENGTH, &name));
  CK(napi_create_async_work(env, nullptr, name, step_execute, step_complete, w, &w->work));
  CK(napi_queue_async_work(env, w->work));
  return promise;
}

// readbacks --------------------------------------------------------------
// readImage / readResidual(ctx, view, width, height) -> Float32Array(h*w*3)
napi_value read_rgb(napi_env env, napi_callback_info info, sv_status (*fn)(sv_ctx *, int, float *),
                    const char *what, size_t comps) {
  Args a;
  if (!get_args(env, info, a, 4)) return nullptr;
  sv_ctx *c = ctx_of(env, a.v[0]);
  if (!c) return nullptr;
  int32_t v;
  uint32_t w, h;
  if (!get_i32(env, a.v[1], &v) || !get_u32(env, a.v[2], &w) || !get_u32(env, a.v[3], &h)) return nullptr;
  float *p = nullptr;
  napi_value arr = new_typed<float>(env, napi_float32_array, (size_t)w * h * comps, &p);
  if (!arr) return nullptr;
  SV(c, fn(c, v, p), what);
  return arr;
}
napi_value ReadImage(napi_env env, napi_callback_info info) { return read_rgb(env, info, sv_read_image, "sv_read_image", 3); }
napi_value ReadResidual(napi_env env, napi_callback_info info) {
  return read_rgb(env, info, sv_read_residual, "sv_read_residual", 3);
}
napi_value ReadAlpha(napi_env env, napi_callback_info info) { return read_rgb(env, info, sv_read_alpha, "sv_read_alpha", 1); }

napi_value ReadLastIds(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 4)) return nullptr;
  sv_ctx *c = ctx_of(env, a.v[0]);
  if (!c) return nullptr;
  int32_t v;
  uint32_t w, h;
  if (!get_i32(env, a.v[1], &v) || !get_u32(env, a.v[2], &w) || !get_u32(env, a.v[3], &h)) return nullptr;
  uint32_t *p = nullptr;
  napi_value arr = new_typed<uint32_t>(env, napi_uint32_array, (size_t)w * h, &p);
  if (!arr) return nullptr;
  SV(c, sv_read_last_ids(c, v, p), "sv_read_last_ids");
  return arr;
}

napi_value ReadGrads(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 1)) return nullptr;
  sv_ctx *c = ctx_of(env, a.v[0]);
  if (!c) return nullptr;
  const size_t n = sv_get_count(c);
  float *p[5];
  napi_value o = params_to_js(env, n, p);
  if (!o) return nullptr;
  float *geom = nullptr;
  napi_value g = new_typed<float>(env, napi_float32_array, 5 * n, &geom);
  if (!g) return nullptr;
  CK(napi_set_named_property(env, o, "geom", g));
  SV(c, sv_read_grads(c, p[0], p[1], p[2], p[3], p[4], geom), "sv_read_grads");
  return o;
}

napi_value ReadKeys(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 3)) return nullptr;
  sv_ctx *c = ctx_of(env, a.v[0]);
  if (!c) return nullptr;
  int32_t v;
  double n;
  if (!get_i32(env, a.v[1], &v) || !get_f64(env, a.v[2], &n)) return nullptr;
  uint32_t *k = nullptr, *val = nullptr;
  napi_value ka = new_typed<uint32_t>(env, napi_uint32_array, (size_t)n, &k);
  napi_value va = new_typed<uint32_t>(env, napi_uint32_array, (size_t)n, &val);
  if (!ka || !va) return nullptr;
  SV(c, sv_read_keys(c, v, k, val), "sv_read_keys");
  napi_value o;
  CK(napi_create_object(env, &o));
  CK(napi_set_named_property(env, o, "keys", ka));
  CK(napi_set_named_property(env, o, "values", va));
  return o;
}

// prune / densify ----------------------------------------------------------
napi_value PruneOpacity(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 2)) return nullptr;
  sv_ctx *c = ctx_of(env, a.v[0]);
  if (!c) return nullptr;
  double thr;
  if (!get_f64(env, a.v[1], &thr)) return nullptr;
  size_t removed = 0;
  SV(c, sv_prune_opacity(c, (float)thr, &removed), "sv_prune_opacity");
  napi_value r;
  CK(napi_create_double(env, (double)removed, &r));
  return r;
}

napi_value PruneMask(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 2)) return nullptr;
  sv_ctx *c = ctx_of(env, a.v[0]);
  if (!c) return nullptr;
  const uint8_t *m = typed<uint8_t>(env, a.v[1], napi_uint8_array, sv_get_count(c));
  if (!m) return nullptr;
  size_t removed = 0;
  SV(c, sv_prune_mask(c, m, &removed), "sv_prune_mask");
  napi_value r;
  CK(napi_create_double(env, (double)removed, &r));
  return r;
}

napi_value ResetOpacity(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 2)) return nullptr;
  sv_ctx *c = ctx_of(env, a.v[0]);
  if (!c) return nullptr;
  double mx;
  if (!get_f64(env, a.v[1], &mx)) return nullptr;
  size_t n = 0;
  SV(c, sv_reset_opacity(c, (float)mx, &n), "sv_reset_opacity");
  napi_value r;
  CK(napi_create_double(env, (double)n, &r));
  return r;
}

// knnInit(ctx, sceneExtent) -> Float32Array density (initializeKnn)
napi_value KnnInit(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 2)) return nullptr;
  sv_ctx *c = ctx_of(env, a.v[0]);
  if (!c) return nullptr;
  double e;
  if (!get_f64(env, a.v[1], &e)) return nullptr;
  const size_t n = sv_get_count(c);
  float *d = nullptr;
  napi_value arr = new_typed<float>(env, napi_float32_array, n, &d);
  if (!arr) return nullptr;
  SV(c, sv_knn_init(c, (float)e, n ? d : nullptr), "sv_knn_init");
  return arr;
}

// multiViewScores(ctx, views) -> {importance: Uint32Array, pruning: Float32Array,
// sampleCount} (generateMultiViewScores)
napi_value MultiViewScores(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 2)) return nullptr;
  sv_ctx *c = ctx_of(env, a.v[0]);
  if (!c) return nullptr;
  std::vector<int> vs;
  if (!views_from_js(env, a.v[1], vs)) return nullptr;
  const size_t n = sv_get_count(c);
  uint32_t *imp = nullptr;
  float *pru = nullptr;
  napi_value ai = new_typed<uint32_t>(env, napi_uint32_array, n, &imp);
  napi_value ap = new_typed<float>(env, napi_float32_array, n, &pru);
  if (!ai || !ap) return nullptr;
  int ns = 0;
  SV(c, sv_multiview_scores(c, vs.data(), (int)vs.size(), imp, pru, &ns), "sv_multiview_scores");
  napi_value o, k;
  CK(napi_create_object(env, &o));
  CK(napi_set_named_property(env, o, "importance", ai));
  CK(napi_set_named_property(env, o, "pruning", ap));
  CK(napi_create_int32(env, ns, &k));
  CK(napi_set_named_property(env, o, "sampleCount", k));
  return o;
}

// renderCamera(ctx, camera) -> {image: Float32Array(h*w*3), alpha: Float32Array(h*w), width, height}
napi_value RenderCamera(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 2)) return nullptr;
  sv_ctx *c = ctx_of(env, a.v[0]);
  if (!c) return nullptr;
  sv_camera cam;
  if (!camera_from_js(env, a.v[1], &cam)) return nullptr;
  const uint32_t w = (uint32_t)std::lround(cam.viewport[0]), h = (uint32_t)std::lround(cam.viewport[1]);
  float *rgb = nullptr, *al = nullptr;
  napi_value img = new_typed<float>(env, napi_float32_array, (size_t)w * h * 3, &rgb);
  napi_value alp = new_typed<float>(env, napi_float32_array, (size_t)w * h, &al);
  if (!img || !alp) return nullptr;
  SV(c, sv_render_camera(c, &cam, rgb, al), "sv_render_camera");
  napi_value o, vw, vh;
  CK(napi_create_object(env, &o));
  CK(napi_set_named_property(env, o, "image", img));
  CK(napi_set_named_property(env, o, "alpha", alp));
  CK(napi_create_uint32(env, w, &vw));
  CK(napi_create_uint32(env, h, &vh));
  CK(napi_set_named_property(env, o, "width", vw));
  CK(napi_set_named_property(env, o, "height", vh));
  return o;
}

napi_value Densify(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 3)) return nullptr;
  sv_ctx *c = ctx_of(env, a.v[0]);
  if (!c) return nullptr;
  const uint8_t *act = typed<uint8_t>(env, a.v[1], napi_uint8_array, sv_get_count(c));
  double seed;
  if (!act || !get_f64(env, a.v[2], &seed)) return nullptr;
  size_t added = 0;
  SV(c, sv_densify(c, act, (uint64_t)seed, &added), "sv_densify");
  napi_value r;
  CK(napi_create_double(env, (double)added, &r));
  return r;
}

napi_value DensifyDecide(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 5)) return nullptr;
  sv_ctx *c = ctx_of(env, a.v[0]);
  if (!c) return nullptr;
  std::vector<int> vs;
  double pct, split, max_new;
  if (!views_from_js(env, a.v[1], vs) || !get_f64(env, a.v[2], &pct) || !get_f64(env, a.v[3], &split) ||
      !get_f64(env, a.v[4], &max_new))
    return nullptr;
  uint8_t *act = nullptr;
  napi_value arr = new_typed<uint8_t>(env, napi_uint8_array, sv_get_count(c), &act);
  if (!arr) return nullptr;
  size_t nc = 0, ns = 0;
  SV(c, sv_densify_decide(c, vs.data(), (int)vs.size(), (float)pct, (float)split,
                          max_new > 0 ? (size_t)max_new : 0, act, &nc, &ns),
     "sv_densify_decide");
  napi_value o, v;
  CK(napi_create_object(env, &o));
  CK(napi_set_named_property(env, o, "actions", arr));
  CK(napi_create_double(env, (double)nc, &v));
  CK(napi_set_named_property(env, o, "clones", v));
  CK(napi_create_double(env, (double)ns, &v));
  CK(napi_set_named_property(env, o, "splits", v));
  return o;
}

// densifyDecideGrad(ctx, rule {gradThreshold, gradAbsThreshold, sizeThreshold,
// minVisibility, importanceThreshold, pixelScale}, importance Uint32Array,
// sampleCount, maxNew) -> {actions, clones, splits}
napi_value DensifyDecideGrad(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 5)) return nullptr;
  sv_ctx *c = ctx_of(env, a.v[0]);
  if (!c) return nullptr;
  sv_densify_rule r{};
  const struct {
    const char *name;
    double *dst;
  } keys[] = {{"gradThreshold", &r.grad_threshold},         {"gradAbsThreshold", &r.grad_abs_threshold},
              {"sizeThreshold", &r.size_threshold},         {"minVisibility", &r.min_visibility},
              {"importanceThreshold", &r.importance_threshold}, {"pixelScale", &r.pixel_scale}};
  for (const auto &k : keys) {
    napi_value v = prop(env, a.v[1], k.name);
    if (!v || !get_f64(env, v, k.dst)) {
      napi_throw_type_error(env, nullptr, "densify rule needs gradThreshold, gradAbsThreshold, sizeThreshold, "
                                          "minVisibility, importanceThreshold and pixelScale");
      return nullptr;
    }
  }
  const size_t n = sv_get_count(c);
  const uint32_t *imp = typed<uint32_t>(env, a.v[2], napi_uint32_array, n);
  if (!imp) return nullptr;
  uint32_t samples = 0;
  double max_new = 0;
  if (!get_u32(env, a.v[3], &samples) || !get_f64(env, a.v[4], &max_new)) return nullptr;
  uint8_t *act = nullptr;
  napi_value arr = new_typed<uint8_t>(env, napi_uint8_array, n, &act);
  if (!arr) return nullptr;
  size_t nc = 0, ns = 0;
  SV(c, sv_densify_decide_grad(c, &r, imp, samples, max_new > 0 ? (size_t)max_new : 0, act, &nc, &ns),
     "sv_densify_decide_grad");
  napi_value o, v;
  CK(napi_create_object(env, &o));
  CK(napi_set_named_property(env, o, "actions", arr));
  CK(napi_create_double(env, (double)nc, &v));
  CK(napi_set_named_property(env, o, "clones", v));
  CK(napi_create_double(env, (double)ns, &v));
  CK(napi_set_named_property(env, o, "splits", v));
  return o;
}

// timing -------------------------------------------------------------------
napi_value EnableTiming(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 2)) return nullptr;
  sv_ctx *c = ctx_of(env, a.v[0]);
  if (!c) return nullptr;
  bool on = false;
  napi_get_value_bool(env, a.v[1], &on);
  SV(c, sv_enable_timing(c, on ? 1 : 0), "sv_enable_timing");
  return nullptr;
}

napi_value Timings(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 1)) return nullptr;
  sv_ctx *c = ctx_of(env, a.v[0]);
  if (!c) return nullptr;
  sv_kernel_time t[64];
  int n = 64;
  SV(c, sv_timings(c, t, &n), "sv_timings");
  napi_value arr;
  CK(napi_create_array(env, &arr));
  for (int i = 0, j = 0; i < n && i < 64; ++i) {
    if (!t[i].launches) continue;
    napi_value o, v;
    CK(napi_create_object(env, &o));
    CK(napi_create_string_utf8(env, t[i].name, NAPI_AUTO_LENGTH, &v));
    CK(napi_set_named_property(env, o, "name", v));
    CK(napi_create_double(env, t[i].total_ms, &v));
    CK(napi_set_named_property(env, o, "totalMs", v));
    CK(napi_create_double(env, (double)t[i].launches, &v));
    CK(napi_set_named_property(env, o, "launches", v));
    CK(napi_set_element(env, arr, (uint32_t)j++, o));
  }
  return arr;
}

napi_value Autotune(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 3)) return nullptr;
  sv_ctx *c = ctx_of(env, a.v[0]);
  if (!c) return nullptr;
  std::vector<int> vs;
  int32_t steps;
  if (!views_from_js(env, a.v[1], vs) || !get_i32(env, a.v[2], &steps)) return nullptr;
  int n = 0;
  SV(c, sv_autotune(c, vs.data(), (int)vs.size(), steps, nullptr, &n), "sv_autotune");
  std::vector<sv_tune_entry> t(n);
  SV(c, sv_autotune(c, vs.data(), (int)vs.size(), steps, t.data(), &n), "sv_autotune");
  napi_value arr;
  CK(napi_create_array(env, &arr));
  for (int i = 0; i < n; ++i) {
    napi_value o, v;
    CK(napi_create_object(env, &o));
    const uint32_t f[5] = {t[i].batch, t[i].persistent, t[i].xcd_remap, t[i].tile_size, t[i].split_backward};
    const char *names[5] = {"batch", "persistent", "xcd_remap", "tile", "splitBackward"};
    for (int k = 0; k < 5; ++k) {
      CK(napi_create_uint32(env, f[k], &v));
      CK(napi_set_named_property(env, o, names[k], v));
    }
    CK(napi_create_double(env, t[i].ms_per_step, &v));
    CK(napi_set_named_property(env, o, "msPerStep", v));
    CK(napi_set_element(env, arr, (uint32_t)i, o));
  }
  return arr;
}

napi_value checkpoint(napi_env env, napi_callback_info info, sv_status (*fn)(sv_ctx *, const char *),
                      const char *what) {
  Args a;
  if (!get_args(env, info, a, 2)) return nullptr;
  sv_ctx *c = ctx_of(env, a.v[0]);
  if (!c) return nullptr;
  char path[4096];
  size_t len = 0;
  if (napi_get_value_string_utf8(env, a.v[1], path, sizeof(path), &len) != napi_ok || len + 1 >= sizeof(path)) {
    napi_throw_type_error(env, nullptr, "path must be a string");
    return nullptr;
  }
  SV(c, fn(c, path), what);
  return nullptr;
}
napi_value SaveCheckpoint(napi_env env, napi_callback_info info) {
  return checkpoint(env, info, sv_save_checkpoint, "sv_save_checkpoint");
}
napi_value LoadCheckpoint(napi_env env, napi_callback_info info) {
  return checkpoint(env, info, sv_load_checkpoint, "sv_load_checkpoint");
}

napi_value Sync(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 1)) return nullptr;
  sv_ctx *c = ctx_of(env, a.v[0]);
  if (!c) return nullptr;
  SV(c, sv_sync(c), "sv_sync");
  return nullptr;
}

// seeded scene helpers (no device) ----------------------------------------
napi_value SceneGaussians(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 3)) return nullptr;
  double seed, n, half;
  if (!get_f64(env, a.v[0], &seed) || !get_f64(env, a.v[1], &n) || !get_f64(env, a.v[2], &half)) return nullptr;
  float *p[5];
  napi_value o = params_to_js(env, (size_t)n, p);
  if (!o) return nullptr;
  SV(nullptr, sv_scene_gaussians((uint64_t)seed, (size_t)n, (float)half, p[0], p[1], p[2], p[3], p[4]),
     "sv_scene_gaussians");
  return o;
}

napi_value SceneJitter(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 3)) return nullptr;
  double seed, sigma;
  size_t len = 0;
  const float *src = typed<float>(env, a.v[1], napi_float32_array, 0, &len);
  if (!src || !get_f64(env, a.v[0], &seed) || !get_f64(env, a.v[2], &sigma)) return nullptr;
  float *p = nullptr;
  napi_value arr = new_typed<float>(env, napi_float32_array, len, &p);
  if (!arr) return nullptr;
  std::memcpy(p, src, len * sizeof(float));
  SV(nullptr, sv_scene_jitter((uint64_t)seed, len / 3, (float)sigma, p), "sv_scene_jitter");
  return arr;
}

napi_value SceneCamera(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a, 4)) return nullptr;
  float eye[3], tgt[3];
  for (uint32_t i = 0; i < 3; ++i) {
    napi_value e;
    double d;
    napi_get_element(env, a.v[0], i, &e);
    if (!get_f64(env, e, &d)) return nullptr;
    eye[i] = (float)d;
    napi_get_element(env, a.v[1], i, &e);
    if (!get_f64(env, e, &d)) return nullptr;
    tgt[i] = (float)d;
  }
  uint32_t w, h;
  double fovy = 60.0;
  if (!get_u32(env, a.v[2], &w) || !get_u32(env, a.v[3], &h)) return nullptr;
  if (a.n > 4) get_f64(env, a.v[4], &fovy);
  sv_camera cam;
  SV(nullptr, sv_scene_camera(eye, tgt, w, h, (float)fovy, &cam), "sv_scene_camera");
  napi_value o;
  CK(napi_create_object(env, &o));
  struct {
    const char *name;
    const float *src;
    size_t n;
  } f[] = {{"view", cam.view, 16}, {"view_inv", cam.view_inv, 16}, {"proj", cam.proj, 16},
           {"proj_inv", cam.proj_inv, 16}, {"viewport", cam.viewport, 2}, {"focal", cam.focal, 2}};
  for (auto &x : f) {
    float *p = nullptr;
    napi_value arr = new_typed<float>(env, napi_float32_array, x.n, &p);
    if (!arr) return nullptr;
    std::memcpy(p, x.src, x.n * sizeof(float));
    CK(napi_set_named_property(env, o, x.name, arr));
  }
  return o;
}

napi_value Init(napi_env env, napi_value exports) {
  struct {
    const char *name;
    napi_callback fn;
  } fns[] = {{"version", Version},         {"create", Create},           {"destroy", Destroy},
             {"getConfig", GetConfig},     {"setConfig", SetConfig},     {"setGaussians", SetGaussians},
             {"getGaussians", GetGaussians}, {"setMoments", SetMoments}, {"getMoments", GetMoments},
             {"getCount", GetCount},       {"setViews", SetViews},       {"setTarget", SetTarget},
             {"setResidual", SetResidual}, {"buildTiles", BuildTiles},   {"forward", Forward},
             {"loss", Loss},               {"backward", Backward},       {"backwardFixed", BackwardFixed},
             {"zeroGrad", ZeroGrad},       {"adam", Adam},               {"step", Step},
             {"stepAsync", StepAsync},     {"readImage", ReadImage},     {"readResidual", ReadResidual},
             {"readAlpha", ReadAlpha},     {"readLastIds", ReadLastIds}, {"readGrads", ReadGrads},
             {"readKeys", ReadKeys},       {"pruneOpacity", PruneOpacity}, {"pruneMask", PruneMask},
             {"densify", Densify},         {"resetOpacity", ResetOpacity}, {"knnInit", KnnInit},
             {"multiViewScores", MultiViewScores}, {"renderCamera", RenderCamera},         {"densifyDecide", DensifyDecide}, {"densifyDecideGrad", DensifyDecideGrad}, {"enableTiming", EnableTiming},
             {"timings", Timings},         {"sync", Sync},               {"autotune", Autotune},
             {"saveCheckpoint", SaveCheckpoint}, {"loadCheckpoint", LoadCheckpoint},               {"sceneGaussians", SceneGaussians},
             {"sceneJitter", SceneJitter}, {"sceneCamera", SceneCamera}};
  for (auto &f : fns) {
    napi_value v;
    if (napi_create_function(env, f.name, NAPI_AUTO_LENGTH, f.fn, nullptr, &v) != napi_ok ||
        napi_set_named_property(env, exports, f.name, v) != napi_ok)
      return nullptr;
  }
  napi_value k;
  const struct {
    const char *name;
    int v;
  } consts[] = {{"MODE_PLAIN", SV_MODE_PLAIN},
                {"MODE_BATCHED", SV_MODE_BATCHED},
                {"MODE_SEQUENTIAL", SV_MODE_SEQUENTIAL},
                {"WF_REFERENCE_ORDER", SV_WF_REFERENCE_ORDER},
                {"WF_SKIP_098", SV_WF_SKIP_098}};
  for (auto &x : consts) {
    napi_create_int32(env, x.v, &k);
    napi_set_named_property(env, exports, x.name, k);
  }
  return exports;
}

}  // namespace

NAPI_MODULE(skyfall_sv, Init)
