// pike_napi.cc — Node N-API addon over the C ABI in include/pk.h.
//
// This is the JavaScript host's only way into the GPU: the pass classes in
// pike_amd/js call these functions where the reference's passes call WebGPU
// (device.createBuffer / createComputePipeline / dispatchWorkgroups / queue
// submission).  Every failing call throws an Error whose message is
// pk_last_error() and whose `code` is the PK_E_* value.  Handles are tagged
// externals; they are released by the explicit destroy functions (the
// reference never frees its GPU objects either), so garbage collection
// order can never free a context before the scenes that use it.
#include <node_api.h>

#include <cstring>
#include <algorithm>
#include <string>
#include <vector>

#include "../../include/pk.h"

namespace {

enum Tag : uint32_t { kCtx = 0x70636b31, kBvh = 0x70626b31, kScene = 0x70736b31, kFrame = 0x70666b31 };

struct Handle {
  uint32_t tag;
  void* ptr;
};

#define NAPI_OK(call)                                              \
  do {                                                             \
    if ((call) != napi_ok) {                                       \
      napi_throw_error(env, nullptr, "N-API call failed: " #call); \
      return nullptr;                                              \
    }                                                              \
  } while (0)

napi_value throw_pk(napi_env env, int rc, const char* what) {
  std::string msg = std::string(what) + ": " + pk_last_error();
  napi_value err, m, code;
  napi_create_string_utf8(env, msg.c_str(), msg.size(), &m);
  napi_create_error(env, nullptr, m, &err);
  napi_create_int32(env, rc, &code);
  napi_set_named_property(env, err, "code", code);
  napi_throw(env, err);
  return nullptr;
}

#define PK_CALL(expr, what)                                   \
  do {                                                        \
    int rc_ = (expr);                                         \
    if (rc_ != PK_OK) return throw_pk(env, rc_, what);        \
  } while (0)

struct Args {
  size_t argc = 8;
  napi_value argv[8];
  napi_value self;
};

bool get_args(napi_env env, napi_callback_info info, Args* a, size_t need) {
  if (napi_get_cb_info(env, info, &a->argc, a->argv, &a->self, nullptr) != napi_ok) return false;
  if (a->argc < need) {
    napi_throw_type_error(env, nullptr, "too few arguments");
    return false;
  }
  return true;
}

napi_value wrap(napi_env env, uint32_t tag, void* p) {
  auto* h = new Handle{tag, p};
  napi_value v;
  if (napi_create_external(env, h, [](napi_env, void* d, void*) { delete static_cast<Handle*>(d); }, nullptr, &v) !=
      napi_ok)
    return nullptr;
  return v;
}

void* unwrap(napi_env env, napi_value v, uint32_t tag, const char* what) {
  void* d = nullptr;
  if (napi_get_value_external(env, v, &d) != napi_ok || !d || static_cast<Handle*>(d)->tag != tag) {
    napi_throw_type_error(env, nullptr, (std::string("expected a ") + what + " handle").c_str());
    return nullptr;
  }
  void* p = static_cast<Handle*>(d)->ptr;
  if (!p) napi_throw_error(env, nullptr, (std::string(what) + " handle was destroyed").c_str());
  return p;
}

void clear(napi_env env, napi_value v) {
  void* d = nullptr;
  if (napi_get_value_external(env, v, &d) == napi_ok && d) static_cast<Handle*>(d)->ptr = nullptr;
}

// typed array view -> data pointer and element count
bool typed(napi_env env, napi_value v, napi_typedarray_type want, void** data, size_t* len, const char* what) {
  bool is = false;
  napi_is_typedarray(env, v, &is);
  if (!is) {
    napi_throw_type_error(env, nullptr, (std::string(what) + " must be a typed array").c_str());
    return false;
  }
  napi_typedarray_type t;
  napi_value ab;
  size_t off;
  if (napi_get_typedarray_info(env, v, &t, len, data, &ab, &off) != napi_ok) return false;
  if (t != want) {
    napi_throw_type_error(env, nullptr, (std::string(what) + " has the wrong element type").c_str());
    return false;
  }
  return true;
}

bool get_u32(napi_env env, napi_value obj, const char* key, uint32_t* out, uint32_t def) {
  bool has = false;
  napi_has_named_property(env, obj, key, &has);
  if (!has) {
    *out = def;
    return true;
  }
  napi_value v;
  napi_get_named_property(env, obj, key, &v);
  return napi_get_value_uint32(env, v, out) == napi_ok;
}

bool get_f32(napi_env env, napi_value obj, const char* key, float* out, float def) {
  bool has = false;
  napi_has_named_property(env, obj, key, &has);
  if (!has) {
    *out = def;
    return true;
  }
  napi_value v;
  double d;
  napi_get_named_property(env, obj, key, &v);
  if (napi_get_value_double(env, v, &d) != napi_ok) return false;
  *out = (float)d;
  return true;
}

bool get_vec3(napi_env env, napi_value obj, const char* key, float out[3]) {
  napi_value arr;
  if (napi_get_named_property(env, obj, key, &arr) != napi_ok) return false;
  for (uint32_t i = 0; i < 3; ++i) {
    napi_value e;
    double d;
    if (napi_get_element(env, arr, i, &e) != napi_ok || napi_get_value_double(env, e, &d) != napi_ok) return false;
    out[i] = (float)d;
  }
  return true;
}

napi_value make_typed(napi_env env, napi_typedarray_type t, size_t n, size_t elem, void** data) {
  napi_value ab, ta;
  if (napi_create_arraybuffer(env, n * elem, data, &ab) != napi_ok) return nullptr;
  if (napi_create_typedarray(env, t, n, ab, 0, &ta) != napi_ok) return nullptr;
  return ta;
}

void set_num(napi_env env, napi_value obj, const char* key, double v) {
  napi_value x;
  napi_create_double(env, v, &x);
  napi_set_named_property(env, obj, key, x);
}

napi_value undefined(napi_env env) {
  napi_value u;
  napi_get_undefined(env, &u);
  return u;
}

// ---------------------------------------------------------------------------
napi_value Version(napi_env env, napi_callback_info) {
  napi_value s;
  const char* v = pk_version();
  napi_create_string_utf8(env, v, std::strlen(v), &s);
  return s;
}

napi_value CtxCreate(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, &a, 0)) return nullptr;
  int32_t dev = 0;
  if (a.argc > 0) napi_get_value_int32(env, a.argv[0], &dev);
  pk_ctx* c = nullptr;
  PK_CALL(pk_ctx_create(dev, &c), "pk_ctx_create");
  return wrap(env, kCtx, c);
}

napi_value CtxDestroy(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, &a, 1)) return nullptr;
  void* d = nullptr;
  napi_get_value_external(env, a.argv[0], &d);
  if (d && static_cast<Handle*>(d)->tag == kCtx && static_cast<Handle*>(d)->ptr) {
    pk_ctx_destroy((pk_ctx*)static_cast<Handle*>(d)->ptr);
    clear(env, a.argv[0]);
  }
  return undefined(env);
}

napi_value Sync(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, &a, 1)) return nullptr;
  auto* c = (pk_ctx*)unwrap(env, a.argv[0], kCtx, "context");
  if (!c) return nullptr;
  PK_CALL(pk_sync(c), "pk_sync");
  return undefined(env);
}

napi_value BvhBuild(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, &a, 2)) return nullptr;
  void *pos, *tri;
  size_t np, nt;
  if (!typed(env, a.argv[0], napi_float64_array, &pos, &np, "positions")) return nullptr;
  if (!typed(env, a.argv[1], napi_int32_array, &tri, &nt, "triangles")) return nullptr;
  int32_t threads = 0;
  if (a.argc > 2) napi_get_value_int32(env, a.argv[2], &threads);
  bool spatial = false;  // bvhBuild(positions, triangles, threads, spatialSplits)
  if (a.argc > 3) napi_get_value_bool(env, a.argv[3], &spatial);
  pk_bvh* b = nullptr;
  PK_CALL(pk_bvh_build_ex((const double*)pos, (uint32_t)(np / 3), (const int32_t*)tri, (uint32_t)(nt / 4), threads,
                          spatial ? PK_BVH_SPATIAL_SPLITS : 0u, &b),
          "pk_bvh_build_ex");
  return wrap(env, kBvh, b);
}

bool scene_desc(napi_env env, napi_value o, pk_scene_desc* d);

// bvhBuildInstanced(desc, meshes, instances, threads, spatialSplits):
// meshes [{positions: Float64Array, triangles: Int32Array, attributes?:
// Uint32Array}], instances [{mesh, material (-1 default), transform: 12
// numbers, row-major 3x4}]
napi_value BvhBuildInstanced(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, &a, 3)) return nullptr;
  pk_scene_desc d;
  if (!scene_desc(env, a.argv[0], &d)) return nullptr;
  uint32_t nm = 0, ni = 0;
  if (napi_get_array_length(env, a.argv[1], &nm) != napi_ok ||
      napi_get_array_length(env, a.argv[2], &ni) != napi_ok) {
    napi_throw_type_error(env, nullptr, "bvhBuildInstanced: meshes and instances must be arrays");
    return nullptr;
  }
  std::vector<pk_mesh_desc> meshes(nm);
  for (uint32_t m = 0; m < nm; ++m) {
    napi_value o, v;
    void* p;
    size_t n;
    napi_get_element(env, a.argv[1], m, &o);
    pk_mesh_desc& md = meshes[m];
    std::memset(&md, 0, sizeof(md));
    if (napi_get_named_property(env, o, "positions", &v) != napi_ok ||
        !typed(env, v, napi_float64_array, &p, &n, "mesh.positions"))
      return nullptr;
    md.positions = (const double*)p;
    md.num_vertices = (uint32_t)(n / 3);
    if (napi_get_named_property(env, o, "triangles", &v) != napi_ok ||
        !typed(env, v, napi_int32_array, &p, &n, "mesh.triangles"))
      return nullptr;
    md.triangles = (const int32_t*)p;
    md.num_triangles = (uint32_t)(n / 4);
    bool has = false;
    napi_has_named_property(env, o, "attributes", &has);
    if (has) {
      napi_get_named_property(env, o, "attributes", &v);
      if (!typed(env, v, napi_uint32_array, &p, &n, "mesh.attributes")) return nullptr;
      if (n != (size_t)md.num_vertices * 4) {
        napi_throw_range_error(env, nullptr, "mesh.attributes must hold 4 words per vertex");
        return nullptr;
      }
      md.attributes = (const uint32_t*)p;
    }
  }
  std::vector<pk_instance_desc> insts(ni);
  for (uint32_t i = 0; i < ni; ++i) {
    napi_value o, v;
    napi_get_element(env, a.argv[2], i, &o);
    pk_instance_desc& id = insts[i];
    std::memset(&id, 0, sizeof(id));
    get_u32(env, o, "mesh", &id.mesh, 0);
    double mat = -1;
    bool has = false;
    napi_has_named_property(env, o, "material", &has);
    if (has) {
      napi_get_named_property(env, o, "material", &v);
      napi_get_value_double(env, v, &mat);
    }
    id.material = (int32_t)mat;
    uint32_t len = 0;
    if (napi_get_named_property(env, o, "transform", &v) != napi_ok) return nullptr;
    bool is_typed = false;
    napi_is_typedarray(env, v, &is_typed);
    if (is_typed) {
      void* p;
      size_t n;
      if (!typed(env, v, napi_float64_array, &p, &n, "instance.transform")) return nullptr;
      len = (uint32_t)n;
      if (len == 12) std::memcpy(id.transform, p, sizeof(id.transform));
    } else if (napi_get_array_length(env, v, &len) == napi_ok && len == 12) {
      for (uint32_t k = 0; k < 12; ++k) {
        napi_value e;
        napi_get_element(env, v, k, &e);
        napi_get_value_double(env, e, &id.transform[k]);
      }
    }
    if (len != 12) {
      napi_throw_range_error(env, nullptr, "instance.transform must hold 12 numbers (row-major 3x4)");
      return nullptr;
    }
  }
  int32_t threads = 0;
  if (a.argc > 3) napi_get_value_int32(env, a.argv[3], &threads);
  bool spatial = false;
  if (a.argc > 4) napi_get_value_bool(env, a.argv[4], &spatial);
  pk_bvh* b = nullptr;
  PK_CALL(pk_bvh_build_instanced(&d, meshes.data(), nm, insts.data(), ni, threads,
                                 spatial ? PK_BVH_SPATIAL_SPLITS : 0u, &b),
          "pk_bvh_build_instanced");
  return wrap(env, kBvh, b);
}

// bvhNormalizer(bvh) -> Float64Array [cx, cy, cz, scale]
napi_value BvhNormalizer(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, &a, 1)) return nullptr;
  auto* b = (pk_bvh*)unwrap(env, a.argv[0], kBvh, "bvh");
  if (!b) return nullptr;
  void* p;
  napi_value out = make_typed(env, napi_float64_array, 4, 8, &p);
  if (!out) return nullptr;
  PK_CALL(pk_bvh_normalizer(b, (double*)p), "pk_bvh_normalizer");
  return out;
}

napi_value BvhInfo(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, &a, 1)) return nullptr;
  auto* b = (pk_bvh*)unwrap(env, a.argv[0], kBvh, "bvh");
  if (!b) return nullptr;
  int64_t v[6];
  PK_CALL(pk_bvh_info(b, v), "pk_bvh_info");
  napi_value o;
  napi_create_object(env, &o);
  const char* keys[6] = {"binaryNodes", "wideNodes", "leafTriangles", "binaryDepth", "wideDepth", "largestLeaf"};
  for (int i = 0; i < 6; ++i) set_num(env, o, keys[i], (double)v[i]);
  return o;
}

napi_value BvhExportReference(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, &a, 3)) return nullptr;
  auto* b = (pk_bvh*)unwrap(env, a.argv[0], kBvh, "bvh");
  if (!b) return nullptr;
  uint32_t nv = 0, nt = 0;
  napi_get_value_uint32(env, a.argv[1], &nv);
  napi_get_value_uint32(env, a.argv[2], &nt);
  int64_t v[6];
  PK_CALL(pk_bvh_info(b, v), "pk_bvh_info");
  void *nodes, *tris, *verts, *norm;
  napi_value tn = make_typed(env, napi_int32_array, (size_t)v[0] * 5, 4, &nodes);
  // leaf triangle references (more than the triangles with spatial splits)
  napi_value tt = make_typed(env, napi_int32_array, (size_t)std::max<int64_t>(v[2], nt) * 4, 4, &tris);
  napi_value tv = make_typed(env, napi_float32_array, (size_t)nv * 4, 4, &verts);
  napi_value tm = make_typed(env, napi_float64_array, 4, 8, &norm);
  if (!tn || !tt || !tv || !tm) return nullptr;
  PK_CALL(pk_bvh_export_reference(b, (int32_t*)nodes, (int32_t*)tris, (float*)verts, (double*)norm),
          "pk_bvh_export_reference");
  napi_value o;
  napi_create_object(env, &o);
  napi_set_named_property(env, o, "nodes", tn);
  napi_set_named_property(env, o, "triangles", tt);
  napi_set_named_property(env, o, "vertices", tv);
  napi_set_named_property(env, o, "normalizer", tm);
  return o;
}

napi_value BvhDestroy(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, &a, 1)) return nullptr;
  void* d = nullptr;
  napi_get_value_external(env, a.argv[0], &d);
  if (d && static_cast<Handle*>(d)->tag == kBvh && static_cast<Handle*>(d)->ptr) {
    pk_bvh_destroy((pk_bvh*)static_cast<Handle*>(d)->ptr);
    clear(env, a.argv[0]);
  }
  return undefined(env);
}

// desc object -> pk_scene_desc (pointers stay valid while the JS arrays live,
// i.e. for the duration of the call that copies them)
bool scene_desc(napi_env env, napi_value o, pk_scene_desc* d) {
  std::memset(d, 0, sizeof(*d));
  napi_value v;
  void* p;
  size_t n;
  if (napi_get_named_property(env, o, "positions", &v) != napi_ok ||
      !typed(env, v, napi_float64_array, &p, &n, "desc.positions"))
    return false;
  d->positions = (const double*)p;
  d->num_vertices = (uint32_t)(n / 3);
  if (napi_get_named_property(env, o, "triangles", &v) != napi_ok ||
      !typed(env, v, napi_int32_array, &p, &n, "desc.triangles"))
    return false;
  d->triangles = (const int32_t*)p;
  d->num_triangles = (uint32_t)(n / 4);
  bool has = false;
  napi_has_named_property(env, o, "attributes", &has);
  if (has) {
    napi_get_named_property(env, o, "attributes", &v);
    if (!typed(env, v, napi_uint32_array, &p, &n, "desc.attributes")) return false;
    d->attributes = (const uint32_t*)p;
  }
  napi_has_named_property(env, o, "materials", &has);
  if (has) {
    napi_get_named_property(env, o, "materials", &v);
    if (!typed(env, v, napi_int32_array, &p, &n, "desc.materials")) return false;
    d->materials = (const int32_t*)p;
    d->num_materials = (uint32_t)(n / 4);
  }
  napi_has_named_property(env, o, "atlas", &has);
  if (has) {
    napi_value t;
    napi_get_named_property(env, o, "atlas", &t);
    get_u32(env, t, "width", &d->tex_width, 0);
    get_u32(env, t, "height", &d->tex_height, 0);
    get_u32(env, t, "layers", &d->tex_layers, 0);
    get_u32(env, t, "format", &d->tex_format, PK_TEX_RGBA8);
    napi_get_named_property(env, t, "data", &v);
    if (!typed(env, v, d->tex_format == PK_TEX_RGBA8 ? napi_uint8_array : napi_float32_array, &p, &n, "atlas.data"))
      return false;
    d->tex_data = p;
  }
  napi_has_named_property(env, o, "env", &has);
  if (has) {
    napi_value e;
    napi_get_named_property(env, o, "env", &e);
    napi_valuetype et;
    napi_typeof(env, e, &et);
    if (et == napi_object) {
      get_u32(env, e, "width", &d->env_width, 0);
      get_u32(env, e, "height", &d->env_height, 0);
      get_u32(env, e, "format", &d->env_format, PK_ENV_RGBE8);
      get_u32(env, e, "sampler", &d->env_sampler, PK_ENV_SAMPLER_BINS);
      napi_get_named_property(env, e, "data", &v);
      if (!typed(env, v, d->env_format == PK_ENV_RGBE8 ? napi_uint8_array : napi_float32_array, &p, &n, "env.data"))
        return false;
      d->env_data = p;
    }
  }
  return true;
}

napi_value SceneUpload(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, &a, 3)) return nullptr;
  auto* c = (pk_ctx*)unwrap(env, a.argv[0], kCtx, "context");
  if (!c) return nullptr;
  auto* b = (pk_bvh*)unwrap(env, a.argv[2], kBvh, "bvh");
  if (!b) return nullptr;
  pk_scene_desc d;
  if (!scene_desc(env, a.argv[1], &d)) return nullptr;
  pk_scene* s = nullptr;
  PK_CALL(pk_scene_upload(c, &d, b, &s), "pk_scene_upload");
  return wrap(env, kScene, s);
}

napi_value SceneDestroy(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, &a, 1)) return nullptr;
  void* d = nullptr;
  napi_get_value_external(env, a.argv[0], &d);
  if (d && static_cast<Handle*>(d)->tag == kScene && static_cast<Handle*>(d)->ptr) {
    pk_scene_destroy((pk_scene*)static_cast<Handle*>(d)->ptr);
    clear(env, a.argv[0]);
  }
  return undefined(env);
}

napi_value EnvTables(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, &a, 1)) return nullptr;
  pk_scene_desc d;
  if (!scene_desc(env, a.argv[0], &d)) return nullptr;
  int64_t inf[4];
  PK_CALL(pk_env_tables(&d, inf, nullptr, nullptr, nullptr), "pk_env_tables");
  void *bins, *coords, *pdf;
  napi_value tb = make_typed(env, napi_int32_array, (size_t)inf[0] * 2, 4, &bins);
  napi_value tc = make_typed(env, napi_int32_array, (size_t)inf[1] * 2, 4, &coords);
  napi_value tp = make_typed(env, napi_float32_array, (size_t)(inf[2] * inf[3]), 4, &pdf);
  if (!tb || !tc || !tp) return nullptr;
  PK_CALL(pk_env_tables(&d, inf, (int32_t*)bins, (int32_t*)coords, (float*)pdf), "pk_env_tables");
  napi_value o;
  napi_create_object(env, &o);
  napi_set_named_property(env, o, "bins", tb);
  napi_set_named_property(env, o, "coords", tc);
  napi_set_named_property(env, o, "pdf", tp);
  return o;
}

napi_value FrameCreate(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, &a, 3)) return nullptr;
  auto* c = (pk_ctx*)unwrap(env, a.argv[0], kCtx, "context");
  if (!c) return nullptr;
  uint32_t w = 0, h = 0;
  napi_get_value_uint32(env, a.argv[1], &w);
  napi_get_value_uint32(env, a.argv[2], &h);
  pk_frame* f = nullptr;
  PK_CALL(pk_frame_create(c, w, h, &f), "pk_frame_create");
  return wrap(env, kFrame, f);
}

napi_value FrameDestroy(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, &a, 1)) return nullptr;
  void* d = nullptr;
  napi_get_value_external(env, a.argv[0], &d);
  if (d && static_cast<Handle*>(d)->tag == kFrame && static_cast<Handle*>(d)->ptr) {
    pk_frame_destroy((pk_frame*)static_cast<Handle*>(d)->ptr);
    clear(env, a.argv[0]);
  }
  return undefined(env);
}

#define FRAME_ARG(i)                                                  \
  auto* f = (pk_frame*)unwrap(env, a.argv[i], kFrame, "frame");       \
  if (!f) return nullptr
#define SCENE_ARG(i)                                                  \
  auto* s = (pk_scene*)unwrap(env, a.argv[i], kScene, "scene");       \
  if (!s) return nullptr

napi_value FrameReset(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, &a, 1)) return nullptr;
  FRAME_ARG(0);
  PK_CALL(pk_frame_reset(f), "pk_frame_reset");
  return undefined(env);
}

// CameraState fields as CameraPass.setCameraState receives them
// (render/passes/camera_pass.js:118-125, pike.js:138-147)
napi_value FrameSetCamera(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, &a, 2)) return nullptr;
  FRAME_ARG(0);
  pk_camera cam;
  napi_value o = a.argv[1];
  if (!get_vec3(env, o, "pos", cam.origin) || !get_vec3(env, o, "dir", cam.dir)) {
    napi_throw_type_error(env, nullptr, "camera state needs pos and dir arrays");
    return nullptr;
  }
  get_f32(env, o, "fov", &cam.fov, 1.0f);
  get_f32(env, o, "focalDepth", &cam.focal_depth, 0.5f);
  get_f32(env, o, "apertureSize", &cam.aperture, 0.0f);
  get_f32(env, o, "distortion", &cam.distortion, 0.0f);
  get_f32(env, o, "bokeh", &cam.bokeh, 0.5f);
  PK_CALL(pk_frame_set_camera(f, &cam), "pk_frame_set_camera");
  return undefined(env);
}

napi_value FrameSetEnvTheta(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, &a, 2)) return nullptr;
  FRAME_ARG(0);
  double t = 0;
  napi_get_value_double(env, a.argv[1], &t);
  PK_CALL(pk_frame_set_env_theta(f, (float)t), "pk_frame_set_env_theta");
  return undefined(env);
}

napi_value FrameBegin(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, &a, 2)) return nullptr;
  FRAME_ARG(0);
  uint32_t sample = 0;
  napi_get_value_uint32(env, a.argv[1], &sample);
  PK_CALL(pk_frame_begin(f, sample), "pk_frame_begin");
  return undefined(env);
}

napi_value Raygen(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, &a, 1)) return nullptr;
  FRAME_ARG(0);
  PK_CALL(pk_raygen(f), "pk_raygen");
  return undefined(env);
}

napi_value TraceClosest(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, &a, 3)) return nullptr;
  FRAME_ARG(0);
  SCENE_ARG(1);
  uint32_t bounce = 0;
  napi_get_value_uint32(env, a.argv[2], &bounce);
  PK_CALL(pk_trace_closest(f, s, bounce), "pk_trace_closest");
  return undefined(env);
}

napi_value Shade(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, &a, 3)) return nullptr;
  FRAME_ARG(0);
  SCENE_ARG(1);
  uint32_t bounce = 0;
  napi_get_value_uint32(env, a.argv[2], &bounce);
  PK_CALL(pk_shade(f, s, bounce), "pk_shade");
  return undefined(env);
}

napi_value TraceShadow(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, &a, 2)) return nullptr;
  FRAME_ARG(0);
  SCENE_ARG(1);
  PK_CALL(pk_trace_shadow(f, s), "pk_trace_shadow");
  return undefined(env);
}

napi_value Accumulate(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, &a, 1)) return nullptr;
  FRAME_ARG(0);
  PK_CALL(pk_accumulate(f), "pk_accumulate");
  return undefined(env);
}

napi_value Render(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, &a, 3)) return nullptr;
  FRAME_ARG(0);
  SCENE_ARG(1);
  pk_render_params p;
  napi_value o = a.argv[2];
  get_u32(env, o, "width", &p.width, 0);
  get_u32(env, o, "height", &p.height, 0);
  get_u32(env, o, "maxDepth", &p.max_depth, 4);
  get_u32(env, o, "firstSample", &p.first_sample, 0);
  get_u32(env, o, "numFrames", &p.num_frames, 1);
  get_f32(env, o, "envTheta", &p.env_theta, 0.0f);
  get_u32(env, o, "useGraph", &p.use_graph, 1);
  PK_CALL(pk_render(f, s, &p), "pk_render");
  return undefined(env);
}

napi_value ReadImage(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, &a, 3)) return nullptr;
  FRAME_ARG(0);
  uint32_t w = 0, h = 0;
  napi_get_value_uint32(env, a.argv[1], &w);
  napi_get_value_uint32(env, a.argv[2], &h);
  bool color = false;
  if (a.argc > 3) napi_get_value_bool(env, a.argv[3], &color);
  void* data;
  napi_value t = make_typed(env, napi_float32_array, (size_t)w * h * 4, 4, &data);
  if (!t) return nullptr;
  if (color) PK_CALL(pk_read_color(f, (float*)data, (size_t)w * h * 16), "pk_read_color");
  else PK_CALL(pk_read_image(f, (float*)data, (size_t)w * h * 16), "pk_read_image");
  return t;
}

napi_value Counters(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, &a, 1)) return nullptr;
  FRAME_ARG(0);
  pk_counters c;
  PK_CALL(pk_read_counters(f, &c), "pk_read_counters");
  napi_value o;
  napi_create_object(env, &o);
  set_num(env, o, "raysTraced", (double)c.rays_traced);
  set_num(env, o, "shadowRays", (double)c.shadow_rays);
  set_num(env, o, "hits", (double)c.hits);
  set_num(env, o, "stackOverflow", (double)c.stack_overflow);
  set_num(env, o, "frames", (double)c.frames);
  return o;
}

napi_value Autotune(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, &a, 3)) return nullptr;
  FRAME_ARG(0);
  SCENE_ARG(1);
  uint32_t depth = 4, frames = 2;
  napi_get_value_uint32(env, a.argv[2], &depth);
  if (a.argc > 3) napi_get_value_uint32(env, a.argv[3], &frames);
  float ms[2];
  PK_CALL(pk_autotune_render(f, s, depth, frames, ms), "pk_autotune_render");
  napi_value o;
  napi_create_object(env, &o);
  set_num(env, o, "msBefore", ms[0]);
  set_num(env, o, "msAfter", ms[1]);
  return o;
}

napi_value TraceRays(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, &a, 3)) return nullptr;
  auto* c = (pk_ctx*)unwrap(env, a.argv[0], kCtx, "context");
  if (!c) return nullptr;
  SCENE_ARG(1);
  void* rays;
  size_t n;
  if (!typed(env, a.argv[2], napi_float32_array, &rays, &n, "rays")) return nullptr;
  bool any = false;
  if (a.argc > 3) napi_get_value_bool(env, a.argv[3], &any);
  const size_t nr = n / 8;
  void *ids, *t, *bary;
  napi_value ti = make_typed(env, napi_int32_array, nr, 4, &ids);
  napi_value tt = make_typed(env, napi_float32_array, nr, 4, &t);
  napi_value tb = make_typed(env, napi_uint32_array, nr, 4, &bary);
  if (!ti || !tt || !tb) return nullptr;
  PK_CALL(pk_trace_rays(c, s, (const float*)rays, (int64_t)nr, any ? 1 : 0, (int32_t*)ids, (float*)t, (uint32_t*)bary),
          "pk_trace_rays");
  napi_value o;
  napi_create_object(env, &o);
  napi_set_named_property(env, o, "ids", ti);
  napi_set_named_property(env, o, "t", tt);
  napi_set_named_property(env, o, "bary", tb);
  return o;
}

napi_value FrameSetSamplesPerPass(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, &a, 2)) return nullptr;
  FRAME_ARG(0);
  uint32_t samples = 1;
  napi_get_value_uint32(env, a.argv[1], &samples);
  PK_CALL(pk_frame_set_samples_per_pass(f, samples), "pk_frame_set_samples_per_pass");
  return undefined(env);
}

napi_value AutotuneSamples(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, &a, 3)) return nullptr;
  FRAME_ARG(0);
  SCENE_ARG(1);
  uint32_t depth = 4, max_samples = 8;
  napi_get_value_uint32(env, a.argv[2], &depth);
  if (a.argc > 3) napi_get_value_uint32(env, a.argv[3], &max_samples);
  float ms[2];
  PK_CALL(pk_autotune_samples_per_pass(f, s, depth, max_samples, ms), "pk_autotune_samples_per_pass");
  uint32_t chosen = 1;
  PK_CALL(pk_frame_get_samples_per_pass(f, &chosen), "pk_frame_get_samples_per_pass");
  napi_value o;
  napi_create_object(env, &o);
  set_num(env, o, "msPerSampleBefore", ms[0]);
  set_num(env, o, "msPerSampleAfter", ms[1]);
  set_num(env, o, "samplesPerPass", chosen);
  return o;
}

// autotuneScheduler(frame, scene, depth, samples): keeps the faster pass
// scheduler in the frame's context
napi_value AutotuneScheduler(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, &a, 3)) return nullptr;
  FRAME_ARG(0);
  SCENE_ARG(1);
  uint32_t depth = 4, samples = 8;
  napi_get_value_uint32(env, a.argv[2], &depth);
  if (a.argc > 3) napi_get_value_uint32(env, a.argv[3], &samples);
  float ms[2];
  PK_CALL(pk_autotune_scheduler(f, s, depth, samples, ms), "pk_autotune_scheduler");
  napi_value o;
  napi_create_object(env, &o);
  set_num(env, o, "msPerSampleBefore", ms[0]);
  set_num(env, o, "msPerSampleAfter", ms[1]);
  return o;
}

// ctxSetScheduler(ctx, "wavefront" | "paths"); ctxScheduler(ctx)
napi_value CtxSetScheduler(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, &a, 2)) return nullptr;
  auto* c = (pk_ctx*)unwrap(env, a.argv[0], kCtx, "context");
  if (!c) return nullptr;
  char name[16] = {0};
  size_t len = 0;
  napi_get_value_string_utf8(env, a.argv[1], name, sizeof(name), &len);
  const std::string n(name, len);
  if (n != "wavefront" && n != "paths") {
    napi_throw_error(env, nullptr, "ctxSetScheduler: scheduler must be \"wavefront\" or \"paths\"");
    return nullptr;
  }
  PK_CALL(pk_ctx_set_scheduler(c, n == "paths" ? PK_SCHED_PATHS : PK_SCHED_WAVEFRONT), "pk_ctx_set_scheduler");
  return undefined(env);
}

napi_value CtxScheduler(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, &a, 1)) return nullptr;
  auto* c = (pk_ctx*)unwrap(env, a.argv[0], kCtx, "context");
  if (!c) return nullptr;
  int sched = 0;
  PK_CALL(pk_ctx_get_scheduler(c, &sched), "pk_ctx_get_scheduler");
  napi_value v;
  napi_create_string_utf8(env, sched == PK_SCHED_PATHS ? "paths" : "wavefront", NAPI_AUTO_LENGTH, &v);
  return v;
}

// sceneWriteParams(scene, which, offset, Float32Array, alsoSampling)
napi_value SceneWriteParams(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, &a, 4)) return nullptr;
  SCENE_ARG(0);
  uint32_t which = 0, offset = 0;
  napi_get_value_uint32(env, a.argv[1], &which);
  napi_get_value_uint32(env, a.argv[2], &offset);
  void* data;
  size_t n;
  if (!typed(env, a.argv[3], napi_float32_array, &data, &n, "params")) return nullptr;
  bool also = false;
  if (a.argc > 4) napi_get_value_bool(env, a.argv[4], &also);
  PK_CALL(pk_scene_write_params(s, (int)which, offset, (const float*)data, n, 0, also ? 1 : 0),
          "pk_scene_write_params");
  return undefined(env);
}

// sceneReadParams(scene, which) -> Float32Array of the whole set
napi_value SceneReadParams(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, &a, 2)) return nullptr;
  SCENE_ARG(0);
  uint32_t which = 0;
  napi_get_value_uint32(env, a.argv[1], &which);
  uint32_t pi[6];
  PK_CALL(pk_scene_param_info(s, pi), "pk_scene_param_info");
  const size_t n = which == PK_PARAM_ATLAS ? (size_t)pi[0] * pi[1] * pi[2] * 4 : (size_t)pi[4] * pi[5] * 4;
  void* data;
  napi_value t = make_typed(env, napi_float32_array, n, 4, &data);
  if (!t) return nullptr;
  PK_CALL(pk_scene_read_params(s, (int)which, 0, (float*)data, n), "pk_scene_read_params");
  return t;
}

// adjoint(frame, scene, params, dImage Float32Array) -> {atlas, env} gradients
napi_value Adjoint(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, &a, 4)) return nullptr;
  FRAME_ARG(0);
  SCENE_ARG(1);
  pk_render_params p;
  napi_value o = a.argv[2];
  get_u32(env, o, "width", &p.width, 0);
  get_u32(env, o, "height", &p.height, 0);
  get_u32(env, o, "maxDepth", &p.max_depth, 4);
  get_u32(env, o, "firstSample", &p.first_sample, 0);
  get_u32(env, o, "numFrames", &p.num_frames, 1);
  get_f32(env, o, "envTheta", &p.env_theta, 0.0f);
  p.use_graph = 0;
  void* dimg;
  size_t n;
  if (!typed(env, a.argv[3], napi_float32_array, &dimg, &n, "dImage")) return nullptr;
  if (n < (size_t)p.width * p.height * 4) {
    napi_throw_range_error(env, nullptr, "dImage is smaller than width*height*4");
    return nullptr;
  }
  uint32_t pi[6];
  PK_CALL(pk_scene_param_info(s, pi), "pk_scene_param_info");
  void *ga, *ge;
  napi_value ta = make_typed(env, napi_float32_array, (size_t)pi[0] * pi[1] * pi[2] * 4, 4, &ga);
  napi_value te = make_typed(env, napi_float32_array, (size_t)pi[4] * pi[5] * 4, 4, &ge);
  if (!ta || !te) return nullptr;
  pk_grad_buffers g{(float*)ga, pi[4] ? (float*)ge : nullptr, 0, 0};
  PK_CALL(pk_adjoint(f, s, &p, (const float*)dimg, 0, &g), "pk_adjoint");
  napi_value r;
  napi_create_object(env, &r);
  napi_set_named_property(env, r, "atlas", ta);
  napi_set_named_property(env, r, "env", te);
  return r;
}

napi_value Init(napi_env env, napi_value exports) {
  struct Fn {
    const char* name;
    napi_callback cb;
  } fns[] = {
      {"version", Version},           {"ctxCreate", CtxCreate},
      {"ctxDestroy", CtxDestroy},     {"sync", Sync},
      {"bvhBuild", BvhBuild},         {"bvhInfo", BvhInfo},
      {"bvhExportReference", BvhExportReference}, {"bvhDestroy", BvhDestroy},
      {"sceneUpload", SceneUpload},   {"sceneDestroy", SceneDestroy},
      {"envTables", EnvTables},       {"frameCreate", FrameCreate},
      {"frameDestroy", FrameDestroy}, {"frameReset", FrameReset},
      {"frameSetCamera", FrameSetCamera}, {"frameSetEnvTheta", FrameSetEnvTheta},
      {"frameBegin", FrameBegin},     {"raygen", Raygen},
      {"traceClosest", TraceClosest}, {"shade", Shade},
      {"traceShadow", TraceShadow},   {"accumulate", Accumulate},
      {"render", Render},             {"readImage", ReadImage},
      {"counters", Counters},         {"autotune", Autotune},
      {"traceRays", TraceRays},
      {"frameSetSamplesPerPass", FrameSetSamplesPerPass}, {"autotuneSamples", AutotuneSamples},
      {"autotuneScheduler", AutotuneScheduler}, {"ctxSetScheduler", CtxSetScheduler},
      {"ctxScheduler", CtxScheduler},
      {"sceneWriteParams", SceneWriteParams}, {"sceneReadParams", SceneReadParams},
      {"adjoint", Adjoint},
      {"bvhBuildInstanced", BvhBuildInstanced}, {"bvhNormalizer", BvhNormalizer},
  };
  for (const Fn& f : fns) {
    napi_value v;
    NAPI_OK(napi_create_function(env, f.name, NAPI_AUTO_LENGTH, f.cb, nullptr, &v));
    NAPI_OK(napi_set_named_property(env, exports, f.name, v));
  }
  return exports;
}

}  // namespace

NAPI_MODULE(NODE_GYP_MODULE_NAME, Init)
