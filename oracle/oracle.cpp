// oracle.cpp — TEST INFRASTRUCTURE ONLY (see oracle.h).
//
// A CPU restatement of apbodnar/Pike's render path, templated on the real
// type: T = float is the FP32 restatement that the HIP path must match bit
// for bit (same arithmetic contract, pike_amd/csrc/pk_math.h), T = double is
// the FP64 oracle of the same estimator (std:: math, common random numbers).
// Built with g++ -ffp-contract=off by oracle/Makefile.
#include "oracle.h"

#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>
#include <future>
#include <limits>
#include <memory>
#include <numeric>
#include <thread>
#include <vector>

#include "../pike_amd/csrc/pk_math.h"

namespace {

constexpr double kInf = std::numeric_limits<double>::infinity();

// Planted mistakes for the tests that show each closed-form pin has teeth
// (or_set_mutation; 0 = the faithful restatement).  Each OR_MUT_* names one
// wrong variant of one line below; the test suite checks that the pin aimed
// at it fails while it is planted.
int g_mut = 0;
inline bool mut(int k) { return g_mut == k; }
enum {
  OR_MUT_G1_SWAP = 1,          // Smith G1 takes the microfacet normal's slope instead of w's
  OR_MUT_LAMBERT_INV_PI,       // evalLambert without 1/pi
  OR_MUT_ONB_SIGN,             // branchlessONB: b1.z sign flipped
  OR_MUT_JITTER_ORDER,         // createPrimaryRay draws the y jitter first
  OR_MUT_D_PI,                 // GGX D without pi
  OR_MUT_SCHLICK_POW,          // Schlick (1-cos)^4
  OR_MUT_PH_BALANCE,           // balance heuristic instead of power heuristic
  OR_MUT_LAMBERT_SAMPLE,       // cosine sampling with r = u instead of sqrt(u)
  OR_MUT_VNDF_BLEND,           // sampleGGX: s = 0.5 (1 - vh.z)
  OR_MUT_ENV_PDF_SIN,          // envPdf without the 1/sin(phi) Jacobian
  OR_MUT_ENV_CENTER,           // sampleEnv at the texel corner, not its centre
  OR_MUT_ACC_DENOM,            // accumulate divides by samples + 2
  OR_MUT_NORMAL_MAP,           // normal map xy not scaled by 2
  OR_MUT_ORIGIN_SIGN,          // hit origin pushed past the surface
  OR_MUT_ASPECT,               // camera without the W/H aspect factor
  OR_MUT_SPEC_PDF4,            // specularPdf without the 1/4 Jacobian
  OR_MUT_ENV_COLOR_V,          // envColor v from acos(-y)
  OR_MUT_TEX_HALF,             // bilinear lookup without the half-texel offset
  OR_MUT_EMISSION,             // emission scale 1 instead of 50
  OR_MUT_NEE_SPEC_WEIGHT,      // NEE specular MIS weight against the Lambert pdf
  OR_MUT_BOUNCE_WEIGHT,        // bounce ray without its MIS weight
  OR_MUT_EVAL_SPEC_COS,        // evalSpecular divides by wi.z instead of wo.z
  OR_MUT_SAMPLE_ENV_SIN,       // sampleEnv pdf without 1/sin(phi)
  OR_MUT_METAL_CHANNEL,        // metallic read from G instead of B
  OR_MUT_ALPHA_SQUARE,         // GGX alpha = roughness instead of roughness^2
  OR_MUT_DISTORTION,           // radial distortion r (1 - k r^2)
  OR_MUT_G1_SIGN,              // G1 without the back-facing test
  OR_MUT_FRESNEL_METAL,        // f = mix(schlick, 1, 1 - metallic)
  OR_MUT_THETA_SIGN,           // envColor rotates by -env_theta
  OR_MUT_COUNT
};

// ---------------------------------------------------------------------------
// PCG hash, restated from camera.wgsl:46-56 (identical in shade_hit.wgsl:153-163)
// ---------------------------------------------------------------------------
struct Rng {
  uint32_t seed;
  uint32_t hash() {
    uint32_t state = seed;
    seed = seed * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
  }
  template <class T> T rand();
};
template <> float Rng::rand<float>() { return (float)hash() / 4294967296.0f; }
template <> double Rng::rand<double>() { return (double)hash() / 4294967296.0; }

// ---------------------------------------------------------------------------
// Real-type dispatch: float uses the shared f32 contract, double uses libm.
// ---------------------------------------------------------------------------
template <class T> struct M;
template <> struct M<float> {
  static float sqrt(float x) { return pkm::sqrt_(x); }
  static float sin(float x) { return pkm::sinf_(x); }
  static float cos(float x) { return pkm::cosf_(x); }
  static float tan(float x) { return pkm::tanf_(x); }
  static float atan2(float y, float x) { return pkm::atan2f_(y, x); }
  static float acos(float x) { return pkm::acosf_(x); }
  static float pow(float x, float y) { return pkm::powf_(x, y); }
  static float floor(float x) { return std::floor(x); }
  static constexpr float pi() { return pkm::kPi; }
  static constexpr float tau() { return pkm::kTau; }
  static constexpr float inv_pi() { return pkm::kInvPi; }
};
template <> struct M<double> {
  static double sqrt(double x) { return std::sqrt(x); }
  static double sin(double x) { return std::sin(x); }
  static double cos(double x) { return std::cos(x); }
  static double tan(double x) { return std::tan(x); }
  static double atan2(double y, double x) { return std::atan2(y, x); }
  static double acos(double x) { return std::acos(x); }
  static double pow(double x, double y) { return std::pow(x, y); }
  static double floor(double x) { return std::floor(x); }
  static constexpr double pi() { return 3.141592653589793; }
  static constexpr double tau() { return 6.283185307179586; }
  static constexpr double inv_pi() { return 0.3183098861837907; }
};
// min/max as "b < a ? b : a" — the contract both sides use.
template <class T> inline T mn(T a, T b) { return b < a ? b : a; }
template <class T> inline T mx(T a, T b) { return a < b ? b : a; }

template <class T> struct V3 {
  T x, y, z;
};
template <class T> inline V3<T> v3(T x, T y, T z) { return V3<T>{x, y, z}; }
template <class T> inline V3<T> operator+(V3<T> a, V3<T> b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
template <class T> inline V3<T> operator-(V3<T> a, V3<T> b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
template <class T> inline V3<T> operator*(V3<T> a, V3<T> b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
template <class T> inline V3<T> operator*(V3<T> a, T s) { return v3(a.x * s, a.y * s, a.z * s); }
template <class T> inline V3<T> operator*(T s, V3<T> a) { return v3(s * a.x, s * a.y, s * a.z); }
template <class T> inline V3<T> operator-(V3<T> a) { return v3(-a.x, -a.y, -a.z); }
// WGSL dot, evaluated left to right
template <class T> inline T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> inline V3<T> cross(V3<T> a, V3<T> b) {
  return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
template <class T> inline T length(V3<T> a) { return M<T>::sqrt(dot(a, a)); }
template <class T> inline V3<T> normalize(V3<T> a) {
  T l = length(a);
  return v3(a.x / l, a.y / l, a.z / l);
}
template <class T> inline V3<T> reflect(V3<T> e1, V3<T> e2) {
  T k = T(2) * dot(e2, e1);
  return e1 - e2 * k;
}
template <class T> inline T mix(T a, T b, T t) { return a * (T(1) - t) + b * t; }
template <class T> struct M3 {
  V3<T> c0, c1, c2;
};
template <class T> inline V3<T> mat_vec(const M3<T>& m, V3<T> v) {
  return v3(m.c0.x * v.x + m.c1.x * v.y + m.c2.x * v.z, m.c0.y * v.x + m.c1.y * v.y + m.c2.y * v.z,
            m.c0.z * v.x + m.c1.z * v.y + m.c2.z * v.z);
}
template <class T> inline V3<T> vec_mat(V3<T> v, const M3<T>& m) {
  return v3(dot(v, m.c0), dot(v, m.c1), dot(v, m.c2));
}

// ---------------------------------------------------------------------------
// BVH build: literal restatement of bvh.js:12-43, 80-107, 119-215
// ---------------------------------------------------------------------------
struct Box {
  double lo[3] = {kInf, kInf, kInf};
  double hi[3] = {-kInf, -kInf, -kInf};
};
// Math.min / Math.max on finite values (primitives.js:13-26, vector.js:186-192)
inline void box_add(Box& a, const Box& b) {
  for (int k = 0; k < 3; ++k) {
    a.lo[k] = std::min(a.lo[k], b.lo[k]);
    a.hi[k] = std::max(a.hi[k], b.hi[k]);
  }
}
// primitives.js:35-40
inline double box_area(const Box& b) {
  double xl = std::max(b.hi[0] - b.lo[0], 0.0);
  double yl = std::max(b.hi[1] - b.lo[1], 0.0);
  double zl = std::max(b.hi[2] - b.lo[2], 0.0);
  return (xl * yl + xl * zl + yl * zl) * 2;
}

struct ONode {
  Box bounds;
  uint32_t count = 0;
  bool leaf = false;
  std::vector<uint32_t> leaf_tris;  // triangle ids in leaf order (axisRefs[split.axis])
  std::unique_ptr<ONode> left, right;
};

struct Builder {
  const std::vector<Box>* ref_box;
  const std::vector<std::array<double, 3>>* centroid;
  std::vector<uint8_t> in_left;  // membership flags (stands in for bvh.js:89 Set)
  int degenerate = 0;

  std::unique_ptr<ONode> build(std::array<std::vector<uint32_t>, 3> axis_refs, int depth, int par_depth,
                               int* max_depth) {
    auto node = std::make_unique<ONode>();
    const size_t n = axis_refs[0].size();
    node->count = (uint32_t)n;
    for (uint32_t r : axis_refs[0]) box_add(node->bounds, (*ref_box)[r]);  // Node ctor :123-126
    *max_depth = std::max(*max_depth, depth);
    // #computeSplit (:170-215), full sweep on axes 0,1,2, first minimum wins
    const double parent_area = box_area(node->bounds);
    double best_cost = kInf;
    int best_axis = -1;
    size_t best_index = 0;
    std::vector<double> front(n), back(n);
    for (int axis = 0; axis < 3; ++axis) {
      const std::vector<uint32_t>& refs = axis_refs[axis];
      Box bf, bb;
      for (size_t i = 0; i < n; ++i) {
        box_add(bf, (*ref_box)[refs[i]]);
        box_add(bb, (*ref_box)[refs[n - 1 - i]]);
        front[i] = box_area(bf);
        back[i] = box_area(bb);
      }
      for (size_t i = 0; i < n; ++i) {
        double saf = front[i];
        double sab = back[n - 1 - i];
        double cost = 0.5 + (saf / parent_area) * (double)(i + 1) + (sab / parent_area) * (double)(n - 1 - i);
        if (cost < best_cost) {
          best_cost = cost;
          best_index = i;
          best_axis = axis;
        }
      }
    }
    if (best_axis < 0) {
      // Every cost was NaN (zero parent area): bvh.js throws here
      // (split.axisRefs[-1]); the restatement makes a leaf and records it.
      degenerate = 1;
      node->leaf = true;
      node->leaf_tris = axis_refs[0];
      return node;
    }
    if (best_cost > (double)n) {  // :137-138 -> #setLeaf (:217-223)
      node->leaf = true;
      node->leaf_tris = std::move(axis_refs[best_axis]);
      return node;
    }
    // BVH.#splitNode (:80-107)
    std::array<std::vector<uint32_t>, 3> left, right;
    const std::vector<uint32_t>& split_refs = axis_refs[best_axis];
    left[best_axis].assign(split_refs.begin(), split_refs.begin() + best_index + 1);
    right[best_axis].assign(split_refs.begin() + best_index + 1, split_refs.end());
    for (uint32_t r : left[best_axis]) in_left[r] = 1;
    for (int axis = 0; axis < 3; ++axis) {
      if (axis == best_axis) continue;
      left[axis].reserve(left[best_axis].size());
      right[axis].reserve(right[best_axis].size());
      for (uint32_t r : axis_refs[axis]) (in_left[r] ? left[axis] : right[axis]).push_back(r);
    }
    for (uint32_t r : left[best_axis]) in_left[r] = 0;
    for (auto& v : axis_refs) std::vector<uint32_t>().swap(v);
    int dl = 0, dr = 0;
    if (par_depth > 0 && n > 65536) {
      auto fut = std::async(std::launch::async, [&]() { return build(std::move(left), depth + 1, par_depth - 1, &dl); });
      node->right = build(std::move(right), depth + 1, par_depth - 1, &dr);
      node->left = fut.get();
    } else {
      node->left = build(std::move(left), depth + 1, 0, &dl);
      node->right = build(std::move(right), depth + 1, 0, &dr);
    }
    *max_depth = std::max(*max_depth, std::max(dl, dr));
    return node;
  }
};

// JS ToInt16 (ECMAScript 7.1.8): truncate, then wrap modulo 2^16
inline int16_t to_int16(double x) {
  if (!std::isfinite(x)) return 0;
  double t = std::trunc(x);
  double m = std::fmod(t, 65536.0);
  if (m < 0) m += 65536.0;
  uint32_t u = (uint32_t)m;
  return (int16_t)(uint16_t)u;
}
// trace_pass.js:213-218
inline uint32_t quantize_vec2(double a, double b) {
  int16_t s0 = to_int16(a * (a > 0 ? 32767 : 32768));
  int16_t s1 = to_int16(b * (b > 0 ? 32767 : 32768));
  return (uint32_t)(uint16_t)s0 | ((uint32_t)(uint16_t)s1 << 16);
}

struct RefNode {  // BVHNodeStruct (structs.js:160-171)
  int32_t child, tris;
  uint32_t x, y, z;
};

// Canonical-traversal BVH: binary tree with conservative float64 boxes.
struct CNode {
  double lo[3], hi[3];
  int32_t left = -1, right = -1;  // internal children
  uint32_t first = 0, count = 0;  // leaf range into ctris
};

}  // namespace

// ---------------------------------------------------------------------------
// scene
// ---------------------------------------------------------------------------
struct or_scene {
  uint32_t nv = 0, nt = 0, nm = 0;
  std::vector<double> pos;       // world f64 [nv][3]
  std::vector<uint32_t> attr;    // [nv][4]
  std::vector<int32_t> tri;      // [nt][4]
  std::vector<int32_t> mat;      // [nm][4]
  uint32_t tw = 0, th = 0, tl = 0, tfmt = 0;
  std::vector<uint8_t> tex8;
  std::vector<float> tex32;
  std::vector<float> tex32_eval;  // evaluation atlas (empty: the sampling atlas)
  uint32_t ew = 0, eh = 0, efmt = 0, esampler = 0;
  std::vector<uint8_t> env8;
  std::vector<float> env32;
  // bvh
  std::unique_ptr<ONode> root;
  int64_t bvh_nodes = 0, bvh_leaf_tris = 0, bvh_depth = 0, bvh_largest_leaf = 0;
  int degenerate = 0;
  // serialized reference layout
  std::vector<RefNode> rnodes;
  std::vector<int32_t> rtris;   // leaf order [n][4]
  std::vector<uint32_t> rleaf_to_orig;
  std::vector<float> vf;        // normalized f32 [nv][3]
  std::vector<double> vd;       // normalized f64 [nv][3]
  double norm_center[3] = {0, 0, 0}, norm_scale = 1;
  // canonical BVH (f32 boxes, f64 boxes)
  std::vector<CNode> cnodes32, cnodes64;
  std::vector<uint32_t> ctris;  // original ids in leaf order
  // env
  std::vector<std::array<int32_t, 2>> ebins, ecoords;
  std::vector<float> epdf32;
  std::vector<double> epdf64;
  std::vector<float> erad32;    // decoded radiance rgb
  std::vector<double> erad64;
  // marginal (rows) / conditional (columns) CDF tables of the CDF sampler
  std::vector<float> row_cdf, col_cdf;
  // instancing: per-mesh sub-scenes (the scene's own geometry first, when it
  // has triangles) and the placements, each with its scene -> mesh map
  struct Inst {
    uint32_t mesh = 0;
    int32_t material = -1;
    double t[12] = {0};
    double w2o64[12] = {0}, o2w64[9] = {0};
    float w2o32[12] = {0}, o2w32[9] = {0};
    int32_t id_base = 0;
  };
  bool instanced = false, built = false;
  std::vector<std::unique_ptr<or_scene>> meshes;  // [0]: the scene's own triangles, when it has any
  std::vector<Inst> insts;                        // the placements (of the other meshes)
  int32_t own = -1;                               // index of the scene's own triangles in meshes
  // a normalizer imposed from outside (the scene's, for its own triangles)
  bool force_norm = false;
  double force_center[3] = {0, 0, 0}, force_scale = 1;
};

namespace {

// makeNormalizer (trace_pass.js:202-211)
void make_normalizer(or_scene* s) {
  const Box& b = s->root->bounds;
  double span[3], longest;
  for (int k = 0; k < 3; ++k) span[k] = b.hi[k] - b.lo[k];
  longest = std::max(std::max(span[0], span[1]), span[2]);
  for (int k = 0; k < 3; ++k) s->norm_center[k] = b.lo[k] + span[k] * 0.5;
  s->norm_scale = 2 / longest;
  if (s->force_norm) {
    for (int k = 0; k < 3; ++k) s->norm_center[k] = s->force_center[k];
    s->norm_scale = s->force_scale;
  }
  s->vf.resize((size_t)s->nv * 3);
  s->vd.resize((size_t)s->nv * 3);
  for (size_t i = 0; i < (size_t)s->nv * 3; ++i) {
    double v = (s->pos[i] - s->norm_center[i % 3]) * s->norm_scale;
    s->vd[i] = v;
    s->vf[i] = (float)v;
  }
}

// BFS flatten (bvh.js:46-74) + pack (trace_pass.js:233-259)
void serialize(or_scene* s) {
  std::vector<const ONode*> order;
  std::vector<int32_t> left_index;
  std::deque<const ONode*> q;
  q.push_back(s->root.get());
  while (!q.empty()) {
    const ONode* n = q.front();
    q.pop_front();
    order.push_back(n);
    if (!n->leaf) {
      q.push_back(n->left.get());
      q.push_back(n->right.get());
    }
  }
  // children were enqueued in pairs, so the left child of the k-th interior
  // node in BFS order sits at 2k+1 (odd index), the right one at 2k+2.
  s->rnodes.resize(order.size());
  s->rtris.clear();
  s->rleaf_to_orig.clear();
  int32_t interior = 0;
  int32_t tri_index = 0;
  for (size_t i = 0; i < order.size(); ++i) {
    const ONode* n = order[i];
    double lo[3], hi[3];
    for (int k = 0; k < 3; ++k) {
      lo[k] = (n->bounds.lo[k] - s->norm_center[k]) * s->norm_scale;
      hi[k] = (n->bounds.hi[k] - s->norm_center[k]) * s->norm_scale;
    }
    RefNode& r = s->rnodes[i];
    if (n->leaf) {
      r.child = -1;
      // maskTriIndex: (numTris << 24) | index in int32 arithmetic
      r.tris = (int32_t)(((uint32_t)n->leaf_tris.size() << 24) | (uint32_t)tri_index);
      for (uint32_t t : n->leaf_tris) {
        for (int c = 0; c < 4; ++c) s->rtris.push_back(s->tri[(size_t)t * 4 + c]);
        s->rleaf_to_orig.push_back(t);
      }
      tri_index += (int32_t)n->leaf_tris.size();
    } else {
      r.child = 2 * interior + 1;
      r.tris = -1;
      ++interior;
    }
    r.x = quantize_vec2(hi[0], lo[0]);
    r.y = quantize_vec2(hi[1], lo[1]);
    r.z = quantize_vec2(hi[2], lo[2]);
  }
}

// Conservative BVH for the canonical closest hit: same tree, boxes recomputed
// from the normalized vertices the intersector uses, widened by 1e-6.
template <class VT>
void build_canonical(const or_scene* s, const std::vector<VT>& verts, std::vector<CNode>& out,
                     std::vector<uint32_t>* leaf_tris) {
  out.clear();
  if (leaf_tris) leaf_tris->clear();
  struct Item {
    const ONode* n;
    int32_t idx;
  };
  std::vector<Item> stack;
  out.push_back(CNode());
  stack.push_back({s->root.get(), 0});
  std::vector<uint32_t> local;
  uint32_t tri_cursor = 0;
  // first pass: structure (DFS, leaves get consecutive triangle ranges)
  std::vector<std::pair<int32_t, const ONode*>> leaves;
  while (!stack.empty()) {
    Item it = stack.back();
    stack.pop_back();
    if (it.n->leaf) {
      out[it.idx].first = tri_cursor;
      out[it.idx].count = (uint32_t)it.n->leaf_tris.size();
      tri_cursor += out[it.idx].count;
      leaves.push_back({it.idx, it.n});
      if (leaf_tris)
        for (uint32_t t : it.n->leaf_tris) leaf_tris->push_back(t);
    } else {
      int32_t l = (int32_t)out.size();
      out.push_back(CNode());
      int32_t r = (int32_t)out.size();
      out.push_back(CNode());
      out[it.idx].left = l;
      out[it.idx].right = r;
      stack.push_back({it.n->right.get(), r});
      stack.push_back({it.n->left.get(), l});
    }
  }
  // boxes bottom-up: children always have larger indices than parents
  for (auto& c : out)
    for (int k = 0; k < 3; ++k) {
      c.lo[k] = kInf;
      c.hi[k] = -kInf;
    }
  for (auto& lf : leaves) {
    CNode& c = out[lf.first];
    for (uint32_t t : lf.second->leaf_tris)
      for (int v = 0; v < 3; ++v) {
        uint32_t vi = (uint32_t)s->tri[(size_t)t * 4 + v];
        for (int k = 0; k < 3; ++k) {
          double p = (double)verts[(size_t)vi * 3 + k];
          c.lo[k] = std::min(c.lo[k], p);
          c.hi[k] = std::max(c.hi[k], p);
        }
      }
  }
  for (int32_t i = (int32_t)out.size() - 1; i >= 0; --i) {
    CNode& c = out[i];
    if (c.left >= 0) {
      for (int k = 0; k < 3; ++k) {
        c.lo[k] = std::min(out[c.left].lo[k], out[c.right].lo[k]);
        c.hi[k] = std::max(out[c.left].hi[k], out[c.right].hi[k]);
      }
    }
  }
  for (auto& c : out)
    for (int k = 0; k < 3; ++k) {
      c.lo[k] -= 1e-6;
      c.hi[k] += 1e-6;
    }
}

// ---------------------------------------------------------------------------
// Ray/triangle (trace.wgsl:97-114; trace_shadow.wgsl uses EPSILON 1e-7)
// ---------------------------------------------------------------------------
template <class T>
inline T ray_triangle(V3<T> o, V3<T> d, V3<T> p0, V3<T> p1, V3<T> p2, T eps, T max_t, T* uo, T* vo, uint32_t* bary) {
  V3<T> e1 = p1 - p0;
  V3<T> e2 = p2 - p0;
  V3<T> p = cross(d, e2);
  T det = dot(e1, p);
  if ((det < 0 ? -det : det) < eps) return max_t;
  T inv_det = T(1) / det;
  V3<T> tv = o - p0;
  T u = dot(tv, p) * inv_det;
  if (u < T(0) || u > T(1)) return max_t;
  V3<T> q = cross(tv, e1);
  T v = dot(d, q) * inv_det;
  if (v < T(0) || u + v > T(1)) return max_t;
  T dist = dot(e2, q) * inv_det;
  if (bary) {
    // pack2x16unorm
    T cu = mn(mx(u, T(0)), T(1)), cv = mn(mx(v, T(0)), T(1));
    uint32_t lo = (uint32_t)(T(0.5) + T(65535) * cu);
    uint32_t hi = (uint32_t)(T(0.5) + T(65535) * cv);
    *bary = lo | (hi << 16);
  }
  *uo = u;
  *vo = v;
  return dist > eps ? dist : max_t;
}

template <class T> inline V3<T> vert(const or_scene* s, uint32_t vi);
template <> inline V3<float> vert<float>(const or_scene* s, uint32_t vi) {
  return v3(s->vf[vi * 3 + 0], s->vf[vi * 3 + 1], s->vf[vi * 3 + 2]);
}
template <> inline V3<double> vert<double>(const or_scene* s, uint32_t vi) {
  return v3(s->vd[vi * 3 + 0], s->vd[vi * 3 + 1], s->vd[vi * 3 + 2]);
}

template <class T> struct HitT {
  int32_t id = -1;  // original triangle id (baked numbering with instancing)
  T t;
  uint32_t bary = 0;
  int32_t inst = -1, local = -1;  // instancing: placement and mesh triangle
};

// Literal trace.wgsl intersectScene (:148-189) over the 20-byte nodes.
struct LitResult {
  int32_t index = -1;  // leaf-order index
  float t = 1e5f;
  uint32_t bary = 0;
  bool overflow = false;
};
inline float unpack_snorm16_lo(uint32_t w) { return pkm::unpack_snorm16(w & 0xffffu); }
inline float unpack_snorm16_hi(uint32_t w) { return pkm::unpack_snorm16(w >> 16); }
inline float lit_box(const RefNode& n, V3<float> o, V3<float> d) {
  const float kMaxT = 100000.0f;
  V3<float> inv = v3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
  V3<float> r1 = v3(unpack_snorm16_lo(n.x), unpack_snorm16_lo(n.y), unpack_snorm16_lo(n.z));
  V3<float> r2 = v3(unpack_snorm16_hi(n.x), unpack_snorm16_hi(n.y), unpack_snorm16_hi(n.z));
  V3<float> t1 = (r1 - o) * inv;
  V3<float> t2 = (r2 - o) * inv;
  V3<float> mnv = v3(mn(t1.x, t2.x), mn(t1.y, t2.y), mn(t1.z, t2.z));
  V3<float> mxv = v3(mx(t1.x, t2.x), mx(t1.y, t2.y), mx(t1.z, t2.z));
  float tmax = mn(mn(mxv.x, mxv.y), mxv.z);
  float tmin = mx(mx(mnv.x, mnv.y), mnv.z);
  return (tmax >= tmin && tmax > 0.0f) ? tmin : kMaxT;
}
LitResult literal_trace(const or_scene* s, V3<float> o, V3<float> d, bool any_hit) {
  const float eps = any_hit ? 0.0000001f : 0.00000001f;
  const float kMaxT = 100000.0f;
  LitResult res;
  int32_t stack[32];
  int sptr = 0;
  auto push = [&](int32_t v) {
    if (sptr >= 32) res.overflow = true;
    stack[std::min(sptr, 31)] = v;  // robust-access clamp
    ++sptr;
  };
  auto pop = [&]() -> int32_t {
    --sptr;
    return stack[std::max(0, std::min(sptr, 31))];
  };
  push(-1);
  int32_t idx = 0;
  while (true) {
    if (idx <= -1) break;
    const RefNode& cur = s->rnodes[idx];
    if (cur.tris > -1) {
      int32_t size = cur.tris >> 24;
      int32_t base = cur.tris & 0x00ffffff;
      for (int32_t i = 0; i < size; ++i) {
        const int32_t* tr = &s->rtris[(size_t)(base + i) * 4];
        float u, v;
        uint32_t bm = 0;
        float r = ray_triangle<float>(o, d, vert<float>(s, tr[0]), vert<float>(s, tr[1]), vert<float>(s, tr[2]), eps,
                                      kMaxT, &u, &v, &bm);
        if (r < res.t) {
          res.index = base + i;
          res.t = r;
          res.bary = bm;
        }
      }
      if (any_hit && res.index != -1) return res;
    } else {
      int32_t li = cur.child, ri = cur.child + 1;
      float lh = lit_box(s->rnodes[li], o, d);
      float rh = lit_box(s->rnodes[ri], o, d);
      if (lh < res.t && rh < res.t) {
        int32_t deferred;
        if (lh > rh) {
          idx = ri;
          deferred = li;
        } else {
          idx = li;
          deferred = ri;
        }
        push(deferred);
        continue;
      } else {
        if (lh < res.t) {
          idx = li;
          continue;
        }
        if (rh < res.t) {
          idx = ri;
          continue;
        }
      }
    }
    idx = pop();
  }
  return res;
}

// Canonical closest hit: argmin over (t, original id) with t < MAX_T.
template <class T>
HitT<T> canonical_trace(const or_scene* s, const std::vector<CNode>& nodes, V3<T> o, V3<T> d, bool any_hit) {
  const T eps = any_hit ? T(0.0000001f) : T(0.00000001f);
  const T kMaxT = T(100000.0);
  HitT<T> best;
  best.t = kMaxT;
  double od[3] = {(double)o.x, (double)o.y, (double)o.z};
  double inv[3];
  double dd[3] = {(double)d.x, (double)d.y, (double)d.z};
  for (int k = 0; k < 3; ++k) {
    double c = dd[k];
    if (std::fabs(c) < 1e-30) c = std::copysign(1e-30, c);
    inv[k] = 1.0 / c;
  }
  auto box = [&](const CNode& c, double* tnear) -> bool {
    double t0 = -kInf, t1 = kInf;
    for (int k = 0; k < 3; ++k) {
      double a = (c.lo[k] - od[k]) * inv[k], b = (c.hi[k] - od[k]) * inv[k];
      if (a > b) std::swap(a, b);
      t0 = std::max(t0, a);
      t1 = std::min(t1, b);
    }
    *tnear = t0;
    if (t1 < t0 * (1 + 1e-9) - 1e-9) return false;
    if (t1 < -1e-6) return false;
    double limit = (double)best.t * (1 + 1e-3) + 1e-6;
    return t0 <= limit;
  };
  std::vector<int32_t> stack;
  stack.reserve(64);
  double tn;
  if (!box(nodes[0], &tn)) return best;
  stack.push_back(0);
  while (!stack.empty()) {
    int32_t ni = stack.back();
    stack.pop_back();
    const CNode& c = nodes[ni];
    if (c.left < 0) {
      for (uint32_t k = 0; k < c.count; ++k) {
        uint32_t id = s->ctris[c.first + k];
        const int32_t* tr = &s->tri[(size_t)id * 4];
        T u, v;
        uint32_t bm = 0;
        T r = ray_triangle<T>(o, d, vert<T>(s, tr[0]), vert<T>(s, tr[1]), vert<T>(s, tr[2]), eps, kMaxT, &u, &v, &bm);
        if (r < kMaxT && (r < best.t || (r == best.t && (int32_t)id < best.id))) {
          best.t = r;
          best.id = (int32_t)id;
          best.bary = bm;
          if (any_hit) return best;
        }
      }
      continue;
    }
    double tl, tr;
    bool hl = box(nodes[c.left], &tl), hr = box(nodes[c.right], &tr);
    if (hl && hr) {
      if (tl <= tr) {
        stack.push_back(c.right);
        stack.push_back(c.left);
      } else {
        stack.push_back(c.left);
        stack.push_back(c.right);
      }
    } else if (hl) {
      stack.push_back(c.left);
    } else if (hr) {
      stack.push_back(c.right);
    }
  }
  return best;
}

template <class T> HitT<T> brute_trace(const or_scene* s, V3<T> o, V3<T> d, bool any_hit) {
  const T eps = any_hit ? T(0.0000001f) : T(0.00000001f);
  const T kMaxT = T(100000.0);
  HitT<T> best;
  best.t = kMaxT;
  for (uint32_t id = 0; id < s->nt; ++id) {
    const int32_t* tr = &s->tri[(size_t)id * 4];
    T u, v;
    uint32_t bm = 0;
    T r = ray_triangle<T>(o, d, vert<T>(s, tr[0]), vert<T>(s, tr[1]), vert<T>(s, tr[2]), eps, kMaxT, &u, &v, &bm);
    if (r < best.t) {
      best.t = r;
      best.id = (int32_t)id;
      best.bary = bm;
      if (any_hit) return best;
    }
  }
  return best;
}

template <class T> const std::vector<CNode>& nodes_of(const or_scene* s);
template <> const std::vector<CNode>& nodes_of<float>(const or_scene* s) { return s->cnodes32; }
template <> const std::vector<CNode>& nodes_of<double>(const or_scene* s) { return s->cnodes64; }

// Scene -> mesh map of a placement, in T: each output is ((a x + b y) + c z)
// (+ d for points), the order the device evaluates.
template <class T> const T* w2o_of(const or_scene::Inst& I);
template <> const float* w2o_of<float>(const or_scene::Inst& I) { return I.w2o32; }
template <> const double* w2o_of<double>(const or_scene::Inst& I) { return I.w2o64; }
template <class T> const T* o2w_of(const or_scene::Inst& I);
template <> const float* o2w_of<float>(const or_scene::Inst& I) { return I.o2w32; }
template <> const double* o2w_of<double>(const or_scene::Inst& I) { return I.o2w64; }

template <class T> V3<T> map_point(const T* m, V3<T> p) {
  return v3(((m[0] * p.x + m[1] * p.y) + m[2] * p.z) + m[3], ((m[4] * p.x + m[5] * p.y) + m[6] * p.z) + m[7],
            ((m[8] * p.x + m[9] * p.y) + m[10] * p.z) + m[11]);
}
template <class T> V3<T> map_dir(const T* m, V3<T> p) {
  return v3((m[0] * p.x + m[1] * p.y) + m[2] * p.z, (m[4] * p.x + m[5] * p.y) + m[6] * p.z,
            (m[8] * p.x + m[9] * p.y) + m[10] * p.z);
}

// Closest (or any) hit over the whole scene.  With instancing: argmin over
// (t, baked id) of every placement's canonical hit, each found with the ray
// mapped into that mesh's space (t is the same parameter in both spaces).
// No placement is culled by its box, so the minimum is exact.
template <class T> HitT<T> scene_trace(const or_scene* s, V3<T> o, V3<T> d, bool any_hit) {
  if (!s->instanced) return canonical_trace<T>(s, nodes_of<T>(s), o, d, any_hit);
  HitT<T> best;
  best.t = T(100000.0);
  if (s->own >= 0) {  // the scene's own triangles, in the scene's space
    const or_scene* w = s->meshes[(size_t)s->own].get();
    best = canonical_trace<T>(w, nodes_of<T>(w), o, d, any_hit);
    best.local = best.id;
    if (any_hit && best.id >= 0) return best;
  }
  for (size_t i = 0; i < s->insts.size(); ++i) {
    const or_scene::Inst& I = s->insts[i];
    const or_scene* m = s->meshes[I.mesh].get();
    const T* w = w2o_of<T>(I);
    HitT<T> h = canonical_trace<T>(m, nodes_of<T>(m), map_point<T>(w, o), map_dir<T>(w, d), any_hit);
    if (h.id < 0) continue;
    const int32_t gid = h.id + I.id_base;
    if (h.t < best.t || (h.t == best.t && gid < best.id)) {
      best = h;
      best.id = gid;
      best.inst = (int32_t)i;
      best.local = h.id;
      if (any_hit) return best;
    }
  }
  return best;
}

template <class T> HitT<T> scene_brute(const or_scene* s, V3<T> o, V3<T> d, bool any_hit) {
  if (!s->instanced) return brute_trace<T>(s, o, d, any_hit);
  HitT<T> best;
  best.t = T(100000.0);
  if (s->own >= 0) {
    best = brute_trace<T>(s->meshes[(size_t)s->own].get(), o, d, any_hit);
    best.local = best.id;
    if (any_hit && best.id >= 0) return best;
  }
  for (size_t i = 0; i < s->insts.size(); ++i) {
    const or_scene::Inst& I = s->insts[i];
    const T* w = w2o_of<T>(I);
    HitT<T> h = brute_trace<T>(s->meshes[I.mesh].get(), map_point<T>(w, o), map_dir<T>(w, d), any_hit);
    if (h.id < 0) continue;
    const int32_t gid = h.id + I.id_base;
    if (h.t < best.t || (h.t == best.t && gid < best.id)) {
      best = h;
      best.id = gid;
      best.inst = (int32_t)i;
      best.local = h.id;
      if (any_hit) return best;
    }
  }
  return best;
}

template <class F> void parallel_for(int64_t n, int threads, F f) {
  if (threads <= 1 || n < 64) {
    for (int64_t i = 0; i < n; ++i) f(i);
    return;
  }
  std::vector<std::thread> ts;
  std::atomic<int64_t> next(0);
  for (int t = 0; t < threads; ++t)
    ts.emplace_back([&]() {
      while (true) {
        int64_t b = next.fetch_add(64);
        if (b >= n) break;
        int64_t e = std::min(n, b + 64);
        for (int64_t i = b; i < e; ++i) f(i);
      }
    });
  for (auto& t : ts) t.join();
}

// ---------------------------------------------------------------------------
// Environment tables (env_sampler.js:31-157)
// ---------------------------------------------------------------------------
void build_env(or_scene* s) {
  const uint32_t W = s->ew, H = s->eh;
  const size_t T = (size_t)W * H;
  s->erad32.assign(T * 3, 0.0f);
  s->erad64.assign(T * 3, 0.0);
  std::vector<double> rad(T);
  for (size_t i = 0; i < T; ++i) {
    if (s->efmt == PK_ENV_RGBE8) {
      const uint8_t* c = &s->env8[i * 4];
      // _getRadiance (:43-48): 2^(a-128) * luma(rgb) / 255 on 0..255 values
      double power = std::pow(2.0, (double)c[3] - 128);
      double luma = 0.2126 * c[0] + 0.7152 * c[1] + 0.0722 * c[2];
      rad[i] = power * luma / 255.0;
      // shade_miss.wgsl:41-42 decode: rgb * 2^(a*255 - 128) on unorm values
      float a = (float)c[3] / 255.0f;
      float p2 = pkm::exp2f_(a * 255.0f - 128.0f);
      for (int k = 0; k < 3; ++k) {
        s->erad32[i * 3 + k] = ((float)c[k] / 255.0f) * p2;
        s->erad64[i * 3 + k] = ((double)c[k] / 255.0) * std::ldexp(1.0, (int)c[3] - 128);
      }
    } else {
      const float* c = &s->env32[i * 3];
      rad[i] = 0.2126 * (double)c[0] + 0.7152 * (double)c[1] + 0.0722 * (double)c[2];
      for (int k = 0; k < 3; ++k) {
        s->erad32[i * 3 + k] = c[k];
        s->erad64[i * 3 + k] = (double)c[k];
      }
    }
  }
  // _sortPixels (:59-72): row-major scan, running total/max, stable descending sort
  double total = 0, brightest = 0;
  std::vector<uint32_t> order(T);
  for (size_t i = 0; i < T; ++i) {
    brightest = std::max(rad[i], brightest);
    total += rad[i];
    order[i] = (uint32_t)i;
  }
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return rad[b] < rad[a]; });
  // createLuminanceHistogram (:74-94)
  // Math.min/Math.max propagate NaN (an all-black map): then no bin ever
  // closes and the single bin spans every texel.
  double ratio = total / brightest;
  double bin_size = std::isnan(ratio) ? ratio : total / std::min(256.0, std::max(32.0, 2 * std::ceil(ratio)));
  s->ebins.clear();
  double current = 0;
  for (size_t i = 0; i < T; ++i) {
    if (current >= bin_size) {
      int32_t prev = s->ebins.empty() ? 0 : s->ebins.back()[1];
      s->ebins.push_back({prev, (int32_t)i});
      current = 0;
    }
    current += rad[order[i]];
  }
  int32_t last = s->ebins.empty() ? 0 : s->ebins.back()[1];  // bvh.js would throw on empty
  s->ebins.push_back({last, (int32_t)T});
  s->ecoords.resize(T);
  for (size_t i = 0; i < T; ++i) s->ecoords[i] = {(int32_t)(order[i] % W), (int32_t)(order[i] / W)};
  // createPdfTexture (:124-157)
  s->epdf32.assign(T, 0.0f);
  s->epdf64.assign(T, 0.0);
  const double nb = (double)s->ebins.size();
  for (auto& b : s->ebins) {
    for (int32_t i = b[0]; i < b[1]; ++i) {
      const auto& c = s->ecoords[i];
      size_t idx = (size_t)c[0] + (size_t)W * c[1];
      double bin_pdf = 1 / nb;
      double coord_pdf = (double)T / (double)(b[1] - b[0]);
      double sphere_pdf = 1 / (3.141592653589793 * 3.141592653589793 * 2);
      double pdf = bin_pdf * coord_pdf * sphere_pdf;
      s->epdf64[idx] = pdf;
      s->epdf32[idx] = (float)pdf;
    }
  }
  if (s->esampler != PK_ENV_SAMPLER_CDF) return;
  // CDF sampler (the product's device binary-search table; no reference
  // counterpart): texel weight = luminance * sin(theta at the row center);
  // rows by their sums, columns within a row by the weights; a texel's pdf is
  // the product of its two f32 CDF steps times T / (2 pi^2) (solid angle).
  const double pi = 3.141592653589793;
  std::vector<double> wgt(T), rows(H, 0.0);
  for (uint32_t y = 0; y < H; ++y) {
    const double sy = std::sin(pi * ((double)y + 0.5) / (double)H);
    for (uint32_t x = 0; x < W; ++x) {
      wgt[(size_t)y * W + x] = rad[(size_t)y * W + x] * sy;
      rows[y] += wgt[(size_t)y * W + x];
    }
  }
  const double all = std::accumulate(rows.begin(), rows.end(), 0.0);
  s->row_cdf.assign(H + 1, 0.0f);
  s->col_cdf.assign((size_t)H * (W + 1), 0.0f);
  double below = 0;
  for (uint32_t y = 0; y < H; ++y) {
    s->row_cdf[y] = (float)(below / all);
    below += rows[y];
    float* cc = &s->col_cdf[(size_t)y * (W + 1)];
    double left = 0;
    for (uint32_t x = 0; x < W; ++x) {
      cc[x] = rows[y] > 0 ? (float)(left / rows[y]) : (float)x / (float)W;
      left += wgt[(size_t)y * W + x];
    }
    cc[W] = 1.0f;
  }
  s->row_cdf[H] = 1.0f;
  for (uint32_t y = 0; y < H; ++y) {
    const double py = (double)s->row_cdf[y + 1] - (double)s->row_cdf[y];
    const float* cc = &s->col_cdf[(size_t)y * (W + 1)];
    for (uint32_t x = 0; x < W; ++x) {
      const double px = (double)cc[x + 1] - (double)cc[x];
      const double pdf = py * px * (double)T / (2 * pi * pi);
      s->epdf64[(size_t)y * W + x] = pdf;
      s->epdf32[(size_t)y * W + x] = (float)pdf;
    }
  }
}

// Forward-mode dual number for the reference gradients: a value and its
// derivatives with respect to up to kDualN seeded parameters.  S = double is
// the FP64 reference; S = float keeps the value arithmetic of the FP32
// restatement (so its paths are bit-identical to the product's).
constexpr int kDualN = 16;
template <class S> struct DualT {
  S v = 0;
  S g[kDualN] = {};
  DualT() = default;
  DualT(S x) : v(x) {}  // NOLINT: constants and real values convert implicitly
};
template <class S> inline DualT<S> operator+(const DualT<S>& a, const DualT<S>& b) {
  DualT<S> r(a.v + b.v);
  for (int i = 0; i < kDualN; ++i) r.g[i] = a.g[i] + b.g[i];
  return r;
}
template <class S> inline DualT<S> operator-(const DualT<S>& a, const DualT<S>& b) {
  DualT<S> r(a.v - b.v);
  for (int i = 0; i < kDualN; ++i) r.g[i] = a.g[i] - b.g[i];
  return r;
}
template <class S> inline DualT<S> operator-(const DualT<S>& a) {
  DualT<S> r(-a.v);
  for (int i = 0; i < kDualN; ++i) r.g[i] = -a.g[i];
  return r;
}
template <class S> inline DualT<S> operator*(const DualT<S>& a, const DualT<S>& b) {
  DualT<S> r(a.v * b.v);
  for (int i = 0; i < kDualN; ++i) r.g[i] = a.g[i] * b.v + a.v * b.g[i];
  return r;
}
template <class S> inline DualT<S> operator/(const DualT<S>& a, const DualT<S>& b) {
  DualT<S> r(a.v / b.v);
  for (int i = 0; i < kDualN; ++i) r.g[i] = (a.g[i] - r.v * b.g[i]) / b.v;
  return r;
}
// mixed real / dual operands (template deduction does not convert)
template <class S> inline DualT<S> operator+(S a, const DualT<S>& b) { return DualT<S>(a) + b; }
template <class S> inline DualT<S> operator+(const DualT<S>& a, S b) { return a + DualT<S>(b); }
template <class S> inline DualT<S> operator-(S a, const DualT<S>& b) { return DualT<S>(a) - b; }
template <class S> inline DualT<S> operator-(const DualT<S>& a, S b) { return a - DualT<S>(b); }
template <class S> inline DualT<S> operator*(S a, const DualT<S>& b) { return DualT<S>(a) * b; }
template <class S> inline DualT<S> operator*(const DualT<S>& a, S b) { return a * DualT<S>(b); }
template <class S> inline DualT<S> operator/(S a, const DualT<S>& b) { return DualT<S>(a) / b; }
template <class S> inline DualT<S> operator/(const DualT<S>& a, S b) { return a / DualT<S>(b); }
template <class S> inline DualT<S>& operator+=(DualT<S>& a, const DualT<S>& b) { return a = a + b; }
template <class S> inline bool operator<(const DualT<S>& a, const DualT<S>& b) { return a.v < b.v; }
template <class S> inline bool operator<=(const DualT<S>& a, S b) { return a.v <= b; }
template <class S> inline bool operator>(const DualT<S>& a, const DualT<S>& b) { return a.v > b.v; }
// square root of a value type: the f32 contract, libm, or the dual rule
inline float usqrt(float x) { return M<float>::sqrt(x); }
inline double usqrt(double x) { return std::sqrt(x); }
template <class S> inline DualT<S> usqrt(const DualT<S>& a) {
  DualT<S> r(usqrt(a.v));
  const S k = r.v > S(0) ? S(0.5) / r.v : S(0);
  for (int i = 0; i < kDualN; ++i) r.g[i] = a.g[i] * k;
  return r;
}
// mixed real / dual vector arithmetic of the path loop
template <class S> inline V3<DualT<S>> operator*(V3<S> a, V3<DualT<S>> b) {
  return v3(DualT<S>(a.x) * b.x, DualT<S>(a.y) * b.y, DualT<S>(a.z) * b.z);
}
template <class S> inline V3<DualT<S>> operator*(V3<DualT<S>> a, S s) {
  return v3(a.x * DualT<S>(s), a.y * DualT<S>(s), a.z * DualT<S>(s));
}

// Parameters whose derivatives a dual evaluation carries: float offsets into
// the PK_PARAM_ATLAS / PK_PARAM_ENV layouts mapped to dual slots.
struct Seeds {
  std::vector<std::pair<int64_t, int>> atlas, env;  // sorted by offset
  static int find(const std::vector<std::pair<int64_t, int>>& v, int64_t off) {
    auto it = std::lower_bound(v.begin(), v.end(), std::make_pair(off, -1));
    return it != v.end() && it->first == off ? it->second : -1;
  }
};

// ---------------------------------------------------------------------------
// Shading (shade_hit.wgsl), camera (camera.wgsl), miss (shade_miss.wgsl)
//
// T is the real type of geometry and sampling; U is the type of the values
// that depend on the differentiable parameters (base color, GGX evaluation
// roughness, environment radiance) and of everything they multiply (BSDF
// weights, throughput, radiance).  U = T renders; U = Dual differentiates.
// Sampling reads the sampling atlas, evaluation the evaluation atlas
// (detached sampling, as in the product).
// ---------------------------------------------------------------------------
template <class T> struct Ctx {
  const or_scene* s;
  const std::vector<CNode>* nodes;
  T env_theta;
  bool sample_env;
  const Seeds* seeds;
};

template <class T> inline T texel(const or_scene* s, const std::vector<float>& t32, int layer, int x, int y, int c) {
  size_t i = (((size_t)layer * s->th + (size_t)y) * s->tw + (size_t)x) * 4 + c;
  if (s->tfmt == PK_TEX_RGBA8) return (T)((float)s->tex8[i] / 255.0f);
  return (T)t32[i];
}
template <>
inline double texel<double>(const or_scene* s, const std::vector<float>& t32, int layer, int x, int y, int c) {
  size_t i = (((size_t)layer * s->th + (size_t)y) * s->tw + (size_t)x) * 4 + c;
  if (s->tfmt == PK_TEX_RGBA8) return (double)s->tex8[i] / 255.0;
  return (double)t32[i];
}
inline int wrap(int i, int n) {
  int r = i % n;
  return r < 0 ? r + n : r;
}
// textureSampleLevel with a linear filter and repeat addressing, level 0,
// of the sampling atlas
template <class T> void sample_tex(const or_scene* s, int layer, T u, T v, T out[4]) {
  const T half = mut(OR_MUT_TEX_HALF) ? T(0) : T(0.5);
  T x = u * (T)s->tw - half, y = v * (T)s->th - half;
  T fx0 = M<T>::floor(x), fy0 = M<T>::floor(y);
  T fx = x - fx0, fy = y - fy0;
  int x0 = wrap((int)fx0, (int)s->tw), x1 = wrap((int)fx0 + 1, (int)s->tw);
  int y0 = wrap((int)fy0, (int)s->th), y1 = wrap((int)fy0 + 1, (int)s->th);
  for (int c = 0; c < 4; ++c) {
    T a = texel<T>(s, s->tex32, layer, x0, y0, c), b = texel<T>(s, s->tex32, layer, x1, y0, c);
    T cc = texel<T>(s, s->tex32, layer, x0, y1, c), dd = texel<T>(s, s->tex32, layer, x1, y1, c);
    T top = a + (b - a) * fx;
    T bot = cc + (dd - cc) * fx;
    out[c] = top + (bot - top) * fy;
  }
}
// evaluation texel as a value type (seeded when differentiating)
template <class T, class U> U eval_texel(const Ctx<T>& c, int layer, int x, int y, int ch);
template <> float eval_texel<float, float>(const Ctx<float>& c, int layer, int x, int y, int ch) {
  const or_scene* s = c.s;
  return texel<float>(s, s->tex32_eval.empty() ? s->tex32 : s->tex32_eval, layer, x, y, ch);
}
template <> double eval_texel<double, double>(const Ctx<double>& c, int layer, int x, int y, int ch) {
  const or_scene* s = c.s;
  return texel<double>(s, s->tex32_eval.empty() ? s->tex32 : s->tex32_eval, layer, x, y, ch);
}
template <class T>
DualT<T> seeded_texel(const Ctx<T>& c, int layer, int x, int y, int ch) {
  DualT<T> r(eval_texel<T, T>(c, layer, x, y, ch));
  const int64_t off = ((((int64_t)layer * c.s->th + y) * c.s->tw + x) * 4 + ch);
  const int slot = c.seeds ? Seeds::find(c.seeds->atlas, off) : -1;
  if (slot >= 0) r.g[slot] = T(1);
  return r;
}
template <> DualT<float> eval_texel<float, DualT<float>>(const Ctx<float>& c, int layer, int x, int y, int ch) {
  return seeded_texel<float>(c, layer, x, y, ch);
}
template <> DualT<double> eval_texel<double, DualT<double>>(const Ctx<double>& c, int layer, int x, int y, int ch) {
  return seeded_texel<double>(c, layer, x, y, ch);
}
// the same lookup of the evaluation atlas
template <class T, class U> void sample_tex_eval(const Ctx<T>& c, int layer, T u, T v, U out[4]) {
  const or_scene* s = c.s;
  T x = u * (T)s->tw - T(0.5), y = v * (T)s->th - T(0.5);
  T fx0 = M<T>::floor(x), fy0 = M<T>::floor(y);
  T fx = x - fx0, fy = y - fy0;
  int x0 = wrap((int)fx0, (int)s->tw), x1 = wrap((int)fx0 + 1, (int)s->tw);
  int y0 = wrap((int)fy0, (int)s->th), y1 = wrap((int)fy0 + 1, (int)s->th);
  for (int ch = 0; ch < 4; ++ch) {
    U a = eval_texel<T, U>(c, layer, x0, y0, ch), b = eval_texel<T, U>(c, layer, x1, y0, ch);
    U cc = eval_texel<T, U>(c, layer, x0, y1, ch), dd = eval_texel<T, U>(c, layer, x1, y1, ch);
    U top = a + (b - a) * fx;
    U bot = cc + (dd - cc) * fx;
    out[ch] = top + (bot - top) * fy;
  }
}

template <class T, class U> V3<U> env_radiance(const Ctx<T>& c, size_t idx);
template <> V3<float> env_radiance<float, float>(const Ctx<float>& c, size_t i) {
  const or_scene* s = c.s;
  return v3(s->erad32[i * 3], s->erad32[i * 3 + 1], s->erad32[i * 3 + 2]);
}
template <> V3<double> env_radiance<double, double>(const Ctx<double>& c, size_t i) {
  const or_scene* s = c.s;
  return v3(s->erad64[i * 3], s->erad64[i * 3 + 1], s->erad64[i * 3 + 2]);
}
template <class T> V3<DualT<T>> seeded_env(const Ctx<T>& c, size_t i) {
  V3<T> e = env_radiance<T, T>(c, i);
  DualT<T> r[3] = {DualT<T>(e.x), DualT<T>(e.y), DualT<T>(e.z)};
  for (int k = 0; k < 3; ++k) {
    const int slot = c.seeds ? Seeds::find(c.seeds->env, (int64_t)i * 4 + k) : -1;
    if (slot >= 0) r[k].g[slot] = T(1);
  }
  return v3(r[0], r[1], r[2]);
}
template <> V3<DualT<float>> env_radiance<float, DualT<float>>(const Ctx<float>& c, size_t i) {
  return seeded_env<float>(c, i);
}
template <> V3<DualT<double>> env_radiance<double, DualT<double>>(const Ctx<double>& c, size_t i) {
  return seeded_env<double>(c, i);
}
template <class T> T env_pdf_tex(const or_scene* s, size_t i);
template <> float env_pdf_tex<float>(const or_scene* s, size_t i) { return s->epdf32[i]; }
template <> double env_pdf_tex<double>(const or_scene* s, size_t i) { return s->epdf64[i]; }

// shade_miss.wgsl envColor (:37-43): nearest sampler, repeat on u and v
template <class T, class U> V3<U> env_color(const Ctx<T>& c, V3<T> dir) {
  const or_scene* s = c.s;
  if (s->efmt == PK_ENV_NONE) return v3(U(0), U(0), U(0));
  T u = (mut(OR_MUT_THETA_SIGN) ? -c.env_theta : c.env_theta) + M<T>::atan2(dir.z, dir.x) / M<T>::tau();
  T v = M<T>::acos(mut(OR_MUT_ENV_COLOR_V) ? -dir.y : dir.y) * M<T>::inv_pi();
  int ix = wrap((int)M<T>::floor(u * (T)s->ew), (int)s->ew);
  int iy = wrap((int)M<T>::floor(v * (T)s->eh), (int)s->eh);
  return env_radiance<T, U>(c, (size_t)iy * s->ew + (size_t)ix);
}
// shade_hit.wgsl envPdf (:202-210); the texel index is clamped into the map
template <class T> T env_pdf(const Ctx<T>& c, V3<T> dir) {
  const or_scene* s = c.s;
  T u = std::fmod(T(1) + c.env_theta + M<T>::atan2(dir.z, dir.x) / M<T>::tau(), T(1));
  T v = M<T>::acos(dir.y) * M<T>::inv_pi();
  int ix = (int)(u * (T)s->ew), iy = (int)(v * (T)s->eh);
  ix = std::min(std::max(ix, 0), (int)s->ew - 1);
  iy = std::min(std::max(iy, 0), (int)s->eh - 1);
  T phi = v * M<T>::pi();
  T sin_phi = mut(OR_MUT_ENV_PDF_SIN) ? T(1) : M<T>::sin(phi);
  return env_pdf_tex<T>(s, (size_t)iy * s->ew + (size_t)ix) / sin_phi;
}

template <class T> struct Sample {
  V3<T> wi;
  T pdf;
};
// sampling side of the surface interaction
template <class T> struct SI {
  M3<T> basis;
  V3<T> wo, origin, normal, emission;
  T metallic, opacity, alpha;
};
// evaluation side (base color, specular color, GGX alpha)
template <class U> struct EV {
  V3<U> base, spec;
  U alpha;
};
// branchlessONB (:166-177), Duff et al. 2017
template <class T> M3<T> onb(V3<T> n) {
  T sign = n.z > T(0) ? T(1) : T(-1);
  T a = T(-1) / (sign + n.z);
  T b = n.x * n.y * a;
  V3<T> b1 = v3(T(1) + sign * n.x * n.x * a, sign * b, mut(OR_MUT_ONB_SIGN) ? sign * n.x : -sign * n.x);
  V3<T> b2 = v3(b, sign + n.y * n.y * a, -n.y);
  return M3<T>{b1, b2, n};
}
template <class T, class U> U ggx_d(V3<T> m, U au, U av) {
  U auv = au * av;
  U tx = m.x / au, ty = m.y / av;
  U e = tx * tx + ty * ty + m.z * m.z;
  return U(1) / ((mut(OR_MUT_D_PI) ? T(1) : M<T>::pi()) * auv * e * e);
}
template <class T, class U> U ggx_g1(V3<T> w, V3<T> m, U au, U av) {
  const V3<T> ws = mut(OR_MUT_G1_SWAP) ? m : w;
  U ax = ws.x * au, ay = ws.y * av;
  U axy2 = ax * ax + ay * ay;
  U tan_theta = axy2 / (ws.z * ws.z);
  U r = U(2) / (U(1) + usqrt(U(1) + tan_theta));
  if (mut(OR_MUT_G1_SIGN)) return r;
  return dot(w, m) * dot(w, v3(T(0), T(0), T(1))) <= T(0) ? U(0) : r;
}
template <class T> T lambert_pdf(V3<T> wi) { return mx(wi.z, T(0.0000001f)) * M<T>::inv_pi(); }
template <class T> T specular_pdf(const SI<T>& si, V3<T> m) {
  return mx(T(0.0000001f), ggx_d<T, T>(m, si.alpha, si.alpha) * ggx_g1<T, T>(si.wo, m, si.alpha, si.alpha) /
                               ((mut(OR_MUT_SPEC_PDF4) ? T(1) : T(4)) * si.wo.z));
}
template <class T, class U> V3<U> eval_lambert(const EV<U>& e, const Sample<T>& s) {
  T k = mx(T(0.0000001f), s.wi.z);
  const T ip = mut(OR_MUT_LAMBERT_INV_PI) ? T(1) : M<T>::inv_pi();
  return v3(e.base.x * ip * k / s.pdf, e.base.y * ip * k / s.pdf, e.base.z * ip * k / s.pdf);
}
template <class T, class U> V3<U> eval_specular(const SI<T>& si, const EV<U>& e, const Sample<T>& s) {
  V3<T> h = normalize(si.wo + s.wi);
  U d = ggx_d<T, U>(h, e.alpha, e.alpha);
  U g = ggx_g1<T, U>(s.wi, h, e.alpha, e.alpha) * ggx_g1<T, U>(si.wo, h, e.alpha, e.alpha);
  U k = mx(d * g / (T(4) * (mut(OR_MUT_EVAL_SPEC_COS) ? s.wi.z : si.wo.z) * s.pdf), U(0));
  return e.spec * k;
}
template <class T> T power_heuristic(T a, T b) {
  if (mut(OR_MUT_PH_BALANCE)) return a / (a + b);
  T a2 = a * a;
  return a2 / (a2 + b * b);
}
template <class T> T schlick(T cos_theta, T ior) {
  T r0 = (T(1) - ior) / (T(1) + ior);
  r0 *= r0;
  T tmp = T(1) - cos_theta;
  T tmp2 = tmp * tmp;
  return r0 + (T(1) - r0) * tmp2 * tmp2 * (mut(OR_MUT_SCHLICK_POW) ? T(1) : tmp);
}
// sampleGGX (:295-316), Heitz 2018 VNDF
template <class T> V3<T> sample_ggx(const SI<T>& si, Rng& rng) {
  T au = si.alpha, av = si.alpha;
  V3<T> vh = normalize(v3(au * si.wo.x, av * si.wo.y, si.wo.z));
  T lensq = vh.x * vh.x + vh.y * vh.y;
  V3<T> t1v = lensq > T(0) ? v3(-vh.y, vh.x, T(0)) * (T(1) / M<T>::sqrt(lensq)) : v3(T(1), T(0), T(0));
  V3<T> t2v = cross(vh, t1v);
  T r = M<T>::sqrt(rng.rand<T>());
  T phi = M<T>::tau() * rng.rand<T>();
  T t1 = r * M<T>::cos(phi);
  T t2 = r * M<T>::sin(phi);
  T sv = T(0.5) * (mut(OR_MUT_VNDF_BLEND) ? T(1) - vh.z : T(1) + vh.z);
  t2 = (T(1) - sv) * M<T>::sqrt(T(1) - t1 * t1) + sv * t2;
  V3<T> nh = t1v * t1 + t2v * t2 + vh * M<T>::sqrt(mx(T(0), T(1) - t1 * t1 - t2 * t2));
  return normalize(v3(au * nh.x, av * nh.y, mx(T(0), nh.z)));
}
template <class T> Sample<T> sample_lambert(Rng& rng) {
  T r = mut(OR_MUT_LAMBERT_SAMPLE) ? rng.rand<T>() : M<T>::sqrt(rng.rand<T>());
  T phi = M<T>::tau() * rng.rand<T>();
  T x = r * M<T>::cos(phi);
  T y = r * M<T>::sin(phi);
  T z = M<T>::sqrt(mx(T(0), T(1) - x * x - y * y));
  V3<T> d = v3(x, y, z);
  return Sample<T>{d, lambert_pdf(d)};
}
// sampleEnv (:213-227); the CDF sampler draws its two uniforms from the same
// two hashes and binary-searches the f32 tables (largest index whose CDF
// value is <= the uniform)
template <class T> Sample<T> sample_env(const Ctx<T>& c, const M3<T>& basis, Rng& rng) {
  const or_scene* s = c.s;
  int32_t cx, cy;
  if (s->esampler == PK_ENV_SAMPLER_CDF) {
    const T u1 = mn(rng.rand<T>(), (T)0x1.fffffep-1f);
    const T u2 = mn(rng.rand<T>(), (T)0x1.fffffep-1f);
    auto search = [](const float* cdf, int n, T u) {
      int lo = 0, hi = n;
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if ((T)cdf[mid] <= u) lo = mid;
        else hi = mid;
      }
      return lo;
    };
    cy = search(s->row_cdf.data(), (int)s->eh, u1);
    cx = search(&s->col_cdf[(size_t)cy * (s->ew + 1)], (int)s->ew, u2);
  } else {
    uint32_t bi = rng.hash() % (uint32_t)s->ebins.size();
    const auto& bin = s->ebins[bi];
    int32_t ci = (int32_t)(rng.hash() % (uint32_t)(bin[1] - bin[0])) + bin[0];
    cx = s->ecoords[ci][0];
    cy = s->ecoords[ci][1];
  }
  const std::array<int32_t, 2> coord = {cx, cy};
  const T half = mut(OR_MUT_ENV_CENTER) ? T(0) : T(0.5);
  T u = -c.env_theta + ((half + (T)coord[0]) / (T)s->ew);
  T v = (half + (T)coord[1]) / (T)s->eh;
  T theta = u * M<T>::tau();
  T phi = v * M<T>::pi();
  T sin_phi = M<T>::sin(phi);
  V3<T> dir = v3(M<T>::cos(theta) * sin_phi, M<T>::cos(phi), M<T>::sin(theta) * sin_phi);
  T pdf = env_pdf_tex<T>(s, (size_t)coord[1] * s->ew + (size_t)coord[0]) /
          (mut(OR_MUT_SAMPLE_ENV_SIN) ? T(1) : sin_phi);
  return Sample<T>{vec_mat(dir, basis), pdf};
}

template <class T> V3<T> attr_vec(const or_scene* s, uint32_t vi, int word) {
  uint32_t w = s->attr[(size_t)vi * 4 + word];
  auto un = [](uint32_t q) { return (T)mx((float)(int8_t)(uint8_t)q / 127.0f, -1.0f); };
  return v3(un(w & 0xff), un((w >> 8) & 0xff), un((w >> 16) & 0xff));
}
template <> V3<double> attr_vec<double>(const or_scene* s, uint32_t vi, int word) {
  uint32_t w = s->attr[(size_t)vi * 4 + word];
  auto un = [](uint32_t q) { return mx((double)(int8_t)(uint8_t)q / 127.0, -1.0); };
  return v3(un(w & 0xff), un((w >> 8) & 0xff), un((w >> 16) & 0xff));
}
template <class T> T attr_uv(const or_scene* s, uint32_t vi, int comp) {
  uint32_t w = s->attr[(size_t)vi * 4 + 3];
  uint32_t q = comp == 0 ? (w & 0xffff) : (w >> 16);
  return (T)mx((float)(int16_t)(uint16_t)q / 32767.0f, -1.0f);
}
template <> double attr_uv<double>(const or_scene* s, uint32_t vi, int comp) {
  uint32_t w = s->attr[(size_t)vi * 4 + 3];
  uint32_t q = comp == 0 ? (w & 0xffff) : (w >> 16);
  return mx((double)(int16_t)(uint16_t)q / 32767.0, -1.0);
}

// getSurfaceInteraction (:179-200) with interpolateVertexAttribute (:363-394)
template <class T, class U> SI<T> surface(const Ctx<T>& c, V3<T> ro, V3<T> rd, const HitT<T>& h, EV<U>* ev) {
  const or_scene* s = c.s;
  // geometry of the hit: the scene's, or the placement's mesh
  const or_scene* g = s;
  const or_scene::Inst* inst = nullptr;
  int32_t tid = h.id;
  if (s->instanced && h.inst >= 0) {
    inst = &s->insts[(size_t)h.inst];
    g = s->meshes[inst->mesh].get();
    tid = h.local;
  }
  const int32_t* tr = &g->tri[(size_t)tid * 4];
  T bu = (T)(float)(h.bary & 0xffff) / T(65535), bv = (T)(float)(h.bary >> 16) / T(65535);
  V3<T> bary = v3(T(1) - bu - bv, bu, bv);
  auto interp = [&](int word) {
    M3<T> m{attr_vec<T>(g, tr[0], word), attr_vec<T>(g, tr[1], word), attr_vec<T>(g, tr[2], word)};
    return normalize(mat_vec(m, bary));
  };
  V3<T> tangent = interp(0), bitangent = interp(1), normal = interp(2);
  if (inst) {
    // mesh -> scene: tangents by the linear part, the normal by the
    // transpose of the scene -> mesh linear part
    const T* L = o2w_of<T>(*inst);
    const T* A = w2o_of<T>(*inst);
    auto lin = [&](V3<T> v) {
      return v3((L[0] * v.x + L[1] * v.y) + L[2] * v.z, (L[3] * v.x + L[4] * v.y) + L[5] * v.z,
                (L[6] * v.x + L[7] * v.y) + L[8] * v.z);
    };
    tangent = normalize(lin(tangent));
    bitangent = normalize(lin(bitangent));
    normal = normalize(v3((A[0] * normal.x + A[4] * normal.y) + A[8] * normal.z,
                          (A[1] * normal.x + A[5] * normal.y) + A[9] * normal.z,
                          (A[2] * normal.x + A[6] * normal.y) + A[10] * normal.z));
  }
  T uvx = attr_uv<T>(g, tr[0], 0) * bary.x + attr_uv<T>(g, tr[1], 0) * bary.y + attr_uv<T>(g, tr[2], 0) * bary.z;
  T uvy = attr_uv<T>(g, tr[0], 1) * bary.x + attr_uv<T>(g, tr[1], 1) * bary.y + attr_uv<T>(g, tr[2], 1) * bary.z;
  const int32_t mat_id = (inst && inst->material >= 0) ? inst->material : tr[3];
  const int32_t* mi = &s->mat[(size_t)mat_id * 4];
  T nm[4], dc[4], mr[4], em[4];
  sample_tex<T>(s, mi[2], uvx, uvy, nm);
  const T nms = mut(OR_MUT_NORMAL_MAP) ? T(1) : T(2);
  V3<T> map_normal = v3((nm[0] - T(0.5)) * nms, (nm[1] - T(0.5)) * nms, (nm[2] - T(0)) * T(1));
  SI<T> si;
  si.normal = normalize(mat_vec(M3<T>{tangent, bitangent, normal}, map_normal));
  const T eps40 = T(0.0000001f) * T(40);
  si.origin = ro + rd * (mut(OR_MUT_ORIGIN_SIGN) ? h.t + eps40 : h.t - eps40);
  si.basis = onb(si.normal);
  si.wo = vec_mat(-rd, si.basis);
  sample_tex<T>(s, mi[0], uvx, uvy, dc);
  si.opacity = dc[3];
  sample_tex<T>(s, mi[1], uvx, uvy, mr);
  si.metallic = mut(OR_MUT_METAL_CHANNEL) ? mr[1] : mr[2];
  si.alpha = mut(OR_MUT_ALPHA_SQUARE) ? mr[1] : mr[1] * mr[1];
  sample_tex<T>(s, mi[3], uvx, uvy, em);
  si.emission = v3(em[0], em[1], em[2]);
  U dce[4], mre[4];
  sample_tex_eval<T, U>(c, mi[0], uvx, uvy, dce);
  sample_tex_eval<T, U>(c, mi[1], uvx, uvy, mre);
  ev->base = v3(dce[0], dce[1], dce[2]);
  ev->alpha = mre[1] * mre[1];
  auto mixu = [&](U b) { return T(1) * (T(1) - si.metallic) + b * si.metallic; };
  ev->spec = v3(mixu(ev->base.x), mixu(ev->base.y), mixu(ev->base.z));
  return si;
}

template <class T> struct Ray {
  V3<T> o, d;
};
// createPrimaryRay (camera.wgsl:58-73)
template <class T> Ray<T> primary_ray(const pk_camera& cam, uint32_t px, uint32_t py, uint32_t W, uint32_t H, Rng& rng) {
  const T k = (T)cam.distortion;
  T resx = (T)W, resy = (T)H;
  T jx = rng.rand<T>();
  T jy = rng.rand<T>();
  if (mut(OR_MUT_JITTER_ORDER)) std::swap(jx, jy);
  T ux = (T(2) * (((T)px + jx) / resx) - T(1)) * (mut(OR_MUT_ASPECT) ? T(1) : resx / resy);
  T uy = (T(2) * (((T)py + jy) / resy) - T(1)) * T(-1);
  T rd = M<T>::sqrt(ux * ux + uy * uy);
  T ru = rd * (mut(OR_MUT_DISTORTION) ? T(1) - k * rd * rd : T(1) + k * rd * rd);
  T l2 = M<T>::sqrt(ux * ux + uy * uy);
  ux = ru * (ux / l2);
  uy = ru * (uy / l2);
  V3<T> up = v3(T(0), T(1), T(0));
  V3<T> dir = v3((T)cam.dir[0], (T)cam.dir[1], (T)cam.dir[2]);
  V3<T> org = v3((T)cam.origin[0], (T)cam.origin[1], (T)cam.origin[2]);
  T tf = M<T>::tan((T)cam.fov * T(0.5));
  V3<T> bx = normalize(cross(dir, up)) * tf;
  V3<T> by = normalize(cross(bx, dir)) * tf;
  T theta = rng.rand<T>() * M<T>::tau();
  T pw = M<T>::pow(rng.rand<T>(), (T)cam.bokeh);
  V3<T> dof = (bx * M<T>::cos(theta) + by * M<T>::sin(theta)) * (T)cam.aperture * pw;
  V3<T> screen = bx * ux + by * uy + dir + org;
  V3<T> d = normalize((screen + dof * (T)cam.focal_depth) - (org + dof));
  return Ray<T>{dof + org, d};
}

template <class U> struct PixelOut {
  U c[4] = {U(0), U(0), U(0), U(0)};
  uint64_t rays = 0, shadow = 0, hits = 0;
};

template <class T>
HitT<T> closest(const Ctx<T>& c, V3<T> o, V3<T> d) { return scene_trace<T>(c.s, o, d, false); }

template <class T, class U>
void render_pixel(const Ctx<T>& c, const pk_camera& cam, uint32_t px, uint32_t py, uint32_t W, uint32_t H,
                  uint32_t sample, uint32_t depth, PixelOut<U>& out) {
  const uint32_t pixel = py * W + px;
  Rng rng;
  rng.seed = (px * 1973u + py * 9277u + sample * 26699u) | 1u;  // camera.wgsl:86-87
  rng.seed = rng.hash();
  Ray<T> ray = primary_ray<T>(cam, px, py, W, H, rng);
  V3<U> thr = v3(U(1), U(1), U(1));
  auto add = [&](V3<U> L) {
    out.c[0] += L.x;
    out.c[1] += L.y;
    out.c[2] += L.z;
    out.c[3] += U(1);
  };
  const T eps = T(0.0000001f);
  for (uint32_t b = 0; b < depth; ++b) {
    out.rays++;
    HitT<T> h = closest<T>(c, ray.o, ray.d);
    if (h.id < 0) {  // shade_miss.wgsl:53-57
      add(env_color<T, U>(c, ray.d) * thr);
      return;
    }
    out.hits++;
    // shade_hit.wgsl main (:411-483); the seed uses the bounce index where the
    // reference uses the hit's atomic compaction slot (order-dependent).
    Rng sr;
    sr.seed = (b * 1973u + pixel * 9277u + sample * 26699u) | 1u;
    sr.seed = sr.hash();
    EV<U> ev;
    SI<T> si = surface<T, U>(c, ray.o, ray.d, h, &ev);
    if (sr.rand<T>() > si.opacity) {
      V3<T> o2 = ray.o + ray.d * (h.t + eps * T(40));
      ray = Ray<T>{o2, ray.d};
      continue;
    }
    if (si.emission.x * T(1) + si.emission.y * T(1) + si.emission.z * T(1) > T(0)) {
      T att = mx(dot(si.normal, mat_vec(si.basis, si.wo)), T(0));
      const T scale = mut(OR_MUT_EMISSION) ? T(1) : T(50);
      add(v3(scale, scale, scale) * si.emission * thr * att);
      return;
    }
    V3<T> m = sample_ggx<T>(si, sr);
    T f = mix(schlick(mx(dot(si.wo, m), T(0)), T(1.5)), T(1),
              mut(OR_MUT_FRESNEL_METAL) ? T(1) - si.metallic : si.metallic);
    Sample<T> bs;
    V3<U> bsdf;
    if (sr.rand<T>() > f) {
      bs = sample_lambert<T>(sr);
      bsdf = eval_lambert<T, U>(ev, bs);
    } else {
      V3<T> wi = reflect(-si.wo, m);
      bs = Sample<T>{wi, specular_pdf(si, m)};
      bsdf = eval_specular<T, U>(si, ev, bs);
    }
    bool bounce = false;
    Ray<T> next;
    V3<U> next_thr;
    if (bs.wi.z > T(0)) {
      V3<T> dir = mat_vec(si.basis, bs.wi);
      T weight = c.sample_env && !mut(OR_MUT_BOUNCE_WEIGHT) ? power_heuristic(bs.pdf, env_pdf<T>(c, dir)) : T(1);
      next = Ray<T>{si.origin, dir};
      next_thr = bsdf * thr * weight;
      bounce = true;
    }
    if (c.sample_env) {
      Sample<T> es = sample_env<T>(c, si.basis, sr);
      if (dot(es.wi, m) > T(0) && es.wi.z > T(0)) {
        // createMaterialSampleRay (:339-348)
        T lw = power_heuristic(es.pdf, lambert_pdf(es.wi));
        V3<U> el = eval_lambert<T, U>(ev, es);
        V3<U> scale = v3((T(1) - f) * el.x * lw, (T(1) - f) * el.y * lw, (T(1) - f) * el.z * lw);
        V3<T> hh = normalize(si.wo + es.wi);
        T sw = power_heuristic(es.pdf, mut(OR_MUT_NEE_SPEC_WEIGHT) ? lambert_pdf(es.wi) : specular_pdf(si, hh));
        V3<U> evs = eval_specular<T, U>(si, ev, es);
        scale = scale + v3(f * evs.x * sw, f * evs.y * sw, f * evs.z * sw);
        V3<T> sdir = mat_vec(si.basis, es.wi);
        V3<U> sthr = scale * thr;
        out.shadow++;
        HitT<T> oh = scene_trace<T>(c.s, si.origin, sdir, true);
        if (oh.id < 0) add(env_color<T, U>(c, sdir) * sthr);
      }
    }
    if (!bounce) return;
    ray = next;
    thr = next_thr;
  }
}

// accumulate.wgsl:26-31: acc = (max(c, 0) + acc * s) / (s + 1), s = the
// frame's sample index; the accumulator is an f32 buffer (FP64 evaluates the
// update in double and rounds once).
template <class T> float accumulate(T c, float acc, uint32_t sample) {
  const uint32_t den = sample + (mut(OR_MUT_ACC_DENOM) ? 2u : 1u);
  // max(color, 0) with the GPU max's NaN rule: a NaN color adds 0
  if (sizeof(T) == 4) {
    const float col = (float)c > 0.0f ? (float)c : 0.0f;
    return (col + acc * (float)sample) / (float)den;
  }
  const double col = (double)c > 0.0 ? (double)c : 0.0;
  return (float)((col + (double)acc * (double)sample) / (double)den);
}

template <class T>
int render_t(const or_scene* s, const pk_camera* cam, const or_render_params* p, float* accum, float* color_out,
             uint64_t* counters) {
  Ctx<T> c;
  c.s = s;
  c.nodes = sizeof(T) == 4 ? &s->cnodes32 : &s->cnodes64;
  c.env_theta = (T)p->env_theta;
  c.sample_env = s->efmt != PK_ENV_NONE;
  c.seeds = nullptr;
  const uint32_t W = p->width, H = p->height;
  uint32_t x0 = p->x0, x1 = p->x1, y0 = p->y0, y1 = p->y1;
  if (x1 == 0 && y1 == 0) {
    x0 = 0;
    y0 = 0;
    x1 = W;
    y1 = H;
  }
  const uint32_t ww = x1 - x0, wh = y1 - y0;
  std::vector<uint64_t> rays((size_t)ww * wh), shadow((size_t)ww * wh), hits((size_t)ww * wh);
  for (uint32_t f = 0; f < p->num_frames; ++f) {
    const uint32_t sample = p->first_sample + f;
    parallel_for((int64_t)ww * wh, p->num_threads, [&](int64_t i) {
      uint32_t px = x0 + (uint32_t)(i % ww), py = y0 + (uint32_t)(i / ww);
      PixelOut<T> o;
      render_pixel<T, T>(c, *cam, px, py, W, H, sample, p->max_depth, o);
      rays[i] += o.rays;
      shadow[i] += o.shadow;
      hits[i] += o.hits;
      // accumulate.wgsl:26-30 in f32 (the accumulator is an f32 buffer)
      float* a = &accum[(size_t)i * 4];
      for (int k = 0; k < 3; ++k) a[k] = accumulate<T>((T)o.c[k], a[k], sample);
      a[3] = 1.0f;
      if (color_out && f + 1 == p->num_frames)
        for (int k = 0; k < 4; ++k) color_out[(size_t)i * 4 + k] = (float)o.c[k];
    });
  }
  if (counters) {
    counters[0] = std::accumulate(rays.begin(), rays.end(), (uint64_t)0);
    counters[1] = std::accumulate(shadow.begin(), shadow.end(), (uint64_t)0);
    counters[2] = std::accumulate(hits.begin(), hits.end(), (uint64_t)0);
  }
  return 0;
}

// Reference gradient: the frame's estimator evaluated with dual numbers
// (real type T for geometry, sampling and values), kDualN parameters per
// pass over the image.  Derivatives are summed in f64.
template <class T>
int gradient_t(const or_scene* s, const pk_camera* cam, const or_render_params* p, const float* d_image,
               const int32_t* which, const int64_t* index, int64_t n, double* out, double* value) {
  using D = DualT<T>;
  const uint32_t W = p->width, H = p->height;
  const int64_t npx = (int64_t)W * H;
  const double inv_n = 1.0 / (double)p->num_frames;
  // one pass at least, for the objective value
  for (int64_t base = 0; base < std::max<int64_t>(n, 1); base += kDualN) {
    const int cnt = (int)std::min<int64_t>(kDualN, n - base);
    Seeds seeds;
    for (int k = 0; k < cnt; ++k) {
      auto& v = which[base + k] == PK_PARAM_ENV ? seeds.env : seeds.atlas;
      v.push_back({index[base + k], k});
    }
    std::sort(seeds.atlas.begin(), seeds.atlas.end());
    std::sort(seeds.env.begin(), seeds.env.end());
    Ctx<T> c;
    c.s = s;
    c.nodes = sizeof(T) == 4 ? &s->cnodes32 : &s->cnodes64;
    c.env_theta = (T)p->env_theta;
    c.sample_env = s->efmt != PK_ENV_NONE;
    c.seeds = &seeds;
    std::vector<double> part((size_t)npx * std::max(cnt, 1), 0.0), val((size_t)npx, 0.0);
    parallel_for(npx, p->num_threads, [&](int64_t i) {
      const uint32_t px = (uint32_t)(i % W), py = (uint32_t)(i / W);
      const float* g = &d_image[(size_t)i * 4];
      if (g[0] == 0.0f && g[1] == 0.0f && g[2] == 0.0f) return;
      for (uint32_t f = 0; f < p->num_frames; ++f) {
        PixelOut<D> o;
        render_pixel<T, D>(c, *cam, px, py, W, H, p->first_sample + f, p->max_depth, o);
        // accumulate drops a NaN channel of a sample (adds 0), so it
        // contributes neither value nor gradient
        double gk[3];
        for (int ch = 0; ch < 3; ++ch) gk[ch] = o.c[ch].v == o.c[ch].v ? (double)g[ch] : 0.0;
        for (int k = 0; k < cnt; ++k)
          part[(size_t)i * cnt + k] += (gk[0] * (double)o.c[0].g[k] + gk[1] * (double)o.c[1].g[k] +
                                        gk[2] * (double)o.c[2].g[k]) *
                                       inv_n;
        val[i] += ((gk[0] != 0.0 ? gk[0] * (double)o.c[0].v : 0.0) + (gk[1] != 0.0 ? gk[1] * (double)o.c[1].v : 0.0) +
                   (gk[2] != 0.0 ? gk[2] * (double)o.c[2].v : 0.0)) *
                  inv_n;
      }
    });
    for (int k = 0; k < cnt; ++k) {
      double acc = 0;
      for (int64_t i = 0; i < npx; ++i) acc += part[(size_t)i * cnt + k];
      out[base + k] = acc;
    }
    if (value && base == 0) *value = std::accumulate(val.begin(), val.end(), 0.0);
  }
  return 0;
}

}  // namespace

// ---------------------------------------------------------------------------
// Probes: single functions of the restatement evaluated on given inputs, for
// the closed-form tests (or_probe).  T = double is the FP64 oracle, float
// the FP32 restatement; both instantiate the same templates.
// ---------------------------------------------------------------------------
template <class T> V3<T> in3(const double* p) { return v3((T)p[0], (T)p[1], (T)p[2]); }
template <class T> void out3(double* o, V3<T> v) {
  o[0] = (double)v.x;
  o[1] = (double)v.y;
  o[2] = (double)v.z;
}

template <class T> int probe_one(const or_scene* s, int fn, const double* in, double* out) {
  Ctx<T> c;
  c.s = s;
  c.nodes = nullptr;
  c.env_theta = T(0);
  c.sample_env = s && s->efmt != PK_ENV_NONE;
  c.seeds = nullptr;
  auto need_scene = [&]() { return s != nullptr; };
  switch (fn) {
    case OR_PROBE_GGX_D: out[0] = (double)ggx_d<T, T>(in3<T>(in), (T)in[3], (T)in[3]); return 0;
    case OR_PROBE_GGX_G1: out[0] = (double)ggx_g1<T, T>(in3<T>(in), in3<T>(in + 3), (T)in[6], (T)in[6]); return 0;
    case OR_PROBE_LAMBERT_PDF: out[0] = (double)lambert_pdf<T>(in3<T>(in)); return 0;
    case OR_PROBE_SPECULAR_PDF: {
      SI<T> si{};
      si.wo = in3<T>(in);
      si.alpha = (T)in[6];
      out[0] = (double)specular_pdf<T>(si, in3<T>(in + 3));
      return 0;
    }
    case OR_PROBE_EVAL_LAMBERT: {
      EV<T> e{};
      e.base = in3<T>(in);
      out3(out, eval_lambert<T, T>(e, Sample<T>{in3<T>(in + 3), (T)in[6]}));
      return 0;
    }
    case OR_PROBE_EVAL_SPECULAR: {
      SI<T> si{};
      si.wo = in3<T>(in);
      EV<T> e{};
      e.alpha = (T)in[7];
      e.spec = in3<T>(in + 8);
      out3(out, eval_specular<T, T>(si, e, Sample<T>{in3<T>(in + 3), (T)in[6]}));
      return 0;
    }
    case OR_PROBE_POWER_HEURISTIC: out[0] = (double)power_heuristic<T>((T)in[0], (T)in[1]); return 0;
    case OR_PROBE_SCHLICK: out[0] = (double)schlick<T>((T)in[0], (T)in[1]); return 0;
    case OR_PROBE_ONB: {
      const M3<T> b = onb<T>(in3<T>(in));
      out3(out, b.c0);
      out3(out + 3, b.c1);
      out3(out + 6, b.c2);
      return 0;
    }
    case OR_PROBE_SAMPLE_GGX: {
      SI<T> si{};
      si.wo = in3<T>(in);
      si.alpha = (T)in[3];
      Rng rng{(uint32_t)in[4]};
      out3(out, sample_ggx<T>(si, rng));
      return 0;
    }
    case OR_PROBE_SAMPLE_LAMBERT: {
      Rng rng{(uint32_t)in[0]};
      const Sample<T> smp = sample_lambert<T>(rng);
      out3(out, smp.wi);
      out[3] = (double)smp.pdf;
      return 0;
    }
    case OR_PROBE_ENV_COLOR: {
      if (!need_scene()) return -1;
      c.env_theta = (T)in[3];
      out3(out, env_color<T, T>(c, in3<T>(in)));
      return 0;
    }
    case OR_PROBE_ENV_PDF: {
      if (!need_scene() || s->efmt == PK_ENV_NONE) return -1;
      c.env_theta = (T)in[3];
      out[0] = (double)env_pdf<T>(c, in3<T>(in));
      return 0;
    }
    case OR_PROBE_SAMPLE_ENV: {
      if (!need_scene() || s->efmt == PK_ENV_NONE) return -1;
      c.env_theta = (T)in[1];
      Rng rng{(uint32_t)in[0]};
      const M3<T> id{v3(T(1), T(0), T(0)), v3(T(0), T(1), T(0)), v3(T(0), T(0), T(1))};
      const Sample<T> smp = sample_env<T>(c, id, rng);
      out3(out, smp.wi);
      out[3] = (double)smp.pdf;
      return 0;
    }
    case OR_PROBE_TEXTURE: {
      if (!need_scene() || s->tw == 0) return -1;
      T r[4];
      sample_tex<T>(s, (int)in[0], (T)in[1], (T)in[2], r);
      for (int k = 0; k < 4; ++k) out[k] = (double)r[k];
      return 0;
    }
    case OR_PROBE_PRIMARY_RAY: {
      pk_camera cam{};
      for (int k = 0; k < 3; ++k) {
        cam.origin[k] = (float)in[k];
        cam.dir[k] = (float)in[3 + k];
      }
      cam.fov = (float)in[6];
      cam.focal_depth = (float)in[7];
      cam.aperture = (float)in[8];
      cam.distortion = (float)in[9];
      cam.bokeh = (float)in[10];
      const uint32_t px = (uint32_t)in[11], py = (uint32_t)in[12], W = (uint32_t)in[13], H = (uint32_t)in[14];
      const uint32_t sample = (uint32_t)in[15];
      // the camera's seed and discarded first hash (camera.wgsl:86-87)
      Rng rng;
      rng.seed = (px * 1973u + py * 9277u + sample * 26699u) | 1u;
      rng.seed = rng.hash();
      const Ray<T> r = primary_ray<T>(cam, px, py, W, H, rng);
      out3(out, r.o);
      out3(out + 3, r.d);
      return 0;
    }
    case OR_PROBE_SURFACE: {
      if (!need_scene()) return -1;
      c.nodes = sizeof(T) == 4 ? &s->cnodes32 : &s->cnodes64;
      const V3<T> o = in3<T>(in), d = in3<T>(in + 3);
      const HitT<T> h = closest<T>(c, o, d);
      for (int k = 0; k < OR_PROBE_SURFACE_OUT; ++k) out[k] = 0;
      out[0] = h.id < 0 ? -1.0 : (double)h.t;
      out[1] = (double)h.id;
      if (h.id < 0) return 0;
      EV<T> ev;
      const SI<T> si = surface<T, T>(c, o, d, h, &ev);
      out3(out + 2, si.normal);
      out3(out + 5, si.origin);
      out3(out + 8, si.wo);
      out3(out + 11, si.basis.c0);
      out3(out + 14, si.basis.c1);
      out3(out + 17, si.basis.c2);
      out[20] = (double)si.metallic;
      out[21] = (double)si.opacity;
      out[22] = (double)si.alpha;
      out3(out + 23, si.emission);
      out3(out + 26, ev.base);
      out3(out + 29, ev.spec);
      out[32] = (double)ev.alpha;
      out[33] = (double)(h.bary & 0xffffu) / 65535.0;
      out[34] = (double)(h.bary >> 16) / 65535.0;
      return 0;
    }
    case OR_PROBE_ACCUMULATE: out[0] = (double)accumulate<T>((T)in[0], (float)in[1], (uint32_t)in[2]); return 0;
    default: return -1;
  }
}

// ---------------------------------------------------------------------------
// C API
// ---------------------------------------------------------------------------
extern "C" {

int or_probe(const or_scene* s, int fn, int precision, const double* in, int64_t in_stride, int64_t n, double* out,
             int64_t out_stride) {
  if (fn < 0 || fn >= OR_PROBE_COUNT || n < 0) return -1;
  if (fn == OR_PROBE_SURFACE && s && (s->instanced ? !s->built : s->cnodes64.empty())) return -1;
  for (int64_t i = 0; i < n; ++i) {
    const int rc = precision == 32 ? probe_one<float>(s, fn, in + i * in_stride, out + i * out_stride)
                                   : probe_one<double>(s, fn, in + i * in_stride, out + i * out_stride);
    if (rc) return rc;
  }
  return 0;
}

void or_set_mutation(int k) { g_mut = (k > 0 && k < OR_MUT_COUNT) ? k : 0; }
int or_mutation_count(void) { return OR_MUT_COUNT; }

or_scene* or_scene_create(const pk_scene_desc* d) {
  auto* s = new or_scene();
  s->nv = d->num_vertices;
  s->nt = d->num_triangles;
  s->nm = d->num_materials;
  s->pos.assign(d->positions, d->positions + (size_t)s->nv * 3);
  if (d->attributes) s->attr.assign(d->attributes, d->attributes + (size_t)s->nv * 4);
  else s->attr.assign((size_t)s->nv * 4, 0u);
  s->tri.assign(d->triangles, d->triangles + (size_t)s->nt * 4);
  if (d->materials) s->mat.assign(d->materials, d->materials + (size_t)s->nm * 4);
  s->tw = d->tex_width;
  s->th = d->tex_height;
  s->tl = d->tex_layers;
  s->tfmt = d->tex_format;
  size_t ntex = (size_t)s->tw * s->th * s->tl * 4;
  if (d->tex_data && ntex) {
    if (s->tfmt == PK_TEX_RGBA8) s->tex8.assign((const uint8_t*)d->tex_data, (const uint8_t*)d->tex_data + ntex);
    else s->tex32.assign((const float*)d->tex_data, (const float*)d->tex_data + ntex);
  }
  s->ew = d->env_width;
  s->eh = d->env_height;
  s->efmt = d->env_format;
  s->esampler = d->env_sampler;
  size_t nenv = (size_t)s->ew * s->eh;
  if (s->efmt == PK_ENV_RGBE8) s->env8.assign((const uint8_t*)d->env_data, (const uint8_t*)d->env_data + nenv * 4);
  if (s->efmt == PK_ENV_RGB32F) s->env32.assign((const float*)d->env_data, (const float*)d->env_data + nenv * 3);
  if (s->efmt != PK_ENV_NONE) build_env(s);
  return s;
}

void or_scene_destroy(or_scene* s) { delete s; }

int or_scene_set_instances(or_scene* s, const pk_mesh_desc* meshes, uint32_t num_meshes,
                           const pk_instance_desc* instances, uint32_t num_instances) {
  s->instanced = true;
  s->meshes.clear();
  s->insts.clear();
  auto sub = [](uint32_t nv, const double* pos, const uint32_t* attr, uint32_t nt, const int32_t* tri) {
    pk_scene_desc d;
    std::memset(&d, 0, sizeof(d));
    d.num_vertices = nv;
    d.positions = pos;
    d.attributes = attr;
    d.num_triangles = nt;
    d.triangles = tri;
    d.env_format = PK_ENV_NONE;
    return std::unique_ptr<or_scene>(or_scene_create(&d));
  };
  uint32_t first = 0;
  s->own = -1;
  if (s->nt > 0) {  // the scene's own triangles (traced in the scene's space, not placed)
    s->meshes.push_back(sub(s->nv, s->pos.data(), s->attr.data(), s->nt, s->tri.data()));
    s->own = 0;
    first = 1;
  }
  for (uint32_t m = 0; m < num_meshes; ++m)
    s->meshes.push_back(sub(meshes[m].num_vertices, meshes[m].positions, meshes[m].attributes,
                            meshes[m].num_triangles, meshes[m].triangles));
  for (uint32_t i = 0; i < num_instances; ++i) {
    if (instances[i].mesh >= num_meshes) return -1;
    or_scene::Inst I;
    I.mesh = first + instances[i].mesh;
    I.material = instances[i].material;
    std::memcpy(I.t, instances[i].transform, sizeof(I.t));
    s->insts.push_back(I);
  }
  return (s->insts.empty() && s->own < 0) ? -1 : 0;
}

int or_bvh_build(or_scene* s, int num_threads);

// Instanced build: each mesh's BVH and normalizer as for a scene of its own;
// the scene's own triangles in the scene's normalized space (the normalizer
// over them and the placements' boxes); per placement the scene -> mesh map
// A = (s_m / s) L^-1 with L^-1 = adj(L) / det(L), offset
// b = s_m ((L^-1 (c - t)) - c_m), and mesh -> scene directions (s / s_m) L,
// composed in double and rounded to f32 once.
static int build_instanced(or_scene* s, int num_threads) {
  // the scene's box: its own triangles' vertices and the placements' meshes'
  // triangle vertices mapped by x = ((T0 x + T1 y) + T2 z) + T3
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (size_t v = 0; v < (size_t)s->nt * 3; ++v) {
    const double* p = &s->pos[(size_t)s->tri[(v / 3) * 4 + v % 3] * 3];
    for (int k = 0; k < 3; ++k) {
      lo[k] = std::min(lo[k], p[k]);
      hi[k] = std::max(hi[k], p[k]);
    }
  }
  for (const auto& I : s->insts) {
    const or_scene& m = *s->meshes[I.mesh];
    if (m.nt == 0) return -1;
    for (size_t v = 0; v < (size_t)m.nt * 3; ++v) {
      const double* p = &m.pos[(size_t)m.tri[(v / 3) * 4 + v % 3] * 3];
      for (int k = 0; k < 3; ++k) {
        const double x = ((I.t[k * 4] * p[0] + I.t[k * 4 + 1] * p[1]) + I.t[k * 4 + 2] * p[2]) + I.t[k * 4 + 3];
        lo[k] = std::min(lo[k], x);
        hi[k] = std::max(hi[k], x);
      }
    }
  }
  double span[3];
  for (int k = 0; k < 3; ++k) span[k] = hi[k] - lo[k];
  const double longest = std::max(std::max(span[0], span[1]), span[2]);
  for (int k = 0; k < 3; ++k) s->norm_center[k] = lo[k] + span[k] * 0.5;
  s->norm_scale = 2 / longest;
  const double sc = s->norm_scale;
  // the scene's own triangles keep the scene's normalizer (the baked scene's
  // f32 vertices); every mesh gets its own
  for (size_t mi = 0; mi < s->meshes.size(); ++mi) {
    or_scene* m = s->meshes[mi].get();
    if ((int32_t)mi == s->own) {
      m->force_norm = true;
      for (int k = 0; k < 3; ++k) m->force_center[k] = s->norm_center[k];
      m->force_scale = sc;
    }
    if (m->nt > 0 && or_bvh_build(m, num_threads) < -1) return -2;
  }
  int32_t id_base = (int32_t)s->nt;
  for (auto& I : s->insts) {
    const or_scene& m = *s->meshes[I.mesh];
    const double* t = I.t;
    const double a = t[0], b = t[1], c = t[2], d = t[4], e = t[5], f = t[6], g = t[8], h = t[9], i = t[10];
    const double A = e * i - f * h, B = f * g - d * i, C = d * h - e * g;
    const double det = (a * A + b * B) + c * C;
    if (!(std::fabs(det) > 0.0)) return -1;
    const double r = 1.0 / det;
    const double Li[9] = {A * r, (c * h - b * i) * r, (b * f - c * e) * r,
                          B * r, (a * i - c * g) * r, (c * d - a * f) * r,
                          C * r, (b * g - a * h) * r, (a * e - b * d) * r};
    const double sm = m.norm_scale, k1 = sm / sc, k2 = sc / sm;
    const double u[3] = {s->norm_center[0] - t[3], s->norm_center[1] - t[7], s->norm_center[2] - t[11]};
    for (int row = 0; row < 3; ++row) {
      for (int col = 0; col < 3; ++col) {
        I.w2o64[row * 4 + col] = k1 * Li[row * 3 + col];
        I.o2w64[row * 3 + col] = k2 * t[row * 4 + col];
      }
      const double v = (Li[row * 3] * u[0] + Li[row * 3 + 1] * u[1]) + Li[row * 3 + 2] * u[2];
      I.w2o64[row * 4 + 3] = sm * (v - m.norm_center[row]);
    }
    for (int k = 0; k < 12; ++k) I.w2o32[k] = (float)I.w2o64[k];
    for (int k = 0; k < 9; ++k) I.o2w32[k] = (float)I.o2w64[k];
    I.id_base = id_base;
    id_base += (int32_t)m.nt;
  }
  s->built = true;
  return 0;
}

int or_bvh_build(or_scene* s, int num_threads) {
  if (s->instanced) return build_instanced(s, num_threads);
  const uint32_t nt = s->nt;
  std::vector<Box> ref_box(nt);
  std::vector<std::array<double, 3>> centroid(nt);
  for (uint32_t t = 0; t < nt; ++t) {
    Box b;
    for (int v = 0; v < 3; ++v) {  // BoundingBox.addTriangle (primitives.js:17-23)
      const double* p = &s->pos[(size_t)s->tri[(size_t)t * 4 + v] * 3];
      for (int k = 0; k < 3; ++k) {
        b.lo[k] = std::min(p[k], b.lo[k]);
        b.hi[k] = std::max(p[k], b.hi[k]);
      }
    }
    ref_box[t] = b;
    for (int k = 0; k < 3; ++k) centroid[t][k] = (b.lo[k] + b.hi[k]) * 0.5;  // centroid getter (:28-33)
  }
  // Node.sortRefs (bvh.js:150-168): stable sort by centroid per axis
  std::array<std::vector<uint32_t>, 3> axis_refs;
  for (int axis = 0; axis < 3; ++axis) {
    axis_refs[axis].resize(nt);
    std::iota(axis_refs[axis].begin(), axis_refs[axis].end(), 0u);
    std::stable_sort(axis_refs[axis].begin(), axis_refs[axis].end(),
                     [&](uint32_t a, uint32_t b) { return centroid[a][axis] < centroid[b][axis]; });
  }
  Builder b;
  b.ref_box = &ref_box;
  b.centroid = &centroid;
  b.in_left.assign(nt, 0);
  int depth = 0;
  int par = 0;
  for (int t = std::max(1, num_threads); t > 1; t >>= 1) ++par;
  s->root = b.build(std::move(axis_refs), 1, par, &depth);
  s->degenerate = b.degenerate;
  // stats
  int64_t nodes = 0, leaf_tris = 0, largest = 0;
  std::vector<const ONode*> st{s->root.get()};
  while (!st.empty()) {
    const ONode* n = st.back();
    st.pop_back();
    ++nodes;
    if (n->leaf) {
      leaf_tris += (int64_t)n->leaf_tris.size();
      largest = std::max(largest, (int64_t)n->leaf_tris.size());
    } else {
      st.push_back(n->left.get());
      st.push_back(n->right.get());
    }
  }
  s->bvh_nodes = nodes;
  s->bvh_leaf_tris = leaf_tris;
  s->bvh_depth = depth;
  s->bvh_largest_leaf = largest;
  make_normalizer(s);
  serialize(s);
  build_canonical<float>(s, s->vf, s->cnodes32, &s->ctris);
  build_canonical<double>(s, s->vd, s->cnodes64, nullptr);
  return s->degenerate ? -1 : 0;
}

void or_bvh_info(const or_scene* s, int64_t* info) {
  info[0] = s->bvh_nodes;
  info[1] = s->bvh_leaf_tris;
  info[2] = s->bvh_depth;
  info[3] = s->bvh_largest_leaf;
}

int or_bvh_serialize(const or_scene* s, int32_t* nodes5, int32_t* tris4, float* verts4, double* norm4) {
  if (!s->root) return -1;
  for (size_t i = 0; i < s->rnodes.size(); ++i) {
    const RefNode& r = s->rnodes[i];
    nodes5[i * 5 + 0] = r.child;
    nodes5[i * 5 + 1] = r.tris;
    nodes5[i * 5 + 2] = (int32_t)r.x;
    nodes5[i * 5 + 3] = (int32_t)r.y;
    nodes5[i * 5 + 4] = (int32_t)r.z;
  }
  std::memcpy(tris4, s->rtris.data(), s->rtris.size() * 4);
  for (size_t i = 0; i < s->nv; ++i) {
    verts4[i * 4 + 0] = s->vf[i * 3 + 0];
    verts4[i * 4 + 1] = s->vf[i * 3 + 1];
    verts4[i * 4 + 2] = s->vf[i * 3 + 2];
    verts4[i * 4 + 3] = 0.0f;
  }
  if (norm4) {
    for (int k = 0; k < 3; ++k) norm4[k] = s->norm_center[k];
    norm4[3] = s->norm_scale;
  }
  return 0;
}

int or_trace(const or_scene* s, const float* rays, int64_t n, int mode, int any_hit, int num_threads, int32_t* ids,
             float* t, uint32_t* bary, uint32_t* flags) {
  if (s->instanced ? (!s->built || mode == 0) : !s->root) return -1;
  parallel_for(n, num_threads, [&](int64_t i) {
    const float* r = &rays[i * 8];
    V3<float> o = v3(r[0], r[1], r[2]), d = v3(r[4], r[5], r[6]);
    uint32_t fl = 0;
    if (mode == 0) {
      LitResult lr = literal_trace(s, o, d, any_hit != 0);
      ids[i] = lr.index < 0 ? -1 : (int32_t)s->rleaf_to_orig[lr.index];
      t[i] = lr.t;
      bary[i] = lr.bary;
      fl = lr.overflow ? 1u : 0u;
    } else {
      HitT<float> h = mode == 1 ? scene_brute<float>(s, o, d, any_hit != 0)
                                : scene_trace<float>(s, o, d, any_hit != 0);
      ids[i] = h.id;
      t[i] = h.t;
      bary[i] = h.bary;
    }
    if (flags) flags[i] = fl;
  });
  return 0;
}

int or_env_info(const or_scene* s, int64_t* info) {
  info[0] = (int64_t)s->ebins.size();
  info[1] = (int64_t)s->ecoords.size();
  info[2] = s->ew;
  info[3] = s->eh;
  return 0;
}

int or_env_tables(const or_scene* s, int32_t* bins2, int32_t* coords2, float* pdf) {
  for (size_t i = 0; i < s->ebins.size(); ++i) {
    bins2[i * 2] = s->ebins[i][0];
    bins2[i * 2 + 1] = s->ebins[i][1];
  }
  for (size_t i = 0; i < s->ecoords.size(); ++i) {
    coords2[i * 2] = s->ecoords[i][0];
    coords2[i * 2 + 1] = s->ecoords[i][1];
  }
  std::memcpy(pdf, s->epdf32.data(), s->epdf32.size() * 4);
  return 0;
}

int or_render(const or_scene* s, const pk_camera* cam, const or_render_params* p, float* accum, float* color_out,
              uint64_t* counters) {
  if (s->instanced ? !s->built : !s->root) return -1;
  if (p->precision == 64) return render_t<double>(s, cam, p, accum, color_out, counters);
  return render_t<float>(s, cam, p, accum, color_out, counters);
}

int or_scene_write_params(or_scene* s, int which, int64_t offset, const float* src, int64_t count,
                          int also_sampling) {
  if (which == PK_PARAM_ATLAS) {
    if (s->tfmt != PK_TEX_RGBA32F) return -1;
    const int64_t n = (int64_t)s->tex32.size();
    if (offset < 0 || count < 0 || offset + count > n) return -1;
    if (!also_sampling && s->tex32_eval.empty()) s->tex32_eval = s->tex32;
    for (int64_t i = 0; i < count; ++i) {
      if (!s->tex32_eval.empty()) s->tex32_eval[offset + i] = src[i];
      if (also_sampling) s->tex32[offset + i] = src[i];
    }
    return 0;
  }
  if (which == PK_PARAM_ENV) {
    if (s->efmt == PK_ENV_NONE) return -1;
    const int64_t n = (int64_t)s->ew * s->eh * 4;
    if (offset < 0 || count < 0 || offset + count > n) return -1;
    for (int64_t i = 0; i < count; ++i) {
      const int64_t o = offset + i, texel = o / 4, c = o % 4;
      if (c == 3) continue;
      s->erad32[(size_t)texel * 3 + c] = src[i];
      s->erad64[(size_t)texel * 3 + c] = (double)src[i];
    }
    return 0;
  }
  return -1;
}

int or_gradient(const or_scene* s, const pk_camera* cam, const or_render_params* p, const float* d_image,
                const int32_t* which, const int64_t* index, int64_t n, double* out, double* value) {
  if ((s->instanced ? !s->built : s->cnodes64.empty()) || p->num_frames == 0) return -1;
  if (p->precision == 32) return gradient_t<float>(s, cam, p, d_image, which, index, n, out, value);
  return gradient_t<double>(s, cam, p, d_image, which, index, n, out, value);
}

void or_math_eval(int fn, const float* x, const float* y, float* out, int64_t n) {
  for (int64_t i = 0; i < n; ++i) {
    switch (fn) {
      case 0: out[i] = pkm::sinf_(x[i]); break;
      case 1: out[i] = pkm::cosf_(x[i]); break;
      case 2: out[i] = pkm::tanf_(x[i]); break;
      case 3: out[i] = pkm::atan2f_(x[i], y[i]); break;
      case 4: out[i] = pkm::acosf_(x[i]); break;
      case 5: out[i] = pkm::log2f_(x[i]); break;
      case 6: out[i] = pkm::exp2f_(x[i]); break;
      case 7: out[i] = pkm::powf_(x[i], y[i]); break;
      case 8: out[i] = pkm::sqrt_(x[i]); break;
      case 9: out[i] = x[i] / y[i]; break;
      case 10: {
        uint32_t seed = pkm::f2u(x[i]);
        out[i] = pkm::u2f(pkm::pcg_next(&seed));
        break;
      }
      // the kernels' integer unpacking (x = the packed integer's value)
      case 11: out[i] = pkm::unpack_unorm16((uint32_t)x[i]); break;
      case 12: out[i] = pkm::unpack_snorm16((uint32_t)(int32_t)x[i] & 0xffffu); break;
      case 13: out[i] = pkm::unpack_snorm8((uint32_t)(int32_t)x[i] & 0xffu); break;
      case 14: out[i] = pkm::unpack_unorm8((uint32_t)x[i]); break;
      default: out[i] = 0;
    }
  }
}

}  // extern "C"
