// pk_internal.h — host-side structures shared by the runtime translation units.
#pragma once
#include <stdint.h>

#include <exception>
#include <new>
#include <string>
#include <vector>

#include "../../include/pk.h"
#include "pk_layout.h"

// Child box of a wide node in normalized space (f32, rounded outward).
struct FBox {
  float lo[3], hi[3];
};

// Binary BVH node (render/bvh.js Node): float64 world-space bounds.
struct BNode {
  double lo[3], hi[3];
  int32_t left = -1, right = -1;  // children (internal nodes)
  uint32_t begin = 0, count = 0;  // leaf: range into pk_bvh::leaf_tris
  uint8_t leaf = 0;
  uint8_t split_axis = 0;
  // Oversized leaves (count > kLeafMax) get virtual halves that only the wide
  // collapse follows; the reference export still sees the original leaf.
  int32_t vleft = -1, vright = -1;
  int32_t inst = -1;  // a placement's leaf (instanced top level), else -1
};

struct pk_bvh {
  uint32_t nv = 0, nt = 0;
  std::vector<int32_t> tris;      // scene.indices [nt][4]
  std::vector<BNode> nodes;       // root at 0
  std::vector<uint32_t> leaf_tris;
  int degenerate = 0;
  int binary_depth = 0, wide_depth = 0, largest_leaf = 0;
  int64_t num_binary = 0;         // nodes of the bvh.js tree (before virtual splits)
  bool spatial_splits = false;    // built with SBVH spatial splits (leaves hold fragments)
  bool sah_collapse = false;      // wide collapse by the SAH-optimal dynamic program
  int64_t num_spatial_splits = 0;
  double norm_center[3] = {0, 0, 0};
  double norm_scale = 1;
  std::vector<float> verts;       // normalized f32 [nv][3]
  // wide layout
  std::vector<pk::WNode> wnodes;
  std::vector<pk::TriRecord> wtris;
  std::vector<int32_t> wtri_index;  // leaf order -> {i0,i1,i2,mat} [n][4]
  // instancing (pk_bvh_build_instanced): wnodes/wtris/wtri_index hold the
  // top-level nodes followed by every mesh's; nt/nv are the scene's own
  bool instanced = false;
  std::vector<pk::InstRecord> inst;   // the placements
  uint32_t world_leaves = 0;          // top-level triangle records (the scene's own), first in wtris
  std::vector<FBox> inst_boxes;       // normalized boxes of placement leaves (by placement), for the collapse
  std::vector<uint32_t> inst_attrs;   // the meshes' vertex attributes [n][4], after the scene's own
  int64_t baked_triangles = 0;        // triangles of the equivalent baked scene
};

// Conservative 8-bit quantization of wide-node child boxes (pk_bvh.cpp).
void pk_quantize_node(pk::WNode& w, const FBox* cb, int nc);

// The 4-wide layout of a linked single-level 8-wide layout (pk_bvh.cpp):
// node i < n8.size() stands for 8-wide node i, sub-nodes follow.
void pk_make_wide4(const std::vector<pk::WNode>& n8, std::vector<pk::WNode4>& n4);

// Environment importance-sampling tables (render/env_sampler.js).
struct EnvTables {
  uint32_t width = 0, height = 0;
  std::vector<float> radiance;      // decoded rgb f32 [w*h][3] (shade_miss lookup)
  std::vector<int32_t> bins;        // [nbins][2] {h0, h1}
  std::vector<int32_t> coords;      // [w*h][2] {x, y}, radiance-descending
  std::vector<float> pdf;           // [w*h]
  // marginal/conditional CDF sampler
  std::vector<float> row_cdf;       // [h+1]
  std::vector<float> col_cdf;       // [h][w+1]
  // cutpoint guides: row_guide[k] = largest y with row_cdf[y] <= k/h,
  // col_guide[y][k] = largest x with col_cdf[y][x] <= k/w (k = 0..h / 0..w)
  std::vector<uint32_t> row_guide;  // [h+1]
  std::vector<uint32_t> col_guide;  // [h][w+1]
};

int pk_build_env_tables(const pk_scene_desc* d, EnvTables* out, std::string* err);
void pk_set_error(const std::string& msg);

// Every C entry point runs its body through pk_guard: no C++ exception leaves
// the C ABI.  Host memory (or threads) running out ends the call with
// PK_E_OOM and the error text; what the call had allocated may stay allocated.
template <class F> int pk_guard(const char* what, F&& body) {
  try {
    return body();
  } catch (const std::bad_alloc&) {
    try {
      pk_set_error(std::string(what) + ": out of host memory");
    } catch (...) {
    }
  } catch (const std::exception& e) {
    try {
      pk_set_error(std::string(what) + ": " + e.what());
    } catch (...) {
    }
  }
  return PK_E_OOM;
}
