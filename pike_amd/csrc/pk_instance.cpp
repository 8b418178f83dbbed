// pk_instance.cpp — two-level BVH for instanced scenes.
//
// The reference bakes every prop into one triangle list (scene.js:159-170)
// and lists instancing as a TODO (README.md:20).  Here each mesh gets its own
// wide BVH, built by pk_bvh_build_ex in the mesh's own normalized space.  The
// top level is the scene's own triangles' BVH, in the scene's normalized
// space exactly as a baked scene would have it, with the placements inserted
// as leaves of their own (one placement per leaf); a scene with no triangles
// of its own gets an SAH over the placements' boxes instead.
//
// Transforms, all composed in double and rounded to f32 once:
//   scene normalizer  x_n = (x - c) * s   (trace_pass.js:202-211 over the
//                     union of the instances' scene-space boxes)
//   mesh normalizer   x_m = (x_o - c_m) * s_m   (the mesh BVH's own)
//   instance          x = L x_o + t
// so the ray map scene -> mesh space is x_m = A x_n + b with
//   A = (s_m / s) L^-1,   b_i = s_m * ((L^-1 (c - t))_i - c_m,i),
// L^-1 = adj(L) / det(L), and shading maps mesh-space directions back with
// (s / s_m) L.  The oracle composes the same expressions in the same order.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <deque>
#include <exception>
#include <memory>
#include <new>
#include <numeric>
#include <queue>
#include <string>
#include <vector>

#include "pk_internal.h"

namespace {

struct Inv3 {
  double m[9];  // row-major
};

// adjugate / determinant; false if singular or not finite
bool invert3(const double* t, Inv3* out) {
  const double a = t[0], b = t[1], c = t[2];
  const double d = t[4], e = t[5], f = t[6];
  const double g = t[8], h = t[9], i = t[10];
  const double A = e * i - f * h;
  const double B = f * g - d * i;
  const double C = d * h - e * g;
  const double det = (a * A + b * B) + c * C;
  if (!(std::fabs(det) > 0.0) || !std::isfinite(det)) return false;
  const double r = 1.0 / det;
  out->m[0] = A * r;
  out->m[1] = (c * h - b * i) * r;
  out->m[2] = (b * f - c * e) * r;
  out->m[3] = B * r;
  out->m[4] = (a * i - c * g) * r;
  out->m[5] = (c * d - a * f) * r;
  out->m[6] = C * r;
  out->m[7] = (b * g - a * h) * r;
  out->m[8] = (a * e - b * d) * r;
  for (double v : out->m)
    if (!std::isfinite(v)) return false;
  return true;
}

struct DBox {
  double lo[3], hi[3];
};

// Top-level binary node over instance boxes.
struct TNode {
  FBox box;
  int32_t left = -1, right = -1;
  int32_t inst = -1;  // leaf: instance index
};

double farea(const FBox& b) {
  const double x = std::max((double)b.hi[0] - b.lo[0], 0.0), y = std::max((double)b.hi[1] - b.lo[1], 0.0),
               z = std::max((double)b.hi[2] - b.lo[2], 0.0);
  return x * y + x * z + y * z;
}

void fadd(FBox& a, const FBox& b) {
  for (int k = 0; k < 3; ++k) {
    a.lo[k] = std::min(a.lo[k], b.lo[k]);
    a.hi[k] = std::max(a.hi[k], b.hi[k]);
  }
}

FBox fempty() {
  FBox b;
  for (int k = 0; k < 3; ++k) {
    b.lo[k] = INFINITY;
    b.hi[k] = -INFINITY;
  }
  return b;
}

// SAH over instance boxes: full sweep on all three axes (centroid order),
// one instance per leaf.  Instances are few (thousands), so O(n log^2 n);
// an explicit work list keeps degenerate (chain-like) trees off the call
// stack.
int build_top(std::vector<TNode>& out, const std::vector<FBox>& boxes, std::vector<uint32_t> items) {
  struct Work {
    int32_t node;
    std::vector<uint32_t> items;
    int depth;
  };
  int max_depth = 0;
  out.emplace_back();
  std::vector<Work> work;
  work.push_back({0, std::move(items), 1});
  std::vector<double> right_area;
  while (!work.empty()) {
    Work w = std::move(work.back());
    work.pop_back();
    max_depth = std::max(max_depth, w.depth);
    FBox all = fempty();
    for (uint32_t i : w.items) fadd(all, boxes[i]);
    out[w.node].box = all;
    const size_t n = w.items.size();
    if (n == 1) {
      out[w.node].inst = (int32_t)w.items[0];
      continue;
    }
    double best = INFINITY;
    int best_axis = 0;
    size_t best_split = n / 2;
    std::vector<uint32_t> sorted[3];
    right_area.assign(n, 0.0);
    for (int a = 0; a < 3; ++a) {
      sorted[a] = w.items;
      std::stable_sort(sorted[a].begin(), sorted[a].end(), [&](uint32_t x, uint32_t y) {
        return (boxes[x].lo[a] + boxes[x].hi[a]) < (boxes[y].lo[a] + boxes[y].hi[a]);
      });
      FBox r = fempty();
      for (size_t i = n; i-- > 1;) {
        fadd(r, boxes[sorted[a][i]]);
        right_area[i] = farea(r);
      }
      FBox l = fempty();
      for (size_t i = 1; i < n; ++i) {
        fadd(l, boxes[sorted[a][i - 1]]);
        const double cost = farea(l) * (double)i + right_area[i] * (double)(n - i);
        if (cost < best) {
          best = cost;
          best_axis = a;
          best_split = i;
        }
      }
    }
    const int32_t l = (int32_t)out.size();
    out.emplace_back();
    const int32_t r = (int32_t)out.size();
    out.emplace_back();
    out[w.node].left = l;
    out[w.node].right = r;
    const std::vector<uint32_t>& sa = sorted[best_axis];
    work.push_back({r, std::vector<uint32_t>(sa.begin() + best_split, sa.end()), w.depth + 1});
    work.push_back({l, std::vector<uint32_t>(sa.begin(), sa.begin() + best_split), w.depth + 1});
  }
  return max_depth;
}

// Collapse into wide nodes (open the largest-area internal child until 8
// children); instance leaves become kInstChild children.
int collapse_top(const std::vector<TNode>& t, std::vector<pk::WNode>& w) {
  struct Item {
    int32_t node;
    uint32_t wnode;
    int depth;
  };
  std::deque<Item> q;
  w.emplace_back();
  q.push_back({0, 0, 1});
  int depth = 1;
  while (!q.empty()) {
    const Item it = q.front();
    q.pop_front();
    depth = std::max(depth, it.depth);
    int32_t kids[pk::kWide];
    int nk = 0;
    if (t[it.node].inst >= 0) {
      kids[nk++] = it.node;  // a single instance at the root
    } else {
      kids[nk++] = t[it.node].left;
      kids[nk++] = t[it.node].right;
      while (nk < pk::kWide) {
        int best = -1;
        double best_a = -1;
        for (int i = 0; i < nk; ++i)
          if (t[kids[i]].inst < 0 && farea(t[kids[i]].box) > best_a) {
            best_a = farea(t[kids[i]].box);
            best = i;
          }
        if (best < 0) break;
        const int32_t x = kids[best];
        for (int i = nk; i > best + 1; --i) kids[i] = kids[i - 1];
        kids[best] = t[x].left;
        kids[best + 1] = t[x].right;
        ++nk;
      }
    }
    FBox cb[pk::kWide];
    pk::WNode node;
    std::memset(&node, 0, sizeof(node));
    node.nchild = (uint8_t)nk;
    for (int i = 0; i < nk; ++i) {
      cb[i] = t[kids[i]].box;
      node.meta[i] = 0;
      if (t[kids[i]].inst >= 0) {
        node.child[i] = pk::kInstChild | (uint32_t)t[kids[i]].inst;
      } else {
        const uint32_t wi = (uint32_t)w.size();
        w.emplace_back();
        q.push_back({kids[i], wi, it.depth + 1});
        node.child[i] = wi;
      }
    }
    pk_quantize_node(node, cb, nk);
    w[it.wnode] = node;
  }
  return depth;
}

struct MeshIn {
  uint32_t nv, nt;
  const double* pos;
  const uint32_t* attrs;
  const int32_t* tris;
};

// Inserts placement leaves into a binary BVH in place: each goes beside the
// node where it grows the tree's total surface area least (the insertion
// search of Bittner et al. 2013: best-first over the area growth inherited
// from the ancestors, bounded below by the placement's own area), then the
// ancestors are refit.  The root stays at index 0.
void insert_placements(std::vector<BNode>& nodes, const std::vector<DBox>& boxes) {
  auto area = [](const double* lo, const double* hi) {
    const double x = std::max(hi[0] - lo[0], 0.0), y = std::max(hi[1] - lo[1], 0.0), z = std::max(hi[2] - lo[2], 0.0);
    return x * y + x * z + y * z;
  };
  auto merged = [](const BNode& n, const DBox& b, double* lo, double* hi) {
    for (int k = 0; k < 3; ++k) {
      lo[k] = std::min(n.lo[k], b.lo[k]);
      hi[k] = std::max(n.hi[k], b.hi[k]);
    }
  };
  std::vector<int32_t> parent(nodes.size(), -1);
  for (size_t i = 0; i < nodes.size(); ++i)
    if (!nodes[i].leaf) {
      parent[(size_t)nodes[i].left] = (int32_t)i;
      parent[(size_t)nodes[i].right] = (int32_t)i;
    }
  for (size_t p = 0; p < boxes.size(); ++p) {
    const DBox& b = boxes[p];
    const double ab = area(b.lo, b.hi);
    using Item = std::pair<double, int32_t>;  // inherited growth, node
    std::priority_queue<Item, std::vector<Item>, std::greater<Item>> pq;
    pq.push({0.0, 0});
    double best_cost = INFINITY;
    int32_t best = 0;
    while (!pq.empty()) {
      const Item it = pq.top();
      pq.pop();
      if (it.first + ab >= best_cost) break;
      const BNode& x = nodes[(size_t)it.second];
      double lo[3], hi[3];
      merged(x, b, lo, hi);
      const double u = area(lo, hi);
      if (it.first + u < best_cost) {
        best_cost = it.first + u;
        best = it.second;
      }
      if (!x.leaf) {
        const double inherited = it.first + u - area(x.lo, x.hi);
        if (inherited + ab < best_cost) {
          pq.push({inherited, x.left});
          pq.push({inherited, x.right});
        }
      }
    }
    BNode leaf;
    for (int k = 0; k < 3; ++k) {
      leaf.lo[k] = b.lo[k];
      leaf.hi[k] = b.hi[k];
    }
    leaf.leaf = 1;
    leaf.count = 0;
    leaf.inst = (int32_t)p;
    const int32_t li = (int32_t)nodes.size();
    nodes.push_back(leaf);
    parent.push_back(-1);
    int32_t sib = best;
    if (sib == 0) {  // beside the root: the old root moves out of index 0
      const int32_t moved = (int32_t)nodes.size();
      nodes.push_back(nodes[0]);
      parent.push_back(0);
      if (!nodes[(size_t)moved].leaf) {
        parent[(size_t)nodes[(size_t)moved].left] = moved;
        parent[(size_t)nodes[(size_t)moved].right] = moved;
      }
      sib = moved;
      BNode root;
      merged(nodes[(size_t)moved], b, root.lo, root.hi);
      root.leaf = 0;
      root.left = moved;
      root.right = li;
      nodes[0] = root;
      parent[(size_t)li] = 0;
      continue;
    }
    const int32_t gp = parent[(size_t)sib];
    const int32_t pi = (int32_t)nodes.size();
    BNode pn;
    merged(nodes[(size_t)sib], b, pn.lo, pn.hi);
    pn.leaf = 0;
    pn.left = sib;
    pn.right = li;
    nodes.push_back(pn);
    parent.push_back(gp);
    if (nodes[(size_t)gp].left == sib) nodes[(size_t)gp].left = pi;
    else nodes[(size_t)gp].right = pi;
    parent[(size_t)sib] = pi;
    parent[(size_t)li] = pi;
    for (int32_t a = gp; a >= 0; a = parent[(size_t)a]) {
      BNode& n = nodes[(size_t)a];
      const BNode& l = nodes[(size_t)n.left];
      const BNode& r = nodes[(size_t)n.right];
      for (int k = 0; k < 3; ++k) {
        n.lo[k] = std::min(l.lo[k], r.lo[k]);
        n.hi[k] = std::max(l.hi[k], r.hi[k]);
      }
    }
  }
}

// Every child reference of the merged wide layout points inside it: internal
// children at nodes, leaf ranges inside the triangle records, placement
// children at placements (top level only), and each node's leaf mask agrees
// with its leaf counts -- the kernels follow them unchecked.
bool wide_layout_ok(const pk_bvh& b, uint32_t top_nodes) {
  const size_t nn = b.wnodes.size(), nt = b.wtris.size();
  for (size_t i = 0; i < nn; ++i) {
    const pk::WNode& w = b.wnodes[i];
    if (w.nchild == 0 || w.nchild > pk::kWide) return false;
    uint32_t lm = 0;
    for (int c = 0; c < w.nchild; ++c)
      if (w.meta[c]) lm |= 1u << c;
    if (w.leafmask != lm) return false;
    for (int c = 0; c < w.nchild; ++c) {
      const uint32_t ch = w.child[c];
      if (w.meta[c]) {
        if ((size_t)ch + w.meta[c] > nt) return false;
      } else if (ch & pk::kInstChild) {
        if (i >= top_nodes || (ch & ~pk::kInstChild) >= b.inst.size()) return false;
      } else if (ch >= nn) {
        return false;
      }
    }
  }
  for (const pk::InstRecord& r : b.inst)
    if (r.root < top_nodes || r.root >= nn) return false;
  return true;
}

}  // namespace

void pk_bvh_make_wide(pk_bvh* bvh, int threads);

static int build_instanced(const pk_scene_desc* scene, const pk_mesh_desc* meshes, uint32_t num_meshes,
                           const pk_instance_desc* instances, uint32_t num_instances, int num_threads, uint32_t flags,
                           pk_bvh** out) {
  if (!scene || !out || (num_meshes && !meshes) || (num_instances && !instances)) {
    pk_set_error("pk_bvh_build_instanced: null argument");
    return PK_E_INVALID;
  }
  *out = nullptr;
  const bool has_world = scene->num_triangles > 0;
  if (!has_world && num_instances == 0) {
    pk_set_error("pk_bvh_build_instanced: no triangles and no instances");
    return PK_E_INVALID;
  }
  if (has_world && (!scene->positions || !scene->triangles)) {
    pk_set_error("pk_bvh_build_instanced: scene triangles without positions");
    return PK_E_INVALID;
  }
  if (num_instances >= (1u << 31)) {
    pk_set_error("pk_bvh_build_instanced: more than 2^31 instances");
    return PK_E_LIMIT;
  }
  std::vector<MeshIn> mesh(num_meshes);
  for (uint32_t m = 0; m < num_meshes; ++m) {
    const pk_mesh_desc& d = meshes[m];
    mesh[m] = {d.num_vertices, d.num_triangles, d.positions, d.attributes, d.triangles};
  }
  struct Inst {
    uint32_t mesh;
    int32_t material;
    double t[12];
  };
  std::vector<Inst> inst(num_instances);
  std::vector<char> used(num_meshes, 0);
  for (uint32_t i = 0; i < num_instances; ++i) {
    const pk_instance_desc& d = instances[i];
    if (d.mesh >= num_meshes) {
      pk_set_error("pk_bvh_build_instanced: instance " + std::to_string(i) + " references a missing mesh");
      return PK_E_INVALID;
    }
    inst[i].mesh = d.mesh;
    inst[i].material = d.material;
    std::memcpy(inst[i].t, d.transform, sizeof(inst[i].t));
    for (double v : inst[i].t)
      if (!std::isfinite(v)) {
        pk_set_error("pk_bvh_build_instanced: instance " + std::to_string(i) + " has a non-finite transform");
        return PK_E_INVALID;
      }
    used[d.mesh] = 1;
  }
  // the scene's own triangles' BVH and the meshes' (each normalized on its own
  // for now; the scene's own is renormalized below)
  std::unique_ptr<pk_bvh> world;
  if (has_world) {
    pk_bvh* b = nullptr;
    const int rc = pk_bvh_build_ex(scene->positions, scene->num_vertices, scene->triangles, scene->num_triangles,
                                   num_threads, flags, &b);
    if (rc) return rc;
    world.reset(b);
  }
  std::vector<std::unique_ptr<pk_bvh>> mb(num_meshes);
  for (uint32_t m = 0; m < num_meshes; ++m) {
    if (!used[m]) continue;
    pk_bvh* b = nullptr;
    const int rc = pk_bvh_build_ex(mesh[m].pos, mesh[m].nv, mesh[m].tris, mesh[m].nt, num_threads, flags, &b);
    if (rc) return rc;  // the error names the problem (empty mesh, bad index, ...)
    mb[m].reset(b);
  }
  // scene-space boxes of the placements over their meshes' triangle vertices;
  // the scene's box also holds its own triangles (the root of their BVH)
  const size_t ni = inst.size();
  std::vector<DBox> wbox(ni);
  DBox all;
  for (int k = 0; k < 3; ++k) {
    all.lo[k] = has_world ? world->nodes[0].lo[k] : INFINITY;
    all.hi[k] = has_world ? world->nodes[0].hi[k] : -INFINITY;
  }
  for (size_t i = 0; i < ni; ++i) {
    const MeshIn& m = mesh[inst[i].mesh];
    const double* T = inst[i].t;
    DBox& b = wbox[i];
    for (int k = 0; k < 3; ++k) {
      b.lo[k] = INFINITY;
      b.hi[k] = -INFINITY;
    }
    for (size_t t = 0; t < (size_t)m.nt * 3; ++t) {
      const double* p = m.pos + (size_t)m.tris[(t / 3) * 4 + t % 3] * 3;
      for (int k = 0; k < 3; ++k) {
        const double x = ((T[k * 4] * p[0] + T[k * 4 + 1] * p[1]) + T[k * 4 + 2] * p[2]) + T[k * 4 + 3];
        b.lo[k] = std::min(b.lo[k], x);
        b.hi[k] = std::max(b.hi[k], x);
      }
    }
    for (int k = 0; k < 3; ++k) {
      all.lo[k] = std::min(all.lo[k], b.lo[k]);
      all.hi[k] = std::max(all.hi[k], b.hi[k]);
    }
  }
  auto* bvh = new pk_bvh();
  std::unique_ptr<pk_bvh> guard(bvh);
  bvh->instanced = true;
  bvh->spatial_splits = (flags & PK_BVH_SPATIAL_SPLITS) != 0;
  bvh->sah_collapse = (flags & PK_BVH_SAH_COLLAPSE) != 0;
  bvh->nt = scene->num_triangles;
  bvh->nv = scene->num_vertices;
  {
    double span[3];
    for (int k = 0; k < 3; ++k) span[k] = all.hi[k] - all.lo[k];
    const double longest = std::max(std::max(span[0], span[1]), span[2]);
    if (!(longest > 0)) {
      pk_set_error("pk_bvh_build_instanced: the scene has no extent");
      return PK_E_INVALID;
    }
    for (int k = 0; k < 3; ++k) bvh->norm_center[k] = all.lo[k] + span[k] * 0.5;
    bvh->norm_scale = 2 / longest;
  }
  const double s = bvh->norm_scale;
  const double* c = bvh->norm_center;
  const uint32_t world_leaves = has_world ? (uint32_t)world->wtris.size() : 0u;
  // placement records: transforms, baked id and virtual leaf numbering (the
  // scene's own triangles and their records come first)
  std::vector<FBox> nbox(ni);
  bvh->inst.resize(ni);
  uint64_t id_base = scene->num_triangles, leaf_base = world_leaves;
  for (size_t i = 0; i < ni; ++i) {
    const pk_bvh& m = *mb[inst[i].mesh];
    Inv3 L;
    if (!invert3(inst[i].t, &L)) {
      pk_set_error("pk_bvh_build_instanced: instance transform " + std::to_string(i) + " is not invertible");
      return PK_E_INVALID;
    }
    pk::InstRecord& r = bvh->inst[i];
    std::memset(&r, 0, sizeof(r));
    const double sm = m.norm_scale;
    const double k1 = sm / s, k2 = s / sm;
    const double u[3] = {c[0] - inst[i].t[3], c[1] - inst[i].t[7], c[2] - inst[i].t[11]};
    for (int row = 0; row < 3; ++row) {
      for (int col = 0; col < 3; ++col) {
        r.w2o[row * 4 + col] = (float)(k1 * L.m[row * 3 + col]);
        r.o2w[row * 3 + col] = (float)(k2 * inst[i].t[row * 4 + col]);
      }
      const double v = (L.m[row * 3] * u[0] + L.m[row * 3 + 1] * u[1]) + L.m[row * 3 + 2] * u[2];
      r.w2o[row * 4 + 3] = (float)(sm * (v - m.norm_center[row]));
    }
    r.material = inst[i].material;
    r.id_delta = (uint32_t)id_base;
    r.leaf_base = (uint32_t)leaf_base;
    id_base += mesh[inst[i].mesh].nt;
    leaf_base += m.wtris.size();
    if (id_base >= (1ull << 31) || leaf_base >= 0xffffffffull) {
      pk_set_error("pk_bvh_build_instanced: the baked scene would exceed 2^31 triangles");
      return PK_E_LIMIT;
    }
    // normalized box, rounded outward to f32
    for (int k = 0; k < 3; ++k) {
      const double lo = (wbox[i].lo[k] - c[k]) * s, hi = (wbox[i].hi[k] - c[k]) * s;
      float fl = (float)lo, fh = (float)hi;
      if ((double)fl > lo) fl = std::nextafter(fl, -INFINITY);
      if ((double)fh < hi) fh = std::nextafter(fh, INFINITY);
      nbox[i].lo[k] = fl;
      nbox[i].hi[k] = fh;
    }
  }
  bvh->baked_triangles = (int64_t)id_base;
  // top level
  int top_depth = 0, top_wide_depth = 0;
  if (has_world) {
    // the scene's own triangles in the scene's normalized space (the same f32
    // vertices a baked scene has), placements inserted beside them
    pk_bvh& w = *world;
    for (int k = 0; k < 3; ++k) w.norm_center[k] = c[k];
    w.norm_scale = s;
    for (size_t i = 0; i < (size_t)w.nv * 3; ++i) w.verts[i] = (float)((scene->positions[i] - c[i % 3]) * s);
    w.inst_boxes = nbox;
    insert_placements(w.nodes, wbox);
    pk_bvh_make_wide(&w, num_threads);
    bvh->wnodes.swap(w.wnodes);
    bvh->wtris.swap(w.wtris);
    bvh->wtri_index.swap(w.wtri_index);
    top_depth = w.binary_depth + 1;
    top_wide_depth = w.wide_depth;
    bvh->num_binary += w.num_binary + 2 * (int64_t)ni;
    bvh->largest_leaf = w.largest_leaf;
    bvh->num_spatial_splits += w.num_spatial_splits;
    bvh->degenerate |= w.degenerate;
  } else {
    std::vector<TNode> top;
    std::vector<uint32_t> items(ni);
    std::iota(items.begin(), items.end(), 0u);
    top_depth = build_top(top, nbox, std::move(items));
    top_wide_depth = collapse_top(top, bvh->wnodes);
    bvh->num_binary += (int64_t)top.size();
  }
  bvh->world_leaves = (uint32_t)bvh->wtris.size();
  if (bvh->world_leaves != world_leaves) {
    pk_set_error("pk_bvh_build_instanced: internal error (top-level triangle records changed)");
    return PK_E_INVALID;
  }
  const uint32_t top_nodes = (uint32_t)bvh->wnodes.size();
  // append the meshes' wide BVHs; mesh vertices follow the scene's own
  std::vector<uint32_t> root(num_meshes, 0), toff(num_meshes, 0);
  uint32_t vnext = scene->num_vertices;
  int blas_depth = 0, blas_wide_depth = 0;
  for (uint32_t m = 0; m < num_meshes; ++m) {
    if (!used[m]) continue;
    const pk_bvh& b = *mb[m];
    const uint32_t noff = (uint32_t)bvh->wnodes.size(), to = (uint32_t)bvh->wtris.size();
    root[m] = noff;
    toff[m] = to;
    for (pk::WNode w : b.wnodes) {
      pk::wnode_rebase(w, noff, to);
      bvh->wnodes.push_back(w);
    }
    bvh->wtris.insert(bvh->wtris.end(), b.wtris.begin(), b.wtris.end());
    for (size_t i = 0; i < b.wtri_index.size(); i += 4) {
      for (int v = 0; v < 3; ++v) bvh->wtri_index.push_back(b.wtri_index[i + v] + (int32_t)vnext);
      bvh->wtri_index.push_back(b.wtri_index[i + 3]);
    }
    const MeshIn& mi = mesh[m];
    if (mi.attrs) bvh->inst_attrs.insert(bvh->inst_attrs.end(), mi.attrs, mi.attrs + (size_t)mi.nv * 4);
    else bvh->inst_attrs.insert(bvh->inst_attrs.end(), (size_t)mi.nv * 4, 0u);
    vnext += mi.nv;
    bvh->num_binary += b.num_binary;
    blas_depth = std::max(blas_depth, b.binary_depth);
    blas_wide_depth = std::max(blas_wide_depth, b.wide_depth);
    bvh->largest_leaf = std::max(bvh->largest_leaf, b.largest_leaf);
    bvh->num_spatial_splits += b.num_spatial_splits;
    bvh->degenerate |= b.degenerate;
  }
  if (bvh->wnodes.size() >= pk::kInstChild || bvh->wtris.size() >= pk::kInstChild) {
    pk_set_error("pk_bvh_build_instanced: more than 2^31 wide nodes or triangle records");
    return PK_E_LIMIT;
  }
  for (size_t i = 0; i < ni; ++i) {
    pk::InstRecord& r = bvh->inst[i];
    r.root = root[inst[i].mesh];
    r.leaf_delta = r.leaf_base - toff[inst[i].mesh];
  }
  if (!wide_layout_ok(*bvh, top_nodes)) {
    pk_set_error("pk_bvh_build_instanced: internal error (a child reference leaves the layout)");
    return PK_E_INVALID;
  }
  bvh->binary_depth = top_depth + blas_depth;
  bvh->wide_depth = top_wide_depth + blas_wide_depth;
  *out = guard.release();
  return PK_OK;
}

extern "C" int pk_bvh_build_instanced(const pk_scene_desc* scene, const pk_mesh_desc* meshes, uint32_t num_meshes,
                                      const pk_instance_desc* instances, uint32_t num_instances, int num_threads,
                                      uint32_t flags, pk_bvh** out) {
  return pk_guard("pk_bvh_build_instanced", [&]() {
    return build_instanced(scene, meshes, num_meshes, instances, num_instances, num_threads, flags, out);
  });
}
