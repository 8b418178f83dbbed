// pk_bvh.cpp — SAH binary BVH with render/bvh.js's split decisions, its
// reference flattening (trace_pass.js) and the collapse into the 128-byte
// wide layout the gfx950 kernels traverse.  Host code, no device needed.
//
// Decision contract kept from bvh.js (float64 throughout):
//   * refs presorted by centroid per axis, stable (bvh.js:21, 150-168);
//   * full sweep on axes 0,1,2; cost = 0.5 + A(front_i)/A(P)*(i+1)
//     + A(back_i)/A(P)*(N-1-i), where front_i is the union of refs[0..i] and
//     back_i the union of refs[i..N-1] (bvh.js:184-208 — the back area
//     includes ref i, as the reference computes it); first strict minimum;
//   * leaf iff best cost > N (bvh.js:137), leaf order = refs on the best axis;
//   * children inherit each axis order by stable filtering (bvh.js:80-107).
// The implementation is range-based (three index arrays partitioned in
// place, O(N) scratch per level) instead of bvh.js's per-node copies, and
// builds independent subtrees on threads; the tree is the same.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <deque>
#include <future>
#include <mutex>
#include <numeric>
#include <thread>

#include "pk_internal.h"

namespace {

struct Box {
  double lo[3], hi[3];
};
inline void box_empty(Box& b) {
  for (int k = 0; k < 3; ++k) {
    b.lo[k] = INFINITY;
    b.hi[k] = -INFINITY;
  }
}
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

// Chunked node store with lock-free allocation for the threaded build.
class NodePool {
 public:
  explicit NodePool(size_t max_nodes) : pages_((max_nodes >> kShift) + 1) {}
  uint32_t alloc() {
    uint32_t i = next_.fetch_add(1);
    size_t p = i >> kShift;
    if (!pages_[p].load(std::memory_order_acquire)) {
      std::lock_guard<std::mutex> g(mu_);
      if (!pages_[p].load(std::memory_order_relaxed)) pages_[p].store(new BNode[kPage], std::memory_order_release);
    }
    return i;
  }
  BNode& at(uint32_t i) { return pages_[i >> kShift].load(std::memory_order_relaxed)[i & (kPage - 1)]; }
  uint32_t size() const { return next_.load(); }
  ~NodePool() {
    for (auto& p : pages_) delete[] p.load();
  }

 private:
  static constexpr int kShift = 16;
  static constexpr size_t kPage = (size_t)1 << kShift;
  std::vector<std::atomic<BNode*>> pages_;
  std::atomic<uint32_t> next_{0};
  std::mutex mu_;
};

struct Build {
  const std::vector<Box>* ref_box;
  uint32_t* axis[3];           // three index arrays of length nt
  uint32_t* leaf_out;          // leaf_tris, written per leaf range
  std::vector<uint8_t> in_left;
  NodePool* pool;
  std::atomic<int> degenerate{0};
  std::atomic<int> max_depth{0};

  void sweep(int a, uint32_t b, uint32_t n, double parent_area, double* best_cost, uint32_t* best_index,
             std::vector<double>& suf) {
    const uint32_t* refs = axis[a] + b;
    const std::vector<Box>& rb = *ref_box;
    Box bb;
    box_empty(bb);
    for (uint32_t i = n; i-- > 0;) {
      box_add(bb, rb[refs[i]]);
      suf[i] = box_area(bb);  // union of refs[i..n-1]
    }
    Box bf;
    box_empty(bf);
    double bc = INFINITY;
    uint32_t bi = 0;
    for (uint32_t i = 0; i < n; ++i) {
      box_add(bf, rb[refs[i]]);
      double saf = box_area(bf);
      double cost = 0.5 + (saf / parent_area) * (double)(i + 1) + (suf[i] / parent_area) * (double)(n - 1 - i);
      if (cost < bc) {
        bc = cost;
        bi = i;
      }
    }
    *best_cost = bc;
    *best_index = bi;
  }

  // Builds the subtree over [b, b+n) of all three arrays into node `ni`.
  void node(uint32_t ni, uint32_t b, uint32_t n, int depth, int par_depth) {
    int md = max_depth.load();
    while (depth > md && !max_depth.compare_exchange_weak(md, depth)) {
    }
    const std::vector<Box>& rb = *ref_box;
    Box bounds;
    box_empty(bounds);
    for (uint32_t i = 0; i < n; ++i) box_add(bounds, rb[axis[0][b + i]]);
    BNode& nd = pool->at(ni);
    for (int k = 0; k < 3; ++k) {
      nd.lo[k] = bounds.lo[k];
      nd.hi[k] = bounds.hi[k];
    }
    const double parent_area = box_area(bounds);
    double cost[3];
    uint32_t index[3];
    if (n > (1u << 20)) {
      std::vector<double> s0(n), s1(n), s2(n);
      std::thread t1([&]() { sweep(1, b, n, parent_area, &cost[1], &index[1], s1); });
      std::thread t2([&]() { sweep(2, b, n, parent_area, &cost[2], &index[2], s2); });
      sweep(0, b, n, parent_area, &cost[0], &index[0], s0);
      t1.join();
      t2.join();
    } else {
      std::vector<double> suf(n);
      for (int a = 0; a < 3; ++a) sweep(a, b, n, parent_area, &cost[a], &index[a], suf);
    }
    double best_cost = INFINITY;
    int best_axis = -1;
    uint32_t best_index = 0;
    for (int a = 0; a < 3; ++a)
      if (cost[a] < best_cost) {
        best_cost = cost[a];
        best_axis = a;
        best_index = index[a];
      }
    if (best_axis < 0 || best_cost > (double)n) {
      if (best_axis < 0) {
        degenerate.store(1);
        best_axis = 0;
      }
      nd.leaf = 1;
      nd.split_axis = (uint8_t)best_axis;
      nd.begin = b;
      nd.count = n;
      std::memcpy(leaf_out + b, axis[best_axis] + b, (size_t)n * 4);
      return;
    }
    nd.split_axis = (uint8_t)best_axis;
    const uint32_t nl = best_index + 1;
    uint32_t* split = axis[best_axis] + b;
    for (uint32_t i = 0; i < nl; ++i) in_left[split[i]] = 1;
    std::vector<uint32_t> tmp(n - nl);
    for (int a = 0; a < 3; ++a) {
      if (a == best_axis) continue;
      uint32_t* arr = axis[a] + b;
      uint32_t li = 0, ri = 0;
      for (uint32_t i = 0; i < n; ++i) {
        uint32_t r = arr[i];
        if (in_left[r]) arr[li++] = r;
        else tmp[ri++] = r;
      }
      std::memcpy(arr + nl, tmp.data(), (size_t)ri * 4);
    }
    for (uint32_t i = 0; i < nl; ++i) in_left[split[i]] = 0;
    std::vector<uint32_t>().swap(tmp);
    uint32_t l = pool->alloc(), r = pool->alloc();
    {
      BNode& p = pool->at(ni);
      p.left = (int32_t)l;
      p.right = (int32_t)r;
    }
    if (par_depth > 0 && n > 32768) {
      auto fut = std::async(std::launch::async, [&]() { node(l, b, nl, depth + 1, par_depth - 1); });
      node(r, b + nl, n - nl, depth + 1, par_depth - 1);
      fut.get();
    } else {
      node(l, b, nl, depth + 1, 0);
      node(r, b + nl, n - nl, depth + 1, 0);
    }
  }
};

// JS ToInt16 (ECMAScript ToInt16: truncate, wrap modulo 2^16)
inline int16_t js_to_int16(double x) {
  if (!std::isfinite(x)) return 0;
  double t = std::trunc(x);
  double m = std::fmod(t, 65536.0);
  if (m < 0) m += 65536.0;
  return (int16_t)(uint16_t)(uint32_t)m;
}
// quantizeVec2ToInt (trace_pass.js:213-218): {max, min} -> lo16 | hi16 << 16
inline uint32_t quantize_pair(double a, double b) {
  int16_t s0 = js_to_int16(a * (a > 0 ? 32767 : 32768));
  int16_t s1 = js_to_int16(b * (b > 0 ? 32767 : 32768));
  return (uint32_t)(uint16_t)s0 | ((uint32_t)(uint16_t)s1 << 16);
}

struct FBox {
  float lo[3], hi[3];
};

// Conservative 8-bit quantization of child boxes relative to the node box.
void quantize_node(pk::WNode& w, const FBox* cb, int nc) {
  const double m = pk::kBoxMargin;
  double plo[3] = {INFINITY, INFINITY, INFINITY}, phi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int i = 0; i < nc; ++i)
    for (int k = 0; k < 3; ++k) {
      plo[k] = std::min(plo[k], (double)cb[i].lo[k]);
      phi[k] = std::max(phi[k], (double)cb[i].hi[k]);
    }
  float origin[3];
  int ex[3];
  uint8_t* qlo[3] = {w.qlox, w.qloy, w.qloz};
  uint8_t* qhi[3] = {w.qhix, w.qhiy, w.qhiz};
  for (int k = 0; k < 3; ++k) {
    double lo_m = plo[k] - m;
    float o = (float)lo_m;
    if ((double)o > lo_m) o = std::nextafter(o, -INFINITY);
    origin[k] = o;
    double extent = (phi[k] + m) - (double)o;
    int e;
    std::frexp(extent / 254.0, &e);  // extent/254 < 2^e
    for (;;) {
      double s = std::ldexp(1.0, e);
      bool ok = true;
      for (int i = 0; i < nc && ok; ++i) {
        double a = std::floor(((double)cb[i].lo[k] - m - (double)o) / s);
        double b = std::ceil(((double)cb[i].hi[k] + m - (double)o) / s);
        if (a < 0) a = 0;
        while (a > 0 && (double)o + a * s > (double)cb[i].lo[k] - m) a -= 1;
        while ((double)o + b * s < (double)cb[i].hi[k] + m) b += 1;
        if (b > 255) ok = false;
        qlo[k][i] = (uint8_t)a;
        qhi[k][i] = (uint8_t)std::min(b, 255.0);
      }
      if (ok) break;
      ++e;
    }
    ex[k] = e;
    for (int i = nc; i < pk::kWide; ++i) {
      qlo[k][i] = 255;
      qhi[k][i] = 0;
    }
  }
  w.ox = origin[0];
  w.oy = origin[1];
  w.oz = origin[2];
  w.ex = (uint8_t)(ex[0] + 127);
  w.ey = (uint8_t)(ex[1] + 127);
  w.ez = (uint8_t)(ex[2] + 127);
}

}  // namespace

void pk_bvh_make_wide(pk_bvh* bvh);

extern "C" {

int pk_bvh_build(const double* positions, uint32_t nv, const int32_t* triangles, uint32_t nt, int num_threads,
                 pk_bvh** out) {
  if (!positions || !triangles || !out || nt == 0 || nv == 0) {
    pk_set_error("pk_bvh_build: empty or null input");
    return PK_E_INVALID;
  }
  if (nt >= (1u << 31) || nv >= (1u << 31)) {
    pk_set_error("pk_bvh_build: more than 2^31 triangles or vertices");
    return PK_E_LIMIT;
  }
  for (size_t i = 0; i < (size_t)nt; ++i)
    for (int c = 0; c < 3; ++c)
      if (triangles[i * 4 + c] < 0 || (uint32_t)triangles[i * 4 + c] >= nv) {
        pk_set_error("pk_bvh_build: triangle " + std::to_string(i) + " references a vertex out of range");
        return PK_E_INVALID;
      }
  if (num_threads <= 0) num_threads = (int)std::max(1u, std::thread::hardware_concurrency());
  auto* bvh = new pk_bvh();
  bvh->nv = nv;
  bvh->nt = nt;
  bvh->tris.assign(triangles, triangles + (size_t)nt * 4);
  // reference boxes and centroids (primitives.js:17-33)
  std::vector<Box> ref_box(nt);
  std::vector<double> cen((size_t)nt * 3);
  for (uint32_t t = 0; t < nt; ++t) {
    Box b;
    box_empty(b);
    for (int v = 0; v < 3; ++v) {
      const double* p = positions + (size_t)triangles[(size_t)t * 4 + v] * 3;
      for (int k = 0; k < 3; ++k) {
        b.lo[k] = std::min(p[k], b.lo[k]);
        b.hi[k] = std::max(p[k], b.hi[k]);
      }
    }
    ref_box[t] = b;
    for (int k = 0; k < 3; ++k) cen[(size_t)t * 3 + k] = (b.lo[k] + b.hi[k]) * 0.5;
  }
  std::vector<uint32_t> arr((size_t)nt * 3);
  uint32_t* ax[3] = {arr.data(), arr.data() + nt, arr.data() + 2 * (size_t)nt};
  {
    std::vector<std::thread> ts;
    for (int a = 0; a < 3; ++a)
      ts.emplace_back([&, a]() {
        std::iota(ax[a], ax[a] + nt, 0u);
        std::stable_sort(ax[a], ax[a] + nt,
                         [&](uint32_t x, uint32_t y) { return cen[(size_t)x * 3 + a] < cen[(size_t)y * 3 + a]; });
      });
    for (auto& t : ts) t.join();
  }
  std::vector<double>().swap(cen);
  bvh->leaf_tris.resize(nt);
  NodePool pool((size_t)nt * 2 + 2);
  Build b;
  b.ref_box = &ref_box;
  for (int a = 0; a < 3; ++a) b.axis[a] = ax[a];
  b.leaf_out = bvh->leaf_tris.data();
  b.in_left.assign(nt, 0);
  b.pool = &pool;
  int par = 0;
  for (int t = num_threads; t > 1; t >>= 1) ++par;
  uint32_t root = pool.alloc();
  b.node(root, 0, nt, 1, par + 1);
  bvh->degenerate = b.degenerate.load();
  bvh->binary_depth = b.max_depth.load();
  // compact: renumber nodes in allocation order (root is 0)
  const uint32_t nn = pool.size();
  bvh->nodes.resize(nn);
  for (uint32_t i = 0; i < nn; ++i) bvh->nodes[i] = pool.at(i);
  int largest = 0;
  for (auto& n : bvh->nodes)
    if (n.leaf) largest = std::max(largest, (int)n.count);
  bvh->largest_leaf = largest;
  bvh->num_binary = (int64_t)nn;
  // normalizer (trace_pass.js:202-211) and f32 vertices (structs.js setFloat32)
  const BNode& r = bvh->nodes[0];
  double span[3];
  for (int k = 0; k < 3; ++k) span[k] = r.hi[k] - r.lo[k];
  double longest = std::max(std::max(span[0], span[1]), span[2]);
  for (int k = 0; k < 3; ++k) bvh->norm_center[k] = r.lo[k] + span[k] * 0.5;
  bvh->norm_scale = 2 / longest;
  bvh->verts.resize((size_t)nv * 3);
  for (size_t i = 0; i < (size_t)nv * 3; ++i)
    bvh->verts[i] = (float)((positions[i] - bvh->norm_center[i % 3]) * bvh->norm_scale);
  pk_bvh_make_wide(bvh);
  *out = bvh;
  return PK_OK;
}

int pk_bvh_destroy(pk_bvh* bvh) {
  delete bvh;
  return PK_OK;
}

int pk_bvh_info(const pk_bvh* bvh, int64_t* info) {
  if (!bvh || !info) {
    pk_set_error("pk_bvh_info: null argument");
    return PK_E_INVALID;
  }
  info[0] = bvh->num_binary;
  info[1] = (int64_t)bvh->wnodes.size();
  info[2] = (int64_t)bvh->wtris.size();
  info[3] = bvh->binary_depth;
  info[4] = bvh->wide_depth;
  info[5] = bvh->largest_leaf;
  return PK_OK;
}

int pk_bvh_export_reference(const pk_bvh* bvh, int32_t* nodes5, int32_t* tris4, float* verts4, double* norm4) {
  if (!bvh) {
    pk_set_error("pk_bvh_export_reference: null bvh");
    return PK_E_INVALID;
  }
  // BFS (bvh.js:46-74): siblings adjacent, left child of the k-th interior
  // node at 2k+1.
  std::deque<int32_t> q;
  q.push_back(0);
  int64_t out = 0, interior = 0, tri_index = 0;
  while (!q.empty()) {
    int32_t ni = q.front();
    q.pop_front();
    const BNode& n = bvh->nodes[ni];
    double lo[3], hi[3];
    for (int k = 0; k < 3; ++k) {
      lo[k] = (n.lo[k] - bvh->norm_center[k]) * bvh->norm_scale;
      hi[k] = (n.hi[k] - bvh->norm_center[k]) * bvh->norm_scale;
    }
    int32_t* w = nodes5 ? nodes5 + out * 5 : nullptr;
    if (n.leaf) {
      if (w) {
        w[0] = -1;
        w[1] = (int32_t)(((uint32_t)n.count << 24) | (uint32_t)tri_index);
      }
      if (tris4)
        for (uint32_t i = 0; i < n.count; ++i)
          std::memcpy(tris4 + (tri_index + i) * 4, &bvh->tris[(size_t)bvh->leaf_tris[n.begin + i] * 4], 16);
      tri_index += n.count;
    } else {
      if (w) {
        w[0] = (int32_t)(2 * interior + 1);
        w[1] = -1;
      }
      ++interior;
      q.push_back(n.left);
      q.push_back(n.right);
    }
    if (w) {
      w[2] = (int32_t)quantize_pair(hi[0], lo[0]);
      w[3] = (int32_t)quantize_pair(hi[1], lo[1]);
      w[4] = (int32_t)quantize_pair(hi[2], lo[2]);
    }
    ++out;
  }
  if (verts4)
    for (size_t i = 0; i < bvh->nv; ++i) {
      verts4[i * 4 + 0] = bvh->verts[i * 3 + 0];
      verts4[i * 4 + 1] = bvh->verts[i * 3 + 1];
      verts4[i * 4 + 2] = bvh->verts[i * 3 + 2];
      verts4[i * 4 + 3] = 0.0f;
    }
  if (norm4) {
    for (int k = 0; k < 3; ++k) norm4[k] = bvh->norm_center[k];
    norm4[3] = bvh->norm_scale;
  }
  return PK_OK;
}

int pk_bvh_export_wide(const pk_bvh* bvh, void* nodes, void* tris, int32_t* tri_index) {
  if (!bvh) {
    pk_set_error("pk_bvh_export_wide: null bvh");
    return PK_E_INVALID;
  }
  if (nodes) std::memcpy(nodes, bvh->wnodes.data(), bvh->wnodes.size() * sizeof(pk::WNode));
  if (tris) std::memcpy(tris, bvh->wtris.data(), bvh->wtris.size() * sizeof(pk::TriRecord));
  if (tri_index) std::memcpy(tri_index, bvh->wtri_index.data(), bvh->wtri_index.size() * sizeof(int32_t));
  return PK_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Wide collapse: each wide node takes its binary node's two children and
// repeatedly opens the internal child with the largest surface area until it
// holds 8 children (or only leaves remain).  Wide nodes are numbered in
// breadth-first order so the top levels share lines; leaf triangles are laid
// out in the order the collapse reaches them.
// ---------------------------------------------------------------------------
void pk_bvh_make_wide(pk_bvh* bvh) {
  std::vector<BNode>& nodes = bvh->nodes;
  // split oversized leaves into virtual halves for the collapse only
  const size_t n0 = nodes.size();
  for (size_t i = 0; i < n0; ++i) {
    std::vector<size_t> work{i};
    while (!work.empty()) {
      size_t j = work.back();
      work.pop_back();
      if (!nodes[j].leaf || nodes[j].count <= (uint32_t)pk::kLeafMax) continue;
      BNode a = nodes[j], c = nodes[j];
      a.count = nodes[j].count / 2;
      c.begin = nodes[j].begin + a.count;
      c.count = nodes[j].count - a.count;
      a.vleft = a.vright = c.vleft = c.vright = -1;
      nodes[j].vleft = (int32_t)nodes.size();
      nodes.push_back(a);
      nodes[j].vright = (int32_t)nodes.size();
      nodes.push_back(c);
      work.push_back((size_t)nodes[j].vleft);
      work.push_back((size_t)nodes[j].vright);
    }
  }
  auto is_internal = [&](int32_t i) { return !nodes[i].leaf || nodes[i].vleft >= 0; };
  auto kid_l = [&](int32_t i) { return nodes[i].leaf ? nodes[i].vleft : nodes[i].left; };
  auto kid_r = [&](int32_t i) { return nodes[i].leaf ? nodes[i].vright : nodes[i].right; };
  // f32 boxes over the normalized vertices, bottom-up
  std::vector<FBox> fb(nodes.size());
  const std::vector<float>& V = bvh->verts;
  std::vector<int32_t> order;
  order.reserve(nodes.size());
  {
    std::vector<int32_t> st{0};
    while (!st.empty()) {
      int32_t i = st.back();
      st.pop_back();
      order.push_back(i);
      if (is_internal(i)) {
        st.push_back(kid_l(i));
        st.push_back(kid_r(i));
      }
    }
  }
  for (size_t oi = order.size(); oi-- > 0;) {
    int32_t i = order[oi];
    FBox& b = fb[i];
    for (int k = 0; k < 3; ++k) {
      b.lo[k] = INFINITY;
      b.hi[k] = -INFINITY;
    }
    if (is_internal(i)) {
      const FBox& l = fb[kid_l(i)];
      const FBox& r = fb[kid_r(i)];
      for (int k = 0; k < 3; ++k) {
        b.lo[k] = std::min(l.lo[k], r.lo[k]);
        b.hi[k] = std::max(l.hi[k], r.hi[k]);
      }
    } else {
      const BNode& n = nodes[i];
      for (uint32_t t = 0; t < n.count; ++t) {
        const int32_t* tr = &bvh->tris[(size_t)bvh->leaf_tris[n.begin + t] * 4];
        for (int v = 0; v < 3; ++v)
          for (int k = 0; k < 3; ++k) {
            float p = V[(size_t)tr[v] * 3 + k];
            b.lo[k] = std::min(b.lo[k], p);
            b.hi[k] = std::max(b.hi[k], p);
          }
      }
    }
  }
  auto area = [&](int32_t i) {
    const FBox& b = fb[i];
    double x = std::max((double)b.hi[0] - b.lo[0], 0.0), y = std::max((double)b.hi[1] - b.lo[1], 0.0),
           z = std::max((double)b.hi[2] - b.lo[2], 0.0);
    return x * y + x * z + y * z;
  };
  bvh->wnodes.clear();
  bvh->wtris.clear();
  bvh->wtri_index.clear();
  bvh->wtris.reserve(bvh->nt);
  bvh->wtri_index.reserve((size_t)bvh->nt * 4);
  struct QItem {
    int32_t bnode;
    uint32_t wnode;
    int depth;
  };
  std::deque<QItem> q;
  bvh->wnodes.emplace_back();
  q.push_back({0, 0, 1});
  int wdepth = 1;
  auto emit_leaf = [&](int32_t li, uint32_t* first) {
    const BNode& n = nodes[li];
    *first = (uint32_t)bvh->wtris.size();
    for (uint32_t t = 0; t < n.count; ++t) {
      uint32_t id = bvh->leaf_tris[n.begin + t];
      const int32_t* tr = &bvh->tris[(size_t)id * 4];
      const float* p0 = &V[(size_t)tr[0] * 3];
      const float* p1 = &V[(size_t)tr[1] * 3];
      const float* p2 = &V[(size_t)tr[2] * 3];
      pk::TriRecord rec;
      rec.v0x = p0[0];
      rec.v0y = p0[1];
      rec.v0z = p0[2];
      rec.id = id;
      rec.e1x = p1[0] - p0[0];
      rec.e1y = p1[1] - p0[1];
      rec.e1z = p1[2] - p0[2];
      rec.pad0 = 0;
      rec.e2x = p2[0] - p0[0];
      rec.e2y = p2[1] - p0[1];
      rec.e2z = p2[2] - p0[2];
      rec.pad1 = 0;
      bvh->wtris.push_back(rec);
      for (int c = 0; c < 4; ++c) bvh->wtri_index.push_back(tr[c]);
    }
  };
  while (!q.empty()) {
    QItem it = q.front();
    q.pop_front();
    wdepth = std::max(wdepth, it.depth);
    int32_t kids[pk::kWide];
    int nk = 0;
    if (is_internal(it.bnode)) {
      kids[nk++] = kid_l(it.bnode);
      kids[nk++] = kid_r(it.bnode);
      while (nk < pk::kWide) {
        int best = -1;
        double best_a = -1;
        for (int i = 0; i < nk; ++i)
          if (is_internal(kids[i]) && area(kids[i]) > best_a) {
            best_a = area(kids[i]);
            best = i;
          }
        if (best < 0) break;
        int32_t x = kids[best];
        for (int i = nk; i > best + 1; --i) kids[i] = kids[i - 1];
        kids[best] = kid_l(x);
        kids[best + 1] = kid_r(x);
        ++nk;
      }
    } else {
      kids[nk++] = it.bnode;  // root is a leaf
    }
    FBox cb[pk::kWide];
    pk::WNode w;
    std::memset(&w, 0, sizeof(w));
    w.nchild = (uint8_t)nk;
    for (int i = 0; i < nk; ++i) {
      cb[i] = fb[kids[i]];
      if (is_internal(kids[i])) {
        uint32_t wi = (uint32_t)bvh->wnodes.size();
        bvh->wnodes.emplace_back();
        q.push_back({kids[i], wi, it.depth + 1});
        w.child[i] = wi;
        w.meta[i] = 0;
      } else {
        uint32_t first;
        emit_leaf(kids[i], &first);
        w.child[i] = first;
        w.meta[i] = (uint8_t)nodes[kids[i]].count;
      }
    }
    quantize_node(w, cb, nk);
    bvh->wnodes[it.wnode] = w;
  }
  bvh->wide_depth = wdepth;
}
