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
#include <iterator>
#include <cfloat>
#include <cstdio>
#include <cstdlib>
#include <atomic>
#include <cmath>
#include <cstring>
#include <deque>
#include <future>
#include <memory>
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

// Chunked store with lock-free allocation and stable addresses for the
// threaded builds (nodes; SBVH references).
template <class T> class Pool {
 public:
  explicit Pool(size_t max_items) : pages_((max_items >> kShift) + 1) {}
  uint32_t alloc() {
    uint32_t i = next_.fetch_add(1);
    ensure_page(i >> kShift);
    return i;
  }
  // n consecutive items, or UINT32_MAX (nothing taken) if they do not fit
  uint32_t try_alloc(uint32_t n) {
    uint32_t i = next_.load();
    do {
      if (n == 0 || (((size_t)i + n - 1) >> kShift) >= pages_.size()) return UINT32_MAX;
    } while (!next_.compare_exchange_weak(i, i + n));
    for (size_t p = i >> kShift; p <= (((size_t)i + n - 1) >> kShift); ++p) ensure_page(p);
    return i;
  }
  void ensure_page(size_t p) {
    if (!pages_[p].load(std::memory_order_acquire)) {
      std::lock_guard<std::mutex> g(mu_);
      if (!pages_[p].load(std::memory_order_relaxed)) pages_[p].store(new T[kPage], std::memory_order_release);
    }
  }
  size_t capacity() const { return pages_.size() << kShift; }
  T& at(uint32_t i) { return pages_[i >> kShift].load(std::memory_order_relaxed)[i & (kPage - 1)]; }
  const T& at(uint32_t i) const { return pages_[i >> kShift].load(std::memory_order_relaxed)[i & (kPage - 1)]; }
  uint32_t size() const { return next_.load(); }
  ~Pool() {
    for (auto& p : pages_) delete[] p.load();
  }

 private:
  static constexpr int kShift = 16;
  static constexpr size_t kPage = (size_t)1 << kShift;
  std::vector<std::atomic<T*>> pages_;
  std::atomic<uint32_t> next_{0};
  std::mutex mu_;
};
using NodePool = Pool<BNode>;

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

// ---------------------------------------------------------------------------
// SBVH: bvh.js's spatial-split path (Node constructor :119-143,
// getSpatialSplit :231-274, BoundingBox.axisIntersectTriangle,
// primitives.js:52-80), which the reference ships disabled
// (checkSpatialSplit = false), made live.  Decisions restated:
//   * the object split as in the plain build; a spatial split is tried when
//     the area of (front box at the best index) ∩ (back box there) exceeds
//     kSplitAlpha of the root box area;
//   * the node is cut into kSpatialBins equal slabs along the object split's
//     axis; every reference gives one fragment per slab its box spans: the
//     bounds of the triangle's part inside the slab (axisIntersectTriangle:
//     edge/plane crossings and vertices within the slab widened by 2 ulp of
//     1), here also intersected with the reference's own box;
//   * a spatial split wins if cheaper than the object split; leaf iff the
//     chosen cost > N; a leaf keeps the node's own references.
// One deliberate change: bvh.js sweeps the fragments as if each were a
// separate reference, so a triangle crossing m slabs counts m times and
// spatial splits essentially never win.  Planes between slabs are evaluated
// with the binned SBVH counts instead (Stich, Friedrich, Dietrich 2009): a
// reference counts once on each side it reaches, with its box clipped to
// that side; cost 0.5 + A(L)/A(P) N_L + A(R)/A(P) N_R as bvh.js's.
// Children of a spatial split are re-sorted by centroid on every axis.  A
// triangle can sit in several leaves.
// ---------------------------------------------------------------------------
constexpr double kSplitAlpha = 1.0e-5;  // bvh.js:4
constexpr int kSpatialBins = 64;        // Node.NUM_BINS

struct SRef {
  uint32_t tri;
  Box box;
};

struct SBuild {
  const double* pos;
  const int32_t* tris;
  NodePool* pool;
  Pool<SRef>* refs;
  std::vector<uint8_t> mark;  // per reference id: in the left child
  double root_area = 0;
  std::mutex leaf_mu;
  std::vector<uint32_t> leaf_out;
  std::atomic<int> degenerate{0};
  std::atomic<int> max_depth{0};
  std::atomic<int64_t> spatial{0};

  static double area_of(const Box& b) { return box_area(b); }

  template <class BoxOf>
  static void sweep_boxes(uint32_t n, BoxOf box_of, double parent_area, double* best_cost, uint32_t* best_index) {
    std::vector<double> suf(n);
    Box bb;
    box_empty(bb);
    for (uint32_t i = n; i-- > 0;) {
      box_add(bb, box_of(i));
      suf[i] = box_area(bb);
    }
    Box bf;
    box_empty(bf);
    double bc = INFINITY;
    uint32_t bi = 0;
    for (uint32_t i = 0; i < n; ++i) {
      box_add(bf, box_of(i));
      double cost = 0.5 + (box_area(bf) / parent_area) * (double)(i + 1) + (suf[i] / parent_area) * (double)(n - 1 - i);
      if (cost < bc) {
        bc = cost;
        bi = i;
      }
    }
    *best_cost = bc;
    *best_index = bi;
  }
  void sweep(const std::vector<uint32_t>& l, double parent_area, double* best_cost, uint32_t* best_index) {
    sweep_boxes((uint32_t)l.size(), [&](uint32_t i) -> const Box& { return refs->at(l[i]).box; }, parent_area,
                best_cost, best_index);
  }

  // axisIntersectTriangle: bounds of the triangle's part inside slab [p0, p1]
  // on axis a; returns false if it has no point there.
  bool clip(uint32_t t, int a, double p0, double p1, Box* out) const {
    const double lo = p0 - DBL_EPSILON * 2, hi = p1 + DBL_EPSILON * 2;
    const double* v[3];
    for (int k = 0; k < 3; ++k) v[k] = pos + (size_t)tris[(size_t)t * 4 + k] * 3;
    Box b;
    box_empty(b);
    auto add = [&](const double* p) {
      for (int k = 0; k < 3; ++k) {
        b.lo[k] = std::min(p[k], b.lo[k]);
        b.hi[k] = std::max(p[k], b.hi[k]);
      }
    };
    auto add_lerp = [&](const double* x, const double* y, double s) {
      double q[3];
      for (int k = 0; k < 3; ++k) q[k] = x[k] + (y[k] - x[k]) * s;
      add(q);
    };
    bool any = false;
    for (int i = 0; i < 3; ++i) {
      const double* x = v[i];
      const double* y = v[(i + 1) % 3];
      const double q0 = x[a], q1 = y[a];
      const double tmin = (lo - q0) / (q1 - q0);
      if (tmin >= 0 && tmin <= 1) {
        add_lerp(x, y, tmin);
        any = true;
      }
      const double tmax = (hi - q0) / (q1 - q0);
      if (tmax >= 0 && tmax <= 1) {
        add_lerp(x, y, tmax);
        any = true;
      }
      if (q0 >= lo && q0 <= hi) {
        add(x);
        any = true;
      }
    }
    *out = b;
    return any;
  }

  void node(uint32_t ni, std::vector<uint32_t> ax[3], int depth, int par) {
    int md = max_depth.load();
    while (depth > md && !max_depth.compare_exchange_weak(md, depth)) {
    }
    const uint32_t n = (uint32_t)ax[0].size();
    Box bounds;
    box_empty(bounds);
    for (uint32_t r : ax[0]) box_add(bounds, refs->at(r).box);
    {
      BNode& nd = pool->at(ni);
      for (int k = 0; k < 3; ++k) {
        nd.lo[k] = bounds.lo[k];
        nd.hi[k] = bounds.hi[k];
      }
    }
    const double parent_area = box_area(bounds);
    double best_cost = INFINITY;
    int best_axis = -1;
    uint32_t best_index = 0;
    for (int a = 0; a < 3; ++a) {
      double c;
      uint32_t i;
      sweep(ax[a], parent_area, &c, &i);
      if (c < best_cost) {
        best_cost = c;
        best_axis = a;
        best_index = i;
      }
    }
    // a spatial split's children: whole references kept, and new halves
    std::vector<uint32_t> sp_left, sp_right, new_left, new_right;
    bool spatial_split = false;
    double split_cost = best_cost;
    uint32_t split_index = best_index;
    if (best_axis >= 0) {
      // overlap of the object split's two child boxes (back box includes ref i)
      Box f, b;
      box_empty(f);
      box_empty(b);
      const std::vector<uint32_t>& l = ax[best_axis];
      for (uint32_t i = 0; i <= best_index; ++i) box_add(f, refs->at(l[i]).box);
      for (uint32_t i = best_index; i < n; ++i) box_add(b, refs->at(l[i]).box);
      for (int k = 0; k < 3; ++k) {
        f.lo[k] = std::max(f.lo[k], b.lo[k]);
        f.hi[k] = std::min(f.hi[k], b.hi[k]);
      }
      if (root_area > 0 && box_area(f) / root_area > kSplitAlpha) {
        const int a = best_axis;
        const double lo = bounds.lo[a], span = bounds.hi[a] - lo;
        if (span != 0) {
          // per reference: fragment boxes of the slabs it reaches (clipped
          // triangle, within the reference's own box); per slab: bounds,
          // references entering and leaving there
          struct Cut {
            uint32_t ref;
            int first, last;
            size_t part;  // parts[part + i - first]: fragment in slab i
          };
          std::vector<Cut> cuts;
          std::vector<Box> parts;
          cuts.reserve(n);
          parts.reserve(n + n / 2);
          Box tmp[kSpatialBins];
          Box slab_box[kSpatialBins];
          uint32_t enter[kSpatialBins] = {0}, leave[kSpatialBins] = {0};
          for (auto& sb : slab_box) box_empty(sb);
          for (uint32_t r : l) {
            const SRef& ref = refs->at(r);
            const double rmin = ref.box.lo[a] - lo, rmax = ref.box.hi[a] - lo;
            // a box starting on the node's upper face maps to slab 64: clamp
            const int start = std::min(std::max(0, (int)std::floor((rmin / span) * kSpatialBins)), kSpatialBins - 1);
            const int stop = std::max(std::min((int)std::floor((rmax / span) * kSpatialBins), kSpatialBins - 1), start);
            Cut c{r, -1, -1, parts.size()};
            if (start == stop) {
              // within one slab: its fragment there is the reference itself
              // (the slab's clip of the triangle contains the reference box)
              c.first = c.last = start;
              parts.push_back(ref.box);
              box_add(slab_box[start], ref.box);
              ++enter[start];
              ++leave[start];
              cuts.push_back(c);
              continue;
            }
            for (int i = start; i <= stop; ++i) {
              const double p0 = lo + (double)i * span / kSpatialBins;
              const double p1 = lo + (double)(i + 1) * span / kSpatialBins;
              Box fb;
              if (!clip(ref.tri, a, p0, p1, &fb)) fb = Box{{INFINITY, INFINITY, INFINITY}, {-INFINITY, -INFINITY, -INFINITY}};
              for (int k = 0; k < 3; ++k) {
                fb.lo[k] = std::max(fb.lo[k], ref.box.lo[k]);
                fb.hi[k] = std::min(fb.hi[k], ref.box.hi[k]);
              }
              const bool empty = !(fb.lo[0] <= fb.hi[0] && fb.lo[1] <= fb.hi[1] && fb.lo[2] <= fb.hi[2]);
              if (!empty) {
                if (c.first < 0) c.first = i;
                c.last = i;
                box_add(slab_box[i], fb);
              }
              tmp[i - start] = fb;
            }
            if (c.first < 0) {  // degenerate clip: keep the reference whole in its first slab
              c.first = c.last = start;
              parts.push_back(ref.box);
              box_add(slab_box[start], ref.box);
            } else {
              parts.insert(parts.end(), tmp + (c.first - start), tmp + (c.last - start + 1));
            }
            ++enter[c.first];
            ++leave[c.last];
            cuts.push_back(std::move(c));
          }
          // sweep the 63 inner planes: cost 0.5 + A(L)/A(P) N_L + A(R)/A(P) N_R
          double suf_area[kSpatialBins + 1];
          uint32_t suf_n[kSpatialBins + 1];
          {
            Box bb;
            box_empty(bb);
            uint32_t cnt = 0;
            suf_area[kSpatialBins] = 0;
            suf_n[kSpatialBins] = 0;
            for (int i = kSpatialBins; i-- > 0;) {
              box_add(bb, slab_box[i]);
              cnt += leave[i];
              suf_area[i] = box_area(bb);
              suf_n[i] = cnt;
            }
          }
          Box bf;
          box_empty(bf);
          uint32_t nl = 0;
          double sc = INFINITY;
          int plane = -1;
          for (int k = 1; k < kSpatialBins; ++k) {
            box_add(bf, slab_box[k - 1]);
            nl += enter[k - 1];
            if (nl == 0 || suf_n[k] == 0) continue;
            const double cost = 0.5 + (box_area(bf) / parent_area) * (double)nl +
                                (suf_area[k] / parent_area) * (double)suf_n[k];
            if (cost < sc) {
              sc = cost;
              plane = k;
            }
          }
          if (plane > 0 && sc < best_cost) {
            // references wholly on one side keep their id and box; a
            // straddling one becomes two, each clipped to its side
            std::vector<SRef> halves;
            for (const Cut& c : cuts) {
              if (c.last < plane) {
                sp_left.push_back(c.ref);
              } else if (c.first >= plane) {
                sp_right.push_back(c.ref);
              } else {
                Box lb, rb;
                box_empty(lb);
                box_empty(rb);
                for (int i = c.first; i <= c.last; ++i) box_add(i < plane ? lb : rb, parts[c.part + i - c.first]);
                const uint32_t t = refs->at(c.ref).tri;
                halves.push_back(SRef{t, lb});
                halves.push_back(SRef{t, rb});
              }
            }
            const uint32_t first = halves.empty() ? 0u : refs->try_alloc((uint32_t)halves.size());
            if (first != UINT32_MAX) {
              for (size_t i = 0; i < halves.size(); ++i) {
                refs->at(first + (uint32_t)i) = halves[i];
                (i % 2 ? new_right : new_left).push_back(first + (uint32_t)i);
              }
              spatial_split = true;
              split_cost = sc;
            } else {
              sp_left.clear();
              sp_right.clear();
              new_left.clear();
              new_right.clear();
            }
          }
        }
      }
    }
    if (best_axis < 0 || split_cost > (double)n) {
      if (best_axis < 0) {
        degenerate.store(1);
        best_axis = 0;
      }
      std::lock_guard<std::mutex> g(leaf_mu);
      BNode& nd = pool->at(ni);
      nd.leaf = 1;
      nd.split_axis = (uint8_t)best_axis;
      nd.begin = (uint32_t)leaf_out.size();
      nd.count = n;
      for (uint32_t r : ax[best_axis]) leaf_out.push_back(refs->at(r).tri);
      return;
    }
    pool->at(ni).split_axis = (uint8_t)best_axis;
    std::vector<uint32_t> lax[3], rax[3];
    if (spatial_split) {
      // whole references keep their centroid order in every axis list (a
      // subsequence of a sorted list is sorted); the new halves are sorted
      // by centroid and merged in (stable: whole references first on ties)
      spatial.fetch_add(1);
      for (uint32_t r : sp_left) mark_at(r) = 1;
      for (uint32_t r : sp_right) mark_at(r) = 2;
      for (int k = 0; k < 3; ++k) {
        auto by_centroid = [&](uint32_t x, uint32_t y) {
          const Box& bx = refs->at(x).box;
          const Box& by = refs->at(y).box;
          return (bx.lo[k] + bx.hi[k]) * 0.5 < (by.lo[k] + by.hi[k]) * 0.5;
        };
        std::vector<uint32_t> kl, kr;
        kl.reserve(sp_left.size());
        kr.reserve(sp_right.size());
        for (uint32_t r : ax[k]) {
          const uint8_t m = mark_at(r);
          if (m == 1) kl.push_back(r);
          else if (m == 2) kr.push_back(r);
        }
        std::vector<uint32_t>().swap(ax[k]);
        std::stable_sort(new_left.begin(), new_left.end(), by_centroid);
        std::stable_sort(new_right.begin(), new_right.end(), by_centroid);
        lax[k].reserve(kl.size() + new_left.size());
        rax[k].reserve(kr.size() + new_right.size());
        std::merge(kl.begin(), kl.end(), new_left.begin(), new_left.end(), std::back_inserter(lax[k]), by_centroid);
        std::merge(kr.begin(), kr.end(), new_right.begin(), new_right.end(), std::back_inserter(rax[k]), by_centroid);
      }
      for (uint32_t r : sp_left) mark_at(r) = 0;
      for (uint32_t r : sp_right) mark_at(r) = 0;
    } else {
      const std::vector<uint32_t>& split = ax[best_axis];
      for (uint32_t i = 0; i <= split_index; ++i) mark_at(split[i]) = 1;
      for (int k = 0; k < 3; ++k) {
        for (uint32_t r : ax[k]) (mark_at(r) ? lax[k] : rax[k]).push_back(r);
        std::vector<uint32_t>().swap(ax[k]);
      }
      for (uint32_t r : lax[0]) mark_at(r) = 0;
    }
    uint32_t l = pool->alloc(), r = pool->alloc();
    {
      BNode& p = pool->at(ni);
      p.left = (int32_t)l;
      p.right = (int32_t)r;
    }
    if (par > 0 && n > 32768) {
      auto fut = std::async(std::launch::async, [&]() { node(l, lax, depth + 1, par - 1); });
      node(r, rax, depth + 1, par - 1);
      fut.get();
    } else {
      node(l, lax, depth + 1, 0);
      node(r, rax, depth + 1, 0);
    }
  }

  uint8_t& mark_at(uint32_t id) { return mark[id]; }
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
  uint8_t lm = 0;
  for (int i = 0; i < nc; ++i)
    if (w.meta[i]) lm |= (uint8_t)(1u << i);
  w.leafmask = lm;
  w.ox = origin[0];
  w.oy = origin[1];
  w.oz = origin[2];
  w.ex = (uint8_t)(ex[0] + 127);
  w.ey = (uint8_t)(ex[1] + 127);
  w.ez = (uint8_t)(ex[2] + 127);
  pk::wnode_link(w);
}

}  // namespace

void pk_quantize_node(pk::WNode& w, const FBox* cb, int nc) { quantize_node(w, cb, nc); }

namespace {
// Slots [b, e) of 8-wide node w as a 4-wide node (e - b <= 4) in w's frame.
pk::WNode4 wide4_part(const pk::WNode& w, int b, int e) {
  pk::WNode4 n;
  std::memset(&n, 0, sizeof(n));
  n.ox = w.ox, n.oy = w.oy, n.oz = w.oz;
  n.ex = w.ex, n.ey = w.ey, n.ez = w.ez;
  n.nchild = (uint8_t)(e - b);
  uint32_t ib = w.ibase, tb = w.tbase;
  for (int c = 0; c < b; ++c) {
    if (w.meta[c]) tb += w.meta[c];
    else ++ib;
  }
  n.ibase = ib;
  n.tbase = tb;
  for (int i = 0; i < 4; ++i) {
    const int c = b + i;
    const bool on = c < e;
    n.meta[i] = on ? w.meta[c] : 0;
    n.qlox[i] = on ? w.qlox[c] : 255, n.qhix[i] = on ? w.qhix[c] : 0;
    n.qloy[i] = on ? w.qloy[c] : 255, n.qhiy[i] = on ? w.qhiy[c] : 0;
    n.qloz[i] = on ? w.qloz[c] : 255, n.qhiz[i] = on ? w.qhiz[c] : 0;
    if (on && w.meta[c]) n.leafmask |= (uint8_t)(1u << i);
  }
  return n;
}

// Surface area (in quantization steps of w's frame) of the box around
// slots [b, e).
double wide4_area(const pk::WNode& w, int b, int e) {
  const double sx = std::ldexp(1.0, (int)w.ex - 127), sy = std::ldexp(1.0, (int)w.ey - 127),
               sz = std::ldexp(1.0, (int)w.ez - 127);
  int lo[3] = {255, 255, 255}, hi[3] = {0, 0, 0};
  for (int c = b; c < e; ++c) {
    lo[0] = std::min(lo[0], (int)w.qlox[c]), hi[0] = std::max(hi[0], (int)w.qhix[c]);
    lo[1] = std::min(lo[1], (int)w.qloy[c]), hi[1] = std::max(hi[1], (int)w.qhiy[c]);
    lo[2] = std::min(lo[2], (int)w.qloz[c]), hi[2] = std::max(hi[2], (int)w.qhiz[c]);
  }
  const double dx = (hi[0] - lo[0]) * sx, dy = (hi[1] - lo[1]) * sy, dz = (hi[2] - lo[2]) * sz;
  return dx * dy + dy * dz + dz * dx;
}
}  // namespace

void pk_make_wide4(const std::vector<pk::WNode>& n8, std::vector<pk::WNode4>& n4) {
  n4.assign(n8.size(), pk::WNode4{});
  for (size_t i = 0; i < n8.size(); ++i) {
    const pk::WNode& w = n8[i];
    const int n = w.nchild;
    if (n <= 4) {
      n4[i] = wide4_part(w, 0, n);
      continue;
    }
    // one cut s (slots [0, s) and [s, n), at most four each) with the least
    // summed surface area: a ray that enters the node visits each sub-node
    // in proportion to its area
    int best_s = 4;
    double best = INFINITY;
    for (int s = n - 4; s <= 4; ++s) {
      const double a = wide4_area(w, 0, s) + wide4_area(w, s, n);
      if (a < best) {
        best = a;
        best_s = s;
      }
    }
    const uint32_t sub = (uint32_t)n4.size();
    pk::WNode4 root;
    std::memset(&root, 0, sizeof(root));
    root.ox = w.ox, root.oy = w.oy, root.oz = w.oz;
    root.ex = w.ex, root.ey = w.ey, root.ez = w.ez;
    root.nchild = 2;
    root.ibase = sub;
    const int cut[3] = {0, best_s, n};
    for (int h = 0; h < 4; ++h) {
      uint8_t lx = 255, hx = 0, ly = 255, hy = 0, lz = 255, hz = 0;
      if (h < 2)
        for (int c = cut[h]; c < cut[h + 1]; ++c) {
          lx = std::min(lx, w.qlox[c]), hx = std::max(hx, w.qhix[c]);
          ly = std::min(ly, w.qloy[c]), hy = std::max(hy, w.qhiy[c]);
          lz = std::min(lz, w.qloz[c]), hz = std::max(hz, w.qhiz[c]);
        }
      root.qlox[h] = lx, root.qhix[h] = hx, root.qloy[h] = ly, root.qhiy[h] = hy;
      root.qloz[h] = lz, root.qhiz[h] = hz;
    }
    n4[i] = root;
    n4.push_back(wide4_part(w, 0, best_s));
    n4.push_back(wide4_part(w, best_s, n));
  }
}

void pk_bvh_make_wide(pk_bvh* bvh);

extern "C" {

int pk_bvh_build(const double* positions, uint32_t nv, const int32_t* triangles, uint32_t nt, int num_threads,
                 pk_bvh** out) {
  return pk_bvh_build_ex(positions, nv, triangles, nt, num_threads, 0, out);
}

int pk_bvh_build_ex(const double* positions, uint32_t nv, const int32_t* triangles, uint32_t nt, int num_threads,
                    uint32_t flags, pk_bvh** out) {
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
  int par = 0;
  for (int t = num_threads; t > 1; t >>= 1) ++par;
  bvh->spatial_splits = (flags & PK_BVH_SPATIAL_SPLITS) != 0;
  bvh->sah_collapse = (flags & PK_BVH_SAH_COLLAPSE) != 0;
  // SBVH references: the triangles, then fragments (each spatial split
  // adds at most 64 per reference it cuts; the pool bounds the total)
  const size_t ref_cap = std::min((size_t)nt * 8 + (1u << 20), (size_t)UINT32_MAX - (1u << 17));
  std::unique_ptr<NodePool> pool(new NodePool(bvh->spatial_splits ? ref_cap * 2 + 2 : (size_t)nt * 2 + 2));
  uint32_t root = pool->alloc();
  if (!bvh->spatial_splits) {
    bvh->leaf_tris.resize(nt);
    Build b;
    b.ref_box = &ref_box;
    for (int a = 0; a < 3; ++a) b.axis[a] = ax[a];
    b.leaf_out = bvh->leaf_tris.data();
    b.in_left.assign(nt, 0);
    b.pool = pool.get();
    b.node(root, 0, nt, 1, par + 1);
    bvh->degenerate = b.degenerate.load();
    bvh->binary_depth = b.max_depth.load();
  } else {
    Pool<SRef> refs(ref_cap);
    for (uint32_t t = 0; t < nt; ++t) refs.at(refs.alloc()) = SRef{t, ref_box[t]};
    SBuild b;
    b.pos = positions;
    b.tris = triangles;
    b.pool = pool.get();
    b.refs = &refs;
    b.mark.assign(refs.capacity(), 0);
    Box rb;
    box_empty(rb);
    for (uint32_t t = 0; t < nt; ++t) box_add(rb, ref_box[t]);
    b.root_area = box_area(rb);
    std::vector<uint32_t> lists[3];
    for (int a = 0; a < 3; ++a) lists[a].assign(ax[a], ax[a] + nt);
    std::vector<uint32_t>().swap(arr);
    b.node(root, lists, 1, par + 1);
    bvh->degenerate = b.degenerate.load();
    bvh->binary_depth = b.max_depth.load();
    bvh->leaf_tris.swap(b.leaf_out);
    bvh->num_spatial_splits = b.spatial.load();
  }
  NodePool& pool_ref = *pool;
  // compact: renumber nodes in allocation order (root is 0)
  const uint32_t nn = pool_ref.size();
  bvh->nodes.resize(nn);
  for (uint32_t i = 0; i < nn; ++i) bvh->nodes[i] = pool_ref.at(i);
  pool.reset();
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

int pk_bvh_normalizer(const pk_bvh* bvh, double* norm4) {
  if (!bvh || !norm4) {
    pk_set_error("pk_bvh_normalizer: null argument");
    return PK_E_INVALID;
  }
  for (int k = 0; k < 3; ++k) norm4[k] = bvh->norm_center[k];
  norm4[3] = bvh->norm_scale;
  return PK_OK;
}

int pk_bvh_export_reference(const pk_bvh* bvh, int32_t* nodes5, int32_t* tris4, float* verts4, double* norm4) {
  if (!bvh) {
    pk_set_error("pk_bvh_export_reference: null bvh");
    return PK_E_INVALID;
  }
  if (bvh->instanced) {
    pk_set_error("pk_bvh_export_reference: an instanced BVH has no reference layout");
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

int pk_bvh_export_wide4(const pk_bvh* bvh, void* nodes64, int64_t* count) {
  if (!bvh || !count) {
    pk_set_error("pk_bvh_export_wide4: null argument");
    return PK_E_INVALID;
  }
  if (bvh->instanced) {
    pk_set_error("pk_bvh_export_wide4: an instanced BVH has no 4-wide layout");
    return PK_E_INVALID;
  }
  std::vector<pk::WNode4> n4;
  pk_make_wide4(bvh->wnodes, n4);
  if (nodes64) {
    if (*count < (int64_t)n4.size()) {
      pk_set_error("pk_bvh_export_wide4: output holds fewer than " + std::to_string(n4.size()) + " nodes");
      return PK_E_LIMIT;
    }
    std::memcpy(nodes64, n4.data(), n4.size() * sizeof(pk::WNode4));
  }
  *count = (int64_t)n4.size();
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
    } else if (nodes[i].inst >= 0) {
      b = bvh->inst_boxes[(size_t)nodes[i].inst];  // a placement's leaf (instanced top level)
    } else if (bvh->spatial_splits) {
      // a leaf's references are clipped fragments: its double box, mapped
      // to the normalized space and rounded outward to f32
      const BNode& n = nodes[i];
      for (int k = 0; k < 3; ++k) {
        const double lo = (n.lo[k] - bvh->norm_center[k]) * bvh->norm_scale;
        const double hi = (n.hi[k] - bvh->norm_center[k]) * bvh->norm_scale;
        float fl = (float)lo, fh = (float)hi;
        if ((double)fl > lo) fl = std::nextafter(fl, -INFINITY);
        if ((double)fh < hi) fh = std::nextafter(fh, INFINITY);
        b.lo[k] = fl;
        b.hi[k] = fh;
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
  // SAH-optimal collapse (Ylitie et al. 2017): for every binary node x and
  // slot budget k, the least SAH cost of representing x's subtree with at most
  // k children of the wide node above it -- as one child (a wide node of its
  // own, or a leaf holding its whole subtree when that has at most
  // kMergeLeaf triangles) or dissolved into its two children's budgets.
  // Costs: kCostNode per node visit, kCostTri per triangle test, weighted by
  // surface area.
  constexpr int kMergeLeaf = 4;
  constexpr float kCostNode = 1.0f, kCostTri = 1.2f;
  enum : uint8_t { kAsLeaf = 0, kAsNode = 1, kSplit = 2, kAsPlacement = 3 };
  const size_t nb = nodes.size();
  std::vector<float> slot_cost, dist_cost;
  std::vector<uint8_t> slot_how, split_at;
  std::vector<uint32_t> tri_count;
  if (bvh->sah_collapse) {
    slot_cost.assign(nb * pk::kWide, 0.0f);  // [x][k-1]
    dist_cost.assign(nb * pk::kWide, 0.0f);
    slot_how.assign(nb * pk::kWide, kAsNode);
    split_at.assign(nb * pk::kWide, 1);
    tri_count.assign(nb, 0);
    for (size_t oi = order.size(); oi-- > 0;) {  // children before parents
      const int32_t x = order[oi];
      const float ax = (float)area(x);
      float* sc = &slot_cost[(size_t)x * pk::kWide];
      uint8_t* sh = &slot_how[(size_t)x * pk::kWide];
      if (!is_internal(x)) {
        const bool placement = nodes[x].inst >= 0;
        tri_count[x] = placement ? 0xffffffffu : nodes[x].count;
        const float c = placement ? ax * kCostNode : ax * kCostTri * (float)nodes[x].count;
        for (int k = 0; k < pk::kWide; ++k) {
          sc[k] = c;
          sh[k] = placement ? kAsPlacement : kAsLeaf;
        }
        continue;
      }
      const int32_t l = kid_l(x), r = kid_r(x);
      const uint32_t tl = tri_count[l], tr = tri_count[r];
      tri_count[x] = (tl == 0xffffffffu || tr == 0xffffffffu) ? 0xffffffffu : tl + tr;
      const float* sl = &slot_cost[(size_t)l * pk::kWide];
      const float* sr = &slot_cost[(size_t)r * pk::kWide];
      float* dc = &dist_cost[(size_t)x * pk::kWide];
      uint8_t* sa = &split_at[(size_t)x * pk::kWide];
      for (int k = 2; k <= pk::kWide; ++k) {
        float best = INFINITY;
        int bj = 1;
        for (int j = 1; j < k; ++j) {
          const float c = sl[j - 1] + sr[k - j - 1];
          if (c < best) {
            best = c;
            bj = j;
          }
        }
        dc[k - 1] = best;
        sa[k - 1] = (uint8_t)bj;
      }
      const float as_node = ax * kCostNode + dc[pk::kWide - 1];
      const bool mergeable = tri_count[x] <= (uint32_t)kMergeLeaf;
      const float as_leaf = mergeable ? ax * kCostTri * (float)tri_count[x] : INFINITY;
      const float one = std::min(as_node, as_leaf);
      sc[0] = one;
      sh[0] = as_leaf < as_node ? kAsLeaf : kAsNode;
      for (int k = 2; k <= pk::kWide; ++k) {
        if (dc[k - 1] < one) {
          sc[k - 1] = dc[k - 1];
          sh[k - 1] = kSplit;
        } else {
          sc[k - 1] = one;
          sh[k - 1] = sh[0];
        }
      }
    }
  }
  // the children of the wide node for internal binary node b under the
  // collapse's decisions: (binary node, how)
  std::vector<std::pair<int32_t, uint8_t>> pending;
  auto expand = [&](int32_t b, std::vector<std::pair<int32_t, uint8_t>>& out) {
    pending.clear();
    const int j = split_at[(size_t)b * pk::kWide + pk::kWide - 1];
    std::vector<std::pair<int32_t, int>> work{{kid_r(b), pk::kWide - j}, {kid_l(b), j}};
    while (!work.empty()) {
      const auto [x, k] = work.back();
      work.pop_back();
      const uint8_t how = slot_how[(size_t)x * pk::kWide + k - 1];
      if (how == kSplit) {
        const int jj = split_at[(size_t)x * pk::kWide + k - 1];
        work.push_back({kid_r(x), k - jj});
        work.push_back({kid_l(x), jj});
      } else {
        out.push_back({x, how});
      }
    }
  };
  // a leaf child holding every triangle of binary subtree x, in order
  auto emit_subtree = [&](int32_t x, uint32_t* first, uint32_t* count) {
    *first = (uint32_t)bvh->wtris.size();
    std::vector<int32_t> st{x};
    while (!st.empty()) {
      const int32_t y = st.back();
      st.pop_back();
      if (is_internal(y)) {
        st.push_back(kid_r(y));
        st.push_back(kid_l(y));
      } else {
        uint32_t f;
        emit_leaf(y, &f);
      }
    }
    *count = (uint32_t)bvh->wtris.size() - *first;
  };
  std::vector<std::pair<int32_t, uint8_t>> chosen;
  while (!q.empty()) {
    QItem it = q.front();
    q.pop_front();
    wdepth = std::max(wdepth, it.depth);
    if (bvh->sah_collapse && is_internal(it.bnode)) {
      chosen.clear();
      expand(it.bnode, chosen);
      FBox cb[pk::kWide];
      pk::WNode w;
      std::memset(&w, 0, sizeof(w));
      const int nk = (int)chosen.size();
      w.nchild = (uint8_t)nk;
      for (int i = 0; i < nk; ++i) {
        const int32_t x = chosen[i].first;
        cb[i] = fb[x];
        if (chosen[i].second == kAsNode) {
          const uint32_t wi = (uint32_t)bvh->wnodes.size();
          bvh->wnodes.emplace_back();
          q.push_back({x, wi, it.depth + 1});
          w.child[i] = wi;
          w.meta[i] = 0;
        } else if (chosen[i].second == kAsPlacement) {
          w.child[i] = pk::kInstChild | (uint32_t)nodes[x].inst;
          w.meta[i] = 0;
        } else {
          uint32_t first, count;
          emit_subtree(x, &first, &count);
          w.child[i] = first;
          w.meta[i] = (uint8_t)count;
        }
      }
      quantize_node(w, cb, nk);
      bvh->wnodes[it.wnode] = w;
      continue;
    }
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
      } else if (nodes[kids[i]].inst >= 0) {
        w.child[i] = pk::kInstChild | (uint32_t)nodes[kids[i]].inst;
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
