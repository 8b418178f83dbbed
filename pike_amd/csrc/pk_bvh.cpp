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
// place, O(N) scratch per level) instead of bvh.js's per-node copies; it
// splits a large node's passes (unions, sweeps, binning, partitions) over
// threads and hands subtrees to threads as they free up.  The tree is the
// same for every thread count.
#include <algorithm>
#include <iterator>
#include <cfloat>
#include <cstdio>
#include <cstdlib>
#include <atomic>
#include <cmath>
#include <cstring>
#include <deque>
#include <exception>
#include <future>
#include <memory>
#include <mutex>
#include <new>
#include <numeric>
#include <system_error>
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

// fn(c) for c in [0, parts), part 0 on the calling thread, the others on
// threads of their own (on the calling thread too when no thread can be
// started); an exception from any part is rethrown after all end.
template <class F> void run_parts(int parts, F&& fn) {
  if (parts <= 1) {
    fn(0);
    return;
  }
  std::vector<std::exception_ptr> err(parts);
  auto guarded = [&](int c) {
    try {
      fn(c);
    } catch (...) {
      err[c] = std::current_exception();
    }
  };
  std::vector<std::thread> ts;
  ts.reserve(parts - 1);
  for (int c = 1; c < parts; ++c) {
    try {
      ts.emplace_back(guarded, c);
    } catch (const std::system_error&) {
      for (int d = c; d < parts; ++d) guarded(d);
      break;
    }
  }
  guarded(0);
  for (auto& t : ts) t.join();
  for (auto& e : err)
    if (e) std::rethrow_exception(e);
}
// first item of part c when n items are cut into `parts` consecutive parts
inline size_t part_begin(size_t n, int parts, int c) { return n * (size_t)c / (size_t)parts; }

// A node's passes over its references run on up to `width` threads when it
// holds at least two grains of them.  min and max are exact, so unions taken
// part by part are the serial unions, and sweeps cut into parts keep the
// serial first minimum: the tree does not depend on the thread count.
constexpr size_t kParGrain = 1u << 15;
inline int parts_for(int width, size_t n) {
  return width <= 1 || n < 2 * kParGrain ? 1 : (int)std::min<size_t>((size_t)width, n / kParGrain);
}

// Subtree tasks: a split builds its left subtree on a thread of its own while
// fewer than `threads` threads are working; a thread that waits for the half
// it handed off does not count as working.  Which thread builds a subtree
// changes node numbering only, never the tree.
struct Tasks {
  int threads = 1;
  std::atomic<int> working{1};
  template <class L, class R> void fork(size_t n, L&& left, R&& right) {
    if (n <= 32768 || working.load(std::memory_order_relaxed) >= threads) {
      left();
      right();
      return;
    }
    working.fetch_add(1);
    std::future<void> fut;
    try {
      fut = std::async(std::launch::async, [&]() {
        left();
        working.fetch_sub(1);
      });
    } catch (const std::system_error&) {  // no thread to be had: both halves here
      working.fetch_sub(1);
      left();
      right();
      return;
    }
    right();
    working.fetch_sub(1);
    fut.get();
    working.fetch_add(1);
  }
};

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

// The object-split sweep (bvh.js:184-208) over n boxes in sweep order:
// cost_i = 0.5 + A(front_i)/A(P)*(i+1) + A(back_i)/A(P)*(n-1-i), first strict
// minimum.  Part c of `parts` starts from the union of the boxes before it
// and the union of the boxes after it, so each cost is the serial one.
template <class BoxOf>
void sweep_parts(size_t n, BoxOf box_of, double parent_area, int parts, double* best_cost, uint32_t* best_index) {
  if (parts <= 1) {
    // small sweeps reuse one buffer per thread; large ones take their own
    thread_local std::vector<double> small;
    std::vector<double> large;
    if (n <= kParGrain && small.size() < n) small.resize(kParGrain);
    if (n > kParGrain) large.resize(n);
    double* suf = n <= kParGrain ? small.data() : large.data();
    Box bb;
    box_empty(bb);
    for (size_t i = n; i-- > 0;) {
      box_add(bb, box_of(i));
      suf[i] = box_area(bb);
    }
    Box bf;
    box_empty(bf);
    double bc = INFINITY;
    uint32_t bi = 0;
    for (size_t i = 0; i < n; ++i) {
      box_add(bf, box_of(i));
      const double cst = 0.5 + (box_area(bf) / parent_area) * (double)(i + 1) + (suf[i] / parent_area) * (double)(n - 1 - i);
      if (cst < bc) {
        bc = cst;
        bi = (uint32_t)i;
      }
    }
    *best_cost = bc;
    *best_index = bi;
    return;
  }
  std::vector<Box> before(parts), after(parts);
  for (int c = 0; c < parts; ++c) {
    box_empty(before[c]);
    box_empty(after[c]);
  }
  if (parts > 1) {
    std::vector<Box> tot(parts);
    run_parts(parts, [&](int c) {
      box_empty(tot[c]);
      for (size_t i = part_begin(n, parts, c), e = part_begin(n, parts, c + 1); i < e; ++i) box_add(tot[c], box_of(i));
    });
    for (int c = 1; c < parts; ++c) {
      before[c] = before[c - 1];
      box_add(before[c], tot[c - 1]);
    }
    for (int c = parts - 1; c-- > 0;) {
      after[c] = after[c + 1];
      box_add(after[c], tot[c + 1]);
    }
  }
  std::vector<double> cost(parts);
  std::vector<uint32_t> index(parts);
  run_parts(parts, [&](int c) {
    const size_t b = part_begin(n, parts, c), e = part_begin(n, parts, c + 1);
    std::vector<double> suf(e - b);
    Box bb = after[c];
    for (size_t i = e; i-- > b;) {
      box_add(bb, box_of(i));
      suf[i - b] = box_area(bb);  // union of boxes i..n-1
    }
    Box bf = before[c];
    double bc = INFINITY;
    uint32_t bi = 0;
    for (size_t i = b; i < e; ++i) {
      box_add(bf, box_of(i));
      const double cst =
          0.5 + (box_area(bf) / parent_area) * (double)(i + 1) + (suf[i - b] / parent_area) * (double)(n - 1 - i);
      if (cst < bc) {
        bc = cst;
        bi = (uint32_t)i;
      }
    }
    cost[c] = bc;
    index[c] = bi;
  });
  double bc = INFINITY;
  uint32_t bi = 0;
  for (int c = 0; c < parts; ++c)
    if (cost[c] < bc) {
      bc = cost[c];
      bi = index[c];
    }
  *best_cost = bc;
  *best_index = bi;
}

struct Build {
  const std::vector<Box>* ref_box;
  uint32_t* axis[3];           // three index arrays of length nt
  uint32_t* leaf_out;          // leaf_tris, written per leaf range
  std::vector<uint8_t> in_left;
  NodePool* pool;
  std::atomic<int> degenerate{0};
  std::atomic<int> max_depth{0};
  Tasks tasks;

  // Builds the subtree over [b, b+n) of all three arrays into node `ni`;
  // width = threads for this node's own passes.  Returns the subtree's node
  // count (kept in an internal node's `begin` until the compaction).
  uint32_t node(uint32_t ni, uint32_t b, uint32_t n, int depth, int width) {
    int md = max_depth.load();
    while (depth > md && !max_depth.compare_exchange_weak(md, depth)) {
    }
    const std::vector<Box>& rb = *ref_box;
    const int parts = parts_for(width, n);
    Box bounds;
    {
      std::vector<Box> u(parts);
      run_parts(parts, [&](int c) {
        box_empty(u[c]);
        for (size_t i = part_begin(n, parts, c), e = part_begin(n, parts, c + 1); i < e; ++i)
          box_add(u[c], rb[axis[0][b + i]]);
      });
      box_empty(bounds);
      for (const Box& x : u) box_add(bounds, x);
    }
    BNode& nd = pool->at(ni);
    for (int k = 0; k < 3; ++k) {
      nd.lo[k] = bounds.lo[k];
      nd.hi[k] = bounds.hi[k];
    }
    const double parent_area = box_area(bounds);
    double best_cost = INFINITY;
    int best_axis = -1;
    uint32_t best_index = 0;
    for (int a = 0; a < 3; ++a) {
      const uint32_t* refs = axis[a] + b;
      double c;
      uint32_t i;
      sweep_parts(n, [&](size_t j) -> const Box& { return rb[refs[j]]; }, parent_area, parts, &c, &i);
      if (c < best_cost) {
        best_cost = c;
        best_axis = a;
        best_index = i;
      }
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
      return 1;
    }
    nd.split_axis = (uint8_t)best_axis;
    const uint32_t nl = best_index + 1;
    uint32_t* split = axis[best_axis] + b;
    auto set_left = [&](uint8_t v) {
      run_parts(parts, [&](int c) {
        for (size_t i = part_begin(nl, parts, c), e = part_begin(nl, parts, c + 1); i < e; ++i) in_left[split[i]] = v;
      });
    };
    set_left(1);
    // the other two arrays, partitioned stably (on two threads when the
    // node's passes are split)
    int other[2], no = 0;
    for (int a = 0; a < 3; ++a)
      if (a != best_axis) other[no++] = a;
    auto partition = [&](int c) {
      uint32_t* arr = axis[other[c]] + b;
      std::vector<uint32_t> tmp(n - nl);
      uint32_t li = 0, ri = 0;
      for (uint32_t i = 0; i < n; ++i) {
        uint32_t r = arr[i];
        if (in_left[r]) arr[li++] = r;
        else tmp[ri++] = r;
      }
      std::memcpy(arr + nl, tmp.data(), (size_t)ri * 4);
    };
    if (parts > 1) run_parts(2, partition);
    else
      for (int c = 0; c < 2; ++c) partition(c);
    set_left(0);
    uint32_t l = pool->alloc(), r = pool->alloc();
    {
      BNode& p = pool->at(ni);
      p.left = (int32_t)l;
      p.right = (int32_t)r;
    }
    const int cw = std::max(1, width / 2);
    uint32_t sl = 0, sr = 0;
    tasks.fork(n, [&]() { sl = node(l, b, nl, depth + 1, cw); }, [&]() { sr = node(r, b + nl, n - nl, depth + 1, cw); });
    pool->at(ni).begin = 1 + sl + sr;
    return 1 + sl + sr;
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
  Tasks tasks;

  // a reference's fragments: parts[part + i - first] is its fragment in slab i
  struct Cut {
    uint32_t ref;
    int first, last;
    size_t part;
  };
  // the slab binning of a run of references: per reference its fragments,
  // per slab the bounds and the references entering and leaving there
  struct Bins {
    std::vector<Cut> cuts;
    std::vector<Box> parts;
    Box slab_box[kSpatialBins];
    uint32_t enter[kSpatialBins] = {0}, leave[kSpatialBins] = {0};
  };

  const Box& box_of(uint32_t r) const { return refs->at(r).box; }

  Box union_of(const std::vector<uint32_t>& l, size_t b, size_t e, int parts) const {
    if (parts <= 1) {
      Box r;
      box_empty(r);
      for (size_t i = b; i < e; ++i) box_add(r, box_of(l[i]));
      return r;
    }
    std::vector<Box> u(parts);
    run_parts(parts, [&](int c) {
      box_empty(u[c]);
      for (size_t i = b + part_begin(e - b, parts, c), ie = b + part_begin(e - b, parts, c + 1); i < ie; ++i)
        box_add(u[c], box_of(l[i]));
    });
    Box r;
    box_empty(r);
    for (const Box& x : u) box_add(r, x);
    return r;
  }

  void sweep(const std::vector<uint32_t>& l, double parent_area, int parts, double* best_cost,
             uint32_t* best_index) const {
    sweep_parts(l.size(), [&](size_t i) -> const Box& { return box_of(l[i]); }, parent_area, parts, best_cost,
                best_index);
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

  // Bins l[b, e) into kSpatialBins slabs of [lo, lo + span) along axis a.
  void bin(const std::vector<uint32_t>& l, size_t b, size_t e, int a, double lo, double span, Bins& out) const {
    out.cuts.reserve(e - b);
    out.parts.reserve((e - b) + (e - b) / 2);
    for (auto& sb : out.slab_box) box_empty(sb);
    Box tmp[kSpatialBins];
    for (size_t j = b; j < e; ++j) {
      const uint32_t r = l[j];
      const SRef& ref = refs->at(r);
      const double rmin = ref.box.lo[a] - lo, rmax = ref.box.hi[a] - lo;
      // a box starting on the node's upper face maps to slab 64: clamp
      const int start = std::min(std::max(0, (int)std::floor((rmin / span) * kSpatialBins)), kSpatialBins - 1);
      const int stop = std::max(std::min((int)std::floor((rmax / span) * kSpatialBins), kSpatialBins - 1), start);
      Cut c{r, -1, -1, out.parts.size()};
      if (start == stop) {
        // within one slab: its fragment there is the reference itself
        // (the slab's clip of the triangle contains the reference box)
        c.first = c.last = start;
        out.parts.push_back(ref.box);
        box_add(out.slab_box[start], ref.box);
        ++out.enter[start];
        ++out.leave[start];
        out.cuts.push_back(c);
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
          box_add(out.slab_box[i], fb);
        }
        tmp[i - start] = fb;
      }
      if (c.first < 0) {  // degenerate clip: keep the reference whole in its first slab
        c.first = c.last = start;
        out.parts.push_back(ref.box);
        box_add(out.slab_box[start], ref.box);
      } else {
        out.parts.insert(out.parts.end(), tmp + (c.first - start), tmp + (c.last - start + 1));
      }
      ++out.enter[c.first];
      ++out.leave[c.last];
      out.cuts.push_back(c);
    }
  }

  // Builds the subtree over the three centroid-sorted lists ax into node ni;
  // width = threads for this node's own passes.  Returns the subtree's node
  // count (kept in an internal node's `begin` until the compaction).
  uint32_t node(uint32_t ni, std::vector<uint32_t> ax[3], int depth, int width) {
    int md = max_depth.load();
    while (depth > md && !max_depth.compare_exchange_weak(md, depth)) {
    }
    const uint32_t n = (uint32_t)ax[0].size();
    const int parts = parts_for(width, n);
    // per-axis passes run on three threads when the node's passes are split
    auto each_axis = [&](auto&& f) {
      if (parts > 1) run_parts(3, f);
      else
        for (int k = 0; k < 3; ++k) f(k);
    };
    const Box bounds = union_of(ax[0], 0, n, parts);
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
      sweep(ax[a], parent_area, parts, &c, &i);
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
      const std::vector<uint32_t>& l = ax[best_axis];
      Box f = union_of(l, 0, (size_t)best_index + 1, parts);
      const Box b = union_of(l, best_index, n, parts);
      for (int k = 0; k < 3; ++k) {
        f.lo[k] = std::max(f.lo[k], b.lo[k]);
        f.hi[k] = std::min(f.hi[k], b.hi[k]);
      }
      if (root_area > 0 && box_area(f) / root_area > kSplitAlpha) {
        const int a = best_axis;
        const double lo = bounds.lo[a], span = bounds.hi[a] - lo;
        if (span != 0) {
          std::vector<Bins> bins(parts);
          run_parts(parts, [&](int c) { bin(l, part_begin(n, parts, c), part_begin(n, parts, c + 1), a, lo, span, bins[c]); });
          Box slab_box[kSpatialBins];
          uint32_t enter[kSpatialBins] = {0}, leave[kSpatialBins] = {0};
          for (auto& sb : slab_box) box_empty(sb);
          for (const Bins& bn : bins)
            for (int i = 0; i < kSpatialBins; ++i) {
              box_add(slab_box[i], bn.slab_box[i]);
              enter[i] += bn.enter[i];
              leave[i] += bn.leave[i];
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
            struct Side {
              std::vector<uint32_t> left, right;
              std::vector<SRef> halves;
            };
            std::vector<Side> side(parts);
            run_parts(parts, [&](int c) {
              Side& s = side[c];
              const Bins& bn = bins[c];
              for (const Cut& cu : bn.cuts) {
                if (cu.last < plane) {
                  s.left.push_back(cu.ref);
                } else if (cu.first >= plane) {
                  s.right.push_back(cu.ref);
                } else {
                  Box lb, rb;
                  box_empty(lb);
                  box_empty(rb);
                  for (int i = cu.first; i <= cu.last; ++i) box_add(i < plane ? lb : rb, bn.parts[cu.part + i - cu.first]);
                  const uint32_t t = refs->at(cu.ref).tri;
                  s.halves.push_back(SRef{t, lb});
                  s.halves.push_back(SRef{t, rb});
                }
              }
            });
            std::vector<Bins>().swap(bins);
            size_t nh = 0, nsl = 0, nsr = 0;
            for (const Side& s : side) {
              nh += s.halves.size();
              nsl += s.left.size();
              nsr += s.right.size();
            }
            sp_left.reserve(nsl);
            sp_right.reserve(nsr);
            for (const Side& s : side) {
              sp_left.insert(sp_left.end(), s.left.begin(), s.left.end());
              sp_right.insert(sp_right.end(), s.right.begin(), s.right.end());
            }
            const uint32_t first = nh == 0 ? 0u : refs->try_alloc((uint32_t)nh);
            if (first != UINT32_MAX) {
              uint32_t id = first;
              for (const Side& s : side)
                for (size_t i = 0; i < s.halves.size(); ++i, ++id) {
                  refs->at(id) = s.halves[i];
                  (i % 2 ? new_right : new_left).push_back(id);
                }
              spatial_split = true;
              split_cost = sc;
            } else {
              std::vector<uint32_t>().swap(sp_left);
              std::vector<uint32_t>().swap(sp_right);
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
      return 1;
    }
    pool->at(ni).split_axis = (uint8_t)best_axis;
    // sets mark[id] = v for every id of list s, on the node's threads
    auto set_marks = [&](const std::vector<uint32_t>& s, size_t count, uint8_t v) {
      run_parts(parts, [&](int c) {
        for (size_t i = part_begin(count, parts, c), e = part_begin(count, parts, c + 1); i < e; ++i) mark[s[i]] = v;
      });
    };
    std::vector<uint32_t> lax[3], rax[3];
    if (spatial_split) {
      // whole references keep their centroid order in every axis list (a
      // subsequence of a sorted list is sorted); the new halves are sorted
      // by centroid and merged in (stable: whole references first on ties).
      // The halves' order for axis k is a stable re-sort of their order for
      // axis k-1.
      spatial.fetch_add(1);
      set_marks(sp_left, sp_left.size(), 1);
      set_marks(sp_right, sp_right.size(), 2);
      auto by_centroid = [&](int k) {
        return [this, k](uint32_t x, uint32_t y) {
          const Box& bx = refs->at(x).box;
          const Box& by = refs->at(y).box;
          return (bx.lo[k] + bx.hi[k]) * 0.5 < (by.lo[k] + by.hi[k]) * 0.5;
        };
      };
      std::vector<uint32_t> nls[3], nrs[3];
      for (int k = 0; k < 3; ++k) {
        std::stable_sort(new_left.begin(), new_left.end(), by_centroid(k));
        std::stable_sort(new_right.begin(), new_right.end(), by_centroid(k));
        nls[k] = new_left;
        nrs[k] = new_right;
      }
      each_axis([&](int k) {
        std::vector<uint32_t> kl, kr;
        kl.reserve(sp_left.size());
        kr.reserve(sp_right.size());
        for (uint32_t r : ax[k]) {
          const uint8_t m = mark[r];
          if (m == 1) kl.push_back(r);
          else if (m == 2) kr.push_back(r);
        }
        std::vector<uint32_t>().swap(ax[k]);
        lax[k].reserve(kl.size() + nls[k].size());
        rax[k].reserve(kr.size() + nrs[k].size());
        std::merge(kl.begin(), kl.end(), nls[k].begin(), nls[k].end(), std::back_inserter(lax[k]), by_centroid(k));
        std::merge(kr.begin(), kr.end(), nrs[k].begin(), nrs[k].end(), std::back_inserter(rax[k]), by_centroid(k));
      });
      set_marks(sp_left, sp_left.size(), 0);
      set_marks(sp_right, sp_right.size(), 0);
    } else {
      const size_t nl = (size_t)split_index + 1;
      set_marks(ax[best_axis], nl, 1);
      each_axis([&](int k) {
        lax[k].reserve(nl);
        rax[k].reserve(n - nl);
        for (uint32_t r : ax[k]) (mark[r] ? lax[k] : rax[k]).push_back(r);
        std::vector<uint32_t>().swap(ax[k]);
      });
      set_marks(lax[0], nl, 0);
    }
    uint32_t l = pool->alloc(), r = pool->alloc();
    {
      BNode& p = pool->at(ni);
      p.left = (int32_t)l;
      p.right = (int32_t)r;
    }
    const int cw = std::max(1, width / 2);
    uint32_t sl = 0, sr = 0;
    tasks.fork(n, [&]() { sl = node(l, lax, depth + 1, cw); }, [&]() { sr = node(r, rax, depth + 1, cw); });
    pool->at(ni).begin = 1 + sl + sr;
    return 1 + sl + sr;
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

void pk_bvh_make_wide(pk_bvh* bvh, int threads);

static int build_bvh(const double* positions, uint32_t nv, const int32_t* triangles, uint32_t nt, int num_threads,
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
    for (int c = 0; c < 3; ++c) {
      if (triangles[i * 4 + c] < 0 || (uint32_t)triangles[i * 4 + c] >= nv) {
        pk_set_error("pk_bvh_build: triangle " + std::to_string(i) + " references a vertex out of range");
        return PK_E_INVALID;
      }
      const double* p = positions + (size_t)triangles[i * 4 + c] * 3;
      if (!std::isfinite(p[0]) || !std::isfinite(p[1]) || !std::isfinite(p[2])) {
        pk_set_error("pk_bvh_build: triangle " + std::to_string(i) + " has a vertex with a non-finite coordinate");
        return PK_E_INVALID;
      }
    }
  if (num_threads <= 0) num_threads = (int)std::max(1u, std::thread::hardware_concurrency());
  auto* bvh = new pk_bvh();
  std::unique_ptr<pk_bvh> guard(bvh);
  bvh->nv = nv;
  bvh->nt = nt;
  bvh->tris.assign(triangles, triangles + (size_t)nt * 4);
  // reference boxes and centroids (primitives.js:17-33)
  std::vector<Box> ref_box(nt);
  std::vector<double> cen((size_t)nt * 3);
  const int tparts = parts_for(num_threads, nt);
  run_parts(tparts, [&](int c) {
    for (size_t t = part_begin(nt, tparts, c), e = part_begin(nt, tparts, c + 1); t < e; ++t) {
      Box b;
      box_empty(b);
      for (int v = 0; v < 3; ++v) {
        const double* p = positions + (size_t)triangles[t * 4 + v] * 3;
        for (int k = 0; k < 3; ++k) {
          b.lo[k] = std::min(p[k], b.lo[k]);
          b.hi[k] = std::max(p[k], b.hi[k]);
        }
      }
      ref_box[t] = b;
      for (int k = 0; k < 3; ++k) cen[t * 3 + k] = (b.lo[k] + b.hi[k]) * 0.5;
    }
  });
  std::vector<uint32_t> arr((size_t)nt * 3);
  uint32_t* ax[3] = {arr.data(), arr.data() + nt, arr.data() + 2 * (size_t)nt};
  // per axis: triangles by centroid, ties by index (a stable sort of
  // 0..nt-1), as (centroid, index) pairs sorted in parts on threads of their
  // own and merged pairwise (stable: the earlier part first on ties)
  run_parts(3, [&](int a) {
    typedef std::pair<double, uint32_t> Key;
    std::vector<Key> kv(nt);
    for (uint32_t t = 0; t < nt; ++t) kv[t] = Key(cen[(size_t)t * 3 + a], t);
    auto less = [](const Key& x, const Key& y) { return x.first < y.first; };
    const int parts = parts_for(std::max(1, num_threads / 3), nt);
    run_parts(parts, [&](int c) {
      std::stable_sort(kv.begin() + part_begin(nt, parts, c), kv.begin() + part_begin(nt, parts, c + 1), less);
    });
    for (int w = 1; w < parts; w *= 2) {
      const int merges = (parts - w + 2 * w - 1) / (2 * w);
      run_parts(merges, [&](int m) {
        const int c = m * 2 * w;
        std::inplace_merge(kv.begin() + part_begin(nt, parts, c), kv.begin() + part_begin(nt, parts, c + w),
                           kv.begin() + part_begin(nt, parts, std::min(c + 2 * w, parts)), less);
      });
    }
    for (uint32_t i = 0; i < nt; ++i) ax[a][i] = kv[i].second;
  });
  std::vector<double>().swap(cen);
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
    b.tasks.threads = num_threads;
    b.node(root, 0, nt, 1, num_threads);
    bvh->degenerate = b.degenerate.load();
    bvh->binary_depth = b.max_depth.load();
  } else {
    Pool<SRef> refs(ref_cap);
    refs.try_alloc(nt);  // ids 0..nt-1: the triangles
    run_parts(tparts, [&](int c) {
      for (size_t t = part_begin(nt, tparts, c), e = part_begin(nt, tparts, c + 1); t < e; ++t)
        refs.at((uint32_t)t) = SRef{(uint32_t)t, ref_box[t]};
    });
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
    b.tasks.threads = num_threads;
    b.node(root, lists, 1, num_threads);
    bvh->degenerate = b.degenerate.load();
    bvh->binary_depth = b.max_depth.load();
    bvh->leaf_tris.swap(b.leaf_out);
    bvh->num_spatial_splits = b.spatial.load();
  }
  NodePool& pool_ref = *pool;
  // compact: nodes renumbered depth-first (root 0, a node's left subtree
  // before its right one), so that the passes over the tree walk memory in
  // order whichever thread built which subtree.  Subtree sizes give every
  // node's number up front: subtrees are copied on threads.
  const uint32_t nn = pool_ref.size();
  bvh->nodes.resize(nn);
  {
    auto size_of = [&](uint32_t id) { return pool_ref.at(id).leaf ? 1u : pool_ref.at(id).begin; };
    // copies pool subtree `id` to numbers p.. ; stops at subtrees of at most
    // `grain` nodes when `rest` is given, listing them there
    auto copy = [&](uint32_t id, uint32_t p, uint32_t grain, std::vector<std::pair<uint32_t, uint32_t>>* rest) {
      std::vector<std::pair<uint32_t, uint32_t>> st{{id, p}};
      while (!st.empty()) {
        const auto [x, q] = st.back();
        st.pop_back();
        if (rest && size_of(x) <= grain) {
          rest->emplace_back(x, q);
          continue;
        }
        BNode& d = bvh->nodes[q];
        d = pool_ref.at(x);
        if (!d.leaf) {
          const uint32_t l = (uint32_t)d.left, r = (uint32_t)d.right;
          d.begin = 0;
          d.left = (int32_t)(q + 1);
          d.right = (int32_t)(q + 1 + size_of(l));
          st.emplace_back(r, (uint32_t)d.right);
          st.emplace_back(l, (uint32_t)d.left);
        }
      }
    };
    std::vector<std::pair<uint32_t, uint32_t>> jobs;
    const int cparts = parts_for(num_threads, nn);
    copy(0, 0, cparts > 1 ? nn / (uint32_t)(8 * cparts) : nn, &jobs);
    run_parts(cparts, [&](int c) {
      for (size_t j = part_begin(jobs.size(), cparts, c), e = part_begin(jobs.size(), cparts, c + 1); j < e; ++j)
        copy(jobs[j].first, jobs[j].second, 0, nullptr);
    });
  }
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
  pk_bvh_make_wide(bvh, num_threads);
  *out = guard.release();
  return PK_OK;
}

extern "C" {

int pk_bvh_build(const double* positions, uint32_t nv, const int32_t* triangles, uint32_t nt, int num_threads,
                 pk_bvh** out) {
  return pk_bvh_build_ex(positions, nv, triangles, nt, num_threads, 0, out);
}

// (a build that runs out of host memory frees what it had built)
int pk_bvh_build_ex(const double* positions, uint32_t nv, const int32_t* triangles, uint32_t nt, int num_threads,
                    uint32_t flags, pk_bvh** out) {
  return pk_guard("pk_bvh_build", [&]() { return build_bvh(positions, nv, triangles, nt, num_threads, flags, out); });
}

int pk_bvh_destroy(pk_bvh* bvh) {
  return pk_guard("pk_bvh_destroy", [&]() -> int {
    delete bvh;
    return PK_OK;
  });
}

int pk_bvh_info(const pk_bvh* bvh, int64_t* info) {
  return pk_guard("pk_bvh_info", [&]() -> int {
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
  });
}

int pk_bvh_normalizer(const pk_bvh* bvh, double* norm4) {
  return pk_guard("pk_bvh_normalizer", [&]() -> int {
    if (!bvh || !norm4) {
      pk_set_error("pk_bvh_normalizer: null argument");
      return PK_E_INVALID;
    }
    for (int k = 0; k < 3; ++k) norm4[k] = bvh->norm_center[k];
    norm4[3] = bvh->norm_scale;
    return PK_OK;
  });
}

int pk_bvh_export_reference(const pk_bvh* bvh, int32_t* nodes5, int32_t* tris4, float* verts4, double* norm4) {
  return pk_guard("pk_bvh_export_reference", [&]() -> int {
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
  });
}

int pk_bvh_export_wide4(const pk_bvh* bvh, void* nodes64, int64_t* count) {
  return pk_guard("pk_bvh_export_wide4", [&]() -> int {
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
  });
}

int pk_bvh_export_wide(const pk_bvh* bvh, void* nodes, void* tris, int32_t* tri_index) {
  return pk_guard("pk_bvh_export_wide", [&]() -> int {
    if (!bvh) {
      pk_set_error("pk_bvh_export_wide: null bvh");
      return PK_E_INVALID;
    }
    if (nodes) std::memcpy(nodes, bvh->wnodes.data(), bvh->wnodes.size() * sizeof(pk::WNode));
    if (tris) std::memcpy(tris, bvh->wtris.data(), bvh->wtris.size() * sizeof(pk::TriRecord));
    if (tri_index) std::memcpy(tri_index, bvh->wtri_index.data(), bvh->wtri_index.size() * sizeof(int32_t));
    return PK_OK;
  });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Wide collapse: each wide node takes its binary node's two children and
// repeatedly opens the internal child with the largest surface area until it
// holds 8 children (or only leaves remain).  Wide nodes are numbered in
// breadth-first order so the top levels share lines; leaf triangles are laid
// out in the order the collapse reaches them.
// ---------------------------------------------------------------------------
void pk_bvh_make_wide(pk_bvh* bvh, int threads) {
  std::vector<BNode>& nodes = bvh->nodes;
  // split oversized leaves into virtual halves for the collapse only
  const size_t n0 = nodes.size();
  std::vector<size_t> work;
  for (size_t i = 0; i < n0; ++i) {
    if (!nodes[i].leaf || nodes[i].count <= (uint32_t)pk::kLeafMax) continue;
    work.assign(1, i);
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
  // The bottom-up passes (these boxes, the collapse's dynamic program) visit
  // the subtrees below a cut on threads, then the nodes above the cut.
  if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
  const int tparts = parts_for(threads, nodes.size());
  std::vector<int32_t> top, roots;  // above the cut (parents first); subtrees below it
  {
    int cut = 0;
    while (tparts > 1 && (1 << cut) < 8 * tparts) ++cut;
    std::vector<std::pair<int32_t, int>> st{{0, 0}};
    while (!st.empty()) {
      const auto [x, d] = st.back();
      st.pop_back();
      if (d >= cut || !is_internal(x)) {
        roots.push_back(x);
        continue;
      }
      top.push_back(x);
      st.push_back({kid_l(x), d + 1});
      st.push_back({kid_r(x), d + 1});
    }
  }
  auto for_bottom_up = [&](const auto& visit) {  // visit(x) after visit(children of x)
    run_parts(tparts, [&](int c) {
      std::vector<int32_t> order, st;
      for (size_t j = part_begin(roots.size(), tparts, c), e = part_begin(roots.size(), tparts, c + 1); j < e; ++j) {
        order.clear();
        st.assign(1, roots[j]);
        while (!st.empty()) {
          const int32_t i = st.back();
          st.pop_back();
          order.push_back(i);
          if (is_internal(i)) {
            st.push_back(kid_l(i));
            st.push_back(kid_r(i));
          }
        }
        for (size_t oi = order.size(); oi-- > 0;) visit(order[oi]);
      }
    });
    for (size_t t = top.size(); t-- > 0;) visit(top[t]);
  };
  for_bottom_up([&](int32_t i) {
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
  });
  auto area = [&](int32_t i) {
    const FBox& b = fb[i];
    double x = std::max((double)b.hi[0] - b.lo[0], 0.0), y = std::max((double)b.hi[1] - b.lo[1], 0.0),
           z = std::max((double)b.hi[2] - b.lo[2], 0.0);
    return x * y + x * z + y * z;
  };
  bvh->wnodes.clear();
  bvh->wtris.clear();
  bvh->wtri_index.clear();
  // the walk below numbers the triangle records; (binary leaf, its first
  // record) pairs are filled in afterwards, on threads
  std::vector<std::pair<int32_t, uint32_t>> fills;
  uint32_t num_records = 0;
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
    *first = num_records;
    fills.emplace_back(li, num_records);
    num_records += nodes[li].count;
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
    for_bottom_up([&](int32_t x) {
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
        return;
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
    });
  }
  // the children of the wide node for internal binary node b under the
  // collapse's decisions: (binary node, how)
  std::vector<std::pair<int32_t, uint8_t>> pending;
  std::vector<std::pair<int32_t, int>> expand_work;
  auto expand = [&](int32_t b, std::vector<std::pair<int32_t, uint8_t>>& out) {
    pending.clear();
    const int j = split_at[(size_t)b * pk::kWide + pk::kWide - 1];
    std::vector<std::pair<int32_t, int>>& work = expand_work;
    work.assign({{kid_r(b), pk::kWide - j}, {kid_l(b), j}});
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
  std::vector<int32_t> subtree_stack;
  // a leaf child holding every triangle of binary subtree x, in order
  auto emit_subtree = [&](int32_t x, uint32_t* first, uint32_t* count) {
    *first = num_records;
    std::vector<int32_t>& st = subtree_stack;
    st.assign(1, x);
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
    *count = num_records - *first;
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
  // triangle records: vertex 0 and the two edges of the normalized f32
  // triangle, and its original index
  bvh->wtris.resize(num_records);
  bvh->wtri_index.resize((size_t)num_records * 4);
  const int parts = parts_for(threads, num_records) > 1 ? std::min<int>(threads, (int)fills.size()) : 1;
  run_parts(parts, [&](int c) {
    for (size_t f = part_begin(fills.size(), parts, c), fe = part_begin(fills.size(), parts, c + 1); f < fe; ++f) {
      const BNode& n = nodes[fills[f].first];
      for (uint32_t t = 0; t < n.count; ++t) {
        const uint32_t id = bvh->leaf_tris[n.begin + t];
        const int32_t* tr = &bvh->tris[(size_t)id * 4];
        const float* p0 = &V[(size_t)tr[0] * 3];
        const float* p1 = &V[(size_t)tr[1] * 3];
        const float* p2 = &V[(size_t)tr[2] * 3];
        pk::TriRecord& rec = bvh->wtris[fills[f].second + t];
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
        for (int k = 0; k < 4; ++k) bvh->wtri_index[((size_t)fills[f].second + t) * 4 + k] = tr[k];
      }
    }
  });
}
