// pk_layout.h — device data layout shared by the host runtime and the kernels.
//
// Scene geometry lives in the normalized space trace_pass.js:202-211 maps the
// scene into (longest axis spans [-1, 1]).  Everything the traversal reads is
// packed to HBM-line granularity:
//   * wide BVH node: 128 bytes = one L2/HBM line, 8 children with 8-bit
//     quantized boxes (conservative), explicit child references;
//   * triangle record: 48 bytes = {v0, id}, {e1, 0}, {e2, 0} in leaf order,
//     e1 = v1 - v0 and e2 = v2 - v0 precomputed in f32 — the same rounding the
//     reference intersector performs (trace.wgsl:98-99), so the Möller–Trumbore
//     result is bit-identical while saving two vertex gathers per test.
#pragma once
#include <stdint.h>

namespace pk {

constexpr int kWide = 8;            // children per node
constexpr int kLeafMax = 255;       // triangles per leaf child (u8 count)
constexpr float kMaxT = 100000.0f;  // trace.wgsl:2
constexpr float kEpsTrace = 0.00000001f;   // trace.wgsl:1
constexpr float kEpsShadow = 0.0000001f;   // trace_shadow.wgsl:1, shade_hit.wgsl:1
// Box quantization margin in normalized units (the scene spans 2.0): covers
// the slab test's f32 rounding for ray origins within a few scene radii.
constexpr double kBoxMargin = 1.0 / 131072.0;

// 128-byte wide node.  Child i (i < nchild) has box
//   lo = origin + qlo[i] * 2^(e-127),  hi = origin + qhi[i] * 2^(e-127)
// per axis; meta[i] == 0 marks an internal child (child[i] = node index),
// meta[i] = k > 0 a leaf of k triangles starting at triangle child[i];
// leafmask bit i = (meta[i] != 0), so traversal reads it with one byte
// extract instead of testing eight bytes.
//
// The builders lay a node's children out in slot order: its internal
// children are the consecutive nodes ibase, ibase + 1, ... and its leaf
// children's triangle ranges follow each other from record tbase.  So the
// traversal finds child i from the header alone -- node ibase + (internal
// slots below i), triangles tbase + (leaf counts below i) -- without loading
// or selecting among the eight child words.  Placement children of an
// instanced top level (instmask bit i) are outside the runs: the traversal
// reads their child word, the only one it needs.  The child words stay for
// that and for host-side consumers.
struct alignas(128) WNode {
  float ox, oy, oz;       // 0
  uint8_t ex, ey, ez;     // 12: biased exponents
  uint8_t nchild;         // 15
  uint32_t child[8];      // 16
  uint8_t meta[8];        // 48
  uint8_t qlox[8];        // 56
  uint8_t qhix[8];        // 64
  uint8_t qloy[8];        // 72
  uint8_t qhiy[8];        // 80
  uint8_t qloz[8];        // 88
  uint8_t qhiz[8];        // 96
  uint8_t leafmask;       // 104
  uint8_t instmask;       // 105: placement children (instanced top level)
  uint8_t pad8[2];        // 106
  uint32_t ibase;         // 108: first internal child (0 without one)
  uint32_t tbase;         // 112: first triangle record of the first leaf child (0 without one)
  uint32_t pad[3];        // 116..127
};
static_assert(sizeof(WNode) == 128, "wide node must be one 128-byte line");

// Sets ibase/tbase and instmask from the child words.
inline void wnode_link(WNode& w) {
  bool have_i = false, have_t = false;
  w.ibase = 0;
  w.tbase = 0;
  w.instmask = 0;
  for (int c = 0; c < w.nchild; ++c) {
    if (w.meta[c]) {
      if (!have_t) w.tbase = w.child[c];
      have_t = true;
    } else if (w.child[c] & 0x80000000u) {
      w.instmask |= (uint8_t)(1u << c);
    } else {
      if (!have_i) w.ibase = w.child[c];
      have_i = true;
    }
  }
}

// True when the child words follow the slot-order runs from ibase/tbase and
// instmask marks exactly the placement children.
inline bool wnode_linked(const WNode& w) {
  uint32_t ni = w.ibase, nt = w.tbase;
  uint32_t im = 0;
  for (int c = 0; c < w.nchild; ++c) {
    if (w.meta[c]) {
      if (w.child[c] != nt) return false;
      nt += w.meta[c];
    } else if (w.child[c] & 0x80000000u) {
      im |= 1u << c;
    } else {
      if (w.child[c] != ni) return false;
      ++ni;
    }
  }
  return im == w.instmask;
}

// Moves a node into a merged layout whose nodes start at noff and triangle
// records at toff (single-level BVHs: no placement children).
inline void wnode_rebase(WNode& w, uint32_t noff, uint32_t toff) {
  for (int c = 0; c < w.nchild; ++c) w.child[c] += w.meta[c] ? toff : noff;
  w.ibase += noff;
  w.tbase += toff;
}

// 64-byte 4-wide node of a single-level BVH: the same fields as WNode for
// four children, without child words (children come from ibase/tbase only).
// The 4-wide layout is derived from the 8-wide one and shares its triangle
// and shading records: node i (< the 8-wide node count) stands for 8-wide
// node i -- its children when it has at most four, else two sub-nodes that
// split its slots at one cut -- and the sub-nodes follow after.  Every box is
// the 8-wide node's quantized box (a sub-node's box is the byte-wise union
// of its slots' boxes in the same frame), so nothing is quantized again.
struct alignas(64) WNode4 {
  float ox, oy, oz;       // 0
  uint8_t ex, ey, ez;     // 12
  uint8_t nchild;         // 15
  uint8_t meta[4];        // 16
  uint8_t qlox[4];        // 20
  uint8_t qhix[4];        // 24
  uint8_t qloy[4];        // 28
  uint8_t qhiy[4];        // 32
  uint8_t qloz[4];        // 36
  uint8_t qhiz[4];        // 40
  uint8_t leafmask;       // 44
  uint8_t pad8[3];        // 45
  uint32_t ibase;         // 48
  uint32_t tbase;         // 52
  uint32_t pad[2];        // 56..63
};
static_assert(sizeof(WNode4) == 64, "4-wide node must be 64 bytes");

struct TriRecord {        // 48 bytes
  float v0x, v0y, v0z;
  uint32_t id;            // original scene.indices position
  float e1x, e1y, e1z;
  uint32_t pad0;
  float e2x, e2y, e2z;
  uint32_t pad1;
};
static_assert(sizeof(TriRecord) == 48, "triangle record must be 48 bytes");

// Instancing.  A top-level wide node's child with kInstChild set (and meta
// 0) is instance child & ~kInstChild; entering it maps the ray into the
// instance's mesh space (where that mesh's wide BVH and triangle records
// live) and pushes kReturnMarker, whose pop maps the ray back.
constexpr uint32_t kInstChild = 0x80000000u;
constexpr uint32_t kReturnMarker = 0xffffffffu;

// 128 bytes: the traversal reads the first 64, shading the rest.
struct alignas(64) InstRecord {
  float w2o[12];        // scene (normalized) -> mesh space, row-major 3x4
  uint32_t root;        // the mesh's root wide node
  uint32_t id_delta;    // baked triangle id = mesh triangle id + id_delta
  uint32_t leaf_delta;  // virtual leaf index = triangle record index + leaf_delta
  int32_t material;     // -1: the triangles' own
  float o2w[9];         // mesh -> scene linear part, row-major 3x3
  uint32_t leaf_base;   // first virtual leaf index of this instance
  uint32_t pad[6];
};
static_assert(sizeof(InstRecord) == 128, "instance record must be 128 bytes");

}  // namespace pk
