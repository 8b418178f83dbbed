// pk_device.h — gfx950 device code: wide-BVH traversal, intersection and
// shading helpers used by the wavefront kernels in pk_kernels.hip.
//
// The f32 arithmetic follows the contract in pk_math.h (no contraction,
// correctly rounded divide/sqrt, project-written transcendentals), so the
// values written here equal the CPU restatement's bit for bit.
#pragma once
#include <hip/hip_runtime.h>

#include "pk_layout.h"
#include "pk_math.h"

namespace pk {

using pkm::v3;

struct DevScene {
  const WNode* nodes;
  const WNode4* nodes4;      // the derived 4-wide layout (single-level scenes; else null)
  const float4* tris;        // TriRecord as 3 float4, leaf order
  // Shading record per leaf-order triangle, 64 bytes: the quantized
  // attributes {tangent, bitangent, normal, uv} of its vertices i0, i1, i2
  // and {material, 0, 0, 0}, so a hit reads one contiguous half line
  // instead of an index record and three vertex records.
  const uint4* shade_rec;
  const int4* materials;     // {diffMap, metRoughMap, normMap, emitMap}
  const void* tex;
  uint32_t tw, th, tl, tfmt;
  // Evaluation copy of the atlas for differentiation with detached sampling:
  // BSDF values (base color; GGX D and G roughness) read it, while sampling
  // decisions and pdfs (opacity, metalness, sampling roughness) read `tex`.
  // It aliases `tex` unless parameters were written.  Environment radiance
  // is only ever evaluated (sampling reads the frozen importance tables), so
  // it needs no second copy.
  const void* tex_eval;
  const float4* env_rad;     // decoded radiance (rgb, 0)
  const int2* env_bins;
  const int2* env_coords;
  const float* env_pdf;
  const float* row_cdf;
  const float* col_cdf;
  const uint32_t* row_guide;  // cutpoint guides of the two CDFs
  const uint32_t* col_guide;
  uint32_t ew, eh, nbins;
  int env_mode;              // 0 none, 1 bins, 2 cdf
  // instancing: nodes[0] is the top level; hits carry virtual leaf indices
  const InstRecord* inst;
  const uint32_t* inst_leaf_base;  // ascending, one per instance
  uint32_t num_inst;               // 0: a single-level BVH
  uint32_t world_leaves;           // leaves below this are the scene's own triangles
  // traversal stack entries per ray (<= kStackMax), set per launch from the
  // context; a ray that needs more flags an overflow
  uint32_t stack_limit;
  // leaf phase trigger, in 16ths of the wave's tracing lanes that hold
  // deferred leaves (8: half), set per launch from the kernel class's shape
  uint32_t leaf16;
};

__device__ inline v3 ld3(float4 a) { return pkm::mk(a.x, a.y, a.z); }

// Möller–Trumbore exactly as trace.wgsl:97-114, with e1/e2 precomputed.
// Returns the distance if accepted (eps < t < kMaxT), otherwise kMaxT.
__device__ inline float intersect_tri(const float4* tris, uint32_t k, v3 o, v3 d, float eps, uint32_t* id,
                                      uint32_t* bary) {
  float4 a = tris[3 * k], b = tris[3 * k + 1], c = tris[3 * k + 2];
  v3 p0 = ld3(a), e1 = ld3(b), e2 = ld3(c);
  *id = __float_as_uint(a.w);
  v3 p = pkm::cross(d, e2);
  float det = pkm::dot(e1, p);
  if (pkm::abs_(det) < eps) return kMaxT;
  float inv_det = 1.0f / det;
  v3 tv = pkm::sub(o, p0);
  float u = pkm::dot(tv, p) * inv_det;
  if (u < 0.0f || u > 1.0f) return kMaxT;
  v3 q = pkm::cross(tv, e1);
  float v = pkm::dot(d, q) * inv_det;
  if (v < 0.0f || u + v > 1.0f) return kMaxT;
  float dist = pkm::dot(e2, q) * inv_det;
  *bary = pkm::pack2x16unorm(u, v);
  return dist > eps ? dist : kMaxT;
}

// Per-lane traversal stack: kLds entries in LDS ([entry][lane], so the
// 64 lanes of a wave hit 64 distinct banks), the rest spilled to a per-wave
// region in global memory.  An entry is a node index and the entry distance
// of its box (for culling at pop time).  The frame and ray-query traversal
// kernels come in 8, 12 and 16 LDS entries per lane (4, 6 and 8 KB per wave:
// deeper stacks spill less, shallower ones let more waves share a CU's
// 160 KB); the path megakernel, the adjoint and the counting and instanced
// variants use kStackLds.  Every wave's spill region is kSpillStride entries
// per lane, enough for the shallowest LDS part.
constexpr int kStackLds = 12;
constexpr int kStackLdsMin = 8;
constexpr int kStackMax = 64;
constexpr int kSpillStride = kStackMax - kStackLdsMin;
// Occupancy target of the traversal kernels.  With the 8-entry LDS part,
// 8 waves per SIMD (64 VGPRs): 32 one-wave blocks per CU use 32 * 5 KB (the
// 4 KB stack part and the cooperative leaf phase's 1 KB) of the 160 KB LDS,
// and the single-level kernels keep their traversal state in 64 VGPRs (a few
// loop invariants are reloaded from scratch).  The 12- and 16-entry parts
// fit at most 22 and 17 waves per CU in LDS, so they target 6 waves per SIMD
// (80 VGPRs).
constexpr int kTraceWavesPerSimd = 6;
constexpr int trace_waves_per_simd(int lds) { return lds <= kStackLdsMin ? 8 : kTraceWavesPerSimd; }

// Culling bound: a box entered beyond the best hit by more than the
// intersector's rounding cannot hold a better (t, id) hit.
__device__ inline float cull_bound(float tbest) { return fmaf(tbest, 0x1.004p+0f, 0x1.0p-20f); }

struct TraverseStats {
  uint32_t nodes, tris;
};

// Pointers into LDS (address space 3): accesses through them are ds_
// operations, never flat ones that have to test which memory they reach.
using lds_u32 = __attribute__((address_space(3))) uint32_t*;
using lds_f32 = __attribute__((address_space(3))) float*;
using lds_u64 = __attribute__((address_space(3))) unsigned long long*;

struct Stack {
  lds_u32 lds_node;     // [kLds][64]
  lds_f32 lds_t;        // [kLds][64]
  uint2* spill;         // [kSpillStride][64] for this wave
  int lane;
  int limit;            // entries a ray may hold, in [1, kStackMax]
  // [128] per wave in LDS for a cooperative leaf phase (the wave-level trace
  // loops; null elsewhere): [lane] each lane's best (t bits, triangle id),
  // [64 + lane] the record and barycentrics of that hit
  lds_u64 key;
};

// State of one ray's traversal, kept in registers between steps.  With
// instancing, o/d are the ray in the space of the BVH being walked and
// wo/wd the scene-space ray; id_delta/leaf_delta turn the current mesh's
// triangle ids and record indices into baked ids and virtual leaves.  The
// single-level kernels never touch those fields, so they cost no registers.
struct RayState {
  v3 o, d;
  float invx, invy, invz;
  float tbest;
  // best hit: leaf-order triangle record (0xffffffff: none), packed
  // barycentrics and, for the instanced traversal only, the baked triangle
  // id (its record indices are virtual); the single-level traversal reads the
  // id back from the record (ray_hit_id), so it holds no register for it
  uint32_t best_id, best_leaf, best_bary;
  uint32_t node;
  int sp;
  bool have;       // `node` is to be visited
  bool overflow;
  v3 wo, wd;
  uint32_t id_delta, leaf_delta;
  // leaf children found by a node visit and not yet intersected: the node
  // and its hit leaf-child mask (0: none)
  uint32_t pnode, pmask;
};

__device__ inline void ray_set_inv(RayState& r) {
  const v3 d = r.d;
  r.invx = 1.0f / (pkm::abs_(d.x) < 1e-30f ? pkm::copysign_(1e-30f, d.x) : d.x);
  r.invy = 1.0f / (pkm::abs_(d.y) < 1e-30f ? pkm::copysign_(1e-30f, d.y) : d.y);
  r.invz = 1.0f / (pkm::abs_(d.z) < 1e-30f ? pkm::copysign_(1e-30f, d.z) : d.z);
}

// x' = A x + b with A, b the rows of a 3x4 (f32, no contraction); the
// oracle evaluates the same sums in the same order
__device__ inline v3 xform_point(const float* m, v3 p) {
  return pkm::mk(((m[0] * p.x + m[1] * p.y) + m[2] * p.z) + m[3], ((m[4] * p.x + m[5] * p.y) + m[6] * p.z) + m[7],
                 ((m[8] * p.x + m[9] * p.y) + m[10] * p.z) + m[11]);
}
__device__ inline v3 xform_dir(const float* m, v3 v) {
  return pkm::mk((m[0] * v.x + m[1] * v.y) + m[2] * v.z, (m[4] * v.x + m[5] * v.y) + m[6] * v.z,
                 (m[8] * v.x + m[9] * v.y) + m[10] * v.z);
}

template <bool kAnyHit>
__device__ inline void ray_init(RayState& r, v3 o, v3 d) {
  r.o = o;
  r.d = d;
  ray_set_inv(r);
  r.tbest = kMaxT;
  r.best_id = 0xffffffffu;
  r.best_leaf = 0xffffffffu;
  r.best_bary = 0;
  r.node = 0;
  r.sp = 0;
  r.have = true;
  r.overflow = false;
  r.id_delta = 0;
  r.leaf_delta = 0;
  r.pnode = 0;
  r.pmask = 0;
}

// Culling bound of a ray, recomputed from tbest where it is used (one fma)
// rather than held in a register; any-hit rays never tighten it.
template <bool kAnyHit>
__device__ inline float ray_cull(const RayState& r) {
  return kAnyHit ? kMaxT * 2.0f : cull_bound(r.tbest);
}

__device__ inline bool ray_missed(const RayState& r) { return r.best_leaf == 0xffffffffu; }

// Original triangle id of the best hit (0xffffffff: none).
template <bool kInst>
__device__ inline uint32_t ray_hit_id(const DevScene& S, const RayState& r) {
  if (kInst) return r.best_id;
  return ray_missed(r) ? 0xffffffffu : __float_as_uint(S.tris[3 * (size_t)r.best_leaf].w);
}

// A push beyond the limit drops the entry (its subtree is not visited) and
// flags the ray; the traversal still ends normally.
template <int kLds>
__device__ inline void stack_push(const Stack& stk, RayState& r, uint32_t node, float t) {
  const int lane = stk.lane;
  if (r.sp >= stk.limit) {
    r.overflow = true;
  } else if (r.sp < kLds) {
    stk.lds_node[r.sp * 64 + lane] = node;
    stk.lds_t[r.sp * 64 + lane] = t;
    ++r.sp;
  } else {
    stk.spill[(r.sp - kLds) * 64 + lane] = make_uint2(node, __float_as_uint(t));
    ++r.sp;
  }
}

__device__ inline void cswap(uint32_t& a, uint32_t& b) {
  const uint32_t lo = min(a, b), hi = max(a, b);
  a = lo;
  b = hi;
}

// First triangle record of leaf child c (c < 8) of a linked node: tbase
// plus the leaf counts of the slots below c (internal slots count 0), summed
// bytewise with v_sad_u8 against zero.
__device__ inline uint32_t leaf_first(uint32_t tbase, uint32_t meta0, uint32_t meta1, uint32_t c) {
  const uint32_t b = 8u * c;
  const uint32_t lo = b >= 32u ? meta0 : meta0 & ((1u << b) - 1u);
  const uint32_t hi = b <= 32u ? 0u : meta1 & ((1u << (b - 32u)) - 1u);
  return __builtin_amdgcn_sad_u8(lo, 0u, __builtin_amdgcn_sad_u8(hi, 0u, tbase));
}

// Intersect the triangles of the leaf children in `lm` of one node (first(c)
// gives leaf child c's first triangle record, meta0/meta1 the leaf counts),
// one loop over all of them so the wave iterates as often as its busiest
// lane has triangles.  Returns true for an any-hit ray that found its hit.
template <bool kAnyHit, bool kCount, bool kInst, class First>
__device__ inline bool test_leaves(const DevScene& S, RayState& r, First first, uint32_t meta0, uint32_t meta1,
                                   uint32_t lm, TraverseStats* st) {
  const float eps = kAnyHit ? kEpsShadow : kEpsTrace;
  uint32_t tri = 0, tend = 0;
  while (true) {
    if (tri == tend) {
      if (lm == 0) break;
      const int c = __builtin_ctz(lm);
      lm &= lm - 1;
      tri = first((uint32_t)c);
      tend = tri + (((c < 4 ? meta0 : meta1) >> ((c & 3) * 8)) & 0xffu);
    }
    if (kCount) st->tris++;
    uint32_t id, bary = 0;
    const float t = intersect_tri(S.tris, tri, r.o, r.d, eps, &id, &bary);
    if (kInst) id += r.id_delta;
    // t == tbest < kMaxT: a best hit exists, so ray_hit_id reads a record
    if (t < kMaxT && (t < r.tbest || (t == r.tbest && id < ray_hit_id<kInst>(S, r)))) {
      r.tbest = t;
      if (kInst) r.best_id = id;
      r.best_leaf = kInst ? tri + r.leaf_delta : tri;
      r.best_bary = bary;
      if (kAnyHit) return true;
    }
    ++tri;
  }
  return false;
}

// Bytes of the slots in the 4-bit mask m4 set to 0xff, the others 0.
__device__ inline uint32_t slot_bytes(uint32_t m4) { return ((m4 * 0x00204081u) & 0x01010101u) * 0xffu; }

// Lanes below this one in the 64-bit mask m.
__device__ inline uint32_t lanes_below(unsigned long long m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Cooperative leaf phase of the traversal kernels, run by all 64 lanes of
// the wave together (idle lanes included).  The triangles of the leaf
// children each pending lane deferred are numbered lane by lane (in slot
// order within a lane) and dealt out to the wave 64 at a time: lane w of a
// round tests item base + w against its owner's ray.  Every candidate hit
// takes the minimum of its owner's key in LDS, (t bits, triangle id) as one
// 64-bit word — for t > 0 the order of the canonical (t, id) comparison —
// and the lane holding the minimum after the round writes its record and
// barycentrics beside it.  The tested set is exactly the one the lane-by-
// lane loop tests, and the minimum over it does not depend on how it is
// dealt, so the hit is the same bit for bit; lanes stay busy instead of
// waiting for the lane with the most triangles.  Any-hit rays take the
// least hit of the phase (the same whichever way it is dealt).  `pend`
// marks the lanes whose ray holds deferred leaves; they leave with none.
// With instancing an owner's ray, id and leaf deltas are those of the mesh
// it waits in (a lane with deferred leaves visits no node), and keys carry
// baked ids.  Returns true on a pending any-hit lane that found a hit.
template <bool kAnyHit, bool kCount, bool kInst, int kW>
__device__ inline bool coop_leaves(const DevScene& S, RayState& r, bool pend, const Stack& stk, TraverseStats* st) {
  const uint32_t lane = (uint32_t)stk.lane;
  const lds_u64 key = stk.key;
  const lds_u64 aux = stk.key + 64;
  const unsigned long long init =
      kAnyHit ? ~0ull : ((unsigned long long)__float_as_uint(r.tbest) << 32) | 0xffffffffull;
  // the pending lane's node: its leaf-count bytes and first triangle record
  uint32_t m0 = 0, m1 = 0, tb = 0, lm = 0;
  if (pend) {
    lm = r.pmask;
    r.pmask = 0;
    if (kW == 4) {
      const uint32_t* w4 = reinterpret_cast<const uint32_t*>(S.nodes4 + r.pnode);
      m0 = w4[4];   // bytes 16..19
      tb = w4[13];  // byte 52
    } else {
      const uint2* w8 = reinterpret_cast<const uint2*>(S.nodes + r.pnode);
      const uint2 M = w8[6];  // bytes 48..55
      m0 = M.x;
      m1 = M.y;
      tb = reinterpret_cast<const uint32_t*>(w8)[28];  // byte 112
    }
    key[lane] = init;
  }
  // the pending children's counts (other bytes zeroed) and their sum
  const uint32_t p0 = m0 & slot_bytes(lm & 0xfu), p1 = m1 & slot_bytes(lm >> 4);
  const uint32_t n = __builtin_amdgcn_sad_u8(p0, 0u, __builtin_amdgcn_sad_u8(p1, 0u, 0u));
  // E: items of the lanes below (exclusive prefix), T: items of the wave,
  // summed bit plane by bit plane of n
  uint32_t E = 0, T = 0;
  for (uint32_t b = 0; __ballot((n >> b) != 0u); ++b) {
    const unsigned long long bm = __ballot((n >> b) & 1u);
    E += lanes_below(bm) << b;
    T += (uint32_t)__popcll(bm) << b;
  }
  const float eps = kAnyHit ? kEpsShadow : kEpsTrace;
  for (uint32_t base = 0; base < T; base += 64u) {
    const uint32_t w = base + lane;
    // owner: the last lane whose first item is at most w (a lane without
    // items shares its E with the next lane, so the last one has items)
    int o = 0;
#pragma unroll
    for (int step = 32; step > 0; step >>= 1) {
      const uint32_t e = (uint32_t)__shfl((int)E, o + step);
      o = e <= w ? o + step : o;
    }
    const uint32_t q0 = w - (uint32_t)__shfl((int)E, o);
    const uint32_t op0 = (uint32_t)__shfl((int)p0, o), op1 = (uint32_t)__shfl((int)p1, o);
    // the owner's pending child holding item q0: halve the slot range by
    // pending counts (the four low slots, then two, then one)
    uint32_t q = q0, c = 0, word = op0;
    const uint32_t h4 = __builtin_amdgcn_sad_u8(op0, 0u, 0u);
    if (q >= h4) {
      q -= h4;
      c = 4;
      word = op1;
    }
    const uint32_t h2 = __builtin_amdgcn_sad_u8(word & 0xffffu, 0u, 0u);
    if (q >= h2) {
      q -= h2;
      c += 2;
      word >>= 16;
    }
    if (q >= (word & 0xffu)) {
      q -= word & 0xffu;
      c += 1;
    }
    const uint32_t om0 = (uint32_t)__shfl((int)m0, o), om1 = (uint32_t)__shfl((int)m1, o);
    const uint32_t rec = leaf_first((uint32_t)__shfl((int)tb, o), om0, om1, c) + q;
    const v3 ro = pkm::mk(__shfl(r.o.x, o), __shfl(r.o.y, o), __shfl(r.o.z, o));
    const v3 rd = pkm::mk(__shfl(r.d.x, o), __shfl(r.d.y, o), __shfl(r.d.z, o));
    uint32_t id_delta = 0, leaf_delta = 0;
    if (kInst) {
      id_delta = (uint32_t)__shfl((int)r.id_delta, o);
      leaf_delta = (uint32_t)__shfl((int)r.leaf_delta, o);
    }
    if (w < T) {
      if (kCount) st->tris++;
      uint32_t id, bary = 0;
      const float t = intersect_tri(S.tris, rec, ro, rd, eps, &id, &bary);
      if (t < kMaxT) {
        const unsigned long long mine =
            ((unsigned long long)__float_as_uint(t) << 32) | (unsigned long long)(id + id_delta);
        __hip_atomic_fetch_min(key + o, mine, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        // (a wave's LDS operations complete in order: every lane's minimum
        // of this round is in)
        if (key[o] == mine) aux[o] = ((unsigned long long)bary << 32) | (rec + leaf_delta);
      }
    }
  }
  if (!pend) return false;
  const unsigned long long k = key[lane];
  if (k == init) return false;
  const float t = __uint_as_float((uint32_t)(k >> 32));
  // a hit at the best distance replaces the best only with a lower id
  if (!kAnyHit && t == r.tbest && !ray_missed(r) && (uint32_t)k >= ray_hit_id<kInst>(S, r)) return false;
  const unsigned long long a = aux[lane];
  if (kInst) r.best_id = (uint32_t)k;
  r.tbest = t;
  r.best_leaf = (uint32_t)a;
  r.best_bary = (uint32_t)(a >> 32);
  return kAnyHit;
}

template <bool kAnyHit, bool kCount, bool kInst, int kLds, int kW>
__device__ inline bool node_step(const DevScene& S, RayState& r, const Stack& stk, TraverseStats* st);

// One traversal step, wave-cooperative (every lane of the wave that still
// traces calls it together): either a leaf phase, in which the lanes holding
// deferred leaf children intersect their triangles, or a node step, in
// which the other lanes visit their current node (box-test its 8 children,
// defer the hit leaf children, continue with the nearest hit internal child
// and push the others far-to-near) or pop the next live stack entry.  The
// canonical closest hit does not depend on the visiting order, so deferring
// leaves changes only how much work the wave does, not the result.
// With kInst, visiting an instance child maps the ray into its mesh space,
// pushes a return marker (entry distance 0, so it is never culled) and goes
// on at the mesh's root; popping the marker maps the ray back.  Distances
// are the same parameter in both spaces (the map is affine and the direction
// is not renormalized), so tbest and the culling bound carry over.
// kW is the node width the single-level traversal walks: 8 (WNode) or 4
// (WNode4, the derived layout); the instanced traversal walks 8-wide nodes.
// Returns true when the ray is finished.
template <bool kAnyHit, bool kCount, bool kInst = false, int kLds = kStackLds, int kW = 8>
__device__ inline bool ray_step(const DevScene& S, RayState& r, const Stack& stk, TraverseStats* st) {
  static_assert(kW == 8 || (kW == 4 && !kInst), "4-wide nodes are single-level only");
  // Leaf phase.  A node visit defers its hit leaf children; the lane then
  // waits while the others visit nodes, until at least S.leaf16 / 16 of the
  // wave's tracing lanes hold deferred leaves (or no lane has a node left to
  // visit), and those lanes intersect them together: on incoherent rays the
  // triangle loop would otherwise run with a few lanes of the wave active.
  const bool pend = r.pmask != 0u;
  const unsigned long long pm = __ballot(pend);
  if (pm) {
    const unsigned long long inner = __ballot(!pend && (r.have || r.sp > 0));
    if (inner == 0ull || 16u * (uint32_t)__popcll(pm) >= S.leaf16 * (uint32_t)__popcll(__ballot(true))) {
      if (pend) {
        const uint4* lp = reinterpret_cast<const uint4*>(S.nodes + r.pnode);
        const uint32_t lm = r.pmask;
        r.pmask = 0;
        if (kW == 4) {
          const uint32_t* w4 = reinterpret_cast<const uint32_t*>(S.nodes4 + r.pnode);
          const uint32_t meta = w4[4], tbase = w4[13];  // bytes 16..19 and 52
          if (test_leaves<kAnyHit, kCount, kInst>(
                  S, r, [&](uint32_t c) { return leaf_first(tbase, meta, 0u, c); }, meta, 0u, lm, st))
            return true;
        } else {
          const uint2 M = reinterpret_cast<const uint2*>(lp)[6];          // meta, bytes 48..55
          const uint32_t tbase = reinterpret_cast<const uint32_t*>(lp)[28];  // byte 112
          if (test_leaves<kAnyHit, kCount, kInst>(
                  S, r, [&](uint32_t c) { return leaf_first(tbase, M.x, M.y, c); }, M.x, M.y, lm, st))
            return true;
        }
      }
      return !r.have && r.sp == 0 && r.pmask == 0u;
    }
  }
  if (pend) return false;  // waits for the leaf phase
  return node_step<kAnyHit, kCount, kInst, kLds, kW>(S, r, stk, st);
}

// The node step of a traversal step: pop the next live stack entry unless a
// node is at hand, then visit it (box-test its children, defer the hit leaf
// children, continue with the nearest hit internal child and push the others
// far-to-near).  The ray holds no deferred leaves.  Returns true when the
// ray is finished.
template <bool kAnyHit, bool kCount, bool kInst, int kLds, int kW>
__device__ inline bool node_step(const DevScene& S, RayState& r, const Stack& stk, TraverseStats* st) {
  const int lane = stk.lane;
  if (!r.have) {
    const float cull = ray_cull<kAnyHit>(r);
    while (r.sp > 0) {
      --r.sp;
      uint32_t n;
      float tn;
      if (r.sp < kLds) {
        n = stk.lds_node[r.sp * 64 + lane];
        tn = stk.lds_t[r.sp * 64 + lane];
      } else {
        const uint2 e = stk.spill[(r.sp - kLds) * 64 + lane];
        n = e.x;
        tn = __uint_as_float(e.y);
      }
      if (tn <= cull) {
        r.node = n;
        r.have = true;
        break;
      }
    }
    if (!r.have) return true;
  }
  if (kInst) {
    if (r.node == kReturnMarker) {
      r.o = r.wo;
      r.d = r.wd;
      ray_set_inv(r);
      r.id_delta = 0;
      r.leaf_delta = 0;
      r.have = false;
      return r.sp == 0;
    }
    if (r.node & kInstChild) {
      if (r.sp >= stk.limit) {  // no room for the way back: skip it, flagged
        r.overflow = true;
        r.have = false;
        return r.sp == 0;
      }
      stack_push<kLds>(stk, r, kReturnMarker, 0.0f);
      const float4* ip = reinterpret_cast<const float4*>(S.inst + (r.node & ~kInstChild));
      const float4 m0 = ip[0], m1 = ip[1], m2 = ip[2];
      const uint4 ref = reinterpret_cast<const uint4*>(ip)[3];
      const float m[12] = {m0.x, m0.y, m0.z, m0.w, m1.x, m1.y, m1.z, m1.w, m2.x, m2.y, m2.z, m2.w};
      r.wo = r.o;
      r.wd = r.d;
      r.o = xform_point(m, r.o);
      r.d = xform_dir(m, r.d);
      ray_set_inv(r);
      r.node = ref.x;
      r.id_delta = ref.y;
      r.leaf_delta = ref.z;
    }
  }
  r.have = false;
  if (kCount) st->nodes++;
  const uint32_t visited = r.node;
  // header (origin, exponents, child count) and, per axis and ray direction
  // sign, the near and far plane bytes (kW / 4 words each), leaf mask and ibase
  constexpr int kWords = kW / 4;
  const uint4* np = kW == 8 ? reinterpret_cast<const uint4*>(S.nodes + visited)
                            : reinterpret_cast<const uint4*>(S.nodes4 + visited);
  const uint4 L0 = np[0];
  uint32_t qlx[kWords], qhx[kWords], qly[kWords], qhy[kWords], qlz[kWords], qhz[kWords];
  uint32_t leafmask, ibase, instmask = 0;
  if constexpr (kW == 8) {
    const uint4 L3 = np[3], L4 = np[4], L5 = np[5], L6 = np[6];
    qlx[0] = L3.z, qlx[1] = L3.w, qhx[0] = L4.x, qhx[1] = L4.y;
    qly[0] = L4.z, qly[1] = L4.w, qhy[0] = L5.x, qhy[1] = L5.y;
    qlz[0] = L5.z, qlz[1] = L5.w, qhz[0] = L6.x, qhz[1] = L6.y;
    leafmask = L6.z & 0xffu;
    if constexpr (kInst) instmask = (L6.z >> 8) & 0xffu;
    ibase = L6.w;
  } else {
    const uint4 L1 = np[1], L2 = np[2];
    qlx[0] = L1.y, qhx[0] = L1.z, qly[0] = L1.w, qhy[0] = L2.x, qlz[0] = L2.y, qhz[0] = L2.z;
    leafmask = L2.w & 0xffu;
    ibase = reinterpret_cast<const uint32_t*>(np)[12];
  }
  const float ox = __uint_as_float(L0.x), oy = __uint_as_float(L0.y), oz = __uint_as_float(L0.z);
  const uint32_t ew = L0.w;
  const int nchild = (int)(ew >> 24);
  const float sx = __uint_as_float((ew & 0xffu) << 23);
  const float sy = __uint_as_float(((ew >> 8) & 0xffu) << 23);
  const float sz = __uint_as_float(((ew >> 16) & 0xffu) << 23);
  const float ax = sx * r.invx, ay = sy * r.invy, az = sz * r.invz;
  const float bx = (ox - r.o.x) * r.invx, by = (oy - r.o.y) * r.invy, bz = (oz - r.o.z) * r.invz;
  const bool negx = r.invx < 0.0f, negy = r.invy < 0.0f, negz = r.invz < 0.0f;
  // near/far plane bytes per axis by ray direction sign
  uint32_t nxw[kWords], fxw[kWords], nyw[kWords], fyw[kWords], nzw[kWords], fzw[kWords];
#pragma unroll
  for (int i = 0; i < kWords; ++i) {
    nxw[i] = negx ? qhx[i] : qlx[i], fxw[i] = negx ? qlx[i] : qhx[i];
    nyw[i] = negy ? qhy[i] : qly[i], fyw[i] = negy ? qly[i] : qhy[i];
    nzw[i] = negz ? qhz[i] : qlz[i], fzw[i] = negz ? qlz[i] : qhz[i];
  }
  // child node of an internal slot: ibase + the internal slots below it; a
  // placement child of an instanced top level (outside the runs) reads its
  // child word
  const uint32_t inner = ~leafmask & ~instmask & ((1u << nchild) - 1u) & 0xffu;
  auto node_of = [&](uint32_t slot) {
    if (kInst && ((instmask >> slot) & 1u))
      return reinterpret_cast<const uint32_t*>(S.nodes + visited)[4u + slot];
    return ibase + (uint32_t)__popc(inner & ((1u << slot) - 1u));
  };
  const float cull = ray_cull<kAnyHit>(r);
  float tn[kW];
  uint32_t hitmask = 0;
#pragma unroll
  for (int c = 0; c < kW; ++c) {
    const int sh = (c & 3) * 8;
    const uint32_t wnx = nxw[c >> 2], wfx = fxw[c >> 2];
    const uint32_t wny = nyw[c >> 2], wfy = fyw[c >> 2];
    const uint32_t wnz = nzw[c >> 2], wfz = fzw[c >> 2];
    const float tnx = fmaf((float)((wnx >> sh) & 0xffu), ax, bx);
    const float tny = fmaf((float)((wny >> sh) & 0xffu), ay, by);
    const float tnz = fmaf((float)((wnz >> sh) & 0xffu), az, bz);
    const float tfx = fmaf((float)((wfx >> sh) & 0xffu), ax, bx);
    const float tfy = fmaf((float)((wfy >> sh) & 0xffu), ay, by);
    const float tfz = fmaf((float)((wfz >> sh) & 0xffu), az, bz);
    const float tnear = fmaxf(fmaxf(tnx, tny), fmaxf(tnz, 0.0f));
    const float tfar = fminf(fminf(tfx, tfy), tfz);
    tn[c] = tnear;
    const bool h = tnear <= tfar && tnear <= cull;
    hitmask |= h ? (1u << c) : 0u;
  }
  hitmask &= (1u << nchild) - 1u;  // the slots past the node's children
  // hit leaf children: intersected in the next leaf phase
  const uint32_t lmask = hitmask & leafmask;
  if (lmask) {
    r.pnode = visited;
    r.pmask = lmask;
  }
  if (kAnyHit) {
    // occlusion asks for any hit, not the nearest (the bound never
    // tightens): the hit internal children in slot order, unsorted, the
    // lowest visited next and the rest pushed so that lower slots pop first
    uint32_t im = hitmask & ~leafmask;
    if (im == 0) return r.sp == 0 && r.pmask == 0u;
    r.node = node_of((uint32_t)__builtin_ctz(im));
    r.have = true;
    im &= im - 1u;
    while (im) {
      const int c = 31 - __builtin_clz(im);
      im &= ~(1u << c);
      stack_push<kLds>(stk, r, node_of((uint32_t)c), 0.0f);
    }
    return false;
  }
  // internal children, sorted near-to-far: key = entry distance bits (>= 0,
  // so order-preserving) with the slot in the low 3 bits
  // (a hit child's entry distance is within the culling bound already)
  const uint32_t imask = hitmask & ~leafmask;
  if (imask == 0u) return r.sp == 0 && r.pmask == 0u;
  const int m = __popc(imask);
  uint32_t k[kW];
#pragma unroll
  for (int c = 0; c < kW; ++c)
    k[c] = (imask >> c) & 1u ? ((__float_as_uint(tn[c]) & ~7u) | (uint32_t)c) : 0xffffffffu;
  if constexpr (kW == 8) {
    cswap(k[0], k[2]); cswap(k[1], k[3]); cswap(k[4], k[6]); cswap(k[5], k[7]);
    cswap(k[0], k[4]); cswap(k[1], k[5]); cswap(k[2], k[6]); cswap(k[3], k[7]);
    cswap(k[0], k[1]); cswap(k[2], k[3]); cswap(k[4], k[5]); cswap(k[6], k[7]);
    cswap(k[2], k[4]); cswap(k[3], k[5]);
    cswap(k[1], k[4]); cswap(k[3], k[6]);
    cswap(k[1], k[2]); cswap(k[3], k[4]); cswap(k[5], k[6]);
  } else {
    cswap(k[0], k[1]); cswap(k[2], k[3]);
    cswap(k[0], k[2]); cswap(k[1], k[3]);
    cswap(k[1], k[2]);
  }
  r.node = node_of(k[0] & 7u);
  r.have = true;
#pragma unroll
  for (int i = kW - 1; i >= 1; --i) {
    if (i < m) {
      const uint32_t nx = node_of(k[i] & 7u);
      const float t = __uint_as_float(k[i] & ~7u);  // rounded down: conservative
      stack_push<kLds>(stk, r, nx, t);
    }
  }
  return false;
}

// ---------------------------------------------------------------------------
// Shading helpers (shade_hit.wgsl / shade_miss.wgsl semantics)
// ---------------------------------------------------------------------------
__device__ inline int wrapi(int i, int n) {
  int r = i % n;
  return r < 0 ? r + n : r;
}

// shade_miss.wgsl envColor (:37-43): nearest sampler, repeat addressing.
// Returns the texel index through *texel (for gradients).
__device__ inline uint32_t env_texel(const DevScene& S, float theta, v3 dir) {
  float u = theta + pkm::atan2f_(dir.z, dir.x) / pkm::kTau;
  float v = pkm::acosf_(dir.y) * pkm::kInvPi;
  int ix = wrapi((int)floorf(u * (float)S.ew), (int)S.ew);
  int iy = wrapi((int)floorf(v * (float)S.eh), (int)S.eh);
  return (uint32_t)iy * S.ew + (uint32_t)ix;
}
__device__ inline v3 env_color(const DevScene& S, float theta, v3 dir) {
  if (S.env_mode == 0) return pkm::mk(0.0f, 0.0f, 0.0f);
  float4 L = S.env_rad[env_texel(S, theta, dir)];
  return pkm::mk(L.x, L.y, L.z);
}

// envPdf (shade_hit.wgsl:202-210); texel index clamped into the map
__device__ inline float env_pdf(const DevScene& S, float theta, v3 dir) {
  float a = 1.0f + theta + pkm::atan2f_(dir.z, dir.x) / pkm::kTau;
  float u = a - truncf(a);
  float v = pkm::acosf_(dir.y) * pkm::kInvPi;
  int ix = (int)(u * (float)S.ew), iy = (int)(v * (float)S.eh);
  ix = min(max(ix, 0), (int)S.ew - 1);
  iy = min(max(iy, 0), (int)S.eh - 1);
  float sin_phi = pkm::sinf_(v * pkm::kPi);
  return S.env_pdf[(size_t)iy * S.ew + ix] / sin_phi;
}

// The four channels of one texel with one load: an RGBA8 texel is one
// little-endian word (channel c in byte c), an RGBA32F texel one float4.
__device__ inline float4 texel4(const DevScene& S, const void* tex, int layer, int x, int y) {
  const size_t i = ((size_t)layer * S.th + (size_t)y) * S.tw + (size_t)x;
  if (S.tfmt == 0) {
    const uint32_t w = ((const uint32_t*)tex)[i];
    return make_float4(pkm::unpack_unorm8(w & 0xffu), pkm::unpack_unorm8((w >> 8) & 0xffu),
                       pkm::unpack_unorm8((w >> 16) & 0xffu), pkm::unpack_unorm8(w >> 24));
  }
  return ((const float4*)tex)[i];
}

// Bilinear footprint of one texture lookup (the same index arithmetic as
// sample_tex): texel indices within a layer, corners (x0,y0) (x1,y0)
// (x0,y1) (x1,y1), and the interpolation weights of those corners.
struct Footprint {
  uint32_t idx[4];
  float w[4];
};
__device__ inline Footprint tex_footprint(const DevScene& S, float u, float v) {
  const float x = u * (float)S.tw - 0.5f, y = v * (float)S.th - 0.5f;
  const float fx0 = floorf(x), fy0 = floorf(y);
  const float fx = x - fx0, fy = y - fy0;
  const int x0 = wrapi((int)fx0, (int)S.tw), x1 = wrapi((int)fx0 + 1, (int)S.tw);
  const int y0 = wrapi((int)fy0, (int)S.th), y1 = wrapi((int)fy0 + 1, (int)S.th);
  Footprint fp;
  fp.idx[0] = (uint32_t)(y0 * (int)S.tw + x0);
  fp.idx[1] = (uint32_t)(y0 * (int)S.tw + x1);
  fp.idx[2] = (uint32_t)(y1 * (int)S.tw + x0);
  fp.idx[3] = (uint32_t)(y1 * (int)S.tw + x1);
  fp.w[0] = (1.0f - fx) * (1.0f - fy);
  fp.w[1] = fx * (1.0f - fy);
  fp.w[2] = (1.0f - fx) * fy;
  fp.w[3] = fx * fy;
  return fp;
}

// textureSampleLevel, linear filter, repeat, level 0, from atlas `tex`
__device__ inline float4 sample_tex(const DevScene& S, const void* tex, int layer, float u, float v) {
  // 1x1 RGBA8 layers (constant-color materials): the four taps are one
  // texel, and for its unorm8 values (finite, >= +0) the blend below
  // returns that texel bit for bit (a + (a - a) * f = a), so it is read once.
  if (S.tw == 1 && S.th == 1 && S.tfmt == 0) return texel4(S, tex, layer, 0, 0);
  float x = u * (float)S.tw - 0.5f, y = v * (float)S.th - 0.5f;
  float fx0 = floorf(x), fy0 = floorf(y);
  float fx = x - fx0, fy = y - fy0;
  int x0 = wrapi((int)fx0, (int)S.tw), x1 = wrapi((int)fx0 + 1, (int)S.tw);
  int y0 = wrapi((int)fy0, (int)S.th), y1 = wrapi((int)fy0 + 1, (int)S.th);
  const float4 A = texel4(S, tex, layer, x0, y0), B = texel4(S, tex, layer, x1, y0);
  const float4 C = texel4(S, tex, layer, x0, y1), D = texel4(S, tex, layer, x1, y1);
  const float ta[4] = {A.x, A.y, A.z, A.w}, tb[4] = {B.x, B.y, B.z, B.w};
  const float tc[4] = {C.x, C.y, C.z, C.w}, td[4] = {D.x, D.y, D.z, D.w};
  float r[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const float a = ta[c], b = tb[c], cc = tc[c], dd = td[c];
    float top = a + (b - a) * fx;
    float bot = cc + (dd - cc) * fx;
    r[c] = top + (bot - top) * fy;
  }
  return make_float4(r[0], r[1], r[2], r[3]);
}

}  // namespace pk
