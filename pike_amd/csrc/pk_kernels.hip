// pk_kernels.hip — the wavefront path-tracing kernels for gfx950.
//
// One frame (render/pike.js:120-162) is
//   raygen -> { trace_closest -> shade -> trace_shadow } x depth -> accumulate
// with path state in SoA queues in HBM.  Differences from the reference's
// WGSL dispatch structure, all MI355X-driven:
//   * ray, hit and shadow queues are SoA float4/uint4 arrays written through
//     wave-ballot compaction (one atomic per wave, coalesced stores);
//   * miss shading is fused into the two trace kernels (no miss queue), and
//     each pixel's color is written only by the thread that owns its path in
//     that stage, so accumulation is race-free and deterministic
//     (shade_miss.wgsl:57 is a racy +=);
//   * the trace kernels are persistent: waves pull 64-ray batches from a
//     device counter until the queue drains, so launch sizes never depend on
//     the data-dependent ray counts;
//   * counters are reset by the kernel that runs before their consumer, so a
//     whole frame is kernels only and can be captured in one hipGraph.
#include "pk_device.h"
#include <type_traits>

#include "pk_kernels.h"

namespace pk {

__device__ inline uint32_t wave_sum(uint32_t v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// Wave-level compaction: returns this lane's slot in a queue whose counter is
// `counter` (valid only where pred is true).
__device__ inline uint32_t wave_append(uint32_t* counter, bool pred) {
  const unsigned long long m = __ballot(pred);
  const uint32_t total = (uint32_t)__popcll(m);
  const int lane = (int)__lane_id();
  uint32_t base = 0;
  if (total) {
    const int leader = __ffsll((long long)m) - 1;
    if (lane == leader) base = atomicAdd(counter, total);
    base = __shfl(base, leader);
  }
  const unsigned long long below = m & ((1ull << lane) - 1ull);
  return base + (uint32_t)__popcll(below);
}

// Two appends aggregated over a block of kWaves waves: per-wave counts meet
// in LDS and each queue takes one global atomic per block.  Appends to one
// counter serialize at the L2, so per-wave atomics from every wave of a
// 16M-thread launch made k_shade atomic-bound; a block pays a quarter of
// them.  Every thread of the block must call it (it synchronizes the block).
template <int kWaves>
__device__ inline void block_append2(uint32_t* c0, bool p0, uint32_t* c1, bool p1, uint32_t* slot0,
                                     uint32_t* slot1) {
  __shared__ uint32_t wcnt[2][kWaves];
  __shared__ uint32_t base[2];
  const int lane = (int)__lane_id(), wave = (int)(threadIdx.x >> 6);
  const unsigned long long m0 = __ballot(p0), m1 = __ballot(p1);
  if (lane == 0) {
    wcnt[0][wave] = (uint32_t)__popcll(m0);
    wcnt[1][wave] = (uint32_t)__popcll(m1);
  }
  __syncthreads();
  if (threadIdx.x == 0 || threadIdx.x == 64) {  // the two queues' atomics from two waves
    const int q = threadIdx.x == 0 ? 0 : 1;
    uint32_t t = 0;
    for (int w = 0; w < kWaves; ++w) t += wcnt[q][w];
    base[q] = t ? atomicAdd(q == 0 ? c0 : c1, t) : 0u;
  }
  __syncthreads();
  uint32_t o0 = base[0], o1 = base[1];
  for (int w = 0; w < wave; ++w) {
    o0 += wcnt[0][w];
    o1 += wcnt[1][w];
  }
  const unsigned long long below = (1ull << lane) - 1ull;
  *slot0 = o0 + (uint32_t)__popcll(m0 & below);
  *slot1 = o1 + (uint32_t)__popcll(m1 & below);
}

// ---------------------------------------------------------------------------
// camera.wgsl main (:75-92) / createPrimaryRay (:58-73)
// ---------------------------------------------------------------------------
// The image plane's axes scaled by tan(fov / 2): uniform over a frame.
__device__ inline void camera_basis(const pk_camera& cam, v3* bx, v3* by) {
  const v3 up = pkm::mk(0.0f, 1.0f, 0.0f);
  const v3 dir = pkm::mk(cam.dir[0], cam.dir[1], cam.dir[2]);
  const float tf = pkm::tanf_(cam.fov * 0.5f);
  *bx = pkm::scale(pkm::normalize(pkm::cross(dir, up)), tf);
  *by = pkm::scale(pkm::normalize(pkm::cross(*bx, dir)), tf);
}

// createPrimaryRay for pixel (px, py) and `sample`, given the camera's
// basis (camera_basis).
__device__ inline void camera_ray_b(uint32_t px, uint32_t py, uint32_t W, uint32_t H, uint32_t sample,
                                    const pk_camera& cam, v3 bx, v3 by, v3* o_out, v3* d_out) {
  uint32_t seed = (px * 1973u + py * 9277u + sample * 26699u) | 1u;
  seed = pkm::pcg_next(&seed);
  const float k = cam.distortion;
  const float resx = (float)W, resy = (float)H;
  const float jx = pkm::pcg_rand(&seed);
  const float jy = pkm::pcg_rand(&seed);
  float ux = (2.0f * (((float)px + jx) / resx) - 1.0f) * (resx / resy);
  float uy = (2.0f * (((float)py + jy) / resy) - 1.0f) * -1.0f;
  const float rd = pkm::sqrt_(ux * ux + uy * uy);
  const float ru = rd * (1.0f + k * rd * rd);
  const float l2 = pkm::sqrt_(ux * ux + uy * uy);
  ux = ru * (ux / l2);
  uy = ru * (uy / l2);
  const v3 dir = pkm::mk(cam.dir[0], cam.dir[1], cam.dir[2]);
  const v3 org = pkm::mk(cam.origin[0], cam.origin[1], cam.origin[2]);
  const float theta = pkm::pcg_rand(&seed) * pkm::kTau;
  const float pw = pkm::powf_(pkm::pcg_rand(&seed), cam.bokeh);
  float st, ct;
  pkm::sincosf_(theta, &st, &ct);
  const v3 dof = pkm::scale(pkm::scale(pkm::add(pkm::scale(bx, ct), pkm::scale(by, st)), cam.aperture), pw);
  const v3 screen = pkm::add(pkm::add(pkm::add(pkm::scale(bx, ux), pkm::scale(by, uy)), dir), org);
  *d_out = pkm::normalize(pkm::sub(pkm::add(screen, pkm::scale(dof, cam.focal_depth)), pkm::add(org, dof)));
  *o_out = pkm::add(dof, org);
}

__device__ inline void camera_ray(uint32_t px, uint32_t py, uint32_t W, uint32_t H, uint32_t sample,
                                  const pk_camera& cam, v3* o_out, v3* d_out) {
  v3 bx, by;
  camera_basis(cam, &bx, &by);
  camera_ray_b(px, py, W, H, sample, cam, bx, by, o_out, d_out);
}

// One thread per path of the pass: path idx = pixel * ns + s traces sample
// P.sample + s of that pixel (ns = the pass's samples, fewer than S at the
// end of a run).
__global__ void __launch_bounds__(256) k_raygen(FrameDev F) {
  const uint32_t W = F.W, H = F.H, B = W * H;
  const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  const FrameParams P = *F.params;
  const uint32_t ns = pass_samples(F, P);
  const uint32_t n = B * ns;
  if (idx == 0) {
    F.cnt[kCntRays0] = n;
    F.cnt[kCntHits] = 0;
    F.cnt[kCntWorkTrace] = 0;
  }
  if (idx >= n) return;
  // paths are pixel-major (the samples of a pixel are neighbours in the
  // queues and in the color buffer)
  const uint32_t pixel = idx / ns, s = idx - pixel * ns;
  v3 o, d;
  camera_ray_b(pixel % W, pixel / W, W, H, P.sample + s, P.cam, pkm::mk(P.cbx[0], P.cbx[1], P.cbx[2]),
               pkm::mk(P.cby[0], P.cby[1], P.cby[2]), &o, &d);
  F.qo[0][idx] = make_float4(o.x, o.y, o.z, __uint_as_float(idx));
  F.qd[0][idx] = make_float4(d.x, d.y, d.z, 0.0f);
  // (camera rays carry throughput 1: bounce 0 does not read F.qt)
  F.color[(size_t)pixel * F.S + s] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// Batch ray generation: one thread per (global pixel p, sample s) of the
// round, path id p * S + s; samples at or beyond their view's spp emit
// nothing, the rest are compacted into ray queue 0 (d.w carries the view).
// The queue counter is zeroed by k_batch_round_begin before the round.
__global__ void __launch_bounds__(256) k_raygen_batch(FrameDev F) {
  const uint32_t B = F.W, S = F.S;
  const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  const FrameParams P = *F.params;
  bool emit = false;
  uint32_t view = 0;
  v3 o = pkm::mk(0.0f, 0.0f, 0.0f), d = o;
  if (idx < B * S) {
    const uint32_t p = idx / S, sm = idx - p * S;
    view = view_of(F, p);
    const ViewDev& V = F.views[view];
    // this path's color slot (p * S + sm; slots of samples beyond a view's
    // spp are never folded)
    F.color[idx] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (P.sample + sm < V.spp) {
      const uint32_t lp = p - V.offset;
      camera_ray(lp % V.W, lp / V.W, V.W, V.H, P.sample + sm, V.cam, &o, &d);
      emit = true;
    }
  }
  const uint32_t slot = wave_append(&F.cnt[kCntRays0], emit);
  if (emit) {
    F.qo[0][slot] = make_float4(o.x, o.y, o.z, __uint_as_float(idx));
    F.qd[0][slot] = make_float4(d.x, d.y, d.z, __uint_as_float(view));
  }
}

// Zeroes the queue counters a batch round's raygen and first trace use.
__global__ void k_batch_round_begin(FrameDev F) {
  F.cnt[kCntRays0] = 0;
  F.cnt[kCntHits] = 0;
  F.cnt[kCntWorkTrace] = 0;
}

// Color slot of path pid: passes number paths pixel-major (pid = pixel * ns
// + s; batch pixels are global, ns = the round's S) and keep each pixel's S
// color slots together (pixel * S + s), the layout k_accumulate folds: the
// paths a wave finishes together are neighbours, so their read-modify-write
// of the color slots touch a few cache lines instead of one line (and page)
// each.
__device__ inline uint32_t color_slot(const FrameDev& F, uint32_t ns, uint32_t pid) {
  if (ns == F.S) return pid;
  const uint32_t pixel = pid / ns;
  return pixel * F.S + (pid - pixel * ns);
}

__device__ inline void add_color(const FrameDev& F, uint32_t ns, uint32_t pid, v3 L) {
  const uint32_t c_at = color_slot(F, ns, pid);
  float4 c = F.color[c_at];
  c.x += L.x;
  c.y += L.y;
  c.z += L.z;
  c.w += 1.0f;
  F.color[c_at] = c;
}

// ---------------------------------------------------------------------------
// trace.wgsl main (:191-213) + shade_miss.wgsl (:45-57) for the misses
// ---------------------------------------------------------------------------
// Persistent traversal loop shared by the three trace kernels.  Each lane
// owns one ray at a time; a lane that finishes stays idle until at least
// `refill_min` lanes of the wave are idle, then the wave compacts the
// finished results (flush) and refills the idle lanes from the slots it has
// claimed, claiming kClaimChunk (2 kClaimChunk in the largest closest-hit
// launches) more with one atomic when those run out.  This keeps lanes busy
// on incoherent rays while the queue cursor, whose atomics serialize at the
// L2, sees one atomic per 256 or 512 rays; each wave also
// works through contiguous (neighbouring-pixel) slots.  Every wave leaves the
// loop once the queue is exhausted and its own lanes are done.
constexpr uint32_t kClaimChunk = 256;  // slots per cursor atomic (>= 64)

// One wave-wide traversal step for the lanes whose `active` ray still
// traces: the leaf phase runs on the whole wave (coop_leaves), the node
// steps on the active lanes without deferred leaves, with ray_step's
// trigger.  Instanced any-hit rays keep ray_step's lane-by-lane leaf phase:
// their kernels have no registers to spare for the cooperative one
// (measured 8% slower on the instanced 1080p workload).  Every lane of the
// wave calls it together; returns true on an active lane whose ray is
// finished.
template <bool kAnyHit, bool kCount, bool kInst, int kLds, int kNW>
__device__ inline bool wave_step(const DevScene& S, RayState& r, bool active, const Stack& stk, TraverseStats* st) {
  if constexpr (kInst && kAnyHit) {
    return active && ray_step<kAnyHit, kCount, kInst, kLds, kNW>(S, r, stk, st);
  } else {
    const bool pend = active && r.pmask != 0u;
    const unsigned long long pm = __ballot(pend);
    bool leaf = false;
    if (pm) {
      const unsigned long long inner = __ballot(active && !pend && (r.have || r.sp > 0));
      leaf = inner == 0ull || 16u * (uint32_t)__popcll(pm) >= S.leaf16 * (uint32_t)__popcll(__ballot(active));
    }
    if (leaf) {
      const bool hit = coop_leaves<kAnyHit, kCount, kInst, kNW>(S, r, pend, stk, st);
      return active && (hit || (!r.have && r.sp == 0));
    }
    if (active && !pend) return node_step<kAnyHit, kCount, kInst, kLds, kNW>(S, r, stk, st);
    return false;
  }
}

template <bool kAnyHit, bool kCount, bool kInst, int kLds, int kNW, class Load, class Finish, class Flush>
__device__ inline void trace_loop(const DevScene& S, uint32_t n, uint32_t* cursor, uint32_t refill_min,
                                  const Stack& stk, TraverseStats* st, uint32_t* overflow, Load load,
                                  Finish finish, Flush flush) {
  const int lane = stk.lane;
  RayState r;
  r.best_leaf = 0xffffffffu;
  bool active = false, done = false, exhausted = false;
  uint32_t slot = 0;
  // slots this wave has claimed and not yet started (wave-uniform); claims
  // take `claim` slots per atomic, since atomics on one counter
  // serialize at the L2
  uint32_t pool_lo = 0, pool_hi = 0;
  bool claimed_all = false;
  // closest-hit launches of a quarter billion rays or more claim twice as
  // many slots per atomic: their tails are short next to their work
  const uint32_t claim = !kAnyHit && n >= (1u << 28) ? 2u * kClaimChunk : kClaimChunk;
  while (true) {
    if (wave_step<kAnyHit, kCount, kInst, kLds, kNW>(S, r, active, stk, st)) {
      active = false;
      done = true;
      if (r.overflow) *overflow = 1;
      finish(slot, r);
    }
    const unsigned long long idle = __ballot(!active);
    const uint32_t nidle = (uint32_t)__popcll(idle);
    if (nidle < refill_min && nidle < 64u) continue;
    flush(done, slot, r);
    done = false;
    if (exhausted) {
      if (nidle == 64u) break;
      continue;
    }
    const uint32_t have = pool_hi - pool_lo;
    uint32_t fresh = 0;
    bool got = false;
    if (have < nidle && !claimed_all) {
      uint32_t b = 0;
      if (lane == 0) b = atomicAdd(cursor, claim);
      fresh = (uint32_t)__builtin_amdgcn_readfirstlane((int)__shfl(b, 0));  // wave-uniform: a scalar
      got = true;
      claimed_all = fresh >= n || n - fresh <= claim;
    }
    if (!active) {
      // idle lanes take the pool's slots first, then the fresh chunk's
      const uint32_t rank = (uint32_t)__popcll(idle & ((1ull << lane) - 1ull));
      const uint32_t s2 = rank < have ? pool_lo + rank : (got ? fresh + (rank - have) : 0xffffffffu);
      if (s2 < n) {
        slot = s2;
        load(slot, r);
        active = true;
      }
    }
    if (have >= nidle) {
      pool_lo += nidle;
    } else if (got) {
      pool_lo = fresh + (nidle - have);
      pool_hi = fresh + claim;
    } else {
      pool_lo = pool_hi;
    }
    pool_hi = min(pool_hi, n);
    pool_lo = min(pool_lo, pool_hi);
    if (claimed_all && pool_lo >= pool_hi) {
      exhausted = true;
      if (__ballot(active) == 0ull) break;
    }
  }
}

// The traversal kernels run kWpb independent waves per block (the block
// size is a launch-shape choice of the autotuner), each with its own LDS
// stack part of kLds entries per lane, its own spill region and, with the
// cooperative leaf phase (kCoop), its 128 words of leaf-phase keys.
#define PK_TRACE_WAVE_SETUP(kCoop)                                                                                \
  __shared__ uint32_t lds_node[kWpb][kLds * 64];                                                                 \
  __shared__ float lds_t[kWpb][kLds * 64];                                                                       \
  __shared__ unsigned long long lds_key[kWpb][(kCoop) ? 128 : 1];                                                \
  const int lane = (int)(threadIdx.x & 63u);                                                                     \
  const int wave = kWpb == 1 ? 0 : (int)(threadIdx.x >> 6);                                                      \
  const size_t wave_id = (size_t)blockIdx.x * kWpb + (size_t)wave

template <bool kCount, bool kInst, int kLds, int kWpb, int kNW = 8>
__global__ void __launch_bounds__(64 * kWpb, trace_waves_per_simd(kLds)) k_trace_closest(DevScene S, FrameDev F, int bounce,
                                                                                uint32_t refill_min) {
  PK_TRACE_WAVE_SETUP(true);
  const int in = bounce & 1;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    F.cnt[ray_counter(in ^ 1)] = 0;
    F.cnt[kCntShadow] = 0;
  }
  const uint32_t n = F.cnt[ray_counter(in)];
  const float theta = F.params->env_theta;
  const Stack stk{(lds_u32)lds_node[wave], (lds_f32)lds_t[wave], F.spill + wave_id * kSpillStride * 64, lane, (int)S.stack_limit,
                  (lds_u64)lds_key[wave]};
  TraverseStats st = {0, 0};
  // traced/hits are wave-uniform (counted from ballots at each flush, which
  // every finished ray passes once), so they cost no per-lane registers
  uint32_t traced = 0, hits = 0, overflow = 0;
  const float4* qo = F.qo[in];
  const float4* qd = F.qd[in];
  const uint32_t ns = F.views ? F.S : pass_samples(F, *F.params);
  uint32_t res_lo = 0, res_hi = 0;  // this wave's unused run of hit slots (wave-uniform)
  trace_loop<false, kCount, kInst, kLds, kNW>(
      S, n, &F.cnt[kCntWorkTrace], refill_min, stk, &st, &overflow,
      [&](uint32_t slot, RayState& r) {
        const float4 o4 = qo[slot], d4 = qd[slot];
        ray_init<false>(r, pkm::mk(o4.x, o4.y, o4.z), pkm::mk(d4.x, d4.y, d4.z));
        if (F.views) r.node = F.views[__float_as_uint(d4.w)].root;
      },
      [&](uint32_t, const RayState&) {},
      [&](bool done, uint32_t slot, const RayState& r) {
        traced += (uint32_t)__popcll(__ballot(done));
        // misses are shaded here, by every lane that finished one since the
        // last refill together, not in the step each one finished in (where
        // the wave would run the environment lookup for a lane or two)
        if (done && ray_missed(r)) {  // shade_miss.wgsl envColor x throughput
          const uint32_t pid = __float_as_uint(qo[slot].w);  // reread: no register held through the traversal
          const float4 t4 = bounce == 0 ? make_float4(1.0f, 1.0f, 1.0f, 0.0f) : F.qt[in][slot];
          add_color(F, ns, pid, pkm::mul(env_color(S, theta, r.d), pkm::mk(t4.x, t4.y, t4.z)));
        }
        // hit records go to this wave's reserved run of the hit queue, in
        // the order the wave finishes them (its claimed chunk's pixel order)
        const bool is_hit = done && !ray_missed(r);
        const unsigned long long m = __ballot(is_hit);
        const uint32_t cnt = (uint32_t)__popcll(m);
        hits += cnt;
        if (cnt == 0) return;
        const uint32_t rank = (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
        const uint32_t rem = res_hi - res_lo;
        uint32_t at = res_lo + rank;
        if (rem < cnt) {  // the run ends: continue in a fresh one
          uint32_t b = 0;
          if (lane == 0) b = atomicAdd(&F.cnt[kCntHits], kHitRun);
          const uint32_t fresh = (uint32_t)__builtin_amdgcn_readfirstlane((int)__shfl(b, 0));
          if (rank >= rem) at = fresh + (rank - rem);
          res_lo = fresh + (cnt - rem);
          res_hi = fresh + kHitRun;
        } else {
          res_lo += cnt;
        }
        if (is_hit) F.hits[at] = make_uint4(slot, r.best_leaf, __float_as_uint(r.tbest), r.best_bary);
      });
  // the unused end of the wave's run: holes k_shade skips
  for (uint32_t i = res_lo + (uint32_t)lane; i < res_hi; i += 64u) F.hits[i] = make_uint4(0xffffffffu, 0u, 0u, 0u);
  if (kCount) {
    st.nodes = wave_sum(st.nodes);
    st.tris = wave_sum(st.tris);
  }
  overflow = wave_sum(overflow);
  if (lane == 0) {
    atomicAdd(&F.stats[kStatRays], (unsigned long long)traced);
    atomicAdd(&F.stats[kStatHits], (unsigned long long)hits);
    if (kCount) {
      atomicAdd(&F.stats[kStatNodes], (unsigned long long)st.nodes);
      atomicAdd(&F.stats[kStatTris], (unsigned long long)st.tris);
    }
    if (overflow) atomicOr(&F.cnt[kCntOverflow], 1u);
  }
}

// ---------------------------------------------------------------------------
// shade_hit.wgsl main (:407-483)
// ---------------------------------------------------------------------------
// Surface interaction.  Sampling decisions and pdfs read metallic, opacity
// and alpha from the sampling atlas; BSDF values read base_e, spec_e and
// alpha_e from the evaluation atlas (the same texels unless parameters were
// written), so derivatives flow only through evaluation (detached sampling).
struct SI {
  pkm::m3 basis;
  v3 wo, origin, normal, emission;
  v3 base_e, spec_e;
  float metallic, opacity, alpha;
  float alpha_e, rough_e;  // alpha_e = rough_e^2
};

__device__ inline float ggx_d(v3 m, float au, float av) {
  float auv = au * av;
  float tx = m.x / au, ty = m.y / av;
  float e = tx * tx + ty * ty + m.z * m.z;
  return 1.0f / (pkm::kPi * auv * e * e);
}
__device__ inline float ggx_g1(v3 w, v3 m, float au, float av) {
  float ax = w.x * au, ay = w.y * av;
  float axy2 = ax * ax + ay * ay;
  float tan_theta = axy2 / (w.z * w.z);
  float r = 2.0f / (1.0f + pkm::sqrt_(1.0f + tan_theta));
  return pkm::dot(w, m) * pkm::dot(w, pkm::mk(0.0f, 0.0f, 1.0f)) <= 0.0f ? 0.0f : r;
}
__device__ inline float lambert_pdf(v3 wi) { return pkm::max_(wi.z, kEpsShadow) * pkm::kInvPi; }
__device__ inline float specular_pdf(const SI& si, v3 m) {
  return pkm::max_(kEpsShadow, ggx_d(m, si.alpha, si.alpha) * ggx_g1(si.wo, m, si.alpha, si.alpha) / (4.0f * si.wo.z));
}
__device__ inline v3 eval_lambert(const SI& si, v3 wi, float pdf) {
  float k = pkm::max_(kEpsShadow, wi.z);
  return pkm::mk(si.base_e.x * pkm::kInvPi * k / pdf, si.base_e.y * pkm::kInvPi * k / pdf,
                 si.base_e.z * pkm::kInvPi * k / pdf);
}
__device__ inline v3 eval_specular(const SI& si, v3 wi, float pdf) {
  v3 h = pkm::normalize(pkm::add(si.wo, wi));
  float d = ggx_d(h, si.alpha_e, si.alpha_e);
  float g = ggx_g1(wi, h, si.alpha_e, si.alpha_e) * ggx_g1(si.wo, h, si.alpha_e, si.alpha_e);
  float k = pkm::max_(d * g / (4.0f * si.wo.z * pdf), 0.0f);
  return pkm::scale(si.spec_e, k);
}

// Derivatives with respect to (isotropic) alpha of the GGX terms:
//   d ln D / d alpha = 4 s / (alpha^3 e) - 2 / alpha,
//     s = m.x^2 + m.y^2, e = s / alpha^2 + m.z^2;
//   dG1 / d alpha = -G1^2 alpha q / (2 sqrt(1 + alpha^2 q)),
//     q = (w.x^2 + w.y^2) / w.z^2.
__device__ inline float ggx_d_dalpha(v3 m, float a, float d) {
  const float s = m.x * m.x + m.y * m.y;
  const float e = s / (a * a) + m.z * m.z;
  return d * (4.0f * s / (a * a * a * e) - 2.0f / a);
}
__device__ inline float ggx_g1_dalpha(v3 w, float a, float g1) {
  if (g1 == 0.0f) return 0.0f;
  const float q = (w.x * w.x + w.y * w.y) / (w.z * w.z);
  return -0.5f * g1 * g1 * a * q / pkm::sqrt_(1.0f + a * a * q);
}
// eval_specular = spec_e * k; returns k and dk / d alpha_e
__device__ inline void specular_k_grad(const SI& si, v3 wi, float pdf, float* k, float* dk) {
  const v3 h = pkm::normalize(pkm::add(si.wo, wi));
  const float a = si.alpha_e;
  const float d = ggx_d(h, a, a);
  const float gi = ggx_g1(wi, h, a, a), go = ggx_g1(si.wo, h, a, a);
  const float den = 4.0f * si.wo.z * pdf;
  const float kk = d * (gi * go) / den;
  *k = pkm::max_(kk, 0.0f);
  *dk = kk > 0.0f ? (ggx_d_dalpha(h, a, d) * (gi * go) +
                     d * (ggx_g1_dalpha(wi, a, gi) * go + gi * ggx_g1_dalpha(si.wo, a, go))) / den
                  : 0.0f;
}
__device__ inline float power_heuristic(float a, float b) {
  float a2 = a * a;
  return a2 / (a2 + b * b);
}
__device__ inline float schlick(float cos_theta, float ior) {
  float r0 = (1.0f - ior) / (1.0f + ior);
  r0 *= r0;
  float tmp = 1.0f - cos_theta;
  float tmp2 = tmp * tmp;
  return r0 + (1.0f - r0) * tmp2 * tmp2 * tmp;
}
__device__ inline pkm::m3 onb(v3 n) {
  float sign = n.z > 0.0f ? 1.0f : -1.0f;
  float a = -1.0f / (sign + n.z);
  float b = n.x * n.y * a;
  pkm::m3 m;
  m.c0 = pkm::mk(1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x);
  m.c1 = pkm::mk(b, sign + n.y * n.y * a, -n.y);
  m.c2 = n;
  return m;
}
__device__ inline v3 sample_ggx(const SI& si, uint32_t* seed) {
  float au = si.alpha, av = si.alpha;
  v3 vh = pkm::normalize(pkm::mk(au * si.wo.x, av * si.wo.y, si.wo.z));
  float lensq = vh.x * vh.x + vh.y * vh.y;
  v3 t1v = lensq > 0.0f ? pkm::scale(pkm::mk(-vh.y, vh.x, 0.0f), 1.0f / pkm::sqrt_(lensq)) : pkm::mk(1.0f, 0.0f, 0.0f);
  v3 t2v = pkm::cross(vh, t1v);
  float r = pkm::sqrt_(pkm::pcg_rand(seed));
  float phi = pkm::kTau * pkm::pcg_rand(seed);
  float sp, cp;
  pkm::sincosf_(phi, &sp, &cp);
  float t1 = r * cp;
  float t2 = r * sp;
  float sv = 0.5f * (1.0f + vh.z);
  t2 = (1.0f - sv) * pkm::sqrt_(1.0f - t1 * t1) + sv * t2;
  v3 nh = pkm::add(pkm::add(pkm::scale(t1v, t1), pkm::scale(t2v, t2)),
                   pkm::scale(vh, pkm::sqrt_(pkm::max_(0.0f, 1.0f - t1 * t1 - t2 * t2))));
  return pkm::normalize(pkm::mk(au * nh.x, av * nh.y, pkm::max_(0.0f, nh.z)));
}


// Largest i in [0, n) with cdf[i] <= u (cdf[0] = 0 <= u < 1), searched
// between the cutpoint guides of the cells around k = (int)(u * n): the
// guides of cells k-1 and k+2 bracket the answer even when u * n rounds
// across a cell boundary, so the result is the full binary search's.
__device__ inline int guided_search(const float* cdf, const uint32_t* guide, uint32_t n, float u) {
  const int k = (int)(u * (float)n);
  int lo = (int)guide[max(k - 1, 0)];
  int hi = (int)guide[min(k + 2, (int)n)] + 1;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (cdf[mid] <= u) lo = mid;
    else hi = mid;
  }
  return lo;
}

// sampleEnv (shade_hit.wgsl:213-227) for the bins sampler; the CDF sampler
// draws two uniform numbers from the same two hashes and binary-searches the
// marginal and conditional tables.
__device__ inline void sample_env(const DevScene& S, float theta, const pkm::m3& basis, uint32_t* seed, v3* wi,
                                  float* pdf) {
  int cx, cy;
  if (S.env_mode == 1) {
    uint32_t bi = pkm::pcg_next(seed) % S.nbins;
    int2 bin = S.env_bins[bi];
    // bins are never empty (env_sampler.js closes a bin only after adding a
    // texel); the max() only keeps a malformed table from dividing by zero
    int ci = (int)(pkm::pcg_next(seed) % (uint32_t)max(bin.y - bin.x, 1)) + bin.x;
    int2 c = S.env_coords[ci];
    cx = c.x;
    cy = c.y;
  } else {
    const float u1 = pkm::min_(pkm::pcg_rand(seed), 0x1.fffffep-1f);
    const float u2 = pkm::min_(pkm::pcg_rand(seed), 0x1.fffffep-1f);
    cy = guided_search(S.row_cdf, S.row_guide, S.eh, u1);
    const size_t row = (size_t)cy * (S.ew + 1);
    cx = guided_search(S.col_cdf + row, S.col_guide + row, S.ew, u2);
  }
  float u = -theta + ((0.5f + (float)cx) / (float)S.ew);
  float v = (0.5f + (float)cy) / (float)S.eh;
  float th = u * pkm::kTau;
  float ph = v * pkm::kPi;
  float sph, cph, sth, cth;
  pkm::sincosf_(ph, &sph, &cph);
  pkm::sincosf_(th, &sth, &cth);
  v3 dir = pkm::mk(cth * sph, cph, sth * sph);
  *pdf = S.env_pdf[(size_t)cy * S.ew + cx] / sph;
  *wi = pkm::vec_mul(dir, basis);
}

// Result of shading one closest hit (shade_hit.wgsl main :411-483).
struct ShadeOut {
  bool emit;     // emissive surface: L is the contribution and the path ends
  bool scatter;  // the BSDF branch ran (the vertex depends on the parameters)
  bool bounce;   // continuation ray (also the alpha pass-through)
  bool shadow;   // NEE shadow ray
  v3 L;
  v3 b_o, b_d, b_thr;
  v3 s_o, s_d, s_thr;
};

// What the adjoint needs about a scattering vertex: where the evaluation
// parameters were read and the derivatives of the two weights it applied,
// `bsdf` (the continuation's BSDF value, before the MIS weight) and the NEE
// weight (before the path throughput).  Base-color derivatives are per
// channel (d w_c / d base_c), roughness derivatives are taken with respect
// to alpha_e.
struct ShadeGrad {
  float u, v;
  int layer_base, layer_rough;
  float rough;
  v3 bsdf;
  float db_base;
  v3 db_alpha;
  float dn_base;
  v3 dn_alpha;
};

// deterministic shading seed: the bounce index in place of the reference's
// compaction slot, the pixel index local to its view
__device__ inline uint32_t shade_seed(uint32_t bounce, uint32_t pixel, uint32_t sample) {
  uint32_t seed = (bounce * 1973u + pixel * 9277u + sample * 26699u) | 1u;
  return pkm::pcg_next(&seed);
}

template <bool kGrad>
__device__ inline void shade_hit(const DevScene& S, float theta, v3 ro, v3 rd, float t, uint32_t leaf,
                                 uint32_t bary_mask, uint32_t seed, v3 thr, ShadeOut& o, ShadeGrad* g) {
  o.emit = o.scatter = o.bounce = o.shadow = false;
  // getSurfaceInteraction (:179-200), interpolateVertexAttribute (:363-394)
  // With instancing `leaf` is a virtual leaf index: below world_leaves it is
  // one of the scene's own triangle records, else the placement holding it is
  // the last one whose leaf_base is <= leaf.
  const InstRecord* inst = nullptr;
  uint32_t rec = leaf;
  if (S.num_inst && leaf >= S.world_leaves) {
    uint32_t lo = 0, hi = S.num_inst;
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (S.inst_leaf_base[mid] <= leaf) lo = mid;
      else hi = mid;
    }
    inst = S.inst + lo;
    rec = leaf - inst->leaf_delta;
  }
  const uint4* R = S.shade_rec + 4 * (size_t)rec;
  const uint4 a0 = R[0], a1 = R[1], a2 = R[2];
  const int mat = (inst && inst->material >= 0) ? inst->material : (int)R[3].x;
  const float bu = pkm::unpack_unorm16(bary_mask & 0xffffu), bv = pkm::unpack_unorm16(bary_mask >> 16);
  const v3 bary = pkm::mk(1.0f - bu - bv, bu, bv);
  v3 tng, btg, nrm;
  {
    pkm::m3 m;
    m.c0 = pkm::unpack4x8snorm_xyz(a0.x);
    m.c1 = pkm::unpack4x8snorm_xyz(a1.x);
    m.c2 = pkm::unpack4x8snorm_xyz(a2.x);
    tng = pkm::normalize(pkm::mat_mul(m, bary));
    m.c0 = pkm::unpack4x8snorm_xyz(a0.y);
    m.c1 = pkm::unpack4x8snorm_xyz(a1.y);
    m.c2 = pkm::unpack4x8snorm_xyz(a2.y);
    btg = pkm::normalize(pkm::mat_mul(m, bary));
    m.c0 = pkm::unpack4x8snorm_xyz(a0.z);
    m.c1 = pkm::unpack4x8snorm_xyz(a1.z);
    m.c2 = pkm::unpack4x8snorm_xyz(a2.z);
    nrm = pkm::normalize(pkm::mat_mul(m, bary));
  }
  if (inst) {
    // mesh -> scene: tangents by the linear part, the normal by the inverse
    // transpose (the transpose of the scene -> mesh linear part)
    const float* L = inst->o2w;
    const float* A = inst->w2o;
    auto lin = [&](v3 v) {
      return pkm::mk((L[0] * v.x + L[1] * v.y) + L[2] * v.z, (L[3] * v.x + L[4] * v.y) + L[5] * v.z,
                     (L[6] * v.x + L[7] * v.y) + L[8] * v.z);
    };
    tng = pkm::normalize(lin(tng));
    btg = pkm::normalize(lin(btg));
    nrm = pkm::normalize(pkm::mk((A[0] * nrm.x + A[4] * nrm.y) + A[8] * nrm.z,
                                 (A[1] * nrm.x + A[5] * nrm.y) + A[9] * nrm.z,
                                 (A[2] * nrm.x + A[6] * nrm.y) + A[10] * nrm.z));
  }
  const uint32_t w0 = a0.w, w1 = a1.w, w2 = a2.w;
  const float uvx = pkm::unpack_snorm16(w0 & 0xffffu) * bary.x + pkm::unpack_snorm16(w1 & 0xffffu) * bary.y +
                    pkm::unpack_snorm16(w2 & 0xffffu) * bary.z;
  const float uvy = pkm::unpack_snorm16(w0 >> 16) * bary.x + pkm::unpack_snorm16(w1 >> 16) * bary.y +
                    pkm::unpack_snorm16(w2 >> 16) * bary.z;
  const int4 mi = S.materials[mat];
  const float4 nm = sample_tex(S, S.tex, mi.z, uvx, uvy);
  const v3 map_normal = pkm::mk((nm.x - 0.5f) * 2.0f, (nm.y - 0.5f) * 2.0f, (nm.z - 0.0f) * 1.0f);
  SI si;
  {
    pkm::m3 tbn;
    tbn.c0 = tng;
    tbn.c1 = btg;
    tbn.c2 = nrm;
    si.normal = pkm::normalize(pkm::mat_mul(tbn, map_normal));
  }
  const float eps40 = kEpsShadow * 40.0f;
  si.origin = pkm::add(ro, pkm::scale(rd, t - eps40));
  si.basis = onb(si.normal);
  si.wo = pkm::vec_mul(pkm::neg(rd), si.basis);
  const float4 dc = sample_tex(S, S.tex, mi.x, uvx, uvy);
  si.opacity = dc.w;
  const float4 mr = sample_tex(S, S.tex, mi.y, uvx, uvy);
  si.metallic = mr.z;
  si.alpha = mr.y * mr.y;
  // evaluation parameters: the same texels unless an evaluation copy is set
  const bool split = S.tex_eval != S.tex;
  const float4 dce = split ? sample_tex(S, S.tex_eval, mi.x, uvx, uvy) : dc;
  const float4 mre = split ? sample_tex(S, S.tex_eval, mi.y, uvx, uvy) : mr;
  si.base_e = pkm::mk(dce.x, dce.y, dce.z);
  si.rough_e = mre.y;
  si.alpha_e = mre.y * mre.y;
  const float4 em = sample_tex(S, S.tex, mi.w, uvx, uvy);
  si.emission = pkm::mk(em.x, em.y, em.z);
  si.spec_e = pkm::mix3(pkm::mk(1.0f, 1.0f, 1.0f), si.base_e, si.metallic);
  if (pkm::pcg_rand(&seed) > si.opacity) {
    // alpha pass-through (:426-431)
    o.b_o = pkm::add(ro, pkm::scale(rd, t + kEpsShadow * 40.0f));
    o.b_d = rd;
    o.b_thr = thr;
    o.bounce = true;
    return;
  }
  if (si.emission.x * 1.0f + si.emission.y * 1.0f + si.emission.z * 1.0f > 0.0f) {
    // emission (:434-438)
    const float att = pkm::max_(pkm::dot(si.normal, pkm::mat_mul(si.basis, si.wo)), 0.0f);
    o.L = pkm::scale(pkm::mul(pkm::mul(pkm::mk(50.0f, 50.0f, 50.0f), si.emission), thr), att);
    o.emit = true;
    return;
  }
  o.scatter = true;
  if (kGrad) {
    g->u = uvx;
    g->v = uvy;
    g->layer_base = mi.x;
    g->layer_rough = mi.y;
    g->rough = si.rough_e;
  }
  const v3 m = sample_ggx(si, &seed);
  const float f = pkm::mix(schlick(pkm::max_(pkm::dot(si.wo, m), 0.0f), 1.5f), 1.0f, si.metallic);
  v3 wi, bsdf;
  float pdf;
  if (pkm::pcg_rand(&seed) > f) {
    // sampleLambert (:243-253)
    const float r = pkm::sqrt_(pkm::pcg_rand(&seed));
    const float phi = pkm::kTau * pkm::pcg_rand(&seed);
    float sp, cp;
    pkm::sincosf_(phi, &sp, &cp);
    const float x = r * cp, y = r * sp;
    const float z = pkm::sqrt_(pkm::max_(0.0f, 1.0f - x * x - y * y));
    wi = pkm::mk(x, y, z);
    pdf = lambert_pdf(wi);
    bsdf = eval_lambert(si, wi, pdf);
    if (kGrad) {
      g->db_base = pkm::kInvPi * pkm::max_(kEpsShadow, wi.z) / pdf;
      g->db_alpha = pkm::mk(0.0f, 0.0f, 0.0f);
    }
  } else {
    wi = pkm::reflect(pkm::neg(si.wo), m);
    pdf = specular_pdf(si, m);
    bsdf = eval_specular(si, wi, pdf);
    if (kGrad) {
      float k, dk;
      specular_k_grad(si, wi, pdf, &k, &dk);
      g->db_base = si.metallic * k;
      g->db_alpha = pkm::scale(si.spec_e, dk);
    }
  }
  if (kGrad) {
    g->bsdf = bsdf;
    g->dn_base = 0.0f;
    g->dn_alpha = pkm::mk(0.0f, 0.0f, 0.0f);
  }
  if (wi.z > 0.0f) {
    const v3 dir = pkm::mat_mul(si.basis, wi);
    const float weight = S.env_mode ? power_heuristic(pdf, env_pdf(S, theta, dir)) : 1.0f;
    o.b_o = si.origin;
    o.b_d = dir;
    o.b_thr = pkm::scale(pkm::mul(bsdf, thr), weight);
    o.bounce = true;
  }
  if (S.env_mode) {
    v3 ewi;
    float epdf;
    sample_env(S, theta, si.basis, &seed, &ewi, &epdf);
    if (pkm::dot(ewi, m) > 0.0f && ewi.z > 0.0f) {
      // createMaterialSampleRay (:339-348)
      const float lw = power_heuristic(epdf, lambert_pdf(ewi));
      const v3 el = eval_lambert(si, ewi, epdf);
      v3 sc = pkm::mk((1.0f - f) * el.x * lw, (1.0f - f) * el.y * lw, (1.0f - f) * el.z * lw);
      // specular_pdf(si, hh) and eval_specular(si, ewi, epdf) share the
      // half vector, and also D and G1(wo) unless the evaluation roughness
      // differs from the sampling one (a scene-wide, wave-uniform test)
      const v3 hh = pkm::normalize(pkm::add(si.wo, ewi));
      const float d_s = ggx_d(hh, si.alpha, si.alpha);
      const float g1o_s = ggx_g1(si.wo, hh, si.alpha, si.alpha);
      const float sw = power_heuristic(epdf, pkm::max_(kEpsShadow, d_s * g1o_s / (4.0f * si.wo.z)));
      float d_e = d_s, g1o_e = g1o_s;
      if (split) {
        d_e = ggx_d(hh, si.alpha_e, si.alpha_e);
        g1o_e = ggx_g1(si.wo, hh, si.alpha_e, si.alpha_e);
      }
      const v3 ev = pkm::scale(
          si.spec_e, pkm::max_(d_e * (ggx_g1(ewi, hh, si.alpha_e, si.alpha_e) * g1o_e) / (4.0f * si.wo.z * epdf), 0.0f));
      sc = pkm::add(sc, pkm::mk(f * ev.x * sw, f * ev.y * sw, f * ev.z * sw));
      o.s_o = si.origin;
      o.s_d = pkm::mat_mul(si.basis, ewi);
      o.s_thr = pkm::mul(sc, thr);
      o.shadow = true;
      if (kGrad) {
        float ks, dks;
        specular_k_grad(si, ewi, epdf, &ks, &dks);
        g->dn_base = (1.0f - f) * lw * (pkm::kInvPi * pkm::max_(kEpsShadow, ewi.z) / epdf) + f * sw * si.metallic * ks;
        g->dn_alpha = pkm::scale(si.spec_e, f * sw * dks);
      }
    }
  }
}

constexpr int kShadeWaves = 4;  // k_shade blocks: 256 threads

// 7 waves per SIMD (70 VGPRs, no scratch); 8 (64 VGPRs) spills 32 bytes per
// lane and measured 10% slower.
__global__ void __launch_bounds__(kShadeWaves * 64, 7) k_shade(DevScene S, FrameDev F, int bounce) {
  const int in = bounce & 1;
  const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx == 0) F.cnt[kCntWorkShadow] = 0;
  const uint32_t nh = F.cnt[kCntHits];  // hit records and holes
  // The grid covers the queue's capacity (fixed when the pass is captured);
  // most bounces fill a small part of it, and the blocks past the records
  // leave at once (block-uniform: no thread of theirs reaches a barrier).
  if (blockIdx.x * blockDim.x >= nh) return;
  const FrameParams P = *F.params;
  const uint32_t ns = F.views ? F.S : pass_samples(F, P);
  ShadeOut o;
  o.emit = o.scatter = o.bounce = o.shadow = false;
  uint32_t pid = 0, view_bits = 0;
  // compact the block's hit records in LDS (their order kept), dropping the
  // holes at the ends of trace_closest's per-wave runs
  __shared__ uint4 s_hit[kShadeWaves * 64];
  __shared__ uint32_t s_wcnt[kShadeWaves];
  {
    const uint4 h = idx < nh ? F.hits[idx] : make_uint4(0xffffffffu, 0u, 0u, 0u);
    const bool is_hit = h.x != 0xffffffffu;
    const unsigned long long m = __ballot(is_hit);
    const int lane = (int)__lane_id(), wave = (int)(threadIdx.x >> 6);
    if (lane == 0) s_wcnt[wave] = (uint32_t)__popcll(m);
    __syncthreads();
    uint32_t at = (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
    for (int w = 0; w < wave; ++w) at += s_wcnt[w];
    if (is_hit) s_hit[at] = h;
    __syncthreads();
  }
  uint32_t nb = 0;
#pragma unroll
  for (int w = 0; w < kShadeWaves; ++w) nb += s_wcnt[w];
  if (threadIdx.x < nb) {
    const uint4 hr = s_hit[threadIdx.x];
    const uint32_t slot = hr.x;
    const float4 o4 = F.qo[in][slot], d4 = F.qd[in][slot];
    pid = __float_as_uint(o4.w);
    view_bits = __float_as_uint(d4.w);
    // the pixel index the seeds use is local to the view; a single-view
    // pass holds samples P.sample .. P.sample + ns - 1 of every pixel,
    // numbered pixel-major
    const uint32_t s = pid % ns;
    const uint32_t pixel = F.views ? pid / ns - F.views[view_bits].offset : pid / ns;
    const float4 t4 = bounce == 0 ? make_float4(1.0f, 1.0f, 1.0f, 0.0f) : F.qt[in][slot];  // camera rays: 1
    shade_hit<false>(S, P.env_theta, pkm::mk(o4.x, o4.y, o4.z), pkm::mk(d4.x, d4.y, d4.z), __uint_as_float(hr.z),
                     hr.y, hr.w, shade_seed((uint32_t)bounce, pixel, P.sample + s), pkm::mk(t4.x, t4.y, t4.z), o,
                     nullptr);
    if (o.emit) add_color(F, ns, pid, o.L);
  }
  // (every thread of the block reaches this point: block-wide append)
  uint32_t bs, ss;
  block_append2<kShadeWaves>(&F.cnt[ray_counter(in ^ 1)], o.bounce, &F.cnt[kCntShadow], o.shadow, &bs, &ss);
  if (o.bounce) {
    F.qo[in ^ 1][bs] = make_float4(o.b_o.x, o.b_o.y, o.b_o.z, __uint_as_float(pid));
    F.qd[in ^ 1][bs] = make_float4(o.b_d.x, o.b_d.y, o.b_d.z, __uint_as_float(view_bits));
    F.qt[in ^ 1][bs] = make_float4(o.b_thr.x, o.b_thr.y, o.b_thr.z, 0.0f);
  }
  if (o.shadow) {
    F.so[ss] = make_float4(o.s_o.x, o.s_o.y, o.s_o.z, __uint_as_float(pid));
    F.sd[ss] = make_float4(o.s_d.x, o.s_d.y, o.s_d.z, __uint_as_float(view_bits));
    F.sthr[ss] = make_float4(o.s_thr.x, o.s_thr.y, o.s_thr.z, 0.0f);
  }
}

// ---------------------------------------------------------------------------
// trace_shadow.wgsl main (:192-207) + shade_miss for unoccluded shadow rays
// ---------------------------------------------------------------------------
template <bool kCount, bool kInst, int kLds, int kWpb, int kNW = 8>
__global__ void __launch_bounds__(64 * kWpb, trace_waves_per_simd(kLds)) k_trace_shadow(DevScene S, FrameDev F,
                                                                               uint32_t refill_min) {
  PK_TRACE_WAVE_SETUP(!kInst);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    F.cnt[kCntHits] = 0;
    F.cnt[kCntWorkTrace] = 0;
  }
  const uint32_t n = F.cnt[kCntShadow];
  const uint32_t ns = F.views ? F.S : pass_samples(F, *F.params);
  const float theta = F.params->env_theta;
  const Stack stk{(lds_u32)lds_node[wave], (lds_f32)lds_t[wave], F.spill + wave_id * kSpillStride * 64, lane, (int)S.stack_limit,
                  (lds_u64)lds_key[wave]};
  TraverseStats st = {0, 0};
  uint32_t traced = 0, overflow = 0;  // traced: wave-uniform, counted at each flush
  trace_loop<true, kCount, kInst, kLds, kNW>(
      S, n, &F.cnt[kCntWorkShadow], refill_min, stk, &st, &overflow,
      [&](uint32_t slot, RayState& r) {
        const float4 o4 = F.so[slot], d4 = F.sd[slot];
        ray_init<true>(r, pkm::mk(o4.x, o4.y, o4.z), pkm::mk(d4.x, d4.y, d4.z));
        if (F.views) r.node = F.views[__float_as_uint(d4.w)].root;
      },
      [&](uint32_t, const RayState&) {},
      [&](bool done, uint32_t slot, const RayState& r) {
        traced += (uint32_t)__popcll(__ballot(done));
        // unoccluded rays are shaded at the refill, together (as misses are
        // in k_trace_closest): env radiance x NEE throughput
        if (done && ray_missed(r)) {
          const uint32_t pid = __float_as_uint(F.so[slot].w);
          const float4 t4 = F.sthr[slot];
          add_color(F, ns, pid, pkm::mul(env_color(S, theta, r.d), pkm::mk(t4.x, t4.y, t4.z)));
        }
      });
  if (kCount) {
    st.nodes = wave_sum(st.nodes);
    st.tris = wave_sum(st.tris);
  }
  overflow = wave_sum(overflow);
  if (lane == 0) {
    atomicAdd(&F.stats[kStatShadow], (unsigned long long)traced);
    if (kCount) {
      atomicAdd(&F.stats[kStatShadowNodes], (unsigned long long)st.nodes);
      atomicAdd(&F.stats[kStatShadowTris], (unsigned long long)st.tris);
    }
    if (overflow) atomicOr(&F.cnt[kCntOverflow], 1u);
  }
}

// ---------------------------------------------------------------------------
// Path megakernel: a pass in one persistent launch.  Each lane owns one path
// at a time and takes it through every bounce (closest hit, miss / emission /
// shading, NEE shadow ray) in registers, summing its contributions in the
// order the wavefront kernels add them to its color slot; when the path ends
// it writes that slot (rgb sum, contribution count) once and takes the next
// path of the pass.  The per-bounce queues (rays, hits, shadow rays,
// throughput) never reach memory, and lanes whose path ended refill between
// bounces, so waves stay full while path lengths differ.  Path numbering,
// seeds and arithmetic are the wavefront's (pixel-major samples; a batch
// numbers global pixels, with the view's camera and root, and has no path
// for samples beyond the view's spp), so the color buffer
// is bit-identical to the one raygen / trace_closest / shade / trace_shadow
// produce.  The work cursor (kCntWorkTrace) is zeroed before the launch.
// ---------------------------------------------------------------------------
template <bool kInst, bool kBatch>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(kMegaWavesPerSimd)))
    k_render_paths(DevScene S, FrameDev F, uint32_t depth) {
  __shared__ uint32_t lds_node[kStackLds * 64];
  __shared__ float lds_t[kStackLds * 64];
  const int lane = (int)threadIdx.x;
  __shared__ unsigned long long lds_key[128];
  const Stack stk{(lds_u32)lds_node, (lds_f32)lds_t, F.spill + (size_t)blockIdx.x * kSpillStride * 64, lane, (int)S.stack_limit,
                  (lds_u64)lds_key};
  const FrameParams P = *F.params;
  const uint32_t ns = kBatch ? F.S : pass_samples(F, P);
  const uint32_t B = F.W * F.H;
  const uint32_t n = B * ns;
  uint32_t* const cursor = &F.cnt[kCntWorkTrace];
  TraverseStats st = {0, 0};
  uint32_t traced = 0, hits = 0, shadows = 0, overflow = 0;
  const v3 zero = pkm::mk(0.0f, 0.0f, 0.0f);
  bool has = false;
  uint32_t pixel = 0, sample = 0, slot = 0, root = 0, b = 0;
  v3 o = zero, d = zero, thr = zero, L = zero;
  float cnt = 0.0f;
  // paths claimed by this wave and not yet started (wave-uniform)
  uint32_t pool_lo = 0, pool_hi = 0;
  bool claimed_all = n == 0;
  while (true) {
    // lanes without a path take the next ones of the pool (again while a
    // batch path of a finished view left a lane idle)
    for (unsigned long long idle = __ballot(!has); idle && !(claimed_all && pool_lo >= pool_hi);
         idle = __ballot(!has)) {
      const uint32_t nidle = (uint32_t)__popcll(idle);
      const uint32_t have = pool_hi - pool_lo;
      uint32_t fresh = 0;
      bool got = false;
      if (have < nidle && !claimed_all) {
        uint32_t c0 = 0;
        if (lane == 0) c0 = atomicAdd(cursor, kClaimChunk);
        fresh = (uint32_t)__builtin_amdgcn_readfirstlane((int)__shfl(c0, 0));
        got = true;
        claimed_all = fresh >= n || n - fresh <= kClaimChunk;
      }
      if (!has) {
        const uint32_t rank = (uint32_t)__popcll(idle & ((1ull << lane) - 1ull));
        const uint32_t p = rank < have ? pool_lo + rank : (got ? fresh + (rank - have) : 0xffffffffu);
        if (p < n) {
          if (kBatch) {
            const uint32_t gp = p / ns, sm = p - gp * ns;
            const ViewDev& V = F.views[view_of(F, gp)];
            if (P.sample + sm < V.spp) {
              has = true;
              pixel = gp - V.offset;
              sample = P.sample + sm;
              slot = p;  // gp * S + sm
              root = V.root;
              camera_ray(pixel % V.W, pixel / V.W, V.W, V.H, sample, V.cam, &o, &d);
            }
          } else {
            has = true;
            pixel = p / ns;
            const uint32_t s = p - pixel * ns;
            sample = P.sample + s;
            slot = color_slot(F, ns, p);
            camera_ray_b(pixel % F.W, pixel / F.W, F.W, F.H, sample, P.cam, pkm::mk(P.cbx[0], P.cbx[1], P.cbx[2]),
                         pkm::mk(P.cby[0], P.cby[1], P.cby[2]), &o, &d);
          }
          b = 0;
          thr = pkm::mk(1.0f, 1.0f, 1.0f);
          L = zero;
          cnt = 0.0f;
        }
      }
      if (have >= nidle) {
        pool_lo += nidle;
      } else if (got) {
        pool_lo = fresh + (nidle - have);
        pool_hi = fresh + kClaimChunk;
      } else {
        pool_lo = pool_hi;
      }
      pool_hi = min(pool_hi, n);
      pool_lo = min(pool_lo, pool_hi);
    }
    if (__ballot(has) == 0ull) break;
    // closest hit (trace.wgsl) and miss shading (shade_miss.wgsl)
    RayState r;
    r.best_leaf = 0xffffffffu;
    r.d = d;
    if (has) {
      ray_init<false>(r, o, d);
      if (kBatch) r.node = root;
    }
    for (bool tr = has; __ballot(tr);)
      if (wave_step<false, false, kInst, kStackLds, 8>(S, r, tr, stk, &st)) tr = false;
    if (has) {
      if (r.overflow) overflow = 1;
      traced++;
    }
    bool end = false;
    ShadeOut so;
    so.emit = so.scatter = so.bounce = so.shadow = false;
    if (has) {
      if (ray_missed(r)) {
        const v3 c = pkm::mul(env_color(S, P.env_theta, r.d), thr);
        L.x += c.x;
        L.y += c.y;
        L.z += c.z;
        cnt += 1.0f;
        end = true;
      } else {
        hits++;
        shade_hit<false>(S, P.env_theta, o, d, r.tbest, r.best_leaf, r.best_bary, shade_seed(b, pixel, sample), thr,
                         so, nullptr);
        if (so.emit) {
          L.x += so.L.x;
          L.y += so.L.y;
          L.z += so.L.z;
          cnt += 1.0f;
        }
      }
    }
    // NEE shadow ray (trace_shadow.wgsl), unoccluded: env radiance x NEE throughput
    RayState q;
    q.best_leaf = 0xffffffffu;
    if (so.shadow) {
      ray_init<true>(q, so.s_o, so.s_d);
      if (kBatch) q.node = root;
    }
    for (bool tr = so.shadow; __ballot(tr);)
      if (wave_step<true, false, kInst, kStackLds, 8>(S, q, tr, stk, &st)) tr = false;
    if (so.shadow) {
      if (q.overflow) overflow = 1;
      shadows++;
      if (ray_missed(q)) {
        const v3 c = pkm::mul(env_color(S, P.env_theta, q.d), so.s_thr);
        L.x += c.x;
        L.y += c.y;
        L.z += c.z;
        cnt += 1.0f;
      }
    }
    if (has && !end) {
      // the continuation of the last bounce is never traced
      if (so.bounce && b + 1u < depth) {
        o = so.b_o;
        d = so.b_d;
        thr = so.b_thr;
        b++;
      } else {
        end = true;
      }
    }
    if (end) {
      F.color[slot] = make_float4(L.x, L.y, L.z, cnt);
      has = false;
    }
  }
  traced = wave_sum(traced);
  hits = wave_sum(hits);
  shadows = wave_sum(shadows);
  overflow = wave_sum(overflow);
  if (lane == 0) {
    atomicAdd(&F.stats[kStatRays], (unsigned long long)traced);
    atomicAdd(&F.stats[kStatHits], (unsigned long long)hits);
    atomicAdd(&F.stats[kStatShadow], (unsigned long long)shadows);
    if (overflow) atomicOr(&F.cnt[kCntOverflow], 1u);
  }
}

// ---------------------------------------------------------------------------
// accumulate.wgsl main (:20-34): running mean over frames.  A pass of S
// samples folds them in sample order with the same per-sample update, so the
// mean equals S one-sample frames bit for bit.
// ---------------------------------------------------------------------------
// A pixel's S color slots are adjacent, so one thread per pixel reading its
// own would stride S * 16 bytes across the wave: a block stages its kAccPix
// pixels' slots in LDS, kAccChunk samples at a time, with loads that read
// each pixel's chunk contiguously, and every thread folds its own pixel in
// sample order.  A pixel's row in LDS is padded by one slot (kAccRow), so
// the fold's 16-lane phases of 16-byte reads fall on distinct banks.
constexpr uint32_t kAccPix = 256, kAccChunk = 8, kAccRow = kAccChunk + 1;
__global__ void __launch_bounds__(256) k_accumulate(FrameDev F) {
  __shared__ float4 tile[kAccPix * kAccRow];
  const uint32_t B = F.W * F.H, S = F.S;
  const uint32_t pix0 = blockIdx.x * kAccPix;
  const uint32_t t = threadIdx.x;
  const FrameParams P = *F.params;
  // this thread's pixel and its sample count this pass (batch: a view's
  // pixels fold the round's samples below its spp; views whose budget is
  // spent keep their mean)
  const uint32_t idx = pix0 + t;
  uint32_t ns = 0;
  float4 a = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (idx < B) {
    ns = pass_samples(F, P);
    if (F.views) {
      const uint32_t spp = F.views[view_of(F, idx)].spp;
      ns = P.sample < spp ? min(S, spp - P.sample) : 0u;
    }
    a = F.accum[idx];
  }
  const uint32_t nmax = F.views ? S : pass_samples(F, P);
  for (uint32_t c0 = 0; c0 < nmax; c0 += kAccChunk) {
    for (uint32_t e = t; e < kAccPix * kAccChunk; e += blockDim.x) {
      const uint32_t pl = e / kAccChunk, k = e - pl * kAccChunk;
      if (pix0 + pl < B && c0 + k < nmax) tile[pl * kAccRow + k] = F.color[(size_t)(pix0 + pl) * S + c0 + k];
    }
    __syncthreads();
    const uint32_t kend = ns > c0 ? min(kAccChunk, ns - c0) : 0u;
    for (uint32_t k = 0; k < kend; ++k) {
      const float sm = (float)(P.sample + c0 + k);
      const float s1 = (float)(P.sample + c0 + k + 1u);
      const float4 c = tile[t * kAccRow + k];
      a.x = (pkm::max0_(c.x) + a.x * sm) / s1;
      a.y = (pkm::max0_(c.y) + a.y * sm) / s1;
      a.z = (pkm::max0_(c.z) + a.z * sm) / s1;
    }
    __syncthreads();
  }
  if (idx < B && ns > 0) {
    a.w = 1.0f;
    F.accum[idx] = a;
  }
}

// Frame bookkeeping kernels (no memset/memcpy nodes, so frames can be
// captured and replayed as one graph).
__global__ void k_begin(FrameDev F, uint32_t sample, float theta, pk_camera cam, uint32_t limit) {
  F.params->sample = sample;
  F.params->env_theta = theta;
  F.params->cam = cam;
  F.params->limit = limit;
  v3 bx, by;
  camera_basis(cam, &bx, &by);
  F.params->cbx[0] = bx.x, F.params->cbx[1] = bx.y, F.params->cbx[2] = bx.z;
  F.params->cby[0] = by.x, F.params->cby[1] = by.y, F.params->cby[2] = by.z;
}
__global__ void k_advance(FrameDev F) {
  const FrameParams P = *F.params;
  F.stats[kStatFrames] += (unsigned long long)(F.views ? 1u : pass_samples(F, P));
  F.params->sample = P.sample + F.S;
}
// Shading records (DevScene::shade_rec) from the leaf-order {i0, i1, i2,
// mat} records and the vertex attributes, one record per thread.
__global__ void __launch_bounds__(256) k_build_shade_records(const int4* tri, const uint4* attrs, uint4* rec,
                                                             uint32_t n) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= n) return;
  const int4 t = tri[i];
  uint4* r = rec + 4 * (size_t)i;
  r[0] = attrs[t.x];
  r[1] = attrs[t.y];
  r[2] = attrs[t.z];
  r[3] = make_uint4((uint32_t)t.w, 0u, 0u, 0u);
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_fill_f4(float4* p, float4 v, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void __launch_bounds__(256) k_copy_f4(float4* dst, const float4* src, uint32_t n) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) dst[i] = src[i];
}

// ---------------------------------------------------------------------------
// Path-replay adjoint (differentiation of the frame's image)
// ---------------------------------------------------------------------------
__device__ inline float wave_sum_f(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// Rounds of key peeling before the remaining lanes add directly.
constexpr int kScatterPeel = 4;

// Adds v[0..C) to dst[key * 4 + c0 + c] for every lane where `valid` holds.
// Lanes sharing a key are summed across the wave first: a segmented
// reduction that peels one key per round (the key of the lowest pending
// lane), so a texel hit by many lanes (coherent camera paths, the bright
// environment texels most NEE rays pick) costs one atomic per wave.  After
// kScatterPeel rounds the lanes still pending add on their own.  Called by
// the whole wave.
template <int C>
__device__ inline void wave_scatter_add(float* dst, uint32_t key, bool valid, const float* v, int c0) {
  const int lane = (int)__lane_id();
  // a path that turned non-finite (its sample is dropped by accumulate's
  // NaN rule) scatters nothing from the point where it did
#pragma unroll
  for (int c = 0; c < C; ++c) valid = valid && pkm::abs_(v[c]) < __builtin_inff();
  unsigned long long pending = __ballot(valid);
  for (int round = 0; pending; ++round) {
    const bool mine_pending = (pending >> lane) & 1ull;
    if (round == kScatterPeel) {
      if (mine_pending)
#pragma unroll
        for (int c = 0; c < C; ++c) unsafeAtomicAdd(dst + (size_t)key * 4 + c0 + c, v[c]);
      break;
    }
    const int leader = __ffsll((long long)pending) - 1;
    const uint32_t k = __shfl(key, leader);
    const bool mine = mine_pending && key == k;
    const unsigned long long grp = __ballot(mine);
    if (__popcll(grp) == 1) {
      if (lane == leader)
#pragma unroll
        for (int c = 0; c < C; ++c) unsafeAtomicAdd(dst + (size_t)k * 4 + c0 + c, v[c]);
    } else {
      float s[C];
#pragma unroll
      for (int c = 0; c < C; ++c) s[c] = wave_sum_f(mine ? v[c] : 0.0f);
      if (lane == leader)
#pragma unroll
        for (int c = 0; c < C; ++c) unsafeAtomicAdd(dst + (size_t)k * 4 + c0 + c, s[c]);
    }
    pending &= ~grp;
  }
}

// One thread per path (pixel, sample); persistent one-wave blocks pull
// 64-path batches from A.cursor.  Each path is replayed once from its seeds
// with the same arithmetic as the wavefront frame, summing its radiance L
// (equal to the frame's color for that sample).  With delta = dLoss/dI / N:
//   environment texel e of a miss or an unoccluded NEE ray:
//     dLoss/dE_e += delta * (throughput of that contribution),
//   scattered as the contribution is met;
//   base color / alpha at a vertex with throughput thr:
//     NEE:   delta * E * thr * d(nee weight)/d(theta)
//     bounce: delta * rem * d(bsdf)/d(theta) / bsdf
//   and d alpha / d roughness = 2 * roughness, scattered bilinearly to the
//   4 texels of the lookup.  rem, the radiance that flows through a vertex's
//   continuation weight, is L minus the contributions met up to and
//   including that vertex, so the vertex terms wait in A.vtx (kAdjVtxWords
//   words per vertex, lane-interleaved per wave) until the path has ended,
//   and are then scattered with rem = L - c_0 - ... - c_b, subtracted in
//   path order.
template <bool kInst>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) k_adjoint(DevScene S, AdjointArgs A) {
  __shared__ uint32_t lds_node[kStackLds * 64];
  __shared__ float lds_t[kStackLds * 64];
  const int lane = (int)threadIdx.x;
  __shared__ unsigned long long lds_key[128];
  const Stack stk{(lds_u32)lds_node, (lds_f32)lds_t, A.spill + (size_t)blockIdx.x * kSpillStride * 64, lane, (int)S.stack_limit,
                  (lds_u64)lds_key};
  float* const vtx = A.vtx + (size_t)blockIdx.x * A.depth * kAdjVtxWords * 64 + lane;
  const uint32_t B = A.W * A.H;
  const uint32_t total = B * A.num_frames;
  const uint32_t plane = S.tw * S.th;
  TraverseStats st = {0, 0};
  uint32_t overflow = 0;
  const v3 zero = pkm::mk(0.0f, 0.0f, 0.0f);
  while (true) {
    uint32_t base = 0;
    if (lane == 0) base = atomicAdd(A.cursor, 64u);
    base = __shfl(base, 0);
    if (base >= total) break;
    const uint32_t i = base + (uint32_t)lane;
    uint32_t pixel = 0, sample = 0;
    v3 delta = zero;
    bool live = i < total;
    if (live) {
      pixel = i % B;
      sample = A.first_sample + i / B;
      const float4 g4 = A.d_image[pixel];
      delta = pkm::mk(g4.x * A.inv_frames, g4.y * A.inv_frames, g4.z * A.inv_frames);
      live = g4.x != 0.0f || g4.y != 0.0f || g4.z != 0.0f;
    }
    v3 sum = zero;
    uint32_t nv = 0;  // vertices this lane recorded
    bool alive = live;
    v3 o = zero, d = zero, thr = pkm::mk(1.0f, 1.0f, 1.0f);
    if (alive) camera_ray(pixel % A.W, pixel / A.W, A.W, A.H, sample, A.cam, &o, &d);
    for (uint32_t b = 0; b < A.depth; ++b) {
      if (__ballot(alive) == 0ull) break;
      float* const rec = vtx + (size_t)b * kAdjVtxWords * 64;
      const bool at_vertex = alive;
      v3 c = zero;  // this vertex's contribution (at most one of the three)
      // closest hit (trace.wgsl) and miss shading (shade_miss.wgsl)
      RayState r;
      r.tbest = kMaxT;
      r.best_leaf = 0xffffffffu;
      r.best_bary = 0;
      if (alive) ray_init<false>(r, o, d);
      for (bool tr = alive; __ballot(tr);)
        if (wave_step<false, false, kInst, kStackLds, 8>(S, r, tr, stk, &st)) tr = false;
      if (alive && r.overflow) overflow = 1;
      const bool miss = alive && ray_missed(r);
      if (miss) {
        c = pkm::mul(env_color(S, A.theta, d), thr);
        alive = false;
      }
      {
        const bool valid = miss && S.env_mode != 0;
        const float ev[3] = {delta.x * thr.x, delta.y * thr.y, delta.z * thr.z};
        wave_scatter_add<3>(A.g_env, valid ? env_texel(S, A.theta, d) : 0u, valid, ev, 0);
      }
      // shading (shade_hit.wgsl)
      ShadeOut so;
      so.emit = so.scatter = so.bounce = so.shadow = false;
      ShadeGrad sg;
      sg.u = sg.v = sg.rough = 0.0f;
      sg.layer_base = sg.layer_rough = 0;
      const bool hit = alive;
      if (hit) {
        shade_hit<true>(S, A.theta, o, d, r.tbest, r.best_leaf, r.best_bary, shade_seed(b, pixel, sample), thr, so,
                        &sg);
        if (so.emit) {
          c = so.L;
          alive = false;
        }
      }
      // NEE shadow ray (trace_shadow.wgsl) and its environment term
      bool vis = false;
      {
        RayState q;
        q.best_leaf = 0xffffffffu;
        const bool sh = hit && so.shadow;
        if (sh) ray_init<true>(q, so.s_o, so.s_d);
        for (bool tr = sh; __ballot(tr);)
          if (wave_step<true, false, kInst, kStackLds, 8>(S, q, tr, stk, &st)) tr = false;
        if (sh) {
          if (q.overflow) overflow = 1;
          vis = ray_missed(q);
        }
      }
      v3 E = zero;
      if (vis) {
        E = env_color(S, A.theta, so.s_d);
        c = pkm::mul(E, so.s_thr);
      }
      {
        const float ev[3] = {delta.x * so.s_thr.x, delta.y * so.s_thr.y, delta.z * so.s_thr.z};
        wave_scatter_add<3>(A.g_env, vis ? env_texel(S, A.theta, so.s_d) : 0u, vis, ev, 0);
      }
      if (at_vertex) {
        sum.x += c.x;
        sum.y += c.y;
        sum.z += c.z;
        const bool scat = hit && so.scatter;
        // the last vertex's continuation is never traced
        const bool cont = scat && so.bounce && b + 1 < A.depth;
        rec[0 * 64] = c.x;
        rec[1 * 64] = c.y;
        rec[2 * 64] = c.z;
        rec[3 * 64] = __uint_as_float((scat ? 1u : 0u) | (cont ? 2u : 0u));
        if (scat) {
          v3 ga = zero;
          float gr = 0.0f;
          if (vis) {
            const v3 w = pkm::mul(pkm::mul(delta, E), thr);
            ga = pkm::scale(w, sg.dn_base);
            gr = pkm::dot(w, sg.dn_alpha);
          }
          rec[4 * 64] = ga.x;
          rec[5 * 64] = ga.y;
          rec[6 * 64] = ga.z;
          rec[7 * 64] = gr;
          rec[8 * 64] = sg.rough;
          rec[9 * 64] = sg.u;
          rec[10 * 64] = sg.v;
          rec[11 * 64] = __int_as_float(sg.layer_base);
          rec[12 * 64] = __int_as_float(sg.layer_rough);
          if (cont) {
            rec[13 * 64] = sg.bsdf.x;
            rec[14 * 64] = sg.bsdf.y;
            rec[15 * 64] = sg.bsdf.z;
            rec[16 * 64] = sg.db_base;
            rec[17 * 64] = sg.db_alpha.x;
            rec[18 * 64] = sg.db_alpha.y;
            rec[19 * 64] = sg.db_alpha.z;
          }
        }
        nv = b + 1;
      }
      if (hit && !so.emit) {
        if (so.bounce) {
          o = so.b_o;
          d = so.b_d;
          thr = so.b_thr;
        } else {
          alive = false;
        }
      }
    }
    // the path has ended: scatter the vertex terms, subtracting each
    // contribution from rem in path order
    v3 rem = sum;
    for (uint32_t b = 0; __ballot(b < nv) != 0ull; ++b) {
      const float* const rec = vtx + (size_t)b * kAdjVtxWords * 64;
      bool scat = false;
      v3 ga = zero;
      float gr = 0.0f, u = 0.0f, v = 0.0f;
      int layer_base = 0, layer_rough = 0;
      if (b < nv) {
        rem.x -= rec[0 * 64];
        rem.y -= rec[1 * 64];
        rem.z -= rec[2 * 64];
        const uint32_t flags = __float_as_uint(rec[3 * 64]);
        scat = (flags & 1u) != 0u;
        if (scat) {
          ga = pkm::mk(rec[4 * 64], rec[5 * 64], rec[6 * 64]);
          gr = rec[7 * 64];
          const float rough = rec[8 * 64];
          u = rec[9 * 64];
          v = rec[10 * 64];
          layer_base = __float_as_int(rec[11 * 64]);
          layer_rough = __float_as_int(rec[12 * 64]);
          if (flags & 2u) {
            const v3 bsdf = pkm::mk(rec[13 * 64], rec[14 * 64], rec[15 * 64]);
            const float db_base = rec[16 * 64];
            const v3 db_alpha = pkm::mk(rec[17 * 64], rec[18 * 64], rec[19 * 64]);
            const float q0 = bsdf.x != 0.0f ? delta.x * rem.x / bsdf.x : 0.0f;
            const float q1 = bsdf.y != 0.0f ? delta.y * rem.y / bsdf.y : 0.0f;
            const float q2 = bsdf.z != 0.0f ? delta.z * rem.z / bsdf.z : 0.0f;
            ga = pkm::add(ga, pkm::mk(q0 * db_base, q1 * db_base, q2 * db_base));
            gr += q0 * db_alpha.x + q1 * db_alpha.y + q2 * db_alpha.z;
          }
          gr *= 2.0f * rough;
        }
      }
      const Footprint fp = tex_footprint(S, u, v);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const float va[3] = {ga.x * fp.w[j], ga.y * fp.w[j], ga.z * fp.w[j]};
        wave_scatter_add<3>(A.g_atlas, (uint32_t)layer_base * plane + fp.idx[j], scat, va, 0);
        const float vr[1] = {gr * fp.w[j]};
        wave_scatter_add<1>(A.g_atlas, (uint32_t)layer_rough * plane + fp.idx[j], scat, vr, 1);
      }
    }
  }
  overflow = wave_sum(overflow);
  if (lane == 0 && overflow) atomicOr(A.overflow, 1u);
}

// L2 image loss: loss += sum over pixels and rgb of (img - target)^2 (in
// f64), grad = 2 (img - target).
__global__ void __launch_bounds__(256) k_loss_l2(const float4* img, const float4* target, float4* grad,
                                                 double* loss, uint32_t n) {
  __shared__ double part[4];
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  double l = 0.0;
  if (i < n) {
    const float4 a = img[i], t = target[i];
    const float dx = a.x - t.x, dy = a.y - t.y, dz = a.z - t.z;
    grad[i] = make_float4(2.0f * dx, 2.0f * dy, 2.0f * dz, 0.0f);
    l = (double)dx * dx + (double)dy * dy + (double)dz * dz;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) l += __shfl_xor(l, off);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = l;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(loss, part[0] + part[1] + part[2] + part[3]);
}

// ---------------------------------------------------------------------------
// Standalone ray queries (traversal microbenchmark, pk_trace_rays)
// ---------------------------------------------------------------------------
template <bool kAnyHit, bool kCount, bool kInst, int kLds, int kWpb, int kNW = 8>
__global__ void __launch_bounds__(64 * kWpb, trace_waves_per_simd(kLds)) k_trace_rays(DevScene S, const float4* rays,
                                                                             uint32_t n, int32_t* ids, float* tout,
                                                                             uint32_t* bary, uint32_t* cursor,
                                                                             uint2* spill_all,
                                                                             unsigned long long* stats,
                                                                             uint32_t* overflow_flag,
                                                                             uint32_t refill_min) {
  PK_TRACE_WAVE_SETUP(!(kInst && kAnyHit));
  const Stack stk{(lds_u32)lds_node[wave], (lds_f32)lds_t[wave], spill_all + wave_id * kSpillStride * 64, lane, (int)S.stack_limit,
                  (lds_u64)lds_key[wave]};
  TraverseStats st = {0, 0};
  uint32_t overflow = 0;
  trace_loop<kAnyHit, kCount, kInst, kLds, kNW>(
      S, n, cursor, refill_min, stk, &st, &overflow,
      [&](uint32_t i, RayState& r) {
        const float4 o4 = rays[2 * (size_t)i], d4 = rays[2 * (size_t)i + 1];
        ray_init<kAnyHit>(r, pkm::mk(o4.x, o4.y, o4.z), pkm::mk(d4.x, d4.y, d4.z));
      },
      [&](uint32_t, const RayState&) {},
      [&](bool done, uint32_t i, const RayState& r) {  // results written at the refill, together
        if (done) {
          ids[i] = (int32_t)ray_hit_id<kInst>(S, r);
          tout[i] = r.tbest;
          bary[i] = r.best_bary;
        }
      });
  if (kCount) {
    st.nodes = wave_sum(st.nodes);
    st.tris = wave_sum(st.tris);
  }
  overflow = wave_sum(overflow);
  if (lane == 0) {
    if (kCount) {
      atomicAdd(&stats[0], (unsigned long long)st.nodes);
      atomicAdd(&stats[1], (unsigned long long)st.tris);
    }
    if (overflow) atomicOr(overflow_flag, 1u);
  }
}

// ---------------------------------------------------------------------------
// host launch wrappers
// ---------------------------------------------------------------------------
static inline uint32_t cdiv(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

void launch_begin(const FrameDev& F, uint32_t sample, float theta, const pk_camera& cam, hipStream_t s,
                  uint32_t limit) {
  hipLaunchKernelGGL(k_begin, dim3(1), dim3(1), 0, s, F, sample, theta, cam, limit);
}
void launch_advance(const FrameDev& F, hipStream_t s) { hipLaunchKernelGGL(k_advance, dim3(1), dim3(1), 0, s, F); }
void launch_raygen(const FrameDev& F, hipStream_t s) {
  if (F.views) {
    hipLaunchKernelGGL(k_batch_round_begin, dim3(1), dim3(1), 0, s, F);
    hipLaunchKernelGGL(k_raygen_batch, dim3(cdiv(F.W * F.H * F.S, 256)), dim3(256), 0, s, F);
  } else {
    hipLaunchKernelGGL(k_raygen, dim3(cdiv(F.W * F.H * F.S, 256)), dim3(256), 0, s, F);
  }
}
// the scene as one launch sees it: the context's stack limit and the
// class's leaf phase trigger stamped in
static inline DevScene with_limit(const DevScene& S0, uint32_t stack_limit, uint32_t leaf16 = 8u) {
  DevScene S = S0;
  S.stack_limit = stack_limit;
  S.leaf16 = leaf16;
  return S;
}

// Calls f(kLds, kWpb, kNW) (integral constants) for the launch shape's
// variant: 8-wide nodes with 8, 12 or 16 LDS entries and 1, 2 or 4 waves per
// block, 4-wide nodes with 8 or 12 entries and one wave per block (the
// context validates shapes; anything else takes 12 entries, one wave).
template <int v>
using ic = std::integral_constant<int, v>;
template <class Fn>
static inline void with_variant(const TraceLaunch& L, Fn f) {
  const uint32_t w = L.waves_per_block;
  if (L.node_width == 4) {
    if (L.stack_lds == 8) f(ic<8>(), ic<1>(), ic<4>());
    else f(ic<12>(), ic<1>(), ic<4>());
    return;
  }
  switch (L.stack_lds) {
    case 8:
      if (w == 2) f(ic<8>(), ic<2>(), ic<8>());
      else if (w == 4) f(ic<8>(), ic<4>(), ic<8>());
      else f(ic<8>(), ic<1>(), ic<8>());
      break;
    case 16:
      if (w == 2) f(ic<16>(), ic<2>(), ic<8>());
      else if (w == 4) f(ic<16>(), ic<4>(), ic<8>());
      else f(ic<16>(), ic<1>(), ic<8>());
      break;
    default:
      if (w == 2) f(ic<12>(), ic<2>(), ic<8>());
      else if (w == 4) f(ic<12>(), ic<4>(), ic<8>());
      else f(ic<12>(), ic<1>(), ic<8>());
      break;
  }
}

void launch_trace_closest(const DevScene& S0, const FrameDev& F0, int bounce, const TraceLaunch& L, hipStream_t s) {
  const DevScene S = with_limit(S0, L.stack_limit, L.leaf16);
  FrameDev F = F0;
  F.spill = L.spill;
  const bool inst = S.num_inst != 0;
  if (L.count && inst) hipLaunchKernelGGL((k_trace_closest<true, true, kStackLds, 1>), dim3(L.fixed_grid), dim3(64), 0, s, S, F, bounce, L.refill);
  else if (L.count) hipLaunchKernelGGL((k_trace_closest<true, false, kStackLds, 1>), dim3(L.fixed_grid), dim3(64), 0, s, S, F, bounce, L.refill);
  else if (inst) hipLaunchKernelGGL((k_trace_closest<false, true, kStackLds, 1>), dim3(L.fixed_grid), dim3(64), 0, s, S, F, bounce, L.refill);
  else
    with_variant(L, [&](auto lds, auto wpb, auto nw) {
      constexpr int kL = decltype(lds)::value, kW = decltype(wpb)::value, kN = decltype(nw)::value;
      hipLaunchKernelGGL((k_trace_closest<false, false, kL, kW, kN>), dim3(L.grid / kW), dim3(64 * kW), 0, s, S,
                         F, bounce, L.refill);
    });
}
void launch_shade(const DevScene& S, const FrameDev& F, int bounce, hipStream_t s) {
  hipLaunchKernelGGL(k_shade, dim3(cdiv(F.hit_cap, kShadeWaves * 64)), dim3(kShadeWaves * 64), 0, s, S, F, bounce);
}
void launch_trace_shadow(const DevScene& S0, const FrameDev& F0, const TraceLaunch& L, hipStream_t s) {
  const DevScene S = with_limit(S0, L.stack_limit, L.leaf16);
  FrameDev F = F0;
  F.spill = L.spill;
  const bool inst = S.num_inst != 0;
  if (L.count && inst) hipLaunchKernelGGL((k_trace_shadow<true, true, kStackLds, 1>), dim3(L.fixed_grid), dim3(64), 0, s, S, F, L.refill);
  else if (L.count) hipLaunchKernelGGL((k_trace_shadow<true, false, kStackLds, 1>), dim3(L.fixed_grid), dim3(64), 0, s, S, F, L.refill);
  else if (inst) hipLaunchKernelGGL((k_trace_shadow<false, true, kStackLds, 1>), dim3(L.fixed_grid), dim3(64), 0, s, S, F, L.refill);
  else
    with_variant(L, [&](auto lds, auto wpb, auto nw) {
      constexpr int kL = decltype(lds)::value, kW = decltype(wpb)::value, kN = decltype(nw)::value;
      hipLaunchKernelGGL((k_trace_shadow<false, false, kL, kW, kN>), dim3(L.grid / kW), dim3(64 * kW), 0, s, S,
                         F, L.refill);
    });
}
void launch_render_paths(const DevScene& S0, const FrameDev& F0, uint32_t depth, const TraceLaunch& L,
                         hipStream_t s) {
  const DevScene S = with_limit(S0, L.stack_limit, L.leaf16);
  FrameDev F = F0;
  F.spill = L.spill;
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, F.cnt + kCntWorkTrace, 0u, 1u);
  // batches hold no instances (pk_scene_upload_batch rejects them)
  if (F.views) hipLaunchKernelGGL((k_render_paths<false, true>), dim3(L.grid), dim3(64), 0, s, S, F, depth);
  else if (S.num_inst) hipLaunchKernelGGL((k_render_paths<true, false>), dim3(L.grid), dim3(64), 0, s, S, F, depth);
  else hipLaunchKernelGGL((k_render_paths<false, false>), dim3(L.grid), dim3(64), 0, s, S, F, depth);
}
void launch_accumulate(const FrameDev& F, hipStream_t s) {
  hipLaunchKernelGGL(k_accumulate, dim3(cdiv(F.W * F.H, kAccPix)), dim3(256), 0, s, F);
}
void launch_build_shade_records(const int4* tri, const uint4* attrs, uint4* rec, uint32_t n, hipStream_t s) {
  if (n) hipLaunchKernelGGL(k_build_shade_records, dim3(cdiv(n, 256)), dim3(256), 0, s, tri, attrs, rec, n);
}
void launch_fill_u32(uint32_t* p, uint32_t v, uint32_t n, hipStream_t s) {
  if (n) hipLaunchKernelGGL(k_fill_u32, dim3(cdiv(n, 256)), dim3(256), 0, s, p, v, n);
}
void launch_fill_f4(float4* p, float4 v, uint32_t n, hipStream_t s) {
  if (n) hipLaunchKernelGGL(k_fill_f4, dim3(cdiv(n, 256)), dim3(256), 0, s, p, v, n);
}
void launch_copy_f4(float4* dst, const float4* src, uint32_t n, hipStream_t s) {
  if (n) hipLaunchKernelGGL(k_copy_f4, dim3(min(cdiv(n, 256), 8192u)), dim3(256), 0, s, dst, src, n);
}
void launch_adjoint(const DevScene& S0, const AdjointArgs& A, uint32_t grid, uint32_t stack_limit, hipStream_t s) {
  const DevScene S = with_limit(S0, stack_limit);
  if (S.num_inst) hipLaunchKernelGGL(k_adjoint<true>, dim3(grid), dim3(64), 0, s, S, A);
  else hipLaunchKernelGGL(k_adjoint<false>, dim3(grid), dim3(64), 0, s, S, A);
}
void launch_loss_l2(const float4* img, const float4* target, float4* grad, double* loss, uint32_t n, hipStream_t s) {
  if (n) hipLaunchKernelGGL(k_loss_l2, dim3(cdiv(n, 256)), dim3(256), 0, s, img, target, grad, loss, n);
}
void launch_trace_rays(const DevScene& S0, const float4* rays, uint32_t n, int any_hit, int32_t* ids, float* t,
                       uint32_t* bary, uint32_t* cursor, uint2* spill, unsigned long long* stats,
                       uint32_t* overflow, const TraceLaunch& L, hipStream_t s) {
  const DevScene S = with_limit(S0, L.stack_limit, L.leaf16);
#define PK_LAUNCH_RAYS(A, C, I)                                                                                    \
  hipLaunchKernelGGL((k_trace_rays<A, C, I, kStackLds, 1>), dim3(L.fixed_grid), dim3(64), 0, s, S, rays, n, ids, t, bary,   \
                     cursor, spill, stats, overflow, L.refill)
  const int variant = (any_hit ? 4 : 0) | (L.count ? 2 : 0) | (S.num_inst ? 1 : 0);
  switch (variant) {
    case 0:
      with_variant(L, [&](auto lds, auto wpb, auto nw) {
        constexpr int kL = decltype(lds)::value, kW = decltype(wpb)::value, kN = decltype(nw)::value;
        hipLaunchKernelGGL((k_trace_rays<false, false, false, kL, kW, kN>), dim3(L.grid / kW), dim3(64 * kW), 0,
                           s, S, rays, n, ids, t, bary, cursor, spill, stats, overflow, L.refill);
      });
      break;
    case 1: PK_LAUNCH_RAYS(false, false, true); break;
    case 2: PK_LAUNCH_RAYS(false, true, false); break;
    case 3: PK_LAUNCH_RAYS(false, true, true); break;
    case 4:
      with_variant(L, [&](auto lds, auto wpb, auto nw) {
        constexpr int kL = decltype(lds)::value, kW = decltype(wpb)::value, kN = decltype(nw)::value;
        hipLaunchKernelGGL((k_trace_rays<true, false, false, kL, kW, kN>), dim3(L.grid / kW), dim3(64 * kW), 0,
                           s, S, rays, n, ids, t, bary, cursor, spill, stats, overflow, L.refill);
      });
      break;
    case 5: PK_LAUNCH_RAYS(true, false, true); break;
    case 6: PK_LAUNCH_RAYS(true, true, false); break;
    default: PK_LAUNCH_RAYS(true, true, true); break;
  }
#undef PK_LAUNCH_RAYS
}

}  // namespace pk
