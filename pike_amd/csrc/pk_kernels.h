// pk_kernels.h — device frame state and the host launch wrappers of
// pk_kernels.hip (the runtime in pk_runtime.hip calls only these).
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/pk.h"
#include "pk_device.h"

namespace pk {

// Device counters (uint32), each in a 4 KB page of its own: atomics and
// loads on one cache line serialize at the L2, so a queue length that a
// kernel's waves read must not share a line with the counters another queue's
// appends are hammering.
constexpr int kCounterStride = 1024;  // words
enum : int {
  kCntRays0 = 0 * kCounterStride,      // ray queue 0 length
  kCntRays1 = 1 * kCounterStride,      // ray queue 1 length
  kCntHits = 2 * kCounterStride,       // hit queue length (records and holes)
  kCntShadow = 3 * kCounterStride,     // shadow queue length
  kCntWorkTrace = 4 * kCounterStride,  // trace_closest work cursor
  kCntWorkShadow = 5 * kCounterStride, // trace_shadow work cursor
  kCntOverflow = 6 * kCounterStride,   // stack overflow flag
  kNumCounters = 7 * kCounterStride
};
// trace_closest reserves the hit queue in runs of kHitRun records per wave
// (one atomic per run instead of one per refill); the unused end of a
// wave's last run is filled with holes, so the queue holds at most
// kHitRun - 1 holes per wave of the launch beyond the hits.
constexpr uint32_t kHitRun = 256;
// counter of ray queue q (0 or 1)
__host__ __device__ inline int ray_counter(int q) { return q ? kCntRays1 : kCntRays0; }
// device statistics (uint64)
enum : int {
  kStatRays = 0,
  kStatShadow = 1,
  kStatHits = 2,
  kStatNodes = 3,
  kStatTris = 4,
  kStatFrames = 5,
  kStatShadowNodes = 6,
  kStatShadowTris = 7,
  kNumStats = 8
};

struct FrameParams {
  uint32_t sample;  // first sample of the current pass (RenderState.samples)
  float env_theta;
  pk_camera cam;
  uint32_t limit;   // samples >= limit are not traced (end of a pk_render run)
  float cbx[3], cby[3];  // the camera's basis (camera_basis), formed once by k_begin
};

// One camera view of a batch: its pixels occupy [offset, offset + W*H) of
// the batch's global path index space; rays of this view start at `root`,
// the wide-BVH root of its scene inside the merged scene.
struct ViewDev {
  uint32_t W, H, offset, spp;
  uint32_t root, scene, pad0, pad1;
  pk_camera cam;
};

struct FrameDev {
  float4* qo[2];   // ray queues: origin.xyz, pixel id bits
  float4* qd[2];   //             direction.xyz
  float4* qt[2];   //             path throughput (bounce rays; camera rays carry 1 and have none)
  uint4* hits;     // {ray slot, leaf triangle, t bits, bary mask}; x = ~0: a hole
  float4* so;      // shadow queue: origin.xyz, pixel id bits
  float4* sd;      //               direction.xyz
  float4* sthr;    //               NEE throughput
  float4* color;   // RenderState.colorBuffer (rgb sum, contribution count)
  float4* accum;   // running mean
  uint32_t* cnt;
  unsigned long long* stats;
  FrameParams* params;
  uint2* spill;    // traversal stack spill, per wave
  uint32_t W, H;   // single view: the resolution; batch: W = total paths, H = 1
  const ViewDev* views;  // batch views sorted by offset (nullptr: single view)
  uint32_t nviews;
  // Samples per pass: a single-view pass traces samples sample .. sample+S-1
  // of every pixel as one wavefront, path id = pixel * S + s (batches use 1).
  uint32_t S;
  uint32_t hit_cap;  // records the hit queue holds (paths + holes)
};

// samples the current pass traces (fewer than S at the end of a run)
__device__ inline uint32_t pass_samples(const FrameDev& F, const FrameParams& P) {
  return P.limit > P.sample ? min(F.S, P.limit - P.sample) : 0u;
}

// view of global path index p (batch mode): views are sorted by offset
__device__ inline uint32_t view_of(const FrameDev& F, uint32_t p) {
  uint32_t lo = 0, hi = F.nviews;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (F.views[mid].offset <= p) lo = mid;
    else hi = mid;
  }
  return lo;
}

void launch_begin(const FrameDev& F, uint32_t sample, float theta, const pk_camera& cam, hipStream_t s,
                  uint32_t limit = 0xffffffffu);
void launch_advance(const FrameDev& F, hipStream_t s);
void launch_raygen(const FrameDev& F, hipStream_t s);
// persistent trace launch shape: blocks of one wave, lane refill threshold
// grid counts waves (grid / waves_per_block blocks); the frame and ray-query
// traversal kernels come in stack_lds = 8, 12 or 16 LDS entries and 1, 2 or
// 4 waves per block (the counting and instanced variants: 12 and 1)
struct TraceLaunch {
  uint32_t grid;        // waves of the shape's variant (a multiple of waves_per_block)
  // waves of the counting and instanced variants (kStackLds entries, one
  // wave per block, 6 waves per SIMD): the shape's waves per CU capped at the
  // 24 those keep resident, so no wave of a persistent launch waits for a slot
  uint32_t fixed_grid;
  uint32_t refill;
  bool count;
  uint2* spill;  // grid * kSpillStride * 64 entries
  uint32_t stack_limit;  // traversal stack entries per ray, in [1, kStackMax]
  uint32_t stack_lds;
  uint32_t waves_per_block;
  uint32_t leaf16;       // leaf phase trigger in 16ths of the tracing lanes
  uint32_t node_width;   // 8 (WNode) or 4 (WNode4, single-level scenes)
};

void launch_trace_closest(const DevScene& S, const FrameDev& F, int bounce, const TraceLaunch& L, hipStream_t s);
void launch_shade(const DevScene& S, const FrameDev& F, int bounce, hipStream_t s);
void launch_trace_shadow(const DevScene& S, const FrameDev& F, const TraceLaunch& L, hipStream_t s);
void launch_accumulate(const FrameDev& F, hipStream_t s);
// Single-view pass as one persistent path megakernel (k_render_paths): the
// same color buffer as raygen + depth x (trace_closest, shade, trace_shadow).
// Waves per SIMD its registers are compiled for; L.grid must not exceed
// 4 * kMegaWavesPerSimd one-wave blocks per CU (the spill area's size).
constexpr int kMegaWavesPerSimd = 4;
void launch_render_paths(const DevScene& S, const FrameDev& F, uint32_t depth, const TraceLaunch& L, hipStream_t s);
// DevScene::shade_rec from leaf-order {i0, i1, i2, mat} and vertex attributes
// (every index already checked against the attribute count).
void launch_build_shade_records(const int4* tri, const uint4* attrs, uint4* rec, uint32_t n, hipStream_t s);
void launch_fill_u32(uint32_t* p, uint32_t v, uint32_t n, hipStream_t s);
void launch_fill_f4(float4* p, float4 v, uint32_t n, hipStream_t s);
void launch_copy_f4(float4* dst, const float4* src, uint32_t n, hipStream_t s);
// Path-replay adjoint of a frame's image (k_adjoint).
// words a path vertex's gradient terms occupy in AdjointArgs::vtx
constexpr int kAdjVtxWords = 20;
struct AdjointArgs {
  const float4* d_image;  // dLoss/dImage per pixel (rgb)
  float* g_atlas;         // gradient in the atlas layout [layer][y][x][4]
  float* g_env;           // gradient in the environment layout [y][x][4]
  uint32_t W, H, first_sample, num_frames, depth;
  float inv_frames, theta;
  pk_camera cam;
  uint32_t* cursor;       // zeroed before the launch
  uint2* spill;           // grid * kSpillStride * 64 entries
  float* vtx;             // grid * depth * kAdjVtxWords * 64 floats
  uint32_t* overflow;
};
void launch_adjoint(const DevScene& S, const AdjointArgs& A, uint32_t grid, uint32_t stack_limit, hipStream_t s);
void launch_loss_l2(const float4* img, const float4* target, float4* grad, double* loss, uint32_t n, hipStream_t s);
void launch_trace_rays(const DevScene& S, const float4* rays, uint32_t n, int any_hit, int32_t* ids, float* t,
                       uint32_t* bary, uint32_t* cursor, uint2* spill, unsigned long long* stats, uint32_t* overflow,
                       const TraceLaunch& L, hipStream_t s);

}  // namespace pk
