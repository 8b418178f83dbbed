// pk_runtime.hip — the C ABI of include/pk.h: context, scene upload, frame
// state, per-pass entry points, whole-frame rendering (optionally replayed
// through one captured hipGraph) and standalone ray queries.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "pk_internal.h"
#include "pk_kernels.h"

namespace {
thread_local std::string g_err;

// Scratch of pk_trace_rays_noalloc: the work cursor (first 256 bytes) and
// the traversal stack spill area of up to kNoAllocMaxWaves one-wave blocks
// (8 per CU on a 256-CU MI355X), 58.7 MB, held by every device the library
// runs on until the process ends.  Host code never uses the array's name as
// an address: pk_ctx_create resolves each device's copy once with
// hipGetSymbolAddress, and only that address is passed to kernels.
constexpr uint32_t kNoAllocMaxWaves = 2048;
constexpr size_t kNoAllocCursorBytes = 256;
constexpr size_t kNoAllocScratchBytes =
    kNoAllocCursorBytes + (size_t)kNoAllocMaxWaves * pk::kSpillStride * 64 * sizeof(uint2);
__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on `device` (current), resolved once.
int scratch_address(int device, unsigned char** out) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (device < 0 || device >= kMaxDevices) {
    g_err = "device index beyond the scratch table";
    return PK_E_LIMIT;
  }
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      g_err = std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e);
      return PK_E_HIP;
    }
    g_scratch_ptr[device] = static_cast<unsigned char*>(p);
  }
  *out = g_scratch_ptr[device];
  return PK_OK;
}
}  // namespace

void pk_set_error(const std::string& msg) { g_err = msg; }

#define PK_HIP(call)                                                                                 \
  do {                                                                                                 \
    hipError_t e_ = (call);                                                                            \
    if (e_ != hipSuccess) {                                                                            \
      pk_set_error(std::string(#call) + " failed: " + hipGetErrorString(e_));                          \
      return e_ == hipErrorOutOfMemory ? PK_E_OOM : PK_E_HIP;                                          \
    }                                                                                                  \
  } while (0)

struct pk_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int num_cus = 0;
  struct Shape {
    uint32_t waves;   // persistent traversal waves per CU
    uint32_t refill;  // idle lanes that trigger a wave-level ray refill
    uint32_t lds = pk::kStackLds;  // LDS stack entries per lane (8, 12, 16)
    uint32_t wpb = 1;              // waves per block (1, 2, 4)
    uint32_t leaf16 = 8;           // leaf phase trigger, 16ths of the tracing lanes
    uint32_t width = 8;            // BVH node width (8, or 4 with 8/12 entries and one wave per block)
  } shape[PK_TUNE_CLASSES];
  uint32_t tuning_gen = 0;  // bumped by pk_ctx_set_tuning; captured graphs keep the old shape
  uint32_t stack_limit = pk::kStackMax;  // traversal stack entries per ray
  // device memory of this context's scenes, frames and their scratch:
  // bytes in use, the cap (0: none) and each allocation's size
  size_t mem_used = 0, mem_cap = 0;
  std::unordered_map<void*, size_t> mem_sizes;
  int sched = PK_SCHED_WAVEFRONT;
  uint32_t* cursors = nullptr;  // per-launch work cursors for ray queries
  uint32_t ncursors = 0;
  uint2* spill = nullptr;
  unsigned long long* stats = nullptr;
  uint32_t* overflow = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  unsigned char* scratch = nullptr;  // this device's pk_trace_rays_noalloc scratch
};

// Device allocations of one owner (a scene, a frame, one call's
// temporaries), counted against the owning context's memory cap.
struct AllocList {
  pk_ctx* ctx = nullptr;
  std::vector<void*> ptrs;
};

struct pk_scene {
  pk_ctx* ctx = nullptr;
  pk::DevScene dev{};
  AllocList allocs;
  int64_t env_info[4] = {0, 0, 0, 0};
  std::vector<uint32_t> roots{0};  // wide-BVH root of each merged sub-scene
  float* tex_eval = nullptr;       // evaluation atlas, allocated by the first parameter write
  uint32_t gen = 0;                // bumped when `dev` changes; captured graphs hold a copy of it
};

struct pk_frame {
  pk_ctx* ctx = nullptr;
  uint32_t W = 0, H = 0;
  pk::FrameDev dev{};
  AllocList allocs;
  AllocList path_allocs;  // per-path queues, sized W * H * samples per pass
  pk_camera cam{};
  bool has_camera = false;
  float theta = 0;
  bool begun = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  const pk_scene* graph_scene = nullptr;
  uint32_t graph_scene_gen = 0;
  uint32_t graph_depth = 0;
  uint32_t graph_gen = 0;
  float graph_theta = 0;
  bool count_nodes = false;
  // batch frames: views over one merged scene, paths laid out view after view
  std::vector<pk::ViewDev> views;
  pk_scene* batch_scene = nullptr;
  // adjoint scratch, allocated on first use
  float4* adj_dimage = nullptr;
  float* adj_atlas = nullptr;
  size_t adj_atlas_n = 0;
  float* adj_env = nullptr;
  size_t adj_env_n = 0;
  uint32_t* adj_cursor = nullptr;
  float* adj_vtx = nullptr;  // per-wave path vertex records
  size_t adj_vtx_n = 0;
  double* loss = nullptr;
};

namespace {

// Waves per CU at most of a persistent traversal launch; the spill area and
// the hit queue's slack are sized for it.  The wavefront kernels' LDS (an
// 8-entry stack part and the leaf phase's 1 KB per wave) fits 32 per CU, the
// path megakernel and the adjoint launch fewer.
constexpr uint32_t kMaxWavesPerCu = 40;
constexpr uint32_t kLdsBytesPerCu = 160u * 1024u;
// waves per CU that a shape's LDS fits: whole blocks of wpb waves, each with
// its stack part (lds entries x 64 lanes x 8 bytes) and the leaf phase's
// 128 words of 8 bytes
uint32_t lds_max_waves(uint32_t lds, uint32_t wpb) {
  const uint32_t per_block = (lds * 64u * 8u + 128u * 8u) * wpb;
  return std::min(kMaxWavesPerCu, kLdsBytesPerCu / per_block * wpb);
}
// One-wave adjoint blocks per CU: k_adjoint is compiled for 4 waves per SIMD
// (128 VGPRs, the rest spilled to scratch), which measured 1.35x faster on
// the 512x512 differentiation workload than the 2 waves per SIMD its
// registers need unspilled.
constexpr uint32_t kAdjointWavesPerCu = 16;
static_assert(kAdjointWavesPerCu <= kMaxWavesPerCu, "adjoint grid exceeds the spill area");
static_assert(4 * pk::kMegaWavesPerSimd <= kMaxWavesPerCu, "path megakernel grid exceeds the spill area");

pk::TraceLaunch shape_launch(const pk_ctx* c, int cls, bool count) {
  const pk_ctx::Shape& sh = c->shape[cls];
  // whole blocks: the grid (in waves) is a multiple of the waves per block
  const uint32_t grid = std::max(sh.wpb, (uint32_t)c->num_cus * sh.waves / sh.wpb * sh.wpb);
  const uint32_t fixed_grid =
      (uint32_t)c->num_cus * std::min({sh.waves, 4u * pk::kTraceWavesPerSimd, lds_max_waves(pk::kStackLds, 1)});
  return pk::TraceLaunch{grid,   fixed_grid, sh.refill, count,     c->spill,
                         c->stack_limit, sh.lds, sh.wpb, sh.leaf16, sh.width};
}
pk::TraceLaunch trace_launch(const pk_frame* f, int cls) { return shape_launch(f->ctx, cls, f->count_nodes); }
int closest_class(uint32_t bounce) { return bounce == 0 ? PK_TUNE_PRIMARY : PK_TUNE_SECONDARY; }

template <class T> int dev_alloc(AllocList& list, T** p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  const size_t bytes = count * sizeof(T);
  pk_ctx* c = list.ctx;
  if (c && c->mem_cap && c->mem_used + bytes > c->mem_cap) {
    pk_set_error("device memory cap of " + std::to_string(c->mem_cap) + " bytes reached: " +
                 std::to_string(c->mem_used) + " in use, " + std::to_string(bytes) + " requested");
    return PK_E_OOM;
  }
  hipError_t e = hipMalloc((void**)p, bytes);
  if (e != hipSuccess) {
    pk_set_error(std::string("hipMalloc of ") + std::to_string(bytes) + " bytes failed: " + hipGetErrorString(e));
    *p = nullptr;
    return PK_E_OOM;
  }
  list.ptrs.push_back(*p);
  if (c) {
    c->mem_used += bytes;
    c->mem_sizes[*p] = bytes;
  }
  return PK_OK;
}

template <class T> int upload(AllocList& list, T** p, const void* src, size_t count, hipStream_t s) {
  int rc = dev_alloc(list, p, count);
  if (rc) return rc;
  if (src && count) PK_HIP(hipMemcpyAsync(*p, src, count * sizeof(T), hipMemcpyHostToDevice, s));
  return PK_OK;
}

void free_all(AllocList& list) {
  pk_ctx* c = list.ctx;
  for (void* p : list.ptrs) {
    (void)hipFree(p);
    if (c) {
      auto it = c->mem_sizes.find(p);
      if (it != c->mem_sizes.end()) {
        c->mem_used -= it->second;
        c->mem_sizes.erase(it);
      }
    }
  }
  list.ptrs.clear();
}

int ensure_cursors(pk_ctx* c, uint32_t n) {
  if (c->ncursors >= n) return PK_OK;
  if (c->cursors) (void)hipFree(c->cursors);
  c->cursors = nullptr;
  PK_HIP(hipMalloc((void**)&c->cursors, (size_t)n * sizeof(uint32_t)));
  c->ncursors = n;
  return PK_OK;
}

}  // namespace

extern "C" {

const char* pk_last_error(void) { return g_err.c_str(); }
const char* pk_version(void) { return "pike_amd 0.1 (gfx950)"; }

int pk_ctx_create(int device, pk_ctx** out) {
  return pk_guard("pk_ctx_create", [&]() -> int {
    if (!out) {
      pk_set_error("pk_ctx_create: null output");
      return PK_E_INVALID;
    }
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
      pk_set_error("pk_ctx_create: no HIP device is visible");
      return PK_E_NO_DEVICE;
    }
    if (device < 0 || device >= n) {
      pk_set_error("pk_ctx_create: device index out of range");
      return PK_E_INVALID;
    }
    PK_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    PK_HIP(hipGetDeviceProperties(&prop, device));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
      pk_set_error(std::string("pk_ctx_create: device is ") + prop.gcnArchName + ", this library is built for gfx950");
      return PK_E_NO_DEVICE;
    }
    auto* c = new pk_ctx();
    c->device = device;
    c->num_cus = prop.multiProcessorCount;
    // Defaults measured on MI355X (1080p, 3.1M triangles, depth 8): the
    // 8-entry LDS stack part at 32 waves per CU (the traversal kernels fit 64
    // VGPRs, 8 waves per SIMD); camera rays keep whole waves together (refill
    // 64), bounce rays refill at half a wave, shadow rays and incoherent ray
    // queries early.
    c->shape[PK_TUNE_PRIMARY] = {32, 64, 8, 1, 8};
    c->shape[PK_TUNE_SECONDARY] = {32, 32, 8, 1, 8};
    c->shape[PK_TUNE_SHADOW] = {32, 8, 8, 1, 8};
    c->shape[PK_TUNE_RAYS_CLOSEST] = {32, 16, 8, 1, 8};
    c->shape[PK_TUNE_RAYS_ANY] = {32, 24, 8, 1, 8};
    c->shape[PK_TUNE_PATHS] = {4 * pk::kMegaWavesPerSimd, 64};
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess)
      e = hipMalloc((void**)&c->spill, (size_t)c->num_cus * kMaxWavesPerCu * pk::kSpillStride * 64 * sizeof(uint2));
    if (e == hipSuccess) e = hipMalloc((void**)&c->stats, 8 * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMalloc((void**)&c->overflow, 4 * sizeof(uint32_t));
    if (e == hipSuccess) e = hipEventCreate(&c->ev0);
    if (e == hipSuccess) e = hipEventCreate(&c->ev1);
    if (e != hipSuccess) {
      pk_set_error(std::string("pk_ctx_create: ") + hipGetErrorString(e));
      pk_ctx_destroy(c);
      return PK_E_HIP;
    }
    int rc = ensure_cursors(c, 64);
    if (!rc) rc = scratch_address(device, &c->scratch);
    if (rc) {
      pk_ctx_destroy(c);
      return rc;
    }
    *out = c;
    return PK_OK;
  });
}

int pk_ctx_destroy(pk_ctx* c) {
  return pk_guard("pk_ctx_destroy", [&]() -> int {
    if (!c) return PK_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->spill) (void)hipFree(c->spill);
    if (c->stats) (void)hipFree(c->stats);
    if (c->overflow) (void)hipFree(c->overflow);
    if (c->cursors) (void)hipFree(c->cursors);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return PK_OK;
  });
}

int pk_sync(pk_ctx* c) {
  return pk_guard("pk_sync", [&]() -> int {
    if (!c) {
      pk_set_error("pk_sync: null context");
      return PK_E_INVALID;
    }
    PK_HIP(hipStreamSynchronize(c->stream));
    return PK_OK;
  });
}

// largest waves_per_cu a class allows with the given stack part and block
static uint32_t class_max_waves(int cls, uint32_t lds, uint32_t wpb) {
  return cls == PK_TUNE_PATHS ? 4u * pk::kMegaWavesPerSimd : lds_max_waves(lds, wpb);
}

int pk_ctx_set_launch(pk_ctx* c, int cls, const pk_launch_shape* sh) {
  return pk_guard("pk_ctx_set_launch", [&]() -> int {
    if (!c || !sh || cls < 0 || cls >= PK_TUNE_CLASSES) {
      pk_set_error("pk_ctx_set_launch: null argument or unknown kernel class");
      return PK_E_INVALID;
    }
    const bool lds_ok = sh->stack_lds == 8 || sh->stack_lds == 12 || sh->stack_lds == 16;
    const bool wpb_ok = sh->waves_per_block == 1 || sh->waves_per_block == 2 || sh->waves_per_block == 4;
    const bool width_ok = sh->node_width == 8 ||
                          (sh->node_width == 4 && sh->stack_lds != 16 && sh->waves_per_block == 1);
    if (!lds_ok || !wpb_ok || !width_ok ||
        (cls == PK_TUNE_PATHS && (sh->stack_lds != 12 || sh->waves_per_block != 1 || sh->node_width != 8))) {
      pk_set_error("pk_ctx_set_launch: stack_lds must be 8, 12 or 16, waves_per_block 1, 2 or 4 and node_width 8, "
                   "or 4 with stack_lds 8 or 12 and 1 wave per block (the path megakernel: 12, 1 and 8)");
      return PK_E_INVALID;
    }
    const uint32_t wmax = class_max_waves(cls, sh->stack_lds, sh->waves_per_block);
    if (sh->waves_per_cu < sh->waves_per_block || sh->waves_per_cu > wmax || sh->refill_min < 1 ||
        sh->refill_min > 64 || sh->leaf_phase16 < 1 || sh->leaf_phase16 > 16) {
      pk_set_error("pk_ctx_set_launch: waves_per_cu must be in [waves_per_block, " + std::to_string(wmax) +
                   "] for this stack part and block, refill_min in [1, 64], leaf_phase16 in [1, 16]");
      return PK_E_LIMIT;
    }
    c->shape[cls] = {sh->waves_per_cu, sh->refill_min, sh->stack_lds, sh->waves_per_block, sh->leaf_phase16,
                     sh->node_width};
    c->tuning_gen++;
    return PK_OK;
  });
}

int pk_ctx_get_launch(const pk_ctx* c, int cls, pk_launch_shape* sh) {
  return pk_guard("pk_ctx_get_launch", [&]() -> int {
    if (!c || !sh || cls < 0 || cls >= PK_TUNE_CLASSES) {
      pk_set_error("pk_ctx_get_launch: null argument or unknown kernel class");
      return PK_E_INVALID;
    }
    const pk_ctx::Shape& s = c->shape[cls];
    *sh = pk_launch_shape{s.waves, s.refill, s.lds, s.wpb, s.leaf16, s.width};
    return PK_OK;
  });
}

int pk_ctx_set_tuning(pk_ctx* c, int cls, uint32_t waves_per_cu, uint32_t refill_min) {
  return pk_guard("pk_ctx_set_tuning", [&]() -> int {
    if (!c || cls < 0 || cls >= PK_TUNE_CLASSES) {
      pk_set_error("pk_ctx_set_tuning: null context or unknown kernel class");
      return PK_E_INVALID;
    }
    pk_launch_shape sh{waves_per_cu, refill_min, c->shape[cls].lds, c->shape[cls].wpb, c->shape[cls].leaf16,
                       c->shape[cls].width};
    return pk_ctx_set_launch(c, cls, &sh);
  });
}

int pk_ctx_set_scheduler(pk_ctx* c, int sched) {
  return pk_guard("pk_ctx_set_scheduler", [&]() -> int {
    if (!c || (sched != PK_SCHED_WAVEFRONT && sched != PK_SCHED_PATHS)) {
      pk_set_error("pk_ctx_set_scheduler: null context or unknown scheduler");
      return PK_E_INVALID;
    }
    if (sched != c->sched) c->tuning_gen++;
    c->sched = sched;
    return PK_OK;
  });
}

int pk_ctx_get_scheduler(const pk_ctx* c, int* sched) {
  return pk_guard("pk_ctx_get_scheduler", [&]() -> int {
    if (!c || !sched) {
      pk_set_error("pk_ctx_get_scheduler: null argument");
      return PK_E_INVALID;
    }
    *sched = c->sched;
    return PK_OK;
  });
}

int pk_ctx_get_tuning(const pk_ctx* c, int cls, uint32_t* waves_per_cu, uint32_t* refill_min, uint32_t* num_cus) {
  return pk_guard("pk_ctx_get_tuning", [&]() -> int {
    if (!c || cls < 0 || cls >= PK_TUNE_CLASSES) {
      pk_set_error("pk_ctx_get_tuning: null context or unknown kernel class");
      return PK_E_INVALID;
    }
    if (waves_per_cu) *waves_per_cu = c->shape[cls].waves;
    if (refill_min) *refill_min = c->shape[cls].refill;
    if (num_cus) *num_cus = (uint32_t)c->num_cus;
    return PK_OK;
  });
}

int pk_ctx_set_stack_limit(pk_ctx* c, uint32_t entries) {
  return pk_guard("pk_ctx_set_stack_limit", [&]() -> int {
    if (!c || entries < 1 || entries > (uint32_t)pk::kStackMax) {
      pk_set_error("pk_ctx_set_stack_limit: null context or entries outside [1, " + std::to_string(pk::kStackMax) +
                   "]");
      return PK_E_INVALID;
    }
    if (entries != c->stack_limit) ++c->tuning_gen;  // captured graphs hold the old limit
    c->stack_limit = entries;
    return PK_OK;
  });
}

int pk_ctx_get_stack_limit(const pk_ctx* c, uint32_t* entries) {
  return pk_guard("pk_ctx_get_stack_limit", [&]() -> int {
    if (!c || !entries) {
      pk_set_error("pk_ctx_get_stack_limit: null argument");
      return PK_E_INVALID;
    }
    *entries = c->stack_limit;
    return PK_OK;
  });
}

int pk_ctx_set_memory_limit(pk_ctx* c, uint64_t bytes) {
  return pk_guard("pk_ctx_set_memory_limit", [&]() -> int {
    if (!c) {
      pk_set_error("pk_ctx_set_memory_limit: null context");
      return PK_E_INVALID;
    }
    c->mem_cap = (size_t)bytes;
    return PK_OK;
  });
}

int pk_ctx_memory_used(const pk_ctx* c, uint64_t* bytes) {
  return pk_guard("pk_ctx_memory_used", [&]() -> int {
    if (!c || !bytes) {
      pk_set_error("pk_ctx_memory_used: null argument");
      return PK_E_INVALID;
    }
    *bytes = (uint64_t)c->mem_used;
    return PK_OK;
  });
}

// ---------------------------------------------------------------------------
// device buffers
// ---------------------------------------------------------------------------
int pk_buffer_create(pk_ctx* c, size_t bytes, void** p) {
  return pk_guard("pk_buffer_create", [&]() -> int {
    if (!c || !p) {
      pk_set_error("pk_buffer_create: null argument");
      return PK_E_INVALID;
    }
    *p = nullptr;
    PK_HIP(hipSetDevice(c->device));
    PK_HIP(hipMalloc(p, bytes ? bytes : 1));
    return PK_OK;
  });
}

int pk_buffer_destroy(pk_ctx* c, void* p) {
  return pk_guard("pk_buffer_destroy", [&]() -> int {
    if (!c) {
      pk_set_error("pk_buffer_destroy: null context");
      return PK_E_INVALID;
    }
    if (p) {
      PK_HIP(hipStreamSynchronize(c->stream));
      PK_HIP(hipFree(p));
    }
    return PK_OK;
  });
}

int pk_buffer_write(pk_ctx* c, void* dst, const void* src, size_t bytes) {
  return pk_guard("pk_buffer_write", [&]() -> int {
    if (!c || (bytes && (!dst || !src))) {
      pk_set_error("pk_buffer_write: null argument");
      return PK_E_INVALID;
    }
    if (bytes) {
      PK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
      PK_HIP(hipStreamSynchronize(c->stream));
    }
    return PK_OK;
  });
}

int pk_buffer_read(pk_ctx* c, void* dst, const void* src, size_t bytes) {
  return pk_guard("pk_buffer_read", [&]() -> int {
    if (!c || (bytes && (!dst || !src))) {
      pk_set_error("pk_buffer_read: null argument");
      return PK_E_INVALID;
    }
    if (bytes) {
      PK_HIP(hipStreamSynchronize(c->stream));
      PK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
      PK_HIP(hipStreamSynchronize(c->stream));
    }
    return PK_OK;
  });
}

int pk_buffer_copy(pk_ctx* c, void* dst, const void* src, size_t bytes) {
  return pk_guard("pk_buffer_copy", [&]() -> int {
    if (!c || (bytes && (!dst || !src))) {
      pk_set_error("pk_buffer_copy: null argument");
      return PK_E_INVALID;
    }
    if (bytes % 16 == 0 && ((uintptr_t)dst % 16) == 0 && ((uintptr_t)src % 16) == 0 && bytes / 16 < (1ull << 32)) {
      // a plain streaming copy kernel (also the reference point for HBM byte counters)
      pk::launch_copy_f4((float4*)dst, (const float4*)src, (uint32_t)(bytes / 16), c->stream);
      PK_HIP(hipGetLastError());
    } else if (bytes) {
      PK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c->stream));
    }
    return PK_OK;
  });
}

// ---------------------------------------------------------------------------
// scene
// ---------------------------------------------------------------------------
int pk_scene_upload(pk_ctx* c, const pk_scene_desc* d, const pk_bvh* bvh, pk_scene** out) {
  return pk_guard("pk_scene_upload", [&]() -> int {
    if (!c || !d || !bvh || !out) {
      pk_set_error("pk_scene_upload: null argument");
      return PK_E_INVALID;
    }
    *out = nullptr;
    if (d->num_triangles != bvh->nt || d->num_vertices != bvh->nv) {
      pk_set_error("pk_scene_upload: the BVH was built for a different triangle/vertex count");
      return PK_E_INVALID;
    }
    const uint32_t nm = d->materials ? d->num_materials : 1;
    const uint32_t layers = d->tex_data ? d->tex_layers : 1;
    if (d->tex_data && (d->tex_width == 0 || d->tex_height == 0 || d->tex_layers == 0)) {
      pk_set_error("pk_scene_upload: texture atlas has a zero dimension");
      return PK_E_INVALID;
    }
    if (d->tex_data && d->tex_format != PK_TEX_RGBA8 && d->tex_format != PK_TEX_RGBA32F) {
      pk_set_error("pk_scene_upload: unknown texture format");
      return PK_E_INVALID;
    }
    for (uint32_t t = 0; t < d->num_triangles; ++t) {
      int32_t m = d->triangles[(size_t)t * 4 + 3];
      if (d->materials && (m < 0 || (uint32_t)m >= nm)) {
        pk_set_error("pk_scene_upload: triangle " + std::to_string(t) + " has material id out of range");
        return PK_E_INVALID;
      }
    }
    if (bvh->instanced) {
      // mesh triangles and instance overrides must name existing materials
      for (size_t i = 3; i < bvh->wtri_index.size(); i += 4)
        if (bvh->wtri_index[i] < 0 || (uint32_t)bvh->wtri_index[i] >= nm) {
          pk_set_error("pk_scene_upload: a mesh triangle has material id out of range");
          return PK_E_INVALID;
        }
      for (size_t i = 0; i < bvh->inst.size(); ++i)
        if (bvh->inst[i].material >= (int32_t)nm) {
          pk_set_error("pk_scene_upload: instance material " + std::to_string(bvh->inst[i].material) + " out of range");
          return PK_E_INVALID;
        }
    }
    {
      // the traversal finds children from ibase/tbase (and instmask) alone
      for (size_t i = 0; i < bvh->wnodes.size(); ++i)
        if (!pk::wnode_linked(bvh->wnodes[i])) {
          pk_set_error("pk_scene_upload: wide node " + std::to_string(i) +
                       " does not keep its children in slot-order runs from ibase/tbase");
          return PK_E_INVALID;
        }
    }
    if (d->materials)
      for (uint32_t i = 0; i < nm * 4; ++i)
        if (d->materials[i] < 0 || (uint32_t)d->materials[i] >= layers) {
          pk_set_error("pk_scene_upload: material " + std::to_string(i / 4) + " references a missing atlas layer");
          return PK_E_INVALID;
        }
    EnvTables env;
    std::string err;
    int rc = pk_build_env_tables(d, &env, &err);
    if (rc) {
      pk_set_error("pk_scene_upload: " + err);
      return rc;
    }
    PK_HIP(hipSetDevice(c->device));
    auto* s = new pk_scene();
    s->ctx = c;
    s->allocs.ctx = c;
    pk::DevScene& D = s->dev;
    hipStream_t st = c->stream;
    AllocList tmp{c, {}};  // the index records and attributes the shading records are built from
    auto fail = [&](int code) {
      (void)hipStreamSynchronize(st);
      free_all(tmp);
      free_all(s->allocs);
      delete s;
      return code;
    };
    pk::WNode* nodes;
    float4* tris;
    int4* tidx;
    if ((rc = upload(s->allocs, &nodes, bvh->wnodes.data(), bvh->wnodes.size(), st))) return fail(rc);
    if ((rc = upload(s->allocs, &tris, bvh->wtris.data(), bvh->wtris.size() * 3, st))) return fail(rc);
    if ((rc = upload(tmp, &tidx, bvh->wtri_index.data(), bvh->wtri_index.size() / 4, st))) return fail(rc);
    D.nodes = nodes;
    D.tris = tris;
    // the 4-wide layout of a single-level BVH (node_width 4 shapes walk it);
    // n4 stays alive until the stream synchronizes below
    std::vector<pk::WNode4> n4;
    D.nodes4 = nullptr;
    if (!bvh->instanced) {
      pk_make_wide4(bvh->wnodes, n4);
      pk::WNode4* nodes4;
      if ((rc = upload(s->allocs, &nodes4, n4.data(), n4.size(), st))) return fail(rc);
      D.nodes4 = nodes4;
    }
    std::vector<uint32_t> all_attr;
    const void* attr_src = d->attributes;
    size_t attr_count = d->num_vertices;
    if (!attr_src || bvh->instanced) {
      // the scene's own vertices (zero if absent), then the meshes'
      if (d->attributes) all_attr.assign(d->attributes, d->attributes + (size_t)d->num_vertices * 4);
      else all_attr.assign((size_t)d->num_vertices * 4, 0u);
      all_attr.insert(all_attr.end(), bvh->inst_attrs.begin(), bvh->inst_attrs.end());
      if (all_attr.empty()) all_attr.assign(4, 0u);
      attr_src = all_attr.data();
      attr_count = all_attr.size() / 4;
    }
    // every vertex index of a leaf record must name an attribute record
    for (size_t i = 0; i + 3 < bvh->wtri_index.size(); i += 4)
      for (int v = 0; v < 3; ++v)
        if (bvh->wtri_index[i + v] < 0 || (size_t)bvh->wtri_index[i + v] >= attr_count) {
          pk_set_error("pk_scene_upload: triangle vertex index beyond the vertex attributes");
          return fail(PK_E_INVALID);
        }
    uint4* attrs;
    if ((rc = upload(tmp, &attrs, attr_src, attr_count, st))) return fail(rc);
    const size_t nrec = bvh->wtri_index.size() / 4;
    uint4* srec;
    if ((rc = dev_alloc(s->allocs, &srec, 4 * nrec))) return fail(rc);
    pk::launch_build_shade_records(tidx, attrs, srec, (uint32_t)nrec, st);
    PK_HIP(hipGetLastError());
    D.shade_rec = srec;
    D.inst = nullptr;
    D.inst_leaf_base = nullptr;
    D.num_inst = 0;
    D.world_leaves = bvh->instanced ? bvh->world_leaves : 0u;
    std::vector<uint32_t> bases;  // lives until the stream is synchronized below
    if (bvh->instanced) {
      bases.resize(bvh->inst.size());
      for (size_t i = 0; i < bases.size(); ++i) bases[i] = bvh->inst[i].leaf_base;
      pk::InstRecord* ir;
      uint32_t* lb;
      if ((rc = upload(s->allocs, &ir, bvh->inst.data(), bvh->inst.size(), st))) return fail(rc);
      if ((rc = upload(s->allocs, &lb, bases.data(), bases.size(), st))) return fail(rc);
      D.inst = ir;
      D.inst_leaf_base = lb;
      D.num_inst = (uint32_t)bvh->inst.size();
    }
    const int32_t default_mat[4] = {0, 0, 0, 0};
    int4* mats;
    if ((rc = upload(s->allocs, &mats, d->materials ? (const void*)d->materials : (const void*)default_mat, nm, st)))
      return fail(rc);
    D.materials = mats;
    if (d->tex_data) {
      const size_t texels = (size_t)d->tex_width * d->tex_height * d->tex_layers;
      if (d->tex_format == PK_TEX_RGBA8) {
        uint8_t* t8;
        if ((rc = upload(s->allocs, &t8, d->tex_data, texels * 4, st))) return fail(rc);
        D.tex = t8;
      } else {
        float* t32;
        if ((rc = upload(s->allocs, &t32, d->tex_data, texels * 4, st))) return fail(rc);
        D.tex = t32;
      }
      D.tw = d->tex_width;
      D.th = d->tex_height;
      D.tl = d->tex_layers;
      D.tfmt = d->tex_format;
    } else {
      static const uint8_t white[4] = {255, 255, 255, 255};
      uint8_t* t8;
      if ((rc = upload(s->allocs, &t8, white, 4, st))) return fail(rc);
      D.tex = t8;
      D.tw = D.th = D.tl = 1;
      D.tfmt = PK_TEX_RGBA8;
    }
    D.tex_eval = D.tex;  // parameters not yet written: evaluation reads the sampling atlas
    D.env_mode = 0;
    if (d->env_format != PK_ENV_NONE) {
      const size_t T = (size_t)env.width * env.height;
      std::vector<float> rad4(T * 4);
      for (size_t i = 0; i < T; ++i) {
        rad4[i * 4 + 0] = env.radiance[i * 3 + 0];
        rad4[i * 4 + 1] = env.radiance[i * 3 + 1];
        rad4[i * 4 + 2] = env.radiance[i * 3 + 2];
        rad4[i * 4 + 3] = 0.0f;
      }
      float4* er;
      int2 *eb, *ec;
      float *ep, *rc_cdf = nullptr, *cc_cdf = nullptr;
      uint32_t *rguide = nullptr, *cguide = nullptr;
      if ((rc = upload(s->allocs, &er, rad4.data(), T, st))) return fail(rc);
      if ((rc = upload(s->allocs, &eb, env.bins.data(), env.bins.size() / 2, st))) return fail(rc);
      if ((rc = upload(s->allocs, &ec, env.coords.data(), env.coords.size() / 2, st))) return fail(rc);
      if ((rc = upload(s->allocs, &ep, env.pdf.data(), env.pdf.size(), st))) return fail(rc);
      if (d->env_sampler == PK_ENV_SAMPLER_CDF) {
        if ((rc = upload(s->allocs, &rc_cdf, env.row_cdf.data(), env.row_cdf.size(), st))) return fail(rc);
        if ((rc = upload(s->allocs, &cc_cdf, env.col_cdf.data(), env.col_cdf.size(), st))) return fail(rc);
        if ((rc = upload(s->allocs, &rguide, env.row_guide.data(), env.row_guide.size(), st))) return fail(rc);
        if ((rc = upload(s->allocs, &cguide, env.col_guide.data(), env.col_guide.size(), st))) return fail(rc);
      }
      D.env_rad = er;
      D.env_bins = eb;
      D.env_coords = ec;
      D.env_pdf = ep;
      D.row_cdf = rc_cdf;
      D.col_cdf = cc_cdf;
      D.row_guide = rguide;
      D.col_guide = cguide;
      D.ew = env.width;
      D.eh = env.height;
      D.nbins = (uint32_t)(env.bins.size() / 2);
      D.env_mode = d->env_sampler == PK_ENV_SAMPLER_CDF ? 2 : 1;
      s->env_info[0] = (int64_t)D.nbins;
      s->env_info[1] = (int64_t)(env.coords.size() / 2);
      s->env_info[2] = env.width;
      s->env_info[3] = env.height;
    }
    hipError_t e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
      pk_set_error(std::string("pk_scene_upload: ") + hipGetErrorString(e));
      return fail(PK_E_HIP);
    }
    free_all(tmp);
    *out = s;
    return PK_OK;
  });
}

int pk_scene_destroy(pk_scene* s) {
  return pk_guard("pk_scene_destroy", [&]() -> int {
    if (!s) return PK_OK;
    (void)hipSetDevice(s->ctx->device);
    (void)hipStreamSynchronize(s->ctx->stream);
    free_all(s->allocs);
    delete s;
    return PK_OK;
  });
}

int pk_scene_env_info(const pk_scene* s, int64_t* info4) {
  return pk_guard("pk_scene_env_info", [&]() -> int {
    if (!s || !info4) {
      pk_set_error("pk_scene_env_info: null argument");
      return PK_E_INVALID;
    }
    std::memcpy(info4, s->env_info, sizeof(s->env_info));
    return PK_OK;
  });
}

// ---------------------------------------------------------------------------
// frame
// ---------------------------------------------------------------------------
// Per-path queue buffers for `paths` paths (replacing any previous ones).
// Allocates per-path queues for `paths` paths into `list` and the pointers
// of F (the frame's current queues are not touched); on failure everything
// allocated here is freed again.
static int alloc_paths_into(pk_frame* f, size_t paths, AllocList& list, pk::FrameDev& F) {
  int rc = PK_OK;
  auto fail = [&](int code) {
    free_all(list);
    return code;
  };
  for (int q = 0; q < 2; ++q) {
    if ((rc = dev_alloc(list, &F.qo[q], paths))) return fail(rc);
    if ((rc = dev_alloc(list, &F.qd[q], paths))) return fail(rc);
    if ((rc = dev_alloc(list, &F.qt[q], paths))) return fail(rc);
  }
  // the hit queue also holds the holes at the end of each trace_closest
  // wave's last run (fewer than kHitRun per wave of the largest grid)
  const size_t hit_cap = paths + (size_t)f->ctx->num_cus * kMaxWavesPerCu * pk::kHitRun;
  if (hit_cap >= (1ull << 32)) {
    pk_set_error("frame: hit queue beyond 2^32 records");
    return fail(PK_E_LIMIT);
  }
  if ((rc = dev_alloc(list, &F.hits, hit_cap))) return fail(rc);
  F.hit_cap = (uint32_t)hit_cap;
  if ((rc = dev_alloc(list, &F.so, paths))) return fail(rc);
  if ((rc = dev_alloc(list, &F.sd, paths))) return fail(rc);
  if ((rc = dev_alloc(list, &F.sthr, paths))) return fail(rc);
  if ((rc = dev_alloc(list, &F.color, paths))) return fail(rc);
  pk::launch_fill_f4(F.color, make_float4(0, 0, 0, 0), (uint32_t)paths, f->ctx->stream);
  return PK_OK;
}

int pk_frame_create(pk_ctx* c, uint32_t W, uint32_t H, pk_frame** out) {
  return pk_guard("pk_frame_create", [&]() -> int {
    if (!c || !out) {
      pk_set_error("pk_frame_create: null argument");
      return PK_E_INVALID;
    }
    *out = nullptr;
    if (W == 0 || H == 0 || (uint64_t)W * H >= (1ull << 28)) {
      pk_set_error("pk_frame_create: resolution must be nonzero and below 2^28 pixels");
      return PK_E_LIMIT;
    }
    PK_HIP(hipSetDevice(c->device));
    auto* f = new pk_frame();
    f->ctx = c;
    f->allocs.ctx = c;
    f->path_allocs.ctx = c;
    f->W = W;
    f->H = H;
    const size_t B = (size_t)W * H;
    pk::FrameDev& F = f->dev;
    F.W = W;
    F.H = H;
    int rc = 0;
    auto fail = [&](int code) {
      free_all(f->path_allocs);
      free_all(f->allocs);
      delete f;
      return code;
    };
    F.S = 1;
    if ((rc = alloc_paths_into(f, B, f->path_allocs, F))) return fail(rc);
    if ((rc = dev_alloc(f->allocs, &F.accum, B))) return fail(rc);
    if ((rc = dev_alloc(f->allocs, &F.cnt, pk::kNumCounters))) return fail(rc);
    if ((rc = dev_alloc(f->allocs, &F.stats, pk::kNumStats))) return fail(rc);
    if ((rc = dev_alloc(f->allocs, &F.params, 1))) return fail(rc);
    F.spill = c->spill;
    pk::launch_fill_f4(F.accum, make_float4(0, 0, 0, 0), (uint32_t)B, c->stream);
    pk::launch_fill_u32(F.cnt, 0u, pk::kNumCounters, c->stream);
    pk::launch_fill_u32((uint32_t*)F.stats, 0u, pk::kNumStats * 2, c->stream);
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
      pk_set_error(std::string("pk_frame_create: ") + hipGetErrorString(e));
      return fail(PK_E_HIP);
    }
    *out = f;
    return PK_OK;
  });
}

static void drop_graph(pk_frame* f) {
  if (f->exec) (void)hipGraphExecDestroy(f->exec);
  if (f->graph) (void)hipGraphDestroy(f->graph);
  f->exec = nullptr;
  f->graph = nullptr;
  f->graph_scene = nullptr;
}

int pk_frame_destroy(pk_frame* f) {
  return pk_guard("pk_frame_destroy", [&]() -> int {
    if (!f) return PK_OK;
    (void)hipSetDevice(f->ctx->device);
    (void)hipStreamSynchronize(f->ctx->stream);
    drop_graph(f);
    free_all(f->path_allocs);
    free_all(f->allocs);
    delete f;
    return PK_OK;
  });
}

int pk_frame_get_samples_per_pass(const pk_frame* f, uint32_t* samples) {
  return pk_guard("pk_frame_get_samples_per_pass", [&]() -> int {
    if (!f || !samples) {
      pk_set_error("pk_frame_get_samples_per_pass: null argument");
      return PK_E_INVALID;
    }
    *samples = f->dev.S;
    return PK_OK;
  });
}

int pk_frame_set_samples_per_pass(pk_frame* f, uint32_t samples) {
  return pk_guard("pk_frame_set_samples_per_pass", [&]() -> int {
    if (!f || samples == 0) {
      pk_set_error("pk_frame_set_samples_per_pass: null frame or zero samples");
      return PK_E_INVALID;
    }
    const uint64_t paths = (uint64_t)f->W * f->H * samples;
    if (paths >= (1ull << 30)) {
      pk_set_error("pk_frame_set_samples_per_pass: width*height*samples must stay below 2^30 paths");
      return PK_E_LIMIT;
    }
    if (samples == f->dev.S) return PK_OK;
    PK_HIP(hipSetDevice(f->ctx->device));
    PK_HIP(hipStreamSynchronize(f->ctx->stream));
    // the new queues first: on failure the frame keeps its queues and its
    // samples per pass
    AllocList fresh{f->ctx, {}};
    pk::FrameDev F = f->dev;
    int rc = alloc_paths_into(f, (size_t)paths, fresh, F);
    if (rc) return rc;
    drop_graph(f);
    free_all(f->path_allocs);
    f->path_allocs = std::move(fresh);
    F.S = samples;
    f->dev = F;
    PK_HIP(hipStreamSynchronize(f->ctx->stream));
    return PK_OK;
  });
}

int pk_frame_reset(pk_frame* f) {
  return pk_guard("pk_frame_reset", [&]() -> int {
    if (!f) {
      pk_set_error("pk_frame_reset: null frame");
      return PK_E_INVALID;
    }
    const uint32_t B = f->W * f->H;
    pk::launch_fill_f4(f->dev.accum, make_float4(0, 0, 0, 0), B, f->ctx->stream);
    pk::launch_fill_f4(f->dev.color, make_float4(0, 0, 0, 0), B * f->dev.S, f->ctx->stream);
    pk::launch_fill_u32((uint32_t*)f->dev.stats, 0u, pk::kNumStats * 2, f->ctx->stream);
    pk::launch_fill_u32(f->dev.cnt, 0u, pk::kNumCounters, f->ctx->stream);
    PK_HIP(hipGetLastError());
    return PK_OK;
  });
}

int pk_frame_set_camera(pk_frame* f, const pk_camera* cam) {
  return pk_guard("pk_frame_set_camera", [&]() -> int {
    if (!f || !cam) {
      pk_set_error("pk_frame_set_camera: null argument");
      return PK_E_INVALID;
    }
    f->cam = *cam;
    f->has_camera = true;
    return PK_OK;
  });
}

int pk_frame_set_env_theta(pk_frame* f, float theta) {
  return pk_guard("pk_frame_set_env_theta", [&]() -> int {
    if (!f) {
      pk_set_error("pk_frame_set_env_theta: null frame");
      return PK_E_INVALID;
    }
    f->theta = theta;
    return PK_OK;
  });
}

int pk_frame_begin(pk_frame* f, uint32_t sample) {
  return pk_guard("pk_frame_begin", [&]() -> int {
    if (!f) {
      pk_set_error("pk_frame_begin: null frame");
      return PK_E_INVALID;
    }
    if (!f->has_camera) {
      pk_set_error("pk_frame_begin: no camera set (pk_frame_set_camera)");
      return PK_E_STATE;
    }
    pk::launch_begin(f->dev, sample, f->theta, f->cam, f->ctx->stream);
    f->begun = true;
    PK_HIP(hipGetLastError());
    return PK_OK;
  });
}

static int check_pass(pk_frame* f, pk_scene* s, const char* what) {
  if (!f || (s == nullptr && what[0] != 'r' && what[0] != 'a')) {
    pk_set_error(std::string(what) + ": null argument");
    return PK_E_INVALID;
  }
  if (s && s->ctx != f->ctx) {
    pk_set_error(std::string(what) + ": scene and frame belong to different contexts");
    return PK_E_INVALID;
  }
  if (!f->begun) {
    pk_set_error(std::string(what) + ": frame not begun (pk_frame_begin)");
    return PK_E_STATE;
  }
  return PK_OK;
}

int pk_raygen(pk_frame* f) {
  return pk_guard("pk_raygen", [&]() -> int {
    int rc = check_pass(f, nullptr, "raygen");
    if (rc) return rc;
    pk::launch_raygen(f->dev, f->ctx->stream);
    PK_HIP(hipGetLastError());
    return PK_OK;
  });
}

int pk_trace_closest(pk_frame* f, pk_scene* s, uint32_t bounce) {
  return pk_guard("pk_trace_closest", [&]() -> int {
    int rc = check_pass(f, s, "pk_trace_closest");
    if (rc) return rc;
    pk::launch_trace_closest(s->dev, f->dev, (int)bounce, trace_launch(f, closest_class(bounce)), f->ctx->stream);
    PK_HIP(hipGetLastError());
    return PK_OK;
  });
}

int pk_shade(pk_frame* f, pk_scene* s, uint32_t bounce) {
  return pk_guard("pk_shade", [&]() -> int {
    int rc = check_pass(f, s, "pk_shade");
    if (rc) return rc;
    pk::launch_shade(s->dev, f->dev, (int)bounce, f->ctx->stream);
    PK_HIP(hipGetLastError());
    return PK_OK;
  });
}

int pk_trace_shadow(pk_frame* f, pk_scene* s) {
  return pk_guard("pk_trace_shadow", [&]() -> int {
    int rc = check_pass(f, s, "pk_trace_shadow");
    if (rc) return rc;
    pk::launch_trace_shadow(s->dev, f->dev, trace_launch(f, PK_TUNE_SHADOW), f->ctx->stream);
    PK_HIP(hipGetLastError());
    return PK_OK;
  });
}

int pk_accumulate(pk_frame* f) {
  return pk_guard("pk_accumulate", [&]() -> int {
    int rc = check_pass(f, nullptr, "accumulate");
    if (rc) return rc;
    pk::launch_accumulate(f->dev, f->ctx->stream);
    PK_HIP(hipGetLastError());
    return PK_OK;
  });
}

static void enqueue_frame(pk_frame* f, pk_scene* s, uint32_t depth, hipStream_t st) {
  if (f->ctx->sched == PK_SCHED_PATHS && !f->count_nodes && depth > 0) {
    pk::launch_render_paths(s->dev, f->dev, depth, trace_launch(f, PK_TUNE_PATHS), st);
    pk::launch_accumulate(f->dev, st);
    pk::launch_advance(f->dev, st);
    return;
  }
  pk::launch_raygen(f->dev, st);
  for (uint32_t b = 0; b < depth; ++b) {
    pk::launch_trace_closest(s->dev, f->dev, (int)b, trace_launch(f, closest_class(b)), st);
    pk::launch_shade(s->dev, f->dev, (int)b, st);
    pk::launch_trace_shadow(s->dev, f->dev, trace_launch(f, PK_TUNE_SHADOW), st);
  }
  pk::launch_accumulate(f->dev, st);
  pk::launch_advance(f->dev, st);
}

int pk_render(pk_frame* f, pk_scene* s, const pk_render_params* p) {
  return pk_guard("pk_render", [&]() -> int {
    if (!f || !s || !p) {
      pk_set_error("pk_render: null argument");
      return PK_E_INVALID;
    }
    if (p->width != f->W || p->height != f->H) {
      pk_set_error("pk_render: params resolution differs from the frame's");
      return PK_E_INVALID;
    }
    if (!f->has_camera) {
      pk_set_error("pk_render: no camera set (pk_frame_set_camera)");
      return PK_E_STATE;
    }
    if (p->max_depth > 64) {
      pk_set_error("pk_render: max_depth above 64");
      return PK_E_LIMIT;
    }
    if (p->first_sample + p->num_frames < p->first_sample) {
      pk_set_error("pk_render: first_sample + num_frames overflows");
      return PK_E_LIMIT;
    }
    hipStream_t st = f->ctx->stream;
    f->theta = p->env_theta;
    // one pass traces S samples per pixel; the last pass stops at the limit
    const uint32_t passes = (p->num_frames + f->dev.S - 1) / f->dev.S;
    pk::launch_begin(f->dev, p->first_sample, p->env_theta, f->cam, st, p->first_sample + p->num_frames);
    f->begun = true;
    if (!p->use_graph) {
      for (uint32_t i = 0; i < passes; ++i) enqueue_frame(f, s, p->max_depth, st);
      PK_HIP(hipGetLastError());
      return PK_OK;
    }
    if (!f->exec || f->graph_scene != s || f->graph_scene_gen != s->gen || f->graph_depth != p->max_depth ||
        f->graph_gen != f->ctx->tuning_gen) {
      drop_graph(f);
      PK_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      enqueue_frame(f, s, p->max_depth, st);
      hipError_t e = hipStreamEndCapture(st, &f->graph);
      if (e != hipSuccess) {
        pk_set_error(std::string("pk_render: graph capture failed: ") + hipGetErrorString(e));
        return PK_E_HIP;
      }
      PK_HIP(hipGraphInstantiate(&f->exec, f->graph, nullptr, nullptr, 0));
      f->graph_scene = s;
      f->graph_scene_gen = s->gen;
      f->graph_depth = p->max_depth;
      f->graph_gen = f->ctx->tuning_gen;
    }
    for (uint32_t i = 0; i < passes; ++i) PK_HIP(hipGraphLaunch(f->exec, st));
    return PK_OK;
  });
}

int pk_read_image(pk_frame* f, float* rgba, size_t bytes) {
  return pk_guard("pk_read_image", [&]() -> int {
    if (!f || !rgba || bytes < (size_t)f->W * f->H * 16) {
      pk_set_error("pk_read_image: null argument or buffer smaller than width*height*16");
      return PK_E_INVALID;
    }
    PK_HIP(hipStreamSynchronize(f->ctx->stream));
    PK_HIP(hipMemcpy(rgba, f->dev.accum, (size_t)f->W * f->H * 16, hipMemcpyDeviceToHost));
    return PK_OK;
  });
}

int pk_read_image_device(pk_frame* f, float* d_rgba, size_t bytes) {
  return pk_guard("pk_read_image_device", [&]() -> int {
    if (!f || !d_rgba || bytes < (size_t)f->W * f->H * 16) {
      pk_set_error("pk_read_image_device: null argument or buffer smaller than width*height*16");
      return PK_E_INVALID;
    }
    PK_HIP(hipSetDevice(f->ctx->device));
    PK_HIP(hipMemcpyAsync(d_rgba, f->dev.accum, (size_t)f->W * f->H * 16, hipMemcpyDeviceToDevice, f->ctx->stream));
    return PK_OK;
  });
}

int pk_read_color(pk_frame* f, float* rgba, size_t bytes) {
  return pk_guard("pk_read_color", [&]() -> int {
    if (!f || !rgba || bytes < (size_t)f->W * f->H * 16) {
      pk_set_error("pk_read_color: null argument or buffer smaller than width*height*16");
      return PK_E_INVALID;
    }
    PK_HIP(hipStreamSynchronize(f->ctx->stream));
    // sample 0 of each pixel: a pixel's S color slots are adjacent
    const size_t S = f->dev.S;
    if (S == 1) PK_HIP(hipMemcpy(rgba, f->dev.color, (size_t)f->W * f->H * 16, hipMemcpyDeviceToHost));
    else PK_HIP(hipMemcpy2D(rgba, 16, f->dev.color, 16 * S, 16, (size_t)f->W * f->H, hipMemcpyDeviceToHost));
    return PK_OK;
  });
}

int pk_read_counters(pk_frame* f, pk_counters* out) {
  return pk_guard("pk_read_counters", [&]() -> int {
    if (!f || !out) {
      pk_set_error("pk_read_counters: null argument");
      return PK_E_INVALID;
    }
    PK_HIP(hipStreamSynchronize(f->ctx->stream));
    unsigned long long st[pk::kNumStats];
    uint32_t overflow = 0;
    PK_HIP(hipMemcpy(st, f->dev.stats, sizeof(st), hipMemcpyDeviceToHost));
    PK_HIP(hipMemcpy(&overflow, f->dev.cnt + pk::kCntOverflow, sizeof(overflow), hipMemcpyDeviceToHost));
    out->rays_traced = st[pk::kStatRays];
    out->shadow_rays = st[pk::kStatShadow];
    out->hits = st[pk::kStatHits];
    out->nodes_visited = st[pk::kStatNodes];
    out->tris_tested = st[pk::kStatTris];
    out->shadow_nodes = st[pk::kStatShadowNodes];
    out->shadow_tris = st[pk::kStatShadowTris];
    out->frames = (uint32_t)st[pk::kStatFrames];
    out->stack_overflow = overflow;
    return PK_OK;
  });
}

int pk_frame_count_nodes(pk_frame* f, int enable) {
  return pk_guard("pk_frame_count_nodes", [&]() -> int {
    if (!f) {
      pk_set_error("pk_frame_count_nodes: null frame");
      return PK_E_INVALID;
    }
    if ((bool)enable != f->count_nodes) drop_graph(f);
    f->count_nodes = enable != 0;
    return PK_OK;
  });
}

// ---------------------------------------------------------------------------
// standalone ray queries
// ---------------------------------------------------------------------------
static int trace_rays_launch(pk_ctx* c, pk_scene* s, const float* d_rays, int64_t n, int any_hit, int32_t* d_ids,
                             float* d_t, uint32_t* d_bary, int count, uint32_t cursor_index) {
  pk::launch_trace_rays(s->dev, (const float4*)d_rays, (uint32_t)n, any_hit, d_ids, d_t, d_bary,
                        c->cursors + cursor_index, c->spill, c->stats, c->overflow,
                        shape_launch(c, any_hit ? PK_TUNE_RAYS_ANY : PK_TUNE_RAYS_CLOSEST, count != 0), c->stream);
  PK_HIP(hipGetLastError());
  return PK_OK;
}

static int check_rays(pk_ctx* c, pk_scene* s, int64_t n, const char* what) {
  if (!c || !s) {
    pk_set_error(std::string(what) + ": null context or scene");
    return PK_E_INVALID;
  }
  if (s->ctx != c) {
    pk_set_error(std::string(what) + ": scene belongs to another context");
    return PK_E_INVALID;
  }
  if (n < 0 || n >= (1ll << 31)) {
    pk_set_error(std::string(what) + ": ray count must be in [0, 2^31)");
    return PK_E_LIMIT;
  }
  return PK_OK;
}

int pk_trace_rays_device(pk_ctx* c, pk_scene* s, const float* d_rays, int64_t n, int any_hit, int32_t* d_ids,
                         float* d_t, uint32_t* d_bary, int count_nodes) {
  return pk_guard("pk_trace_rays_device", [&]() -> int {
    int rc = check_rays(c, s, n, "pk_trace_rays_device");
    if (rc) return rc;
    if (n == 0) return PK_OK;
    if (!d_rays || !d_ids || !d_t || !d_bary) {
      pk_set_error("pk_trace_rays_device: null buffer");
      return PK_E_INVALID;
    }
    pk::launch_fill_u32(c->cursors, 0u, 1, c->stream);
    pk::launch_fill_u32((uint32_t*)c->stats, 0u, 4, c->stream);
    pk::launch_fill_u32(c->overflow, 0u, 1, c->stream);
    return trace_rays_launch(c, s, d_rays, n, any_hit, d_ids, d_t, d_bary, count_nodes, 0);
  });
}

// pk_trace_rays_device with its work cursor and traversal stack spill area
// in the file-scope scratch array (this device's copy, resolved by
// pk_ctx_create) instead of context memory: no allocation, no blocking call.
// The grid is capped at the kNoAllocMaxWaves one-wave blocks the scratch
// holds stacks for.  Overflow is reported through pk_ctx_trace_stats, as
// after pk_trace_rays_device.  Calls that overlap on one device race on the
// shared scratch.
int pk_trace_rays_noalloc(pk_scene* s, const float* d_rays, int64_t n, int any_hit, int32_t* d_ids, float* d_t,
                          uint32_t* d_bary) {
  return pk_guard("pk_trace_rays_noalloc", [&]() -> int {
    if (!s) {
      pk_set_error("pk_trace_rays_noalloc: null scene");
      return PK_E_INVALID;
    }
    pk_ctx* c = s->ctx;
    int rc = check_rays(c, s, n, "pk_trace_rays_noalloc");
    if (rc) return rc;
    if (n == 0) return PK_OK;
    if (!d_rays || !d_ids || !d_t || !d_bary) {
      pk_set_error("pk_trace_rays_noalloc: null buffer");
      return PK_E_INVALID;
    }
    if (!c->scratch) {
      pk_set_error("pk_trace_rays_noalloc: the context has no scratch address");
      return PK_E_STATE;
    }
    PK_HIP(hipSetDevice(c->device));
    uint32_t* cursor = reinterpret_cast<uint32_t*>(c->scratch);
    uint2* spill = reinterpret_cast<uint2*>(c->scratch + kNoAllocCursorBytes);
    pk::TraceLaunch L = shape_launch(c, any_hit ? PK_TUNE_RAYS_ANY : PK_TUNE_RAYS_CLOSEST, false);
    L.grid = std::min(L.grid, kNoAllocMaxWaves);
    L.fixed_grid = std::min(L.fixed_grid, kNoAllocMaxWaves);
    L.spill = spill;
    // cleared by kernels (no memset nodes if a caller captures this into a graph)
    pk::launch_fill_u32(cursor, 0u, 1, c->stream);
    pk::launch_fill_u32(c->overflow, 0u, 1, c->stream);
    pk::launch_trace_rays(s->dev, (const float4*)d_rays, (uint32_t)n, any_hit, d_ids, d_t, d_bary, cursor, spill,
                          c->stats, c->overflow, L, c->stream);
    PK_HIP(hipGetLastError());
    return PK_OK;
  });
}

int pk_ctx_trace_stats(pk_ctx* c, uint64_t* stats2) {
  return pk_guard("pk_ctx_trace_stats", [&]() -> int {
    if (!c || !stats2) {
      pk_set_error("pk_ctx_trace_stats: null argument");
      return PK_E_INVALID;
    }
    PK_HIP(hipStreamSynchronize(c->stream));
    unsigned long long st[2];
    uint32_t ov = 0;
    PK_HIP(hipMemcpy(st, c->stats, sizeof(st), hipMemcpyDeviceToHost));
    PK_HIP(hipMemcpy(&ov, c->overflow, 4, hipMemcpyDeviceToHost));
    stats2[0] = st[0];
    stats2[1] = st[1];
    if (ov) {
      pk_set_error("traversal stack overflow");
      return PK_E_OVERFLOW;
    }
    return PK_OK;
  });
}

int pk_trace_rays(pk_ctx* c, pk_scene* s, const float* rays, int64_t n, int any_hit, int32_t* ids, float* t,
                  uint32_t* bary) {
  return pk_guard("pk_trace_rays", [&]() -> int {
    int rc = check_rays(c, s, n, "pk_trace_rays");
    if (rc) return rc;
    if (n == 0) return PK_OK;
    if (!rays || !ids || !t || !bary) {
      pk_set_error("pk_trace_rays: null buffer");
      return PK_E_INVALID;
    }
    PK_HIP(hipSetDevice(c->device));
    AllocList tmp{c, {}};
    float4* d_rays;
    int32_t* d_ids;
    float* d_t;
    uint32_t* d_b;
    auto fail = [&](int code) {
      (void)hipStreamSynchronize(c->stream);
      free_all(tmp);
      return code;
    };
    if ((rc = upload(tmp, &d_rays, rays, (size_t)n * 2, c->stream))) return fail(rc);
    if ((rc = dev_alloc(tmp, &d_ids, (size_t)n))) return fail(rc);
    if ((rc = dev_alloc(tmp, &d_t, (size_t)n))) return fail(rc);
    if ((rc = dev_alloc(tmp, &d_b, (size_t)n))) return fail(rc);
    if ((rc = pk_trace_rays_device(c, s, (const float*)d_rays, n, any_hit, d_ids, d_t, d_b, 0))) return fail(rc);
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = hipMemcpy(ids, d_ids, (size_t)n * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(t, d_t, (size_t)n * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(bary, d_b, (size_t)n * 4, hipMemcpyDeviceToHost);
    uint32_t ov = 0;
    if (e == hipSuccess) e = hipMemcpy(&ov, c->overflow, 4, hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
      pk_set_error(std::string("pk_trace_rays: ") + hipGetErrorString(e));
      return fail(PK_E_HIP);
    }
    free_all(tmp);
    if (ov) {
      pk_set_error("pk_trace_rays: traversal stack overflow");
      return PK_E_OVERFLOW;
    }
    return PK_OK;
  });
}

int pk_time_trace_rays(pk_ctx* c, pk_scene* s, const float* d_rays, int64_t n, int any_hit, int32_t* d_ids,
                       float* d_t, uint32_t* d_bary, int iters, float* ms_out) {
  return pk_guard("pk_time_trace_rays", [&]() -> int {
    int rc = check_rays(c, s, n, "pk_time_trace_rays");
    if (rc) return rc;
    if (iters <= 0 || !ms_out || !d_rays || !d_ids || !d_t || !d_bary) {
      pk_set_error("pk_time_trace_rays: bad iteration count or null buffer");
      return PK_E_INVALID;
    }
    if ((rc = ensure_cursors(c, (uint32_t)iters))) return rc;
    pk::launch_fill_u32(c->cursors, 0u, (uint32_t)iters, c->stream);
    pk::launch_fill_u32(c->overflow, 0u, 1, c->stream);
    PK_HIP(hipEventRecord(c->ev0, c->stream));
    for (int i = 0; i < iters; ++i)
      if ((rc = trace_rays_launch(c, s, d_rays, n, any_hit, d_ids, d_t, d_bary, 0, (uint32_t)i))) return rc;
    PK_HIP(hipEventRecord(c->ev1, c->stream));
    PK_HIP(hipEventSynchronize(c->ev1));
    float ms = 0;
    PK_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    *ms_out = ms / (float)iters;
    return PK_OK;
  });
}

// Kernel-level timing of whole frames: events around each stage on the
// context stream; ms5 receives total milliseconds of raygen, trace_closest,
// shade, trace_shadow and accumulate over num_frames frames.
int pk_render_profile(pk_frame* f, pk_scene* s, const pk_render_params* p, double* ms5) {
  return pk_guard("pk_render_profile", [&]() -> int {
    if (!f || !s || !p || !ms5) {
      pk_set_error("pk_render_profile: null argument");
      return PK_E_INVALID;
    }
    if (!f->has_camera) {
      pk_set_error("pk_render_profile: no camera set");
      return PK_E_STATE;
    }
    hipStream_t st = f->ctx->stream;
    std::vector<hipEvent_t> ev;
    const uint32_t per_pass = 2 + 3 * p->max_depth;  // launches timed per pass
    const uint32_t passes = (p->num_frames + f->dev.S - 1) / f->dev.S;
    const size_t nev = (size_t)passes * per_pass * 2;
    ev.resize(nev);
    for (auto& e : ev) PK_HIP(hipEventCreate(&e));
    for (int k = 0; k < 5; ++k) ms5[k] = 0;
    pk::launch_begin(f->dev, p->first_sample, p->env_theta, f->cam, st, p->first_sample + p->num_frames);
    f->begun = true;
    size_t ei = 0;
    std::vector<int> kind;
    for (uint32_t fr = 0; fr < passes; ++fr) {
      auto timed = [&](int k, auto fn) {
        (void)hipEventRecord(ev[ei++], st);
        fn();
        (void)hipEventRecord(ev[ei++], st);
        kind.push_back(k);
      };
      timed(0, [&]() { pk::launch_raygen(f->dev, st); });
      for (uint32_t b = 0; b < p->max_depth; ++b) {
        timed(1, [&]() { pk::launch_trace_closest(s->dev, f->dev, (int)b, trace_launch(f, closest_class(b)), st); });
        timed(2, [&]() { pk::launch_shade(s->dev, f->dev, (int)b, st); });
        timed(3, [&]() { pk::launch_trace_shadow(s->dev, f->dev, trace_launch(f, PK_TUNE_SHADOW), st); });
      }
      timed(4, [&]() { pk::launch_accumulate(f->dev, st); });
      pk::launch_advance(f->dev, st);
    }
    hipError_t e = hipStreamSynchronize(st);
    for (size_t i = 0; i < kind.size() && e == hipSuccess; ++i) {
      float ms = 0;
      e = hipEventElapsedTime(&ms, ev[2 * i], ev[2 * i + 1]);
      ms5[kind[i]] += ms;
    }
    for (auto& x : ev) (void)hipEventDestroy(x);
    if (e != hipSuccess) {
      pk_set_error(std::string("pk_render_profile: ") + hipGetErrorString(e));
      return PK_E_HIP;
    }
    return PK_OK;
  });
}

// ---------------------------------------------------------------------------
// Batches: many scenes and camera views traced through one global wavefront.
// The scenes are merged into one device scene whose indices are global
// (child and leaf references, vertex ids, material ids, atlas layers), so
// the kernels need only a per-ray root node; each view's pixels own a range
// of the global path index space.
// ---------------------------------------------------------------------------
int pk_scene_upload_batch(pk_ctx* c, const pk_scene_desc* descs, const pk_bvh* const* bvhs, uint32_t count,
                          pk_scene** out) {
  return pk_guard("pk_scene_upload_batch", [&]() -> int {
    if (!c || !descs || !bvhs || !out || count == 0) {
      pk_set_error("pk_scene_upload_batch: null argument or no scenes");
      return PK_E_INVALID;
    }
    *out = nullptr;
    const pk_scene_desc& d0 = descs[0];
    pk_bvh merged;
    pk_scene_desc md = d0;
    std::vector<uint32_t> attrs, roots;
    std::vector<int32_t> tris, mats;
    std::vector<uint8_t> tex8;
    std::vector<float> tex32;
    const size_t texel_bytes = d0.tex_format == PK_TEX_RGBA8 ? 4 : 16;
    uint32_t voff = 0, moff = 0, loff = 0;
    for (uint32_t k = 0; k < count; ++k) {
      const pk_scene_desc& d = descs[k];
      const pk_bvh* b = bvhs[k];
      if (b && b->instanced) {
        pk_set_error("pk_scene_upload_batch: scene " + std::to_string(k) + " has an instanced BVH; batches take "
                     "single-level BVHs");
        return PK_E_INVALID;
      }
      if (!b || d.num_triangles != b->nt || d.num_vertices != b->nv || !d.materials || !d.tex_data) {
        pk_set_error("pk_scene_upload_batch: scene " + std::to_string(k) +
                     " lacks materials/atlas or its BVH was built for other geometry");
        return PK_E_INVALID;
      }
      if (d.tex_width != d0.tex_width || d.tex_height != d0.tex_height || d.tex_format != d0.tex_format) {
        pk_set_error("pk_scene_upload_batch: every scene's atlas needs the same layer size and format");
        return PK_E_INVALID;
      }
      const uint32_t noff = (uint32_t)merged.wnodes.size(), toff = (uint32_t)merged.wtris.size();
      roots.push_back(noff);
      for (pk::WNode w : b->wnodes) {
        pk::wnode_rebase(w, noff, toff);
        merged.wnodes.push_back(w);
      }
      merged.wtris.insert(merged.wtris.end(), b->wtris.begin(), b->wtris.end());
      for (size_t i = 0; i < b->wtri_index.size(); i += 4) {
        for (int v = 0; v < 3; ++v) merged.wtri_index.push_back(b->wtri_index[i + v] + (int32_t)voff);
        merged.wtri_index.push_back(b->wtri_index[i + 3] + (int32_t)moff);
      }
      for (size_t t = 0; t < (size_t)d.num_triangles; ++t) {
        for (int v = 0; v < 3; ++v) tris.push_back(d.triangles[t * 4 + v] + (int32_t)voff);
        tris.push_back(d.triangles[t * 4 + 3] + (int32_t)moff);
      }
      if (d.attributes) attrs.insert(attrs.end(), d.attributes, d.attributes + (size_t)d.num_vertices * 4);
      else attrs.insert(attrs.end(), (size_t)d.num_vertices * 4, 0u);
      for (size_t i = 0; i < (size_t)d.num_materials * 4; ++i) mats.push_back(d.materials[i] + (int32_t)loff);
      const size_t layer_bytes = (size_t)d.tex_width * d.tex_height * texel_bytes;
      if (texel_bytes == 4)
        tex8.insert(tex8.end(), (const uint8_t*)d.tex_data, (const uint8_t*)d.tex_data + layer_bytes * d.tex_layers);
      else
        tex32.insert(tex32.end(), (const float*)d.tex_data,
                     (const float*)d.tex_data + layer_bytes / 4 * d.tex_layers);
      voff += d.num_vertices;
      moff += d.num_materials;
      loff += d.tex_layers;
    }
    merged.nt = (uint32_t)merged.wtris.size();
    merged.nv = voff;
    md.num_vertices = voff;
    md.num_triangles = merged.nt;
    md.num_materials = moff;
    md.tex_layers = loff;
    md.attributes = attrs.data();
    md.triangles = tris.data();
    md.materials = mats.data();
    md.tex_data = texel_bytes == 4 ? (const void*)tex8.data() : (const void*)tex32.data();
    // the upload reads no positions: the wide layout already holds every
    // scene's normalized geometry
    md.positions = nullptr;
    int rc = pk_scene_upload(c, &md, &merged, out);
    if (rc) return rc;
    (*out)->roots = roots;
    return PK_OK;
  });
}

int pk_scene_roots(const pk_scene* s, uint32_t* roots, uint32_t count) {
  return pk_guard("pk_scene_roots", [&]() -> int {
    if (!s || (count && !roots)) {
      pk_set_error("pk_scene_roots: null argument");
      return PK_E_INVALID;
    }
    for (uint32_t i = 0; i < count && i < s->roots.size(); ++i) roots[i] = s->roots[i];
    return (int)s->roots.size();
  });
}

int pk_batch_create(pk_ctx* c, pk_scene* s, const pk_view* views, uint32_t nviews, pk_frame** out) {
  return pk_guard("pk_batch_create", [&]() -> int {
    if (!c || !s || !views || !out || nviews == 0) {
      pk_set_error("pk_batch_create: null argument or no views");
      return PK_E_INVALID;
    }
    *out = nullptr;
    std::vector<pk::ViewDev> vd(nviews);
    uint64_t total = 0;
    for (uint32_t i = 0; i < nviews; ++i) {
      const pk_view& v = views[i];
      if (v.scene >= s->roots.size() || v.width == 0 || v.height == 0 || v.spp == 0) {
        pk_set_error("pk_batch_create: view " + std::to_string(i) + " has a bad scene index, size or sample count");
        return PK_E_INVALID;
      }
      vd[i] = pk::ViewDev{v.width, v.height, (uint32_t)total, v.spp, s->roots[v.scene], v.scene, 0, 0, v.camera};
      total += (uint64_t)v.width * v.height;
    }
    if (total >= (1ull << 28)) {
      pk_set_error("pk_batch_create: more than 2^28 pixels in one batch");
      return PK_E_LIMIT;
    }
    pk_frame* f = nullptr;
    int rc = pk_frame_create(c, (uint32_t)total, 1, &f);
    if (rc) return rc;
    pk::ViewDev* dv = nullptr;
    if ((rc = dev_alloc(f->allocs, &dv, nviews))) {
      pk_frame_destroy(f);
      return rc;
    }
    hipError_t e = hipMemcpy(dv, vd.data(), nviews * sizeof(pk::ViewDev), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      pk_set_error(std::string("pk_batch_create: ") + hipGetErrorString(e));
      pk_frame_destroy(f);
      return PK_E_HIP;
    }
    f->dev.views = dv;
    f->dev.nviews = nviews;
    f->views = vd;
    f->batch_scene = s;
    f->cam = views[0].camera;
    f->has_camera = true;
    *out = f;
    return PK_OK;
  });
}

// A batch render runs rounds of S samples (the frame's samples per pass):
// round r traces samples r*S .. r*S+S-1 of every view pixel whose view has
// that many.
static void enqueue_batch(pk_frame* f, pk_scene* s, uint32_t depth, float theta, hipStream_t st) {
  uint32_t max_spp = 0;
  for (const auto& v : f->views) max_spp = std::max(max_spp, v.spp);
  const uint32_t S = f->dev.S, rounds = (max_spp + S - 1) / S;
  const bool paths = f->ctx->sched == PK_SCHED_PATHS && !f->count_nodes;
  for (uint32_t r = 0; r < rounds; ++r) {
    pk::launch_begin(f->dev, r * S, theta, f->cam, st);
    if (paths) {
      pk::launch_render_paths(s->dev, f->dev, depth, trace_launch(f, PK_TUNE_PATHS), st);
      pk::launch_accumulate(f->dev, st);
      continue;
    }
    pk::launch_raygen(f->dev, st);
    for (uint32_t b = 0; b < depth; ++b) {
      pk::launch_trace_closest(s->dev, f->dev, (int)b, trace_launch(f, closest_class(b)), st);
      pk::launch_shade(s->dev, f->dev, (int)b, st);
      pk::launch_trace_shadow(s->dev, f->dev, trace_launch(f, PK_TUNE_SHADOW), st);
    }
    pk::launch_accumulate(f->dev, st);
  }
}

int pk_batch_render(pk_frame* f, uint32_t depth, float theta, int use_graph) {
  return pk_guard("pk_batch_render", [&]() -> int {
    if (!f || !f->batch_scene) {
      pk_set_error("pk_batch_render: not a batch (pk_batch_create)");
      return PK_E_INVALID;
    }
    if (depth == 0 || depth > 64) {
      pk_set_error("pk_batch_render: depth outside [1, 64]");
      return PK_E_LIMIT;
    }
    pk_scene* s = f->batch_scene;
    hipStream_t st = f->ctx->stream;
    f->theta = theta;
    f->begun = true;
    if (!use_graph) {
      enqueue_batch(f, s, depth, theta, st);
      PK_HIP(hipGetLastError());
      return PK_OK;
    }
    if (!f->exec || f->graph_depth != depth || f->graph_gen != f->ctx->tuning_gen || f->graph_theta != theta ||
        f->graph_scene_gen != s->gen) {
      drop_graph(f);
      PK_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      enqueue_batch(f, s, depth, theta, st);
      hipError_t e = hipStreamEndCapture(st, &f->graph);
      if (e != hipSuccess) {
        pk_set_error(std::string("pk_batch_render: graph capture failed: ") + hipGetErrorString(e));
        return PK_E_HIP;
      }
      PK_HIP(hipGraphInstantiate(&f->exec, f->graph, nullptr, nullptr, 0));
      f->graph_scene = s;
      f->graph_scene_gen = s->gen;
      f->graph_depth = depth;
      f->graph_gen = f->ctx->tuning_gen;
      f->graph_theta = theta;
    }
    PK_HIP(hipGraphLaunch(f->exec, st));
    return PK_OK;
  });
}

int pk_batch_read_view(pk_frame* f, uint32_t view, float* rgba, size_t bytes) {
  return pk_guard("pk_batch_read_view", [&]() -> int {
    if (!f || !f->batch_scene || view >= f->views.size() || !rgba) {
      pk_set_error("pk_batch_read_view: not a batch, view out of range or null buffer");
      return PK_E_INVALID;
    }
    const pk::ViewDev& v = f->views[view];
    const size_t n = (size_t)v.W * v.H * 16;
    if (bytes < n) {
      pk_set_error("pk_batch_read_view: buffer smaller than width*height*16");
      return PK_E_INVALID;
    }
    PK_HIP(hipStreamSynchronize(f->ctx->stream));
    PK_HIP(hipMemcpy(rgba, f->dev.accum + v.offset, n, hipMemcpyDeviceToHost));
    return PK_OK;
  });
}

// ---------------------------------------------------------------------------
// Differentiable parameters and the path-replay adjoint
// ---------------------------------------------------------------------------
static size_t param_floats(const pk_scene* s, int which) {
  const pk::DevScene& D = s->dev;
  if (which == PK_PARAM_ATLAS) return (size_t)D.tw * D.th * D.tl * 4;
  if (which == PK_PARAM_ENV) return D.env_mode ? (size_t)D.ew * D.eh * 4 : 0;
  return 0;
}

int pk_scene_param_info(const pk_scene* s, uint32_t* info6) {
  return pk_guard("pk_scene_param_info", [&]() -> int {
    if (!s || !info6) {
      pk_set_error("pk_scene_param_info: null argument");
      return PK_E_INVALID;
    }
    const pk::DevScene& D = s->dev;
    info6[0] = D.tw;
    info6[1] = D.th;
    info6[2] = D.tl;
    info6[3] = D.tfmt;
    info6[4] = D.env_mode ? D.ew : 0;
    info6[5] = D.env_mode ? D.eh : 0;
    return PK_OK;
  });
}

int pk_scene_write_params(pk_scene* s, int which, size_t offset, const float* src, size_t count, int src_on_device,
                          int also_sampling) {
  return pk_guard("pk_scene_write_params", [&]() -> int {
    if (!s || (count && !src) || (which != PK_PARAM_ATLAS && which != PK_PARAM_ENV)) {
      pk_set_error("pk_scene_write_params: null argument or unknown parameter set");
      return PK_E_INVALID;
    }
    if (which == PK_PARAM_ATLAS && s->dev.tfmt != PK_TEX_RGBA32F) {
      pk_set_error("pk_scene_write_params: atlas parameters need a scene uploaded with an RGBA32F atlas");
      return PK_E_INVALID;
    }
    const size_t n = param_floats(s, which);
    if (n == 0 || offset > n || count > n - offset) {
      pk_set_error("pk_scene_write_params: range [" + std::to_string(offset) + ", +" + std::to_string(count) +
                   ") outside the " + std::to_string(n) + " parameter floats (or the scene has no environment)");
      return PK_E_INVALID;
    }
    PK_HIP(hipSetDevice(s->ctx->device));
    hipStream_t st = s->ctx->stream;
    const hipMemcpyKind kind = src_on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    if (which == PK_PARAM_ATLAS) {
      float* samp = (float*)s->dev.tex;
      if (!also_sampling && !s->tex_eval) {
        int rc = dev_alloc(s->allocs, &s->tex_eval, n);
        if (rc) return rc;
        PK_HIP(hipMemcpyAsync(s->tex_eval, samp, n * sizeof(float), hipMemcpyDeviceToDevice, st));
        s->dev.tex_eval = s->tex_eval;
        s->gen++;
      }
      if (count && s->tex_eval)
        PK_HIP(hipMemcpyAsync(s->tex_eval + offset, src, count * sizeof(float), kind, st));
      if (count && also_sampling) PK_HIP(hipMemcpyAsync(samp + offset, src, count * sizeof(float), kind, st));
    } else if (count) {
      PK_HIP(hipMemcpyAsync((float*)s->dev.env_rad + offset, src, count * sizeof(float), kind, st));
    }
    if (!src_on_device) PK_HIP(hipStreamSynchronize(st));
    return PK_OK;
  });
}

int pk_scene_read_params(pk_scene* s, int which, size_t offset, float* dst, size_t count) {
  return pk_guard("pk_scene_read_params", [&]() -> int {
    if (!s || (count && !dst) || (which != PK_PARAM_ATLAS && which != PK_PARAM_ENV)) {
      pk_set_error("pk_scene_read_params: null argument or unknown parameter set");
      return PK_E_INVALID;
    }
    if (which == PK_PARAM_ATLAS && s->dev.tfmt != PK_TEX_RGBA32F) {
      pk_set_error("pk_scene_read_params: atlas parameters need a scene uploaded with an RGBA32F atlas");
      return PK_E_INVALID;
    }
    const size_t n = param_floats(s, which);
    if (n == 0 || offset > n || count > n - offset) {
      pk_set_error("pk_scene_read_params: range outside the parameter floats (or the scene has no environment)");
      return PK_E_INVALID;
    }
    PK_HIP(hipSetDevice(s->ctx->device));
    PK_HIP(hipStreamSynchronize(s->ctx->stream));
    const float* base = which == PK_PARAM_ENV ? (const float*)s->dev.env_rad : (const float*)s->dev.tex_eval;
    if (count) PK_HIP(hipMemcpy(dst, base + offset, count * sizeof(float), hipMemcpyDeviceToHost));
    return PK_OK;
  });
}

int pk_loss_l2(pk_frame* f, const float* d_target, float* d_grad, double* loss) {
  return pk_guard("pk_loss_l2", [&]() -> int {
    if (!f || !d_target || !d_grad || f->batch_scene) {
      pk_set_error("pk_loss_l2: null argument or a batch frame");
      return PK_E_INVALID;
    }
    PK_HIP(hipSetDevice(f->ctx->device));
    hipStream_t st = f->ctx->stream;
    if (!f->loss) {
      int rc = dev_alloc(f->allocs, &f->loss, 1);
      if (rc) return rc;
    }
    pk::launch_fill_u32((uint32_t*)f->loss, 0u, 2, st);
    pk::launch_loss_l2(f->dev.accum, (const float4*)d_target, (float4*)d_grad, f->loss, f->W * f->H, st);
    PK_HIP(hipGetLastError());
    if (loss) {
      PK_HIP(hipStreamSynchronize(st));
      PK_HIP(hipMemcpy(loss, f->loss, sizeof(double), hipMemcpyDeviceToHost));
    }
    return PK_OK;
  });
}

int pk_adjoint(pk_frame* f, pk_scene* s, const pk_render_params* p, const float* d_image, int d_image_on_device,
               const pk_grad_buffers* g) {
  return pk_guard("pk_adjoint", [&]() -> int {
    if (!f || !s || !p || !d_image || !g) {
      pk_set_error("pk_adjoint: null argument");
      return PK_E_INVALID;
    }
    if (f->batch_scene || s->roots.size() != 1) {
      pk_set_error("pk_adjoint: batch frames and merged scenes are not differentiable");
      return PK_E_INVALID;
    }
    if (p->width != f->W || p->height != f->H) {
      pk_set_error("pk_adjoint: params resolution differs from the frame's");
      return PK_E_INVALID;
    }
    if (!f->has_camera) {
      pk_set_error("pk_adjoint: no camera set (pk_frame_set_camera)");
      return PK_E_STATE;
    }
    if (p->max_depth == 0 || p->max_depth > 64 || p->num_frames == 0) {
      pk_set_error("pk_adjoint: max_depth outside [1, 64] or no frames");
      return PK_E_LIMIT;
    }
    const uint64_t paths = (uint64_t)f->W * f->H * p->num_frames;
    if (paths >= (1ull << 32) - 64) {
      pk_set_error("pk_adjoint: width*height*num_frames must stay below 2^32");
      return PK_E_LIMIT;
    }
    PK_HIP(hipSetDevice(f->ctx->device));
    hipStream_t st = f->ctx->stream;
    const size_t na = param_floats(s, PK_PARAM_ATLAS), ne = param_floats(s, PK_PARAM_ENV);
    int rc;
    if (!f->adj_cursor && (rc = dev_alloc(f->allocs, &f->adj_cursor, 1))) return rc;
    const float4* dimg = (const float4*)d_image;
    if (!d_image_on_device) {
      if (!f->adj_dimage && (rc = dev_alloc(f->allocs, &f->adj_dimage, (size_t)f->W * f->H))) return rc;
      PK_HIP(hipMemcpyAsync(f->adj_dimage, d_image, (size_t)f->W * f->H * 16, hipMemcpyHostToDevice, st));
      dimg = f->adj_dimage;
    }
    // gradient targets: the caller's device buffers, or frame scratch
    float* ga = g->on_device ? g->atlas : nullptr;
    float* ge = g->on_device ? g->env : nullptr;
    if (!ga) {
      if (f->adj_atlas_n < na) {
        if ((rc = dev_alloc(f->allocs, &f->adj_atlas, na))) return rc;
        f->adj_atlas_n = na;
      }
      ga = f->adj_atlas;
    }
    if (!ge) {
      if (f->adj_env_n < ne || !f->adj_env) {
        if ((rc = dev_alloc(f->allocs, &f->adj_env, ne ? ne : 4))) return rc;
        f->adj_env_n = ne ? ne : 4;
      }
      ge = f->adj_env;
    }
    // start values: zero, or the caller's gradients when accumulating
    const bool host_acc = !g->on_device && g->accumulate;
    if (host_acc && g->atlas) PK_HIP(hipMemcpyAsync(ga, g->atlas, na * 4, hipMemcpyHostToDevice, st));
    if (host_acc && g->env && ne) PK_HIP(hipMemcpyAsync(ge, g->env, ne * 4, hipMemcpyHostToDevice, st));
    const bool keep_a = g->accumulate && g->atlas, keep_e = g->accumulate && g->env;
    if (!keep_a) pk::launch_fill_u32((uint32_t*)ga, 0u, (uint32_t)na, st);
    if (ne && !keep_e) pk::launch_fill_u32((uint32_t*)ge, 0u, (uint32_t)ne, st);
    pk::launch_fill_u32(f->adj_cursor, 0u, 1, st);
    pk::AdjointArgs A;
    A.d_image = dimg;
    A.g_atlas = ga;
    A.g_env = ge;
    A.W = f->W;
    A.H = f->H;
    A.first_sample = p->first_sample;
    A.num_frames = p->num_frames;
    A.depth = p->max_depth;
    A.inv_frames = 1.0f / (float)p->num_frames;
    A.theta = p->env_theta;
    A.cam = f->cam;
    A.cursor = f->adj_cursor;
    A.spill = f->ctx->spill;
    A.overflow = f->dev.cnt + pk::kCntOverflow;
    // the spill area holds kMaxWavesPerCu >= kAdjointWavesPerCu waves per CU
    const uint32_t grid = (uint32_t)f->ctx->num_cus * kAdjointWavesPerCu;
    const size_t nvtx = (size_t)grid * p->max_depth * pk::kAdjVtxWords * 64;
    if (f->adj_vtx_n < nvtx) {
      if ((rc = dev_alloc(f->allocs, &f->adj_vtx, nvtx))) return rc;
      f->adj_vtx_n = nvtx;
    }
    A.vtx = f->adj_vtx;
    pk::launch_adjoint(s->dev, A, grid, f->ctx->stack_limit, st);
    PK_HIP(hipGetLastError());
    if (!g->on_device) {
      PK_HIP(hipStreamSynchronize(st));
      if (g->atlas) PK_HIP(hipMemcpy(g->atlas, ga, na * 4, hipMemcpyDeviceToHost));
      if (g->env && ne) PK_HIP(hipMemcpy(g->env, ge, ne * 4, hipMemcpyDeviceToHost));
    }
    return PK_OK;
  });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Autotuner.  The launch shape that wins depends on ray coherence: camera and
// shadow rays toward a sun are coherent and want whole waves refilled
// together, incoherent bounces and ray soups want early per-lane refills and
// more waves to hide memory latency.  A coordinate search times each class's
// kernels in isolation (HIP events on the context stream) and keeps the best.
// ---------------------------------------------------------------------------
namespace {

const uint32_t kTuneWaves[] = {8, 12, 16, 20, 24, 28, 32, 40};
const uint32_t kTuneRefill[] = {8, 16, 24, 32, 40, 48, 56, 64};
const uint32_t kTuneLeaf16[] = {2, 4, 6, 8, 10, 12, 14};

// The launch-shape dimensions after the refill threshold and the launch
// width: the leaf phase trigger, then the LDS stack part (each depth at the
// best launch width and at its own widest launch), then the waves per block,
// then (single-level scenes) the 4-wide node layout.  `trial` times one
// shape and keeps the fastest in `best`.
template <class Trial>
int tune_shape_dims(pk_ctx::Shape& best, bool with_width4, Trial trial) {
  int rc;
  for (uint32_t l : kTuneLeaf16) {
    pk_ctx::Shape t = best;
    t.leaf16 = l;
    if ((rc = trial(t))) return rc;
  }
  const pk_ctx::Shape base = best;
  for (uint32_t lds : {8u, 12u, 16u}) {
    if (lds == base.lds) continue;
    const uint32_t wmax = lds_max_waves(lds, 1);
    for (uint32_t w : {std::min(base.waves, wmax), wmax}) {
      pk_ctx::Shape t = base;
      t.lds = lds;
      t.wpb = 1;
      t.waves = w;
      if ((rc = trial(t))) return rc;
    }
  }
  const pk_ctx::Shape base2 = best;
  for (uint32_t wpb : {2u, 4u}) {
    pk_ctx::Shape t = base2;
    t.wpb = wpb;
    t.waves = std::max(wpb, std::min(base2.waves, lds_max_waves(base2.lds, wpb)) / wpb * wpb);
    if ((rc = trial(t))) return rc;
  }
  if (with_width4) {
    // 4-wide nodes: one wave per block, the 8- or 12-entry stack part
    const pk_ctx::Shape base3 = best;
    for (uint32_t lds : {8u, 12u}) {
      pk_ctx::Shape t = base3;
      t.width = 4;
      t.wpb = 1;
      t.lds = lds;
      t.waves = std::min(base3.waves, lds_max_waves(lds, 1));
      if ((rc = trial(t))) return rc;
    }
  }
  return PK_OK;
}

// Milliseconds per frame spent in each render traversal class.
int time_render_classes(pk_frame* f, pk_scene* s, uint32_t depth, uint32_t frames, double ms[3]) {
  hipStream_t st = f->ctx->stream;
  const size_t per_frame = 2 * (size_t)depth;
  std::vector<hipEvent_t> ev(2 * per_frame * frames);
  for (auto& e : ev) PK_HIP(hipEventCreate(&e));
  std::vector<int> cls;
  size_t ei = 0;
  pk::launch_begin(f->dev, 0, f->theta, f->cam, st);
  for (uint32_t fr = 0; fr < frames; ++fr) {
    pk::launch_raygen(f->dev, st);
    for (uint32_t b = 0; b < depth; ++b) {
      (void)hipEventRecord(ev[ei++], st);
      pk::launch_trace_closest(s->dev, f->dev, (int)b, trace_launch(f, closest_class(b)), st);
      (void)hipEventRecord(ev[ei++], st);
      cls.push_back(closest_class(b));
      pk::launch_shade(s->dev, f->dev, (int)b, st);
      (void)hipEventRecord(ev[ei++], st);
      pk::launch_trace_shadow(s->dev, f->dev, trace_launch(f, PK_TUNE_SHADOW), st);
      (void)hipEventRecord(ev[ei++], st);
      cls.push_back(PK_TUNE_SHADOW);
    }
    pk::launch_accumulate(f->dev, st);
    pk::launch_advance(f->dev, st);
  }
  hipError_t e = hipStreamSynchronize(st);
  for (int k = 0; k < 3; ++k) ms[k] = 0;
  for (size_t i = 0; i < cls.size() && e == hipSuccess; ++i) {
    float t = 0;
    e = hipEventElapsedTime(&t, ev[2 * i], ev[2 * i + 1]);
    ms[cls[i]] += t / frames;
  }
  for (auto& x : ev) (void)hipEventDestroy(x);
  if (e != hipSuccess) {
    pk_set_error(std::string("autotune: ") + hipGetErrorString(e));
    return PK_E_HIP;
  }
  return PK_OK;
}

}  // namespace

extern "C" int pk_autotune_render(pk_frame* f, pk_scene* s, uint32_t depth, uint32_t frames, float* ms2) {
  return pk_guard("pk_autotune_render", [&]() -> int {
    if (!f || !s || depth == 0 || depth > 64 || frames == 0) {
      pk_set_error("pk_autotune_render: null argument, depth outside [1, 64] or zero frames");
      return PK_E_INVALID;
    }
    if (!f->has_camera) {
      pk_set_error("pk_autotune_render: no camera set");
      return PK_E_STATE;
    }
    pk_ctx* c = f->ctx;
    const bool count = f->count_nodes;
    f->count_nodes = false;
    double base[3];
    int rc = time_render_classes(f, s, depth, frames, base);
    if (rc) return rc;
    const double before = base[0] + base[1] + base[2];
    // per class: the refill threshold at the widest launch, then the launch
    // width at that threshold
    for (int cls = PK_TUNE_PRIMARY; cls <= PK_TUNE_SHADOW; ++cls) {
      pk_ctx::Shape best = c->shape[cls];
      double best_ms = base[cls];
      auto trial_shape = [&](const pk_ctx::Shape& sh) -> int {
        c->shape[cls] = sh;
        double ms[3];
        const int e = time_render_classes(f, s, depth, frames, ms);
        if (!e && ms[cls] < best_ms) {
          best_ms = ms[cls];
          best = c->shape[cls];
        }
        return e;
      };
      auto trial = [&](uint32_t w, uint32_t r) -> int {
        pk_ctx::Shape t = best;
        t.waves = std::min(w, lds_max_waves(t.lds, t.wpb)) / t.wpb * t.wpb;
        t.refill = r;
        return trial_shape(t);
      };
      for (uint32_t r : kTuneRefill)
        if ((rc = trial(best.waves, r))) return rc;
      const uint32_t r_best = best.refill;
      for (uint32_t w : kTuneWaves)
        if ((rc = trial(w, r_best))) return rc;
      if ((rc = tune_shape_dims(best, s->dev.nodes4 != nullptr, trial_shape))) return rc;
      c->shape[cls] = best;
    }
    double after[3];
    if ((rc = time_render_classes(f, s, depth, frames, after))) return rc;
    c->tuning_gen++;
    f->count_nodes = count;
    if (ms2) {
      ms2[0] = (float)before;
      ms2[1] = (float)(after[0] + after[1] + after[2]);
    }
    return pk_frame_reset(f);
  });
}

// Queue capacity: samples per pass in 1, 2, 4, ... up to max_samples (and
// below kTunePathCap paths), each timed over max_samples samples of
// graph-replayed passes; keeps the fastest per sample.
constexpr uint64_t kTunePathCap = 1ull << 30;  // the frame's own limit

extern "C" int pk_autotune_samples_per_pass(pk_frame* f, pk_scene* s, uint32_t depth, uint32_t max_samples,
                                            float* ms2) {
  return pk_guard("pk_autotune_samples_per_pass", [&]() -> int {
    if (!f || !s || depth == 0 || depth > 64 || max_samples == 0 || f->batch_scene) {
      pk_set_error("pk_autotune_samples_per_pass: null argument, depth outside [1, 64], zero samples or a batch");
      return PK_E_INVALID;
    }
    if (!f->has_camera) {
      pk_set_error("pk_autotune_samples_per_pass: no camera set");
      return PK_E_STATE;
    }
    pk_ctx* c = f->ctx;
    hipStream_t st = c->stream;
    const uint32_t start = f->dev.S;
    uint32_t best_S = start;
    float best = 0, at_start = 0;
    pk_render_params p{f->W, f->H, depth, 0, max_samples, f->theta, 1};
    int rc = PK_OK;
    for (uint32_t S = 1; S <= max_samples; S *= 2) {
      if ((uint64_t)f->W * f->H * S >= kTunePathCap) break;
      // down to one sample first: a resize holds the old queues and the new
      // ones together, and the largest sizes only fit alone
      if (S > 2 && (rc = pk_frame_set_samples_per_pass(f, 1))) break;
      if ((rc = pk_frame_set_samples_per_pass(f, S))) {
        if (rc == PK_E_OOM && S > 1) rc = PK_OK;  // device memory bounds the capacity: the search ends there
        break;
      }
      pk_render_params warm = p;
      warm.num_frames = S;
      if ((rc = pk_render(f, s, &warm))) break;
      PK_HIP(hipEventRecord(c->ev0, st));
      if ((rc = pk_render(f, s, &p))) break;
      PK_HIP(hipEventRecord(c->ev1, st));
      PK_HIP(hipEventSynchronize(c->ev1));
      float ms = 0;
      PK_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
      ms /= (float)max_samples;
      if (S == start) at_start = ms;
      if (best == 0 || ms < best) {
        best = ms;
        best_S = S;
      }
    }
    int rc2 = best_S == f->dev.S ? PK_OK : pk_frame_set_samples_per_pass(f, 1);
    if (!rc2) rc2 = pk_frame_set_samples_per_pass(f, best_S);
    if (rc || rc2) return rc ? rc : rc2;
    if (ms2) {
      ms2[0] = at_start;
      ms2[1] = best;
    }
    return pk_frame_reset(f);
  });
}

// Pass scheduler: times `samples` samples of graph-replayed passes under each
// scheduler (after one warm pass, which also records the graph) and keeps
// the faster.  Which wins depends on the scene: the wavefront kernels refill
// lanes inside traversal and run more waves per SIMD, the megakernel saves
// the queues' memory traffic and the per-bounce launches.
extern "C" int pk_autotune_scheduler(pk_frame* f, pk_scene* s, uint32_t depth, uint32_t samples, float* ms2) {
  return pk_guard("pk_autotune_scheduler", [&]() -> int {
    if (!f || !s || depth == 0 || depth > 64 || samples == 0 || f->batch_scene) {
      pk_set_error("pk_autotune_scheduler: null argument, depth outside [1, 64], zero samples or a batch");
      return PK_E_INVALID;
    }
    if (!f->has_camera) {
      pk_set_error("pk_autotune_scheduler: no camera set");
      return PK_E_STATE;
    }
    pk_ctx* c = f->ctx;
    hipStream_t st = c->stream;
    const int start = c->sched;
    float ms[2] = {0, 0};
    pk_render_params p{f->W, f->H, depth, 0, samples, f->theta, 1};
    int rc = PK_OK;
    for (int sched : {PK_SCHED_WAVEFRONT, PK_SCHED_PATHS}) {
      if ((rc = pk_ctx_set_scheduler(c, sched))) break;
      pk_render_params warm = p;
      warm.num_frames = f->dev.S;
      if ((rc = pk_render(f, s, &warm))) break;
      PK_HIP(hipEventRecord(c->ev0, st));
      if ((rc = pk_render(f, s, &p))) break;
      PK_HIP(hipEventRecord(c->ev1, st));
      PK_HIP(hipEventSynchronize(c->ev1));
      PK_HIP(hipEventElapsedTime(&ms[sched], c->ev0, c->ev1));
      ms[sched] /= (float)samples;
    }
    if (rc) {
      (void)pk_ctx_set_scheduler(c, start);
      return rc;
    }
    const int best = ms[PK_SCHED_PATHS] < ms[PK_SCHED_WAVEFRONT] ? PK_SCHED_PATHS : PK_SCHED_WAVEFRONT;
    if ((rc = pk_ctx_set_scheduler(c, best))) return rc;
    if (ms2) {
      ms2[0] = ms[start];
      ms2[1] = ms[best];
    }
    return pk_frame_reset(f);
  });
}

extern "C" int pk_autotune_rays(pk_ctx* c, pk_scene* s, const float* d_rays, int64_t n, int any_hit, int32_t* d_ids,
                                float* d_t, uint32_t* d_bary, float* ms2) {
  return pk_guard("pk_autotune_rays", [&]() -> int {
    int rc = check_rays(c, s, n, "pk_autotune_rays");
    if (rc) return rc;
    const int cls = any_hit ? PK_TUNE_RAYS_ANY : PK_TUNE_RAYS_CLOSEST;
    float before = 0;
    if ((rc = pk_time_trace_rays(c, s, d_rays, n, any_hit, d_ids, d_t, d_bary, 2, &before))) return rc;
    pk_ctx::Shape best = c->shape[cls];
    float best_ms = before;
    auto trial_shape = [&](const pk_ctx::Shape& sh) -> int {
      c->shape[cls] = sh;
      float ms = 0;
      const int e = pk_time_trace_rays(c, s, d_rays, n, any_hit, d_ids, d_t, d_bary, 2, &ms);
      if (!e && ms < best_ms) {
        best_ms = ms;
        best = c->shape[cls];
      }
      return e;
    };
    for (uint32_t w : kTuneWaves)
      for (uint32_t r : {1u, 8u, 16u, 24u, 32u, 48u, 64u}) {
        pk_ctx::Shape t = best;
        t.waves = std::min(w, lds_max_waves(t.lds, t.wpb)) / t.wpb * t.wpb;
        t.refill = r;
        if ((rc = trial_shape(t))) return rc;
      }
    if ((rc = tune_shape_dims(best, s->dev.nodes4 != nullptr, trial_shape))) return rc;
    c->shape[cls] = best;
    c->tuning_gen++;
    if (ms2) {
      ms2[0] = before;
      ms2[1] = best_ms;
    }
    return PK_OK;
  });
}
