/* pk.h — C ABI of the MI355X wavefront path tracer (libpike_amd.so).
 *
 * The JavaScript render passes (pike_amd/js, through the N-API addon in
 * pike_amd/napi) and the Python harness (pike_amd/__init__.py, ctypes) call
 * only these entry points.  Every call returns PK_OK (0) or a negative PK_E_*
 * code; pk_last_error() gives the message of the last failure on the calling
 * thread.  Handles are opaque.  Host arrays passed in are copied (or read to
 * completion) before the call returns.  Device work is asynchronous on the
 * context's stream; pk_sync() and the pk_read_* calls are the blocking ones.
 * A context is thread-compatible, not thread-safe (one JS thread drives it,
 * as in the reference).
 *
 * Reference interfaces each group replaces (apbodnar/Pike, paths relative to
 * the repository root):
 *   context            render/pike.js:92-118 (initWebGpu: adapter + device)
 *   scene upload       render/passes/trace_pass.js:220-263 (initBVHBuffer),
 *                      render/passes/shade_hit_pass.js:58-81,92-167 (atlas,
 *                      attributes, materials, env tables),
 *                      render/env_sampler.js:7-157 (EnvironmentGenerator)
 *   BVH build          render/bvh.js:7-108 (BVH), :110-275 (Node)
 *   frame / passes     render/passes/camera_pass.js:3-77 (RenderState),
 *                      :79-221 (CameraPass), trace_pass.js:265-276,
 *                      trace_shadow_pass.js:42-52, shade_hit_pass.js:169-181,
 *                      shade_miss_pass.js:68-80, accumulate_pass.js:75-85,
 *                      render/pike.js:120-162 (tick)
 */
#ifndef PK_H_
#define PK_H_
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Every entry point returns one of these (pk_last_error() holds the text of
 * the last failure on the calling thread); none lets a C++ exception escape:
 * host memory or threads running out return PK_E_OOM. */
#define PK_OK 0
#define PK_E_INVALID (-1)     /* bad argument or handle */
#define PK_E_NO_DEVICE (-2)   /* HIP device missing or unusable */
#define PK_E_OOM (-3)         /* device or host allocation failed */
#define PK_E_HIP (-4)         /* a HIP runtime call failed */
#define PK_E_LIMIT (-5)       /* size beyond a documented limit */
#define PK_E_STATE (-6)       /* call out of order (e.g. trace before raygen) */
#define PK_E_OVERFLOW (-7)    /* traversal stack overflow flagged by a kernel */

/* texture atlas formats (render/util/texture_packer.js:118-136 produces RGBA8) */
#define PK_TEX_RGBA8 0
#define PK_TEX_RGBA32F 1
/* environment formats: RGBE-in-alpha RGBA8 as env_sampler.js:8-48 reads it,
 * or linear float RGB */
#define PK_ENV_NONE 0
#define PK_ENV_RGBE8 1
#define PK_ENV_RGB32F 2
/* environment importance sampler */
#define PK_ENV_SAMPLER_BINS 0 /* equal-radiance bins, env_sampler.js:74-157 */
#define PK_ENV_SAMPLER_CDF 1  /* marginal/conditional CDF, device binary search */

typedef struct pk_ctx pk_ctx;
typedef struct pk_scene pk_scene;
typedef struct pk_frame pk_frame;
typedef struct pk_bvh pk_bvh;

/* Scene in the reference's host representation (render/scene.js:120-181):
 * positions are world-space float64 (attributes[].pos), triangles are
 * scene.indices {i0, i1, i2, matId}, attributes are the quantized
 * {tangent, bitangent, normal (snorm8x4), uv (snorm16x2)} words, materials are
 * {diffMap, metRoughMap, normMap, emitMap} atlas layer ids.  The atlas is
 * tex_layers layers of tex_width x tex_height texels. */
typedef struct pk_scene_desc {
  uint32_t num_vertices;
  const double* positions;    /* [num_vertices][3] */
  const uint32_t* attributes; /* [num_vertices][4] */
  uint32_t num_triangles;
  const int32_t* triangles;   /* [num_triangles][4] */
  uint32_t num_materials;
  const int32_t* materials;   /* [num_materials][4] */
  uint32_t tex_width, tex_height, tex_layers, tex_format;
  const void* tex_data;       /* RGBA8: 4 B/texel, RGBA32F: 16 B/texel */
  uint32_t env_width, env_height, env_format, env_sampler;
  const void* env_data;       /* RGBE8: 4 B/texel, RGB32F: 12 B/texel */
} pk_scene_desc;

/* CameraState (render/util/structs.js:235-249, camera.wgsl:5-14), in the
 * normalized scene space the tracer works in. */
typedef struct pk_camera {
  float origin[3];
  float dir[3];
  float fov;          /* radians */
  float focal_depth;
  float aperture;
  float distortion;
  float bokeh;
} pk_camera;

typedef struct pk_render_params {
  uint32_t width, height;
  uint32_t max_depth;     /* numBounces (render/pike.js:20) */
  uint32_t first_sample;  /* RenderState.samples at the first frame */
  uint32_t num_frames;    /* one sample per pixel per frame */
  float env_theta;        /* RenderState.env_theta */
  uint32_t use_graph;     /* replay frames through one captured hipGraph */
} pk_render_params;

/* Per-frame device counters, read back by pk_read_counters (the reference
 * keeps num_hits/num_misses/num_rays/numShadowRays in RenderState). */
typedef struct pk_counters {
  uint64_t rays_traced;    /* closest-hit rays over the frames rendered */
  uint64_t shadow_rays;    /* any-hit rays */
  uint64_t hits;
  uint64_t nodes_visited;  /* closest-hit kernel; only when node counting is enabled */
  uint64_t tris_tested;
  uint64_t shadow_nodes;   /* any-hit kernel */
  uint64_t shadow_tris;
  uint32_t stack_overflow; /* nonzero if any ray overflowed its stack */
  uint32_t frames;
} pk_counters;

const char* pk_last_error(void);
const char* pk_version(void);

/* ---- context (render/pike.js:92-97) ---- */
int pk_ctx_create(int device, pk_ctx** out);
int pk_ctx_destroy(pk_ctx* ctx);
int pk_sync(pk_ctx* ctx);
/* Launch shape of the persistent traversal kernels, per kernel class (no
 * reference counterpart; chosen by the autotuner): waves_per_cu persistent
 * waves per CU (at most what the class's LDS allows, below), refill_min in
 * [1, 64] idle lanes that trigger a wave's ray refill (64 = refill only when
 * the whole wave is idle, which keeps coherent rays together). */
#define PK_TUNE_PRIMARY 0      /* closest hit, bounce 0 (camera rays) */
#define PK_TUNE_SECONDARY 1    /* closest hit, bounces >= 1 */
#define PK_TUNE_SHADOW 2       /* any hit, NEE shadow rays */
#define PK_TUNE_RAYS_CLOSEST 3 /* pk_trace_rays*, closest hit */
#define PK_TUNE_RAYS_ANY 4     /* pk_trace_rays*, any hit */
#define PK_TUNE_PATHS 5        /* path megakernel (PK_SCHED_PATHS); waves_per_cu in [1, 16], refill_min unused */
#define PK_TUNE_CLASSES 6
int pk_ctx_set_tuning(pk_ctx* ctx, int kernel_class, uint32_t waves_per_cu, uint32_t refill_min);
int pk_ctx_get_tuning(const pk_ctx* ctx, int kernel_class, uint32_t* waves_per_cu, uint32_t* refill_min,
                      uint32_t* num_cus);
/* The whole launch shape of a kernel class (no reference counterpart; the
 * reference's trace shaders have a fixed 128-thread workgroup and a 24+8
 * entry stack, trace.wgsl:5-9, 191-213).  pk_ctx_set_tuning changes the
 * first two fields and keeps the others.
 *   waves_per_cu     persistent traversal waves per CU, at most what the LDS
 *                    allows: each wave holds its stack part and 1 KB for the
 *                    cooperative leaf phase (8 entries: 32; 12: 22 with 1 or
 *                    2 waves per block, 20 with 4; 16: 17 with 1, 16 with 2
 *                    or 4)
 *   refill_min       idle lanes that trigger a wave's ray refill, [1, 64]
 *   stack_lds        traversal stack entries per lane kept in LDS (8, 12 or
 *                    16; the rest of the stack_limit entries spill to global
 *                    memory); the counting and instanced variants use 12
 *   waves_per_block  waves of one block (1, 2 or 4), each independent
 *   leaf_phase16     the wave runs its deferred leaf tests together once
 *                    this many 16ths of its tracing lanes hold some, [1, 16]
 *   node_width       BVH nodes the traversal walks: 8 (128-byte nodes) or 4
 *                    (the 64-byte layout pk_bvh_export_wide4 describes; with
 *                    stack_lds 8 or 12 and 1 wave per block); instanced
 *                    scenes and the counting variants walk 8-wide nodes
 * The path megakernel class (PK_TUNE_PATHS) takes stack_lds 12, 1 wave per
 * block and node_width 8.  Every choice traces the same hits and renders the
 * same image. */
typedef struct pk_launch_shape {
  uint32_t waves_per_cu;
  uint32_t refill_min;
  uint32_t stack_lds;
  uint32_t waves_per_block;
  uint32_t leaf_phase16;
  uint32_t node_width;
} pk_launch_shape;
int pk_ctx_set_launch(pk_ctx* ctx, int kernel_class, const pk_launch_shape* shape);
int pk_ctx_get_launch(const pk_ctx* ctx, int kernel_class, pk_launch_shape* shape);
/* Traversal stack entries per ray, in [1, 64] (default 64: 12 in LDS, the
 * rest in a per-wave global spill area).  The reference's stack holds 32
 * entries and has no overflow check (trace.wgsl:5-9, 134-146).  A ray that
 * needs more entries than the limit skips the subtrees it cannot push; the
 * kernels end normally and report it: pk_read_counters().stack_overflow,
 * PK_E_OVERFLOW from pk_trace_rays / pk_ctx_trace_stats.  Lower limits cap
 * the work per ray (and let tests force the overflow path). */
int pk_ctx_set_stack_limit(pk_ctx* ctx, uint32_t entries);
int pk_ctx_get_stack_limit(const pk_ctx* ctx, uint32_t* entries);
/* Device memory cap of the context's scenes, frames and their scratch (no
 * reference counterpart: WebGPU reports allocation failure through its
 * validation errors only).  An allocation that would pass `bytes` fails with
 * PK_E_OOM and leaves the call's objects unchanged; 0 removes the cap.
 * pk_ctx_memory_used gives the bytes currently allocated. */
int pk_ctx_set_memory_limit(pk_ctx* ctx, uint64_t bytes);
int pk_ctx_memory_used(const pk_ctx* ctx, uint64_t* bytes);
/* Pass scheduler of frames (no reference counterpart: the
 * reference always runs its passes as wavefront compute dispatches,
 * render/pike.js:135-155).  PK_SCHED_WAVEFRONT: raygen, then per bounce
 * trace_closest / shade / trace_shadow over queues in device memory.
 * PK_SCHED_PATHS: one persistent megakernel in which each lane carries a
 * path through every bounce in registers.  Both give bit-identical color
 * buffers, for single-view frames and batches (pk_batch_render) alike; node
 * counting (pk_frame_count_nodes) and pk_render_profile always use the
 * wavefront kernels.  Captured graphs are re-recorded after a change. */
#define PK_SCHED_WAVEFRONT 0
#define PK_SCHED_PATHS 1
int pk_ctx_set_scheduler(pk_ctx* ctx, int scheduler);
int pk_ctx_get_scheduler(const pk_ctx* ctx, int* scheduler);
/* Autotuner: coordinate search over the launch shape of each render kernel
 * class (refill_min in {8, 16, ..., 64} at the current width, then
 * waves_per_cu in {8, 12, ..., 32, 40} (capped by the LDS) at the best
 * threshold, then leaf_phase16 in {2, 4, ..., 14}, then stack_lds 8 / 12 / 16
 * at the best width and at that depth's widest launch, then 2 and 4 waves
 * per block, then for single-level scenes the 4-wide nodes), timing that
 * class's kernels with HIP events over `frames` profiled frames of this
 * scene, camera and resolution; keeps the fastest shape per class in the
 * context.  ms2 receives the traversal milliseconds per frame before and
 * after. */
int pk_autotune_render(pk_frame* frame, pk_scene* scene, uint32_t max_depth, uint32_t frames, float* ms2);
/* Queue capacity: chooses the frame's samples per pass (1, 2, 4, ... up to
 * max_samples, below 2^30 paths and within device memory: a size whose
 * queues cannot be allocated ends the search) by timing max_samples samples of
 * graph-replayed passes of this scene, camera and resolution; leaves the
 * fastest set and the frame reset.  ms2 = milliseconds per sample at the
 * frame's previous setting (0 if it was not among the candidates) and at the
 * chosen one. */
int pk_autotune_samples_per_pass(pk_frame* frame, pk_scene* scene, uint32_t max_depth, uint32_t max_samples,
                                 float* ms2);
/* Pass scheduler: times `samples` samples of graph-replayed passes of this
 * scene, camera and resolution under PK_SCHED_WAVEFRONT and PK_SCHED_PATHS
 * (with the context's launch shapes and the frame's samples per pass) and
 * leaves the faster set in the context and the frame reset.  ms2 =
 * milliseconds per sample under the scheduler set before and the chosen one. */
int pk_autotune_scheduler(pk_frame* frame, pk_scene* scene, uint32_t max_depth, uint32_t samples, float* ms2);
/* The same for the standalone ray queries on device-resident rays. */
int pk_autotune_rays(pk_ctx* ctx, pk_scene* scene, const float* d_rays, int64_t n, int any_hit, int32_t* d_ids,
                     float* d_t, uint32_t* d_bary, float* ms2);

/* ---- device buffers (GPUDevice.createBuffer / queue.writeBuffer /
 * mapAsync in the reference's WebGPU use, e.g. structs.js:148-157) ---- */
int pk_buffer_create(pk_ctx* ctx, size_t bytes, void** device_ptr);
int pk_buffer_destroy(pk_ctx* ctx, void* device_ptr);
int pk_buffer_write(pk_ctx* ctx, void* device_dst, const void* host_src, size_t bytes);  /* synchronous */
int pk_buffer_read(pk_ctx* ctx, void* host_dst, const void* device_src, size_t bytes);   /* synchronous */
int pk_buffer_copy(pk_ctx* ctx, void* device_dst, const void* device_src, size_t bytes);  /* async on the stream */

/* ---- BVH (render/bvh.js) : host-side, no device needed ----
 * Builds the SAH binary BVH with bvh.js's decisions (full sweep, float64,
 * first minimum, leaf iff cost > N) and collapses it into the 128-byte wide
 * layout the kernels traverse.  num_threads <= 0: every hardware thread; the
 * result does not depend on the thread count.  PK_E_INVALID for an empty
 * input, a vertex index out of range, or a referenced vertex with a
 * non-finite coordinate. */
int pk_bvh_build(const double* positions, uint32_t num_vertices, const int32_t* triangles,
                 uint32_t num_triangles, int num_threads, pk_bvh** out);
/* The same with options.  PK_BVH_SPATIAL_SPLITS makes bvh.js's SBVH path
 * live (Node :119-143, getSpatialSplit :231-274, the check the reference
 * leaves disabled): spatial splits over 64 slabs when the object split's
 * children overlap by more than 1e-5 of the root area.  Leaves then hold
 * clipped fragments, so a triangle can appear in several leaves; the
 * closest/any hit answer is unchanged. */
#define PK_BVH_SPATIAL_SPLITS 1u
/* PK_BVH_SAH_COLLAPSE collapses the binary tree into the 8-wide layout by the
 * SAH-optimal dynamic program (a subtree becomes a wide node, a leaf of up to
 * 8 triangles, or is spread over its parent's child slots) instead of
 * greedily opening the largest child.  Hits are unchanged. */
#define PK_BVH_SAH_COLLAPSE 2u
int pk_bvh_build_ex(const double* positions, uint32_t num_vertices, const int32_t* triangles,
                    uint32_t num_triangles, int num_threads, uint32_t flags, pk_bvh** out);
int pk_bvh_destroy(pk_bvh* bvh);
/* sizes: [0] binary nodes, [1] wide nodes, [2] leaf triangles, [3] binary depth,
 * [4] wide depth, [5] largest leaf */
int pk_bvh_info(const pk_bvh* bvh, int64_t* info6);
/* The reference's flattened buffers (trace_pass.js:220-263): BFS nodes of 5
 * words {child_base_idx, triangles, x, y, z} with snorm16 boxes, triangles
 * {i0,i1,i2,mat} in leaf order, normalized float32 vertices padded to 16 B,
 * and the normalizer (center xyz, scale) it applied. */
int pk_bvh_export_reference(const pk_bvh* bvh, int32_t* nodes5, int32_t* tris4, float* verts4,
                            double* normalizer4);
/* The wide layout itself (no reference counterpart; for inspection and the
 * CPU traversal cost model): wide_nodes * 128 bytes, leaf_triangles * 48
 * bytes of triangle records, and leaf_triangles * 4 words {i0,i1,i2,mat}. */
int pk_bvh_export_wide(const pk_bvh* bvh, void* nodes128, void* tris48, int32_t* tri_index4);
/* The 4-wide layout derived from a single-level BVH's wide layout (no
 * reference counterpart; the layout the node_width 4 traversal reads):
 * *count 64-byte nodes.  With nodes64 null, only *count is set; otherwise
 * *count is the capacity of nodes64 in nodes on entry. */
int pk_bvh_export_wide4(const pk_bvh* bvh, void* nodes64, int64_t* count);
/* The normalizer (center xyz, scale) of any BVH, instanced ones included. */
int pk_bvh_normalizer(const pk_bvh* bvh, double* normalizer4);

/* ---- instancing: a two-level BVH (the reference bakes every prop into one
 * triangle list, scene.js:159-170, and lists instancing as a TODO,
 * README.md:20).  A mesh is geometry in its own object space; an instance
 * places a mesh in the scene with an affine transform and may override its
 * material.  One BVH is built per mesh and one over the instances' boxes;
 * the scene's own triangles (pk_scene_desc positions/triangles, possibly
 * none) form an implicit first instance with the identity transform.
 * Triangle ids are those of the equivalent baked scene: the scene's own
 * triangles first, then each instance's mesh triangles in instance order.
 * Limits: fewer than 2^31 baked triangles and 2^31 instances. ---- */
typedef struct pk_mesh_desc {
  uint32_t num_vertices;
  const double* positions;    /* object space [num_vertices][3] */
  const uint32_t* attributes; /* [num_vertices][4] object-space tangent frame + uv, or NULL */
  uint32_t num_triangles;
  const int32_t* triangles;   /* [num_triangles][4] {i0, i1, i2, mat} */
} pk_mesh_desc;

typedef struct pk_instance_desc {
  uint32_t mesh;              /* index into the meshes */
  int32_t material;           /* material of all its triangles; -1: the mesh's own */
  double transform[12];       /* object -> scene space, row-major 3x4 (invertible) */
} pk_instance_desc;

/* Builds the two-level BVH; pk_scene_upload then takes it with the scene
 * desc (its materials, atlas, environment and the attributes of its own
 * vertices).  flags as for pk_bvh_build_ex (applied to every mesh).  The
 * reference layout export is not available for instanced BVHs. */
int pk_bvh_build_instanced(const pk_scene_desc* scene, const pk_mesh_desc* meshes, uint32_t num_meshes,
                           const pk_instance_desc* instances, uint32_t num_instances, int num_threads,
                           uint32_t flags, pk_bvh** out);

/* ---- scene (TracePass / ShadeHitPass / ShadeMissPass resources) ---- */
int pk_scene_upload(pk_ctx* ctx, const pk_scene_desc* desc, const pk_bvh* bvh, pk_scene** out);
int pk_scene_destroy(pk_scene* scene);
/* environment sampling tables as built (for inspection): bins count, coords count */
int pk_scene_env_info(const pk_scene* scene, int64_t* info4);
/* Host-only environment tables (EnvironmentGenerator, env_sampler.js:74-157:
 * createHistogramBuffer / createEnvCoordBuffer / createPdfTexture).  Call with
 * NULL arrays to get info4 = {bins, coords, width, height}; then bins2 is
 * [bins][2] {h0, h1}, coords2 [coords][2] {x, y}, pdf [height][width]. */
int pk_env_tables(const pk_scene_desc* desc, int64_t* info4, int32_t* bins2, int32_t* coords2, float* pdf);

/* ---- frame: RenderState + per-pass entry points ---- */
int pk_frame_create(pk_ctx* ctx, uint32_t width, uint32_t height, pk_frame** out);
int pk_frame_destroy(pk_frame* frame);
int pk_frame_reset(pk_frame* frame);                       /* RenderState.resetSamples */
int pk_frame_set_camera(pk_frame* frame, const pk_camera* cam);  /* CameraPass.setCameraState */
int pk_frame_set_env_theta(pk_frame* frame, float theta); /* RenderState.setEnvRotation */
int pk_frame_begin(pk_frame* frame, uint32_t sample);      /* RenderState.generateCommands + clears */
int pk_raygen(pk_frame* frame);                            /* CameraPass.generateCommands */
int pk_trace_closest(pk_frame* frame, pk_scene* scene, uint32_t bounce);  /* TracePass + miss shading */
int pk_shade(pk_frame* frame, pk_scene* scene, uint32_t bounce);          /* ShadeHitPass */
int pk_trace_shadow(pk_frame* frame, pk_scene* scene);     /* TraceShadowPass + miss shading */
int pk_accumulate(pk_frame* frame);                        /* AccumulatePass */
/* A whole run of frames (tick() x num_frames), optionally graph-replayed. */
int pk_render(pk_frame* frame, pk_scene* scene, const pk_render_params* params);
/* Samples per pass (no reference counterpart; default 1): each pass of
 * pk_render and of the per-pass entry points traces `samples` consecutive
 * samples of every pixel as one wavefront (width*height*samples paths, below
 * 2^30; 176 bytes of queues per path), so small frames fill the GPU.  The accumulated image equals that of
 * one-sample frames bit for bit.  On a batch frame (pk_batch_create) it sets
 * the samples per round of pk_batch_render (total view pixels * samples
 * paths).  Reallocates the frame's queues. */
int pk_frame_set_samples_per_pass(pk_frame* frame, uint32_t samples);
int pk_frame_get_samples_per_pass(const pk_frame* frame, uint32_t* samples);
/* Running-mean image, float RGBA, width*height*16 bytes. */
int pk_read_image(pk_frame* frame, float* rgba, size_t bytes);
/* The same image copied to device memory (asynchronous, context stream). */
int pk_read_image_device(pk_frame* frame, float* d_rgba, size_t bytes);
/* The current frame's color buffer (RenderState.colorBuffer) before
 * accumulation: with several samples per pass, the first sample's. */
int pk_read_color(pk_frame* frame, float* rgba, size_t bytes);
int pk_read_counters(pk_frame* frame, pk_counters* out);
/* Count BVH nodes visited and triangles tested in this frame's trace kernels
 * (adds per-lane counters; the uncounted kernels are the timed ones). */
int pk_frame_count_nodes(pk_frame* frame, int enable);
/* Kernel-level timing of num_frames whole frames (no graph): HIP events
 * around every launch on the context stream.  ms5 receives the total
 * milliseconds spent in raygen, trace_closest, shade, trace_shadow and
 * accumulate.  Renders exactly what pk_render renders. */
int pk_render_profile(pk_frame* frame, pk_scene* scene, const pk_render_params* params, double* ms5);

/* ---- batches: many scenes and camera views in one launch (no reference
 * counterpart: Pike renders one scene and one camera).  The scenes are merged
 * into one device scene (every scene keeps its own normalized space and BVH;
 * they must share the atlas layer size/format, and the environment of
 * descs[0] lights all of them).  A batch frame lays every view's pixels out
 * in one global wavefront; pk_batch_render traces all sample rounds of all
 * views (round r covers the views with spp > r) and, with use_graph,
 * replays them as one captured hipGraph launch.  Per-pixel results equal
 * pk_render of that view alone (seeds depend on the view-local pixel, the
 * sample and the bounce only). ---- */
typedef struct pk_view {
  uint32_t scene;  /* index into the merged scenes */
  uint32_t width, height, spp;
  pk_camera camera; /* in that scene's normalized space */
} pk_view;
int pk_scene_upload_batch(pk_ctx* ctx, const pk_scene_desc* descs, const pk_bvh* const* bvhs, uint32_t count,
                          pk_scene** out);
/* Root wide node of each merged scene; returns the number of scenes. */
int pk_scene_roots(const pk_scene* scene, uint32_t* roots, uint32_t count);
int pk_batch_create(pk_ctx* ctx, pk_scene* merged, const pk_view* views, uint32_t nviews, pk_frame** out);
int pk_batch_render(pk_frame* batch, uint32_t max_depth, float env_theta, int use_graph);
/* Running mean of one view, width*height RGBA float. */
int pk_batch_read_view(pk_frame* batch, uint32_t view, float* rgba, size_t bytes);

/* ---- standalone ray queries (traversal microbenchmark) ----
 * rays: [n][8] floats {ox, oy, oz, 0, dx, dy, dz, 0} in normalized space.
 * Outputs per ray: original triangle id (-1 on miss), t, bary mask
 * (pack2x16unorm(u, v)).  any_hit selects occlusion (trace_shadow.wgsl
 * semantics: epsilon 1e-7, returns whether anything is hit; id is -1 or the
 * id of some hit triangle). */
int pk_trace_rays(pk_ctx* ctx, pk_scene* scene, const float* rays, int64_t n, int any_hit,
                  int32_t* ids, float* t, uint32_t* bary);
/* Device-resident variant: pointers are device addresses (hipMalloc'd or
 * torch tensors); asynchronous on the context stream.  count_nodes adds the
 * per-ray node/triangle totals to the context's traversal counters. */
int pk_trace_rays_device(pk_ctx* ctx, pk_scene* scene, const float* d_rays, int64_t n, int any_hit,
                         int32_t* d_ids, float* d_t, uint32_t* d_bary, int count_nodes);
/* Allocation-free variant of pk_trace_rays_device (no counterpart in the
 * reference).  It is for callers that cannot allocate device memory: code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc.  The kernels' work cursor and traversal stack spill area live in
 * a file-scope device array of the library (58.7 MB per device, held until
 * the process ends) instead of context memory; the call allocates nothing,
 * does not block and launches on the scene's context stream.  The launch
 * uses the context's ray-query shape with at most 2048 one-wave blocks (the
 * stacks the array holds).  Same limits (n < 2^31), results and overflow
 * reporting (pk_ctx_trace_stats) as pk_trace_rays_device, without node
 * counting.  Calls that overlap on one device share that array and race:
 * order them on one stream. */
int pk_trace_rays_noalloc(pk_scene* scene, const float* d_rays, int64_t n, int any_hit, int32_t* d_ids, float* d_t,
                          uint32_t* d_bary);
/* Traversal statistics of the last counted pk_trace_rays_device call:
 * [0] nodes visited, [1] triangles tested. */
int pk_ctx_trace_stats(pk_ctx* ctx, uint64_t* stats2);
/* Time the device-resident trace kernel: runs it `iters` times with events on
 * the context stream and returns the mean kernel milliseconds. */
int pk_time_trace_rays(pk_ctx* ctx, pk_scene* scene, const float* d_rays, int64_t n, int any_hit,
                       int32_t* d_ids, float* d_t, uint32_t* d_bary, int iters, float* ms_out);

/* ---- differentiation (no reference counterpart: Pike renders only; the
 * parameters are the ones shade_hit.wgsl:179-200 reads from its atlas and
 * shade_miss.wgsl:37-43 from its environment) ----
 * Parameters: PK_PARAM_ATLAS is the texture atlas in its own layout,
 * [layers][height][width] RGBA f32 (base color = RGB of a material's diffuse
 * layer, roughness = G of its metal-roughness layer, alpha = roughness^2);
 * PK_PARAM_ENV is the environment radiance, [height][width] RGBA f32 (A
 * unused).  Sampling is detached: sampling decisions and pdfs (opacity,
 * metalness, the GGX sampling roughness, the environment's importance tables)
 * keep the values given at upload, while BSDF values and environment radiance
 * read the current parameters.  A frame rendered from the current parameters
 * is therefore a smooth function of them along fixed paths, which central
 * finite differences over common seeds and the adjoint both differentiate. */
#define PK_PARAM_ATLAS 0
#define PK_PARAM_ENV 1
/* info6 = {tex width, tex height, tex layers, tex format, env width, env height} (env 0x0 if none). */
int pk_scene_param_info(const pk_scene* scene, uint32_t* info6);
/* Writes count floats of a parameter set starting at float `offset`
 * (host or device source).  Atlas writes need an RGBA32F atlas; the first one
 * with also_sampling == 0 gives the scene its own evaluation atlas (a copy of
 * the sampling atlas), after which the two diverge.  also_sampling != 0 also
 * writes the sampling atlas (sampling follows the parameters).  Environment
 * writes change radiance only (the importance tables stay frozen). */
int pk_scene_write_params(pk_scene* scene, int which, size_t offset, const float* src, size_t count,
                          int src_on_device, int also_sampling);
/* Reads the current (evaluation) parameters to host memory. */
int pk_scene_read_params(pk_scene* scene, int which, size_t offset, float* dst, size_t count);
/* L2 image loss of the frame's running-mean image against a device target
 * (width*height RGBA f32): writes d_grad = 2 (image - target) (A = 0, device)
 * and, if loss != NULL, blocks and returns sum over pixels and RGB of
 * (image - target)^2. */
int pk_loss_l2(pk_frame* frame, const float* d_target, float* d_grad, double* loss);
typedef struct pk_grad_buffers {
  float* atlas;    /* PK_PARAM_ATLAS layout, or NULL */
  float* env;      /* PK_PARAM_ENV layout, or NULL */
  int on_device;   /* 1: device pointers (asynchronous); 0: host pointers (the call blocks) */
  int accumulate;  /* 1: add to the buffers' contents; 0: overwrite them */
} pk_grad_buffers;
/* Path-replay adjoint: gradients of sum_p sum_c d_image[p][c] * I[p][c],
 * where I is the mean over the samples first_sample .. first_sample +
 * num_frames - 1 of the frame's paths (with first_sample 0, the image
 * pk_render produces on a reset frame with the same params), with respect to
 * both parameter sets.  Replaying other samples than the forward's gives an
 * independent gradient estimate.  d_image is width*height RGBA f32
 * (A ignored), host or device.  Each path is recomputed from its seeds
 * (nothing is stored by the forward render); device gradients are summed
 * with atomics, so their low bits vary run to run. */
int pk_adjoint(pk_frame* frame, pk_scene* scene, const pk_render_params* params, const float* d_image,
               int d_image_on_device, const pk_grad_buffers* grads);

#ifdef __cplusplus
}
#endif
#endif
