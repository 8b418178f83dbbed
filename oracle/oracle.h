/* oracle.h — TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of apbodnar/Pike's wavefront path-tracing path, used by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
 * checker.  The product (libpike_amd.so) never links, loads or calls it.
 *
 * What it restates (reference file:line):
 *   BVH build           render/bvh.js:7-215, render/primitives.js:3-44
 *   flatten/quantize    render/passes/trace_pass.js:166-170, 202-263, bvh.js:46-74
 *   closest hit         render/shader/trace.wgsl:83-213
 *   any hit             render/shader/trace_shadow.wgsl:1, 146-207
 *   camera              render/shader/camera.wgsl:46-92
 *   shading             render/shader/shade_hit.wgsl:153-483
 *   miss shading        render/shader/shade_miss.wgsl:37-57
 *   accumulate          render/shader/accumulate.wgsl:20-34
 *   env tables          render/env_sampler.js:31-157
 *   frame loop          render/pike.js:120-162
 * Deliberate, documented deviations from the reference (its defects) are the
 * same ones the product makes: deterministic shading seeds, race-free
 * accumulation, every miss shaded, correct last pixel.
 */
#ifndef ORACLE_H_
#define ORACLE_H_
#include <stdint.h>
#include "../include/pk.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_scene or_scene;

or_scene* or_scene_create(const pk_scene_desc* desc);
/* Instancing (pk_bvh_build_instanced semantics), before or_bvh_build: the
 * scene's own triangles become the first placement with the identity. */
int or_scene_set_instances(or_scene* s, const pk_mesh_desc* meshes, uint32_t num_meshes,
                           const pk_instance_desc* instances, uint32_t num_instances);
void or_scene_destroy(or_scene* s);
/* Literal bvh.js build (float64 SAH); returns 0 or -1 on a degenerate node. */
int or_bvh_build(or_scene* s, int num_threads);
/* [0] nodes, [1] leaf triangles, [2] depth, [3] largest leaf */
void or_bvh_info(const or_scene* s, int64_t* info4);
/* BFS flatten + normalize + snorm16 quantize exactly as trace_pass.js does. */
int or_bvh_serialize(const or_scene* s, int32_t* nodes5, int32_t* tris4, float* verts4, double* normalizer4);
/* Ray queries. rays: [n][8] {ox,oy,oz,_,dx,dy,dz,_} (normalized space, f32).
 * mode 0: literal reference traversal of the serialized 20-byte nodes
 *         (ids are original triangle ids mapped through leaf order),
 * mode 1: brute force over all triangles, canonical argmin (t, id),
 * mode 2: canonical closest hit through a conservative float64-box BVH.
 * any_hit selects trace_shadow semantics. flags[i] bit0 = literal stack overflow. */
int or_trace(const or_scene* s, const float* rays, int64_t n, int mode, int any_hit, int num_threads,
             int32_t* ids, float* t, uint32_t* bary, uint32_t* flags);
/* Environment tables (env_sampler.js). info: [0] bins, [1] coords, [2] width, [3] height */
int or_env_info(const or_scene* s, int64_t* info4);
int or_env_tables(const or_scene* s, int32_t* bins2, int32_t* coords2, float* pdf);

typedef struct or_render_params {
  uint32_t width, height, max_depth, first_sample, num_frames;
  float env_theta;
  int precision;   /* 32: FP32 restatement bit-for-bit with the product, 64: FP64 oracle */
  int num_threads;
  /* optional pixel window [x0, x1) x [y0, y1) (all zero = whole image) */
  uint32_t x0, x1, y0, y1;
} or_render_params;

/* Renders num_frames frames; writes the running mean (RGBA f32, A=1) over
 * the window and, if color_out != NULL, the last frame's color buffer (rgb sum
 * + contribution count).  Counters: [0] closest-hit rays, [1] shadow rays, [2] hits. */
int or_render(const or_scene* s, const pk_camera* cam, const or_render_params* p, float* accum_inout,
              float* color_out, uint64_t* counters3);

/* Parameters, as pk_scene_write_params: PK_PARAM_ATLAS writes the evaluation
 * atlas (RGBA32F scenes; with also_sampling the sampling atlas too),
 * PK_PARAM_ENV the environment radiance ([height][width] RGBA, A ignored).
 * Returns 0 or -1 on a bad range. */
int or_scene_write_params(or_scene* s, int which, int64_t offset, const float* src, int64_t count,
                          int also_sampling);
/* Reference gradient of sum_p sum_c d_image[p][c] * I[p][c] (I = the mean
 * image over the rendered frames, d_image [H][W] RGBA with A ignored) with
 * respect to n parameters given as float offsets index[i] into the
 * PK_PARAM_* layout which[i]: the same detached-sampling estimator evaluated
 * with forward-mode dual numbers (no hand-derived derivative), in FP64
 * (p->precision 64) or with the FP32 restatement's value arithmetic (32: the
 * product's paths bit for bit).  If value is not NULL it receives the
 * objective.  The pixel window is ignored. */
int or_gradient(const or_scene* s, const pk_camera* cam, const or_render_params* p, const float* d_image,
                const int32_t* which, const int64_t* index, int64_t n, double* out, double* value);

/* pk_math contract probes for tests */
void or_math_eval(int fn, const float* x, const float* y, float* out, int64_t n);

/* Single functions of the shading / camera / accumulate restatement
 * (shade_hit.wgsl:166-361, shade_miss.wgsl:37-43, camera.wgsl:58-73,
 * accumulate.wgsl:26-30) evaluated on n inputs, for the closed-form tests.
 * precision 64: the FP64 oracle, 32: the FP32 restatement.  Inputs per item
 * (doubles, in_stride apart) -> outputs (out_stride apart):
 *   GGX_D            m.xyz, alpha                                  -> D
 *   GGX_G1           w.xyz, m.xyz, alpha                           -> G1(w, m)
 *   LAMBERT_PDF      wi.xyz                                        -> pdf
 *   SPECULAR_PDF     wo.xyz, m.xyz, alpha                          -> pdf of the reflected direction
 *   EVAL_LAMBERT     base.rgb, wi.xyz, pdf                         -> rgb
 *   EVAL_SPECULAR    wo.xyz, wi.xyz, pdf, alpha, spec.rgb          -> rgb
 *   POWER_HEURISTIC  a, b                                          -> weight
 *   SCHLICK          cos, ior                                      -> F
 *   ONB              n.xyz                                         -> b1.xyz, b2.xyz, n.xyz (columns)
 *   SAMPLE_GGX       wo.xyz, alpha, rng seed                       -> m.xyz
 *   SAMPLE_LAMBERT   rng seed                                      -> wi.xyz, pdf
 *   ENV_COLOR        dir.xyz, env_theta     (scene)                -> rgb
 *   ENV_PDF          dir.xyz, env_theta     (scene)                -> pdf
 *   SAMPLE_ENV       rng seed, env_theta    (scene; identity frame)-> wi.xyz, pdf
 *   TEXTURE          layer, u, v            (scene; sampling atlas)-> rgba
 *   PRIMARY_RAY      origin.xyz, dir.xyz, fov, focal depth, aperture, distortion, bokeh,
 *                    px, py, W, H, sample                          -> origin.xyz, dir.xyz
 *   SURFACE          o.xyz, d.xyz           (scene with BVH)       -> OR_PROBE_SURFACE_OUT values:
 *                    [0] t (-1: miss), [1] id, [2..4] shading normal, [5..7] origin,
 *                    [8..10] wo (local), [11..19] basis columns, [20] metallic,
 *                    [21] opacity, [22] alpha, [23..25] emission, [26..28] base,
 *                    [29..31] spec, [32] evaluation alpha, [33..34] barycentric u, v
 *   ACCUMULATE       color, accumulator, sample index              -> new accumulator
 * Returns 0, or -1 on an unknown function, a missing scene or bad input. */
enum {
  OR_PROBE_GGX_D = 0,
  OR_PROBE_GGX_G1,
  OR_PROBE_LAMBERT_PDF,
  OR_PROBE_SPECULAR_PDF,
  OR_PROBE_EVAL_LAMBERT,
  OR_PROBE_EVAL_SPECULAR,
  OR_PROBE_POWER_HEURISTIC,
  OR_PROBE_SCHLICK,
  OR_PROBE_ONB,
  OR_PROBE_SAMPLE_GGX,
  OR_PROBE_SAMPLE_LAMBERT,
  OR_PROBE_ENV_COLOR,
  OR_PROBE_ENV_PDF,
  OR_PROBE_SAMPLE_ENV,
  OR_PROBE_TEXTURE,
  OR_PROBE_PRIMARY_RAY,
  OR_PROBE_SURFACE,
  OR_PROBE_ACCUMULATE,
  OR_PROBE_COUNT
};
#define OR_PROBE_SURFACE_OUT 35
int or_probe(const or_scene* s, int fn, int precision, const double* in, int64_t in_stride, int64_t n, double* out,
             int64_t out_stride);
/* Test-only planted mistakes (one wrong line of the restatement each, see
 * OR_MUT_* in oracle.cpp): the pin tests show that each closed-form check
 * fails while its mutation is planted.  0 restores the faithful code. */
void or_set_mutation(int k);
int or_mutation_count(void);

#ifdef __cplusplus
}
#endif
#endif
