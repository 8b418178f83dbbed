"""ctypes declarations of include/pk.h (libpike_amd.so).

The library is built in-tree by ``pike_amd/csrc/Makefile`` (see
``__graft_entry__.build``).  Loading it does not touch the GPU; calls that
need a device fail with ``PK_E_NO_DEVICE`` when none is visible.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpike_amd.so")

PK_OK = 0
PK_E_INVALID = -1
PK_E_NO_DEVICE = -2
PK_E_OOM = -3
PK_E_HIP = -4
PK_E_LIMIT = -5
PK_E_STATE = -6
PK_E_OVERFLOW = -7

PK_TEX_RGBA8 = 0
PK_TEX_RGBA32F = 1
PK_ENV_NONE = 0
PK_ENV_RGBE8 = 1
PK_ENV_RGB32F = 2
PK_ENV_SAMPLER_BINS = 0
PK_ENV_SAMPLER_CDF = 1
PK_PARAM_ATLAS = 0
PK_BVH_SPATIAL_SPLITS = 1
PK_BVH_SAH_COLLAPSE = 2
PK_PARAM_ENV = 1

c_void_p = ctypes.c_void_p
c_int = ctypes.c_int
c_uint32 = ctypes.c_uint32
c_int64 = ctypes.c_int64
c_float = ctypes.c_float
c_size_t = ctypes.c_size_t


class MeshDesc(ctypes.Structure):
    _fields_ = [
        ("num_vertices", c_uint32),
        ("positions", c_void_p),
        ("attributes", c_void_p),
        ("num_triangles", c_uint32),
        ("triangles", c_void_p),
    ]


class InstanceDesc(ctypes.Structure):
    _fields_ = [
        ("mesh", c_uint32),
        ("material", ctypes.c_int32),
        ("transform", ctypes.c_double * 12),
    ]


class SceneDesc(ctypes.Structure):
    _fields_ = [
        ("num_vertices", c_uint32),
        ("positions", c_void_p),
        ("attributes", c_void_p),
        ("num_triangles", c_uint32),
        ("triangles", c_void_p),
        ("num_materials", c_uint32),
        ("materials", c_void_p),
        ("tex_width", c_uint32),
        ("tex_height", c_uint32),
        ("tex_layers", c_uint32),
        ("tex_format", c_uint32),
        ("tex_data", c_void_p),
        ("env_width", c_uint32),
        ("env_height", c_uint32),
        ("env_format", c_uint32),
        ("env_sampler", c_uint32),
        ("env_data", c_void_p),
    ]


class LaunchShape(ctypes.Structure):
    _fields_ = [("waves_per_cu", c_uint32), ("refill_min", c_uint32), ("stack_lds", c_uint32),
                ("waves_per_block", c_uint32), ("leaf_phase16", c_uint32), ("node_width", c_uint32)]


class Camera(ctypes.Structure):
    _fields_ = [
        ("origin", c_float * 3),
        ("dir", c_float * 3),
        ("fov", c_float),
        ("focal_depth", c_float),
        ("aperture", c_float),
        ("distortion", c_float),
        ("bokeh", c_float),
    ]


class View(ctypes.Structure):
    _fields_ = [
        ("scene", c_uint32),
        ("width", c_uint32),
        ("height", c_uint32),
        ("spp", c_uint32),
        ("camera", Camera),
    ]


class RenderParams(ctypes.Structure):
    _fields_ = [
        ("width", c_uint32),
        ("height", c_uint32),
        ("max_depth", c_uint32),
        ("first_sample", c_uint32),
        ("num_frames", c_uint32),
        ("env_theta", c_float),
        ("use_graph", c_uint32),
    ]


class Counters(ctypes.Structure):
    _fields_ = [
        ("rays_traced", ctypes.c_uint64),
        ("shadow_rays", ctypes.c_uint64),
        ("hits", ctypes.c_uint64),
        ("nodes_visited", ctypes.c_uint64),
        ("tris_tested", ctypes.c_uint64),
        ("shadow_nodes", ctypes.c_uint64),
        ("shadow_tris", ctypes.c_uint64),
        ("stack_overflow", c_uint32),
        ("frames", c_uint32),
    ]


# name -> (restype, argtypes); every entry point declared in include/pk.h
class GradBuffers(ctypes.Structure):
    _fields_ = [
        ("atlas", c_void_p),
        ("env", c_void_p),
        ("on_device", c_int),
        ("accumulate", c_int),
    ]


SIGNATURES = {
    "pk_last_error": (ctypes.c_char_p, []),
    "pk_version": (ctypes.c_char_p, []),
    "pk_ctx_create": (c_int, [c_int, ctypes.POINTER(c_void_p)]),
    "pk_ctx_destroy": (c_int, [c_void_p]),
    "pk_sync": (c_int, [c_void_p]),
    "pk_ctx_set_tuning": (c_int, [c_void_p, c_int, c_uint32, c_uint32]),
    "pk_ctx_set_scheduler": (c_int, [c_void_p, c_int]),
    "pk_ctx_get_scheduler": (c_int, [c_void_p, ctypes.POINTER(c_int)]),
    "pk_ctx_get_tuning": (c_int, [c_void_p, c_int, ctypes.POINTER(c_uint32), ctypes.POINTER(c_uint32),
                                  ctypes.POINTER(c_uint32)]),
    "pk_trace_rays_noalloc": (c_int, [c_void_p, c_void_p, c_int64, c_int, c_void_p, c_void_p, c_void_p]),
    "pk_ctx_set_launch": (c_int, [c_void_p, c_int, ctypes.POINTER(LaunchShape)]),
    "pk_ctx_get_launch": (c_int, [c_void_p, c_int, ctypes.POINTER(LaunchShape)]),
    "pk_ctx_set_stack_limit": (c_int, [c_void_p, c_uint32]),
    "pk_ctx_get_stack_limit": (c_int, [c_void_p, ctypes.POINTER(c_uint32)]),
    "pk_ctx_set_memory_limit": (c_int, [c_void_p, ctypes.c_uint64]),
    "pk_ctx_memory_used": (c_int, [c_void_p, ctypes.POINTER(ctypes.c_uint64)]),
    "pk_autotune_render": (c_int, [c_void_p, c_void_p, c_uint32, c_uint32, c_void_p]),
    "pk_autotune_rays": (c_int, [c_void_p, c_void_p, c_void_p, c_int64, c_int, c_void_p, c_void_p, c_void_p,
                                 c_void_p]),
    "pk_buffer_create": (c_int, [c_void_p, c_size_t, ctypes.POINTER(c_void_p)]),
    "pk_buffer_destroy": (c_int, [c_void_p, c_void_p]),
    "pk_buffer_write": (c_int, [c_void_p, c_void_p, c_void_p, c_size_t]),
    "pk_buffer_read": (c_int, [c_void_p, c_void_p, c_void_p, c_size_t]),
    "pk_buffer_copy": (c_int, [c_void_p, c_void_p, c_void_p, c_size_t]),
    "pk_bvh_build":(c_int, [c_void_p, c_uint32, c_void_p, c_uint32, c_int, ctypes.POINTER(c_void_p)]),
    "pk_bvh_build_ex": (c_int, [c_void_p, c_uint32, c_void_p, c_uint32, c_int, c_uint32, ctypes.POINTER(c_void_p)]),
    "pk_bvh_destroy": (c_int, [c_void_p]),
    "pk_bvh_info": (c_int, [c_void_p, c_void_p]),
    "pk_bvh_export_reference": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "pk_bvh_export_wide": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "pk_bvh_export_wide4": (c_int, [c_void_p, c_void_p, ctypes.POINTER(ctypes.c_int64)]),
    "pk_bvh_normalizer": (c_int, [c_void_p, c_void_p]),
    "pk_bvh_build_instanced": (c_int, [c_void_p, c_void_p, c_uint32, c_void_p, c_uint32, c_int, c_uint32,
                                       ctypes.POINTER(c_void_p)]),
    "pk_scene_upload": (c_int, [c_void_p, ctypes.POINTER(SceneDesc), c_void_p, ctypes.POINTER(c_void_p)]),
    "pk_scene_destroy": (c_int, [c_void_p]),
    "pk_scene_env_info": (c_int, [c_void_p, c_void_p]),
    "pk_env_tables": (c_int, [ctypes.POINTER(SceneDesc), c_void_p, c_void_p, c_void_p, c_void_p]),
    "pk_frame_create": (c_int, [c_void_p, c_uint32, c_uint32, ctypes.POINTER(c_void_p)]),
    "pk_frame_destroy": (c_int, [c_void_p]),
    "pk_frame_reset": (c_int, [c_void_p]),
    "pk_frame_set_camera": (c_int, [c_void_p, ctypes.POINTER(Camera)]),
    "pk_frame_set_env_theta": (c_int, [c_void_p, c_float]),
    "pk_frame_begin": (c_int, [c_void_p, c_uint32]),
    "pk_raygen": (c_int, [c_void_p]),
    "pk_trace_closest": (c_int, [c_void_p, c_void_p, c_uint32]),
    "pk_shade": (c_int, [c_void_p, c_void_p, c_uint32]),
    "pk_trace_shadow": (c_int, [c_void_p, c_void_p]),
    "pk_accumulate": (c_int, [c_void_p]),
    "pk_render": (c_int, [c_void_p, c_void_p, ctypes.POINTER(RenderParams)]),
    "pk_read_image": (c_int, [c_void_p, c_void_p, c_size_t]),
    "pk_read_color": (c_int, [c_void_p, c_void_p, c_size_t]),
    "pk_read_counters": (c_int, [c_void_p, ctypes.POINTER(Counters)]),
    "pk_frame_count_nodes": (c_int, [c_void_p, c_int]),
    "pk_render_profile": (c_int, [c_void_p, c_void_p, ctypes.POINTER(RenderParams), c_void_p]),
    "pk_scene_upload_batch": (c_int, [c_void_p, ctypes.POINTER(SceneDesc), ctypes.POINTER(c_void_p), c_uint32,
                                      ctypes.POINTER(c_void_p)]),
    "pk_scene_roots": (c_int, [c_void_p, c_void_p, c_uint32]),
    "pk_batch_create": (c_int, [c_void_p, c_void_p, ctypes.POINTER(View), c_uint32, ctypes.POINTER(c_void_p)]),
    "pk_batch_render": (c_int, [c_void_p, c_uint32, c_float, c_int]),
    "pk_batch_read_view": (c_int, [c_void_p, c_uint32, c_void_p, c_size_t]),
    "pk_trace_rays": (c_int, [c_void_p, c_void_p, c_void_p, c_int64, c_int, c_void_p, c_void_p, c_void_p]),
    "pk_trace_rays_device": (c_int, [c_void_p, c_void_p, c_void_p, c_int64, c_int, c_void_p, c_void_p, c_void_p,
                                     c_int]),
    "pk_ctx_trace_stats": (c_int, [c_void_p, c_void_p]),
    "pk_time_trace_rays": (c_int, [c_void_p, c_void_p, c_void_p, c_int64, c_int, c_void_p, c_void_p, c_void_p,
                                   c_int, ctypes.POINTER(c_float)]),
    "pk_read_image_device": (c_int, [c_void_p, c_void_p, ctypes.c_size_t]),
    "pk_frame_set_samples_per_pass": (c_int, [c_void_p, ctypes.c_uint32]),
    "pk_frame_get_samples_per_pass": (c_int, [c_void_p, ctypes.POINTER(ctypes.c_uint32)]),
    "pk_autotune_samples_per_pass": (c_int, [c_void_p, c_void_p, ctypes.c_uint32, ctypes.c_uint32,
                                             ctypes.POINTER(c_float)]),
    "pk_autotune_scheduler": (c_int, [c_void_p, c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(c_float)]),
    "pk_scene_param_info": (c_int, [c_void_p, c_void_p]),
    "pk_scene_write_params": (c_int, [c_void_p, c_int, ctypes.c_size_t, c_void_p, ctypes.c_size_t, c_int, c_int]),
    "pk_scene_read_params": (c_int, [c_void_p, c_int, ctypes.c_size_t, c_void_p, ctypes.c_size_t]),
    "pk_loss_l2": (c_int, [c_void_p, c_void_p, c_void_p, ctypes.POINTER(ctypes.c_double)]),
    "pk_adjoint": (c_int, [c_void_p, c_void_p, ctypes.POINTER(RenderParams), c_void_p, c_int,
                           ctypes.POINTER(GradBuffers)]),
}

_lib = None


def load():
    """Load libpike_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `make -C pike_amd/csrc` or __graft_entry__.build()")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib
