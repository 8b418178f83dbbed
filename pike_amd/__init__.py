"""pike_amd — MI355X-native wavefront path tracer (Python harness).

Thin, typed wrappers over the C ABI in include/pk.h.  Every call goes to
libpike_amd.so (HIP kernels for gfx950); there is no CPU fallback: a missing
library raises ImportError on use, a missing GPU raises PikeError with
PK_E_NO_DEVICE.

The object model mirrors the reference's render passes
(render/passes/*.js): a Context (the WebGPU device, pike.js:92-97), a BVH
(bvh.js), a Scene (TracePass/ShadeHitPass/ShadeMissPass resources), and a
Frame (RenderState + CameraPass state) whose methods are the passes'
generateCommands.
"""
import ctypes
import weakref

import numpy as np

from . import _ffi
from ._ffi import Camera, Counters, RenderParams  # noqa: F401

__all__ = ["PikeError", "Context", "BVH", "Scene", "BatchScene", "Batch", "Frame", "DeviceBuffer", "lib"]


class PikeError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{msg} (code {code})")
        self.code = code


def lib():
    return _ffi.load()


def _check(rc, what):
    if rc != _ffi.PK_OK:
        err = lib().pk_last_error()
        raise PikeError(rc, f"{what}: {err.decode() if err else 'error'}")


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def _dev(x):
    """Device address of a DeviceBuffer, an integer pointer, or None."""
    if x is None:
        return None
    return ctypes.c_void_p(x.ptr if isinstance(x, DeviceBuffer) else int(x))


class Context:
    """A HIP device context (one stream)."""

    def __init__(self, device=0):
        h = ctypes.c_void_p()
        _check(lib().pk_ctx_create(device, ctypes.byref(h)), "pk_ctx_create")
        self.h = h
        self._children = weakref.WeakSet()  # scenes, frames, buffers: destroyed before the context

    def _adopt(self, child):
        self._children.add(child)

    def sync(self):
        _check(lib().pk_sync(self.h), "pk_sync")

    TUNE_CLASSES = {"primary": 0, "secondary": 1, "shadow": 2, "rays_closest": 3, "rays_any": 4, "paths": 5}
    SCHEDULERS = {"wavefront": 0, "paths": 1}

    def set_scheduler(self, name):
        """How single-view frames run a pass: "wavefront" (per-bounce kernels
        over device queues) or "paths" (one path megakernel); the images are
        bit-identical."""
        _check(lib().pk_ctx_set_scheduler(self.h, self.SCHEDULERS[name]), "pk_ctx_set_scheduler")

    def scheduler(self):
        v = ctypes.c_int()
        _check(lib().pk_ctx_get_scheduler(self.h, ctypes.byref(v)), "pk_ctx_get_scheduler")
        return {c: n for n, c in self.SCHEDULERS.items()}[v.value]

    def set_tuning(self, kernel_class, waves_per_cu, refill_min, stack_lds=None, waves_per_block=None,
                   leaf_phase16=None, node_width=None):
        """Launch shape of a kernel class; the fields left None keep their value."""
        cls = self.TUNE_CLASSES[kernel_class] if isinstance(kernel_class, str) else kernel_class
        sh = _ffi.LaunchShape()
        _check(lib().pk_ctx_get_launch(self.h, cls, ctypes.byref(sh)), "pk_ctx_get_launch")
        sh.waves_per_cu, sh.refill_min = waves_per_cu, refill_min
        if stack_lds is not None:
            sh.stack_lds = stack_lds
        if waves_per_block is not None:
            sh.waves_per_block = waves_per_block
        if leaf_phase16 is not None:
            sh.leaf_phase16 = leaf_phase16
        if node_width is not None:
            sh.node_width = node_width
        _check(lib().pk_ctx_set_launch(self.h, cls, ctypes.byref(sh)), "pk_ctx_set_launch")

    def tuning(self):
        out = {}
        for name, cls in self.TUNE_CLASSES.items():
            sh = _ffi.LaunchShape()
            _check(lib().pk_ctx_get_launch(self.h, cls, ctypes.byref(sh)), "pk_ctx_get_launch")
            out[name] = dict(waves_per_cu=sh.waves_per_cu, refill_min=sh.refill_min, stack_lds=sh.stack_lds,
                             waves_per_block=sh.waves_per_block, leaf_phase16=sh.leaf_phase16,
                             node_width=sh.node_width)
        return out

    def set_stack_limit(self, entries):
        """Traversal stack entries per ray (1..64, default 64); rays needing
        more are flagged as overflowed and skip what they cannot push."""
        _check(lib().pk_ctx_set_stack_limit(self.h, entries), "pk_ctx_set_stack_limit")

    def stack_limit(self):
        v = ctypes.c_uint32()
        _check(lib().pk_ctx_get_stack_limit(self.h, ctypes.byref(v)), "pk_ctx_get_stack_limit")
        return v.value

    def set_memory_limit(self, nbytes):
        """Cap on the device memory of this context's scenes and frames (0: none)."""
        _check(lib().pk_ctx_set_memory_limit(self.h, int(nbytes)), "pk_ctx_set_memory_limit")

    def memory_used(self):
        v = ctypes.c_uint64()
        _check(lib().pk_ctx_memory_used(self.h, ctypes.byref(v)), "pk_ctx_memory_used")
        return v.value

    def autotune_rays(self, scene, rays_ptr, n, any_hit, ids_ptr, t_ptr, bary_ptr):
        ms = (ctypes.c_float * 2)()
        _check(lib().pk_autotune_rays(self.h, scene.h, rays_ptr, n, int(any_hit), ids_ptr, t_ptr, bary_ptr, ms),
               "pk_autotune_rays")
        return float(ms[0]), float(ms[1])

    def close(self):
        if self.h:
            for child in list(self._children):
                try:
                    child.close()
                except AttributeError:  # construction failed before the handle existed
                    pass
            lib().pk_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def trace_rays(self, scene, rays, any_hit=False):
        """Closest (or any) hit of rays (n, 8) float32 {o, _, d, _} in the
        normalized space.  Returns (ids int32, t float32, bary uint32)."""
        rays = np.ascontiguousarray(rays, dtype=np.float32)
        n = rays.shape[0]
        ids = np.empty(n, np.int32)
        t = np.empty(n, np.float32)
        bary = np.empty(n, np.uint32)
        _check(lib().pk_trace_rays(self.h, scene.h, _ptr(rays), n, int(any_hit), _ptr(ids), _ptr(t), _ptr(bary)),
               "pk_trace_rays")
        return ids, t, bary

    def trace_rays_device(self, scene, rays_ptr, n, any_hit, ids_ptr, t_ptr, bary_ptr, count_nodes=False):
        _check(lib().pk_trace_rays_device(self.h, scene.h, rays_ptr, n, int(any_hit), ids_ptr, t_ptr, bary_ptr,
                                          int(count_nodes)), "pk_trace_rays_device")

    def trace_stats(self):
        s = (ctypes.c_uint64 * 2)()
        _check(lib().pk_ctx_trace_stats(self.h, s), "pk_ctx_trace_stats")
        return int(s[0]), int(s[1])

    def time_trace_rays(self, scene, rays_ptr, n, any_hit, ids_ptr, t_ptr, bary_ptr, iters):
        ms = ctypes.c_float()
        _check(lib().pk_time_trace_rays(self.h, scene.h, rays_ptr, n, int(any_hit), ids_ptr, t_ptr, bary_ptr, iters,
                                        ctypes.byref(ms)), "pk_time_trace_rays")
        return float(ms.value)


class DeviceBuffer:
    """A device allocation owned by a Context (createBuffer/writeBuffer/mapAsync)."""

    def __init__(self, ctx, nbytes):
        self.ctx = ctx
        ctx._adopt(self)
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        _check(lib().pk_buffer_create(ctx.h, self.nbytes, ctypes.byref(p)), "pk_buffer_create")
        self.ptr = p.value

    @classmethod
    def from_array(cls, ctx, arr):
        arr = np.ascontiguousarray(arr)
        b = cls(ctx, arr.nbytes)
        b.write(arr)
        return b

    def write(self, arr):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        _check(lib().pk_buffer_write(self.ctx.h, self.ptr, _ptr(arr), arr.nbytes), "pk_buffer_write")

    def read(self, dtype, count):
        out = np.empty(count, dtype)
        assert out.nbytes <= self.nbytes
        _check(lib().pk_buffer_read(self.ctx.h, _ptr(out), self.ptr, out.nbytes), "pk_buffer_read")
        return out

    def close(self):
        if self.ptr and self.ctx.h:
            lib().pk_buffer_destroy(self.ctx.h, self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BVH:
    """bvh.js SAH binary BVH + its 128-byte wide collapse (host side)."""

    def __init__(self, positions, triangles, threads=0, spatial_splits=False, sah_collapse=False):
        pos = np.ascontiguousarray(positions, dtype=np.float64)
        tri = np.ascontiguousarray(triangles, dtype=np.int32)
        h = ctypes.c_void_p()
        flags = (_ffi.PK_BVH_SPATIAL_SPLITS if spatial_splits else 0) | (_ffi.PK_BVH_SAH_COLLAPSE if sah_collapse else 0)
        _check(lib().pk_bvh_build_ex(_ptr(pos), pos.shape[0], _ptr(tri), tri.shape[0], threads, flags,
                                     ctypes.byref(h)), "pk_bvh_build_ex")
        self.h = h
        self.num_vertices = pos.shape[0]
        self.num_triangles = tri.shape[0]

    def info(self):
        a = np.zeros(6, np.int64)
        _check(lib().pk_bvh_info(self.h, _ptr(a)), "pk_bvh_info")
        keys = ["binary_nodes", "wide_nodes", "leaf_triangles", "binary_depth", "wide_depth", "largest_leaf"]
        return dict(zip(keys, (int(x) for x in a)))

    def export_reference(self):
        """The reference's flattened buffers (trace_pass.js:220-263)."""
        info = self.info()
        nodes = np.zeros((info["binary_nodes"], 5), np.int32)
        tris = np.zeros((max(info["leaf_triangles"], self.num_triangles), 4), np.int32)
        verts = np.zeros((self.num_vertices, 4), np.float32)
        norm = np.zeros(4, np.float64)
        _check(lib().pk_bvh_export_reference(self.h, _ptr(nodes), _ptr(tris), _ptr(verts), _ptr(norm)),
               "pk_bvh_export_reference")
        return nodes, tris, verts, norm

    WIDE_NODE_DTYPE = np.dtype([("origin", "<f4", 3), ("exp", "u1", 3), ("nchild", "u1"), ("child", "<u4", 8),
                                ("meta", "u1", 8), ("qlox", "u1", 8), ("qhix", "u1", 8), ("qloy", "u1", 8),
                                ("qhiy", "u1", 8), ("qloz", "u1", 8), ("qhiz", "u1", 8), ("leafmask", "u1"),
                                ("instmask", "u1"), ("pad8", "u1", 2), ("ibase", "<u4"), ("tbase", "<u4"),
                                ("pad", "<u4", 3)])
    WIDE4_NODE_DTYPE = np.dtype([("origin", "<f4", 3), ("exp", "u1", 3), ("nchild", "u1"), ("meta", "u1", 4),
                                 ("qlox", "u1", 4), ("qhix", "u1", 4), ("qloy", "u1", 4), ("qhiy", "u1", 4),
                                 ("qloz", "u1", 4), ("qhiz", "u1", 4), ("leafmask", "u1"), ("pad8", "u1", 3),
                                 ("ibase", "<u4"), ("tbase", "<u4"), ("pad", "<u4", 2)])
    TRI_DTYPE = np.dtype([("v0", "<f4", 3), ("id", "<u4"), ("e1", "<f4", 3), ("pad0", "<u4"), ("e2", "<f4", 3),
                          ("pad1", "<u4")])

    def export_wide(self):
        """(nodes, triangle records, triangle {i0,i1,i2,mat}) of the 128-byte layout."""
        info = self.info()
        nodes = np.zeros(info["wide_nodes"], self.WIDE_NODE_DTYPE)
        tris = np.zeros(info["leaf_triangles"], self.TRI_DTYPE)
        tidx = np.zeros((info["leaf_triangles"], 4), np.int32)
        assert nodes.itemsize == 128 and tris.itemsize == 48
        _check(lib().pk_bvh_export_wide(self.h, _ptr(nodes), _ptr(tris), _ptr(tidx)), "pk_bvh_export_wide")
        return nodes, tris, tidx

    def export_wide4(self):
        """Nodes of the 64-byte 4-wide layout derived from the 128-byte one
        (same triangle records)."""
        n = ctypes.c_int64(0)
        _check(lib().pk_bvh_export_wide4(self.h, None, ctypes.byref(n)), "pk_bvh_export_wide4")
        nodes = np.zeros(n.value, self.WIDE4_NODE_DTYPE)
        assert nodes.itemsize == 64
        _check(lib().pk_bvh_export_wide4(self.h, _ptr(nodes), ctypes.byref(n)), "pk_bvh_export_wide4")
        return nodes

    def normalizer(self):
        norm = np.zeros(4, np.float64)
        _check(lib().pk_bvh_normalizer(self.h, _ptr(norm)), "pk_bvh_normalizer")
        return norm

    @classmethod
    def for_scene(cls, data, threads=0, spatial_splits=False, sah_collapse=False):
        """The BVH a SceneData needs: single-level, or two-level when it has
        instances (pk_bvh_build_instanced)."""
        if not data.instances:
            return cls(data.positions, data.triangles, threads, spatial_splits, sah_collapse)
        desc = data.desc()
        meshes, insts, keep = data.instance_descs()  # noqa: F841 (arrays stay alive through the call)
        h = ctypes.c_void_p()
        flags = (_ffi.PK_BVH_SPATIAL_SPLITS if spatial_splits else 0) | (_ffi.PK_BVH_SAH_COLLAPSE if sah_collapse else 0)
        _check(lib().pk_bvh_build_instanced(ctypes.byref(desc), meshes, len(data.meshes), insts, len(data.instances),
                                            threads, flags, ctypes.byref(h)), "pk_bvh_build_instanced")
        b = cls.__new__(cls)
        b.h = h
        b.num_vertices = int(np.asarray(data.positions).shape[0])
        b.num_triangles = data.num_triangles
        b.instanced = True
        return b

    def close(self):
        if self.h:
            lib().pk_bvh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Scene:
    """Device-resident scene (geometry, wide BVH, materials, atlas, env tables)."""

    def __init__(self, ctx, data, bvh):
        self.ctx = ctx
        ctx._adopt(self)
        self.desc = data.desc()
        self._data = data
        h = ctypes.c_void_p()
        _check(lib().pk_scene_upload(ctx.h, ctypes.byref(self.desc), bvh.h, ctypes.byref(h)), "pk_scene_upload")
        self.h = h

    def env_info(self):
        a = np.zeros(4, np.int64)
        _check(lib().pk_scene_env_info(self.h, _ptr(a)), "pk_scene_env_info")
        return dict(bins=int(a[0]), coords=int(a[1]), width=int(a[2]), height=int(a[3]))

    # differentiable parameters (pk_scene_write_params): the atlas as
    # [layers][height][width][4] f32 and the environment as [height][width][4]
    def param_info(self):
        a = np.zeros(6, np.uint32)
        _check(lib().pk_scene_param_info(self.h, _ptr(a)), "pk_scene_param_info")
        return dict(tex_width=int(a[0]), tex_height=int(a[1]), tex_layers=int(a[2]), tex_format=int(a[3]),
                    env_width=int(a[4]), env_height=int(a[5]))

    def param_count(self, which):
        i = self.param_info()
        if which == _ffi.PK_PARAM_ATLAS:
            return i["tex_width"] * i["tex_height"] * i["tex_layers"] * 4
        return i["env_width"] * i["env_height"] * 4

    def write_params(self, which, values=None, offset=0, also_sampling=False, device_ptr=None, count=None):
        """Writes floats of a parameter set from host `values` or from a device
        pointer (`device_ptr`, `count` floats; the caller has synchronized the
        producer of that memory)."""
        if device_ptr is not None:
            src, n, dev = ctypes.c_void_p(int(device_ptr)), int(count), 1
        else:
            v = np.ascontiguousarray(values, np.float32).ravel()
            src, n, dev, keep = _ptr(v), v.size, 0, v  # noqa: F841 (keep v alive through the call)
        _check(lib().pk_scene_write_params(self.h, which, offset, src, n, dev, int(also_sampling)),
               "pk_scene_write_params")

    def read_params(self, which, count=None, offset=0):
        n = self.param_count(which) - offset if count is None else count
        out = np.empty(n, np.float32)
        _check(lib().pk_scene_read_params(self.h, which, offset, _ptr(out), n), "pk_scene_read_params")
        return out

    def close(self):
        if self.h:
            lib().pk_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BatchScene(Scene):
    """Several scenes merged into one device scene (pk_scene_upload_batch)."""

    def __init__(self, ctx, datas, bvhs):  # noqa: super().__init__ is not used: one upload for all
        self.ctx = ctx
        ctx._adopt(self)
        self._datas = list(datas)
        descs = (_ffi.SceneDesc * len(datas))()
        for i, d in enumerate(datas):
            descs[i] = d.desc()
        self.descs = descs
        handles = (ctypes.c_void_p * len(bvhs))(*[b.h.value for b in bvhs])
        h = ctypes.c_void_p()
        _check(lib().pk_scene_upload_batch(ctx.h, descs, handles, len(datas), ctypes.byref(h)),
               "pk_scene_upload_batch")
        self.h = h
        self.num_scenes = len(datas)


class Batch:
    """Many views over a BatchScene, traced through one global wavefront."""

    def __init__(self, ctx, scene, views):
        """views: list of (scene index, width, height, spp, Camera)."""
        self.ctx = ctx
        ctx._adopt(self)
        self.scene = scene
        arr = (_ffi.View * len(views))()
        for i, (si, w, hh, spp, cam) in enumerate(views):
            arr[i] = _ffi.View(si, w, hh, spp, cam)
        self.views = [(si, w, hh, spp) for si, w, hh, spp, _ in views]
        h = ctypes.c_void_p()
        _check(lib().pk_batch_create(ctx.h, scene.h, arr, len(views), ctypes.byref(h)), "pk_batch_create")
        self.h = h

    def render(self, max_depth=4, env_theta=0.0, use_graph=True):
        _check(lib().pk_batch_render(self.h, max_depth, float(env_theta), int(use_graph)), "pk_batch_render")

    def set_samples_per_round(self, samples):
        """Samples per view pixel traced in one round of render() (default
        1); the images are identical for any value."""
        _check(lib().pk_frame_set_samples_per_pass(self.h, int(samples)), "pk_frame_set_samples_per_pass")

    @property
    def samples_per_round(self):
        v = ctypes.c_uint32()
        _check(lib().pk_frame_get_samples_per_pass(self.h, ctypes.byref(v)), "pk_frame_get_samples_per_pass")
        return v.value

    def read_view(self, i):
        _, w, hh, _ = self.views[i]
        out = np.empty((hh, w, 4), np.float32)
        _check(lib().pk_batch_read_view(self.h, i, _ptr(out), out.nbytes), "pk_batch_read_view")
        return out

    def reset(self):
        _check(lib().pk_frame_reset(self.h), "pk_frame_reset")

    def counters(self):
        c = Counters()
        _check(lib().pk_read_counters(self.h, ctypes.byref(c)), "pk_read_counters")
        return {k: int(getattr(c, k)) for k, _ in Counters._fields_}

    def close(self):
        if self.h:
            lib().pk_frame_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Frame:
    """RenderState + camera + ray/hit queues for one resolution."""

    def __init__(self, ctx, width, height):
        self.ctx = ctx
        ctx._adopt(self)
        self.width, self.height = width, height
        h = ctypes.c_void_p()
        _check(lib().pk_frame_create(ctx.h, width, height, ctypes.byref(h)), "pk_frame_create")
        self.h = h

    def set_camera(self, cam):
        self._cam = cam
        _check(lib().pk_frame_set_camera(self.h, ctypes.byref(cam)), "pk_frame_set_camera")

    def set_env_theta(self, theta):
        _check(lib().pk_frame_set_env_theta(self.h, float(theta)), "pk_frame_set_env_theta")

    def set_samples_per_pass(self, samples):
        """Trace `samples` samples of every pixel per pass (one wavefront)."""
        _check(lib().pk_frame_set_samples_per_pass(self.h, int(samples)), "pk_frame_set_samples_per_pass")

    @property
    def samples_per_pass(self):
        v = ctypes.c_uint32()
        _check(lib().pk_frame_get_samples_per_pass(self.h, ctypes.byref(v)), "pk_frame_get_samples_per_pass")
        return v.value

    def reset(self):
        _check(lib().pk_frame_reset(self.h), "pk_frame_reset")

    def count_nodes(self, enable=True):
        _check(lib().pk_frame_count_nodes(self.h, int(enable)), "pk_frame_count_nodes")

    # per-pass entry points (CameraPass / TracePass / ShadeHitPass / ...)
    def begin(self, sample):
        _check(lib().pk_frame_begin(self.h, sample), "pk_frame_begin")

    def raygen(self):
        _check(lib().pk_raygen(self.h), "pk_raygen")

    def trace_closest(self, scene, bounce):
        _check(lib().pk_trace_closest(self.h, scene.h, bounce), "pk_trace_closest")

    def shade(self, scene, bounce):
        _check(lib().pk_shade(self.h, scene.h, bounce), "pk_shade")

    def trace_shadow(self, scene):
        _check(lib().pk_trace_shadow(self.h, scene.h), "pk_trace_shadow")

    def accumulate(self):
        _check(lib().pk_accumulate(self.h), "pk_accumulate")

    def render(self, scene, max_depth=4, first_sample=0, num_frames=1, env_theta=0.0, use_graph=True):
        p = RenderParams(self.width, self.height, max_depth, first_sample, num_frames, env_theta, int(use_graph))
        _check(lib().pk_render(self.h, scene.h, ctypes.byref(p)), "pk_render")

    def autotune(self, scene, max_depth=4, frames=2):
        """Choose traversal launch shapes for this scene/camera/resolution
        (resets the accumulation).  Returns (ms/frame before, after)."""
        ms = (ctypes.c_float * 2)()
        _check(lib().pk_autotune_render(self.h, scene.h, max_depth, frames, ms), "pk_autotune_render")
        return float(ms[0]), float(ms[1])

    def autotune_samples(self, scene, max_depth=4, max_samples=8):
        """Choose this frame's samples per pass (queue capacity) by timing;
        returns (ms/sample before, after, chosen samples per pass)."""
        ms = (ctypes.c_float * 2)()
        _check(lib().pk_autotune_samples_per_pass(self.h, scene.h, max_depth, max_samples, ms),
               "pk_autotune_samples_per_pass")
        return float(ms[0]), float(ms[1])

    def autotune_scheduler(self, scene, max_depth=4, samples=8):
        """Choose the context's pass scheduler ("wavefront" or "paths") by
        timing `samples` samples under each (resets the accumulation);
        returns (ms/sample before, after)."""
        ms = (ctypes.c_float * 2)()
        _check(lib().pk_autotune_scheduler(self.h, scene.h, max_depth, samples, ms), "pk_autotune_scheduler")
        return float(ms[0]), float(ms[1])

    def profile(self, scene, max_depth=4, first_sample=0, num_frames=1, env_theta=0.0):
        p = RenderParams(self.width, self.height, max_depth, first_sample, num_frames, env_theta, 0)
        ms = np.zeros(5, np.float64)
        _check(lib().pk_render_profile(self.h, scene.h, ctypes.byref(p), _ptr(ms)), "pk_render_profile")
        return dict(zip(["raygen", "trace_closest", "shade", "trace_shadow", "accumulate"], ms.tolist()))

    def read_image(self):
        out = np.empty((self.height, self.width, 4), np.float32)
        _check(lib().pk_read_image(self.h, _ptr(out), out.nbytes), "pk_read_image")
        return out

    def read_image_device(self, dst):
        """Copies the running-mean image to device memory (asynchronous)."""
        _check(lib().pk_read_image_device(self.h, _dev(dst), self.width * self.height * 16), "pk_read_image_device")

    # differentiation
    def loss_l2(self, target, grad, want_loss=True):
        """L2 loss of the running-mean image against a device target; writes
        2 (image - target) to the device buffer `grad`.  target/grad are
        DeviceBuffers or device pointers.  Returns the loss (blocking) or None."""
        loss = ctypes.c_double()
        _check(lib().pk_loss_l2(self.h, _dev(target), _dev(grad), ctypes.byref(loss) if want_loss else None),
               "pk_loss_l2")
        return loss.value if want_loss else None

    def adjoint(self, scene, d_image, max_depth=4, first_sample=0, num_frames=1, env_theta=0.0):
        """Path-replay adjoint with host arrays: d_image (H, W, 4) f32; returns
        (atlas gradient [layers][h][w][4], environment gradient [h][w][4])."""
        info = scene.param_info()
        d = np.ascontiguousarray(d_image, np.float32)
        assert d.shape == (self.height, self.width, 4)
        ga = np.empty((info["tex_layers"], info["tex_height"], info["tex_width"], 4), np.float32)
        ge = np.empty((info["env_height"], info["env_width"], 4), np.float32)
        g = _ffi.GradBuffers(ga.ctypes.data, ge.ctypes.data if ge.size else None, 0, 0)
        p = RenderParams(self.width, self.height, max_depth, first_sample, num_frames, env_theta, 0)
        _check(lib().pk_adjoint(self.h, scene.h, ctypes.byref(p), _ptr(d), 0, ctypes.byref(g)), "pk_adjoint")
        return ga, ge

    def adjoint_device(self, scene, d_image, atlas_grad, env_grad, max_depth=4, first_sample=0, num_frames=1,
                       env_theta=0.0, accumulate=False):
        """Asynchronous adjoint on device memory (DeviceBuffers or pointers)."""
        g = _ffi.GradBuffers(_dev(atlas_grad), _dev(env_grad), 1, int(accumulate))
        p = RenderParams(self.width, self.height, max_depth, first_sample, num_frames, env_theta, 0)
        _check(lib().pk_adjoint(self.h, scene.h, ctypes.byref(p), _dev(d_image), 1, ctypes.byref(g)),
               "pk_adjoint")

    def read_color(self):
        out = np.empty((self.height, self.width, 4), np.float32)
        _check(lib().pk_read_color(self.h, _ptr(out), out.nbytes), "pk_read_color")
        return out

    def counters(self):
        c = Counters()
        _check(lib().pk_read_counters(self.h, ctypes.byref(c)), "pk_read_counters")
        return {k: int(getattr(c, k)) for k, _ in Counters._fields_}

    def close(self):
        if self.h:
            lib().pk_frame_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
