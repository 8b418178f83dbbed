"""TEST INFRASTRUCTURE ONLY — ctypes wrapper of the oracle (oracle/liboracle.so).

Imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
as the checker of the HIP path; the product never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        build()
    lib = ctypes.CDLL(LIB)
    vp, i32, i64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64
    lib.or_scene_create.restype = vp
    lib.or_scene_create.argtypes = [vp]
    lib.or_scene_destroy.argtypes = [vp]
    lib.or_scene_set_instances.restype = i32
    lib.or_scene_set_instances.argtypes = [vp, vp, ctypes.c_uint32, vp, ctypes.c_uint32]
    lib.or_bvh_build.restype = i32
    lib.or_bvh_build.argtypes = [vp, i32]
    lib.or_bvh_info.argtypes = [vp, vp]
    lib.or_bvh_serialize.restype = i32
    lib.or_bvh_serialize.argtypes = [vp, vp, vp, vp, vp]
    lib.or_trace.restype = i32
    lib.or_trace.argtypes = [vp, vp, i64, i32, i32, i32, vp, vp, vp, vp]
    lib.or_env_info.restype = i32
    lib.or_env_info.argtypes = [vp, vp]
    lib.or_env_tables.restype = i32
    lib.or_env_tables.argtypes = [vp, vp, vp, vp]
    lib.or_render.restype = i32
    lib.or_render.argtypes = [vp, vp, vp, vp, vp, vp]
    lib.or_math_eval.argtypes = [i32, vp, vp, vp, i64]
    lib.or_scene_write_params.restype = i32
    lib.or_scene_write_params.argtypes = [vp, i32, i64, vp, i64, i32]
    lib.or_gradient.restype = i32
    lib.or_gradient.argtypes = [vp, vp, vp, vp, vp, vp, i64, vp, vp]
    lib.or_probe.restype = i32
    lib.or_probe.argtypes = [vp, i32, i32, vp, i64, i64, vp, i64]
    lib.or_set_mutation.argtypes = [i32]
    lib.or_mutation_count.restype = i32
    _lib = lib
    return lib


class RenderParams(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32), ("max_depth", ctypes.c_uint32),
                ("first_sample", ctypes.c_uint32), ("num_frames", ctypes.c_uint32), ("env_theta", ctypes.c_float),
                ("precision", ctypes.c_int), ("num_threads", ctypes.c_int), ("x0", ctypes.c_uint32),
                ("x1", ctypes.c_uint32), ("y0", ctypes.c_uint32), ("y1", ctypes.c_uint32)]


def _p(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


MATH_FNS = {"sin": 0, "cos": 1, "tan": 2, "atan2": 3, "acos": 4, "log2": 5, "exp2": 6, "pow": 7, "sqrt": 8,
            "div": 9, "pcg": 10, "unorm16": 11, "snorm16": 12, "snorm8": 13, "unorm8": 14}


def math_eval(fn, x, y=None):
    x = np.ascontiguousarray(x, np.float32)
    y = np.ascontiguousarray(x if y is None else y, np.float32)
    out = np.empty_like(x)
    load().or_math_eval(MATH_FNS[fn], _p(x), _p(y), _p(out), x.size)
    return out


PROBES = ["ggx_d", "ggx_g1", "lambert_pdf", "specular_pdf", "eval_lambert", "eval_specular", "power_heuristic",
          "schlick", "onb", "sample_ggx", "sample_lambert", "env_color", "env_pdf", "sample_env", "texture",
          "primary_ray", "surface", "accumulate"]
PROBE_OUT = {"eval_lambert": 3, "eval_specular": 3, "onb": 9, "sample_ggx": 3, "sample_lambert": 4, "env_color": 3,
             "sample_env": 4, "texture": 4, "primary_ray": 6, "surface": 35}


def probe(fn, inputs, precision=64, scene=None):
    """Evaluates one restated function (oracle.h or_probe) on each row of
    `inputs`; returns [n, outputs] float64."""
    x = np.ascontiguousarray(np.atleast_2d(np.asarray(inputs, np.float64)))
    nout = PROBE_OUT.get(fn, 1)
    out = np.zeros((x.shape[0], nout), np.float64)
    h = scene.h if scene is not None else None
    rc = load().or_probe(h, PROBES.index(fn), precision, _p(x), x.shape[1], x.shape[0], _p(out), nout)
    if rc != 0:
        raise ValueError(f"or_probe({fn}) failed")
    return out


MUTATIONS = ["g1_swap", "lambert_inv_pi", "onb_sign", "jitter_order", "d_pi", "schlick_pow", "ph_balance",
             "lambert_sample", "vndf_blend", "env_pdf_sin", "env_center", "acc_denom", "normal_map", "origin_sign",
             "aspect", "spec_pdf4", "env_color_v", "tex_half", "emission", "nee_spec_weight", "bounce_weight",
             "eval_spec_cos", "sample_env_sin", "metal_channel", "alpha_square", "distortion", "g1_sign",
             "fresnel_metal", "theta_sign"]


def set_mutation(name):
    """Plants one of the test-only mistakes of oracle.cpp (None: none)."""
    lib = load()
    assert lib.or_mutation_count() == len(MUTATIONS) + 1
    lib.or_set_mutation(0 if name is None else MUTATIONS.index(name) + 1)


class OracleScene:
    """The restatement's view of a scene (same pk_scene_desc as the product)."""

    def __init__(self, data, threads=None):
        self.lib = load()
        self.data = data
        self.desc = data.desc()
        self.h = self.lib.or_scene_create(ctypes.byref(self.desc))
        self.threads = threads or min(16, os.cpu_count() or 1)
        self.built = False
        if data.instances:
            meshes, insts, keep = data.instance_descs()
            assert self.lib.or_scene_set_instances(self.h, ctypes.byref(meshes), len(data.meshes), ctypes.byref(insts),
                                                   len(data.instances)) == 0
            self._keep = (meshes, insts, keep)

    def build_bvh(self):
        rc = self.lib.or_bvh_build(self.h, self.threads)
        self.built = True
        return rc

    def info(self):
        a = np.zeros(4, np.int64)
        self.lib.or_bvh_info(self.h, _p(a))
        return dict(nodes=int(a[0]), leaf_triangles=int(a[1]), depth=int(a[2]), largest_leaf=int(a[3]))

    def serialize(self):
        info = self.info()
        nodes = np.zeros((info["nodes"], 5), np.int32)
        tris = np.zeros((self.data.num_triangles, 4), np.int32)
        verts = np.zeros((self.data.positions.shape[0], 4), np.float32)
        norm = np.zeros(4, np.float64)
        assert self.lib.or_bvh_serialize(self.h, _p(nodes), _p(tris), _p(verts), _p(norm)) == 0
        return nodes, tris, verts, norm

    def trace(self, rays, mode, any_hit=False):
        """mode: 'literal' (reference traversal), 'brute', 'canonical'."""
        rays = np.ascontiguousarray(rays, np.float32)
        n = rays.shape[0]
        ids = np.empty(n, np.int32)
        t = np.empty(n, np.float32)
        bary = np.empty(n, np.uint32)
        flags = np.empty(n, np.uint32)
        m = {"literal": 0, "brute": 1, "canonical": 2}[mode]
        assert self.lib.or_trace(self.h, _p(rays), n, m, int(any_hit), self.threads, _p(ids), _p(t), _p(bary),
                                 _p(flags)) == 0
        return ids, t, bary, flags

    def env_tables(self):
        info = np.zeros(4, np.int64)
        self.lib.or_env_info(self.h, _p(info))
        bins = np.zeros((int(info[0]), 2), np.int32)
        coords = np.zeros((int(info[1]), 2), np.int32)
        pdf = np.zeros((int(info[3]), int(info[2])), np.float32)
        self.lib.or_env_tables(self.h, _p(bins), _p(coords), _p(pdf))
        return bins, coords, pdf

    def render(self, cam, width, height, max_depth, first_sample=0, num_frames=1, env_theta=0.0, precision=32,
               window=None, accum=None):
        """Returns (running mean RGBA, last frame color RGBA, counters)."""
        x0, x1, y0, y1 = window if window else (0, 0, 0, 0)
        ww, wh = (x1 - x0, y1 - y0) if window else (width, height)
        if accum is None:
            accum = np.zeros((wh, ww, 4), np.float32)
        color = np.zeros((wh, ww, 4), np.float32)
        cnt = np.zeros(3, np.uint64)
        p = RenderParams(width, height, max_depth, first_sample, num_frames, env_theta, precision, self.threads,
                         x0, x1, y0, y1)
        assert self.lib.or_render(self.h, ctypes.byref(cam), ctypes.byref(p), _p(accum), _p(color), _p(cnt)) == 0
        return accum, color, cnt

    def write_params(self, which, values, offset=0, also_sampling=False):
        """Writes float parameters (PK_PARAM_ATLAS = 0 / PK_PARAM_ENV = 1 layouts)."""
        v = np.ascontiguousarray(values, np.float32).ravel()
        assert self.lib.or_scene_write_params(self.h, which, offset, _p(v), v.size, int(also_sampling)) == 0

    def gradient(self, cam, width, height, max_depth, d_image, params, first_sample=0, num_frames=1,
                 env_theta=0.0, precision=64):
        """FP64 gradient of J = sum(d_image * image) w.r.t. params = [(which, float offset), ...];
        returns (gradient, J)."""
        d = np.ascontiguousarray(d_image, np.float32).reshape(height, width, 4)
        which = np.ascontiguousarray([w for w, _ in params] or [0], np.int32)
        index = np.ascontiguousarray([i for _, i in params] or [0], np.int64)
        out = np.zeros(max(len(params), 1), np.float64)
        value = np.zeros(1, np.float64)
        p = RenderParams(width, height, max_depth, first_sample, num_frames, env_theta, precision, self.threads,
                         0, 0, 0, 0)
        assert self.lib.or_gradient(self.h, ctypes.byref(cam), ctypes.byref(p), _p(d), _p(which), _p(index),
                                    len(params), _p(out), _p(value)) == 0
        return out[:len(params)], float(value[0])

    def close(self):
        if self.h:
            self.lib.or_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
