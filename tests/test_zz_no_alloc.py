"""pk_trace_rays_noalloc (the allocation-free variant of
pk_trace_rays_device, for callers that cannot allocate device memory) gives
bit-identical closest hits and equal occlusion on two calls in a row, the
second smaller than the first (the shared scratch keeps the previous call's
contents; the work cursor must be cleared), and allocates nothing through
the library."""
import ctypes

import numpy as np
import pytest

import pike_amd
from pike_amd import scenes

pytestmark = pytest.mark.gpu


def _call(name, *args):
    rc = getattr(pike_amd.lib(), name)(*args)
    assert rc == 0, f"{name}: {pike_amd.lib().pk_last_error()}"


def test_noalloc_equals_device_entry_point():
    lib = pike_amd.lib()
    lib.pk_trace_rays_noalloc.restype = ctypes.c_int
    lib.pk_trace_rays_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int,
                                          ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    ctx = pike_amd.Context(0)
    try:
        data = scenes.config1(seed=6)
        bvh = pike_amd.BVH(data.positions, data.triangles)
        scene = pike_amd.Scene(ctx, data, bvh)
        rays_h = scenes.incoherent_rays(300_000, bvh.normalizer(), seed=61)
        rays = pike_amd.DeviceBuffer.from_array(ctx, rays_h)
        n1, n2 = rays_h.shape[0], 70_001
        out = {k: pike_amd.DeviceBuffer(ctx, 4 * n1) for k in ("ids", "t", "b", "ids2", "t2", "b2")}
        used = ctx.memory_used()
        for n in (n1, n2):
            for any_hit in (0, 1):
                _call("pk_trace_rays_device", ctx.h, scene.h, rays.ptr, n, any_hit, out["ids"].ptr, out["t"].ptr,
                      out["b"].ptr, 0)
                # poison the outputs so a call that skips rays cannot pass
                for k in ("ids2", "t2", "b2"):
                    out[k].write(np.full(n1, 0x7fc00001, np.uint32))
                _call("pk_trace_rays_noalloc", scene.h, rays.ptr, n, any_hit, out["ids2"].ptr, out["t2"].ptr,
                      out["b2"].ptr)
                ctx.sync()
                ids, ids2 = out["ids"].read(np.int32, n), out["ids2"].read(np.int32, n)
                if any_hit:
                    np.testing.assert_array_equal(ids >= 0, ids2 >= 0)
                else:
                    np.testing.assert_array_equal(ids, ids2)
                    np.testing.assert_array_equal(out["t"].read(np.uint32, n), out["t2"].read(np.uint32, n))
                    np.testing.assert_array_equal(out["b"].read(np.uint32, n), out["b2"].read(np.uint32, n))
                # rays past n are untouched
                assert np.all(out["ids2"].read(np.uint32, n1)[n:] == 0x7fc00001)
                nodes = (ctypes.c_uint64 * 2)()
                _call("pk_ctx_trace_stats", ctx.h, nodes)  # no overflow reported
        assert ctx.memory_used() == used
        # argument checks fail cleanly
        assert lib.pk_trace_rays_noalloc(scene.h, rays.ptr, -1, 0, out["ids2"].ptr, out["t2"].ptr,
                                         out["b2"].ptr) != 0
        assert lib.pk_trace_rays_noalloc(scene.h, None, 10, 0, out["ids2"].ptr, out["t2"].ptr, out["b2"].ptr) != 0
        assert lib.pk_trace_rays_noalloc(None, rays.ptr, 10, 0, out["ids2"].ptr, out["t2"].ptr, out["b2"].ptr) != 0
    finally:
        ctx.close()
