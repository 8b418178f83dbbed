"""Failure paths of the traversal and the allocator, forced from the host so
that every kernel still runs to its normal end:

* a traversal stack limit below what the rays need (pk_ctx_set_stack_limit):
  the reference's 32-entry stack has no check (trace.wgsl:134-146, an
  out-of-bounds write); here the push is dropped, the ray flagged, and the
  call reports PK_E_OVERFLOW / counters.stack_overflow;
* rays with NaN, infinite or zero components: every box test and every
  Moller-Trumbore test of such a ray is false (NaN compares false; fmaxf and
  fminf pass the other operand, so the entry distance stays finite while the
  exit distance is NaN), so the defined answer is a miss, as in the oracle;
* a device memory cap (pk_ctx_set_memory_limit): frame and scene creation
  beyond it fail with PK_E_OOM, free what they had, and the context works on.
"""
import numpy as np
import pytest

import pike_amd
from oracle.oracle import OracleScene
from pike_amd import _ffi, scenes

pytestmark = pytest.mark.gpu


def _scene(ctx, data):
    bvh = pike_amd.BVH(data.positions, data.triangles)
    return bvh, pike_amd.Scene(ctx, data, bvh)


def test_stack_limit_forces_a_flagged_overflow():
    ctx = pike_amd.Context(0)
    try:
        data = scenes.config1(seed=2)
        bvh, scene = _scene(ctx, data)
        rays = scenes.incoherent_rays(40_000, bvh.normalizer(), seed=77)
        assert ctx.stack_limit() == 64
        ids, t, b = ctx.trace_rays(scene, rays)
        for limit in (1, 3):
            ctx.set_stack_limit(limit)
            assert ctx.stack_limit() == limit
            for any_hit in (False, True):
                with pytest.raises(pike_amd.PikeError) as e:
                    ctx.trace_rays(scene, rays, any_hit=any_hit)
                assert e.value.code == _ffi.PK_E_OVERFLOW
            # a frame finishes, flags the overflow, and stays finite
            fr = pike_amd.Frame(ctx, 48, 32)
            fr.set_camera(scenes.normalized_camera(data.camera, bvh.normalizer()))
            fr.render(scene, max_depth=4, num_frames=2, use_graph=True)
            assert fr.counters()["stack_overflow"] == 1
            assert np.isfinite(fr.read_image()).all()
            fr.close()
        for bad in (0, 65):
            with pytest.raises(pike_amd.PikeError):
                ctx.set_stack_limit(bad)
        ctx.set_stack_limit(64)
        ids2, t2, b2 = ctx.trace_rays(scene, rays)
        np.testing.assert_array_equal(ids, ids2)
        np.testing.assert_array_equal(t.view(np.uint32), t2.view(np.uint32))
    finally:
        ctx.close()


def test_nan_inf_and_zero_rays_miss(ctx):
    data = scenes.config1(seed=4, num_spheres=6)
    bvh, scene = _scene(ctx, data)
    orc = OracleScene(data)
    orc.build_bvh()
    good = scenes.incoherent_rays(2000, bvh.normalizer(), seed=5)
    nan, inf = np.nan, np.inf
    bad = np.zeros((9, 8), np.float32)
    bad[:, 4:7] = (0, -1, 0)
    bad[0, 4:7] = (nan, -1, 0)
    bad[1, 0:3] = (nan, 0, 0)
    bad[2, 4:7] = (inf, -1, 0)
    bad[3, 0:3] = (0, inf, 0)
    bad[4, 0:3] = (0, -inf, 0)
    bad[4, 4:7] = (0, 1, 0)
    bad[5, 4:7] = (-inf, -inf, -inf)
    bad[6, 4:7] = (0, 0, 0)
    bad[7, 4:7] = (nan, nan, nan)
    bad[8, 0:3] = (inf, -inf, nan)
    rays = np.concatenate([good[:1000], bad, good[1000:]])
    for any_hit in (False, True):
        ids, t, b = ctx.trace_rays(scene, rays, any_hit=any_hit)
        ids_o, t_o, b_o, _ = orc.trace(rays, "brute", any_hit=any_hit)
        assert np.all(ids[1000:1009] == -1) and np.all(t[1000:1009] == np.float32(1e5))
        if any_hit:  # occlusion only: the reported id of an any-hit query is order-dependent
            np.testing.assert_array_equal(ids >= 0, ids_o >= 0)
        else:
            np.testing.assert_array_equal(ids, ids_o)
            np.testing.assert_array_equal(t.view(np.uint32), t_o.view(np.uint32))
            np.testing.assert_array_equal(b, b_o)


def test_memory_cap_fails_cleanly():
    ctx = pike_amd.Context(0)
    try:
        data = scenes.config1(seed=1, num_spheres=8)
        used0 = ctx.memory_used()
        ctx.set_memory_limit(used0 + (1 << 20))
        with pytest.raises(pike_amd.PikeError) as e:
            pike_amd.Frame(ctx, 1920, 1080)
        assert e.value.code == _ffi.PK_E_OOM and "cap" in str(e.value)
        assert ctx.memory_used() == used0
        bvh = pike_amd.BVH(data.positions, data.triangles)
        ctx.set_memory_limit(used0 + 1024)
        with pytest.raises(pike_amd.PikeError) as e:
            pike_amd.Scene(ctx, data, bvh)
        assert e.value.code == _ffi.PK_E_OOM
        assert ctx.memory_used() == used0
        # a frame's queues beyond the cap: the resize fails, the frame keeps
        # its queues and samples per pass and renders on
        ctx.set_memory_limit(0)
        scene = pike_amd.Scene(ctx, data, bvh)
        fr = pike_amd.Frame(ctx, 64, 48)
        fr.set_camera(scenes.normalized_camera(data.camera, bvh.normalizer()))
        fr.set_samples_per_pass(2)
        used1 = ctx.memory_used()
        ctx.set_memory_limit(used1 + (1 << 16))
        with pytest.raises(pike_amd.PikeError) as e:
            fr.set_samples_per_pass(64)
        assert e.value.code == _ffi.PK_E_OOM
        assert fr.samples_per_pass == 2 and ctx.memory_used() == used1
        fr.render(scene, max_depth=3, num_frames=4)
        ctx.set_memory_limit(0)
        assert np.isfinite(fr.read_image()).all() and fr.read_image()[..., :3].max() > 0
        # the samples-per-pass search ends at the first size whose queues do
        # not fit (1 and 2 fit, 4 needs the old and new queues together)
        fr.set_samples_per_pass(1)
        u1 = ctx.memory_used()
        fr.set_samples_per_pass(2)
        per_sample = ctx.memory_used() - u1
        fr.set_samples_per_pass(1)
        ctx.set_memory_limit(u1 + 3 * per_sample)
        fr.autotune_samples(scene, max_depth=3, max_samples=64)
        assert fr.samples_per_pass in (1, 2)
        ctx.set_memory_limit(0)
        used = ctx.memory_used()
        fr.close()
        scene.close()
        assert ctx.memory_used() < used
    finally:
        ctx.close()
