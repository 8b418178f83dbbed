"""The two pass schedulers of single-view frames, PK_SCHED_WAVEFRONT
(raygen, then per bounce trace_closest / shade / trace_shadow over device
queues) and PK_SCHED_PATHS (one path megakernel), give bit-identical color
buffers and running means: same path numbering, seeds, arithmetic and
contribution order.  The wavefront side is itself pinned against the oracle
in test_gpu_render."""
import numpy as np
import pytest

import pike_amd
from pike_amd import scenes

pytestmark = pytest.mark.gpu


def _render(ctx, data, bvh, sched, W, H, depth, spp, per_pass, first=0, graph=True):
    ctx.set_scheduler(sched)
    scene = pike_amd.Scene(ctx, data, bvh)
    fr = pike_amd.Frame(ctx, W, H)
    fr.set_camera(scenes.normalized_camera(data.camera, bvh.normalizer()))
    fr.set_samples_per_pass(per_pass)
    fr.render(scene, max_depth=depth, first_sample=first, num_frames=spp, use_graph=graph)
    out = fr.read_image(), fr.read_color(), fr.counters()
    ctx.set_scheduler("wavefront")
    return out


@pytest.mark.parametrize("depth,spp,per_pass,graph", [(1, 3, 1, False), (4, 8, 4, True), (8, 6, 4, True),
                                                     (4, 45, 40, True)])
def test_paths_scheduler_is_bit_identical(ctx, depth, spp, per_pass, graph):
    data = scenes.config1(seed=5)
    bvh = pike_amd.BVH(data.positions, data.triangles)
    a = _render(ctx, data, bvh, "wavefront", 80, 56, depth, spp, per_pass, first=3, graph=graph)
    b = _render(ctx, data, bvh, "paths", 80, 56, depth, spp, per_pass, first=3, graph=graph)
    np.testing.assert_array_equal(a[0].view(np.uint32), b[0].view(np.uint32))
    np.testing.assert_array_equal(a[1].view(np.uint32), b[1].view(np.uint32))
    for k in ("rays_traced", "shadow_rays", "frames"):
        assert a[2][k] == b[2][k], k
    assert b[2]["stack_overflow"] == 0


def test_paths_scheduler_on_the_1080p_scene_type(ctx):
    """The headline workload's scene (GGX spheres, textured ground, importance-
    sampled environment) at a reduced size, with the tail of a run that ends
    inside a pass."""
    data = scenes.config2(cells=64, num_spheres=20, subdiv=2, env_size=(64, 32))
    bvh = pike_amd.BVH(data.positions, data.triangles)
    a = _render(ctx, data, bvh, "wavefront", 96, 54, 8, 11, 4)
    b = _render(ctx, data, bvh, "paths", 96, 54, 8, 11, 4)
    np.testing.assert_array_equal(a[0].view(np.uint32), b[0].view(np.uint32))
    np.testing.assert_array_equal(a[1].view(np.uint32), b[1].view(np.uint32))


def test_paths_scheduler_instanced(ctx):
    data = scenes.instanced_scene(9, num_instances=10)
    bvh = pike_amd.BVH.for_scene(data)
    a = _render(ctx, data, bvh, "wavefront", 64, 48, 5, 4, 2)
    b = _render(ctx, data, bvh, "paths", 64, 48, 5, 4, 2)
    np.testing.assert_array_equal(a[0].view(np.uint32), b[0].view(np.uint32))


def test_scheduler_api(ctx):
    assert ctx.scheduler() == "wavefront"
    ctx.set_scheduler("paths")
    assert ctx.scheduler() == "paths"
    ctx.set_scheduler("wavefront")
    with pytest.raises(KeyError):
        ctx.set_scheduler("other")
    assert ctx.tuning()["paths"]["waves_per_cu"] == 16
    with pytest.raises(pike_amd.PikeError):
        ctx.set_tuning("paths", 17, 64)


def test_autotune_scheduler_leaves_a_consistent_context(ctx):
    data = scenes.config1(seed=6)
    bvh = pike_amd.BVH(data.positions, data.triangles)
    scene = pike_amd.Scene(ctx, data, bvh)
    fr = pike_amd.Frame(ctx, 64, 48)
    fr.set_camera(scenes.normalized_camera(data.camera, bvh.normalizer()))
    fr.set_samples_per_pass(2)
    before, after = fr.autotune_scheduler(scene, max_depth=4, samples=4)
    assert after <= before and after > 0
    chosen = ctx.scheduler()
    fr.render(scene, max_depth=4, num_frames=4)
    img = fr.read_image()
    ctx.set_scheduler("paths" if chosen == "wavefront" else "wavefront")
    fr.reset()
    fr.render(scene, max_depth=4, num_frames=4)
    np.testing.assert_array_equal(img.view(np.uint32), fr.read_image().view(np.uint32))
    ctx.set_scheduler("wavefront")
