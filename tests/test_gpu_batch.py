"""Batched multi-scene, multi-view tracing (one global wavefront, one graph
launch) renders exactly what each view renders alone through pk_render:
seeds depend only on the view-local pixel, the sample and the bounce, and
the traversal result is the canonical closest hit in each scene's BVH."""
import numpy as np
import pytest

import pike_amd
from oracle.oracle import OracleScene
from pike_amd import scenes

pytestmark = pytest.mark.gpu


def _small_batch():
    env = scenes.lognormal_sun_env(64, 32, seed=5, sigma=1.0, sun_deg=6.0, sun_scale=30.0)
    sc = [scenes.mixed_scene(100 + k, n, env=env) for k, n in enumerate([1200, 5000, 20000])]
    views = [(0, 40, 30, 3), (1, 64, 48, 1), (2, 32, 32, 4), (1, 24, 40, 2), (0, 16, 16, 5)]
    return sc, views


@pytest.mark.parametrize("use_graph,sched,per_round", [(True, "wavefront", 1), (False, "wavefront", 1),
                                                       (True, "paths", 1), (False, "paths", 1),
                                                       (True, "wavefront", 2), (True, "wavefront", 3),
                                                       (False, "paths", 2), (True, "paths", 4)])
def test_batch_equals_per_view_render(ctx, use_graph, sched, per_round):
    """Under either pass scheduler for the batch and with rounds of several
    samples (views whose spp is not a multiple end inside a round); the
    single views always render through the wavefront kernels."""
    sc, views = _small_batch()
    bvhs = [pike_amd.BVH(s.positions, s.triangles) for s in sc]
    merged = pike_amd.BatchScene(ctx, sc, bvhs)
    cams = []
    for i, (k, w, h, spp) in enumerate(views):
        az = 0.7 * i
        cam = scenes.camera((0.5 + 0.8 * np.cos(az), 0.4, 0.5 + 0.8 * np.sin(az)), (0.5, 0.05, 0.5), 1.0)
        cams.append(scenes.normalized_camera(cam, bvhs[k].normalizer()))
    batch = pike_amd.Batch(ctx, merged, [(k, w, h, spp, cams[i]) for i, (k, w, h, spp) in enumerate(views)])
    batch.set_samples_per_round(per_round)
    assert batch.samples_per_round == per_round
    ctx.set_scheduler(sched)
    try:
        batch.render(max_depth=4, use_graph=use_graph)
    finally:
        ctx.set_scheduler("wavefront")
    assert batch.counters()["stack_overflow"] == 0
    for i, (k, w, h, spp) in enumerate(views):
        scene = pike_amd.Scene(ctx, sc[k], bvhs[k])
        fr = pike_amd.Frame(ctx, w, h)
        fr.set_camera(cams[i])
        fr.render(scene, max_depth=4, first_sample=0, num_frames=spp)
        np.testing.assert_array_equal(batch.read_view(i).view(np.uint32), fr.read_image().view(np.uint32),
                                      err_msg=f"view {i}")


_ORACLE_VIEWS = {}


def _oracle_view(sc, k, cam, w, h, spp):
    """FP32 oracle render of one view (cached across parametrizations)."""
    key = (k, tuple(cam.origin), tuple(cam.dir), w, h, spp)
    if key not in _ORACLE_VIEWS:
        orc = OracleScene(sc[k])
        orc.build_bvh()
        _ORACLE_VIEWS[key] = orc.render(cam, w, h, 4, 0, spp, precision=32)[0]
    return _ORACLE_VIEWS[key]


@pytest.mark.parametrize("sched,per_round", [("wavefront", 1), ("wavefront", 3), ("paths", 2)])
def test_batch_views_equal_fp32_oracle(ctx, sched, per_round):
    """Every view of one graph-launched batch (rounds of one and several
    samples per view pixel) equals the oracle's FP32 render of that view
    alone, bit for bit: view-local pixel seeds, canonical hits in the view's
    own scene and the frame's accumulate order."""
    sc, views = _small_batch()
    bvhs = [pike_amd.BVH(s.positions, s.triangles) for s in sc]
    merged = pike_amd.BatchScene(ctx, sc, bvhs)
    cams = []
    for i, (k, w, h, spp) in enumerate(views):
        az = 0.7 * i + 0.3
        cam = scenes.camera((0.5 + 0.9 * np.cos(az), 0.45, 0.5 + 0.9 * np.sin(az)), (0.5, 0.05, 0.5), 1.0)
        cams.append(scenes.normalized_camera(cam, bvhs[k].normalizer()))
    batch = pike_amd.Batch(ctx, merged, [(k, w, h, spp, cams[i]) for i, (k, w, h, spp) in enumerate(views)])
    batch.set_samples_per_round(per_round)
    ctx.set_scheduler(sched)
    try:
        batch.render(max_depth=4, use_graph=True)
    finally:
        ctx.set_scheduler("wavefront")
    assert batch.counters()["stack_overflow"] == 0
    for i, (k, w, h, spp) in enumerate(views):
        ref = _oracle_view(sc, k, cams[i], w, h, spp)
        np.testing.assert_array_equal(batch.read_view(i).view(np.uint32), ref.view(np.uint32), err_msg=f"view {i}")


def test_batch_rejects_bad_views(ctx):
    sc, _ = _small_batch()
    bvhs = [pike_amd.BVH(s.positions, s.triangles) for s in sc]
    merged = pike_amd.BatchScene(ctx, sc, bvhs)
    cam = scenes.normalized_camera(scenes.camera((0.5, 0.5, 1.5), (0.5, 0.0, 0.5)), bvhs[0].normalizer())
    with pytest.raises(pike_amd.PikeError):
        pike_amd.Batch(ctx, merged, [(7, 8, 8, 1, cam)])
    with pytest.raises(pike_amd.PikeError):
        pike_amd.Batch(ctx, merged, [(0, 8, 8, 0, cam)])
