"""Every launch shape of the traversal kernels traces the same hits and
renders the same image: node width 8 or 4 (the 64-byte layout derived from
the 128-byte one), LDS stack part of 8, 12 or 16 entries, 1, 2 or 4 waves
per block, and the leaf-phase trigger.  Hits are compared bit for bit with
the oracle's canonical closest hit, frames with the oracle's FP32 render,
batches with per-view renders; invalid shapes are refused."""
import numpy as np
import pytest

import pike_amd
from pike_amd import scenes
from oracle.oracle import OracleScene
from tests.test_gpu_render import _compare
from tests.test_gpu_trace import _assert_same

pytestmark = pytest.mark.gpu

# (node_width, stack_lds, waves_per_block, leaf_phase16)
SHAPES = [(8, 8, 1, 8), (8, 8, 2, 4), (8, 8, 4, 8), (8, 12, 1, 12), (8, 12, 2, 8), (8, 12, 4, 16),
          (8, 16, 1, 8), (8, 16, 2, 6), (8, 16, 4, 8), (4, 8, 1, 8), (4, 8, 1, 16), (4, 12, 1, 4)]
TRACE_CLASSES = ("primary", "secondary", "shadow")


def _waves(lds, wpb):
    """A launch width every block of the stack part allows (a wave's LDS holds
    its stack part and the leaf phase's 1 KB: 16 waves per CU fit the
    16-entry part in blocks of up to 4 waves, 20 the 12-entry part)."""
    return (16 if lds == 16 else 20) // wpb * wpb


@pytest.fixture
def shaped(ctx):
    """The session context, its launch shapes restored afterwards."""
    saved = ctx.tuning()
    yield ctx
    for cls, sh in saved.items():
        ctx.set_tuning(cls, sh["waves_per_cu"], sh["refill_min"], sh["stack_lds"], sh["waves_per_block"],
                       sh["leaf_phase16"], sh["node_width"])


def _apply(ctx, classes, shape, refill=16):
    width, lds, wpb, leaf = shape
    for cls in classes:
        ctx.set_tuning(cls, _waves(lds, wpb), refill, lds, wpb, leaf, width)
        t = ctx.tuning()[cls]
        assert (t["node_width"], t["stack_lds"], t["waves_per_block"], t["leaf_phase16"]) == shape


@pytest.fixture(scope="module")
def soup():
    data = scenes.config1(seed=4)
    bvh = pike_amd.BVH(data.positions, data.triangles)
    orc = OracleScene(data)
    orc.build_bvh()
    rays = scenes.incoherent_rays(30_000, bvh.normalizer(), seed=41)
    return data, bvh, orc, rays, orc.trace(rays, "canonical"), orc.trace(rays, "canonical", any_hit=True)


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "w%d_lds%d_wpb%d_leaf%d" % s)
def test_ray_queries_every_shape(shaped, soup, shape):
    data, bvh, orc, rays, (ids_o, t_o, b_o, _), (occ_o, _, _, _) = soup
    scene = pike_amd.Scene(shaped, data, bvh)
    _apply(shaped, ("rays_closest", "rays_any"), shape)
    ids, t, bary = shaped.trace_rays(scene, rays)
    _assert_same(ids, t, bary, ids_o, t_o, b_o)
    occ, _, _ = shaped.trace_rays(scene, rays, any_hit=True)
    np.testing.assert_array_equal(occ >= 0, occ_o >= 0)


@pytest.mark.parametrize("shape", [(4, 8, 1, 8), (4, 12, 1, 12), (8, 16, 4, 8)], ids=lambda s: "w%d_lds%d_wpb%d_leaf%d" % s)
def test_spatial_split_soup_every_width(shaped, shape):
    """Spatial splits put one triangle in several leaves; the 4-wide layout
    keeps their references."""
    data = scenes.config3(seed=19, num_tris=60_000)
    bvh = pike_amd.BVH(data.positions, data.triangles, spatial_splits=True, sah_collapse=True)
    scene = pike_amd.Scene(shaped, data, bvh)
    orc = OracleScene(data)
    orc.build_bvh()
    rays = scenes.incoherent_rays(40_000, bvh.normalizer(), seed=20)
    _apply(shaped, ("rays_closest", "rays_any"), shape)
    ids, t, bary = shaped.trace_rays(scene, rays)
    ids_o, t_o, b_o, _ = orc.trace(rays, "canonical")
    _assert_same(ids, t, bary, ids_o, t_o, b_o)
    occ, _, _ = shaped.trace_rays(scene, rays, any_hit=True)
    np.testing.assert_array_equal(occ >= 0, ids_o >= 0)


@pytest.mark.parametrize("shape", [(4, 8, 1, 8), (4, 12, 1, 4), (8, 16, 2, 8), (8, 12, 4, 12)],
                         ids=lambda s: "w%d_lds%d_wpb%d_leaf%d" % s)
def test_render_every_width_matches_fp32_oracle(shaped, shape):
    data = scenes.config1(seed=1)
    bvh = pike_amd.BVH(data.positions, data.triangles)
    scene = pike_amd.Scene(shaped, data, bvh)
    orc = OracleScene(data)
    orc.build_bvh()
    cam = scenes.normalized_camera(data.camera, bvh.normalizer())
    W, H, depth, frames = 48, 40, 4, 2
    _apply(shaped, TRACE_CLASSES, shape)
    fr = pike_amd.Frame(shaped, W, H)
    fr.set_camera(cam)
    fr.render(scene, max_depth=depth, first_sample=0, num_frames=frames, use_graph=True)
    img, col = fr.read_image(), fr.read_color()
    acc_o, col_o, _ = orc.render(cam, W, H, depth, 0, frames, precision=32)
    _compare(col, col_o, "last frame color")
    _compare(img, acc_o, "running mean")
    np.testing.assert_array_equal(col[..., 3], col_o[..., 3])
    assert fr.counters()["stack_overflow"] == 0


def test_batch_width4_equals_width8(shaped):
    """Batched scenes: each view starts at its scene's root, which has the
    same index in both layouts."""
    from tests.test_gpu_batch import _small_batch

    sc, views = _small_batch()
    bvhs = [pike_amd.BVH(s.positions, s.triangles) for s in sc]
    merged = pike_amd.BatchScene(shaped, sc, bvhs)
    cams = []
    for i, (k, w, h, spp) in enumerate(views):
        cam = scenes.camera((0.5 + 0.8 * np.cos(0.9 * i), 0.45, 0.5 + 0.8 * np.sin(0.9 * i)), (0.5, 0.05, 0.5), 1.0)
        cams.append(scenes.normalized_camera(cam, bvhs[k].normalizer()))
    out = {}
    for shape in [(8, 8, 1, 8), (4, 8, 1, 8)]:
        _apply(shaped, TRACE_CLASSES, shape)
        batch = pike_amd.Batch(shaped, merged, [(k, w, h, spp, cams[i]) for i, (k, w, h, spp) in enumerate(views)])
        batch.render(max_depth=4, use_graph=True)
        out[shape[0]] = [batch.read_view(i).copy() for i in range(len(views))]
    for i in range(len(views)):
        np.testing.assert_array_equal(out[4][i].view(np.uint32), out[8][i].view(np.uint32), err_msg=f"view {i}")


def test_instanced_scene_under_every_width(shaped):
    """Instanced scenes walk 8-wide nodes in the instanced kernels (12 LDS
    entries, one wave per block) whatever the classes' shapes ask; hits and
    frames stay the oracle's."""
    data = scenes.instanced_scene(7, num_instances=20)
    bvh = pike_amd.BVH.for_scene(data)
    scene = pike_amd.Scene(shaped, data, bvh)
    orc = OracleScene(data, threads=16)
    orc.build_bvh()
    cam = scenes.normalized_camera(data.camera, bvh.normalizer())
    rays = scenes.incoherent_rays(30_000, bvh.normalizer(), seed=71)
    ids_o, t_o, b_o, _ = orc.trace(rays, "canonical")
    ref, _, _ = orc.render(cam, 40, 32, 3, 0, 2, precision=32)
    for shape in [(4, 8, 1, 8), (8, 16, 4, 12)]:
        _apply(shaped, TRACE_CLASSES + ("rays_closest", "rays_any"), shape)
        ids, t, bary = shaped.trace_rays(scene, rays)
        _assert_same(ids, t, bary, ids_o, t_o, b_o)
        fr = pike_amd.Frame(shaped, 40, 32)
        fr.set_camera(cam)
        fr.render(scene, max_depth=3, first_sample=0, num_frames=2, use_graph=True)
        np.testing.assert_array_equal(fr.read_image().view(np.uint32), ref.view(np.uint32))


def test_invalid_shapes_are_refused(shaped):
    bad = [dict(node_width=4, stack_lds=16, waves_per_block=1),   # 4-wide: 8 or 12 LDS entries
           dict(node_width=4, stack_lds=8, waves_per_block=2),    # 4-wide: one wave per block
           dict(node_width=6, stack_lds=8, waves_per_block=1),
           dict(node_width=8, stack_lds=10, waves_per_block=1),
           dict(node_width=8, stack_lds=8, waves_per_block=3)]
    for b in bad:
        with pytest.raises(pike_amd.PikeError):
            shaped.set_tuning("secondary", 16, 16, b["stack_lds"], b["waves_per_block"], 8, b["node_width"])
    with pytest.raises(pike_amd.PikeError):
        shaped.set_tuning("paths", 16, 64, 12, 1, 8, 4)    # the path megakernel walks 8-wide nodes
    with pytest.raises(pike_amd.PikeError):
        shaped.set_tuning("primary", 27, 16, 12, 1, 8, 8)  # 12 LDS entries fit 26 waves per CU
    with pytest.raises(pike_amd.PikeError):
        shaped.set_tuning("shadow", 24, 16, 8, 1, 17, 8)   # leaf_phase16 in [1, 16]
    # a refused shape leaves the class as it was
    before = shaped.tuning()["secondary"]
    with pytest.raises(pike_amd.PikeError):
        shaped.set_tuning("secondary", 16, 16, 16, 1, 8, 4)
    assert shaped.tuning()["secondary"] == before
