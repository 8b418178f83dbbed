"""Traversal and rendering over a BVH built with spatial splits
(PK_BVH_SPATIAL_SPLITS): a triangle may sit in several leaves under
clipped boxes, and the closest hit must still be the canonical one,
argmin (t, original triangle id) over every triangle with the f32
Möller–Trumbore of trace.wgsl:97-114 — bit-exact ids, t bits and bary masks
against the oracle, and whole images bit-identical to the plain SAH BVH's."""
import numpy as np
import pytest

import pike_amd
from pike_amd import scenes
from oracle.oracle import OracleScene

from .test_bvh import _data, _slivers
from .test_gpu_trace import _assert_same

pytestmark = pytest.mark.gpu


def _sbvh_scene(ctx, data):
    bvh = pike_amd.BVH(data.positions, data.triangles, threads=8, spatial_splits=True)
    assert bvh.info()["leaf_triangles"] > data.num_triangles, "no spatial split was taken"
    return bvh, pike_amd.Scene(ctx, data, bvh)


def _check_canonical(ctx, data, nrays, seed):
    bvh, scene = _sbvh_scene(ctx, data)
    orc = OracleScene(data, threads=8)
    orc.build_bvh()
    rays = scenes.incoherent_rays(nrays, bvh.normalizer(), seed=seed)
    ids, t, bary = ctx.trace_rays(scene, rays)
    ids_o, t_o, b_o, _ = orc.trace(rays, "canonical")
    assert (ids_o >= 0).mean() > 0.05
    _assert_same(ids, t, bary, ids_o, t_o, b_o)
    occ, t_occ, _ = ctx.trace_rays(scene, rays, any_hit=True)
    occ_o, _, _, _ = orc.trace(rays, "canonical", any_hit=True)
    np.testing.assert_array_equal(occ >= 0, occ_o >= 0)
    hit = np.flatnonzero(occ >= 0)
    assert np.all(t_occ[hit] > 1e-7)


def test_soup_under_spatial_splits_matches_canonical(ctx):
    """The traversal benchmark's scene type: a random triangle soup."""
    _check_canonical(ctx, scenes.config3(seed=21, num_tris=200_000), 100_000, 22)


def test_slivers_under_spatial_splits_match_canonical(ctx):
    """Long thin triangles, where most references are clipped fragments."""
    pos, tri = _slivers(20_000, 5)
    _check_canonical(ctx, _data(pos, tri), 60_000, 6)


def test_render_with_spatial_splits_equals_plain_bvh(ctx):
    """Shading depends only on the closest hit, so the frame over the SBVH is
    bit-identical to the frame over the SAH BVH."""
    data = scenes.config1(seed=3)
    imgs = []
    for spatial in (False, True):
        bvh = pike_amd.BVH(data.positions, data.triangles, spatial_splits=spatial)
        scene = pike_amd.Scene(ctx, data, bvh)
        cam = scenes.normalized_camera(data.camera, bvh.normalizer())
        fr = pike_amd.Frame(ctx, 96, 64)
        fr.set_camera(cam)
        fr.render(scene, max_depth=4, first_sample=0, num_frames=4, use_graph=True)
        imgs.append(fr.read_image())
        assert fr.counters()["stack_overflow"] == 0
    np.testing.assert_array_equal(imgs[0].view(np.uint32), imgs[1].view(np.uint32))


@pytest.mark.parametrize("spatial", [False, True])
def test_sah_collapse_matches_canonical(ctx, spatial):
    """The SAH-optimal wide collapse (PK_BVH_SAH_COLLAPSE) changes the tree,
    not the answer: hits equal the oracle's canonical ones."""
    data = scenes.config3(seed=31, num_tris=150_000)
    bvh = pike_amd.BVH(data.positions, data.triangles, threads=8, spatial_splits=spatial, sah_collapse=True)
    scene = pike_amd.Scene(ctx, data, bvh)
    orc = OracleScene(data, threads=8)
    orc.build_bvh()
    rays = scenes.incoherent_rays(80_000, bvh.normalizer(), seed=32)
    ids, t, bary = ctx.trace_rays(scene, rays)
    ids_o, t_o, b_o, _ = orc.trace(rays, "canonical")
    _assert_same(ids, t, bary, ids_o, t_o, b_o)
    occ, _, _ = ctx.trace_rays(scene, rays, any_hit=True)
    occ_o, _, _, _ = orc.trace(rays, "canonical", any_hit=True)
    np.testing.assert_array_equal(occ >= 0, occ_o >= 0)
