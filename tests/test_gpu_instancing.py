"""Instanced scenes on the GPU (two-level BVH): the traversal, the frame and
the adjoint against the oracle's instanced restatement (bit-exact hits and
pixels, adjoint against forward-mode dual numbers), and the instanced frame
against the baked scene's (scene.js:159-170)."""
import numpy as np
import pytest

import pike_amd
from pike_amd import scenes
from oracle.oracle import OracleScene

from .test_gpu_trace import _assert_same

pytestmark = pytest.mark.gpu


def _setup(ctx, data, spatial=False):
    bvh = pike_amd.BVH.for_scene(data, spatial_splits=spatial)
    scene = pike_amd.Scene(ctx, data, bvh)
    orc = OracleScene(data, threads=16)
    orc.build_bvh()
    cam = scenes.normalized_camera(data.camera, bvh.normalizer())
    return bvh, scene, orc, cam


@pytest.mark.parametrize("seed,ground,spatial", [(3, True, False), (4, False, False), (5, True, True)])
def test_instanced_hits_match_canonical(ctx, seed, ground, spatial):
    data = scenes.instanced_scene(seed, num_instances=40, ground=ground)
    bvh, scene, orc, _ = _setup(ctx, data, spatial)
    rays = scenes.incoherent_rays(100_000, bvh.normalizer(), seed=seed + 50)
    ids, t, bary = ctx.trace_rays(scene, rays)
    ids_o, t_o, b_o, _ = orc.trace(rays, "canonical")
    assert (ids_o >= 0).mean() > 0.2
    _assert_same(ids, t, bary, ids_o, t_o, b_o)
    occ, t_occ, _ = ctx.trace_rays(scene, rays, any_hit=True)
    occ_o, _, _, _ = orc.trace(rays, "canonical", any_hit=True)
    np.testing.assert_array_equal(occ >= 0, occ_o >= 0)
    hit = np.flatnonzero(occ >= 0)
    assert np.all(t_occ[hit] > 1e-7)


@pytest.mark.parametrize("use_graph", [False, True])
def test_instanced_render_matches_fp32_oracle(ctx, use_graph):
    data = scenes.instanced_scene(6, num_instances=30)
    bvh, scene, orc, cam = _setup(ctx, data)
    W, H, depth, frames = 64, 48, 4, 3
    fr = pike_amd.Frame(ctx, W, H)
    fr.set_camera(cam)
    fr.render(scene, max_depth=depth, first_sample=0, num_frames=frames, use_graph=use_graph)
    img = fr.read_image()
    ref, col_o, _ = orc.render(cam, W, H, depth, 0, frames, precision=32)
    np.testing.assert_array_equal(img.view(np.uint32), ref.view(np.uint32))
    np.testing.assert_array_equal(fr.read_color()[..., 3], col_o[..., 3])
    assert fr.counters()["stack_overflow"] == 0


def test_instanced_terrain_scene_matches_fp32_oracle(ctx):
    """The 1080p scene's structure (terrain + icosphere instances, CDF
    environment sampling) at a small size."""
    data = scenes.config2(cells=48, num_spheres=30, subdiv=2, env_size=(128, 64), instanced=True)
    bvh, scene, orc, cam = _setup(ctx, data)
    W, H, depth, frames = 48, 32, 5, 2
    fr = pike_amd.Frame(ctx, W, H)
    fr.set_camera(cam)
    fr.set_samples_per_pass(2)
    fr.render(scene, max_depth=depth, first_sample=0, num_frames=frames)
    ref, _, _ = orc.render(cam, W, H, depth, 0, frames, precision=32)
    np.testing.assert_array_equal(fr.read_image().view(np.uint32), ref.view(np.uint32))


def test_instanced_frame_matches_baked_frame(ctx):
    """Same seeds, the same triangles hit: the instanced and baked frames
    differ only by the re-quantized tangent frames and f32 rounding."""
    data = scenes.instanced_scene(9, num_instances=30)
    W, H, depth, spp = 64, 48, 4, 64
    imgs = []
    for d in (data, data.bake()):
        bvh = pike_amd.BVH.for_scene(d)
        scene = pike_amd.Scene(ctx, d, bvh)
        fr = pike_amd.Frame(ctx, W, H)
        fr.set_camera(scenes.normalized_camera(d.camera, bvh.normalizer()))
        fr.render(scene, max_depth=depth, num_frames=spp)
        imgs.append(fr.read_image()[..., :3].astype(np.float64))
    rel_l1 = np.abs(imgs[0] - imgs[1]).sum() / np.abs(imgs[1]).sum()
    print("instanced vs baked mean relative L1", rel_l1)
    assert rel_l1 < 1e-2


def test_instanced_adjoint_matches_dual_numbers(ctx):
    """Float atlas + float environment on an instanced scene: pk_adjoint
    against the oracle's forward-mode FP32 gradients (the shading frames go
    through the instance transforms)."""
    data = scenes.instanced_scene(10, num_instances=24)
    data.tex = (data.tex.astype(np.float32) / 255.0)
    rng = np.random.default_rng(4)
    data.env = rng.uniform(0.3, 1.5, (16, 32, 3)).astype(np.float32)
    bvh, scene, orc, cam = _setup(ctx, data)
    W, H, depth, spp = 32, 24, 3, 2
    d_image = rng.uniform(-1, 1, (H, W, 4)).astype(np.float32)
    fr = pike_amd.Frame(ctx, W, H)
    fr.set_camera(cam)
    ga, ge = fr.adjoint(scene, d_image, max_depth=depth, num_frames=spp)
    ATLAS, ENV = pike_amd._ffi.PK_PARAM_ATLAS, pike_amd._ffi.PK_PARAM_ENV
    fa, fe = ga.ravel(), ge.ravel()
    pa = rng.choice(np.flatnonzero(fa), 24, replace=False)
    pe = rng.choice(np.flatnonzero(fe), 16, replace=False)
    params = [(ATLAS, int(i)) for i in pa] + [(ENV, int(i)) for i in pe]
    g = np.array([fa[i] for i in pa] + [fe[i] for i in pe], np.float64)
    g32, _ = orc.gradient(cam, W, H, depth, d_image, params, num_frames=spp, precision=32)
    scale = np.abs(g32).max()
    err = np.abs(g - g32)
    assert np.all(err <= 2e-3 * np.abs(g32) + 1e-5 * scale), np.column_stack([g, g32])[err > 2e-3 * np.abs(g32)]


def test_batches_reject_instanced_bvhs(ctx):
    data = scenes.instanced_scene(11, num_instances=3)
    bvh = pike_amd.BVH.for_scene(data)
    with pytest.raises(pike_amd.PikeError):
        pike_amd.BatchScene(ctx, [data], [bvh])
