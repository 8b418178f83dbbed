"""Per-pixel parity of the full wavefront frame (raygen -> {closest hit ->
shade -> shadow any-hit} x depth -> accumulate) against the oracle's FP32
restatement under the same seeded PCG streams.  The contract is 1e-4
relative per pixel; with the shared arithmetic contract the values are
expected to be bit-identical, which the tests also report."""
import numpy as np
import pytest

import pike_amd
from pike_amd import scenes
from oracle.oracle import OracleScene

pytestmark = pytest.mark.gpu

REL_TOL = 1e-4


def _setup(ctx, data):
    bvh = pike_amd.BVH(data.positions, data.triangles)
    scene = pike_amd.Scene(ctx, data, bvh)
    orc = OracleScene(data)
    orc.build_bvh()
    cam = scenes.normalized_camera(data.camera, bvh.normalizer())
    return bvh, scene, orc, cam


def _rel_err(a, b):
    return np.abs(a - b) / np.maximum(np.maximum(np.abs(a), np.abs(b)), 1e-6)


def _compare(img_g, img_o, what):
    err = _rel_err(img_g[..., :3].astype(np.float64), img_o[..., :3].astype(np.float64))
    worst = np.unravel_index(np.argmax(err), err.shape)
    assert err.max() <= REL_TOL, f"{what}: max rel err {err.max():.3e} at {worst}: " \
                                 f"gpu {img_g[worst[:2]]} oracle {img_o[worst[:2]]}"
    return int(np.sum(img_g[..., :3].view(np.uint32) != img_o[..., :3].view(np.uint32)))


@pytest.mark.parametrize("use_graph", [False, True])
def test_render_config1_matches_fp32_oracle(ctx, use_graph):
    data = scenes.config1(seed=1)
    bvh, scene, orc, cam = _setup(ctx, data)
    W, H, depth, frames = 64, 48, 4, 3
    fr = pike_amd.Frame(ctx, W, H)
    fr.set_camera(cam)
    fr.render(scene, max_depth=depth, first_sample=0, num_frames=frames, use_graph=use_graph)
    img = fr.read_image()
    col = fr.read_color()
    acc_o, col_o, cnt_o = orc.render(cam, W, H, depth, 0, frames, precision=32)
    _compare(col, col_o, "last frame color")
    _compare(img, acc_o, "running mean")
    # contribution counts (color.a) are integers and must agree exactly
    np.testing.assert_array_equal(col[..., 3], col_o[..., 3])
    c = fr.counters()
    assert c["stack_overflow"] == 0
    assert c["frames"] == frames


def test_per_pass_api_equals_render(ctx):
    """CameraPass/TracePass/ShadeHitPass/TraceShadowPass/AccumulatePass entry
    points called one by one give the same image as pk_render."""
    data = scenes.config1(seed=2, num_spheres=16)
    bvh, scene, orc, cam = _setup(ctx, data)
    W, H, depth = 40, 30, 3
    a = pike_amd.Frame(ctx, W, H)
    a.set_camera(cam)
    a.render(scene, max_depth=depth, first_sample=5, num_frames=1, use_graph=False)
    b = pike_amd.Frame(ctx, W, H)
    b.set_camera(cam)
    b.begin(5)
    b.raygen()
    for k in range(depth):
        b.trace_closest(scene, k)
        b.shade(scene, k)
        b.trace_shadow(scene)
    b.accumulate()
    np.testing.assert_array_equal(a.read_color(), b.read_color())
    acc_o, col_o, _ = orc.render(cam, W, H, depth, 5, 1, precision=32)
    _compare(b.read_color(), col_o, "per-pass color")


def test_render_float_env_bins_sampler(ctx):
    """Float HDR environment with the reference's equal-radiance bins."""
    data = scenes.config1(seed=6, num_spheres=10)
    data.env = scenes.lognormal_sun_env(128, 64, seed=3, sigma=1.0, sun_deg=8.0, sun_scale=20.0)
    bvh, scene, orc, cam = _setup(ctx, data)
    W, H = 48, 32
    fr = pike_amd.Frame(ctx, W, H)
    fr.set_camera(cam)
    fr.render(scene, max_depth=5, first_sample=0, num_frames=2, env_theta=0.25)
    acc_o, col_o, _ = orc.render(cam, W, H, 5, 0, 2, env_theta=0.25, precision=32)
    _compare(fr.read_color(), col_o, "float env color")


def test_empty_scene_region_is_env(ctx):
    """A camera looking away from all geometry sees exactly env radiance."""
    data = scenes.config1(seed=1, num_spheres=4)
    bvh, scene, orc, cam = _setup(ctx, data)
    d = np.array([0.0, 0.9, 0.4]) / np.linalg.norm([0.0, 0.9, 0.4])
    cam.dir[0], cam.dir[1], cam.dir[2] = d
    cam.origin[1] = 3.0
    fr = pike_amd.Frame(ctx, 16, 16)
    fr.set_camera(cam)
    fr.render(scene, max_depth=2, num_frames=1)
    img = fr.read_image()
    np.testing.assert_array_equal(img[..., :3], np.ones_like(img[..., :3]))


@pytest.mark.parametrize("samples,frames", [(4, 8), (3, 7), (16, 5), (40, 45), (64, 70)])
def test_samples_per_pass_equals_one_sample_frames(ctx, samples, frames):
    """A pass of S samples per pixel (one wavefront of W*H*S paths) folds its
    samples in order: the image equals one-sample frames bit for bit, also
    when the last pass is partial, and equals the FP32 oracle.  S above 32
    makes k_accumulate fold more than one LDS chunk of 32 samples per pixel,
    with a partial last chunk."""
    env = scenes.lognormal_sun_env(64, 32, seed=5, sigma=1.0, sun_deg=6.0, sun_scale=30.0)
    data = scenes.mixed_scene(7, 4000, env=env)
    data.camera = scenes.camera((0.5, 0.5, 1.4), (0.5, 0.05, 0.45), 1.0)
    bvh, scene, orc, cam = _setup(ctx, data)
    W, H, depth = 40, 24, 4
    ref = pike_amd.Frame(ctx, W, H)
    ref.set_camera(cam)
    ref.render(scene, max_depth=depth, first_sample=2, num_frames=frames)
    fr = pike_amd.Frame(ctx, W, H)
    fr.set_camera(cam)
    fr.set_samples_per_pass(samples)
    for use_graph in (True, False):
        fr.reset()
        fr.render(scene, max_depth=depth, first_sample=2, num_frames=frames, use_graph=use_graph)
        np.testing.assert_array_equal(fr.read_image().view(np.uint32), ref.read_image().view(np.uint32))
        c = fr.counters()
        assert c["frames"] == frames and c["stack_overflow"] == 0
    acc_o, _, _ = orc.render(cam, W, H, depth, 2, frames, precision=32, accum=np.zeros((H, W, 4), np.float32))
    # the run started at sample 2 on a cleared accumulator, as the frames did
    np.testing.assert_array_equal(fr.read_image().view(np.uint32), acc_o.view(np.uint32))


@pytest.mark.parametrize("W,H,samples", [(1, 1, 1), (1, 1, 5), (37, 23, 1), (37, 23, 3), (129, 3, 2)])
def test_ragged_frames_match_fp32_oracle(ctx, W, H, samples):
    """Frames whose pixel and path counts fill no launch block (one pixel; 37x23
    and 129x3 against 256-thread raygen/accumulate blocks and 64-lane waves),
    with one- and multi-sample passes: every pixel, the last included (the
    reference's color buffer drops it, A.5; its accumulate bounds check is
    off by one, A.6), equals the FP32 oracle bit for bit."""
    env = scenes.lognormal_sun_env(64, 32, seed=3, sigma=1.0, sun_deg=6.0, sun_scale=30.0)
    data = scenes.mixed_scene(5, 3000, env=env)
    data.camera = scenes.camera((0.5, 0.5, 1.4), (0.5, 0.05, 0.45), 1.0)
    bvh, scene, orc, cam = _setup(ctx, data)
    depth, frames = 4, 5
    fr = pike_amd.Frame(ctx, W, H)
    fr.set_camera(cam)
    fr.set_samples_per_pass(samples)
    fr.render(scene, max_depth=depth, first_sample=0, num_frames=frames)
    c = fr.counters()
    assert c["frames"] == frames and c["stack_overflow"] == 0
    img = fr.read_image()
    assert img.shape == (H, W, 4)
    acc_o, _, _ = orc.render(cam, W, H, depth, 0, frames, precision=32)
    np.testing.assert_array_equal(img.view(np.uint32), acc_o.view(np.uint32))
    assert np.all(np.isfinite(img)) and img[..., :3].max() > 0


def _sun_scene(sampler):
    env = scenes.lognormal_sun_env(128, 64, seed=8, sigma=1.5, sun_deg=4.0, sun_scale=1e3)
    data = scenes.mixed_scene(11, 6000, env=env)
    data.env_sampler = sampler
    data.camera = scenes.camera((0.5, 0.45, 1.3), (0.5, 0.05, 0.45), 1.0)
    return data


def test_render_cdf_sampler_matches_fp32_oracle(ctx):
    """The marginal/conditional CDF sampler (device binary search) against the
    oracle's restatement: bit-identical frames."""
    data = _sun_scene(pike_amd._ffi.PK_ENV_SAMPLER_CDF)
    bvh, scene, orc, cam = _setup(ctx, data)
    W, H, depth, frames = 48, 32, 5, 3
    fr = pike_amd.Frame(ctx, W, H)
    fr.set_camera(cam)
    fr.render(scene, max_depth=depth, first_sample=0, num_frames=frames)
    acc_o, _, _ = orc.render(cam, W, H, depth, 0, frames, precision=32)
    np.testing.assert_array_equal(fr.read_image().view(np.uint32), acc_o.view(np.uint32))


@pytest.mark.timeout(600)
@pytest.mark.parametrize("sampler", [pike_amd._ffi.PK_ENV_SAMPLER_BINS, pike_amd._ffi.PK_ENV_SAMPLER_CDF])
def test_converged_image_matches_fp64_oracle(ctx, sampler):
    """4096 samples per pixel on the GPU against the FP64 oracle of the same
    estimator (common random numbers): mean relative L1 < 1e-3."""
    data = _sun_scene(sampler)
    bvh, scene, orc, cam = _setup(ctx, data)
    W, H, depth, spp = 24, 16, 4, 4096
    fr = pike_amd.Frame(ctx, W, H)
    fr.set_camera(cam)
    fr.set_samples_per_pass(256)
    fr.render(scene, max_depth=depth, first_sample=0, num_frames=spp)
    gpu = fr.read_image()[..., :3].astype(np.float64)
    print("fp64 oracle render", flush=True)
    ref, _, _ = orc.render(cam, W, H, depth, 0, spp, precision=64)
    ref = ref[..., :3].astype(np.float64)
    rel = np.abs(gpu - ref).sum() / np.abs(ref).sum()
    print("converged image mean relative L1 vs FP64:", rel)
    assert rel < 1e-3, rel


def test_quarter_billion_path_pass_equals_small_passes(ctx):
    """A pass of 2^28 paths (512x512 at 1024 samples: 47 GB of queues) takes
    the closest-hit kernel's large-launch claim size; its image equals that of
    16-sample passes bit for bit."""
    env = scenes.lognormal_sun_env(64, 32, seed=4, sigma=1.0, sun_deg=6.0, sun_scale=30.0)
    data = scenes.mixed_scene(9, 3000, env=env)
    data.camera = scenes.camera((0.5, 0.5, 1.4), (0.5, 0.05, 0.45), 1.0)
    bvh, scene, _, cam = _setup(ctx, data)
    W = H = 512
    imgs = []
    for S in (1024, 16):
        fr = pike_amd.Frame(ctx, W, H)
        fr.set_camera(cam)
        fr.set_samples_per_pass(S)
        fr.render(scene, max_depth=2, first_sample=0, num_frames=1024)
        c = fr.counters()
        assert c["frames"] == 1024 and c["stack_overflow"] == 0
        imgs.append(fr.read_image())
        fr.close()
    assert W * H * 1024 >= 1 << 28
    np.testing.assert_array_equal(imgs[0].view(np.uint32), imgs[1].view(np.uint32))
