"""The JavaScript host (pike_amd/js over the N-API addon) drives the same C
ABI as the Python harness: its BVH buffers, environment tables and rendered
images must equal the Python path's bit for bit."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import pike_amd
from pike_amd import scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "pike_amd", "js", "cli.js")
NODE = shutil.which("node")

pytestmark = pytest.mark.skipif(NODE is None, reason="node is not installed")


def node(*args, timeout=300):
    r = subprocess.run([NODE, *args], capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, r.stderr
    return r


def test_addon_loads_and_fails_loudly_without_device():
    code = ("const p=require(%r); const n=p.native.load(); console.log(n.version());"
            "try { n.ctxCreate(0); console.log('device'); } catch (e) { console.log('error', e.code); }"
            % os.path.join(ROOT, "pike_amd", "js", "index.js"))
    out = node("-e", code).stdout.split()
    assert out[0] == "pike_amd"
    if not os.path.exists("/dev/kfd"):
        assert out[-2:] == ["error", "-2"]


def test_js_bvh_export_equals_python(tmp_path):
    data = scenes.config1(seed=3)
    data.save(tmp_path / "scene")
    node(CLI, "bvh", str(tmp_path / "scene"), str(tmp_path / "out"))
    b = pike_amd.BVH(data.positions, data.triangles)
    nodes, tris, verts, norm = b.export_reference()
    np.testing.assert_array_equal(np.fromfile(tmp_path / "out" / "nodes.i32", np.int32), nodes.ravel())
    np.testing.assert_array_equal(np.fromfile(tmp_path / "out" / "tris.i32", np.int32), tris.ravel())
    np.testing.assert_array_equal(np.fromfile(tmp_path / "out" / "verts.f32", np.uint32), verts.view(np.uint32).ravel())
    np.testing.assert_array_equal(np.fromfile(tmp_path / "out" / "normalizer.f64", np.float64), norm)


def test_js_env_tables_equal_python(tmp_path):
    data = scenes.config1(seed=1, num_spheres=2)
    data.env = scenes.lognormal_sun_env(96, 48, seed=2)
    data.save(tmp_path / "scene")
    node(CLI, "env", str(tmp_path / "scene"), str(tmp_path / "out"))
    from tests.test_env import product_tables
    bins, coords, pdf = product_tables(data)
    np.testing.assert_array_equal(np.fromfile(tmp_path / "out" / "bins.i32", np.int32), bins.ravel())
    np.testing.assert_array_equal(np.fromfile(tmp_path / "out" / "coords.i32", np.int32), coords.ravel())
    np.testing.assert_array_equal(np.fromfile(tmp_path / "out" / "pdf.f32", np.uint32), pdf.view(np.uint32).ravel())


@pytest.mark.gpu
def test_js_render_equals_python_render(tmp_path, ctx):
    """PikeRenderer.tick() (pike.js:120-162 pass order, recorded and submitted
    through the JS command encoder) renders the same image as pk_render."""
    data = scenes.config1(seed=2, num_spheres=20)
    data.save(tmp_path / "scene")
    W, H, frames, depth = 48, 32, 3, 4
    out = tmp_path / "img.f32"
    node(CLI, "render", str(tmp_path / "scene"), str(out), str(W), str(H), str(frames), str(depth))
    img_js = np.fromfile(out, np.float32).reshape(H, W, 4)
    bvh = pike_amd.BVH(data.positions, data.triangles)
    scene = pike_amd.Scene(ctx, data, bvh)
    fr = pike_amd.Frame(ctx, W, H)
    fr.set_camera(scenes.normalized_camera(data.camera, bvh.normalizer()))
    fr.render(scene, max_depth=depth, first_sample=0, num_frames=frames, use_graph=False)
    np.testing.assert_array_equal(img_js.view(np.uint32), fr.read_image().view(np.uint32))


@pytest.mark.gpu
def test_js_adjoint_equals_python_adjoint(tmp_path, ctx):
    """pk_adjoint through the N-API addon (node cli.js adjoint) gives the
    Python harness's gradients (up to the atomics' summation order)."""
    data = scenes.config5(cells=24, num_spheres=6, tex_res=32, env_size=(32, 16))
    data.save(tmp_path / "scene")
    W, H, spp, depth = 32, 24, 2, 3
    d = np.random.default_rng(2).uniform(-1, 1, (H, W, 4)).astype(np.float32)
    d.tofile(tmp_path / "d.f32")
    node(CLI, "adjoint", str(tmp_path / "scene"), str(tmp_path / "g"), str(tmp_path / "d.f32"), str(W), str(H),
         str(spp), str(depth))
    ga_js = np.fromfile(tmp_path / "g" / "atlas.f32", np.float32)
    ge_js = np.fromfile(tmp_path / "g" / "env.f32", np.float32)
    bvh = pike_amd.BVH(data.positions, data.triangles)
    scene = pike_amd.Scene(ctx, data, bvh)
    fr = pike_amd.Frame(ctx, W, H)
    fr.set_camera(scenes.normalized_camera(data.camera, bvh.normalizer()))
    ga, ge = fr.adjoint(scene, d, max_depth=depth, num_frames=spp)
    scale = max(np.abs(ga).max(), np.abs(ge).max())
    assert scale > 0
    np.testing.assert_allclose(ga_js, ga.ravel(), rtol=1e-4, atol=1e-6 * scale)
    np.testing.assert_allclose(ge_js, ge.ravel(), rtol=1e-4, atol=1e-6 * scale)
