"""A textured glTF scene, loaded by the JavaScript host (scene.js,
gltf_loader.js, texture_packer.js: the roles of render/scene.js:43-70,
104-181 and render/util/texture_packer.js:87-136) and rendered through the
drop-in JS passes (PikeRenderer.tick over the N-API addon), equals the
oracle's FP32 restatement bit for bit.

The quad's material has RGBA8 textures larger than 1x1 for base colour (with
partial alpha, so some paths pass through), metal-roughness and a non-flat
normal map; the tetrahedron's constant material gets the packer's flat
normal layer (128, 128, 255), which tilts the shading normal by
2 (128/255 - 1/2) along T + B (the reference quirk kept for parity).
"""
import json
import math
import shutil

import numpy as np
import pytest

import pike_amd
from oracle.oracle import OracleScene, probe
from pike_amd import scenes
from tests import pngtools
from tests.test_js_scene import CLI, Asset, node

pytestmark = pytest.mark.skipif(shutil.which("node") is None, reason="node is not installed")


class TexturedAsset(Asset):
    """Asset's quad with metal-roughness and normal textures added to its
    material and partial alpha in its base colour."""

    def __init__(self, seed):
        super().__init__(seed)
        rng = np.random.default_rng(seed + 100)
        self.tex[..., 3] = rng.integers(150, 256, self.tex.shape[:2])
        self.png = pngtools.write(self.tex, 6, 8)
        self.mr = rng.integers(0, 256, (8, 8, 4)).astype(np.uint8)
        self.mr[..., 3] = 255
        self.nrm = np.zeros((4, 8, 4), np.uint8)
        self.nrm[..., 0] = rng.integers(60, 200, (4, 8))
        self.nrm[..., 1] = rng.integers(60, 200, (4, 8))
        self.nrm[..., 2] = rng.integers(170, 256, (4, 8))
        self.nrm[..., 3] = 255

    def gltf(self, buffer_uri=None, image_uri=None, image_view=False):
        doc, blob = super().gltf(buffer_uri, image_uri, image_view)
        doc["images"] += [dict(uri="mr.png"), dict(uri="nrm.png")]
        doc["textures"] += [dict(source=1), dict(source=2)]
        doc["materials"][0]["pbrMetallicRoughness"]["metallicRoughnessTexture"] = dict(index=1)
        doc["materials"][0]["normalTexture"] = dict(index=2)
        return doc, blob

    def write_gltf(self, d):
        super().write_gltf(d)
        (d / "mr.png").write_bytes(pngtools.write(self.mr, 6, 8))
        (d / "nrm.png").write_bytes(pngtools.write(self.nrm, 6, 8))


def _snorm8_words(words):
    b = np.stack([(words >> (8 * k)) & 0xFF for k in range(3)], axis=-1).astype(np.uint8).view(np.int8)
    return np.maximum(b.astype(np.float64) / 127.0, -1.0)


@pytest.fixture(scope="module")
def textured(tmp_path_factory):
    d = tmp_path_factory.mktemp("textured")
    a = TexturedAsset(21)
    a.write_gltf(d)
    env = np.random.default_rng(4).integers(0, 256, (8, 16, 4)).astype(np.uint8)
    env[..., 3] = np.random.default_rng(5).integers(124, 132, (8, 16))
    (d / "env.png").write_bytes(pngtools.write(env, 6, 8))
    quad_m, tet_m = a.matrices()
    q_c, q_n = quad_m[:3, 3], quad_m[:3, :3] @ np.array([0, 0, 1.0])
    q_n /= np.linalg.norm(q_n)
    t_c = (tet_m @ np.array([0.25, 0.25, 0.25, 1.0]))[:3]
    views = {"quad": dict(origin=(q_c + 5.0 * q_n + [0.3, 0.2, 0.0]).tolist(), target=q_c.tolist(), fov=0.9),
             "tet": dict(origin=(t_c + [2.0, 1.5, 2.5]).tolist(), target=t_c.tolist(), fov=0.8)}
    for name, cam in views.items():
        (d / f"{name}.json").write_text(json.dumps(dict(props=[dict(uri="a.gltf")], environment="env.png",
                                                        atlasRes=16, camera=cam)))
    node(CLI, "load-scene", str(d / "quad.json"), str(d / "raw"))
    return d, scenes.SceneData.load(str(d / "raw")), t_c


def test_loaded_atlas_is_textured(textured):
    d, s, _ = textured
    assert s.tex.dtype == np.uint8 and s.tex.shape[1] > 1 and s.tex.shape[2] > 1
    m = s.materials
    textured_mat = [k for k in range(len(m)) if s.tex[m[k, 2]].reshape(-1, 4).std(axis=0)[:2].max() > 0]
    assert len(textured_mat) == 1, "exactly one material carries a non-flat normal map"
    k = textured_mat[0]
    for layer in (m[k, 0], m[k, 1], m[k, 2]):
        assert s.tex[layer].reshape(-1, 4).std(axis=0).max() > 0, "texture layer is not constant"
    assert s.tex[m[k, 0], ..., 3].min() < 255, "base colour carries partial alpha"
    flat = [k2 for k2 in range(len(m)) if k2 != k]
    for k2 in flat:
        assert np.all(s.tex[m[k2, 2]] == np.array([128, 128, 255, 255], np.uint8))


@pytest.mark.gpu
@pytest.mark.parametrize("view", ["quad", "tet"])
def test_loaded_textured_scene_matches_fp32_oracle(textured, tmp_path, view):
    d, s, _ = textured
    W, H, frames, depth = 48, 40, 3, 4
    node(CLI, "render-scene", str(d / f"{view}.json"), str(tmp_path / "img.png"), str(W), str(H), str(frames),
         str(depth), str(tmp_path / "img.f32"))
    img_js = np.fromfile(tmp_path / "img.f32", np.float32).reshape(H, W, 4)
    cam_world = json.loads((d / f"{view}.json").read_text())["camera"]
    cam = scenes.camera(cam_world["origin"], cam_world["target"], cam_world["fov"])
    bvh = pike_amd.BVH(s.positions, s.triangles)
    cam_n = scenes.normalized_camera(cam, bvh.normalizer())
    orc = OracleScene(s)
    orc.build_bvh()
    ref, _, _ = orc.render(cam_n, W, H, depth, 0, frames, precision=32)
    assert img_js[..., :3].std() > 0
    np.testing.assert_array_equal(img_js.view(np.uint32), ref.view(np.uint32))


def test_flat_normal_layer_tilts_the_shading_normal(textured):
    """On the constant-material tetrahedron the shading normal is the
    interpolated normal tilted by 2 (128/255 - 1/2) (T + B), about 0.0055
    rad for an orthonormal frame (FP64 oracle, whose FP32 twin the GPU
    matches bit for bit above)."""
    d, s, t_c = textured
    orc = OracleScene(s)
    orc.build_bvh()
    nrm = orc.serialize()[3]
    o = (np.asarray(t_c) + [2.0, 1.5, 2.5] - nrm[:3]) * nrm[3]
    tgt = (np.asarray(t_c) - nrm[:3]) * nrm[3]
    dirn = (tgt - o) / np.linalg.norm(tgt - o)
    r = probe("surface", [[*o, *dirn]], precision=64, scene=orc)[0]
    assert r[0] > 0, "the ray hits the tetrahedron"
    tri = s.triangles[int(r[1])]
    assert np.all(s.tex[s.materials[tri[3], 2]] == np.array([128, 128, 255, 255], np.uint8))
    b = np.array([1 - r[33] - r[34], r[33], r[34]])

    def interp(word):
        v = b @ _snorm8_words(s.attributes[tri[:3], word].astype(np.int64))
        return v / np.linalg.norm(v)

    T, B, N = interp(0), interp(1), interp(2)
    tilt = 2 * (128 / 255 - 0.5)
    expect = N + tilt * (T + B)
    np.testing.assert_allclose(r[2:5], expect / np.linalg.norm(expect), atol=1e-12)
    ang = math.acos(min(1.0, float(np.dot(r[2:5], N))))
    assert 0.003 < ang < 0.008, ang
