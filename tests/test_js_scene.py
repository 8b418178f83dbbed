"""Scene-JSON + glTF loading in the JavaScript host (pike_amd/js/scene.js,
gltf_loader.js, texture_packer.js, png.js): the role of render/scene.js,
render/util/gltf_loader.js and texture_packer.js, feeding pk_scene_upload,
plus tonemapped PNG output.

Checked against what the formats and transforms fix, with assets built here:
* PNG decoding of independently written files: every color type, bit depths
  1-16, all five row filters, Adam7, palettes with tRNS;
* PNG encoding read back by an independent reader;
* baked positions equal the node hierarchy (TRS and matrix nodes) and the
  prop transforms applied in float64; normals (given and computed) and
  tangents within snorm8 quantization of the expected directions; uv words
  equal the reference's truncating snorm16 quantizer; triangle offsets,
  material dedupe and atlas layer order; constant-color layers exact;
  texture layers within 1/255 of a bilinear (sRGB-decoded) resample; the
  environment equals the PNG's bytes;
* interleaved accessors with byteStride, normalized texcoords, non-indexed
  primitives; .gltf (external + data: URIs) and .glb give identical scenes;
* on the GPU: a loaded scene renders through PikeRenderer exactly as the
  Python harness renders the same baked scene.
"""
import base64
import json
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

import pike_amd
from pike_amd import scenes
from tests import pngtools

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "pike_amd", "js", "cli.js")
INDEX = os.path.join(ROOT, "pike_amd", "js", "index.js")
NODE = shutil.which("node")

pytestmark = pytest.mark.skipif(NODE is None, reason="node is not installed")


def node(*args, timeout=300):
    r = subprocess.run([NODE, *args], capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, r.stderr
    return r


# ---------------------------------------------------------------------------
# PNG
# ---------------------------------------------------------------------------
def _expected_rgba(samples, color, depth, palette=None, trns=None):
    h, w = samples.shape[:2]
    s = samples.reshape(h, w, -1).astype(np.int64)
    scale = (lambda v: v >> 8) if depth == 16 else (lambda v: np.floor(v * 255 / ((1 << depth) - 1) + 0.5))
    out = np.zeros((h, w, 4), np.uint8)
    if color == 3:
        pal = np.asarray(palette, np.uint8).reshape(-1, 3)
        out[..., :3] = pal[s[..., 0]]
        alpha = np.full(len(pal), 255, np.uint8)
        if trns is not None:
            alpha[:len(trns)] = trns
        out[..., 3] = alpha[s[..., 0]]
        return out
    if color in (0, 4):
        g = scale(s[..., 0]) if depth != 8 else s[..., 0]
        out[..., 0] = out[..., 1] = out[..., 2] = g
        out[..., 3] = (scale(s[..., 1]) if depth != 8 else s[..., 1]) if color == 4 else 255
        return out
    for c in range(3):
        out[..., c] = scale(s[..., c]) if depth != 8 else s[..., c]
    out[..., 3] = (scale(s[..., 3]) if depth != 8 else s[..., 3]) if color == 6 else 255
    return out


PNG_CASES = [(2, 8, False), (6, 8, False), (6, 8, True), (0, 8, False), (0, 1, False), (0, 2, True),
             (0, 4, False), (4, 8, True), (3, 8, False), (3, 2, True), (3, 4, False), (2, 16, False),
             (6, 16, True), (0, 16, False), (4, 16, False)]


def test_png_decoder_on_independent_files(tmp_path):
    rng = np.random.default_rng(1)
    files, expected = [], []
    for k, (color, depth, interlace) in enumerate(PNG_CASES):
        h, w = 13, 11
        ch = pngtools.CHANNELS[color]
        palette = trns = None
        if color == 3:
            n = 1 << min(depth, 4)
            palette = rng.integers(0, 256, (n, 3))
            trns = list(rng.integers(0, 256, n // 2))
            samples = rng.integers(0, n, (h, w, 1))
        else:
            samples = rng.integers(0, 1 << depth, (h, w, ch))
        f = tmp_path / f"c{k}.png"
        f.write_bytes(pngtools.write(samples, color, depth, palette, trns, interlace, start_filter=k))
        files.append(str(f))
        expected.append(_expected_rgba(samples, color, depth, palette, trns))
    code = ("const {png}=require(%r); const fs=require('fs');"
            "for (const f of %s) { const d=png.decode(fs.readFileSync(f));"
            "fs.writeFileSync(f+'.rgba', Buffer.from(d.data.buffer)); }" % (INDEX, json.dumps(files)))
    node("-e", code)
    for f, e, case in zip(files, expected, PNG_CASES):
        got = np.fromfile(f + ".rgba", np.uint8).reshape(e.shape)
        np.testing.assert_array_equal(got, e, err_msg=str(case))


def test_png_encoder_round_trip(tmp_path):
    img = np.random.default_rng(2).integers(0, 256, (7, 9, 4)).astype(np.uint8)
    img.tofile(tmp_path / "in.rgba")
    code = ("const {png}=require(%r); const fs=require('fs'); const b=fs.readFileSync(%r);"
            "fs.writeFileSync(%r, png.encode(9, 7, new Uint8Array(b.buffer, b.byteOffset, b.length)));"
            % (INDEX, str(tmp_path / "in.rgba"), str(tmp_path / "out.png")))
    node("-e", code)
    np.testing.assert_array_equal(pngtools.read_rgba8((tmp_path / "out.png").read_bytes()), img)


# ---------------------------------------------------------------------------
# glTF assets
# ---------------------------------------------------------------------------
def _quat(axis, angle):
    a = np.asarray(axis, np.float64) / np.linalg.norm(axis)
    return np.concatenate([a * np.sin(angle / 2), [np.cos(angle / 2)]])


def _trs(t, q, s):
    x, y, z, w = q
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                  [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                  [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
    M = np.eye(4)
    M[:3, :3] = R * np.asarray(s)[None, :]
    M[:3, 3] = t
    return M


class Asset:
    """A glTF 2.0 document with two meshes under a node hierarchy:
    * a quad (indexed UINT16, interleaved POSITION+NORMAL with byteStride,
      float TEXCOORD_0, TANGENT) with a base-color texture and default
      metal/roughness factors, under a TRS parent and a TRS child;
    * a tetrahedron (non-indexed, no NORMAL, normalized UNSIGNED_SHORT uv)
      with constant factors, under a matrix node."""

    def __init__(self, seed):
        rng = np.random.default_rng(seed)
        self.quad_pos = np.array([[-1, -1, 0], [1, -1, 0], [1, 1, 0], [-1, 1, 0]], np.float32)
        self.quad_nrm = np.tile(np.array([0, 0, 1], np.float32), (4, 1))
        self.quad_uv = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float32) * 0.9 + 0.05
        self.quad_tan = np.tile(np.array([1, 0, 0, -1], np.float32), (4, 1))
        self.quad_idx = np.array([0, 1, 2, 0, 2, 3], np.uint16)
        self.tet_pos = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 0], [0, 1, 0], [0, 0, 1],
                                 [0, 0, 0], [0, 0, 1], [1, 0, 0], [1, 0, 0], [0, 0, 1], [0, 1, 0]], np.float32)
        self.tet_uv = rng.integers(0, 65536, (12, 2)).astype(np.uint16)
        self.tex = rng.integers(0, 256, (6, 8, 4)).astype(np.uint8)
        self.png = pngtools.write(self.tex, 6, 8)
        self.parent = dict(translation=[1.0, 2.0, 3.0], rotation=_quat([1, 2, 3], 0.7).tolist(), scale=[2.0, 2.0, 2.0])
        self.child = dict(translation=[0.5, 0.0, 0.0], rotation=_quat([0, 0, 1], 0.3).tolist(), scale=[1.0, 1.0, 1.0])
        M = _trs([-1.0, 0.5, 0.25], _quat([0, 1, 0], 1.1), [1.5, 1.5, 1.5])
        self.tet_matrix = M.T.reshape(-1).tolist()  # column-major
        self.base_color = [0.2, 0.4, 0.6, 1.0]
        self.rough, self.metal = 0.3, 0.0

    def matrices(self):
        P = _trs(self.parent["translation"], self.parent["rotation"], self.parent["scale"])
        C = _trs(self.child["translation"], self.child["rotation"], self.child["scale"])
        T = np.array(self.tet_matrix).reshape(4, 4).T
        return P @ C, T

    def _buffer(self):
        inter = np.concatenate([self.quad_pos, self.quad_nrm], axis=1).astype(np.float32)  # stride 24
        parts = [inter.tobytes(), self.quad_uv.tobytes(), self.quad_tan.tobytes(), self.quad_idx.tobytes() + b"\0\0",
                 self.tet_pos.tobytes(), self.tet_uv.tobytes()]
        offsets, blob = [], b""
        for p in parts:
            offsets.append(len(blob))
            blob += p
            blob += b"\0" * (-len(blob) % 4)
        return blob, offsets, [len(p) for p in parts]

    def gltf(self, buffer_uri=None, image_uri=None, image_view=False):
        blob, off, ln = self._buffer()
        views = [dict(buffer=0, byteOffset=off[0], byteLength=ln[0], byteStride=24),
                 dict(buffer=0, byteOffset=off[1], byteLength=ln[1]),
                 dict(buffer=0, byteOffset=off[2], byteLength=ln[2]),
                 dict(buffer=0, byteOffset=off[3], byteLength=12),
                 dict(buffer=0, byteOffset=off[4], byteLength=ln[4]),
                 dict(buffer=0, byteOffset=off[5], byteLength=ln[5])]
        if image_view:
            views.append(dict(buffer=0, byteOffset=len(blob), byteLength=len(self.png)))
            blob = blob + self.png + b"\0" * (-len(self.png) % 4)
        acc = [dict(bufferView=0, byteOffset=0, componentType=5126, count=4, type="VEC3"),
               dict(bufferView=0, byteOffset=12, componentType=5126, count=4, type="VEC3"),
               dict(bufferView=1, componentType=5126, count=4, type="VEC2"),
               dict(bufferView=2, componentType=5126, count=4, type="VEC4"),
               dict(bufferView=3, componentType=5123, count=6, type="SCALAR"),
               dict(bufferView=4, componentType=5126, count=12, type="VEC3"),
               dict(bufferView=5, componentType=5123, count=12, type="VEC2", normalized=True)]
        image = dict(bufferView=6, mimeType="image/png") if image_view else dict(uri=image_uri)
        doc = dict(
            asset=dict(version="2.0"), scene=0, scenes=[dict(nodes=[0, 2])],
            nodes=[dict(children=[1], **self.parent), dict(mesh=0, **self.child), dict(mesh=1, matrix=self.tet_matrix)],
            meshes=[dict(name="quad", primitives=[dict(attributes=dict(POSITION=0, NORMAL=1, TEXCOORD_0=2, TANGENT=3),
                                                       indices=4, material=0)]),
                    dict(name="tet", primitives=[dict(attributes=dict(POSITION=5, TEXCOORD_0=6), material=1)])],
            materials=[dict(name="textured", pbrMetallicRoughness=dict(baseColorTexture=dict(index=0))),
                       dict(name="flat", pbrMetallicRoughness=dict(baseColorFactor=self.base_color,
                                                                   roughnessFactor=self.rough,
                                                                   metallicFactor=self.metal))],
            textures=[dict(source=0)], images=[image], accessors=acc, bufferViews=views,
            buffers=[dict(byteLength=len(blob)) if buffer_uri is None else dict(byteLength=len(blob), uri=buffer_uri)])
        return doc, blob

    def write_gltf(self, d):
        os.makedirs(d, exist_ok=True)
        doc, blob = self.gltf(buffer_uri="data:application/octet-stream;base64," + base64.b64encode(b"").decode(),
                              image_uri="tex.png")
        doc["buffers"][0]["uri"] = "a.bin"
        open(os.path.join(d, "a.bin"), "wb").write(blob)
        open(os.path.join(d, "tex.png"), "wb").write(self.png)
        json.dump(doc, open(os.path.join(d, "a.gltf"), "w"))

    def write_glb(self, d):
        doc, blob = self.gltf(image_view=True)
        js = json.dumps(doc).encode()
        js += b" " * (-len(js) % 4)
        body = struct.pack("<II", len(js), 0x4E4F534A) + js + struct.pack("<II", len(blob), 0x004E4942) + blob
        open(os.path.join(d, "b.glb"), "wb").write(struct.pack("<III", 0x46546C67, 2, 12 + len(body)) + body)


def _rotate_arbitrary(v, axis, angle):
    x, y, z = axis
    s, c = np.sin(angle), np.cos(angle)
    oc = 1 - c
    L = np.array([[oc * x * x + c, oc * x * y - z * s, oc * z * x + y * s, 0],
                  [oc * x * y + z * s, oc * y * y + c, oc * y * z - x * s, 0],
                  [oc * z * x - y * s, oc * y * z + x * s, oc * z * z + c, 0], [0, 0, 0, 1]])
    return (L.T @ np.append(v, 1.0))[:3]   # listed row by row, applied column-major


def _decode_snorm8(words):
    b = np.stack([(words >> (8 * k)) & 0xFF for k in range(3)], axis=-1).astype(np.uint8).view(np.int8)
    return np.maximum(b.astype(np.float64) / 127.0, -1.0)


def _uv_word(uv):
    q = [int(np.trunc(u * (32767 if u > 0 else 32768))) & 0xFFFF for u in uv]
    return q[0] | (q[1] << 16)


@pytest.fixture(scope="module")
def assets(tmp_path_factory):
    d = tmp_path_factory.mktemp("gltf")
    a = Asset(5)
    a.write_gltf(str(d))
    a.write_glb(str(d))
    env = np.random.default_rng(8).integers(0, 256, (8, 16, 4)).astype(np.uint8)
    env[..., 3] = np.random.default_rng(9).integers(120, 136, (8, 16))
    (d / "env.png").write_bytes(pngtools.write(env, 6, 8))
    xf = dict(rotate=[dict(axis=[0, 1, 0], angle=0.5), dict(axis=[1, 0, 0], angle=-0.2)], scale=0.5,
              translate=[0.0, 0.1, -0.3])
    scene = dict(props=[dict(uri="a.gltf", transforms=xf), dict(uri="b.glb", skipMeshes=[])], environment="env.png",
                 atlasRes=16, camera=dict(origin=[0.5, 1.5, 6.0], target=[0.5, 1.0, 0.0], fov=0.9))
    (d / "scene.json").write_text(json.dumps(scene))
    for name, props in (("only_gltf", [dict(uri="a.gltf")]), ("only_glb", [dict(uri="b.glb")])):
        (d / f"{name}.json").write_text(json.dumps(dict(props=props, atlasRes=16)))
    return d, a, env, xf


def _load(d, name, out):
    node(CLI, "load-scene", str(d / name), str(out))
    return scenes.SceneData.load(str(out))


def test_loader_bakes_the_node_hierarchy_and_prop_transforms(assets, tmp_path):
    d, a, env, xf = assets
    s = _load(d, "scene.json", tmp_path / "raw")
    quad_M, tet_M = a.matrices()

    def xform(p, rotation_only=False):
        for r in xf["rotate"]:
            p = _rotate_arbitrary(p, r["axis"], r["angle"])
        return p if rotation_only else p * xf["scale"] + np.array(xf["translate"])

    exp = []
    for M, P, with_xf in ((quad_M, a.quad_pos, True), (tet_M, a.tet_pos, True), (quad_M, a.quad_pos, False),
                          (tet_M, a.tet_pos, False)):
        for p in P.astype(np.float64):
            w = (M @ np.append(p, 1.0))[:3]
            exp.append(xform(w) if with_xf else w)
    np.testing.assert_allclose(s.positions, np.array(exp), rtol=0, atol=1e-12)
    # triangles: quad (indexed), tet (in order), per prop with vertex offsets
    q = a.quad_idx.reshape(-1, 3).astype(np.int64)
    t = np.arange(12).reshape(-1, 3)
    tri = np.concatenate([q, t + 4, q + 16, t + 20])
    np.testing.assert_array_equal(s.triangles[:, :3], tri)
    np.testing.assert_array_equal(s.triangles[:, 3], [0, 0, 1, 1, 1, 1, 2, 2, 3, 3, 3, 3])
    # normals: given (quad) and computed face normals (tet), world space
    nq = np.linalg.inv(quad_M[:3, :3]).T @ np.array([0, 0, 1.0])
    nrm = []
    for with_xf in (True, False):
        for k in range(4):
            nrm.append(nq)
        for f in range(4):
            local = a.tet_pos[3 * f:3 * f + 3].astype(np.float64)
            fn = np.cross(local[1] - local[0], local[2] - local[0])
            n = np.linalg.inv(tet_M[:3, :3]).T @ fn
            nrm += [n] * 3
        if with_xf:
            nrm[-16:] = [xform(n, True) for n in nrm[-16:]]
    nrm = np.array([n / np.linalg.norm(n) for n in nrm])
    got = _decode_snorm8(s.attributes[:, 2])
    assert np.abs(got - nrm).max() <= 1.01 / 127
    # tangents: quad tangent x (w = -1), transformed as a direction
    tq = quad_M[:3, :3] @ np.array([1.0, 0, 0])
    tq /= np.linalg.norm(tq)
    t_xf = xform(tq, True)
    for i, tt in ((0, t_xf), (16, tq)):
        np.testing.assert_allclose(_decode_snorm8(s.attributes[i:i + 4, 0]), np.tile(tt, (4, 1)), atol=1.01 / 127)
        nn = nrm[i]
        b = -np.cross(nn, tt)
        np.testing.assert_allclose(_decode_snorm8(s.attributes[i:i + 4, 1]), np.tile(b / np.linalg.norm(b), (4, 1)),
                                   atol=2.02 / 127)
    # uv words: float uvs, and normalized uint16 uvs scaled to [0, 1]
    uvs = list(a.quad_uv.astype(np.float64)) + list(np.float64(a.tet_uv) / 65535.0)
    words = [_uv_word(u) for u in uvs] * 2
    np.testing.assert_array_equal(s.attributes[:, 3], np.array(words, np.uint32))


def test_loader_materials_atlas_and_environment(assets, tmp_path):
    d, a, env, xf = assets
    s = _load(d, "scene.json", tmp_path / "raw")
    # layers in _createSceneMaterial order, constant colors deduplicated by
    # value, textures by (file, image)
    np.testing.assert_array_equal(s.materials, [[0, 1, 2, 3], [4, 5, 2, 3], [6, 1, 2, 3], [4, 5, 2, 3]])
    assert s.tex.shape == (7, 6, 6, 4)   # atlasRes 16 > largest texture height 6

    def color(c):
        return np.floor(np.array(list(c[:3]) + [1.0]) * 255 + 0.5).astype(np.uint8)

    for layer, c in ((1, [0, 1, 1]), (2, [0.5, 0.5, 1]), (3, [0, 0, 0]), (4, a.base_color),
                     (5, [0, a.rough, a.metal])):
        assert np.all(s.tex[layer] == color(c)), layer
    # texture layers: bilinear resample of the sRGB-decoded image at texel centers
    img = a.tex.astype(np.float64) / 255
    lin = np.where(img <= 0.04045, img / 12.92, ((img + 0.055) / 1.055) ** 2.4)
    lin[..., 3] = img[..., 3]
    H, W = img.shape[:2]
    res = 6
    exp = np.zeros((res, res, 4))
    for y in range(res):
        fy = (y + 0.5) / res * H - 0.5
        y0 = int(np.floor(fy))
        ty = fy - y0
        for x in range(res):
            fx = (x + 0.5) / res * W - 0.5
            x0 = int(np.floor(fx))
            tx = fx - x0

            def tex(xx, yy):
                return lin[min(max(yy, 0), H - 1), xx % W]

            top = tex(x0, y0) * (1 - tx) + tex(x0 + 1, y0) * tx
            bot = tex(x0, y0 + 1) * (1 - tx) + tex(x0 + 1, y0 + 1) * tx
            exp[y, x] = top * (1 - ty) + bot * ty
    for layer in (0, 6):
        assert np.abs(s.tex[layer].astype(np.float64) - exp * 255).max() <= 0.5 + 1e-6, layer
    np.testing.assert_array_equal(s.env, env)


def test_gltf_and_glb_containers_agree(assets, tmp_path):
    d, a, env, xf = assets
    g = _load(d, "only_gltf.json", tmp_path / "g")
    b = _load(d, "only_glb.json", tmp_path / "b")
    for k in ("positions", "attributes", "triangles", "materials", "tex"):
        np.testing.assert_array_equal(getattr(g, k), getattr(b, k), err_msg=k)


@pytest.mark.gpu
def test_loaded_scene_renders_like_the_python_harness(assets, tmp_path, ctx):
    d, a, env, xf = assets
    s = _load(d, "scene.json", tmp_path / "raw")
    W, H, frames, depth = 40, 30, 2, 3
    node(CLI, "render-scene", str(d / "scene.json"), str(tmp_path / "img.png"), str(W), str(H), str(frames),
         str(depth), str(tmp_path / "img.f32"))
    img_js = np.fromfile(tmp_path / "img.f32", np.float32).reshape(H, W, 4)
    bvh = pike_amd.BVH(s.positions, s.triangles)
    scene = pike_amd.Scene(ctx, s, bvh)
    fr = pike_amd.Frame(ctx, W, H)
    fr.set_camera(scenes.normalized_camera(s.camera, bvh.normalizer()))
    fr.render(scene, max_depth=depth, num_frames=frames, use_graph=False)
    np.testing.assert_array_equal(img_js.view(np.uint32), fr.read_image().view(np.uint32))
    out = pngtools.read_rgba8((tmp_path / "img.png").read_bytes())
    assert out.shape == (H, W, 4) and out[..., :3].std() > 0


def test_instanced_loading_places_the_baked_geometry(assets, tmp_path):
    """--instanced: props sharing a glTF file become placements of one mesh
    baked in the file's own space; applying each placement's 3x4 transform
    (the prop's rotations, scale and translation) reproduces the baked
    scene's vertices, triangles and materials."""
    d, a, env, xf = assets
    scene = json.loads((d / "scene.json").read_text())
    scene["props"] = [dict(uri="a.gltf", transforms=xf), dict(uri="a.gltf"), dict(uri="b.glb", transforms=xf)]
    (d / "inst.json").write_text(json.dumps(scene))
    baked = _load(d, "inst.json", tmp_path / "baked")
    node(CLI, "load-scene", str(d / "inst.json"), str(tmp_path / "inst"), "--instanced")
    inst = scenes.SceneData.load(str(tmp_path / "inst"))
    assert inst.num_triangles == 0 and len(inst.meshes) == 2 and len(inst.instances) == 3
    assert [i.mesh for i in inst.instances] == [0, 0, 1]
    b = inst.bake()
    np.testing.assert_allclose(b.positions, baked.positions, rtol=0, atol=1e-12)
    np.testing.assert_array_equal(b.triangles, baked.triangles)
    np.testing.assert_array_equal(inst.materials, baked.materials)
    np.testing.assert_array_equal(inst.tex, baked.tex)
    # the two-level BVH spans the baked scene
    np.testing.assert_allclose(pike_amd.BVH.for_scene(inst).normalizer(),
                               pike_amd.BVH(baked.positions, baked.triangles).normalizer(), rtol=1e-12)


@pytest.mark.gpu
def test_instanced_scene_renders_through_the_js_host(assets, tmp_path, ctx):
    """render-scene --instanced (two-level BVH through the N-API addon)
    equals the Python harness on the same instanced scene bit for bit."""
    d, a, env, xf = assets
    scene = json.loads((d / "scene.json").read_text())
    scene["props"] = [dict(uri="a.gltf", transforms=xf), dict(uri="a.gltf"), dict(uri="b.glb")]
    (d / "inst_r.json").write_text(json.dumps(scene))
    node(CLI, "load-scene", str(d / "inst_r.json"), str(tmp_path / "inst"), "--instanced")
    s = scenes.SceneData.load(str(tmp_path / "inst"))
    W, H, frames, depth = 40, 30, 2, 3
    node(CLI, "render-scene", str(d / "inst_r.json"), str(tmp_path / "img.png"), str(W), str(H), str(frames),
         str(depth), str(tmp_path / "img.f32"), "--instanced")
    img_js = np.fromfile(tmp_path / "img.f32", np.float32).reshape(H, W, 4)
    bvh = pike_amd.BVH.for_scene(s)
    sc = pike_amd.Scene(ctx, s, bvh)
    fr = pike_amd.Frame(ctx, W, H)
    fr.set_camera(scenes.normalized_camera(s.camera, bvh.normalizer()))
    fr.render(sc, max_depth=depth, num_frames=frames, use_graph=False)
    np.testing.assert_array_equal(img_js.view(np.uint32), fr.read_image().view(np.uint32))
    assert img_js[..., :3].std() > 0
