"""Procedural scenes and rays for the five benchmark configurations.

Everything is generated from a counter-based PCG hash (the same hash the
renderer uses, Jarzynski & Olano 2020), so a (config, seed) pair always gives
the same bytes.  Scenes are produced in the reference's host representation
(render/scene.js:120-181): float64 world-space positions, ``scene.indices``
{i0, i1, i2, matId}, quantized vertex attributes (snorm8 tangent/bitangent/
normal, snorm16 uv, truncated the way JS Int8Array/Int16Array stores do,
scene.js:104-118) and an RGBA8 (or RGBA32F) texture atlas whose constant
colors are 1x1 layers, as texture_packer.js:99-116 packs them.
"""
from dataclasses import dataclass, field
import ctypes
import math

import numpy as np

from . import _ffi

U32 = np.uint32


# ---------------------------------------------------------------------------
# counter-based random numbers
# ---------------------------------------------------------------------------
def pcg_hash(x):
    """pcg_hash (RXS-M-XS output of the PCG LCG state) on uint32 arrays."""
    x = np.asarray(x, dtype=np.uint32)
    with np.errstate(over="ignore"):
        state = x
        word = ((state >> ((state >> U32(28)) + U32(4))) ^ state) * U32(277803737)
        return (word >> U32(22)) ^ word


class Rng:
    """Stream of uniform doubles: u = (53 random bits) / 2^53 from two hashes
    of a running counter mixed with the seed key."""

    def __init__(self, seed):
        self.key = int(pcg_hash(np.uint32(seed & 0xFFFFFFFF) ^ np.uint32(0x9E3779B9)))
        self.ctr = 0

    def u32(self, n):
        i = (np.arange(self.ctr, self.ctr + n, dtype=np.uint64) & np.uint64(0xFFFFFFFF)).astype(np.uint32)
        self.ctr += n
        with np.errstate(over="ignore"):
            return pcg_hash(pcg_hash(i) ^ np.uint32(self.key))

    def uniform(self, n, lo=0.0, hi=1.0):
        a = self.u32(n).astype(np.uint64) >> np.uint64(5)
        b = self.u32(n).astype(np.uint64) >> np.uint64(6)
        u = (a.astype(np.float64) * 67108864.0 + b.astype(np.float64)) / 9007199254740992.0
        return lo + (hi - lo) * u

    def normal(self, n):
        u1 = self.uniform(n)
        u2 = self.uniform(n)
        return np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * math.pi * u2)

    def unit_vectors(self, n):
        z = self.uniform(n, -1.0, 1.0)
        phi = self.uniform(n, 0.0, 2.0 * math.pi)
        r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
        return np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1)


# ---------------------------------------------------------------------------
# quantization as the reference's typed-array stores perform it
# ---------------------------------------------------------------------------
def _js_int(x, bits):
    """ECMAScript ToInt8/ToInt16: truncate toward zero, wrap modulo 2^bits."""
    x = np.where(np.isfinite(x), np.trunc(x), 0.0)
    m = np.mod(x, float(1 << bits)).astype(np.int64)
    return np.where(m >= (1 << (bits - 1)), m - (1 << bits), m)


def quantize_snorm8x3(v):
    """scene.js:104-111 (_quantizeVec3ToInt): bytes = v * (v > 0 ? 127 : 128)."""
    q = _js_int(np.where(v > 0, v * 127.0, v * 128.0), 8) & 0xFF
    return (q[:, 0] | (q[:, 1] << 8) | (q[:, 2] << 16)).astype(np.uint32)


def quantize_snorm16x2(v):
    """scene.js:113-118 (_quantizeVec2ToInt)."""
    q = _js_int(np.where(v > 0, v * 32767.0, v * 32768.0), 16) & 0xFFFF
    return (q[:, 0] | (q[:, 1] << 16)).astype(np.uint32)


def _normalize(v):
    n = np.linalg.norm(v, axis=-1, keepdims=True)
    return v / np.where(n > 0, n, 1.0)


def vertex_attributes(normal, tangent, uv):
    """Quantized {tangent, bitangent, normal, uv} words; the bitangent is
    normalize(cross(normal, tangent)) as scene.js:163 computes it."""
    n = _normalize(normal)
    t = _normalize(tangent)
    b = _normalize(np.cross(n, t))
    out = np.zeros((len(n), 4), dtype=np.uint32)
    out[:, 0] = quantize_snorm8x3(t)
    out[:, 1] = quantize_snorm8x3(b)
    out[:, 2] = quantize_snorm8x3(n)
    out[:, 3] = quantize_snorm16x2(uv)
    return out


def unorm8(c):
    """Float color to an RGBA8 channel (GL clear-color conversion, round to nearest)."""
    return np.clip(np.floor(np.asarray(c, dtype=np.float64) * 255.0 + 0.5), 0, 255).astype(np.uint8)


# ---------------------------------------------------------------------------
# meshes
# ---------------------------------------------------------------------------
_ICO_CACHE = {}


def icosphere(subdiv):
    """Unit icosphere: 20*4^s triangles, 10*4^s+2 vertices."""
    if subdiv in _ICO_CACHE:
        return _ICO_CACHE[subdiv]
    p = (1.0 + math.sqrt(5.0)) / 2.0
    verts = [(-1, p, 0), (1, p, 0), (-1, -p, 0), (1, -p, 0), (0, -1, p), (0, 1, p), (0, -1, -p), (0, 1, -p),
             (p, 0, -1), (p, 0, 1), (-p, 0, -1), (-p, 0, 1)]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
             (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    verts = [tuple(np.array(v, dtype=np.float64) / np.linalg.norm(v)) for v in verts]
    for _ in range(subdiv):
        cache = {}
        new_faces = []

        def mid(a, b):
            key = (a, b) if a < b else (b, a)
            if key not in cache:
                m = (np.array(verts[a]) + np.array(verts[b])) * 0.5
                verts.append(tuple(m / np.linalg.norm(m)))
                cache[key] = len(verts) - 1
            return cache[key]

        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            new_faces += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = new_faces
    v = np.array(verts, dtype=np.float64)
    f = np.array(faces, dtype=np.int64)
    _ICO_CACHE[subdiv] = (v, f)
    return v, f


def _sphere_attrs(unit):
    n = unit
    t = np.cross(np.array([0.0, 1.0, 0.0]), n)
    bad = np.linalg.norm(t, axis=1) < 1e-6
    t[bad] = np.array([1.0, 0.0, 0.0])
    uv = np.stack([np.arctan2(n[:, 2], n[:, 0]) / (2 * math.pi) + 0.5, np.arccos(np.clip(n[:, 1], -1, 1)) / math.pi],
                  axis=1)
    return vertex_attributes(n, t, uv)


def _value_noise(x, z, octave_seed, cells):
    """Smooth value noise on a `cells` lattice over [0, 1]^2."""
    gx, gz = x * cells, z * cells
    ix, iz = np.floor(gx).astype(np.int64), np.floor(gz).astype(np.int64)
    fx, fz = gx - ix, gz - iz
    sx, sz = fx * fx * (3 - 2 * fx), fz * fz * (3 - 2 * fz)

    def h(i, j):
        k = (i * 73856093) ^ (j * 19349663) ^ (octave_seed * 83492791)
        return pcg_hash((k & 0xFFFFFFFF).astype(np.uint32)).astype(np.float64) / 4294967296.0

    a, b = h(ix, iz), h(ix + 1, iz)
    c, d = h(ix, iz + 1), h(ix + 1, iz + 1)
    return (a + (b - a) * sx) + ((c + (d - c) * sx) - (a + (b - a) * sx)) * sz


def fbm_height(x, z, seed, octaves=6, amplitude=0.2):
    h = np.zeros_like(x)
    amp, total = 1.0, 0.0
    for o in range(octaves):
        h += amp * _value_noise(x, z, seed * 131 + o, 4 * (2 ** o))
        total += amp
        amp *= 0.5
    return amplitude * (h / total)


def terrain(cells, seed, amplitude=0.2):
    """(cells+1)^2 vertex heightfield over x, z in [0, 1] with 2*cells^2 triangles."""
    g = np.linspace(0.0, 1.0, cells + 1)
    X, Z = np.meshgrid(g, g, indexing="xy")
    x, z = X.ravel(), Z.ravel()
    y = fbm_height(x, z, seed, 6, amplitude)
    pos = np.stack([x, y, z], axis=1)
    Y = y.reshape(cells + 1, cells + 1)
    dx = np.gradient(Y, g, axis=1).ravel()
    dz = np.gradient(Y, g, axis=0).ravel()
    n = _normalize(np.stack([-dx, np.ones_like(dx), -dz], axis=1))
    t = _normalize(np.stack([np.ones_like(dx), dx, np.zeros_like(dx)], axis=1))
    attrs = vertex_attributes(n, t, np.stack([x, z], axis=1))
    i = np.arange(cells)
    I, J = np.meshgrid(i, i, indexing="xy")
    v00 = (J * (cells + 1) + I).ravel()
    v10, v01, v11 = v00 + 1, v00 + cells + 1, v00 + cells + 2
    tris = np.concatenate([np.stack([v00, v01, v10], 1), np.stack([v10, v01, v11], 1)], axis=0)
    return pos, attrs, tris


# ---------------------------------------------------------------------------
# scene container
# ---------------------------------------------------------------------------
class Mesh:
    """Object-space geometry for instancing."""

    def __init__(self, positions, attributes, triangles):
        self.positions = np.asarray(positions, np.float64)
        self.attributes = np.asarray(attributes, np.uint32)
        self.triangles = np.asarray(triangles, np.int32)


class Instance:
    """A placement of mesh `mesh`: object -> scene affine `transform` (3x4),
    `material` for all its triangles (-1: the mesh's own)."""

    def __init__(self, mesh, transform, material=-1):
        self.mesh = int(mesh)
        self.transform = np.asarray(transform, np.float64).reshape(3, 4)
        self.material = int(material)


def _unpack_snorm8x3(words):
    b = np.stack([(words >> s) & 0xFF for s in (0, 8, 16)], 1).astype(np.uint8).view(np.int8).astype(np.float64)
    return np.maximum(b / 127.0, -1.0)


def transform_attributes(attrs, L):
    """Tangent/bitangent by L, normal by inv(L)^T, normalized and re-quantized
    (uv unchanged)."""
    a = np.asarray(attrs, np.uint32)
    if a.size == 0:
        return a.reshape(0, 4)
    out = a.copy()
    N = np.linalg.inv(L).T
    for word, M in ((0, L), (1, L), (2, N)):
        v = _unpack_snorm8x3(a[:, word]) @ M.T
        out[:, word] = quantize_snorm8x3(_normalize(v))
    return out


@dataclass
class SceneData:
    positions: np.ndarray          # (nv, 3) float64 world space
    attributes: np.ndarray         # (nv, 4) uint32
    triangles: np.ndarray          # (nt, 4) int32 {i0, i1, i2, mat}
    materials: np.ndarray          # (nm, 4) int32 atlas layers
    tex: np.ndarray                # (layers, h, w, 4) uint8 or float32
    env: np.ndarray = None         # (h, w, 4) uint8 RGBE or (h, w, 3) float32
    env_sampler: int = _ffi.PK_ENV_SAMPLER_BINS
    camera: dict = field(default_factory=dict)   # world-space camera
    name: str = ""
    meshes: list = field(default_factory=list)     # [Mesh] instanced geometry (object space)
    instances: list = field(default_factory=list)  # [Instance] placements of the meshes
    _keep: list = field(default_factory=list, repr=False)

    @property
    def num_triangles(self):
        return int(self.triangles.shape[0])

    @property
    def baked_triangles(self):
        """Triangles of the equivalent baked scene (own + every instance's)."""
        return self.num_triangles + sum(int(self.meshes[i.mesh].triangles.shape[0]) for i in self.instances)

    def instance_descs(self):
        """ctypes pk_mesh_desc / pk_instance_desc arrays (+ the arrays they point into)."""
        keep = []
        meshes = (_ffi.MeshDesc * max(1, len(self.meshes)))()
        for k, m in enumerate(self.meshes):
            pos = np.ascontiguousarray(m.positions, np.float64)
            attr = np.ascontiguousarray(m.attributes, np.uint32)
            tri = np.ascontiguousarray(m.triangles, np.int32)
            keep += [pos, attr, tri]
            meshes[k].num_vertices = pos.shape[0]
            meshes[k].positions = pos.ctypes.data
            meshes[k].attributes = attr.ctypes.data
            meshes[k].num_triangles = tri.shape[0]
            meshes[k].triangles = tri.ctypes.data
        insts = (_ffi.InstanceDesc * max(1, len(self.instances)))()
        for k, it in enumerate(self.instances):
            insts[k].mesh = it.mesh
            insts[k].material = it.material
            t = np.asarray(it.transform, np.float64).reshape(12)
            for j in range(12):
                insts[k].transform[j] = float(t[j])
        return meshes, insts, keep

    def bake(self):
        """The equivalent single-level scene (what scene.js:159-170 builds):
        every instance's mesh transformed into the scene, in instance order,
        after the scene's own triangles.  Tangents map by the linear part,
        normals by its inverse transpose, re-quantized to snorm8."""
        pos, attr, tri = [np.asarray(self.positions, np.float64)], [np.asarray(self.attributes, np.uint32)], \
            [np.asarray(self.triangles, np.int32)]
        nv = pos[0].shape[0]
        for it in self.instances:
            m = self.meshes[it.mesh]
            T = np.asarray(it.transform, np.float64).reshape(3, 4)
            p = np.asarray(m.positions, np.float64)
            # ((T0 x + T1 y) + T2 z) + T3 per row, as the instanced build evaluates it
            w = np.stack([((T[k, 0] * p[:, 0] + T[k, 1] * p[:, 1]) + T[k, 2] * p[:, 2]) + T[k, 3]
                          for k in range(3)], 1)
            pos.append(w)
            attr.append(transform_attributes(m.attributes, T[:, :3]))
            t = np.asarray(m.triangles, np.int32).copy()
            t[:, :3] += nv
            if it.material >= 0:
                t[:, 3] = it.material
            tri.append(t)
            nv += p.shape[0]
        return SceneData(positions=np.concatenate(pos), attributes=np.concatenate(attr), triangles=np.concatenate(tri),
                         materials=self.materials, tex=self.tex, env=self.env, env_sampler=self.env_sampler,
                         camera=dict(self.camera), name=self.name + "-baked")

    @classmethod
    def load(cls, directory):
        """Reads the raw format `save` (and pike_amd/js saveRawScene) writes."""
        import json
        import os

        m = json.load(open(os.path.join(directory, "scene.json")))
        dt = {"f64": np.float64, "f32": np.float32, "i32": np.int32, "u32": np.uint32, "u8": np.uint8}

        def get(entry):
            return np.fromfile(os.path.join(directory, entry["file"]), dt[entry["dtype"]])

        a = m["atlas"]
        tex = get(a["data"]).reshape(a["layers"], a["height"], a["width"], 4)
        env, sampler = None, _ffi.PK_ENV_SAMPLER_BINS
        if m.get("env"):
            e = m["env"]
            env = get(e["data"]).reshape(e["height"], e["width"], -1)
            sampler = int(e.get("sampler", sampler))
        cam = {k: (np.array(v) if isinstance(v, list) else v) for k, v in (m.get("camera") or {}).items()}
        meshes = [Mesh(get(x["positions"]).reshape(-1, 3), get(x["attributes"]).reshape(-1, 4),
                       get(x["triangles"]).reshape(-1, 4)) for x in m.get("meshes", [])]
        instances = [Instance(i["mesh"], np.array(i["transform"], np.float64), i.get("material", -1))
                     for i in m.get("instances", [])]
        return cls(positions=get(m["positions"]).reshape(-1, 3), attributes=get(m["attributes"]).reshape(-1, 4),
                   triangles=get(m["triangles"]).reshape(-1, 4), materials=get(m["materials"]).reshape(-1, 4),
                   tex=tex, env=env, env_sampler=sampler, camera=cam, name=m.get("name", ""), meshes=meshes,
                   instances=instances)

    def save(self, directory):
        """Raw arrays + scene.json manifest (read by pike_amd/js/raw_scene.js)."""
        import json
        import os

        os.makedirs(directory, exist_ok=True)

        def put(name, arr, dtype):
            arr = np.ascontiguousarray(arr, dtype=dtype)
            arr.tofile(os.path.join(directory, name))
            code = {np.float64: "f64", np.float32: "f32", np.int32: "i32", np.uint32: "u32", np.uint8: "u8"}
            return {"file": name, "dtype": code[np.dtype(dtype).type]}

        tex = np.ascontiguousarray(self.tex)
        m = {
            "name": self.name,
            "positions": put("positions.bin", self.positions, np.float64),
            "attributes": put("attributes.bin", self.attributes, np.uint32),
            "triangles": put("triangles.bin", self.triangles, np.int32),
            "materials": put("materials.bin", self.materials, np.int32),
            "atlas": {"width": int(tex.shape[2]), "height": int(tex.shape[1]), "layers": int(tex.shape[0]),
                      "format": _ffi.PK_TEX_RGBA8 if tex.dtype == np.uint8 else _ffi.PK_TEX_RGBA32F,
                      "data": put("atlas.bin", tex, tex.dtype)},
            "camera": {k: (v.tolist() if hasattr(v, "tolist") else v) for k, v in self.camera.items()},
        }
        if self.env is not None:
            env = np.ascontiguousarray(self.env)
            m["env"] = {"width": int(env.shape[1]), "height": int(env.shape[0]),
                        "format": _ffi.PK_ENV_RGBE8 if env.dtype == np.uint8 else _ffi.PK_ENV_RGB32F,
                        "sampler": int(self.env_sampler), "data": put("env.bin", env, env.dtype)}
        if self.instances:
            m["meshes"] = [{"positions": put(f"mesh{k}_positions.bin", x.positions, np.float64),
                            "attributes": put(f"mesh{k}_attributes.bin", x.attributes, np.uint32),
                            "triangles": put(f"mesh{k}_triangles.bin", x.triangles, np.int32)}
                           for k, x in enumerate(self.meshes)]
            m["instances"] = [{"mesh": i.mesh, "material": i.material, "transform": i.transform.ravel().tolist()}
                              for i in self.instances]
        with open(os.path.join(directory, "scene.json"), "w") as f:
            json.dump(m, f)

    def desc(self):
        """ctypes pk_scene_desc over this scene's arrays (kept alive by self)."""
        pos = np.ascontiguousarray(self.positions, dtype=np.float64)
        attr = np.ascontiguousarray(self.attributes, dtype=np.uint32)
        tri = np.ascontiguousarray(self.triangles, dtype=np.int32)
        mat = np.ascontiguousarray(self.materials, dtype=np.int32)
        tex = np.ascontiguousarray(self.tex)
        d = _ffi.SceneDesc()
        d.num_vertices = pos.shape[0]
        d.positions = pos.ctypes.data
        d.attributes = attr.ctypes.data
        d.num_triangles = tri.shape[0]
        d.triangles = tri.ctypes.data
        d.num_materials = mat.shape[0]
        d.materials = mat.ctypes.data
        d.tex_layers, d.tex_height, d.tex_width = tex.shape[0], tex.shape[1], tex.shape[2]
        d.tex_format = _ffi.PK_TEX_RGBA8 if tex.dtype == np.uint8 else _ffi.PK_TEX_RGBA32F
        d.tex_data = tex.ctypes.data
        keep = [pos, attr, tri, mat, tex]
        if self.env is None:
            d.env_format = _ffi.PK_ENV_NONE
        else:
            env = np.ascontiguousarray(self.env)
            d.env_height, d.env_width = env.shape[0], env.shape[1]
            d.env_format = _ffi.PK_ENV_RGBE8 if env.dtype == np.uint8 else _ffi.PK_ENV_RGB32F
            d.env_data = env.ctypes.data
            d.env_sampler = self.env_sampler
            keep.append(env)
        self._keep = keep
        return d


class AtlasPacker:
    """texture_packer.js addColor semantics: one 1x1 RGBA8 layer per distinct
    color (deduplicated by value), alpha forced to 1 by gl.clearColor."""

    def __init__(self):
        self.layers = []
        self.keys = {}

    def add_color(self, rgb):
        q = tuple(int(v) for v in unorm8(list(rgb)[:3])) + (255,)
        if q not in self.keys:
            self.keys[q] = len(self.layers)
            self.layers.append(q)
        return self.keys[q]

    def material(self, base, roughness, metallic, emission=(0.0, 0.0, 0.0)):
        """scene.js:43-70 (_createSceneMaterial) for constant-color materials:
        metRough texel (0, roughness, metallic), flat normal (0.5, 0.5, 1)."""
        return [self.add_color(base), self.add_color((0.0, roughness, metallic)),
                self.add_color((0.5, 0.5, 1.0)), self.add_color(emission)]

    def atlas(self):
        return np.array(self.layers, dtype=np.uint8).reshape(len(self.layers), 1, 1, 4)


def constant_env_rgbe(width=64, height=32, value=1.0):
    """RGBE-in-alpha env (env_sampler.js:43-48): rgb8 * 2^(a-128) / 255."""
    e = np.zeros((height, width, 4), dtype=np.uint8)
    exp = int(math.ceil(math.log2(value))) if value > 1 else 0
    m = int(round(value / (2.0 ** exp) * 255))
    e[..., 0:3] = m
    e[..., 3] = 128 + exp
    return e


def lognormal_sun_env(width, height, seed, sigma=1.5, sun_deg=0.5, sun_scale=1e4, sun_dir=(0.45, 0.64, 0.62)):
    """Float RGB env: per-texel log-normal radiance plus a sun disk."""
    rng = Rng(seed)
    L = np.exp(sigma * rng.normal(width * height)).reshape(height, width)
    mean = float(L.mean())
    v = (np.arange(height) + 0.5) / height
    u = (np.arange(width) + 0.5) / width
    V, Uu = np.meshgrid(v, u, indexing="ij")
    th, ph = Uu * 2 * math.pi, V * math.pi
    d = np.stack([np.cos(th) * np.sin(ph), np.cos(ph), np.sin(th) * np.sin(ph)], axis=-1)
    s = np.array(sun_dir, dtype=np.float64)
    s /= np.linalg.norm(s)
    cosang = d @ s
    disk = cosang >= math.cos(math.radians(sun_deg * 0.5))
    L = np.where(disk, sun_scale * mean, L)
    tint = np.array([1.0, 0.97, 0.92])
    return (L[..., None] * tint[None, None, :]).astype(np.float32)


def camera(origin, target, fov=1.0, focal_depth=0.5, aperture=0.0, distortion=0.0, bokeh=0.5):
    o = np.array(origin, dtype=np.float64)
    d = np.array(target, dtype=np.float64) - o
    return dict(origin=o, dir=d / np.linalg.norm(d), fov=fov, focal_depth=focal_depth, aperture=aperture,
                distortion=distortion, bokeh=bokeh)


def normalized_camera(cam, normalizer):
    """The tracer works in trace_pass.js's normalized space: x' = (x - c) * s."""
    c = _ffi.Camera()
    center, s = np.asarray(normalizer[:3], dtype=np.float64), float(normalizer[3])
    o = (np.asarray(cam["origin"], dtype=np.float64) - center) * s
    for k in range(3):
        c.origin[k] = float(o[k])
        c.dir[k] = float(cam["dir"][k])
    c.fov = cam["fov"]
    c.focal_depth = cam["focal_depth"]
    c.aperture = cam["aperture"]
    c.distortion = cam["distortion"]
    c.bokeh = cam["bokeh"]
    return c


def _merge(parts):
    pos, attr, tris = [], [], []
    off = 0
    for p, a, t, m in parts:
        pos.append(p)
        attr.append(a)
        tt = np.empty((len(t), 4), dtype=np.int64)
        tt[:, :3] = t + off
        tt[:, 3] = m
        tris.append(tt)
        off += len(p)
    return (np.concatenate(pos).astype(np.float64), np.concatenate(attr).astype(np.uint32),
            np.concatenate(tris).astype(np.int32))


# ---------------------------------------------------------------------------
# configurations
# ---------------------------------------------------------------------------
C1_GROUND = dict(y=-0.1, extent=(-1.0, 2.0), base=(0.5, 0.5, 0.5), roughness=0.8, metallic=0.0)
C1_CAMERA = dict(origin=(0.5, 0.55, 2.2), target=(0.5, 0.25, 0.45), fov=1.0)


def config1(seed=1, num_spheres=64, subdiv=2, env_size=(64, 32)):
    """64 icospheres (subdivision 2) at U[0,1]^3 with radii U[0.02, 0.1], a
    2-triangle ground quad, per-mesh albedo U[0.2, 0.8]^3, roughness
    U[0.05, 1], metalness Bernoulli(0.3), constant white environment."""
    rng = Rng(seed)
    uv_, f = icosphere(subdiv)
    centers = rng.uniform(num_spheres * 3).reshape(num_spheres, 3)
    radii = rng.uniform(num_spheres, 0.02, 0.1)
    albedo = rng.uniform(num_spheres * 3, 0.2, 0.8).reshape(num_spheres, 3)
    rough = rng.uniform(num_spheres, 0.05, 1.0)
    metal = rng.uniform(num_spheres) < 0.3
    packer = AtlasPacker()
    mats = []
    parts = []
    sphere_attr = _sphere_attrs(uv_)
    for i in range(num_spheres):
        mats.append(packer.material(albedo[i], rough[i], 1.0 if metal[i] else 0.0))
        parts.append((centers[i] + radii[i] * uv_, sphere_attr, f, i))
    g = C1_GROUND
    lo, hi = g["extent"]
    qp = np.array([[lo, g["y"], lo], [hi, g["y"], lo], [hi, g["y"], hi], [lo, g["y"], hi]], dtype=np.float64)
    qa = vertex_attributes(np.tile([0.0, 1.0, 0.0], (4, 1)), np.tile([1.0, 0.0, 0.0], (4, 1)),
                           np.array([[0, 0], [1, 0], [1, 1], [0, 1]], dtype=np.float64))
    mats.append(packer.material(g["base"], g["roughness"], g["metallic"]))
    parts.append((qp, qa, np.array([[0, 2, 1], [0, 3, 2]]), num_spheres))
    pos, attr, tris = _merge(parts)
    return SceneData(positions=pos, attributes=attr, triangles=tris, materials=np.array(mats, dtype=np.int32),
                     tex=packer.atlas(), env=constant_env_rgbe(*env_size),
                     camera=camera(C1_CAMERA["origin"], C1_CAMERA["target"], C1_CAMERA["fov"]), name="c1")


C2_CAMERA = dict(origin=(0.5, 0.42, 1.2), target=(0.5, 0.02, 0.4), fov=1.0)


def config2(seed=2, cells=1024, num_spheres=200, subdiv=4, env_size=(2048, 1024), env_sampler=None, instanced=False):
    """Terrain of cells^2 quads (fBm, 6 octaves, amplitude 0.2) plus
    num_spheres icospheres (subdivision 4); log-normal (sigma 1.5) HDR
    environment with a 0.5-degree sun at 1e4 x mean radiance, importance
    sampled through the marginal/conditional CDF (pass
    env_sampler=PK_ENV_SAMPLER_BINS for env_sampler.js's bins).  With
    instanced=True the spheres are instances of one unit icosphere mesh
    (uniform scale + translation, own material each) over the terrain; bake()
    of that scene places the same vertices as the baked default."""
    rng = Rng(seed)
    tp, ta, tt = terrain(cells, seed, 0.2)
    packer = AtlasPacker()
    mats = [packer.material((0.45, 0.40, 0.33), 0.9, 0.0)]
    parts = [(tp, ta, tt, 0)]
    uv_, f = icosphere(subdiv)
    sa = _sphere_attrs(uv_)
    cx = rng.uniform(num_spheres, 0.05, 0.95)
    cz = rng.uniform(num_spheres, 0.05, 0.95)
    r = rng.uniform(num_spheres, 0.01, 0.04)
    cy = fbm_height(cx, cz, seed, 6, 0.2) + 0.8 * r
    albedo = rng.uniform(num_spheres * 3, 0.2, 0.8).reshape(num_spheres, 3)
    rough = rng.uniform(num_spheres, 0.05, 1.0)
    metal = rng.uniform(num_spheres) < 0.3
    instances = []
    for i in range(num_spheres):
        mats.append(packer.material(albedo[i], rough[i], 1.0 if metal[i] else 0.0))
        c = np.array([cx[i], cy[i], cz[i]])
        if instanced:
            instances.append(Instance(0, [[r[i], 0, 0, c[0]], [0, r[i], 0, c[1]], [0, 0, r[i], c[2]]], i + 1))
        else:
            parts.append((c + r[i] * uv_, sa, f, i + 1))
    pos, attr, tris = _merge(parts)
    meshes = []
    if instanced:
        mp, ma, mt = _merge([(uv_, sa, f, 0)])
        meshes.append(Mesh(mp, ma, mt))
    env = lognormal_sun_env(env_size[0], env_size[1], seed + 17)
    return SceneData(positions=pos, attributes=attr, triangles=tris, materials=np.array(mats, dtype=np.int32),
                     tex=packer.atlas(), env=env,
                     env_sampler=_ffi.PK_ENV_SAMPLER_CDF if env_sampler is None else env_sampler,
                     camera=camera(C2_CAMERA["origin"], C2_CAMERA["target"], C2_CAMERA["fov"]),
                     name="c2-instanced" if instanced else "c2", meshes=meshes, instances=instances)


def config3(seed=3, num_tris=10_000_000):
    """Triangle soup: v0 ~ U[0,1]^3, two edges of length U[0.001, 0.01] in
    uniform directions."""
    rng = Rng(seed)
    v0 = rng.uniform(num_tris * 3).reshape(num_tris, 3)
    e1 = rng.unit_vectors(num_tris) * rng.uniform(num_tris, 0.001, 0.01)[:, None]
    e2 = rng.unit_vectors(num_tris) * rng.uniform(num_tris, 0.001, 0.01)[:, None]
    pos = np.empty((num_tris * 3, 3), dtype=np.float64)
    pos[0::3] = v0
    pos[1::3] = v0 + e1
    pos[2::3] = v0 + e2
    tris = np.zeros((num_tris, 4), dtype=np.int32)
    tris[:, 0] = np.arange(0, 3 * num_tris, 3)
    tris[:, 1] = tris[:, 0] + 1
    tris[:, 2] = tris[:, 0] + 2
    packer = AtlasPacker()
    mats = [packer.material((0.5, 0.5, 0.5), 0.5, 0.0)]
    attr = np.zeros((num_tris * 3, 4), dtype=np.uint32)
    return SceneData(positions=pos, attributes=attr, triangles=tris, materials=np.array(mats, dtype=np.int32),
                     tex=packer.atlas(), env=None, name="c3")


def incoherent_rays(num_rays, normalizer, seed=33):
    """Origins U[0,1]^3 (world), directions uniform on S^2; returned as
    (n, 8) float32 {o, 0, d, 0} in normalized space (f64 map, one f32 round)."""
    rng = Rng(seed)
    o = rng.uniform(num_rays * 3).reshape(num_rays, 3)
    d = rng.unit_vectors(num_rays)
    center, s = np.asarray(normalizer[:3]), float(normalizer[3])
    rays = np.zeros((num_rays, 8), dtype=np.float32)
    rays[:, 0:3] = ((o - center) * s).astype(np.float32)
    rays[:, 4:7] = d.astype(np.float32)
    return rays


def config5(seed=5, cells=300, num_spheres=50, tex_res=512, env_size=(512, 256)):
    """~200k triangles (terrain + icospheres) with 512^2 float albedo (RGB)
    and roughness textures and a 512x256 float environment."""
    rng = Rng(seed)
    tp, ta, tt = terrain(cells, seed, 0.2)
    uv_, f = icosphere(2)
    sa = _sphere_attrs(uv_)
    cx = rng.uniform(num_spheres, 0.05, 0.95)
    cz = rng.uniform(num_spheres, 0.05, 0.95)
    r = rng.uniform(num_spheres, 0.02, 0.05)
    cy = fbm_height(cx, cz, seed, 6, 0.2) + 0.8 * r
    parts = [(tp, ta, tt, 0)] + [(np.array([cx[i], cy[i], cz[i]]) + r[i] * uv_, sa, f, 0) for i in range(num_spheres)]
    pos, attr, tris = _merge(parts)
    g = (np.arange(tex_res) + 0.5) / tex_res
    X, Y = np.meshgrid(g, g, indexing="xy")
    tex = np.zeros((4, tex_res, tex_res, 4), dtype=np.float32)
    for c in range(3):
        tex[0, ..., c] = 0.2 + 0.6 * _value_noise(X.ravel(), Y.ravel(), seed * 7 + c, 16).reshape(X.shape)
    tex[0, ..., 3] = 1.0
    tex[1, ..., 1] = 0.2 + 0.7 * _value_noise(X.ravel(), Y.ravel(), seed * 7 + 5, 8).reshape(X.shape)
    tex[1, ..., 3] = 1.0
    tex[2, ..., 0:3] = (0.5, 0.5, 1.0)
    tex[2, ..., 3] = 1.0
    tex[3, ..., 3] = 1.0
    env = lognormal_sun_env(env_size[0], env_size[1], seed + 17, sigma=1.0, sun_deg=4.0, sun_scale=50.0)
    return SceneData(positions=pos, attributes=attr, triangles=tris, materials=np.array([[0, 1, 2, 3]], np.int32),
                     tex=tex, env=env, camera=camera((0.5, 0.5, 1.4), (0.5, 0.05, 0.45), 1.0), name="c5")


C4_RESOLUTIONS = ((128, 128), (256, 256), (512, 384), (1024, 1024))


def mixed_scene(seed, target_tris, env=None):
    """Terrain + subdivision-2 icospheres with about target_tris triangles
    (a quarter to two thirds in the terrain, the rest in spheres)."""
    rng = Rng(seed)
    frac = rng.uniform(1, 0.25, 0.67)[0]
    cells = max(4, int(math.sqrt(target_tris * frac / 2)))
    tp, ta, tt = terrain(cells, seed, 0.2)
    packer = AtlasPacker()
    mats = [packer.material(rng.uniform(3, 0.3, 0.6), rng.uniform(1, 0.6, 1.0)[0], 0.0)]
    parts = [(tp, ta, tt, 0)]
    uv_, f = icosphere(2)
    sa = _sphere_attrs(uv_)
    n = max(1, (target_tris - len(tt)) // len(f))
    cx, cz = rng.uniform(n, 0.05, 0.95), rng.uniform(n, 0.05, 0.95)
    r = rng.uniform(n, 0.01, 0.05)
    cy = fbm_height(cx, cz, seed, 6, 0.2) + 0.8 * r
    albedo = rng.uniform(n * 3, 0.2, 0.8).reshape(n, 3)
    rough = rng.uniform(n, 0.05, 1.0)
    metal = rng.uniform(n) < 0.3
    for i in range(n):
        mats.append(packer.material(albedo[i], rough[i], 1.0 if metal[i] else 0.0))
        parts.append((np.array([cx[i], cy[i], cz[i]]) + r[i] * uv_, sa, f, i + 1))
    pos, attr, tris = _merge(parts)
    return SceneData(positions=pos, attributes=attr, triangles=tris, materials=np.array(mats, dtype=np.int32),
                     tex=packer.atlas(), env=env, name=f"mixed{seed}")


def config4(seed=4, num_scenes=32, num_views=64, tri_range=(1_000, 500_000), spp_range=(8, 64), env_size=(512, 256)):
    """Batching: scenes with triangle counts log-uniform in tri_range and
    views with resolutions from C4_RESOLUTIONS and spp in spp_range; one
    shared log-normal + sun environment.  Returns (scenes, views) with views
    as (scene index, width, height, spp, world-space camera dict)."""
    rng = Rng(seed)
    lo, hi = math.log(tri_range[0]), math.log(tri_range[1])
    counts = np.exp(rng.uniform(num_scenes, lo, hi)).astype(np.int64)
    env = lognormal_sun_env(env_size[0], env_size[1], seed + 17, sigma=1.0, sun_deg=2.0, sun_scale=100.0)
    sc = [mixed_scene(seed * 1000 + k, int(counts[k]), env=env) for k in range(num_scenes)]
    views = []
    res_idx = rng.u32(num_views) % len(C4_RESOLUTIONS)
    spp = spp_range[0] + (rng.u32(num_views) % (spp_range[1] - spp_range[0] + 1))
    az = rng.uniform(num_views, 0.0, 2 * math.pi)
    for v in range(num_views):
        k = v % num_scenes
        w, h = C4_RESOLUTIONS[int(res_idx[v])]
        o = (0.5 + 0.75 * math.cos(az[v]), 0.45, 0.5 + 0.75 * math.sin(az[v]))
        views.append((k, w, h, int(spp[v]), camera(o, (0.5, 0.05, 0.5), 1.0)))
    return sc, views


def _rotation(axis, angle):
    """Rodrigues rotation matrix about a unit axis."""
    x, y, z = axis
    c, s = np.cos(angle), np.sin(angle)
    C = 1 - c
    return np.array([[c + x * x * C, x * y * C - z * s, x * z * C + y * s],
                     [y * x * C + z * s, c + y * y * C, y * z * C - x * s],
                     [z * x * C - y * s, z * y * C + x * s, c + z * z * C]])


def box_mesh():
    """Unit cube [-1, 1]^3 with per-face normals/tangents/uvs (24 vertices)."""
    pos, nrm, tng, uv, tri = [], [], [], [], []
    for axis in range(3):
        for sign in (-1.0, 1.0):
            n = np.zeros(3)
            n[axis] = sign
            u = np.zeros(3)
            u[(axis + 1) % 3] = 1.0
            v = np.cross(n, u)
            base = len(pos)
            for a, b in ((-1, -1), (1, -1), (1, 1), (-1, 1)):
                pos.append(n + a * u + b * v)
                nrm.append(n)
                tng.append(u)
                uv.append(((a + 1) / 2, (b + 1) / 2))
            tri += [[base, base + 1, base + 2], [base, base + 2, base + 3]]
    p, a, t = _merge([(np.array(pos), vertex_attributes(np.array(nrm), np.array(tng), np.array(uv)),
                       np.array(tri), 0)])
    return Mesh(p, a, t)


def instanced_scene(seed, num_instances=24, subdiv=2, ground=True, env=True):
    """Parity scene for instancing: an icosphere and a box mesh placed with
    random rotations, non-uniform scales (one mirrored, one sheared) and
    translations in the unit cube, per-instance materials (some left to the
    mesh's own), over the ground quad of config1."""
    rng = np.random.default_rng(seed)
    packer = AtlasPacker()
    mats = [packer.material((0.6, 0.55, 0.5), 0.7, 0.0)]
    uv_, f = icosphere(subdiv)
    sp, sa, st = _merge([(uv_, _sphere_attrs(uv_), f, 0)])
    meshes = [Mesh(sp, sa, st), box_mesh()]
    instances = []
    for i in range(num_instances):
        ax = rng.normal(size=3)
        R = _rotation(ax / np.linalg.norm(ax), rng.uniform(0, 2 * np.pi))
        S = np.diag(rng.uniform(0.03, 0.12, 3))
        if i == 1:
            S[0, 0] = -S[0, 0]            # mirrored
        M = R @ S
        if i == 2:
            M = M @ np.array([[1.0, 0.4, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]])  # sheared
        T = np.concatenate([M, rng.uniform(0.1, 0.9, (3, 1))], 1)
        mat = -1
        if rng.uniform() < 0.75:
            mats.append(packer.material(rng.uniform(0.2, 0.8, 3), rng.uniform(0.05, 1.0),
                                        1.0 if rng.uniform() < 0.3 else 0.0))
            mat = len(mats) - 1
        instances.append(Instance(i % 2, T, mat))
    if ground:
        g = C1_GROUND
        lo, hi = g["extent"]
        qp = np.array([[lo, g["y"], lo], [hi, g["y"], lo], [hi, g["y"], hi], [lo, g["y"], hi]], dtype=np.float64)
        qa = vertex_attributes(np.tile([0.0, 1.0, 0.0], (4, 1)), np.tile([1.0, 0.0, 0.0], (4, 1)),
                               np.array([[0, 0], [1, 0], [1, 1], [0, 1]], dtype=np.float64))
        pos, attr, tris = _merge([(qp, qa, np.array([[0, 2, 1], [0, 3, 2]]), 0)])
    else:
        pos, attr, tris = np.zeros((0, 3)), np.zeros((0, 4), np.uint32), np.zeros((0, 4), np.int32)
    return SceneData(positions=pos, attributes=attr, triangles=tris, materials=np.array(mats, dtype=np.int32),
                     tex=packer.atlas(), env=constant_env_rgbe(64, 32) if env else None,
                     camera=camera(C1_CAMERA["origin"], C1_CAMERA["target"], C1_CAMERA["fov"]),
                     name="instanced", meshes=meshes, instances=instances)


def random_scene(seed, num_spheres=6, subdiv=1, env=True):
    """Small scene for parity tests: a few spheres, a ground quad, the
    constant environment."""
    s = config1(seed=seed, num_spheres=num_spheres, subdiv=subdiv)
    if not env:
        s.env = None
    return s
