"""BVH build (render/bvh.js) and flattening (trace_pass.js:166-263): the
product's range-based builder must reproduce the oracle's literal
restatement byte for byte, satisfy hand-worked trees, and its 128-byte wide
collapse must be conservative."""
import numpy as np
import pytest

import pike_amd
from pike_amd import scenes
from oracle.oracle import OracleScene


def _soup(n, seed, spread=1.0):
    rng = scenes.Rng(seed)
    v0 = rng.uniform(n * 3).reshape(n, 3) * spread
    pos = np.repeat(v0, 3, axis=0)
    pos[1::3] += rng.uniform(n * 3, -0.05, 0.05).reshape(n, 3)
    pos[2::3] += rng.uniform(n * 3, -0.05, 0.05).reshape(n, 3)
    tri = np.zeros((n, 4), np.int32)
    tri[:, 0] = np.arange(0, 3 * n, 3)
    tri[:, 1] = tri[:, 0] + 1
    tri[:, 2] = tri[:, 0] + 2
    return pos, tri


def _data(pos, tri):
    packer = scenes.AtlasPacker()
    mats = np.array([packer.material((0.5, 0.5, 0.5), 0.5, 0.0)], np.int32)
    tri = tri.copy()
    tri[:, 3] = 0
    return scenes.SceneData(positions=pos, attributes=np.zeros((len(pos), 4), np.uint32), triangles=tri,
                            materials=mats, tex=packer.atlas())


def _compare(data):
    b = pike_amd.BVH(data.positions, data.triangles, threads=4)
    o = OracleScene(data, threads=4)
    o.build_bvh()
    n1, t1, v1, m1 = b.export_reference()
    n2, t2, v2, m2 = o.serialize()
    np.testing.assert_array_equal(n1, n2)
    np.testing.assert_array_equal(t1, t2)
    np.testing.assert_array_equal(v1.view(np.uint32), v2.view(np.uint32))
    np.testing.assert_array_equal(m1, m2)
    return b, n1, t1


@pytest.mark.parametrize("seed", [1, 2])
def test_config1_matches_oracle(seed):
    b, nodes, tris = _compare(scenes.config1(seed=seed))
    assert b.info()["binary_nodes"] == nodes.shape[0]


@pytest.mark.parametrize("n,seed", [(1, 1), (2, 2), (3, 3), (5, 4), (17, 5), (100, 6), (3000, 7)])
def test_random_soups_match_oracle(n, seed):
    _compare(_data(*_soup(n, seed)))


def test_heightfield_matches_oracle():
    pos, attr, tri = scenes.terrain(40, seed=3)
    t = np.zeros((len(tri), 4), np.int32)
    t[:, :3] = tri
    _compare(_data(pos, t))


def test_duplicate_and_coplanar_triangles_match_oracle():
    pos, tri = _soup(50, 9)
    pos = np.concatenate([pos, pos])           # every triangle twice
    tri2 = tri.copy()
    tri2[:, :3] += len(pos) // 2
    _compare(_data(pos, np.concatenate([tri, tri2])))
    pos, tri = _soup(64, 10)
    pos[:, 1] = 0.25                           # all in one plane (flat box axis)
    _compare(_data(pos, tri))


def test_two_separated_triangles_split_into_two_leaves():
    """Hand-worked bvh.js case: A in x [0,1], B in x [10,11], both spanning
    y, z in [0,1].  A(P)=46; on x, i=0: 0.5 + 6/46*1 + 46/46*1 = 1.63 < 2
    (the back area is the union of refs[0..1]); axes y and z tie and lose
    (strict <).  Children are single-triangle leaves (cost 1.5 > 1)."""
    pos = np.array([[0, 0, 0], [1, 1, 0], [0, 1, 1], [10, 0, 0], [11, 1, 0], [10, 1, 1]], np.float64)
    tri = np.array([[0, 1, 2, 0], [3, 4, 5, 0]], np.int32)
    b, nodes, tris = _compare(_data(pos, tri))
    assert nodes.shape[0] == 3
    assert nodes[0, 0] == 1 and nodes[0, 1] == -1           # root: left child 1, interior
    assert nodes[1, 0] == -1 and nodes[1, 1] == (1 << 24) | 0
    assert nodes[2, 0] == -1 and nodes[2, 1] == (1 << 24) | 1
    np.testing.assert_array_equal(tris[:, :3], [[0, 1, 2], [3, 4, 5]])
    # normalizer: center (5.5, 0.5, 0.5), scale 2/11; A's box x in [-1, -9/11]
    # snorm16 stores max*32767 (trunc) low, min*32768 high: -0.8181..*32768 -> -26810
    lo16, hi16 = nodes[1, 2] & 0xFFFF, (nodes[1, 2] >> 16) & 0xFFFF
    assert np.int16(np.uint16(lo16)) == -26810 and np.int16(np.uint16(hi16)) == -32768


def test_identical_triangles_make_one_leaf():
    """Two coincident triangles: every split costs 2.5 > N=2 -> one leaf."""
    pos = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]] * 2, np.float64)
    tri = np.array([[0, 1, 2, 0], [3, 4, 5, 0]], np.int32)
    b, nodes, tris = _compare(_data(pos, tri))
    assert nodes.shape[0] == 1 and nodes[0, 1] == (2 << 24)


def test_leaf_triangles_are_a_permutation():
    data = scenes.config1(seed=4)
    b = pike_amd.BVH(data.positions, data.triangles)
    _, tris, _, _ = b.export_reference()
    key = lambda a: np.sort(a.view([("", a.dtype)] * 4).ravel())  # noqa: E731
    np.testing.assert_array_equal(key(np.ascontiguousarray(tris)), key(np.ascontiguousarray(data.triangles)))


def _decode_child_boxes(n):
    s = np.array([2.0 ** (int(e) - 127) for e in n["exp"]])
    org = n["origin"].astype(np.float64)
    lo = np.stack([org[0] + n["qlox"] * s[0], org[1] + n["qloy"] * s[1], org[2] + n["qloz"] * s[2]], 1)
    hi = np.stack([org[0] + n["qhix"] * s[0], org[1] + n["qhiy"] * s[1], org[2] + n["qhiz"] * s[2]], 1)
    return lo, hi


def test_wide_layout_is_conservative_and_complete():
    data = scenes.config1(seed=5)
    b = pike_amd.BVH(data.positions, data.triangles)
    nodes, tris, tidx = b.export_wide()
    # leafmask bit c = (meta[c] != 0) over the node's children, zero past them
    bits = (nodes["meta"] != 0) & (np.arange(8)[None, :] < nodes["nchild"][:, None])
    np.testing.assert_array_equal(nodes["leafmask"], (bits * (1 << np.arange(8))).sum(1))
    assert (nodes["leafmask"] != 0).any() and (nodes["leafmask"] != 255).any()
    # slot-order runs: internal children are the nodes ibase, ibase + 1, ...
    # and leaf children's triangle ranges follow each other from tbase, so the
    # single-level traversal finds child c as ibase + (internal slots below
    # c) or tbase + (leaf counts below c)
    for n in nodes:
        k = int(n["nchild"])
        internal = [c for c in range(k) if n["meta"][c] == 0]
        leaves = [c for c in range(k) if n["meta"][c] != 0]
        assert [int(n["child"][c]) for c in internal] == [int(n["ibase"]) + i for i in range(len(internal))]
        first = int(n["tbase"])
        for c in leaves:
            assert int(n["child"][c]) == first
            first += int(n["meta"][c])
    _, _, verts, _ = b.export_reference()
    V = verts[:, :3].astype(np.float64)
    seen = np.zeros(len(tris), np.int64)
    # every subtree's vertices lie inside each quantized child box (+ margin)
    def subtree_tris(ni):
        out, st = [], [ni]
        while st:
            n = nodes[st.pop()]
            for c in range(n["nchild"]):
                if n["meta"][c]:
                    out.append(np.arange(n["child"][c], n["child"][c] + n["meta"][c]))
                else:
                    st.append(n["child"][c])
        return np.concatenate(out)

    for ni in range(0, len(nodes), max(1, len(nodes) // 300)):
        n = nodes[ni]
        lo, hi = _decode_child_boxes(n)
        for c in range(n["nchild"]):
            ts = (np.arange(n["child"][c], n["child"][c] + n["meta"][c]) if n["meta"][c]
                  else subtree_tris(n["child"][c]))
            pts = V[tidx[ts, :3].ravel()]
            assert np.all(pts >= lo[c] + 1e-6) and np.all(pts <= hi[c] - 1e-6)
    for n in nodes:
        for c in range(n["nchild"]):
            if n["meta"][c]:
                seen[n["child"][c]:n["child"][c] + n["meta"][c]] += 1
    assert np.all(seen == 1)
    assert sorted(tris["id"].tolist()) == list(range(data.num_triangles))
    # the triangle records hold v0 and the f32 edges of the normalized vertices
    i0, i1, i2 = tidx[:, 0], tidx[:, 1], tidx[:, 2]
    np.testing.assert_array_equal(tris["v0"], verts[i0, :3])
    np.testing.assert_array_equal(tris["e1"], verts[i1, :3] - verts[i0, :3])
    np.testing.assert_array_equal(tris["e2"], verts[i2, :3] - verts[i0, :3])
    assert np.all(nodes["nchild"] >= 1) and np.all(nodes["nchild"] <= 8)


@pytest.mark.parametrize("spatial", [False, True])
def test_wide4_layout_is_conservative_and_complete(spatial):
    """The 4-wide layout: node i stands for 8-wide node i (its children, or
    two sub-nodes splitting its slots at one cut), every box holds its
    subtree, each triangle record is reached exactly once, and children come
    from ibase/tbase alone."""
    data = scenes.config1(seed=6)
    b = pike_amd.BVH(data.positions, data.triangles, spatial_splits=spatial, sah_collapse=True)
    n8, tris, tidx = b.export_wide()
    n4 = b.export_wide4()
    split = n8["nchild"] > 4
    assert len(n4) == len(n8) + 2 * int(split.sum()) and split.any() and (~split).any()
    assert np.all(n4["nchild"] >= 1) and np.all(n4["nchild"] <= 4)
    bits = (n4["meta"] != 0) & (np.arange(4)[None, :] < n4["nchild"][:, None])
    np.testing.assert_array_equal(n4["leafmask"], (bits * (1 << np.arange(4))).sum(1))
    # nodes without a split keep the 8-wide node's slots, frame and runs
    for i in np.flatnonzero(~split)[:200]:
        k = int(n8["nchild"][i])
        assert n4["nchild"][i] == k and n4["ibase"][i] == n8["ibase"][i] and n4["tbase"][i] == n8["tbase"][i]
        for f in ("meta", "qlox", "qhix", "qloy", "qhiy", "qloz", "qhiz"):
            np.testing.assert_array_equal(n4[f][i][:k], n8[f][i][:k])

    def kids(n):
        out, ni, ti = [], int(n["ibase"]), int(n["tbase"])
        for c in range(int(n["nchild"])):
            if n["meta"][c]:
                out.append(("leaf", ti, int(n["meta"][c])))
                ti += int(n["meta"][c])
            else:
                out.append(("node", ni, 0))
                ni += 1
        return out

    _, _, verts, _ = b.export_reference()
    V = verts[:, :3].astype(np.float64)
    seen = np.zeros(len(tris), np.int64)

    def subtree_tris(ni):
        out, st = [], [ni]
        while st:
            for kind, x, cnt in kids(n4[st.pop()]):
                if kind == "leaf":
                    out.append(np.arange(x, x + cnt))
                else:
                    st.append(x)
        return np.concatenate(out)

    st = [0]
    visited = 0
    while st:
        ni = st.pop()
        visited += 1
        n = n4[ni]
        lo, hi = _decode_child_boxes(n)
        for c, (kind, x, cnt) in enumerate(kids(n)):
            if kind == "leaf":
                seen[x:x + cnt] += 1
                ts = np.arange(x, x + cnt)
            else:
                st.append(x)
                ts = subtree_tris(x) if ni % 7 == 0 else None
            # (spatial splits clip references: their vertices may leave the box)
            if ts is not None and not spatial:
                pts = V[tidx[ts, :3].ravel()]
                assert np.all(pts >= lo[c] + 1e-6) and np.all(pts <= hi[c] - 1e-6)
    assert visited == len(n4)
    assert np.all(seen == 1)


def test_bvh_build_rejects_bad_input():
    pos = np.zeros((3, 3))
    with pytest.raises(pike_amd.PikeError) as e:
        pike_amd.BVH(pos, np.array([[0, 1, 5, 0]], np.int32))
    assert e.value.code == -1 and "out of range" in str(e.value)
    with pytest.raises(pike_amd.PikeError):
        pike_amd.BVH(pos, np.zeros((0, 4), np.int32))
    for bad in (np.nan, np.inf, -np.inf):
        p = np.zeros((4, 3))
        p[1, 0] = 1.0
        p[2, 1] = 1.0
        p[3, 2] = bad               # vertex 3: referenced by the second triangle only
        t = np.array([[0, 1, 2, 0], [0, 1, 3, 0]], np.int32)
        with pytest.raises(pike_amd.PikeError) as e:
            pike_amd.BVH(p, t)
        assert e.value.code == -1 and "non-finite" in str(e.value) and "triangle 1" in str(e.value)
        pike_amd.BVH(p, t[:1])      # an unreferenced non-finite vertex is not an error


def test_build_out_of_host_memory_fails_cleanly():
    """A build that runs out of host memory returns PK_E_OOM with a message
    instead of ending the process (no exception crosses the C ABI): a child
    process limits its address space to 48 MB above what it holds, then
    builds a ~1M-triangle SBVH, and afterwards a small scene again."""
    import subprocess
    import sys
    code = r"""
import resource, sys
import numpy as np
import pike_amd
from pike_amd import scenes
pos, _, tri = scenes.terrain(700, seed=1)
t = np.zeros((len(tri), 4), np.int32)
t[:, :3] = tri
pos = np.ascontiguousarray(pos, np.float64)
small_p, _, small_t = scenes.terrain(8, seed=1)
st = np.zeros((len(small_t), 4), np.int32)
st[:, :3] = small_t
pike_amd.lib()
vm = [int(l.split()[1]) * 1024 for l in open("/proc/self/status") if l.startswith("VmSize:")][0]
resource.setrlimit(resource.RLIMIT_AS, (vm + (48 << 20), resource.getrlimit(resource.RLIMIT_AS)[1]))
try:
    pike_amd.BVH(pos, t, threads=4, spatial_splits=True, sah_collapse=True)
    print("built")
except pike_amd.PikeError as e:
    print("code", e.code, str(e))
b = pike_amd.BVH(small_p, st, threads=2)
print("small", b.info()["leaf_triangles"])
"""
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300,
                       cwd=str(__import__("pathlib").Path(__file__).resolve().parents[1]))
    assert r.returncode == 0, r.stderr[-2000:]
    assert "code -3" in r.stdout and "pk_bvh_build" in r.stdout, r.stdout
    assert "small 128" in r.stdout, r.stdout


def _exports(b):
    return [a.view(np.uint8) for a in b.export_wide() + b.export_reference()]


@pytest.mark.parametrize("spatial", [False, True])
def test_build_does_not_depend_on_thread_count(spatial):
    """Node passes split over threads (unions and sweeps in parts) and
    subtrees handed to threads as they free up give the serial tree: every
    byte of both layouts equal for 1, 2, 3 and 8 threads.  Sizes pass the
    threshold of split passes (65,536 references) at the root; the terrain
    has many equal centroid coordinates (ties), the slivers spatial splits."""
    pos, _, tri = scenes.terrain(200, seed=5)       # 80,000 triangles
    t = np.zeros((len(tri), 4), np.int32)
    t[:, :3] = tri
    cases = [(pos, t, (1, 2, 3, 8))] + ([_slivers(70000, 4) + ((1, 8),)] if spatial else [])
    for p, tt, counts in cases:
        ref = None
        for threads in counts:
            b = pike_amd.BVH(p, tt, threads=threads, spatial_splits=spatial, sah_collapse=True)
            ex = _exports(b)
            if ref is None:
                ref = ex
                assert b.info()["binary_nodes"] > 2 * 32768 // 8
                continue
            for x, y in zip(ref, ex):
                np.testing.assert_array_equal(x, y)


def _slivers(n, seed):
    """Long thin triangles (one ~0.3-0.6 edge, ~0.002 wide) at random
    orientations in the unit cube: the case spatial splits are for."""
    rng = np.random.default_rng(seed)
    a = rng.uniform(0.2, 0.8, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    w = np.cross(d, rng.normal(size=(n, 3)))
    w /= np.linalg.norm(w, axis=1, keepdims=True)
    b = a + d * rng.uniform(0.3, 0.6, (n, 1))
    c = a + w * 0.002
    pos = np.stack([a, b, c], 1).reshape(-1, 3)
    tri = np.concatenate([np.arange(3 * n).reshape(-1, 3), np.zeros((n, 1), np.int64)], 1).astype(np.int32)
    return pos, tri


def _reached_ids(nodes, tris, o, d, tmax):
    """Triangle ids in the leaves whose quantized boxes the ray segment
    [0, tmax] touches (a float64 model of the wide traversal without the
    near-to-far order: closest-hit traversal can only find a hit through
    such leaves)."""
    inv = np.where(np.abs(d) > 1e-30, 1.0 / np.where(d == 0, 1, d), np.inf)
    ids, stack = set(), [0]
    while stack:
        n = nodes[stack.pop()]
        lo, hi = _decode_child_boxes(n)
        for c in range(n["nchild"]):
            with np.errstate(invalid="ignore"):
                t0, t1 = (lo[c] - o) * inv, (hi[c] - o) * inv
            tn = np.nanmax(np.minimum(t0, t1))
            tf = np.nanmin(np.maximum(t0, t1))
            if tn <= tf and tf >= 0 and tn <= tmax * (1 + 1e-9) + 1e-9:
                if n["meta"][c]:
                    ids.update(tris["id"][n["child"][c]:n["child"][c] + n["meta"][c]].tolist())
                else:
                    stack.append(n["child"][c])
    return ids


def test_spatial_splits_keep_every_closest_hit_reachable():
    pos, tri = _slivers(1500, 3)
    data = _data(pos, tri)
    b = pike_amd.BVH(data.positions, data.triangles, spatial_splits=True)
    info = b.info()
    assert info["leaf_triangles"] > data.num_triangles  # fragments in several leaves
    nodes, wt, tidx = b.export_wide()
    assert set(wt["id"].tolist()) == set(range(data.num_triangles))
    o = OracleScene(data, threads=8)
    o.build_bvh()
    nm = b.normalizer()
    rays = scenes.incoherent_rays(3000, nm, seed=4)
    ids, t, _, _ = o.trace(rays, "brute")
    hits = np.flatnonzero(ids >= 0)
    assert hits.size > 300
    for r in hits:
        reach = _reached_ids(nodes, wt, rays[r, :3].astype(np.float64), rays[r, 4:7].astype(np.float64), float(t[r]))
        assert int(ids[r]) in reach, r
    # the reference export holds every leaf reference; its triangles are the scene's
    _, rtris, _, _ = b.export_reference()
    assert rtris.shape[0] == info["leaf_triangles"]
    assert {tuple(x) for x in rtris.tolist()} == {tuple(x) for x in data.triangles.tolist()}


def test_spatial_splits_reduce_sah_cost_on_slivers():
    """bvh.js's cost model on the same triangles: with spatial splits the
    tree traverses fewer boxes per ray (measured with the model above)."""
    pos, tri = _slivers(800, 5)
    data = _data(pos, tri)
    plain = pike_amd.BVH(data.positions, data.triangles)
    sbvh = pike_amd.BVH(data.positions, data.triangles, spatial_splits=True)
    rays = scenes.incoherent_rays(300, plain.normalizer(), seed=6)
    o = OracleScene(data, threads=8)
    o.build_bvh()
    ids, t, _, _ = o.trace(rays, "brute")
    tmax = np.where(ids >= 0, t, 10.0)

    def tested(b):
        nodes, wt, _ = b.export_wide()
        return sum(len(_reached_ids(nodes, wt, rays[r, :3].astype(np.float64), rays[r, 4:7].astype(np.float64),
                                    float(tmax[r]))) for r in range(len(rays)))

    assert tested(sbvh) < tested(plain)


@pytest.mark.parametrize("spatial", [False, True])
def test_sah_collapse_layout_and_reachability(spatial):
    """PK_BVH_SAH_COLLAPSE: every triangle in exactly one leaf record (or its
    fragments, with spatial splits), every child reference in range, merged
    leaves of at most 4 triangles, and every closest hit reachable through
    the quantized boxes its ray passes."""
    pos, tri = _slivers(800, 11) if spatial else (None, None)
    data = _data(pos, tri) if spatial else scenes.config1(seed=12, num_spheres=20)
    b = pike_amd.BVH(data.positions, data.triangles, spatial_splits=spatial, sah_collapse=True)
    g = pike_amd.BVH(data.positions, data.triangles, spatial_splits=spatial)
    nodes, wt, _ = b.export_wide()
    assert b.info()["leaf_triangles"] == g.info()["leaf_triangles"]
    assert nodes["meta"].max() <= max(4, g.info()["largest_leaf"])
    if not spatial:
        np.testing.assert_array_equal(np.sort(wt["id"]), np.arange(data.num_triangles))
    else:
        assert set(wt["id"].tolist()) == set(range(data.num_triangles))
    for n in nodes:
        for c in range(int(n["nchild"])):
            ch, m = int(n["child"][c]), int(n["meta"][c])
            assert (ch + m <= len(wt)) if m else (ch < len(nodes))
    o = OracleScene(data, threads=8)
    o.build_bvh()
    rays = scenes.incoherent_rays(2000, b.normalizer(), seed=13)
    ids, t, _, _ = o.trace(rays, "brute")
    hits = np.flatnonzero(ids >= 0)
    assert hits.size > 100
    for r in hits:
        reach = _reached_ids(nodes, wt, rays[r, :3].astype(np.float64), rays[r, 4:7].astype(np.float64), float(t[r]))
        assert int(ids[r]) in reach, r
