"""Closest-hit and any-hit parity of the gfx950 traversal (pk_trace_rays)
against the oracle's canonical closest hit: argmin (t, original triangle id)
over every triangle with the reference's f32 Möller–Trumbore
(trace.wgsl:97-114).  Bit-exact ids, t bits and bary masks."""
import numpy as np
import pytest

import pike_amd
from pike_amd import scenes
from oracle.oracle import OracleScene

pytestmark = pytest.mark.gpu


def _setup(ctx, data):
    bvh = pike_amd.BVH(data.positions, data.triangles)
    scene = pike_amd.Scene(ctx, data, bvh)
    orc = OracleScene(data)
    orc.build_bvh()
    return bvh, scene, orc


def _camera_rays(n, seed):
    """Rays from outside the scene towards it (camera-like)."""
    rng = scenes.Rng(seed)
    o = np.stack([rng.uniform(n, -0.3, 0.3), rng.uniform(n, 0.2, 0.8), np.full(n, 1.8)], 1)
    tgt = np.stack([rng.uniform(n, -0.8, 0.8), rng.uniform(n, -0.6, 0.2), rng.uniform(n, -0.8, 0.8)], 1)
    d = tgt - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((n, 8), np.float32)
    r[:, 0:3] = o
    r[:, 4:7] = d
    return r


def _assert_same(ids_g, t_g, b_g, ids_o, t_o, b_o):
    bad = np.flatnonzero((ids_g != ids_o) | (t_g.view(np.uint32) != t_o.view(np.uint32)) | (b_g != b_o))
    assert bad.size == 0, f"{bad.size} rays differ, first {bad[:5]}: gpu {ids_g[bad[:5]]} {t_g[bad[:5]]} " \
                          f"oracle {ids_o[bad[:5]]} {t_o[bad[:5]]}"


@pytest.mark.parametrize("seed", [1, 7])
def test_closest_hit_config1_incoherent(ctx, seed):
    data = scenes.config1(seed=seed)
    bvh, scene, orc = _setup(ctx, data)
    rays = scenes.incoherent_rays(50_000, bvh.normalizer(), seed=seed + 100)
    ids, t, bary = ctx.trace_rays(scene, rays)
    ids_o, t_o, b_o, _ = orc.trace(rays, "canonical")
    assert (ids >= 0).mean() > 0.3
    _assert_same(ids, t, bary, ids_o, t_o, b_o)


def test_closest_hit_camera_rays_vs_brute_force(ctx):
    data = scenes.config1(seed=3, num_spheres=12)
    bvh, scene, orc = _setup(ctx, data)
    rays = _camera_rays(4000, 9)
    ids, t, bary = ctx.trace_rays(scene, rays)
    ids_b, t_b, b_b, _ = orc.trace(rays, "brute")
    _assert_same(ids, t, bary, ids_b, t_b, b_b)


def test_any_hit_matches_occlusion(ctx):
    data = scenes.config1(seed=2)
    bvh, scene, orc = _setup(ctx, data)
    rays = scenes.incoherent_rays(50_000, bvh.normalizer(), seed=77)
    ids, t, _ = ctx.trace_rays(scene, rays, any_hit=True)
    ids_o, _, _, _ = orc.trace(rays, "canonical", any_hit=True)
    np.testing.assert_array_equal(ids >= 0, ids_o >= 0)
    # a reported occluder really is hit (epsilon 1e-7 semantics)
    hit = np.flatnonzero(ids >= 0)
    assert np.all(t[hit] > 1e-7) and np.all(t[hit] < 1e5)


def test_triangle_soup_against_canonical(ctx):
    data = scenes.config3(seed=11, num_tris=200_000)
    bvh, scene, orc = _setup(ctx, data)
    rays = scenes.incoherent_rays(100_000, bvh.normalizer(), seed=12)
    ids, t, bary = ctx.trace_rays(scene, rays)
    ids_o, t_o, b_o, _ = orc.trace(rays, "canonical")
    _assert_same(ids, t, bary, ids_o, t_o, b_o)
    occ, _, _ = ctx.trace_rays(scene, rays, any_hit=True)
    occ_o, _, _, _ = orc.trace(rays, "canonical", any_hit=True)
    np.testing.assert_array_equal(occ >= 0, occ_o >= 0)


def test_edge_cases(ctx):
    data = scenes.config1(seed=4, num_spheres=4)
    bvh, scene, orc = _setup(ctx, data)
    rays = np.zeros((6, 8), np.float32)
    rays[0, 4:7] = (0, 0, 1)          # axis-aligned, zero components
    rays[1, 4:7] = (0, -1, 0)         # straight down at the ground
    rays[2, 0:3] = (0, 5, 0)          # far outside, pointing away
    rays[2, 4:7] = (0, 1, 0)
    rays[3, 4:7] = (1e-30, 1, 0)      # tiny component
    rays[4, 0:3] = (0.1, 0.1, 0.1)
    rays[4, 4:7] = (-0.577, -0.577, -0.577)
    rays[5, 4:7] = (0, 0, -1)
    ids, t, bary = ctx.trace_rays(scene, rays)
    ids_b, t_b, b_b, _ = orc.trace(rays, "brute")
    _assert_same(ids, t, bary, ids_b, t_b, b_b)
    # empty input is a no-op
    e = ctx.trace_rays(scene, np.zeros((0, 8), np.float32))
    assert e[0].size == 0


def test_node_counters_match_between_runs(ctx):
    """Counting kernels report deterministic totals (the traversal is
    deterministic per ray)."""
    data = scenes.config1(seed=5)
    bvh, scene, orc = _setup(ctx, data)
    host_rays = scenes.incoherent_rays(20_000, bvh.normalizer(), seed=5)
    n = host_rays.shape[0]
    rays = pike_amd.DeviceBuffer.from_array(ctx, host_rays)
    ids = pike_amd.DeviceBuffer(ctx, 4 * n)
    t = pike_amd.DeviceBuffer(ctx, 4 * n)
    b = pike_amd.DeviceBuffer(ctx, 4 * n)
    stats = []
    for _ in range(2):
        ctx.trace_rays_device(scene, rays.ptr, n, False, ids.ptr, t.ptr, b.ptr, True)
        stats.append(ctx.trace_stats())
    assert stats[0] == stats[1] and stats[0][0] > n
    ids_o, t_o, b_o, _ = orc.trace(host_rays, "canonical")
    _assert_same(ids.read(np.int32, n), t.read(np.float32, n), b.read(np.uint32, n), ids_o, t_o, b_o)
    # the timed (uncounted) kernel gives the same answers
    ms = ctx.time_trace_rays(scene, rays.ptr, n, False, ids.ptr, t.ptr, b.ptr, 2)
    assert ms > 0
    _assert_same(ids.read(np.int32, n), t.read(np.float32, n), b.read(np.uint32, n), ids_o, t_o, b_o)


@pytest.mark.parametrize("refill", [1, 64])
def test_overlapping_boxes_deep_stacks(ctx, refill):
    """Large random triangles spanning the cube: every node's children
    overlap, so rays push far more entries than the LDS part of the stack
    holds (the global spill area is used) and most lanes carry deferred leaf
    children at once; under the extreme refill thresholds (a lane refilled
    as soon as it is idle, or only once the whole wave is) the hits stay the
    canonical ones."""
    rng = np.random.default_rng(11)
    n = 3000
    pos = rng.uniform(0.0, 1.0, (3 * n, 3))
    tri = np.concatenate([np.arange(3 * n).reshape(-1, 3), np.zeros((n, 1), np.int64)], 1).astype(np.int32)
    from .test_bvh import _data
    data = _data(pos, tri)
    bvh, scene, orc = _setup(ctx, data)
    rays = scenes.incoherent_rays(20_000, bvh.normalizer(), seed=12)
    old = ctx.tuning()
    try:
        for cls in ("rays_closest", "rays_any"):
            ctx.set_tuning(cls, old[cls]["waves_per_cu"], refill)
        ids, t, bary = ctx.trace_rays(scene, rays)
        occ, _, _ = ctx.trace_rays(scene, rays, any_hit=True)
    finally:
        for cls in ("rays_closest", "rays_any"):
            ctx.set_tuning(cls, old[cls]["waves_per_cu"], old[cls]["refill_min"])
    ids_o, t_o, b_o, _ = orc.trace(rays, "canonical")
    assert (ids_o >= 0).mean() > 0.5
    _assert_same(ids, t, bary, ids_o, t_o, b_o)
    occ_o, _, _, _ = orc.trace(rays, "canonical", any_hit=True)
    np.testing.assert_array_equal(occ >= 0, occ_o >= 0)


@pytest.mark.parametrize("leaf16", [1, 16])
def test_stacked_duplicates_large_leaves(ctx, leaf16):
    """Stacks of identical triangles (equal t for every copy, so the lowest
    id must win) that no split separates: leaves of up to 255 triangles and
    leaf phases whose items outnumber the wave's 64 lanes several times, one
    lane's items spanning several rounds.  Closest hits equal the canonical
    ones (ties resolved by id) and occlusion the oracle's, with the leaf
    phase triggered by a single deferring lane or only by all of them."""
    rng = np.random.default_rng(21)
    pos, tri = [], []
    stacks = []  # centroid of each stack
    for k in (1, 7, 33, 64, 130, 255, 300):
        c = rng.uniform(0.1, 0.9, 3)
        v = c + rng.uniform(-0.04, 0.04, (3, 3))
        stacks.append(v.mean(0))
        for _ in range(k):
            base = len(pos)
            pos.extend(v)
            tri.append((base, base + 1, base + 2, 0))
    bg = rng.uniform(0.0, 1.0, (600, 3))
    for i in range(200):
        base = len(pos)
        pos.extend(bg[3 * i:3 * i + 3])
        tri.append((base, base + 1, base + 2, 0))
    from .test_bvh import _data
    data = _data(np.asarray(pos, np.float64), np.asarray(tri, np.int32))
    bvh, scene, orc = _setup(ctx, data)
    assert bvh.info()["largest_leaf"] >= 64
    # rays through the stacks' centroids from random origins, and incoherent ones
    cents = np.asarray(stacks)
    n = 30_000
    org = rng.uniform(-0.5, 1.5, (n, 3))
    tgt = cents[rng.integers(0, len(cents), n)] + rng.normal(0, 0.01, (n, 3))
    norm = bvh.normalizer()
    o = (org - norm[:3]) * norm[3]
    d = tgt - org
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = o
    rays[:, 4:7] = d
    rays = np.concatenate([rays, scenes.incoherent_rays(10_000, norm, seed=22)])
    old = ctx.tuning()
    try:
        for cls in ("rays_closest", "rays_any"):
            sh = old[cls]
            ctx.set_tuning(cls, sh["waves_per_cu"], sh["refill_min"], sh["stack_lds"], sh["waves_per_block"], leaf16,
                           sh["node_width"])
        ids, t, bary = ctx.trace_rays(scene, rays)
        occ, _, _ = ctx.trace_rays(scene, rays, any_hit=True)
    finally:
        for cls in ("rays_closest", "rays_any"):
            sh = old[cls]
            ctx.set_tuning(cls, sh["waves_per_cu"], sh["refill_min"], sh["stack_lds"], sh["waves_per_block"],
                           sh["leaf_phase16"], sh["node_width"])
    ids_o, t_o, b_o, _ = orc.trace(rays, "canonical")
    assert (ids_o >= 0).mean() > 0.5
    _assert_same(ids, t, bary, ids_o, t_o, b_o)
    occ_o, _, _, _ = orc.trace(rays, "canonical", any_hit=True)
    np.testing.assert_array_equal(occ >= 0, occ_o >= 0)
