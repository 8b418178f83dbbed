"""Pins on the oracle itself (CPU only):
* its canonical closest hit equals brute force over every triangle;
* every disagreement of the literal reference traversal (trace.wgsl over the
  snorm16-quantized 20-byte nodes) with the canonical hit is explained by a
  quantized box that excludes the true hit point (boxes are truncated toward
  zero, trace_pass.js:213-218) or by an equal-t tie;
* the FP32 restatement converges to the FP64 oracle of the same estimator
  (common random numbers) within 1e-3 mean relative L1.
"""
import numpy as np
import pytest

from pike_amd import scenes
import pike_amd
from oracle.oracle import OracleScene


@pytest.fixture(scope="module")
def c1():
    data = scenes.config1(seed=1)
    o = OracleScene(data, threads=8)
    o.build_bvh()
    return data, o


def test_canonical_equals_brute_force(c1):
    data, o = c1
    nm = o.serialize()[3]
    rays = scenes.incoherent_rays(4000, nm, seed=8)
    a = o.trace(rays, "brute")
    b = o.trace(rays, "canonical")
    for x, y in zip(a[:3], b[:3]):
        np.testing.assert_array_equal(x.view(np.uint32), y.view(np.uint32))
    occ_b = o.trace(rays, "brute", any_hit=True)[0] >= 0
    occ_c = o.trace(rays, "canonical", any_hit=True)[0] >= 0
    np.testing.assert_array_equal(occ_b, occ_c)


def _unorm16_box(word):
    lo = np.int16(np.uint16(word & 0xFFFF))
    hi = np.int16(np.uint16((word >> 16) & 0xFFFF))
    return max(hi / 32767.0, -1.0), max(lo / 32767.0, -1.0)   # (min plane, max plane)


def test_literal_mismatches_are_explained(c1):
    data, o = c1
    nodes, tris, verts, nm = o.serialize()
    rays = scenes.incoherent_rays(60_000, nm, seed=21)
    lit = o.trace(rays, "literal")
    can = o.trace(rays, "canonical")
    bad = np.flatnonzero(lit[0] != can[0])
    assert bad.size < 0.001 * len(rays)
    # parent map and leaf of every leaf-order triangle
    parent = np.full(len(nodes), -1)
    for i, n in enumerate(nodes):
        if n[1] == -1:
            parent[n[0]] = i
            parent[n[0] + 1] = i
    leaf_of = np.empty(len(tris), np.int64)
    for i, n in enumerate(nodes):
        if n[1] > -1:
            base, cnt = n[1] & 0xFFFFFF, n[1] >> 24
            leaf_of[base:base + cnt] = i
    key = {tuple(t[:3]): k for k, t in enumerate(tris)}
    explained = 0
    for r in bad:
        o_, d_ = rays[r, :3].astype(np.float64), rays[r, 4:7].astype(np.float64)
        if can[0][r] < 0:
            # literal found a hit the canonical argmin rejects: only possible as a tie
            assert lit[1][r] == can[1][r] or lit[1][r] >= can[1][r]
            explained += 1
            continue
        t_ = float(can[1][r])
        p = o_ + t_ * d_
        tri = tuple(data.triangles[can[0][r], :3])
        n = leaf_of[key[tri]]
        outside = False
        while n >= 0:
            for ax in range(3):
                lo, hi = _unorm16_box(int(np.uint32(nodes[n, 2 + ax])))
                if p[ax] < lo or p[ax] > hi:
                    outside = True
            n = parent[n]
        tie = lit[0][r] >= 0 and lit[1][r] == can[1][r]
        assert outside or tie, f"ray {r}: literal {lit[0][r]} {lit[1][r]} canonical {can[0][r]} {can[1][r]}"
        explained += 1
    assert explained == bad.size


def test_fp32_restatement_converges_to_fp64(c1):
    """Common random numbers: same seeds and hash, u32 -> real per precision.
    Mean relative L1 of the running means after 256 samples per pixel."""
    data, o = c1
    b = pike_amd.BVH(data.positions, data.triangles)
    cam = scenes.normalized_camera(data.camera, b.normalizer())
    W, H, depth, spp = 24, 16, 4, 256
    a32, _, _ = o.render(cam, W, H, depth, 0, spp, precision=32)
    a64, _, _ = o.render(cam, W, H, depth, 0, spp, precision=64)
    rel = np.abs(a32[..., :3] - a64[..., :3]).sum() / np.abs(a64[..., :3]).sum()
    assert rel < 1e-3, rel
