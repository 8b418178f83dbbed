"""Instancing (two-level BVH, pk_bvh_build_instanced) on the CPU: the builder's
normalizer and validation, and the oracle's instanced traversal pinned
against brute force and against the baked scene (what scene.js:159-170
builds), whose hits a correct scene -> mesh map must reproduce."""
import numpy as np
import pytest

import pike_amd
from pike_amd import scenes
from oracle.oracle import OracleScene


@pytest.mark.parametrize("ground", [True, False])
def test_normalizer_equals_the_baked_scenes(ground):
    d = scenes.instanced_scene(3, ground=ground)
    b = pike_amd.BVH.for_scene(d)
    bk = d.bake()
    np.testing.assert_array_equal(b.normalizer(), pike_amd.BVH(bk.positions, bk.triangles).normalizer())
    assert bk.num_triangles == d.baked_triangles
    info = b.info()
    # one triangle record per mesh triangle used (not per placement)
    used = {i.mesh for i in d.instances}
    assert info["leaf_triangles"] == d.num_triangles + sum(d.meshes[m].triangles.shape[0] for m in used)


def test_config2_instanced_bakes_to_the_same_vertices():
    d = scenes.config2(cells=32, num_spheres=12, subdiv=2, env_size=(32, 16), instanced=True)
    flat = scenes.config2(cells=32, num_spheres=12, subdiv=2, env_size=(32, 16))
    bk = d.bake()
    np.testing.assert_array_equal(bk.positions, flat.positions)
    np.testing.assert_array_equal(bk.triangles, flat.triangles)
    np.testing.assert_array_equal(pike_amd.BVH.for_scene(d).normalizer(),
                                  pike_amd.BVH(flat.positions, flat.triangles).normalizer())


@pytest.mark.parametrize("seed,ground", [(3, True), (4, False)])
def test_oracle_instanced_canonical_equals_brute_force(seed, ground):
    d = scenes.instanced_scene(seed, num_instances=12, ground=ground)
    o = OracleScene(d, threads=8)
    o.build_bvh()
    nm = pike_amd.BVH.for_scene(d).normalizer()
    rays = scenes.incoherent_rays(6000, nm, seed=seed)
    for any_hit in (False, True):
        i1, t1, b1, _ = o.trace(rays, "canonical", any_hit=any_hit)
        i2, t2, b2, _ = o.trace(rays, "brute", any_hit=any_hit)
        if any_hit:
            np.testing.assert_array_equal(i1 >= 0, i2 >= 0)
        else:
            assert (i1 >= 0).mean() > 0.1
            np.testing.assert_array_equal(i1, i2)
            np.testing.assert_array_equal(t1.view(np.uint32), t2.view(np.uint32))
            np.testing.assert_array_equal(b1, b2)


@pytest.mark.parametrize("seed,ground", [(3, True), (5, False)])
def test_oracle_instanced_hits_match_the_baked_scene(seed, ground):
    """Rotations, non-uniform and mirrored scales and a shear: the same
    triangles are hit (baked ids), at the same distance up to f32 rounding
    in two coordinate frames (grazing rays amplify it)."""
    d = scenes.instanced_scene(seed, ground=ground)
    o = OracleScene(d, threads=8)
    o.build_bvh()
    bk = d.bake()
    ob = OracleScene(bk, threads=8)
    ob.build_bvh()
    nm = pike_amd.BVH.for_scene(d).normalizer()
    rays = scenes.incoherent_rays(20000, nm, seed=seed + 10)
    i1, t1, _, _ = o.trace(rays, "canonical")
    i2, t2, _, _ = ob.trace(rays, "canonical")
    assert (i1 >= 0).mean() > 0.2
    assert (i1 == i2).mean() > 0.9995
    # the scene's own triangles are traced in the scene's space, as baked:
    # their hits are bit-identical
    own = (i2 >= 0) & (i2 < d.num_triangles) & (i1 == i2)
    if d.num_triangles:
        assert own.sum() > 100
        np.testing.assert_array_equal(t1[own].view(np.uint32), t2[own].view(np.uint32))
    both = (i1 >= 0) & (i1 == i2)
    rel = np.abs(t1[both].astype(np.float64) - t2[both]) / t2[both]
    assert np.median(rel) < 1e-6 and np.quantile(rel, 0.99) < 1e-4, (np.median(rel), np.quantile(rel, 0.99))
    o1, _, _, _ = o.trace(rays, "canonical", any_hit=True)
    o2, _, _, _ = ob.trace(rays, "canonical", any_hit=True)
    assert ((o1 >= 0) == (o2 >= 0)).mean() > 0.9995


def test_bake_maps_tangent_frames():
    """Tangents follow the linear part, normals its inverse transpose: the
    baked normal stays perpendicular to the baked tangent."""
    d = scenes.instanced_scene(6, num_instances=6)
    bk = d.bake()
    a = bk.attributes[d.num_triangles and 4:]  # skip the ground quad's 4 vertices
    t = scenes._unpack_snorm8x3(a[:, 0])
    n = scenes._unpack_snorm8x3(a[:, 2])
    cos = np.abs(np.sum(t * n, 1)) / (np.linalg.norm(t, axis=1) * np.linalg.norm(n, axis=1))
    assert np.quantile(cos, 0.99) < 0.05


def test_builder_rejects_bad_instances():
    d = scenes.instanced_scene(7, num_instances=3)
    bad = [scenes.Instance(5, np.eye(3, 4)),                                    # missing mesh
           scenes.Instance(0, np.zeros((3, 4))),                                # singular
           scenes.Instance(0, np.full((3, 4), np.nan))]                         # not finite
    for inst in bad:
        d2 = scenes.SceneData(positions=d.positions, attributes=d.attributes, triangles=d.triangles,
                              materials=d.materials, tex=d.tex, meshes=d.meshes, instances=[inst])
        with pytest.raises(pike_amd.PikeError):
            pike_amd.BVH.for_scene(d2)
    # no instances and no triangles of its own
    d3 = scenes.SceneData(positions=np.zeros((0, 3)), attributes=np.zeros((0, 4), np.uint32),
                          triangles=np.zeros((0, 4), np.int32), materials=d.materials, tex=d.tex,
                          meshes=d.meshes, instances=[])
    with pytest.raises(pike_amd.PikeError):
        pike_amd.BVH(d3.positions, d3.triangles)


def test_instanced_bvh_has_no_reference_layout():
    b = pike_amd.BVH.for_scene(scenes.instanced_scene(8, num_instances=2))
    with pytest.raises(pike_amd.PikeError):
        b.export_reference()
