"""The C ABI on the CPU: the library loads and exports every symbol include/raytracing_amd.h declares; scenes built by
name flatten exactly as the oracle's independent restatement of the reference flattens them; the BVH is the twin's,
node for node; every check that can fail without the GPU fails before touching it, with the reference's error kinds.
"""
import ctypes as C
import os
import re

import numpy as np
import pytest

import oracle
from tests.flatten import product_vector

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SCENES = ["materials", "weekend", "bouncing", "checker", "earth", "perlin", "quads", "light", "cornell", "smoke",
          "final", "aluminum", "glass", "furnace"]


def header_functions():
    text = open(os.path.join(ROOT, "include", "raytracing_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rt_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported(rt):
    names = header_functions()
    assert len(names) >= 30
    lib = rt.library()
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert lib.rt_abi_version() == 4


def test_scene_names(rt):
    assert set(SCENES) <= set(rt.scene_names())


@pytest.mark.parametrize("name", SCENES)
def test_flattening_matches_oracle(rt, name):
    s = rt.build_scene(name, width=96, samples=16, earth_width=256)
    ours = product_vector(s.desc)
    theirs = oracle.flatten(name, 96, 16, -1, 0, 256)
    assert ours.shape == theirs.shape
    assert np.array_equal(ours, theirs)


@pytest.mark.parametrize("name,sampling", [("cornell", "uniform"), ("glass", "lights"), ("final", "material"),
                                           ("light", "mixture")])
def test_flattening_with_strategy(rt, name, sampling):
    s = rt.build_scene(name, width=64, samples=4, sampling=sampling, earth_width=256)
    theirs = oracle.flatten(name, 64, 4, rt.SAMPLING[sampling], 0, 256)
    assert np.array_equal(product_vector(s.desc), theirs)


def test_scene_counts_match_the_reference():
    # final: 20x20 boxes of 6 quads + the lamp; 1000 cloud spheres + 6 spheres; two media (src/scenes.ixx:380-458)
    import raytracing_amd as rt

    s = rt.build_scene("final", earth_width=256)
    assert (s.desc.quad_count, s.desc.sphere_count, s.desc.medium_count) == (2401, 1006, 2)
    assert (s.width, s.height, s.desc.camera.sqrt_spp, s.desc.camera.max_depth) == (1080, 1080, 32, 40)
    w = rt.build_scene("weekend")
    assert (w.width, w.height, w.desc.camera.sqrt_spp) == (1920, 1080, 22)
    g = rt.build_scene("glass", width=600, samples=1024)
    assert (g.desc.quad_count, g.desc.sphere_count, g.desc.light_count) == (12, 1, 2)
    assert [g.desc.lights[i].weight for i in range(2)] == [0.5, 0.5]


@pytest.mark.parametrize("name", ["weekend", "glass", "final", "bouncing", "materials", "smoke", "light", "cornell"])
def test_bvh_matches_twin(rt, name):
    s = rt.build_scene(name, width=64, samples=16, earth_width=256)
    ours = rt.scene_bvh(s)
    theirs = oracle.twin_bvh(name, 64, 16, 256)
    for a, b in zip(ours[:4], theirs[:4]):
        assert np.array_equal(a, b)
    assert ours[4] == theirs[4]


def test_bvh_node_counts():
    # the reference's counts at leaf size 1: 2n - 1 nodes (weekend 485 spheres, final 3407 primitives)
    import raytracing_amd as rt

    assert rt.scene_bvh(rt.build_scene("weekend"))[0].shape[0] == 969
    assert rt.scene_bvh(rt.build_scene("final", earth_width=256))[0].shape[0] == 6813


@pytest.mark.parametrize("name", ["weekend", "final", "cornell", "smoke"])
def test_surface_area_tree_holds_every_primitive(rt, name):
    s = rt.build_scene(name, width=64, samples=16, earth_width=256)
    boxes, ints, sphere_ids, quad_ids, depth = rt.scene_bvh(s, "sah")
    median = rt.scene_bvh(s, "median")
    # the same primitives: each sphere and quad of the scene once, in one leaf of its kind
    assert sorted(sphere_ids) == sorted(median[2]) and sorted(quad_ids) == sorted(median[3])
    leaf = ints[:, 1] != 0
    counts, kinds = ints[leaf, 1] & 255, ints[leaf, 1] >> 8
    assert counts.sum() == len(sphere_ids) + len(quad_ids)
    assert counts.max() <= 8 and depth <= 64
    for kind in range(3):
        starts = np.sort(ints[leaf][kinds == kind, 0])
        sizes = counts[kinds == kind][np.argsort(ints[leaf][kinds == kind, 0])]
        assert np.array_equal(starts, np.concatenate([[0], np.cumsum(sizes)[:-1]])) if len(starts) else True
    # every inner node's box holds both children's; fewer nodes than the reference's leaf-per-primitive tree of a
    # scene of more than 8 shapes
    inner = np.nonzero(~leaf)[0]
    for child in (inner + 1, ints[inner, 0]):
        assert np.all(boxes[child, :3] >= boxes[inner, :3]) and np.all(boxes[child, 3:] <= boxes[inner, 3:])
    if counts.sum() > 8:
        assert len(boxes) < len(median[0])


def test_unknown_tree_is_refused(rt):
    s = rt.build_scene("weekend", width=32, samples=1)
    with pytest.raises(ValueError, match="tree"):
        rt.Session(s, tree=4)
    with pytest.raises(ValueError, match="tree"):
        rt.render(s, tree=-1)
    with pytest.raises(ValueError, match="tree"):
        rt.scene_bvh(s, 9)


def test_twin_traversal_equals_brute_force():
    import raytracing_amd as rt
    from tests.rays import scene_rays

    # the final scene's bounce rays meet coplanar faces of adjacent boxes from inside at exactly the same distance:
    # the tie rule (the lower scene index wins) makes the walk's answer the brute-force search's even then
    ties = 0
    for name in ("weekend", "glass", "bouncing", "final", "cornell"):
        s = rt.build_scene(name, width=64, samples=16, earth_width=256)
        o, d, tm, f, inw = scene_rays(s, 20000, 5, 64, 16, earth_width=256)
        ids, t, _ = oracle.twin_closest_hit(name, o, d, tm, f, inw, 64, 16, 256)
        bids, bt, tied = oracle.twin_brute_force(name, o, d, tm, f, inw, 64, 16, 256)
        assert np.array_equal(t.view(np.uint32), bt.view(np.uint32))
        assert np.array_equal(ids, bids)
        ties += int((tied > 1).sum())
    assert ties > 100, "the ray sets should hold ties"


def test_strata_overrides(rt):
    s = rt.build_scene("weekend", width=64, samples=50, strata=4)
    assert s.passes == 50 and s.desc.camera.sqrt_spp == 4
    s2 = rt.build_scene("weekend", width=64, samples=1000)
    assert s2.passes == 32 * 32


def test_errors_before_the_gpu(rt):
    with pytest.raises(ValueError):
        rt.build_scene("no-such-scene")
    with pytest.raises(ValueError):
        rt.build_scene("weekend", strata=50000)
    s = rt.build_scene("weekend", width=32, samples=4)
    bad = rt.Scene.from_buffer_copy(s.desc)
    bad.camera.image_width = 0
    with pytest.raises(RuntimeError, match="pixel"):
        rt.Session(bad)
    mats = (rt.MaterialDesc * s.desc.material_count)()
    C.memmove(mats, s.desc.materials, C.sizeof(mats))
    mats[0].texture = 10**6
    bad2 = rt.Scene.from_buffer_copy(s.desc)
    bad2.materials = mats
    with pytest.raises(RuntimeError, match="texture"):
        rt.Session(bad2)
    with pytest.raises(ValueError):
        rt.Session(s, launch_seconds=2.0)
    with pytest.raises(ValueError):
        rt.Session(s, block_size=100)
    # a server's step sizes are checked before it makes its streams
    with pytest.raises(ValueError, match="paths_per_step"):
        rt.Server(paths_per_step=-1)
    with pytest.raises(ValueError, match="passes_per_step"):
        rt.Server(passes_per_step=1025)


def test_convexity_of_media_boundaries(rt):
    s = rt.build_scene("smoke", width=32, samples=4)
    assert s.desc.medium_count == 2
    assert all(s.desc.media[i].boundary == 1 and s.desc.media[i].quad_count == 6 for i in range(2))
    quads = (rt.QuadDesc * 6)()
    C.memmove(quads, s.desc.media[0].quads, C.sizeof(quads))
    quads[0].q.x += 50.0  # a face moved off the box: the quads no longer bound a convex shape
    media = (rt.MediumDesc * 2)()
    C.memmove(media, s.desc.media, C.sizeof(media))
    media[0].quads = quads
    bad = rt.Scene.from_buffer_copy(s.desc)
    bad.media = media
    with pytest.raises(RuntimeError, match="convex"):
        rt.Session(bad)
