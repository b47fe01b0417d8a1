"""GPU parity probes against the oracle, bit for bit.

- The reference's own traversal (rt_session_reference_hits): the session's binary median tree walked in the
  reference's order with its intervals (src/gpu/bvh.cuh:576-706, sphere.cuh:116, quad.cuh:70) gives the twin's
  walk-order ids, distances and sides on the seven book scenes' seeded ray sets, tied rays counted and included, and
  the hand-built tie cases' answers; the product's own walk gives the twin's scene-order answers on the same cases.
- The per-sample generator (rt_sample_numbers): raw words, float and double draws equal the twin's for seeds of
  several rounds (seed + r 2^32), pixels and passes; the draws' bit layout is the reference's (random.cuh:76-90).
- The device's direction samplers (the shading's own): a uniform unit vector, the concentric map's uniform disk
  point and the cosine lobe about a normal, against their distributions' closed-form moments and CDFs.
- Shadow rays (rt_session_occluded): both shadow walks, the wavefront's occluded and the quality megakernel's any-hit
  closest_hit, against a loop over every primitive (oracle.twin_custom_occluded) on the reference's any-hit check's
  scene and rays (tests/lamp_tests.cu:1618-1900): 200 still and moving spheres and parallelograms, rays leaving a
  primitive or starting anywhere, finite and infinite t_max, a random or the closest primitive skipped; on the
  surface-area tree, the reference's median tree and packed nodes. No ray may differ but where the oracle marks the
  answer as within rounding of t_max (band) or of another decision (near), and those stay rare.
"""
import zlib

import numpy as np
import pytest

import oracle
from tests.custom_scene import CustomScene
from tests.rays import scene_rays
from tests.shadow_rays import t7_rays, t7_scene
from tests.test_tie_rules import DOWN, FAR_QUADS, FAR_SPHERES, SQUARE

pytestmark = pytest.mark.gpu

BOOKS = ["weekend", "bouncing", "glass", "final", "cornell", "smoke", "light"]


@pytest.mark.parametrize("name", BOOKS)
def test_reference_walk_bit_exact(rt, name):
    scene = rt.build_scene(name, width=64, samples=16, earth_width=512)
    o, d, tm, f, inw = scene_rays(scene, 60000, seed=zlib.crc32(name.encode()) % 1000, width=64, spp=16,
                                  earth_width=512)
    with rt.Session(scene, integrator="megakernel", tree="median") as s:
        ids, t, ent = s.reference_hits(o, d, tm, f, inw)
        pids, pt, pent = s.closest_hit(o, d, tm, f, inw)
    wids, wt, went = oracle.twin_closest_hit_walk_order(name, o, d, tm, f, inw, 64, 16, 512)
    _, _, tied = oracle.twin_brute_force(name, o, d, tm, f, inw, 64, 16, 512)
    assert (wids >= 0).sum() > len(wids) // 4, "the ray set should hit something"
    mism = np.nonzero((ids != wids) | (t.view(np.uint32) != wt.view(np.uint32)) | (ent != went))[0]
    assert len(mism) == 0, f"{len(mism)} of {len(ids)} rays differ, first {mism[:5]}: gpu {ids[mism[:5]]} twin {wids[mism[:5]]}"
    # the product's rule differs from the reference's on tied rays only
    assert np.array_equal(pt.view(np.uint32), t.view(np.uint32))
    differ = np.count_nonzero(pids != ids)
    print(f"{name}: {len(ids)} rays, {np.count_nonzero(tied > 1)} tied (all included above); the product's "
          f"lower-index rule decides {differ} of them otherwise ({100 * differ / len(ids):.2f} % of the rays)")
    assert np.all(tied[pids != ids] > 1)
    if name == "final":
        assert np.count_nonzero(tied > 1) > 100, "final's ray set holds many ties"


def _custom(rt, spheres, quads):
    scene = CustomScene(rt, spheres, quads)
    _, _, sid, qid, _ = rt.scene_bvh(scene, "median")
    return scene, sid, qid


@pytest.mark.parametrize("case,spheres,quads,reference_pick,product_pick", [
    ("equal spheres, one leaf", [[0, 0, -2, 1], [0, 0, -2, 1]], [], ("s", 0), ("s", 0)),
    ("equal quads, one leaf", [], [SQUARE, SQUARE], ("q", 1), ("q", 0)),
    ("sphere and quad", [[0, 0, -2, 1]], [SQUARE], ("q", 0), ("s", 0)),
    ("sibling quads", [], [SQUARE, SQUARE] + FAR_QUADS, ("q", 1), ("q", 0)),
    ("sibling spheres", [[0, 0, -2, 1], [0, 0, -2, 1]] + FAR_SPHERES, [], ("s", 0), ("s", 0)),
])
def test_hand_built_ties(rt, case, spheres, quads, reference_pick, product_pick):
    scene, sid, qid = _custom(rt, spheres, quads)
    o = np.float32(DOWN[0])
    d = np.float32(DOWN[1])
    args = (o, d, np.zeros(1, np.float32), np.full(1, -1, np.int32), np.zeros(1, np.uint8))
    with rt.Session(scene, integrator="megakernel", tree="median") as s:
        ref_ids, ref_t, _ = s.reference_hits(*args)
        prod_ids, prod_t, _ = s.closest_hit(*args)
    pick = lambda p: sid[p[1]] if p[0] == "s" else qid[p[1]]
    assert ref_t[0] == 6.0 and prod_t[0] == 6.0, case
    assert ref_ids[0] == pick(reference_pick), f"{case}: the reference's walk picks {reference_pick}"
    assert prod_ids[0] == pick(product_pick), f"{case}: the product's rule picks {product_pick}"
    # and the twin agrees on both rules (ids through each tree's own numbering of the inputs)
    for walk, want in ((True, reference_pick), (False, product_pick)):
        tid, tt, _, tsid, tqid = oracle.twin_custom_closest_hit(spheres, quads, DOWN[0], DOWN[1], walk)
        assert tt[0] == 6.0 and tid[0] == (tsid[want[1]] if want[0] == "s" else tqid[want[1]])


def test_sample_numbers_bit_exact(rt):
    rng = np.random.default_rng(17)
    n, draws = 4096, 24
    pixels = rng.integers(0, 1920 * 1080, n).astype(np.uint64)
    passes = rng.integers(0, 1024, n).astype(np.uint64)
    for seed in (0, 1, 5 + (3 << 32), 0xFFFFFFFFFFFFFFFF):
        words, floats, doubles, _ = rt.sample_numbers(seed, pixels, passes, draws)
        for k in range(0, n, 97):
            px, ps = int(pixels[k]), int(passes[k])
            assert np.array_equal(words[k], oracle.sample_words(seed, px, ps, draws)), (seed, px, ps)
            assert np.array_equal(floats[k], oracle.sample_floats(seed, px, ps, draws)), (seed, px, ps)
            assert np.array_equal(doubles[k].view(np.uint64), oracle.sample_doubles(seed, px, ps, draws).view(np.uint64))
        # the draws' bit layout, on every item: a float is the top 24 bits of one word, a double's top 32 of its 53
        # bits are the first of its two words
        assert np.array_equal((floats * 2.0 ** 24).astype(np.uint64), words.astype(np.uint64) >> 8)
        assert np.array_equal(np.floor(doubles[:, : draws // 2] * 2.0 ** 32).astype(np.uint64),
                              words[:, 0:draws:2].astype(np.uint64))
        assert floats.min() >= 0 and floats.max() < 1 and doubles.min() >= 0 and doubles.max() < 1
    # neighbouring pixels and passes get unrelated streams
    w, _, _, _ = rt.sample_numbers(9, np.arange(64, dtype=np.uint64), np.zeros(64, np.uint64), 1)
    assert len(np.unique(w)) == 64


def test_direction_samplers_distributions(rt):
    """rt_sample_numbers' directions over 2^18 items: the uniform unit vector (|v| = 1, E[z] = 0, E[z^2] = 1/3), the
    uniform disk point (r <= 1, z = 0, E[x] = 0, E[r^2] = 1/2, P(r^2 <= t) = t) and the cosine lobe about +z (|v| = 1,
    z > 0, E[z] = 2/3, E[z^2] = 1/2, P(z >= t) = 1 - t^2). A sampler with a wrong density (a uniform hemisphere for
    the cosine lobe, r uniform for the disk) misses these by hundreds of standard errors; each check allows 5."""
    n = 1 << 18
    _, _, _, dirs = rt.sample_numbers(77, np.arange(n, dtype=np.uint64), np.full(n, 3, np.uint64), 1)
    unit, disk, cosine = dirs[:, 0].astype(np.float64), dirs[:, 1].astype(np.float64), dirs[:, 2].astype(np.float64)

    def close(values, expect, label):
        se = values.std() / np.sqrt(values.size)
        assert abs(values.mean() - expect) < 5 * se + 1e-9, f"{label}: {values.mean():.6f} vs {expect:.6f} (se {se:.2e})"

    assert np.allclose(np.linalg.norm(unit, axis=1), 1.0, atol=1e-5)
    close(unit[:, 2], 0.0, "uniform E[z]")
    close(unit[:, 2] ** 2, 1 / 3, "uniform E[z^2]")
    close(unit[:, 0] * unit[:, 1], 0.0, "uniform E[xy]")
    r2 = disk[:, 0] ** 2 + disk[:, 1] ** 2
    assert np.all(r2 <= 1.0 + 1e-6) and np.all(disk[:, 2] == 0)
    close(disk[:, 0], 0.0, "disk E[x]")
    close(r2, 0.5, "disk E[r^2]")
    for t in (0.1, 0.25, 0.5, 0.9):
        close((r2 <= t).astype(np.float64), t, f"disk P(r^2 <= {t})")
    assert np.allclose(np.linalg.norm(cosine, axis=1), 1.0, atol=1e-5) and np.all(cosine[:, 2] > 0)
    close(cosine[:, 2], 2 / 3, "cosine E[z]")
    close(cosine[:, 2] ** 2, 0.5, "cosine E[z^2]")
    close(cosine[:, 0], 0.0, "cosine E[x]")
    for t in (0.2, 0.5, 0.8, 0.95):
        close((cosine[:, 2] >= t).astype(np.float64), 1 - t * t, f"cosine P(z >= {t})")


@pytest.mark.parametrize("tree,nodes,combination", [("default", "float", 0), ("median", "float", 0),
                                                     ("default", "packed", 2)])
def test_shadow_rays_against_brute_force(rt, tree, nodes, combination):
    spheres, quads = t7_scene()
    scene = CustomScene(rt, spheres, quads, combination=combination)
    _, _, sid, qid, _ = rt.scene_bvh(scene, tree)
    session_id = np.concatenate([sid, qid])            # input index -> the session tree's global id
    input_index = np.full(session_id.max() + 1, -1, np.int32)
    input_index[session_id] = np.arange(len(session_id), dtype=np.int32)
    to_session = lambda idx: np.where(idx >= 0, session_id[np.maximum(idx, 0)], -1).astype(np.int32)
    n = 400000
    o, d, tm, fi, inw, tmax, ski, closest = t7_rays(spheres, quads, n, seed=7691 + len(tree) + 7 * combination)
    with rt.Session(scene, integrator="megakernel", tree=tree, nodes=nodes) as s:
        # the closest hit as the skipped primitive: the lamp a light sample aimed at
        hid, _, _ = s.closest_hit(o, d, tm, to_session(fi), inw)
        ski = np.where(closest, np.where(hid >= 0, input_index[np.maximum(hid, 0)], -1), ski).astype(np.int32)
        blocked, walked = s.occluded(o, d, tm, to_session(fi), inw, tmax, to_session(ski))
    want, band, near, _, _ = oracle.twin_custom_occluded(spheres, quads, o, d, tm, fi, inw, tmax, ski)
    assert 0.2 < want.mean() < 0.8, "the ray set should hold blocked and open rays"
    assert band.mean() < 0.01 and near.mean() < 0.01
    for walk, got in (("occluded", blocked), ("any-hit closest_hit", walked)):
        differ = got != want
        hard = differ & (band == 0) & (near == 0)
        print(f"{tree}/{nodes}, {walk}: {int(want.sum())} of {n} rays blocked; {int(band.sum())} in the t_max band "
              f"({int((differ & (band == 1)).sum())} differ), {int((near & ~band.astype(bool)).sum())} near another "
              f"decision ({int((differ & (near == 1) & (band == 0)).sum())} differ)")
        first = np.nonzero(hard)[0][:5]
        assert hard.sum() == 0, f"{walk}: {int(hard.sum())} rays differ from brute force, first {first}"
        assert differ.sum() <= max(5, n // 10000), f"{walk}: too many rounding-zone differences"


@pytest.mark.parametrize("tree,nodes", [("default", "float"), ("default", "packed"), ("two_level", "auto")])
def test_triangle_shadow_rays_agree_with_closest_hits(rt, tree, nodes):
    """Config 3's triangles (2.62 M, the blobs baked into world space or walked through their 64 instances): a shadow
    ray is blocked exactly when the same tree's closest hit (itself checked against brute force and across trees in
    test_gpu_quality.py) lies within t_max -- short of it by 0.1 % or beyond it by 0.1 %, from the camera and from the
    hit points onward, leaving the triangle hit -- and both shadow walks agree when the hit triangle is skipped"""
    scene = rt.build_scene("triangles", width=64, samples=1, sampling="mis")
    cam = scene.desc.camera
    rng = np.random.default_rng(2024)
    n = 60000
    center = np.array([cam.center.x, cam.center.y, cam.center.z])
    p00 = np.array([cam.pixel00_loc.x, cam.pixel00_loc.y, cam.pixel00_loc.z])
    du = np.array([cam.pixel_delta_u.x, cam.pixel_delta_u.y, cam.pixel_delta_u.z])
    dv = np.array([cam.pixel_delta_v.x, cam.pixel_delta_v.y, cam.pixel_delta_v.z])
    targets = p00 + rng.uniform(0, cam.image_width, n)[:, None] * du + rng.uniform(0, cam.image_height, n)[:, None] * dv
    o = np.repeat(center[None], n, axis=0).astype(np.float32)
    d = (targets - center).astype(np.float32)
    none = np.full(n, -1, np.int32)
    with rt.Session(scene, integrator="megakernel", tree=tree, nodes=nodes) as s:
        ids, t, _ = s.closest_hit(o, d, np.zeros(n, np.float32), none, np.zeros(n, np.uint8))
        hit = ids >= 0
        assert hit.mean() > 0.5
        # onward from the hit points, leaving the primitive hit
        bo = (o[hit].astype(np.float64) + t[hit, None].astype(np.float64) * d[hit]).astype(np.float32)
        bd = rng.normal(size=(len(bo), 3)).astype(np.float32)
        O = np.concatenate([o, bo])
        D = np.concatenate([d, bd])
        F = np.concatenate([none, ids[hit]]).astype(np.int32)
        m = len(O)
        tm = np.zeros(m, np.float32)
        inw = np.zeros(m, np.uint8)
        ids2, t2, _ = s.closest_hit(O, D, tm, F, inw)
        met = ids2 >= 0
        assert 0.2 < met.mean() < 0.95
        noskip = np.full(m, -1, np.int32)
        for factor in (0.999, 1.001):
            t_max = np.where(met, t2 * np.float32(factor), rng.uniform(1, 100, m)).astype(np.float32)
            want = met & (factor > 1)
            blocked, walked = s.occluded(O, D, tm, F, inw, t_max, noskip)
            for walk, got in (("occluded", blocked), ("any-hit closest_hit", walked)):
                bad = np.nonzero(got.astype(bool) != want)[0]
                assert len(bad) == 0, f"{walk}, t_max = {factor} x the closest hit: {len(bad)} of {m} differ, first {bad[:5]}"
        # the hit triangle skipped: whatever lies behind it within 2x its distance
        t_max = np.where(met, 2 * t2, np.inf).astype(np.float32)
        blocked, walked = s.occluded(O, D, tm, F, inw, t_max, np.where(met, ids2, -1).astype(np.int32))
        assert np.array_equal(blocked, walked)
        assert 0 < blocked.sum() < met.sum()
        print(f"{tree}/{nodes}: {m} rays, {int(met.sum())} hit; behind the hit within 2x: {int(blocked.sum())}")
