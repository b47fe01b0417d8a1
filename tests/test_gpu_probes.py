"""GPU parity probes against the oracle, bit for bit.

- The reference's own traversal (rt_session_reference_hits): the session's binary median tree walked in the
  reference's order with its intervals (src/gpu/bvh.cuh:576-706, sphere.cuh:116, quad.cuh:70) gives the twin's
  walk-order ids, distances and sides on the seven book scenes' seeded ray sets, tied rays counted and included, and
  the hand-built tie cases' answers; the product's own walk gives the twin's scene-order answers on the same cases.
- The per-sample generator (rt_sample_numbers): raw words, float and double draws equal the twin's for seeds of
  several rounds (seed + r 2^32), pixels and passes; the draws' bit layout is the reference's (random.cuh:76-90).
- The device's direction samplers (the shading's own): a uniform unit vector, the concentric map's uniform disk
  point and the cosine lobe about a normal, against their distributions' closed-form moments and CDFs.
"""
import zlib

import numpy as np
import pytest

import oracle
from tests.custom_scene import CustomScene
from tests.rays import scene_rays
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
