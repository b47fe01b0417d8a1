"""GPU parity of the packed tree (RT_NODES_PACKED: 64-byte nodes of 8-bit planes). A packed box holds its float box, so
the walks over packed nodes enter every child the float walks enter and find the same closest hit: the same primitive
at the same distance (bit for bit) on the same side, ties included, and every integrator renders the same bytes on
either layout.
"""
import zlib

import numpy as np
import pytest

import oracle
from tests.rays import scene_rays

pytestmark = pytest.mark.gpu


def _same_hits(a, b):
    ids_a, t_a, ent_a = a
    ids_b, t_b, ent_b = b
    mism = np.nonzero((ids_a != ids_b) | (t_a.view(np.uint32) != t_b.view(np.uint32)) | (ent_a != ent_b))[0]
    return mism


@pytest.mark.parametrize("name", ["weekend", "bouncing", "glass", "final", "cornell", "smoke", "light"])
def test_packed_walk_bit_exact_against_the_twin(rt, name):
    """The book scenes under -p mis (so that the quality integrator's kernels, which walk packed trees, render them),
    on the reference's median tree packed: every ray's closest hit equals the twin's, ties included"""
    scene = rt.build_scene(name, width=64, samples=16, earth_width=512, sampling="mis")
    o, d, tm, f, inw = scene_rays(scene, 60000, seed=3 + zlib.crc32(name.encode()) % 1000, width=64, spp=16,
                                  earth_width=512)
    with rt.Session(scene, integrator="megakernel", tree="median", nodes="packed") as s:
        got = s.closest_hit(o, d, tm, f, inw)
    twin = oracle.twin_closest_hit(name, o, d, tm, f, inw, 64, 16, 512)
    assert (twin[0] >= 0).sum() > len(twin[0]) // 4, "the ray set should hit something"
    mism = _same_hits(got, twin)
    assert len(mism) == 0, f"{len(mism)} of {len(o)} rays differ, first {mism[:5]}"


def _triangle_rays(rt, scene, n, seed):
    """Camera rays, rays from random points of the scene's box, and bounce rays from the camera rays' hits (found by
    the float walk), leaving the primitive they start on"""
    rng = np.random.default_rng(seed)
    cam = scene.desc.camera
    center = np.array([cam.center.x, cam.center.y, cam.center.z])
    p00 = np.array([cam.pixel00_loc.x, cam.pixel00_loc.y, cam.pixel00_loc.z])
    du = np.array([cam.pixel_delta_u.x, cam.pixel_delta_u.y, cam.pixel_delta_u.z])
    dv = np.array([cam.pixel_delta_v.x, cam.pixel_delta_v.y, cam.pixel_delta_v.z])
    k = n // 3
    targets = p00 + rng.uniform(0, cam.image_width, k)[:, None] * du + rng.uniform(0, cam.image_height, k)[:, None] * dv
    o1 = np.repeat(center[None], k, axis=0)
    d1 = targets - o1
    lo, hi = targets.min(axis=0), targets.max(axis=0)
    o2 = lo + rng.uniform(0, 1, (k, 3)) * (hi - lo)
    d2 = rng.normal(size=(k, 3))
    return np.concatenate([o1, o2]).astype(np.float32), np.concatenate([d1, d2]).astype(np.float32)


def test_packed_walk_equals_float_walk_on_config3(rt):
    """Config 3's 2.62 M triangles (the tree packed by default): the packed walk's closest hits equal the float walk's
    bit for bit on camera rays, rays from inside the scene and bounce rays off the first hits"""
    scene = rt.build_scene("triangles", width=192, samples=1, sampling="mis")
    o, d = _triangle_rays(rt, scene, 60000, 11)
    n = len(o)
    times = np.zeros(n, np.float32)
    from_ids = np.full(n, -1, np.int32)
    inward = np.zeros(n, np.uint8)
    with rt.Session(scene, integrator="megakernel", nodes="float") as s:
        first = s.closest_hit(o, d, times, from_ids, inward)
        # bounce rays: from the hits, in random directions, leaving the primitive hit
        hit = first[0] >= 0
        rng = np.random.default_rng(12)
        bo = (o[hit].astype(np.float64) + first[1][hit, None].astype(np.float64) * d[hit]).astype(np.float32)
        bd = rng.normal(size=bo.shape).astype(np.float32)
        bfrom = first[0][hit]
        o2 = np.concatenate([o, bo])
        d2 = np.concatenate([d, bd])
        t2 = np.zeros(len(o2), np.float32)
        f2 = np.concatenate([from_ids, bfrom])
        i2 = np.zeros(len(o2), np.uint8)
        float_hits = s.closest_hit(o2, d2, t2, f2, i2)
        assert s.kernel().name  # (the session is a quality one)
    with rt.Session(scene, integrator="megakernel", nodes="packed") as s:
        packed_hits = s.closest_hit(o2, d2, t2, f2, i2)
    assert hit.mean() > 0.5
    mism = _same_hits(packed_hits, float_hits)
    assert len(mism) == 0, f"{len(mism)} of {len(o2)} rays differ, first {mism[:5]}"


@pytest.mark.parametrize("name,sampling", [("triangles", "mis"), ("cornell-ggx", "mis"), ("veach", "nee"),
                                           ("final", "mis")])
def test_packed_tree_renders_the_same_bytes(rt, name, sampling):
    """Every integrator (the queue megakernel, the plain megakernel, the wavefront) on the packed tree renders the
    float tree's sums bit for bit"""
    width, spp = (64, 9) if name == "triangles" else (96, 16)
    scene = rt.build_scene(name, width=width, samples=spp, strata=3 if spp == 9 else 4, sampling=sampling,
                           earth_width=512)
    with rt.Session(scene, seed=4, integrator="megakernel", nodes="float") as s:
        s.render_to(spp)
        reference = s.sums()
        counts = s.counters()[:2]
    for integrator in ("megakernel", "plain", "wavefront"):
        with rt.Session(scene, seed=4, integrator=integrator, nodes="packed") as s:
            s.render_to(spp)
            got = s.sums()
            assert s.counters()[:2] == counts, integrator
        mism = np.count_nonzero(got.view(np.uint64) != reference.view(np.uint64))
        assert mism == 0, f"{integrator}: {mism} of {got.size} sums differ"


def test_packed_nodes_refused_where_no_kernel_walks_them(rt):
    """Packed nodes are for the quality integrator's float32 one-level trees"""
    books = rt.build_scene("weekend", width=32, samples=1)
    with pytest.raises(ValueError, match="packed nodes"):
        rt.Session(books, nodes="packed")
    quality = rt.build_scene("cornell", width=32, samples=1, sampling="mis")
    with pytest.raises(ValueError):
        rt.Session(quality, nodes="packed", precision="float64")
    with pytest.raises(ValueError, match="node layout"):
        rt.Session(quality, nodes=7)
