"""The brute-force shadow-ray oracle (oracle.twin_custom_occluded), pinned on the CPU before it judges the GPU's two
shadow walks (tests/test_gpu_probes.py):

- known answers of src/gpu/shadow.cuh:14-25's rules: a sphere's interval open at t_max, a quad's closed, the skipped
  primitive never blocks, the primitive a ray leaves never blocks it unless the ray goes into a sphere (its far root);
- invariants on the brute-force check's random rays: nothing blocks below t_min, the answer can only turn on as t_max
  grows and only turn off as a primitive is skipped;
- agreement with the twin's closest-hit walk (itself pinned against brute force and the reference's tie rules) on
  still spheres and quads: a ray is blocked exactly when its closest hit lies before t_max, away from the rays the
  oracle marks as rounding-sensitive (band, near).
"""
import numpy as np
import pytest

import oracle
from tests.shadow_rays import t7_rays, t7_scene

SPHERE = [0.0, 0.0, -5.0, 1.0, 0.0, 0.0, 0.0]            # entered at t = 4 by a ray from the origin down -z
SQUARE_AT_4 = [-1.0, -1.0, -4.0, 2.0, 0.0, 0.0, 0.0, 2.0, 0.0]  # met at t = 4 exactly by the same ray
DOWN = (np.zeros((1, 3), np.float32), np.array([[0, 0, -1]], np.float32))


def _one(spheres, quads, t_max, origin=DOWN[0], direction=DOWN[1], from_index=-1, inward=0, skip=-1, time=0.0):
    blocked, band, near, _, _ = oracle.twin_custom_occluded(
        np.asarray(spheres, np.float64).reshape(-1, 7), np.asarray(quads, np.float64).reshape(-1, 9),
        origin, direction, np.float32([time]), [from_index], [inward], np.float32([t_max]), [skip])
    return int(blocked[0]), int(band[0]), int(near[0])


def test_sphere_interval_is_open_at_t_max():
    assert _one([SPHERE], [], 4.0)[0] == 0
    assert _one([SPHERE], [], np.nextafter(np.float32(4), np.float32(5)))[0] == 1
    assert _one([SPHERE], [], np.inf)[0] == 1
    assert _one([SPHERE], [], 3.0)[0] == 0


def test_quad_interval_is_closed_at_t_max():
    assert _one([], [SQUARE_AT_4], 4.0)[0] == 1
    assert _one([], [SQUARE_AT_4], np.nextafter(np.float32(4), np.float32(3)))[0] == 0


def test_the_skipped_primitive_never_blocks():
    assert _one([SPHERE], [SQUARE_AT_4], np.inf, skip=0)[0] == 1   # the quad still blocks
    assert _one([SPHERE], [SQUARE_AT_4], np.inf, skip=1)[0] == 1   # the sphere still blocks
    assert _one([SPHERE], [], np.inf, skip=0)[0] == 0
    assert _one([], [SQUARE_AT_4], np.inf, skip=0)[0] == 0


def test_the_primitive_left_blocks_only_a_ray_into_a_sphere():
    top = np.float32([[0, 0, -4]])  # on the sphere, which the ray down -z crosses to its far root at t = 2
    assert _one([SPHERE], [], np.inf, origin=top, from_index=0, inward=0)[0] == 0
    assert _one([SPHERE], [], 2.5, origin=top, from_index=0, inward=1)[0] == 1
    assert _one([SPHERE], [], 1.5, origin=top, from_index=0, inward=1)[0] == 0
    on_quad = np.float32([[0.25, 0.5, -4]])
    assert _one([], [SQUARE_AT_4], np.inf, origin=on_quad, direction=np.float32([[0.3, 0, 1]]), from_index=0)[0] == 0
    assert _one([], [SQUARE_AT_4], np.inf, origin=on_quad, direction=np.float32([[0.3, 0, -1]]), from_index=0)[0] == 0


def test_a_moving_sphere_blocks_where_it_is_at_the_ray_time():
    moving = [0.0, 0.0, -5.0, 1.0, 10.0, 0.0, 0.0]  # at x = 10 t
    assert _one([moving], [], np.inf, time=0.0)[0] == 1
    assert _one([moving], [], np.inf, time=0.5)[0] == 0
    assert _one([moving], [], np.inf, origin=np.float32([[5, 0, 0]]), time=0.5)[0] == 1


@pytest.fixture(scope="module")
def t7():
    spheres, quads = t7_scene()
    rays = t7_rays(spheres, quads, 60000, seed=7681)
    return spheres, quads, rays


def test_invariants_of_the_random_rays(t7):
    spheres, quads, (o, d, tm, fi, inw, tmax, ski, _) = t7
    blocked, band, near, _, _ = oracle.twin_custom_occluded(spheres, quads, o, d, tm, fi, inw, tmax, ski)
    assert 0.2 < blocked.mean() < 0.8, "the ray set should hold blocked and open rays"
    assert band.mean() < 0.01 and near.mean() < 0.01
    # nothing lies in an empty interval
    tiny, _, _, _, _ = oracle.twin_custom_occluded(spheres, quads, o, d, tm, fi, inw, np.full_like(tmax, 0.001), ski)
    assert tiny.sum() == 0
    # a longer shadow ray is blocked wherever the shorter one is
    longer, _, _, _, _ = oracle.twin_custom_occluded(spheres, quads, o, d, tm, fi, inw, 2 * tmax, ski)
    assert np.all(longer >= blocked) and np.any(longer > blocked)
    # skipping a primitive can only open a ray
    none, _, _, _, _ = oracle.twin_custom_occluded(spheres, quads, o, d, tm, fi, inw, tmax, np.full_like(ski, -1))
    assert np.all(none >= blocked) and np.any(none > blocked)


def test_agrees_with_the_twins_closest_hit_walk(t7):
    spheres, quads, (o, d, _, _, _, tmax, _, _) = t7
    still = spheres[:, :4]
    n = len(o)
    ids, t, _, sid, qid = oracle.twin_custom_closest_hit(still, quads, o, d, walk_order=False)
    blocked, band, near, osid, oqid = oracle.twin_custom_occluded(
        np.concatenate([still, np.zeros((len(still), 3))], axis=1), quads, o, d, np.zeros(n, np.float32),
        np.full(n, -1), np.zeros(n, np.uint8), tmax, np.full(n, -1))
    assert np.array_equal(sid, osid) and np.array_equal(qid, oqid)
    quad_hit = ids >= len(still)
    expect = (ids >= 0) & ((t < tmax) | (quad_hit & (t == tmax)))
    sure = (band == 0) & (near == 0)
    assert sure.mean() > 0.99
    assert np.array_equal(blocked[sure].astype(bool), expect[sure])
    # and band marks a hit at t_max itself
    at = (ids >= 0) & sure
    b2, band2, _, _, _ = oracle.twin_custom_occluded(
        np.concatenate([still, np.zeros((len(still), 3))], axis=1), quads, o[at], d[at],
        np.zeros(at.sum(), np.float32), np.full(at.sum(), -1), np.zeros(at.sum(), np.uint8), t[at], np.full(at.sum(), -1))
    assert np.all(band2 == 1)
    assert np.array_equal(b2.astype(bool), quad_hit[at])
