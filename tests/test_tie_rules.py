"""The twin's two tie rules, pinned on the CPU before the GPU probes are checked against them.

The reference's closest-hit walk (src/gpu/bvh.cuh:576-706) keeps the first sphere it meets at a distance (its test is
open, t < closest: src/gpu/sphere.cuh:116) and the last quad (closed, t <= closest: src/gpu/quad.cuh:70), visiting the
nearer child first and the first child when both are entered at the same distance. The twin restates that walk
(walk order) beside the product's rule (scene order: the lower scene index wins a tie).

Pins:
- hand-built scenes whose tie outcome those intervals and that order fix, each chosen so that a plausible slip (a
  closed sphere test, an open quad test, the second child first on equal entries, or the scene-order rule itself)
  gives another answer;
- the seven book scenes' seeded ray sets: away from ties both rules equal the brute-force search over every
  primitive, and on tied rays the distance is the same and only the primitive may differ.
"""
import zlib

import numpy as np
import pytest

import oracle
from tests.rays import scene_rays

DOWN = [[0.0, 0.0, 5.0]], [[0.0, 0.0, -1.0]]  # a ray down the z axis: it meets z = -1 at t = 6
SQUARE = [-1, -1, -1, 2, 0, 0, 0, 2, 0]       # the quad z = -1, x and y in [-1, 1], its normal +z
FAR_QUADS = [[10.0 * k, 5, 5, 1, 0, 0, 0, 1, 0] for k in range(1, 9)]
FAR_SPHERES = [[10.0 * k, 5, 5, 0.5] for k in range(1, 9)]


def _hit(spheres, quads, walk_order, origins=DOWN[0], directions=DOWN[1]):
    ids, t, ent, sid, qid = oracle.twin_custom_closest_hit(spheres, quads, origins, directions, walk_order)
    return int(ids[0]), float(t[0]), list(sid), list(qid)


def test_two_equal_spheres_in_a_leaf_first_met_stays():
    # two shapes: one leaf of both, tested in scene order; the open test keeps the first
    for walk in (True, False):
        hid, t, sid, _ = _hit([[0, 0, -2, 1], [0, 0, -2, 1]], [], walk)
        assert t == 6.0 and hid == sid[0]


def test_two_equal_quads_in_a_leaf_last_met_wins():
    hid, t, _, qid = _hit([], [SQUARE, SQUARE], True)
    assert t == 6.0 and hid == qid[1], "the closed quad test lets the second quad replace the first"
    hid, _, _, qid = _hit([], [SQUARE, SQUARE], False)
    assert hid == qid[0], "scene order keeps the lower index"


def test_a_quad_wins_a_tie_against_a_sphere():
    # the sphere of radius 1 at z = -2 and the square at z = -1 are both met at t = 6
    hid, t, sid, qid = _hit([[0, 0, -2, 1]], [SQUARE], True)
    assert t == 6.0 and hid == qid[0]
    hid, _, sid, _ = _hit([[0, 0, -2, 1]], [SQUARE], False)
    assert hid == sid[0]


def test_coincident_quads_in_sibling_leaves():
    # ten shapes: leaves of one; the two squares sort first on x and become sibling leaves with equal boxes, entered at
    # the same distance: the first child is visited first and the second one, pushed, replaces it (closed test)
    hid, t, _, qid = _hit([], [SQUARE, SQUARE] + FAR_QUADS, True)
    assert t == 6.0 and hid == qid[1]
    hid, _, _, qid = _hit([], [SQUARE, SQUARE] + FAR_QUADS, False)
    assert hid == qid[0]


def test_coincident_spheres_in_sibling_leaves():
    # the same with two equal spheres: the first child's sphere is met first and the open test keeps it
    hid, t, sid, _ = _hit([[0, 0, -2, 1], [0, 0, -2, 1]] + FAR_SPHERES, [], True)
    assert t == 6.0 and hid == sid[0]


BOOKS = ["weekend", "bouncing", "glass", "final", "cornell", "smoke", "light"]


def _rays(rt, name):
    scene = rt.build_scene(name, width=64, samples=16, earth_width=512)
    return scene_rays(scene, 20000, seed=zlib.crc32(name.encode()) % 1000, width=64, spp=16, earth_width=512)


@pytest.mark.parametrize("name", BOOKS)
def test_walk_order_equals_brute_force_away_from_ties(rt, name):
    o, d, tm, f, inw = _rays(rt, name)
    bids, bt, tied = oracle.twin_brute_force(name, o, d, tm, f, inw, 64, 16, 512)
    wids, wt, _ = oracle.twin_closest_hit_walk_order(name, o, d, tm, f, inw, 64, 16, 512)
    sids, st, _ = oracle.twin_closest_hit(name, o, d, tm, f, inw, 64, 16, 512)
    assert np.array_equal(wt.view(np.uint32), bt.view(np.uint32)), "the closest distance doesn't depend on the rule"
    single = tied <= 1
    assert np.array_equal(wids[single], bids[single])
    assert np.array_equal(sids, bids), "scene order is the brute-force search's rule"
    differ = np.count_nonzero(wids != sids)
    print(f"{name}: {np.count_nonzero(tied > 1)} tied rays of {len(tied)}, {differ} decided differently by the "
          f"reference's walk order")
    assert differ <= np.count_nonzero(tied > 1)
    if name == "final":
        assert differ > 0, "final's adjacent ground boxes give ties the two rules decide differently"
