"""The packed BVH nodes (RT_NODES_PACKED) on the host: every packed child box holds its float box (its lower planes
rounded down to the node's grid, its upper ones up) and is no looser than that rounding; every grid point is a float
(the corner a multiple of a power-of-two step, below 2^24 steps from zero), so that the kernels unpack a plane exactly
and their slab test on it, monotonic in the plane, never rejects a child the float box would enter; missing children
stay empty; the references are the float nodes'. A packing that rounded a plane the wrong way, put the corner off the
grid, or mixed up axes or children fails here; the GPU tests check the walks' closest hits on either layout against
each other and the twin.
"""
import numpy as np
import pytest

NO_CHILD = 0x7FFFFFFF


def _check_packing(wide, packed):
    assert len(wide) == len(packed) > 0
    assert np.array_equal(wide["ref"], packed["ref"])
    step = packed["step"].astype(np.float64)
    mantissa, _ = np.frexp(step)
    assert np.all(mantissa == 0.5), "every step is a power of two"
    assert np.all(step >= np.float64(np.finfo(np.float32).tiny))
    valid = wide["ref"] != NO_CHILD  # (node, child)
    assert np.all(valid.any(axis=1))
    origin = packed["origin"].astype(np.float64)
    # the corner is a multiple of the step, and every grid point corner + q step (q <= 255) a float
    assert np.all(np.floor(origin / step) == origin / step)
    assert np.all(np.abs(origin / step) + 255 < 2 ** 24)
    for q in (0, 1, 254, 255):
        g = origin + q * step
        assert np.array_equal(g.astype(np.float32).astype(np.float64), g)
    o = origin[:, :, None]
    s = step[:, :, None]
    lo_p = o + packed["lo"].astype(np.float64) * s  # (node, axis, child)
    hi_p = o + packed["hi"].astype(np.float64) * s
    lo_f = wide["lo"].astype(np.float64)
    hi_f = wide["hi"].astype(np.float64)
    v = np.broadcast_to(valid[:, None, :], lo_p.shape)
    s4 = np.broadcast_to(s, lo_p.shape)
    # conservative, and no looser than rounding to the grid
    assert np.all(lo_p[v] <= lo_f[v]) and np.all(hi_p[v] >= hi_f[v])
    # (the grid's next points, exact in double, against the float planes)
    assert np.all(lo_p[v] + s4[v] > lo_f[v]) and np.all(hi_p[v] - s4[v] < hi_f[v])
    # the grid spans the node's box in at most 254 steps, with the smallest such step unless the coordinates'
    # precision sets it
    node_lo = np.where(v, lo_f, np.inf).min(axis=2)
    node_hi = np.where(v, hi_f, -np.inf).max(axis=2)
    extent = node_hi - node_lo
    assert np.all(extent <= 254 * step)
    ulp_bound = np.ldexp(1.0, np.frexp(np.maximum(np.abs(node_lo), np.abs(node_hi)))[1] - 1 - 22)
    coarse = step > np.maximum(ulp_bound, np.float64(np.finfo(np.float32).tiny))
    assert np.all(extent[coarse] > 127 * step[coarse]), "the step is the smallest power of two that fits"
    # missing children: empty boxes
    m = ~valid
    assert np.all(packed["lo"][np.broadcast_to(m[:, None, :], lo_p.shape)] == 255)
    assert np.all(packed["hi"][np.broadcast_to(m[:, None, :], lo_p.shape)] == 0)


@pytest.mark.parametrize("name", ["cornell", "final", "weekend", "veach", "sun-test"])
def test_packed_boxes_hold_the_float_boxes(rt, name):
    scene = rt.build_scene(name, width=64, samples=4, earth_width=64)
    wide, packed = rt.scene_wide_nodes(scene)
    _check_packing(wide, packed)


def test_packed_boxes_of_the_triangle_scene(rt):
    """Config 3's tree (2.62 M triangles, ~730 K wide nodes): the tree that is packed by default."""
    scene = rt.build_scene("triangles", width=64, samples=4)
    wide, packed = rt.scene_wide_nodes(scene)
    assert wide.nbytes > 16 << 20, "large enough to be packed by default"
    _check_packing(wide, packed)
