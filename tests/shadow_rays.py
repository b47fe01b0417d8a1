"""The shadow rays' brute-force check's scene and rays (the reference's T7, tests/lamp_tests.cu:1618-1900, restated
with numpy's generator): 200 primitives in [-10, 10]^3 -- 80 still spheres of radius 0.2 to 2, 40 moving ones
(a displacement of up to 2 per axis), 80 parallelograms of sides 0.5 to 4 in random directions at least 11.5 degrees
apart -- and rays of which half leave a random primitive and half start anywhere in the box, with a random time and
direction, t_max infinite for a fifth of them and log-uniform in [0.01, 50) otherwise, and a primitive skipped for
two fifths: a random one, or (skip_closest) the one the ray would hit first."""
import numpy as np

SPHERES = 120
MOVING_FROM = 80  # spheres from this index on move
QUADS = 80


def _unit(rng, n):
    v = rng.normal(size=(n, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def t7_scene(seed: int = 7680):
    """(spheres, quads): x, y, z, r, dx, dy, dz rows and q, u, v rows, in float64"""
    rng = np.random.default_rng(seed)
    spheres = np.zeros((SPHERES, 7))
    spheres[:, 0:3] = rng.uniform(-10, 10, (SPHERES, 3))
    spheres[:, 3] = rng.uniform(0.2, 2, SPHERES)
    spheres[MOVING_FROM:, 4:7] = rng.uniform(-2, 2, (SPHERES - MOVING_FROM, 3))
    quads = np.zeros((QUADS, 9))
    for k in range(QUADS):
        a = _unit(rng, 1)[0]
        b = _unit(rng, 1)[0]
        while np.linalg.norm(np.cross(a, b)) < 0.2:
            b = _unit(rng, 1)[0]
        quads[k, 0:3] = rng.uniform(-10, 10, 3)
        quads[k, 3:6] = rng.uniform(0.5, 4) * a
        quads[k, 6:9] = rng.uniform(0.5, 4) * b
    return spheres, quads


def t7_rays(spheres, quads, n: int, seed: int):
    """(origins, directions, times, from_index, inward, t_max, skip_index, skip_closest): from and skip name input
    primitives (spheres, then quads) or -1; skip_closest marks the rays whose skip is to be the closest hit"""
    rng = np.random.default_rng(seed)
    count = len(spheres) + len(quads)
    times = rng.uniform(0, 1, n).astype(np.float32)
    on = rng.uniform(0, 1, n) < 0.5
    from_index = np.where(on, rng.integers(0, count, n), -1).astype(np.int32)
    origins = rng.uniform(-10, 10, (n, 3))
    outward = _unit(rng, n)
    directions = _unit(rng, n).astype(np.float32)
    on_sphere = on & (from_index < len(spheres))
    s = spheres[from_index[on_sphere]]
    centers = s[:, 0:3] + times[on_sphere, None].astype(np.float64) * s[:, 4:7]
    origins[on_sphere] = centers + s[:, 3:4] * outward[on_sphere]
    on_quad = on & (from_index >= len(spheres))
    q = quads[from_index[on_quad] - len(spheres)]
    ab = rng.uniform(0, 1, (n, 2))[on_quad]
    origins[on_quad] = q[:, 0:3] + ab[:, 0:1] * q[:, 3:6] + ab[:, 1:2] * q[:, 6:9]
    inward = (on_sphere & (np.einsum("ij,ij->i", directions.astype(np.float64), outward) < 0)).astype(np.uint8)
    finite = rng.uniform(0, 1, n) < 0.8
    t_max = np.where(finite, 0.01 * np.exp(rng.uniform(0, 1, n) * np.log(5000.0)), np.inf).astype(np.float32)
    pick = rng.uniform(0, 1, n)
    skip_index = np.where((pick >= 0.6) & (pick < 0.8), rng.integers(0, count, n), -1).astype(np.int32)
    skip_closest = pick >= 0.8
    return (origins.astype(np.float32), directions, times, from_index, inward, t_max, skip_index, skip_closest)
