"""Seeded ray sets for the closest-hit parity tests: camera rays, rays from random points, and bounce rays that start on
the primitive a first ray hit (with `from` and `inward` set as a path would set them)."""
import numpy as np

import oracle


def scene_rays(scene, n: int, seed: int, width: int = 0, spp: int = 0, earth_width: int = 0):
    rng = np.random.default_rng(seed)
    cam = scene.desc.camera
    center = np.array([cam.center.x, cam.center.y, cam.center.z])
    p00 = np.array([cam.pixel00_loc.x, cam.pixel00_loc.y, cam.pixel00_loc.z])
    du = np.array([cam.pixel_delta_u.x, cam.pixel_delta_u.y, cam.pixel_delta_u.z])
    dv = np.array([cam.pixel_delta_v.x, cam.pixel_delta_v.y, cam.pixel_delta_v.z])
    k = n // 3
    # camera rays through random points of the image
    u = rng.uniform(0, cam.image_width, k)
    v = rng.uniform(0, cam.image_height, k)
    targets = p00 + u[:, None] * du + v[:, None] * dv
    o1 = np.repeat(center[None], k, axis=0)
    d1 = targets - o1
    # rays from random points in a box around the camera's view, in random directions
    lo = np.minimum(center, targets.min(axis=0))
    hi = np.maximum(center, targets.max(axis=0))
    span = hi - lo
    o2 = lo - 0.5 * span + rng.uniform(0, 1, (k, 3)) * 2 * span
    d2 = rng.normal(size=(k, 3))
    origins = np.concatenate([o1, o2]).astype(np.float32)
    directions = np.concatenate([d1, d2]).astype(np.float32)
    times = rng.uniform(0, 1, len(origins)).astype(np.float32)
    from_ids = np.full(len(origins), -1, np.int32)
    inward = np.zeros(len(origins), np.uint8)
    # bounce rays: from the hit points of the first rays, in random directions, leaving the hit primitive
    ids, t, _ = oracle.twin_closest_hit(scene.name, origins, directions, times, from_ids, inward, width, spp,
                                        earth_width)
    hit = ids >= 0
    p = (origins[hit].astype(np.float64) + t[hit, None].astype(np.float64) * directions[hit]).astype(np.float32)
    m = min(len(p), n - len(origins))
    bo = p[:m]
    bd = rng.normal(size=(m, 3)).astype(np.float32)
    bt = times[hit][:m]
    bf = ids[hit][:m]
    bi = rng.integers(0, 2, m).astype(np.uint8)
    return (np.concatenate([origins, bo]), np.concatenate([directions, bd]), np.concatenate([times, bt]),
            np.concatenate([from_ids, bf]), np.concatenate([inward, bi]))
