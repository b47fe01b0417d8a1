"""The product's flattened scene (rt_scene) as the same vector of numbers oracle.flatten() writes
(oracle/cpu_reference.cpp, oracle_flatten), so the two can be compared bit for bit."""
import numpy as np


def v3(v):
    return [v.x, v.y, v.z]


def sphere(s):
    return v3(s.center) + [s.radius] + v3(s.displacement) + [s.uv_cos, s.uv_sin, s.material]


def quad(q):
    return v3(q.q) + v3(q.u) + v3(q.v) + [q.material]


def product_vector(desc) -> np.ndarray:
    v = [desc.sphere_count, desc.quad_count, desc.material_count, desc.texture_count, desc.medium_count,
         desc.light_count, desc.image_count, desc.perlin_count]
    c = desc.camera
    v += [c.image_width, c.image_height, c.sqrt_spp, c.max_depth]
    v += v3(c.center) + v3(c.pixel00_loc) + v3(c.pixel_delta_u) + v3(c.pixel_delta_v)
    v += [c.defocus, c.sampling]
    v += v3(c.defocus_disk_u) + v3(c.defocus_disk_v) + v3(c.sky_bottom) + v3(c.sky_top)
    for i in range(desc.sphere_count):
        v += sphere(desc.spheres[i])
    for i in range(desc.quad_count):
        v += quad(desc.quads[i])
    for i in range(desc.material_count):
        m = desc.materials[i]
        v += [m.kind, m.texture] + v3(m.albedo) + [m.fuzz, m.refraction_index]
    for i in range(desc.texture_count):
        t = desc.textures[i]
        v += [t.kind, t.even, t.odd, t.image, t.perlin] + v3(t.rgb) + [t.inv_scale, t.scale]
    for i in range(desc.medium_count):
        m = desc.media[i]
        v += [m.boundary, m.material, m.neg_inv_density] + sphere(m.sphere) + [m.quad_count]
        for k in range(m.quad_count):
            v += quad(m.quads[k])
    for i in range(desc.light_count):
        l = desc.lights[i]
        v += [l.kind] + quad(l.quad) + v3(l.center) + [l.radius, l.weight]
    for i in range(desc.image_count):
        im = desc.images[i]
        n = im.width * im.height * 3
        px = np.ctypeslib.as_array(im.rgb, shape=(n,)).astype(np.float64)
        weights = (np.arange(1, n + 1) % 9973).astype(np.float64)
        # the same order of additions as the oracle's loop
        acc = 0.0
        wacc = 0.0
        for x, wt in zip(px.tolist(), weights.tolist()):
            acc += x
            wacc += wt * x
        v += [im.width, im.height, acc, wacc]
    for i in range(desc.perlin_count):
        p = desc.perlins[i]
        for k in range(256):
            v += v3(p.random_vectors[k])
        v += list(p.perm_x) + list(p.perm_y) + list(p.perm_z)
    return np.array(v, dtype=np.float64)
