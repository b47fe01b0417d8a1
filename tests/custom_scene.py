"""A scene of given spheres and quads through the C ABI's rt_scene (for the tie tests' hand-built cases and the shadow
rays' brute-force check): every primitive a gray lambertian, a small camera looking down -z. Spheres are x, y, z, r
rows, or x, y, z, r, dx, dy, dz rows for moving ones; combination: the scene's rt_light_combination (0, the books'
material samples; 2, -p mis, renders it with the quality integrator)."""
import ctypes as C

import numpy as np


class CustomScene:
    def __init__(self, rt, spheres, quads, combination=0):
        self.rt = rt
        sp = np.asarray(spheres, np.float64)
        sp = sp.reshape(-1, sp.shape[-1] if sp.size else 4)
        if sp.shape[1] == 4:
            sp = np.concatenate([sp, np.zeros((len(sp), 3))], axis=1)
        qd = np.asarray(quads, np.float64).reshape(-1, 9)
        self._spheres = (rt.SphereDesc * max(1, len(sp)))()
        for k, (x, y, z, r, dx, dy, dz) in enumerate(sp):
            self._spheres[k] = rt.SphereDesc(rt.Vec3(x, y, z), r, rt.Vec3(dx, dy, dz), 1.0, 0.0, 0, 0)
        self._quads = (rt.QuadDesc * max(1, len(qd)))()
        for k, v in enumerate(qd):
            self._quads[k] = rt.QuadDesc(rt.Vec3(*v[0:3]), rt.Vec3(*v[3:6]), rt.Vec3(*v[6:9]), 0, 0)
        self._materials = (rt.MaterialDesc * 1)(rt.MaterialDesc(0, 0, rt.Vec3(0.5, 0.5, 0.5), 0.0, 1.0))
        self._textures = (rt.TextureDesc * 1)(rt.TextureDesc(0, 0, 0, 0, 0, 0, rt.Vec3(0.5, 0.5, 0.5), 1.0, 1.0))
        cam = rt.CameraDesc()
        cam.image_width, cam.image_height, cam.sqrt_spp, cam.max_depth = 8, 8, 1, 4
        cam.center = rt.Vec3(0, 0, 5)
        cam.pixel00_loc = rt.Vec3(-0.5, 0.5, 4)
        cam.pixel_delta_u = rt.Vec3(0.125, 0, 0)
        cam.pixel_delta_v = rt.Vec3(0, -0.125, 0)
        cam.defocus, cam.sampling = 0, rt.SAMPLING["material"]
        cam.sky_bottom, cam.sky_top = rt.Vec3(1, 1, 1), rt.Vec3(0.5, 0.7, 1.0)
        d = rt.Scene()
        d.spheres, d.sphere_count = C.cast(self._spheres, C.POINTER(rt.SphereDesc)), len(sp)
        d.quads, d.quad_count = C.cast(self._quads, C.POINTER(rt.QuadDesc)), len(qd)
        d.materials, d.material_count = C.cast(self._materials, C.POINTER(rt.MaterialDesc)), 1
        d.textures, d.texture_count = C.cast(self._textures, C.POINTER(rt.TextureDesc)), 1
        d.camera = cam
        d.combination = combination
        self.desc = d
        self.sphere_count, self.quad_count = len(sp), len(qd)
