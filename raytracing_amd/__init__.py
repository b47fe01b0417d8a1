"""MI355X path tracer: Python bindings over the C ABI of ``librt_amd.so`` (include/raytracing_amd.h).

The API mirrors the reference's renderer interface (DaniilK19/RayTracing ``src/gpu/renderer.h``):
``build_scene`` is ``build_gpu_scene`` (``src/gpu_scene.ixx:1333-1445``), ``render`` is ``gpu::render``
(``renderer.h:526``), ``Session`` is ``gpu::session`` (``renderer.h:585-673``). Errors raise the exception types the
reference throws: ``RuntimeError`` (GPU and scene errors), ``ValueError`` (std::invalid_argument: options out of range)
and ``RenderLogicError`` (std::logic_error: state misuse).

The library is the only path: importing this package without the built ``librt_amd.so`` raises ``ImportError``.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# RT_AMD_LIBRARY names another build of the same library (a variant built for a measurement); it must exist too
LIBRARY_PATH = os.environ.get("RT_AMD_LIBRARY") or os.path.join(_HERE, "librt_amd.so")

if not os.path.exists(LIBRARY_PATH):
    raise ImportError(
        f"{LIBRARY_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()' or make -C raytracing_amd)"
    )
_lib = C.CDLL(LIBRARY_PATH)


class RenderLogicError(RuntimeError):
    """std::logic_error of the reference: the session is used in a state that doesn't allow the call."""


# --- the C structs (include/raytracing_amd.h) ------------------------------------------------------------------------
class Vec3(C.Structure):
    _fields_ = [("x", C.c_double), ("y", C.c_double), ("z", C.c_double)]


class TextureDesc(C.Structure):
    _fields_ = [("kind", C.c_int32), ("even", C.c_int32), ("odd", C.c_int32), ("image", C.c_int32),
                ("perlin", C.c_int32), ("reserved", C.c_int32), ("rgb", Vec3), ("inv_scale", C.c_double),
                ("scale", C.c_double)]


class ImageDesc(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("rgb", C.POINTER(C.c_float))]


class PerlinDesc(C.Structure):
    _fields_ = [("random_vectors", Vec3 * 256), ("perm_x", C.c_int32 * 256), ("perm_y", C.c_int32 * 256),
                ("perm_z", C.c_int32 * 256)]


class MaterialDesc(C.Structure):
    _fields_ = [("kind", C.c_int32), ("texture", C.c_int32), ("albedo", Vec3), ("fuzz", C.c_double),
                ("refraction_index", C.c_double)]


class SphereDesc(C.Structure):
    _fields_ = [("center", Vec3), ("radius", C.c_double), ("displacement", Vec3), ("uv_cos", C.c_double),
                ("uv_sin", C.c_double), ("material", C.c_int32), ("reserved", C.c_int32)]


class QuadDesc(C.Structure):
    _fields_ = [("q", Vec3), ("u", Vec3), ("v", Vec3), ("material", C.c_int32), ("reserved", C.c_int32)]


class MediumDesc(C.Structure):
    _fields_ = [("boundary", C.c_int32), ("material", C.c_int32), ("neg_inv_density", C.c_double),
                ("sphere", SphereDesc), ("quads", C.POINTER(QuadDesc)), ("quad_count", C.c_int64)]


class LightDesc(C.Structure):
    _fields_ = [("kind", C.c_int32), ("reserved", C.c_int32), ("quad", QuadDesc), ("center", Vec3),
                ("radius", C.c_double), ("weight", C.c_double)]


class CameraDesc(C.Structure):
    _fields_ = [("image_width", C.c_int32), ("image_height", C.c_int32), ("sqrt_spp", C.c_int32),
                ("max_depth", C.c_int32), ("center", Vec3), ("pixel00_loc", Vec3), ("pixel_delta_u", Vec3),
                ("pixel_delta_v", Vec3), ("defocus", C.c_int32), ("sampling", C.c_int32), ("defocus_disk_u", Vec3),
                ("defocus_disk_v", Vec3), ("sky_bottom", Vec3), ("sky_top", Vec3)]


class MicrofacetDesc(C.Structure):
    _fields_ = [("kind", C.c_int32), ("reserved", C.c_int32), ("roughness", C.c_double)]


class EnvironmentDesc(C.Structure):
    _fields_ = [("image", ImageDesc), ("yaw_degrees", C.c_double), ("scale", C.c_double)]


class MeshDesc(C.Structure):
    _fields_ = [("positions", C.POINTER(C.c_float)), ("vertex_count", C.c_int64),
                ("indices", C.POINTER(C.c_uint32)), ("triangle_count", C.c_int64), ("material", C.c_int32),
                ("reserved", C.c_int32)]


class MeshInstanceDesc(C.Structure):
    _fields_ = [("mesh", C.c_int32), ("reserved", C.c_int32), ("object_to_world", C.c_double * 12)]


class Scene(C.Structure):
    _fields_ = [("spheres", C.POINTER(SphereDesc)), ("sphere_count", C.c_int64),
                ("materials", C.POINTER(MaterialDesc)), ("material_count", C.c_int64),
                ("camera", CameraDesc),
                ("textures", C.POINTER(TextureDesc)), ("texture_count", C.c_int64),
                ("images", C.POINTER(ImageDesc)), ("image_count", C.c_int64),
                ("perlins", C.POINTER(PerlinDesc)), ("perlin_count", C.c_int64),
                ("quads", C.POINTER(QuadDesc)), ("quad_count", C.c_int64),
                ("media", C.POINTER(MediumDesc)), ("medium_count", C.c_int64),
                ("lights", C.POINTER(LightDesc)), ("light_count", C.c_int64),
                ("combination", C.c_int32), ("reserved", C.c_int32),
                ("microfacets", C.POINTER(MicrofacetDesc)), ("microfacet_count", C.c_int64),
                ("environment", C.POINTER(EnvironmentDesc)),
                ("meshes", C.POINTER(MeshDesc)), ("mesh_count", C.c_int64),
                ("instances", C.POINTER(MeshInstanceDesc)), ("instance_count", C.c_int64)]


class SceneOverrides(C.Structure):
    _fields_ = [("image_width", C.c_int32), ("image_height", C.c_int32), ("samples_per_pixel", C.c_int32),
                ("strata", C.c_int32), ("sampling", C.c_int32), ("max_depth", C.c_int32),
                ("earth_width", C.c_int32), ("earth_seed", C.c_uint32)]


class SessionOptions(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("precision", C.c_int32), ("integrator", C.c_int32),
                ("launch_seconds", C.c_double), ("block_size", C.c_int32), ("wavefront_slots", C.c_int32),
                ("wavefront_graph", C.c_int32), ("wavefront_drain", C.c_int32), ("megakernel_waves", C.c_int32),
                ("tree", C.c_int32), ("nodes", C.c_int32), ("kernel", C.c_int32), ("wait", C.c_int32)]


class RenderReport(C.Structure):
    _fields_ = [("passes", C.c_int64), ("rows", C.c_int32), ("launches", C.c_int32), ("gpu_seconds", C.c_double),
                ("longest_launch_seconds", C.c_double), ("bounces", C.c_int64)]


class KernelInfo(C.Structure):
    _fields_ = [("name", C.c_char * 32), ("registers", C.c_int32), ("local_bytes", C.c_int32),
                ("lds_bytes", C.c_int32), ("threads_per_block", C.c_int32), ("blocks_per_cu", C.c_int32),
                ("compute_units", C.c_int32)]


PROGRESS_FN = C.CFUNCTYPE(None, C.c_int64, C.c_int64, C.c_void_p)


class RenderOptions(C.Structure):
    _fields_ = [("session", SessionOptions), ("passes", C.c_int64), ("progress", PROGRESS_FN),
                ("progress_user", C.c_void_p), ("tone_map", C.c_int32), ("exposure", C.c_double),
                ("rgba", C.POINTER(C.c_uint8))]


class RenderStats(C.Structure):
    _fields_ = [("seconds", C.c_double), ("kernel_seconds", C.c_double), ("rays", C.c_uint64),
                ("shadow_rays", C.c_uint64), ("kernel", KernelInfo)]


class Traffic(C.Structure):
    _fields_ = [("paths", C.c_uint64), ("rays", C.c_uint64), ("shadow_rays", C.c_uint64), ("events", C.c_uint64 * 17),
                ("event_bytes", C.c_double * 17), ("bytes_per_path", C.c_double),
                ("phase_lane_cycles", C.c_uint64 * 4)]


TRAFFIC_PHASES = ["queue", "walk", "media", "shade"]
TRAFFIC_EVENTS = ["node", "sphere", "moving_sphere", "quad", "medium", "boundary_quad", "hit", "frame", "texture",
                  "texel", "perlin", "light", "triangle", "lamp", "env_texel", "alias_bin", "instance"]


class ServerOptions(C.Structure):
    _fields_ = [("memory_budget_bytes", C.c_uint64), ("max_active", C.c_int32), ("passes_per_step", C.c_int32),
                ("paths_per_step", C.c_int64)]


class RequestStatus(C.Structure):
    _fields_ = [("id", C.c_int64), ("state", C.c_int32), ("reserved", C.c_int32), ("passes_done", C.c_int64),
                ("passes_total", C.c_int64), ("milestone", C.c_int64), ("submit_time", C.c_double),
                ("start_time", C.c_double), ("milestone_time", C.c_double), ("finish_time", C.c_double)]


class ServerMemory(C.Structure):
    _fields_ = [("budget_bytes", C.c_uint64), ("reserved_bytes", C.c_uint64), ("held_bytes", C.c_uint64)]


class DeviceInfo(C.Structure):
    _fields_ = [("name", C.c_char * 128), ("arch", C.c_char * 64), ("compute_units", C.c_int32),
                ("reserved", C.c_int32), ("memory_gib", C.c_double), ("runtime_version", C.c_int32),
                ("driver_version", C.c_int32)]


# Every function the header declares, with its signature
_P = C.POINTER
_SIGNATURES = {
    "rt_last_error": (C.c_char_p, []),
    "rt_abi_version": (C.c_int, []),
    "rt_scene_names": (C.c_char_p, []),
    "rt_scene_build": (C.c_int, [C.c_char_p, _P(SceneOverrides), _P(C.c_void_p)]),
    "rt_scene_get": (_P(Scene), [C.c_void_p]),
    "rt_scene_passes": (C.c_int64, [C.c_void_p]),
    "rt_scene_warnings": (C.c_char_p, [C.c_void_p]),
    "rt_scene_free": (None, [C.c_void_p]),
    "rt_session_create": (C.c_int, [_P(Scene), _P(SessionOptions), _P(C.c_void_p)]),
    "rt_session_destroy": (None, [C.c_void_p]),
    "rt_session_set_camera": (C.c_int, [C.c_void_p, _P(CameraDesc)]),
    "rt_session_render_to": (C.c_int, [C.c_void_p, C.c_int64, C.c_int32, _P(RenderReport)]),
    "rt_session_camera": (C.c_int, [C.c_void_p, _P(CameraDesc)]),
    "rt_session_has_view": (C.c_int, [C.c_void_p, _P(C.c_int32)]),
    "rt_session_render_to_progress": (C.c_int, [C.c_void_p, C.c_int64, C.c_int32, _P(RenderReport), PROGRESS_FN,
                                                C.c_void_p]),
    "rt_session_render_to_static": (C.c_int, [C.c_void_p, C.c_int64, _P(RenderReport)]),
    "rt_session_render_for": (C.c_int, [C.c_void_p, C.c_double, C.c_int64, _P(RenderReport)]),
    "rt_session_state": (C.c_int, [C.c_void_p, _P(C.c_int64), _P(C.c_int32)]),
    "rt_session_pixels": (C.c_int, [C.c_void_p, _P(C.c_double), C.c_size_t]),
    "rt_session_sums": (C.c_int, [C.c_void_p, _P(C.c_double), C.c_size_t]),
    "rt_session_display_rgba8": (C.c_int, [C.c_void_p, _P(C.c_uint8), C.c_size_t]),
    "rt_session_display_toned": (C.c_int, [C.c_void_p, C.c_int32, C.c_double, _P(C.c_uint8), C.c_size_t]),
    "rt_session_counters": (C.c_int, [C.c_void_p, _P(C.c_uint64), _P(C.c_uint64), _P(C.c_double)]),
    "rt_session_kernel_info": (C.c_int, [C.c_void_p, _P(KernelInfo)]),
    "rt_render": (C.c_int, [_P(Scene), _P(RenderOptions), _P(C.c_double), C.c_size_t, _P(RenderStats)]),
    "rt_session_closest_hit": (C.c_int, [C.c_void_p, _P(C.c_float), _P(C.c_float), _P(C.c_float), _P(C.c_int32),
                                         _P(C.c_uint8), C.c_int64, _P(C.c_int32), _P(C.c_float), _P(C.c_uint8)]),
    "rt_session_occluded": (C.c_int, [C.c_void_p, _P(C.c_float), _P(C.c_float), _P(C.c_float), _P(C.c_int32),
                                      _P(C.c_uint8), _P(C.c_float), _P(C.c_int32), C.c_int64, _P(C.c_uint8),
                                      _P(C.c_uint8)]),
    "rt_session_media_events": (C.c_int, [C.c_void_p, _P(C.c_float), _P(C.c_float), _P(C.c_float), _P(C.c_float),
                                          C.c_int64, C.c_uint64, _P(C.c_int32), _P(C.c_float), _P(C.c_float)]),
    "rt_session_light_samples": (C.c_int, [C.c_void_p, _P(C.c_float), C.c_int64, C.c_uint64, _P(C.c_float),
                                           _P(C.c_float), _P(C.c_float), _P(C.c_float), _P(C.c_int32),
                                           _P(C.c_float)]),
    "rt_session_reference_hits": (C.c_int, [C.c_void_p, _P(C.c_float), _P(C.c_float), _P(C.c_float),
                                            _P(C.c_int32), _P(C.c_uint8), C.c_int64, _P(C.c_int32), _P(C.c_float),
                                            _P(C.c_uint8)]),
    "rt_sample_numbers": (C.c_int, [C.c_uint64, _P(C.c_uint64), _P(C.c_uint64), C.c_int64, C.c_int32,
                                    _P(C.c_uint32), _P(C.c_float), _P(C.c_double), _P(C.c_float)]),
    "rt_tone_probe": (C.c_int, [C.c_int32, C.c_double, _P(C.c_float), _P(C.c_float), C.c_int64, _P(C.c_uint8)]),
    "rt_scatter_probe": (C.c_int, [C.c_int32, C.c_float, C.c_float, _P(C.c_float), C.c_int64, _P(C.c_float),
                                   _P(C.c_float), C.c_uint64, _P(C.c_float), _P(C.c_float), _P(C.c_float),
                                   _P(C.c_float), _P(C.c_float)]),
    "rt_session_traffic": (C.c_int, [C.c_void_p, C.c_int64, _P(Traffic)]),
    "rt_scene_wide_nodes": (C.c_int, [_P(Scene), C.c_int32, C.c_void_p, C.c_void_p, _P(C.c_int64)]),
    "rt_scene_bvh": (C.c_int, [_P(Scene), C.c_int32, _P(C.c_float), _P(C.c_int32), _P(C.c_int64), _P(C.c_int32),
                               _P(C.c_int32), _P(C.c_int32)]),
    "rt_scene_lamps": (C.c_int, [_P(Scene), _P(C.c_float), _P(C.c_int64)]),
    "rt_environment_tables": (C.c_int, [_P(EnvironmentDesc), _P(C.c_float), _P(C.c_float), _P(C.c_float),
                                        _P(C.c_int32), _P(C.c_int32), _P(C.c_double)]),
    "rt_server_create": (C.c_int, [_P(ServerOptions), _P(C.c_void_p)]),
    "rt_server_destroy": (None, [C.c_void_p]),
    "rt_server_submit": (C.c_int, [C.c_void_p, _P(Scene), C.c_uint64, C.c_int64, _P(C.c_int64)]),
    "rt_server_add_scene": (C.c_int, [C.c_void_p, _P(Scene), _P(C.c_int64)]),
    "rt_server_remove_scene": (C.c_int, [C.c_void_p, C.c_int64]),
    "rt_server_submit_view": (C.c_int, [C.c_void_p, C.c_int64, _P(CameraDesc), C.c_uint64, C.c_int64,
                                        _P(C.c_int64)]),
    "rt_server_step": (C.c_int, [C.c_void_p, _P(C.c_int32)]),
    "rt_server_status": (C.c_int, [C.c_void_p, C.c_int64, _P(RequestStatus)]),
    "rt_server_memory_use": (C.c_int, [C.c_void_p, _P(ServerMemory)]),
    "rt_server_pixels": (C.c_int, [C.c_void_p, C.c_int64, _P(C.c_double), C.c_size_t]),
    "rt_server_release": (C.c_int, [C.c_void_p, C.c_int64]),
    "rt_query_device": (C.c_int, [_P(DeviceInfo)]),
    "rt_query_device_wait": (C.c_int, [_P(DeviceInfo), C.c_int32]),
    "rt_copy_probe": (C.c_int, [C.c_uint64, C.c_int32, _P(C.c_double)]),
}
for _name, (_res, _args) in _SIGNATURES.items():
    _f = getattr(_lib, _name)
    _f.restype = _res
    _f.argtypes = _args

RT_OK, RT_ERROR_RUNTIME, RT_ERROR_INVALID_ARGUMENT, RT_ERROR_LOGIC = 0, 1, 2, 3
SAMPLING = {"uniform": 0, "material": 1, "lights": 2, "mixture": 3, "nee": 4, "mis": 5}
MATERIAL_KINDS = {"lambertian": 0, "metal": 1, "dielectric": 2, "diffuse_light": 3, "isotropic": 4}
INTEGRATORS = {"auto": 0, "megakernel": 1, "wavefront": 2, "plain": 3}
PRECISIONS = {"float32": 0, "fp32": 0, "float64": 1, "fp64": 1, "float64_all": 2}
TREES = {"default": 0, "median": 1, "sah": 2, "two_level": 3}
NODES = {"auto": 0, "float": 1, "packed": 2}
KERNELS = {"auto": 0, "all_books": 1, "quality": 2}
WAITS = {"spin": 0, "sleep": 1}
TONE_MAPS = {"books": 0, "none": 1, "agx": 2, "neutral": 3}


def library() -> C.CDLL:
    return _lib


def last_error() -> str:
    return _lib.rt_last_error().decode()


def _check(status: int) -> None:
    if status == RT_OK:
        return
    message = last_error()
    if status == RT_ERROR_INVALID_ARGUMENT:
        raise ValueError(message)
    if status == RT_ERROR_LOGIC:
        raise RenderLogicError(message)
    raise RuntimeError(message)


def scene_names() -> list[str]:
    return _lib.rt_scene_names().decode().split(",")


class BuiltScene:
    """A scene built by name and flattened (build_gpu_scene): ``desc`` is the rt_scene view over its tables. A name that
    ends in .gltf or .glb is a glTF 2.0 model's file, rendered in a studio; ``warnings`` lists what the model holds
    that isn't rendered."""

    def __init__(self, name: str, width: int = 0, height: int = 0, samples: int = 0, strata: int = 0,
                 sampling: Optional[str] = None, max_depth: int = 0, earth_width: int = 0, earth_seed: int = 0):
        o = SceneOverrides(width, height, samples, strata, -1 if sampling is None else SAMPLING[sampling], max_depth,
                           earth_width, earth_seed)
        handle = C.c_void_p()
        _check(_lib.rt_scene_build(name.encode(), C.byref(o), C.byref(handle)))
        self._handle = handle
        self.name = name
        self.desc: Scene = _lib.rt_scene_get(handle).contents
        self.passes: int = _lib.rt_scene_passes(handle)
        self.warnings: list = [w for w in _lib.rt_scene_warnings(handle).decode().split("\n") if w]

    @property
    def width(self) -> int:
        return self.desc.camera.image_width

    @property
    def height(self) -> int:
        return self.desc.camera.image_height

    def __del__(self):
        if getattr(self, "_handle", None):
            _lib.rt_scene_free(self._handle)
            self._handle = None


def build_scene(name: str, **overrides) -> BuiltScene:
    return BuiltScene(name, **overrides)


def _session_options(seed=0, precision="float32", integrator="auto", launch_seconds=0.0, block_size=0,
                     wavefront_slots=0, wavefront_graph=0, wavefront_drain=0, megakernel_waves=0,
                     tree="default", nodes="auto", kernel="auto", wait="spin") -> SessionOptions:
    return SessionOptions(seed, PRECISIONS[precision], INTEGRATORS[integrator],
                          launch_seconds, block_size, wavefront_slots, wavefront_graph, wavefront_drain,
                          megakernel_waves, TREES[tree] if isinstance(tree, str) else int(tree),
                          NODES[nodes] if isinstance(nodes, str) else int(nodes),
                          KERNELS[kernel] if isinstance(kernel, str) else int(kernel),
                          WAITS[wait] if isinstance(wait, str) else int(wait))


@dataclass
class RenderResult:
    """gpu::render_result (renderer.h:475-492)"""
    pixels: np.ndarray  # (height, width, 3) float64, rows from the top
    seconds: float
    kernel_seconds: float
    rays: int
    shadow_rays: int
    kernel: str = ""             # the kernel that rendered
    rgba: np.ndarray = None      # (height, width, 4) display bytes, for a tone map other than the books'


def _progress_fn(progress):
    """A C progress callback calling progress(done, total), or a null one"""
    if progress is None:
        return PROGRESS_FN()
    return PROGRESS_FN(lambda done, total, _user: progress(done, total))


def render(scene: BuiltScene, passes: int = 0, progress=None, tone_map: str = "books", exposure: float = 0.0,
           **options) -> RenderResult:
    """gpu::render: a session rendered to ``passes`` (0: the scene's own samples); progress(done, total), if given, is
    called as passes complete; a tone map other than the books' also gives the display's bytes (render_result::rgba)."""
    desc = scene.desc
    h, w = desc.camera.image_height, desc.camera.image_width
    out = np.zeros((h, w, 3), dtype=np.float64)
    tm = TONE_MAPS[tone_map] if isinstance(tone_map, str) else int(tone_map)
    rgba = np.zeros((h, w, 4), np.uint8) if tm != 0 else None
    callback = _progress_fn(progress)  # (kept alive through the call)
    ro = RenderOptions(_session_options(**options), passes if passes else scene.passes, callback, None, tm, exposure,
                       rgba.ctypes.data_as(_P(C.c_uint8)) if rgba is not None else None)
    stats = RenderStats()
    _check(_lib.rt_render(C.byref(desc), C.byref(ro), out.ctypes.data_as(_P(C.c_double)), out.size, C.byref(stats)))
    return RenderResult(out, stats.seconds, stats.kernel_seconds, stats.rays, stats.shadow_rays,
                        stats.kernel.name.decode(), rgba)


class Session:
    """gpu::session: the scene stays on the GPU and the image of one view gains passes on demand."""

    def __init__(self, scene, **options):
        desc = getattr(scene, "desc", scene)  # a built scene, anything holding an rt_scene as .desc, or the struct
        self._keep = scene
        handle = C.c_void_p()
        _check(_lib.rt_session_create(C.byref(desc), C.byref(_session_options(**options)), C.byref(handle)))
        self._h = handle
        self.camera = CameraDesc.from_buffer_copy(desc.camera)

    def close(self):
        if getattr(self, "_h", None):
            _lib.rt_session_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_camera(self, camera: CameraDesc) -> None:
        _check(_lib.rt_session_set_camera(self._h, C.byref(camera)))
        self.camera = CameraDesc.from_buffer_copy(camera)

    def view(self) -> CameraDesc:
        """The last view set successfully (session::camera)"""
        c = CameraDesc()
        _check(_lib.rt_session_camera(self._h, C.byref(c)))
        return c

    def has_view(self) -> bool:
        v = C.c_int32()
        _check(_lib.rt_session_has_view(self._h, C.byref(v)))
        return bool(v.value)

    def render_to(self, passes: int, rows: int = 0, progress=None) -> RenderReport:
        """Renders until the image holds ``passes`` passes and ``rows`` rows of the next; progress(done, total), if
        given, is called as passes complete."""
        r = RenderReport()
        if progress is None:
            _check(_lib.rt_session_render_to(self._h, passes, rows, C.byref(r)))
        else:
            callback = _progress_fn(progress)
            _check(_lib.rt_session_render_to_progress(self._h, passes, rows, C.byref(r), callback, None))
        return r

    def render_to_static(self, passes: int) -> RenderReport:
        """render_to's whole passes without device allocation (the library's static scratch)"""
        r = RenderReport()
        _check(_lib.rt_session_render_to_static(self._h, passes, C.byref(r)))
        return r

    def render_for(self, seconds: float, pass_limit: int) -> RenderReport:
        r = RenderReport()
        _check(_lib.rt_session_render_for(self._h, seconds, pass_limit, C.byref(r)))
        return r

    def state(self) -> tuple[int, int]:
        p, r = C.c_int64(), C.c_int32()
        _check(_lib.rt_session_state(self._h, C.byref(p), C.byref(r)))
        return p.value, r.value

    def _image(self, fn) -> np.ndarray:
        out = np.zeros((self.camera.image_height, self.camera.image_width, 3), dtype=np.float64)
        _check(fn(self._h, out.ctypes.data_as(_P(C.c_double)), out.size))
        return out

    def pixels(self) -> np.ndarray:
        return self._image(_lib.rt_session_pixels)

    def sums(self) -> np.ndarray:
        return self._image(_lib.rt_session_sums)

    def display(self, tone_map: str = "books", exposure: float = 0.0) -> np.ndarray:
        """RGBA bytes: the books' (sqrt, clamp 0.999, x256) or a tone map's ("none", "agx", "neutral") at an exposure
        in stops"""
        out = np.zeros((self.camera.image_height, self.camera.image_width, 4), dtype=np.uint8)
        if tone_map == "books" and exposure == 0:
            _check(_lib.rt_session_display_rgba8(self._h, out.ctypes.data_as(_P(C.c_uint8)), out.size))
        else:
            op = TONE_MAPS[tone_map] if isinstance(tone_map, str) else int(tone_map)
            _check(_lib.rt_session_display_toned(self._h, op, exposure, out.ctypes.data_as(_P(C.c_uint8)), out.size))
        return out

    def counters(self) -> tuple[int, int, float]:
        rays, shadow, seconds = C.c_uint64(), C.c_uint64(), C.c_double()
        _check(_lib.rt_session_counters(self._h, C.byref(rays), C.byref(shadow), C.byref(seconds)))
        return rays.value, shadow.value, seconds.value

    def kernel(self) -> KernelInfo:
        k = KernelInfo()
        _check(_lib.rt_session_kernel_info(self._h, C.byref(k)))
        return k

    def traffic(self, passes: int) -> Traffic:
        """The algorithmic scene traffic of `passes` passes, counted by the instrumented megakernel."""
        t = Traffic()
        _check(_lib.rt_session_traffic(self._h, passes, C.byref(t)))
        return t

    def closest_hit(self, origins, directions, times, from_ids, inward):
        """The session's BVH traversal on given rays: (ids, t, entering)."""
        return self._hits(_lib.rt_session_closest_hit, origins, directions, times, from_ids, inward)

    def reference_hits(self, origins, directions, times, from_ids, inward):
        """The reference's own walk of the session's binary tree, with its tie rule: (ids, t, entering)."""
        return self._hits(_lib.rt_session_reference_hits, origins, directions, times, from_ids, inward)

    def media_events(self, origins, directions, times, t_surface, seed: int = 0):
        """The media's event along each ray up to t_surface: (medium index or -1, t or inf, transmittance)"""
        o = np.ascontiguousarray(origins, dtype=np.float32)
        d = np.ascontiguousarray(directions, dtype=np.float32)
        tm = np.ascontiguousarray(times, dtype=np.float32)
        ts = np.ascontiguousarray(t_surface, dtype=np.float32)
        n = len(tm)
        medium = np.zeros(n, np.int32)
        t = np.zeros(n, np.float32)
        tr = np.zeros(n, np.float32)
        f = lambda a: a.ctypes.data_as(_P(C.c_float))
        _check(_lib.rt_session_media_events(self._h, f(o), f(d), f(tm), f(ts), n, seed,
                                            medium.ctypes.data_as(_P(C.c_int32)), f(t), f(tr)))
        return medium, t, tr

    def light_samples(self, points, seed: int = 0) -> dict:
        """One light-set draw from each point: wi, distance, radiance, density, skip (-1 the map, -2 failed) and
        hit_density (the density a scattered ray finding the same light is given)"""
        p = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
        n = len(p)
        out = {"wi": np.zeros((n, 3), np.float32), "distance": np.zeros(n, np.float32),
               "radiance": np.zeros((n, 3), np.float32), "density": np.zeros(n, np.float32),
               "skip": np.zeros(n, np.int32), "hit_density": np.zeros(n, np.float32)}
        f = lambda a: a.ctypes.data_as(_P(C.c_float))
        _check(_lib.rt_session_light_samples(self._h, f(p), n, seed, f(out["wi"]), f(out["distance"]),
                                             f(out["radiance"]), f(out["density"]),
                                             out["skip"].ctypes.data_as(_P(C.c_int32)), f(out["hit_density"])))
        return out

    def occluded(self, origins, directions, times, from_ids, inward, t_max, skip):
        """Shadow rays through both shadow walks: (blocked, walked), each 1 where something but `skip` lies in
        (0.001, t_max)."""
        o = np.ascontiguousarray(origins, dtype=np.float32)
        d = np.ascontiguousarray(directions, dtype=np.float32)
        tm = np.ascontiguousarray(times, dtype=np.float32)
        f = np.ascontiguousarray(from_ids, dtype=np.int32)
        inw = np.ascontiguousarray(inward, dtype=np.uint8)
        tx = np.ascontiguousarray(t_max, dtype=np.float32)
        sk = np.ascontiguousarray(skip, dtype=np.int32)
        n = len(tm)
        blocked = np.zeros(n, np.uint8)
        walked = np.zeros(n, np.uint8)
        _check(_lib.rt_session_occluded(self._h, o.ctypes.data_as(_P(C.c_float)), d.ctypes.data_as(_P(C.c_float)),
                                        tm.ctypes.data_as(_P(C.c_float)), f.ctypes.data_as(_P(C.c_int32)),
                                        inw.ctypes.data_as(_P(C.c_uint8)), tx.ctypes.data_as(_P(C.c_float)),
                                        sk.ctypes.data_as(_P(C.c_int32)), n, blocked.ctypes.data_as(_P(C.c_uint8)),
                                        walked.ctypes.data_as(_P(C.c_uint8))))
        return blocked, walked

    def _hits(self, fn, origins, directions, times, from_ids, inward):
        o = np.ascontiguousarray(origins, dtype=np.float32)
        d = np.ascontiguousarray(directions, dtype=np.float32)
        tm = np.ascontiguousarray(times, dtype=np.float32)
        f = np.ascontiguousarray(from_ids, dtype=np.int32)
        inw = np.ascontiguousarray(inward, dtype=np.uint8)
        n = len(tm)
        ids = np.zeros(n, np.int32)
        t = np.zeros(n, np.float32)
        ent = np.zeros(n, np.uint8)
        _check(fn(self._h, o.ctypes.data_as(_P(C.c_float)), d.ctypes.data_as(_P(C.c_float)),
                  tm.ctypes.data_as(_P(C.c_float)), f.ctypes.data_as(_P(C.c_int32)), inw.ctypes.data_as(_P(C.c_uint8)),
                  n, ids.ctypes.data_as(_P(C.c_int32)), t.ctypes.data_as(_P(C.c_float)),
                  ent.ctypes.data_as(_P(C.c_uint8))))
        return ids, t, ent


def sample_numbers(seed: int, pixels, passes, draws: int):
    """The per-sample generator's outputs on the device for items (pixels[k], passes[k]): (words, floats, doubles,
    directions), words/floats/doubles of shape (n, draws), directions (n, 3, 3): a unit vector, a disk point and a
    cosine-weighted direction about +z"""
    px = np.ascontiguousarray(pixels, dtype=np.uint64)
    ps = np.ascontiguousarray(passes, dtype=np.uint64)
    n = len(px)
    words = np.zeros((n, draws), np.uint32)
    floats = np.zeros((n, draws), np.float32)
    doubles = np.zeros((n, draws), np.float64)
    dirs = np.zeros((n, 3, 3), np.float32)
    _check(_lib.rt_sample_numbers(seed, px.ctypes.data_as(_P(C.c_uint64)), ps.ctypes.data_as(_P(C.c_uint64)), n,
                                  draws, words.ctypes.data_as(_P(C.c_uint32)), floats.ctypes.data_as(_P(C.c_float)),
                                  doubles.ctypes.data_as(_P(C.c_double)), dirs.ctypes.data_as(_P(C.c_float))))
    return words, floats, doubles, dirs


def tone_probe(tone_map, colours, exposure: float = 0.0, dither=None) -> np.ndarray:
    """Linear colours (n x 3) through a tone map ("none", "agx", "neutral") at an exposure, quantized with a dither
    per colour (None: none), on device 0: n x 3 bytes"""
    c = np.ascontiguousarray(colours, dtype=np.float32).reshape(-1, 3)
    d = None if dither is None else np.ascontiguousarray(dither, dtype=np.float32)
    out = np.zeros((len(c), 3), np.uint8)
    _check(_lib.rt_tone_probe(TONE_MAPS[tone_map] if isinstance(tone_map, str) else int(tone_map), exposure,
                              c.ctypes.data_as(_P(C.c_float)), d.ctypes.data_as(_P(C.c_float)) if d is not None else None,
                              len(c), out.ctypes.data_as(_P(C.c_uint8))))
    return out


LOBES = {"lambert": 0, "phase": 1, "conductor": 2, "gloss": 3, "metal": 4, "glass_entering": 5, "glass_leaving": 6}


def scatter_probe(lobe, v, l, color=(1.0, 1.0, 1.0), alpha: float = 0.0, roughness: float = 0.0, seed: int = 0):
    """The quality path's scattering lobe on given cases in the tangent frame (normal +z), on device 0: a dict of
    value (f cos_l, n x 3) and density of l seen from v, and drawn, weight (f cos / density) and drawn_density of a
    direction drawn from v by the generator of (seed, case, 0)"""
    v = np.ascontiguousarray(v, dtype=np.float32).reshape(-1, 3)
    l = np.ascontiguousarray(l, dtype=np.float32).reshape(-1, 3)
    if len(v) != len(l):
        raise ValueError("as many directions l as v")
    n = len(v)
    c = (C.c_float * 3)(*color)
    out = {"value": np.zeros((n, 3), np.float32), "density": np.zeros(n, np.float32),
           "drawn": np.zeros((n, 3), np.float32), "weight": np.zeros((n, 3), np.float32),
           "drawn_density": np.zeros(n, np.float32)}
    f = lambda a: a.ctypes.data_as(_P(C.c_float))
    _check(_lib.rt_scatter_probe(LOBES[lobe] if isinstance(lobe, str) else int(lobe), alpha, roughness, c, n, f(v),
                                 f(l), seed, f(out["value"]), f(out["density"]), f(out["drawn"]), f(out["weight"]),
                                 f(out["drawn_density"])))
    return out


def scene_bvh(scene: BuiltScene, tree: str = "median"):
    """The float32 BVH this library builds for a scene with the splits of `tree` ("default", "median" for the
    reference's, "sah"), in the reference's node layout (host only)."""
    count = C.c_int64()
    depth = C.c_int32()
    kind = TREES[tree] if isinstance(tree, str) else int(tree)
    _check(_lib.rt_scene_bvh(C.byref(scene.desc), kind, None, None, C.byref(count), None, None, C.byref(depth)))
    boxes = np.zeros((count.value, 6), np.float32)
    ints = np.zeros((count.value, 2), np.int32)
    sphere_ids = np.zeros(scene.desc.sphere_count, np.int32)
    quad_ids = np.zeros(scene.desc.quad_count, np.int32)
    _check(_lib.rt_scene_bvh(C.byref(scene.desc), kind, boxes.ctypes.data_as(_P(C.c_float)),
                             ints.ctypes.data_as(_P(C.c_int32)), C.byref(count),
                             sphere_ids.ctypes.data_as(_P(C.c_int32)), quad_ids.ctypes.data_as(_P(C.c_int32)),
                             C.byref(depth)))
    return boxes, ints, sphere_ids, quad_ids, depth.value


WIDE_NODE = np.dtype([("lo", np.float32, (3, 4)), ("hi", np.float32, (3, 4)), ("ref", np.int32, 4),
                      ("unused", np.int32, 4)])
PACKED_NODE = np.dtype([("origin", np.float32, 3), ("step", np.float32, 3), ("lo", np.uint8, (3, 4)),
                        ("hi", np.uint8, (3, 4)), ("ref", np.int32, 4)])


def scene_wide_nodes(scene: BuiltScene, tree: str = "default"):
    """The four-wide tree the kernels walk (float32, host only): its 128-byte float nodes and the same nodes packed
    into 64 bytes as RT_NODES_PACKED lays them out, as numpy structured arrays (WIDE_NODE, PACKED_NODE)."""
    count = C.c_int64()
    kind = TREES[tree] if isinstance(tree, str) else int(tree)
    _check(_lib.rt_scene_wide_nodes(C.byref(scene.desc), kind, None, None, C.byref(count)))
    wide = np.zeros(count.value, WIDE_NODE)
    packed = np.zeros(count.value, PACKED_NODE)
    assert WIDE_NODE.itemsize == 128 and PACKED_NODE.itemsize == 64
    _check(_lib.rt_scene_wide_nodes(C.byref(scene.desc), kind, wide.ctypes.data, packed.ctypes.data,
                                    C.byref(count)))
    return wide, packed


LAMP_FIELDS = 20  # floats per lamp record: corner[3], size, u[3], kind, v[3], primitive, normal[3], 0, radiance[3], 0


def scene_lamps(scene: BuiltScene) -> np.ndarray:
    """The quality integrator's light set (lamps.cuh build_lamps), one row of 20 float32 per lamp (host only)."""
    count = C.c_int64()
    _check(_lib.rt_scene_lamps(C.byref(scene.desc), None, C.byref(count)))
    out = np.zeros((count.value, LAMP_FIELDS), np.float32)
    _check(_lib.rt_scene_lamps(C.byref(scene.desc), out.ctypes.data_as(_P(C.c_float)), C.byref(count)))
    return out


@dataclass
class EnvironmentTables:
    texels: np.ndarray   # (height, width, 4): r, g, b, density per steradian
    row_cos: np.ndarray  # (height + 1,)
    bin_q: np.ndarray    # (2^log2_bins,) coins
    bin_alias: np.ndarray
    log2_bins: int
    total: float


def environment_tables(rgb: np.ndarray, yaw_degrees: float = 0.0, scale: float = 1.0) -> EnvironmentTables:
    """An environment map's host tables (environment_data.h build_environment) for an (H, W, 3) float32 image."""
    rgb = np.ascontiguousarray(rgb, dtype=np.float32)
    h, w = rgb.shape[:2]
    env = EnvironmentDesc(ImageDesc(w, h, rgb.ctypes.data_as(_P(C.c_float))), yaw_degrees, scale)
    log2 = C.c_int32()
    total = C.c_double()
    _check(_lib.rt_environment_tables(C.byref(env), None, None, None, None, C.byref(log2), C.byref(total)))
    bins = (1 << log2.value) if log2.value > 0 else 0
    texels = np.zeros((h, w, 4), np.float32)
    row_cos = np.zeros(h + 1, np.float32)
    q = np.zeros(max(bins, 1), np.float32)
    alias = np.zeros(max(bins, 1), np.int32)
    _check(_lib.rt_environment_tables(C.byref(env), texels.ctypes.data_as(_P(C.c_float)),
                                      row_cos.ctypes.data_as(_P(C.c_float)), q.ctypes.data_as(_P(C.c_float)),
                                      alias.ctypes.data_as(_P(C.c_int32)), C.byref(log2), C.byref(total)))
    return EnvironmentTables(texels, row_cos, q[:bins], alias[:bins], log2.value, total.value)


def query_device(wait=None) -> DeviceInfo:
    """Describes device 0 and starts the runtime; wait ("spin" or "sleep") also sets how the device's own waits wait"""
    info = DeviceInfo()
    if wait is None:
        _check(_lib.rt_query_device(C.byref(info)))
    else:
        _check(_lib.rt_query_device_wait(C.byref(info), WAITS[wait] if isinstance(wait, str) else int(wait)))
    return info


def copy_probe(nbytes: int, repeats: int = 5) -> float:
    """Average seconds of a plain device-to-device copy kernel over `nbytes` (the HBM counters' calibration)."""
    seconds = C.c_double()
    _check(_lib.rt_copy_probe(nbytes, repeats, C.byref(seconds)))
    return seconds.value


class Server:
    """Progressive renders of mixed requests on one device, admitted by a memory budget (continuous batching)."""

    def __init__(self, memory_budget_bytes=0, max_active=0, passes_per_step=0, paths_per_step=0):
        h = C.c_void_p()
        _check(_lib.rt_server_create(C.byref(ServerOptions(memory_budget_bytes, max_active, passes_per_step,
                                                           paths_per_step)),
                                     C.byref(h)))
        self._h = h
        self._scenes = {}

    def close(self):
        if getattr(self, "_h", None):
            _lib.rt_server_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def submit(self, scene: BuiltScene, seed: int, passes: int) -> int:
        rid = C.c_int64()
        _check(_lib.rt_server_submit(self._h, C.byref(scene.desc), seed, passes, C.byref(rid)))
        self._scenes[rid.value] = (scene.width, scene.height)
        return rid.value

    def add_scene(self, scene: BuiltScene) -> int:
        """Makes a scene resident (uploaded once, shared by its views)"""
        sid = C.c_int64()
        _check(_lib.rt_server_add_scene(self._h, C.byref(scene.desc), C.byref(sid)))
        return sid.value

    def remove_scene(self, scene_id: int) -> None:
        _check(_lib.rt_server_remove_scene(self._h, scene_id))

    def submit_view(self, scene_id: int, camera: CameraDesc, seed: int, passes: int) -> int:
        rid = C.c_int64()
        _check(_lib.rt_server_submit_view(self._h, scene_id, C.byref(camera), seed, passes, C.byref(rid)))
        self._scenes[rid.value] = (camera.image_width, camera.image_height)
        return rid.value

    def step(self) -> int:
        active = C.c_int32()
        _check(_lib.rt_server_step(self._h, C.byref(active)))
        return active.value

    def memory(self) -> ServerMemory:
        m = ServerMemory()
        _check(_lib.rt_server_memory_use(self._h, C.byref(m)))
        return m

    def status(self, rid: int) -> RequestStatus:
        st = RequestStatus()
        _check(_lib.rt_server_status(self._h, rid, C.byref(st)))
        return st

    def pixels(self, rid: int) -> np.ndarray:
        w, h = self._scenes[rid]
        out = np.zeros((h, w, 3), np.float64)
        _check(_lib.rt_server_pixels(self._h, rid, out.ctypes.data_as(_P(C.c_double)), out.size))
        return out

    def release(self, rid: int) -> None:
        _check(_lib.rt_server_release(self._h, rid))
