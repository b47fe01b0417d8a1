"""ORACLE — test infrastructure only.

The checker the tests, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg use: a C++ restatement of the
reference's CPU renderer (cpu_model.hpp, cpu_reference.cpp), the GPU-semantics twin for bit-exact closest hits and
random numbers (gpu_twin.cpp), and ImageCompare (image_compare.py). Nothing in the product imports this package.

Parity status per function (the checks that pin each are in tests/test_oracle.py and tests/test_abi.py):
  random numbers (mt19937_64, PCG32, SplitMix64)   pinned by published known answers
  scene building and flattening                    pinned against the product's independent builder, bit for bit, and
                                                   the reference's documented counts and values
  CPU renderer (ray_color)                         pinned by the furnace closed form (exactly 1), CPU-vs-CPU
                                                   consistency, a planted fault, and the reference's published mean
                                                   radiance of final (docs/gpu.md:66)
  twin closest_hit, scene-order ties               pinned by brute force over all primitives
  twin closest_hit, the reference's walk-order ties pinned by brute force away from ties, by the tied set on tied
                                                   rays, and by hand-built scenes whose tie outcome the reference's
                                                   intervals and near-first order fix (tests/test_oracle.py)
  twin sample words, floats and doubles            pinned by the PCG32 demo stream and SplitMix64 known answers, and
                                                   the draws' bit layout (floor(f 2^24) = w >> 8, floor(d 2^32) = w)
  ImageCompare                                     pinned by synthetic images: equal expectations pass, a 0.1% bias and
                                                   a shifted image fail
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def _load() -> C.CDLL:
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    P = C.POINTER
    lib.oracle_last_error.restype = C.c_char_p
    lib.oracle_render.restype = C.c_int
    lib.oracle_render.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint32, C.c_int, C.c_int,
                                  C.c_uint32, C.c_int, C.c_int, P(C.c_double), C.c_int64, P(C.c_int32),
                                  P(C.c_uint64)]
    lib.oracle_flatten.restype = C.c_int64
    lib.oracle_flatten.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint32,
                                   P(C.c_double), C.c_int64]
    lib.oracle_mt19937_64.argtypes = [C.c_uint64, C.c_int64, P(C.c_uint64)]
    lib.oracle_random_doubles.argtypes = [C.c_uint64, C.c_int64, P(C.c_double)]
    lib.oracle_noise_image.argtypes = [C.c_int, C.c_uint32, P(C.c_float)]
    lib.twin_last_error.restype = C.c_char_p
    lib.twin_pcg32.argtypes = [C.c_uint64, C.c_uint64, C.c_int64, P(C.c_uint32)]
    lib.twin_mix64.restype = C.c_uint64
    lib.twin_mix64.argtypes = [C.c_uint64]
    lib.twin_sample_floats.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, C.c_int64, P(C.c_float)]
    lib.twin_bvh.restype = C.c_int
    lib.twin_bvh.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, P(C.c_float), P(C.c_int32), P(C.c_int64),
                             P(C.c_int32), P(C.c_int32), P(C.c_int64), P(C.c_int32)]
    lib.twin_closest_hit.restype = C.c_int
    lib.twin_closest_hit.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, P(C.c_float), P(C.c_float),
                                     P(C.c_float), P(C.c_int32), P(C.c_uint8), C.c_int64, P(C.c_int32),
                                     P(C.c_float), P(C.c_uint8), P(C.c_uint64)]
    lib.twin_closest_hit_walk_order.restype = C.c_int
    lib.twin_closest_hit_walk_order.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, P(C.c_float), P(C.c_float),
                                                P(C.c_float), P(C.c_int32), P(C.c_uint8), C.c_int64, P(C.c_int32),
                                                P(C.c_float), P(C.c_uint8)]
    lib.twin_custom_closest_hit.restype = C.c_int
    lib.twin_custom_closest_hit.argtypes = [P(C.c_double), C.c_int64, P(C.c_double), C.c_int64, C.c_int,
                                            P(C.c_float), P(C.c_float), C.c_int64, P(C.c_int32), P(C.c_float),
                                            P(C.c_uint8), P(C.c_int32), P(C.c_int32)]
    lib.twin_sample_doubles.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, C.c_int64, P(C.c_double)]
    lib.twin_sample_words.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, C.c_int64, P(C.c_uint32)]
    lib.twin_brute_force.restype = C.c_int
    lib.twin_brute_force.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, P(C.c_float), P(C.c_float),
                                     P(C.c_float), P(C.c_int32), P(C.c_uint8), C.c_int64, P(C.c_int32),
                                     P(C.c_float), P(C.c_int32)]
    lib.twin_custom_occluded.restype = C.c_int
    lib.twin_custom_occluded.argtypes = [P(C.c_double), C.c_int64, P(C.c_double), C.c_int64, P(C.c_float),
                                         P(C.c_float), P(C.c_float), P(C.c_int32), P(C.c_uint8), P(C.c_float),
                                         P(C.c_int32), C.c_int64, P(C.c_uint8), P(C.c_uint8), P(C.c_uint8),
                                         P(C.c_int32), P(C.c_int32)]
    return lib


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


def _ptr(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def render(name: str, width: int = 0, spp: int = 0, seed: int = 0, sampling: int = -1, max_depth: int = 0,
           threads: int = 0, rows: tuple | None = None, earth_width: int = 0, earth_seed: int = 0, height: int = 0):
    """The reference's CPU render (double, mt19937_64 per row): (image (h, w, 3), rays). rows=(begin, end) renders
    only those rows, each exactly as in the whole image. height > 0: that many rows, the view widened or narrowed
    (the reference's --height)."""
    L = lib()
    dims = np.zeros(3, np.int32)
    if L.oracle_render(name.encode(), width, height, spp, sampling, max_depth, seed, threads, earth_width, earth_seed,
                       0, 0,
                       None, 0, _ptr(dims, C.c_int32), None):
        raise RuntimeError(L.oracle_last_error().decode())
    w, h = int(dims[0]), int(dims[1])
    r0, r1 = rows if rows else (0, h)
    out = np.zeros((r1 - r0, w, 3), np.float64)
    rays = C.c_uint64()
    if L.oracle_render(name.encode(), width, height, spp, sampling, max_depth, seed, threads, earth_width, earth_seed,
                       r0, r1,
                       _ptr(out, C.c_double), out.size, _ptr(dims, C.c_int32), C.byref(rays)):
        raise RuntimeError(L.oracle_last_error().decode())
    return out, rays.value


def frame(name: str, width: int = 0, spp: int = 0):
    """(width, height, sqrt_spp) of a scene with the overrides"""
    L = lib()
    dims = np.zeros(3, np.int32)
    if L.oracle_render(name.encode(), width, 0, spp, -1, 0, 0, 1, 0, 0, 0, 0, None, 0, _ptr(dims, C.c_int32), None):
        raise RuntimeError(L.oracle_last_error().decode())
    return tuple(int(x) for x in dims)


def flatten(name: str, width: int = 0, spp: int = 0, sampling: int = -1, max_depth: int = 0, earth_width: int = 0,
            earth_seed: int = 0) -> np.ndarray:
    L = lib()
    n = L.oracle_flatten(name.encode(), width, spp, sampling, max_depth, earth_width, earth_seed, None, 0)
    if n < 0:
        raise RuntimeError(L.oracle_last_error().decode())
    out = np.zeros(n, np.float64)
    L.oracle_flatten(name.encode(), width, spp, sampling, max_depth, earth_width, earth_seed, _ptr(out, C.c_double), n)
    return out


def mt19937_64(seed: int, n: int) -> np.ndarray:
    out = np.zeros(n, np.uint64)
    lib().oracle_mt19937_64(seed, n, _ptr(out, C.c_uint64))
    return out


def random_doubles(seed: int, n: int) -> np.ndarray:
    out = np.zeros(n, np.float64)
    lib().oracle_random_doubles(seed, n, _ptr(out, C.c_double))
    return out


def noise_image(width: int, seed: int) -> np.ndarray:
    out = np.zeros((width // 2, width, 3), np.float32)
    lib().oracle_noise_image(width, seed, _ptr(out, C.c_float))
    return out


def pcg32(initstate: int, initseq: int, n: int) -> np.ndarray:
    out = np.zeros(n, np.uint32)
    lib().twin_pcg32(initstate, initseq, n, _ptr(out, C.c_uint32))
    return out


def mix64(z: int) -> int:
    return int(lib().twin_mix64(z))


def sample_floats(seed: int, pixel: int, pass_: int, n: int) -> np.ndarray:
    out = np.zeros(n, np.float32)
    lib().twin_sample_floats(seed, pixel, pass_, n, _ptr(out, C.c_float))
    return out


def twin_bvh(name: str, width: int = 0, spp: int = 0, earth_width: int = 0):
    L = lib()
    count = C.c_int64()
    counts = np.zeros(2, np.int64)
    depth = C.c_int32()
    if L.twin_bvh(name.encode(), width, spp, earth_width, None, None, C.byref(count), None, None,
                  _ptr(counts, C.c_int64), C.byref(depth)):
        raise RuntimeError(L.twin_last_error().decode())
    boxes = np.zeros((count.value, 6), np.float32)
    ints = np.zeros((count.value, 2), np.int32)
    sid = np.zeros(counts[0], np.int32)
    qid = np.zeros(counts[1], np.int32)
    L.twin_bvh(name.encode(), width, spp, earth_width, _ptr(boxes, C.c_float), _ptr(ints, C.c_int32), C.byref(count),
               _ptr(sid, C.c_int32), _ptr(qid, C.c_int32), _ptr(counts, C.c_int64), C.byref(depth))
    return boxes, ints, sid, qid, depth.value


def twin_closest_hit(name: str, origins, directions, times, from_ids, inward, width: int = 0, spp: int = 0,
                     earth_width: int = 0, stats: bool = False):
    L = lib()
    o = np.ascontiguousarray(origins, np.float32)
    d = np.ascontiguousarray(directions, np.float32)
    tm = np.ascontiguousarray(times, np.float32)
    f = np.ascontiguousarray(from_ids, np.int32)
    inw = np.ascontiguousarray(inward, np.uint8)
    n = len(tm)
    ids = np.zeros(n, np.int32)
    t = np.zeros(n, np.float32)
    ent = np.zeros(n, np.uint8)
    st = np.zeros(4, np.uint64)
    if L.twin_closest_hit(name.encode(), width, spp, earth_width, _ptr(o, C.c_float), _ptr(d, C.c_float),
                          _ptr(tm, C.c_float), _ptr(f, C.c_int32), _ptr(inw, C.c_uint8), n, _ptr(ids, C.c_int32),
                          _ptr(t, C.c_float), _ptr(ent, C.c_uint8), _ptr(st, C.c_uint64) if stats else None):
        raise RuntimeError(L.twin_last_error().decode())
    return (ids, t, ent, st) if stats else (ids, t, ent)


def twin_closest_hit_walk_order(name: str, origins, directions, times, from_ids, inward, width: int = 0,
                                spp: int = 0, earth_width: int = 0):
    """twin_closest_hit under the reference's tie rule: the binary near-first walk keeps the first sphere met at a
    distance (open interval) and the last quad (closed interval), src/gpu/bvh.cuh:576-706, sphere.cuh:116, quad.cuh:70"""
    L = lib()
    o = np.ascontiguousarray(origins, np.float32)
    d = np.ascontiguousarray(directions, np.float32)
    tm = np.ascontiguousarray(times, np.float32)
    f = np.ascontiguousarray(from_ids, np.int32)
    inw = np.ascontiguousarray(inward, np.uint8)
    n = len(tm)
    ids = np.zeros(n, np.int32)
    t = np.zeros(n, np.float32)
    ent = np.zeros(n, np.uint8)
    if L.twin_closest_hit_walk_order(name.encode(), width, spp, earth_width, _ptr(o, C.c_float), _ptr(d, C.c_float),
                                     _ptr(tm, C.c_float), _ptr(f, C.c_int32), _ptr(inw, C.c_uint8), n,
                                     _ptr(ids, C.c_int32), _ptr(t, C.c_float), _ptr(ent, C.c_uint8)):
        raise RuntimeError(L.twin_last_error().decode())
    return ids, t, ent


def twin_custom_closest_hit(spheres, quads, origins, directions, walk_order: bool):
    """Closest hits in a scene of plain spheres (x, y, z, r rows) and quads (q, u, v rows): (ids, t, entering,
    sphere_ids, quad_ids), ids in the twin tree's numbering"""
    L = lib()
    sp = np.ascontiguousarray(np.asarray(spheres, np.float64).reshape(-1, 4))
    qd = np.ascontiguousarray(np.asarray(quads, np.float64).reshape(-1, 9))
    o = np.ascontiguousarray(origins, np.float32).reshape(-1, 3)
    d = np.ascontiguousarray(directions, np.float32).reshape(-1, 3)
    n = len(o)
    ids = np.zeros(n, np.int32)
    t = np.zeros(n, np.float32)
    ent = np.zeros(n, np.uint8)
    sid = np.zeros(len(sp), np.int32)
    qid = np.zeros(len(qd), np.int32)
    if L.twin_custom_closest_hit(_ptr(sp, C.c_double), len(sp), _ptr(qd, C.c_double), len(qd), 1 if walk_order else 0,
                                 _ptr(o, C.c_float), _ptr(d, C.c_float), n, _ptr(ids, C.c_int32), _ptr(t, C.c_float),
                                 _ptr(ent, C.c_uint8), _ptr(sid, C.c_int32), _ptr(qid, C.c_int32)):
        raise RuntimeError(L.twin_last_error().decode())
    return ids, t, ent, sid, qid


def twin_custom_occluded(spheres, quads, origins, directions, times, from_index, inward, t_max, skip_index):
    """Shadow rays by brute force in a scene of spheres (x, y, z, r, dx, dy, dz rows; a zero displacement stands still)
    and quads (q, u, v rows); from_index and skip_index name input primitives (spheres, then quads) or -1:
    (blocked, band, near, sphere_ids, quad_ids), the ids in the twin tree's numbering"""
    L = lib()
    sp = np.ascontiguousarray(np.asarray(spheres, np.float64).reshape(-1, 7))
    qd = np.ascontiguousarray(np.asarray(quads, np.float64).reshape(-1, 9))
    o = np.ascontiguousarray(origins, np.float32).reshape(-1, 3)
    d = np.ascontiguousarray(directions, np.float32).reshape(-1, 3)
    tm = np.ascontiguousarray(times, np.float32)
    f = np.ascontiguousarray(from_index, np.int32)
    inw = np.ascontiguousarray(inward, np.uint8)
    tx = np.ascontiguousarray(t_max, np.float32)
    sk = np.ascontiguousarray(skip_index, np.int32)
    n = len(o)
    blocked = np.zeros(n, np.uint8)
    band = np.zeros(n, np.uint8)
    near = np.zeros(n, np.uint8)
    sid = np.zeros(len(sp), np.int32)
    qid = np.zeros(len(qd), np.int32)
    if L.twin_custom_occluded(_ptr(sp, C.c_double), len(sp), _ptr(qd, C.c_double), len(qd), _ptr(o, C.c_float),
                              _ptr(d, C.c_float), _ptr(tm, C.c_float), _ptr(f, C.c_int32), _ptr(inw, C.c_uint8),
                              _ptr(tx, C.c_float), _ptr(sk, C.c_int32), n, _ptr(blocked, C.c_uint8),
                              _ptr(band, C.c_uint8), _ptr(near, C.c_uint8), _ptr(sid, C.c_int32), _ptr(qid, C.c_int32)):
        raise RuntimeError(L.twin_last_error().decode())
    return blocked, band, near, sid, qid


def sample_doubles(seed: int, pixel: int, pass_: int, n: int) -> np.ndarray:
    out = np.zeros(n, np.float64)
    lib().twin_sample_doubles(seed, pixel, pass_, n, _ptr(out, C.c_double))
    return out


def sample_words(seed: int, pixel: int, pass_: int, n: int) -> np.ndarray:
    out = np.zeros(n, np.uint32)
    lib().twin_sample_words(seed, pixel, pass_, n, _ptr(out, C.c_uint32))
    return out


def twin_brute_force(name: str, origins, directions, times, from_ids, inward, width: int = 0, spp: int = 0,
                     earth_width: int = 0):
    L = lib()
    o = np.ascontiguousarray(origins, np.float32)
    d = np.ascontiguousarray(directions, np.float32)
    tm = np.ascontiguousarray(times, np.float32)
    f = np.ascontiguousarray(from_ids, np.int32)
    inw = np.ascontiguousarray(inward, np.uint8)
    n = len(tm)
    ids = np.zeros(n, np.int32)
    t = np.zeros(n, np.float32)
    tied = np.zeros(n, np.int32)
    if L.twin_brute_force(name.encode(), width, spp, earth_width, _ptr(o, C.c_float), _ptr(d, C.c_float),
                          _ptr(tm, C.c_float), _ptr(f, C.c_int32), _ptr(inw, C.c_uint8), n, _ptr(ids, C.c_int32),
                          _ptr(t, C.c_float), _ptr(tied, C.c_int32)):
        raise RuntimeError(L.twin_last_error().decode())
    return ids, t, tied
