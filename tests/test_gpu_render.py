"""GPU parity tests: the HIP kernels through the C ABI against the oracle.

- closest-hit ids, distances and sides of seeded ray sets, bit for bit against the GPU-semantics twin;
- the furnace scene (src/gpu_scenes.ixx:33-55): every pixel exactly 1.0, in every kernel and precision;
- ImageCompare of GPU renders against the CPU restatement of the reference renderer, with the reference's re-run rule
  (scripts/verify/compare.ps1: a second, independent pair of seeds decides when the first pair fails);
- byte reproducibility: a fixed seed gives the same sums whatever the launch grouping or session history.
"""
import os
import zlib

import numpy as np
import pytest

import oracle
from oracle import image_compare
from tests.rays import scene_rays

pytestmark = pytest.mark.gpu


def _compare_with_rerun(name, width, spp, sampling=None, precision="float32", integrator="auto", depth=0,
                        earth_width=512):
    import raytracing_amd as rt

    samp = -1 if sampling is None else rt.SAMPLING[sampling]
    verdicts = []
    for cpu_seed, gpu_seed in ((0, 1), (2, 3)):
        scene = rt.build_scene(name, width=width, samples=spp, sampling=sampling, max_depth=depth,
                               earth_width=earth_width)
        gpu = rt.render(scene, seed=gpu_seed, precision=precision, integrator=integrator).pixels
        cpu, _ = oracle.render(name, width=width, spp=spp, seed=cpu_seed, sampling=samp, max_depth=depth,
                               earth_width=earth_width)
        v = image_compare.compare(cpu, gpu)
        verdicts.append(v)
        print(f"{name} {width}w {spp}spp {precision} {integrator}: CPU {cpu.mean():.6f} GPU {gpu.mean():.6f}: "
              f"{v.summary()}", flush=True)
        if v.consistent:
            return v
    raise AssertionError(f"{name}: GPU and CPU differ beyond noise twice: " + " | ".join(v.summary() for v in verdicts))


@pytest.mark.parametrize("name", ["weekend", "bouncing", "glass", "final", "cornell", "smoke", "light"])
def test_closest_hit_bit_exact(rt, name):
    """The GPU walks the reference's median tree four-wide; the twin walks it two-wide in the reference's order. Every
    ray's primitive, distance (bits) and side are the same, ties included (the lower scene index wins in both)."""
    scene = rt.build_scene(name, width=64, samples=16, earth_width=512)
    o, d, tm, f, inw = scene_rays(scene, 60000, seed=zlib.crc32(name.encode()) % 1000, width=64, spp=16, earth_width=512)
    with rt.Session(scene, integrator="megakernel", tree="median") as s:
        ids, t, ent = s.closest_hit(o, d, tm, f, inw)
    tids, tt, tent = oracle.twin_closest_hit(name, o, d, tm, f, inw, 64, 16, 512)
    assert (tids >= 0).sum() > len(tids) // 4, "the ray set should hit something"
    mism = np.nonzero((ids != tids) | (t.view(np.uint32) != tt.view(np.uint32)) | (ent != tent))[0]
    assert len(mism) == 0, f"{len(mism)} of {len(ids)} rays differ, first {mism[:5]}: gpu {ids[mism[:5]]} {t[mism[:5]]} twin {tids[mism[:5]]} {tt[mism[:5]]}"


def _to_tree(rt, scene, ids, frm: str, to: str):
    """Global primitive ids of one tree as the ids of the same spheres and quads in another (-1 stays -1)."""
    _, _, s_from, q_from, _ = rt.scene_bvh(scene, frm)
    _, _, s_to, q_to, _ = rt.scene_bvh(scene, to)
    table = np.full(len(s_from) + len(q_from) + 1, -1, np.int64)
    table[np.concatenate([s_from, q_from])] = np.concatenate([s_to, q_to])
    return np.where(ids >= 0, table[np.maximum(ids, 0)], -1).astype(np.int32)


@pytest.mark.parametrize("name", ["weekend", "final", "cornell", "smoke"])
def test_surface_area_tree_finds_the_same_hits(rt, name):
    """The surface-area tree tests other boxes in another order, but every ray's closest hit is the same primitive at
    the same distance (bit for bit) on the same side as in the reference's median tree, ties included."""
    scene = rt.build_scene(name, width=64, samples=16, earth_width=512)
    o, d, tm, f, inw = scene_rays(scene, 60000, seed=7 + zlib.crc32(name.encode()) % 1000, width=64, spp=16,
                                  earth_width=512)
    with rt.Session(scene, integrator="megakernel", tree="median") as s:
        ids_m, t_m, ent_m = s.closest_hit(o, d, tm, f, inw)
    with rt.Session(scene, integrator="megakernel", tree="sah") as s:
        ids_s, t_s, ent_s = s.closest_hit(o, d, tm, _to_tree(rt, scene, f, "median", "sah"), inw)
    assert np.array_equal(t_m.view(np.uint32), t_s.view(np.uint32))
    assert np.array_equal(ids_m, _to_tree(rt, scene, ids_s, "sah", "median"))
    assert np.array_equal(ent_m, ent_s)


@pytest.mark.parametrize("precision", ["float32", "float64"])
@pytest.mark.parametrize("integrator", ["megakernel", "auto"])
def test_furnace_exactly_one(rt, precision, integrator):
    scene = rt.build_scene("furnace", width=240, samples=16)
    img = rt.render(scene, seed=7, precision=precision, integrator=integrator).pixels
    assert np.all(img == 1.0), f"{np.count_nonzero(img != 1.0)} values differ from 1, e.g. {img[img != 1.0][:5]}"


def test_weekend_config0_vs_cpu(rt):
    """Config 0: weekend, 400x225, 64 spp, depth 50, against the CPU render."""
    _compare_with_rerun("weekend", 400, 64)


@pytest.mark.parametrize("name,width,spp,sampling", [
    ("materials", 160, 64, None), ("bouncing", 160, 64, None), ("checker", 160, 64, None), ("earth", 160, 64, None),
    ("perlin", 160, 64, None), ("quads", 120, 64, None), ("light", 160, 64, None), ("cornell", 120, 64, None),
    ("smoke", 120, 64, None), ("final", 120, 64, None), ("aluminum", 120, 64, None), ("glass", 120, 64, None),
    ("glass", 120, 64, "material"), ("cornell", 120, 64, "lights"), ("cornell", 120, 64, "uniform"),
    ("final", 120, 64, "material"),
])
def test_scene_vs_cpu(rt, name, width, spp, sampling):
    _compare_with_rerun(name, width, spp, sampling)


def test_final_benchmark_texture_vs_cpu(rt):
    """The final scene with the benchmark's own 2048x1024 noise texture (the other tests use a 512-wide one), at a
    reduced resolution: ImageCompare against the CPU"""
    _compare_with_rerun("final", 160, 64, earth_width=2048)


def test_fp64_vs_cpu(rt):
    _compare_with_rerun("final", 100, 64, precision="float64")
    _compare_with_rerun("weekend", 160, 64, precision="float64")


def test_same_seed_same_bytes(rt):
    scene = rt.build_scene("final", width=96, samples=16, earth_width=512)
    with rt.Session(scene, seed=5, launch_seconds=0.5) as a:
        a.render_to(16)
        sa = a.sums()
    with rt.Session(scene, seed=5, launch_seconds=0.0001) as b:
        for p in (1, 2, 5, 9, 16):
            b.render_to(p)
        sb = b.sums()
    with rt.Session(scene, seed=5) as c:
        c.render_to(7, 40)  # a band of rows of pass 7
        c.render_to(16)
        sc = c.sums()
    assert np.array_equal(sa.view(np.uint64), sb.view(np.uint64))
    assert np.array_equal(sa.view(np.uint64), sc.view(np.uint64))
    with rt.Session(scene, seed=6) as d:
        d.render_to(16)
        assert not np.array_equal(sa, d.sums())


def test_progress_callbacks(rt):
    """render_options::progress and session::render_to's on_pass (src/gpu/renderer.h:423-427): called after launches
    that complete passes, with the passes done (increasing, the last the target) and the total; the image is the one
    rendered without them"""
    scene = rt.build_scene("final", width=96, samples=16, earth_width=512)
    calls = []
    # (the megakernel alone: the autotuner's first launches may take all 16 passes at once)
    r = rt.render(scene, seed=5, integrator="megakernel", launch_seconds=0.0001,
                  progress=lambda done, total: calls.append((done, total)))
    assert len(calls) >= 2 and calls[-1] == (16, 16)
    assert all(total == 16 for _, total in calls)
    assert all(a[0] < b[0] for a, b in zip(calls, calls[1:]))
    plain = rt.render(scene, seed=5)
    assert np.array_equal(r.pixels.view(np.uint64), plain.pixels.view(np.uint64))
    with rt.Session(scene, seed=5, integrator="megakernel", launch_seconds=0.0001) as s:
        seen = []
        s.render_to(7, 40, progress=lambda done, total: seen.append((done, total)))  # pass 7's band of rows
        assert seen and seen[-1] == (7, 7)
        seen.clear()
        s.render_to(16, progress=lambda done, total: seen.append((done, total)))
        assert seen[0][0] == 8 and seen[-1] == (16, 16), seen  # the band's pass completes first
        assert np.array_equal(s.pixels().view(np.uint64), plain.pixels.view(np.uint64))


def test_rounds_of_strata(rt):
    """More passes than one grid: round r draws seed + r * 2^32 (src/gpu/renderer.h:564-584)"""
    scene = rt.build_scene("weekend", width=64, samples=50, strata=4)
    assert scene.passes == 50 and scene.desc.camera.sqrt_spp == 4
    r = rt.render(scene, seed=0)
    with rt.Session(scene, seed=0) as s:
        s.render_to(16)
        first_round = s.sums()
        s.render_to(50)
        assert np.array_equal(s.pixels(), r.pixels)
    with rt.Session(scene, seed=1 << 32) as s2:
        s2.render_to(16)  # round 1 of seed 0 is round 0 of seed 2^32
        round1 = s2.sums()
    with rt.Session(scene, seed=0) as s3:
        s3.render_to(32)
        assert np.allclose(s3.sums(), first_round + round1, rtol=0, atol=1e-12)


def test_session_errors(rt):
    scene = rt.build_scene("glass", width=32, samples=4)
    s = rt.Session(scene)
    with pytest.raises(rt.RenderLogicError):
        s.pixels()
    s.render_to(2)
    with pytest.raises(ValueError):
        s.render_to(1)
    cam = rt.CameraDesc.from_buffer_copy(scene.desc.camera)
    cam.sampling = rt.SAMPLING["material"]
    with pytest.raises(ValueError):
        s.set_camera(cam)
    # session::camera and has_view: a refused camera keeps the old view; a new one starts from no samples
    assert s.has_view() and bytes(s.view()) == bytes(scene.desc.camera)
    wider = rt.CameraDesc.from_buffer_copy(scene.desc.camera)
    wider.image_width, wider.image_height = 48, 40
    s.set_camera(wider)
    assert s.has_view() and s.view().image_width == 48 and s.state() == (0, 0)
    bad = rt.CameraDesc.from_buffer_copy(wider)
    bad.image_width = 0
    with pytest.raises(RuntimeError, match="at least one pixel"):  # (the reference's std::runtime_error)
        s.set_camera(bad)
    assert s.has_view() and bytes(s.view()) == bytes(wider)
    s.render_to(1)
    assert s.pixels().shape == (40, 48, 3)
    s.close()


def test_quality_scene_refusals(rt):
    """What the quality integrator refuses (tests/routing_tests.cu:53-67): float64, and the lights list's strategies
    (uniform, lights, mixture) on a scene that needs it (GGX here); the books' kernels take both"""
    ggx = rt.build_scene("cornell-ggx", width=32, samples=4)
    with pytest.raises(ValueError, match="float32 only"):
        rt.Session(ggx, precision="float64")
    ggx.desc.camera.sampling = rt.SAMPLING["uniform"]
    with pytest.raises(ValueError, match="lights list"):
        rt.Session(ggx)
    for strategy in ("lights", "mixture"):  # without a lights list the scene itself is refused first
        ggx.desc.camera.sampling = rt.SAMPLING[strategy]
        with pytest.raises(RuntimeError, match="need a lights list"):
            rt.Session(ggx)
    books = rt.build_scene("cornell", width=32, samples=4, sampling="mixture")
    with rt.Session(books, precision="float64") as s:
        s.render_to(1)


def test_display_bytes(rt):
    scene = rt.build_scene("weekend", width=80, samples=4)
    with rt.Session(scene) as s:
        s.render_to(4)
        px = s.pixels()
        rgba = s.display()
    expect = (256 * np.clip(np.sqrt(np.maximum(px, 0)), 0, 0.999)).astype(np.uint8)
    assert np.array_equal(rgba[:, :, :3], expect)
    assert np.all(rgba[:, :, 3] == 255)


def test_traffic_counts(rt):
    scene = rt.build_scene("final", width=128, samples=16, earth_width=512)
    with rt.Session(scene, integrator="megakernel") as s:
        t = s.traffic(4)
        rays, _, _ = s.counters()
    assert t.paths == 4 * 128 * 128
    assert t.rays > t.paths
    events = dict(zip(rt.TRAFFIC_EVENTS, list(t.events)))
    assert events["node"] > t.rays  # several nodes per ray
    assert events["medium"] >= t.rays  # final's fog is tested by every ray
    assert 100 < t.bytes_per_path < 1e6
    assert rays == 0  # the view's counters are untouched


@pytest.mark.parametrize("name", ["weekend", "final", "glass", "cornell", "smoke", "bouncing", "earth", "perlin",
                                  "light", "materials"])
def test_wavefront_equals_megakernel(rt, name):
    """The two integrators trace the same paths with the same random numbers: the same sums, bit for bit, and the
    same ray counts"""
    scene = rt.build_scene(name, width=96, samples=16, earth_width=512)
    with rt.Session(scene, seed=3, integrator="megakernel") as a:
        a.render_to(16)
        sa = a.sums()
        ra = a.counters()[0]
    with rt.Session(scene, seed=3, integrator="wavefront") as b:
        assert b.kernel().name.decode().startswith("wavefront")
        b.render_to(5)
        b.render_to(16)
        sb = b.sums()
        rb = b.counters()[0]
    assert ra == rb
    mism = np.count_nonzero(sa.view(np.uint64) != sb.view(np.uint64))
    assert mism == 0, f"{mism} of {sa.size} sums differ; max |diff| {np.abs(sa - sb).max()}"


@pytest.mark.parametrize("name,width,spp,seed", [("weekend", 1080, 47, 500), ("glass", 600, 64, 7)])
def test_wavefront_equals_megakernel_at_size(rt, name, width, spp, seed):
    """Many paths (weekend: 55 M, where the two once parted on 5 pixels: a first-book scene's matte lobe, which always
    goes on, shaded by the wavefront with the third book's rule, which ends a path drawn below the surface); the same
    sums bit for bit"""
    scene = rt.build_scene(name, width=width, samples=1)
    with rt.Session(scene, seed=seed, integrator="megakernel") as a:
        a.render_to(spp)
        sa = a.sums()
    with rt.Session(scene, seed=seed, integrator="wavefront") as b:
        b.render_to(spp)
        sb = b.sums()
    mism = np.count_nonzero(np.any(sa.view(np.uint64) != sb.view(np.uint64), axis=-1))
    assert mism == 0, f"{mism} pixels differ"


@pytest.mark.parametrize("name", ["weekend", "final", "cornell"])
@pytest.mark.parametrize("drain,graph", [(-1, 0), (1 << 30, 0), (0, -1), (5000, 2)])
def test_wavefront_drain_and_graph_equal(rt, name, drain, graph):
    """The drain (slots finishing their passes on their own), direct launches and short graphs: the same sums"""
    scene = rt.build_scene(name, width=96, samples=16, earth_width=512)
    with rt.Session(scene, seed=8, integrator="megakernel") as a:
        a.render_to(16)
        sa = a.sums()
        ra = a.counters()[0]
    with rt.Session(scene, seed=8, integrator="wavefront", wavefront_drain=drain, wavefront_graph=graph) as b:
        r = b.render_to(16)
        sb = b.sums()
        rb = b.counters()[0]
    assert r.bounces > 0
    assert ra == rb
    assert np.array_equal(sa.view(np.uint64), sb.view(np.uint64))


def test_wavefront_small_pool(rt):
    """A pool smaller than the image: the launches go in bands of rows, and the image is the same"""
    scene = rt.build_scene("final", width=64, samples=16, earth_width=512)
    with rt.Session(scene, seed=2, integrator="megakernel") as a:
        a.render_to(16)
        sa = a.sums()
    with rt.Session(scene, seed=2, integrator="wavefront", wavefront_slots=64 * 10) as b:
        b.render_to(16)
        sb = b.sums()
    assert np.array_equal(sa.view(np.uint64), sb.view(np.uint64))


@pytest.mark.parametrize("name,width", [("final", 37), ("weekend", 100), ("cornell", 50)])
def test_megakernel_queue_shapes(rt, name, width):
    """The megakernel's chunk queue (16x4 tiles x groups of 4 passes): image sides that aren't multiples of the tile,
    pass counts that aren't multiples of the group, launches of one pass up to all of them, a band of rows, every
    register budget and block size: the same sums as the wavefront, bit for bit, and the same ray counts"""
    scene = rt.build_scene(name, width=width, samples=7, strata=4, earth_width=512)
    assert scene.passes == 7
    with rt.Session(scene, seed=11, integrator="wavefront") as w:
        w.render_to(7)
        reference = w.sums()
        rays = w.counters()[0]
    variants = [dict(), dict(megakernel_waves=5, block_size=128), dict(megakernel_waves=6, block_size=256),
                dict(block_size=192, launch_seconds=0.0001)]
    for options in variants:
        with rt.Session(scene, seed=11, integrator="megakernel", **options) as m:
            m.render_to(1)
            m.render_to(3, 5)  # a band of the third pass's rows
            m.render_to(7)
            got = m.sums()
            assert m.counters()[0] == rays, options
        mism = np.count_nonzero(got.view(np.uint64) != reference.view(np.uint64))
        assert mism == 0, f"{options}: {mism} of {got.size} sums differ"


def test_megakernel_fp64_launch_splits(rt):
    """FP64 (megakernel only): one launch or a launch per pass, the same bytes"""
    scene = rt.build_scene("final", width=37, samples=6, strata=4, earth_width=512)
    with rt.Session(scene, seed=4, precision="float64", launch_seconds=0.5) as a:
        a.render_to(6)
        sa = a.sums()
    with rt.Session(scene, seed=4, precision="float64", launch_seconds=0.0001) as b:
        for p in (1, 2, 6):
            b.render_to(p)
        sb = b.sums()
    assert np.array_equal(sa.view(np.uint64), sb.view(np.uint64))


@pytest.mark.parametrize("name,sampling,precision", [("weekend", None, "float32"), ("final", None, "float32"),
                                                     ("cornell", None, "float32"), ("final", None, "float64"),
                                                     ("cornell-ggx", "mis", "float32"),
                                                     ("triangles", "mis", "float32")])
def test_plain_equals_megakernel(rt, name, sampling, precision):
    """The plain megakernel (the reference's mapping: a thread per pixel of a 16x8 block, a launch's passes in its
    loop, src/gpu/renderer.cu:340-397) and the persistent queue megakernel: the same sums bit for bit and the same ray
    and shadow-ray counts, on an image side that isn't a multiple of the block, over launches of several passes and a
    band of rows"""
    scene = rt.build_scene(name, width=37, samples=9, strata=3, sampling=sampling, earth_width=512)
    with rt.Session(scene, seed=5, integrator="megakernel", precision=precision) as a:
        a.render_to(9)
        sa = a.sums()
        ca = a.counters()
    with rt.Session(scene, seed=5, integrator="plain", precision=precision) as b:
        assert b.kernel().name.decode().startswith("plain"), b.kernel().name
        b.render_to(2)
        b.render_to(4, 7)  # a band of the fifth pass's rows
        b.render_to(9)
        sb = b.sums()
        cb = b.counters()
    assert ca[0] == cb[0] and ca[1] == cb[1], f"rays {ca[:2]} vs {cb[:2]}"
    mism = np.count_nonzero(sa.view(np.uint64) != sb.view(np.uint64))
    assert mism == 0, f"{mism} of {sa.size} sums differ; max |diff| {np.abs(sa - sb).max()}"


def test_all_books_kernel_on_first_book_scenes(rt):
    """RT_KERNEL_ALL_BOOKS (the reference's all_books_kernel, src/gpu/renderer.h:428-431): the three books' shading
    renders the first book's scenes (still spheres of the basic materials under -p material) to the same image, bit for
    bit, with the cost of its generality"""
    for name, width, spp in (("weekend", 96, 16), ("furnace", 64, 16)):
        scene = rt.build_scene(name, width=width, samples=spp, sampling="material")
        sums = {}
        for kernel, expect in (("auto", "first book"), ("all_books", "all books")):
            with rt.Session(scene, seed=7, integrator="megakernel", kernel=kernel) as s:
                s.render_to(spp)
                sums[kernel] = s.sums()
                assert s.kernel().name.decode().startswith(expect), s.kernel().name
        differ = np.count_nonzero(sums["auto"].view(np.uint64) != sums["all_books"].view(np.uint64))
        assert differ == 0, f"{name}: {differ} of {sums['auto'].size} sums differ"


def test_quality_kernel_on_books_scenes(rt):
    """RT_KERNEL_QUALITY (the reference's quality_kernel): the quality integrator renders the books' scenes under
    -p material to the books' image within its noise (ImageCompare, with the reference's re-run rule)"""
    for name, width, spp in (("cornell", 96, 64), ("final", 96, 64), ("smoke", 96, 64), ("weekend", 96, 32)):
        scene = rt.build_scene(name, width=width, samples=spp, earth_width=512, sampling="material")
        verdicts = []
        for seeds in ((1, 2), (3, 4)):
            imgs = []
            for kernel, seed in zip(("auto", "quality"), seeds):
                with rt.Session(scene, seed=seed, integrator="megakernel", kernel=kernel) as s:
                    s.render_to(spp)
                    imgs.append(s.pixels())
                    assert ("quality" in s.kernel().name.decode()) == (kernel == "quality")
            v = image_compare.compare(imgs[0], imgs[1])
            verdicts.append(v)
            if v.consistent:
                break
        assert verdicts[-1].consistent, f"{name}: " + " | ".join(v.summary() for v in verdicts)


def test_kernel_refusals(rt):
    with pytest.raises(ValueError, match="can't both render"):
        rt.Session(rt.build_scene("cornell", width=32, samples=4, sampling="mis"), kernel="all_books")
    with pytest.raises(ValueError, match="float only"):
        rt.Session(rt.build_scene("weekend", width=32, samples=4), kernel="quality", precision="float64")
    with pytest.raises(ValueError):
        rt.Session(rt.build_scene("cornell", width=32, samples=4, sampling="mixture"), kernel="quality")


SLEEPING_SESSION = """
import numpy as np, sys
sys.path.insert(0, {root!r})
import raytracing_amd as rt
info = rt.query_device(wait="sleep")
scene = rt.build_scene("weekend", width=64, samples=16)
with rt.Session(scene, seed=3, integrator="megakernel", wait="sleep") as s:
    s.render_to(16)
    a = s.sums()
with rt.Session(scene, seed=3, integrator="megakernel") as s:
    s.render_to(16)
    b = s.sums()
assert np.array_equal(a.view(np.uint64), b.view(np.uint64))
print("sleeping session ok", info.compute_units)
"""


def test_wait_modes(rt):
    """session_options::wait and query_device(wait) (src/gpu/renderer.h:505-540): a session that sleeps renders the
    same sums as one that spins, once the device's waits sleep too (set before the runtime starts: in a process of its
    own), and is refused with a logic error before that"""
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, "-c", SLEEPING_SESSION.format(root=root)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0 and "sleeping session ok" in r.stdout, r.stdout + r.stderr
    with pytest.raises(rt.RenderLogicError, match="sleep"):
        rt.Session(rt.build_scene("weekend", width=32, samples=4), wait="sleep")
    with pytest.raises(ValueError):
        rt.Session(rt.build_scene("weekend", width=32, samples=4), wait=5)


def test_render_for(rt):
    """session::render_for (src/gpu/renderer.h:606-614): about `seconds` of GPU time, never past pass_limit, at least
    one complete pass, nothing once the limit is reached, no launch shorter than a quarter of a full one after the
    first; and the launch planner keeps launches near their planned length (src/gpu/renderer.cu:201-308)"""
    scene = rt.build_scene("final", width=240, samples=4096, earth_width=512)
    with rt.Session(scene, seed=1, integrator="megakernel", launch_seconds=0.02) as s:
        r = s.render_for(0.0, 1000)
        assert r.passes >= 1 and r.rows == 0, "at least one complete pass"
        before = s.state()[0]
        r = s.render_for(0.25, 10 ** 6)
        assert r.passes > before
        assert 0.25 * 0.75 <= r.gpu_seconds <= 0.25 + 2 * r.longest_launch_seconds, (r.gpu_seconds, r.launches)
        assert r.longest_launch_seconds < 4 * 0.02, f"a launch of {r.longest_launch_seconds:.3f} s planned at 0.02"
        r = s.render_for(1.0, s.state()[0])
        assert r.launches == 0, "a pass limit the image has reached does nothing"
        limit = s.state()[0] + 3
        r = s.render_for(10.0, limit)
        assert r.passes == limit and s.state() == (limit, 0), "never past the pass limit"
        for seconds, limit in ((-1.0, 10), (4000.0, 10), (1.0, -1)):
            with pytest.raises(ValueError):
                s.render_for(seconds, limit)
