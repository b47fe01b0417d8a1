"""GPU tests of the quality integrator (light samples weighed by MIS, GGX, the environment map, triangles), through
the C ABI. What pins each part:

- closed forms: env-furnace and mesh-furnace are exactly 1 in every pixel under -p material; env-furnace's background
  stays exactly 1 under -p mis and its sphere is 1 on average; ggx-furnace's spheres keep all the light (Turquin's
  compensation), gloss-furnace's keep a little less (src/gpu_scenes_materials.ixx:78-116);
- three estimators of one integral: -p material, -p nee and -p mis draw their paths differently and must give the same
  image up to noise (ImageCompare), on cornell-ggx (GGX under a quad lamp), veach (GGX under sphere lamps), sun-test
  (the environment map's light samples) and the triangle scene; a wrong density, weight or lamp shape breaks one;
- the books' CPU renderer: the glass scene under -p mis against the CPU's -p mixture render (ImageCompare);
- brute force: the closest triangle of seeded rays against Moller-Trumbore over every triangle in double.
"""
import numpy as np
import pytest

import oracle
from oracle import image_compare

pytestmark = pytest.mark.gpu


def _render(rt, name, width, spp, sampling=None, seed=0, depth=0, integrator="auto"):
    scene = rt.build_scene(name, width=width, samples=spp, sampling=sampling, max_depth=depth)
    return rt.render(scene, seed=seed, integrator=integrator)


def _consistent(a, b, label):
    v = image_compare.compare(a, b)
    print(f"{label}: {a.mean():.6f} vs {b.mean():.6f}: {v.summary()}", flush=True)
    return v


def _agree_with_rerun(rt, name, width, spp, s1, s2, depth=0):
    verdicts = []
    for seeds in ((11, 12), (13, 14)):
        a = _render(rt, name, width, spp, s1, seeds[0], depth).pixels
        b = _render(rt, name, width, spp, s2, seeds[1], depth).pixels
        v = _consistent(a, b, f"{name} {s1} vs {s2} {width}w {spp}spp")
        verdicts.append(v)
        if v.consistent:
            return
    raise AssertionError(f"{name}: {s1} and {s2} differ beyond noise twice: " + " | ".join(v.summary() for v in verdicts))


def test_env_furnace_exactly_one(rt):
    r = _render(rt, "env-furnace", 240, 16, "material", seed=3)
    img = r.pixels
    assert np.all(img == 1.0), f"{np.count_nonzero(img != 1.0)} values differ from 1, e.g. {img[img != 1.0][:5]}"
    assert r.shadow_rays == 0


def test_env_furnace_mis(rt):
    scene = rt.build_scene("env-furnace", width=240, samples=64, sampling="mis")
    with rt.Session(scene, seed=5) as s:
        s.render_to(scene.passes)
        img = s.pixels()
        rays, shadow, _ = s.counters()
    # camera rays that miss see the map alone: exactly 1
    background = np.all(img == 1.0, axis=2)
    assert background.mean() > 0.5
    sphere = img[~background]
    z = (sphere.mean() - 1.0) / (sphere.std() / np.sqrt(sphere.size))
    print(f"sphere mean {sphere.mean():.6f}, z {z:.2f}, rays {rays}, shadow rays {shadow}")
    assert abs(z) < 4.5 and abs(sphere.mean() - 1.0) < 2e-3
    assert shadow > 0


def test_mesh_furnace_exactly_one(rt):
    for sampling in ("material", "mis"):
        img = _render(rt, "mesh-furnace", 240, 16, sampling, seed=9).pixels
        assert np.all(img == 1.0), f"{sampling}: {np.count_nonzero(img != 1.0)} values differ from 1"


def _centre(img, x_world, half_width=5.0):
    """The 5x5 pixels around the image of a sphere centred at (x_world, 0, 0), seen by the furnaces' camera"""
    h, w, _ = img.shape
    # vfov 6 from 60 away: the frame's half height is 60 tan(3 deg)
    half_h = 60 * np.tan(np.radians(3.0))
    px = w / 2 + x_world / half_h * (h / 2)
    i, j = int(round(px - 0.5)), h // 2
    return img[j - 2:j + 3, i - 2:i + 3]


def test_ggx_furnace_keeps_the_light(rt):
    img = _render(rt, "ggx-furnace", 480, 1024, seed=21).pixels
    for k, x in enumerate([-4.0, -2.0, 0.0, 2.0, 4.0]):
        patch = _centre(img, x)
        print(f"sphere {k}: {patch.mean():.5f} +- {patch.std() / 5:.5f}")
        assert abs(patch.mean() - 1.0) < 0.01, f"sphere {k} keeps {patch.mean()}"


def test_gloss_furnace(rt):
    img = _render(rt, "gloss-furnace", 480, 1024, seed=22).pixels
    for k, x in enumerate([-3.0, -1.0, 1.0, 3.0]):
        patch = _centre(img, x)
        print(f"sphere {k}: {patch.mean():.5f}")
        assert 0.9 < patch.mean() <= 1.0 + 0.01, f"sphere {k} keeps {patch.mean()}"


@pytest.mark.parametrize("name,width,spp", [("cornell-ggx", 96, 256), ("veach", 128, 256), ("sun-test", 128, 256)])
def test_nee_mis_agree(rt, name, width, spp):
    _agree_with_rerun(rt, name, width, spp, "nee", "mis")


@pytest.mark.parametrize("name,width,spp", [("cornell-ggx", 96, 512), ("cornell", 96, 256), ("glass", 96, 256)])
def test_material_mis_agree(rt, name, width, spp):
    _agree_with_rerun(rt, name, width, spp, "material", "mis")


def test_glass_mis_vs_cpu(rt):
    """The precision config's scene under -p mis against the CPU renderer's -p mixture image"""
    verdicts = []
    for cpu_seed, gpu_seed in ((0, 1), (2, 3)):
        gpu = _render(rt, "glass", 120, 256, "mis", seed=gpu_seed).pixels
        cpu, _ = oracle.render("glass", width=120, spp=64, seed=cpu_seed)
        v = _consistent(cpu, gpu, "glass cpu mixture vs gpu mis")
        verdicts.append(v)
        if v.consistent:
            return
    raise AssertionError("glass -p mis differs from the CPU beyond noise twice: " +
                         " | ".join(v.summary() for v in verdicts))


def test_triangles_small(rt):
    """The triangle scene, small: finite, lit, and -p nee agrees with -p mis"""
    r = _render(rt, "triangles", 160, 16, seed=1)
    img = r.pixels
    assert np.all(np.isfinite(img)) and img.mean() > 0.05
    assert r.shadow_rays > 0
    _agree_with_rerun(rt, "triangles", 96, 64, "nee", "mis")


@pytest.mark.parametrize("name,sampling", [("cornell-ggx", "mis"), ("cornell-ggx", "nee"), ("sun-test", "mis"),
                                           ("env-furnace", "mis"), ("veach", "mis"), ("glass", "mis"),
                                           ("final", "mis"), ("triangles", "mis")])
def test_wavefront_equals_megakernel_quality(rt, name, sampling):
    """The quality integrator in both forms: the same sums bit for bit, the same ray and shadow-ray counts"""
    scene = rt.build_scene(name, width=64, samples=16, sampling=sampling, earth_width=512)
    with rt.Session(scene, seed=6, integrator="megakernel") as a:
        assert a.kernel().name.decode().startswith("quality")
        a.render_to(16)
        sa = a.sums()
        ca = a.counters()
    with rt.Session(scene, seed=6, integrator="wavefront") as b:
        assert b.kernel().name.decode().startswith("wavefront")
        b.render_to(7)
        b.render_to(16)
        sb = b.sums()
        cb = b.counters()
    assert ca[0] == cb[0] and ca[1] == cb[1], f"rays {ca[:2]} vs {cb[:2]}"
    mism = np.count_nonzero(sa.view(np.uint64) != sb.view(np.uint64))
    assert mism == 0, f"{mism} of {sa.size} sums differ; max |diff| {np.abs(sa - sb).max()}"


def test_quality_same_seed_same_bytes(rt):
    a = _render(rt, "cornell-ggx", 64, 16, "mis", seed=4).pixels
    b = _render(rt, "cornell-ggx", 64, 16, "mis", seed=4).pixels
    assert a.tobytes() == b.tobytes()


def _mesh_scene_triangles(scene):
    d = scene.desc
    tris = []
    instances = [(d.instances[k].mesh, np.array(d.instances[k].object_to_world).reshape(3, 4))
                 for k in range(d.instance_count)] or [(m, None) for m in range(d.mesh_count)]
    for mesh, m in instances:
        md = d.meshes[mesh]
        pos = np.ctypeslib.as_array(md.positions, shape=(md.vertex_count * 3,)).reshape(-1, 3).astype(np.float64)
        idx = np.ctypeslib.as_array(md.indices, shape=(md.triangle_count * 3,)).reshape(-1, 3)
        if m is not None:
            pos = pos @ m[:, :3].T + m[:, 3]
        tris.append(pos[idx])
    return np.concatenate(tris)


def _brute_force(tris, o, d):
    """The nearest t in [0.001, inf) over every triangle, Moller-Trumbore in double"""
    best = np.full(len(o), np.inf)
    v0, e1, e2 = tris[:, 0], tris[:, 1] - tris[:, 0], tris[:, 2] - tris[:, 0]
    for k in range(0, len(tris), 512):
        a, b, c = v0[k:k + 512], e1[k:k + 512], e2[k:k + 512]
        pv = np.cross(d[:, None, :], c[None])
        det = np.einsum("tk,rtk->rt", b, pv)
        with np.errstate(divide="ignore", invalid="ignore"):
            inv = 1 / det
            s = o[:, None, :] - a[None]
            u = np.einsum("rtk,rtk->rt", s, pv) * inv
            qv = np.cross(s, b[None])
            v = np.einsum("rk,rtk->rt", d, qv) * inv
            t = np.einsum("tk,rtk->rt", c, qv) * inv
        ok = (np.abs(det) > 1e-30) & (u >= 0) & (u <= 1) & (v >= 0) & (u + v <= 1) & (t >= 0.001)
        best = np.minimum(best, np.where(ok, t, np.inf).min(axis=1))
    return best


def test_triangle_closest_hit_brute_force(rt):
    scene = rt.build_scene("mesh-furnace", width=64, samples=1)
    tris = _mesh_scene_triangles(scene)
    assert len(tris) == 2 * 5120
    rng = np.random.default_rng(77)
    n = 4000
    o = np.stack([rng.uniform(-3, 3, n), rng.uniform(-2, 2, n), rng.uniform(3, 6, n)], 1)
    target = np.stack([rng.uniform(-2.4, 2.4, n), rng.uniform(-1.2, 1.2, n), rng.uniform(-1, 1, n)], 1)
    d = target - o
    with rt.Session(scene, integrator="megakernel") as s:
        ids, t, ent = s.closest_hit(o.astype(np.float32), d.astype(np.float32), np.zeros(n, np.float32),
                                    np.full(n, -1, np.int32), np.zeros(n, np.uint8))
    bf = _brute_force(tris, o.astype(np.float32).astype(np.float64), d.astype(np.float32).astype(np.float64))
    hit_gpu = ids >= 0
    hit_bf = np.isfinite(bf)
    assert hit_bf.mean() > 0.3
    disagree = np.count_nonzero(hit_gpu != hit_bf)
    assert disagree <= 2, f"{disagree} rays hit in one and miss in the other"
    both = hit_gpu & hit_bf
    rel = np.abs(t[both] - bf[both]) / bf[both]
    assert np.quantile(rel, 0.999) < 1e-5 and rel.max() < 1e-3, f"t differs: max {rel.max()}"
    assert np.all(ent[both] == 1), "rays from outside meet the spheres' fronts"


@pytest.mark.parametrize("waves", [0, 5, 6, 7])
def test_split_stack_equals_wavefront(rt, waves):
    """The triangle scene's tree is deep (36 stack levels): the quality megakernel keeps its stacks' deeper levels in
    HBM (more of them at a tighter register budget, since more waves share the LDS); the sums are the wavefront's,
    bit for bit"""
    scene = rt.build_scene("triangles", width=48, samples=9, strata=3, sampling="mis")
    with rt.Session(scene, seed=2, integrator="wavefront") as w:
        w.render_to(9)
        reference = w.sums()
        counts = w.counters()[:2]
    with rt.Session(scene, seed=2, integrator="megakernel", megakernel_waves=waves) as m:
        m.render_to(4)
        m.render_to(9)
        got = m.sums()
        assert m.counters()[:2] == counts
    mism = np.count_nonzero(got.view(np.uint64) != reference.view(np.uint64))
    assert mism == 0, f"{mism} of {got.size} sums differ"


def test_instanced_meshes_find_the_baked_hits(rt):
    """Config 3's two blobs, placed 32 times each, walked through their instances (each mesh's tree once, in its own
    coordinates: RT_TREE_TWO_LEVEL) against every instance baked into world coordinates (the default tree): the same
    rays hit and miss alike, at the same distances to float rounding, on the same sides; and renders of both agree by
    ImageCompare"""
    scene = rt.build_scene("triangles", width=96, samples=16, sampling="mis")
    cam = scene.desc.camera
    rng = np.random.default_rng(123)
    n = 30000
    center = np.array([cam.center.x, cam.center.y, cam.center.z])
    p00 = np.array([cam.pixel00_loc.x, cam.pixel00_loc.y, cam.pixel00_loc.z])
    du = np.array([cam.pixel_delta_u.x, cam.pixel_delta_u.y, cam.pixel_delta_u.z])
    dv = np.array([cam.pixel_delta_v.x, cam.pixel_delta_v.y, cam.pixel_delta_v.z])
    targets = p00 + rng.uniform(0, cam.image_width, n)[:, None] * du + rng.uniform(0, cam.image_height, n)[:, None] * dv
    o = np.repeat(center[None], n, axis=0).astype(np.float32)
    d = (targets - center).astype(np.float32)
    t0 = np.zeros(n, np.float32)
    none = np.full(n, -1, np.int32)
    out = np.zeros(n, np.uint8)
    hits = {}
    for tree in ("two_level", "default"):
        with rt.Session(scene, integrator="megakernel", tree=tree) as s:
            hits[tree] = s.closest_hit(o, d, t0, none, out)
    (ia, ta, ea), (ib, tb, eb) = hits["two_level"], hits["default"]
    both = (ia >= 0) & (ib >= 0)
    assert both.mean() > 0.5
    assert np.count_nonzero((ia >= 0) != (ib >= 0)) <= 3
    rel = np.abs(ta[both] - tb[both]) / tb[both]
    assert np.quantile(rel, 0.999) < 1e-5 and rel.max() < 1e-3, f"t differs: max {rel.max()}"
    assert np.count_nonzero(ea[both] != eb[both]) <= 3
    print(f"{both.sum()} rays hit both; {np.count_nonzero(ia[both] != ib[both])} by other ids (instanced triangles)")
    verdicts = []
    for seeds in ((41, 42), (43, 44)):
        imgs = []
        for tree, seed in zip(("two_level", "default"), seeds):
            with rt.Session(scene, seed=seed, integrator="megakernel", tree=tree) as s:
                s.render_to(16)
                imgs.append(s.pixels())
        v = _consistent(imgs[0], imgs[1], "triangles two-level vs baked")
        verdicts.append(v)
        if v.consistent:
            return
    raise AssertionError("two-level and baked trees differ beyond noise twice: " + " | ".join(v.summary() for v in verdicts))


def test_fp64_quality_furnaces_exactly_one(rt):
    """The quality integrator in float64 (RT_FLOAT64_ALL, the RMSE protocol's oracle): the furnaces' closed forms hold
    exactly as in float32"""
    for name, sampling in (("env-furnace", "material"), ("mesh-furnace", "material"), ("mesh-furnace", "mis")):
        scene = rt.build_scene(name, width=160, samples=16, sampling=sampling)
        img = rt.render(scene, seed=4, precision="float64_all").pixels
        assert np.all(img == 1.0), f"{name} {sampling}: {np.count_nonzero(img != 1.0)} values differ from 1"


@pytest.mark.parametrize("name,sampling,width,spp", [("triangles", "mis", 96, 64), ("cornell-ggx", "mis", 96, 256),
                                                     ("sun-test", "nee", 96, 64), ("glass", "mis", 96, 256)])
def test_fp64_quality_agrees_with_fp32(rt, name, sampling, width, spp):
    """The float64 quality integrator draws other numbers (53-bit draws) in double arithmetic: its render of the same
    scene must not differ from the float32 one by more than noise (ImageCompare, with the re-run rule)"""
    scene = rt.build_scene(name, width=width, samples=spp, sampling=sampling)
    verdicts = []
    for seeds in ((21, 22), (23, 24)):
        a = rt.render(scene, seed=seeds[0], precision="float32", integrator="megakernel").pixels
        b = rt.render(scene, seed=seeds[1], precision="float64_all", integrator="megakernel").pixels
        assert np.all(np.isfinite(b))
        v = _consistent(a, b, f"{name} {sampling} fp32 vs fp64")
        verdicts.append(v)
        if v.consistent:
            return
    raise AssertionError(f"{name}: float32 and float64 differ beyond noise twice: " +
                         " | ".join(v.summary() for v in verdicts))


def test_fp64_refused_where_the_reference_refuses(rt):
    """RT_FLOAT64 keeps the reference's refusal of the quality integrator (tests/routing_tests.cu:53-67); only the
    explicit RT_FLOAT64_ALL renders it in float64; the wavefront is float32 only"""
    ggx = rt.build_scene("cornell-ggx", width=32, samples=4)
    with pytest.raises(ValueError, match="float32 only"):
        rt.Session(ggx, precision="float64")
    with rt.Session(ggx, precision="float64_all") as s:
        s.render_to(2)
        assert "fp64" in s.kernel().name.decode() or s.kernel().name
    with pytest.raises(ValueError, match="float32 only"):
        rt.Session(ggx, precision="float64_all", integrator="wavefront")
