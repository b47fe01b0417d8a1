"""GPU tests of the serving loop (config 4): requests of mixed scenes and resolutions admitted by a memory budget and
rendered together, with their kernels overlapping on the device. Each request is a session of its own (its own
stream, image and seed), and a pixel's samples are summed in pass order, so a result is the same bytes whether the
request ran alone or in the mix: the stated tolerance is zero.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run(server, ids):
    while server.step() > 0:
        pass
    while any(server.status(i).state < 2 for i in ids):
        server.step()


def _view(rt, name, width, height):
    return rt.build_scene(name, width=width, height=height, samples=1)


def test_same_alone_and_in_mix(rt):
    a = _view(rt, "final", 96, 96)
    mix = [_view(rt, "weekend", 160, 90), _view(rt, "glass", 64, 64), _view(rt, "cornell-ggx", 80, 80),
           _view(rt, "sun-test", 128, 72)]
    alone = rt.Server()
    sid = alone.add_scene(a)
    ra = alone.submit_view(sid, a.desc.camera, 5, 24)
    _run(alone, [ra])
    img_alone = alone.pixels(ra)
    st = alone.status(ra)
    assert st.state == 2 and st.passes_done == 24 and st.milestone == 24
    alone.close()

    server = rt.Server(max_active=4, passes_per_step=3)
    others = [server.submit(s, 11 + k, 16 + 8 * k) for k, s in enumerate(mix)]
    sid = server.add_scene(a)
    ra = server.submit_view(sid, a.desc.camera, 5, 24)
    _run(server, others + [ra])
    img_mix = server.pixels(ra)
    for rid in others:
        st = server.status(rid)
        assert st.state == 2, f"request {rid} state {st.state}"
    assert np.array_equal(img_alone.view(np.uint64), img_mix.view(np.uint64))
    # a request alone equals a plain session of the same view and seed
    with rt.Session(mix[0], seed=11) as s:
        s.render_to(16)
        ref = s.pixels()
    assert np.array_equal(ref.view(np.uint64), server.pixels(others[0]).view(np.uint64))
    server.close()


def test_views_of_one_resident_scene(rt):
    """Two resolutions of one resident scene: each equals its own session"""
    server = rt.Server()
    small, large = _view(rt, "cornell", 48, 48), _view(rt, "cornell", 96, 64)
    sid = server.add_scene(small)
    r1 = server.submit_view(sid, small.desc.camera, 1, 8)
    r2 = server.submit_view(sid, large.desc.camera, 2, 8)
    server.remove_scene(sid)  # stays resident until its requests are released
    _run(server, [r1, r2])
    for rid, view, seed in ((r1, small, 1), (r2, large, 2)):
        with rt.Session(view, seed=seed) as s:
            s.render_to(8)
            assert np.array_equal(s.pixels().view(np.uint64), server.pixels(rid).view(np.uint64))
    server.release(r1)
    server.release(r2)
    with pytest.raises(ValueError):
        server.submit_view(sid, small.desc.camera, 1, 8)
    server.close()


def test_admission_by_memory_budget(rt):
    """A request waits while the budget is full and starts once an earlier one is released; the budget is set from
    the server's own accounting so that one view fits beside the scene and two don't"""
    view = _view(rt, "weekend", 200, 112)
    probe = rt.Server(max_active=8)
    sid = probe.add_scene(view)
    scene_bytes = probe.memory().reserved_bytes
    probe.submit_view(sid, view.desc.camera, 1, 4)
    probe.step()
    request_bytes = probe.memory().reserved_bytes - scene_bytes
    probe.close()
    assert request_bytes > 0
    server = rt.Server(memory_budget_bytes=scene_bytes + request_bytes + request_bytes // 2, max_active=8)
    sid = server.add_scene(view)
    first = server.submit_view(sid, view.desc.camera, 1, 4)
    second = server.submit_view(sid, view.desc.camera, 2, 4)
    server.step()
    assert server.status(first).state >= 1
    assert server.status(second).state == 0, "the second view doesn't fit beside the first"
    _run(server, [first])
    server.release(first)
    _run(server, [second])
    assert server.status(second).state == 2
    server.close()


def test_memory_held_within_budget(rt):
    """What the active requests' sessions actually hold (image, staging, stacks, counters) never exceeds what the
    server reserved for them, nor the budget, at any step of a mix that the budget admits only part of at once"""
    views = [_view(rt, "final", 256, 256), _view(rt, "weekend", 512, 288), _view(rt, "triangles", 192, 108),
             _view(rt, "glass", 128, 128)]
    probe = rt.Server()
    scene_ids = [probe.add_scene(v) for v in views]
    scenes = probe.memory().reserved_bytes
    probe.close()
    # (a view of the triangle scene holds ~3 MB a level of far stack for its deep tree, over a hundred MB: one such
    # view fits at a time beside a few others)
    server = rt.Server(memory_budget_bytes=scenes + (256 << 20), max_active=8, passes_per_step=8)
    scene_ids = [server.add_scene(v) for v in views]
    ids = [server.submit_view(scene_ids[k % 4], views[k % 4].desc.camera, k, 8 + 4 * k) for k in range(10)]
    peak_active = 0
    while any(server.status(i).state < 2 for i in ids):
        active = server.step()
        m = server.memory()
        assert m.held_bytes <= m.reserved_bytes <= m.budget_bytes, (m.held_bytes, m.reserved_bytes, m.budget_bytes)
        peak_active = max(peak_active, active)
        for i in ids:
            if server.status(i).state == 2:
                server.release(i)
                ids.remove(i)
    assert peak_active >= 1
    server.close()


def test_milestones_are_powers_of_two(rt):
    view = _view(rt, "glass", 48, 48)
    server = rt.Server(passes_per_step=100)
    rid = server.submit(view, 3, 40)
    seen = []
    while server.status(rid).state < 2:
        server.step()
        seen.append(server.status(rid).milestone)
    assert seen == [1, 2, 4, 8, 16, 32, 40]
    server.close()


def test_steps_raised_by_paths(rt):
    """A small view's steps are raised to paths_per_step paths (ceil(paths / pixels) passes), still stopping at every
    power-of-two milestone"""
    view = _view(rt, "glass", 64, 64)
    server = rt.Server(passes_per_step=4, paths_per_step=64 * 64 * 16 - 100)  # ceil: 16 passes a step
    rid = server.submit(view, 3, 100)
    seen = []
    while server.status(rid).state < 2:
        server.step()
        seen.append(server.status(rid).passes_done)
    assert seen == [1, 2, 4, 8, 16, 32, 48, 64, 80, 96, 100]
    server.close()
    with rt.Session(view, seed=3) as s:
        s.render_to(100)
        alone = s.pixels()
    server = rt.Server(paths_per_step=64 * 64 * 16)
    rid = server.submit(view, 3, 100)
    _run(server, [rid])
    assert np.array_equal(alone.view(np.uint64), server.pixels(rid).view(np.uint64))
    server.close()


@pytest.mark.parametrize("paths_per_step", [0, 1 << 22])
def test_benchmark_mix_same_alone_and_in_mix(rt, paths_per_step):
    """Requests drawn from the serving benchmark's own mix (its four scenes resident, each scene twice, each request
    one of the four resolutions 256x256 to 1920x1080, a log-uniform spp in [16, 1024] and its own seed; seeded, with
    the paths of one request capped at 2^28 so the test stays short): every result in the mix equals its alone-session
    bytes, with steps of a fixed pass count and with steps raised to 4 M paths for small views"""
    scenes = ["weekend", "glass", "final", "triangles"]
    sizes = [(256, 256), (512, 512), (1080, 1080), (1920, 1080)]
    rng = np.random.default_rng(44)
    picks = []
    for k in range(8):
        name, (w, h) = scenes[k % 4], sizes[rng.integers(4)]  # every scene twice
        spp = int(round(np.exp(rng.uniform(np.log(16), np.log(1024)))))
        spp = max(1, min(spp, (1 << 28) // (w * h)))
        picks.append((name, w, h, spp, 500 + k))
    print(picks)
    server = rt.Server(max_active=4, passes_per_step=8, paths_per_step=paths_per_step)
    resident = {name: server.add_scene(rt.build_scene(name, width=256, height=256, samples=1)) for name in scenes}
    views = {(name, w, h): rt.build_scene(name, width=w, height=h, samples=1) for name, w, h, _, _ in picks}
    ids = [server.submit_view(resident[name], views[(name, w, h)].desc.camera, seed, spp)
           for name, w, h, spp, seed in picks]
    _run(server, ids)
    for rid, (name, w, h, spp, seed) in zip(ids, picks):
        st = server.status(rid)
        assert st.state == 2 and st.passes_done == spp, (name, w, h, spp, st.state, st.passes_done)
        with rt.Session(views[(name, w, h)], seed=seed) as s:
            s.render_to(spp)
            alone = s.pixels()
        mixed = server.pixels(rid)
        assert np.array_equal(alone.view(np.uint64), mixed.view(np.uint64)), (name, w, h, spp)
    server.close()
