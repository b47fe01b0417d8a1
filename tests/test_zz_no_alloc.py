"""rt_session_render_to_static, the entry point that allocates no device memory: its sums equal rt_session_render_to's
bit for bit, on two calls in a row (the second on a smaller image, over the scratch the first call left), with an
image too large for its scratch refused cleanly before any launch."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _sums(rt, scene, passes, static, **options):
    with rt.Session(scene, seed=3, **options) as s:
        r = s.render_to_static(passes) if static else s.render_to(passes)
        assert s.state() == (passes, 0)
        return s.sums(), s.counters()[0], r


@pytest.mark.parametrize("options", [{"integrator": "megakernel"}, {"integrator": "auto"},
                                     {"integrator": "megakernel", "precision": "float64"}])
def test_static_equals_render_to(rt, options):
    large = rt.build_scene("final", width=256, samples=16, earth_width=512)
    small = rt.build_scene("cornell", width=96, samples=9)
    for scene, passes in ((large, 16), (small, 9)):
        want, rays, _ = _sums(rt, scene, passes, False, **options)
        got, got_rays, report = _sums(rt, scene, passes, True, **options)
        assert report.passes == passes and report.launches >= 1
        assert got_rays == rays
        assert np.array_equal(want.view(np.uint64), got.view(np.uint64)), f"{scene.name}: sums differ"


def test_static_continues_a_render(rt):
    scene = rt.build_scene("weekend", width=160, samples=16)
    with rt.Session(scene, seed=5, integrator="megakernel") as a, rt.Session(scene, seed=5, integrator="megakernel") as b:
        a.render_to(16)
        b.render_to(5)
        b.render_to_static(11)
        b.render_to(16)
        assert np.array_equal(a.sums().view(np.uint64), b.sums().view(np.uint64))


def test_too_large_for_the_scratch(rt):
    scene = rt.build_scene("weekend", width=2400, height=1800, samples=1)
    with rt.Session(scene, integrator="megakernel") as s:
        with pytest.raises(ValueError, match="static scratch"):
            s.render_to_static(1)
        assert s.state() == (0, 0)
