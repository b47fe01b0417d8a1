"""The quality path's light samples on the GPU (rt_session_light_samples), each against the geometry it states, in
float64 (the reference checks the same draws in tests/lamp_tests.cu T2-T4 and tests/env_tests.cu env_sampling):

- a quad lamp (cornell): the drawn point lies on the lamp, uniform over its area (chi-square over 8 x 8 cells of its
  parameters), with the density d^2 / (cos_l A) per solid angle; points behind the lamp draw nothing;
- sphere lamps (veach's four): the light chosen uniformly among the lamps, the drawn point the near crossing of the
  sphere, 1 - cos theta uniform in [0, 1 - cos theta_max] about the centre's direction (Kolmogorov-Smirnov), phi uniform,
  the density 1 / (2 pi (1 - cos theta_max)) over the lamps' count;
- the environment map (sun-test, triangles' sky with its sun): the drawn direction's radiance and density are its
  texel's (equirectangular, +y up, turned by the yaw), and the texels are drawn in proportion to density x solid angle
  (chi-square over blocks of texels), so that radiance / density is constant where the map is sampled;
- everywhere: the density a scattered ray finding the same light is given (what MIS weighs a BSDF sample by) is the
  draw's own density.
"""
import math

import numpy as np
import pytest
from scipy import stats

pytestmark = pytest.mark.gpu

LIGHT_QUAD, LIGHT_SPHERE = 0, 1  # rt_light_kind


def _lamps(rt, scene):
    L = rt.scene_lamps(scene).astype(np.float64)
    ints = rt.scene_lamps(scene).view(np.int32)
    return [dict(corner=L[k, 0:3], size=L[k, 3], u=L[k, 4:7], kind=int(ints[k, 7]), v=L[k, 8:11],
                 primitive=int(ints[k, 11]), normal=L[k, 12:15], radiance=L[k, 16:19]) for k in range(len(L))]


def _close(got, want, rtol, what):
    err = np.abs(got - want) / np.maximum(np.abs(want), 1e-30)
    assert np.quantile(err, 0.999) < rtol and err.max() < 20 * rtol, f"{what}: relative error max {err.max():.3g}"


def test_quad_lamp_samples(rt):
    scene = rt.build_scene("cornell", width=32, samples=1, sampling="mis")
    (lamp,) = _lamps(rt, scene)
    assert lamp["kind"] == LIGHT_QUAD
    rng = np.random.default_rng(3)
    n = 200000
    x = rng.uniform([10, 10, 10], [545, 540, 545], (n, 3)).astype(np.float32)
    with rt.Session(scene) as s:
        out = s.light_samples(x, seed=4)
    ok = out["skip"] >= 0
    x64 = x.astype(np.float64)
    # points above the lamp's plane (the lamp faces down) draw nothing
    above = np.einsum("ij,j->i", x64 - lamp["corner"], lamp["normal"]) <= 0
    assert np.array_equal(ok, ~above)
    assert np.all(out["skip"][ok] == lamp["primitive"])
    wi = out["wi"][ok].astype(np.float64)
    d = out["distance"][ok].astype(np.float64)
    p = x64[ok] + d[:, None] * wi
    # the point's parameters on the lamp: p = corner + a u + b v
    uv = np.stack([lamp["u"], lamp["v"]], 1)
    ab, *_ = np.linalg.lstsq(uv, (p - lamp["corner"]).T, rcond=None)
    a, b = ab
    assert np.all((a > -1e-4) & (a < 1 + 1e-4) & (b > -1e-4) & (b < 1 + 1e-4))
    off_plane = np.abs(np.einsum("ij,j->i", p - lamp["corner"], lamp["normal"]))
    assert off_plane.max() < 1e-3
    cos_l = -wi @ lamp["normal"]
    _close(out["density"][ok].astype(np.float64), d * d / (cos_l * lamp["size"]), 2e-4, "quad lamp density")
    _close(out["hit_density"][ok].astype(np.float64), out["density"][ok].astype(np.float64), 1e-5, "hit density")
    assert np.all(out["radiance"][ok] == lamp["radiance"].astype(np.float32))
    cells = (np.minimum((np.clip(a, 0, 1) * 8).astype(int), 7) * 8 + np.minimum((np.clip(b, 0, 1) * 8).astype(int), 7))
    counts = np.bincount(cells, minlength=64)
    chi2 = stats.chisquare(counts)
    assert chi2.pvalue > 1e-4, f"points not uniform on the lamp: p = {chi2.pvalue:.2e}"


def test_sphere_lamp_samples(rt):
    scene = rt.build_scene("veach", width=32, samples=1, sampling="mis")
    lamps = _lamps(rt, scene)
    assert len(lamps) == 4 and all(l["kind"] == LIGHT_SPHERE for l in lamps)
    n = 400000
    x1 = np.array([0.3, 2.0, 6.0])
    x = np.tile(x1, (n, 1)).astype(np.float32)
    with rt.Session(scene) as s:
        out = s.light_samples(x, seed=9)
    assert np.all(out["skip"] >= 0)
    by_id = {l["primitive"]: l for l in lamps}
    counts = np.array([np.count_nonzero(out["skip"] == l["primitive"]) for l in lamps])
    assert stats.chisquare(counts).pvalue > 1e-4, f"lamps not chosen uniformly: {counts}"
    xf = x1.astype(np.float32).astype(np.float64)
    for l in lamps:
        sel = out["skip"] == l["primitive"]
        wi = out["wi"][sel].astype(np.float64)
        d = out["distance"][sel].astype(np.float64)
        c, r = l["corner"], l["size"]
        p = xf + d[:, None] * wi
        dist_c = np.linalg.norm(p - c, axis=1)
        assert np.all(np.abs(dist_c - r) < 1e-3 * r + 1e-5), f"radius {r}: points off the sphere"
        assert np.all(np.einsum("ij,ij->i", p - c, wi) <= 1e-4 * r), "the near crossing"
        to_c = c - xf
        dc = np.linalg.norm(to_c)
        one_minus_cos_max = 1 - math.sqrt(1 - (r / dc) ** 2)
        _close(out["density"][sel].astype(np.float64), np.full(sel.sum(), 0.25 / (2 * math.pi * one_minus_cos_max)),
               2e-4, f"sphere lamp {r} density")
        _close(out["hit_density"][sel].astype(np.float64), out["density"][sel].astype(np.float64), 1e-5,
               "hit density")
        # 1 - cos theta about the centre's direction, over its largest value: uniform (the direction renormalised in
        # float64 first, so that its float rounding moves cos theta at second order only)
        cos_t = (wi / np.linalg.norm(wi, axis=1, keepdims=True)) @ (to_c / dc)
        u = (1 - cos_t) / one_minus_cos_max
        assert u.min() > -1e-3 and u.max() < 1 + 1e-2
        ks = stats.kstest(np.clip(u, 0, 1), "uniform")
        assert ks.pvalue > 1e-4, f"radius {r}: 1 - cos theta not uniform (p = {ks.pvalue:.2e})"
        assert np.all(out["radiance"][sel] == by_id[l["primitive"]]["radiance"].astype(np.float32))


def _environment(rt, scene):
    env = scene.desc.environment.contents
    w, h = env.image.width, env.image.height
    rgb = np.ctypeslib.as_array(env.image.rgb, shape=(h * w * 3,)).reshape(h, w, 3).copy()
    return rgb, env.yaw_degrees, env.scale


def _texel_of(wi, w, h, yaw_degrees):
    c, s = math.cos(math.radians(yaw_degrees)), math.sin(math.radians(yaw_degrees))
    mx = c * wi[:, 0] + s * wi[:, 2]
    mz = c * wi[:, 2] - s * wi[:, 0]
    u = 0.5 + np.arctan2(mz, mx) / (2 * math.pi)
    t = np.arctan2(np.hypot(mx, mz), wi[:, 1]) / math.pi
    i = np.clip((u * w).astype(int), 0, w - 1)
    j = np.clip((t * h).astype(int), 0, h - 1)
    return j, i


@pytest.mark.parametrize("name", ["sun-test", "triangles"])
def test_map_samples(rt, name):
    scene = rt.build_scene(name, width=32, samples=1, sampling="mis")
    assert len(_lamps(rt, scene)) == 0
    rgb, yaw, scale = _environment(rt, scene)
    tables = rt.environment_tables(rgb, yaw, scale)
    h, w = rgb.shape[:2]
    n = 1000000
    x = np.zeros((n, 3), np.float32)
    with rt.Session(scene) as s:
        out = s.light_samples(x, seed=21)
    ok = out["skip"] == -1
    assert ok.mean() > 0.999
    assert np.all(np.isinf(out["distance"][ok]))
    wi = out["wi"][ok].astype(np.float64)
    assert np.allclose(np.linalg.norm(wi, axis=1), 1, atol=2e-6)
    j, i = _texel_of(wi, w, h, yaw)
    texel = tables.texels[j, i]
    # the drawn direction's texel (but for directions within rounding of a texel border)
    same = out["density"][ok] == texel[:, 3]
    assert same.mean() > 0.999, f"{(~same).sum()} draws read another texel's density"
    assert np.array_equal(out["hit_density"][ok], out["density"][ok])
    assert np.allclose(out["radiance"][ok][same], scale * texel[same, :3], rtol=1e-6, atol=0)
    # texels drawn in proportion to density x solid angle, over blocks of 16 x 16 texels
    omega = (tables.row_cos[:-1] - tables.row_cos[1:])[:, None] * (2 * math.pi / w)
    prob = (tables.texels[:, :, 3] * omega).astype(np.float64)
    assert abs(prob.sum() - 1) < 1e-4
    bj, bi = h // 16, w // 16
    expected = prob.reshape(bj, 16, bi, 16).sum(axis=(1, 3)).reshape(-1) * ok.sum()
    observed = np.bincount((j // 16) * bi + (i // 16), minlength=bj * bi)
    keep = expected >= 5
    chi2 = np.sum((observed[keep] - expected[keep]) ** 2 / expected[keep])
    p = stats.chi2.sf(chi2, keep.sum() - 1)
    assert observed[~keep].sum() < 5 * max(expected[~keep].sum(), 1) + 20
    assert p > 1e-4, f"{name}: texel blocks not drawn by density x solid angle (p = {p:.2e})"


def test_light_probe_refusals(rt):
    scene = rt.build_scene("cornell", width=32, samples=1)  # -p material: no light samples
    with rt.Session(scene) as s:
        with pytest.raises(ValueError):
            s.light_samples(np.zeros((1, 3)))
