"""The media on the GPU (rt_session_media_events), against their geometry and the exponential law, in float64 (the
reference checks the same in tests/lamp_tests.cu's media suite: chord, limits, grazing, particles, visibility):

- each medium's chord along a ray: the two roots of its sphere, or the first and last crossings of its convex body's
  faces, within (0.001, the surfaces' hit); the transmittance a shadow ray is weighted by is exp(-sum rho_k L_k) over
  those chords' lengths (final's ball of smoke inside its fog, smoke's two boxes);
- the paths' events: a collision is reported only inside the chord of the medium named, and along a fixed ray the
  first collision's distance follows the superposed media's law, F(s) = 1 - exp(-integral of rho up to s) (Kolmogorov-
  Smirnov, conditioned on a collision), while the share of rays with none is the transmittance (binomial).
"""
import math

import numpy as np
import pytest
from scipy import stats

pytestmark = pytest.mark.gpu


def _media(scene):
    out = []
    d = scene.desc
    for k in range(d.medium_count):
        m = d.media[k]
        rho = -1.0 / m.neg_inv_density
        if m.quad_count == 0:
            c = np.array([m.sphere.center.x, m.sphere.center.y, m.sphere.center.z])
            out.append(dict(rho=rho, sphere=(c, m.sphere.radius)))
        else:
            quads = [(np.array([q.q.x, q.q.y, q.q.z]), np.array([q.u.x, q.u.y, q.u.z]), np.array([q.v.x, q.v.y, q.v.z]))
                     for q in (m.quads[i] for i in range(m.quad_count))]
            out.append(dict(rho=rho, quads=quads))
    return out


def _chord(medium, o, d):
    """(enter, leave) in t along o + t d, or None"""
    if "sphere" in medium:
        c, r = medium["sphere"]
        a = d @ d
        oc = o - c
        b = d @ oc
        disc = b * b - a * (oc @ oc - r * r)
        if disc <= 0:
            return None
        s = math.sqrt(disc)
        lo, hi = (-b - s) / a, (-b + s) / a
        return (lo, hi) if hi - lo > 1e-4 else None
    ts = []
    for q, u, v in medium["quads"]:
        n = np.cross(u, v)
        den = n @ d
        if abs(den) < 1e-12:
            continue
        t = n @ (q - o) / den
        p = o + t * d - q
        w = n / (n @ n)
        al, be = w @ np.cross(p, v), w @ np.cross(u, p)
        if -1e-9 <= al <= 1 + 1e-9 and -1e-9 <= be <= 1 + 1e-9:
            ts.append(t)
    if len(ts) < 2 or max(ts) - min(ts) < 1e-4:
        return None
    return min(ts), max(ts)


def _optical_depth(media, o, d, t_hi, t_lo=0.001):
    speed = math.sqrt(d @ d)
    depth = 0.0
    for m in media:
        ch = _chord(m, o, d)
        if ch is None:
            continue
        a, b = max(ch[0], t_lo, 0.0), min(ch[1], t_hi)
        if b > a:
            depth += m["rho"] * (b - a) * speed
    return depth


@pytest.mark.parametrize("name", ["final", "smoke"])
def test_chords_and_transmittance(rt, name):
    scene = rt.build_scene(name, width=32, samples=1)
    media = _media(scene)
    rng = np.random.default_rng(5)
    n = 20000
    # rays from around the media toward points inside them, and in random directions
    targets = []
    for m in media:
        if "sphere" in m:
            c, r = m["sphere"]
            r = min(r, 400)
            targets.append(c + rng.uniform(-r, r, (n // (2 * len(media)), 3)))
        else:
            pts = np.array([q for q, u, v in m["quads"]] + [q + u + v for q, u, v in m["quads"]])
            lo, hi = pts.min(0), pts.max(0)
            targets.append(rng.uniform(lo, hi, (n // (2 * len(media)), 3)))
    targets = np.concatenate(targets)
    k = len(targets)
    o = rng.uniform(-100, 650, (n, 3))
    d = np.concatenate([targets - o[:k], rng.normal(size=(n - k, 3))])
    d = (d / np.linalg.norm(d, axis=1, keepdims=True) * rng.uniform(0.5, 2, (n, 1))).astype(np.float32)
    o = o.astype(np.float32)
    tm = np.zeros(n, np.float32)
    t_surface = rng.uniform(50, 2000, n).astype(np.float32) / np.linalg.norm(d, axis=1)
    with rt.Session(scene, integrator="megakernel") as s:
        medium, t, tr = s.media_events(o, d, tm, t_surface, seed=3)
    o64, d64, ts64 = o.astype(np.float64), d.astype(np.float64), t_surface.astype(np.float64)
    want = np.array([math.exp(-_optical_depth(media, o64[i], d64[i], ts64[i])) for i in range(n)])
    crossing = want < 1
    assert crossing.mean() > 0.3
    err = np.abs(tr.astype(np.float64) - want) / want
    assert np.quantile(err, 0.999) < 1e-4 and err.max() < 1e-2, f"{name}: transmittance off by {err.max():.3g}"
    # a collision lies inside the named medium's chord, within (0.001, t_surface)
    hit = medium >= 0
    assert hit.sum() > 100
    for i in np.nonzero(hit)[0]:
        ch = _chord(media[medium[i]], o64[i], d64[i])
        assert ch is not None, f"ray {i}: an event in a medium it doesn't cross"
        tol = 1e-4 * (1 + abs(ch[1]))
        assert max(ch[0], 0.001) - tol <= t[i] <= min(ch[1], ts64[i]) + tol, f"ray {i}: t {t[i]} outside {ch}"


@pytest.mark.parametrize("through_ball", [True, False])
def test_first_collision_law(rt, through_ball):
    """final's ball of smoke (rho 0.2, radius 70) inside its fog (rho 1e-4): one ray through the ball's middle, or one
    through the fog alone, 400 K free flights each"""
    scene = rt.build_scene("final", width=32, samples=1)
    media = _media(scene)
    ball = next(k for k, m in enumerate(media) if "sphere" in m and m["sphere"][1] < 1000)
    c, r = media[ball]["sphere"]
    n = 400000
    if through_ball:
        o1 = np.array([c[0] + 3.0, c[1] - 5.0, c[2] - 300.0], np.float32)
        d1 = np.array([0.0, 0.02, 1.0], np.float32)
    else:
        o1 = np.array([c[0], c[1] + 200.0, c[2]], np.float32)
        d1 = np.array([0.6, 0.0, -0.8], np.float32)
    t_surface = np.float32(600.0)
    o = np.tile(o1, (n, 1))
    d = np.tile(d1, (n, 1))
    with rt.Session(scene, integrator="megakernel") as s:
        medium, t, tr = s.media_events(o, d, np.zeros(n, np.float32), np.full(n, t_surface), seed=17)
    o64, d64 = o1.astype(np.float64), d1.astype(np.float64)
    total = _optical_depth(media, o64, d64, float(t_surface))
    p_none = math.exp(-total)
    assert abs(tr[0] - p_none) < 1e-5 * p_none + 1e-7
    ch = _chord(media[ball], o64, d64)
    assert (ch is not None) == through_ball
    none = np.count_nonzero(medium < 0)
    assert stats.binomtest(none, n, p_none).pvalue > 1e-4, f"{none} of {n} without a collision, expected {p_none:.4f}"
    hit_t = t[medium >= 0].astype(np.float64)
    # the law conditioned on a collision: the optical depth up to s, piecewise linear, from each medium's chord
    spans = []
    for m in media:
        mc = _chord(m, o64, d64)
        if mc is not None:
            spans.append((max(mc[0], 0.001, 0.0), min(mc[1], float(t_surface)), m["rho"] * math.sqrt(d64 @ d64)))
    depth = lambda x: sum(k * np.clip(np.asarray(x) - a, 0, max(b - a, 0)) for a, b, k in spans)
    ks = stats.kstest(hit_t, lambda x: (1 - np.exp(-depth(x))) / (1 - p_none))
    assert ks.pvalue > 1e-4, f"first collisions off the law: KS p = {ks.pvalue:.2e}"
    # the collisions name the medium whose chord they are in: the ball's inside it, the fog's elsewhere
    named = medium[medium >= 0]
    if through_ball:
        in_ball = (hit_t >= ch[0] - 1e-3) & (hit_t <= ch[1] + 1e-3)
        assert np.all(named[~in_ball] != ball)
        assert np.mean(named[in_ball] == ball) > 0.99
    else:
        assert np.all(named != ball)
