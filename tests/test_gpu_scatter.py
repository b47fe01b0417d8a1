"""The quality path's scattering lobes on the GPU (rt_scatter_probe), each against the mathematics it states, in the
tangent frame (the normal +z). The reference's own checks of the same lobes are tests/ggx_tests.cu (formulas,
reciprocity, lobes, conductor, ggx_energy compensated), tests/gloss_tests.cu and tests/bsdf_tests.cu (draws,
distributions); the expected values here are computed in float64 from the published formulas, not from the kernels'
code:

- GGX (Walter et al. 2007; Heitz 2014): D(h) = a^2 / (pi (cos^2 theta_h (a^2 - 1) + 1)^2), Smith's
  Lambda(c) = (sqrt(1 + a^2 tan^2 theta) - 1) / 2, the height-correlated G2 = 1 / (1 + Lambda(v) + Lambda(l)),
  f cos_l = F D G2 / (4 cos_v) with Schlick's F at v.h, and the visible-normal density of the reflected direction
  G1(v) D / (4 cos_v); the conductor times Turquin's 1 + F0 (1 - E) / E, E the energy table (the reference's own
  table, tests/golden/ggx_energy_ref.json) read bilinearly at (sqrt(cos_v), r);
- the glossy coat: the GGX layer of F0 0.04 plus Ashikhmin and Shirley's coupled matte base
  (28 / (23 pi)) rho (1 - 0.04) (1 - (1 - cos_l / 2)^5) (1 - (1 - cos_v / 2)^5) cos_l, drawn half by the layer's
  visible normals and half by the cosine;
- Lambert: rho cos_l / pi, density cos_l / pi; the isotropic phase function: rho / (4 pi), 1 / (4 pi);
- a draw's weight is f cos / density at the drawn direction and its density the evaluated one; the drawn directions
  follow the density (a chi-square test over cells of the hemisphere, expected counts by quadrature of the float64
  density); the lobes are reciprocal; a white conductor's compensated energy is 1 within the table's interpolation,
  and its single-scattering energy (the draws' weights over the compensation) is the table's E;
- below alpha 1e-3 the conductor is a mirror: the drawn direction is v reflected, weighted by Schlick's F, density 0;
- the books' delta lobes: a fuzzed metal's direction is the unit mirror direction plus fuzz times a point uniform on
  the unit sphere, absorbed below the surface, weighted by the albedo; glass reflects with Schlick's probability
  (binomial test per angle), always beyond the critical angle, and otherwise refracts by Snell's law in the plane of
  incidence (the reference's bsdf draws and distributions).
"""
import json
import math
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = np.array(json.load(open(os.path.join(HERE, "golden", "ggx_energy_ref.json")))["values"]).reshape(32, 32)
COAT_F0 = 0.04
ALPHAS = [0.01, 0.05, 0.2, 0.5, 1.0]


def _hemisphere(rng, n, min_z=0.02):
    """Directions uniform on the upper hemisphere with z >= min_z, float32-exact"""
    d = rng.normal(size=(4 * n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d[:, 2] = np.abs(d[:, 2])
    d = d[d[:, 2] >= min_z][:n]
    return d.astype(np.float32).astype(np.float64)


def _D(h, a2):
    c2 = h[:, 2] ** 2
    return a2 / (math.pi * (c2 * (a2 - 1) + 1) ** 2)


def _lam(c, a2):
    t2 = (1 - c * c) / (c * c)
    return (np.sqrt(1 + a2 * t2) - 1) / 2


def _schlick(f0, c):
    return f0[None, :] + (1 - f0[None, :]) * ((1 - c) ** 5)[:, None]


def _ggx(v, l, alpha, f0):
    """(f cos_l, density) of GGX reflection, float64"""
    a2 = alpha * alpha
    h = v + l
    h /= np.linalg.norm(h, axis=1, keepdims=True)
    d = _D(h, a2)
    lv, ll = _lam(v[:, 2], a2), _lam(l[:, 2], a2)
    g2 = 1 / (1 + lv + ll)
    g1 = 1 / (1 + lv)
    f = _schlick(np.asarray(f0, np.float64), np.sum(v * h, axis=1))
    value = f * (d * g2 / (4 * v[:, 2]))[:, None]
    density = g1 * d / (4 * v[:, 2])
    return value, density


def _energy(cos_v, r):
    x = np.sqrt(np.clip(cos_v, 0, 1)) * 31
    y = np.clip(r, 0, 1) * 31
    i = np.minimum(x.astype(int), 30)
    j = np.minimum(int(y), 30)
    fx, fy = x - i, y - j
    below = (1 - fx) * TABLE[j, i] + fx * TABLE[j, i + 1]
    above = (1 - fx) * TABLE[j + 1, i] + fx * TABLE[j + 1, i + 1]
    return (1 - fy) * below + fy * above


def _compensation(f0, cos_v, r):
    e = np.minimum(_energy(cos_v, r), 1.0)
    return 1 + np.asarray(f0)[None, :] * ((1 - e) / e)[:, None]


def _coat(v, l, alpha, rho):
    value, density = _ggx(v, l, alpha, [COAT_F0] * 3)
    t = lambda c: 1 - (1 - c / 2) ** 5
    base = 28 / (23 * math.pi) * (1 - COAT_F0) * t(l[:, 2]) * t(v[:, 2]) * l[:, 2]
    return value + base[:, None] * np.asarray(rho)[None, :], 0.5 * (density + l[:, 2] / math.pi)


def _expected(lobe, v, l, alpha, color):
    if lobe == "conductor":
        value, density = _ggx(v, l, alpha, color)
        return value * _compensation(color, v[:, 2], math.sqrt(alpha)), density
    if lobe == "gloss":
        return _coat(v, l, alpha, color)
    if lobe == "lambert":
        c = np.maximum(l[:, 2], 0) / math.pi
        return c[:, None] * np.asarray(color)[None, :], c
    return np.tile(np.asarray(color) / (4 * math.pi), (len(v), 1)), np.full(len(v), 1 / (4 * math.pi))


def _close(got, want, rtol, what):
    err = np.abs(got - want) / np.maximum(np.abs(want), 1e-30)
    worst = np.unravel_index(np.argmax(err), err.shape)
    assert np.quantile(err, 0.999) < rtol and err.max() < 10 * rtol, \
        f"{what}: relative error max {err.max():.3g} at {worst} (got {got[worst]}, want {want[worst]})"


@pytest.mark.parametrize("lobe,alpha", [(lobe, a) for lobe in ("conductor", "gloss") for a in ALPHAS] +
                         [("lambert", 0.0), ("phase", 0.0)])
def test_value_and_density_formulas(rt, lobe, alpha):
    rng = np.random.default_rng(int(alpha * 1000) + len(lobe))
    v = _hemisphere(rng, 20000)
    # (a lobe without roughness ignores alpha)
    l = _hemisphere(rng, 20000)
    color = [0.9, 0.6, 0.3]
    out = rt.scatter_probe(lobe, v, l, color=color, alpha=alpha, roughness=math.sqrt(alpha), seed=1)
    value, density = _expected(lobe, v, l, alpha, color)
    # float32 arithmetic: D's peak is sharp for small alpha (relative error grows as the half vector's angle shrinks)
    rtol = 2e-4 if alpha >= 0.05 else 2e-3
    _close(out["density"].astype(np.float64), density, rtol, f"{lobe} density")
    _close(out["value"].astype(np.float64), value, rtol, f"{lobe} value")


@pytest.mark.parametrize("lobe", ["conductor", "gloss"])
@pytest.mark.parametrize("alpha", ALPHAS)
def test_reciprocity(rt, lobe, alpha):
    """f(v, l) = f(l, v): the conductor's single-scattering lobe (its value over its compensation, which depends on
    v alone) and the coat"""
    rng = np.random.default_rng(7 + int(alpha * 100))
    v = _hemisphere(rng, 20000, 0.05)
    l = _hemisphere(rng, 20000, 0.05)
    color = [1.0, 0.8, 0.5]
    a = rt.scatter_probe(lobe, v, l, color=color, alpha=alpha, roughness=math.sqrt(alpha))["value"].astype(np.float64)
    b = rt.scatter_probe(lobe, l, v, color=color, alpha=alpha, roughness=math.sqrt(alpha))["value"].astype(np.float64)
    if lobe == "conductor":
        a /= _compensation(color, v[:, 2], math.sqrt(alpha))
        b /= _compensation(color, l[:, 2], math.sqrt(alpha))
    fa = a / l[:, 2:3]
    fb = b / v[:, 2:3]
    _close(fa, fb, 2e-4 if alpha >= 0.05 else 3e-3, f"{lobe} reciprocity")


@pytest.mark.parametrize("lobe", ["conductor", "gloss", "lambert", "phase"])
@pytest.mark.parametrize("alpha", [0.05, 0.5])
def test_draws_are_consistent_with_evaluation(rt, lobe, alpha):
    """A drawn direction's density is the lobe's density there, and its weight is f cos over that density"""
    rng = np.random.default_rng(99)
    n = 50000
    v = _hemisphere(rng, n, 0.05)
    color = [0.9, 0.6, 0.3]
    kw = dict(color=color, alpha=alpha, roughness=math.sqrt(alpha), seed=5)
    d = rt.scatter_probe(lobe, v, v, **kw)
    ok = d["drawn_density"] > 0
    assert ok.mean() > 0.5
    drawn = d["drawn"][ok]
    assert np.allclose(np.linalg.norm(drawn, axis=1), 1, atol=1e-5)
    e = rt.scatter_probe(lobe, v[ok], drawn, **kw)
    _close(e["density"].astype(np.float64), d["drawn_density"][ok].astype(np.float64), 1e-4, f"{lobe} drawn density")
    _close(d["weight"][ok].astype(np.float64) * e["density"][:, None].astype(np.float64),
           e["value"].astype(np.float64), 2e-4, f"{lobe} drawn weight")


def _chi_square_p(observed, expected):
    from scipy import stats
    keep = expected >= 5
    chi2 = np.sum((observed[keep] - expected[keep]) ** 2 / expected[keep])
    return stats.chi2.sf(chi2, keep.sum() - 1), keep.sum()


@pytest.mark.parametrize("lobe,alpha,view_z", [("conductor", 0.05, 0.9), ("conductor", 0.3, 0.5),
                                               ("conductor", 0.8, 0.2), ("gloss", 0.2, 0.7), ("gloss", 0.6, 0.3),
                                               ("lambert", 0.0, 0.6)])
def test_draws_follow_the_density(rt, lobe, alpha, view_z):
    """Drawn directions from one view, binned in (cos theta, phi), against the float64 density's integral over each
    cell (midpoint quadrature on a 48 x 48 grid per cell: below 1e-3 relative error on the sharpest lobe's cell)"""
    n = 400000
    v1 = np.array([math.sqrt(1 - view_z ** 2), 0.0, view_z], np.float32).astype(np.float64)
    v = np.tile(v1, (n, 1))
    d = rt.scatter_probe(lobe, v, v, color=[1, 1, 1], alpha=alpha, roughness=math.sqrt(alpha), seed=11)
    ok = d["drawn_density"] > 0
    w = d["drawn"][ok].astype(np.float64)
    nc, nphi, sub = 12, 24, 48
    ci = np.minimum((w[:, 2] * nc).astype(int), nc - 1)
    pj = np.minimum(((np.arctan2(w[:, 1], w[:, 0]) + math.pi) / (2 * math.pi) * nphi).astype(int), nphi - 1)
    observed = np.bincount(ci * nphi + pj, minlength=nc * nphi).astype(np.float64)
    # the density's integral over each cell: dw = dcos dphi
    cs = (np.arange(nc * sub) + 0.5) / (nc * sub)
    ps = (np.arange(nphi * sub) + 0.5) / (nphi * sub) * 2 * math.pi - math.pi
    C, P = np.meshgrid(cs, ps, indexing="ij")
    s = np.sqrt(1 - C ** 2)
    dirs = np.stack([s * np.cos(P), s * np.sin(P), C], -1).reshape(-1, 3)
    _, dens = _expected(lobe, np.tile(v1, (len(dirs), 1)), dirs, max(alpha, 1e-3), [1, 1, 1])
    cell = dens.reshape(nc, sub, nphi, sub).sum(axis=(1, 3)) * (1 / (nc * sub)) * (2 * math.pi / (nphi * sub))
    expected = cell.reshape(-1) * n
    # the drawn share: what the density gives the upper hemisphere (a conductor's reflections below it end the path)
    assert abs(ok.sum() - expected.sum()) < 5 * math.sqrt(expected.sum()) + 0.002 * n
    p, cells = _chi_square_p(observed, expected * ok.sum() / expected.sum())
    assert p > 1e-4, f"{lobe} alpha {alpha} view {view_z}: chi-square p = {p:.2e} over {cells} cells"


@pytest.mark.parametrize("alpha", [0.02, 0.1, 0.3, 0.6, 1.0])
@pytest.mark.parametrize("view_z", [0.15, 0.4, 0.9])
def test_white_conductor_energy(rt, alpha, view_z):
    """F0 = 1: the draws' mean weight is the compensated energy, 1 within the table's bilinear reading; over the
    compensation it is E_ss, the table's value (the reference's ggx_energy compensated and factor checks)"""
    n = 400000
    v = np.tile(np.array([math.sqrt(1 - view_z ** 2), 0.0, view_z]), (n, 1))
    r = math.sqrt(alpha)
    d = rt.scatter_probe("conductor", v, v, color=[1, 1, 1], alpha=alpha, roughness=r, seed=3)
    wt = d["weight"][:, 0].astype(np.float64)
    mean, se = wt.mean(), wt.std() / math.sqrt(n)
    comp = _compensation([1.0, 1.0, 1.0], np.array([view_z]), r)[0, 0]
    e_ss = (wt / comp).mean()
    # (the table's bilinear reading is within 0.5 % of E_ss by quadrature at these nodes' neighbours)
    assert abs(e_ss / _energy(np.array([view_z]), r)[0] - 1) < 5 * se + 6e-3, f"E_ss {e_ss} vs table"
    assert abs(mean - 1.0) < 5 * se + 8e-3, f"compensated energy {mean} +- {se}"


def test_mirror_below_the_smallest_alpha(rt):
    rng = np.random.default_rng(5)
    v = _hemisphere(rng, 1000, 0.01)
    f0 = [0.95, 0.64, 0.54]
    d = rt.scatter_probe("conductor", v, v, color=f0, alpha=5e-4, roughness=math.sqrt(5e-4), seed=2)
    assert np.all(d["drawn_density"] == 0)
    assert np.allclose(d["drawn"], v * np.array([-1, -1, 1]), atol=1e-6)
    _close(d["weight"].astype(np.float64), _schlick(np.array(f0), v[:, 2]), 1e-5, "mirror weight")
    e = rt.scatter_probe("conductor", v, d["drawn"], color=f0, alpha=5e-4, roughness=math.sqrt(5e-4))
    assert np.all(e["density"] == 0) and np.all(e["value"] == 0)


def test_probe_refusals(rt):
    v = np.array([[0, 0, 1.0]])
    with pytest.raises(ValueError):
        rt.scatter_probe(7, v, v)
    with pytest.raises(ValueError):
        rt.scatter_probe("conductor", v, v, roughness=2.0)


@pytest.mark.parametrize("fuzz", [0.0, 0.3, 0.9])
def test_books_metal(rt, fuzz):
    """The books' fuzzed metal (src/gpu/material.cuh: metal): the mirror direction, unit, plus fuzz times a point
    uniform on the unit sphere, absorbed when that leaves below the surface; the weight is the albedo"""
    rng = np.random.default_rng(31)
    n = 200000
    v = _hemisphere(rng, n, 0.05)
    v[: n // 2] = np.array([0.0, 0.0, 1.0])  # half at normal incidence, for the sphere point's distribution
    albedo = [0.8, 0.6, 0.2]
    d = rt.scatter_probe("metal", v, v, color=albedo, roughness=fuzz, seed=8)
    ok = np.linalg.norm(d["drawn"], axis=1) > 0
    mirror = v * np.array([-1, -1, 1])
    u = (d["drawn"][ok].astype(np.float64) - mirror[ok]) / max(fuzz, 1e-30)
    assert np.all(d["drawn_density"] == 0)
    assert np.all(d["weight"][ok] == np.float32(albedo))
    if fuzz == 0:
        assert ok.all() and np.allclose(d["drawn"], mirror, atol=1e-6)
        return
    assert np.allclose(np.linalg.norm(u, axis=1), 1, atol=2e-5 / fuzz)
    assert np.all(d["drawn"][ok][:, 2] > 0)
    # absorbed exactly where the fuzzed direction would go below the surface: never at normal incidence for fuzz < 1
    normal = ok[: n // 2]
    assert normal.all() if fuzz < 1 else normal.mean() > 0.99
    uz = u[: normal.sum(), 2]
    assert stats_p(np.clip(uz, -1, 1), "uniform", (-1, 2)) > 1e-4, "the sphere point's z is uniform in [-1, 1]"


def stats_p(x, dist, args):
    from scipy import stats
    return stats.kstest(x, dist, args=args).pvalue


@pytest.mark.parametrize("side,ior", [("glass_entering", 1.5), ("glass_leaving", 1.5), ("glass_entering", 1.33)])
def test_books_glass(rt, side, ior):
    """The books' dielectric: the mirror direction with Schlick's probability (r0 = ((1 - eta) / (1 + eta))^2, eta the
    ratio of indices across the interface), always beyond the critical angle, else Snell's refraction: in the plane of
    incidence, sin theta_t = eta sin theta_i, into the other side; weight 1"""
    n_views, per = 12, 40000
    eta = 1 / ior if side == "glass_entering" else ior
    for k in range(n_views):
        cos_i = 0.05 + 0.95 * k / (n_views - 1)
        v1 = np.array([math.sqrt(1 - cos_i ** 2), 0.0, cos_i], np.float32).astype(np.float64)
        v = np.tile(v1, (per, 1))
        d = rt.scatter_probe(side, v, v, alpha=ior, seed=100 + k)
        w = d["drawn"].astype(np.float64)
        assert np.all(d["weight"] == 1)
        reflected = w[:, 2] > 0
        mirror = v1 * np.array([-1, -1, 1])
        assert np.allclose(w[reflected], mirror, atol=1e-6)
        ci = v1[2]
        tir = 1 - eta ** 2 * (1 - ci ** 2) < 0
        if tir:
            assert reflected.all()
            continue
        wt = w[~reflected]
        wt /= np.linalg.norm(wt, axis=1, keepdims=True)
        assert np.allclose(wt[:, 1], 0, atol=1e-6)                   # the plane of incidence
        assert np.allclose(np.hypot(wt[:, 0], wt[:, 1]), eta * math.sqrt(1 - ci ** 2), atol=2e-6)  # Snell
        assert np.all(wt[:, 0] < 0 + 1e-7) or ci > 0.999               # onward, across the normal's plane
        r0 = ((1 - eta) / (1 + eta)) ** 2
        schlick = r0 + (1 - r0) * (1 - ci) ** 5
        from scipy import stats
        p = stats.binomtest(int(reflected.sum()), per, schlick).pvalue
        assert p > 1e-4, f"cos {ci:.3f}: {reflected.mean():.4f} reflected, Schlick {schlick:.4f}"
