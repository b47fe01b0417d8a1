"""Tone maps beyond the books' bytes on the GPU (rt_session_display_toned), against the same curves evaluated in double
from their published equations (sRGB, IEC 61966-2-1; AgX's inset/outset matrices and closed-form curve; Khronos PBR
Neutral; interleaved gradient noise): every byte within one level of the double reference, nearly all equal, on
renders with lamps bright enough to reach each curve's shoulder. And the refusals the reference makes."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def srgb(v):
    v = np.clip(v, 0, 1)
    return np.where(v >= 1, 1.0, np.where(v <= 0.0031308, 12.92 * v, 1.055 * np.power(v, 1 / 2.4) - 0.055))


def agx(c):
    inset = np.array([[0.842479062253094, 0.0784335999999992, 0.0792237451477643],
                      [0.0423282422610123, 0.878468636469772, 0.0791661274605434],
                      [0.0423756549057051, 0.0784336, 0.879142973793104]])
    outset = np.array([[1.19687900512017, -0.0980208811401368, -0.0990297440797205],
                       [-0.0528968517574562, 1.15190312990417, -0.0989611768448433],
                       [-0.0529716355144438, -0.0980434501171241, 1.15107367264116]])
    x = c @ inset.T
    with np.errstate(divide="ignore"):
        pos = np.clip((np.log2(x) + 12.47393) / (4.026069 + 12.47393), 0, 1)
    t = 10 / 16.5
    d = pos - t
    m = np.abs(d)
    low = 0.5 + 2 * d * np.power(1 + 59.507875 * m ** 3, -1 / 3)
    high = 0.5 + 2 * d * np.power(1 + 69.86278913545539 * m ** 3.25, -4 / 13)
    curve = np.where(pos <= t, low, high)
    return np.clip(curve @ outset.T, 0, 1)


def neutral(c):
    f90, ks, kd = 0.04, 0.76, 0.15
    x = c.min(axis=-1, keepdims=True)
    f = np.where(x <= 2 * f90, x - 6.25 * x * x, f90)
    s = c - f
    p = s.max(axis=-1, keepdims=True)
    with np.errstate(divide="ignore", invalid="ignore"):
        pn = 1 - (1 - ks) ** 2 / (p + 1 - 2 * ks)
        g = 1 / (kd * (p - pn) + 1)
        out = s * (pn / p * g) + pn * (1 - g)
    return srgb(np.where(p <= ks, s, out))


def reference(sums, passes, op, exposure):
    h, w, _ = sums.shape
    light = (sums * (1.0 / passes)).astype(np.float32).astype(np.float64) * float(np.float32(2.0 ** exposure))
    light = np.where(light > 0, np.minimum(light, 3e38), 0.0)
    enc = {"none": srgb, "agx": agx, "neutral": neutral}[op](light)
    y, x = np.mgrid[0:h, 0:w].astype(np.float64)
    inner = 0.06711056 * x + 0.00583715 * y
    dither = np.mod(52.9829189 * np.mod(inner, 1.0), 1.0) - 0.5
    q = np.floor(255 * enc + 0.5 + dither[..., None])
    q = np.where(enc <= 0, 0, np.where(enc >= 1, 255, np.clip(q, 0, 255)))
    return q.astype(np.uint8)


@pytest.mark.parametrize("name", ["cornell", "final"])
@pytest.mark.parametrize("op,exposure", [("none", 0.0), ("none", 1.5), ("agx", 0.0), ("agx", -2.0), ("neutral", 0.0),
                                         ("neutral", 3.0)])
def test_tone_maps_match_double(rt, name, op, exposure):
    scene = rt.build_scene(name, width=96, samples=4, earth_width=256)
    with rt.Session(scene, seed=2) as s:
        s.render_to(scene.passes)
        sums = s.sums()
        got = s.display(op, exposure)
    want = reference(sums, scene.passes, op, exposure)
    assert np.all(got[..., 3] == 255)
    diff = np.abs(got[..., :3].astype(int) - want.astype(int))
    assert diff.max() <= 1, f"{op} {exposure}: largest difference {diff.max()}"
    assert (diff == 0).mean() > 0.995, f"{op} {exposure}: {(diff == 0).mean():.4f} equal"
    assert got[..., :3].std() > 5  # an image, not a constant


def test_books_and_refusals(rt):
    scene = rt.build_scene("cornell", width=64, samples=4)
    with rt.Session(scene) as s:
        s.render_to(scene.passes)
        assert np.array_equal(s.display(), s.display("books", 0.0))
        assert np.array_equal(s.display(), s.display(0, 0.0))
        for bad in [("books", 1.0), ("none", 25.0), ("agx", float("nan")), (7, 0.0)]:
            with pytest.raises(ValueError):
                s.display(*bad)
    with rt.Session(scene, precision="float64") as s:
        s.render_to(1)
        with pytest.raises(ValueError, match="float32"):
            s.display("agx", 0.0)


def test_render_display_option(rt):
    """render_options::display (src/gpu/renderer.h:444-448): a tone map other than the books' fills the result's rgba
    with the display's bytes after the last pass, the same bytes a session's display gives; the pixels don't depend on
    it; the result names the kernel that rendered"""
    scene = rt.build_scene("cornell", width=64, samples=16)
    plain = rt.render(scene, seed=2, integrator="megakernel")
    assert plain.rgba is None and plain.kernel.startswith("all books")
    toned = rt.render(scene, seed=2, integrator="megakernel", tone_map="agx", exposure=0.5)
    assert np.array_equal(toned.pixels.view(np.uint64), plain.pixels.view(np.uint64))
    with rt.Session(scene, seed=2, integrator="megakernel") as s:
        s.render_to(scene.passes)
        want = s.display("agx", 0.5)
    assert np.array_equal(toned.rgba, want)
    assert np.all(toned.rgba[..., 3] == 255)


# The reference's grey bytes (tests/tone_tests.cu:1523-1550, the tone maps' published greys): a grey's byte with no
# dither; "none" exact, AgX and PBR Neutral within one level (their channels may round apart)
REFERENCE_GREYS = {
    "none": (0, [(0.001, 3), (0.01, 25), (0.05, 63), (0.18, 118), (0.5, 188)]),
    "agx": (1, [(0.001, 4), (0.01, 25), (0.05, 72), (0.18, 127), (0.5, 173), (1.0, 201), (2.0, 223), (4.0, 239),
                (16.0, 255), (100.0, 255)]),
    "neutral": (1, [(0.18, 105), (1.0, 240), (4.0, 253)]),
}


@pytest.mark.parametrize("op", ["none", "agx", "neutral"])
def test_reference_grey_bytes(rt, op):
    allowed, greys = REFERENCE_GREYS[op]
    v = np.array([[g, g, g] for g, _ in greys], np.float32)
    got = rt.tone_probe(op, v).astype(int)
    want = np.array([b for _, b in greys])[:, None]
    assert np.all(np.abs(got - want) <= allowed), f"{op}: {got.tolist()} against {want.ravel().tolist()}"


def test_quantize_with_dither(rt):
    """round(255 v + dither) with the dither in [-0.5, 0.5), black and white never dithered (src/gpu/tonemap.cu):
    the sRGB OETF's bytes in float64, but for values within rounding of a step"""
    rng = np.random.default_rng(12)
    n = 200000
    lin = rng.uniform(0, 1.2, (n, 3)).astype(np.float32)
    lin[:100] = 0.0
    lin[100:200] = 1.0
    dither = rng.uniform(-0.5, 0.5, n).astype(np.float32)
    got = rt.tone_probe("none", lin, dither=dither).astype(int)
    x = np.clip(lin.astype(np.float64), 0, 1)
    e = np.where(x <= 0.0031308, 12.92 * x, 1.055 * x ** (1 / 2.4) - 0.055)
    raw = 255 * e + 0.5 + dither[:, None]
    want = np.clip(np.floor(raw), 0, 255)
    want = np.where(x <= 0, 0, np.where(x >= 1, 255, want))
    near = np.abs(raw - np.round(raw)) < 1e-3
    assert np.all(got[:100] == 0) and np.all(got[100:200] == 255)
    assert np.all((got == want) | near), f"{np.count_nonzero((got != want) & ~near)} bytes off"
    assert np.count_nonzero(got != want) < 1e-3 * got.size


def test_tone_probe_refusals(rt):
    grey = [[0.5, 0.5, 0.5]]
    with pytest.raises(ValueError):
        rt.tone_probe("books", grey)  # the books' bytes are the display kernel's, not a tone map's
    with pytest.raises(ValueError):
        rt.tone_probe("agx", grey, exposure=float("nan"))
