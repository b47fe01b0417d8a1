"""The oracle, pinned before it is trusted (CPU only).

- random numbers against published known answers: mt19937_64's 10000th output (C++ [rand.predef]), PCG32's demo
  stream for seed 42 / sequence 54 (pcg-random.org, pcg32-demo), SplitMix64's first outputs from seed 0;
- uniform doubles as generate_canonical defines them for a 64-bit engine (x / 2^64);
- the CPU renderer: the furnace scene's closed form (every pixel exactly 1, src/gpu_scenes.ixx:33-55), two
  independent renders consistent under ImageCompare, a planted 1% fault rejected, rows rendered alone equal to the
  same rows of the whole image, and final's mean radiance near the reference's published 0.442442 (docs/gpu.md:66);
- ImageCompare on synthetic images with known answers: equal expectations pass, a 0.2% bias and a one-pixel shift fail.
"""
import numpy as np
import pytest

import oracle
from oracle import image_compare


def test_mt19937_64_known_answer():
    # The 10000th invocation of a default-constructed std::mt19937_64 produces 9981545732273789042
    assert int(oracle.mt19937_64(5489, 10000)[-1]) == 9981545732273789042


def test_random_double_is_generate_canonical():
    raw = oracle.mt19937_64(5489, 8).astype(np.float64)
    assert np.array_equal(oracle.random_doubles(5489, 8), raw * 2.0 ** -64)


def test_pcg32_demo_stream():
    # pcg32_srandom_r(&rng, 42u, 54u), then pcg32_random_r: pcg32-demo's published first outputs
    expected = [0xA15C02B7, 0x7B47F409, 0xBA1D3330, 0x83D2F293, 0xBFA4784B, 0xCBED606E]
    assert [int(x) for x in oracle.pcg32(42, 54, 6)] == expected


def test_splitmix64_known_answers():
    # splitmix64 from state 0: the state grows by the golden gamma, and the finalizer gives these outputs
    gamma = 0x9E3779B97F4A7C15
    assert oracle.mix64(gamma) == 16294208416658607535
    assert oracle.mix64((2 * gamma) % 2 ** 64) == 7960286522194355700


def test_sample_floats_are_top_24_bits():
    # random_real<float>: the top 24 bits of one output times 2^-24; the generator of (seed, pixel, pass) is
    # pcg32 seeded with mix64 hashes (src/gpu/random.cuh:62-90)
    seed, pixel, pass_ = 3, 1234, 7
    m = oracle.mix64
    initstate = m(m(m(seed) ^ pixel) ^ pass_)
    initseq = m(m(m(seed ^ 0x9E3779B97F4A7C15) ^ pass_) ^ pixel)
    raw = oracle.pcg32(initstate, initseq, 5)
    assert np.array_equal(oracle.sample_floats(seed, pixel, pass_, 5), (raw >> 8).astype(np.float32) * 2.0 ** -24)


def test_sample_words_and_doubles():
    # the raw words are pcg32 of the mix64 hashes; a double is 53 bits of two words, the first one high
    # (src/gpu/random.cuh:62-90): floor(d 2^32) is the first word, and d 2^53 is an integer below 2^53
    seed, pixel, pass_ = 5 + (2 << 32), 777, 31
    m = oracle.mix64
    raw = oracle.pcg32(m(m(m(seed) ^ pixel) ^ pass_), m(m(m(seed ^ 0x9E3779B97F4A7C15) ^ pass_) ^ pixel), 8)
    assert np.array_equal(oracle.sample_words(seed, pixel, pass_, 8), raw)
    d = oracle.sample_doubles(seed, pixel, pass_, 4)
    assert np.array_equal(np.floor(d * 2.0 ** 32).astype(np.uint64), raw[0::2].astype(np.uint64))
    scaled = d * 2.0 ** 53
    assert np.all(scaled == np.floor(scaled)) and np.all(scaled < 2.0 ** 53)
    assert np.array_equal(scaled.astype(np.uint64) & np.uint64(2 ** 21 - 1), raw[1::2].astype(np.uint64) >> np.uint64(11))


def test_noise_image_stream():
    img = oracle.noise_image(64, 9)
    raw = oracle.pcg32(0x853C49E6748FEA9B ^ 9, 0xDA3E39CB94B95BDB, 2)
    lut = (np.arange(256) / 255.0) ** 2.2
    first = [lut[raw[0] & 255], lut[(raw[0] >> 8) & 255], lut[(raw[0] >> 16) & 255]]
    assert np.array_equal(img[0, 0], np.array(first, np.float32))


def test_furnace_is_exactly_one():
    img, rays = oracle.render("furnace", width=96, spp=16)
    assert img.shape == (54, 96, 3)
    assert np.all(img == 1.0)
    assert rays >= 96 * 54 * 16


def test_rows_alone_equal_rows_of_whole_image():
    whole, _ = oracle.render("weekend", width=64, spp=4, seed=3)
    part, _ = oracle.render("weekend", width=64, spp=4, seed=3, rows=(10, 20))
    assert np.array_equal(whole[10:20], part)


def test_two_cpu_renders_are_consistent_and_a_fault_is_caught():
    a, _ = oracle.render("cornell", width=96, spp=64, seed=0)
    b, _ = oracle.render("cornell", width=96, spp=64, seed=1)
    v = image_compare.compare(a, b)
    assert v.consistent, v.summary()
    planted = image_compare.compare(a, b * 1.01)
    assert not planted.consistent, planted.summary()


def _se_of_mean(a, b):
    """The standard error of one render's mean radiance, from two independent renders of the same settings: the
    spread of their channel-averaged pixel differences over sqrt(2 N)"""
    d = a.mean(axis=2) - b.mean(axis=2)
    return float(d.std() / np.sqrt(2 * d.size))


def test_materials_mean_radiance_published():
    """docs/image-compare.md:20-27: `RayTracing materials -w 480 -s 64` gives mean radiance 0.266767, with --seed 1
    0.266774 (the reference's CPU renderer). The oracle draws the reference's own random numbers (mt19937_64 per row,
    seeded (seed << 32) | row, in the reference's order), so its renders of the same settings and seeds give the
    published means to their six digits, where two seeds differ by a standard error of 4.5e-5: a dropped term, a
    draw out of order or a wrong density moves them by far more. (Allowed: the published rounding, 5e-7, and as much
    again for the platform's sin/cos/atan2, which the reference takes from MSVC.)"""
    a, _ = oracle.render("materials", width=480, spp=64, seed=0)
    b, _ = oracle.render("materials", width=480, spp=64, seed=1)
    assert a.shape == (270, 480, 3)
    se = _se_of_mean(a, b)
    for img, published in ((a, 0.266767), (b, 0.266774)):
        print(f"materials: {img.mean():.7f} vs {published} (se of a render {se:.2e})")
        assert abs(img.mean() - published) < 1e-6


def test_cornell_mixture_mean_published():
    """docs/image-compare.md:108-114: cornell at 480x270 (the 1:1 camera widened), 1024 rays per pixel, -p mixture,
    mean radiance 0.087674 (the books' GPU kernel; -p mis 0.087681). The oracle's mixture at 64 rays per pixel, two
    seeds averaged, within 4 standard errors of both renders' noise together."""
    a, _ = oracle.render("cornell", width=480, height=270, spp=64, seed=0)
    b, _ = oracle.render("cornell", width=480, height=270, spp=64, seed=1)
    assert a.shape == (270, 480, 3)
    se = _se_of_mean(a, b)
    se_published = se * np.sqrt(64 / 1024)
    mean = (a.mean() + b.mean()) / 2
    print(f"cornell: {mean:.6f} vs 0.087674 (se {se / np.sqrt(2):.2e}, published {se_published:.2e})")
    assert abs(mean - 0.087674) < 4 * np.sqrt(se * se / 2 + se_published ** 2)


@pytest.mark.parametrize("name,published", [("cornell", 4.15), ("light", 1.59), ("final", 3.68)])
def test_rays_per_sample_published(name, published):
    """docs/quality.md:239-243: rays per sample at 480x270 under -p material, 4.15 (cornell), 1.59 (light), 3.68
    (final). A ray count doesn't depend on the albedo, so final's noise texture doesn't change it. Within half a unit
    of the published last digit, plus 0.2 % for this render's own noise and the published render's."""
    img, rays = oracle.render(name, width=480, height=270, spp=16, seed=0, sampling=1, earth_width=512)
    per_sample = rays / (480 * 270 * 16)
    print(f"{name}: {per_sample:.4f} rays per sample vs {published}")
    assert abs(per_sample - published) <= 0.005 + 0.002 * published


def test_image_compare_synthetic():
    rng = np.random.default_rng(1)
    base = rng.uniform(0.2, 0.8, (96, 128, 3))
    a = base + rng.normal(0, 0.05, base.shape)
    b = base + rng.normal(0, 0.05, base.shape)
    assert image_compare.compare(a, b).consistent
    assert not image_compare.compare(a, b * 1.002 + 0.002).consistent
    # an edge moved by one pixel: brighter on one side of it, as a sub-pixel shift of the image makes it
    step_a = np.zeros_like(base)
    step_a[:, 40:] = 1.0
    step_b = np.zeros_like(base)
    step_b[:, 41:] = 1.0
    assert not image_compare.compare(a + step_a, b + step_b).consistent
    with pytest.raises(ValueError):
        image_compare.compare(a, a)


def test_image_compare_thresholds():
    # the reference's thresholds (src/image_compare.cpp:403-458): |z| 3.42 for 4 rows at 1/1600 each, 3.02 for
    # Stouffer's sum at 1/400, and 4.85 for the largest of 2040 tiles
    assert abs(image_compare.normal_quantile(1 / 1600) - 3.42) < 0.005
    assert abs(image_compare.normal_quantile(1 / 400) - 3.02) < 0.005
    assert abs(image_compare.normal_quantile(1 / 400 / 2040) - 4.85) < 0.005


def test_pfm_round_trip(tmp_path):
    img = np.random.default_rng(0).uniform(0, 3, (7, 11, 3)).astype(np.float32).astype(np.float64)
    p = tmp_path / "x.pfm"
    image_compare.write_pfm(str(p), img)
    assert np.array_equal(image_compare.read_pfm(str(p)), img)
