"""rt_image_compare, the library's ImageCompare command (src/image_compare.cpp), against the oracle's restatement of
the same statistics (oracle/image_compare.py, scipy's incomplete beta and normal quantile): the same z-scores, tile
counts, limits and verdicts on images that agree up to noise, images with a bias over the whole image or in one
region, and the inputs it must refuse. CPU only."""
import os
import re
import subprocess

import numpy as np
import pytest

from oracle import image_compare

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TOOL = os.path.join(ROOT, "raytracing_amd", "rt_image_compare")


def run(tmp_path, a, b, *extra):
    pa, pb = str(tmp_path / "a.pfm"), str(tmp_path / "b.pfm")
    image_compare.write_pfm(pa, a)
    image_compare.write_pfm(pb, b)
    r = subprocess.run([TOOL, pa, pb, *extra], capture_output=True, text=True, timeout=120)
    return r.returncode, r.stdout


def numbers(out):
    whole = [float(m) for m in re.findall(r"^\s+(?:red|green|blue|average)\s+\S+\s+z ([-+0-9.e]+)", out, re.M)]
    tiles = re.search(r"(\d+) tested, all tiles z ([-+0-9.e]+), largest \|z\| ([0-9.e+-]+), (\d+) with", out)
    checks = {m[0].strip(): (float(m[1]), float(m[2])) for m in
              re.findall(r"^\s+(whole image|all tiles|largest tile|tile count)\s+([0-9.e+-]+) \(limit ([0-9.e+-]+)\)",
                         out, re.M)}
    return whole, tiles, checks


def pair(seed, h=150, w=200, sigma=0.2):
    rng = np.random.default_rng(seed)
    truth = 0.2 + rng.uniform(0, 1, (h, w, 3))
    return truth, truth + rng.normal(0, sigma, truth.shape), truth + rng.normal(0, sigma, truth.shape)


@pytest.mark.parametrize("case", ["noise", "bias", "region", "partial-tiles"])
def test_same_numbers_and_verdict_as_the_oracle(tmp_path, case):
    truth, a, b = pair(5 if case != "partial-tiles" else 6, *( (97, 131) if case == "partial-tiles" else (150, 200)))
    if case == "bias":
        b = b * 1.02
    if case == "region":
        b[40:70, 60:100] += 0.15
    code, out = run(tmp_path, a, b)
    v = image_compare.compare(image_compare.read_pfm(str(tmp_path / "a.pfm")),
                              image_compare.read_pfm(str(tmp_path / "b.pfm")))
    whole, tiles, checks = numbers(out)
    assert np.allclose(whole, v.whole_image_z, rtol=1e-6, atol=1e-9), (whole, v.whole_image_z)
    assert int(tiles[1]) == v.tested_tiles and int(tiles[4]) == v.tiles_over_3
    assert abs(float(tiles[2]) - v.all_tiles_z) < 1e-6 * max(1, abs(v.all_tiles_z))
    assert abs(float(tiles[3]) - v.largest_tile_z) < 1e-6 * max(1, v.largest_tile_z)
    for name, c in v.checks.items():
        assert abs(checks[name][0] - c["value"]) < 1e-6 * max(1, c["value"]), name
        assert abs(checks[name][1] - c["limit"]) < 1e-6 * max(1, c["limit"]), name
    assert code == (0 if v.consistent else 1), out
    assert (case in ("noise", "partial-tiles")) == v.consistent


def test_refusals(tmp_path):
    _, a, b = pair(7, 40, 50)
    assert run(tmp_path, a, a)[0] == 2  # identical: not two independent renders
    assert run(tmp_path, a, b[:, :49])[0] == 2  # different sizes
    bad = b.copy()
    bad[3, 4, 1] = np.nan
    assert run(tmp_path, a, bad)[0] == 2
    r = subprocess.run([TOOL, str(tmp_path / "missing.pfm"), str(tmp_path / "a.pfm")], capture_output=True, text=True)
    assert r.returncode == 2


def test_tile_option_and_big_endian(tmp_path):
    _, a, b = pair(8, 64, 64)
    code, out = run(tmp_path, a, b, "--tile", "16")
    assert "Tiles of 16x16 pixels: 16 tested" in out
    # a big-endian file (positive scale) reads the same
    pb = tmp_path / "b_be.pfm"
    with open(pb, "wb") as f:
        f.write(b"PF\n64 64\n1.0\n")
        f.write(np.ascontiguousarray(b[::-1].astype(">f4")).tobytes())
    r = subprocess.run([TOOL, str(tmp_path / "a.pfm"), str(pb), "--tile", "16"], capture_output=True, text=True)
    assert r.returncode == code and r.stdout.split("\n", 2)[2] == out.split("\n", 2)[2]
