"""The quality integrator's host tables on the CPU: the GGX energy table, the environment map's densities and alias
table, the light set, and the triangle scene's geometry and tree.

Pins:
- GGX E_ss: against the reference's compiled-in table (tests/golden/ggx_energy_ref.json, written from
  src/gpu/ggx_energy.cuh by tests/golden/make_ggx_energy.py), and against an independent Monte Carlo estimate that
  samples l uniformly over the hemisphere and evaluates D G2 / (4 cos_v) (no visible-normal sampling, no V form);
- environment: a map of radiance 1 has the density 1/(4 pi) everywhere and S = 4 pi; every alias table's implied
  chances equal the cells' shares to within the coins' grid, for maps of eight orders of magnitude of dynamic range;
- lamps: the Cornell lamp's area, normal and radiance; Veach's four spheres; the glass ball, which is in the books'
  lights list but doesn't glow, is no lamp (src/gpu/lamps.cuh:1-8).
"""
import json
import math
import os
import re

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def _compiled_table():
    text = open(os.path.join(ROOT, "raytracing_amd", "csrc", "kernels", "ggx_energy_table.hpp")).read()
    body = text[text.index("= {") + 3:text.rindex("};")]
    body = re.sub(r"/\*.*?\*/", " ", body, flags=re.S)
    return np.array([float(t.strip().rstrip("f")) for t in body.split(",") if t.strip()]).reshape(32, 32)


def test_ggx_table_matches_reference_table():
    ref = np.array(json.load(open(os.path.join(HERE, "golden", "ggx_energy_ref.json")))["values"]).reshape(32, 32)
    mine = _compiled_table()
    assert np.all(mine[0] == 1.0) and np.all(mine[:, 0] == 1.0)
    assert np.all((mine > 0) & (mine <= 1.0 + 1e-6))
    diff = np.abs(mine - ref)
    # the reference's own quadrature is least accurate at its sharpest, most grazing node (alpha = mu = 0.001), where
    # an independent estimate sides with this table (test_ggx_table_against_monte_carlo)
    diff[1, 1] = 0
    assert diff.max() < 1.5e-4, f"largest difference {diff.max()} at {np.unravel_index(diff.argmax(), diff.shape)}"
    assert abs(mine[1, 1] - ref[1, 1]) < 1e-3


def _energy_uniform_hemisphere(cos_v, alpha, n, seed):
    """E_ss with F = 1 by uniform hemisphere sampling: f cos_l / p = D G2 / (4 cos_v) * 2 pi"""
    rng = np.random.default_rng(seed)
    u1, u2 = rng.random(n), rng.random(n)
    cl = u1  # uniform over the hemisphere: cos theta uniform
    sl = np.sqrt(1 - cl * cl)
    phi = 2 * np.pi * u2
    l = np.stack([sl * np.cos(phi), sl * np.sin(phi), cl], 1)
    v = np.array([math.sqrt(1 - cos_v * cos_v), 0.0, cos_v])
    h = l + v
    h /= np.linalg.norm(h, axis=1, keepdims=True)
    ch = h[:, 2]
    a2 = alpha * alpha
    d = a2 / (np.pi * (ch * ch * (a2 - 1) + 1) ** 2)
    lam = lambda c: (np.sqrt(1 + a2 * (1 - c * c) / (c * c)) - 1) / 2
    g2 = 1 / (1 + lam(cos_v) + lam(cl))
    w = d * g2 / (4 * cos_v) * 2 * np.pi
    return w.mean(), w.std() / math.sqrt(n)


@pytest.mark.parametrize("i,j", [(31, 8), (31, 20), (31, 31), (16, 16), (8, 24), (24, 12)])
def test_ggx_table_against_monte_carlo(i, j):
    mine = _compiled_table()
    r = j / 31
    mu = (i / 31) ** 2
    e, sigma = _energy_uniform_hemisphere(mu, r * r, 4_000_000, seed=100 * i + j)
    assert abs(mine[j, i] - e) < 5 * sigma + 2e-4, f"node ({i}, {j}): table {mine[j, i]}, estimate {e} +- {sigma}"


def _implied_chances(q, alias, bins):
    chance = q.astype(np.float64).copy()
    np.add.at(chance, alias, 1.0 - q.astype(np.float64))
    return chance / bins


def test_environment_white_map(rt):
    t = rt.environment_tables(np.ones((8, 16, 3), np.float32))
    assert abs(t.total - 4 * math.pi) < 1e-12
    assert np.all(t.texels[..., :3] == 1.0)
    assert np.allclose(t.texels[..., 3], np.float32(1 / (4 * math.pi)), rtol=0, atol=0)
    assert t.row_cos[0] == 1.0 and t.row_cos[-1] == -1.0 and np.all(np.diff(t.row_cos) < 0)
    assert t.log2_bins == 7
    chance = _implied_chances(t.bin_q, t.bin_alias, 128)
    omega = np.repeat(2 * np.pi / 16 * -np.diff(t.row_cos.astype(np.float64)), 16)
    assert np.allclose(chance, omega / (4 * np.pi), atol=2 ** -24)


def test_environment_black_map(rt):
    t = rt.environment_tables(np.zeros((4, 8, 3), np.float32))
    assert t.total == 0 and t.log2_bins == 0 and len(t.bin_q) == 0
    assert np.all(t.texels[..., 3] == 0)


@pytest.mark.parametrize("h,w,seed", [(1, 1, 0), (3, 5, 1), (64, 128, 2), (512, 1024, 3)])
def test_environment_alias_table_exact(rt, h, w, seed):
    rng = np.random.default_rng(seed)
    rgb = (rng.random((h, w, 3)) ** 8 * 10.0 ** rng.uniform(-3, 1, (h, w, 1))).astype(np.float32)
    if h * w > 4:
        rgb[rng.integers(0, h), rng.integers(0, w)] = 6e4  # a sun
        rgb[0, : min(w, 3)] = 0  # dark cells: never drawn
    t = rt.environment_tables(rgb)
    bins = 1 << t.log2_bins
    assert bins >= h * w and bins < 2 * max(h * w, 2)
    # the coins lie on the 2^-24 grid, so the comparison random_real < q has exactly the chance q
    assert np.all(t.bin_q * 2 ** 24 == np.round(t.bin_q * 2 ** 24))
    assert np.all((t.bin_q >= 0) & (t.bin_q <= 1))
    rgb64 = rgb.astype(np.float64)
    y = 0.2126 * rgb64[..., 0] + 0.7152 * rgb64[..., 1] + 0.0722 * rgb64[..., 2]
    cos = t.row_cos.astype(np.float64)
    border = np.cos(np.pi * np.arange(h + 1) / h)
    border[0], border[-1] = 1, -1
    omega = 2 * np.pi / w * (border[:-1] - border[1:])
    share = (y * omega[:, None]).ravel()
    total = share.sum()
    assert abs(t.total - total) <= 1e-12 * total
    assert np.allclose(cos, border.astype(np.float32))
    chance = _implied_chances(t.bin_q, t.bin_alias, bins)
    expected = np.zeros(bins)
    expected[: h * w] = share / total
    # each cell's chance is its share up to the rounding of its own coin (one step of 2^-24 per bin)
    assert np.all(np.abs(chance - expected) <= 1.01 * 2 ** -24 / bins * 2 + 1e-15)
    lit = expected > 0
    assert np.all(chance[lit] > 0), "every cell with light can be drawn"
    assert np.all(chance[~lit] == 0), "a dark cell or padding is never drawn"
    assert np.all(expected[t.bin_alias[t.bin_q < 1]] > 0), "every alias is a cell with light"
    # the texels' densities: Y / S per steradian
    assert np.allclose(t.texels[..., 3], (y / total).astype(np.float32), rtol=1e-6, atol=0)


def test_lamps_cornell(rt):
    s = rt.build_scene("cornell", width=32, samples=1)
    lamps = rt.scene_lamps(s)
    assert lamps.shape == (1, 20)
    lamp = lamps[0]
    assert lamp[7].view(np.int32) == 0  # a quad
    assert lamp[3] == np.float32(130 * 105)
    assert np.all(lamp[12:15] == np.float32([0, -1, 0]))
    assert np.all(lamp[16:19] == 15)
    _, _, _, quad_ids, _ = rt.scene_bvh(s, "default")
    assert lamp[11].view(np.int32) == quad_ids[5], "the lamp is the world's sixth quad"


def test_lamps_glass_ball_is_no_lamp(rt):
    s = rt.build_scene("glass", width=32, samples=1)
    assert rt.scene_lamps(s).shape[0] == 1


def test_lamps_veach(rt):
    s = rt.build_scene("veach", width=32, samples=1)
    lamps = rt.scene_lamps(s)
    assert lamps.shape[0] == 4
    radii = sorted(lamps[:, 3])
    assert np.allclose(radii, [0.05, 0.15, 0.4, 1.0])
    for l in lamps:
        assert l[7].view(np.int32) == 1  # a sphere
        y = 0.2126 * l[16] + 0.7152 * l[17] + 0.0722 * l[18]
        assert abs(y - 4 / l[3] ** 2) < 1e-3 * y


@pytest.mark.slow
def test_triangle_scene_tree(rt):
    s = rt.build_scene("triangles", width=64, samples=1)
    d = s.desc
    tris = sum(d.meshes[d.instances[k].mesh].triangle_count for k in range(d.instance_count))
    assert tris == 20 * 4 ** 8 + 64 * 20 * 4 ** 5
    # one level (the default): every instance's triangles baked into the scene's tree
    boxes, ints, _, _, depth = rt.scene_bvh(s, "default")
    leaf = ints[:, 1] != 0
    counts = ints[leaf, 1] & 255
    kinds = ints[leaf, 1] >> 8
    assert counts[kinds == 3].sum() == tris and counts[kinds == 2].sum() == 1
    assert counts.max() <= 8 and depth <= 64
    # two levels: the icosphere, placed once, baked; the two blobs, placed 32 times each, one leaf per instance
    # (kind 4, one instance a leaf)
    uses = np.bincount([d.instances[k].mesh for k in range(d.instance_count)], minlength=d.mesh_count)
    baked = sum(d.meshes[m].triangle_count for m in range(d.mesh_count) if uses[m] == 1)
    assert baked == 20 * 4 ** 8 and sorted(uses) == [1, 32, 32]
    b2, i2, _, _, depth2 = rt.scene_bvh(s, "two_level")
    leaf2 = i2[:, 1] != 0
    counts2, kinds2 = i2[leaf2, 1] & 255, i2[leaf2, 1] >> 8
    assert counts2[kinds2 == 3].sum() == baked and counts2[kinds2 == 2].sum() == 1
    assert np.all(counts2[kinds2 == 4] == 1) and np.count_nonzero(kinds2 == 4) == 64
    assert len(b2) < len(boxes) and depth2 <= depth
    # every inner node's box holds both children's
    inner = np.nonzero(~leaf)[0]
    first, second = inner + 1, ints[inner, 0]
    for child in (first, second):
        assert np.all(boxes[child, :3] >= boxes[inner, :3]) and np.all(boxes[child, 3:] <= boxes[inner, 3:])
