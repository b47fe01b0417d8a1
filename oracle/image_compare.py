"""ORACLE — test infrastructure only. ImageCompare restated in numpy (DaniilK19/RayTracing src/image_compare.cpp).

Are two renders the same picture up to Monte Carlo noise? d = B - A per pixel and channel, tested four ways, each
with a false-alarm rate of 1/400 (src/image_compare.cpp:403-458):
  1. whole image: the mean of d per channel (r, g, b, average), Student's t turned into an equally rare z;
  2. all tiles: Stouffer's sum of the 32x32 tiles' z (on the channel average), clamped to +-40, over sqrt(T);
  3. the largest tile's |z| against the Bonferroni threshold for T tiles;
  4. the number of tiles with |z| > 3 against a binomial threshold.
Plus the sign test over pixels that differ by more than 0.1% (src/image_compare.cpp:263-297).

The regularized incomplete beta function and the normal quantile come from scipy (the same functions the reference
evaluates by Lentz's continued fraction and by bisection, src/image_compare.cpp:65-170).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
from scipy import special, stats

ROUNDING_MARGIN = 1e-6
SIGN_TEST_THRESHOLD = 1000 * ROUNDING_MARGIN
FALSE_ALARM_RATE = 1.0 / 400
TILE = 32


def normal_tail(z):
    return special.erfc(np.asarray(z) / np.sqrt(2.0))


def normal_quantile(p: float) -> float:
    """The z with P(|Z| > z) = p (src/image_compare.cpp:71-92)"""
    if not p > 0:
        return float("inf")
    return float(stats.norm.isf(p / 2.0))


def normal_equivalent(t, dof):
    """The z exactly as rare as Student's t with dof degrees of freedom (src/image_compare.cpp:158-170)"""
    t = np.asarray(t, dtype=np.float64)
    dof = np.asarray(dof, dtype=np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        x = dof / (dof + t * t)
        p = special.betainc(dof / 2.0, 0.5, x)
        z = np.where(p > 0, stats.norm.isf(np.maximum(p, 1e-300) / 2.0), np.sqrt(dof * np.log1p(t * t / dof)))
        z = np.where(np.isinf(t), np.inf, z)
    return np.copysign(z, t)


def test_mean(count, mean, var, level):
    """mean / SE beyond the rounding margin, as a normal z (src/image_compare.cpp:241-253); vectorized"""
    count = np.asarray(count, dtype=np.float64)
    se = np.sqrt(var / count)
    beyond = np.abs(mean) - ROUNDING_MARGIN * level
    with np.errstate(divide="ignore", invalid="ignore"):
        t = np.where(se > 0, beyond / se, np.inf)
    z = np.where(beyond > 0, normal_equivalent(t, count - 1), 0.0)
    return np.copysign(z, mean), se


def binomial_tail(n: int, q: float, k: int) -> float:
    return float(stats.binom.sf(k - 1, n, q))


@dataclass
class Verdict:
    consistent: bool
    checks: dict = field(default_factory=dict)
    whole_image_z: np.ndarray = None
    all_tiles_z: float = 0.0
    largest_tile_z: float = 0.0
    tiles_over_3: int = 0
    tested_tiles: int = 0
    sign_z: float = 0.0
    mean_a: float = 0.0
    mean_b: float = 0.0

    def summary(self) -> str:
        failed = [k for k, v in self.checks.items() if v["failed"]]
        parts = [f"{k} {v['value']:.2f} (limit {v['limit']:.2f})" for k, v in self.checks.items()]
        return ("consistent" if self.consistent else f"DIFFERENT ({', '.join(failed)})") + ": " + "; ".join(parts)


def compare(a: np.ndarray, b: np.ndarray, tile: int = TILE) -> Verdict:
    """a, b: (height, width, 3) linear images. Raises ValueError for inputs ImageCompare refuses."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if a.shape != b.shape or a.ndim != 3 or a.shape[2] != 3:
        raise ValueError("the images have different sizes")
    if a.shape[0] * a.shape[1] < 2:
        raise ValueError("an image of one pixel has no spread to compare with")
    if not (np.isfinite(a).all() and np.isfinite(b).all()):
        raise ValueError("values that are NaN or infinite")
    h, w, _ = a.shape
    d = b - a
    level = (np.abs(a) + np.abs(b)) / 2
    d4 = np.concatenate([d, d.mean(axis=2, keepdims=True)], axis=2)  # r, g, b, average
    level4 = np.concatenate([level, level.mean(axis=2, keepdims=True)], axis=2)
    n = h * w
    if np.all(d4 == 0):
        raise ValueError("A and B are identical: the comparison needs two independent renders")
    flat = d4.reshape(n, 4)
    means = flat.mean(axis=0)
    var = flat.var(axis=0, ddof=1)
    levels = level4.reshape(n, 4).mean(axis=0)
    whole_z, _ = test_mean(n, means, var, levels)

    avg = d4[:, :, 3]
    lvl = level4[:, :, 3]
    thr = SIGN_TEST_THRESHOLD * lvl
    b_brighter = int((avg > thr).sum())
    a_brighter = int((avg < -thr).sum())
    differing = a_brighter + b_brighter
    sign_z = (b_brighter - a_brighter) / np.sqrt(differing) if differing else 0.0

    # tiles, row by row; partial at the right and bottom edges
    zs = []
    for y0 in range(0, h, tile):
        for x0 in range(0, w, tile):
            td = avg[y0:y0 + tile, x0:x0 + tile].ravel()
            if td.size < 2:
                continue
            rms = np.sqrt(np.mean(td * td))
            if rms == 0:
                continue  # identical in A and B: nothing to test
            tl = lvl[y0:y0 + tile, x0:x0 + tile].ravel()
            z, _ = test_mean(td.size, td.mean(), td.var(ddof=1), tl.mean())
            zs.append(float(z))
    zs = np.array(zs)
    tested = len(zs)
    checks = {}
    whole_limit = normal_quantile(FALSE_ALARM_RATE / 4)
    largest_whole = float(np.max(np.abs(whole_z)))
    checks["whole image"] = {"value": largest_whole, "limit": whole_limit, "failed": largest_whole > whole_limit}
    all_tiles = largest = 0.0
    over3 = 0
    if tested:
        all_tiles = float(np.clip(zs, -40, 40).sum() / np.sqrt(tested))
        largest = float(np.max(np.abs(zs)))
        over3 = int((np.abs(zs) > 3).sum())
        all_limit = normal_quantile(FALSE_ALARM_RATE)
        largest_limit = normal_quantile(FALSE_ALARM_RATE / tested)
        q = float(normal_tail(3.0))
        k = max(1, int(tested * q))
        while binomial_tail(tested, q, k) > FALSE_ALARM_RATE:
            k += 1
        checks["all tiles"] = {"value": abs(all_tiles), "limit": all_limit, "failed": abs(all_tiles) > all_limit}
        checks["largest tile"] = {"value": largest, "limit": largest_limit, "failed": largest > largest_limit}
        checks["tile count"] = {"value": float(over3), "limit": float(k), "failed": over3 >= k}
    consistent = not any(c["failed"] for c in checks.values())
    return Verdict(consistent, checks, whole_z, all_tiles, largest, over3, tested, float(sign_z),
                   float(a.mean()), float(b.mean()))


def rmse(a: np.ndarray, b: np.ndarray) -> float:
    """sqrt of the mean over pixels and channels of (a - b)^2"""
    return float(np.sqrt(np.mean((np.asarray(a, np.float64) - np.asarray(b, np.float64)) ** 2)))


def write_pfm(path: str, img: np.ndarray) -> None:
    """PF, the size, -1.0, rows from the bottom (src/image.ixx:157-187)"""
    h, w, _ = img.shape
    with open(path, "wb") as f:
        f.write(f"PF\n{w} {h}\n-1.0\n".encode())
        f.write(np.ascontiguousarray(img[::-1].astype("<f4")).tobytes())


def read_pfm(path: str) -> np.ndarray:
    with open(path, "rb") as f:
        magic = f.readline().strip()
        if magic != b"PF":
            raise ValueError("not a color PFM file")
        w, h = map(int, f.readline().split())
        scale = float(f.readline())
        data = np.frombuffer(f.read(), dtype="<f4" if scale < 0 else ">f4", count=w * h * 3)
    return data.reshape(h, w, 3)[::-1].astype(np.float64)
