"""GPU tests of the command-line programs (the reference's RayTracingGPU, src/main_gpu.cpp:104-294,362-640, and its
verification tools, docs/image-compare.md):

- `rt_gpu final -w 120 -s 64 -o x.pfm` passes `rt_image_compare` against the CPU oracle's render of the same size
  (the re-run rule: a second pair of seeds decides when the first differs);
- `rt_gpu furnace -o f.pfm` passes `rt_pfm_stats stats f.pfm --equal 1 --share 1`;
- `rt_gpu -o x.png` writes the books' display bytes: the PNG decodes to the session's own display() of the same
  render;
- the refusals, each with exit code 2 and a message, before anything is written: a quality scene in float64
  (tests/routing_tests.cu:53-67), an unknown scene, strategy, integrator or option, malformed or out-of-range
  numbers; --integrator plain gives the wavefront's image bytes.
"""
import os
import struct
import subprocess
import zlib

import numpy as np
import pytest

import oracle
from oracle import image_compare

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RT_GPU = os.path.join(ROOT, "raytracing_amd", "rt_gpu")
COMPARE = os.path.join(ROOT, "raytracing_amd", "rt_image_compare")
STATS = os.path.join(ROOT, "raytracing_amd", "rt_pfm_stats")


def run(*args):
    r = subprocess.run([str(a) for a in args], capture_output=True, text=True, timeout=120)
    print(" ".join(str(a) for a in args), "->", r.returncode, "\n", r.stdout[-600:], flush=True)
    return r.returncode, r.stdout


def test_final_pfm_against_cpu(tmp_path):
    results = []
    for gpu_seed, cpu_seed in ((1, 0), (3, 2)):
        out = tmp_path / f"gpu{gpu_seed}.pfm"
        code, text = run(RT_GPU, "final", "-w", 120, "-s", 64, "--seed", gpu_seed, "-o", out)
        assert code == 0 and "Saved" in text
        # the passes left, counted down (src/main_gpu.cpp:516-550; the text mode reads each "\r" as a line end)
        assert "Passes remaining: 64" in text
        cpu, _ = oracle.render("final", width=120, spp=64, seed=cpu_seed)
        ref = tmp_path / f"cpu{cpu_seed}.pfm"
        image_compare.write_pfm(str(ref), cpu)
        code, text = run(COMPARE, ref, out)
        assert code in (0, 1), text
        results.append(code)
        if code == 0:
            return
    raise AssertionError(f"rt_gpu's final differs from the CPU's beyond noise twice: {results}")


def test_furnace_pfm_exactly_one(tmp_path):
    out = tmp_path / "f.pfm"
    code, _ = run(RT_GPU, "furnace", "-w", 96, "-s", 16, "-o", out)
    assert code == 0
    code, text = run(STATS, "stats", out, "--equal", 1, "--share", 1)
    assert code == 0, text


def _read_png(path):
    data = open(path, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, idat, size = 8, b"", None
    while pos < len(data):
        n, kind = struct.unpack(">I4s", data[pos:pos + 8])
        body = data[pos + 8:pos + 8 + n]
        assert zlib.crc32(kind + body) == struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])[0], kind
        if kind == b"IHDR":
            size = struct.unpack(">II", body[:8])
            assert body[8:10] == b"\x08\x02"  # 8 bits, RGB
        elif kind == b"IDAT":
            idat += body
        pos += 12 + n
    w, h = size
    raw = zlib.decompress(idat)
    rows = np.frombuffer(raw, np.uint8).reshape(h, 1 + 3 * w)
    assert np.all(rows[:, 0] == 0)  # no filter
    return rows[:, 1:].reshape(h, w, 3)


def test_png_is_the_display_bytes(rt, tmp_path):
    out = tmp_path / "w.png"
    code, _ = run(RT_GPU, "weekend", "-w", 64, "-s", 4, "--seed", 3, "--integrator", "megakernel", "-o", out)
    assert code == 0
    png = _read_png(out)
    scene = rt.build_scene("weekend", width=64, samples=4)
    with rt.Session(scene, seed=3, integrator="megakernel") as s:
        s.render_to(scene.passes)
        display = s.display()[:, :, :3]
    assert png.shape == display.shape
    assert np.array_equal(png, display)


@pytest.mark.parametrize("args,message", [
    (["cornell", "-p", "mis", "--fp64"], "float32 only"),
    (["no-such-scene"], "unknown scene"),
    (["weekend", "-p", "bogus"], "unknown strategy"),
    (["weekend", "--integrator", "bogus"], "unknown integrator"),
    (["weekend", "-w", "abc"], "expected an integer"),
    (["weekend", "-s", "0"], "expected an integer"),
    (["weekend", "--exposure", "nan"], "expected a number"),
    (["weekend", "--wavefront"], "unknown option"),
    (["weekend", "--all-books-kernel", "--quality-kernel"], "--quality-kernel needs the quality kernel"),
    (["cornell", "-p", "mis", "--all-books-kernel"], "can't both render"),
    (["weekend", "--quality-kernel", "--fp64"], "float only"),
])
def test_refusals(tmp_path, args, message):
    out = tmp_path / "x.pfm"
    code, text = run(RT_GPU, *args, "-o", out)
    assert code == 2 and message in text, text
    assert not out.exists()


def test_all_books_kernel_same_image(tmp_path):
    """--all-books-kernel renders a first book's scene to the same image, bit for bit (src/main_gpu.cpp:78-80)"""
    a, b = tmp_path / "a.pfm", tmp_path / "b.pfm"
    code, text = run(RT_GPU, "weekend", "-w", 64, "-s", 16, "-p", "material", "-o", a)
    assert code == 0 and "Kernel: first book" in text
    code, text = run(RT_GPU, "weekend", "-w", 64, "-s", 16, "-p", "material", "--all-books-kernel", "-o", b)
    assert code == 0 and "Kernel: all books" in text
    code, text = run(STATS, "same", a, b)
    assert code == 0, text


def test_plain_integrator_same_image(tmp_path):
    a, b = tmp_path / "a.pfm", tmp_path / "b.pfm"
    assert run(RT_GPU, "cornell", "-w", 64, "-s", 16, "--integrator", "plain", "-o", a)[0] == 0
    assert run(RT_GPU, "cornell", "-w", 64, "-s", 16, "--integrator", "wavefront", "-o", b)[0] == 0
    code, text = run(STATS, "same", a, b)
    assert code == 0, text
