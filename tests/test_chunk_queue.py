"""The megakernel's work queue on the host (raytracing_amd/csrc/kernels/chunk_queue.hpp): for image widths, row bands
and pass counts on and off the 16x4 tile and the pass group, every pixel of the launch's rows and every pass is one
item of exactly one chunk, and each item has its own staging place inside the launch's staging area. (The GPU tests
check the same through the images: the megakernel's sums equal the wavefront's bit for bit.)"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROGRAM = r"""
#include <cstdio>
#include <vector>
#include "raytracing_amd/csrc/kernels/chunk_queue.hpp"
using namespace rt::dev;
int main()
{
    const int shapes[][6] = {
        // width, first_row, row_count, first_pass, pass_count, group
        {16, 0, 4, 0, 1, 2}, {37, 0, 37, 0, 7, 2}, {1080, 0, 1080, 0, 57, 2}, {100, 10, 13, 5, 3, 4},
        {1, 0, 1, 0, 1, 1}, {17, 3, 5, 0, 9, 8}, {64, 0, 3, 2, 2, 2}, {33, 7, 1, 0, 5, 3},
    };
    int bad = 0;
    for (const auto& sh : shapes)
    {
        const int width = sh[0];
        const launch_work work{sh[3], sh[4], sh[1], sh[2]};
        const int group = sh[5];
        const uint32_t tiles_x = (width + 15) / 16;
        const uint32_t tiles = tiles_x * ((work.row_count + 3) / 4);
        const uint32_t groups = (work.pass_count + group - 1) / group;
        const chunk_queue q{nullptr, tiles_x, tiles, tiles * groups, group, width, work};
        const size_t places = static_cast<size_t>(width) * work.row_count * work.pass_count;
        std::vector<int> seen(places, 0);
        for (uint32_t c = 0; c <= q.chunks; ++c)
        {
            const chunk_span s = span_of(q, c);
            if (c == q.chunks && s.items != 0)
            {
                ++bad;
            }
            for (uint32_t k = 0; k < s.items; ++k)
            {
                int i, j, pass;
                item_of(s, k, work.first_row, i, j, pass);
                if (i < 0 || i >= width || j < work.first_row || j >= work.first_row + work.row_count || pass < 0 ||
                    pass >= work.pass_count)
                {
                    ++bad;
                    continue;
                }
                const int32_t at = staged_at(q, i, j, pass);
                const int32_t expect = (pass * work.row_count + (j - work.first_row)) * width + i;
                if (at != expect || at < 0 || static_cast<size_t>(at) >= places)
                {
                    ++bad;
                    continue;
                }
                ++seen[at];
            }
        }
        for (const int n : seen)
        {
            bad += n != 1;
        }
    }
    std::printf("%d\n", bad);
    return bad != 0;
}
"""


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host C++ compiler")
def test_every_item_once(tmp_path):
    src = tmp_path / "queue.cpp"
    src.write_text(PROGRAM)
    exe = tmp_path / "queue"
    subprocess.run(["g++", "-std=c++20", "-O1", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", f"-I{ROOT}",
                    str(src), "-o", str(exe)], check=True, capture_output=True, timeout=300)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and r.stdout.strip() == "0", r.stdout + r.stderr
