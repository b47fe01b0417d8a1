// The megakernel's work queue (megakernel.hip): a launch's paths as chunks of a 16x4 tile over a group of passes,
// the items of a chunk, and where each item's sample is staged. Host and device code: the host checks the mapping
// (every pixel of the launch's rows and every pass exactly once, each at its own staging place).
#pragma once

#include <cstdint>

#include "../device_types.hpp"

namespace rt::dev
{
__host__ __device__ constexpr int imin(int a, int b) { return a < b ? a : b; }

// A launch's paths as a queue of chunks. Chunk c is one 16x4 tile of the launch's rows (tile c % tiles, row-major)
// over a group of the launch's passes (group c / tiles): pass-major, so every tile's first passes come before any
// tile's later ones, the costly tiles (long paths) are spread over the launch, and the launch ends on short chunks
// of every kind instead of on its last rows.
struct chunk_queue
{
    uint32_t* next_chunk; // the queue's head: zero when the launch starts
    uint32_t tiles_x;     // tiles in a row of tiles
    uint32_t tiles;       // tiles in the launch's rows
    uint32_t chunks;      // tiles x groups
    int32_t group;        // passes in a chunk (the last group's may be fewer)
    int32_t width;        // the image's width
    launch_work work;
};

// A chunk's tile and passes (the same in every lane); no items for a chunk past the queue's end
struct chunk_span
{
    int32_t x0, y0;  // the tile's first pixel, y0 counted from the launch's first row
    int32_t wx, wy;  // the tile's pixels in the image: 16x4 but at the image's right and the rows' bottom edge
    int32_t pass0;   // the chunk's first pass, counted from the launch's first
    uint32_t items;  // pixels x passes
};

__host__ __device__ inline chunk_span span_of(const chunk_queue& q, uint32_t c)
{
    chunk_span s{0, 0, 0, 0, 0, 0};
    if (c >= q.chunks)
    {
        return s;
    }
    const uint32_t g = c / q.tiles;
    const uint32_t t = c - g * q.tiles;
    const uint32_t ty = t / q.tiles_x;
    s.x0 = static_cast<int32_t>(16 * (t - ty * q.tiles_x));
    s.y0 = static_cast<int32_t>(4 * ty);
    s.wx = imin(16, q.width - s.x0);
    s.wy = imin(4, q.work.row_count - s.y0);
    s.pass0 = static_cast<int32_t>(g) * q.group;
    s.items = static_cast<uint32_t>(s.wx * s.wy * imin(q.group, q.work.pass_count - s.pass0));
    return s;
}

// Item k of a chunk, pass-major over the tile's pixels in row-major order: its pixel (i, j) and pass (counted from
// the launch's first)
__host__ __device__ inline void item_of(const chunk_span& s, uint32_t k, int first_row, int& i, int& j, int& pass_local)
{
    uint32_t p, at, row;
    if (s.wx == 16 && s.wy == 4)
    {
        p = k >> 6;
        at = k & 63;
        row = at >> 4;
    }
    else
    {
        const uint32_t n = static_cast<uint32_t>(s.wx * s.wy);
        p = k / n;
        at = k - p * n;
        row = at / static_cast<uint32_t>(s.wx);
    }
    i = s.x0 + static_cast<int>(at - row * static_cast<uint32_t>(s.wx));
    j = first_row + s.y0 + static_cast<int>(row);
    pass_local = s.pass0 + static_cast<int>(p);
}

// Where a sample is staged: pass-major over the launch's pixels (add_staged reads a pass's samples of consecutive
// pixels together)
__host__ __device__ inline int32_t staged_at(const chunk_queue& q, int i, int j, int pass_local)
{
    return (pass_local * q.work.row_count + (j - q.work.first_row)) * q.width + i;
}

}
