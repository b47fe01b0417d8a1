// The wavefront integrator's pool of path slots and its launch functions (kernels in wavefront.hip, driver in
// wavefront.cpp)
#pragma once

#include <cstdint>

#include <hip/hip_runtime.h>

#include "../device_types.hpp"
#include "kernel_api.hpp"

namespace rt::dev
{
// The shading classes, in the order the shade kernel takes their queues
enum wf_class : int
{
    wf_class_miss = 0,
    wf_class_light,
    wf_class_matte,
    wf_class_textured,
    wf_class_marble,
    wf_class_smoke,
    wf_class_metal,
    wf_class_glass,
    wf_class_ggx,   // quality integrator: a GGX conductor
    wf_class_gloss, // quality integrator: a glossy coat
    wf_classes
};

// Every list (the active slots, each class's queue) is cut into wf_stripes stripes: block b of a kernel appends to
// stripe b % wf_stripes, with one device-scope atomic per block and list, so that no counter is the target of more
// than a few hundred atomics per kernel (the eight XCDs' L2s don't share a line: a device-scope atomic is served
// beyond them, and atomics on one address queue there). A stripe's items lie in its own segment of the list.
constexpr int wf_stripes = 32;
constexpr int wf_block = 256; // threads per block of every pool kernel

// The counters of one parity: the active list's stripes, then each class's stripes (class-major)
constexpr int wf_counter_active = 0;
constexpr int wf_counter_class = wf_stripes;
constexpr int wf_counter_stride = wf_stripes * (1 + wf_classes);
// The prefix sums a bounce reads: the active stripes' (wf_stripes + 1), then the class stripes'
constexpr int wf_prefix_active = 0;
constexpr int wf_prefix_class = wf_stripes + 1;
constexpr int wf_prefix_size = wf_prefix_class + wf_classes * wf_stripes + 1;

// The slots' state, structure of arrays, 16-byte records
struct wf_pool
{
    float4* ray_o;  // origin, time
    float4* ray_d;  // direction, the primitive the ray leaves (int bits)
    float4* thr;    // throughput, rays left | inward << 30 | specular << 31 (int bits)
    float4* qst;    // quality integrator only (else nullptr): the light gathered so far, the last bsdf density
    uint4* rng;     // pcg32 state and increment
    int4* hit;      // the closest event: primitive id, t (bits), entering
    int32_t* pass;  // the pass the slot's path belongs to
    uint32_t* active[2]; // wf_stripes segments of `segment` slots
    uint32_t* queue;     // wf_classes x wf_stripes segments of `segment` slots
    uint32_t* counters;  // 2 parities x wf_counter_stride
    uint32_t* prefix;    // wf_prefix_size
    unsigned long long* tallies; // wf_stripes x 2: rays and shadow rays, added to the session's counters per launch
    uint32_t capacity;
    uint32_t segment;    // slots per stripe segment: ceil(blocks / wf_stripes) * wf_block
};
}

namespace rt
{
// Blocks of a pool kernel over `capacity` slots
inline uint32_t wf_blocks(uint32_t capacity) { return (capacity + dev::wf_block - 1) / dev::wf_block; }

hipError_t wf_launch_init(const dev::camera_rec<float>& cam, const dev::scene_view<float>& sc, uint64_t seed,
                          const dev::launch_work& work, const dev::wf_pool& pool, hipStream_t stream);
// One bounce of parity `parity`: the active list's prefix sums, extend, the queues' prefix sums, shade
hipError_t wf_launch_bounce(const dev::camera_rec<float>& cam, const dev::scene_view<float>& sc, uint64_t seed,
                            const dev::launch_work& work, const dev::wf_pool& pool, int parity, double* sums,
                            bool quality, hipStream_t stream);
// The drain: the slots of the parity-0 active list finish their paths and their pixels' remaining passes, a thread
// each (mode: the scene's kernel_mode; count: at least the list's length)
hipError_t wf_launch_drain(int mode, const dev::camera_rec<float>& cam, const dev::scene_view<float>& sc, uint64_t seed,
                           const dev::launch_work& work, const dev::wf_pool& pool, uint32_t count, double* sums,
                           hipStream_t stream);
// Adds the launch's ray counts to the session's counters (rays: [0] rays, [1] shadow rays) and clears them
hipError_t wf_launch_finish(const dev::wf_pool& pool, unsigned long long* rays, hipStream_t stream);
// which: 0 extend, 1 shade, 2 the quality integrator's shade
kernel_attributes wf_attributes(int depth, int which);
}
