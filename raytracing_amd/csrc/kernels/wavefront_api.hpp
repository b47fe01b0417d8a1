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

// Every list (the active slots, each class's queue, the free slots) is cut into wf_stripes stripes: block b of a
// kernel appends to stripe b % wf_stripes, with one device-scope atomic per block and list, so that no counter is the
// target of more than a few hundred atomics per kernel (the eight XCDs' L2s don't share a line: a device-scope atomic
// is served beyond them, and atomics on one address queue there). A stripe's items lie in its own segment of the
// list. The active list has one more segment, after the stripes: the slots refilled with new work at the bounce's
// start.
constexpr int wf_stripes = 32;
constexpr int wf_block = 256; // threads per block of every pool kernel

// The counters: per parity, the active list's stripes, then each class's stripes (class-major); after both
// parities, the free list's stripes and the launch's work state
constexpr int wf_counter_active = 0;
constexpr int wf_counter_class = wf_stripes;
constexpr int wf_counter_stride = wf_stripes * (1 + wf_classes);
constexpr int wf_counter_free = 2 * wf_counter_stride;
constexpr int wf_counter_next = wf_counter_free + wf_stripes; // the next work item to hand out
constexpr int wf_counter_items = wf_counter_next + 1;         // the launch's work items
constexpr int wf_counter_refill = wf_counter_next + 2;        // slots refilled at this bounce's start
constexpr int wf_counter_refill_base = wf_counter_next + 3;   // their first work item
constexpr int wf_counters = wf_counter_next + 4;
// The prefix sums a bounce reads: the active list's (stripes and the refill segment: wf_stripes + 2), the queues'
// (class-major), the free list's
constexpr int wf_prefix_active = 0;
constexpr int wf_prefix_class = wf_stripes + 2;
constexpr int wf_prefix_free = wf_prefix_class + wf_classes * wf_stripes + 1;
constexpr int wf_prefix_size = wf_prefix_free + wf_stripes + 1;

// What changes from one launch of the pool to the next, in device memory: the bounce kernels read it there, so that
// the captured bounce graph stays valid for every launch of the pool and is captured once
struct wf_launch
{
    uint64_t seed;
    launch_work work;
};

// The slots' state, structure of arrays, 16-byte records. A slot renders work items: item k of a launch is pass
// first_pass + k / pixels of the launch's pixel k % pixels (pixels in row order from its first row); its sample goes
// to samples[k], and the launch ends by adding each pixel's samples to its sums in pass order.
struct wf_pool
{
    const wf_launch* launch; // the current launch's seed and work
    float4* ray_o;   // origin, time
    float4* ray_d;   // direction, the primitive the ray leaves (int bits)
    float4* thr;     // throughput, rays left | inward << 30 | specular << 31 (int bits)
    float4* qst;     // quality integrator only (else nullptr): the light gathered so far, the last bsdf density
    uint4* rng;      // pcg32 state and increment
    int4* hit;       // the closest event: primitive id, t (bits), entering
    uint32_t* item;  // the work item the slot's path renders
    float4* samples; // per work item of the launch: its path's sample
    uint32_t* active[2]; // wf_stripes segments of `segment` slots, then the refill segment (`capacity` slots)
    uint32_t* free;      // wf_stripes segments of `segment` slots
    uint32_t* queue;     // wf_classes x wf_stripes segments of `segment` slots
    uint32_t* counters;  // wf_counters
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

// A launch's start: `items` work items (passes x pixels of work), every slot free, no slot active
hipError_t wf_launch_init(const dev::wf_pool& pool, uint32_t items, hipStream_t stream);
// One bounce of parity `parity`: plan (refills, the active list's prefix sums), refill, extend, the queues' prefix
// sums, shade (mode: the scene's kernel_mode, whose shading the megakernel's equals). The launch's seed and work are
// read from pool.launch.
hipError_t wf_launch_bounce(const dev::camera_rec<float>& cam, const dev::scene_view<float>& sc,
                            const dev::wf_pool& pool, int parity, int mode, hipStream_t stream);
// The drain: the slots of the parity-0 active list finish their paths, a thread each (mode: the scene's kernel_mode;
// count: at least the list's length). Only when every work item has been handed out.
hipError_t wf_launch_drain(int mode, const dev::camera_rec<float>& cam, const dev::scene_view<float>& sc, uint64_t seed,
                           const dev::launch_work& work, const dev::wf_pool& pool, uint32_t count, hipStream_t stream);
// The launch's end: each pixel's samples added to its sums in pass order, and the ray counts to the session's
// counters (rays: [0] rays, [1] shadow rays)
hipError_t wf_launch_finish(const dev::camera_rec<float>& cam, const dev::launch_work& work, const dev::wf_pool& pool,
                            double* sums, unsigned long long* rays, hipStream_t stream);
// which: 0 extend, 1 shade, 2 the quality integrator's shade
kernel_attributes wf_attributes(int depth, int which);
}
