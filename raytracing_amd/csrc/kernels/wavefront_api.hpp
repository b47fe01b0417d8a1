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

// The counters of one parity: [0] the active list's length, [1 + c] class c's queue length, then the two kernels'
// work counters
constexpr int wf_counter_class = 1;
constexpr int wf_counter_extend_work = wf_counter_class + wf_classes;
constexpr int wf_counter_shade_work = wf_counter_extend_work + 1;
constexpr int wf_counter_stride = 16;

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
    uint32_t* active[2];
    uint32_t* queue; // wf_classes segments of `capacity` slots
    uint32_t* counters; // 2 parities x wf_counter_stride
    uint32_t capacity;
};
}

namespace rt
{
hipError_t wf_launch_init(const dev::camera_rec<float>& cam, const dev::scene_view<float>& sc, uint64_t seed,
                          const dev::launch_work& work, const dev::wf_pool& pool, hipStream_t stream);
// rays: [0] the rays traced, [1] the shadow rays (quality)
hipError_t wf_launch_bounce(const dev::camera_rec<float>& cam, const dev::scene_view<float>& sc, uint64_t seed,
                            const dev::launch_work& work, const dev::wf_pool& pool, int parity, double* sums,
                            unsigned long long* rays, int extend_blocks, int shade_blocks, bool quality,
                            hipStream_t stream);
// The drain: the `count` slots of the parity-0 active list finish their paths and their pixels' remaining passes, a
// thread each (mode: the scene's kernel_mode)
hipError_t wf_launch_drain(int mode, const dev::camera_rec<float>& cam, const dev::scene_view<float>& sc, uint64_t seed,
                           const dev::launch_work& work, const dev::wf_pool& pool, uint32_t count, double* sums,
                           unsigned long long* rays, hipStream_t stream);
// which: 0 extend, 1 shade, 2 the quality integrator's shade
kernel_attributes wf_attributes(int depth, int which);
}
