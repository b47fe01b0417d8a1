// The host's view of the kernels: launch functions with plain arguments, defined in the .hip files. The host runtime
// (session.cpp) calls HIP only through these and the HIP runtime API.
#pragma once

#include <cstdint>

#include <hip/hip_runtime.h>

#include "../device_types.hpp"

namespace rt
{
struct kernel_attributes
{
    int registers = 0;
    int local_bytes = 0;
    int lds_bytes = 0;
    int blocks_per_cu = 0;
};

template <typename Real>
size_t stack_lds_bytes(int depth, int threads);

// The megakernel's staging buffer holds one sample (dev::staged_sample<Real>) per pixel of a launch's rows per pass; a
// launch over a whole image of width x height stages at most this many samples per pass
size_t megakernel_staging_per_pass(int width, int height);

// waves: the occupancy the kernel's register budget targets (0: the compiler's choice, 5, 6 or 7 waves per SIMD; the
// FP64 and counting instances take 0 whatever is asked). staging: staging_samples samples; a launch of more passes
// than fit is split into several. queue: one device word, the chunk queue's head (cleared by the launch itself).
// block_threads: 64, 128, 192 or 256 (one to four waves a block; the waves work independently). far_buffer (may be
// null): far_bytes of device memory where the quality megakernel keeps its stacks' deeper levels when the whole
// stacks in LDS would cost occupancy (megakernel_far_stack_bytes(depth) bytes always suffice). grid_ways: the
// persistent grid takes a 1/grid_ways share of the waves the chip holds (several launches on different streams then
// run side by side, as the serving loop's do).
template <typename Real>
hipError_t launch_megakernel(int mode, const dev::camera_rec<Real>& cam, const dev::scene_view<Real>& sc,
                             uint64_t seed, const dev::launch_work& work, double* sums, void* staging,
                             size_t staging_samples, uint32_t* queue, unsigned long long* rays, int block_threads,
                             hipStream_t stream, unsigned long long* tallies = nullptr, int waves = 0,
                             void* far_buffer = nullptr, size_t far_bytes = 0, int grid_ways = 1);
size_t megakernel_far_stack_bytes(int depth);
// The static scratch of the no-allocation entry point on the current device (its address there, and its size in
// *bytes), or null if the runtime can't give its address
void* static_scratch(size_t* bytes);
bool megakernel_waves_supported(int waves);

// The plain megakernel (the reference's mapping, src/gpu/renderer.cu:340-397,421-451): one thread per pixel of a
// plain_block_x x plain_block_y block, the launch's passes in the thread's loop, added straight to the sums. The
// baseline the batched integrators are measured against; its sums equal theirs bit for bit.
constexpr int plain_block_x = 16;
constexpr int plain_block_y = 8;
template <typename Real>
hipError_t launch_plain_megakernel(int mode, const dev::camera_rec<Real>& cam, const dev::scene_view<Real>& sc,
                                   uint64_t seed, const dev::launch_work& work, double* sums, unsigned long long* rays,
                                   hipStream_t stream);
template <typename Real>
kernel_attributes plain_megakernel_attributes(int mode, int depth, int walk = dev::walk_all);

template <typename Real>
hipError_t launch_closest_hit(const dev::scene_view<Real>& sc, const Real* origins, const Real* directions,
                              const Real* times, const int32_t* from, const uint8_t* inward, int64_t n, int32_t* index,
                              Real* t, uint8_t* entering, hipStream_t stream);

// The reference's own closest-hit walk over the binary tree (probes.hip): binary_nodes holds node_count nodes of
// binary_node_bytes() each in the reference's layout (bvh_build.hpp ref_node<float>); overflow: one device word, counts
// rays whose walk outgrew the probe's stack
hipError_t launch_reference_walk(const dev::scene_view<float>& sc, const void* binary_nodes, int node_count,
                                 const float* origins, const float* directions, const float* times,
                                 const int32_t* from, const uint8_t* inward, int64_t n, int32_t* index, float* t,
                                 uint8_t* entering, int32_t* overflow, hipStream_t stream);
size_t binary_node_bytes();
// The media's events and transmittance along n rays (probes.hip, media_kernel)
hipError_t launch_media_probe(const dev::scene_view<float>& sc, const float* origins, const float* directions,
                              const float* times, const float* t_surface, int64_t n, uint64_t seed, int32_t* medium,
                              float* t, float* transmitted, hipStream_t stream);
// The quality path's light samples from n points (probes.hip, light_kernel)
hipError_t launch_light_probe(const dev::scene_view<float>& sc, const float* points, int64_t n, uint64_t seed,
                              float* wi, float* distance, float* radiance, float* density, int32_t* skip,
                              float* hit_density, hipStream_t stream);
// The scattering lobe `lobe` (rt_lobe) on n cases in the tangent frame (probes.hip, scatter_kernel)
hipError_t launch_scatter_probe(int lobe, float alpha, float roughness, const float color[3], int64_t n,
                                const float* v, const float* l, uint64_t seed, float* value, float* density,
                                float* drawn, float* weight, float* drawn_density, hipStream_t stream);
// Both shadow walks' answers for n rays (probes.hip): blocked from occluded, walked from closest_hit's any-hit mode
hipError_t launch_occluded(const dev::scene_view<float>& sc, const float* origins, const float* directions,
                           const float* times, const int32_t* from, const uint8_t* inward, const float* t_max,
                           const int32_t* skip, int64_t n, uint8_t* blocked, uint8_t* walked, hipStream_t stream);
// The per-sample generator's outputs for n (pixel, pass) items (probes.hip); any output may be null
hipError_t launch_sample_numbers(uint64_t seed, const uint64_t* pixels, const uint64_t* passes, int64_t n, int draws,
                                 uint32_t* words, float* floats, double* doubles, float* directions,
                                 hipStream_t stream);

// A tone map other than the books' (op: rt_tone_map) of the image's sums, width pixels a row
hipError_t launch_tone_map(const double* sums, int width, int64_t pixels, double inverse_passes, int op,
                           float exposure_scale, uint8_t* rgba, hipStream_t stream);
hipError_t launch_display(const double* sums, int64_t pixels, double inverse_passes, uint8_t* rgba,
                          hipStream_t stream);
// n linear colours through the tone map `op` and quantized with the given dither (tone.hip, tone_probe_kernel)
hipError_t launch_tone_probe(const float* colours, const float* dither, int64_t n, int op, float exposure_scale,
                             uint8_t* out, hipStream_t stream);

hipError_t launch_copy(const void* in, void* out, size_t bytes, hipStream_t stream);

// walk: the quality kernels' walk of the scene's tree (dev::tree_walk: a two-level tree's walks can enter instances, a
// packed tree's read packed nodes); marble: the scene has a noise texture (the three books' kernel for it)
template <typename Real>
kernel_attributes megakernel_attributes(int mode, int block_threads, int depth, int waves = 0, int walk = dev::walk_all,
                                        bool marble = false);
}
