// A render that goes on (gpu::session, src/gpu/renderer.h:568-673): the scene stays in the GPU's memory, and the image
// of one view gains passes on demand. Host code; it reaches the GPU through the HIP runtime API and the launch
// functions of kernels/kernel_api.hpp.
#pragma once

#include <cstdint>
#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "../../include/raytracing_amd.h"
#include "scene_tables.hpp"

namespace rt
{
struct accumulation
{
    int64_t passes = 0;
    int32_t rows = 0;
};

class session
{
public:
    virtual ~session() = default;
    virtual void set_camera(const rt_camera_desc& camera) = 0;
    // on_pass (may be empty): called on the host after each launch that completes passes, with the passes done
    virtual rt_render_report render_to(int64_t passes, int32_t rows,
                                       const std::function<void(int64_t done)>& on_pass = {}) = 0;
    virtual rt_render_report render_for(double seconds, int64_t pass_limit) = 0;
    // render_to's whole passes on the megakernel with its staging and queue head in the library's static scratch:
    // allocates no device memory (the caller under graph capture or in a loop that can't allocate)
    virtual rt_render_report render_to_static(int64_t passes) = 0;
    virtual accumulation state() const = 0;
    // The last view set successfully, and whether the session has one (a set_camera whose image the GPU couldn't
    // hold leaves none until another succeeds)
    virtual rt_camera_desc camera() const = 0;
    virtual bool has_view() const = 0;
    virtual void pixels(double* out, std::size_t count) = 0;
    virtual void sums(double* out, std::size_t count) = 0;
    virtual void display(uint8_t* out, std::size_t count) = 0;
    // tone_map: an rt_tone_map; exposure in stops
    virtual void display_toned(int32_t tone_map, double exposure, uint8_t* out, std::size_t count) = 0;
    virtual void counters(uint64_t* rays, uint64_t* shadow_rays, double* kernel_seconds) const = 0;
    virtual rt_kernel_info kernel() const = 0;
    virtual void closest_hit(const float* origins, const float* directions, const float* times, const int32_t* from,
                             const uint8_t* inward, int64_t n, int32_t* index, float* t, uint8_t* entering) = 0;
    // A shadow ray's answer from both shadow walks (probes.hip): blocked by occluded, walked by closest_hit's any-hit
    // mode; either output may be null
    virtual void occluded(const float* origins, const float* directions, const float* times, const int32_t* from,
                          const uint8_t* inward, const float* t_max, const int32_t* skip, int64_t n, uint8_t* blocked,
                          uint8_t* walked) = 0;
    // The media's events and transmittance along given rays (probes.hip, media_kernel); float32
    virtual void media_events(const float* origins, const float* directions, const float* times,
                              const float* t_surface, int64_t n, uint64_t seed, int32_t* medium, float* t,
                              float* transmitted) = 0;
    // The light set's samples from given points (probes.hip, light_kernel); quality scenes in float32
    virtual void light_samples(const float* points, int64_t n, uint64_t seed, float* wi, float* distance,
                               float* radiance, float* density, int32_t* skip, float* hit_density) = 0;
    // The reference's own traversal of the session's binary tree, with its tie rule (probes.hip)
    virtual void reference_hits(const float* origins, const float* directions, const float* times, const int32_t* from,
                                const uint8_t* inward, int64_t n, int32_t* index, float* t, uint8_t* entering) = 0;
    virtual std::size_t device_bytes() const = 0;
    virtual rt_traffic traffic(int64_t passes) = 0;

    // Serving (the server steps many sessions at once). The session's launches go to `stream` (a stream the caller
    // owns and may share between sessions; null: the session's own), and its megakernel stages at most staging_passes
    // passes a launch (0: the session's default).
    virtual void serve_on(void* stream, int staging_passes) = 0;
    // Whole passes up to `passes` enqueued on the stream without waiting for them (megakernel launches, each within a
    // round of strata, with a persistent grid of a 1/grid_ways share of the chip); a step already in flight must have
    // been collected first (RT_ERROR_LOGIC otherwise)
    virtual void enqueue_to(int64_t passes, int grid_ways) = 0;
    // Whether the enqueued step has ended; once it has, the image holds its passes (state(), pixels())
    virtual bool poll() = 0;
    // The device memory the session holds beside its (shared) scene: its image, staging, stacks, counters
    virtual std::size_t view_device_bytes() const = 0;
};

// A scene resident on the device: its tables and its arrays (tree, records, textures) in one precision, shared by
// every session rendered from it (the server keeps scenes resident while requests of any resolution come and go);
// tree is an rt_tree, nodes an rt_nodes
class shared_scene;
std::shared_ptr<const shared_scene> upload_scene(const rt_scene& scene, int32_t precision,
                                                 int32_t tree = RT_TREE_DEFAULT, int32_t nodes = RT_NODES_AUTO,
                                                 int32_t kernel = RT_KERNEL_AUTO);
std::size_t device_bytes(const shared_scene& scene);
const rt_scene& scene_of(const shared_scene& scene);

// The device memory a megakernel session of this view of a resident scene holds beside the scene once it has
// rendered, with its staging sized for staging_passes passes a launch: an upper bound of view_device_bytes()
std::size_t view_device_bytes(const shared_scene& scene, const rt_camera_desc& camera, int staging_passes);

// The per-sample generator's outputs on device 0 (the numbers the kernels draw): n items of (pixel, pass) for one
// seed; any output may be null (probes.hip, sample_numbers_kernel)
// n linear colours through a tone map at an exposure, quantized with a dither each (or none), on device 0
void tone_probe(int32_t tone_map, double exposure, const float* colours, const float* dither, int64_t n,
                uint8_t* out);
// The quality path's scattering lobe `lobe` (rt_lobe) on n cases, on device 0 (probes.hip, scatter_kernel)
void scatter_probe(int32_t lobe, float alpha, float roughness, const float color[3], int64_t n, const float* v,
                   const float* l, uint64_t seed, float* value, float* density, float* drawn, float* weight,
                   float* drawn_density);
void sample_numbers(uint64_t seed, const uint64_t* pixels, const uint64_t* passes, int64_t n, int draws,
                    uint32_t* words, float* floats, double* doubles, float* directions);

// Validates the scene (before any GPU work), then builds its BVH and uploads it
std::unique_ptr<session> make_session(const rt_scene& scene, const rt_session_options& options);
// A session over a resident scene, with its own view (camera), image and stream
std::unique_ptr<session> make_session(std::shared_ptr<const shared_scene> scene, const rt_camera_desc& camera,
                                      const rt_session_options& options);
}
