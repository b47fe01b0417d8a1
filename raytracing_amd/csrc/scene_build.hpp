// The scene-building API: named scenes built and flattened into scene_tables (see scene_build.cpp)
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "scene_tables.hpp"

namespace rt
{
struct scene_options
{
    int earth_width = 2048;
    std::uint32_t earth_seed = 0;
};

struct image_pixels
{
    int width = 0, height = 0;
    std::vector<float> rgb;
};

struct built_scene
{
    scene_tables tables;
    std::int64_t passes = 0;           // sqrt_spp^2, or the samples asked for with strata
    std::vector<std::string> warnings; // a model's: what it holds that isn't rendered, one line each
};

// The scene called `name` with the overrides applied, flattened (build_gpu_scene, src/gpu_scene.ixx:1333-1445). A name
// that ends in .gltf or .glb is a model's file, rendered in a studio (gltf.hpp).
// Throws std::invalid_argument for an unknown name or options out of range, std::runtime_error for a scene that can't
// be flattened.
built_scene build_scene(const std::string& name, const rt_scene_overrides& overrides);

// The seeded RGB8 noise image that stands in for the reference's assets/earthmap.jpg, as linear floats
image_pixels noise_image(int width, int height, std::uint32_t seed);

// Whether quads bound one convex shape (gpu::is_convex, src/gpu/renderer.h:163-168)
bool quads_are_convex(const rt_quad_desc* quads, std::size_t count);

const char* scene_names();
}
