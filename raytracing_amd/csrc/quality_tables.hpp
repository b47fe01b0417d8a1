// The quality integrator's host tables: the lamps (the world's glowing quads and still spheres), the environment map's
// densities and alias table, and the materials' GGX side entries, built in double and rounded once for the kernels
// (src/gpu/lamp_table.cu:60-134, src/gpu/environment_table.cu:81-225, src/gpu/microfacets_on_device.cuh).
#pragma once

#include <cstdint>
#include <span>
#include <vector>

#include "device_types.hpp"
#include "scene_tables.hpp"

namespace rt
{
struct lamp_set
{
    std::vector<dev::emitter_rec> lamps;
    std::vector<int32_t> lamp_of; // per surface primitive (spheres, moving spheres, quads, triangles): its lamp or -1
};

// Every quad, and every sphere that stands still, whose material is a diffuse light of one solid color, spheres first,
// each in the scene's order. sphere_ids and quad_ids are the BVH's global ids of the scene's spheres and quads;
// surface_count is how many surface primitives the BVH holds. std::invalid_argument for ids that don't fit.
lamp_set build_lamp_set(const rt_scene& scene, std::span<const int32_t> sphere_ids, std::span<const int32_t> quad_ids,
                        std::size_t surface_count);

// The lamps a scene has, without the BVH (count_lamps, src/gpu/renderer.h:400)
int count_lamps(const rt_scene& scene);

struct environment_tables
{
    int width = 0, height = 0, log2_bins = 0;
    double total = 0;                  // S: the map's luminance over the sphere, sum of Y times each cell's solid angle
    std::vector<dev::env_texel> texels; // radiance, and the density Y/S per steradian of the cell's directions
    std::vector<dev::alias_bin> bins;  // 2^log2_bins bins (the cells, then weightless padding); empty for a black map
    std::vector<float> row_cos;        // height + 1 row borders, cos(pi j / height), 1 first and -1 last
};

// The tables of an equirectangular map (+y up, row 0 at +y): cell (i, j) spans phi in [2 pi i/W - pi, 2 pi (i+1)/W - pi)
// and theta in [pi j/H, pi (j+1)/H), so a direction drawn uniformly in (phi, cos theta) inside the cell picked with
// chance Y*Omega/S has the density Y/S per steradian. std::invalid_argument for a map out of range.
environment_tables build_environment_tables(const rt_environment_desc& env);

// The alias table of weights (any scale, each >= 0, their sum > 0) padded with weightless bins to `bins` (a power of
// two >= weights.size()): every coin a multiple of 2^-24, the grid of random_real<float>, so the coin's chance is
// exactly what is stored
std::vector<dev::alias_bin> build_alias_table(const std::vector<double>& weights, std::size_t bins);

// One entry per material (kind, alpha = r^2 rounded once, r), or none when the scene gives none
std::vector<dev::microfacet_rec> build_microfacets(const rt_scene& scene);

// Whether a scene needs the quality integrator: light samples (nee or mis), a GGX side entry, an environment map or
// triangles
bool needs_quality(const rt_scene& scene);
}
