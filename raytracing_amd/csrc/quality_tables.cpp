// The quality integrator's host tables (see quality_tables.hpp)
#include "quality_tables.hpp"

#include <algorithm>
#include <bit>
#include <cmath>
#include <numbers>
#include <stdexcept>

#include "bvh_build.hpp"

namespace rt
{
namespace
{
bool still(const rt_sphere_desc& s) { return s.displacement.x == 0 && s.displacement.y == 0 && s.displacement.z == 0; }

// A diffuse light of one solid color (src/gpu/lamp_table.cu:24-39)
bool glows_evenly(const rt_scene& scene, int32_t material)
{
    if (material < 0 || material >= scene.material_count)
    {
        throw std::invalid_argument("a primitive refers to a material that doesn't exist");
    }
    const rt_material_desc& m = scene.materials[material];
    if (m.kind != RT_MATERIAL_DIFFUSE_LIGHT)
    {
        return false;
    }
    if (m.texture < 0 || m.texture >= scene.texture_count)
    {
        throw std::invalid_argument("a lamp refers to a texture that doesn't exist");
    }
    return scene.textures[m.texture].kind == RT_TEXTURE_SOLID;
}

void put(float* out, const rt_vec3& v)
{
    out[0] = static_cast<float>(v.x);
    out[1] = static_cast<float>(v.y);
    out[2] = static_cast<float>(v.z);
}

// Rec. 709 luminance: the weight of a cell of the map
double luminance(const float* rgb) { return 0.2126 * rgb[0] + 0.7152 * rgb[1] + 0.0722 * rgb[2]; }

constexpr double coin_grid = 16777216.0; // 2^24: random_real<float> is k / 2^24
}

lamp_set build_lamp_set(const rt_scene& scene, std::span<const int32_t> sphere_ids, std::span<const int32_t> quad_ids,
                        std::size_t surface_count)
{
    if (sphere_ids.size() != static_cast<std::size_t>(scene.sphere_count) ||
        quad_ids.size() != static_cast<std::size_t>(scene.quad_count))
    {
        throw std::invalid_argument("the lamps need one global id per sphere and per quad");
    }
    lamp_set set;
    set.lamp_of.assign(surface_count, -1);
    const auto add = [&set](const dev::emitter_rec& e) {
        if (e.primitive < 0 || static_cast<std::size_t>(e.primitive) >= set.lamp_of.size())
        {
            throw std::invalid_argument("a lamp's global id lies outside the scene's primitives");
        }
        set.lamp_of[static_cast<std::size_t>(e.primitive)] = static_cast<int32_t>(set.lamps.size());
        set.lamps.push_back(e);
    };
    for (int64_t k = 0; k < scene.sphere_count; ++k)
    {
        const rt_sphere_desc& s = scene.spheres[k];
        if (!glows_evenly(scene, s.material) || !still(s))
        {
            continue;
        }
        dev::emitter_rec e{};
        put(e.corner, s.center);
        e.size = static_cast<float>(s.radius);
        e.kind = RT_LIGHT_SPHERE;
        e.primitive = sphere_ids[static_cast<std::size_t>(k)];
        put(e.radiance, scene.textures[scene.materials[s.material].texture].rgb);
        add(e);
    }
    for (int64_t k = 0; k < scene.quad_count; ++k)
    {
        const rt_quad_desc& q = scene.quads[k];
        if (!glows_evenly(scene, q.material))
        {
            continue;
        }
        const double nx = q.u.y * q.v.z - q.u.z * q.v.y;
        const double ny = q.u.z * q.v.x - q.u.x * q.v.z;
        const double nz = q.u.x * q.v.y - q.u.y * q.v.x;
        dev::emitter_rec e{};
        put(e.corner, q.q);
        e.size = static_cast<float>(std::sqrt(nx * nx + ny * ny + nz * nz));
        put(e.u, q.u);
        e.kind = RT_LIGHT_QUAD;
        put(e.v, q.v);
        e.primitive = quad_ids[static_cast<std::size_t>(k)];
        const dev::quad_rec<float> rec = make_quad_rec<float>(q); // the same unit normal the hit test uses
        std::copy(rec.normal, rec.normal + 3, e.normal);
        put(e.radiance, scene.textures[scene.materials[q.material].texture].rgb);
        add(e);
    }
    return set;
}

int count_lamps(const rt_scene& scene)
{
    int n = 0;
    for (int64_t k = 0; k < scene.sphere_count; ++k)
    {
        n += glows_evenly(scene, scene.spheres[k].material) && still(scene.spheres[k]) ? 1 : 0;
    }
    for (int64_t k = 0; k < scene.quad_count; ++k)
    {
        n += glows_evenly(scene, scene.quads[k].material) ? 1 : 0;
    }
    return n;
}

std::vector<dev::alias_bin> build_alias_table(const std::vector<double>& weights, std::size_t bins)
{
    if (bins < weights.size() || !std::has_single_bit(bins))
    {
        throw std::invalid_argument("an alias table's bins must be a power of two, at least the weights");
    }
    double sum = 0;
    for (const double w : weights)
    {
        if (!(w >= 0 && std::isfinite(w)))
        {
            throw std::invalid_argument("alias table weights must be finite numbers >= 0");
        }
        sum += w;
    }
    if (!(sum > 0))
    {
        throw std::invalid_argument("an alias table needs some weight");
    }
    // each bin's share of one bin's worth: they add up to the bin count
    std::vector<double> share(bins, 0.0);
    for (std::size_t c = 0; c < weights.size(); ++c)
    {
        share[c] = weights[c] * static_cast<double>(bins) / sum;
    }
    std::vector<dev::alias_bin> table(bins);
    // Walker's method in Vose's linear form: two work lists in index order, read from their heads. A bin below one
    // bin's worth keeps its share (its coin) and takes the rest from a bin above, which may then drop below and join
    // the first list.
    std::vector<uint32_t> under, over;
    for (uint32_t c = 0; c < bins; ++c)
    {
        (share[c] < 1 ? under : over).push_back(c);
    }
    std::size_t u = 0, o = 0;
    double drift = 0; // the coins' rounding so far, in grid steps: fed into the next coin so it stays within half
    while (u < under.size() && o < over.size())
    {
        const uint32_t small = under[u++];
        const uint32_t big = over[o];
        double q = 0;
        if (share[small] > 0)
        {
            // the grid point below or above the exact coin, whichever leaves the smaller drift; a cell with light
            // keeps at least one step, so it can always be drawn
            const double exact = share[small] * coin_grid;
            const double below = std::floor(exact);
            double k = std::fabs(drift + (below - exact)) <= std::fabs(drift + (below + 1 - exact)) ? below : below + 1;
            k = std::max(k, 1.0);
            drift += k - exact;
            q = k / coin_grid;
        }
        table[small] = {static_cast<float>(q), static_cast<int32_t>(big)};
        share[big] -= 1 - q; // exact: q is on the 2^-24 grid
        if (share[big] < 1)
        {
            ++o;
            under.push_back(big);
        }
    }
    // what is left holds one bin's worth up to rounding: it keeps itself
    for (; u < under.size(); ++u)
    {
        table[under[u]] = {1.0f, static_cast<int32_t>(under[u])};
    }
    for (; o < over.size(); ++o)
    {
        table[over[o]] = {1.0f, static_cast<int32_t>(over[o])};
    }
    return table;
}

environment_tables build_environment_tables(const rt_environment_desc& env)
{
    const rt_image_desc& map = env.image;
    if (map.width < 1 || map.height < 1 || map.width > 2048 || map.height > 1024 || map.rgb == nullptr)
    {
        throw std::invalid_argument("an environment map must have from 1x1 to 2048x1024 texels");
    }
    const int w = map.width, h = map.height;
    const std::size_t cells = static_cast<std::size_t>(w) * static_cast<std::size_t>(h);
    for (std::size_t k = 0; k < 3 * cells; ++k)
    {
        if (!(std::isfinite(map.rgb[k]) && map.rgb[k] >= 0))
        {
            throw std::invalid_argument("an environment map's radiance must be a finite number >= 0 in every texel");
        }
    }
    environment_tables t;
    t.width = w;
    t.height = h;
    std::vector<double> border(static_cast<std::size_t>(h) + 1);
    for (int j = 0; j <= h; ++j)
    {
        border[static_cast<std::size_t>(j)] = j == 0 ? 1.0 : (j == h ? -1.0 : std::cos(std::numbers::pi * j / h));
    }
    // a row's cells share one solid angle: (2 pi / W)(cos theta_j - cos theta_j+1); the rows add up to 4 pi exactly
    std::vector<double> omega(static_cast<std::size_t>(h));
    for (int j = 0; j < h; ++j)
    {
        omega[static_cast<std::size_t>(j)] =
            2 * std::numbers::pi / w * (border[static_cast<std::size_t>(j)] - border[static_cast<std::size_t>(j) + 1]);
    }
    for (int j = 0; j < h; ++j)
    {
        double row = 0;
        for (int i = 0; i < w; ++i)
        {
            row += luminance(map.rgb + 3 * (static_cast<std::size_t>(j) * w + i));
        }
        t.total += row * omega[static_cast<std::size_t>(j)];
    }
    t.row_cos.assign(border.begin(), border.end());
    t.texels.resize(cells);
    for (std::size_t k = 0; k < cells; ++k)
    {
        const float* rgb = map.rgb + 3 * k;
        const float pdf = t.total > 0 ? static_cast<float>(luminance(rgb) / t.total) : 0.0f;
        t.texels[k] = {rgb[0], rgb[1], rgb[2], pdf};
    }
    if (t.total > 0)
    {
        // the bin is the top log2_bins bits of one random number: at least one bit, so the shift stays below 32
        t.log2_bins = std::max(1, static_cast<int>(std::bit_width(cells - 1)));
        std::vector<double> weights(cells);
        for (std::size_t k = 0; k < cells; ++k)
        {
            weights[k] = luminance(map.rgb + 3 * k) * omega[k / static_cast<std::size_t>(w)];
        }
        t.bins = build_alias_table(weights, std::size_t{1} << t.log2_bins);
    }
    return t;
}

std::vector<dev::microfacet_rec> build_microfacets(const rt_scene& scene)
{
    std::vector<dev::microfacet_rec> out;
    for (int64_t k = 0; k < scene.microfacet_count; ++k)
    {
        const rt_microfacet_desc& d = scene.microfacets[k];
        out.push_back({d.kind, static_cast<float>(d.roughness * d.roughness), static_cast<float>(d.roughness), 0});
    }
    return out;
}

bool needs_quality(const rt_scene& scene)
{
    bool ggx = false;
    for (int64_t k = 0; k < scene.microfacet_count; ++k)
    {
        ggx = ggx || scene.microfacets[k].kind != RT_MICROFACET_NONE;
    }
    return ggx || scene.combination != RT_COMBINE_BSDF_ONLY || scene.environment != nullptr || scene.mesh_count > 0;
}
}
