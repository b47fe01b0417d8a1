// The scene in the GPU's memory (see device_scene.hpp)
#include "device_scene.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <numbers>

#include "quality_tables.hpp"
#include "scene_build.hpp"

namespace rt
{
namespace
{
bool has_texture(const rt_material_desc& m)
{
    return m.kind == RT_MATERIAL_LAMBERTIAN || m.kind == RT_MATERIAL_DIFFUSE_LIGHT || m.kind == RT_MATERIAL_ISOTROPIC;
}
bool solid_texture(const rt_scene& s, const rt_material_desc& m)
{
    return has_texture(m) && s.textures[m.texture].kind == RT_TEXTURE_SOLID;
}
bool texture_needs_uv(const rt_scene& s, int index)
{
    const rt_texture_desc& t = s.textures[index];
    if (t.kind == RT_TEXTURE_IMAGE)
    {
        return true;
    }
    if (t.kind == RT_TEXTURE_CHECKER)
    {
        return texture_needs_uv(s, t.even) || texture_needs_uv(s, t.odd);
    }
    return false;
}
bool moves(const rt_sphere_desc& s) { return s.displacement.x != 0 || s.displacement.y != 0 || s.displacement.z != 0; }
bool same(const rt_vec3& a, const rt_vec3& b) { return a.x == b.x && a.y == b.y && a.z == b.z; }
template <typename Real>
void store(Real* out, const rt_vec3& v)
{
    out[0] = static_cast<Real>(v.x);
    out[1] = static_cast<Real>(v.y);
    out[2] = static_cast<Real>(v.z);
}
template <typename T>
std::size_t bytes_of(const device_array<T>& a)
{
    return a.bytes();
}
}

void validate_camera(const rt_camera_desc& c)
{
    if (c.image_width < 1 || c.image_height < 1)
    {
        throw std::runtime_error("the image must be at least one pixel wide and high");
    }
    if (c.image_height > 65535 * 8)
    {
        throw std::runtime_error("the image is too tall for the GPU renderer");
    }
    if (c.sqrt_spp < 1 || c.sqrt_spp > 46340)
    {
        throw std::runtime_error("the side of a pixel's grid of samples must be from 1 to 46340");
    }
    if (c.max_depth < 1)
    {
        throw std::runtime_error("a path must be allowed at least one ray");
    }
    if (c.sampling < RT_SAMPLING_UNIFORM || c.sampling > RT_SAMPLING_MIXTURE)
    {
        throw std::runtime_error("unknown sampling strategy");
    }
}

void validate_scene(const rt_scene& s)
{
    validate_camera(s.camera);
    const auto nonempty = [](const void* p, int64_t n) { return n == 0 || p != nullptr; };
    if (!nonempty(s.spheres, s.sphere_count) || !nonempty(s.materials, s.material_count) ||
        !nonempty(s.textures, s.texture_count) || !nonempty(s.images, s.image_count) ||
        !nonempty(s.perlins, s.perlin_count) || !nonempty(s.quads, s.quad_count) ||
        !nonempty(s.media, s.medium_count) || !nonempty(s.lights, s.light_count) ||
        !nonempty(s.microfacets, s.microfacet_count) || !nonempty(s.meshes, s.mesh_count) ||
        !nonempty(s.instances, s.instance_count) || s.sphere_count < 0 || s.material_count < 0 ||
        s.texture_count < 0 || s.image_count < 0 || s.perlin_count < 0 || s.quad_count < 0 || s.medium_count < 0 ||
        s.light_count < 0)
    {
        throw std::runtime_error("a table of the scene has a count but no data");
    }
    for (int64_t p = 0; p < s.perlin_count; ++p)
    {
        for (const int32_t* perm : {s.perlins[p].perm_x, s.perlins[p].perm_y, s.perlins[p].perm_z})
        {
            for (int i = 0; i < 256; ++i)
            {
                if (perm[i] < 0 || perm[i] > 255)
                {
                    throw std::runtime_error("a Perlin permutation has an entry outside 0..255");
                }
            }
        }
    }
    std::size_t texels = 0;
    for (int64_t i = 0; i < s.image_count; ++i)
    {
        const rt_image_desc& im = s.images[i];
        if (im.width < 1 || im.height < 1 || im.rgb == nullptr)
        {
            throw std::runtime_error("an image texture's pixels don't match its size");
        }
        texels += static_cast<std::size_t>(im.width) * im.height;
    }
    if (texels > static_cast<std::size_t>(std::numeric_limits<int>::max()))
    {
        throw std::runtime_error("the images have too many pixels for the GPU renderer");
    }
    for (int64_t i = 0; i < s.texture_count; ++i)
    {
        const rt_texture_desc& t = s.textures[i];
        switch (t.kind)
        {
        case RT_TEXTURE_SOLID:
            break;
        case RT_TEXTURE_CHECKER:
            if (t.even < 0 || t.even >= i || t.odd < 0 || t.odd >= i)
            {
                throw std::runtime_error("a checker texture must refer to textures that come before it");
            }
            break;
        case RT_TEXTURE_IMAGE:
            if (t.image < 0 || t.image >= s.image_count)
            {
                throw std::runtime_error("an image texture refers to an image that doesn't exist");
            }
            break;
        case RT_TEXTURE_NOISE:
            if (t.perlin < 0 || t.perlin >= s.perlin_count)
            {
                throw std::runtime_error("a marble texture refers to Perlin tables that don't exist");
            }
            break;
        default:
            throw std::runtime_error("unknown texture kind");
        }
    }
    for (int64_t i = 0; i < s.material_count; ++i)
    {
        const rt_material_desc& m = s.materials[i];
        if (has_texture(m))
        {
            if (m.texture < 0 || m.texture >= s.texture_count)
            {
                throw std::runtime_error("a material refers to a texture that doesn't exist");
            }
        }
        else if (m.kind != RT_MATERIAL_METAL && m.kind != RT_MATERIAL_DIELECTRIC)
        {
            throw std::runtime_error("unknown material kind");
        }
    }
    const int sampling = s.camera.sampling;
    if ((sampling == RT_SAMPLING_LIGHTS || sampling == RT_SAMPLING_MIXTURE) && s.light_count == 0)
    {
        throw std::runtime_error("the lights and mixture strategies need a lights list that isn't empty");
    }
    for (int64_t i = 0; i < s.light_count; ++i)
    {
        const rt_light_desc& l = s.lights[i];
        const double nx = l.quad.u.y * l.quad.v.z - l.quad.u.z * l.quad.v.y;
        const double ny = l.quad.u.z * l.quad.v.x - l.quad.u.x * l.quad.v.z;
        const double nz = l.quad.u.x * l.quad.v.y - l.quad.u.y * l.quad.v.x;
        const bool has_size = l.kind == RT_LIGHT_QUAD ? nx * nx + ny * ny + nz * nz > 0 : l.radius > 0;
        if (!(l.weight > 0 && l.weight <= 1) || !has_size || (l.kind != RT_LIGHT_QUAD && l.kind != RT_LIGHT_SPHERE))
        {
            throw std::runtime_error("a light must have a weight in (0, 1] and a size");
        }
    }
    for (int64_t i = 0; i < s.sphere_count; ++i)
    {
        if (s.spheres[i].material < 0 || s.spheres[i].material >= s.material_count)
        {
            throw std::runtime_error("a sphere refers to a material that doesn't exist");
        }
    }
    for (int64_t i = 0; i < s.quad_count; ++i)
    {
        if (s.quads[i].material < 0 || s.quads[i].material >= s.material_count)
        {
            throw std::runtime_error("a quad refers to a material that doesn't exist");
        }
    }
    std::size_t boundary_quads = 0;
    for (int64_t i = 0; i < s.medium_count; ++i)
    {
        const rt_medium_desc& m = s.media[i];
        if (m.material < 0 || m.material >= s.material_count || s.materials[m.material].kind != RT_MATERIAL_ISOTROPIC)
        {
            throw std::runtime_error("a volume of smoke must refer to an isotropic material");
        }
        if (m.boundary == RT_BOUNDARY_CONVEX_QUADS)
        {
            if (m.quad_count <= 0 || m.quads == nullptr ||
                !quads_are_convex(m.quads, static_cast<std::size_t>(m.quad_count)))
            {
                throw std::runtime_error("a volume of smoke must be bounded by one sphere or by quads of a convex shape");
            }
            boundary_quads += static_cast<std::size_t>(m.quad_count);
        }
        else if (m.boundary != RT_BOUNDARY_SPHERE)
        {
            throw std::runtime_error("a volume of smoke must be bounded by one sphere or by quads of a convex shape");
        }
    }
    const auto limit = static_cast<std::size_t>(std::numeric_limits<int>::max());
    if (static_cast<std::size_t>(s.sphere_count + s.quad_count + s.medium_count) > limit || boundary_quads > limit)
    {
        throw std::runtime_error("the scene has too many shapes for the GPU renderer");
    }
    if (s.combination < RT_COMBINE_BSDF_ONLY || s.combination > RT_COMBINE_POWER_HEURISTIC)
    {
        throw std::runtime_error("unknown light combination");
    }
    if (s.microfacet_count != 0 && s.microfacet_count != s.material_count)
    {
        throw std::runtime_error("the materials' GGX side entries must be one per material");
    }
    for (int64_t k = 0; k < s.microfacet_count; ++k)
    {
        const rt_microfacet_desc& side = s.microfacets[k];
        if (!(side.roughness >= 0 && side.roughness <= 1))
        {
            throw std::runtime_error("a GGX roughness must be a number from 0 to 1");
        }
        if (side.kind == RT_MICROFACET_CONDUCTOR &&
            (s.materials[k].kind != RT_MATERIAL_METAL || s.materials[k].fuzz != 0))
        {
            throw std::runtime_error("a rough metal (a GGX conductor) must complete a metal material with fuzz 0");
        }
        if (side.kind == RT_MICROFACET_GLOSS &&
            (s.materials[k].kind != RT_MATERIAL_LAMBERTIAN || !(side.roughness * side.roughness >= 1e-3)))
        {
            throw std::runtime_error("a glossy coat must complete a lambertian material, with r^2 at least 0.001");
        }
        if (side.kind < RT_MICROFACET_NONE || side.kind > RT_MICROFACET_GLOSS)
        {
            throw std::runtime_error("unknown kind of GGX side entry");
        }
    }
    if (s.environment)
    {
        const rt_image_desc& map = s.environment->image;
        if (map.width < 1 || map.height < 1 || map.width > 2048 || map.height > 1024 || map.rgb == nullptr)
        {
            throw std::runtime_error("an environment map must have from 1x1 to 2048x1024 texels");
        }
        for (std::size_t i = 0; i < 3ull * map.width * map.height; ++i)
        {
            if (!(std::isfinite(map.rgb[i]) && map.rgb[i] >= 0))
            {
                throw std::runtime_error("an environment map's radiance must be a finite number >= 0 in every texel");
            }
        }
        if (!std::isfinite(s.environment->yaw_degrees) ||
            !(std::isfinite(s.environment->scale) && s.environment->scale >= 0))
        {
            throw std::runtime_error("an environment map's yaw and scale must be finite numbers, the scale >= 0");
        }
    }
    if (s.mesh_count < 0 || s.instance_count < 0)
    {
        throw std::runtime_error("a table of the scene has a count but no data");
    }
    std::size_t triangles = 0;
    for (int64_t i = 0; i < s.mesh_count; ++i)
    {
        const rt_mesh_desc& m = s.meshes[i];
        if (m.vertex_count < 0 || m.triangle_count < 0 || (m.vertex_count > 0 && m.positions == nullptr) ||
            (m.triangle_count > 0 && m.indices == nullptr))
        {
            throw std::runtime_error("a mesh's vertices or triangles don't match their counts");
        }
        if (m.material < 0 || m.material >= s.material_count)
        {
            throw std::runtime_error("a mesh refers to a material that doesn't exist");
        }
        for (int64_t k = 0; k < 3 * m.vertex_count; ++k)
        {
            if (!std::isfinite(m.positions[k]))
            {
                throw std::runtime_error("a mesh's vertex positions must be finite numbers");
            }
        }
        for (int64_t k = 0; k < 3 * m.triangle_count; ++k)
        {
            if (static_cast<int64_t>(m.indices[k]) >= m.vertex_count)
            {
                throw std::runtime_error("a mesh's triangle refers to a vertex that doesn't exist");
            }
        }
        if (s.instance_count == 0)
        {
            triangles += static_cast<std::size_t>(m.triangle_count);
        }
    }
    for (int64_t i = 0; i < s.instance_count; ++i)
    {
        const rt_mesh_instance_desc& in = s.instances[i];
        if (in.mesh < 0 || in.mesh >= s.mesh_count)
        {
            throw std::runtime_error("a mesh instance refers to a mesh that doesn't exist");
        }
        for (const double x : in.object_to_world)
        {
            if (!std::isfinite(x))
            {
                throw std::runtime_error("a mesh instance's transform must hold finite numbers");
            }
        }
        triangles += static_cast<std::size_t>(s.meshes[in.mesh].triangle_count);
    }
    if (triangles > static_cast<std::size_t>(dev::max_leaf_first))
    {
        throw std::runtime_error("the scene has more triangles than the GPU renderer's tree can address");
    }
}

bool first_book_scene(const rt_scene& s)
{
    if (s.quad_count != 0 || s.medium_count != 0 || s.camera.sampling != RT_SAMPLING_MATERIAL || s.mesh_count != 0)
    {
        return false;
    }
    for (int64_t i = 0; i < s.sphere_count; ++i)
    {
        if (moves(s.spheres[i]))
        {
            return false;
        }
    }
    for (int64_t i = 0; i < s.material_count; ++i)
    {
        const rt_material_desc& m = s.materials[i];
        const bool ok = (m.kind == RT_MATERIAL_LAMBERTIAN && solid_texture(s, m)) || m.kind == RT_MATERIAL_METAL ||
                        m.kind == RT_MATERIAL_DIELECTRIC;
        if (!ok)
        {
            return false;
        }
    }
    return true;
}

template <typename Real>
dev::camera_rec<Real> to_camera(const rt_camera_desc& d)
{
    dev::camera_rec<Real> c{};
    c.image_width = d.image_width;
    c.image_height = d.image_height;
    c.sqrt_spp = d.sqrt_spp;
    c.max_depth = d.max_depth;
    c.recip_sqrt_spp = static_cast<Real>(1.0 / d.sqrt_spp);
    store(c.center, d.center);
    // the pixel grid relative to the camera center, subtracted in double before rounding (src/gpu/renderer.cu:529-534)
    store(c.center_to_pixel00, rt_vec3{d.pixel00_loc.x - d.center.x, d.pixel00_loc.y - d.center.y,
                                       d.pixel00_loc.z - d.center.z});
    store(c.pixel_delta_u, d.pixel_delta_u);
    store(c.pixel_delta_v, d.pixel_delta_v);
    c.defocus = d.defocus != 0;
    store(c.defocus_disk_u, d.defocus_disk_u);
    store(c.defocus_disk_v, d.defocus_disk_v);
    store(c.sky_bottom, d.sky_bottom);
    store(c.sky_top, d.sky_top);
    return c;
}

template <typename Real>
device_scene<Real>::device_scene(const rt_scene& s, int32_t tree, int32_t nodes, bool quality_kernel)
    : tree_(build_bvh<Real>(s, leaf_size_for(static_cast<std::size_t>(s.sphere_count + s.quad_count)), tree))
{
    // materials: a texture of one color becomes the material's albedo (src/gpu/renderer.cu:548-601)
    std::vector<dev::material_rec<Real>> materials;
    for (int64_t i = 0; i < s.material_count; ++i)
    {
        const rt_material_desc& d = s.materials[i];
        dev::material_rec<Real> m{};
        m.kind = d.kind;
        const bool textured = has_texture(d) && !solid_texture(s, d);
        m.texture = textured ? d.texture : -1;
        m.flags = (textured && texture_needs_uv(s, d.texture) ? dev::needs_uv : 0u) |
                  (textured && s.textures[d.texture].kind == RT_TEXTURE_CHECKER ? dev::needs_exact_point : 0u);
        store(m.albedo, solid_texture(s, d) ? s.textures[d.texture].rgb : d.albedo);
        m.fuzz = static_cast<Real>(d.fuzz);
        m.refraction_index = static_cast<Real>(d.refraction_index);
        materials.push_back(m);
    }
    std::vector<dev::texture_rec<Real>> textures;
    int noises_so_far = 0;
    for (int64_t i = 0; i < s.texture_count; ++i)
    {
        const rt_texture_desc& d = s.textures[i];
        dev::texture_rec<Real> t{};
        t.kind = d.kind;
        t.even = d.even;
        t.odd = d.odd;
        t.table = d.kind == RT_TEXTURE_NOISE ? noises_so_far++ : d.image;
        store(t.rgb, d.rgb);
        t.inv_scale = static_cast<Real>(d.inv_scale);
        textures.push_back(t);
    }
    std::vector<dev::noise_rec<Real>> noises;
    for (int64_t i = 0; i < s.texture_count; ++i)
    {
        const rt_texture_desc& d = s.textures[i];
        if (d.kind != RT_TEXTURE_NOISE)
        {
            continue;
        }
        const rt_perlin_desc& p = s.perlins[d.perlin];
        dev::noise_rec<Real> n{};
        for (int k = 0; k < 256; ++k)
        {
            store(n.gradients[k], p.random_vectors[k]);
            n.perm_x[k] = static_cast<uint8_t>(p.perm_x[k]);
            n.perm_y[k] = static_cast<uint8_t>(p.perm_y[k]);
            n.perm_z[k] = static_cast<uint8_t>(p.perm_z[k]);
        }
        n.scale = static_cast<Real>(d.scale);
        noises.push_back(n);
    }
    std::vector<dev::image_info> infos;
    std::vector<dev::texel> texels;
    for (int64_t i = 0; i < s.image_count; ++i)
    {
        const rt_image_desc& im = s.images[i];
        infos.push_back({static_cast<int32_t>(texels.size()), im.width, im.height, 0});
        for (std::size_t k = 0; k < static_cast<std::size_t>(im.width) * im.height; ++k)
        {
            texels.push_back({im.rgb[3 * k], im.rgb[3 * k + 1], im.rgb[3 * k + 2], 0.0f});
        }
    }
    // media: each with a box around it over the whole shutter time, rounded outward (src/gpu/renderer.cu:933-975)
    std::vector<dev::medium_rec<Real>> media;
    std::vector<dev::quad_rec<Real>> boundary;
    for (int64_t i = 0; i < s.medium_count; ++i)
    {
        const rt_medium_desc& d = s.media[i];
        dev::medium_rec<Real> m{};
        m.neg_inv_density = static_cast<Real>(d.neg_inv_density);
        m.material = d.material;
        box_d box;
        if (d.boundary == RT_BOUNDARY_SPHERE)
        {
            store(m.center, d.sphere.center);
            m.radius = static_cast<Real>(d.sphere.radius);
            store(m.displacement, d.sphere.displacement);
            for (int k = 0; k < 3; ++k)
            {
                const Real end = m.center[k] + m.displacement[k];
                const double lo = std::min(static_cast<double>(m.center[k]), static_cast<double>(end));
                const double hi = std::max(static_cast<double>(m.center[k]), static_cast<double>(end));
                box.min[k] = std::nextafter(lo - static_cast<double>(m.radius), -INFINITY);
                box.max[k] = std::nextafter(hi + static_cast<double>(m.radius), INFINITY);
            }
        }
        else
        {
            m.first_quad = static_cast<int32_t>(boundary.size());
            m.quad_count = static_cast<int32_t>(d.quad_count);
            for (int64_t q = 0; q < d.quad_count; ++q)
            {
                boundary.push_back(make_quad_rec<Real>(d.quads[q]));
                box.enclose(quad_box(d.quads[q]));
            }
        }
        for (int k = 0; k < 3; ++k)
        {
            m.box_min[k] = round_down<Real>(box.min[k]);
            m.box_max[k] = round_up<Real>(box.max[k]);
        }
        media.push_back(m);
    }
    // the lights list, for the strategies that aim at it (src/gpu/renderer.cu:976-1008)
    std::vector<dev::light_rec<Real>> lights;
    const int sampling = s.camera.sampling;
    if (sampling == RT_SAMPLING_LIGHTS || sampling == RT_SAMPLING_MIXTURE)
    {
        double cdf = 0;
        for (int64_t i = 0; i < s.light_count; ++i)
        {
            const rt_light_desc& d = s.lights[i];
            cdf += d.weight;
            dev::light_rec<Real> l{};
            l.twin = -1;
            if (d.kind == RT_LIGHT_QUAD)
            {
                const double nx = d.quad.u.y * d.quad.v.z - d.quad.u.z * d.quad.v.y;
                const double ny = d.quad.u.z * d.quad.v.x - d.quad.u.x * d.quad.v.z;
                const double nz = d.quad.u.x * d.quad.v.y - d.quad.u.y * d.quad.v.x;
                l.quad = make_quad_rec<Real>(d.quad);
                store(l.u, d.quad.u);
                store(l.v, d.quad.v);
                l.area = static_cast<Real>(std::sqrt(nx * nx + ny * ny + nz * nz));
                for (int64_t q = 0; q < s.quad_count; ++q)
                {
                    if (same(s.quads[q].q, d.quad.q) && same(s.quads[q].u, d.quad.u) && same(s.quads[q].v, d.quad.v))
                    {
                        l.twin = tree_.quad_ids[q];
                        break;
                    }
                }
            }
            else
            {
                store(l.center, d.center);
                l.radius = static_cast<Real>(d.radius);
                for (int64_t q = 0; q < s.sphere_count; ++q)
                {
                    if (!moves(s.spheres[q]) && same(s.spheres[q].center, d.center) && s.spheres[q].radius == d.radius)
                    {
                        l.twin = tree_.sphere_ids[q];
                        break;
                    }
                }
            }
            l.cdf = static_cast<Real>(cdf);
            l.weight = static_cast<Real>(d.weight);
            l.kind = d.kind;
            lights.push_back(l);
        }
    }
    bool motion = false;
    for (int64_t i = 0; i < s.sphere_count; ++i)
    {
        motion = motion || moves(s.spheres[i]);
    }
    for (int64_t i = 0; i < s.medium_count; ++i)
    {
        motion = motion || (s.media[i].boundary == RT_BOUNDARY_SPHERE && moves(s.media[i].sphere));
    }

    // the kernels address a node's planes as the array's base plus a 32-bit byte offset
    if (tree_.wide.size() > std::numeric_limits<uint32_t>::max() / sizeof(dev::wide_node<Real>))
    {
        throw std::runtime_error("the scene's BVH is larger than 4 GiB");
    }
    nodes_ = device_array<dev::wide_node<Real>>(tree_.wide);
    spheres_ = device_array<dev::sphere_rec<Real>>(tree_.spheres);
    moving_ = device_array<dev::moving_sphere_rec<Real>>(tree_.moving_spheres);
    quads_ = device_array<dev::quad_rec<Real>>(tree_.quads);
    prim_materials_ = device_array<int32_t>(tree_.prim_materials);
    prim_ranks_ = device_array<int32_t>(tree_.prim_ranks);
    frames_ = device_array<dev::sphere_frame<Real>>(tree_.sphere_frames);
    materials_ = device_array<dev::material_rec<Real>>(materials);
    textures_ = device_array<dev::texture_rec<Real>>(textures);
    images_ = device_array<dev::image_info>(infos);
    texels_ = device_array<dev::texel>(texels);
    noises_ = device_array<dev::noise_rec<Real>>(noises);
    media_ = device_array<dev::medium_rec<Real>>(media);
    boundary_quads_ = device_array<dev::quad_rec<Real>>(boundary);
    lights_ = device_array<dev::light_rec<Real>>(lights);
    triangles_ = device_array<dev::triangle_rec<Real>>(tree_.triangles);
    instances_ = device_array<dev::instance_rec<Real>>(tree_.instances);
    bytes_ = bytes_of(nodes_) + bytes_of(spheres_) + bytes_of(moving_) + bytes_of(quads_) + bytes_of(prim_materials_) +
             bytes_of(prim_ranks_) + bytes_of(instances_) +
             bytes_of(frames_) + bytes_of(materials_) + bytes_of(textures_) + bytes_of(images_) + bytes_of(texels_) +
             bytes_of(noises_) + bytes_of(media_) + bytes_of(boundary_quads_) + bytes_of(lights_) +
             bytes_of(triangles_);

    dev::scene_view<Real>& v = view_;
    quality_ = quality_kernel || needs_quality(s);
    if (quality_)
    {
        // the quality integrator's tables (src/gpu/quality.cu:112-170; float records, read as Real by the FP64 kernels)
        const lamp_set set = build_lamp_set(s, tree_.sphere_ids, tree_.quad_ids, tree_.prim_materials.size());
        lamps_ = device_array<dev::emitter_rec>(set.lamps);
        lamp_of_ = device_array<int32_t>(set.lamp_of);
        microfacets_ = device_array<dev::microfacet_rec>(build_microfacets(s));
        dev::quality_view& q = v.quality;
        q.combination = s.combination;
        q.lamp_count = static_cast<int32_t>(set.lamps.size());
        q.lamps = lamps_.data();
        q.lamp_of = lamp_of_.data();
        q.microfacets = microfacets_.size() > 0 ? microfacets_.data() : nullptr;
        int lights_sampled = q.lamp_count;
        if (s.environment)
        {
            const environment_tables env = build_environment_tables(*s.environment);
            env_texels_ = device_array<dev::env_texel>(env.texels);
            env_bins_ = device_array<dev::alias_bin>(env.bins);
            env_row_cos_ = device_array<float>(env.row_cos);
            const double yaw = s.environment->yaw_degrees * std::numbers::pi / 180;
            q.env_present = 1;
            q.env_texels = env_texels_.data();
            q.env_bins = env_bins_.size() > 0 ? env_bins_.data() : nullptr;
            q.env_row_cos = env_row_cos_.data();
            q.env_width = env.width;
            q.env_height = env.height;
            q.env_log2_bins = env.log2_bins;
            q.env_cos_yaw = static_cast<float>(std::cos(yaw));
            q.env_sin_yaw = static_cast<float>(std::sin(yaw));
            q.env_scale = static_cast<float>(s.environment->scale);
            lights_sampled += env_bins_.size() > 0 ? 1 : 0;
        }
        q.light_pmf = lights_sampled > 0 ? static_cast<float>(1.0 / lights_sampled) : 0.0f;
        bytes_ += bytes_of(lamps_) + bytes_of(lamp_of_) + bytes_of(microfacets_) + bytes_of(env_texels_) +
                  bytes_of(env_bins_) + bytes_of(env_row_cos_);
    }
    // packed nodes for the quality integrator's one-level trees larger than the L2 caches (or when asked): half the
    // bytes a walk reads per node, and fewer registers to hold a node's planes
    if (nodes != RT_NODES_AUTO && nodes != RT_NODES_FLOAT && nodes != RT_NODES_PACKED)
    {
        throw std::invalid_argument("unknown node layout");
    }
    // (a tree whose root is a leaf has no inner node to pack: its walks read none)
    const bool packable = sizeof(Real) == sizeof(float) && quality_ && tree_.instances.empty();
    if (nodes == RT_NODES_PACKED && !packable)
    {
        throw std::invalid_argument("packed nodes are for float32 scenes on the quality integrator with a one-level tree");
    }
    constexpr std::size_t packed_above = std::size_t{16} << 20;
    if (packable && (nodes == RT_NODES_PACKED || (nodes == RT_NODES_AUTO && nodes_.bytes() > packed_above)))
    {
        if constexpr (sizeof(Real) == sizeof(float))
        {
            packed_ = device_array<dev::packed_node>(pack_nodes(tree_.wide));
            bytes_ += bytes_of(packed_);
        }
    }
    v.nodes = nodes_.data();
    v.packed_nodes = packed_.size() > 0 ? packed_.data() : nullptr;
    v.root_ref = tree_.root_ref;
    v.stack_depth = tree_.nodes.empty() ? 0 : tree_.stack_entries;
    v.spheres = spheres_.data();
    v.moving_spheres = moving_.data();
    v.quads = quads_.data();
    v.triangles = triangles_.data();
    v.instances = instances_.data();
    v.instance_count = static_cast<int32_t>(tree_.instances.size());
    v.moving_base = tree_.moving_base;
    v.quad_base = tree_.quad_base;
    v.triangle_base = tree_.triangle_base;
    v.instanced_base = tree_.instanced_base;
    v.media_base = tree_.media_base;
    v.prim_materials = prim_materials_.data();
    v.prim_ranks = prim_ranks_.data();
    v.sphere_frames = frames_.data();
    v.materials = materials_.data();
    v.textures = textures_.data();
    v.images = images_.data();
    v.texels = texels_.data();
    v.noises = noises_.data();
    v.media = media_.data();
    v.boundary_quads = boundary_quads_.data();
    v.medium_count = static_cast<int32_t>(media.size());
    v.light_count = static_cast<int32_t>(lights.size());
    v.lights = lights_.data();
    v.sampling = sampling;
    v.motion = motion ? 1 : 0;
}

template class device_scene<float>;
template class device_scene<double>;
template dev::camera_rec<float> to_camera<float>(const rt_camera_desc&);
template dev::camera_rec<double> to_camera<double>(const rt_camera_desc&);
}
