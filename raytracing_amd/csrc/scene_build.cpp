// The scene-building API: the books' scenes built as objects (spheres, quads, boxes, lists, object BVHs, translate and
// rotate_y instances, constant media; materials and textures shared by pointer), then flattened into the renderer's
// tables (scene_tables). This is the path by which the reference's programs get a scene
// (src/scenes.ixx, src/gpu_scene.ixx:97-768,1333-1445), kept as this library's API.
//
// The objects here carry only what building and flattening need: their bounding boxes (which order an object BVH) and
// their numbers. They draw random numbers from the scene's own mt19937_64, started from its default seed for every
// build, in the reference's order, so a scene comes out the same as the reference's: the same sphere positions and
// materials, the same Perlin tables.
#include "scene_build.hpp"

#include "gltf.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <numbers>
#include <optional>
#include <random>
#include <span>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <variant>

namespace rt
{
namespace
{
// --- numbers --------------------------------------------------------------------------------------------------------
struct d3
{
    double x = 0, y = 0, z = 0;
};
d3 operator+(d3 a, d3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
d3 operator-(d3 a, d3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
d3 operator*(double t, d3 v) { return {t * v.x, t * v.y, t * v.z}; }
d3 operator*(d3 a, d3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
d3 operator/(d3 v, double t) { return (1 / t) * v; }
d3 neg(d3 v) { return {-v.x, -v.y, -v.z}; }
double dot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
d3 cross(d3 u, d3 v) { return {u.y * v.z - u.z * v.y, u.z * v.x - u.x * v.z, u.x * v.y - u.y * v.x}; }
double length(d3 v) { return std::sqrt(dot(v, v)); }
d3 unit(d3 v) { return v / length(v); }
double radians(double degrees) { return degrees * std::numbers::pi / 180.0; }
rt_vec3 to_rt(d3 v) { return {v.x, v.y, v.z}; }

// The scene's random numbers (src/random.ixx): mt19937_64 from its default seed, doubles in [0, 1) from
// uniform_real_distribution, and the reference's derived forms
class scene_random
{
public:
    double uniform() { return std::uniform_real_distribution<double>(0.0, 1.0)(engine_); }
    double uniform(double min, double max) { return min + (max - min) * uniform(); }
    int integer(int min, int max) { return static_cast<int>(uniform(min, max + 1)); }
    d3 vector()
    {
        const double a = uniform();
        const double b = uniform();
        const double c = uniform();
        return {a, b, c};
    }
    d3 vector(double min, double max)
    {
        const double a = uniform(min, max);
        const double b = uniform(min, max);
        const double c = uniform(min, max);
        return {a, b, c};
    }

private:
    std::mt19937_64 engine_{std::mt19937_64::default_seed};
};

// --- bounding boxes (src/aabb.ixx, src/interval.ixx) ----------------------------------------------------------------
struct span1
{
    double lo = +INFINITY, hi = -INFINITY;
};
span1 join(span1 a, span1 b) { return {std::min(a.lo, b.lo), std::max(a.hi, b.hi)}; }
// A box side no thinner than 0.0001: a flat shape's box is padded by half of it on each side
span1 padded(double a, double b)
{
    span1 s{std::min(a, b), std::max(a, b)};
    if (s.hi - s.lo < 0.0001)
    {
        s = {s.lo - 0.0001 / 2, s.hi + 0.0001 / 2};
    }
    return s;
}
struct box3
{
    span1 axis[3];
    static box3 corners(d3 a, d3 b) { return {{padded(a.x, b.x), padded(a.y, b.y), padded(a.z, b.z)}}; }
    int longest() const
    {
        const double x = axis[0].hi - axis[0].lo, y = axis[1].hi - axis[1].lo, z = axis[2].hi - axis[2].lo;
        if (x > y)
        {
            return x > z ? 0 : 2;
        }
        return y > z ? 1 : 2;
    }
};
box3 join(const box3& a, const box3& b)
{
    return {{join(a.axis[0], b.axis[0]), join(a.axis[1], b.axis[1]), join(a.axis[2], b.axis[2])}};
}
box3 shifted(const box3& b, d3 o)
{
    return {{{b.axis[0].lo + o.x, b.axis[0].hi + o.x}, {b.axis[1].lo + o.y, b.axis[1].hi + o.y},
             {b.axis[2].lo + o.z, b.axis[2].hi + o.z}}};
}

// --- textures and materials, shared by pointer -----------------------------------------------------------------------
struct texture_def;
using texture_ref = std::shared_ptr<const texture_def>;
struct image_data
{
    int width = 0, height = 0;
    std::vector<float> rgb;
};
struct perlin_tables
{
    std::array<d3, 256> gradients;
    std::array<int, 256> perm_x, perm_y, perm_z;
};
struct texture_def
{
    int kind = RT_TEXTURE_SOLID;
    d3 rgb;
    double inv_scale = 1;
    texture_ref even, odd;
    std::shared_ptr<const image_data> image;
    std::shared_ptr<const perlin_tables> perlin;
    double scale = 1;
};
texture_ref solid(d3 c)
{
    auto t = std::make_shared<texture_def>();
    t->rgb = c;
    return t;
}
texture_ref checker(double scale, d3 even, d3 odd)
{
    auto t = std::make_shared<texture_def>();
    t->kind = RT_TEXTURE_CHECKER;
    t->inv_scale = 1.0 / scale;
    t->even = solid(even);
    t->odd = solid(odd);
    return t;
}
// noise_texture: its perlin object draws 256 unit gradients, then three Fisher-Yates permutations of 0..255
// (src/perlin.ixx:18-107)
texture_ref marble(scene_random& rng, double scale)
{
    auto p = std::make_shared<perlin_tables>();
    for (d3& g : p->gradients)
    {
        g = unit(rng.vector(-1, 1));
    }
    for (auto* perm : {&p->perm_x, &p->perm_y, &p->perm_z})
    {
        for (int i = 0; i < 256; ++i)
        {
            (*perm)[i] = i;
        }
        for (int i = 255; i > 0; --i)
        {
            std::swap((*perm)[i], (*perm)[rng.integer(0, i)]);
        }
    }
    auto t = std::make_shared<texture_def>();
    t->kind = RT_TEXTURE_NOISE;
    t->perlin = std::move(p);
    t->scale = scale;
    return t;
}
texture_ref picture(std::shared_ptr<const image_data> image)
{
    auto t = std::make_shared<texture_def>();
    t->kind = RT_TEXTURE_IMAGE;
    t->image = std::move(image);
    return t;
}

struct material_def
{
    int kind = RT_MATERIAL_LAMBERTIAN;
    texture_ref texture;
    d3 albedo;
    double fuzz = 0;
    double refraction_index = 1;
    int microfacet = RT_MICROFACET_NONE; // a GGX side entry (src/gpu_scenes_materials.ixx:35-60)
    double roughness = 0;
};
using material_ref = std::shared_ptr<const material_def>;
material_ref lambertian(texture_ref t)
{
    auto m = std::make_shared<material_def>();
    m->texture = std::move(t);
    return m;
}
material_ref lambertian(d3 c) { return lambertian(solid(c)); }
material_ref metal(d3 albedo, double fuzz)
{
    auto m = std::make_shared<material_def>();
    m->kind = RT_MATERIAL_METAL;
    m->albedo = albedo;
    m->fuzz = std::min(fuzz, 1.0);
    return m;
}
material_ref dielectric(double ri)
{
    auto m = std::make_shared<material_def>();
    m->kind = RT_MATERIAL_DIELECTRIC;
    m->refraction_index = ri;
    return m;
}
material_ref lamp_light(d3 c)
{
    auto m = std::make_shared<material_def>();
    m->kind = RT_MATERIAL_DIFFUSE_LIGHT;
    m->texture = solid(c);
    return m;
}
// A GGX rough metal of reflectance F0 at normal incidence: the books' metal without fuzz, completed by its side entry
material_ref rough_metal(d3 f0, double roughness)
{
    auto m = std::make_shared<material_def>();
    m->kind = RT_MATERIAL_METAL;
    m->albedo = f0;
    m->microfacet = RT_MICROFACET_CONDUCTOR;
    m->roughness = roughness;
    return m;
}
// A glossy coat (GGX, F0 0.04) over a matte color
material_ref glossy(d3 base, double roughness)
{
    auto m = std::make_shared<material_def>();
    m->texture = solid(base);
    m->microfacet = RT_MICROFACET_GLOSS;
    m->roughness = roughness;
    return m;
}

// A triangle mesh in its own coordinates (vertices shared by its triangles) and the material all of it wears
struct mesh_def
{
    std::vector<float> positions;
    std::vector<uint32_t> indices;
    material_ref material;
};

// --- objects --------------------------------------------------------------------------------------------------------
struct object;
using object_ref = std::unique_ptr<object>;
struct sphere_shape
{
    d3 center0, motion; // center at time 0 and its displacement over the shutter time
    double radius;
    material_ref material;
};
struct quad_shape
{
    d3 q, u, v;
    material_ref material;
};
struct group
{
    std::vector<object_ref> members; // a hittable_list
};
struct tree_node
{
    object_ref left, right; // a node of an object BVH
};
struct shifted_object
{
    object_ref inner;
    d3 offset;
};
struct turned_object
{
    object_ref inner;
    double sin_theta, cos_theta;
};
struct medium_object
{
    object_ref boundary;
    double neg_inv_density;
    material_ref phase; // isotropic, made by the medium itself
};
struct object
{
    std::variant<sphere_shape, quad_shape, group, tree_node, shifted_object, turned_object, medium_object> what;
    box3 bounds;
};

object_ref make_sphere(d3 c1, d3 c2, double radius, material_ref m)
{
    auto o = std::make_unique<object>();
    const double r = std::fmax(0.0, radius);
    const d3 motion = c2 - c1;
    const d3 rv{r, r, r};
    const d3 end = c1 + 1.0 * motion; // the center's path at time 1, as the reference's ray::at(1)
    const d3 start = c1 + 0.0 * motion;
    o->bounds = join(box3::corners(start - rv, start + rv), box3::corners(end - rv, end + rv));
    o->what = sphere_shape{c1, motion, r, std::move(m)};
    return o;
}
object_ref make_sphere(d3 c, double radius, material_ref m) { return make_sphere(c, c, radius, std::move(m)); }
object_ref make_quad(d3 q, d3 u, d3 v, material_ref m)
{
    auto o = std::make_unique<object>();
    o->bounds = join(box3::corners(q, q + u + v), box3::corners(q + u, q + v));
    o->what = quad_shape{q, u, v, std::move(m)};
    return o;
}
struct list_builder
{
    group g;
    box3 bounds;
    void add(object_ref o)
    {
        bounds = join(bounds, o->bounds);
        g.members.push_back(std::move(o));
    }
    object_ref finish()
    {
        auto o = std::make_unique<object>();
        o->bounds = bounds;
        o->what = std::move(g);
        return o;
    }
};
// The six faces of a box, their normals outward (src/quad.ixx:116-140)
object_ref make_box(d3 a, d3 b, const material_ref& m)
{
    const d3 mn{std::fmin(a.x, b.x), std::fmin(a.y, b.y), std::fmin(a.z, b.z)};
    const d3 mx{std::fmax(a.x, b.x), std::fmax(a.y, b.y), std::fmax(a.z, b.z)};
    const d3 dx{mx.x - mn.x, 0, 0}, dy{0, mx.y - mn.y, 0}, dz{0, 0, mx.z - mn.z};
    list_builder sides;
    sides.add(make_quad({mn.x, mn.y, mx.z}, dx, dy, m));
    sides.add(make_quad({mx.x, mn.y, mx.z}, neg(dz), dy, m));
    sides.add(make_quad({mx.x, mn.y, mn.z}, neg(dx), dy, m));
    sides.add(make_quad({mn.x, mn.y, mn.z}, dz, dy, m));
    sides.add(make_quad({mn.x, mx.y, mx.z}, dx, neg(dz), m));
    sides.add(make_quad({mn.x, mn.y, mn.z}, dx, dz, m));
    return sides.finish();
}
object_ref make_translate(object_ref inner, d3 offset)
{
    auto o = std::make_unique<object>();
    o->bounds = shifted(inner->bounds, offset);
    o->what = shifted_object{std::move(inner), offset};
    return o;
}
object_ref make_rotate_y(object_ref inner, double angle)
{
    const double s = std::sin(radians(angle));
    const double c = std::cos(radians(angle));
    auto o = std::make_unique<object>();
    const box3 b = inner->bounds;
    box3 turned;
    for (const double x : {b.axis[0].lo, b.axis[0].hi})
    {
        for (const double y : {b.axis[1].lo, b.axis[1].hi})
        {
            for (const double z : {b.axis[2].lo, b.axis[2].hi})
            {
                const d3 corner{c * x + s * z, y, -s * x + c * z};
                turned = join(turned, box3::corners(corner, corner));
            }
        }
    }
    o->bounds = turned;
    o->what = turned_object{std::move(inner), s, c};
    return o;
}
object_ref make_medium(object_ref boundary, double density, d3 albedo)
{
    auto o = std::make_unique<object>();
    o->bounds = boundary->bounds;
    auto phase = std::make_shared<material_def>();
    phase->kind = RT_MATERIAL_ISOTROPIC;
    phase->texture = solid(albedo);
    o->what = medium_object{std::move(boundary), -1 / density, std::move(phase)};
    return o;
}
// An object BVH over a list (src/bvh.ixx:55-84): split in half on the longest axis of the common box, ordered by where
// the members' boxes start on it (a stable sort: equal starts keep the list's order)
object_ref split_tree(std::span<object_ref> members)
{
    if (members.size() == 1)
    {
        return std::move(members.front());
    }
    box3 all;
    for (const auto& m : members)
    {
        all = join(all, m->bounds);
    }
    const int axis = all.longest();
    std::stable_sort(members.begin(), members.end(),
                     [axis](const object_ref& a, const object_ref& b) { return a->bounds.axis[axis].lo < b->bounds.axis[axis].lo; });
    const std::size_t middle = members.size() / 2;
    object_ref left = split_tree(members.first(middle));
    object_ref right = split_tree(members.subspan(middle));
    auto o = std::make_unique<object>();
    o->bounds = join(left->bounds, right->bounds);
    o->what = tree_node{std::move(left), std::move(right)};
    return o;
}
object_ref make_tree(list_builder list)
{
    if (list.g.members.empty())
    {
        return list.finish();
    }
    return split_tree(list.g.members);
}

// --- scenes (src/scenes.ixx) -----------------------------------------------------------------------------------------
struct camera_settings
{
    double aspect_ratio = 1.0;
    int image_width = 100;
    int samples_per_pixel = 10;
    int max_depth = 10;
    int sampling = RT_SAMPLING_MIXTURE;
    double vfov = 90;
    d3 lookfrom{0, 0, 0};
    d3 lookat{0, 0, -1};
    d3 vup{0, 1, 0};
    double defocus_angle = 0;
    double focus_dist = 10;
    d3 sky_bottom{0, 0, 0};
    d3 sky_top{0, 0, 0};
};
constexpr d3 daylight_bottom{1.0, 1.0, 1.0};
constexpr d3 daylight_top{0.5, 0.7, 1.0};

camera_settings cover_camera()
{
    camera_settings c;
    c.aspect_ratio = 16.0 / 9.0;
    c.image_width = 1920;
    c.samples_per_pixel = 500;
    c.max_depth = 50;
    c.vfov = 20;
    c.lookfrom = {13, 2, 3};
    c.lookat = {0, 0, 0};
    c.defocus_angle = 0.6;
    c.focus_dist = 10.0;
    c.sky_bottom = daylight_bottom;
    c.sky_top = daylight_top;
    return c;
}
camera_settings cornell_camera()
{
    camera_settings c;
    c.image_width = 1080;
    c.samples_per_pixel = 1000;
    c.max_depth = 50;
    c.vfov = 40;
    c.lookfrom = {278, 278, -800};
    c.lookat = {278, 278, 0};
    return c;
}

struct environment_def
{
    std::shared_ptr<const image_data> image;
    double yaw_degrees = 0;
    double scale = 1;
};

struct built
{
    list_builder world;
    list_builder lights;
    camera_settings camera;
    std::vector<mesh_def> meshes;
    std::vector<rt_mesh_instance_desc> instances;
    std::optional<environment_def> environment;
    int combination = -1; // the light combination the scene asks for by default (-p), or -1 for the books' strategy
};

list_builder random_spheres(scene_random& rng, bool bouncing)
{
    list_builder world;
    const texture_ref ground = bouncing ? checker(0.32, {0.2, 0.3, 0.1}, {0.9, 0.9, 0.9}) : solid({0.5, 0.5, 0.5});
    world.add(make_sphere({0, -1000, 0}, 1000, lambertian(ground)));
    for (int a = -11; a < 11; ++a)
    {
        for (int b = -11; b < 11; ++b)
        {
            const double choose = rng.uniform();
            const double cx = a + 0.9 * rng.uniform();
            const double cz = b + 0.9 * rng.uniform();
            const d3 center{cx, 0.2, cz};
            if (length(center - d3{4, 0.2, 0}) <= 0.9)
            {
                continue;
            }
            material_ref m;
            d3 center2 = center;
            if (choose < 0.8)
            {
                const d3 c1 = rng.vector();
                const d3 c2 = rng.vector();
                m = lambertian(c1 * c2);
                if (bouncing)
                {
                    center2 = center + d3{0, rng.uniform(0, 0.5), 0};
                }
            }
            else if (choose < 0.95)
            {
                // The reference's metal(color::random(0.5, 1), random_double(0, 0.5)): its compiler evaluates call
                // arguments right to left, so the fuzz comes first in the stream
                const double fuzz = rng.uniform(0, 0.5);
                const d3 albedo = rng.vector(0.5, 1);
                m = metal(albedo, fuzz);
            }
            else
            {
                m = dielectric(1.5);
            }
            world.add(make_sphere(center, center2, 0.2, m));
        }
    }
    world.add(make_sphere({0, 1, 0}, 1.0, dielectric(1.5)));
    world.add(make_sphere({-4, 1, 0}, 1.0, lambertian(d3{0.4, 0.2, 0.1})));
    world.add(make_sphere({4, 1, 0}, 1.0, metal({0.7, 0.6, 0.5}, 0.0)));
    return world;
}
void cornell_walls(list_builder& world)
{
    const auto red = lambertian(d3{0.65, 0.05, 0.05});
    const auto white = lambertian(d3{0.73, 0.73, 0.73});
    const auto green = lambertian(d3{0.12, 0.45, 0.15});
    world.add(make_quad({555, 0, 0}, {0, 555, 0}, {0, 0, 555}, green));
    world.add(make_quad({0, 0, 0}, {0, 555, 0}, {0, 0, 555}, red));
    world.add(make_quad({0, 0, 0}, {555, 0, 0}, {0, 0, 555}, white));
    world.add(make_quad({555, 555, 555}, {-555, 0, 0}, {0, 0, -555}, white));
    world.add(make_quad({0, 0, 555}, {555, 0, 0}, {0, 555, 0}, white));
}
object_ref tall_block(const material_ref& m)
{
    return make_translate(make_rotate_y(make_box({0, 0, 0}, {165, 330, 165}, m), 15), {265, 0, 295});
}
object_ref short_block(const material_ref& m)
{
    return make_translate(make_rotate_y(make_box({0, 0, 0}, {165, 165, 165}, m), -18), {130, 0, 65});
}
object_ref cornell_lamp(material_ref m) { return make_quad({343, 554, 332}, {-130, 0, 0}, {0, 0, -105}, std::move(m)); }

// --- generated content: meshes and skies (the triangle scene and the light-sample tests) -----------------------------
// The unit icosphere: the icosahedron's 20 faces, each split into 4 `level` times, every new vertex (an edge's midpoint)
// pushed out onto the sphere; 20 * 4^level triangles, counter-clockwise seen from outside
struct sphere_mesh
{
    std::vector<d3> vertices;
    std::vector<std::array<uint32_t, 3>> faces;
};
sphere_mesh icosphere(int level)
{
    const double g = (1 + std::sqrt(5.0)) / 2;
    sphere_mesh m;
    m.vertices = {{-1, g, 0}, {1, g, 0}, {-1, -g, 0}, {1, -g, 0}, {0, -1, g}, {0, 1, g},
                  {0, -1, -g}, {0, 1, -g}, {g, 0, -1}, {g, 0, 1}, {-g, 0, -1}, {-g, 0, 1}};
    for (d3& v : m.vertices)
    {
        v = unit(v);
    }
    m.faces = {{0, 11, 5}, {0, 5, 1},  {0, 1, 7},   {0, 7, 10}, {0, 10, 11}, {1, 5, 9}, {5, 11, 4},
               {11, 10, 2}, {10, 7, 6}, {7, 1, 8},   {3, 9, 4},  {3, 4, 2},   {3, 2, 6}, {3, 6, 8},
               {3, 8, 9},  {4, 9, 5},  {2, 4, 11},  {6, 2, 10}, {8, 6, 7},   {9, 8, 1}};
    for (int l = 0; l < level; ++l)
    {
        std::unordered_map<uint64_t, uint32_t> middle;
        const auto mid = [&](uint32_t a, uint32_t b) {
            const uint64_t key = (static_cast<uint64_t>(std::min(a, b)) << 32) | std::max(a, b);
            if (const auto f = middle.find(key); f != middle.end())
            {
                return f->second;
            }
            const auto k = static_cast<uint32_t>(m.vertices.size());
            m.vertices.push_back(unit(0.5 * (m.vertices[a] + m.vertices[b])));
            middle.emplace(key, k);
            return k;
        };
        std::vector<std::array<uint32_t, 3>> next;
        next.reserve(4 * m.faces.size());
        for (const auto& f : m.faces)
        {
            const uint32_t ab = mid(f[0], f[1]), bc = mid(f[1], f[2]), ca = mid(f[2], f[0]);
            next.push_back({f[0], ab, ca});
            next.push_back({f[1], bc, ab});
            next.push_back({f[2], ca, bc});
            next.push_back({ab, bc, ca});
        }
        m.faces = std::move(next);
    }
    return m;
}

// A smooth seeded field on the unit sphere: the mean of `waves` plane waves sin(f (d . p) + phase), directions uniform
// on the sphere, frequencies in [f_lo, f_hi], from std::mt19937_64(seed); values in [-1, 1]
class wave_field
{
public:
    wave_field(uint64_t seed, int waves, double f_lo, double f_hi)
    {
        std::mt19937_64 g(seed);
        std::uniform_real_distribution<double> u(0.0, 1.0);
        for (int k = 0; k < waves; ++k)
        {
            const double z = 2 * u(g) - 1, a = 2 * std::numbers::pi * u(g), s = std::sqrt(1 - z * z);
            waves_.push_back({{s * std::cos(a), s * std::sin(a), z}, f_lo + (f_hi - f_lo) * u(g),
                              2 * std::numbers::pi * u(g)});
        }
    }
    double operator()(d3 p) const
    {
        double sum = 0;
        for (const wave& w : waves_)
        {
            sum += std::sin(w.frequency * dot(w.direction, p) + w.phase);
        }
        return sum / static_cast<double>(waves_.size());
    }

private:
    struct wave
    {
        d3 direction;
        double frequency, phase;
    };
    std::vector<wave> waves_;
};

// An icosphere of radius `radius` around `center`, each vertex moved along its direction by amplitude * field
mesh_def displaced_sphere(int level, const wave_field& field, double amplitude, double radius, d3 center,
                          material_ref material)
{
    const sphere_mesh s = icosphere(level);
    mesh_def m;
    m.material = std::move(material);
    m.positions.reserve(3 * s.vertices.size());
    for (const d3& v : s.vertices)
    {
        const d3 p = center + (radius * (1 + amplitude * field(v))) * v;
        m.positions.insert(m.positions.end(), {static_cast<float>(p.x), static_cast<float>(p.y), static_cast<float>(p.z)});
    }
    m.indices.reserve(3 * s.faces.size());
    for (const auto& f : s.faces)
    {
        m.indices.insert(m.indices.end(), f.begin(), f.end());
    }
    return m;
}

// An equirectangular HDR sky, +y up, row 0 at +y (the environment map's convention): above the horizon a gradient from
// a bright haze to a blue zenith with a faint seeded variation, below it a dim ground, and the sun, a disk of
// `sun_radiance` and angular radius `sun_degrees` at the given elevation and azimuth (degrees; azimuth 0 is +x)
std::shared_ptr<const image_data> sun_sky(int width, int height, uint64_t seed, double elevation, double azimuth,
                                          double sun_degrees, double sun_radiance)
{
    auto img = std::make_shared<image_data>();
    img->width = width;
    img->height = height;
    img->rgb.resize(3 * static_cast<std::size_t>(width) * height);
    const wave_field clouds(seed, 12, 2.0, 6.0);
    const double el = radians(elevation), az = radians(azimuth);
    const d3 sun{std::cos(el) * std::cos(az), std::sin(el), std::cos(el) * std::sin(az)};
    const double cos_disk = std::cos(radians(sun_degrees));
    for (int j = 0; j < height; ++j)
    {
        const double theta = std::numbers::pi * (j + 0.5) / height;
        for (int i = 0; i < width; ++i)
        {
            const double phi = 2 * std::numbers::pi * (i + 0.5) / width - std::numbers::pi;
            const d3 d{std::sin(theta) * std::cos(phi), std::cos(theta), std::sin(theta) * std::sin(phi)};
            d3 c;
            if (d.y >= 0)
            {
                const double up = std::sqrt(d.y);
                const d3 haze{1.1, 1.05, 1.0}, zenith{0.25, 0.45, 0.95};
                c = (1 - up) * haze + up * zenith;
                c = (1 + 0.15 * clouds(d)) * c;
            }
            else
            {
                c = d3{0.18, 0.16, 0.14};
            }
            if (dot(d, sun) >= cos_disk)
            {
                c = d3{sun_radiance, sun_radiance, sun_radiance};
            }
            float* px = img->rgb.data() + 3 * (static_cast<std::size_t>(j) * width + i);
            px[0] = static_cast<float>(c.x);
            px[1] = static_cast<float>(c.y);
            px[2] = static_cast<float>(c.z);
        }
    }
    return img;
}

// A map of radiance 1 everywhere (env-furnace)
std::shared_ptr<const image_data> white_map()
{
    auto img = std::make_shared<image_data>();
    img->width = 16;
    img->height = 8;
    img->rgb.assign(3 * 16 * 8, 1.0f);
    return img;
}

// The furnaces' camera (src/gpu_scenes_materials.ixx:62-75): nearly orthographic from 60 away, a white sky
camera_settings furnace_camera()
{
    camera_settings c;
    c.aspect_ratio = 16.0 / 9.0;
    c.image_width = 480;
    c.samples_per_pixel = 256;
    c.max_depth = 10;
    c.sampling = RT_SAMPLING_MATERIAL;
    c.vfov = 6;
    c.lookfrom = {0, 0, 60};
    c.lookat = {0, 0, 0};
    c.sky_bottom = {1, 1, 1};
    c.sky_top = {1, 1, 1};
    return c;
}

// A glTF model in a studio (RayTracingGPU's FILE.glb / FILE.gltf scene, src/main_gpu.cpp:141-143): its triangle meshes
// placed by its nodes, on a gray ground under the seeded sky with a sun (the triangle scene's), framed by a camera that
// looks at the model's bounds from the front, the right and above. Materials by their factors: metallic ones (>= 0.5) a
// GGX conductor of the base color and roughness, the others a glossy coat over the base color, or plain lambertian at
// roughness 1. 1280 wide (16:9), 64 samples, depth 16, light samples weighed by MIS.
built build_model(const std::string& path, std::vector<std::string>& warnings)
{
    const gltf_model model = load_gltf(path);
    warnings = model.warnings;
    if (model.placements.empty())
    {
        std::string why = path + ": the model has no triangles to render";
        for (const std::string& w : model.warnings)
        {
            why += "; " + w;
        }
        throw std::invalid_argument(why);
    }
    built s;
    std::vector<material_ref> materials;
    for (const gltf_material& m : model.materials)
    {
        const d3 base{m.base_color[0], m.base_color[1], m.base_color[2]};
        materials.push_back(m.metallic >= 0.5  ? rough_metal(base, m.roughness)
                            : m.roughness < 1 ? glossy(base, m.roughness)
                                              : lambertian(base));
    }
    const material_ref fallback = rough_metal({1, 1, 1}, 1.0); // glTF's default material: metallic 1, roughness 1
    for (const gltf_mesh& m : model.meshes)
    {
        s.meshes.push_back({m.positions, m.indices, m.material >= 0 ? materials[static_cast<std::size_t>(m.material)] : fallback});
    }
    // the bounds of every placed vertex
    d3 lo{1e300, 1e300, 1e300}, hi{-1e300, -1e300, -1e300};
    for (const gltf_placement& p : model.placements)
    {
        rt_mesh_instance_desc in{};
        in.mesh = p.mesh;
        std::copy(p.object_to_world, p.object_to_world + 12, in.object_to_world);
        s.instances.push_back(in);
        const std::vector<float>& pos = model.meshes[static_cast<std::size_t>(p.mesh)].positions;
        const double* m = p.object_to_world;
        for (std::size_t k = 0; k + 2 < pos.size(); k += 3)
        {
            const double x = pos[k], y = pos[k + 1], z = pos[k + 2];
            const d3 w{m[0] * x + m[1] * y + m[2] * z + m[3], m[4] * x + m[5] * y + m[6] * z + m[7],
                       m[8] * x + m[9] * y + m[10] * z + m[11]};
            lo = {std::min(lo.x, w.x), std::min(lo.y, w.y), std::min(lo.z, w.z)};
            hi = {std::max(hi.x, w.x), std::max(hi.y, w.y), std::max(hi.z, w.z)};
        }
    }
    const d3 center = 0.5 * (lo + hi);
    const double radius = std::max(0.5 * length(hi - lo), 1e-6);
    const double ground = 40 * radius;
    s.world.add(make_quad({center.x - ground / 2, lo.y, center.z - ground / 2}, {0, 0, ground}, {ground, 0, 0},
                          lambertian(d3{0.5, 0.5, 0.5})));
    s.camera.aspect_ratio = 16.0 / 9.0;
    s.camera.image_width = 1280;
    s.camera.samples_per_pixel = 64;
    s.camera.max_depth = 16;
    s.camera.sampling = RT_SAMPLING_MATERIAL;
    s.camera.vfov = 35;
    s.camera.lookat = center;
    const d3 view{0.55, 0.35, 1.0};
    s.camera.lookfrom = center + (3.2 * radius / length(view)) * view;
    s.environment = environment_def{sun_sky(1024, 512, 7, 35, 60, 0.6, 10000), 0, 1};
    s.combination = RT_COMBINE_POWER_HEURISTIC;
    return s;
}

built build_named(const std::string& name, scene_random& rng, const scene_options& opts)
{
    built s;
    const auto earth_image = [&] {
        const int w = opts.earth_width > 0 ? opts.earth_width : 2048;
        image_pixels img = noise_image(w, w / 2, opts.earth_seed);
        return std::make_shared<const image_data>(image_data{img.width, img.height, std::move(img.rgb)});
    };
    if (name == "materials")
    {
        s.world.add(make_sphere({0.0, -100.5, -1.0}, 100.0, lambertian(d3{0.8, 0.8, 0.0})));
        s.world.add(make_sphere({0.0, 0.0, -1.2}, 0.5, lambertian(d3{0.1, 0.2, 0.5})));
        s.world.add(make_sphere({-1.0, 0.0, -1.0}, 0.5, dielectric(1.50)));
        s.world.add(make_sphere({-1.0, 0.0, -1.0}, 0.4, dielectric(1.00 / 1.50)));
        s.world.add(make_sphere({1.0, 0.0, -1.0}, 0.5, metal({0.8, 0.6, 0.2}, 1.0)));
        s.camera = cover_camera();
        s.camera.lookfrom = {-2, 2, 1};
        s.camera.lookat = {0, 0, -1};
        s.camera.defocus_angle = 10.0;
        s.camera.focus_dist = 3.4;
    }
    else if (name == "weekend" || name == "bouncing")
    {
        s.world = random_spheres(rng, name == "bouncing");
        s.camera = cover_camera();
    }
    else if (name == "checker")
    {
        const auto m = lambertian(checker(0.32, {0.2, 0.3, 0.1}, {0.9, 0.9, 0.9}));
        s.world.add(make_sphere({0, -10, 0}, 10, m));
        s.world.add(make_sphere({0, 10, 0}, 10, m));
        s.camera = cover_camera();
        s.camera.defocus_angle = 0;
    }
    else if (name == "earth")
    {
        s.world.add(make_sphere({0, 0, 0}, 2, lambertian(picture(earth_image()))));
        s.camera = cover_camera();
        s.camera.lookfrom = {0, 0, 12};
        s.camera.defocus_angle = 0;
    }
    else if (name == "perlin")
    {
        const auto m = lambertian(marble(rng, 4));
        s.world.add(make_sphere({0, -1000, 0}, 1000, m));
        s.world.add(make_sphere({0, 2, 0}, 2, m));
        s.camera = cover_camera();
        s.camera.defocus_angle = 0;
    }
    else if (name == "quads")
    {
        s.world.add(make_quad({-3, -2, 5}, {0, 0, -4}, {0, 4, 0}, lambertian(d3{1.0, 0.2, 0.2})));
        s.world.add(make_quad({-2, -2, 0}, {4, 0, 0}, {0, 4, 0}, lambertian(d3{0.2, 1.0, 0.2})));
        s.world.add(make_quad({3, -2, 1}, {0, 0, 4}, {0, 4, 0}, lambertian(d3{0.2, 0.2, 1.0})));
        s.world.add(make_quad({-2, 3, 1}, {4, 0, 0}, {0, 0, 4}, lambertian(d3{1.0, 0.5, 0.0})));
        s.world.add(make_quad({-2, -3, 5}, {4, 0, 0}, {0, 0, -4}, lambertian(d3{0.2, 0.8, 0.8})));
        s.camera.image_width = 1080;
        s.camera.samples_per_pixel = 500;
        s.camera.max_depth = 50;
        s.camera.vfov = 80;
        s.camera.lookfrom = {0, 0, 9};
        s.camera.lookat = {0, 0, 0};
        s.camera.sky_bottom = daylight_bottom;
        s.camera.sky_top = daylight_top;
    }
    else if (name == "light")
    {
        const auto m = lambertian(marble(rng, 4));
        s.world.add(make_sphere({0, -1000, 0}, 1000, m));
        s.world.add(make_sphere({0, 2, 0}, 2, m));
        const auto glow = lamp_light({4, 4, 4});
        s.world.add(make_sphere({0, 7, 0}, 2, glow));
        s.world.add(make_quad({3, 1, -2}, {2, 0, 0}, {0, 2, 0}, glow));
        s.lights.add(make_quad({3, 1, -2}, {2, 0, 0}, {0, 2, 0}, nullptr));
        s.lights.add(make_sphere({0, 7, 0}, 2, nullptr));
        s.camera.aspect_ratio = 16.0 / 9.0;
        s.camera.image_width = 1920;
        s.camera.samples_per_pixel = 1000;
        s.camera.max_depth = 50;
        s.camera.vfov = 20;
        s.camera.lookfrom = {26, 3, 6};
        s.camera.lookat = {0, 2, 0};
    }
    else if (name == "cornell" || name == "aluminum" || name == "glass")
    {
        cornell_walls(s.world);
        s.world.add(cornell_lamp(lamp_light({15, 15, 15})));
        if (name == "cornell")
        {
            const auto white = lambertian(d3{0.73, 0.73, 0.73});
            s.world.add(tall_block(white));
            s.world.add(short_block(white));
            s.lights.add(cornell_lamp(nullptr));
        }
        else if (name == "aluminum")
        {
            s.world.add(tall_block(metal({0.8, 0.85, 0.88}, 0.0)));
            s.world.add(short_block(lambertian(d3{0.73, 0.73, 0.73})));
            s.lights.add(cornell_lamp(nullptr));
        }
        else
        {
            s.world.add(tall_block(lambertian(d3{0.73, 0.73, 0.73})));
            s.world.add(make_sphere({190, 90, 190}, 90, dielectric(1.5)));
            s.lights.add(cornell_lamp(nullptr));
            s.lights.add(make_sphere({190, 90, 190}, 90, nullptr));
        }
        s.camera = cornell_camera();
    }
    else if (name == "smoke")
    {
        cornell_walls(s.world);
        const auto lamp = [](material_ref m) { return make_quad({113, 554, 127}, {330, 0, 0}, {0, 0, 305}, std::move(m)); };
        s.world.add(lamp(lamp_light({7, 7, 7})));
        const auto white = lambertian(d3{0.73, 0.73, 0.73});
        s.world.add(make_medium(tall_block(white), 0.01, {0, 0, 0}));
        s.world.add(make_medium(short_block(white), 0.01, {1, 1, 1}));
        s.lights.add(lamp(nullptr));
        s.camera = cornell_camera();
    }
    else if (name == "final")
    {
        list_builder ground_boxes;
        const auto ground = lambertian(d3{0.48, 0.83, 0.53});
        for (int i = 0; i < 20; ++i)
        {
            for (int j = 0; j < 20; ++j)
            {
                const double w = 100.0;
                const d3 corner0{-1000.0 + i * w, 0.0, -1000.0 + j * w};
                const d3 corner1{corner0.x + w, rng.uniform(1, 101), corner0.z + w};
                ground_boxes.add(make_box(corner0, corner1, ground));
            }
        }
        s.world.add(make_tree(std::move(ground_boxes)));
        const auto lamp = [](material_ref m) { return make_quad({123, 554, 147}, {300, 0, 0}, {0, 0, 265}, std::move(m)); };
        s.world.add(lamp(lamp_light({7, 7, 7})));
        const d3 center1{400, 400, 200};
        s.world.add(make_sphere(center1, center1 + d3{30, 0, 0}, 50, lambertian(d3{0.7, 0.3, 0.1})));
        s.world.add(make_sphere({260, 150, 45}, 50, dielectric(1.5)));
        s.world.add(make_sphere({0, 150, 145}, 50, metal({0.8, 0.8, 0.9}, 1.0)));
        const auto glass = dielectric(1.5);
        s.world.add(make_sphere({360, 150, 145}, 70, glass));
        s.world.add(make_medium(make_sphere({360, 150, 145}, 70, glass), 0.2, {0.2, 0.4, 0.9}));
        s.world.add(make_medium(make_sphere({0, 0, 0}, 5000, glass), 0.0001, {1, 1, 1}));
        s.world.add(make_sphere({400, 200, 400}, 100, lambertian(picture(earth_image()))));
        s.world.add(make_sphere({220, 280, 300}, 80, lambertian(marble(rng, 0.2))));
        list_builder cloud;
        const auto white = lambertian(d3{0.73, 0.73, 0.73});
        for (int i = 0; i < 1000; ++i)
        {
            cloud.add(make_sphere(rng.vector(0, 165), 10, white));
        }
        s.world.add(make_translate(make_rotate_y(make_tree(std::move(cloud)), 15), {-100, 270, 395}));
        s.lights.add(lamp(nullptr));
        s.camera.image_width = 1080;
        s.camera.samples_per_pixel = 1000;
        s.camera.max_depth = 40;
        s.camera.vfov = 40;
        s.camera.lookfrom = {478, 278, -600};
        s.camera.lookat = {278, 278, 0};
    }
    else if (name == "furnace")
    {
        // A white matte sphere under a white sky (src/gpu_scenes.ixx:33-55): every pixel is exactly 1
        s.world.add(make_sphere({0, 0, 0}, 1.0, lambertian(d3{1, 1, 1})));
        s.camera.aspect_ratio = 16.0 / 9.0;
        s.camera.image_width = 480;
        s.camera.samples_per_pixel = 64;
        s.camera.max_depth = 10;
        s.camera.sampling = RT_SAMPLING_MATERIAL;
        s.camera.vfov = 30;
        s.camera.lookfrom = {0, 0, 6};
        s.camera.lookat = {0, 0, 0};
        s.camera.sky_bottom = {1, 1, 1};
        s.camera.sky_top = {1, 1, 1};
    }
    else if (name == "ggx-furnace")
    {
        // Five white metals (F0 = 1) under a white sky, the mirror and GGX at alpha = r^2 = 0.05, 0.3, 0.6, 1
        // (src/gpu_scenes_materials.ixx:86-102): each sphere's centre shows the light its lobe keeps, 1 with
        // Turquin's compensation
        const std::array<double, 5> roughness{0.0, std::sqrt(0.05), std::sqrt(0.3), std::sqrt(0.6), 1.0};
        for (std::size_t i = 0; i < roughness.size(); ++i)
        {
            s.world.add(make_sphere({-4.0 + 2.0 * static_cast<double>(i), 0, 0}, 0.4,
                                    rough_metal({1, 1, 1}, roughness[i])));
        }
        s.camera = furnace_camera();
    }
    else if (name == "gloss-furnace")
    {
        // Four white matte spheres under glossy coats at alpha = 0.01, 0.1, 0.3, 1 (src/gpu_scenes_materials.ixx:104-116)
        const std::array<double, 4> roughness{std::sqrt(0.01), std::sqrt(0.1), std::sqrt(0.3), 1.0};
        for (std::size_t i = 0; i < roughness.size(); ++i)
        {
            s.world.add(make_sphere({-3.0 + 2.0 * static_cast<double>(i), 0, 0}, 0.4, glossy({1, 1, 1}, roughness[i])));
        }
        s.camera = furnace_camera();
    }
    else if (name == "env-furnace")
    {
        // furnace's sphere and camera under a map of radiance 1 and a black sky
        // (src/gpu_scenes_environments.ixx:102-118): every pixel is 1
        s.world.add(make_sphere({0, 0, 0}, 1.0, lambertian(d3{1, 1, 1})));
        s.camera.aspect_ratio = 16.0 / 9.0;
        s.camera.image_width = 480;
        s.camera.samples_per_pixel = 64;
        s.camera.max_depth = 10;
        s.camera.sampling = RT_SAMPLING_MATERIAL;
        s.camera.vfov = 30;
        s.camera.lookfrom = {0, 0, 6};
        s.camera.lookat = {0, 0, 0};
        s.environment = environment_def{white_map(), 0, 1};
    }
    else if (name == "cornell-ggx")
    {
        // The Cornell box with a tall block of rough aluminum (r = 0.3) and a short one glossy white (r = 0.2)
        // (src/gpu_scenes_materials.ixx:181-204): -p nee and -p mis converge to -p material's image
        cornell_walls(s.world);
        s.world.add(cornell_lamp(lamp_light({15, 15, 15})));
        s.world.add(tall_block(rough_metal({0.913, 0.921, 0.925}, 0.3)));
        s.world.add(short_block(glossy({0.73, 0.73, 0.73}, 0.2)));
        s.camera = cornell_camera();
        s.camera.samples_per_pixel = 1024;
        s.camera.sampling = RT_SAMPLING_MATERIAL;
    }
    else if (name == "veach")
    {
        // Veach's four plates under four sphere lamps of one power (src/gpu_scenes_materials.ixx:237-270)
        const d3 eye{0, 2, 15}, lamp_row{0, 2.5, -3.5};
        const std::array<std::pair<d3, double>, 4> plates{
            {{{0, 0.4, -2.4}, 0.05}, {{0, -0.5, -0.9}, 0.1}, {{0, -1.2, 0.7}, 0.2}, {{0, -1.7, 2.2}, 0.35}}};
        for (const auto& [centre, roughness] : plates)
        {
            const d3 n = unit(unit(eye - centre) + unit(lamp_row - centre));
            const d3 u{9, 0, 0};
            const d3 v = unit(cross(n, d3{1, 0, 0}));
            s.world.add(make_quad(centre - 0.5 * u - 0.5 * v, u, v, rough_metal({0.95, 0.93, 0.88}, roughness)));
        }
        const std::array<std::pair<double, d3>, 4> lamps{
            {{0.05, {1.0, 0.5, 0.2}}, {0.15, {1.0, 0.8, 0.55}}, {0.4, {0.8, 0.9, 1.0}}, {1.0, {0.5, 0.7, 1.0}}}};
        for (std::size_t i = 0; i < lamps.size(); ++i)
        {
            const auto& [radius, tint] = lamps[i];
            const double y = 0.2126 * tint.x + 0.7152 * tint.y + 0.0722 * tint.z;
            s.world.add(make_sphere({-3.75 + 2.5 * static_cast<double>(i), 2.5, -3.5}, radius,
                                    lamp_light((4 / (radius * radius) / y) * tint)));
        }
        s.world.add(make_quad({-15, -10, -6}, {30, 0, 0}, {0, 25, 0}, lambertian(d3{0.1, 0.1, 0.1})));
        s.camera.aspect_ratio = 16.0 / 9.0;
        s.camera.image_width = 1920;
        s.camera.samples_per_pixel = 1024;
        s.camera.max_depth = 10;
        s.camera.sampling = RT_SAMPLING_MATERIAL;
        s.camera.vfov = 28;
        s.camera.lookfrom = eye;
        s.camera.lookat = {0, 0.4, 0};
        s.combination = RT_COMBINE_POWER_HEURISTIC;
    }
    else if (name == "sun-test")
    {
        // A grey floor and three matte spheres (albedo 0.8, 0.5, 0.2) in sunlight (src/gpu_scenes_environments.ixx:
        // 172-190), under this library's seeded sun sky (the reference's is a photographed HDR asset)
        s.world.add(make_quad({-100, 0, -100}, {0, 0, 200}, {200, 0, 0}, lambertian(d3{0.5, 0.5, 0.5})));
        const std::array<std::pair<double, double>, 3> balls{{{-2.2, 0.8}, {0, 0.5}, {2.2, 0.2}}};
        for (const auto& [x, albedo] : balls)
        {
            s.world.add(make_sphere({x, 1, 0}, 1.0, lambertian(d3{albedo, albedo, albedo})));
        }
        s.camera.aspect_ratio = 16.0 / 9.0;
        s.camera.image_width = 480;
        s.camera.samples_per_pixel = 1024;
        s.camera.max_depth = 10;
        s.camera.sampling = RT_SAMPLING_MATERIAL;
        s.camera.vfov = 30;
        s.camera.lookfrom = {0, 2, 9};
        s.camera.lookat = {0, 0.9, 0};
        s.environment = environment_def{sun_sky(1024, 512, 48, 48, 30, 0.6, 10000), 0, 1};
        s.combination = RT_COMBINE_POWER_HEURISTIC;
    }
    else if (name == "mesh-furnace")
    {
        // Two icospheres of 5120 triangles under a white sky, one white matte, one a white mirror
        // (src/gpu_scenes.ixx:84-85): every path ends in the sky with its throughput 1, so every pixel is exactly 1
        s.meshes.push_back(displaced_sphere(4, wave_field(1, 1, 1.0, 1.0), 0.0, 1.0, {-1.2, 0, 0},
                                            lambertian(d3{1, 1, 1})));
        s.meshes.push_back(displaced_sphere(4, wave_field(1, 1, 1.0, 1.0), 0.0, 1.0, {1.2, 0, 0}, metal({1, 1, 1}, 0)));
        s.camera.aspect_ratio = 16.0 / 9.0;
        s.camera.image_width = 480;
        s.camera.samples_per_pixel = 64;
        s.camera.max_depth = 50;
        s.camera.sampling = RT_SAMPLING_MATERIAL;
        s.camera.vfov = 30;
        s.camera.lookfrom = {0, 0, 6};
        s.camera.lookat = {0, 0, 0};
        s.camera.sky_bottom = {1, 1, 1};
        s.camera.sky_top = {1, 1, 1};
    }
    else if (name == "triangles")
    {
        // The triangle-heavy scene: a seeded displaced icosphere of 1,310,720 triangles (level 8) in a GGX metal of
        // roughness U[0.05, 0.6], 64 instances of a seeded 20,480-triangle blob (level 5) in matte and glossy-coat
        // materials, on a ground quad, under a seeded 1024x512 HDR sky with a sun disk of radiance 10^4, lit by the
        // map's light samples weighed by MIS. 1920x1080, 256 samples, depth 16.
        std::mt19937_64 g(20250603);
        std::uniform_real_distribution<double> u(0.0, 1.0);
        const double metal_roughness = 0.05 + 0.55 * u(g);
        s.world.add(make_quad({-20, 0, -20}, {0, 0, 40}, {40, 0, 0}, lambertian(d3{0.5, 0.5, 0.5})));
        s.meshes.push_back(displaced_sphere(8, wave_field(11, 24, 3.0, 12.0), 0.05, 1.6, {0, 1.75, 0},
                                            rough_metal({0.955, 0.637, 0.538}, metal_roughness)));
        const wave_field blob_field(12, 10, 2.0, 7.0);
        const d3 matte_color{0.2 + 0.6 * u(g), 0.2 + 0.6 * u(g), 0.2 + 0.6 * u(g)};
        const d3 coat_color{0.1 + 0.6 * u(g), 0.1 + 0.6 * u(g), 0.1 + 0.6 * u(g)};
        const double coat_roughness = 0.1 + 0.4 * u(g);
        s.meshes.push_back(displaced_sphere(5, blob_field, 0.25, 1.0, {0, 0, 0}, lambertian(matte_color)));
        s.meshes.push_back(displaced_sphere(5, blob_field, 0.25, 1.0, {0, 0, 0}, glossy(coat_color, coat_roughness)));
        for (int k = 0; k < 64; ++k)
        {
            // an 8 x 8 grid of cells 2.4 wide, each blob jittered in its cell; the four middle cells move out of the
            // large sphere's way
            const int row = k / 8, column = k % 8;
            const double scale = 0.35 + 0.25 * u(g);
            double x = -8.4 + 2.4 * column + 0.6 * (u(g) - 0.5);
            double z = -8.4 + 2.4 * row + 0.6 * (u(g) - 0.5);
            if (std::fabs(x) < 2.6 && std::fabs(z) < 2.6)
            {
                x *= 2.6;
                z *= 2.6;
            }
            const double yaw = 2 * std::numbers::pi * u(g);
            rt_mesh_instance_desc in{};
            in.mesh = 1 + (k & 1);
            const double c = scale * std::cos(yaw), sn = scale * std::sin(yaw);
            const double m[12] = {c, 0, sn, x, 0, scale, 0, 1.25 * scale, -sn, 0, c, z};
            std::copy(m, m + 12, in.object_to_world);
            s.instances.push_back(in);
        }
        rt_mesh_instance_desc big{};
        big.mesh = 0;
        const double identity[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
        std::copy(identity, identity + 12, big.object_to_world);
        s.instances.push_back(big);
        s.camera.aspect_ratio = 16.0 / 9.0;
        s.camera.image_width = 1920;
        s.camera.samples_per_pixel = 256;
        s.camera.max_depth = 16;
        s.camera.sampling = RT_SAMPLING_MATERIAL;
        s.camera.vfov = 40;
        s.camera.lookfrom = {0, 5.5, 15};
        s.camera.lookat = {0, 1.2, 0};
        s.environment = environment_def{sun_sky(1024, 512, 7, 35, 60, 0.6, 10000), 0, 1};
        s.combination = RT_COMBINE_POWER_HEURISTIC;
    }
    else
    {
        throw std::invalid_argument("unknown scene '" + name + "'");
    }
    return s;
}

// --- flattening (src/gpu_scene.ixx:97-768) ---------------------------------------------------------------------------
// Where an object inside instances ends up: a turn about y, then a shift. Nested instances add up to one such motion.
struct placement
{
    double cos_theta = 1, sin_theta = 0;
    d3 offset{0, 0, 0};
    d3 turn(d3 v) const { return {cos_theta * v.x + sin_theta * v.z, v.y, -sin_theta * v.x + cos_theta * v.z}; }
    d3 place(d3 p) const { return turn(p) + offset; }
    placement inside(const shifted_object& s) const { return {cos_theta, sin_theta, place(s.offset)}; }
    placement inside(const turned_object& t) const
    {
        return {cos_theta * t.cos_theta - sin_theta * t.sin_theta, sin_theta * t.cos_theta + cos_theta * t.sin_theta,
                offset};
    }
};

// Visits every leaf object (a shape or a medium) under `o`, with the placement that brings it into the world
template <typename F>
void walk(const object& o, const placement& where, F&& visit)
{
    if (const auto* g = std::get_if<group>(&o.what))
    {
        for (const auto& m : g->members)
        {
            walk(*m, where, visit);
        }
    }
    else if (const auto* n = std::get_if<tree_node>(&o.what))
    {
        walk(*n->left, where, visit);
        walk(*n->right, where, visit);
    }
    else if (const auto* s = std::get_if<shifted_object>(&o.what))
    {
        walk(*s->inner, where.inside(*s), visit);
    }
    else if (const auto* t = std::get_if<turned_object>(&o.what))
    {
        walk(*t->inner, where.inside(*t), visit);
    }
    else
    {
        visit(o, where);
    }
}

rt_sphere_desc sphere_record(const sphere_shape& s, const placement& where, int material)
{
    rt_sphere_desc d{};
    d.center = to_rt(where.place(s.center0 + 0.0 * s.motion));
    d.radius = s.radius;
    // center_at(1) - center_at(0): exactly 0 for a sphere that stands still
    d.displacement = to_rt(where.turn((s.center0 + 1.0 * s.motion) - (s.center0 + 0.0 * s.motion)));
    d.uv_cos = where.cos_theta;
    d.uv_sin = where.sin_theta;
    d.material = material;
    return d;
}
rt_quad_desc quad_record(const quad_shape& q, const placement& where, int material)
{
    rt_quad_desc d{};
    d.q = to_rt(where.place(q.q));
    d.u = to_rt(where.turn(q.u));
    d.v = to_rt(where.turn(q.v));
    d.material = material;
    return d;
}

class flattener
{
public:
    explicit flattener(scene_tables& out) : out_(out) {}

    void add_world(const object& world)
    {
        walk(world, placement{}, [this](const object& o, const placement& where) { add_leaf(o, where); });
    }

    void add_lights(const object& o, const placement& where, double weight)
    {
        if (const auto* g = std::get_if<group>(&o.what))
        {
            if (g->members.empty())
            {
                throw std::runtime_error("the lights list holds an empty list");
            }
            for (const auto& m : g->members)
            {
                add_lights(*m, where, weight / static_cast<double>(g->members.size()));
            }
        }
        else if (const auto* s = std::get_if<shifted_object>(&o.what))
        {
            add_lights(*s->inner, where.inside(*s), weight);
        }
        else if (const auto* t = std::get_if<turned_object>(&o.what))
        {
            add_lights(*t->inner, where.inside(*t), weight);
        }
        else if (const auto* q = std::get_if<quad_shape>(&o.what))
        {
            const d3 n = cross(q->u, q->v);
            if (!(dot(n, n) > 0))
            {
                throw std::runtime_error("the lights list holds a quad without area");
            }
            rt_light_desc l{};
            l.kind = RT_LIGHT_QUAD;
            l.quad = quad_record(*q, where, 0);
            l.weight = weight;
            out_.lights.push_back(l);
        }
        else if (const auto* b = std::get_if<sphere_shape>(&o.what))
        {
            if (!(b->radius > 0))
            {
                throw std::runtime_error("the lights list holds a sphere without size");
            }
            rt_light_desc l{};
            l.kind = RT_LIGHT_SPHERE;
            l.center = to_rt(where.place(b->center0 + 0.0 * b->motion));
            l.radius = b->radius;
            l.weight = weight;
            out_.lights.push_back(l);
        }
        else
        {
            throw std::runtime_error("the lights list holds an object that can't be aimed at");
        }
    }

    // A mesh's triangles keep their vertices; its material joins the table like any other
    void add_mesh(const mesh_def& m)
    {
        rt_mesh_desc d{};
        d.material = material_index(m.material.get());
        out_.mesh_positions.push_back(m.positions);
        out_.mesh_indices.push_back(m.indices);
        out_.meshes.push_back(d);
    }

    // The GGX side entries, one per material, when some material has one (src/gpu_scene.ixx:700-740)
    void finish_microfacets()
    {
        const bool any = std::any_of(sides_.begin(), sides_.end(),
                                     [](const rt_microfacet_desc& s) { return s.kind != RT_MICROFACET_NONE; });
        out_.microfacets = any ? sides_ : std::vector<rt_microfacet_desc>{};
    }

private:
    scene_tables& out_;
    std::unordered_map<const material_def*, int> materials_;
    std::unordered_map<const texture_def*, int> textures_;
    std::vector<rt_microfacet_desc> sides_;

    void add_leaf(const object& o, const placement& where)
    {
        if (const auto* s = std::get_if<sphere_shape>(&o.what))
        {
            const int m = material_index(s->material.get());
            out_.spheres.push_back(sphere_record(*s, where, m));
        }
        else if (const auto* q = std::get_if<quad_shape>(&o.what))
        {
            const int m = material_index(q->material.get());
            out_.quads.push_back(quad_record(*q, where, m));
        }
        else if (const auto* v = std::get_if<medium_object>(&o.what))
        {
            add_medium(*v, where);
        }
        else
        {
            throw std::runtime_error("an object the renderer doesn't know");
        }
    }
    void add_medium(const medium_object& v, const placement& where)
    {
        rt_medium_desc m{};
        m.material = material_index(v.phase.get());
        m.neg_inv_density = v.neg_inv_density;
        m.sphere.uv_cos = 1;
        int spheres = 0;
        std::vector<rt_quad_desc> quads;
        walk(*v.boundary, where, [&](const object& part, const placement& at) {
            if (const auto* b = std::get_if<sphere_shape>(&part.what))
            {
                ++spheres;
                m.sphere = sphere_record(*b, at, 0);
                m.sphere.uv_cos = 1;
                m.sphere.uv_sin = 0;
            }
            else if (const auto* q = std::get_if<quad_shape>(&part.what))
            {
                quads.push_back(quad_record(*q, at, 0));
            }
            else
            {
                throw std::runtime_error("a medium's boundary holds something other than spheres and quads");
            }
        });
        if (spheres == 1 && quads.empty())
        {
            m.boundary = RT_BOUNDARY_SPHERE;
        }
        else if (spheres == 0 && !quads.empty() && quads_are_convex(quads.data(), quads.size()))
        {
            m.boundary = RT_BOUNDARY_CONVEX_QUADS;
        }
        else
        {
            throw std::runtime_error("a medium's boundary isn't one sphere or a convex shape made of quads");
        }
        out_.medium_quads.push_back(std::move(quads));
        out_.media.push_back(m);
    }
    int material_index(const material_def* m)
    {
        if (!m)
        {
            throw std::runtime_error("an object of the world has no material");
        }
        if (const auto f = materials_.find(m); f != materials_.end())
        {
            return f->second;
        }
        rt_material_desc d{};
        d.kind = m->kind;
        d.texture = -1;
        d.refraction_index = 1;
        switch (m->kind)
        {
        case RT_MATERIAL_METAL:
            d.albedo = to_rt(m->albedo);
            d.fuzz = m->fuzz;
            break;
        case RT_MATERIAL_DIELECTRIC:
            d.refraction_index = m->refraction_index;
            break;
        default:
            d.texture = texture_index(m->texture.get());
            break;
        }
        const int index = static_cast<int>(out_.materials.size());
        out_.materials.push_back(d);
        sides_.push_back({m->microfacet, 0, m->roughness});
        materials_.emplace(m, index);
        return index;
    }
    int texture_index(const texture_def* t)
    {
        if (const auto f = textures_.find(t); f != textures_.end())
        {
            return f->second;
        }
        rt_texture_desc d{};
        d.kind = t->kind;
        d.even = d.odd = d.image = d.perlin = -1;
        d.inv_scale = 1;
        d.scale = 1;
        switch (t->kind)
        {
        case RT_TEXTURE_SOLID:
            d.rgb = to_rt(t->rgb);
            break;
        case RT_TEXTURE_CHECKER:
            d.even = texture_index(t->even.get()); // the two textures first: smaller indices than the checker
            d.odd = texture_index(t->odd.get());
            d.inv_scale = t->inv_scale;
            break;
        case RT_TEXTURE_IMAGE:
            d.image = static_cast<int>(out_.images.size());
            out_.image_pixels.push_back(t->image->rgb);
            out_.images.push_back({t->image->width, t->image->height, nullptr});
            break;
        case RT_TEXTURE_NOISE:
        {
            d.scale = t->scale;
            d.perlin = static_cast<int>(out_.perlins.size());
            rt_perlin_desc p{};
            for (int i = 0; i < 256; ++i)
            {
                p.random_vectors[i] = to_rt(t->perlin->gradients[i]);
                p.perm_x[i] = t->perlin->perm_x[i];
                p.perm_y[i] = t->perlin->perm_y[i];
                p.perm_z[i] = t->perlin->perm_z[i];
            }
            out_.perlins.push_back(p);
            break;
        }
        }
        const int index = static_cast<int>(out_.textures.size());
        out_.textures.push_back(d);
        textures_.emplace(t, index);
        return index;
    }
};

// The camera's frame (src/camera.ixx:112-160): the pixel grid on the focus plane and the lens disk
rt_camera_desc frame_of(const camera_settings& s, int sampling)
{
    rt_camera_desc c{};
    c.image_width = s.image_width;
    c.image_height = std::max(1, static_cast<int>(s.image_width / s.aspect_ratio));
    c.sqrt_spp = std::max(1, static_cast<int>(std::lround(std::sqrt(s.samples_per_pixel))));
    c.max_depth = s.max_depth;
    const d3 center = s.lookfrom;
    const double h = std::tan(radians(s.vfov) / 2);
    const double viewport_height = 2 * h * s.focus_dist;
    const double viewport_width = viewport_height * (static_cast<double>(c.image_width) / c.image_height);
    const d3 w = unit(s.lookfrom - s.lookat);
    const d3 u = unit(cross(s.vup, w));
    const d3 v = cross(w, u);
    const d3 viewport_u = viewport_width * u;
    const d3 viewport_v = viewport_height * neg(v);
    const d3 du = viewport_u / c.image_width;
    const d3 dv = viewport_v / c.image_height;
    const d3 upper_left = center - s.focus_dist * w - viewport_u / 2 - viewport_v / 2;
    const d3 pixel00 = upper_left + 0.5 * (du + dv);
    const double defocus_radius = s.focus_dist * std::tan(radians(s.defocus_angle / 2));
    c.center = to_rt(center);
    c.pixel00_loc = to_rt(pixel00);
    c.pixel_delta_u = to_rt(du);
    c.pixel_delta_v = to_rt(dv);
    c.defocus = s.defocus_angle > 0 ? 1 : 0;
    c.defocus_disk_u = to_rt(defocus_radius * u);
    c.defocus_disk_v = to_rt(defocus_radius * v);
    c.sky_bottom = to_rt(s.sky_bottom);
    c.sky_top = to_rt(s.sky_top);
    c.sampling = sampling;
    return c;
}
}

image_pixels noise_image(int width, int height, std::uint32_t seed)
{
    // Texel k (rows from the top) is one number of a PCG32 stream: red its bits 0-7, green 8-15, blue 16-23, each byte
    // b taken to the linear value (b/255)^2.2 in double, rounded to float: the stand-in for earthmap.jpg
    std::array<float, 256> lut{};
    for (int b = 0; b < 256; ++b)
    {
        lut[b] = static_cast<float>(std::pow(b / 255.0, 2.2));
    }
    const std::uint64_t inc = (0xda3e39cb94b95bdbull << 1u) | 1u;
    std::uint64_t state = 0;
    auto next = [&]() {
        const std::uint64_t old = state;
        state = old * 6364136223846793005ull + inc;
        const auto xs = static_cast<std::uint32_t>(((old >> 18u) ^ old) >> 27u);
        const auto rot = static_cast<std::uint32_t>(old >> 59u);
        return (xs >> rot) | (xs << ((32u - rot) & 31u));
    };
    next();
    state += 0x853c49e6748fea9bull ^ seed;
    next();
    image_pixels img{width, height, std::vector<float>(3ull * width * height)};
    for (std::size_t k = 0; k < static_cast<std::size_t>(width) * height; ++k)
    {
        const std::uint32_t v = next();
        img.rgb[3 * k] = lut[v & 255u];
        img.rgb[3 * k + 1] = lut[(v >> 8) & 255u];
        img.rgb[3 * k + 2] = lut[(v >> 16) & 255u];
    }
    return img;
}

bool quads_are_convex(const rt_quad_desc* quads, std::size_t count)
{
    // Every corner of every quad lies on one side of each quad's plane, or on it within 1e-9 of the largest coordinate
    // (src/gpu/renderer.h:163-168)
    std::vector<d3> corners;
    double largest = 0;
    for (std::size_t i = 0; i < count; ++i)
    {
        const d3 q{quads[i].q.x, quads[i].q.y, quads[i].q.z}, u{quads[i].u.x, quads[i].u.y, quads[i].u.z},
            v{quads[i].v.x, quads[i].v.y, quads[i].v.z};
        for (const d3& c : {q, q + u, q + v, q + u + v})
        {
            corners.push_back(c);
            largest = std::max({largest, std::fabs(c.x), std::fabs(c.y), std::fabs(c.z)});
        }
    }
    const double tolerance = 1e-9 * (1 + largest);
    for (std::size_t i = 0; i < count; ++i)
    {
        const d3 q{quads[i].q.x, quads[i].q.y, quads[i].q.z}, u{quads[i].u.x, quads[i].u.y, quads[i].u.z},
            v{quads[i].v.x, quads[i].v.y, quads[i].v.z};
        const d3 n = cross(u, v);
        const double len = std::sqrt(dot(n, n));
        if (!(len > 0))
        {
            return false;
        }
        const d3 normal = (1 / len) * n;
        bool front = false, back = false;
        for (const d3& c : corners)
        {
            const double height = dot(normal, c - q);
            front = front || height > tolerance;
            back = back || height < -tolerance;
        }
        if (front && back)
        {
            return false;
        }
    }
    return true;
}

built_scene build_scene(const std::string& name, const rt_scene_overrides& overrides)
{
    if (overrides.image_width < 0 || overrides.samples_per_pixel < 0 || overrides.image_height < 0)
    {
        throw std::invalid_argument("the image width, height and the samples per pixel must be positive");
    }
    if (overrides.strata < 0 || overrides.strata > 46340)
    {
        throw std::invalid_argument("the strata must be from 1 to 46340");
    }
    scene_options opts{overrides.earth_width, overrides.earth_seed};
    scene_random rng; // every build starts the stream over: the scene of a fresh process (src/gpu_scene.ixx:1355)
    std::vector<std::string> warnings;
    built s = is_model_path(name) ? build_model(name, warnings) : build_named(name, rng, opts);
    camera_settings& cam = s.camera;
    if (overrides.image_width > 0)
    {
        cam.image_width = overrides.image_width;
    }
    if (overrides.samples_per_pixel > 0)
    {
        cam.samples_per_pixel = overrides.samples_per_pixel;
    }
    const int samples = cam.samples_per_pixel;
    if (overrides.strata > 0)
    {
        cam.samples_per_pixel = overrides.strata * overrides.strata;
    }
    if (overrides.max_depth > 0)
    {
        cam.max_depth = overrides.max_depth;
    }
    if (overrides.image_height > 0)
    {
        cam.aspect_ratio = static_cast<double>(cam.image_width) / overrides.image_height;
    }
    // -p: the books' four strategies set the camera's; nee and mis keep the material strategy for the bounces and
    // ask for light samples (src/gpu_scene.ixx:1375-1440)
    int requested = overrides.sampling;
    int combination = RT_COMBINE_BSDF_ONLY;
    if (requested == -1 && s.combination >= 0)
    {
        requested = s.combination == RT_COMBINE_LIGHTS_ONLY ? 4 : (s.combination == RT_COMBINE_POWER_HEURISTIC ? 5 : 1);
    }
    if (requested >= 0 && requested <= 3)
    {
        cam.sampling = requested;
    }
    else if (requested == 4 || requested == 5)
    {
        cam.sampling = RT_SAMPLING_MATERIAL;
        combination = requested == 4 ? RT_COMBINE_LIGHTS_ONLY : RT_COMBINE_POWER_HEURISTIC;
    }
    else if (requested != -1)
    {
        throw std::invalid_argument("unknown sampling strategy");
    }
    // effective_sampling: lights and mixture need a lights list (src/camera.ixx:64-69)
    int sampling = cam.sampling;
    const bool aims = sampling == RT_SAMPLING_LIGHTS || sampling == RT_SAMPLING_MIXTURE;
    if (aims && s.lights.g.members.empty())
    {
        sampling = RT_SAMPLING_MATERIAL;
    }

    built_scene result;
    scene_tables& t = result.tables;
    flattener f(t);
    const object_ref world = s.world.finish();
    f.add_world(*world);
    t.camera = frame_of(cam, sampling);
    if (sampling == RT_SAMPLING_LIGHTS || sampling == RT_SAMPLING_MIXTURE)
    {
        const object_ref lights = s.lights.finish();
        f.add_lights(*lights, placement{}, 1.0);
    }
    for (const mesh_def& m : s.meshes)
    {
        f.add_mesh(m);
    }
    t.instances = s.instances;
    f.finish_microfacets();
    if (s.environment)
    {
        t.has_environment = true;
        t.environment_pixels = s.environment->image->rgb;
        t.environment.image = {s.environment->image->width, s.environment->image->height, nullptr};
        t.environment.yaw_degrees = s.environment->yaw_degrees;
        t.environment.scale = s.environment->scale;
    }
    t.combination = combination;
    t.relink();
    result.warnings = std::move(warnings);
    result.passes = overrides.strata > 0 ? samples
                                         : static_cast<std::int64_t>(t.camera.sqrt_spp) * t.camera.sqrt_spp;
    return result;
}

const char* scene_names()
{
    return "materials,weekend,bouncing,checker,earth,perlin,quads,light,cornell,smoke,final,aluminum,glass,furnace,"
           "ggx-furnace,gloss-furnace,env-furnace,mesh-furnace,cornell-ggx,veach,sun-test,triangles";
}
}
