// ORACLE — test infrastructure only. A portable C++20 restatement of DaniilK19/RayTracing's CPU ray tracer (the
// reference renderer): its object model, random numbers, materials, textures, scenes, and the flattening of a scene
// into the GPU renderer's tables. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, as the
// checker. The product (raytracing_amd/) never includes, links or calls anything here.
//
// The reference is C++23 modules for MSVC (`import std`); this file restates the same algorithms for GCC 11 /
// libstdc++ 11. Citations are to the reference's files.
//
// Two documented differences from the reference, neither of which changes any distribution:
//   - the object-median BVH sorts with std::stable_sort (src/bvh.ixx:66-67 uses std::ranges::sort, whose order for
//     equal keys is up to the standard library); this fixes the tie order, and so the flattening order of final's
//     ground boxes and cloud, the same way on every platform;
//   - the earth texture is the seeded noise image of noise_image() below instead of assets/earthmap.jpg.
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <memory>
#include <numbers>
#include <optional>
#include <random>
#include <span>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <utility>
#include <variant>
#include <vector>

namespace oracle
{
// --- random (src/random.ixx:9-38) ----------------------------------------------------------------------------------
inline std::mt19937_64& engine()
{
    thread_local std::mt19937_64 generator;
    return generator;
}
inline void seed_random(std::uint64_t seed) { engine().seed(seed); }
inline double random_double() { return std::uniform_real_distribution<double>(0.0, 1.0)(engine()); }
inline double random_double(double min, double max) { return min + (max - min) * random_double(); }
inline int random_int(int min, int max) { return static_cast<int>(random_double(min, max + 1)); }

// --- vec3 (src/vec3.ixx:12-132) -------------------------------------------------------------------------------------
struct vec3
{
    double x = 0, y = 0, z = 0;
    vec3 operator-() const { return {-x, -y, -z}; }
    vec3& operator+=(const vec3& v)
    {
        x += v.x;
        y += v.y;
        z += v.z;
        return *this;
    }
    vec3& operator*=(double t)
    {
        x *= t;
        y *= t;
        z *= t;
        return *this;
    }
    double length_squared() const { return x * x + y * y + z * z; }
    double length() const { return std::sqrt(length_squared()); }
    static vec3 random()
    {
        // a braced list is evaluated left to right: x, then y, then z
        const double a = random_double();
        const double b = random_double();
        const double c = random_double();
        return {a, b, c};
    }
    static vec3 random(double min, double max)
    {
        const double a = random_double(min, max);
        const double b = random_double(min, max);
        const double c = random_double(min, max);
        return {a, b, c};
    }
};
using point3 = vec3;
using color = vec3;
inline vec3 operator+(const vec3& u, const vec3& v) { return {u.x + v.x, u.y + v.y, u.z + v.z}; }
inline vec3 operator-(const vec3& u, const vec3& v) { return {u.x - v.x, u.y - v.y, u.z - v.z}; }
inline vec3 operator*(const vec3& u, const vec3& v) { return {u.x * v.x, u.y * v.y, u.z * v.z}; }
inline vec3 operator*(double t, const vec3& v) { return {t * v.x, t * v.y, t * v.z}; }
inline vec3 operator*(const vec3& v, double t) { return t * v; }
inline vec3 operator/(const vec3& v, double t) { return (1 / t) * v; }
inline double dot(const vec3& u, const vec3& v) { return u.x * v.x + u.y * v.y + u.z * v.z; }
inline vec3 cross(const vec3& u, const vec3& v)
{
    return {u.y * v.z - u.z * v.y, u.z * v.x - u.x * v.z, u.x * v.y - u.y * v.x};
}
inline vec3 unit_vector(const vec3& v) { return v / v.length(); }
inline double degrees_to_radians(double d) { return d * std::numbers::pi / 180.0; }
inline vec3 random_in_unit_disk()
{
    while (true)
    {
        const double a = random_double(-1, 1);
        const double b = random_double(-1, 1);
        const vec3 p{a, b, 0};
        if (p.length_squared() < 1)
        {
            return p;
        }
    }
}
inline vec3 reflect(const vec3& v, const vec3& n) { return v - 2 * dot(v, n) * n; }
inline vec3 refract(const vec3& uv, const vec3& n, double etai_over_etat)
{
    const double cos_theta = std::fmin(dot(-uv, n), 1.0);
    const vec3 r_out_perp = etai_over_etat * (uv + cos_theta * n);
    const vec3 r_out_parallel = -std::sqrt(std::fabs(1.0 - r_out_perp.length_squared())) * n;
    return r_out_perp + r_out_parallel;
}
inline vec3 random_unit_vector()
{
    while (true)
    {
        const vec3 p = vec3::random(-1, 1);
        const double l2 = p.length_squared();
        if (1e-160 < l2 && l2 <= 1)
        {
            return p / std::sqrt(l2);
        }
    }
}
inline vec3 random_cosine_direction()
{
    const double r1 = random_double();
    const double r2 = random_double();
    const double phi = 2 * std::numbers::pi * r1;
    return {std::cos(phi) * std::sqrt(r2), std::sin(phi) * std::sqrt(r2), std::sqrt(1 - r2)};
}

// --- ray, interval, aabb (src/ray.ixx, src/interval.ixx, src/aabb.ixx) --------------------------------------------
constexpr double infinity = std::numeric_limits<double>::infinity();
struct ray
{
    point3 origin;
    vec3 direction;
    double time = 0;
    point3 at(double t) const { return origin + t * direction; }
};
struct interval
{
    double min = +infinity;
    double max = -infinity;
    double size() const { return max - min; }
    bool contains(double x) const { return min <= x && x <= max; }
    bool surrounds(double x) const { return min < x && x < max; }
    double clamp(double x) const { return std::clamp(x, min, max); }
    interval expand(double delta) const { return {min - delta / 2, max + delta / 2}; }
};
inline interval operator+(const interval& i, double d) { return {i.min + d, i.max + d}; }
inline interval enclose(const interval& a, const interval& b) { return {std::min(a.min, b.min), std::max(a.max, b.max)}; }

struct aabb
{
    interval x, y, z;
    static interval axis_span(double a, double b)
    {
        const interval span{std::min(a, b), std::max(a, b)};
        return span.size() < 0.0001 ? span.expand(0.0001) : span;
    }
    static aabb from_corners(const point3& a, const point3& b)
    {
        return {axis_span(a.x, b.x), axis_span(a.y, b.y), axis_span(a.z, b.z)};
    }
    const interval& axis_interval(int axis) const { return axis == 0 ? x : axis == 1 ? y : z; }
    int longest_axis() const
    {
        if (x.size() > y.size())
        {
            return x.size() > z.size() ? 0 : 2;
        }
        return y.size() > z.size() ? 1 : 2;
    }
    static bool hit_slab(const interval& slab, double origin, double direction, interval& t)
    {
        const double inverse = 1.0 / direction;
        double t0 = (slab.min - origin) * inverse;
        double t1 = (slab.max - origin) * inverse;
        if (t0 > t1)
        {
            std::swap(t0, t1);
        }
        t.min = std::max(t.min, t0);
        t.max = std::min(t.max, t1);
        return t.min < t.max;
    }
    bool hit(const ray& r, interval t) const
    {
        return hit_slab(x, r.origin.x, r.direction.x, t) && hit_slab(y, r.origin.y, r.direction.y, t) &&
               hit_slab(z, r.origin.z, r.direction.z, t);
    }
};
inline aabb operator+(const aabb& b, const vec3& o) { return {b.x + o.x, b.y + o.y, b.z + o.z}; }
inline aabb enclose(const aabb& a, const aabb& b) { return {enclose(a.x, b.x), enclose(a.y, b.y), enclose(a.z, b.z)}; }

// --- onb, pdf densities (src/onb.ixx, src/pdf.ixx) -----------------------------------------------------------------
struct onb
{
    vec3 u, v, w;
    explicit onb(const vec3& unit_w) : w(unit_w)
    {
        const vec3 a = std::fabs(w.x) > 0.9 ? vec3{0, 1, 0} : vec3{1, 0, 0};
        v = unit_vector(cross(w, a));
        u = cross(v, w);
    }
    vec3 transform(const vec3& a) const { return a.x * u + a.y * v + a.z * w; }
};
constexpr double uniform_sphere_density = 1 / (4 * std::numbers::pi);
inline double cosine_density(const vec3& unit_normal, const vec3& direction)
{
    const double cosine = dot(unit_normal, unit_vector(direction));
    return cosine > 0 ? cosine / std::numbers::pi : 0.0;
}

// --- textures and Perlin noise (src/texture.ixx, src/perlin.ixx) ---------------------------------------------------
class perlin
{
public:
    static constexpr int point_count = 256;
    perlin()
    {
        for (vec3& v : random_vectors)
        {
            v = unit_vector(vec3::random(-1, 1));
        }
        permute(perm_x);
        permute(perm_y);
        permute(perm_z);
    }
    double noise(const point3& p) const
    {
        const int i = static_cast<int>(std::floor(p.x));
        const int j = static_cast<int>(std::floor(p.y));
        const int k = static_cast<int>(std::floor(p.z));
        const double u = p.x - std::floor(p.x);
        const double v = p.y - std::floor(p.y);
        const double w = p.z - std::floor(p.z);
        const double uu = u * u * (3 - 2 * u);
        const double vv = v * v * (3 - 2 * v);
        const double ww = w * w * (3 - 2 * w);
        double accum = 0;
        for (int di = 0; di < 2; ++di)
        {
            for (int dj = 0; dj < 2; ++dj)
            {
                for (int dk = 0; dk < 2; ++dk)
                {
                    const vec3& g = random_vectors[perm_x[(i + di) & 255] ^ perm_y[(j + dj) & 255] ^ perm_z[(k + dk) & 255]];
                    const vec3 offset{u - di, v - dj, w - dk};
                    accum += (di * uu + (1 - di) * (1 - uu)) * (dj * vv + (1 - dj) * (1 - vv)) *
                             (dk * ww + (1 - dk) * (1 - ww)) * dot(g, offset);
                }
            }
        }
        return accum;
    }
    double turbulence(const point3& p, int depth) const
    {
        double accum = 0;
        point3 scaled = p;
        double weight = 1;
        for (int layer = 0; layer < depth; ++layer)
        {
            accum += weight * noise(scaled);
            weight *= 0.5;
            scaled *= 2;
        }
        return std::fabs(accum);
    }
    std::array<vec3, point_count> random_vectors;
    std::array<int, point_count> perm_x, perm_y, perm_z;

private:
    static void permute(std::array<int, point_count>& perm)
    {
        for (int i = 0; i < point_count; ++i)
        {
            perm[i] = i;
        }
        for (int i = point_count - 1; i > 0; --i)
        {
            const int target = random_int(0, i);
            std::swap(perm[i], perm[target]);
        }
    }
};

struct image
{
    int width = 0, height = 0;
    std::vector<color> pixels;
    const color& at(int x, int y) const { return pixels[static_cast<std::size_t>(y) * width + x]; }
};

// The stand-in for assets/earthmap.jpg: a seeded RGB8 noise image. Texel k (rows from the top) takes one number of a
// PCG32 stream (O'Neill's pcg32_srandom_r(initstate = 0x853c49e6748fea9b ^ seed, initseq = 0xda3e39cb94b95bdb)):
// red its bits 0-7, green 8-15, blue 16-23. A byte b becomes the linear value (b/255)^2.2 computed in double and
// rounded to float, the conversion stbi_loadf applies to LDR images (src/image.ixx:46-48).
inline std::uint32_t pcg32_next(std::uint64_t& state, std::uint64_t inc)
{
    const std::uint64_t old = state;
    state = old * 6364136223846793005ull + inc;
    const auto xorshifted = static_cast<std::uint32_t>(((old >> 18u) ^ old) >> 27u);
    const auto rot = static_cast<std::uint32_t>(old >> 59u);
    return (xorshifted >> rot) | (xorshifted << ((32u - rot) & 31u));
}
inline image noise_image(int width, int height, std::uint32_t seed)
{
    std::array<float, 256> lut{};
    for (int b = 0; b < 256; ++b)
    {
        lut[b] = static_cast<float>(std::pow(b / 255.0, 2.2));
    }
    const std::uint64_t inc = (0xda3e39cb94b95bdbull << 1u) | 1u;
    std::uint64_t state = 0;
    pcg32_next(state, inc);
    state += 0x853c49e6748fea9bull ^ seed;
    pcg32_next(state, inc);
    image img{width, height, {}};
    img.pixels.resize(static_cast<std::size_t>(width) * height);
    for (auto& p : img.pixels)
    {
        const std::uint32_t v = pcg32_next(state, inc);
        p = {lut[v & 255u], lut[(v >> 8) & 255u], lut[(v >> 16) & 255u]};
    }
    return img;
}

class texture
{
public:
    virtual ~texture() = default;
    virtual color value(double u, double v, const point3& p) const = 0;
};
class solid_color : public texture
{
public:
    explicit solid_color(const color& c) : albedo(c) {}
    color value(double, double, const point3&) const override { return albedo; }
    color albedo;
};
class checker_texture : public texture
{
public:
    checker_texture(double scale, std::shared_ptr<const texture> e, std::shared_ptr<const texture> o)
        : inv_scale(1.0 / scale), even(std::move(e)), odd(std::move(o))
    {
    }
    checker_texture(double scale, const color& e, const color& o)
        : checker_texture(scale, std::make_shared<solid_color>(e), std::make_shared<solid_color>(o))
    {
    }
    color value(double u, double v, const point3& p) const override
    {
        const int x = static_cast<int>(std::floor(inv_scale * p.x));
        const int y = static_cast<int>(std::floor(inv_scale * p.y));
        const int z = static_cast<int>(std::floor(inv_scale * p.z));
        return (x + y + z) % 2 == 0 ? even->value(u, v, p) : odd->value(u, v, p);
    }
    double inv_scale;
    std::shared_ptr<const texture> even, odd;
};
class image_texture : public texture
{
public:
    explicit image_texture(image i) : img(std::move(i)) {}
    color value(double u, double v, const point3&) const override
    {
        const interval unit{0, 1};
        u = unit.clamp(u);
        v = 1.0 - unit.clamp(v);
        const int i = std::min(static_cast<int>(u * img.width), img.width - 1);
        const int j = std::min(static_cast<int>(v * img.height), img.height - 1);
        return img.at(i, j);
    }
    image img;
};
class noise_texture : public texture
{
public:
    explicit noise_texture(double s) : scale(s) {}
    color value(double, double, const point3& p) const override
    {
        return color{0.5, 0.5, 0.5} * (1 + std::sin(scale * p.z + 10 * noise.turbulence(p, 7)));
    }
    perlin noise; // its tables are drawn when the texture is made, from the scene's stream
    double scale;
};

// --- hittable, materials (src/hittable.ixx, src/material.ixx) ------------------------------------------------------
class material;
struct hit_record
{
    point3 p;
    vec3 normal;
    const material* mat = nullptr;
    double t = 0, u = 0, v = 0;
    bool front_face = false;
    void set_face_normal(const ray& r, const vec3& outward)
    {
        front_face = dot(r.direction, outward) < 0;
        normal = front_face ? outward : -outward;
    }
};

struct cosine_pdf
{
    onb uvw;
    explicit cosine_pdf(const vec3& n) : uvw(n) {}
    double value(const vec3& d) const { return cosine_density(uvw.w, d); }
    vec3 generate() const { return uvw.transform(random_cosine_direction()); }
};
struct sphere_pdf
{
    double value(const vec3&) const { return uniform_sphere_density; }
    vec3 generate() const { return random_unit_vector(); }
};
struct material_pdf
{
    std::variant<cosine_pdf, sphere_pdf> pdf;
    double value(const vec3& d) const
    {
        return std::visit([&](const auto& p) { return p.value(d); }, pdf);
    }
    vec3 generate() const
    {
        return std::visit([](const auto& p) { return p.generate(); }, pdf);
    }
};
struct diffuse_scatter
{
    color attenuation;
    material_pdf pdf;
};
struct specular_scatter
{
    color attenuation;
    ray scattered;
};
using scatter_record = std::variant<diffuse_scatter, specular_scatter>;

class material
{
public:
    virtual ~material() = default;
    virtual std::optional<scatter_record> scatter(const ray& r, const hit_record& rec) const = 0;
    virtual double scattering_pdf(const ray&, const hit_record&, const ray&) const { return 0; }
    virtual color emitted(const hit_record&) const { return {0, 0, 0}; }
    virtual bool mixes_in_lights() const { return true; }
};
class lambertian : public material
{
public:
    explicit lambertian(const color& a) : lambertian(std::make_shared<solid_color>(a)) {}
    explicit lambertian(std::shared_ptr<const texture> t) : tex(std::move(t)) {}
    std::optional<scatter_record> scatter(const ray&, const hit_record& rec) const override
    {
        return diffuse_scatter{tex->value(rec.u, rec.v, rec.p), {cosine_pdf{rec.normal}}};
    }
    double scattering_pdf(const ray&, const hit_record& rec, const ray& s) const override
    {
        return cosine_density(rec.normal, s.direction);
    }
    std::shared_ptr<const texture> tex;
};
class metal : public material
{
public:
    metal(const color& a, double f) : albedo(a), fuzz(std::min(f, 1.0)) {}
    std::optional<scatter_record> scatter(const ray& r, const hit_record& rec) const override
    {
        const vec3 reflected = unit_vector(reflect(r.direction, rec.normal)) + fuzz * random_unit_vector();
        if (dot(reflected, rec.normal) <= 0)
        {
            return std::nullopt;
        }
        return specular_scatter{albedo, ray{rec.p, reflected, r.time}};
    }
    color albedo;
    double fuzz;
};
class dielectric : public material
{
public:
    explicit dielectric(double ri) : refraction_index(ri) {}
    std::optional<scatter_record> scatter(const ray& r, const hit_record& rec) const override
    {
        const double ri = rec.front_face ? (1.0 / refraction_index) : refraction_index;
        const vec3 ud = unit_vector(r.direction);
        const double cos_theta = std::fmin(dot(-ud, rec.normal), 1.0);
        const double sin_theta = std::sqrt(1.0 - cos_theta * cos_theta);
        const bool cannot_refract = ri * sin_theta > 1.0;
        const vec3 d = cannot_refract || reflectance(cos_theta, ri) > random_double() ? reflect(ud, rec.normal)
                                                                                      : refract(ud, rec.normal, ri);
        return specular_scatter{color{1.0, 1.0, 1.0}, ray{rec.p, d, r.time}};
    }
    static double reflectance(double cosine, double ri)
    {
        double r0 = (1 - ri) / (1 + ri);
        r0 = r0 * r0;
        return r0 + (1 - r0) * std::pow(1 - cosine, 5);
    }
    double refraction_index;
};
class diffuse_light : public material
{
public:
    explicit diffuse_light(const color& e) : diffuse_light(std::make_shared<solid_color>(e)) {}
    explicit diffuse_light(std::shared_ptr<const texture> t) : tex(std::move(t)) {}
    std::optional<scatter_record> scatter(const ray&, const hit_record&) const override { return std::nullopt; }
    color emitted(const hit_record& rec) const override
    {
        return rec.front_face ? tex->value(rec.u, rec.v, rec.p) : color{0, 0, 0};
    }
    std::shared_ptr<const texture> tex;
};
class isotropic : public material
{
public:
    explicit isotropic(std::shared_ptr<const texture> t) : tex(std::move(t)) {}
    std::optional<scatter_record> scatter(const ray&, const hit_record& rec) const override
    {
        return diffuse_scatter{tex->value(rec.u, rec.v, rec.p), {sphere_pdf{}}};
    }
    double scattering_pdf(const ray&, const hit_record&, const ray&) const override { return uniform_sphere_density; }
    bool mixes_in_lights() const override { return false; }
    std::shared_ptr<const texture> tex;
};

class hittable
{
public:
    virtual ~hittable() = default;
    virtual std::optional<hit_record> hit(const ray& r, interval t) const = 0;
    virtual aabb bounding_box() const = 0;
    virtual double pdf_value(const point3&, const vec3&) const { return 0; }
    virtual vec3 random(const point3&) const { return {1, 0, 0}; }
};

class hittable_list : public hittable
{
public:
    void add(std::unique_ptr<hittable> o)
    {
        bbox = enclose(bbox, o->bounding_box());
        objects.push_back(std::move(o));
    }
    aabb bounding_box() const override { return bbox; }
    bool empty() const { return objects.empty(); }
    double pdf_value(const point3& origin, const vec3& d) const override
    {
        if (objects.empty())
        {
            return 0;
        }
        double sum = 0;
        for (const auto& o : objects)
        {
            sum += o->pdf_value(origin, d);
        }
        return sum / static_cast<double>(objects.size());
    }
    vec3 random(const point3& origin) const override
    {
        if (objects.empty())
        {
            return {1, 0, 0};
        }
        const int index = random_int(0, static_cast<int>(objects.size()) - 1);
        return objects[static_cast<std::size_t>(index)]->random(origin);
    }
    std::optional<hit_record> hit(const ray& r, interval t) const override
    {
        std::optional<hit_record> closest;
        for (const auto& o : objects)
        {
            if (auto rec = o->hit(r, t))
            {
                t.max = rec->t;
                closest = rec;
            }
        }
        return closest;
    }
    std::vector<std::unique_ptr<hittable>> objects;
    aabb bbox;
};

// src/sphere.ixx:11-160
class sphere : public hittable
{
public:
    sphere(const point3& c, double r, std::shared_ptr<const material> m) : sphere(c, c, r, std::move(m)) {}
    sphere(const point3& c1, const point3& c2, double r, std::shared_ptr<const material> m)
        : center{c1, c2 - c1}, radius(std::fmax(0.0, r)), mat(std::move(m))
    {
        const vec3 rv{radius, radius, radius};
        const aabb b1 = aabb::from_corners(center.at(0) - rv, center.at(0) + rv);
        const aabb b2 = aabb::from_corners(center.at(1) - rv, center.at(1) + rv);
        bbox = enclose(b1, b2);
    }
    aabb bounding_box() const override { return bbox; }
    point3 center_at(double time) const { return center.at(time); }
    std::optional<hit_record> hit(const ray& r, interval t) const override
    {
        const point3 c = center.at(r.time);
        const vec3 oc = c - r.origin;
        const double a = r.direction.length_squared();
        const double h = dot(r.direction, oc);
        const double cc = oc.length_squared() - radius * radius;
        const double discriminant = h * h - a * cc;
        if (discriminant < 0)
        {
            return std::nullopt;
        }
        const double sqrtd = std::sqrt(discriminant);
        double root = (h - sqrtd) / a;
        if (!t.surrounds(root))
        {
            root = (h + sqrtd) / a;
            if (!t.surrounds(root))
            {
                return std::nullopt;
            }
        }
        hit_record rec;
        rec.t = root;
        rec.p = r.at(rec.t);
        const vec3 outward = (rec.p - c) / radius;
        rec.set_face_normal(r, outward);
        rec.mat = mat.get();
        const double phi = std::atan2(-outward.z, outward.x) + std::numbers::pi;
        const double theta = std::acos(std::clamp(-outward.y, -1.0, 1.0));
        rec.u = phi / (2 * std::numbers::pi);
        rec.v = theta / std::numbers::pi;
        return rec;
    }
    double pdf_value(const point3& origin, const vec3& d) const override
    {
        const double d2 = (center.at(0) - origin).length_squared();
        if (d2 < radius * radius)
        {
            return uniform_sphere_density;
        }
        if (!hit(ray{origin, d}, interval{0.001, infinity}))
        {
            return 0;
        }
        return 1 / (2 * std::numbers::pi * one_minus_cos_theta_max(d2));
    }
    vec3 random(const point3& origin) const override
    {
        const vec3 to_center = center.at(0) - origin;
        const double d2 = to_center.length_squared();
        if (d2 < radius * radius)
        {
            return random_unit_vector();
        }
        const onb uvw{unit_vector(to_center)};
        return uvw.transform(random_to_sphere(one_minus_cos_theta_max(d2)));
    }
    double one_minus_cos_theta_max(double d2) const
    {
        const double s2 = radius * radius / d2;
        return s2 / (1 + std::sqrt(1 - s2));
    }
    static vec3 random_to_sphere(double one_minus_cos_max)
    {
        const double r1 = random_double();
        const double r2 = random_double();
        const double phi = 2 * std::numbers::pi * r1;
        const double omz = r2 * one_minus_cos_max;
        const double sin_theta = std::sqrt(omz * (2 - omz));
        return {std::cos(phi) * sin_theta, std::sin(phi) * sin_theta, 1 - omz};
    }
    ray center;
    double radius;
    std::shared_ptr<const material> mat;
    aabb bbox;
};

// src/quad.ixx:12-140
class quad : public hittable
{
public:
    quad(const point3& q0, const vec3& u0, const vec3& v0, std::shared_ptr<const material> m)
        : q(q0), u(u0), v(v0), mat(std::move(m))
    {
        const vec3 n = cross(u, v);
        normal = unit_vector(n);
        d = dot(normal, q);
        w = n / dot(n, n);
        area = n.length();
        bbox = enclose(aabb::from_corners(q, q + u + v), aabb::from_corners(q + u, q + v));
    }
    aabb bounding_box() const override { return bbox; }
    double pdf_value(const point3& origin, const vec3& dir) const override
    {
        const auto rec = hit(ray{origin, dir}, interval{0.001, infinity});
        if (!rec)
        {
            return 0;
        }
        const double distance_squared = rec->t * rec->t * dir.length_squared();
        const double cosine = std::fabs(dot(dir, normal)) / dir.length();
        return distance_squared / (cosine * area);
    }
    vec3 random(const point3& origin) const override
    {
        const double a = random_double();
        const double b = random_double();
        return unit_vector(q + a * u + b * v - origin);
    }
    std::optional<hit_record> hit(const ray& r, interval t) const override
    {
        const double denom = dot(normal, r.direction);
        if (std::fabs(denom) < 1e-8)
        {
            return std::nullopt;
        }
        const double tt = (d - dot(normal, r.origin)) / denom;
        if (!t.contains(tt))
        {
            return std::nullopt;
        }
        const point3 intersection = r.at(tt);
        const vec3 p = intersection - q;
        const double alpha = dot(w, cross(p, v));
        const double beta = dot(w, cross(u, p));
        const interval unit{0, 1};
        if (!unit.contains(alpha) || !unit.contains(beta))
        {
            return std::nullopt;
        }
        hit_record rec;
        rec.t = tt;
        rec.p = intersection;
        rec.mat = mat.get();
        rec.set_face_normal(r, normal);
        rec.u = alpha;
        rec.v = beta;
        return rec;
    }
    point3 q;
    vec3 u, v;
    std::shared_ptr<const material> mat;
    vec3 normal;
    double d;
    vec3 w;
    double area;
    aabb bbox;
};

inline std::unique_ptr<hittable> box(const point3& a, const point3& b, std::shared_ptr<const material> m)
{
    const point3 mn{std::fmin(a.x, b.x), std::fmin(a.y, b.y), std::fmin(a.z, b.z)};
    const point3 mx{std::fmax(a.x, b.x), std::fmax(a.y, b.y), std::fmax(a.z, b.z)};
    const vec3 dx{mx.x - mn.x, 0, 0};
    const vec3 dy{0, mx.y - mn.y, 0};
    const vec3 dz{0, 0, mx.z - mn.z};
    auto sides = std::make_unique<hittable_list>();
    sides->add(std::make_unique<quad>(point3{mn.x, mn.y, mx.z}, dx, dy, m));
    sides->add(std::make_unique<quad>(point3{mx.x, mn.y, mx.z}, -dz, dy, m));
    sides->add(std::make_unique<quad>(point3{mx.x, mn.y, mn.z}, -dx, dy, m));
    sides->add(std::make_unique<quad>(point3{mn.x, mn.y, mn.z}, dz, dy, m));
    sides->add(std::make_unique<quad>(point3{mn.x, mx.y, mx.z}, dx, -dz, m));
    sides->add(std::make_unique<quad>(point3{mn.x, mn.y, mn.z}, dx, dz, m));
    return sides;
}

// src/instance.ixx:14-130
class translate : public hittable
{
public:
    translate(std::unique_ptr<hittable> o, const vec3& off) : object(std::move(o)), offset(off)
    {
        bbox = object->bounding_box() + offset;
    }
    std::optional<hit_record> hit(const ray& r, interval t) const override
    {
        const ray moved{r.origin - offset, r.direction, r.time};
        auto rec = object->hit(moved, t);
        if (rec)
        {
            rec->p += offset;
        }
        return rec;
    }
    aabb bounding_box() const override { return bbox; }
    double pdf_value(const point3& o, const vec3& d) const override { return object->pdf_value(o - offset, d); }
    vec3 random(const point3& o) const override { return object->random(o - offset); }
    std::unique_ptr<hittable> object;
    vec3 offset;
    aabb bbox;
};
class rotate_y : public hittable
{
public:
    rotate_y(std::unique_ptr<hittable> o, double angle)
        : object(std::move(o)), sin_theta(std::sin(degrees_to_radians(angle))),
          cos_theta(std::cos(degrees_to_radians(angle)))
    {
        const aabb b = object->bounding_box();
        for (const double x : {b.x.min, b.x.max})
        {
            for (const double y : {b.y.min, b.y.max})
            {
                for (const double z : {b.z.min, b.z.max})
                {
                    const point3 c = to_world(point3{x, y, z});
                    bbox = enclose(bbox, aabb::from_corners(c, c));
                }
            }
        }
    }
    std::optional<hit_record> hit(const ray& r, interval t) const override
    {
        const ray rotated{to_object(r.origin), to_object(r.direction), r.time};
        auto rec = object->hit(rotated, t);
        if (rec)
        {
            rec->p = to_world(rec->p);
            rec->normal = to_world(rec->normal);
        }
        return rec;
    }
    aabb bounding_box() const override { return bbox; }
    double pdf_value(const point3& o, const vec3& d) const override
    {
        return object->pdf_value(to_object(o), to_object(d));
    }
    vec3 random(const point3& o) const override { return to_world(object->random(to_object(o))); }
    vec3 to_object(const vec3& v) const
    {
        return {cos_theta * v.x - sin_theta * v.z, v.y, sin_theta * v.x + cos_theta * v.z};
    }
    vec3 to_world(const vec3& v) const
    {
        return {cos_theta * v.x + sin_theta * v.z, v.y, -sin_theta * v.x + cos_theta * v.z};
    }
    std::unique_ptr<hittable> object;
    double sin_theta, cos_theta;
    aabb bbox;
};

// src/constant_medium.ixx:17-90
class constant_medium : public hittable
{
public:
    constant_medium(std::unique_ptr<hittable> b, double density, const color& albedo)
        : boundary(std::move(b)), neg_inv_density(-1 / density),
          phase_function(std::make_shared<isotropic>(std::make_shared<solid_color>(albedo)))
    {
    }
    std::optional<hit_record> hit(const ray& r, interval t) const override
    {
        const auto hit_in = boundary->hit(r, interval{-infinity, infinity});
        if (!hit_in)
        {
            return std::nullopt;
        }
        const auto hit_out = boundary->hit(r, interval{hit_in->t + 0.0001, infinity});
        if (!hit_out)
        {
            return std::nullopt;
        }
        double t_in = std::max(hit_in->t, t.min);
        const double t_out = std::min(hit_out->t, t.max);
        if (t_in >= t_out)
        {
            return std::nullopt;
        }
        t_in = std::max(t_in, 0.0);
        const double length = r.direction.length();
        const double inside = (t_out - t_in) * length;
        const double hit_distance = neg_inv_density * std::log(random_double());
        if (hit_distance > inside)
        {
            return std::nullopt;
        }
        hit_record rec;
        rec.t = t_in + hit_distance / length;
        rec.p = r.at(rec.t);
        rec.normal = vec3{1, 0, 0};
        rec.front_face = true;
        rec.mat = phase_function.get();
        return rec;
    }
    aabb bounding_box() const override { return boundary->bounding_box(); }
    std::unique_ptr<hittable> boundary;
    double neg_inv_density;
    std::shared_ptr<const material> phase_function;
};

// src/bvh.ixx:14-84 (stable sort: see the top of this file)
class bvh_node : public hittable
{
public:
    bvh_node(std::unique_ptr<hittable> l, std::unique_ptr<hittable> r) : left(std::move(l)), right(std::move(r))
    {
        bbox = enclose(left->bounding_box(), right->bounding_box());
    }
    std::optional<hit_record> hit(const ray& r, interval t) const override
    {
        if (!bbox.hit(r, t))
        {
            return std::nullopt;
        }
        const auto hl = left->hit(r, t);
        const auto hr = right->hit(r, interval{t.min, hl ? hl->t : t.max});
        return hr ? hr : hl;
    }
    aabb bounding_box() const override { return bbox; }
    std::unique_ptr<hittable> left, right;
    aabb bbox;
};
inline std::unique_ptr<hittable> build_node(std::span<std::unique_ptr<hittable>> objects)
{
    if (objects.size() == 1)
    {
        return std::move(objects.front());
    }
    aabb b;
    for (const auto& o : objects)
    {
        b = enclose(b, o->bounding_box());
    }
    const int axis = b.longest_axis();
    std::stable_sort(objects.begin(), objects.end(), [axis](const auto& p, const auto& q) {
        return p->bounding_box().axis_interval(axis).min < q->bounding_box().axis_interval(axis).min;
    });
    const std::size_t middle = objects.size() / 2;
    return std::make_unique<bvh_node>(build_node(objects.first(middle)), build_node(objects.subspan(middle)));
}
inline std::unique_ptr<hittable> build_bvh(hittable_list list)
{
    std::vector<std::unique_ptr<hittable>> objects = std::move(list.objects);
    if (objects.empty())
    {
        return std::make_unique<hittable_list>();
    }
    return build_node(objects);
}

// --- camera settings and scenes (src/camera.ixx:74-110, src/scenes.ixx) --------------------------------------------
enum class sampling_strategy
{
    uniform = 0,
    material = 1,
    lights = 2,
    mixture = 3,
};
struct sky
{
    color bottom{0, 0, 0};
    color top{0, 0, 0};
    color in_direction(const vec3& d) const
    {
        const double a = 0.5 * (unit_vector(d).y + 1.0);
        return (1.0 - a) * bottom + a * top;
    }
};
constexpr sky daylight_sky{{1.0, 1.0, 1.0}, {0.5, 0.7, 1.0}};
struct camera_settings
{
    double aspect_ratio = 1.0;
    int image_width = 100;
    int samples_per_pixel = 10;
    int max_depth = 10;
    sampling_strategy sampling = sampling_strategy::mixture;
    std::uint32_t seed = 0;
    double vfov = 90;
    point3 lookfrom{0, 0, 0};
    point3 lookat{0, 0, -1};
    vec3 vup{0, 1, 0};
    double defocus_angle = 0;
    double focus_dist = 10;
    sky background;
};
struct scene
{
    hittable_list world;
    camera_settings camera;
    hittable_list lights;
};

struct scene_options
{
    int earth_width = 2048; // the noise image standing in for earthmap.jpg is earth_width x earth_width / 2
    std::uint32_t earth_seed = 0;
};

inline hittable_list random_spheres(bool bouncing)
{
    hittable_list world;
    std::shared_ptr<texture> ground = std::make_shared<solid_color>(color{0.5, 0.5, 0.5});
    if (bouncing)
    {
        ground = std::make_shared<checker_texture>(0.32, color{0.2, 0.3, 0.1}, color{0.9, 0.9, 0.9});
    }
    world.add(std::make_unique<sphere>(point3{0, -1000, 0}, 1000, std::make_shared<lambertian>(ground)));
    for (int a = -11; a < 11; ++a)
    {
        for (int b = -11; b < 11; ++b)
        {
            const double choose_mat = random_double();
            const double cx = a + 0.9 * random_double();
            const double cz = b + 0.9 * random_double();
            const point3 center{cx, 0.2, cz};
            if ((center - point3{4, 0.2, 0}).length() <= 0.9)
            {
                continue;
            }
            std::shared_ptr<material> m;
            point3 center2 = center;
            if (choose_mat < 0.8)
            {
                // color::random() * color::random(): each channel is the product of two draws, and a product of two
                // numbers doesn't depend on which was drawn first
                const color c1 = color::random();
                const color c2 = color::random();
                m = std::make_shared<lambertian>(c1 * c2);
                if (bouncing)
                {
                    center2 = center + vec3{0, random_double(0, 0.5), 0};
                }
            }
            else if (choose_mat < 0.95)
            {
                // make_shared<metal>(color::random(0.5, 1), random_double(0, 0.5)): MSVC evaluates the arguments
                // of a call from right to left, so the fuzz is drawn before the albedo
                const double fuzz = random_double(0, 0.5);
                const color albedo = color::random(0.5, 1);
                m = std::make_shared<metal>(albedo, fuzz);
            }
            else
            {
                m = std::make_shared<dielectric>(1.5);
            }
            world.add(std::make_unique<sphere>(center, center2, 0.2, m));
        }
    }
    world.add(std::make_unique<sphere>(point3{0, 1, 0}, 1.0, std::make_shared<dielectric>(1.5)));
    world.add(std::make_unique<sphere>(point3{-4, 1, 0}, 1.0, std::make_shared<lambertian>(color{0.4, 0.2, 0.1})));
    world.add(std::make_unique<sphere>(point3{4, 1, 0}, 1.0, std::make_shared<metal>(color{0.7, 0.6, 0.5}, 0.0)));
    return world;
}

inline camera_settings cover_camera()
{
    camera_settings c;
    c.aspect_ratio = 16.0 / 9.0;
    c.image_width = 1920;
    c.samples_per_pixel = 500;
    c.max_depth = 50;
    c.vfov = 20;
    c.lookfrom = {13, 2, 3};
    c.lookat = {0, 0, 0};
    c.vup = {0, 1, 0};
    c.defocus_angle = 0.6;
    c.focus_dist = 10.0;
    c.background = daylight_sky;
    return c;
}

inline void add_cornell_walls(hittable_list& world)
{
    const auto red = std::make_shared<lambertian>(color{0.65, 0.05, 0.05});
    const auto white = std::make_shared<lambertian>(color{0.73, 0.73, 0.73});
    const auto green = std::make_shared<lambertian>(color{0.12, 0.45, 0.15});
    world.add(std::make_unique<quad>(point3{555, 0, 0}, vec3{0, 555, 0}, vec3{0, 0, 555}, green));
    world.add(std::make_unique<quad>(point3{0, 0, 0}, vec3{0, 555, 0}, vec3{0, 0, 555}, red));
    world.add(std::make_unique<quad>(point3{0, 0, 0}, vec3{555, 0, 0}, vec3{0, 0, 555}, white));
    world.add(std::make_unique<quad>(point3{555, 555, 555}, vec3{-555, 0, 0}, vec3{0, 0, -555}, white));
    world.add(std::make_unique<quad>(point3{0, 0, 555}, vec3{555, 0, 0}, vec3{0, 555, 0}, white));
}
inline std::unique_ptr<hittable> tall_block(std::shared_ptr<const material> m)
{
    std::unique_ptr<hittable> b = box(point3{0, 0, 0}, point3{165, 330, 165}, std::move(m));
    b = std::make_unique<rotate_y>(std::move(b), 15);
    return std::make_unique<translate>(std::move(b), vec3{265, 0, 295});
}
inline std::unique_ptr<hittable> short_block(std::shared_ptr<const material> m)
{
    std::unique_ptr<hittable> b = box(point3{0, 0, 0}, point3{165, 165, 165}, std::move(m));
    b = std::make_unique<rotate_y>(std::move(b), -18);
    return std::make_unique<translate>(std::move(b), vec3{130, 0, 65});
}
inline std::unique_ptr<hittable> cornell_lamp(std::shared_ptr<const material> m)
{
    return std::make_unique<quad>(point3{343, 554, 332}, vec3{-130, 0, 0}, vec3{0, 0, -105}, std::move(m));
}
inline camera_settings cornell_camera()
{
    camera_settings c;
    c.aspect_ratio = 1.0;
    c.image_width = 1080;
    c.samples_per_pixel = 1000;
    c.max_depth = 50;
    c.vfov = 40;
    c.lookfrom = {278, 278, -800};
    c.lookat = {278, 278, 0};
    c.vup = {0, 1, 0};
    return c;
}

inline scene build_scene(const std::string& name, const scene_options& opts)
{
    scene s;
    if (name == "materials")
    {
        const auto ground = std::make_shared<lambertian>(color{0.8, 0.8, 0.0});
        const auto center = std::make_shared<lambertian>(color{0.1, 0.2, 0.5});
        const auto left = std::make_shared<dielectric>(1.50);
        const auto bubble = std::make_shared<dielectric>(1.00 / 1.50);
        const auto right = std::make_shared<metal>(color{0.8, 0.6, 0.2}, 1.0);
        s.world.add(std::make_unique<sphere>(point3{0.0, -100.5, -1.0}, 100.0, ground));
        s.world.add(std::make_unique<sphere>(point3{0.0, 0.0, -1.2}, 0.5, center));
        s.world.add(std::make_unique<sphere>(point3{-1.0, 0.0, -1.0}, 0.5, left));
        s.world.add(std::make_unique<sphere>(point3{-1.0, 0.0, -1.0}, 0.4, bubble));
        s.world.add(std::make_unique<sphere>(point3{1.0, 0.0, -1.0}, 0.5, right));
        s.camera = cover_camera();
        s.camera.lookfrom = {-2, 2, 1};
        s.camera.lookat = {0, 0, -1};
        s.camera.defocus_angle = 10.0;
        s.camera.focus_dist = 3.4;
    }
    else if (name == "weekend" || name == "bouncing")
    {
        s.world = random_spheres(name == "bouncing");
        s.camera = cover_camera();
    }
    else if (name == "checker")
    {
        const auto checker = std::make_shared<lambertian>(
            std::make_shared<checker_texture>(0.32, color{0.2, 0.3, 0.1}, color{0.9, 0.9, 0.9}));
        s.world.add(std::make_unique<sphere>(point3{0, -10, 0}, 10, checker));
        s.world.add(std::make_unique<sphere>(point3{0, 10, 0}, 10, checker));
        s.camera = cover_camera();
        s.camera.defocus_angle = 0;
    }
    else if (name == "earth")
    {
        const auto surface = std::make_shared<lambertian>(
            std::make_shared<image_texture>(noise_image(opts.earth_width, opts.earth_width / 2, opts.earth_seed)));
        s.world.add(std::make_unique<sphere>(point3{0, 0, 0}, 2, surface));
        s.camera = cover_camera();
        s.camera.lookfrom = {0, 0, 12};
        s.camera.defocus_angle = 0;
    }
    else if (name == "perlin")
    {
        const auto marble = std::make_shared<lambertian>(std::make_shared<noise_texture>(4));
        s.world.add(std::make_unique<sphere>(point3{0, -1000, 0}, 1000, marble));
        s.world.add(std::make_unique<sphere>(point3{0, 2, 0}, 2, marble));
        s.camera = cover_camera();
        s.camera.defocus_angle = 0;
    }
    else if (name == "quads")
    {
        const auto left_red = std::make_shared<lambertian>(color{1.0, 0.2, 0.2});
        const auto back_green = std::make_shared<lambertian>(color{0.2, 1.0, 0.2});
        const auto right_blue = std::make_shared<lambertian>(color{0.2, 0.2, 1.0});
        const auto upper_orange = std::make_shared<lambertian>(color{1.0, 0.5, 0.0});
        const auto lower_teal = std::make_shared<lambertian>(color{0.2, 0.8, 0.8});
        s.world.add(std::make_unique<quad>(point3{-3, -2, 5}, vec3{0, 0, -4}, vec3{0, 4, 0}, left_red));
        s.world.add(std::make_unique<quad>(point3{-2, -2, 0}, vec3{4, 0, 0}, vec3{0, 4, 0}, back_green));
        s.world.add(std::make_unique<quad>(point3{3, -2, 1}, vec3{0, 0, 4}, vec3{0, 4, 0}, right_blue));
        s.world.add(std::make_unique<quad>(point3{-2, 3, 1}, vec3{4, 0, 0}, vec3{0, 0, 4}, upper_orange));
        s.world.add(std::make_unique<quad>(point3{-2, -3, 5}, vec3{4, 0, 0}, vec3{0, 0, -4}, lower_teal));
        s.camera.aspect_ratio = 1.0;
        s.camera.image_width = 1080;
        s.camera.samples_per_pixel = 500;
        s.camera.max_depth = 50;
        s.camera.vfov = 80;
        s.camera.lookfrom = {0, 0, 9};
        s.camera.lookat = {0, 0, 0};
        s.camera.background = daylight_sky;
    }
    else if (name == "light")
    {
        const auto marble = std::make_shared<lambertian>(std::make_shared<noise_texture>(4));
        s.world.add(std::make_unique<sphere>(point3{0, -1000, 0}, 1000, marble));
        s.world.add(std::make_unique<sphere>(point3{0, 2, 0}, 2, marble));
        const auto sphere_lamp = [](std::shared_ptr<const material> m) {
            return std::make_unique<sphere>(point3{0, 7, 0}, 2, std::move(m));
        };
        const auto quad_lamp = [](std::shared_ptr<const material> m) {
            return std::make_unique<quad>(point3{3, 1, -2}, vec3{2, 0, 0}, vec3{0, 2, 0}, std::move(m));
        };
        const auto light = std::make_shared<diffuse_light>(color{4, 4, 4});
        s.world.add(sphere_lamp(light));
        s.world.add(quad_lamp(light));
        s.lights.add(quad_lamp(nullptr));
        s.lights.add(sphere_lamp(nullptr));
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
        add_cornell_walls(s.world);
        s.world.add(cornell_lamp(std::make_shared<diffuse_light>(color{15, 15, 15})));
        if (name == "cornell")
        {
            const auto white = std::make_shared<lambertian>(color{0.73, 0.73, 0.73});
            s.world.add(tall_block(white));
            s.world.add(short_block(white));
            s.lights.add(cornell_lamp(nullptr));
        }
        else if (name == "aluminum")
        {
            s.world.add(tall_block(std::make_shared<metal>(color{0.8, 0.85, 0.88}, 0.0)));
            s.world.add(short_block(std::make_shared<lambertian>(color{0.73, 0.73, 0.73})));
            s.lights.add(cornell_lamp(nullptr));
        }
        else
        {
            s.world.add(tall_block(std::make_shared<lambertian>(color{0.73, 0.73, 0.73})));
            const auto ball = [](std::shared_ptr<const material> m) {
                return std::make_unique<sphere>(point3{190, 90, 190}, 90, std::move(m));
            };
            s.world.add(ball(std::make_shared<dielectric>(1.5)));
            s.lights.add(cornell_lamp(nullptr));
            s.lights.add(ball(nullptr));
        }
        s.camera = cornell_camera();
    }
    else if (name == "smoke")
    {
        add_cornell_walls(s.world);
        const auto lamp = [](std::shared_ptr<const material> m) {
            return std::make_unique<quad>(point3{113, 554, 127}, vec3{330, 0, 0}, vec3{0, 0, 305}, std::move(m));
        };
        s.world.add(lamp(std::make_shared<diffuse_light>(color{7, 7, 7})));
        const auto white = std::make_shared<lambertian>(color{0.73, 0.73, 0.73});
        s.world.add(std::make_unique<constant_medium>(tall_block(white), 0.01, color{0, 0, 0}));
        s.world.add(std::make_unique<constant_medium>(short_block(white), 0.01, color{1, 1, 1}));
        s.lights.add(lamp(nullptr));
        s.camera = cornell_camera();
    }
    else if (name == "final")
    {
        hittable_list ground_boxes;
        const auto ground = std::make_shared<lambertian>(color{0.48, 0.83, 0.53});
        for (int i = 0; i < 20; ++i)
        {
            for (int j = 0; j < 20; ++j)
            {
                const double w = 100.0;
                const point3 corner0{-1000.0 + i * w, 0.0, -1000.0 + j * w};
                const point3 corner1{corner0.x + w, random_double(1, 101), corner0.z + w};
                ground_boxes.add(box(corner0, corner1, ground));
            }
        }
        s.world.add(build_bvh(std::move(ground_boxes)));
        const auto lamp = [](std::shared_ptr<const material> m) {
            return std::make_unique<quad>(point3{123, 554, 147}, vec3{300, 0, 0}, vec3{0, 0, 265}, std::move(m));
        };
        s.world.add(lamp(std::make_shared<diffuse_light>(color{7, 7, 7})));
        const point3 center1{400, 400, 200};
        const point3 center2 = center1 + vec3{30, 0, 0};
        s.world.add(std::make_unique<sphere>(center1, center2, 50, std::make_shared<lambertian>(color{0.7, 0.3, 0.1})));
        s.world.add(std::make_unique<sphere>(point3{260, 150, 45}, 50, std::make_shared<dielectric>(1.5)));
        s.world.add(std::make_unique<sphere>(point3{0, 150, 145}, 50, std::make_shared<metal>(color{0.8, 0.8, 0.9}, 1.0)));
        const auto glass = std::make_shared<dielectric>(1.5);
        s.world.add(std::make_unique<sphere>(point3{360, 150, 145}, 70, glass));
        s.world.add(std::make_unique<constant_medium>(std::make_unique<sphere>(point3{360, 150, 145}, 70, glass), 0.2,
                                                      color{0.2, 0.4, 0.9}));
        s.world.add(std::make_unique<constant_medium>(std::make_unique<sphere>(point3{0, 0, 0}, 5000, glass), 0.0001,
                                                      color{1, 1, 1}));
        s.world.add(std::make_unique<sphere>(
            point3{400, 200, 400}, 100,
            std::make_shared<lambertian>(
                std::make_shared<image_texture>(noise_image(opts.earth_width, opts.earth_width / 2, opts.earth_seed)))));
        s.world.add(std::make_unique<sphere>(point3{220, 280, 300}, 80,
                                             std::make_shared<lambertian>(std::make_shared<noise_texture>(0.2))));
        hittable_list cloud;
        const auto white = std::make_shared<lambertian>(color{0.73, 0.73, 0.73});
        for (int i = 0; i < 1000; ++i)
        {
            cloud.add(std::make_unique<sphere>(vec3::random(0, 165), 10, white));
        }
        s.world.add(std::make_unique<translate>(std::make_unique<rotate_y>(build_bvh(std::move(cloud)), 15),
                                                vec3{-100, 270, 395}));
        s.lights.add(lamp(nullptr));
        s.camera.aspect_ratio = 1.0;
        s.camera.image_width = 1080;
        s.camera.samples_per_pixel = 1000;
        s.camera.max_depth = 40;
        s.camera.vfov = 40;
        s.camera.lookfrom = {478, 278, -600};
        s.camera.lookat = {278, 278, 0};
        s.camera.vup = {0, 1, 0};
    }
    else if (name == "furnace")
    {
        // src/gpu_scenes.ixx:33-55: a white matte sphere under a white sky; every pixel is exactly 1
        s.world.add(std::make_unique<sphere>(point3{0, 0, 0}, 1.0, std::make_shared<lambertian>(color{1, 1, 1})));
        s.camera.aspect_ratio = 16.0 / 9.0;
        s.camera.image_width = 480;
        s.camera.samples_per_pixel = 64;
        s.camera.max_depth = 10;
        s.camera.sampling = sampling_strategy::material;
        s.camera.vfov = 30;
        s.camera.lookfrom = {0, 0, 6};
        s.camera.lookat = {0, 0, 0};
        s.camera.background = {{1, 1, 1}, {1, 1, 1}};
    }
    else
    {
        throw std::invalid_argument("unknown scene '" + name + "'");
    }
    return s;
}

// --- the camera frame (src/camera.ixx:112-199) -----------------------------------------------------------------------
struct camera_frame
{
    int image_width, image_height, sqrt_spp, max_depth;
    point3 center, pixel00_loc;
    vec3 pixel_delta_u, pixel_delta_v;
    double defocus_angle;
    vec3 defocus_disk_u, defocus_disk_v;
    sky background;
};
inline camera_frame frame_of(const camera_settings& s)
{
    camera_frame f{};
    f.image_width = s.image_width;
    f.image_height = std::max(1, static_cast<int>(s.image_width / s.aspect_ratio));
    f.sqrt_spp = std::max(1, static_cast<int>(std::lround(std::sqrt(s.samples_per_pixel))));
    f.max_depth = s.max_depth;
    f.defocus_angle = s.defocus_angle;
    f.background = s.background;
    f.center = s.lookfrom;
    const double h = std::tan(degrees_to_radians(s.vfov) / 2);
    const double viewport_height = 2 * h * s.focus_dist;
    const double viewport_width = viewport_height * (static_cast<double>(f.image_width) / f.image_height);
    const vec3 w = unit_vector(s.lookfrom - s.lookat);
    const vec3 u = unit_vector(cross(s.vup, w));
    const vec3 v = cross(w, u);
    const vec3 viewport_u = viewport_width * u;
    const vec3 viewport_v = viewport_height * -v;
    f.pixel_delta_u = viewport_u / f.image_width;
    f.pixel_delta_v = viewport_v / f.image_height;
    const point3 upper_left = f.center - s.focus_dist * w - viewport_u / 2 - viewport_v / 2;
    f.pixel00_loc = upper_left + 0.5 * (f.pixel_delta_u + f.pixel_delta_v);
    const double defocus_radius = s.focus_dist * std::tan(degrees_to_radians(s.defocus_angle / 2));
    f.defocus_disk_u = defocus_radius * u;
    f.defocus_disk_v = defocus_radius * v;
    return f;
}
inline sampling_strategy effective_sampling(sampling_strategy requested, const hittable_list& lights)
{
    const bool aims = requested == sampling_strategy::lights || requested == sampling_strategy::mixture;
    return aims && lights.empty() ? sampling_strategy::material : requested;
}

// --- flattening into the GPU renderer's tables (src/gpu_scene.ixx:97-768) -------------------------------------------
struct flat_sphere
{
    vec3 center;
    double radius;
    int material;
    vec3 displacement;
    double uv_cos = 1, uv_sin = 0;
};
struct flat_quad
{
    vec3 q, u, v;
    int material;
};
struct flat_material
{
    int kind; // 0 lambertian, 1 metal, 2 dielectric, 3 diffuse_light, 4 isotropic
    vec3 albedo;
    double fuzz = 0, refraction_index = 1;
    int texture = -1;
};
struct flat_texture
{
    int kind; // 0 solid, 1 checker, 2 image, 3 noise
    vec3 rgb;
    double inv_scale = 1;
    int even = -1, odd = -1, image = -1;
    double scale = 1;
    int perlin = -1;
};
struct flat_medium
{
    int boundary = 0; // 0 sphere, 1 convex quads
    flat_sphere sphere{};
    std::vector<flat_quad> quads;
    double neg_inv_density = 0;
    int material = 0;
};
struct flat_light
{
    int kind = 0; // 0 quad, 1 sphere
    flat_quad quad{};
    vec3 center;
    double radius = 0;
    double weight = 0;
};
struct flat_scene
{
    std::vector<flat_sphere> spheres;
    std::vector<flat_quad> quads;
    std::vector<flat_material> materials;
    std::vector<flat_texture> textures;
    std::vector<const image*> images;
    std::vector<const perlin*> perlins;
    std::vector<flat_medium> media;
    std::vector<flat_light> lights;
    camera_frame frame{};
    sampling_strategy sampling = sampling_strategy::material;
};

struct placement
{
    double cos_theta = 1, sin_theta = 0;
    vec3 offset{0, 0, 0};
    vec3 turn(const vec3& v) const
    {
        return {cos_theta * v.x + sin_theta * v.z, v.y, -sin_theta * v.x + cos_theta * v.z};
    }
    point3 place(const point3& p) const { return turn(p) + offset; }
    placement inside(const translate& t) const { return {cos_theta, sin_theta, place(t.offset)}; }
    placement inside(const rotate_y& r) const
    {
        return {cos_theta * r.cos_theta - sin_theta * r.sin_theta, sin_theta * r.cos_theta + cos_theta * r.sin_theta,
                offset};
    }
};

template <typename F>
void for_each_shape(const hittable& object, const placement& where, const F& on_shape)
{
    if (const auto* list = dynamic_cast<const hittable_list*>(&object))
    {
        for (const auto& child : list->objects)
        {
            for_each_shape(*child, where, on_shape);
        }
        return;
    }
    if (const auto* node = dynamic_cast<const bvh_node*>(&object))
    {
        for_each_shape(*node->left, where, on_shape);
        for_each_shape(*node->right, where, on_shape);
        return;
    }
    if (const auto* moved = dynamic_cast<const translate*>(&object))
    {
        for_each_shape(*moved->object, where.inside(*moved), on_shape);
        return;
    }
    if (const auto* turned = dynamic_cast<const rotate_y*>(&object))
    {
        for_each_shape(*turned->object, where.inside(*turned), on_shape);
        return;
    }
    on_shape(object, where);
}

// Whether quads bound a convex shape (src/gpu/renderer.cu:2209-2240)
inline bool is_convex(const std::vector<flat_quad>& quads)
{
    std::vector<vec3> corners;
    double largest = 0;
    for (const auto& q : quads)
    {
        for (const vec3& c : {q.q, q.q + q.u, q.q + q.v, q.q + q.u + q.v})
        {
            corners.push_back(c);
            largest = std::max({largest, std::fabs(c.x), std::fabs(c.y), std::fabs(c.z)});
        }
    }
    const double tolerance = 1e-9 * (1 + largest);
    for (const auto& q : quads)
    {
        const vec3 n = cross(q.u, q.v);
        const double nl = std::sqrt(dot(n, n));
        if (!(nl > 0))
        {
            return false;
        }
        const vec3 unit = (1 / nl) * n;
        bool above = false, below = false;
        for (const vec3& c : corners)
        {
            const double s = dot(unit, c - q.q);
            above = above || s > tolerance;
            below = below || s < -tolerance;
        }
        if (above && below)
        {
            return false;
        }
    }
    return true;
}

class translator
{
public:
    flat_scene out;

    void add_world(const hittable_list& world)
    {
        for_each_shape(world, placement{}, [this](const hittable& s, const placement& w) { add_shape(s, w); });
    }
    void add_lights(const hittable& object, const placement& where, double weight)
    {
        if (const auto* list = dynamic_cast<const hittable_list*>(&object))
        {
            if (list->objects.empty())
            {
                throw std::runtime_error("an empty lights list");
            }
            for (const auto& child : list->objects)
            {
                add_lights(*child, where, weight / static_cast<double>(list->objects.size()));
            }
            return;
        }
        if (const auto* moved = dynamic_cast<const translate*>(&object))
        {
            add_lights(*moved->object, where.inside(*moved), weight);
            return;
        }
        if (const auto* turned = dynamic_cast<const rotate_y*>(&object))
        {
            add_lights(*turned->object, where.inside(*turned), weight);
            return;
        }
        if (const auto* face = dynamic_cast<const quad*>(&object))
        {
            flat_light l;
            l.kind = 0;
            l.quad = {where.place(face->q), where.turn(face->u), where.turn(face->v), 0};
            l.weight = weight;
            out.lights.push_back(l);
            return;
        }
        if (const auto* ball = dynamic_cast<const sphere*>(&object))
        {
            flat_light l;
            l.kind = 1;
            l.center = where.place(ball->center_at(0));
            l.radius = ball->radius;
            l.weight = weight;
            out.lights.push_back(l);
            return;
        }
        throw std::runtime_error("a light the GPU can't aim at");
    }

private:
    std::unordered_map<const material*, int> material_indices_;
    std::unordered_map<const texture*, int> texture_indices_;

    void add_shape(const hittable& shape, const placement& where)
    {
        if (const auto* ball = dynamic_cast<const sphere*>(&shape))
        {
            flat_sphere s;
            s.center = where.place(ball->center_at(0));
            s.radius = ball->radius;
            s.material = material_index(ball->mat.get());
            s.displacement = where.turn(ball->center_at(1) - ball->center_at(0));
            s.uv_cos = where.cos_theta;
            s.uv_sin = where.sin_theta;
            out.spheres.push_back(s);
            return;
        }
        if (const auto* face = dynamic_cast<const quad*>(&shape))
        {
            const int m = material_index(face->mat.get());
            out.quads.push_back({where.place(face->q), where.turn(face->u), where.turn(face->v), m});
            return;
        }
        if (const auto* volume = dynamic_cast<const constant_medium*>(&shape))
        {
            flat_medium m;
            m.material = material_index(volume->phase_function.get());
            m.neg_inv_density = volume->neg_inv_density;
            int spheres = 0;
            for_each_shape(*volume->boundary, where, [&](const hittable& part, const placement& at) {
                if (const auto* b = dynamic_cast<const sphere*>(&part))
                {
                    ++spheres;
                    m.sphere.center = at.place(b->center_at(0));
                    m.sphere.radius = b->radius;
                    m.sphere.material = 0;
                    m.sphere.displacement = at.turn(b->center_at(1) - b->center_at(0));
                    return;
                }
                if (const auto* q = dynamic_cast<const quad*>(&part))
                {
                    m.quads.push_back({at.place(q->q), at.turn(q->u), at.turn(q->v), 0});
                    return;
                }
                throw std::runtime_error("a boundary of other shapes");
            });
            if (spheres == 1 && m.quads.empty())
            {
                m.boundary = 0;
            }
            else if (spheres == 0 && !m.quads.empty() && is_convex(m.quads))
            {
                m.boundary = 1;
            }
            else
            {
                throw std::runtime_error("a boundary that isn't one sphere or convex quads");
            }
            out.media.push_back(std::move(m));
            return;
        }
        throw std::runtime_error("an object the GPU doesn't know");
    }
    int material_index(const material* mat)
    {
        if (!mat)
        {
            throw std::runtime_error("an object without a material");
        }
        if (const auto f = material_indices_.find(mat); f != material_indices_.end())
        {
            return f->second;
        }
        flat_material d{};
        if (const auto* m = dynamic_cast<const lambertian*>(mat))
        {
            d.kind = 0;
            d.texture = texture_index(*m->tex);
        }
        else if (const auto* m = dynamic_cast<const metal*>(mat))
        {
            d.kind = 1;
            d.albedo = m->albedo;
            d.fuzz = m->fuzz;
        }
        else if (const auto* m = dynamic_cast<const dielectric*>(mat))
        {
            d.kind = 2;
            d.refraction_index = m->refraction_index;
        }
        else if (const auto* m = dynamic_cast<const diffuse_light*>(mat))
        {
            d.kind = 3;
            d.texture = texture_index(*m->tex);
        }
        else if (const auto* m = dynamic_cast<const isotropic*>(mat))
        {
            d.kind = 4;
            d.texture = texture_index(*m->tex);
        }
        const int index = static_cast<int>(out.materials.size());
        out.materials.push_back(d);
        material_indices_.emplace(mat, index);
        return index;
    }
    int texture_index(const texture& tex)
    {
        if (const auto f = texture_indices_.find(&tex); f != texture_indices_.end())
        {
            return f->second;
        }
        flat_texture d{};
        if (const auto* t = dynamic_cast<const solid_color*>(&tex))
        {
            d.kind = 0;
            d.rgb = t->albedo;
        }
        else if (const auto* t = dynamic_cast<const checker_texture*>(&tex))
        {
            const int even = texture_index(*t->even);
            const int odd = texture_index(*t->odd);
            d.kind = 1;
            d.inv_scale = t->inv_scale;
            d.even = even;
            d.odd = odd;
        }
        else if (const auto* t = dynamic_cast<const image_texture*>(&tex))
        {
            d.kind = 2;
            d.image = static_cast<int>(out.images.size());
            out.images.push_back(&t->img);
        }
        else if (const auto* t = dynamic_cast<const noise_texture*>(&tex))
        {
            d.kind = 3;
            d.scale = t->scale;
            d.perlin = static_cast<int>(out.perlins.size());
            out.perlins.push_back(&t->noise);
        }
        const int index = static_cast<int>(out.textures.size());
        out.textures.push_back(d);
        texture_indices_.emplace(&tex, index);
        return index;
    }
};

// A scene built as build_gpu_scene builds it (src/gpu_scene.ixx:1333-1445): from the default seed of mt19937_64,
// with the overrides applied, flattened. The CPU scene is returned too (its objects own the images and tables).
struct built_scene
{
    scene cpu;
    flat_scene flat;
};
inline built_scene build_flat(const std::string& name, int width, int spp, int sampling, int max_depth,
                              const scene_options& opts)
{
    seed_random(std::mt19937_64::default_seed);
    built_scene b{build_scene(name, opts), {}};
    if (width > 0)
    {
        b.cpu.camera.image_width = width;
    }
    if (spp > 0)
    {
        b.cpu.camera.samples_per_pixel = spp;
    }
    if (sampling >= 0 && sampling <= 3)
    {
        b.cpu.camera.sampling = static_cast<sampling_strategy>(sampling);
    }
    if (max_depth > 0)
    {
        b.cpu.camera.max_depth = max_depth;
    }
    translator t;
    t.add_world(b.cpu.world);
    t.out.frame = frame_of(b.cpu.camera);
    t.out.sampling = effective_sampling(b.cpu.camera.sampling, b.cpu.lights);
    if (t.out.sampling == sampling_strategy::lights || t.out.sampling == sampling_strategy::mixture)
    {
        t.add_lights(b.cpu.lights, placement{}, 1.0);
    }
    b.flat = std::move(t.out);
    return b;
}
}
