// The quality integrator's device functions (float): light samples at every bounce (next-event estimation) weighed
// against the material's own samples by multiple importance sampling, GGX rough metals and glossy coats, the lamps of
// the world as a light set, the environment map with its alias table, and shadow rays (any-hit traversal plus the
// media's transmittance). The reference's semantics are cited per function: src/gpu/quality_path.cuh, bsdf.cuh,
// ggx.cuh, mis.cuh, lamps.cuh, light_sampler.cuh, environment.cuh, background.cuh, shadow.cuh.
//
// Under the material combination (-p material) the integrator draws exactly the books' numbers in the books' order,
// so its images equal trace_all_books's; under nee and mis the expected image is the same, with less noise.
#pragma once

#include "ggx_energy_table.hpp"
#include "path_tracer.hpp"

namespace rt::dev
{
// --- multiple importance sampling (src/gpu/mis.cuh:30-62) -----------------------------------------------------------
// Veach's power heuristic, self^2 / (self^2 + other^2), as 1 / (1 + r^2) with r = other / self: no density is squared
__device__ inline float power_weight(float self, float other)
{
    if (!(self > 0.0f))
    {
        return 0.0f;
    }
    const float r = other / self;
    const float w = 1.0f / fmaf(r, r, 1.0f);
    return w == w ? w : 0.0f;
}
__device__ inline float light_weight(int combination, float p_light, float p_bsdf)
{
    return combination == RT_COMBINE_POWER_HEURISTIC ? power_weight(p_light, p_bsdf) : 1.0f;
}
__device__ inline float bsdf_weight(int combination, float p_bsdf, float p_light)
{
    if (combination == RT_COMBINE_BSDF_ONLY || !(p_bsdf > 0.0f) || !(p_light > 0.0f))
    {
        return 1.0f;
    }
    return combination == RT_COMBINE_POWER_HEURISTIC ? power_weight(p_bsdf, p_light) : 0.0f;
}

// --- GGX (src/gpu/ggx.cuh: Walter 2007, Heitz 2014/2018, Dupuy-Benyoub 2023, Turquin 2019) ---------------------------
constexpr float min_ggx_alpha = 1e-3f;

__device__ inline vec3<float> cross(vec3<float> a, vec3<float> b)
{
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
// A(c) = sqrt(c^2 (1 - alpha^2) + alpha^2): Smith's Lambda for GGX is (A(c)/c - 1)/2
__device__ inline float smith_a(float c, float alpha2) { return sqrtf(fmaf(c * c, 1.0f - alpha2, alpha2)); }
// D(h) = alpha^2 / (pi (alpha^2 c_h^2 + sin^2 theta_h)^2), sin^2 from |n x h|^2
__device__ inline float ggx_d(vec3<float> n, vec3<float> h, float alpha2)
{
    const float c = dot(n, h);
    if (!(c > 0.0f))
    {
        return 0.0f;
    }
    const vec3<float> s = cross(n, h);
    const float d = fmaf(alpha2 * c, c, dot(s, s));
    return alpha2 * inverse_pi<float> / (d * d);
}
__device__ inline float ggx_g1(float c, float alpha2) { return 2.0f * c / (c + smith_a(c, alpha2)); }
__device__ inline float ggx_g2(float cv, float cl, float alpha2)
{
    return 2.0f * cv * cl / (cl * smith_a(cv, alpha2) + cv * smith_a(cl, alpha2));
}
// Schlick's F = F0 + (1 - F0)(1 - cos)^5 per channel
__device__ inline vec3<float> schlick(vec3<float> f0, float cosine)
{
    const float x = fmaxf(0.0f, 1.0f - cosine);
    const float x2 = x * x;
    const float x5 = x2 * x2 * x;
    return {fmaf(1.0f - f0.x, x5, f0.x), fmaf(1.0f - f0.y, x5, f0.y), fmaf(1.0f - f0.z, x5, f0.z)};
}
// f(v, l)·c_l in the height-correlated "V form", and the visible-normals density of l: D(h) / (2 (c_v + A_v))
__device__ inline vec3<float> ggx_eval(vec3<float> n, vec3<float> v, vec3<float> l, float alpha, vec3<float> f0,
                                       float& pdf)
{
    pdf = 0.0f;
    const float cv = dot(n, v);
    const float cl = dot(n, l);
    const vec3<float> sum = v + l;
    const float len2 = length_squared(sum);
    if (!(cv > 0.0f) || !(cl > 0.0f) || !(len2 > 0.0f))
    {
        return {0.0f, 0.0f, 0.0f};
    }
    const vec3<float> h = rsqrtf(len2) * sum;
    const float alpha2 = alpha * alpha;
    const float d = ggx_d(n, h, alpha2);
    const float av = smith_a(cv, alpha2);
    const float al = smith_a(cl, alpha2);
    pdf = d / (2.0f * (cv + av));
    return (d * cl / (2.0f * (cl * av + cv * al))) * schlick(f0, dot(v, h));
}
// The visible normals by spherical caps: two numbers (azimuth, then height), whatever the outcome
__device__ inline bool ggx_sample(vec3<float> n, vec3<float> v, float alpha, vec3<float> f0, pcg32& g,
                                  vec3<float>& l, vec3<float>& weight, float& pdf)
{
    const float u1 = random_real<float>(g);
    const float u2 = random_real<float>(g);
    const float cv = dot(n, v);
    if (!(cv > 0.0f))
    {
        l = {0.0f, 0.0f, 0.0f};
        return false;
    }
    // the view squeezed into the hemispheres' space: its tangential part times alpha
    const vec3<float> vs = unit_vector(alpha * (v - cv * n) + cv * n);
    const float zs = dot(n, vs);
    // uniform in the cap of heights (-z_s, 1]
    const float z = fmaf(1.0f - u2, 1.0f + zs, -zs);
    const float sin_theta = sqrtf(fmaxf(0.0f, 1.0f - z * z));
    float sin_phi, cos_phi;
    sincospif(2.0f * u1, &sin_phi, &cos_phi);
    const vec3<float> c = around(n, vec3<float>{sin_theta * cos_phi, sin_theta * sin_phi, z});
    const vec3<float> hs = c + vs;
    const float hn = dot(n, hs);
    const vec3<float> m = alpha * (hs - hn * n) + fmaxf(0.0f, hn) * n;
    const float m2 = length_squared(m);
    if (!(m2 > 0.0f))
    {
        l = {0.0f, 0.0f, 0.0f};
        return false;
    }
    const vec3<float> h = rsqrtf(m2) * m;
    const float vh = dot(v, h);
    l = (2.0f * vh) * h - v;
    const float cl = dot(n, l);
    if (!(cl > 0.0f) || !(vh > 0.0f))
    {
        return false;
    }
    const float alpha2 = alpha * alpha;
    const float av = smith_a(cv, alpha2);
    const float al = smith_a(cl, alpha2);
    weight = (cl * (cv + av) / (cl * av + cv * al)) * schlick(f0, vh);
    pdf = ggx_d(n, h, alpha2) / (2.0f * (cv + av));
    return true;
}
// E_ss(mu, r) of a white conductor, bilinearly from the table (nodes at sqrt(mu) = i/31, r = j/31)
__device__ inline float ggx_energy(float cos_v, float roughness)
{
    constexpr int last = ggx_energy_size - 1;
    const float x = sqrtf(fminf(fmaxf(cos_v, 0.0f), 1.0f)) * static_cast<float>(last);
    const float y = fminf(fmaxf(roughness, 0.0f), 1.0f) * static_cast<float>(last);
    const int i = static_cast<int>(x) < last - 1 ? static_cast<int>(x) : last - 1;
    const int j = static_cast<int>(y) < last - 1 ? static_cast<int>(y) : last - 1;
    const float fx = x - static_cast<float>(i);
    const float fy = y - static_cast<float>(j);
    const float* row = ggx_energy_table + j * ggx_energy_size + i;
    const float below = (1.0f - fx) * row[0] + fx * row[1];
    const float above = (1.0f - fx) * row[ggx_energy_size] + fx * row[ggx_energy_size + 1];
    return (1.0f - fy) * below + fy * above;
}
// Turquin's multiple-scattering factor 1 + F0 (1 - E)/E
__device__ inline vec3<float> ggx_compensation(vec3<float> f0, float cos_v, float roughness)
{
    const float e = fminf(ggx_energy(cos_v, roughness), 1.0f);
    const float lost = (1.0f - e) / e;
    return {fmaf(f0.x, lost, 1.0f), fmaf(f0.y, lost, 1.0f), fmaf(f0.z, lost, 1.0f)};
}

constexpr float gloss_coat_f0 = 0.04f;
constexpr float gloss_diffuse_scale = static_cast<float>(28.0 / (23.0 * 3.14159265358979323846) * (1.0 - 0.04));
__device__ inline float gloss_crossing(float c)
{
    const float x = 1.0f - 0.5f * c;
    const float x2 = x * x;
    return 1.0f - x2 * x2 * x;
}
// The coat (GGX, F0 0.04) plus Ashikhmin-Shirley's coupled diffuse base; the density is the half-half mixture's
__device__ inline vec3<float> gloss_eval(vec3<float> n, vec3<float> v, vec3<float> l, float alpha, vec3<float> rho,
                                         float& pdf)
{
    pdf = 0.0f;
    const float cv = dot(n, v);
    const float cl = dot(n, l);
    if (!(cv > 0.0f) || !(cl > 0.0f))
    {
        return {0.0f, 0.0f, 0.0f};
    }
    float coat_pdf;
    const vec3<float> coat = ggx_eval(n, v, l, alpha, vec3<float>{gloss_coat_f0, gloss_coat_f0, gloss_coat_f0}, coat_pdf);
    pdf = 0.5f * (coat_pdf + cl * inverse_pi<float>);
    const float base = gloss_diffuse_scale * (gloss_crossing(cl) * gloss_crossing(cv)) * cl;
    return coat + base * rho;
}
__device__ inline bool gloss_sample(vec3<float> n, vec3<float> v, float alpha, vec3<float> rho, pcg32& g,
                                    vec3<float>& l, vec3<float>& weight, float& pdf)
{
    if (random_real<float>(g) < 0.5f)
    {
        vec3<float> coat_weight;
        float coat_pdf;
        if (!ggx_sample(n, v, alpha, vec3<float>{gloss_coat_f0, gloss_coat_f0, gloss_coat_f0}, g, l, coat_weight,
                        coat_pdf))
        {
            return false;
        }
    }
    else
    {
        l = around(n, random_cosine_direction<float>(g));
        if (!(dot(n, l) > 0.0f))
        {
            return false;
        }
    }
    const vec3<float> f = gloss_eval(n, v, l, alpha, rho, pdf);
    if (!(pdf > 0.0f))
    {
        return false;
    }
    weight = (1.0f / pdf) * f;
    return true;
}

// --- the BSDF lobes (src/gpu/bsdf.cuh:27-201) --------------------------------------------------------------------------
enum class lobe : int
{
    lambertian,
    isotropic,
    conductor,
    gloss,
    smooth_conductor,
    books_delta,
};
struct bsdf
{
    lobe kind;
    vec3<float> albedo;
    float alpha;
    vec3<float> n;
    vec3<float> wo;
    float roughness;
};
struct bsdf_sample
{
    vec3<float> wi;
    vec3<float> weight;
    float pdf;
    bool delta;
};
__device__ inline bsdf make_bsdf(const material_rec<float>& m, const microfacet_rec* side, vec3<float> c,
                                 const hit_record<float>& rec, vec3<float> ray_direction)
{
    bsdf b{};
    b.n = rec.normal;
    b.albedo = c;
    if (side && side->kind == RT_MICROFACET_CONDUCTOR)
    {
        b.wo = -unit_vector(ray_direction);
        b.kind = side->alpha >= min_ggx_alpha ? lobe::conductor : lobe::smooth_conductor;
        b.albedo = load3(m.albedo);
        b.alpha = side->alpha;
        b.roughness = side->roughness;
        return b;
    }
    if (side && side->kind == RT_MICROFACET_GLOSS)
    {
        b.wo = -unit_vector(ray_direction);
        b.kind = lobe::gloss;
        b.alpha = side->alpha;
        return b;
    }
    b.kind = m.kind == RT_MATERIAL_LAMBERTIAN ? lobe::lambertian
             : m.kind == RT_MATERIAL_ISOTROPIC ? lobe::isotropic
                                               : lobe::books_delta;
    return b;
}
__device__ inline bool is_delta(const bsdf& b) { return b.kind >= lobe::smooth_conductor; }

__device__ inline bool sample_bsdf(const bsdf& b, const material_rec<float>& m, vec3<float> direction,
                                   const hit_record<float>& rec, pcg32& g, bsdf_sample& out)
{
    switch (b.kind)
    {
    case lobe::lambertian:
        out.wi = around(b.n, random_cosine_direction<float>(g));
        if (!(dot(b.n, out.wi) > 0.0f))
        {
            return false;
        }
        out.weight = b.albedo;
        out.pdf = cosine_density(b.n, out.wi);
        out.delta = false;
        return true;
    case lobe::isotropic:
        out.wi = random_unit_vector<float>(g);
        out.weight = b.albedo;
        out.pdf = uniform_sphere_density<float>;
        out.delta = false;
        return true;
    case lobe::books_delta:
    {
        vec3<float> weight{1.0f, 1.0f, 1.0f};
        if (!scatter(m, direction, rec, g, out.wi, weight))
        {
            return false;
        }
        out.weight = weight;
        out.pdf = 0.0f;
        out.delta = true;
        return true;
    }
    case lobe::conductor:
        out.delta = false;
        if (!ggx_sample(b.n, b.wo, b.alpha, b.albedo, g, out.wi, out.weight, out.pdf))
        {
            return false;
        }
        out.weight = out.weight * ggx_compensation(b.albedo, dot(b.n, b.wo), b.roughness);
        return true;
    case lobe::smooth_conductor:
        if (!(dot(b.n, b.wo) > 0.0f))
        {
            return false;
        }
        out.wi = unit_vector(reflect(-b.wo, b.n));
        out.weight = schlick(b.albedo, dot(b.n, b.wo));
        out.pdf = 0.0f;
        out.delta = true;
        return true;
    case lobe::gloss:
        out.delta = false;
        return gloss_sample(b.n, b.wo, b.alpha, b.albedo, g, out.wi, out.weight, out.pdf);
    }
    return false;
}
__device__ inline vec3<float> eval_bsdf(const bsdf& b, vec3<float> wi, float& pdf)
{
    switch (b.kind)
    {
    case lobe::lambertian:
        pdf = cosine_density(b.n, wi);
        return pdf * b.albedo;
    case lobe::isotropic:
        pdf = uniform_sphere_density<float>;
        return uniform_sphere_density<float> * b.albedo;
    case lobe::conductor:
    {
        const vec3<float> single = ggx_eval(b.n, b.wo, wi, b.alpha, b.albedo, pdf);
        return ggx_compensation(b.albedo, dot(b.n, b.wo), b.roughness) * single;
    }
    case lobe::gloss:
        return gloss_eval(b.n, b.wo, wi, b.alpha, b.albedo, pdf);
    default:
        pdf = 0.0f;
        return {0.0f, 0.0f, 0.0f};
    }
}

// --- lamps (src/gpu/lamps.cuh:55-146) ------------------------------------------------------------------------------------
struct lamp_sample
{
    vec3<float> wi;
    float distance;
    float pdf;
};
__device__ inline float quad_lamp_density(float distance_squared, float cos_light, float area)
{
    return distance_squared / (cos_light * area);
}
__device__ inline float sphere_lamp_density(float radius, float distance_squared)
{
    return 1.0f / (2.0f * pi_v<float> * one_minus_cos_theta_max(radius, distance_squared));
}
// Exactly two numbers in every call
__device__ inline bool sample_lamp(const emitter_rec& e, vec3<float> x, pcg32& g, lamp_sample& out)
{
    const float a = random_real<float>(g);
    const float b = random_real<float>(g);
    const vec3<float> corner = load3(e.corner);
    if (e.kind == RT_LIGHT_QUAD)
    {
        const vec3<float> to = (corner - x) + a * load3(e.u) + b * load3(e.v);
        const float d2 = length_squared(to);
        const float inv = rsqrtf(d2);
        out.wi = inv * to;
        out.distance = d2 * inv;
        const float cos_light = -dot(out.wi, load3(e.normal));
        if (!(cos_light > 0.0f))
        {
            return false;
        }
        out.pdf = quad_lamp_density(d2, cos_light, e.size);
        return true;
    }
    const vec3<float> to_center = corner - x;
    const float d2 = sphere_light_distance_squared(to_center);
    const float r2 = e.size * e.size;
    if (!(d2 > r2))
    {
        return false;
    }
    const float omc = one_minus_cos_theta_max(e.size, d2);
    float sin_phi, cos_phi;
    sincospif(2.0f * a, &sin_phi, &cos_phi);
    const float omz = b * omc;
    const float sin2 = omz * (2.0f - omz);
    const float sin_theta = sqrtf(sin2);
    const float cos_theta = 1.0f - omz;
    const float inv = rsqrtf(d2);
    out.wi = around(inv * to_center, vec3<float>{cos_phi * sin_theta, sin_phi * sin_theta, cos_theta});
    const float distance = d2 * inv;
    out.distance = distance * cos_theta - sqrtf(fmaxf(0.0f, fmaf(-d2, sin2, r2)));
    out.pdf = sphere_lamp_density(e.size, d2);
    return true;
}
__device__ inline float lamp_pdf(const emitter_rec& e, vec3<float> x, vec3<float> unit_dir, float distance)
{
    if (e.kind == RT_LIGHT_QUAD)
    {
        const float cos_light = -dot(unit_dir, load3(e.normal));
        return cos_light > 0.0f ? quad_lamp_density(distance * distance, cos_light, e.size) : 0.0f;
    }
    const float d2 = sphere_light_distance_squared(load3(e.corner) - x);
    return d2 > e.size * e.size ? sphere_lamp_density(e.size, d2) : 0.0f;
}

// --- the environment map (src/gpu/environment.cuh, src/gpu/background.cuh) --------------------------------------------
__device__ inline vec3<float> to_map(const quality_view& e, vec3<float> d)
{
    return {e.env_cos_yaw * d.x + e.env_sin_yaw * d.z, d.y, -e.env_sin_yaw * d.x + e.env_cos_yaw * d.z};
}
__device__ inline vec3<float> from_map(const quality_view& e, vec3<float> d)
{
    return {e.env_cos_yaw * d.x - e.env_sin_yaw * d.z, d.y, e.env_sin_yaw * d.x + e.env_cos_yaw * d.z};
}
// u = 0.5 + atan2(m_z, m_x)/(2 pi), t = atan2(sqrt(m_x^2 + m_z^2), m_y)/pi: +y up, seen from inside
__device__ inline void map_coordinates(vec3<float> m, float& u, float& t)
{
    const float phi = atan2f(m.z, m.x);
    const float theta = atan2f(sqrtf(m.x * m.x + m.z * m.z), m.y);
    u = 0.5f + phi * (0.5f * inverse_pi<float>);
    t = theta * inverse_pi<float>;
}
__device__ inline int clamp_index(int k, int n) { return k < 0 ? 0 : (k < n ? k : n - 1); }
// The one map from a direction to its cell: the radiance and the density of a direction come from the same texel
__device__ inline int env_cell(const quality_view& e, vec3<float> world_dir)
{
    float u, t;
    map_coordinates(to_map(e, world_dir), u, t);
    const int i = clamp_index(static_cast<int>(u * static_cast<float>(e.env_width)), e.env_width);
    const int j = clamp_index(static_cast<int>(t * static_cast<float>(e.env_height)), e.env_height);
    return j * e.env_width + i;
}
// A direction at fractions (s_theta in cos theta, s_phi in phi) of cell (i, j): uniform over the cell's solid angle
__device__ inline vec3<float> cell_direction(const quality_view& e, int i, int j, float s_theta, float s_phi)
{
    const float top = e.env_row_cos[j];
    const float bottom = e.env_row_cos[j + 1];
    const float cos_theta = top - s_theta * (top - bottom);
    const float sin_theta = sqrtf(fmaxf(0.0f, (1.0f - cos_theta) * (1.0f + cos_theta)));
    float sin_phi, cos_phi;
    sincospif(2.0f * (static_cast<float>(i) + s_phi) / static_cast<float>(e.env_width) - 1.0f, &sin_phi, &cos_phi);
    return from_map(e, vec3<float>{sin_theta * cos_phi, cos_theta, sin_theta * sin_phi});
}
__device__ inline vec3<float> env_radiance(const quality_view& e, vec3<float> dir, float& pdf)
{
    const env_texel t = e.env_texels[env_cell(e, dir)];
    pdf = t.pdf;
    return e.env_scale * vec3<float>{t.r, t.g, t.b};
}
__device__ inline vec3<float> env_lerp(vec3<float> a, vec3<float> b, float s) { return a + s * (b - a); }
// Bilinear between the four nearest texel centres (columns wrap, rows clamp): what camera rays and rays after a delta
// vertex see
__device__ inline vec3<float> env_radiance_filtered(const quality_view& e, vec3<float> dir)
{
    float u, t;
    map_coordinates(to_map(e, dir), u, t);
    const float x = u * static_cast<float>(e.env_width) - 0.5f;
    const float y = t * static_cast<float>(e.env_height) - 0.5f;
    const float x0 = floorf(x);
    const float y0 = floorf(y);
    const float fx = x - x0;
    const float fy = y - y0;
    const int i = static_cast<int>(x0);
    const int i0 = i < 0 ? i + e.env_width : i;
    const int i1 = i0 + 1 < e.env_width ? i0 + 1 : 0;
    const int j = static_cast<int>(y0);
    const int j0 = clamp_index(j, e.env_height) * e.env_width;
    const int j1 = clamp_index(j + 1, e.env_height) * e.env_width;
    const auto col = [&](int k) {
        const env_texel t = e.env_texels[k];
        return vec3<float>{t.r, t.g, t.b};
    };
    const vec3<float> top = env_lerp(col(j0 + i0), col(j0 + i1), fx);
    const vec3<float> bottom = env_lerp(col(j1 + i0), col(j1 + i1), fx);
    return e.env_scale * env_lerp(top, bottom, fy);
}
__device__ inline bool environment_sampled(const quality_view& e) { return e.env_present && e.env_bins != nullptr; }
__device__ inline vec3<float> background(const quality_view& q, const camera_rec<float>& cam, vec3<float> dir,
                                         bool filtered, float& pdf)
{
    pdf = 0.0f;
    if (!q.env_present)
    {
        return sky(cam, dir);
    }
    if (filtered)
    {
        return env_radiance_filtered(q, dir);
    }
    float texel_pdf;
    const vec3<float> le = env_radiance(q, dir, texel_pdf);
    pdf = environment_sampled(q) ? texel_pdf : 0.0f;
    return le;
}
// Four numbers: the bin (top bits of one output), its coin, then cos theta and phi inside the cell
__device__ inline bool sample_environment(const quality_view& e, pcg32& g, vec3<float>& wi, vec3<float>& le, float& pdf)
{
    if (!environment_sampled(e))
    {
        return false;
    }
    const int bin = static_cast<int>(g.next() >> (32 - e.env_log2_bins));
    const alias_bin b = e.env_bins[bin];
    const int cell = random_real<float>(g) < b.q ? bin : b.alias;
    const int j = cell / e.env_width;
    const int i = cell - j * e.env_width;
    const float s_theta = random_real<float>(g);
    const float s_phi = random_real<float>(g);
    wi = cell_direction(e, i, j, s_theta, s_phi);
    le = env_radiance(e, wi, pdf);
    return pdf > 0.0f;
}

// --- the light sampler (src/gpu/light_sampler.cuh:40-102) ------------------------------------------------------------
struct light_sample
{
    vec3<float> wi;
    float distance;
    vec3<float> le;
    float pdf;
    int skip;
};
__device__ inline int light_count(const quality_view& q)
{
    return q.lamp_count + (environment_sampled(q) ? 1 : 0);
}
// One number picks the light uniformly (lamps first, the map last); then 2 more for a lamp, 4 for the map
template <typename Tally>
__device__ inline bool sample_light(const quality_view& q, vec3<float> x, pcg32& g, light_sample& out, Tally& tl)
{
    const int n = light_count(q);
    if (n == 0)
    {
        return false;
    }
    const float xi = random_real<float>(g);
    int k = static_cast<int>(xi * static_cast<float>(n));
    k = k < n ? k : n - 1;
    if (k < q.lamp_count)
    {
        const emitter_rec e = q.lamps[k];
        tl.add(tally_lamp);
        lamp_sample s;
        if (!sample_lamp(e, x, g, s))
        {
            return false;
        }
        out.wi = s.wi;
        out.distance = s.distance;
        out.le = load3(e.radiance);
        out.pdf = q.light_pmf * s.pdf;
        out.skip = e.primitive;
        return true;
    }
    float pdf;
    tl.add(tally_alias_bin);
    tl.add(tally_env_texel);
    if (!sample_environment(q, g, out.wi, out.le, pdf))
    {
        return false;
    }
    out.distance = inf<float>;
    out.pdf = q.light_pmf * pdf;
    out.skip = -1;
    return true;
}
__device__ inline float lamp_hit_pdf(const quality_view& q, int id, vec3<float> x, vec3<float> unit_dir, float distance)
{
    if (q.lamp_count == 0)
    {
        return 0.0f;
    }
    const int k = q.lamp_of[id];
    return k >= 0 ? q.light_pmf * lamp_pdf(q.lamps[k], x, unit_dir, distance) : 0.0f;
}

// --- shadow rays (src/gpu/shadow.cuh:26-165) -------------------------------------------------------------------------
// Anything in (0.001, t_max) but the primitive `skip`? closest_hit's walk with a fixed t_max, ending at the first hit
template <typename Tally, typename Stack>
__device__ inline bool occluded(const scene_view<float>& sc, const ray<float>& r, float t_max, int from, bool inward,
                                int skip, const Stack& stack, Tally& tl)
{
    const float t_min = 0.001f;
    if (sc.stack_depth == 0)
    {
        return false;
    }
    const float inverse_a = 1.0f / length_squared(r.direction);
    const vec3<float> inv{safe_inverse(r.direction.x), safe_inverse(r.direction.y), safe_inverse(r.direction.z)};
    const vec3<float> os{r.origin.x * inv.x, r.origin.y * inv.y, r.origin.z * inv.z};
    const slab_order so = slab_order_of(inv);
    int size = 0;
    int32_t ref = sc.root_ref;
    while (true)
    {
        if (ref < 0)
        {
            const int kind = leaf_kind(ref);
            const int first = leaf_first(ref);
            const int end = first + leaf_count(ref);
            if (kind == prim_quad)
            {
                for (int i = first; i < end; ++i)
                {
                    const int id = sc.quad_base + i;
                    if (id == from || id == skip)
                    {
                        continue;
                    }
                    tl.add(tally_quad);
                    float t;
                    bool front;
                    if (hit_quad(sc.quads[i], r, t_min, t_max, t, front))
                    {
                        return true;
                    }
                }
            }
            else if (kind == prim_triangle)
            {
                for (int i = first; i < end; ++i)
                {
                    const int id = sc.triangle_base + i;
                    if (id == from || id == skip)
                    {
                        continue;
                    }
                    tl.add(tally_triangle);
                    float t;
                    bool front;
                    if (hit_triangle(sc.triangles[i], r, t_min, t_max, t, front))
                    {
                        return true;
                    }
                }
            }
            else
            {
                const bool moving = kind == prim_moving_sphere;
                for (int i = first; i < end; ++i)
                {
                    const int id = moving ? sc.moving_base + i : i;
                    const bool own = id == from;
                    if ((own && !inward) || id == skip)
                    {
                        continue;
                    }
                    tl.add(moving ? tally_moving : tally_sphere);
                    const sphere<float> s = moving ? moving_at(sc.moving_spheres[i], r.time) : load_sphere(sc.spheres, i);
                    float t;
                    bool entering;
                    if (hit_sphere(s, r, inverse_a, t_min, t_max, own, t, entering))
                    {
                        return true;
                    }
                }
            }
        }
        else
        {
            // any entered child will do: the first one entered is next, the others go on the stack
            tl.add(tally_node);
            float t[wide_children];
            child_entries(sc, ref, so, inv, os, t_min, t_max, t);
            const int4 refs = child_refs(sc, ref);
            const int32_t child[wide_children] = {refs.x, refs.y, refs.z, refs.w};
            int32_t next = no_child;
#pragma unroll
            for (int c = 0; c < wide_children; ++c)
            {
                if (t[c] < inf<float>)
                {
                    if (next == no_child)
                    {
                        next = child[c];
                    }
                    else
                    {
                        stack.store(size++, child[c], t_min);
                    }
                }
            }
            if (next != no_child)
            {
                ref = next;
                continue;
            }
        }
        if (size == 0)
        {
            return false;
        }
        --size;
        ref = stack.ref_at(size);
    }
}
// The chance the segment crosses every medium without a collision: e^(-sum rho L), the expected value of the media's
// free-flight test over the same segments (no random number drawn)
template <typename Tally>
__device__ inline float transmittance(const scene_view<float>& sc, const ray<float>& r, float t_max, Tally& tl)
{
    if (sc.medium_count == 0)
    {
        return 1.0f;
    }
    const float t_min = 0.001f;
    const float inverse_a = 1.0f / length_squared(r.direction);
    const vec3<float> inv{safe_inverse(r.direction.x), safe_inverse(r.direction.y), safe_inverse(r.direction.z)};
    const vec3<float> os{r.origin.x * inv.x, r.origin.y * inv.y, r.origin.z * inv.z};
    const float length = sqrtf(length_squared(r.direction));
    float exponent = 0.0f;
    for (int k = 0; k < sc.medium_count; ++k)
    {
        const medium_rec<float>& m = sc.media[k];
        tl.add(tally_medium);
        if (box_entry(m.box_min, m.box_max, inv, os, t_min, t_max) == inf<float>)
        {
            continue;
        }
        float line_in, line_out;
        if (!boundary_crossings(m, sc.boundary_quads, r, inverse_a, line_in, line_out, tl))
        {
            continue;
        }
        float t_in = fmaxf(line_in, t_min);
        const float t_out = fminf(line_out, t_max);
        if (t_in >= t_out)
        {
            continue;
        }
        t_in = fmaxf(t_in, 0.0f);
        exponent += (t_out - t_in) * length / m.neg_inv_density;
    }
    return expf(exponent);
}

// --- the quality path (src/gpu/quality_path.cuh:27-188) ---------------------------------------------------------------
struct quality_path
{
    path<float> base;
    vec3<float> radiance;
    float bsdf_pdf;
    bool specular;
};

__device__ inline float largest_channel(vec3<float> c) { return fmaxf(fmaxf(c.x, c.y), c.z); }

template <typename Tally>
__device__ inline bool light_contribution(const scene_view<float>& sc, const bsdf& b, vec3<float> throughput,
                                          vec3<float> x, pcg32& g, light_sample& ls, vec3<float>& contribution,
                                          Tally& tl)
{
    if (!sample_light(sc.quality, x, g, ls, tl))
    {
        return false;
    }
    float p_bsdf;
    const vec3<float> f = eval_bsdf(b, ls.wi, p_bsdf);
    if (!(ls.pdf > 0.0f) || !(largest_channel(f) > 0.0f))
    {
        return false;
    }
    contribution = (light_weight(sc.quality.combination, ls.pdf, p_bsdf) / ls.pdf) * (throughput * (f * ls.le));
    const float sum = contribution.x + contribution.y + contribution.z;
    return sum > 0.0f && sum < inf<float>;
}

// What a hit (or a miss) does to a quality path: the part of trace_quality_ray after the closest event
template <typename Tally, typename Stack>
__device__ inline bool shade_quality(quality_path& q, const scene_view<float>& sc, const camera_rec<float>& cam,
                                     const hit<float>& h, pcg32& g, vec3<float>& sample, const Stack& stack,
                                     Tally& tl, uint32_t& shadow_rays)
{
    path<float>& p = q.base;
    const quality_view& qv = sc.quality;
    if (h.index < 0)
    {
        float map_pdf;
        if (qv.env_present)
        {
            tl.add(tally_env_texel, q.specular ? 4u : 1u);
        }
        const vec3<float> le = background(qv, cam, p.r.direction, q.specular, map_pdf);
        const float w =
            q.bsdf_pdf > 0.0f ? bsdf_weight(qv.combination, q.bsdf_pdf, qv.light_pmf * map_pdf) : 1.0f;
        sample = q.radiance + w * (p.throughput * le);
        return true;
    }
    const surface<float> s = surface_at(sc, p.r, h, tl);
    const material_rec<float> m = sc.materials[s.material];
    const bool colored = m.kind == RT_MATERIAL_LAMBERTIAN || m.kind == RT_MATERIAL_ISOTROPIC ||
                         (m.kind == RT_MATERIAL_DIFFUSE_LIGHT && s.rec.front_face);
    vec3<float> c{0.0f, 0.0f, 0.0f};
    if (colored)
    {
        c = s.texture_index < 0 ? load3(m.albedo) : texture_value(sc, s.texture_index, s.u, s.v, s.texture_point, tl);
    }
    if (m.kind == RT_MATERIAL_DIFFUSE_LIGHT)
    {
        float w = 1.0f;
        if (q.bsdf_pdf > 0.0f)
        {
            const float distance = h.t * sqrtf(length_squared(p.r.direction));
            tl.add(tally_lamp);
            w = bsdf_weight(qv.combination, q.bsdf_pdf,
                            lamp_hit_pdf(qv, h.index, p.r.origin, unit_vector(p.r.direction), distance));
        }
        sample = q.radiance + w * (p.throughput * c);
        return true;
    }
    const microfacet_rec* side = qv.microfacets ? &qv.microfacets[s.material] : nullptr;
    const bsdf b = make_bsdf(m, side, c, s.rec, p.r.direction);
    const bool nee = qv.combination != RT_COMBINE_BSDF_ONLY && p.rays_left > 1 && !is_delta(b) &&
                     light_count(qv) > 0 && largest_channel(p.throughput) > 0.0f;
    light_sample ls;
    vec3<float> contribution{0.0f, 0.0f, 0.0f};
    const bool lit = nee && light_contribution(sc, b, p.throughput, s.rec.p, g, ls, contribution, tl);
    bsdf_sample next;
    const bool scattered = sample_bsdf(b, m, p.r.direction, s.rec, g, next);
    if (lit)
    {
        ++shadow_rays;
        const ray<float> shadow{s.rec.p, ls.wi, p.r.time};
        const bool inward = dot(ls.wi, s.outward) < 0.0f;
        const float visible = occluded(sc, shadow, ls.distance, s.from, inward, ls.skip, stack, tl)
                                  ? 0.0f
                                  : transmittance(sc, shadow, ls.distance, tl);
        q.radiance = q.radiance + visible * contribution;
    }
    if (!scattered)
    {
        sample = q.radiance;
        return true;
    }
    p.throughput = p.throughput * next.weight;
    q.bsdf_pdf = nee ? next.pdf : 0.0f;
    q.specular = next.delta;
    if (qv.combination != RT_COMBINE_BSDF_ONLY && p.throughput.x == 0.0f && p.throughput.y == 0.0f &&
        p.throughput.z == 0.0f)
    {
        sample = q.radiance;
        return true;
    }
    if (--p.rays_left == 0)
    {
        sample = q.radiance;
        return true;
    }
    p.r = {s.rec.p, next.wi, p.r.time};
    p.from = s.from;
    p.inward = dot(next.wi, s.outward) < 0.0f;
    return false;
}

template <typename Tally, typename Stack>
__device__ inline bool trace_quality(quality_path& q, const scene_view<float>& sc, const camera_rec<float>& cam,
                                     pcg32& g, vec3<float>& sample, const Stack& stack, Tally& tl,
                                     uint32_t& shadow_rays)
{
    const hit<float> h = closest_event(sc, q.base.r, q.base.from, q.base.inward, g, stack, tl);
    return shade_quality(q, sc, cam, h, g, sample, stack, tl, shadow_rays);
}

__device__ inline quality_path start_quality(const camera_rec<float>& cam, bool motion, int i, int j, int pass,
                                             pcg32& g)
{
    quality_path q;
    q.base = start_path(cam, motion, i, j, pass, g);
    q.radiance = {0.0f, 0.0f, 0.0f};
    q.bsdf_pdf = 0.0f;
    q.specular = true;
    return q;
}
}
