// The quality integrator's device functions (float): next-event estimation at every vertex that isn't a mirror's,
// weighed against the scattering's own samples by multiple importance sampling; GGX rough conductors (with the
// energy lost to single scattering restored) and glossy coats over a matte base; the world's lamps and the
// environment map (alias-table importance sampling) as the light set; shadow rays by an any-hit walk and the media's
// transmittance. Written for this renderer from the published mathematics:
//   - GGX's normal distribution and the height-correlated Smith masking-shadowing (Walter et al. 2007, Heitz 2014),
//     evaluated in the vertex's tangent frame, where sin^2 theta_h = h_x^2 + h_y^2 holds its precision at small alpha;
//   - sampling of the visible normals by spherical caps (Dupuy and Benyoub 2023), in the hemisphere configuration of
//     the stretched view direction, in the same frame;
//   - the multiple-scattering compensation 1 + F0 (1 - E) / E (Turquin 2019), E from this repository's table;
//   - the glossy coat's coupled diffuse base (Ashikhmin and Shirley 2000);
//   - Veach's power heuristic (1995) for the two strategies' weights.
// The semantics an image depends on follow the reference (src/gpu/quality_path<Real>.cuh, bsdf.cuh, lamps.cuh,
// light_sampler.cuh, environment.cuh, background.cuh, shadow.cuh): which vertices take a light sample, the light
// set's uniform choice, which emitter hits count under each combination, the map's cells and densities, the camera's
// filtered view of the map.
#pragma once

#include "books.hpp"
#include "ggx_energy_table.hpp"

namespace rt::dev
{
template <typename Real>
__device__ inline Real max3(vec3<Real> c) { return m_max(m_max(c.x, c.y), c.z); }
// Three floats of a light-set table (lamps, the map) in the kernel's precision
template <typename Real>
__device__ inline vec3<Real> table3(const float* p)
{
    return {static_cast<Real>(p[0]), static_cast<Real>(p[1]), static_cast<Real>(p[2])};
}

// --- combining the two strategies -------------------------------------------------------------------------------------
// The power heuristic's weight of a strategy of density `mine` against one of density `theirs`: mine^2 / (mine^2 +
// theirs^2), formed as 1 / (1 + (theirs / mine)^2) so that no density is squared (densities of distant lamps
// overflow); 0 for a strategy that couldn't have drawn the direction, and 0 where the ratio is NaN
template <typename Real>
__device__ inline Real power_heuristic(Real mine, Real theirs)
{
    if (!(mine > Real(0.0)))
    {
        return Real(0.0);
    }
    const Real ratio = theirs / mine;
    const Real w = Real(1.0) / m_fma(ratio, ratio, Real(1.0));
    return w == w ? w : Real(0.0);
}
// The weights under the scene's combination (src/gpu/mis.cuh:30-62): -p material counts only the scattering's
// samples of emitters, -p nee only the light samples (a scattered ray that meets a light the light set could have
// drawn counts 0), -p mis both, by the power heuristic. Densities of 0 mean "that strategy couldn't": a delta lobe's
// ray (bsdf density 0) keeps all of an emitter it meets.
struct combination
{
    int mode;
    template <typename Real>
    __device__ Real light_sample(Real p_light, Real p_bsdf) const
    {
        return mode == RT_COMBINE_POWER_HEURISTIC ? power_heuristic(p_light, p_bsdf) : Real(1.0);
    }
    template <typename Real>
    __device__ Real scattered_hit(Real p_bsdf, Real p_light) const
    {
        if (mode == RT_COMBINE_BSDF_ONLY || !(p_bsdf > Real(0.0)) || !(p_light > Real(0.0)))
        {
            return Real(1.0);
        }
        return mode == RT_COMBINE_POWER_HEURISTIC ? power_heuristic(p_bsdf, p_light) : Real(0.0);
    }
};

// --- GGX in the tangent frame (z: the normal) --------------------------------------------------------------------------
constexpr float min_ggx_alpha = 1e-3f; // below: the conductor is a perfect mirror

// D(h) = a^2 / (pi (a^2 h_z^2 + h_x^2 + h_y^2)^2) for a unit h with h_z > 0
template <typename Real>
__device__ inline Real ggx_density(vec3<Real> h, Real a2)
{
    if (!(h.z > Real(0.0)))
    {
        return Real(0.0);
    }
    const Real q = m_fma(a2 * h.z, h.z, m_fma(h.x, h.x, h.y * h.y));
    return a2 * one_over_pi<Real> / (q * q);
}
// sqrt(a^2 + (1 - a^2) c^2) = c (1 + 2 Lambda(c)): Smith's masking term of GGX at cos = c
template <typename Real>
__device__ inline Real smith_root(Real c, Real a2) { return m_sqrt(m_fma(Real(1.0) - a2, c * c, a2)); }
// Schlick's Fresnel per channel: F0 + (1 - F0) (1 - cos)^5
template <typename Real>
__device__ inline vec3<Real> fresnel_schlick(vec3<Real> f0, Real cos_d)
{
    const Real x = m_max(Real(0.0), Real(1.0) - cos_d);
    const Real x2 = x * x;
    const Real k = x2 * x2 * x;
    return {m_fma(Real(1.0) - f0.x, k, f0.x), m_fma(Real(1.0) - f0.y, k, f0.y), m_fma(Real(1.0) - f0.z, k, f0.z)};
}

// A GGX reflection of (local) v into l: f(v, l) cos_l = F D G2 / (4 cos_v) with the height-correlated
// G2 = 2 cos_v cos_l / (cos_l A(v) + cos_v A(l)), and the visible-normal density of l, G1(v) D / (4 cos_v) =
// D / (2 (cos_v + A(v)))
template <typename Real>
struct ggx_eval_result
{
    vec3<Real> value; // f cos_l
    Real pdf;
};
template <typename Real>
__device__ inline ggx_eval_result<Real> ggx_reflection(vec3<Real> v, vec3<Real> l, Real alpha, vec3<Real> f0)
{
    ggx_eval_result<Real> out{{Real(0.0), Real(0.0), Real(0.0)}, Real(0.0)};
    const vec3<Real> half = v + l;
    const Real len2 = length_squared(half);
    if (!(v.z > Real(0.0)) || !(l.z > Real(0.0)) || !(len2 > Real(0.0)))
    {
        return out;
    }
    const vec3<Real> h = m_rsqrt(len2) * half;
    const Real a2 = alpha * alpha;
    const Real d = ggx_density(h, a2);
    const Real av = smith_root(v.z, a2);
    const Real al = smith_root(l.z, a2);
    out.pdf = d / (Real(2.0) * (v.z + av));
    out.value = (d * l.z / (Real(2.0) * m_fma(l.z, av, v.z * al))) * fresnel_schlick(f0, dot(v, h));
    return out;
}
// A visible normal of (local) v by spherical caps (Dupuy and Benyoub 2023): v stretched to the hemisphere
// configuration, a point uniform on the cap of the unit sphere above -v_h.z's plane, added to v_h, then unstretched
template <typename Real>
__device__ inline vec3<Real> visible_normal(vec3<Real> v, Real alpha, Real u1, Real u2)
{
    const vec3<Real> vh = unit_vector(vec3<Real>{alpha * v.x, alpha * v.y, v.z});
    const Real z = m_fma(Real(1.0) - u2, Real(1.0) + vh.z, -vh.z);
    const Real sin_theta = m_sqrt(m_max(Real(0.0), Real(1.0) - z * z));
    Real s, c;
    m_sincospi(Real(2.0) * u1, &s, &c);
    const vec3<Real> m{sin_theta * c + vh.x, sin_theta * s + vh.y, z + vh.z};
    return vec3<Real>{alpha * m.x, alpha * m.y, m_max(Real(0.0), m.z)};
}
// One GGX draw from (local) v: the direction l, its weight f cos_l / pdf = F G2 / G1(v), its density; false when the
// reflected direction goes below the surface (two numbers are drawn either way)
template <typename Real>
__device__ inline bool ggx_draw(vec3<Real> v, Real alpha, vec3<Real> f0, pcg32& g, vec3<Real>& l,
                                vec3<Real>& weight, Real& pdf)
{
    const Real u1 = random_real<Real>(g);
    const Real u2 = random_real<Real>(g);
    if (!(v.z > Real(0.0)))
    {
        return false;
    }
    const vec3<Real> m = visible_normal(v, alpha, u1, u2);
    const Real m2 = length_squared(m);
    if (!(m2 > Real(0.0)))
    {
        return false;
    }
    const vec3<Real> h = m_rsqrt(m2) * m;
    const Real vh = dot(v, h);
    l = (Real(2.0) * vh) * h - v;
    if (!(l.z > Real(0.0)) || !(vh > Real(0.0)))
    {
        return false;
    }
    const Real a2 = alpha * alpha;
    const Real av = smith_root(v.z, a2);
    const Real al = smith_root(l.z, a2);
    weight = (l.z * (v.z + av) / m_fma(l.z, av, v.z * al)) * fresnel_schlick(f0, vh);
    pdf = ggx_density(h, a2) / (Real(2.0) * (v.z + av));
    return true;
}

// E(mu, r): the energy a white GGX conductor of roughness r reflects from cos_v = mu in one scattering, bilinear in the
// table's (sqrt(mu), r) grid of 32 x 32 nodes
template <typename Real>
__device__ inline Real single_scatter_energy(Real cos_v, Real roughness)
{
    constexpr int n = ggx_energy_size;
    const Real x = m_sqrt(m_min(m_max(cos_v, Real(0.0)), Real(1.0))) * static_cast<Real>(n - 1);
    const Real y = m_min(m_max(roughness, Real(0.0)), Real(1.0)) * static_cast<Real>(n - 1);
    const int i = min(static_cast<int>(x), n - 2);
    const int j = min(static_cast<int>(y), n - 2);
    const Real fx = x - static_cast<Real>(i);
    const Real fy = y - static_cast<Real>(j);
    const float* e = ggx_energy_table + j * n + i;
    const Real lower = e[0] + fx * (e[1] - e[0]);
    const Real upper = e[n] + fx * (e[n + 1] - e[n]);
    return lower + fy * (upper - lower);
}
// Turquin's factor for the light single scattering loses: 1 + F0 (1 - E) / E per channel
template <typename Real>
__device__ inline vec3<Real> multiple_scatter_factor(vec3<Real> f0, Real cos_v, Real roughness)
{
    const Real e = m_min(single_scatter_energy(cos_v, roughness), Real(1.0));
    const Real lost = (Real(1.0) - e) / e;
    return {m_fma(f0.x, lost, Real(1.0)), m_fma(f0.y, lost, Real(1.0)), m_fma(f0.z, lost, Real(1.0))};
}

// The glossy coat: a GGX dielectric layer of F0 = 0.04 over a matte base of color rho, with Ashikhmin and Shirley's
// coupled base (28 / (23 pi)) rho (1 - F0) (1 - (1 - cos_l / 2)^5) (1 - (1 - cos_v / 2)^5); sampled half the time by
// the coat's visible normals and half by the cosine, so its density is the two's mean
constexpr float coat_f0 = 0.04f;
constexpr float coat_base_scale = static_cast<float>(28.0 / (23.0 * 3.14159265358979323846) * (1.0 - 0.04));
template <typename Real>
__device__ inline Real coat_transmission(Real c)
{
    const Real x = Real(1.0) - Real(0.5) * c;
    const Real x2 = x * x;
    return Real(1.0) - x2 * x2 * x;
}
template <typename Real>
__device__ inline ggx_eval_result<Real> glossy_coat(vec3<Real> v, vec3<Real> l, Real alpha, vec3<Real> rho)
{
    if (!(v.z > Real(0.0)) || !(l.z > Real(0.0)))
    {
        return {{Real(0.0), Real(0.0), Real(0.0)}, Real(0.0)};
    }
    ggx_eval_result<Real> coat = ggx_reflection(v, l, alpha, {coat_f0, coat_f0, coat_f0});
    const Real base = coat_base_scale * (coat_transmission(l.z) * coat_transmission(v.z)) * l.z;
    coat.value = coat.value + base * rho;
    coat.pdf = Real(0.5) * (coat.pdf + l.z * one_over_pi<Real>);
    return coat;
}

// --- the vertex's scattering ------------------------------------------------------------------------------------------
enum class scatter_kind : int
{
    lambert,   // Lambert's cosine lobe (the books' matte)
    phase,     // the isotropic phase function (smoke)
    conductor, // GGX conductor, alpha >= min_ggx_alpha
    coat,      // glossy coat over a matte base
    mirror,    // a conductor too smooth for GGX: Schlick's Fresnel times a perfect mirror
    books,     // the books' metal and glass (delta lobes)
};
template <typename Real>
struct scatterer
{
    scatter_kind kind;
    vec3<Real> color;   // albedo, F0 (conductors) or the base color (coat)
    frame<Real> local;  // the tangent frame around the normal facing the viewer (GGX lobes)
    vec3<Real> v;       // toward the viewer, in `local` (GGX lobes)
    Real alpha;         // GGX: r^2
    Real roughness;     // GGX: r, the energy table's column
};
__device__ inline bool is_delta(scatter_kind k) { return k == scatter_kind::mirror || k == scatter_kind::books; }

// The vertex's scattering from its material and GGX side entry (src/gpu/bsdf.cuh:58-110): a conductor's F0 is the
// material's albedo, a coat's base is the material's color
template <typename Real>
__device__ inline scatterer<Real> make_scatterer(const material_rec<Real>& m, const microfacet_rec* side, vec3<Real> color,
                                           const vertex<Real>& vx, vec3<Real> ray_direction)
{
    scatterer<Real> s{};
    s.color = color;
    const int side_kind = side ? side->kind : RT_MICROFACET_NONE;
    if (side_kind == RT_MICROFACET_CONDUCTOR || side_kind == RT_MICROFACET_GLOSS)
    {
        s.alpha = side->alpha;
        s.roughness = side->roughness;
        s.local = frame_around(vx.n);
        s.v = s.local.to_local(-unit_vector(ray_direction));
        if (side_kind == RT_MICROFACET_GLOSS)
        {
            s.kind = scatter_kind::coat;
        }
        else
        {
            s.kind = s.alpha >= min_ggx_alpha ? scatter_kind::conductor : scatter_kind::mirror;
            s.color = load3(m.albedo);
        }
        return s;
    }
    s.kind = m.kind == RT_MATERIAL_LAMBERTIAN   ? scatter_kind::lambert
             : m.kind == RT_MATERIAL_ISOTROPIC ? scatter_kind::phase
                                               : scatter_kind::books;
    return s;
}

// f cos at world direction wi, and the density the scattering would draw it with (0 for delta lobes)
template <typename Real>
__device__ inline ggx_eval_result<Real> scatter_value(const scatterer<Real>& s, const vertex<Real>& vx, vec3<Real> wi)
{
    switch (s.kind)
    {
    case scatter_kind::lambert:
    {
        const Real pdf = cosine_density(vx.n, wi);
        return {pdf * s.color, pdf};
    }
    case scatter_kind::phase:
        return {one_over_four_pi<Real> * s.color, one_over_four_pi<Real>};
    case scatter_kind::conductor:
    {
        ggx_eval_result<Real> r = ggx_reflection(s.v, s.local.to_local(wi), s.alpha, s.color);
        r.value = multiple_scatter_factor(s.color, s.v.z, s.roughness) * r.value;
        return r;
    }
    case scatter_kind::coat:
        return glossy_coat(s.v, s.local.to_local(wi), s.alpha, s.color);
    default:
        return {{Real(0.0), Real(0.0), Real(0.0)}, Real(0.0)};
    }
}

template <typename Real>
struct scatter_draw
{
    vec3<Real> wi;
    vec3<Real> weight; // f cos / pdf, or the delta lobe's factor
    Real pdf;          // 0 for a delta lobe
};
// A direction drawn from the scattering; false when the path is absorbed
template <typename Real>
__device__ inline bool scatter_sample(const scatterer<Real>& s, const material_rec<Real>& m, vec3<Real> ray_direction,
                                      const vertex<Real>& vx, pcg32& g, scatter_draw<Real>& out)
{
    out.pdf = Real(0.0);
    switch (s.kind)
    {
    case scatter_kind::lambert:
        out.wi = cosine_about(vx.n, g);
        out.weight = s.color;
        out.pdf = cosine_density(vx.n, out.wi);
        return dot(vx.n, out.wi) > Real(0.0);
    case scatter_kind::phase:
        out.wi = random_unit_vector<Real>(g);
        out.weight = s.color;
        out.pdf = one_over_four_pi<Real>;
        return true;
    case scatter_kind::books:
        if (m.kind == RT_MATERIAL_METAL)
        {
            out.weight = load3(m.albedo);
            return fuzzed_mirror(ray_direction, vx.n, m.fuzz, g, out.wi);
        }
        out.weight = {Real(1.0), Real(1.0), Real(1.0)};
        out.wi = dielectric_direction(ray_direction, vx.n, vx.front, m.refraction_index, g);
        return true;
    case scatter_kind::mirror:
    {
        if (!(s.v.z > Real(0.0)))
        {
            return false;
        }
        out.wi = s.local.to_world({-s.v.x, -s.v.y, s.v.z});
        out.weight = fresnel_schlick(s.color, s.v.z);
        return true;
    }
    case scatter_kind::conductor:
    {
        vec3<Real> l;
        if (!ggx_draw(s.v, s.alpha, s.color, g, l, out.weight, out.pdf))
        {
            return false;
        }
        out.wi = s.local.to_world(l);
        out.weight = multiple_scatter_factor(s.color, s.v.z, s.roughness) * out.weight;
        return true;
    }
    case scatter_kind::coat:
    {
        vec3<Real> l;
        if (random_real<Real>(g) < Real(0.5))
        {
            vec3<Real> coat_weight;
            Real coat_pdf;
            if (!ggx_draw(s.v, s.alpha, {coat_f0, coat_f0, coat_f0}, g, l, coat_weight, coat_pdf))
            {
                return false;
            }
        }
        else
        {
            l = s.local.to_local(cosine_about(vx.n, g));
            if (!(l.z > Real(0.0)))
            {
                return false;
            }
        }
        const ggx_eval_result<Real> f = glossy_coat(s.v, l, s.alpha, s.color);
        if (!(f.pdf > Real(0.0)))
        {
            return false;
        }
        out.wi = s.local.to_world(l);
        out.weight = (Real(1.0) / f.pdf) * f.value;
        out.pdf = f.pdf;
        return true;
    }
    }
    return false;
}

// --- the light set: lamps and the environment map ------------------------------------------------------------------
// A lamp (a glowing quad or sphere of the world): points uniform on a quad's area, density s^2 / (A cos theta_l) per
// solid angle; directions uniform in the cone of a sphere seen from outside, 1 / (2 pi (1 - cos theta_max)). Exactly
// two numbers per draw (src/gpu/lamps.cuh:55-146).
template <typename Real>
__device__ inline Real quad_lamp_density(Real d2, Real cos_l, Real area) { return d2 / (cos_l * area); }
template <typename Real>
__device__ inline Real sphere_lamp_density(Real radius, Real d2)
{
    return Real(0.5) * one_over_pi<Real> / cone_drop(radius, d2);
}
template <typename Real>
struct light_draw
{
    vec3<Real> wi;
    Real distance; // to the light's point (infinity for the map)
    vec3<Real> radiance;
    Real pdf;      // per solid angle, the light's choice included
    int skip;       // the lamp's own primitive, which its shadow ray doesn't count (-1: none)
};
template <typename Real>
__device__ inline bool draw_lamp(const emitter_rec& e, vec3<Real> x, pcg32& g, light_draw<Real>& out)
{
    const Real a = random_real<Real>(g);
    const Real b = random_real<Real>(g);
    const vec3<Real> anchor = table3<Real>(e.corner);
    if (e.kind == RT_LIGHT_QUAD)
    {
        const vec3<Real> to = (anchor - x) + a * table3<Real>(e.u) + b * table3<Real>(e.v);
        const Real d2 = length_squared(to);
        const Real inv = m_rsqrt(d2);
        out.wi = inv * to;
        out.distance = d2 * inv;
        const Real cos_l = -dot(out.wi, table3<Real>(e.normal));
        out.pdf = quad_lamp_density(d2, cos_l, static_cast<Real>(e.size));
        return cos_l > Real(0.0);
    }
    const vec3<Real> to_centre = anchor - x;
    const Real d2 = squared_distance(to_centre);
    const Real r2 = static_cast<Real>(e.size) * static_cast<Real>(e.size);
    if (!(d2 > r2))
    {
        return false;
    }
    const Real inv = m_rsqrt(d2);
    // 1 - cos theta uniform in [0, 1 - cos theta_max]: sin^2 and cos from it without cancellation
    const Real drop = b * cone_drop(static_cast<Real>(e.size), d2);
    const Real sin2 = drop * (Real(2.0) - drop);
    const Real cos_t = Real(1.0) - drop;
    const Real sin_t = m_sqrt(sin2);
    Real s, c;
    m_sincospi(Real(2.0) * a, &s, &c);
    out.wi = frame_around(inv * to_centre).to_world({sin_t * c, sin_t * s, cos_t});
    // the near crossing of the sphere along wi: |c - x| cos - sqrt(r^2 - |c - x|^2 sin^2)
    out.distance = d2 * inv * cos_t - m_sqrt(m_max(Real(0.0), m_fma(-d2, sin2, r2)));
    out.pdf = sphere_lamp_density(static_cast<Real>(e.size), d2);
    return true;
}
// The density with which the lamp would have been drawn toward unit_dir from x, its point `distance` away
template <typename Real>
__device__ inline Real lamp_density(const emitter_rec& e, vec3<Real> x, vec3<Real> unit_dir, Real distance)
{
    if (e.kind == RT_LIGHT_QUAD)
    {
        const Real cos_l = -dot(unit_dir, table3<Real>(e.normal));
        return cos_l > Real(0.0) ? quad_lamp_density(distance * distance, cos_l, static_cast<Real>(e.size)) : Real(0.0);
    }
    const Real d2 = squared_distance(table3<Real>(e.corner) - x);
    const Real radius = static_cast<Real>(e.size);
    return d2 > radius * radius ? sphere_lamp_density(radius, d2) : Real(0.0);
}

// The environment map: equirectangular, +y up, turned by the yaw; u = 1/2 + atan2(z, x) / (2 pi) across, t = polar
// angle / pi down. Its cells are exact in solid angle (rows bounded by cosines), and each texel holds its cell's
// radiance and density per steradian. A direction's radiance and density are read from the same texel, whichever way
// the direction was found (src/gpu/environment.cuh:59-135).
template <typename Real>
struct map_point
{
    Real u, t;
};
template <typename Real>
__device__ inline vec3<Real> world_to_map(const quality_view& q, vec3<Real> d)
{
    return {q.env_cos_yaw * d.x + q.env_sin_yaw * d.z, d.y, q.env_cos_yaw * d.z - q.env_sin_yaw * d.x};
}
template <typename Real>
__device__ inline vec3<Real> map_to_world(const quality_view& q, vec3<Real> d)
{
    return {q.env_cos_yaw * d.x - q.env_sin_yaw * d.z, d.y, q.env_sin_yaw * d.x + q.env_cos_yaw * d.z};
}
template <typename Real>
__device__ inline map_point<Real> map_of(const quality_view& q, vec3<Real> world_dir)
{
    const vec3<Real> m = world_to_map(q, world_dir);
    return {Real(0.5) + m_atan2(m.z, m.x) * (Real(0.5) * one_over_pi<Real>),
            m_atan2(m_sqrt(m.x * m.x + m.z * m.z), m.y) * one_over_pi<Real>};
}
template <typename Real>
__device__ inline int texel_index(const quality_view& q, vec3<Real> world_dir)
{
    const map_point<Real> mp = map_of(q, world_dir);
    const int i = min(max(static_cast<int>(mp.u * static_cast<Real>(q.env_width)), 0), q.env_width - 1);
    const int j = min(max(static_cast<int>(mp.t * static_cast<Real>(q.env_height)), 0), q.env_height - 1);
    return j * q.env_width + i;
}
template <typename Real>
__device__ inline vec3<Real> texel_radiance(const quality_view& q, int k, Real& density)
{
    const env_texel x = q.env_texels[k];
    density = x.pdf;
    return static_cast<Real>(q.env_scale) * vec3<Real>{x.r, x.g, x.b};
}
// The map between its four nearest texel centres (columns wrap around, rows stop at the poles): what the camera and
// rays after a mirror see, so that the map looks smooth where it is magnified
template <typename Real>
__device__ inline vec3<Real> map_bilinear(const quality_view& q, vec3<Real> dir)
{
    const map_point<Real> mp = map_of(q, dir);
    const Real x = mp.u * static_cast<Real>(q.env_width) - Real(0.5);
    const Real y = mp.t * static_cast<Real>(q.env_height) - Real(0.5);
    const Real fx0 = m_floor(x), fy0 = m_floor(y);
    const Real sx = x - fx0, sy = y - fy0;
    const int w = q.env_width;
    const int c0 = static_cast<int>(fx0) < 0 ? static_cast<int>(fx0) + w : static_cast<int>(fx0);
    const int c1 = c0 + 1 < w ? c0 + 1 : 0;
    const int r0 = min(max(static_cast<int>(fy0), 0), q.env_height - 1) * w;
    const int r1 = min(max(static_cast<int>(fy0) + 1, 0), q.env_height - 1) * w;
    const auto rgb = [&](int k) {
        const env_texel e = q.env_texels[k];
        return vec3<Real>{e.r, e.g, e.b};
    };
    const vec3<Real> upper = lerp3(rgb(r0 + c0), rgb(r0 + c1), sx);
    const vec3<Real> lower = lerp3(rgb(r1 + c0), rgb(r1 + c1), sx);
    return static_cast<Real>(q.env_scale) * lerp3(upper, lower, sy);
}
__device__ inline bool map_sampled(const quality_view& q) { return q.env_present && q.env_bins != nullptr; }
// What a ray that leaves the scene sees: the books' sky without a map; with one, the bilinear map (a camera ray or a
// ray after a mirror) or its texel and the texel's density (any other ray, whose weight may need it)
template <typename Real>
__device__ inline vec3<Real> escaped(const quality_view& q, const camera_rec<Real>& cam, vec3<Real> dir,
                                      bool smooth, Real& density)
{
    density = Real(0.0);
    if (!q.env_present)
    {
        return sky(cam, dir);
    }
    if (smooth)
    {
        return map_bilinear(q, dir);
    }
    Real texel_density;
    const vec3<Real> le = texel_radiance(q, texel_index(q, dir), texel_density);
    density = map_sampled(q) ? texel_density : Real(0.0);
    return le;
}
// A direction of the map drawn in proportion to radiance x solid angle: the alias table's bin from a word's top bits,
// its coin, then a point uniform in the chosen cell's solid angle (cos theta and phi uniform within the cell). Four
// numbers.
template <typename Real>
__device__ inline bool draw_map(const quality_view& q, pcg32& g, light_draw<Real>& out)
{
    const int bin = static_cast<int>(g.next() >> (32 - q.env_log2_bins));
    const alias_bin entry = q.env_bins[bin];
    const int cell = random_real<Real>(g) < entry.q ? bin : entry.alias;
    const Real in_theta = random_real<Real>(g);
    const Real in_phi = random_real<Real>(g);
    const int row = cell / q.env_width;
    const int column = cell - row * q.env_width;
    const Real cos_top = q.env_row_cos[row];
    const Real cos_theta = cos_top - in_theta * (cos_top - q.env_row_cos[row + 1]);
    const Real sin_theta = m_sqrt(m_max(Real(0.0), (Real(1.0) - cos_theta) * (Real(1.0) + cos_theta)));
    Real s, c;
    m_sincospi(Real(2.0) * (static_cast<Real>(column) + in_phi) / static_cast<Real>(q.env_width) - Real(1.0), &s, &c);
    out.wi = map_to_world(q, vec3<Real>{sin_theta * c, cos_theta, sin_theta * s});
    Real density;
    out.radiance = texel_radiance(q, texel_index(q, out.wi), density);
    out.pdf = density;
    out.distance = inf<Real>;
    out.skip = -1;
    return density > Real(0.0);
}

// The light set: the lamps then the map (when it is sampled), one chosen uniformly by one number (src/gpu/
// light_sampler.cuh:40-102)
__device__ inline int light_set_size(const quality_view& q) { return q.lamp_count + (map_sampled(q) ? 1 : 0); }
template <typename Tally, typename Real>
__device__ inline bool draw_light(const quality_view& q, vec3<Real> x, pcg32& g, light_draw<Real>& out, Tally& tl)
{
    const int n = light_set_size(q);
    const int k = min(static_cast<int>(random_real<Real>(g) * static_cast<Real>(n)), n - 1);
    if (k < q.lamp_count)
    {
        const emitter_rec e = q.lamps[k];
        tl.add(tally_lamp);
        if (!draw_lamp(e, x, g, out))
        {
            return false;
        }
        out.radiance = table3<Real>(e.radiance);
        out.skip = e.primitive;
    }
    else
    {
        tl.add(tally_alias_bin);
        tl.add(tally_env_texel);
        if (!draw_map(q, g, out))
        {
            return false;
        }
    }
    out.pdf *= q.light_pmf;
    return true;
}
// The density with which the light set would have drawn the lamp primitive `id` in unit_dir from x
template <typename Real>
__device__ inline Real lamp_hit_density(const scene_view<Real>& sc, int id, vec3<Real> x, vec3<Real> unit_dir,
                                         Real distance)
{
    const quality_view& q = sc.quality;
    const int k = q.lamp_count > 0 && id < sc.instanced_base ? q.lamp_of[id] : -1; // (an instanced mesh lights nothing)
    return k >= 0 ? q.light_pmf * lamp_density(q.lamps[k], x, unit_dir, distance) : Real(0.0);
}

// --- shadow rays -------------------------------------------------------------------------------------------------------
// Is any primitive but `skip` hit in (0.001, t_max)? The four-wide walk of closest_hit, ending at the first hit, with
// the leaving primitive's self-intersection rule: a sphere it leaves inward can still be hit from inside (its far
// root), any other primitive it leaves cannot. A sphere's interval is open at t_max, a quad's and a triangle's closed
// (src/gpu/shadow.cuh:26-120).
template <typename Tally, typename Real>
__device__ inline bool leaf_blocks(const scene_view<Real>& sc, const ray<Real>& r, int32_t ref, Real inverse_a,
                                   Real t_max, int from, bool inward, int skip, int triangle_shift, Tally& tl)
{
    constexpr Real t_min = Real(0.001);
    const int kind = leaf_kind(ref);
    const int first = leaf_first(ref);
    const int end = first + leaf_count(ref);
    Real t;
    bool side;
    for (int i = first; i < end; ++i)
    {
        if (kind == prim_quad)
        {
            const int id = sc.quad_base + i;
            if (id == from || id == skip)
            {
                continue;
            }
            tl.add(tally_quad);
            if (hit_quad(sc.quads[i], r, t_min, t_max, t, side))
            {
                return true;
            }
        }
        else if (kind == prim_triangle)
        {
            const int id = triangle_shift + i;
            if (id == from || id == skip)
            {
                continue;
            }
            tl.add(tally_triangle);
            if (hit_triangle(sc.triangles[i], r, t_min, t_max, t, side))
            {
                return true;
            }
        }
        else
        {
            const bool moving = kind == prim_moving_sphere;
            const int id = moving ? sc.moving_base + i : i;
            const bool own = id == from;
            if ((own && !inward) || id == skip)
            {
                continue;
            }
            tl.add(moving ? tally_moving : tally_sphere);
            const sphere<Real> s = moving ? moving_at(sc.moving_spheres[i], r.time) : load_sphere(sc.spheres, i);
            if (hit_sphere(s, r, inverse_a, t_min, t_max, own, t, side))
            {
                return true;
            }
        }
    }
    return false;
}
template <int Walk, typename Tally, typename Stack, typename Real>
__device__ inline bool occluded(const scene_view<Real>& sc, const ray<Real>& r, Real t_max, int from, bool inward,
                                int skip, const Stack& stack, Tally& tl)
{
    if (sc.stack_depth == 0)
    {
        return false;
    }
    walk_ray<Real> w;
    w.set(r);
    int depth = 0;
    int floor = 0; // in an instance's mesh: the scene tree's part of the stack
    int triangle_shift = sc.triangle_base;
    int32_t ref = sc.root_ref;
    while (true)
    {
        if (ref < 0)
        {
            if constexpr (Walk == walk_instances)
            {
                if (leaf_kind(ref) == prim_instance)
                {
                    // into the instance's mesh, as closest_hit goes
                    const instance_rec<Real>& in = sc.instances[leaf_first(ref)];
                    tl.add(tally_instance);
                    w.set(into_instance(in.to_object, r));
                    floor = depth;
                    triangle_shift = in.id_base - in.first_triangle;
                    ref = in.root_ref;
                    continue;
                }
            }
            if (leaf_blocks(sc, w.r, ref, w.inverse_a, t_max, from, inward, skip, triangle_shift, tl))
            {
                return true;
            }
        }
        else
        {
            // every entered child must be looked at, in any order: the first goes next, the rest on the stack
            tl.add(tally_node);
            Real t[wide_children];
            int4 refs;
            node_children<Walk>(sc, ref, w.so, w.inv, w.os, Real(0.001), t_max, t, refs);
            const int32_t child[wide_children] = {refs.x, refs.y, refs.z, refs.w};
            int32_t next = no_child;
#pragma unroll
            for (int c = 0; c < wide_children; ++c)
            {
                if (t[c] < inf<Real>)
                {
                    if (next != no_child)
                    {
                        stack.store(depth++, next, Real(0.001));
                    }
                    next = child[c];
                }
            }
            if (next != no_child)
            {
                ref = next;
                continue;
            }
        }
        if constexpr (Walk == walk_instances)
        {
            if (depth == floor && floor > 0)
            {
                // the mesh is done: back to the world's coordinates and the scene tree's entries
                w.set(r);
                floor = 0;
                triangle_shift = sc.triangle_base;
            }
        }
        if (depth == 0)
        {
            return false;
        }
        ref = stack.ref_at(--depth);
    }
}
// The chance that the segment (0.001, t_max) crosses the media without a collision: e^(-sum rho_k L_k) over the
// lengths L_k inside each medium (the expected value of their free-flight tests; no number drawn)
template <typename Tally, typename Real>
__device__ inline Real transmittance(const scene_view<Real>& sc, const ray<Real>& r, Real t_max, Tally& tl)
{
    if (sc.medium_count == 0)
    {
        return Real(1.0);
    }
    const Real inverse_a = Real(1.0) / length_squared(r.direction);
    const vec3<Real> inv{safe_inverse(r.direction.x), safe_inverse(r.direction.y), safe_inverse(r.direction.z)};
    const vec3<Real> os{r.origin.x * inv.x, r.origin.y * inv.y, r.origin.z * inv.z};
    const Real speed = m_sqrt(length_squared(r.direction));
    Real optical_depth = Real(0.0);
    for (int k = 0; k < sc.medium_count; ++k)
    {
        tl.add(tally_medium);
        const medium_rec<Real>& m = sc.media[k];
        Real a, b;
        if (smoke_span(m, sc.boundary_quads, r, inverse_a, inv, os, Real(0.001), t_max, a, b, tl))
        {
            optical_depth -= (b - a) * speed / m.neg_inv_density;
        }
    }
    return m_exp(-optical_depth);
}

// --- one step of a quality path ----------------------------------------------------------------------------------------
// A light sample's shadow ray waiting to be traced as the path's next ray (the queue megakernel's paths): from the
// vertex (the path's ray's origin, leaving the same primitive) toward `direction`, up to `distance`, the sampled lamp
// passed over; `term` is added to the radiance when nothing blocks it
template <typename Real>
struct pending_shadow
{
    vec3<Real> direction;
    vec3<Real> term;
    Real distance;
    int skip;
    bool inward;
    bool waiting;
};

template <typename Real>
struct quality_path
{
    path<Real> base;
    vec3<Real> radiance; // the light gathered so far
    Real bsdf_pdf;       // the last scattering's density, if that vertex took a light sample (else 0)
    bool specular;        // the last vertex was a delta lobe's (or the camera): the map is seen bilinear
    bool ended;           // the path ended at the vertex whose shadow ray waits
    pending_shadow<Real> shadow;
};

template <typename Real>
__device__ inline quality_path<Real> start_quality(const camera_rec<Real>& cam, bool motion, int i, int j, int pass,
                                             pcg32& g)
{
    quality_path<Real> q{};
    q.base = start_path(cam, motion, i, j, pass, g);
    q.specular = true;
    return q;
}

// The light sample's contribution at a vertex before its shadow ray: throughput f Le w / pdf, or false when it adds
// nothing (a density of 0, no reflection toward the light, or a sum that isn't finite and positive)
template <typename Tally, typename Real>
__device__ inline bool light_term(const scene_view<Real>& sc, const scatterer<Real>& s, const vertex<Real>& vx,
                                  vec3<Real> throughput, pcg32& g, light_draw<Real>& ld, vec3<Real>& term, Tally& tl)
{
    if (!draw_light(sc.quality, vx.p, g, ld, tl))
    {
        return false;
    }
    const ggx_eval_result<Real> f = scatter_value(s, vx, ld.wi);
    if (!(ld.pdf > Real(0.0)) || !(max3(f.value) > Real(0.0)))
    {
        return false;
    }
    const Real w = combination{sc.quality.combination}.light_sample(ld.pdf, f.pdf) / ld.pdf;
    term = w * (throughput * (f.value * ld.radiance));
    const Real sum = term.x + term.y + term.z;
    return sum > Real(0.0) && sum < inf<Real>;
}

// What an event (or a miss) does to a quality path (src/gpu/quality_path<Real>.cuh:61-188): an escaped ray or an emitter
// adds its light, weighed against the light set's chance to have drawn it; any other vertex takes a light sample
// (when the path may still trace a ray, the lobe isn't a delta and a light exists), traces its shadow ray, then
// scatters
//
// Deferred: the shadow ray isn't traced here but left waiting in the path (q.shadow) for the caller to trace as its next
// ray (step_quality); a path that ends at this vertex then ends once that ray is traced (q.ended). The random numbers
// are drawn in the same order and the light added in the same order either way: the shadow ray draws none, and its
// light is added before anything the next vertex adds.
template <int Walk, bool Deferred = false, typename Tally, typename Stack, typename Real>
__device__ inline bool shade_quality(quality_path<Real>& q, const scene_view<Real>& sc, const camera_rec<Real>& cam,
                                     const hit<Real>& h, pcg32& g, vec3<Real>& sample, const Stack& stack,
                                     Tally& tl, uint32_t& shadow_rays)
{
    path<Real>& p = q.base;
    const quality_view& qv = sc.quality;
    const combination mix{qv.combination};
    if (h.index < 0)
    {
        if (qv.env_present)
        {
            tl.add(tally_env_texel, q.specular ? 4u : 1u);
        }
        Real map_density;
        const vec3<Real> le = escaped(qv, cam, p.r.direction, q.specular, map_density);
        const Real w = q.bsdf_pdf > Real(0.0) ? mix.scattered_hit(q.bsdf_pdf, qv.light_pmf * map_density) : Real(1.0);
        sample = q.radiance + w * (p.throughput * le);
        return true;
    }
    const vertex<Real> vx = make_vertex<Walk == walk_instances>(sc, p.r, h, tl);
    const material_rec<Real> m = sc.materials[vx.material];
    // the color is read once: matte, smoke and a lamp's front face have one (a coat's base is the matte color)
    const bool lamp = m.kind == RT_MATERIAL_DIFFUSE_LIGHT;
    const bool colored = m.kind == RT_MATERIAL_LAMBERTIAN || m.kind == RT_MATERIAL_ISOTROPIC || (lamp && vx.front);
    const vec3<Real> color = colored ? material_color(sc, m, vx, tl) : vec3<Real>{Real(0.0), Real(0.0), Real(0.0)};
    if (lamp)
    {
        const vec3<Real> le = color;
        Real w = Real(1.0);
        if (q.bsdf_pdf > Real(0.0))
        {
            tl.add(tally_lamp);
            const Real distance = h.t * m_sqrt(length_squared(p.r.direction));
            w = mix.scattered_hit(q.bsdf_pdf,
                                  lamp_hit_density(sc, h.index, p.r.origin, unit_vector(p.r.direction), distance));
        }
        sample = q.radiance + w * (p.throughput * le);
        return true;
    }
    const scatterer<Real> s = make_scatterer(m, qv.microfacets ? &qv.microfacets[vx.material] : nullptr, color, vx,
                                       p.r.direction);
    const bool nee = qv.combination != RT_COMBINE_BSDF_ONLY && p.rays_left > 1 && !is_delta(s.kind) &&
                     light_set_size(qv) > 0 && max3(p.throughput) > Real(0.0);
    light_draw<Real> ld;
    vec3<Real> term;
    const bool lit = nee && light_term(sc, s, vx, p.throughput, g, ld, term, tl);
    scatter_draw<Real> next;
    const bool scattered = scatter_sample(s, m, p.r.direction, vx, g, next);
    if (lit)
    {
        ++shadow_rays;
        const bool inward = dot(ld.wi, vx.outward) < Real(0.0);
        if constexpr (Deferred)
        {
            q.shadow = {ld.wi, term, ld.distance, ld.skip, inward, true};
        }
        else
        {
            const ray<Real> shadow{vx.p, ld.wi, p.r.time};
            if (!occluded<Walk>(sc, shadow, ld.distance, vx.from, inward, ld.skip, stack, tl))
            {
                q.radiance = q.radiance + transmittance(sc, shadow, ld.distance, tl) * term;
            }
        }
    }
    // (a waiting shadow ray leaves the vertex as the next ray would: its origin and primitive go into the path's ray)
    bool ends = !scattered;
    if (scattered)
    {
        p.throughput = p.throughput * next.weight;
        q.bsdf_pdf = nee ? next.pdf : Real(0.0);
        q.specular = is_delta(s.kind);
        ends = (qv.combination != RT_COMBINE_BSDF_ONLY && !(max3(p.throughput) > Real(0.0))) || --p.rays_left == 0;
    }
    if (ends && !(Deferred && lit))
    {
        sample = q.radiance;
        return true;
    }
    p.r = {vx.p, scattered ? next.wi : vx.n, p.r.time}; // (the direction of a path that ended isn't read)
    p.from = vx.from;
    p.inward = scattered && dot(next.wi, vx.outward) < Real(0.0);
    q.ended = ends;
    return false;
}

// One ray of a quality path: its waiting shadow ray, if it has one, else its next ray, both walked by the same code
// (closest_hit's shadow mode), so that the lanes of a wave that trace shadow rays and those that trace their paths'
// next rays walk the tree together. The light the shadow ray carries is added when nothing blocks it; a path whose
// last vertex left a shadow ray ends after it. The same numbers as trace_quality's loop, in the same order (shade_quality,
// Deferred). True when the path's sample is done.
template <int Walk, typename Tally, typename Stack, typename Real>
__device__ inline bool step_quality(quality_path<Real>& q, const scene_view<Real>& sc, const camera_rec<Real>& cam,
                                    pcg32& g, vec3<Real>& sample, const Stack& stack, Tally& tl,
                                    unsigned long long& rays, uint32_t& shadow_rays)
{
    pending_shadow<Real>& sh = q.shadow;
    const bool shadow = sh.waiting;
    // (each choice made on values, component by component: a choice between two members' places would keep the path
    // in scratch memory rather than in registers)
    const vec3<Real> d = q.base.r.direction;
    const vec3<Real> ds = sh.direction;
    const ray<Real> r{q.base.r.origin, {shadow ? ds.x : d.x, shadow ? ds.y : d.y, shadow ? ds.z : d.z},
                       q.base.r.time};
    const bool inward = shadow ? static_cast<bool>(sh.inward) : static_cast<bool>(q.base.inward);
    const hit<Real> h = closest_hit<Real, Walk, Tally, Stack, true>(
        sc, r, q.base.from, inward, stack, tl, shadow ? static_cast<Real>(sh.distance) : inf<Real>,
        shadow ? static_cast<int>(sh.skip) : -1, shadow);
    tl.lap(phase_walk);
    if (shadow)
    {
        sh.waiting = false;
        if (h.index < 0)
        {
            q.radiance = q.radiance + transmittance(sc, r, sh.distance, tl) * sh.term;
        }
        tl.lap(phase_media);
        if (q.ended)
        {
            sample = q.radiance;
            return true;
        }
        return false;
    }
    ++rays;
    const hit<Real> event = media_events(sc, q.base.r, h, g, tl);
    tl.lap(phase_media);
    const bool done = shade_quality<Walk, true>(q, sc, cam, event, g, sample, stack, tl, shadow_rays);
    tl.lap(phase_shade);
    return done;
}

// Walk: the tree's walk (walk_all; walk_instances: a two-level tree, whose walk can enter instances; walk_packed: a
// packed tree), a kernel of its own each
template <int Walk, typename Tally, typename Stack, typename Real>
__device__ inline bool trace_quality(quality_path<Real>& q, const scene_view<Real>& sc, const camera_rec<Real>& cam,
                                     pcg32& g, vec3<Real>& sample, const Stack& stack, Tally& tl,
                                     uint32_t& shadow_rays)
{
    const hit<Real> h = closest_event<Real, Walk>(sc, q.base.r, q.base.from,
                                                                                   q.base.inward, g, stack, tl);
    return shade_quality<Walk>(q, sc, cam, h, g, sample, stack, tl, shadow_rays);
}
}
