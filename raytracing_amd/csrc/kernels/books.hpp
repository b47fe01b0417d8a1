// What a path does at the event its BVH walk found, for the three books' scenes (Shirley, "Ray Tracing in One
// Weekend", "The Next Week", "The Rest of Your Life"): the camera's rays, the vertex a hit makes (point, normals, side,
// texture coordinates), the materials' colors (solid, checker, image, Perlin marble), constant-density smoke, the
// books' lobes (Lambert, isotropic phase, fuzzed mirror, dielectric), the lights list of the third book and one step
// of a path. Written for this renderer from the books' mathematics; the reference GPU renderer's semantics that an
// image depends on are kept and cited per function (which side of a quad a hit point lies on, the lamps' one face,
// the depth rule, the self-intersection data a new ray carries).
//
// Arithmetic as in path_tracer.hpp: -ffp-contract=off, multiply-adds only where written. Nothing here must equal the
// reference bit for bit (the shading is compared by ImageCompare), so the samplers are this file's own: cosine lobes
// as the normal plus a uniform point of the unit sphere, the lens by Shirley and Chiu's concentric map, refraction by
// Snell's law in vector form, Perlin noise by nested interpolation.
#pragma once

#include "path_tracer.hpp"

namespace rt::dev
{
template <typename Real>
constexpr Real one_over_pi = static_cast<Real>(0.318309886183790671537767526745028724);
template <typename Real>
constexpr Real one_over_four_pi = static_cast<Real>(0.0795774715459476678844418816862571810);

template <typename Real>
__device__ inline vec3<Real> lerp3(vec3<Real> a, vec3<Real> b, Real s)
{
    return a + s * (b - a);
}

// --- frames and directions ---------------------------------------------------------------------------------------------
// An orthonormal frame (s, t, w) around a unit vector w, by Duff et al.'s branch-free construction (JCGT 6(1), 2017):
// to_world takes local coordinates with z along w
template <typename Real>
struct frame
{
    vec3<Real> s, t, w;
    __device__ vec3<Real> to_world(vec3<Real> v) const { return v.x * s + v.y * t + v.z * w; }
    __device__ vec3<Real> to_local(vec3<Real> v) const { return {dot(v, s), dot(v, t), dot(v, w)}; }
};
template <typename Real>
__device__ inline frame<Real> frame_around(vec3<Real> w)
{
    const Real sign = m_copysign(Real(1), w.z);
    const Real k = Real(-1) / (sign + w.z);
    const Real xy = w.x * w.y * k;
    return {{Real(1) + sign * k * w.x * w.x, sign * xy, -sign * w.x}, {xy, sign + k * w.y * w.y, -w.y}, w};
}

// A unit direction about the unit normal n with density cos(theta)/pi: the direction from the origin to a uniform
// point of the unit sphere centred at n (Lambert's law in the first book's form; the sum is zero only when the point
// is -n, which a draw of the 2^-24 grid never is, and is taken as n then)
template <typename Real>
__device__ inline vec3<Real> cosine_about(vec3<Real> n, pcg32& g)
{
    const vec3<Real> d = n + random_unit_vector<Real>(g);
    const Real len2 = length_squared(d);
    return len2 > Real(1e-12) ? m_rsqrt(len2) * d : n;
}
template <typename Real>
__device__ inline Real cosine_density(vec3<Real> n, vec3<Real> d)
{
    const Real c = dot(n, d) * m_rsqrt(length_squared(d));
    return c > Real(0) ? c * one_over_pi<Real> : Real(0);
}

// A uniform point of the unit disk from two uniform numbers by Shirley and Chiu's concentric map (1997): the square
// [-1, 1]^2 in concentric squares onto concentric circles, area for area
template <typename Real>
__device__ inline void concentric_disk(Real a, Real b, Real& x, Real& y)
{
    const Real sx = Real(2) * a - Real(1);
    const Real sy = Real(2) * b - Real(1);
    if (sx == Real(0) && sy == Real(0))
    {
        x = Real(0);
        y = Real(0);
        return;
    }
    const bool wide = m_fabs(sx) > m_fabs(sy);
    const Real radius = wide ? sx : sy;
    // the angle in units of pi: (1/4)(sy/sx) on the wide sectors, 1/2 - (1/4)(sx/sy) on the others
    const Real angle = wide ? Real(0.25) * (sy / sx) : Real(0.5) - Real(0.25) * (sx / sy);
    Real s, c;
    m_sincospi(angle, &s, &c);
    x = radius * c;
    y = radius * s;
}

// Directions uniform in the cone about a unit axis whose half-angle has 1 - cos = one_minus_cos_max: 1 - cos(theta)
// uniform in [0, one_minus_cos_max], the azimuth uniform
template <typename Real>
__device__ inline vec3<Real> cone_direction(vec3<Real> axis, Real one_minus_cos_max, Real a, Real b)
{
    const Real drop = b * one_minus_cos_max;
    const Real sin_theta = m_sqrt(drop * (Real(2) - drop));
    Real s, c;
    m_sincospi(Real(2) * a, &s, &c);
    return frame_around(axis).to_world({sin_theta * c, sin_theta * s, Real(1) - drop});
}
// 1 - cos of the half-angle a sphere of radius r subtends from distance sqrt(d2) > r, without cancellation:
// 1 - sqrt(1 - x) = x / (1 + sqrt(1 - x)) with x = r^2 / d2
template <typename Real>
__device__ inline Real cone_drop(Real radius, Real d2)
{
    const Real x = radius * radius / d2;
    return x / (Real(1) + m_sqrt(Real(1) - x));
}
// |c|^2 in the order the sphere lights' tests take it (the boundary between inside and outside is decided by it)
template <typename Real>
__device__ inline Real squared_distance(vec3<Real> c)
{
    return m_fma(c.z, c.z, m_fma(c.x, c.x, c.y * c.y));
}

// --- the camera (thin lens, stratified pixel samples) -------------------------------------------------------------------
template <typename Real>
struct path
{
    ray<Real> r;
    vec3<Real> throughput;
    int rays_left; // rays the path may still trace (the books' depth)
    int from;      // the primitive the ray leaves (-1: none, a camera ray or a ray out of smoke)
    bool inward;   // the ray leaves `from` into its inside
};

// The camera ray of pass `pass` of pixel (i, j): the pass's cell (pass mod k, pass div k) of the pixel's k x k grid
// jittered uniformly, seen from a uniform point of the lens disk (src/gpu/camera.cuh:53-60); a random shutter time
// when something in the scene moves (src/gpu/camera.cuh:444-457). The image plane is kept relative to the camera's
// centre, so the direction is formed before the centre's large coordinates are added.
template <typename Real>
__device__ inline path<Real> start_path(const camera_rec<Real>& cam, bool motion, int i, int j, int pass, pcg32& g)
{
    const int k = cam.sqrt_spp;
    const Real cell_x = static_cast<Real>(pass % k);
    const Real cell_y = static_cast<Real>(pass / k);
    const Real x = static_cast<Real>(i) + ((cell_x + random_real<Real>(g)) * cam.recip_sqrt_spp - Real(0.5));
    const Real y = static_cast<Real>(j) + ((cell_y + random_real<Real>(g)) * cam.recip_sqrt_spp - Real(0.5));
    vec3<Real> aim = load3(cam.center_to_pixel00) + x * load3(cam.pixel_delta_u) + y * load3(cam.pixel_delta_v);
    vec3<Real> eye = load3(cam.center);
    if (cam.defocus)
    {
        const Real a = random_real<Real>(g);
        const Real b = random_real<Real>(g);
        Real lx, ly;
        concentric_disk(a, b, lx, ly);
        const vec3<Real> lens = lx * load3(cam.defocus_disk_u) + ly * load3(cam.defocus_disk_v);
        eye = eye + lens;
        aim = aim - lens;
    }
    path<Real> p;
    p.r = {eye, aim, motion ? random_real<Real>(g) : Real(0)};
    p.throughput = {Real(1), Real(1), Real(1)};
    p.rays_left = cam.max_depth;
    p.from = -1;
    p.inward = false;
    return p;
}

// The books' background: white to blue by the direction's height
template <typename Real>
__device__ inline vec3<Real> sky(const camera_rec<Real>& cam, vec3<Real> direction)
{
    const Real up = Real(0.5) * (m_rsqrt(length_squared(direction)) * direction.y + Real(1));
    return lerp3(load3(cam.sky_bottom), load3(cam.sky_top), up);
}

// --- the vertex at a hit ------------------------------------------------------------------------------------------------
template <typename Real>
struct vertex
{
    vec3<Real> p;       // where the next ray starts
    vec3<Real> n;       // unit normal on the side the ray came from
    vec3<Real> outward; // unit outward normal (zero in smoke)
    vec3<Real> tex_p;   // where a solid texture is read
    Real u, v;          // texture coordinates (0 unless the material reads them)
    int material;
    int from;           // the primitive the next ray leaves: the hit's, or -1 in smoke
    bool front;         // the ray met the outward side
};

// A sphere's texture coordinates (the books' equirectangular map, turned by the sphere's frame): u from the azimuth
// about +y, v from the polar angle measured from -y
template <typename Real>
__device__ inline void sphere_texcoords(vec3<Real> outward, const sphere_frame<Real>& f, Real& u, Real& v)
{
    const Real x = f.cos_theta * outward.x - f.sin_theta * outward.z;
    const Real z = f.sin_theta * outward.x + f.cos_theta * outward.z;
    const Real azimuth = m_atan2(-z, x) + pi_v<Real>;
    const Real polar = m_acos(m_min(m_max(-outward.y, Real(-1)), Real(1)));
    u = azimuth * (Real(0.5) * one_over_pi<Real>);
    v = polar * one_over_pi<Real>;
}

// The hit point moved onto the sphere by one Newton step on F(p) = |p - c|^2 - r^2 along the outward normal
// (grad F = 2 (p - c) ~ 2 r n): p - F(p) / (2 r) n, with F formed as |p|^2 - 2 p.c + (|c|^2 - r^2) from the frame's
// constant k = |c|^2 - r^2 (exact in double on the host). A checker's cell edges then fall where the double-precision
// CPU puts them. Skipped (p kept) when the terms' size makes F meaningless in this precision: the frame's bound.
template <typename Real>
__device__ inline vec3<Real> newton_onto_sphere(vec3<Real> p, vec3<Real> outward, const sphere<Real>& s,
                                                const sphere_frame<Real>& f)
{
    const Real pp = dot(p, p);
    const Real pc = dot(p, s.center);
    if (!(pp + Real(2) * m_fabs(pc) + m_fabs(f.k) < f.limit))
    {
        return p;
    }
    const Real value = pp - Real(2) * pc + f.k;
    return p - (value / (Real(2) * s.radius)) * outward;
}

// A surface's normal and texture coordinates: one function per kind of primitive (the point, vx.p, is set by
// make_vertex)
template <typename Real, typename Tally>
__device__ inline void sphere_vertex(const scene_view<Real>& sc, const ray<Real>& r, const hit<Real>& h, uint32_t flags,
                                     vertex<Real>& vx, Tally& tl)
{
    const bool moving = h.index >= sc.moving_base;
    const sphere<Real> ball =
        moving ? moving_at(sc.moving_spheres[h.index - sc.moving_base], r.time) : load_sphere(sc.spheres, h.index);
    tl.add(moving ? tally_moving : tally_sphere);
    vx.outward = (Real(1) / ball.radius) * (vx.p - ball.center);
    if (flags & (needs_uv | needs_exact_point))
    {
        const sphere_frame<Real> f = sc.sphere_frames[h.index];
        tl.add(tally_frame);
        if (flags & needs_uv)
        {
            sphere_texcoords(vx.outward, f, vx.u, vx.v);
        }
        if (flags & needs_exact_point)
        {
            vx.tex_p = newton_onto_sphere(vx.p, vx.outward, ball, f);
        }
    }
}
template <typename Real, typename Tally>
__device__ inline void quad_vertex(const scene_view<Real>& sc, const ray<Real>& r, const hit<Real>& h, uint32_t flags,
                                   vertex<Real>& vx, Tally& tl)
{
    const quad_rec<Real> face = sc.quads[h.index - sc.quad_base];
    tl.add(tally_quad);
    vx.outward = load3(face.normal);
    if (flags & needs_uv)
    {
        quad_uv(face, r, h.t, vx.u, vx.v);
    }
}
template <typename Real, typename Tally>
__device__ inline void triangle_vertex(const scene_view<Real>& sc, const ray<Real>& r, const hit<Real>& h,
                                       uint32_t flags, vertex<Real>& vx, Tally& tl)
{
    const triangle_rec<Real> tri = sc.triangles[h.index - sc.triangle_base];
    tl.add(tally_triangle);
    vx.outward = load3(tri.normal);
    if (flags & needs_uv)
    {
        Real t, bary[2];
        bool front;
        hit_triangle(tri, r, -inf<Real>, inf<Real>, t, front, bary); // the same test again: its barycentrics
        vx.u = bary[0];
        vx.v = bary[1];
    }
}

// The instance an instanced triangle's id belongs to: the last whose id_base is no larger (instances are sorted by it)
template <typename Real>
__device__ inline int instance_of(const scene_view<Real>& sc, int id)
{
    int lo = 0, hi = sc.instance_count - 1;
    while (lo < hi)
    {
        const int mid = (lo + hi + 1) >> 1;
        if (sc.instances[mid].id_base <= id)
        {
            lo = mid;
        }
        else
        {
            hi = mid - 1;
        }
    }
    return lo;
}
// A surface's material: from the table, or an instanced triangle's from its instance
template <typename Real>
__device__ inline int material_of(const scene_view<Real>& sc, int id)
{
    return id < sc.instanced_base ? sc.prim_materials[id] : sc.instances[instance_of(sc, id)].material;
}
// An instanced triangle: its mesh's record, its normal carried to the world by the inverse transpose of the
// instance's object-to-world map (the transpose of its world-to-object one, A), its barycentrics from the same test in
// the mesh's coordinates
template <typename Real, typename Tally>
__device__ inline void instanced_vertex(const scene_view<Real>& sc, const instance_rec<Real>& in, const ray<Real>& r,
                                        const hit<Real>& h, uint32_t flags, vertex<Real>& vx, Tally& tl)
{
    const triangle_rec<Real> tri = sc.triangles[in.first_triangle + (h.index - in.id_base)];
    tl.add(tally_triangle);
    const Real* a = in.to_object;
    const vec3<Real> n = load3(tri.normal);
    vx.outward = unit_vector(vec3<Real>{m_fma(a[0], n.x, m_fma(a[4], n.y, a[8] * n.z)),
                                        m_fma(a[1], n.x, m_fma(a[5], n.y, a[9] * n.z)),
                                        m_fma(a[2], n.x, m_fma(a[6], n.y, a[10] * n.z))});
    if (flags & needs_uv)
    {
        Real t, bary[2];
        bool front;
        hit_triangle(tri, into_instance(a, r), -inf<Real>, inf<Real>, t, front, bary);
        vx.u = bary[0];
        vx.v = bary[1];
    }
}

// The vertex of any event: a surface's, or a particle of smoke, which has no surface (any normal will do: the phase
// function ignores it) and whose next ray leaves no primitive. Instances: the scene's tree has two levels (a walk that
// can enter an instance found h)
template <bool Instances, typename Real, typename Tally>
__device__ inline vertex<Real> make_vertex(const scene_view<Real>& sc, const ray<Real>& r, const hit<Real>& h, Tally& tl)
{
    tl.add(tally_hit);
    vertex<Real> vx;
    vx.u = Real(0);
    vx.v = Real(0);
    // The point, once for every kind (one computation for the whole wave): where the walk's distance puts it, on a quad
    // the product t*d rounded before the sum (src/gpu/camera.cuh:184-205: the CPU's point, which lands on the quad's
    // plane or on either side of it as often as the CPU's does; a checker's cells and a lamp in its own lights list
    // depend on the side), elsewhere o + t d with one rounding
    const bool on_quad = h.index >= sc.quad_base && h.index < sc.triangle_base;
    vx.p = on_quad ? r.at_unfused(h.t) : r.at(h.t);
    vx.tex_p = vx.p;
    if (h.index >= sc.media_base)
    {
        tl.add(tally_medium);
        vx.n = {Real(1), Real(0), Real(0)};
        vx.outward = {Real(0), Real(0), Real(0)};
        vx.front = true;
        vx.material = sc.media[h.index - sc.media_base].material;
        vx.from = -1;
        return vx;
    }
    vx.from = h.index;
    vx.front = h.entering;
    if (Instances && h.index >= sc.instanced_base)
    {
        tl.add(tally_instance);
        const instance_rec<Real> in = sc.instances[instance_of(sc, h.index)];
        vx.material = in.material;
        instanced_vertex(sc, in, r, h, sc.materials[vx.material].flags, vx, tl);
        vx.n = vx.front ? vx.outward : -vx.outward;
        return vx;
    }
    vx.material = sc.prim_materials[h.index];
    const uint32_t flags = sc.materials[vx.material].flags;
    if (h.index >= sc.triangle_base)
    {
        triangle_vertex(sc, r, h, flags, vx, tl);
    }
    else if (h.index >= sc.quad_base)
    {
        quad_vertex(sc, r, h, flags, vx, tl);
    }
    else
    {
        sphere_vertex(sc, r, h, flags, vx, tl);
    }
    vx.n = vx.front ? vx.outward : -vx.outward;
    return vx;
}

// --- colors: solid, checker, image, marble ------------------------------------------------------------------------------
// Perlin's gradient noise as the second book builds it (random unit gradients at the lattice points, picked by three
// permutation tables, Hermite-smoothed weights): the eight corners' contributions g . (f - corner), interpolated
// along x, then y, then z
template <typename Real, typename Tally>
__device__ inline Real gradient_noise(const noise_rec<Real>& tb, vec3<Real> p, Tally& tl)
{
    tl.add(tally_perlin, 8);
    const Real cx = m_floor(p.x), cy = m_floor(p.y), cz = m_floor(p.z);
    const vec3<Real> f{p.x - cx, p.y - cy, p.z - cz};
    const vec3<Real> smooth{f.x * f.x * (Real(3) - Real(2) * f.x), f.y * f.y * (Real(3) - Real(2) * f.y),
                            f.z * f.z * (Real(3) - Real(2) * f.z)};
    const int ix = static_cast<int>(cx), iy = static_cast<int>(cy), iz = static_cast<int>(cz);
    const int hx[2] = {tb.perm_x[ix & 255], tb.perm_x[(ix + 1) & 255]};
    const int hy[2] = {tb.perm_y[iy & 255], tb.perm_y[(iy + 1) & 255]};
    const int hz[2] = {tb.perm_z[iz & 255], tb.perm_z[(iz + 1) & 255]};
    // the corners' contributions, corner (a, b, c) at index 4a + 2b + c
    Real g[8];
#pragma unroll
    for (int k = 0; k < 8; ++k)
    {
        const int a = k >> 2, b = (k >> 1) & 1, c = k & 1;
        const Real* grad = tb.gradients[hx[a] ^ hy[b] ^ hz[c]];
        g[k] = dot(vec3<Real>{grad[0], grad[1], grad[2]}, vec3<Real>{f.x - Real(a), f.y - Real(b), f.z - Real(c)});
    }
    Real x[4]; // along x: (b, c) at index 2b + c
#pragma unroll
    for (int k = 0; k < 4; ++k)
    {
        x[k] = g[k] + smooth.x * (g[4 + k] - g[k]);
    }
    const Real y0 = x[0] + smooth.y * (x[2] - x[0]);
    const Real y1 = x[1] + smooth.y * (x[3] - x[1]);
    return y0 + smooth.z * (y1 - y0);
}
// The second book's turbulence: |sum over octaves of 2^-o noise(2^o p)|
template <typename Real, typename Tally>
__device__ inline Real turbulence(const noise_rec<Real>& tb, vec3<Real> p, int octaves, Tally& tl)
{
    Real sum = 0;
    Real amplitude = 1;
    for (int o = 0; o < octaves; ++o)
    {
        sum += amplitude * gradient_noise(tb, p, tl);
        amplitude *= Real(0.5);
        p = Real(2) * p;
    }
    return m_fabs(sum);
}
// The marble's octaves of turbulence (the second book's noise_texture, scale z + 10 turb(p, 7))
constexpr int marble_octaves = 7;
// An image's nearest texel to (u, v), v up, both clamped to [0, 1]
template <typename Real>
__device__ inline vec3<Real> image_color(const scene_view<Real>& sc, int table, Real u, Real v)
{
    const image_info im = sc.images[table];
    const Real s = m_min(m_max(u, Real(0)), Real(1));
    const Real t = Real(1) - m_min(m_max(v, Real(0)), Real(1));
    const int column = min(static_cast<int>(s * static_cast<Real>(im.width)), im.width - 1);
    const int row = min(static_cast<int>(t * static_cast<Real>(im.height)), im.height - 1);
    const texel x = sc.texels[im.offset + row * im.width + column];
    return {static_cast<Real>(x.r), static_cast<Real>(x.g), static_cast<Real>(x.b)};
}
// A texture's color at the vertex: a checker picks one of its two textures by the parity of the point's cell (and may
// nest), an image by (u, v), the marble by 1/2 (1 + sin(scale z + 10 turbulence(p))) in gray
// (turb: the marble's turbulence at the vertex when the wave computed it beforehand, wave_turbulence; else null)
template <typename Real, typename Tally>
__device__ inline vec3<Real> texture_color(const scene_view<Real>& sc, int index, const vertex<Real>& vx, Tally& tl,
                                           const Real* turb = nullptr)
{
    texture_rec<Real> tex = sc.textures[index];
    tl.add(tally_texture);
    for (; tex.kind == RT_TEXTURE_CHECKER; tl.add(tally_texture))
    {
        const int cell = static_cast<int>(m_floor(tex.inv_scale * vx.tex_p.x)) +
                         static_cast<int>(m_floor(tex.inv_scale * vx.tex_p.y)) +
                         static_cast<int>(m_floor(tex.inv_scale * vx.tex_p.z));
        tex = sc.textures[cell % 2 == 0 ? tex.even : tex.odd];
    }
    switch (tex.kind)
    {
    case RT_TEXTURE_IMAGE:
        tl.add(tally_texel);
        return image_color(sc, tex.table, vx.u, vx.v);
    case RT_TEXTURE_NOISE:
    {
        const noise_rec<Real>& marble = sc.noises[tex.table];
        const Real phase =
            marble.scale * vx.tex_p.z + Real(10) * (turb ? *turb : turbulence(marble, vx.tex_p, marble_octaves, tl));
        const Real gray = Real(0.5) * (Real(1) + m_sin(phase));
        return {gray, gray, gray};
    }
    default:
        return load3(tex.rgb);
    }
}
// A material's color at the vertex: its albedo, or its texture's color
template <typename Real, typename Tally>
__device__ inline vec3<Real> material_color(const scene_view<Real>& sc, const material_rec<Real>& m,
                                            const vertex<Real>& vx, Tally& tl, const Real* turb = nullptr)
{
    const int index = m.texture; // (named first: `m.texture <` would parse as a template's arguments)
    return index < 0 ? load3(m.albedo) : texture_color(sc, index, vx, tl, turb);
}

// The noise table whose marble gives a material's color at the vertex (the texture chain followed as texture_color
// follows it), or -1 when its color comes from anything else
template <typename Real>
__device__ inline int marble_table(const scene_view<Real>& sc, const material_rec<Real>& m, const vertex<Real>& vx)
{
    const int index = m.texture;
    if (index < 0)
    {
        return -1;
    }
    texture_rec<Real> tex = sc.textures[index];
    while (tex.kind == RT_TEXTURE_CHECKER)
    {
        const int cell = static_cast<int>(m_floor(tex.inv_scale * vx.tex_p.x)) +
                         static_cast<int>(m_floor(tex.inv_scale * vx.tex_p.y)) +
                         static_cast<int>(m_floor(tex.inv_scale * vx.tex_p.z));
        tex = sc.textures[cell % 2 == 0 ? tex.even : tex.odd];
    }
    return tex.kind == RT_TEXTURE_NOISE ? tex.table : -1;
}

// The marble's turbulence for every lane of a wave that asks for one (table >= 0: its noise table, at its point p),
// computed by the whole wave together, one asking lane after the other: lane 8 o + k takes corner k of octave o (56
// lanes for the 7 octaves), the interpolations along x, y and z pair lanes by exchanges, and the octaves are summed in
// order. Every value is formed by the same operations on the same operands as turbulence()'s loop (a point of octave
// o is 2^o p exactly, as repeated doubling gives it), so the result is the same bits; a marble met by one lane of a
// wave costs the wave one lane's corner instead of the 56 corners one after the other. Call with every lane of the
// wave active. Float only.
__device__ inline float wave_turbulence(const scene_view<float>& sc, int table, vec3<float> p)
{
    uint64_t ask = __ballot(table >= 0);
    float mine = 0.0f;
    const int lane = static_cast<int>(__lane_id());
    const int octave = lane >> 3;
    const int a = (lane >> 2) & 1, b = (lane >> 1) & 1, c = lane & 1; // the corner, at index 4a + 2b + c
    while (ask != 0)
    {
        const int who = __builtin_ctzll(ask);
        ask &= ask - 1;
        // the asking lane's table and point, the same in every lane
        const int t = __builtin_amdgcn_readlane(table, who);
        const vec3<float> q{__int_as_float(__builtin_amdgcn_readlane(__float_as_int(p.x), who)),
                            __int_as_float(__builtin_amdgcn_readlane(__float_as_int(p.y), who)),
                            __int_as_float(__builtin_amdgcn_readlane(__float_as_int(p.z), who))};
        const noise_rec<float>& tb = sc.noises[t];
        // this lane's octave's point and its lattice cell, and its corner's contribution (gradient_noise's terms)
        const float scale = static_cast<float>(1 << (octave < marble_octaves ? octave : 0));
        const vec3<float> po{scale * q.x, scale * q.y, scale * q.z};
        const float cx = floorf(po.x), cy = floorf(po.y), cz = floorf(po.z);
        const vec3<float> f{po.x - cx, po.y - cy, po.z - cz};
        const vec3<float> smooth{f.x * f.x * (3.0f - 2.0f * f.x), f.y * f.y * (3.0f - 2.0f * f.y),
                                 f.z * f.z * (3.0f - 2.0f * f.z)};
        const int ix = static_cast<int>(cx), iy = static_cast<int>(cy), iz = static_cast<int>(cz);
        const int h = tb.perm_x[(ix + a) & 255] ^ tb.perm_y[(iy + b) & 255] ^ tb.perm_z[(iz + c) & 255];
        const float* grad = tb.gradients[h];
        const float g = dot(vec3<float>{grad[0], grad[1], grad[2]},
                            vec3<float>{f.x - static_cast<float>(a), f.y - static_cast<float>(b), f.z - static_cast<float>(c)});
        // along x: corner (0, b, c) with (1, b, c); along y: (0, c) with (1, c); along z
        const float gx = __shfl_xor(g, 4, 64);
        const float x = g + smooth.x * (gx - g);
        const float xy = __shfl_xor(x, 2, 64);
        const float y = x + smooth.y * (xy - x);
        const float yz = __shfl_xor(y, 1, 64);
        const float noise = y + smooth.z * (yz - y);
        // the octaves in order (lane 8 o holds octave o's noise)
        float sum = 0.0f;
        float amplitude = 1.0f;
#pragma unroll
        for (int o = 0; o < marble_octaves; ++o)
        {
            sum += amplitude * __int_as_float(__builtin_amdgcn_readlane(__float_as_int(noise), 8 * o));
            amplitude *= 0.5f;
        }
        mine = lane == who ? fabsf(sum) : mine;
    }
    return mine;
}

// --- smoke: constant-density media ------------------------------------------------------------------------------------
// The stretch of the ray's line inside a medium's boundary, [enter, leave]: a sphere's two roots (its centre where it
// is at the ray's time), or for a convex body of quads the first and last of its planes' crossings; false when the
// line misses or only grazes it (src/gpu/constant_medium.cuh:48-80: chords shorter than 1e-4 are no crossing)
template <typename Real, typename Tally>
__device__ inline bool smoke_chord(const medium_rec<Real>& m, const quad_rec<Real>* quads, const ray<Real>& r,
                                   Real inverse_a, Real& enter, Real& leave, Tally& tl)
{
    if (m.quad_count == 0)
    {
        const vec3<Real> centre{m_fma(r.time, m.displacement[0], m.center[0]),
                                m_fma(r.time, m.displacement[1], m.center[1]),
                                m_fma(r.time, m.displacement[2], m.center[2])};
        return sphere_roots(sphere<Real>{centre, m.radius}, r, inverse_a, enter, leave) &&
               leave - enter > Real(0.0001);
    }
    enter = inf<Real>;
    leave = -inf<Real>;
    for (int k = 0; k < m.quad_count; ++k)
    {
        tl.add(tally_boundary_quad);
        Real t;
        bool front;
        if (hit_quad(quads[m.first_quad + k], r, -inf<Real>, inf<Real>, t, front))
        {
            enter = m_min(enter, t);
            leave = m_max(leave, t);
        }
    }
    return leave - enter >= Real(0.0001);
}
// The part of [t_lo, t_hi] inside the medium, or false: its box first (cheap), then the chord
template <typename Real, typename Tally>
__device__ inline bool smoke_span(const medium_rec<Real>& m, const quad_rec<Real>* quads, const ray<Real>& r,
                                  Real inverse_a, vec3<Real> inv, vec3<Real> os, Real t_lo, Real t_hi, Real& a,
                                  Real& b, Tally& tl)
{
    if (box_entry(m.box_min, m.box_max, inv, os, t_lo, t_hi) == inf<Real>)
    {
        return false;
    }
    Real enter, leave;
    if (!smoke_chord(m, quads, r, inverse_a, enter, leave, tl))
    {
        return false;
    }
    a = m_max(m_max(enter, t_lo), Real(0));
    b = m_min(leave, t_hi);
    return m_max(enter, t_lo) < b;
}
// A free flight through the span [a, b] of a medium of density rho: the distance to the next collision is
// exponential, -ln(xi) / rho; the collision's t, or false when the flight leaves the span
template <typename Real>
__device__ inline bool free_flight(const medium_rec<Real>& m, const ray<Real>& r, Real a, Real b, pcg32& g, Real& t)
{
    const Real speed = m_sqrt(length_squared(r.direction)); // distance per unit of t
    const Real flight = m.neg_inv_density * m_log(random_real<Real>(g));
    if (flight > (b - a) * speed)
    {
        return false;
    }
    t = a + flight / speed;
    return true;
}

// The surfaces' closest hit, then each medium in the scene's order against the closest event so far: a collision in
// front of it replaces it (src/gpu/camera.cuh:259-284). A medium's id is media_base + its index.
template <typename Real, typename Tally>
__device__ inline hit<Real> media_events(const scene_view<Real>& sc, const ray<Real>& r, hit<Real> closest, pcg32& g,
                                         Tally& tl)
{
    if (sc.medium_count == 0)
    {
        return closest;
    }
    const Real inverse_a = Real(1) / length_squared(r.direction);
    const vec3<Real> inv{safe_inverse(r.direction.x), safe_inverse(r.direction.y), safe_inverse(r.direction.z)};
    const vec3<Real> os{r.origin.x * inv.x, r.origin.y * inv.y, r.origin.z * inv.z};
    for (int k = 0; k < sc.medium_count; ++k)
    {
        tl.add(tally_medium);
        const medium_rec<Real>& m = sc.media[k];
        Real a, b, t;
        if (smoke_span(m, sc.boundary_quads, r, inverse_a, inv, os, Real(0.001), closest.t, a, b, tl) &&
            free_flight(m, r, a, b, g, t))
        {
            closest = {sc.media_base + k, t, true};
        }
    }
    return closest;
}
template <typename Real, int Kinds = walk_instances, typename Tally, typename Stack>
__device__ inline hit<Real> closest_event(const scene_view<Real>& sc, const ray<Real>& r, int from, bool inward,
                                          pcg32& g, const Stack& stack, Tally& tl)
{
    return media_events(sc, r, closest_hit<Real, Kinds>(sc, r, from, inward, stack, tl), g, tl);
}

// --- the books' lobes ---------------------------------------------------------------------------------------------------
// How a path leaves a vertex of each material kind
enum class books_lobe : int
{
    emitter, // a lamp: the path ends with its light (its front face only)
    matte,   // Lambert's cosine lobe
    smoke,   // the isotropic phase function
    mirror,  // a fuzzed mirror
    glass,   // a dielectric: Fresnel's choice between reflection and refraction
};
__device__ inline books_lobe lobe_of(int kind)
{
    switch (kind)
    {
    case RT_MATERIAL_DIFFUSE_LIGHT:
        return books_lobe::emitter;
    case RT_MATERIAL_ISOTROPIC:
        return books_lobe::smoke;
    case RT_MATERIAL_METAL:
        return books_lobe::mirror;
    case RT_MATERIAL_DIELECTRIC:
        return books_lobe::glass;
    default:
        return books_lobe::matte;
    }
}

template <typename Real>
__device__ inline vec3<Real> mirror_of(vec3<Real> d, vec3<Real> n)
{
    return d - (Real(2) * dot(d, n)) * n;
}

// The mirror's direction, blurred by a uniform point of a sphere of radius `fuzz`; false when that sends it below the
// surface (the light is absorbed)
template <typename Real>
__device__ inline bool fuzzed_mirror(vec3<Real> d, vec3<Real> n, Real fuzz, pcg32& g, vec3<Real>& out)
{
    const vec3<Real> m = mirror_of(d, n);
    out = m_rsqrt(length_squared(m)) * m + fuzz * random_unit_vector<Real>(g);
    return dot(out, n) > Real(0);
}

// Schlick's approximation of the unpolarised Fresnel reflectance at cos(theta_i), relative index eta
template <typename Real>
__device__ inline Real schlick_reflectance(Real cos_i, Real eta)
{
    const Real root = (Real(1) - eta) / (Real(1) + eta);
    const Real f0 = root * root;
    const Real x = Real(1) - cos_i;
    const Real x2 = x * x;
    return f0 + (Real(1) - f0) * (x2 * x2 * x);
}
// A dielectric interface: with unit d, n facing d's origin and eta = n_i / n_t, Snell's law gives the refracted
// direction eta d + (eta cos_i - cos_t) n with cos_t = sqrt(1 - eta^2 (1 - cos_i^2)); when that root is imaginary
// the light is totally reflected, otherwise it is reflected with Schlick's probability
template <typename Real>
__device__ inline vec3<Real> dielectric_direction(vec3<Real> d, vec3<Real> n, bool front, Real ior, pcg32& g)
{
    const Real eta = front ? Real(1) / ior : ior;
    const vec3<Real> ud = m_rsqrt(length_squared(d)) * d;
    const Real cos_i = m_min(-dot(ud, n), Real(1));
    const Real cos_t2 = Real(1) - eta * eta * (Real(1) - cos_i * cos_i);
    if (cos_t2 < Real(0) || schlick_reflectance(cos_i, eta) > random_real<Real>(g))
    {
        return mirror_of(ud, n);
    }
    return eta * ud + (eta * cos_i - m_sqrt(cos_t2)) * n;
}

// --- the third book's lights list ---------------------------------------------------------------------------------------
// Densities per solid angle of directions from x toward a light of the list, and draws of them. A quad light: points
// uniform on its area A seen at distance s under cos theta_l have density s^2 / (A cos theta_l). A sphere light seen
// from outside: directions uniform in the cone it subtends, 1 / (2 pi (1 - cos theta_max)); from inside, the whole
// sphere of directions. A light is "own" when the ray leaves it (its id is the light's twin primitive): an own quad
// can't be hit again, and an own sphere only from inside (src/gpu/lights.cuh:52-226).
template <typename Real>
__device__ inline Real quad_light_density(const light_rec<Real>& l, vec3<Real> x, vec3<Real> d)
{
    Real t;
    bool front;
    if (!hit_quad(l.quad, ray<Real>{x, d, Real(0)}, Real(0.001), inf<Real>, t, front))
    {
        return Real(0);
    }
    const Real d2 = length_squared(d);
    const Real cos_l = m_fabs(dot(d, load3(l.quad.normal))) * m_rsqrt(d2);
    return (t * t * d2) / (cos_l * l.area);
}
template <typename Real>
__device__ inline Real sphere_light_density(const light_rec<Real>& l, vec3<Real> x, vec3<Real> d, bool own)
{
    const vec3<Real> to_centre = load3(l.center) - x;
    const Real d2 = squared_distance(to_centre);
    if (d2 < l.radius * l.radius)
    {
        return one_over_four_pi<Real>;
    }
    if (own && !(dot(d, to_centre) > Real(0)))
    {
        return Real(0);
    }
    Real t;
    bool entering;
    if (!hit_sphere(sphere<Real>{load3(l.center), l.radius}, ray<Real>{x, d, Real(0)}, Real(1) / length_squared(d),
                    Real(0.001), inf<Real>, own, t, entering))
    {
        return Real(0);
    }
    return Real(0.5) * one_over_pi<Real> / cone_drop(l.radius, d2);
}
template <typename Real>
__device__ inline vec3<Real> toward_light(const light_rec<Real>& l, vec3<Real> x, pcg32& g)
{
    const Real a = random_real<Real>(g);
    const Real b = random_real<Real>(g);
    if (l.kind == RT_LIGHT_QUAD)
    {
        const vec3<Real> to = (load3(l.quad.q) - x) + a * load3(l.u) + b * load3(l.v);
        return m_rsqrt(length_squared(to)) * to;
    }
    const vec3<Real> to_centre = load3(l.center) - x;
    const Real d2 = squared_distance(to_centre);
    if (d2 < l.radius * l.radius)
    {
        // from inside, every direction meets the sphere: uniform over the sphere of directions (a, b reused)
        const Real z = Real(1) - Real(2) * a;
        Real s, c;
        m_sincospi(Real(2) * b, &s, &c);
        const Real rxy = m_sqrt(m_max(Real(0), Real(1) - z * z));
        return {rxy * c, rxy * s, z};
    }
    return cone_direction(m_rsqrt(d2) * to_centre, cone_drop(l.radius, d2), a, b);
}
// The list's mixture: sum of weight_k density_k, and a draw of a light by its cumulative weights then a direction
template <typename Real, typename Tally>
__device__ inline Real lights_density(const scene_view<Real>& sc, vec3<Real> x, vec3<Real> d, int from, Tally& tl)
{
    tl.add(tally_light, static_cast<uint32_t>(sc.light_count));
    Real sum = 0;
    for (int k = 0; k < sc.light_count; ++k)
    {
        const light_rec<Real>& l = sc.lights[k];
        const bool own = from >= 0 && l.twin == from;
        if (l.kind == RT_LIGHT_QUAD)
        {
            if (!own)
            {
                sum += l.weight * quad_light_density(l, x, d);
            }
        }
        else
        {
            sum += l.weight * sphere_light_density(l, x, d, own);
        }
    }
    return sum;
}
template <typename Real, typename Tally>
__device__ inline vec3<Real> lights_draw(const scene_view<Real>& sc, vec3<Real> x, pcg32& g, Tally& tl)
{
    tl.add(tally_light);
    const Real pick = random_real<Real>(g);
    int k = 0;
    for (; k + 1 < sc.light_count && !(pick < sc.lights[k].cdf); ++k)
    {
    }
    return toward_light(sc.lights[k], x, g);
}

// A diffuse vertex's next direction under the scene's sampling strategy (the third book's; src/gpu/camera.cuh:
// 304-359), and the factor by which the throughput is multiplied besides the material's color: 1 when the direction
// was drawn from the lobe's own density, p_lobe / p_strategy otherwise, 0 (the path ends) when either density is 0.
//   material: the lobe itself (a matte lobe's draw below the surface ends the path);
//   lights: a direction of the lights list;
//   mixture: lights or the lobe, one half each;
//   uniform: the whole sphere.
// Smoke samples its own (uniform) phase function under every strategy but `lights`.
template <typename Real, typename Tally>
__device__ inline Real diffuse_direction(const scene_view<Real>& sc, bool matte, const vertex<Real>& vx, pcg32& g,
                                         vec3<Real>& d, Tally& tl)
{
    const int strategy = (!matte && sc.sampling != RT_SAMPLING_LIGHTS) ? RT_SAMPLING_MATERIAL : sc.sampling;
    const bool from_lights = strategy == RT_SAMPLING_LIGHTS ||
                             (strategy == RT_SAMPLING_MIXTURE && random_real<Real>(g) < Real(0.5));
    if (from_lights)
    {
        d = lights_draw(sc, vx.p, g, tl);
    }
    else if (!matte || strategy == RT_SAMPLING_UNIFORM)
    {
        d = random_unit_vector<Real>(g);
    }
    else
    {
        d = cosine_about(vx.n, g);
    }
    if (strategy == RT_SAMPLING_MATERIAL)
    {
        return (!matte || dot(vx.n, d) > Real(0)) ? Real(1) : Real(0);
    }
    const Real p_lobe = matte ? cosine_density(vx.n, d) : one_over_four_pi<Real>;
    Real p = one_over_four_pi<Real>;
    if (strategy != RT_SAMPLING_UNIFORM)
    {
        const Real p_lights = lights_density(sc, vx.p, d, vx.from, tl);
        p = strategy == RT_SAMPLING_LIGHTS ? p_lights : Real(0.5) * (p_lights + p_lobe);
    }
    return (p > Real(0) && p_lobe > Real(0)) ? p_lobe / p : Real(0);
}

// --- one step of a books' path ------------------------------------------------------------------------------------------
// The path's next ray leaves the vertex in direction d: the depth rule first (the path that may trace no more rays
// ends black, src/gpu/camera.cuh:420-430), then the ray, the primitive it leaves and the side it leaves into
template <typename Real>
__device__ inline bool continue_path(path<Real>& p, const vertex<Real>& vx, vec3<Real> d, vec3<Real>& sample)
{
    if (--p.rays_left == 0)
    {
        sample = {Real(0), Real(0), Real(0)};
        return true;
    }
    p.r = {vx.p, d, p.r.time};
    p.from = vx.from;
    p.inward = dot(d, vx.outward) < Real(0);
    return false;
}
// Ends the path with `light` reaching the camera through its throughput
template <typename Real>
__device__ inline bool end_path(const path<Real>& p, vec3<Real> light, vec3<Real>& sample)
{
    sample = p.throughput * light;
    return true;
}

// Whether a vertex's lobe reads the material's color (a lamp's back face is black)
__device__ inline bool reads_color(books_lobe lobe, bool front)
{
    return lobe == books_lobe::matte || lobe == books_lobe::smoke || (lobe == books_lobe::emitter && front);
}

// What an event does to the path at its vertex vx of material m (true: the sample is complete, its light in
// `sample`); turb: the vertex's marble turbulence when the wave computed it beforehand (wave_turbulence), else null
template <typename Real, typename Tally>
__device__ inline bool shade_at(path<Real>& p, const scene_view<Real>& sc, const vertex<Real>& vx,
                                const material_rec<Real>& m, pcg32& g, vec3<Real>& sample, Tally& tl,
                                const Real* turb = nullptr)
{
    const vec3<Real> black{Real(0), Real(0), Real(0)};
    const books_lobe lobe = lobe_of(m.kind);
    // the color is read once, for the lobes that have one
    const vec3<Real> color = reads_color(lobe, vx.front) ? material_color(sc, m, vx, tl, turb) : black;
    vec3<Real> d;
    switch (lobe)
    {
    case books_lobe::emitter:
        return end_path(p, color, sample);
    case books_lobe::matte:
    case books_lobe::smoke:
    {
        const Real weight = diffuse_direction(sc, lobe == books_lobe::matte, vx, g, d, tl);
        if (!(weight > Real(0)))
        {
            return end_path(p, black, sample);
        }
        p.throughput = weight * (p.throughput * color);
        break;
    }
    case books_lobe::mirror:
        if (!fuzzed_mirror(p.r.direction, vx.n, m.fuzz, g, d))
        {
            return end_path(p, black, sample);
        }
        p.throughput = p.throughput * load3(m.albedo);
        break;
    case books_lobe::glass:
        d = dielectric_direction(p.r.direction, vx.n, vx.front, m.refraction_index, g);
        break;
    }
    return continue_path(p, vx, d, sample);
}
// What the event h does to the path (true: the sample is complete, its light in `sample`)
template <typename Real, typename Tally>
__device__ inline bool shade(path<Real>& p, const scene_view<Real>& sc, const hit<Real>& h, pcg32& g,
                             vec3<Real>& sample, Tally& tl)
{
    const vertex<Real> vx = make_vertex<false>(sc, p.r, h, tl);
    const material_rec<Real> m = sc.materials[vx.material];
    return shade_at(p, sc, vx, m, g, sample, tl);
}

// One ray of a path with every feature of the three books: the walk, the media, then the event (src/gpu/camera.cuh:
// 364-437); a ray that meets nothing sees the sky
template <typename Real, typename Tally>
__device__ inline bool all_books_after_walk(path<Real>& p, const scene_view<Real>& sc, const camera_rec<Real>& cam,
                                            pcg32& g, vec3<Real>& sample, const hit<Real>& surface_hit, Tally& tl)
{
    const hit<Real> h = media_events(sc, p.r, surface_hit, g, tl);
    tl.lap(phase_media);
    const bool done = h.index < 0 ? end_path(p, sky(cam, p.r.direction), sample) : shade(p, sc, h, g, sample, tl);
    tl.lap(phase_shade);
    return done;
}
template <typename Real, typename Tally, typename Stack>
__device__ inline bool trace_all_books(path<Real>& p, const scene_view<Real>& sc, const camera_rec<Real>& cam,
                                       pcg32& g, vec3<Real>& sample, const Stack& stack, Tally& tl)
{
    const hit<Real> h = closest_hit<Real, walk_books>(sc, p.r, p.from, p.inward, stack, tl);
    tl.lap(phase_walk);
    return all_books_after_walk(p, sc, cam, g, sample, h, tl);
}

// trace_all_books for a lane of a wave whose every lane calls it together (busy: the lane has a path; the others keep
// the wave whole): the walk and the media, then the vertex and its material, then the marble turbulence of every lane
// that needs one, by the whole wave (wave_turbulence), then the shading. The same numbers as trace_all_books, in the
// same order. Float only.
template <typename Tally, typename Stack>
__device__ inline bool trace_all_books_wave(path<float>& p, const scene_view<float>& sc, const camera_rec<float>& cam,
                                            pcg32& g, vec3<float>& sample, const Stack& stack, Tally& tl, bool busy)
{
    hit<float> h{-1, inf<float>, 0};
    if (busy)
    {
        h = closest_hit<float, walk_books>(sc, p.r, p.from, p.inward, stack, tl);
        tl.lap(phase_walk);
        h = media_events(sc, p.r, h, g, tl);
        tl.lap(phase_media);
    }
    vertex<float> vx{};
    material_rec<float> m{};
    int table = -1;
    if (busy && h.index >= 0)
    {
        vx = make_vertex<false>(sc, p.r, h, tl);
        m = sc.materials[vx.material];
        table = reads_color(lobe_of(m.kind), vx.front) ? marble_table(sc, m, vx) : -1;
    }
    const float turb = wave_turbulence(sc, table, vx.tex_p);
    if (table >= 0)
    {
        tl.add(tally_perlin, 8 * marble_octaves);
    }
    if (!busy)
    {
        return false;
    }
    const bool done = h.index < 0 ? end_path(p, sky(cam, p.r.direction), sample)
                                  : shade_at(p, sc, vx, m, g, sample, tl, &turb);
    tl.lap(phase_shade);
    return done;
}

// What the hit h does to a first-book path (spheres that stand still, solid matte, mirrors, glass, the sky): the same
// lobes with only what those scenes need, so the kernel carries no texture, smoke or light code. The matte lobe is the
// first book's scatter, which always goes on (src/gpu/material.cuh: lambertian), where the third book's strategies
// end a path whose direction falls below the surface (diffuse_direction): the two differ on a rare draw, so every
// integrator shades a first-book scene with this function.
template <typename Real, typename Tally>
__device__ inline bool first_book_after_walk(path<Real>& p, const scene_view<Real>& sc, const camera_rec<Real>& cam,
                                             pcg32& g, vec3<Real>& sample, const hit<Real>& h, Tally& tl)
{
    if (h.index < 0)
    {
        return end_path(p, sky(cam, p.r.direction), sample);
    }
    tl.add(tally_hit);
    const sphere<Real> ball = load_sphere(sc.spheres, h.index);
    vertex<Real> vx;
    vx.p = p.r.at(h.t);
    vx.outward = (Real(1) / ball.radius) * (vx.p - ball.center);
    vx.front = h.entering;
    vx.n = vx.front ? vx.outward : -vx.outward;
    vx.from = h.index;
    const material_rec<Real> m = sc.materials[sc.prim_materials[h.index]];
    vec3<Real> d;
    switch (lobe_of(m.kind))
    {
    case books_lobe::mirror:
        if (!fuzzed_mirror(p.r.direction, vx.n, m.fuzz, g, d))
        {
            return end_path(p, vec3<Real>{Real(0), Real(0), Real(0)}, sample);
        }
        p.throughput = p.throughput * load3(m.albedo);
        break;
    case books_lobe::glass:
        d = dielectric_direction(p.r.direction, vx.n, vx.front, m.refraction_index, g);
        break;
    default:
        d = cosine_about(vx.n, g);
        p.throughput = p.throughput * load3(m.albedo);
        break;
    }
    return continue_path(p, vx, d, sample);
}
template <typename Real, typename Tally, typename Stack>
__device__ inline bool trace_first_book(path<Real>& p, const scene_view<Real>& sc, const camera_rec<Real>& cam,
                                        pcg32& g, vec3<Real>& sample, const Stack& stack, Tally& tl)
{
    const hit<Real> h = closest_hit<Real, walk_spheres>(sc, p.r, p.from, p.inward, stack, tl);
    tl.lap(phase_walk);
    const bool done = first_book_after_walk(p, sc, cam, g, sample, h, tl);
    tl.lap(phase_shade);
    return done;
}

// A NaN channel of a sample counts as 0 (src/gpu/renderer.cu:311-316)
template <typename Real>
__device__ inline double without_nan(Real x)
{
    return static_cast<double>(x == x ? x : Real(0));
}
}
