// Parity probes: kernels that expose one part of the path on its own, so that tests can check it against the oracle
// bit for bit (the reference's GpuTests drive its device functions through test kernels the same way,
// tests/test_scene.h).
//
//   - reference_walk_kernel: the closest hit as the reference's own traversal finds it. The product walks a four-wide
//     tree and gives a tie to the primitive of lower scene index; this kernel walks the session's binary tree in the
//     reference's order with the reference's intervals, so that on tied rays too its ids are the reference's.
//   - sample_numbers_kernel: the per-sample generator's raw words, its float and double draws, and the direction
//     samplers built on them, for given (seed, pixel, pass).
//   - occluded_kernel: a shadow ray's answer from both of the product's shadow walks, for given rays, lengths and
//     skipped primitives.
//   - scatter_kernel: the quality path's scattering lobes on their own, in the tangent frame: the value and density of
//     given directions, and directions drawn from them.
//   - media_kernel: the media's events along given rays (a free flight through each medium's chord) and the
//     transmittance a shadow ray is weighted by.
//   - light_kernel: the quality path's light samples from given points (the light set's choice, a lamp's point or the
//     map's direction), with the density a scattered ray finding the same light would be given.
#include <algorithm>
#include <cstdint>

#include "kernel_api.hpp"
#include "books.hpp"
#include "quality.hpp"

namespace rt::dev
{
// The reference's 32-byte node (src/gpu/bvh.cuh:29-38), as bvh_build.hpp's ref_node<float> lays it out: an inner
// node's second child is `index` and its first child follows it; a leaf holds count & 255 primitives of kind
// count >> 8 from `index` on
struct binary_node
{
    float box_min[3];
    int32_t index;
    float box_max[3];
    int32_t count;
};

// The binary near-first walk of src/gpu/bvh.cuh:576-706, in its order: both children's boxes tested against the
// closest hit so far; the nearer one next (the first child when they are entered at the same distance), the farther
// one pushed if entered; a popped node is visited only if it is entered before the closest hit. Spheres are tested on
// the open interval (0.001, closest) and quads on the closed one [0.001, closest] (src/gpu/sphere.cuh:116,
// quad.cuh:70), and a primitive that passes replaces the closest hit: a tie goes to the first sphere and to the last
// quad the walk meets, and to a quad over a sphere. Triangles, which the reference's walk doesn't hold, are tested as
// quads are. The stack is private (the reference's 32 entries; 64 here, with a check).
__global__ void __launch_bounds__(256) reference_walk_kernel(const scene_view<float> sc,
                                                             const binary_node* __restrict__ nodes, int node_count,
                                                             const float* __restrict__ origins,
                                                             const float* __restrict__ directions,
                                                             const float* __restrict__ times,
                                                             const int32_t* __restrict__ from,
                                                             const uint8_t* __restrict__ inward, int64_t n,
                                                             int32_t* __restrict__ index, float* __restrict__ t_out,
                                                             uint8_t* __restrict__ entering,
                                                             int32_t* __restrict__ overflow)
{
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n)
    {
        return;
    }
    const ray<float> r{{origins[3 * k], origins[3 * k + 1], origins[3 * k + 2]},
                       {directions[3 * k], directions[3 * k + 1], directions[3 * k + 2]},
                       times[k]};
    const int own_id = from[k];
    const bool going_in = inward[k] != 0;
    const float t_min = 0.001f;
    hit<float> closest{-1, inf<float>, false};
    if (node_count > 0)
    {
        const float inverse_a = 1.0f / length_squared(r.direction);
        const vec3<float> inv{safe_inverse(r.direction.x), safe_inverse(r.direction.y), safe_inverse(r.direction.z)};
        const vec3<float> os{r.origin.x * inv.x, r.origin.y * inv.y, r.origin.z * inv.z};
        constexpr int capacity = 64;
        int pending[capacity];
        float pending_t[capacity];
        int size = 0;
        int at = 0;
        while (at >= 0)
        {
            const binary_node nd = nodes[at];
            if (nd.count > 0)
            {
                const int kind = nd.count >> 8;
                const int end = nd.index + (nd.count & 255);
                for (int i = nd.index; i < end; ++i)
                {
                    float t;
                    bool side;
                    if (kind == prim_quad || kind == prim_triangle)
                    {
                        const int id = (kind == prim_quad ? sc.quad_base : sc.triangle_base) + i;
                        if (id == own_id)
                        {
                            continue;
                        }
                        const bool met = kind == prim_quad ? hit_quad(sc.quads[i], r, t_min, closest.t, t, side)
                                                           : hit_triangle(sc.triangles[i], r, t_min, closest.t, t, side);
                        if (met)
                        {
                            closest = {id, t, side};
                        }
                    }
                    else
                    {
                        const bool moving = kind == prim_moving_sphere;
                        const int id = moving ? sc.moving_base + i : i;
                        const bool own = id == own_id;
                        if (own && !going_in)
                        {
                            continue;
                        }
                        const sphere<float> s =
                            moving ? moving_at(sc.moving_spheres[i], r.time) : load_sphere(sc.spheres, i);
                        if (hit_sphere(s, r, inverse_a, t_min, closest.t, own, t, side))
                        {
                            closest = {id, t, side};
                        }
                    }
                }
                at = -1;
            }
            else
            {
                const int first = at + 1;
                const int second = nd.index;
                const binary_node a = nodes[first];
                const binary_node b = nodes[second];
                const float ta = box_entry(a.box_min, a.box_max, inv, os, t_min, closest.t);
                const float tb = box_entry(b.box_min, b.box_max, inv, os, t_min, closest.t);
                const bool second_first = tb < ta;
                const float far_t = m_max(ta, tb);
                if (far_t < inf<float>)
                {
                    if (size == capacity)
                    {
                        atomicAdd(overflow, 1);
                        break;
                    }
                    pending[size] = second_first ? first : second;
                    pending_t[size] = far_t;
                    ++size;
                }
                at = m_min(ta, tb) < inf<float> ? (second_first ? second : first) : -1;
            }
            while (at < 0 && size > 0)
            {
                --size;
                if (pending_t[size] < closest.t)
                {
                    at = pending[size];
                }
            }
        }
    }
    index[k] = closest.index;
    t_out[k] = closest.t;
    entering[k] = closest.entering ? 1 : 0;
}

// Item k: the generator of (seed, pixels[k], passes[k]) and what the path draws from it. Each output comes from a
// fresh generator of the item: `draws` raw words, `draws` float draws, `draws` double draws (two words each), and nine
// floats: a uniform unit vector, a uniform point of the unit disk and a cosine-weighted direction about +z
__global__ void __launch_bounds__(256) sample_numbers_kernel(uint64_t seed, const uint64_t* __restrict__ pixels,
                                                             const uint64_t* __restrict__ passes, int64_t n, int draws,
                                                             uint32_t* __restrict__ words, float* __restrict__ floats,
                                                             double* __restrict__ doubles,
                                                             float* __restrict__ directions)
{
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n)
    {
        return;
    }
    const uint64_t pixel = pixels[k];
    const uint64_t pass = passes[k];
    const int64_t base = k * draws;
    if (words)
    {
        pcg32 g = sample_generator(seed, pixel, pass);
        for (int i = 0; i < draws; ++i)
        {
            words[base + i] = g.next();
        }
    }
    if (floats)
    {
        pcg32 g = sample_generator(seed, pixel, pass);
        for (int i = 0; i < draws; ++i)
        {
            floats[base + i] = random_real<float>(g);
        }
    }
    if (doubles)
    {
        pcg32 g = sample_generator(seed, pixel, pass);
        for (int i = 0; i < draws; ++i)
        {
            doubles[base + i] = random_real<double>(g);
        }
    }
    if (directions)
    {
        pcg32 g = sample_generator(seed, pixel, pass);
        const vec3<float> u = random_unit_vector<float>(g);
        g = sample_generator(seed, pixel, pass);
        vec3<float> d{0.0f, 0.0f, 0.0f};
        const float a = random_real<float>(g);
        concentric_disk(a, random_real<float>(g), d.x, d.y);
        g = sample_generator(seed, pixel, pass);
        const vec3<float> c = cosine_about(vec3<float>{0.0f, 0.0f, 1.0f}, g);
        float* out = directions + 9 * k;
        out[0] = u.x;
        out[1] = u.y;
        out[2] = u.z;
        out[3] = d.x;
        out[4] = d.y;
        out[5] = d.z;
        out[6] = c.x;
        out[7] = c.y;
        out[8] = c.z;
    }
}

// Ray k's shadow question (src/gpu/shadow.cuh:26-120): is any primitive but skip[k] hit in (0.001, t_max[k]), with
// the self-intersection rules of the primitive from[k] it leaves? Answered by both shadow walks of the product: the
// wavefront's shading asks occluded (blocked), the quality megakernel walks its shadow rays as closest_hit in its
// any-hit mode (walked). Walk: the scene's tree walk (walk_instances for any float tree, walk_packed for packed nodes).
template <int Walk>
__global__ void __launch_bounds__(256) occluded_kernel(const scene_view<float> sc, const float* __restrict__ origins,
                                                       const float* __restrict__ directions,
                                                       const float* __restrict__ times,
                                                       const int32_t* __restrict__ from,
                                                       const uint8_t* __restrict__ inward,
                                                       const float* __restrict__ t_max,
                                                       const int32_t* __restrict__ skip, int64_t n,
                                                       uint8_t* __restrict__ blocked, uint8_t* __restrict__ walked)
{
    const lds_stack stack = block_stack(sc.stack_depth);
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n)
    {
        return;
    }
    const ray<float> r{{origins[3 * k], origins[3 * k + 1], origins[3 * k + 2]},
                       {directions[3 * k], directions[3 * k + 1], directions[3 * k + 2]},
                       times[k]};
    tally<false> tl;
    blocked[k] = occluded<Walk>(sc, r, t_max[k], from[k], inward[k] != 0, skip[k], stack, tl) ? 1 : 0;
    const hit<float> h = closest_hit<float, Walk, tally<false>, lds_stack, true>(sc, r, from[k], inward[k] != 0, stack,
                                                                                 tl, t_max[k], skip[k], true);
    walked[k] = h.index >= 0 ? 1 : 0;
}

// Case k of one lobe (rt_lobe) in the tangent frame (the normal +z): f cos and the density of the direction l[k] seen
// from v[k] (scatter_value), and a direction drawn from v[k] with the generator of (seed, k, 0) (scatter_sample): the
// direction, its weight f cos / density and its density, or zeros where the draw ends the path
__global__ void __launch_bounds__(256) scatter_kernel(int lobe, float alpha, float roughness, vec3<float> color,
                                                      int64_t n, const float* __restrict__ v,
                                                      const float* __restrict__ l, uint64_t seed,
                                                      float* __restrict__ value, float* __restrict__ density,
                                                      float* __restrict__ drawn, float* __restrict__ weight,
                                                      float* __restrict__ drawn_density)
{
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n)
    {
        return;
    }
    scatterer<float> s{};
    const bool books = lobe >= RT_LOBE_METAL;
    s.kind = lobe == RT_LOBE_LAMBERT     ? scatter_kind::lambert
             : lobe == RT_LOBE_PHASE     ? scatter_kind::phase
             : lobe == RT_LOBE_CONDUCTOR ? (alpha >= min_ggx_alpha ? scatter_kind::conductor : scatter_kind::mirror)
             : books                     ? scatter_kind::books
                                         : scatter_kind::coat;
    s.color = color;
    s.local = {{1.0f, 0.0f, 0.0f}, {0.0f, 1.0f, 0.0f}, {0.0f, 0.0f, 1.0f}};
    s.v = {v[3 * k], v[3 * k + 1], v[3 * k + 2]};
    s.alpha = alpha;
    s.roughness = roughness;
    vertex<float> vx{};
    vx.n = {0.0f, 0.0f, 1.0f}; // facing the viewer
    vx.outward = lobe == RT_LOBE_DIELECTRIC_LEAVING ? vec3<float>{0.0f, 0.0f, -1.0f} : vx.n;
    vx.front = lobe != RT_LOBE_DIELECTRIC_LEAVING;
    // the books' delta lobes read their material: a metal's albedo and fuzz, a dielectric's index
    material_rec<float> m{};
    m.kind = lobe == RT_LOBE_METAL ? RT_MATERIAL_METAL : RT_MATERIAL_DIELECTRIC;
    m.albedo[0] = color.x;
    m.albedo[1] = color.y;
    m.albedo[2] = color.z;
    m.fuzz = roughness;
    m.refraction_index = alpha;
    const ggx_eval_result<float> e = scatter_value(s, vx, vec3<float>{l[3 * k], l[3 * k + 1], l[3 * k + 2]});
    value[3 * k] = e.value.x;
    value[3 * k + 1] = e.value.y;
    value[3 * k + 2] = e.value.z;
    density[k] = e.pdf;
    pcg32 g = sample_generator(seed, static_cast<uint64_t>(k), 0);
    scatter_draw<float> d{};
    const bool ok = scatter_sample(s, m, -s.v, vx, g, d);
    drawn[3 * k] = ok ? d.wi.x : 0.0f;
    drawn[3 * k + 1] = ok ? d.wi.y : 0.0f;
    drawn[3 * k + 2] = ok ? d.wi.z : 0.0f;
    weight[3 * k] = ok ? d.weight.x : 0.0f;
    weight[3 * k + 1] = ok ? d.weight.y : 0.0f;
    weight[3 * k + 2] = ok ? d.weight.z : 0.0f;
    drawn_density[k] = ok ? d.pdf : 0.0f;
}

// Point k's light sample with the generator of (seed, k, 0) (draw_light): its direction, distance, radiance and density
// (the light's choice included), and the lamp primitive it aims at (-1: the map; -2: the draw failed, the outputs
// zero); hit_density: the density the light set gives that direction when a scattered ray finds the light there
// (lamp_hit_density for a lamp, the texel's density for the map, times the choice's chance)
__global__ void __launch_bounds__(256) light_kernel(const scene_view<float> sc, const float* __restrict__ points,
                                                    int64_t n, uint64_t seed, float* __restrict__ wi,
                                                    float* __restrict__ distance, float* __restrict__ radiance,
                                                    float* __restrict__ density, int32_t* __restrict__ skip,
                                                    float* __restrict__ hit_density)
{
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n)
    {
        return;
    }
    const vec3<float> x{points[3 * k], points[3 * k + 1], points[3 * k + 2]};
    pcg32 g = sample_generator(seed, static_cast<uint64_t>(k), 0);
    light_draw<float> d{};
    tally<false> tl;
    const bool ok = draw_light(sc.quality, x, g, d, tl);
    float hit = 0.0f;
    if (ok && d.skip >= 0)
    {
        hit = lamp_hit_density(sc, d.skip, x, d.wi, d.distance);
    }
    else if (ok)
    {
        float texel_density;
        texel_radiance(sc.quality, texel_index(sc.quality, d.wi), texel_density);
        hit = sc.quality.light_pmf * texel_density;
    }
    wi[3 * k] = ok ? d.wi.x : 0.0f;
    wi[3 * k + 1] = ok ? d.wi.y : 0.0f;
    wi[3 * k + 2] = ok ? d.wi.z : 0.0f;
    distance[k] = ok ? d.distance : 0.0f;
    radiance[3 * k] = ok ? d.radiance.x : 0.0f;
    radiance[3 * k + 1] = ok ? d.radiance.y : 0.0f;
    radiance[3 * k + 2] = ok ? d.radiance.z : 0.0f;
    density[k] = ok ? d.pdf : 0.0f;
    skip[k] = ok ? d.skip : -2;
    hit_density[k] = hit;
}

// Ray k against the scene's media up to t_surface[k] (the surfaces' closest hit): the media's event (media_events,
// with the generator of (seed, k, 0)): the medium's index among the scene's media and t, or -1 and infinity; and the
// transmittance of (0.001, t_surface[k]) (transmittance)
__global__ void __launch_bounds__(256) media_kernel(const scene_view<float> sc, const float* __restrict__ origins,
                                                    const float* __restrict__ directions,
                                                    const float* __restrict__ times,
                                                    const float* __restrict__ t_surface, int64_t n, uint64_t seed,
                                                    int32_t* __restrict__ medium, float* __restrict__ t,
                                                    float* __restrict__ transmitted)
{
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n)
    {
        return;
    }
    const ray<float> r{{origins[3 * k], origins[3 * k + 1], origins[3 * k + 2]},
                       {directions[3 * k], directions[3 * k + 1], directions[3 * k + 2]},
                       times[k]};
    pcg32 g = sample_generator(seed, static_cast<uint64_t>(k), 0);
    tally<false> tl;
    const hit<float> h = media_events(sc, r, hit<float>{-1, t_surface[k], false}, g, tl);
    medium[k] = h.index >= 0 ? h.index - sc.media_base : -1;
    t[k] = h.index >= 0 ? h.t : inf<float>;
    transmitted[k] = transmittance(sc, r, t_surface[k], tl);
}
}

namespace rt
{
hipError_t launch_media_probe(const dev::scene_view<float>& sc, const float* origins, const float* directions,
                              const float* times, const float* t_surface, int64_t n, uint64_t seed, int32_t* medium,
                              float* t, float* transmitted, hipStream_t stream)
{
    if (n <= 0)
    {
        return hipSuccess;
    }
    const int threads = 256;
    const dim3 grid(static_cast<unsigned>((n + threads - 1) / threads));
    hipLaunchKernelGGL(dev::media_kernel, grid, dim3(threads), 0, stream, sc, origins, directions, times, t_surface, n,
                       seed, medium, t, transmitted);
    return hipGetLastError();
}

hipError_t launch_light_probe(const dev::scene_view<float>& sc, const float* points, int64_t n, uint64_t seed,
                              float* wi, float* distance, float* radiance, float* density, int32_t* skip,
                              float* hit_density, hipStream_t stream)
{
    if (n <= 0)
    {
        return hipSuccess;
    }
    const int threads = 256;
    const dim3 grid(static_cast<unsigned>((n + threads - 1) / threads));
    hipLaunchKernelGGL(dev::light_kernel, grid, dim3(threads), 0, stream, sc, points, n, seed, wi, distance, radiance,
                       density, skip, hit_density);
    return hipGetLastError();
}

hipError_t launch_scatter_probe(int lobe, float alpha, float roughness, const float color[3], int64_t n,
                                const float* v, const float* l, uint64_t seed, float* value, float* density,
                                float* drawn, float* weight, float* drawn_density, hipStream_t stream)
{
    if (n <= 0)
    {
        return hipSuccess;
    }
    const int threads = 256;
    const dim3 grid(static_cast<unsigned>((n + threads - 1) / threads));
    hipLaunchKernelGGL(dev::scatter_kernel, grid, dim3(threads), 0, stream, lobe, alpha, roughness,
                       dev::vec3<float>{color[0], color[1], color[2]}, n, v, l, seed, value, density, drawn, weight,
                       drawn_density);
    return hipGetLastError();
}

hipError_t launch_occluded(const dev::scene_view<float>& sc, const float* origins, const float* directions,
                           const float* times, const int32_t* from, const uint8_t* inward, const float* t_max,
                           const int32_t* skip, int64_t n, uint8_t* blocked, uint8_t* walked, hipStream_t stream)
{
    if (n <= 0)
    {
        return hipSuccess;
    }
    const int threads = 256;
    const dim3 grid(static_cast<unsigned>((n + threads - 1) / threads));
    const size_t lds = stack_lds_bytes<float>(sc.stack_depth, threads);
    if (dev::tree_walk(sc) == dev::walk_packed)
    {
        hipLaunchKernelGGL((dev::occluded_kernel<dev::walk_packed>), grid, dim3(threads), lds, stream, sc, origins,
                           directions, times, from, inward, t_max, skip, n, blocked, walked);
    }
    else
    {
        hipLaunchKernelGGL((dev::occluded_kernel<dev::walk_instances>), grid, dim3(threads), lds, stream, sc, origins,
                           directions, times, from, inward, t_max, skip, n, blocked, walked);
    }
    return hipGetLastError();
}

hipError_t launch_reference_walk(const dev::scene_view<float>& sc, const void* binary_nodes, int node_count,
                                 const float* origins, const float* directions, const float* times,
                                 const int32_t* from, const uint8_t* inward, int64_t n, int32_t* index, float* t,
                                 uint8_t* entering, int32_t* overflow, hipStream_t stream)
{
    if (n <= 0)
    {
        return hipSuccess;
    }
    const int threads = 256;
    const dim3 grid(static_cast<unsigned>((n + threads - 1) / threads));
    hipLaunchKernelGGL(dev::reference_walk_kernel, grid, dim3(threads), 0, stream, sc,
                       static_cast<const dev::binary_node*>(binary_nodes), node_count, origins, directions, times,
                       from, inward, n, index, t, entering, overflow);
    return hipGetLastError();
}

hipError_t launch_sample_numbers(uint64_t seed, const uint64_t* pixels, const uint64_t* passes, int64_t n, int draws,
                                 uint32_t* words, float* floats, double* doubles, float* directions,
                                 hipStream_t stream)
{
    if (n <= 0)
    {
        return hipSuccess;
    }
    const int threads = 256;
    const dim3 grid(static_cast<unsigned>((n + threads - 1) / threads));
    hipLaunchKernelGGL(dev::sample_numbers_kernel, grid, dim3(threads), 0, stream, seed, pixels, passes, n, draws,
                       words, floats, doubles, directions);
    return hipGetLastError();
}

size_t binary_node_bytes() { return sizeof(dev::binary_node); }
}
