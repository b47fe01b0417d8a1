// The wavefront integrator's kernels (float). A launch's work, passes [first_pass, first_pass + pass_count) of rows
// [first_row, first_row + row_count), is a pool of path slots, one per pixel. A slot traces its pixel's passes one
// after the other: when its path ends, the sample is added to the pixel's double sums and the slot starts the pixel's
// next pass in place. So every pixel has one path in flight and adds its samples in pass order, as the megakernel does,
// and every path draws the same random numbers in the same order: the two integrators' images are the same bytes.
//
// One bounce of every active slot is three kernels:
//   reset:  clears the counters the bounce fills;
//   extend: the closest event of each active slot's ray (BVH walk with an LDS stack, then the media), written to the
//           slot, and the slot appended to the queue of its hit's material class (wave-aggregated atomics: one atomic
//           per class per wave);
//   shade:  the queues in class order (miss, lamp, matte, textured, marble, smoke, metal, glass), so that a wave
//           shades one material at a time: the bounce (or the end of the path, its sample, and the pixel's next pass),
//           and the slot appended to the next bounce's active list.
// extend and shade are persistent: a fixed grid whose waves fetch 64 items at a time from a counter, until the count
// (read from device memory, so the loop needs no host synchronisation and can be captured in a hipGraph) runs out.
#include "kernel_api.hpp"
#include "path_tracer.hpp"
#include "quality.hpp"
#include "wavefront_api.hpp"

namespace rt::dev
{
namespace
{
__device__ inline unsigned long long wsum(unsigned long long v)
{
    for (int offset = 32; offset > 0; offset >>= 1)
    {
        v += __shfl_down(v, offset, 64);
    }
    return v;
}

__device__ inline int lane_id() { return static_cast<int>(threadIdx.x & 63); }

// A slot's path in the SoA pool
struct slot_path
{
    path<float> p;
    pcg32 g;
    bool specular = false; // quality integrator: the last vertex was a mirror's (or the camera)
};

__device__ inline slot_path load_path(const wf_pool& pool, uint32_t s)
{
    const float4 o = pool.ray_o[s];
    const float4 d = pool.ray_d[s];
    const float4 t = pool.thr[s];
    const uint4 r = pool.rng[s];
    slot_path sp;
    sp.p.r.origin = {o.x, o.y, o.z};
    sp.p.r.time = o.w;
    sp.p.r.direction = {d.x, d.y, d.z};
    sp.p.from = __float_as_int(d.w);
    sp.p.throughput = {t.x, t.y, t.z};
    const uint32_t packed = __float_as_uint(t.w);
    sp.p.rays_left = static_cast<int>(packed & 0x3fffffffu);
    sp.p.inward = (packed >> 30) & 1u;
    sp.specular = (packed >> 31) != 0;
    sp.g.state = (static_cast<uint64_t>(r.y) << 32) | r.x;
    sp.g.inc = (static_cast<uint64_t>(r.w) << 32) | r.z;
    return sp;
}

__device__ inline void store_rng(const wf_pool& pool, uint32_t s, const pcg32& g)
{
    pool.rng[s] = make_uint4(static_cast<uint32_t>(g.state), static_cast<uint32_t>(g.state >> 32),
                             static_cast<uint32_t>(g.inc), static_cast<uint32_t>(g.inc >> 32));
}

__device__ inline void store_path(const wf_pool& pool, uint32_t s, const slot_path& sp)
{
    pool.ray_o[s] = make_float4(sp.p.r.origin.x, sp.p.r.origin.y, sp.p.r.origin.z, sp.p.r.time);
    pool.ray_d[s] = make_float4(sp.p.r.direction.x, sp.p.r.direction.y, sp.p.r.direction.z, __int_as_float(sp.p.from));
    const uint32_t packed = static_cast<uint32_t>(sp.p.rays_left) | (sp.p.inward ? 1u << 30 : 0u) |
                            (sp.specular ? 1u << 31 : 0u);
    pool.thr[s] = make_float4(sp.p.throughput.x, sp.p.throughput.y, sp.p.throughput.z, __uint_as_float(packed));
    store_rng(pool, s, sp.g);
}

// Appends `want` (this lane's vote) to a list with one atomic per wave; returns this lane's position
__device__ inline uint32_t wave_append(uint32_t* counter, bool want)
{
    const uint64_t mask = __ballot(want);
    if (mask == 0)
    {
        return 0;
    }
    const int leader = __ffsll(static_cast<unsigned long long>(mask)) - 1;
    uint32_t base = 0;
    if (lane_id() == leader)
    {
        base = atomicAdd(counter, static_cast<uint32_t>(__popcll(mask)));
    }
    base = __shfl(base, leader, 64);
    const uint64_t below = mask & ((uint64_t{1} << lane_id()) - 1);
    return base + static_cast<uint32_t>(__popcll(below));
}

// The next 64 items for this wave, or false when the count is exhausted
__device__ inline bool wave_fetch(uint32_t* work, uint32_t count, uint32_t& base)
{
    uint32_t b = 0;
    if (lane_id() == 0)
    {
        b = atomicAdd(work, 64u);
    }
    base = __shfl(b, 0, 64);
    return base < count;
}

__device__ inline lds_stack<float> thread_stack(int depth)
{
    extern __shared__ __align__(16) unsigned char lds[];
    const int threads = blockDim.x;
    int32_t* refs = reinterpret_cast<int32_t*>(lds);
    float* ts = reinterpret_cast<float*>(lds + ((static_cast<size_t>(depth) * threads * 4 + 15) & ~size_t(15)));
    return {refs + threadIdx.x, ts + threadIdx.x, threads};
}

// The shading class of a hit: which queue it waits in
__device__ inline int hit_class(const scene_view<float>& sc, int index)
{
    if (index < 0)
    {
        return wf_class_miss;
    }
    const int m = index >= sc.media_base ? sc.media[index - sc.media_base].material : sc.prim_materials[index];
    if (sc.quality.microfacets)
    {
        const int side = sc.quality.microfacets[m].kind;
        if (side == RT_MICROFACET_CONDUCTOR)
        {
            return wf_class_ggx;
        }
        if (side == RT_MICROFACET_GLOSS)
        {
            return wf_class_gloss;
        }
    }
    const material_rec<float>& mat = sc.materials[m];
    switch (mat.kind)
    {
    case RT_MATERIAL_DIFFUSE_LIGHT:
        return wf_class_light;
    case RT_MATERIAL_METAL:
        return wf_class_metal;
    case RT_MATERIAL_DIELECTRIC:
        return wf_class_glass;
    case RT_MATERIAL_ISOTROPIC:
        return wf_class_smoke;
    default:
        if (mat.texture < 0)
        {
            return wf_class_matte;
        }
        return sc.textures[mat.texture].kind == RT_TEXTURE_NOISE ? wf_class_marble : wf_class_textured;
    }
}

// The first path of a slot: its pixel's pass `pass`. A quality path starts as the camera's: specular, no light yet.
__device__ inline slot_path start_slot(const camera_rec<float>& cam, const scene_view<float>& sc, uint64_t seed,
                                       uint64_t pixel, int i, int j, int pass, const wf_pool& pool, uint32_t s)
{
    slot_path sp;
    sp.g = sample_generator(seed, pixel, static_cast<uint64_t>(pass));
    sp.p = start_path(cam, sc.motion != 0, i, j, pass, sp.g);
    if (pool.qst)
    {
        sp.specular = true;
        pool.qst[s] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    return sp;
}
}

__global__ void __launch_bounds__(256) wf_init(const camera_rec<float> cam, const scene_view<float> sc,
                                               const uint64_t seed, const launch_work work, wf_pool pool)
{
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t slots = static_cast<uint32_t>(cam.image_width) * static_cast<uint32_t>(work.row_count);
    if (s == 0)
    {
        pool.counters[0] = slots; // the active list of parity 0 holds every slot
    }
    if (s >= slots)
    {
        return;
    }
    const int i = static_cast<int>(s % static_cast<uint32_t>(cam.image_width));
    const int j = work.first_row + static_cast<int>(s / static_cast<uint32_t>(cam.image_width));
    const uint64_t pixel = static_cast<uint64_t>(j) * static_cast<uint64_t>(cam.image_width) + static_cast<uint64_t>(i);
    const slot_path sp = start_slot(cam, sc, seed, pixel, i, j, work.first_pass, pool, s);
    store_path(pool, s, sp);
    pool.pass[s] = work.first_pass;
    pool.active[0][s] = s;
}

// Clears what one bounce of parity p fills: the next active list's count, the class counts and the work counters
__global__ void wf_reset(wf_pool pool, int p)
{
    const int k = threadIdx.x;
    uint32_t* in = pool.counters + wf_counter_stride * p;
    uint32_t* out = pool.counters + wf_counter_stride * (1 - p);
    if (k == 0)
    {
        out[0] = 0;
    }
    if (k >= 1 && k < wf_counter_stride)
    {
        in[k] = 0;
    }
}

__global__ void __launch_bounds__(256) wf_extend(const scene_view<float> sc, wf_pool pool, int p,
                                                 unsigned long long* ray_count)
{
    const lds_stack<float> stack = thread_stack(sc.stack_depth);
    uint32_t* in = pool.counters + wf_counter_stride * p;
    const uint32_t n = in[0];
    const uint32_t* active = pool.active[p];
    unsigned long long rays = 0;
    tally<false> tl;
    uint32_t base;
    while (wave_fetch(in + wf_counter_extend_work, n, base))
    {
        const uint32_t k = base + static_cast<uint32_t>(lane_id());
        const bool valid = k < n;
        int cls = -1;
        uint32_t s = 0;
        if (valid)
        {
            s = active[k];
            slot_path sp = load_path(pool, s);
            const hit<float> h = closest_event(sc, sp.p.r, sp.p.from, sp.p.inward, sp.g, stack, tl);
            pool.hit[s] = make_int4(h.index, __float_as_int(h.t), h.entering ? 1 : 0, 0);
            if (sc.medium_count > 0)
            {
                store_rng(pool, s, sp.g); // the media drew their free flights from the path's generator
            }
            cls = hit_class(sc, h.index);
            ++rays;
        }
        // the slot waits in its class's queue: one atomic per class present in the wave
        for (int c = 0; c < wf_classes; ++c)
        {
            const bool mine = cls == c;
            if (__ballot(mine) == 0)
            {
                continue;
            }
            const uint32_t at = wave_append(in + wf_counter_class + c, mine);
            if (mine)
            {
                pool.queue[static_cast<size_t>(c) * pool.capacity + at] = s;
            }
        }
    }
    rays = wsum(rays);
    if (lane_id() == 0 && rays)
    {
        atomicAdd(ray_count, rays);
    }
}

__global__ void __launch_bounds__(256) wf_shade(const camera_rec<float> cam, const scene_view<float> sc,
                                                const uint64_t seed, const launch_work work, wf_pool pool, int p,
                                                double* __restrict__ sums)
{
    uint32_t* in = pool.counters + wf_counter_stride * p;
    uint32_t* out = pool.counters + wf_counter_stride * (1 - p);
    uint32_t offsets[wf_classes + 1];
    offsets[0] = 0;
    for (int c = 0; c < wf_classes; ++c)
    {
        offsets[c + 1] = offsets[c] + in[wf_counter_class + c];
    }
    const uint32_t total = offsets[wf_classes];
    const int end_pass = work.first_pass + work.pass_count;
    tally<false> tl;
    uint32_t base;
    while (wave_fetch(in + wf_counter_shade_work, total, base))
    {
        const uint32_t g = base + static_cast<uint32_t>(lane_id());
        bool keep = false;
        uint32_t s = 0;
        if (g < total)
        {
            int c = 0;
            while (g >= offsets[c + 1])
            {
                ++c;
            }
            s = pool.queue[static_cast<size_t>(c) * pool.capacity + (g - offsets[c])];
            slot_path sp = load_path(pool, s);
            const int4 hv = pool.hit[s];
            const hit<float> h{hv.x, __int_as_float(hv.y), hv.z != 0};
            vec3<float> sample;
            bool done;
            if (h.index < 0)
            {
                sample = sp.p.throughput * sky(cam, sp.p.r.direction);
                done = true;
            }
            else
            {
                done = shade(sp.p, sc, h, sp.g, sample, tl);
            }
            if (!done)
            {
                store_path(pool, s, sp);
                keep = true;
            }
            else
            {
                const int i = static_cast<int>(s % static_cast<uint32_t>(cam.image_width));
                const int j = work.first_row + static_cast<int>(s / static_cast<uint32_t>(cam.image_width));
                const uint64_t pixel =
                    static_cast<uint64_t>(j) * static_cast<uint64_t>(cam.image_width) + static_cast<uint64_t>(i);
                double* const sum = sums + 3 * pixel;
                sum[0] += without_nan(sample.x);
                sum[1] += without_nan(sample.y);
                sum[2] += without_nan(sample.z);
                const int next = pool.pass[s] + 1;
                if (next < end_pass)
                {
                    pool.pass[s] = next;
                    store_path(pool, s, start_slot(cam, sc, seed, pixel, i, j, next, pool, s));
                    keep = true;
                }
            }
        }
        const uint32_t at = wave_append(out, keep);
        if (keep)
        {
            pool.active[1 - p][at] = s;
        }
    }
}

// The drain: when few slots are left active, each of them (a thread each) traces its current path and its pixel's
// remaining passes of the launch to the end, one after the other, as the megakernel does: the bounces with a nearly
// empty pool would cost a launch each for little work. The sums stay in pass order, so the image is the same bytes.
template <int Mode>
__global__ void __launch_bounds__(128) wf_drain(const camera_rec<float> cam, const scene_view<float> sc,
                                                const uint64_t seed, const launch_work work, wf_pool pool, int p,
                                                double* __restrict__ sums, unsigned long long* __restrict__ ray_count)
{
    const lds_stack<float> stack = thread_stack(sc.stack_depth);
    const uint32_t n = pool.counters[wf_counter_stride * p];
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long rays = 0;
    uint32_t shadow_rays = 0;
    if (k < n)
    {
        const uint32_t s = pool.active[p][k];
        slot_path sp = load_path(pool, s);
        const int i = static_cast<int>(s % static_cast<uint32_t>(cam.image_width));
        const int j = work.first_row + static_cast<int>(s / static_cast<uint32_t>(cam.image_width));
        const uint64_t pixel = static_cast<uint64_t>(j) * static_cast<uint64_t>(cam.image_width) + static_cast<uint64_t>(i);
        double* const sum = sums + 3 * pixel;
        double red = sum[0], green = sum[1], blue = sum[2];
        const int end_pass = work.first_pass + work.pass_count;
        tally<false> tl;
        quality_path q;
        if constexpr (Mode == mode_quality)
        {
            const float4 st = pool.qst[s];
            q.base = sp.p;
            q.radiance = {st.x, st.y, st.z};
            q.bsdf_pdf = st.w;
            q.specular = sp.specular;
        }
        int pass = pool.pass[s];
        while (true)
        {
            ++rays;
            vec3<float> sample;
            bool done;
            if constexpr (Mode == mode_quality)
            {
                done = trace_quality(q, sc, cam, sp.g, sample, stack, tl, shadow_rays);
            }
            else if constexpr (Mode == mode_all_books)
            {
                done = trace_all_books(sp.p, sc, cam, sp.g, sample, stack, tl);
            }
            else
            {
                done = trace_first_book(sp.p, sc, cam, sp.g, sample, stack, tl);
            }
            if (!done)
            {
                continue;
            }
            red += without_nan(sample.x);
            green += without_nan(sample.y);
            blue += without_nan(sample.z);
            if (++pass == end_pass)
            {
                break;
            }
            sp.g = sample_generator(seed, pixel, static_cast<uint64_t>(pass));
            if constexpr (Mode == mode_quality)
            {
                q = start_quality(cam, sc.motion != 0, i, j, pass, sp.g);
            }
            else
            {
                sp.p = start_path(cam, sc.motion != 0, i, j, pass, sp.g);
            }
        }
        sum[0] = red;
        sum[1] = green;
        sum[2] = blue;
    }
    rays = wsum(rays);
    const unsigned long long shadows = wsum(shadow_rays);
    if (lane_id() == 0)
    {
        if (rays)
        {
            atomicAdd(ray_count, rays);
        }
        if (shadows)
        {
            atomicAdd(ray_count + 1, shadows);
        }
    }
}

// The quality integrator's shade: the same queues in class order, each vertex's light sample and its shadow ray
// traced here (an any-hit walk with this thread's LDS stack), in the megakernel's order of random numbers
__global__ void __launch_bounds__(256) wf_shade_quality(const camera_rec<float> cam, const scene_view<float> sc,
                                                        const uint64_t seed, const launch_work work, wf_pool pool,
                                                        int p, double* __restrict__ sums,
                                                        unsigned long long* __restrict__ ray_count)
{
    const lds_stack<float> stack = thread_stack(sc.stack_depth);
    uint32_t* in = pool.counters + wf_counter_stride * p;
    uint32_t* out = pool.counters + wf_counter_stride * (1 - p);
    uint32_t offsets[wf_classes + 1];
    offsets[0] = 0;
    for (int c = 0; c < wf_classes; ++c)
    {
        offsets[c + 1] = offsets[c] + in[wf_counter_class + c];
    }
    const uint32_t total = offsets[wf_classes];
    const int end_pass = work.first_pass + work.pass_count;
    tally<false> tl;
    uint32_t shadow_rays = 0;
    uint32_t base;
    while (wave_fetch(in + wf_counter_shade_work, total, base))
    {
        const uint32_t g = base + static_cast<uint32_t>(lane_id());
        bool keep = false;
        uint32_t s = 0;
        if (g < total)
        {
            int c = 0;
            while (g >= offsets[c + 1])
            {
                ++c;
            }
            s = pool.queue[static_cast<size_t>(c) * pool.capacity + (g - offsets[c])];
            slot_path sp = load_path(pool, s);
            const float4 st = pool.qst[s];
            quality_path q;
            q.base = sp.p;
            q.radiance = {st.x, st.y, st.z};
            q.bsdf_pdf = st.w;
            q.specular = sp.specular;
            const int4 hv = pool.hit[s];
            const hit<float> h{hv.x, __int_as_float(hv.y), hv.z != 0};
            vec3<float> sample;
            if (!shade_quality(q, sc, cam, h, sp.g, sample, stack, tl, shadow_rays))
            {
                sp.p = q.base;
                sp.specular = q.specular;
                pool.qst[s] = make_float4(q.radiance.x, q.radiance.y, q.radiance.z, q.bsdf_pdf);
                store_path(pool, s, sp);
                keep = true;
            }
            else
            {
                const int i = static_cast<int>(s % static_cast<uint32_t>(cam.image_width));
                const int j = work.first_row + static_cast<int>(s / static_cast<uint32_t>(cam.image_width));
                const uint64_t pixel =
                    static_cast<uint64_t>(j) * static_cast<uint64_t>(cam.image_width) + static_cast<uint64_t>(i);
                double* const sum = sums + 3 * pixel;
                sum[0] += without_nan(sample.x);
                sum[1] += without_nan(sample.y);
                sum[2] += without_nan(sample.z);
                const int next = pool.pass[s] + 1;
                if (next < end_pass)
                {
                    pool.pass[s] = next;
                    store_path(pool, s, start_slot(cam, sc, seed, pixel, i, j, next, pool, s));
                    keep = true;
                }
            }
        }
        const uint32_t at = wave_append(out, keep);
        if (keep)
        {
            pool.active[1 - p][at] = s;
        }
    }
    const unsigned long long shadows = wsum(shadow_rays);
    if (lane_id() == 0 && shadows)
    {
        atomicAdd(ray_count + 1, shadows);
    }
}
}

namespace rt
{
hipError_t wf_launch_init(const dev::camera_rec<float>& cam, const dev::scene_view<float>& sc, uint64_t seed,
                          const dev::launch_work& work, const dev::wf_pool& pool, hipStream_t stream)
{
    const uint32_t slots = static_cast<uint32_t>(cam.image_width) * static_cast<uint32_t>(work.row_count);
    hipLaunchKernelGGL(dev::wf_init, dim3((slots + 255) / 256), dim3(256), 0, stream, cam, sc, seed, work, pool);
    return hipGetLastError();
}

hipError_t wf_launch_bounce(const dev::camera_rec<float>& cam, const dev::scene_view<float>& sc, uint64_t seed,
                            const dev::launch_work& work, const dev::wf_pool& pool, int parity, double* sums,
                            unsigned long long* rays, int extend_blocks, int shade_blocks, bool quality,
                            hipStream_t stream)
{
    hipLaunchKernelGGL(dev::wf_reset, dim3(1), dim3(64), 0, stream, pool, parity);
    const size_t lds = stack_lds_bytes<float>(sc.stack_depth, 256);
    hipLaunchKernelGGL(dev::wf_extend, dim3(extend_blocks), dim3(256), lds, stream, sc, pool, parity, rays);
    if (quality)
    {
        hipLaunchKernelGGL(dev::wf_shade_quality, dim3(shade_blocks), dim3(256), lds, stream, cam, sc, seed, work,
                           pool, parity, sums, rays);
    }
    else
    {
        hipLaunchKernelGGL(dev::wf_shade, dim3(shade_blocks), dim3(256), 0, stream, cam, sc, seed, work, pool,
                           parity, sums);
    }
    return hipGetLastError();
}

hipError_t wf_launch_drain(int mode, const dev::camera_rec<float>& cam, const dev::scene_view<float>& sc, uint64_t seed,
                           const dev::launch_work& work, const dev::wf_pool& pool, uint32_t count, double* sums,
                           unsigned long long* rays, hipStream_t stream)
{
    const int threads = 128;
    const dim3 grid((count + threads - 1) / threads);
    const size_t lds = stack_lds_bytes<float>(sc.stack_depth, threads);
    if (mode == dev::mode_quality)
    {
        hipLaunchKernelGGL((dev::wf_drain<dev::mode_quality>), grid, dim3(threads), lds, stream, cam, sc, seed, work,
                           pool, 0, sums, rays);
    }
    else if (mode == dev::mode_all_books)
    {
        hipLaunchKernelGGL((dev::wf_drain<dev::mode_all_books>), grid, dim3(threads), lds, stream, cam, sc, seed, work,
                           pool, 0, sums, rays);
    }
    else
    {
        hipLaunchKernelGGL((dev::wf_drain<dev::mode_first_book>), grid, dim3(threads), lds, stream, cam, sc, seed,
                           work, pool, 0, sums, rays);
    }
    return hipGetLastError();
}

kernel_attributes wf_attributes(int depth, int which)
{
    hipFuncAttributes a{};
    const void* f = which == 0   ? reinterpret_cast<const void*>(&dev::wf_extend)
                    : which == 1 ? reinterpret_cast<const void*>(&dev::wf_shade)
                                 : reinterpret_cast<const void*>(&dev::wf_shade_quality);
    kernel_attributes k{};
    if (hipFuncGetAttributes(&a, f) == hipSuccess)
    {
        k.registers = a.numRegs;
        k.local_bytes = static_cast<int>(a.localSizeBytes);
    }
    k.lds_bytes = which != 1 ? static_cast<int>(stack_lds_bytes<float>(depth, 256)) : 0;
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, f, 256, k.lds_bytes) == hipSuccess)
    {
        k.blocks_per_cu = blocks;
    }
    return k;
}
}
