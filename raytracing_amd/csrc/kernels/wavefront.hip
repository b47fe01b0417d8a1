// The wavefront integrator's kernels (float). A launch's work is a list of work items, pass-major over the launch's
// pixels (item k: pass first_pass + k / pixels of pixel k % pixels), rendered by a pool of path slots: a slot whose
// path ends writes the path's sample to the item's place and takes the next item, so the pool stays full until the
// items run out, whatever the paths' lengths. At the launch's end each pixel's samples are added to its double sums in
// pass order, as the megakernel adds them, and every path draws the same random numbers: the two integrators' images
// are the same bytes.
//
// One bounce of every active slot is five kernels:
//   begin:  the free slots get the next items (as many as there are, or items left), the active list's prefix sums
//           (its stripes, then the refilled slots), and the counters the bounce fills cleared (one block);
//   refill: each refilled slot starts its item's path;
//   extend: the closest event of each active slot's ray (BVH walk with an LDS stack, then the media), written to the
//           slot, and the slot appended to the queue of its hit's material class;
//   sort:   the queues' prefix sums, class by class (one block);
//   shade:  the queues in class order (miss, lamp, matte, textured, marble, smoke, metal, glass, GGX, gloss), so that
//           a wave shades one material at a time: the bounce, or the end of the path (its sample written, the slot
//           appended to the free list); a path that goes on is appended to the next bounce's active list.
// Every thread of refill, extend and shade takes one item by its index (no work counter); appends are counted per
// wave with ballots, per block in LDS, then one device-scope atomic per block and list on a stripe of the list
// (wf_stripes). The lengths live in device memory, so a bounce needs no host synchronisation and a run of bounces is
// captured in a hipGraph. When the items are handed out and few slots are left, the drain kernel finishes them a
// thread each.
#include "kernel_api.hpp"
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
__device__ inline int stripe_of_block() { return static_cast<int>(blockIdx.x % wf_stripes); }

// An exclusive prefix sum of n counters into out[0..n] (out[n]: the total) by one wave; other waves return
__device__ inline void wave_prefix(const uint32_t* in, uint32_t* out, int n)
{
    if (threadIdx.x >= 64)
    {
        return;
    }
    const int lane = lane_id();
    const int per = (n + 63) / 64;
    const int first = lane * per;
    uint32_t mine = 0;
    for (int k = first; k < first + per && k < n; ++k)
    {
        mine += in[k];
    }
    uint32_t scan = mine; // inclusive scan across the wave
    for (int offset = 1; offset < 64; offset <<= 1)
    {
        const uint32_t up = __shfl_up(scan, offset, 64);
        if (lane >= offset)
        {
            scan += up;
        }
    }
    uint32_t run = scan - mine;
    for (int k = first; k < first + per && k < n; ++k)
    {
        out[k] = run;
        run += in[k];
    }
    if (lane == 63)
    {
        out[n] = scan;
    }
}

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

// The segment of item g in a list of `segments` whose prefix sums (segments + 1 of them, the last the length) are at
// `prefix`: the largest k with prefix[k] <= g (empty segments are skipped). Requires g < prefix[segments].
__device__ inline int find_segment(const uint32_t* prefix, int segments, uint32_t g)
{
    int lo = 0, hi = segments - 1;
    while (lo < hi)
    {
        const int mid = (lo + hi + 1) >> 1;
        if (prefix[mid] <= g)
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

// Appends each thread's item to list `list` (none for list < 0) of `Lists` striped lists: counted per wave with a
// ballot, per block in LDS, then one device-scope atomic per list present in the block, on this block's stripe. Every
// thread of the block must call it (it has barriers).
template <int Lists>
__device__ inline void block_append(int list, uint32_t item, uint32_t* counters, uint32_t* storage, uint32_t segment)
{
    __shared__ uint32_t count[Lists];
    __shared__ uint32_t base[Lists];
    const int t = static_cast<int>(threadIdx.x);
    if (t < Lists)
    {
        count[t] = 0;
    }
    __syncthreads();
    uint32_t at = 0;
    for (int c = 0; c < Lists; ++c)
    {
        const uint64_t mask = __ballot(list == c);
        if (mask == 0)
        {
            continue;
        }
        const int leader = __ffsll(static_cast<unsigned long long>(mask)) - 1;
        uint32_t wave_base = 0;
        if (lane_id() == leader)
        {
            wave_base = atomicAdd(&count[c], static_cast<uint32_t>(__popcll(mask)));
        }
        wave_base = __shfl(wave_base, leader, 64);
        if (list == c)
        {
            at = wave_base + static_cast<uint32_t>(__popcll(mask & ((uint64_t{1} << lane_id()) - 1)));
        }
    }
    __syncthreads();
    const int stripe = stripe_of_block();
    if (t < Lists && count[t] > 0)
    {
        base[t] = atomicAdd(&counters[t * wf_stripes + stripe], count[t]);
    }
    __syncthreads();
    if (list >= 0)
    {
        storage[static_cast<size_t>(list * wf_stripes + stripe) * segment + base[list] + at] = item;
    }
}

// Adds a count of every thread of the block to one of the launch's stripe tallies: one atomic per block
__device__ inline void block_tally(unsigned long long* tallies, int which, unsigned long long v)
{
    __shared__ unsigned long long total;
    if (threadIdx.x == 0)
    {
        total = 0;
    }
    __syncthreads();
    v = wsum(v);
    if (lane_id() == 0 && v)
    {
        atomicAdd(&total, v);
    }
    __syncthreads();
    if (threadIdx.x == 0 && total)
    {
        atomicAdd(&tallies[2 * stripe_of_block() + which], total);
    }
}

// Loads `n` prefix sums into LDS (every thread of the block calls it)
__device__ inline void load_prefix(uint32_t* lds, const uint32_t* prefix, int n)
{
    for (int k = static_cast<int>(threadIdx.x); k < n; k += static_cast<int>(blockDim.x))
    {
        lds[k] = prefix[k];
    }
    __syncthreads();
}

__device__ inline lds_stack thread_stack(int entries)
{
    extern __shared__ __align__(16) unsigned char lds[];
    return stack_in(lds, entries, blockDim.x, threadIdx.x);
}

// The shading class of a hit: which queue it waits in
__device__ inline int hit_class(const scene_view<float>& sc, int index)
{
    if (index < 0)
    {
        return wf_class_miss;
    }
    const int m = index >= sc.media_base ? sc.media[index - sc.media_base].material : material_of(sc, index);
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

// A slot's new path: work item `item` of the launch (pass first_pass + item / pixels, pixel item % pixels of the
// launch's rows). A quality path starts as the camera's: specular, no light yet.
__device__ inline slot_path start_item(const camera_rec<float>& cam, const scene_view<float>& sc, uint64_t seed,
                                       const launch_work& work, const wf_pool& pool, uint32_t s, uint32_t item)
{
    const uint32_t width = static_cast<uint32_t>(cam.image_width);
    const uint32_t pixels = width * static_cast<uint32_t>(work.row_count);
    const uint32_t local = item % pixels;
    const int pass = work.first_pass + static_cast<int>(item / pixels);
    const int i = static_cast<int>(local % width);
    const int j = work.first_row + static_cast<int>(local / width);
    const uint64_t pixel = static_cast<uint64_t>(j) * width + static_cast<uint64_t>(i);
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

// The slot behind entry g of the active list of parity p (its segments' prefix sums in LDS)
__device__ inline uint32_t active_slot(const wf_pool& pool, int p, const uint32_t* prefix, uint32_t g)
{
    const int segment = find_segment(prefix, wf_stripes + 1, g);
    return pool.active[p][static_cast<size_t>(segment) * pool.segment + (g - prefix[segment])];
}
}

// A launch's start: its items; every slot on the free list (block b fills its stripe's segment at
// (b / wf_stripes) * wf_block, so the list is written without atomics and its stripes' lengths have a closed form);
// no slot active
__global__ void __launch_bounds__(256) wf_init(wf_pool pool, uint32_t items)
{
    const uint32_t slots = pool.capacity;
    const uint32_t blocks = (slots + wf_block - 1) / wf_block;
    if (blockIdx.x == 0 && threadIdx.x < wf_stripes)
    {
        const uint32_t stripe = threadIdx.x;
        uint32_t length = 0;
        if (stripe < blocks)
        {
            const uint32_t last = blocks - 1;
            const uint32_t mine = (last - stripe) / wf_stripes + 1; // blocks stripe, stripe + S, ... <= last
            length = mine * wf_block;
            if (last % wf_stripes == stripe)
            {
                length -= blocks * wf_block - slots; // the last block is short
            }
        }
        pool.counters[wf_counter_free + stripe] = length;
        pool.counters[wf_counter_active + stripe] = 0; // parity 0: nothing active yet
        if (stripe == 0)
        {
            pool.counters[wf_counter_next] = 0;
            pool.counters[wf_counter_items] = items;
        }
    }
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < slots)
    {
        const uint32_t stripe = blockIdx.x % wf_stripes;
        pool.free[static_cast<size_t>(stripe) * pool.segment + (blockIdx.x / wf_stripes) * wf_block + threadIdx.x] = s;
    }
}

// The start of a bounce of parity p (one block): the free slots get the next items, as many as there are of either;
// the active list's prefix sums (its stripes, then the refilled slots); the queues of parity p, the next active list
// and the free list cleared for the bounce to fill
__global__ void __launch_bounds__(256) wf_begin(wf_pool pool, int p)
{
    uint32_t* in = pool.counters + wf_counter_stride * p;
    uint32_t* out = pool.counters + wf_counter_stride * (1 - p);
    __shared__ uint32_t active_total;
    wave_prefix(pool.counters + wf_counter_free, pool.prefix + wf_prefix_free, wf_stripes);
    wave_prefix(in + wf_counter_active, pool.prefix + wf_prefix_active, wf_stripes);
    __syncthreads();
    if (threadIdx.x == 0)
    {
        const uint32_t free = pool.prefix[wf_prefix_free + wf_stripes];
        const uint32_t next = pool.counters[wf_counter_next];
        const uint32_t left = pool.counters[wf_counter_items] - next;
        const uint32_t refill = free < left ? free : left;
        pool.counters[wf_counter_refill] = refill;
        pool.counters[wf_counter_refill_base] = next;
        pool.counters[wf_counter_next] = next + refill;
        active_total = pool.prefix[wf_prefix_active + wf_stripes];
        pool.prefix[wf_prefix_active + wf_stripes + 1] = active_total + refill;
    }
    __syncthreads();
    const int t = static_cast<int>(threadIdx.x);
    for (int k = t; k < wf_classes * wf_stripes; k += static_cast<int>(blockDim.x))
    {
        in[wf_counter_class + k] = 0;
    }
    if (t < wf_stripes)
    {
        out[wf_counter_active + t] = 0;
        pool.counters[wf_counter_free + t] = 0;
    }
}

// The refilled slots start their items' paths and join the active list's last segment
__global__ void __launch_bounds__(256) wf_refill(const camera_rec<float> cam, const scene_view<float> sc,
                                                 wf_pool pool, int p)
{
    const uint64_t seed = pool.launch->seed;
    const launch_work work = pool.launch->work;
    __shared__ uint32_t prefix[wf_stripes + 1];
    load_prefix(prefix, pool.prefix + wf_prefix_free, wf_stripes + 1);
    const uint32_t refill = pool.counters[wf_counter_refill];
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= refill)
    {
        return;
    }
    const int stripe = find_segment(prefix, wf_stripes, g);
    const uint32_t s = pool.free[static_cast<size_t>(stripe) * pool.segment + (g - prefix[stripe])];
    const uint32_t item = pool.counters[wf_counter_refill_base] + g;
    store_path(pool, s, start_item(cam, sc, seed, work, pool, s, item));
    pool.item[s] = item;
    pool.active[p][static_cast<size_t>(wf_stripes) * pool.segment + g] = s;
}

// The queues' prefix sums, class-major (one wave)
__global__ void __launch_bounds__(64) wf_sort(wf_pool pool, int p)
{
    wave_prefix(pool.counters + wf_counter_stride * p + wf_counter_class, pool.prefix + wf_prefix_class,
                wf_classes * wf_stripes);
}

// Walk: walk_instances (any one-level or two-level tree of float nodes) or walk_packed (a packed tree)
template <int Walk>
__global__ void __launch_bounds__(256) wf_extend(const scene_view<float> sc, wf_pool pool, int p)
{
    __shared__ uint32_t prefix[wf_stripes + 2];
    load_prefix(prefix, pool.prefix + wf_prefix_active, wf_stripes + 2);
    const lds_stack stack = thread_stack(sc.stack_depth);
    const uint32_t n = prefix[wf_stripes + 1];
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (blockIdx.x * blockDim.x >= n)
    {
        return; // the whole block is past the list (uniform across the block: no barrier below is skipped by some)
    }
    int cls = -1;
    uint32_t s = 0;
    if (g < n)
    {
        s = active_slot(pool, p, prefix, g);
        slot_path sp = load_path(pool, s);
        tally<false> tl;
        const hit<float> h = closest_event<float, Walk>(sc, sp.p.r, sp.p.from, sp.p.inward, sp.g, stack, tl);
        pool.hit[s] = make_int4(h.index, __float_as_int(h.t), h.entering ? 1 : 0, 0);
        if (sc.medium_count > 0)
        {
            store_rng(pool, s, sp.g); // the media drew their free flights from the path's generator
        }
        cls = hit_class(sc, h.index);
    }
    block_tally(pool.tallies, 0, cls >= 0 ? 1 : 0);
    block_append<wf_classes>(cls, s, pool.counters + wf_counter_stride * p + wf_counter_class, pool.queue,
                             pool.segment);
}

// A queue entry's slot (the queues' prefix sums in LDS)
__device__ inline uint32_t queued_slot(const wf_pool& pool, const uint32_t* prefix, uint32_t g)
{
    const int list = find_segment(prefix, wf_classes * wf_stripes, g);
    return pool.queue[static_cast<size_t>(list) * pool.segment + (g - prefix[list])];
}

// After shading: a path that goes on joins the next active list; one that ended leaves its sample and joins the free
// list (every thread of the block calls it)
__device__ inline void after_shade(const wf_pool& pool, int p, uint32_t s, bool valid, bool done, vec3<float> sample)
{
    if (valid && done)
    {
        pool.samples[pool.item[s]] = make_float4(sample.x, sample.y, sample.z, 0.0f);
    }
    block_append<1>(valid && !done ? 0 : -1, s, pool.counters + wf_counter_stride * (1 - p) + wf_counter_active,
                    pool.active[1 - p], pool.segment);
    block_append<1>(valid && done ? 0 : -1, s, pool.counters + wf_counter_free, pool.free, pool.segment);
}

// FirstBook: the scene is a first book's (first_book_after_walk, the megakernel's shading of such scenes)
template <bool FirstBook>
__global__ void __launch_bounds__(256) wf_shade(const camera_rec<float> cam, const scene_view<float> sc,
                                                wf_pool pool, int p)
{
    __shared__ uint32_t prefix[wf_classes * wf_stripes + 1];
    load_prefix(prefix, pool.prefix + wf_prefix_class, wf_classes * wf_stripes + 1);
    const uint32_t total = prefix[wf_classes * wf_stripes];
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (blockIdx.x * blockDim.x >= total)
    {
        return;
    }
    bool done = false;
    uint32_t s = 0;
    vec3<float> sample{0.0f, 0.0f, 0.0f};
    if (g < total)
    {
        s = queued_slot(pool, prefix, g);
        slot_path sp = load_path(pool, s);
        const int4 hv = pool.hit[s];
        const hit<float> h{hv.x, __int_as_float(hv.y), hv.z != 0};
        tally<false> tl;
        if constexpr (FirstBook)
        {
            done = first_book_after_walk(sp.p, sc, cam, sp.g, sample, h, tl);
        }
        else if (h.index < 0)
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
        }
    }
    after_shade(pool, p, s, g < total, done, sample);
}

// The quality integrator's shade: the same queues in class order, each vertex's light sample and its shadow ray
// traced here (an any-hit walk with this thread's LDS stack), in the megakernel's order of random numbers
template <int Walk>
__global__ void __launch_bounds__(256) wf_shade_quality(const camera_rec<float> cam, const scene_view<float> sc,
                                                        wf_pool pool, int p)
{
    __shared__ uint32_t prefix[wf_classes * wf_stripes + 1];
    load_prefix(prefix, pool.prefix + wf_prefix_class, wf_classes * wf_stripes + 1);
    const lds_stack stack = thread_stack(sc.stack_depth);
    const uint32_t total = prefix[wf_classes * wf_stripes];
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (blockIdx.x * blockDim.x >= total)
    {
        return;
    }
    bool done = false;
    uint32_t s = 0;
    uint32_t shadow_rays = 0;
    vec3<float> sample{0.0f, 0.0f, 0.0f};
    if (g < total)
    {
        s = queued_slot(pool, prefix, g);
        slot_path sp = load_path(pool, s);
        const float4 st = pool.qst[s];
        quality_path<float> q;
        q.base = sp.p;
        q.radiance = {st.x, st.y, st.z};
        q.bsdf_pdf = st.w;
        q.specular = sp.specular;
        const int4 hv = pool.hit[s];
        const hit<float> h{hv.x, __int_as_float(hv.y), hv.z != 0};
        tally<false> tl;
        done = shade_quality<Walk>(q, sc, cam, h, sp.g, sample, stack, tl, shadow_rays);
        if (!done)
        {
            sp.p = q.base;
            sp.specular = q.specular;
            pool.qst[s] = make_float4(q.radiance.x, q.radiance.y, q.radiance.z, q.bsdf_pdf);
            store_path(pool, s, sp);
        }
    }
    block_tally(pool.tallies, 1, shadow_rays);
    after_shade(pool, p, s, g < total, done, sample);
}

// The drain: once every item is handed out and few slots are left active, each of them (a thread each) traces its
// path to the end, as the megakernel does: the bounces of a nearly empty pool would cost a launch each for little work
template <int Mode, int Walk>
__global__ void __launch_bounds__(128) wf_drain(const camera_rec<float> cam, const scene_view<float> sc,
                                                const uint64_t seed, const launch_work work, wf_pool pool)
{
    __shared__ uint32_t prefix[wf_stripes + 2];
    load_prefix(prefix, pool.prefix + wf_prefix_active, wf_stripes + 2);
    const lds_stack stack = thread_stack(sc.stack_depth);
    const uint32_t n = prefix[wf_stripes + 1];
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long rays = 0;
    uint32_t shadow_rays = 0;
    if (k < n)
    {
        const uint32_t s = active_slot(pool, 0, prefix, k);
        slot_path sp = load_path(pool, s);
        tally<false> tl;
        quality_path<float> q;
        if constexpr (Mode == mode_quality)
        {
            const float4 st = pool.qst[s];
            q.base = sp.p;
            q.radiance = {st.x, st.y, st.z};
            q.bsdf_pdf = st.w;
            q.specular = sp.specular;
        }
        vec3<float> sample;
        while (true)
        {
            ++rays;
            bool done;
            if constexpr (Mode == mode_quality)
            {
                done = trace_quality<Walk>(q, sc, cam, sp.g, sample, stack, tl, shadow_rays);
            }
            else if constexpr (Mode == mode_all_books)
            {
                done = trace_all_books(sp.p, sc, cam, sp.g, sample, stack, tl);
            }
            else
            {
                done = trace_first_book(sp.p, sc, cam, sp.g, sample, stack, tl);
            }
            if (done)
            {
                break;
            }
        }
        pool.samples[pool.item[s]] = make_float4(sample.x, sample.y, sample.z, 0.0f);
    }
    block_tally(pool.tallies, 0, rays);
    block_tally(pool.tallies, 1, shadow_rays);
}

// The launch's end: each pixel's samples added to its sums in pass order (NaN counted as 0, as the megakernel adds
// them); block 0 also adds the ray counts to the session's counters and clears the tallies
__global__ void __launch_bounds__(256) wf_finish(const camera_rec<float> cam, const launch_work work, wf_pool pool,
                                                 double* __restrict__ sums, unsigned long long* rays)
{
    const uint32_t width = static_cast<uint32_t>(cam.image_width);
    const uint32_t pixels = width * static_cast<uint32_t>(work.row_count);
    const uint32_t local = blockIdx.x * blockDim.x + threadIdx.x;
    if (blockIdx.x == 0 && threadIdx.x < 64)
    {
        const int t = static_cast<int>(threadIdx.x);
        unsigned long long mine[2] = {0, 0};
        if (t < wf_stripes)
        {
            mine[0] = pool.tallies[2 * t];
            mine[1] = pool.tallies[2 * t + 1];
            pool.tallies[2 * t] = 0;
            pool.tallies[2 * t + 1] = 0;
        }
        const unsigned long long r = wsum(mine[0]);
        const unsigned long long sh = wsum(mine[1]);
        if (t == 0)
        {
            rays[0] += r;
            rays[1] += sh;
        }
    }
    if (local >= pixels)
    {
        return;
    }
    const uint64_t pixel = static_cast<uint64_t>(work.first_row) * width + local;
    double* const sum = sums + 3 * pixel;
    double red = sum[0], green = sum[1], blue = sum[2];
    for (int q = 0; q < work.pass_count; ++q)
    {
        const float4 v = pool.samples[static_cast<size_t>(q) * pixels + local];
        red += without_nan(v.x);
        green += without_nan(v.y);
        blue += without_nan(v.z);
    }
    sum[0] = red;
    sum[1] = green;
    sum[2] = blue;
}
}

namespace rt
{
hipError_t wf_launch_init(const dev::wf_pool& pool, uint32_t items, hipStream_t stream)
{
    hipLaunchKernelGGL(dev::wf_init, dim3(wf_blocks(pool.capacity)), dim3(dev::wf_block), 0, stream, pool, items);
    return hipGetLastError();
}

hipError_t wf_launch_bounce(const dev::camera_rec<float>& cam, const dev::scene_view<float>& sc,
                            const dev::wf_pool& pool, int parity, int mode, hipStream_t stream)
{
    const dim3 grid(wf_blocks(pool.capacity));
    const size_t lds = stack_lds_bytes<float>(sc.stack_depth, dev::wf_block);
    hipLaunchKernelGGL(dev::wf_begin, dim3(1), dim3(256), 0, stream, pool, parity);
    hipLaunchKernelGGL(dev::wf_refill, grid, dim3(dev::wf_block), 0, stream, cam, sc, pool, parity);
    const bool packed = dev::tree_walk(sc) == dev::walk_packed;
    if (packed)
    {
        hipLaunchKernelGGL(dev::wf_extend<dev::walk_packed>, grid, dim3(dev::wf_block), lds, stream, sc, pool, parity);
    }
    else
    {
        hipLaunchKernelGGL(dev::wf_extend<dev::walk_instances>, grid, dim3(dev::wf_block), lds, stream, sc, pool,
                           parity);
    }
    hipLaunchKernelGGL(dev::wf_sort, dim3(1), dim3(64), 0, stream, pool, parity);
    if (mode == dev::mode_quality && packed)
    {
        hipLaunchKernelGGL(dev::wf_shade_quality<dev::walk_packed>, grid, dim3(dev::wf_block), lds, stream, cam, sc,
                           pool, parity);
    }
    else if (mode == dev::mode_quality)
    {
        hipLaunchKernelGGL(dev::wf_shade_quality<dev::walk_instances>, grid, dim3(dev::wf_block), lds, stream, cam, sc,
                           pool, parity);
    }
    else if (mode == dev::mode_first_book)
    {
        hipLaunchKernelGGL(dev::wf_shade<true>, grid, dim3(dev::wf_block), 0, stream, cam, sc, pool, parity);
    }
    else
    {
        hipLaunchKernelGGL(dev::wf_shade<false>, grid, dim3(dev::wf_block), 0, stream, cam, sc, pool, parity);
    }
    return hipGetLastError();
}

hipError_t wf_launch_drain(int mode, const dev::camera_rec<float>& cam, const dev::scene_view<float>& sc, uint64_t seed,
                           const dev::launch_work& work, const dev::wf_pool& pool, uint32_t count, hipStream_t stream)
{
    hipLaunchKernelGGL(dev::wf_begin, dim3(1), dim3(256), 0, stream, pool, 0);
    const int threads = 128;
    const dim3 grid((count + threads - 1) / threads);
    const size_t lds = stack_lds_bytes<float>(sc.stack_depth, threads);
    if (mode == dev::mode_quality && dev::tree_walk(sc) == dev::walk_packed)
    {
        hipLaunchKernelGGL((dev::wf_drain<dev::mode_quality, dev::walk_packed>), grid, dim3(threads), lds, stream, cam,
                           sc, seed, work, pool);
    }
    else if (mode == dev::mode_quality)
    {
        hipLaunchKernelGGL((dev::wf_drain<dev::mode_quality, dev::walk_instances>), grid, dim3(threads), lds, stream,
                           cam, sc, seed, work, pool);
    }
    else if (mode == dev::mode_all_books)
    {
        hipLaunchKernelGGL((dev::wf_drain<dev::mode_all_books, dev::walk_instances>), grid, dim3(threads), lds, stream,
                           cam, sc, seed, work, pool);
    }
    else
    {
        hipLaunchKernelGGL((dev::wf_drain<dev::mode_first_book, dev::walk_instances>), grid, dim3(threads), lds, stream,
                           cam, sc, seed, work, pool);
    }
    return hipGetLastError();
}

hipError_t wf_launch_finish(const dev::camera_rec<float>& cam, const dev::launch_work& work, const dev::wf_pool& pool,
                            double* sums, unsigned long long* rays, hipStream_t stream)
{
    const uint32_t pixels = static_cast<uint32_t>(cam.image_width) * static_cast<uint32_t>(work.row_count);
    hipLaunchKernelGGL(dev::wf_finish, dim3(wf_blocks(pixels)), dim3(dev::wf_block), 0, stream, cam, work, pool, sums,
                       rays);
    return hipGetLastError();
}

kernel_attributes wf_attributes(int depth, int which)
{
    hipFuncAttributes a{};
    const void* f = which == 0   ? reinterpret_cast<const void*>(&dev::wf_extend<dev::walk_instances>)
                    : which == 1 ? reinterpret_cast<const void*>(&dev::wf_shade<false>)
                                 : reinterpret_cast<const void*>(&dev::wf_shade_quality<dev::walk_instances>);
    kernel_attributes k{};
    if (hipFuncGetAttributes(&a, f) == hipSuccess)
    {
        k.registers = a.numRegs;
        k.local_bytes = static_cast<int>(a.localSizeBytes);
    }
    k.lds_bytes = which != 1 ? static_cast<int>(stack_lds_bytes<float>(depth, dev::wf_block)) : 0;
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, f, dev::wf_block, k.lds_bytes) == hipSuccess)
    {
        k.blocks_per_cu = blocks;
    }
    return k;
}
}
