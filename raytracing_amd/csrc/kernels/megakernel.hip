// The megakernel: every path traced by one thread, one ray per loop iteration, and each pixel's samples added to its
// double sums in pass order (src/gpu/renderer.cu:340-397). It is the plain baseline the wavefront integrator is
// measured against, and the path the FP64 renders take.
//
// MI355X mapping: the grid is as many wave64s as fit the chip at once (persistent); they take a launch's work from a
// queue of chunks, each a 16x4 tile of pixels over a group of passes, with one atomic per chunk, and share each chunk's
// items among their lanes (wave_queue); each sample is staged in HBM and a second kernel adds them to the sums in pass
// order (add_staged). The BVH stack lives in LDS, one column of entries per thread (path_tracer.hpp, lds_stack), sized
// by the tree's worst case at launch; the ray counts are summed across the wave with shuffles, then one atomic per
// wave.
#include <algorithm>
#include <cstdint>
#include <mutex>
#include <vector>

#include "chunk_queue.hpp"
#include "kernel_api.hpp"
#include "quality.hpp"

namespace rt::dev
{
__device__ inline unsigned long long wave_sum(unsigned long long v)
{
    for (int offset = 32; offset > 0; offset >>= 1)
    {
        v += __shfl_down(v, offset, 64);
    }
    return v;
}

static_assert(tally_events == RT_TRAFFIC_EVENTS && tally_phases == RT_TRAFFIC_PHASES,
              "the traffic report lists every tally event and phase");

// The wave's counts: the rays (and shadow rays) summed across the wave, then one atomic per wave
template <bool Count, typename Tally>
__device__ inline void add_counts(unsigned long long rays, unsigned long long shadow_rays,
                                  unsigned long long* __restrict__ ray_count, unsigned long long* __restrict__ tallies,
                                  const Tally& tl)
{
    rays = wave_sum(rays);
    shadow_rays = wave_sum(shadow_rays);
    const bool lane0 = ((threadIdx.y * blockDim.x + threadIdx.x) & 63) == 0;
    if (lane0)
    {
        atomicAdd(ray_count, rays);
        if (shadow_rays)
        {
            atomicAdd(ray_count + 1, shadow_rays);
        }
    }
    if constexpr (Count)
    {
        for (int e = 0; e < tally_events; ++e)
        {
            const unsigned long long total = wave_sum(static_cast<unsigned long long>(tl.n[e]));
            if (lane0)
            {
                atomicAdd(tallies + e, total);
            }
        }
        // each phase's lane-cycles, after the events
        for (int k = 0; k < tally_phases; ++k)
        {
            const unsigned long long total = wave_sum(static_cast<unsigned long long>(tl.cycles[k]));
            if (lane0)
            {
                atomicAdd(tallies + tally_events + k, total);
            }
        }
    }
}

// A wave's place in the queue: the chunk it takes items from and the chunk's next item, the same in every lane. A
// lane whose path ends claims the wave's next item (a ballot and a popcount, no atomic); once the wave's chunk is used
// up, the wave takes the queue's next chunk (one atomic, by its first active lane) and its lanes go on with that
// chunk's items. So no lane waits for the wave's longer paths and no wave waits for another wave's costly tile: the
// lanes stay busy until the queue is empty. Any lane may trace any pixel's pass: a sample depends on the pixel and the
// pass only, and the samples are added to the sums in pass order afterwards (add_staged).
struct wave_queue
{
    uint64_t below;  // the lanes below this one
    chunk_span span; // the wave's chunk
    uint32_t next;   // its next unclaimed item
    bool drained;    // the queue has no chunk left

    __device__ static wave_queue start()
    {
        const int lane = static_cast<int>(threadIdx.x & 63);
        return {(uint64_t{1} << lane) - 1, {0, 0, 0, 0, 0, 0}, 0, false};
    }

    // Claims an item for every lane that has none (needy): true with this lane's pixel and pass when it got one
    __device__ bool claim(const chunk_queue& q, bool needy, int& i, int& j, int& pass_local)
    {
        const uint64_t ask = __ballot(needy);
        const uint32_t n = static_cast<uint32_t>(__popcll(ask));
        const uint32_t rank = static_cast<uint32_t>(__popcll(ask & below));
        const uint32_t avail = span.items - next;
        chunk_span from = span;
        uint32_t k = next + rank;
        if (n > avail)
        {
            chunk_span fresh{0, 0, 0, 0, 0, 0};
            if (!drained)
            {
                uint32_t c = 0;
                if ((__ballot(true) & below) == 0)
                {
                    c = atomicAdd(q.next_chunk, 1u);
                }
                c = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(c)));
                fresh = span_of(q, c);
                drained = c >= q.chunks;
            }
            if (rank >= avail)
            {
                from = fresh;
                k = rank - avail;
            }
            next = min(n - avail, fresh.items);
            span = fresh;
        }
        else
        {
            next += n;
        }
        if (!needy || k >= from.items)
        {
            return false;
        }
        item_of(from, k, q.work.first_row, i, j, pass_local);
        return true;
    }
};

// A staged sample is written once and read once: non-temporal stores and loads, so that the samples don't push the
// tree and the scene's records out of L2 (with cached ones the L2 hit rate of C2 fell from 99 % to 89 %)
template <typename Real>
__device__ inline void stage(staged_sample<Real>* at, const vec3<Real>& s)
{
    typedef Real quad __attribute__((ext_vector_type(4)));
    const quad v = {s.x, s.y, s.z, Real(0)};
    __builtin_nontemporal_store(v, reinterpret_cast<quad*>(at));
}
template <typename Real>
__device__ inline staged_sample<Real> unstage(const staged_sample<Real>* at)
{
    typedef Real quad __attribute__((ext_vector_type(4)));
    const quad v = __builtin_nontemporal_load(reinterpret_cast<const quad*>(at));
    return {v.x, v.y, v.z, Real(0)};
}

// After a megakernel launch: each pixel of the launch's rows adds its staged samples to its double sums in pass order,
// the same additions in the same order as one thread looping over its pixel's passes, so the image is the same bytes
// whichever lane traced which sample (src/gpu/renderer.cu:340-397: the sums of a pixel's passes, NaN as 0)
template <typename Real>
__global__ void __launch_bounds__(256) add_staged(const staged_sample<Real>* __restrict__ staged,
                                                  double* __restrict__ sums, int width, launch_work work)
{
    const int64_t pixels = static_cast<int64_t>(width) * work.row_count;
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= pixels)
    {
        return;
    }
    double* const sum = sums + 3 * (static_cast<int64_t>(work.first_row) * width + k);
    double red = sum[0], green = sum[1], blue = sum[2];
    for (int q = 0; q < work.pass_count; ++q)
    {
        const staged_sample<Real> s = unstage(staged + q * pixels + k);
        red += without_nan(s.x);
        green += without_nan(s.y);
        blue += without_nan(s.z);
    }
    sum[0] = red;
    sum[1] = green;
    sum[2] = blue;
}

// The queue's head, cleared by a kernel before each launch (no memset node where a caller captures the launches)
__global__ void clear_queue(uint32_t* next_chunk) { *next_chunk = 0; }

// What a persistent megakernel reads at every ray: the camera, the scene, the launch's queue and seed. They are the
// kernel's first argument, so they lie at the start of the kernarg segment.
template <typename Real>
struct megakernel_args
{
    camera_rec<Real> cam;
    scene_view<Real> sc;
    chunk_queue q;
    uint64_t seed;
};

// The kernel's first argument, read again from the kernarg segment where this is called (constant memory: scalar
// loads through the scalar cache). The empty asm hides the pointer's value, so the compiler can't hoist the reads to
// the kernel's entry: without it the camera's and the scene's uniform values were all loaded there and held for the
// whole kernel, more than the scalar registers hold, and the spilled ones were read back from vector registers' lanes
// (v_readlane) inside the walk's loop. Read per ray, each value is live only where it is used.
template <typename Args>
__device__ inline const Args& first_kernel_argument()
{
    auto at = __builtin_amdgcn_kernarg_segment_ptr();
    asm volatile("" : "+s"(at));
    return *(const Args*)at;
}

// Waves: the waves per SIMD the register allocation must leave room for (0: the compiler's choice); a tighter budget
// trades spills for occupancy, and the session's autotuner picks per scene. The grid is as many waves as fit the chip
// at once; each lane follows its path ray by ray (src/gpu/renderer.cu:340-397) and claims its next item from its
// wave's place in the queue as the path ends (wave_queue). (Walking the tree in bursts of a few steps per iteration,
// so that a lane whose walk is over shades while a long walk goes on, measured slower here: the walk's state then
// stays live across the shading code, which costs registers.)
// Marble: the scene has a marble (a noise texture); the three books' float kernel then keeps its waves whole until
// every lane is done, so that its lanes evaluate the marble's turbulence together (wave_turbulence needs every lane)
template <typename Real, bool AllBooks, bool Count, int Waves, bool Marble = false>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(Waves > 0 ? Waves : 1, 8)))
megakernel(const megakernel_args<Real> args, staged_sample<Real>* __restrict__ staging,
           unsigned long long* __restrict__ ray_count, unsigned long long* __restrict__ tallies)
{
    const lds_stack stack = block_stack(args.sc.stack_depth);
    tally<Count> tl;
    wave_queue w = wave_queue::start();
    unsigned long long rays = 0;
    int32_t at = -1; // where this lane's path's sample goes, or -1 between items
    path<Real> p{};
    pcg32 g{};
    constexpr bool whole = Marble && AllBooks && sizeof(Real) == sizeof(float);
    while (true)
    {
        const megakernel_args<Real>& a = first_kernel_argument<megakernel_args<Real>>();
        const camera_rec<Real>& cam = a.cam;
        const scene_view<Real>& sc = a.sc;
        const chunk_queue& q = a.q;
        int si, sj, pass_local;
        if (w.claim(q, at < 0, si, sj, pass_local))
        {
            const int pass = q.work.first_pass + pass_local;
            const uint64_t pixel =
                static_cast<uint64_t>(sj) * static_cast<uint64_t>(cam.image_width) + static_cast<uint64_t>(si);
            g = sample_generator(a.seed, pixel, static_cast<uint64_t>(pass));
            p = start_path(cam, AllBooks && sc.motion != 0, si, sj, pass, g);
            at = staged_at(q, si, sj, pass_local);
        }
        const bool busy = at >= 0;
        if (whole ? __ballot(busy) == 0 : !busy)
        {
            if (w.drained)
            {
                break;
            }
            continue; // the wave's chunk ran out for this lane in this round; it claims again
        }
        tl.lap(phase_queue);
        rays += busy ? 1 : 0;
        vec3<Real> sample;
        bool done;
        if constexpr (whole)
        {
            done = trace_all_books_wave(p, sc, cam, g, sample, stack, tl, busy);
        }
        else
        {
            done = AllBooks ? trace_all_books(p, sc, cam, g, sample, stack, tl)
                            : trace_first_book(p, sc, cam, g, sample, stack, tl);
        }
        if (busy && done)
        {
            stage(staging + at, sample);
            at = -1;
        }
    }
    add_counts<Count>(rays, 0, ray_count, tallies, tl);
}

// The deeper levels of the quality megakernel's traversal stacks, in HBM (split_stack): levels [near, depth) of each
// grid thread, refs then distances, one column per thread
struct far_stack
{
    int32_t near;     // levels kept in LDS
    int32_t* refs;    // (depth - near) x stride
    uint16_t* ts;     // (depth - near) x stride
    uint32_t stride;  // threads in the grid
};

// The quality integrator's paths (src/gpu/quality.cu:53-109): the same loop over the queue's items, with light
// samples and their shadow rays at every vertex that isn't a mirror's (quality.hpp). One ray per iteration: a path's
// shadow ray is its own iteration, walked by the same code as the paths' next rays (step_quality), so a wave's lanes
// walk together whichever kind of ray each traces, and the shading's values aren't held across a second walk.
template <typename Real, bool Count, int Walk, typename Stack>
__device__ inline void quality_paths(staged_sample<Real>* __restrict__ staging,
                                     unsigned long long* __restrict__ ray_count,
                                     unsigned long long* __restrict__ tallies, const Stack& stack)
{
    tally<Count> tl;
    wave_queue w = wave_queue::start();
    unsigned long long rays = 0;
    uint32_t shadow_rays = 0;
    int32_t at = -1;
    quality_path<Real> qp{};
    pcg32 g{};
    while (true)
    {
        const megakernel_args<Real>& a = first_kernel_argument<megakernel_args<Real>>();
        const camera_rec<Real>& cam = a.cam;
        const scene_view<Real>& sc = a.sc;
        const chunk_queue& q = a.q;
        int si, sj, pass_local;
        if (w.claim(q, at < 0, si, sj, pass_local))
        {
            const int pass = q.work.first_pass + pass_local;
            const uint64_t pixel =
                static_cast<uint64_t>(sj) * static_cast<uint64_t>(cam.image_width) + static_cast<uint64_t>(si);
            g = sample_generator(a.seed, pixel, static_cast<uint64_t>(pass));
            qp = start_quality(cam, sc.motion != 0, si, sj, pass, g);
            at = staged_at(q, si, sj, pass_local);
        }
        if (at < 0)
        {
            if (w.drained)
            {
                break;
            }
            continue;
        }
        tl.lap(phase_queue);
        vec3<Real> sample;
        if (step_quality<Walk>(qp, sc, cam, g, sample, stack, tl, rays, shadow_rays))
        {
            stage(staging + at, sample);
            at = -1;
        }
    }
    add_counts<Count>(rays, shadow_rays, ray_count, tallies, tl);
}

// Split: the stack's first far.near levels in LDS, the rest in HBM (a deep tree, such as a large mesh's, would
// otherwise hold so much LDS per wave that fewer waves fit than the registers allow)
// Walk: the tree's walk (walk_all, walk_instances or walk_packed). Real: float, or double for the FP64 oracle (at the
// compiler's register budget)
template <typename Real, bool Count, int Waves, bool Split, int Walk>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(Waves > 0 ? Waves : 1, 8)))
quality_megakernel(const megakernel_args<Real> args, staged_sample<Real>* __restrict__ staging,
                   unsigned long long* __restrict__ ray_count, unsigned long long* __restrict__ tallies,
                   const far_stack far)
{
    if constexpr (Split)
    {
        const uint32_t column = blockIdx.x * blockDim.x + threadIdx.x;
        const split_stack stack{block_stack(far.near), far.near, far.refs + column, far.ts + column, far.stride};
        quality_paths<Real, Count, Walk>(staging, ray_count, tallies, stack);
    }
    else
    {
        quality_paths<Real, Count, Walk>(staging, ray_count, tallies, block_stack(args.sc.stack_depth));
    }
}

// The plain megakernel, the reference's own mapping (src/gpu/renderer.cu:340-397,421-451): one thread per pixel of a
// 16x8 block, all of the launch's passes in the thread's loop, one path after the other, each added to the pixel's
// double sums as it ends. No queue and no staging: a wave waits for its longest pixel, and a launch for its costliest
// blocks. It is the baseline the batched integrators are measured against; its sums are the others' bytes (the same
// samples added in the same order).
template <typename Real>
struct plain_args
{
    camera_rec<Real> cam;
    scene_view<Real> sc;
    launch_work work;
    uint64_t seed;
};
template <typename Real, int Mode, int Walk>
__global__ void __launch_bounds__(128) plain_megakernel(const plain_args<Real> args, double* __restrict__ sums,
                                                        unsigned long long* __restrict__ ray_count)
{
    const lds_stack stack = block_stack(args.sc.stack_depth);
    tally<false> tl;
    const launch_work& work = args.work;
    const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    const int j = work.first_row + static_cast<int>(blockIdx.y * blockDim.y + threadIdx.y);
    unsigned long long rays = 0;
    uint32_t shadow_rays = 0;
    if (i < args.cam.image_width && j < work.first_row + work.row_count)
    {
        const uint64_t pixel = static_cast<uint64_t>(j) * static_cast<uint64_t>(args.cam.image_width) + i;
        double* const sum = sums + 3 * pixel;
        double red = sum[0], green = sum[1], blue = sum[2];
        for (int k = 0; k < work.pass_count; ++k)
        {
            // (the camera and the scene read again per pass, as the queue megakernels read them per ray)
            const plain_args<Real>& a = first_kernel_argument<plain_args<Real>>();
            const camera_rec<Real>& cam = a.cam;
            const scene_view<Real>& sc = a.sc;
            const int pass = a.work.first_pass + k;
            pcg32 g = sample_generator(a.seed, pixel, static_cast<uint64_t>(pass));
            vec3<Real> sample;
            if constexpr (Mode == mode_quality)
            {
                quality_path<Real> qp = start_quality(cam, sc.motion != 0, i, j, pass, g);
                do
                {
                    ++rays;
                } while (!trace_quality<Walk>(qp, sc, cam, g, sample, stack, tl, shadow_rays));
            }
            else
            {
                path<Real> p = start_path(cam, Mode == mode_all_books && sc.motion != 0, i, j, pass, g);
                do
                {
                    ++rays;
                } while (!(Mode == mode_all_books ? trace_all_books(p, sc, cam, g, sample, stack, tl)
                                                  : trace_first_book(p, sc, cam, g, sample, stack, tl)));
            }
            red += without_nan(sample.x);
            green += without_nan(sample.y);
            blue += without_nan(sample.z);
        }
        sum[0] = red;
        sum[1] = green;
        sum[2] = blue;
    }
    add_counts<false>(rays, shadow_rays, ray_count, nullptr, tl);
}

// The parity probe: closest_hit on a batch of given rays, with the walk of the scene's kernels (Kinds: walk_instances,
// which walks any one-level or two-level tree, or walk_packed for a packed tree)
template <typename Real, int Kinds>
__global__ void __launch_bounds__(256) closest_hit_kernel(const scene_view<Real> sc, const Real* origins,
                                                          const Real* directions, const Real* times,
                                                          const int32_t* from, const uint8_t* inward, int64_t n,
                                                          int32_t* index, Real* t, uint8_t* entering)
{
    const lds_stack stack = block_stack(sc.stack_depth);
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n)
    {
        return;
    }
    const ray<Real> r{{origins[3 * k], origins[3 * k + 1], origins[3 * k + 2]},
                      {directions[3 * k], directions[3 * k + 1], directions[3 * k + 2]},
                      times[k]};
    tally<false> tl;
    const hit<Real> h = closest_hit<Real, Kinds>(sc, r, from[k], inward[k] != 0, stack, tl);
    index[k] = h.index;
    t[k] = h.t;
    entering[k] = h.entering ? 1 : 0;
}

// Per-pixel means of the sums, and the books' display bytes (src/gpu/display.cu:19-44): sqrt, clamp to [0, 0.999],
// times 256, truncated
__global__ void __launch_bounds__(256) display_kernel(const double* sums, int64_t pixels, double inverse_passes,
                                                      uint8_t* rgba)
{
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= pixels)
    {
        return;
    }
    uint8_t out[4];
    for (int c = 0; c < 3; ++c)
    {
        const double linear = sums[3 * k + c] * inverse_passes;
        const double gamma = linear > 0 ? sqrt(linear) : 0.0;
        const double clamped = gamma < 0.0 ? 0.0 : (gamma > 0.999 ? 0.999 : gamma);
        out[c] = static_cast<uint8_t>(256 * clamped);
    }
    out[3] = 255;
    rgba[4 * k] = out[0];
    rgba[4 * k + 1] = out[1];
    rgba[4 * k + 2] = out[2];
    rgba[4 * k + 3] = out[3];
}
}

namespace rt::dev
{
// The scratch memory of rt_session_render_to_static (the entry point for callers that can't allocate): the chunk
// queue's head in its first 256 bytes, the staged samples after it. One array per GPU, held until the process ends;
// every call on a GPU shares it, so calls that overlap on one GPU race.
constexpr size_t static_scratch_bytes = size_t{64} << 20;
static __device__ __align__(256) unsigned char scratch[static_scratch_bytes];
}

namespace rt
{
template <typename Real>
size_t stack_lds_bytes(int entries, int threads)
{
    return dev::stack_bytes(entries, threads);
}

// The static scratch's device address on the current device: looked up once per device with that device current
// (the array's name is only a host-side stand-in), then kept
void* static_scratch(size_t* bytes)
{
    static std::mutex guard;
    static std::vector<void*> address;
    int device = 0;
    if (hipGetDevice(&device) != hipSuccess || device < 0)
    {
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(guard);
    if (address.size() <= static_cast<size_t>(device))
    {
        address.resize(static_cast<size_t>(device) + 1, nullptr);
    }
    if (address[static_cast<size_t>(device)] == nullptr)
    {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(dev::scratch)) != hipSuccess || p == nullptr)
        {
            return nullptr;
        }
        address[static_cast<size_t>(device)] = p;
    }
    *bytes = dev::static_scratch_bytes;
    return address[static_cast<size_t>(device)];
}

namespace
{
template <typename Real, bool Count, int Waves, int Walk>
const void* quality_megakernel_for(bool split)
{
    return split ? reinterpret_cast<const void*>(&dev::quality_megakernel<Real, Count, Waves, true, Walk>)
                 : reinterpret_cast<const void*>(&dev::quality_megakernel<Real, Count, Waves, false, Walk>);
}

// walk: the quality kernels' walk of the scene's tree (dev::tree_walk)
template <typename Real, bool Count, int Waves>
const void* megakernel_for_waves(int mode, bool split, int walk, bool marble)
{
    if (mode == dev::mode_quality)
    {
        if constexpr (sizeof(Real) == sizeof(float))
        {
            return walk == dev::walk_instances ? quality_megakernel_for<Real, Count, Waves, dev::walk_instances>(split)
                   : walk == dev::walk_packed  ? quality_megakernel_for<Real, Count, Waves, dev::walk_packed>(split)
                                               : quality_megakernel_for<Real, Count, Waves, dev::walk_all>(split);
        }
        else
        {
            // (FP64: the compiler's budget, float nodes)
            return walk == dev::walk_instances ? quality_megakernel_for<Real, Count, 0, dev::walk_instances>(split)
                                               : quality_megakernel_for<Real, Count, 0, dev::walk_all>(split);
        }
    }
    if constexpr (sizeof(Real) == sizeof(float))
    {
        if (mode == dev::mode_all_books && marble)
        {
            return reinterpret_cast<const void*>(&dev::megakernel<Real, true, Count, Waves, true>);
        }
    }
    return mode == dev::mode_all_books ? reinterpret_cast<const void*>(&dev::megakernel<Real, true, Count, Waves>)
                                       : reinterpret_cast<const void*>(&dev::megakernel<Real, false, Count, Waves>);
}

// The instance for a mode and an occupancy target: the counting instances and the FP64 ones at the compiler's choice.
// split: the quality megakernel with its stack's deeper levels in HBM
template <typename Real, bool Count>
// marble: the scene has a noise texture (the three books' kernel that keeps its waves whole)
const void* megakernel_for(int mode, int waves, bool split, int walk, bool marble)
{
    if constexpr (!Count && sizeof(Real) == sizeof(float))
    {
        if (waves == 5)
        {
            return megakernel_for_waves<Real, Count, 5>(mode, split, walk, marble);
        }
        if (waves == 6)
        {
            return megakernel_for_waves<Real, Count, 6>(mode, split, walk, marble);
        }
        if (waves == 7)
        {
            return megakernel_for_waves<Real, Count, 7>(mode, split, walk, marble);
        }
    }
    return megakernel_for_waves<Real, Count, 0>(mode, split, walk, marble);
}

// The stack levels a split quality megakernel keeps in LDS: as many as let the waves its registers allow share a CU's
// LDS (at least 8); the whole stack (no split) when it fits that way already
int near_levels(const void* split_kernel, int depth, int block_threads, int device)
{
    hipFuncAttributes a{};
    int lds_per_cu = 0;
    if (hipFuncGetAttributes(&a, split_kernel) != hipSuccess ||
        hipDeviceGetAttribute(&lds_per_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, device) != hipSuccess ||
        lds_per_cu <= 0)
    {
        return depth;
    }
    const int registers = std::max(8, (a.numRegs + 7) / 8 * 8);
    const int waves_per_cu = 4 * std::min(8, 512 / registers);
    const int blocks_per_cu = std::max(1, waves_per_cu / (block_threads / 64));
    const int levels = lds_per_cu / blocks_per_cu / (block_threads * 6);
    return levels >= depth ? depth : std::max(8, levels);
}
}

bool megakernel_waves_supported(int waves) { return waves == 0 || (waves >= 5 && waves <= 7); }

size_t megakernel_staging_per_pass(int width, int height)
{
    return static_cast<size_t>(width) * static_cast<size_t>(height);
}

size_t megakernel_far_stack_bytes(int depth)
{
    int device = 0, cus = 0;
    if (hipGetDevice(&device) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
    {
        return 0;
    }
    // at most 32 waves a CU, 6 bytes a level
    return static_cast<size_t>(depth) * static_cast<size_t>(cus) * 32 * 64 * 6;
}

// Passes in a queue chunk: a chunk is 64 x 2 paths for a wave, and a launch of C2 has ~530 K chunks (groups of 2
// measured 1 % faster on C2 than 4, and within 1 % on C1 and C3; 8 and 16 6 % and 11 % slower)
constexpr int megakernel_pass_group = 2;

template <typename Real>
hipError_t launch_megakernel(int mode, const dev::camera_rec<Real>& cam, const dev::scene_view<Real>& sc,
                             uint64_t seed, const dev::launch_work& work, double* sums, void* staging,
                             size_t staging_samples, uint32_t* queue, unsigned long long* rays, int block_threads,
                             hipStream_t stream, unsigned long long* tallies, int waves, void* far_buffer,
                             size_t far_bytes, int grid_ways)
{
    if (block_threads % 64 != 0 || grid_ways < 1 ||
        block_threads < 64 || block_threads > 256 || work.pass_count < 1 || work.row_count < 1 ||
        staging_samples > static_cast<size_t>(INT32_MAX))
    {
        return hipErrorInvalidValue;
    }
    // the staging buffer holds a sample per pixel of the launch's rows per pass: a launch of more passes goes as several
    const size_t pixels = static_cast<size_t>(cam.image_width) * static_cast<size_t>(work.row_count);
    const int chunk = static_cast<int>(std::min<size_t>(staging_samples / pixels, static_cast<size_t>(work.pass_count)));
    if (chunk < 1)
    {
        return hipErrorInvalidValue;
    }
    size_t lds = stack_lds_bytes<Real>(sc.stack_depth, block_threads);
    const int walk = dev::tree_walk(sc);
    const bool marble = sc.noises != nullptr;
    const void* f = tallies ? megakernel_for<Real, true>(mode, 0, false, walk, marble)
                            : megakernel_for<Real, false>(mode, waves, false, walk, marble);
    // as many blocks as fit the chip at once
    int device = 0, cus = 0, per_cu = 0;
    hipError_t e = hipGetDevice(&device);
    if (e == hipSuccess)
    {
        e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
    }
    if (e == hipSuccess)
    {
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, block_threads, lds);
    }
    if (e != hipSuccess)
    {
        return e;
    }
    // the quality megakernel on a deep tree: the stack's deeper levels in HBM when the whole stack in LDS would leave
    // room for fewer waves than the registers allow, and the buffer holds them for every thread of the grid
    dev::far_stack far{sc.stack_depth, nullptr, nullptr, 0};
    if (!tallies && mode == dev::mode_quality && far_buffer != nullptr && sc.stack_depth > 0)
    {
        const void* split = megakernel_for<Real, false>(mode, waves, true, walk, marble);
        const int near = near_levels(split, sc.stack_depth, block_threads, device);
        const size_t split_lds = stack_lds_bytes<Real>(near, block_threads);
        int split_per_cu = 0;
        // (worth it for a quarter more waves or better: a split costs a branch at every push and pop, and one wave
        // more per CU measured slower on C3 than the whole stack in LDS)
        if (near < sc.stack_depth &&
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&split_per_cu, split, block_threads, split_lds) == hipSuccess &&
            4 * split_per_cu >= 5 * per_cu)
        {
            const size_t stride = static_cast<size_t>(split_per_cu) * static_cast<size_t>(cus) * block_threads;
            const size_t levels = static_cast<size_t>(sc.stack_depth - near);
            if (levels * stride * 6 <= far_bytes && stride <= UINT32_MAX)
            {
                f = split;
                lds = split_lds;
                per_cu = split_per_cu;
                far = {near, static_cast<int32_t*>(far_buffer),
                       reinterpret_cast<uint16_t*>(static_cast<char*>(far_buffer) + levels * stride * 4),
                       static_cast<uint32_t>(stride)};
            }
        }
    }
    const uint32_t tiles_x = static_cast<uint32_t>((cam.image_width + 15) / 16);
    const uint32_t tiles = tiles_x * static_cast<uint32_t>((work.row_count + 3) / 4);
    auto* staged = static_cast<dev::staged_sample<Real>*>(staging);
    for (int done = 0; done < work.pass_count; done += chunk)
    {
        dev::launch_work w = work;
        w.first_pass = work.first_pass + done;
        w.pass_count = std::min(chunk, work.pass_count - done);
        const int group = std::min(megakernel_pass_group, w.pass_count);
        const uint32_t groups = static_cast<uint32_t>((w.pass_count + group - 1) / group);
        dev::chunk_queue q{queue, tiles_x, tiles, tiles * groups, group, cam.image_width, w};
        const uint32_t waves_per_block = static_cast<uint32_t>(block_threads / 64);
        const uint32_t blocks = std::max<uint32_t>(
            1, std::min<uint32_t>(static_cast<uint32_t>(std::max(1, per_cu * cus / grid_ways)),
                                  (q.chunks + waves_per_block - 1) / waves_per_block));
        hipLaunchKernelGGL(dev::clear_queue, dim3(1), dim3(1), 0, stream, queue);
        dev::megakernel_args<Real> a{cam, sc, q, seed};
        void* args[] = {&a, &staged, &rays, &tallies, &far};
        void* book_args[] = {&a, &staged, &rays, &tallies}; // the books' kernels take no far_stack
        e = hipLaunchKernel(f, dim3(blocks), dim3(block_threads), mode == dev::mode_quality ? args : book_args, lds,
                            stream);
        if (e != hipSuccess)
        {
            return e;
        }
        hipLaunchKernelGGL(dev::add_staged<Real>, dim3(static_cast<unsigned>((pixels + 255) / 256)), dim3(256), 0,
                           stream, staged, sums, cam.image_width, w);
        e = hipGetLastError();
        if (e != hipSuccess)
        {
            return e;
        }
    }
    return hipSuccess;
}

namespace
{
template <typename Real>
const void* plain_megakernel_for(int mode, int walk)
{
    if (mode == dev::mode_quality)
    {
        if constexpr (sizeof(Real) == sizeof(float))
        {
            if (walk == dev::walk_packed)
            {
                return reinterpret_cast<const void*>(&dev::plain_megakernel<float, dev::mode_quality, dev::walk_packed>);
            }
        }
        return walk == dev::walk_instances
                   ? reinterpret_cast<const void*>(&dev::plain_megakernel<Real, dev::mode_quality, dev::walk_instances>)
                   : reinterpret_cast<const void*>(&dev::plain_megakernel<Real, dev::mode_quality, dev::walk_all>);
    }
    return mode == dev::mode_all_books
               ? reinterpret_cast<const void*>(&dev::plain_megakernel<Real, dev::mode_all_books, dev::walk_books>)
               : reinterpret_cast<const void*>(&dev::plain_megakernel<Real, dev::mode_first_book, dev::walk_spheres>);
}
}

template <typename Real>
hipError_t launch_plain_megakernel(int mode, const dev::camera_rec<Real>& cam, const dev::scene_view<Real>& sc,
                                   uint64_t seed, const dev::launch_work& work, double* sums, unsigned long long* rays,
                                   hipStream_t stream)
{
    if (work.pass_count < 1 || work.row_count < 1)
    {
        return hipErrorInvalidValue;
    }
    const dim3 block(plain_block_x, plain_block_y);
    const dim3 grid(static_cast<unsigned>((cam.image_width + plain_block_x - 1) / plain_block_x),
                    static_cast<unsigned>((work.row_count + plain_block_y - 1) / plain_block_y));
    const size_t lds = stack_lds_bytes<Real>(sc.stack_depth, plain_block_x * plain_block_y);
    dev::plain_args<Real> a{cam, sc, work, seed};
    void* args[] = {&a, &sums, &rays};
    return hipLaunchKernel(plain_megakernel_for<Real>(mode, dev::tree_walk(sc)), grid, block, args, lds, stream);
}

template <typename Real>
kernel_attributes plain_megakernel_attributes(int mode, int depth, int walk)
{
    hipFuncAttributes a{};
    const void* f = plain_megakernel_for<Real>(mode, walk);
    kernel_attributes k{};
    if (hipFuncGetAttributes(&a, f) == hipSuccess)
    {
        k.registers = a.numRegs;
        k.local_bytes = static_cast<int>(a.localSizeBytes);
    }
    k.lds_bytes = static_cast<int>(stack_lds_bytes<Real>(depth, plain_block_x * plain_block_y));
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, f, plain_block_x * plain_block_y, k.lds_bytes) ==
        hipSuccess)
    {
        k.blocks_per_cu = blocks;
    }
    return k;
}

template hipError_t launch_plain_megakernel<float>(int, const dev::camera_rec<float>&, const dev::scene_view<float>&,
                                                   uint64_t, const dev::launch_work&, double*, unsigned long long*,
                                                   hipStream_t);
template hipError_t launch_plain_megakernel<double>(int, const dev::camera_rec<double>&,
                                                    const dev::scene_view<double>&, uint64_t, const dev::launch_work&,
                                                    double*, unsigned long long*, hipStream_t);
template kernel_attributes plain_megakernel_attributes<float>(int, int, int);
template kernel_attributes plain_megakernel_attributes<double>(int, int, int);

template <typename Real>
hipError_t launch_closest_hit(const dev::scene_view<Real>& sc, const Real* origins, const Real* directions,
                              const Real* times, const int32_t* from, const uint8_t* inward, int64_t n, int32_t* index,
                              Real* t, uint8_t* entering, hipStream_t stream)
{
    const int threads = 256;
    const dim3 grid(static_cast<unsigned>((n + threads - 1) / threads));
    const size_t lds = stack_lds_bytes<Real>(sc.stack_depth, threads);
    if constexpr (sizeof(Real) == sizeof(float))
    {
        if (dev::tree_walk(sc) == dev::walk_packed)
        {
            hipLaunchKernelGGL((dev::closest_hit_kernel<float, dev::walk_packed>), grid, dim3(threads), lds, stream, sc,
                               origins, directions, times, from, inward, n, index, t, entering);
            return hipGetLastError();
        }
    }
    hipLaunchKernelGGL((dev::closest_hit_kernel<Real, dev::walk_instances>), grid, dim3(threads), lds, stream, sc,
                       origins, directions, times, from, inward, n, index, t, entering);
    return hipGetLastError();
}

hipError_t launch_display(const double* sums, int64_t pixels, double inverse_passes, uint8_t* rgba, hipStream_t stream)
{
    const int threads = 256;
    const dim3 grid(static_cast<unsigned>((pixels + threads - 1) / threads));
    hipLaunchKernelGGL(dev::display_kernel, grid, dim3(threads), 0, stream, sums, pixels, inverse_passes, rgba);
    return hipGetLastError();
}

template <typename Real>
kernel_attributes megakernel_attributes(int mode, int block_threads, int depth, int waves, int walk, bool marble)
{
    hipFuncAttributes a{};
    const void* f = megakernel_for<Real, false>(mode, waves, false, walk, marble);
    kernel_attributes k{};
    if (hipFuncGetAttributes(&a, f) == hipSuccess)
    {
        k.registers = a.numRegs;
        k.local_bytes = static_cast<int>(a.localSizeBytes);
    }
    k.lds_bytes = static_cast<int>(stack_lds_bytes<Real>(depth, block_threads));
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, f, block_threads, k.lds_bytes) == hipSuccess)
    {
        k.blocks_per_cu = blocks;
    }
    return k;
}

template hipError_t launch_megakernel<float>(int, const dev::camera_rec<float>&, const dev::scene_view<float>&,
                                             uint64_t, const dev::launch_work&, double*, void*, size_t, uint32_t*,
                                             unsigned long long*, int, hipStream_t, unsigned long long*, int, void*,
                                             size_t, int);
template hipError_t launch_megakernel<double>(int, const dev::camera_rec<double>&, const dev::scene_view<double>&,
                                              uint64_t, const dev::launch_work&, double*, void*, size_t, uint32_t*,
                                              unsigned long long*, int, hipStream_t, unsigned long long*, int, void*,
                                              size_t, int);
template hipError_t launch_closest_hit<float>(const dev::scene_view<float>&, const float*, const float*, const float*,
                                              const int32_t*, const uint8_t*, int64_t, int32_t*, float*, uint8_t*,
                                              hipStream_t);
template kernel_attributes megakernel_attributes<float>(int, int, int, int, int, bool);
template kernel_attributes megakernel_attributes<double>(int, int, int, int, int, bool);
template size_t stack_lds_bytes<float>(int, int);
template size_t stack_lds_bytes<double>(int, int);
}

namespace rt::dev
{
// A plain copy, one 16-byte word per thread per step: the calibration kernel of the HBM counters
__global__ void __launch_bounds__(256) copy_kernel(const uint4* __restrict__ in, uint4* __restrict__ out, size_t n)
{
    for (size_t k = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; k < n;
         k += static_cast<size_t>(gridDim.x) * blockDim.x)
    {
        out[k] = in[k];
    }
}
}

namespace rt
{
hipError_t launch_copy(const void* in, void* out, size_t bytes, hipStream_t stream)
{
    const size_t n = bytes / 16;
    const unsigned blocks = static_cast<unsigned>(std::min<size_t>((n + 255) / 256, 256 * 64));
    hipLaunchKernelGGL(dev::copy_kernel, dim3(blocks), dim3(256), 0, stream, static_cast<const uint4*>(in),
                       static_cast<uint4*>(out), n);
    return hipGetLastError();
}
}
