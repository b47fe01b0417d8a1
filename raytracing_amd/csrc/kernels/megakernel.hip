// The megakernel: the launch's passes of a block of pixels, every path traced by one thread, one ray per loop
// iteration, and each pixel's samples added to its double sums in pass order (src/gpu/renderer.cu:340-397). It is the
// plain baseline the wavefront integrator is measured against, and the path the FP64 renders take.
//
// MI355X mapping: a block of 16x8 (or 16x16) pixels is two (four) wave64s of 16x4; a wave's lanes share its 64 pixels'
// passes as items (wave_share) and stage their samples until the wave's items are done; the BVH stack lives in LDS, one
// column of entries per thread (path_tracer.hpp, lds_stack), sized by the tree's worst case at launch; the ray counts
// are summed across the wave with shuffles, then one atomic per wave.
#include <algorithm>

#include "kernel_api.hpp"
#include "path_tracer.hpp"
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

__device__ inline lds_stack block_stack(int entries)
{
    extern __shared__ __align__(16) unsigned char lds[];
    return stack_in(lds, entries, blockDim.x * blockDim.y, threadIdx.y * blockDim.x + threadIdx.x);
}

static_assert(tally_events == RT_TRAFFIC_EVENTS, "the traffic report lists every tally event");

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
    }
}

// A wave's share of a launch. Its slots are its 64 threads' pixels that lie in the image and the launch's rows; its
// items are the launch's passes over those slots, pass-major: item k is pass k / count of the (k % count)-th slot. A
// lane whose path ends claims the wave's next item (a ballot and a popcount: the counter is the wave's own, no
// atomic), so the lanes stay busy while the wave has items, however long their pixels' paths are; a lane no longer
// waits for the longest pixel of its wave. Each item's sample goes to the wave's own staging region (no other wave
// touches its cache lines); when the wave's items are done, each lane adds its own pixel's samples to the double sums
// in pass order: the same additions in the same order as one thread looping over its pixel's passes, so the image is
// the same bytes.
struct wave_share
{
    int lane;         // this thread's lane
    int first;        // the wave's first thread in the block
    uint64_t below;   // the lanes below this one
    uint32_t count;   // slots in the image
    uint32_t total;   // items
    uint32_t next;    // the wave's next unclaimed item (the same in every lane)
    const uint8_t* slots; // the wave's slots, in lane order
    bool inside;      // this thread's own pixel is a slot
    int i, j;         // this thread's own pixel
    size_t region;    // the wave's staging region: 64 samples per pass, pass-major

    // Claims an item for every lane that has none (needy); true with this lane's item in (pass_local, slot) when it
    // got one, false when the wave's items are all claimed
    __device__ bool claim(bool needy, uint32_t& pass_local, int& slot)
    {
        const uint64_t ask = __ballot(needy);
        const uint32_t k = next + static_cast<uint32_t>(__popcll(ask & below));
        next += static_cast<uint32_t>(__popcll(ask));
        if (!needy || k >= total)
        {
            return false;
        }
        pass_local = k / count;
        slot = slots[k - pass_local * count];
        return true;
    }
    // A slot's pixel
    __device__ void pixel_of(int slot, const launch_work& work, int& si, int& sj) const
    {
        const int t = first + slot;
        si = static_cast<int>(blockIdx.x * blockDim.x) + t % static_cast<int>(blockDim.x);
        sj = work.first_row + static_cast<int>(blockIdx.y * blockDim.y) + t / static_cast<int>(blockDim.x);
    }
};

template <typename Real>
__device__ inline wave_share share_of(const camera_rec<Real>& cam, const launch_work& work)
{
    __shared__ uint8_t slot_tables[4][64]; // up to 256 threads: four waves
    wave_share w;
    const int tid = static_cast<int>(threadIdx.y * blockDim.x + threadIdx.x);
    w.lane = tid & 63;
    w.first = tid - w.lane;
    w.below = (uint64_t{1} << w.lane) - 1;
    w.i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    w.j = work.first_row + static_cast<int>(blockIdx.y * blockDim.y + threadIdx.y);
    w.inside = w.i < cam.image_width && w.j < work.first_row + work.row_count;
    const uint64_t valid = __ballot(w.inside);
    uint8_t* const table = slot_tables[tid >> 6];
    if (w.inside)
    {
        table[__popcll(valid & w.below)] = static_cast<uint8_t>(w.lane);
    }
    __syncthreads();
    w.slots = table;
    w.count = static_cast<uint32_t>(__popcll(valid));
    w.total = w.count * static_cast<uint32_t>(work.pass_count);
    w.next = 0;
    const size_t waves_per_block = blockDim.x * blockDim.y / 64;
    w.region = ((static_cast<size_t>(blockIdx.y) * gridDim.x + blockIdx.x) * waves_per_block + (tid >> 6)) * 64 *
               static_cast<size_t>(work.pass_count);
    return w;
}

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

// Adds this lane's own pixel's staged samples to its double sums, in pass order (after the wave's loop: every lane of
// the wave is past its last store)
template <typename Real>
__device__ inline void add_samples(const wave_share& w, const camera_rec<Real>& cam, const launch_work& work,
                                   const staged_sample<Real>* __restrict__ staged, double* __restrict__ sums)
{
    __threadfence_block();
    if (!w.inside)
    {
        return;
    }
    double* const sum = sums + 3 * (static_cast<uint64_t>(w.j) * static_cast<uint64_t>(cam.image_width) +
                                    static_cast<uint64_t>(w.i));
    double red = sum[0], green = sum[1], blue = sum[2];
    for (int q = 0; q < work.pass_count; ++q)
    {
        const staged_sample<Real> s = unstage(staged + static_cast<size_t>(q) * 64 + w.lane);
        red += without_nan(s.x);
        green += without_nan(s.y);
        blue += without_nan(s.z);
    }
    sum[0] = red;
    sum[1] = green;
    sum[2] = blue;
}

// Waves: the waves per SIMD the register allocation must leave room for (0: the compiler's choice); a tighter budget
// trades spills for occupancy, and the session's autotuner picks per scene. Each lane follows its path ray by ray
// (src/gpu/renderer.cu:340-397), and a lane whose path ends starts the wave's next item (wave_share). (Walking the
// tree in bursts of a few steps per iteration, so that a lane whose walk is over shades while a long walk goes on,
// measured slower here: the walk's state then stays live across the shading code, which costs registers.)
template <typename Real, bool AllBooks, bool Count, int Waves>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(Waves > 0 ? Waves : 1, 8)))
megakernel(const camera_rec<Real> cam, const scene_view<Real> sc, const uint64_t seed, const launch_work work,
           double* __restrict__ sums, staged_sample<Real>* __restrict__ staging,
           unsigned long long* __restrict__ ray_count, unsigned long long* __restrict__ tallies)
{
    const lds_stack stack = block_stack(sc.stack_depth);
    tally<Count> tl;
    wave_share w = share_of(cam, work);
    staged_sample<Real>* const staged = staging + w.region;
    unsigned long long rays = 0;
    int at = -1; // where this lane's path's sample goes in the wave's region, or -1 between items
    path<Real> p{};
    pcg32 g{};
    while (true)
    {
        uint32_t pass_local;
        int slot;
        if (w.claim(at < 0, pass_local, slot))
        {
            int si, sj;
            w.pixel_of(slot, work, si, sj);
            const int pass = work.first_pass + static_cast<int>(pass_local);
            const uint64_t pixel =
                static_cast<uint64_t>(sj) * static_cast<uint64_t>(cam.image_width) + static_cast<uint64_t>(si);
            g = sample_generator(seed, pixel, static_cast<uint64_t>(pass));
            p = start_path(cam, AllBooks && sc.motion != 0, si, sj, pass, g);
            at = static_cast<int>(pass_local) * 64 + slot;
        }
        else if (at < 0)
        {
            break;
        }
        ++rays;
        vec3<Real> sample;
        const bool done = AllBooks ? trace_all_books(p, sc, cam, g, sample, stack, tl)
                                   : trace_first_book(p, sc, cam, g, sample, stack, tl);
        if (done)
        {
            stage(staged + at, sample);
            at = -1;
        }
    }
    add_samples(w, cam, work, staged, sums);
    add_counts<Count>(rays, 0, ray_count, tallies, tl);
}

// The quality integrator's megakernel (src/gpu/quality.cu:53-109): the same loop over the wave's items, with light
// samples and their shadow rays at every vertex that isn't a mirror's (quality.hpp)
template <bool Count, int Waves>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(Waves > 0 ? Waves : 1, 8)))
quality_megakernel(const camera_rec<float> cam, const scene_view<float> sc, const uint64_t seed,
                   const launch_work work, double* __restrict__ sums, staged_sample<float>* __restrict__ staging,
                   unsigned long long* __restrict__ ray_count, unsigned long long* __restrict__ tallies)
{
    const lds_stack stack = block_stack(sc.stack_depth);
    tally<Count> tl;
    wave_share w = share_of(cam, work);
    staged_sample<float>* const staged = staging + w.region;
    unsigned long long rays = 0;
    uint32_t shadow_rays = 0;
    int at = -1;
    quality_path q{};
    pcg32 g{};
    while (true)
    {
        uint32_t pass_local;
        int slot;
        if (w.claim(at < 0, pass_local, slot))
        {
            int si, sj;
            w.pixel_of(slot, work, si, sj);
            const int pass = work.first_pass + static_cast<int>(pass_local);
            const uint64_t pixel =
                static_cast<uint64_t>(sj) * static_cast<uint64_t>(cam.image_width) + static_cast<uint64_t>(si);
            g = sample_generator(seed, pixel, static_cast<uint64_t>(pass));
            q = start_quality(cam, sc.motion != 0, si, sj, pass, g);
            at = static_cast<int>(pass_local) * 64 + slot;
        }
        else if (at < 0)
        {
            break;
        }
        ++rays;
        vec3<float> sample;
        if (trace_quality(q, sc, cam, g, sample, stack, tl, shadow_rays))
        {
            stage(staged + at, sample);
            at = -1;
        }
    }
    add_samples(w, cam, work, staged, sums);
    add_counts<Count>(rays, shadow_rays, ray_count, tallies, tl);
}

// The parity probe: closest_hit on a batch of given rays, with the traversal of the all-books kernel
template <typename Real>
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
    const hit<Real> h = closest_hit<Real, true>(sc, r, from[k], inward[k] != 0, stack, tl);
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

namespace rt
{
template <typename Real>
size_t stack_lds_bytes(int entries, int threads)
{
    return dev::stack_bytes(entries, threads);
}

namespace
{
template <typename Real, bool Count, int Waves>
const void* megakernel_for_waves(int mode)
{
    if constexpr (sizeof(Real) == sizeof(float))
    {
        if (mode == dev::mode_quality)
        {
            return reinterpret_cast<const void*>(&dev::quality_megakernel<Count, Waves>);
        }
    }
    return mode == dev::mode_all_books ? reinterpret_cast<const void*>(&dev::megakernel<Real, true, Count, Waves>)
                                       : reinterpret_cast<const void*>(&dev::megakernel<Real, false, Count, Waves>);
}

// The instance for a mode and an occupancy target: the counting instances and the FP64 ones at the compiler's choice
template <typename Real, bool Count>
const void* megakernel_for(int mode, int waves)
{
    if constexpr (!Count && sizeof(Real) == sizeof(float))
    {
        if (waves == 5)
        {
            return megakernel_for_waves<Real, Count, 5>(mode);
        }
        if (waves == 6)
        {
            return megakernel_for_waves<Real, Count, 6>(mode);
        }
    }
    return megakernel_for_waves<Real, Count, 0>(mode);
}
}

bool megakernel_waves_supported(int waves) { return waves == 0 || waves == 5 || waves == 6; }

size_t megakernel_staging_per_pass(int width, int height)
{
    const size_t w = (static_cast<size_t>(width) + 15) / 16 * 16;
    const size_t h = (static_cast<size_t>(height) + 15) / 16 * 16;
    return w * h;
}

template <typename Real>
hipError_t launch_megakernel(int mode, const dev::camera_rec<Real>& cam, const dev::scene_view<Real>& sc,
                             uint64_t seed, const dev::launch_work& work, double* sums, void* staging,
                             size_t staging_samples, unsigned long long* rays, int block_w, int block_h,
                             hipStream_t stream, unsigned long long* tallies, int waves)
{
    if ((mode == dev::mode_quality && sizeof(Real) != sizeof(float)) || block_w * block_h % 64 != 0 ||
        block_w * block_h > 256 || work.pass_count < 1 || work.row_count < 1)
    {
        return hipErrorInvalidValue;
    }
    const dim3 block(block_w, block_h);
    const dim3 grid((cam.image_width + block_w - 1) / block_w, (work.row_count + block_h - 1) / block_h);
    const size_t threads = static_cast<size_t>(grid.x) * grid.y * block.x * block.y;
    // the staging buffer holds `threads` samples per pass: a launch of more passes goes as several
    const int chunk = static_cast<int>(std::min<size_t>(staging_samples / threads, static_cast<size_t>(work.pass_count)));
    if (chunk < 1)
    {
        return hipErrorInvalidValue;
    }
    const size_t lds = stack_lds_bytes<Real>(sc.stack_depth, block_w * block_h);
    const void* f = tallies ? megakernel_for<Real, true>(mode, 0) : megakernel_for<Real, false>(mode, waves);
    for (int done = 0; done < work.pass_count; done += chunk)
    {
        dev::camera_rec<Real> c = cam;
        dev::scene_view<Real> v = sc;
        uint64_t s = seed;
        dev::launch_work w = work;
        w.first_pass = work.first_pass + done;
        w.pass_count = std::min(chunk, work.pass_count - done);
        auto* staged = static_cast<dev::staged_sample<Real>*>(staging);
        void* args[] = {&c, &v, &s, &w, &sums, &staged, &rays, &tallies};
        const hipError_t e = hipLaunchKernel(f, grid, block, args, lds, stream);
        if (e != hipSuccess)
        {
            return e;
        }
    }
    return hipSuccess;
}

template <typename Real>
hipError_t launch_closest_hit(const dev::scene_view<Real>& sc, const Real* origins, const Real* directions,
                              const Real* times, const int32_t* from, const uint8_t* inward, int64_t n, int32_t* index,
                              Real* t, uint8_t* entering, hipStream_t stream)
{
    const int threads = 256;
    const dim3 grid(static_cast<unsigned>((n + threads - 1) / threads));
    const size_t lds = stack_lds_bytes<Real>(sc.stack_depth, threads);
    hipLaunchKernelGGL((dev::closest_hit_kernel<Real>), grid, dim3(threads), lds, stream, sc, origins, directions,
                       times, from, inward, n, index, t, entering);
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
kernel_attributes megakernel_attributes(int mode, int block_threads, int depth, int waves)
{
    hipFuncAttributes a{};
    const void* f = megakernel_for<Real, false>(mode, waves);
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
                                             uint64_t, const dev::launch_work&, double*, void*, size_t,
                                             unsigned long long*, int, int, hipStream_t, unsigned long long*, int);
template hipError_t launch_megakernel<double>(int, const dev::camera_rec<double>&, const dev::scene_view<double>&,
                                              uint64_t, const dev::launch_work&, double*, void*, size_t,
                                              unsigned long long*, int, int, hipStream_t, unsigned long long*, int);
template hipError_t launch_closest_hit<float>(const dev::scene_view<float>&, const float*, const float*, const float*,
                                              const int32_t*, const uint8_t*, int64_t, int32_t*, float*, uint8_t*,
                                              hipStream_t);
template kernel_attributes megakernel_attributes<float>(int, int, int, int);
template kernel_attributes megakernel_attributes<double>(int, int, int, int);
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
