// The megakernel: one thread per pixel traces the launch's passes of its pixel, one ray per loop iteration, and adds
// each finished sample to the pixel's double sums in pass order (src/gpu/renderer.cu:340-397). It is the plain
// baseline the wavefront integrator is measured against, and the path the FP64 renders take.
//
// MI355X mapping: a block of 16x8 pixels is two wave64s of 16x4; the BVH stack lives in LDS, one column of entries per
// thread (path_tracer.hpp, lds_stack), sized by the scene's tree depth at launch; the ray counts are summed across the
// wave with DPP/permute shuffles, then one atomic per wave.
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

// Waves: the waves per SIMD the register allocation must leave room for (0: the compiler's choice); a tighter budget
// trades spills for occupancy, and the session's autotuner picks per scene
template <typename Real, bool AllBooks, bool Count, int Waves>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(Waves > 0 ? Waves : 1, 8)))
megakernel(const camera_rec<Real> cam, const scene_view<Real> sc,
                                                  const uint64_t seed, const launch_work work, double* __restrict__ sums,
                                                  unsigned long long* __restrict__ ray_count,
                                                  unsigned long long* __restrict__ tallies)
{
    const lds_stack stack = block_stack(sc.stack_depth);
    tally<Count> tl;
    const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    const int j = work.first_row + static_cast<int>(blockIdx.y * blockDim.y + threadIdx.y);
    unsigned long long rays = 0;
    if (i < cam.image_width && j < work.first_row + work.row_count)
    {
        const uint64_t pixel = static_cast<uint64_t>(j) * static_cast<uint64_t>(cam.image_width) + static_cast<uint64_t>(i);
        double* const sum = sums + 3 * pixel;
        double red = sum[0], green = sum[1], blue = sum[2];
        int pass = work.first_pass;
        const int end_pass = work.first_pass + work.pass_count;
        pcg32 g = sample_generator(seed, pixel, static_cast<uint64_t>(pass));
        path<Real> p = start_path(cam, AllBooks && sc.motion != 0, i, j, pass, g);
        while (true)
        {
            ++rays;
            vec3<Real> sample;
            const bool done = AllBooks ? trace_all_books(p, sc, cam, g, sample, stack, tl)
                                       : trace_first_book(p, sc, cam, g, sample, stack, tl);
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
            g = sample_generator(seed, pixel, static_cast<uint64_t>(pass));
            p = start_path(cam, AllBooks && sc.motion != 0, i, j, pass, g);
        }
        sum[0] = red;
        sum[1] = green;
        sum[2] = blue;
    }
    add_counts<Count>(rays, 0, ray_count, tallies, tl);
}

// The quality integrator's megakernel (src/gpu/quality.cu:53-109): the same loop over the launch's passes, with light
// samples and their shadow rays at every vertex that isn't a mirror's (quality.hpp)
template <bool Count, int Waves>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(Waves > 0 ? Waves : 1, 8)))
quality_megakernel(const camera_rec<float> cam, const scene_view<float> sc,
                                                          const uint64_t seed, const launch_work work,
                                                          double* __restrict__ sums,
                                                          unsigned long long* __restrict__ ray_count,
                                                          unsigned long long* __restrict__ tallies)
{
    const lds_stack stack = block_stack(sc.stack_depth);
    tally<Count> tl;
    const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    const int j = work.first_row + static_cast<int>(blockIdx.y * blockDim.y + threadIdx.y);
    unsigned long long rays = 0;
    uint32_t shadow_rays = 0;
    if (i < cam.image_width && j < work.first_row + work.row_count)
    {
        const uint64_t pixel = static_cast<uint64_t>(j) * static_cast<uint64_t>(cam.image_width) + static_cast<uint64_t>(i);
        double* const sum = sums + 3 * pixel;
        double red = sum[0], green = sum[1], blue = sum[2];
        int pass = work.first_pass;
        const int end_pass = work.first_pass + work.pass_count;
        pcg32 g = sample_generator(seed, pixel, static_cast<uint64_t>(pass));
        quality_path q = start_quality(cam, sc.motion != 0, i, j, pass, g);
        while (true)
        {
            ++rays;
            vec3<float> sample;
            if (!trace_quality(q, sc, cam, g, sample, stack, tl, shadow_rays))
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
            g = sample_generator(seed, pixel, static_cast<uint64_t>(pass));
            q = start_quality(cam, sc.motion != 0, i, j, pass, g);
        }
        sum[0] = red;
        sum[1] = green;
        sum[2] = blue;
    }
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

template <typename Real>
hipError_t launch_megakernel(int mode, const dev::camera_rec<Real>& cam, const dev::scene_view<Real>& sc,
                             uint64_t seed, const dev::launch_work& work, double* sums, unsigned long long* rays,
                             int block_w, int block_h, hipStream_t stream, unsigned long long* tallies, int waves)
{
    if (mode == dev::mode_quality && sizeof(Real) != sizeof(float))
    {
        return hipErrorInvalidValue;
    }
    const dim3 block(block_w, block_h);
    const dim3 grid((cam.image_width + block_w - 1) / block_w, (work.row_count + block_h - 1) / block_h);
    const size_t lds = stack_lds_bytes<Real>(sc.stack_depth, block_w * block_h);
    const void* f = tallies ? megakernel_for<Real, true>(mode, 0) : megakernel_for<Real, false>(mode, waves);
    dev::camera_rec<Real> c = cam;
    dev::scene_view<Real> v = sc;
    uint64_t s = seed;
    dev::launch_work w = work;
    void* args[] = {&c, &v, &s, &w, &sums, &rays, &tallies};
    return hipLaunchKernel(f, grid, block, args, lds, stream);
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
                                             uint64_t, const dev::launch_work&, double*, unsigned long long*, int, int,
                                             hipStream_t, unsigned long long*, int);
template hipError_t launch_megakernel<double>(int, const dev::camera_rec<double>&, const dev::scene_view<double>&,
                                              uint64_t, const dev::launch_work&, double*, unsigned long long*, int, int,
                                              hipStream_t, unsigned long long*, int);
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
