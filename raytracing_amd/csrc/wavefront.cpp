// The wavefront integrator's host driver (see wavefront.hpp)
#include "wavefront.hpp"

#include <algorithm>

namespace rt
{
template <typename Real>
wavefront<Real>::wavefront(const dev::scene_view<Real>& sc, const dev::camera_rec<Real>& cam, int mode,
                           const rt_session_options& options, hipStream_t stream)
    : stream_(stream), mode_(mode), quality_(mode == dev::mode_quality)
{
    if constexpr (!wavefront_supported<Real>())
    {
        throw std::invalid_argument("the wavefront integrator renders in float32 only");
    }
    else
    {
        sc_ = sc;
        cam_ = cam;
        const int slots = options.wavefront_slots;
        const int graph_bounces = options.wavefront_graph;
        if (slots < 0 || options.wavefront_drain < -1)
        {
            throw std::invalid_argument("the wavefront's slots must be 0 or more, its drain count -1 or more");
        }
        if (graph_bounces < -1 || graph_bounces > 64)
        {
            throw std::invalid_argument("the wavefront's bounces per graph must be -1 (none), 0 (default) or 1 to 64");
        }
        use_graph_ = graph_bounces != -1;
        bounces_per_graph_ = graph_bounces > 0 ? 2 * ((graph_bounces + 1) / 2) : 8; // even: each replay ends on parity 0
        const uint64_t pixels = static_cast<uint64_t>(cam.image_width) * static_cast<uint64_t>(cam.image_height);
        const uint64_t want = slots > 0 ? static_cast<uint64_t>(slots) : pixels;
        if (want < static_cast<uint64_t>(cam.image_width))
        {
            throw std::invalid_argument("the wavefront pool needs at least one row of slots");
        }
        if (pixels > 0x7fffffffull)
        {
            throw std::invalid_argument("the image has too many pixels for the wavefront pool");
        }
        // whole rows: a launch's band of rows must fit the pool
        capacity_ = static_cast<uint32_t>(std::min(want, pixels) / cam.image_width * cam.image_width);
        drain_below_ = options.wavefront_drain > 0    ? static_cast<uint32_t>(options.wavefront_drain)
                       : options.wavefront_drain == 0 ? capacity_ / 16
                                                      : 0;
        ray_o_ = device_array<float4>(capacity_);
        ray_d_ = device_array<float4>(capacity_);
        thr_ = device_array<float4>(capacity_);
        if (quality_)
        {
            qst_ = device_array<float4>(capacity_);
        }
        rng_ = device_array<uint4>(capacity_);
        hit_ = device_array<int4>(capacity_);
        item_ = device_array<uint32_t>(capacity_);
        // a launch renders at most max_passes_ passes at a time: their samples wait in a buffer of at most 1 GiB
        max_passes_ = static_cast<int>(std::clamp<uint64_t>((1ull << 30) / (16ull * capacity_), 1, 64));
        samples_ = device_array<float4>(static_cast<std::size_t>(max_passes_) * capacity_);
        // each list is wf_stripes segments, each as long as the items a stripe's blocks can append; the active list
        // has the refill segment after them
        const uint32_t blocks = wf_blocks(capacity_);
        const uint32_t segment = (blocks + dev::wf_stripes - 1) / dev::wf_stripes * dev::wf_block;
        const std::size_t list = static_cast<std::size_t>(segment) * dev::wf_stripes;
        active0_ = device_array<uint32_t>(list + capacity_);
        active1_ = device_array<uint32_t>(list + capacity_);
        free_ = device_array<uint32_t>(list);
        queue_ = device_array<uint32_t>(list * dev::wf_classes);
        counters_ = device_array<uint32_t>(dev::wf_counters);
        prefix_ = device_array<uint32_t>(dev::wf_prefix_size);
        tallies_ = device_array<unsigned long long>(2 * dev::wf_stripes);
        launch_ = device_array<dev::wf_launch>(1);
        check(hipMemsetAsync(counters_.data(), 0, counters_.bytes(), stream_), "clearing counters");
        check(hipMemsetAsync(tallies_.data(), 0, tallies_.bytes(), stream_), "clearing counters");
        check(hipHostMalloc(reinterpret_cast<void**>(&host_counters_), dev::wf_counters * sizeof(uint32_t),
                            hipHostMallocDefault),
              "allocating pinned memory");
        check(hipHostMalloc(reinterpret_cast<void**>(&host_launch_), sizeof(dev::wf_launch), hipHostMallocDefault),
              "allocating pinned memory");
        pool_ = {launch_.data(), ray_o_.data(), ray_d_.data(), thr_.data(), qst_.size() > 0 ? qst_.data() : nullptr, rng_.data(),
                 hit_.data(), item_.data(), samples_.data(), {active0_.data(), active1_.data()}, free_.data(),
                 queue_.data(), counters_.data(), prefix_.data(), tallies_.data(), capacity_, segment};
        bytes_ = ray_o_.bytes() * 3 + qst_.bytes() + rng_.bytes() + hit_.bytes() + item_.bytes() + samples_.bytes() +
                 active0_.bytes() * 2 + free_.bytes() + queue_.bytes() + counters_.bytes() + prefix_.bytes() +
                 tallies_.bytes() + launch_.bytes();
    }
}

template <typename Real>
wavefront<Real>::~wavefront()
{
    (void)hipStreamSynchronize(stream_); // a launch that threw may still run the graph
    if (graph_exec_)
    {
        (void)hipGraphExecDestroy(graph_exec_);
    }
    if (graph_)
    {
        (void)hipGraphDestroy(graph_);
    }
    if (host_counters_)
    {
        (void)hipHostFree(host_counters_);
    }
    if (host_launch_)
    {
        (void)hipHostFree(host_launch_);
    }
}

// One replay's bounces: an even number (parity 0, 1, 0, 1...), so that each replay ends on parity 0
template <typename Real>
hipError_t wavefront<Real>::launch_bounces()
{
    hipError_t launched = hipSuccess;
    for (int b = 0; b < bounces_per_graph_ && launched == hipSuccess; ++b)
    {
        launched = wf_launch_bounce(cam_, sc_, pool_, b & 1, mode_, stream_);
    }
    return launched;
}

template <typename Real>
int64_t wavefront<Real>::render(uint64_t seed, const dev::launch_work& work, double* sums, unsigned long long* rays)
{
    if constexpr (!wavefront_supported<Real>())
    {
        return 0;
    }
    else
    {
        const uint64_t pixels = static_cast<uint64_t>(work.row_count) * static_cast<uint64_t>(cam_.image_width);
        if (pixels > capacity_)
        {
            throw std::invalid_argument("the launch's rows don't fit the wavefront pool");
        }
        int64_t done = 0;
        // at most max_passes_ passes at a time: their samples fit the buffer
        for (int first = 0; first < work.pass_count; first += max_passes_)
        {
            dev::launch_work part = work;
            part.first_pass = work.first_pass + first;
            part.pass_count = std::min(max_passes_, work.pass_count - first);
            done += render_part(seed, part, sums, rays);
        }
        return done;
    }
}

template <typename Real>
int64_t wavefront<Real>::render_part(uint64_t seed, const dev::launch_work& work, double* sums,
                                     unsigned long long* rays)
{
    const uint32_t items =
        static_cast<uint32_t>(work.pass_count) * static_cast<uint32_t>(work.row_count) * static_cast<uint32_t>(cam_.image_width);
    // the launch's seed and work go to device memory before its first kernel; the previous launch has ended (every
    // render_part waits for its work), so the pinned record is free to rewrite
    host_launch_->seed = seed;
    host_launch_->work = work;
    check(hipMemcpyAsync(launch_.data(), host_launch_, sizeof(dev::wf_launch), hipMemcpyHostToDevice, stream_),
          "starting the wavefront's launch");
    check(wf_launch_init(pool_, items, stream_), "starting the wavefront's launch");
    // the bounce loop, captured at the pool's first launch (or launched kernel by kernel), replayed until the items
    // are all handed out and no slot is active
    if (use_graph_ && !graph_exec_)
    {
        check(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal), "capturing the bounce loop");
        const hipError_t launched = launch_bounces();
        const hipError_t ended = hipStreamEndCapture(stream_, &graph_);
        check(launched, "launching the bounce loop's kernels");
        check(ended, "capturing the bounce loop");
        check(hipGraphInstantiate(&graph_exec_, graph_, nullptr, nullptr, 0), "instantiating the bounce loop");
    }
    // every bounce either advances a path or hands out items to every free slot: a safe bound on the rounds
    const long long max_bounces =
        (static_cast<long long>(items) / capacity_ + 2) * static_cast<long long>(cam_.max_depth + 1) * 2 + 64;
    int64_t done = 0;
    while (true)
    {
        if (done > max_bounces)
        {
            throw std::runtime_error("the wavefront loop didn't drain");
        }
        if (use_graph_)
        {
            check(hipGraphLaunch(graph_exec_, stream_), "launching the bounce loop");
        }
        else
        {
            check(launch_bounces(), "launching the bounce loop's kernels");
        }
        done += bounces_per_graph_;
        check(hipMemcpyAsync(host_counters_, counters_.data(), dev::wf_counters * sizeof(uint32_t),
                             hipMemcpyDeviceToHost, stream_),
              "reading the wavefront's counters");
        check(hipStreamSynchronize(stream_), "rendering on the GPU");
        uint64_t active = 0; // the parity-0 active list's stripes: the paths going on into the next round
        for (int k = 0; k < dev::wf_stripes; ++k)
        {
            active += host_counters_[dev::wf_counter_active + k];
        }
        const bool handed_out = host_counters_[dev::wf_counter_next] == host_counters_[dev::wf_counter_items];
        if (handed_out && active == 0)
        {
            break;
        }
        if (handed_out && active < drain_below_)
        {
            check(wf_launch_drain(mode_, cam_, sc_, seed, work, pool_, static_cast<uint32_t>(active), stream_),
                  "launching the wavefront's drain");
            break;
        }
    }
    check(wf_launch_finish(cam_, work, pool_, sums, rays, stream_), "adding the launch's samples");
    check(hipStreamSynchronize(stream_), "rendering on the GPU");
    return done;
}

template <typename Real>
kernel_attributes wavefront<Real>::attributes() const
{
    kernel_attributes a = wf_attributes(sc_.stack_depth, 0);
    const kernel_attributes s = wf_attributes(sc_.stack_depth, quality_ ? 2 : 1);
    a.registers = std::max(a.registers, s.registers);
    a.local_bytes = std::max(a.local_bytes, s.local_bytes);
    return a;
}

template class wavefront<float>;
template class wavefront<double>;
}
