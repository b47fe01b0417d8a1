// The wavefront integrator's host driver: the pool of path slots in device memory, and the bounce loop, captured once
// per pool in a hipGraph of several bounces that is replayed until no slot is active (kernels in
// kernels/wavefront.hip)
#pragma once

#include <cstdint>
#include <memory>
#include <stdexcept>

#include <hip/hip_runtime.h>

#include "device_scene.hpp"
#include "device_types.hpp"
#include "kernels/kernel_api.hpp"
#include "kernels/wavefront_api.hpp"

namespace rt
{
template <typename Real>
constexpr bool wavefront_supported()
{
    return sizeof(Real) == sizeof(float);
}

template <typename Real>
class wavefront
{
public:
    // mode: the scene's kernel_mode; options: the session's (wavefront_slots: the pool's size in paths, 0 for one per
    // pixel; wavefront_graph: bounces per captured hipGraph, 0 for the default, -1 for none; wavefront_drain: the
    // active count below which the pool drains, 0 for the default, -1 for never)
    wavefront(const dev::scene_view<Real>& sc, const dev::camera_rec<Real>& cam, int mode,
              const rt_session_options& options, hipStream_t stream);
    ~wavefront();
    wavefront(const wavefront&) = delete;
    wavefront& operator=(const wavefront&) = delete;

    // The launch's work, added to sums; waits until it is done. Returns the bounce steps it launched.
    int64_t render(uint64_t seed, const dev::launch_work& work, double* sums, unsigned long long* rays);
    kernel_attributes attributes() const;
    std::size_t bytes() const { return bytes_; }
    int bounces_per_graph() const { return bounces_per_graph_; }
    // How many image rows one render call may cover
    int rows_capacity() const { return static_cast<int>(capacity_ / static_cast<uint32_t>(cam_.image_width)); }

private:
    dev::scene_view<float> sc_{};
    dev::camera_rec<float> cam_{};
    hipStream_t stream_ = nullptr;
    int mode_ = dev::mode_all_books;
    bool quality_ = false;
    uint32_t drain_below_ = 0;
    uint32_t capacity_ = 0;
    device_array<float4> ray_o_, ray_d_, thr_, qst_;
    device_array<uint4> rng_;
    device_array<int4> hit_;
    device_array<uint32_t> item_;
    device_array<float4> samples_;
    device_array<uint32_t> active0_, active1_, free_, queue_, counters_, prefix_;
    device_array<unsigned long long> tallies_;
    device_array<dev::wf_launch> launch_;
    uint32_t* host_counters_ = nullptr;    // pinned: the counters after each round
    dev::wf_launch* host_launch_ = nullptr; // pinned: the next launch's seed and work, copied to launch_
    int max_passes_ = 1;                   // passes one launch of the pool renders
    int64_t render_part(uint64_t seed, const dev::launch_work& work, double* sums, unsigned long long* rays);
    hipError_t launch_bounces();
    dev::wf_pool pool_{};
    int bounces_per_graph_ = 8;
    bool use_graph_ = true;
    // the bounce loop's graph: captured at the pool's first launch and replayed by every launch after it (the kernels
    // read the launch's seed and work from launch_), destroyed with the pool
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t graph_exec_ = nullptr;
    std::size_t bytes_ = 0;
};
}
