// The serving loop: progressive renders of mixed requests on one device, admitted by a device-memory budget and
// stepped together (continuous batching). The reference's progressive gpu::session (src/gpu/renderer.h:568-673) is
// the unit of work: a request is a session rendered to power-of-two pass milestones. Scenes stay resident (their tree
// and records uploaded once), and a request is a view of one: its camera (any resolution), its seed and its passes.
#pragma once

#include <chrono>
#include <cstdint>
#include <deque>
#include <map>
#include <memory>
#include <stdexcept>
#include <utility>
#include <vector>

#include <hip/hip_runtime_api.h>

#include "../../include/raytracing_amd.h"
#include "scene_tables.hpp"
#include "session.hpp"

namespace rt
{
class server
{
public:
    explicit server(const rt_server_options& options);
    int64_t add_scene(const rt_scene& scene);
    void remove_scene(int64_t scene_id);
    int64_t submit_view(int64_t scene_id, const rt_camera_desc& camera, uint64_t seed, int64_t passes);
    int64_t submit(const rt_scene& scene, uint64_t seed, int64_t passes); // a scene of its own, removed with it
    // Admits queued requests as memory and slots free up, enqueues the next step of every active request that has
    // none in flight, and returns once one of the steps in flight has ended, with the number of requests still active
    int step();
    rt_server_memory memory() const;
    rt_request_status status(int64_t id) const;
    void pixels(int64_t id, double* out, std::size_t count);
    void release(int64_t id);

private:
    int step_passes(const rt_camera_desc& camera) const;
    struct resident
    {
        std::shared_ptr<const shared_scene> scene;
        std::size_t bytes = 0;
        int users = 0;
        bool removed = false; // removed by its owner: freed when its last request is released
    };
    struct request
    {
        int64_t id = 0;
        int64_t scene_id = 0;
        bool owns_scene = false;
        rt_camera_desc camera{};
        uint64_t seed = 0;
        int64_t passes = 0;
        int step_passes = 0; // the passes of one of its steps, and its staging's
        std::size_t bytes = 0;
        std::unique_ptr<session> work;
        std::size_t stream = 0; // of the server's streams
        int64_t in_flight = 0;  // the passes the image will hold when its step in flight ends (0: none in flight)
        rt_request_status st{};
    };
    // one of the server's streams
    class stream_handle
    {
    public:
        stream_handle()
        {
            if (hipStreamCreateWithFlags(&s_, hipStreamNonBlocking) != hipSuccess)
            {
                throw std::runtime_error("creating a HIP stream");
            }
        }
        ~stream_handle()
        {
            if (s_)
            {
                (void)hipStreamDestroy(s_);
            }
        }
        stream_handle(stream_handle&& o) noexcept : s_(std::exchange(o.s_, nullptr)) {}
        stream_handle(const stream_handle&) = delete;
        stream_handle& operator=(const stream_handle&) = delete;
        hipStream_t get() const { return s_; }

    private:
        hipStream_t s_ = nullptr;
    };
    rt_server_options options_;
    std::vector<stream_handle> streams_;
    std::size_t next_stream_ = 0;
    std::map<int64_t, resident> scenes_;
    std::map<int64_t, std::unique_ptr<request>> requests_; // (destroyed before the streams their steps run on)
    std::deque<int64_t> queue_;
    std::size_t used_bytes_ = 0;
    int64_t next_id_ = 1;
    int64_t next_scene_ = 1;
    std::chrono::steady_clock::time_point start_ = std::chrono::steady_clock::now();
    double now() const;
    void admit();
    void finish_step(request& r);
    void drop_scene_if_unused(int64_t scene_id);
};

}
