// The serving loop: progressive renders of mixed requests on one device, admitted by a device-memory budget and
// stepped together (continuous batching). The reference's progressive gpu::session (src/gpu/renderer.h:568-673) is
// the unit of work: a request is a session rendered to power-of-two pass milestones.
#pragma once

#include <chrono>
#include <cstdint>
#include <deque>
#include <map>
#include <memory>

#include "../../include/raytracing_amd.h"
#include "scene_tables.hpp"
#include "session.hpp"

namespace rt
{
class server
{
public:
    explicit server(const rt_server_options& options);
    int64_t submit(const rt_scene& scene, uint64_t seed, int64_t passes);
    int step(); // returns the number of requests active after the step
    rt_request_status status(int64_t id) const;
    void pixels(int64_t id, double* out, std::size_t count);
    void release(int64_t id);

private:
    struct request
    {
        int64_t id = 0;
        scene_tables scene;
        uint64_t seed = 0;
        int64_t passes = 0;
        std::size_t bytes = 0;
        std::unique_ptr<session> work;
        rt_request_status st{};
    };
    rt_server_options options_;
    std::map<int64_t, std::unique_ptr<request>> requests_;
    std::deque<int64_t> queue_;
    std::size_t used_bytes_ = 0;
    int64_t next_id_ = 1;
    std::chrono::steady_clock::time_point start_ = std::chrono::steady_clock::now();
    double now() const;
    void admit();
};

// The device memory a request needs: its scene's arrays and its image's sums, estimated before anything is allocated
std::size_t request_bytes(const rt_scene& scene);
}
