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
    int step(); // returns the number of requests active after the step
    rt_request_status status(int64_t id) const;
    void pixels(int64_t id, double* out, std::size_t count);
    void release(int64_t id);

private:
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
        std::size_t bytes = 0;
        std::unique_ptr<session> work;
        rt_request_status st{};
    };
    rt_server_options options_;
    std::map<int64_t, resident> scenes_;
    std::map<int64_t, std::unique_ptr<request>> requests_;
    std::deque<int64_t> queue_;
    std::size_t used_bytes_ = 0;
    int64_t next_id_ = 1;
    int64_t next_scene_ = 1;
    std::chrono::steady_clock::time_point start_ = std::chrono::steady_clock::now();
    double now() const;
    void admit();
    void drop_scene_if_unused(int64_t scene_id);
};

// The device memory a view of a scene needs beside the scene itself: its image's sums and the wavefront's slots
std::size_t view_bytes(const rt_camera_desc& camera);
}
