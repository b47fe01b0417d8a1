// The serving loop (see server.hpp)
#include "server.hpp"

#include <algorithm>
#include <stdexcept>
#include <thread>
#include <vector>

#include "device_scene.hpp"

namespace rt
{
std::size_t request_bytes(const rt_scene& s)
{
    const std::size_t pixels = static_cast<std::size_t>(s.camera.image_width) * s.camera.image_height;
    std::size_t texels = 0;
    for (int64_t i = 0; i < s.image_count; ++i)
    {
        texels += static_cast<std::size_t>(s.images[i].width) * s.images[i].height;
    }
    std::size_t prims = static_cast<std::size_t>(s.sphere_count + s.quad_count);
    for (int64_t i = 0; i < s.instance_count; ++i)
    {
        prims += static_cast<std::size_t>(s.meshes[s.instances[i].mesh].triangle_count);
    }
    if (s.instance_count == 0)
    {
        for (int64_t i = 0; i < s.mesh_count; ++i)
        {
            prims += static_cast<std::size_t>(s.meshes[i].triangle_count);
        }
    }
    std::size_t env = 0;
    if (s.environment)
    {
        env = static_cast<std::size_t>(s.environment->image.width) * s.environment->image.height;
    }
    return 24 * pixels                                  // double sums
           + 16 * 8 * pixels                            // the wavefront's slots
           + 16 * texels                                // float4 texels
           + prims * (64 + 48 + 16 + 4)                 // nodes, records, frames, materials ids
           + env * (16 + 16)                            // map texels and alias bins (padded to a power of two)
           + static_cast<std::size_t>(s.texture_count) * 4200 + (1u << 20);
}

server::server(const rt_server_options& options) : options_(options)
{
    if (options_.memory_budget_bytes == 0)
    {
        options_.memory_budget_bytes = 16ull << 30;
    }
    if (options_.max_active <= 0)
    {
        options_.max_active = 8;
    }
    if (options_.passes_per_step <= 0)
    {
        options_.passes_per_step = 4;
    }
}

double server::now() const
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - start_).count();
}

int64_t server::submit(const rt_scene& scene, uint64_t seed, int64_t passes)
{
    validate_scene(scene);
    if (passes <= 0)
    {
        throw std::invalid_argument("a request must ask for at least one pass");
    }
    auto r = std::make_unique<request>();
    r->id = next_id_++;
    r->scene = scene_tables::copy_of(scene);
    r->seed = seed;
    r->passes = passes;
    r->bytes = request_bytes(scene);
    if (r->bytes > options_.memory_budget_bytes)
    {
        throw std::invalid_argument("the request needs more device memory than the server's whole budget");
    }
    r->st.id = r->id;
    r->st.state = 0;
    r->st.passes_total = passes;
    r->st.submit_time = now();
    const int64_t id = r->id;
    requests_.emplace(id, std::move(r));
    queue_.push_back(id);
    return id;
}

void server::admit()
{
    int active = 0;
    for (const auto& [id, r] : requests_)
    {
        active += r->st.state == 1 ? 1 : 0;
    }
    // first come, first served: a request waits until memory and a slot are free, and the ones behind it wait too
    while (!queue_.empty() && active < options_.max_active)
    {
        request& r = *requests_.at(queue_.front());
        if (used_bytes_ + r.bytes > options_.memory_budget_bytes)
        {
            break;
        }
        queue_.pop_front();
        try
        {
            rt_session_options o{};
            o.seed = r.seed;
            const rt_scene view = r.scene.view();
            r.work = make_session(view, o);
            used_bytes_ += r.bytes;
            r.st.state = 1;
            r.st.start_time = now();
            ++active;
        }
        catch (const std::exception&)
        {
            r.st.state = 3;
            r.st.finish_time = now();
        }
    }
}

int server::step()
{
    admit();
    // every active request renders its step at once, each from a host thread of its own on its session's own
    // stream, so their kernels overlap on the device; a request's step ends at its next power-of-two milestone
    std::vector<request*> active;
    for (auto& [id, rp] : requests_)
    {
        if (rp->st.state == 1)
        {
            active.push_back(rp.get());
        }
    }
    const auto run = [this](request& r) {
        const int64_t done = r.st.passes_done;
        int64_t next_milestone = 1;
        while (next_milestone <= done)
        {
            next_milestone *= 2;
        }
        const int64_t target = std::min({r.passes, done + options_.passes_per_step, std::max(next_milestone, done + 1)});
        try
        {
            r.work->render_to(target, 0);
        }
        catch (const std::exception&)
        {
            r.st.state = 3;
            r.st.finish_time = now();
            return;
        }
        r.st.passes_done = target;
        if ((target & (target - 1)) == 0 || target == r.passes)
        {
            r.st.milestone = target;
            r.st.milestone_time = now();
        }
        if (target >= r.passes)
        {
            r.st.state = 2;
            r.st.finish_time = now();
        }
    };
    if (active.size() == 1)
    {
        run(*active[0]);
    }
    else
    {
        std::vector<std::thread> workers;
        workers.reserve(active.size());
        for (request* r : active)
        {
            workers.emplace_back([&run, r] { run(*r); });
        }
        for (std::thread& w : workers)
        {
            w.join();
        }
    }
    int still = 0;
    for (const request* r : active)
    {
        still += r->st.state == 1 ? 1 : 0;
    }
    return still;
}

rt_request_status server::status(int64_t id) const
{
    const auto f = requests_.find(id);
    if (f == requests_.end())
    {
        throw std::invalid_argument("no request with this id");
    }
    return f->second->st;
}

void server::pixels(int64_t id, double* out, std::size_t count)
{
    const auto f = requests_.find(id);
    if (f == requests_.end())
    {
        throw std::invalid_argument("no request with this id");
    }
    if (!f->second->work || f->second->st.passes_done == 0)
    {
        throw std::logic_error("the request has no complete pass yet");
    }
    f->second->work->pixels(out, count);
}

void server::release(int64_t id)
{
    const auto f = requests_.find(id);
    if (f == requests_.end())
    {
        throw std::invalid_argument("no request with this id");
    }
    if (f->second->work)
    {
        used_bytes_ -= f->second->bytes;
    }
    queue_.erase(std::remove(queue_.begin(), queue_.end(), id), queue_.end());
    requests_.erase(f);
}
}
