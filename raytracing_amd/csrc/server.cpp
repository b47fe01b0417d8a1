// The serving loop (see server.hpp)
#include "server.hpp"

#include <algorithm>
#include <stdexcept>
#include <thread>
#include <vector>

#include "device_scene.hpp"

namespace rt
{
namespace
{
constexpr int64_t max_step_passes = 1024; // a step's passes, and so its staging's (the megakernel stages 1024 at most)
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
    if (options_.passes_per_step > max_step_passes || options_.paths_per_step < 0)
    {
        throw std::invalid_argument("passes_per_step at most 1024 and paths_per_step 0 or more");
    }
    // no more streams than the hardware queues a process gets by default (4): more would only contend for them
    const int streams = std::min(4, options_.max_active);
    for (int k = 0; k < streams; ++k)
    {
        streams_.emplace_back();
    }
}

// A step of a view: passes_per_step, or more for a small view, so that each step traces at least paths_per_step paths
// (a step of a 256x256 view at 4 passes is 0.26 M paths: a few launches' fixed costs take most of its time)
int server::step_passes(const rt_camera_desc& camera) const
{
    const int64_t pixels = static_cast<int64_t>(camera.image_width) * camera.image_height;
    const int64_t wanted = options_.paths_per_step > 0 ? (options_.paths_per_step + pixels - 1) / pixels : 0;
    return static_cast<int>(std::clamp<int64_t>(wanted, options_.passes_per_step, max_step_passes));
}

double server::now() const
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - start_).count();
}

int64_t server::add_scene(const rt_scene& scene)
{
    std::shared_ptr<const shared_scene> s = upload_scene(scene, RT_FLOAT32);
    const std::size_t bytes = device_bytes(*s);
    if (used_bytes_ + bytes > options_.memory_budget_bytes)
    {
        throw std::invalid_argument("the scene doesn't fit the server's memory budget beside what is resident");
    }
    used_bytes_ += bytes;
    const int64_t id = next_scene_++;
    scenes_.emplace(id, resident{std::move(s), bytes, 0, false});
    return id;
}

void server::drop_scene_if_unused(int64_t scene_id)
{
    const auto f = scenes_.find(scene_id);
    if (f != scenes_.end() && f->second.removed && f->second.users == 0)
    {
        used_bytes_ -= f->second.bytes;
        scenes_.erase(f);
    }
}

void server::remove_scene(int64_t scene_id)
{
    const auto f = scenes_.find(scene_id);
    if (f == scenes_.end() || f->second.removed)
    {
        throw std::invalid_argument("no resident scene with this id");
    }
    f->second.removed = true;
    drop_scene_if_unused(scene_id);
}

int64_t server::submit_view(int64_t scene_id, const rt_camera_desc& camera, uint64_t seed, int64_t passes)
{
    const auto f = scenes_.find(scene_id);
    if (f == scenes_.end() || f->second.removed)
    {
        throw std::invalid_argument("no resident scene with this id");
    }
    validate_camera(camera);
    if (camera.sampling != scene_of(*f->second.scene).camera.sampling)
    {
        throw std::invalid_argument("a view must keep the scene's sampling strategy");
    }
    if (passes <= 0)
    {
        throw std::invalid_argument("a request must ask for at least one pass");
    }
    auto r = std::make_unique<request>();
    r->id = next_id_++;
    r->scene_id = scene_id;
    r->camera = camera;
    r->seed = seed;
    r->passes = passes;
    r->step_passes = step_passes(camera);
    // what the resident scenes hold stays held while requests may use them: a request that doesn't fit beside them
    // could never be admitted
    std::size_t resident = 0;
    for (const auto& [id, sc] : scenes_)
    {
        resident += sc.bytes;
    }
    if (view_device_bytes(*f->second.scene, camera, r->step_passes) + resident > options_.memory_budget_bytes)
    {
        throw std::invalid_argument("the request needs more device memory than the server's budget leaves beside "
                                    "the resident scenes");
    }
    r->st.id = r->id;
    r->st.state = 0;
    r->st.passes_total = passes;
    r->st.submit_time = now();
    ++f->second.users;
    const int64_t id = r->id;
    requests_.emplace(id, std::move(r));
    queue_.push_back(id);
    return id;
}

int64_t server::submit(const rt_scene& scene, uint64_t seed, int64_t passes)
{
    const int64_t scene_id = add_scene(scene);
    int64_t id = 0;
    try
    {
        id = submit_view(scene_id, scene.camera, seed, passes);
    }
    catch (...)
    {
        remove_scene(scene_id);
        throw;
    }
    requests_.at(id)->owns_scene = true;
    scenes_.at(scene_id).removed = true; // freed with its request
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
        const std::size_t bytes = view_device_bytes(*scenes_.at(r.scene_id).scene, r.camera, r.step_passes);
        if (used_bytes_ + bytes > options_.memory_budget_bytes)
        {
            break;
        }
        queue_.pop_front();
        try
        {
            rt_session_options o{};
            o.seed = r.seed;
            // the megakernel at the compiler's register budget: a request is too short to pay for the autotuner's
            // probes (most of a 16-pass request would be rendered by them, the wavefront's among them)
            o.integrator = RT_INTEGRATOR_MEGAKERNEL;
            r.work = make_session(scenes_.at(r.scene_id).scene, r.camera, o);
            // a stream of the server's few, shared in turn; staging for one step's passes at most
            r.stream = next_stream_;
            next_stream_ = (next_stream_ + 1) % streams_.size();
            r.work->serve_on(streams_[r.stream].get(), r.step_passes);
            r.bytes = bytes;
            used_bytes_ += bytes;
            r.st.state = 1;
            r.st.start_time = now();
            ++active;
        }
        catch (const std::exception&)
        {
            r.work.reset();
            r.st.state = 3;
            r.st.finish_time = now();
        }
    }
}

// A request's step ended: its passes counted, its milestone and completion stamped
void server::finish_step(request& r)
{
    const int64_t target = r.in_flight;
    r.in_flight = 0;
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
}

int server::step()
{
    admit();
    // Every active request without a step in flight gets its next one, enqueued on its stream without waiting: up to
    // the next power-of-two milestone, at most its step's passes. The steps of requests on different streams run
    // side by side on the device, each persistent grid on a share of the chip.
    std::vector<request*> active;
    for (auto& [id, rp] : requests_)
    {
        if (rp->st.state == 1)
        {
            active.push_back(rp.get());
        }
    }
    const int ways = static_cast<int>(std::min(active.size(), streams_.size()));
    for (request* r : active)
    {
        if (r->in_flight != 0)
        {
            continue;
        }
        const int64_t done = r->st.passes_done;
        int64_t next_milestone = 1;
        while (next_milestone <= done)
        {
            next_milestone *= 2;
        }
        const int64_t target = std::min({r->passes, done + r->step_passes, std::max(next_milestone, done + 1)});
        try
        {
            r->work->enqueue_to(target, std::max(1, ways));
            r->in_flight = target;
        }
        catch (const std::exception&)
        {
            r->st.state = 3;
            r->st.finish_time = now();
        }
    }
    // Then wait for the first step to end (any request's, not all of them): its request is done or steps again at the
    // next call, and a request that arrived meanwhile is admitted as soon as memory and a slot are free
    bool ended = active.empty();
    while (!ended)
    {
        for (request* r : active)
        {
            if (r->st.state != 1 || r->in_flight == 0)
            {
                continue;
            }
            bool over = false;
            try
            {
                over = r->work->poll();
            }
            catch (const std::exception&)
            {
                r->in_flight = 0;
                r->st.state = 3;
                r->st.finish_time = now();
                ended = true;
                continue;
            }
            if (over)
            {
                finish_step(*r);
                ended = true;
            }
        }
        if (!ended)
        {
            std::this_thread::sleep_for(std::chrono::microseconds(50));
        }
    }
    int still = 0;
    for (const request* r : active)
    {
        still += r->st.state == 1 ? 1 : 0;
    }
    return still;
}

rt_server_memory server::memory() const
{
    rt_server_memory m{};
    m.budget_bytes = options_.memory_budget_bytes;
    m.reserved_bytes = used_bytes_;
    uint64_t held = 0;
    for (const auto& [id, sc] : scenes_)
    {
        held += sc.bytes;
    }
    for (const auto& [id, r] : requests_)
    {
        held += r->work ? r->work->view_device_bytes() : 0;
    }
    m.held_bytes = held;
    return m;
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
    const int64_t scene_id = f->second->scene_id;
    queue_.erase(std::remove(queue_.begin(), queue_.end(), id), queue_.end());
    requests_.erase(f);
    const auto s = scenes_.find(scene_id);
    if (s != scenes_.end())
    {
        --s->second.users;
        drop_scene_if_unused(scene_id);
    }
}
}
