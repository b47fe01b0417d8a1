// The C ABI (include/raytracing_amd.h): exceptions become status codes and a thread-local message
#include <cstring>
#include <exception>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>

#include <hip/hip_runtime.h>

#include "../../include/raytracing_amd.h"
#include "bvh_build.hpp"
#include "kernels/kernel_api.hpp"
#include "device_scene.hpp"
#include "quality_tables.hpp"
#include "scene_build.hpp"
#include "server.hpp"
#include "session.hpp"

struct rt_scene_handle
{
    rt::built_scene built;
    rt_scene view{};
    std::string warnings; // the model's warnings, one line each
};

struct rt_session
{
    std::unique_ptr<rt::session> impl;
};

struct rt_server
{
    std::unique_ptr<rt::server> impl;
};

namespace
{
thread_local std::string last_error;

template <typename F>
int guarded(F&& f)
{
    try
    {
        f();
        last_error.clear();
        return RT_OK;
    }
    catch (const std::invalid_argument& e)
    {
        last_error = e.what();
        return RT_ERROR_INVALID_ARGUMENT;
    }
    catch (const std::logic_error& e)
    {
        last_error = e.what();
        return RT_ERROR_LOGIC;
    }
    catch (const std::exception& e)
    {
        last_error = e.what();
        return RT_ERROR_RUNTIME;
    }
    catch (...)
    {
        last_error = "unknown error";
        return RT_ERROR_RUNTIME;
    }
}

void need(const void* p, const char* what)
{
    if (!p)
    {
        throw std::invalid_argument(std::string(what) + " must not be NULL");
    }
}
}

extern "C" {

const char* rt_last_error(void) { return last_error.c_str(); }

int rt_abi_version(void) { return RT_ABI_VERSION; }

const char* rt_scene_names(void) { return rt::scene_names(); }

int rt_scene_build(const char* name, const rt_scene_overrides* overrides, rt_scene_handle** out)
{
    return guarded([&] {
        need(name, "name");
        need(out, "out");
        rt_scene_overrides o{};
        o.sampling = -1;
        if (overrides)
        {
            o = *overrides;
        }
        auto h = std::make_unique<rt_scene_handle>();
        h->built = rt::build_scene(name, o);
        h->view = h->built.tables.view();
        for (const std::string& w : h->built.warnings)
        {
            h->warnings += w + "\n";
        }
        *out = h.release();
    });
}

const rt_scene* rt_scene_get(const rt_scene_handle* handle) { return handle ? &handle->view : nullptr; }

int64_t rt_scene_passes(const rt_scene_handle* handle) { return handle ? handle->built.passes : 0; }

const char* rt_scene_warnings(const rt_scene_handle* handle) { return handle ? handle->warnings.c_str() : ""; }

void rt_scene_free(rt_scene_handle* handle) { delete handle; }

int rt_session_create(const rt_scene* scene, const rt_session_options* options, rt_session** out)
{
    return guarded([&] {
        need(scene, "scene");
        need(out, "out");
        rt_session_options o{};
        if (options)
        {
            o = *options;
        }
        auto s = std::make_unique<rt_session>();
        s->impl = rt::make_session(*scene, o);
        *out = s.release();
    });
}

void rt_session_destroy(rt_session* session) { delete session; }

int rt_session_set_camera(rt_session* session, const rt_camera_desc* camera)
{
    return guarded([&] {
        need(session, "session");
        need(camera, "camera");
        session->impl->set_camera(*camera);
    });
}

int rt_session_camera(const rt_session* session, rt_camera_desc* camera)
{
    return guarded([&] {
        need(session, "session");
        need(camera, "camera");
        *camera = session->impl->camera();
    });
}

int rt_session_has_view(const rt_session* session, int32_t* has_view)
{
    return guarded([&] {
        need(session, "session");
        need(has_view, "has_view");
        *has_view = session->impl->has_view() ? 1 : 0;
    });
}

int rt_session_render_to(rt_session* session, int64_t passes, int32_t rows, rt_render_report* report)
{
    return guarded([&] {
        need(session, "session");
        const rt_render_report r = session->impl->render_to(passes, rows);
        if (report)
        {
            *report = r;
        }
    });
}

int rt_session_render_to_progress(rt_session* session, int64_t passes, int32_t rows, rt_render_report* report,
                                  rt_progress_fn progress, void* user)
{
    return guarded([&] {
        need(session, "session");
        std::function<void(int64_t)> on_pass;
        if (progress)
        {
            on_pass = [&](int64_t done) { progress(done, passes, user); };
        }
        const rt_render_report r = session->impl->render_to(passes, rows, on_pass);
        if (report)
        {
            *report = r;
        }
    });
}

int rt_session_render_to_static(rt_session* session, int64_t passes, rt_render_report* report)
{
    return guarded([&] {
        need(session, "session");
        const rt_render_report r = session->impl->render_to_static(passes);
        if (report)
        {
            *report = r;
        }
    });
}

int rt_session_render_for(rt_session* session, double seconds, int64_t pass_limit, rt_render_report* report)
{
    return guarded([&] {
        need(session, "session");
        const rt_render_report r = session->impl->render_for(seconds, pass_limit);
        if (report)
        {
            *report = r;
        }
    });
}

int rt_session_state(const rt_session* session, int64_t* passes, int32_t* rows)
{
    return guarded([&] {
        need(session, "session");
        const rt::accumulation a = session->impl->state();
        if (passes)
        {
            *passes = a.passes;
        }
        if (rows)
        {
            *rows = a.rows;
        }
    });
}

int rt_session_pixels(rt_session* session, double* out, size_t count)
{
    return guarded([&] {
        need(session, "session");
        need(out, "out");
        session->impl->pixels(out, count);
    });
}

int rt_session_sums(rt_session* session, double* out, size_t count)
{
    return guarded([&] {
        need(session, "session");
        need(out, "out");
        session->impl->sums(out, count);
    });
}

int rt_session_display_rgba8(rt_session* session, uint8_t* out, size_t count)
{
    return guarded([&] {
        need(session, "session");
        need(out, "out");
        session->impl->display(out, count);
    });
}

int rt_session_display_toned(rt_session* session, int32_t tone_map, double exposure, uint8_t* out, size_t count)
{
    return guarded([&] {
        need(session, "session");
        need(out, "out");
        session->impl->display_toned(tone_map, exposure, out, count);
    });
}

int rt_session_counters(const rt_session* session, uint64_t* rays, uint64_t* shadow_rays, double* kernel_seconds)
{
    return guarded([&] {
        need(session, "session");
        session->impl->counters(rays, shadow_rays, kernel_seconds);
    });
}

int rt_session_kernel_info(const rt_session* session, rt_kernel_info* info)
{
    return guarded([&] {
        need(session, "session");
        need(info, "info");
        *info = session->impl->kernel();
    });
}

int rt_render(const rt_scene* scene, const rt_render_options* options, double* pixels, size_t count,
              rt_render_stats* stats)
{
    return guarded([&] {
        need(scene, "scene");
        need(pixels, "pixels");
        const auto start = std::chrono::steady_clock::now();
        rt_render_options o{};
        if (options)
        {
            o = *options;
        }
        if (o.passes < 0)
        {
            throw std::invalid_argument("the number of passes must not be negative");
        }
        if (count < 3ull * static_cast<size_t>(scene->camera.image_width) * scene->camera.image_height)
        {
            throw std::invalid_argument("the output buffer is smaller than width x height x 3");
        }
        auto s = rt::make_session(*scene, o.session);
        const int64_t passes =
            o.passes > 0 ? o.passes : static_cast<int64_t>(scene->camera.sqrt_spp) * scene->camera.sqrt_spp;
        std::function<void(int64_t)> on_pass;
        if (o.progress)
        {
            on_pass = [&](int64_t done) { o.progress(done, passes, o.progress_user); };
        }
        if (o.tone_map != RT_TONE_BOOKS && !o.rgba)
        {
            throw std::invalid_argument("a tone map other than the books' needs the rgba buffer it fills");
        }
        s->render_to(passes, 0, on_pass);
        s->pixels(pixels, count);
        // the render's time, before the display's bytes (as the reference times it)
        const double seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
        if (o.tone_map != RT_TONE_BOOKS)
        {
            s->display_toned(o.tone_map, o.exposure, o.rgba,
                             4ull * static_cast<size_t>(scene->camera.image_width) * scene->camera.image_height);
        }
        if (stats)
        {
            s->counters(&stats->rays, &stats->shadow_rays, &stats->kernel_seconds);
            stats->kernel = s->kernel();
            stats->seconds = seconds;
        }
    });
}

int rt_session_closest_hit(rt_session* session, const float* origins, const float* directions, const float* times,
                           const int32_t* from, const uint8_t* inward, int64_t count, int32_t* index, float* t,
                           uint8_t* entering)
{
    return guarded([&] {
        need(session, "session");
        if (count > 0)
        {
            need(origins, "origins");
            need(directions, "directions");
            need(times, "times");
            need(from, "from");
            need(inward, "inward");
            need(index, "index");
            need(t, "t");
            need(entering, "entering");
        }
        session->impl->closest_hit(origins, directions, times, from, inward, count, index, t, entering);
    });
}

int rt_session_occluded(rt_session* session, const float* origins, const float* directions, const float* times,
                        const int32_t* from, const uint8_t* inward, const float* t_max, const int32_t* skip,
                        int64_t count, uint8_t* blocked, uint8_t* walked)
{
    return guarded([&] {
        need(session, "session");
        if (count > 0)
        {
            need(origins, "origins");
            need(directions, "directions");
            need(times, "times");
            need(from, "from");
            need(inward, "inward");
            need(t_max, "t_max");
            need(skip, "skip");
        }
        session->impl->occluded(origins, directions, times, from, inward, t_max, skip, count, blocked, walked);
    });
}

int rt_session_media_events(rt_session* session, const float* origins, const float* directions, const float* times,
                            const float* t_surface, int64_t count, uint64_t seed, int32_t* medium, float* t,
                            float* transmitted)
{
    return guarded([&] {
        need(session, "session");
        if (count > 0)
        {
            need(origins, "origins");
            need(directions, "directions");
            need(times, "times");
            need(t_surface, "t_surface");
            need(medium, "medium");
            need(t, "t");
            need(transmitted, "transmitted");
        }
        session->impl->media_events(origins, directions, times, t_surface, count, seed, medium, t, transmitted);
    });
}

int rt_session_light_samples(rt_session* session, const float* points, int64_t count, uint64_t seed, float* wi,
                             float* distance, float* radiance, float* density, int32_t* skip, float* hit_density)
{
    return guarded([&] {
        need(session, "session");
        if (count > 0)
        {
            need(points, "points");
            need(wi, "wi");
            need(distance, "distance");
            need(radiance, "radiance");
            need(density, "density");
            need(skip, "skip");
            need(hit_density, "hit_density");
        }
        session->impl->light_samples(points, count, seed, wi, distance, radiance, density, skip, hit_density);
    });
}

int rt_session_reference_hits(rt_session* session, const float* origins, const float* directions, const float* times,
                              const int32_t* from, const uint8_t* inward, int64_t count, int32_t* index, float* t,
                              uint8_t* entering)
{
    return guarded([&] {
        need(session, "session");
        if (count > 0)
        {
            need(origins, "origins");
            need(directions, "directions");
            need(times, "times");
            need(from, "from");
            need(inward, "inward");
            need(index, "index");
            need(t, "t");
            need(entering, "entering");
        }
        session->impl->reference_hits(origins, directions, times, from, inward, count, index, t, entering);
    });
}

int rt_sample_numbers(uint64_t seed, const uint64_t* pixels, const uint64_t* passes, int64_t count, int32_t draws,
                      uint32_t* words, float* floats, double* doubles, float* directions)
{
    return guarded([&] { rt::sample_numbers(seed, pixels, passes, count, draws, words, floats, doubles, directions); });
}

int rt_tone_probe(int32_t tone_map, double exposure, const float* colours, const float* dither, int64_t count,
                  uint8_t* out)
{
    return guarded([&] {
        if (count > 0)
        {
            need(colours, "colours");
            need(out, "out");
        }
        rt::tone_probe(tone_map, exposure, colours, dither, count, out);
    });
}

int rt_scatter_probe(int32_t lobe, float alpha, float roughness, const float color[3], int64_t count, const float* v,
                     const float* l, uint64_t seed, float* value, float* density, float* drawn, float* weight,
                     float* drawn_density)
{
    return guarded([&] {
        if (count > 0)
        {
            need(v, "v");
            need(l, "l");
            need(value, "value");
            need(density, "density");
            need(drawn, "drawn");
            need(weight, "weight");
            need(drawn_density, "drawn_density");
        }
        rt::scatter_probe(lobe, alpha, roughness, color, count, v, l, seed, value, density, drawn, weight,
                          drawn_density);
    });
}

int rt_session_traffic(rt_session* session, int64_t passes, rt_traffic* traffic)
{
    return guarded([&] {
        need(session, "session");
        need(traffic, "traffic");
        *traffic = session->impl->traffic(passes);
    });
}

int rt_scene_bvh(const rt_scene* scene, int32_t kind, float* node_boxes, int32_t* node_ints, int64_t* node_count,
                 int32_t* sphere_ids, int32_t* quad_ids, int32_t* depth)
{
    return guarded([&] {
        need(scene, "scene");
        need(node_count, "node_count");
        rt::validate_scene(*scene);
        const auto tree = rt::build_bvh<float>(
            *scene, rt::leaf_size_for(static_cast<size_t>(scene->sphere_count + scene->quad_count)), kind);
        *node_count = static_cast<int64_t>(tree.nodes.size());
        if (depth)
        {
            *depth = tree.depth;
        }
        if (node_boxes && node_ints)
        {
            for (size_t i = 0; i < tree.nodes.size(); ++i)
            {
                const auto& n = tree.nodes[i];
                std::memcpy(node_boxes + 6 * i, n.box_min, 3 * sizeof(float));
                std::memcpy(node_boxes + 6 * i + 3, n.box_max, 3 * sizeof(float));
                node_ints[2 * i] = n.index;
                node_ints[2 * i + 1] = n.count;
            }
        }
        if (sphere_ids)
        {
            std::memcpy(sphere_ids, tree.sphere_ids.data(), tree.sphere_ids.size() * sizeof(int32_t));
        }
        if (quad_ids)
        {
            std::memcpy(quad_ids, tree.quad_ids.data(), tree.quad_ids.size() * sizeof(int32_t));
        }
    });
}

int rt_scene_wide_nodes(const rt_scene* scene, int32_t kind, void* float_nodes, void* packed_nodes, int64_t* count)
{
    return guarded([&] {
        need(scene, "scene");
        need(count, "count");
        rt::validate_scene(*scene);
        const auto tree = rt::build_bvh<float>(
            *scene, rt::leaf_size_for(static_cast<size_t>(scene->sphere_count + scene->quad_count)), kind);
        *count = static_cast<int64_t>(tree.wide.size());
        if (float_nodes)
        {
            std::memcpy(float_nodes, tree.wide.data(), tree.wide.size() * sizeof(rt::dev::wide_node<float>));
        }
        if (packed_nodes)
        {
            const std::vector<rt::dev::packed_node> packed = rt::pack_nodes(tree.wide);
            std::memcpy(packed_nodes, packed.data(), packed.size() * sizeof(rt::dev::packed_node));
        }
    });
}

int rt_scene_lamps(const rt_scene* scene, float* lamps, int64_t* count)
{
    return guarded([&] {
        need(scene, "scene");
        need(count, "count");
        rt::validate_scene(*scene);
        const auto tree = rt::build_bvh<float>(
            *scene, rt::leaf_size_for(static_cast<size_t>(scene->sphere_count + scene->quad_count)), RT_TREE_DEFAULT);
        const rt::lamp_set set = rt::build_lamp_set(*scene, tree.sphere_ids, tree.quad_ids, tree.prim_materials.size());
        *count = static_cast<int64_t>(set.lamps.size());
        if (lamps)
        {
            std::memcpy(lamps, set.lamps.data(), set.lamps.size() * sizeof(rt::dev::emitter_rec));
        }
    });
}

int rt_environment_tables(const rt_environment_desc* environment, float* texels, float* row_cos, float* bin_q,
                          int32_t* bin_alias, int32_t* log2_bins, double* total)
{
    return guarded([&] {
        need(environment, "environment");
        const rt::environment_tables t = rt::build_environment_tables(*environment);
        if (log2_bins)
        {
            *log2_bins = t.bins.empty() ? 0 : t.log2_bins;
        }
        if (total)
        {
            *total = t.total;
        }
        if (texels)
        {
            std::memcpy(texels, t.texels.data(), t.texels.size() * sizeof(rt::dev::env_texel));
        }
        if (row_cos)
        {
            std::memcpy(row_cos, t.row_cos.data(), t.row_cos.size() * sizeof(float));
        }
        for (std::size_t k = 0; k < t.bins.size(); ++k)
        {
            if (bin_q)
            {
                bin_q[k] = t.bins[k].q;
            }
            if (bin_alias)
            {
                bin_alias[k] = t.bins[k].alias;
            }
        }
    });
}

int rt_server_create(const rt_server_options* options, rt_server** out)
{
    return guarded([&] {
        need(out, "out");
        rt_server_options o{};
        if (options)
        {
            o = *options;
        }
        auto s = std::make_unique<rt_server>();
        s->impl = std::make_unique<rt::server>(o);
        *out = s.release();
    });
}

void rt_server_destroy(rt_server* server) { delete server; }

int rt_server_add_scene(rt_server* server, const rt_scene* scene, int64_t* scene_id)
{
    return guarded([&] {
        need(server, "server");
        need(scene, "scene");
        need(scene_id, "scene_id");
        *scene_id = server->impl->add_scene(*scene);
    });
}

int rt_server_remove_scene(rt_server* server, int64_t scene_id)
{
    return guarded([&] {
        need(server, "server");
        server->impl->remove_scene(scene_id);
    });
}

int rt_server_submit_view(rt_server* server, int64_t scene_id, const rt_camera_desc* camera, uint64_t seed,
                          int64_t passes, int64_t* id)
{
    return guarded([&] {
        need(server, "server");
        need(camera, "camera");
        need(id, "id");
        *id = server->impl->submit_view(scene_id, *camera, seed, passes);
    });
}

int rt_server_submit(rt_server* server, const rt_scene* scene, uint64_t seed, int64_t passes, int64_t* id)
{
    return guarded([&] {
        need(server, "server");
        need(scene, "scene");
        need(id, "id");
        *id = server->impl->submit(*scene, seed, passes);
    });
}

int rt_server_step(rt_server* server, int32_t* active)
{
    return guarded([&] {
        need(server, "server");
        const int n = server->impl->step();
        if (active)
        {
            *active = n;
        }
    });
}

int rt_server_memory_use(const rt_server* server, rt_server_memory* memory)
{
    return guarded([&] {
        need(server, "server");
        need(memory, "memory");
        *memory = server->impl->memory();
    });
}

int rt_server_status(const rt_server* server, int64_t id, rt_request_status* status)
{
    return guarded([&] {
        need(server, "server");
        need(status, "status");
        *status = server->impl->status(id);
    });
}

int rt_server_pixels(rt_server* server, int64_t id, double* out, size_t count)
{
    return guarded([&] {
        need(server, "server");
        need(out, "out");
        server->impl->pixels(id, out, count);
    });
}

int rt_server_release(rt_server* server, int64_t id)
{
    return guarded([&] {
        need(server, "server");
        server->impl->release(id);
    });
}

int rt_query_device_wait(rt_device_info* info, int32_t wait)
{
    return guarded([&] {
        if (wait != RT_WAIT_SPIN && wait != RT_WAIT_SLEEP)
        {
            throw std::invalid_argument("unknown wait mode");
        }
        rt::check(hipSetDeviceFlags(wait == RT_WAIT_SLEEP ? hipDeviceScheduleBlockingSync : hipDeviceScheduleSpin),
                  "setting how the device waits");
        if (const int status = rt_query_device(info); status != RT_OK)
        {
            throw std::runtime_error(rt_last_error());
        }
    });
}

int rt_query_device(rt_device_info* info)
{
    return guarded([&] {
        need(info, "info");
        hipDeviceProp_t p{};
        rt::check(hipGetDeviceProperties(&p, 0), "reading the device's properties");
        std::memset(info, 0, sizeof *info);
        std::snprintf(info->name, sizeof info->name, "%s", p.name);
        std::snprintf(info->arch, sizeof info->arch, "%s", p.gcnArchName);
        info->compute_units = p.multiProcessorCount;
        info->memory_gib = static_cast<double>(p.totalGlobalMem) / (1024.0 * 1024.0 * 1024.0);
        int v = 0;
        if (hipRuntimeGetVersion(&v) == hipSuccess)
        {
            info->runtime_version = v;
        }
        if (hipDriverGetVersion(&v) == hipSuccess)
        {
            info->driver_version = v;
        }
        rt::check(hipFree(nullptr), "starting the runtime");
    });
}
}

extern "C" int rt_copy_probe(uint64_t bytes, int32_t repeats, double* seconds)
{
    return guarded([&] {
        need(seconds, "seconds");
        if (bytes < 16 || bytes % 16 != 0 || repeats < 1)
        {
            throw std::invalid_argument("a copy of a positive multiple of 16 bytes, at least once");
        }
        rt::device_array<uint8_t> in(bytes), out(bytes);
        hipStream_t stream = nullptr;
        rt::check(hipStreamCreate(&stream), "creating a stream");
        hipEvent_t start = nullptr, stop = nullptr;
        rt::check(hipEventCreate(&start), "creating an event");
        rt::check(hipEventCreate(&stop), "creating an event");
        rt::check(hipMemsetAsync(in.data(), 1, bytes, stream), "filling the source");
        rt::check(hipEventRecord(start, stream), "recording an event");
        for (int r = 0; r < repeats; ++r)
        {
            rt::check(rt::launch_copy(in.data(), out.data(), bytes, stream), "launching the copy");
        }
        rt::check(hipEventRecord(stop, stream), "recording an event");
        rt::check(hipEventSynchronize(stop), "copying");
        float ms = 0;
        rt::check(hipEventElapsedTime(&ms, start, stop), "timing the copy");
        *seconds = ms / 1000.0 / repeats;
        (void)hipEventDestroy(start);
        (void)hipEventDestroy(stop);
        (void)hipStreamDestroy(stream);
    });
}
