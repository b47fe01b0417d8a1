// ORACLE — test infrastructure only (see cpu_model.hpp). The reference's CPU renderer, restated: camera::render and
// ray_color (src/camera.ixx:204-452), in double precision with one mt19937_64 per image row seeded
// (seed << 32) | row, rejection-sampled directions and a recursive ray_color. Exposed through a small C API for the
// tests (ctypes) and for bench.py's CPU-baseline leg, and as the command-line program rt_cpu (RayTracing's CLI subset,
// src/main.cpp).
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <thread>

#include "cpu_model.hpp"

namespace oracle
{
namespace
{
thread_local std::uint64_t rays_traced = 0;

struct renderer
{
    camera_frame frame;
    double pixel_samples_scale;
    double recip_sqrt_spp;
    sampling_strategy strategy;
    const hittable* world;
    const hittable_list* lights;
    std::uint32_t seed;

    // camera::get_ray (src/camera.ixx:293-301): the stratified offset (x then y), the lens point, then the time
    ray get_ray(int i, int j, int s_i, int s_j) const
    {
        const double px = (s_i + random_double()) * recip_sqrt_spp - 0.5;
        const double py = (s_j + random_double()) * recip_sqrt_spp - 0.5;
        const point3 pixel_sample = frame.pixel00_loc + ((i + px) * frame.pixel_delta_u) + ((j + py) * frame.pixel_delta_v);
        point3 origin = frame.center;
        if (!(frame.defocus_angle <= 0))
        {
            const vec3 p = random_in_unit_disk();
            origin = frame.center + p.x * frame.defocus_disk_u + p.y * frame.defocus_disk_v;
        }
        const double time = random_double();
        return ray{origin, pixel_sample - origin, time};
    }

    template <typename Pdf>
    color sample(const Pdf& pdf, const hit_record& rec, const diffuse_scatter& diffuse, const ray& r, int depth) const
    {
        const ray scattered{rec.p, pdf.generate(), r.time};
        const double pdf_value = pdf.value(scattered.direction);
        const double scattering_pdf = rec.mat->scattering_pdf(r, rec, scattered);
        if (!(pdf_value > 0) || !(scattering_pdf > 0))
        {
            return {0, 0, 0};
        }
        return diffuse.attenuation * (scattering_pdf / pdf_value) * ray_color(scattered, depth - 1);
    }

    struct lights_pdf
    {
        const hittable_list* lights;
        point3 origin;
        double value(const vec3& d) const { return lights->pdf_value(origin, d); }
        vec3 generate() const { return lights->random(origin); }
    };
    template <typename A, typename B>
    struct mixture
    {
        A a;
        B b;
        double value(const vec3& d) const { return 0.5 * a.value(d) + 0.5 * b.value(d); }
        vec3 generate() const { return random_double() < 0.5 ? a.generate() : b.generate(); }
    };

    // ray_color (src/camera.ixx:355-452)
    color ray_color(const ray& r, int depth) const
    {
        if (depth <= 0)
        {
            return {0, 0, 0};
        }
        const auto rec = world->hit(r, interval{0.001, infinity});
        ++rays_traced;
        if (!rec)
        {
            return frame.background.in_direction(r.direction);
        }
        const color emitted = rec->mat->emitted(*rec);
        const auto scatter = rec->mat->scatter(r, *rec);
        if (!scatter)
        {
            return emitted;
        }
        if (const auto* specular = std::get_if<specular_scatter>(&*scatter))
        {
            return emitted + specular->attenuation * ray_color(specular->scattered, depth - 1);
        }
        const auto& diffuse = std::get<diffuse_scatter>(*scatter);
        switch (strategy)
        {
        case sampling_strategy::uniform:
            return emitted + sample(sphere_pdf{}, *rec, diffuse, r, depth);
        case sampling_strategy::material:
            return emitted + sample(diffuse.pdf, *rec, diffuse, r, depth);
        case sampling_strategy::lights:
            return emitted + sample(lights_pdf{lights, rec->p}, *rec, diffuse, r, depth);
        case sampling_strategy::mixture:
            if (!rec->mat->mixes_in_lights())
            {
                return emitted + sample(diffuse.pdf, *rec, diffuse, r, depth);
            }
            return emitted + sample(mixture<lights_pdf, material_pdf>{lights_pdf{lights, rec->p}, diffuse.pdf}, *rec,
                                    diffuse, r, depth);
        }
        return {0, 0, 0};
    }

    static color without_nan(const color& c)
    {
        return {std::isnan(c.x) ? 0.0 : c.x, std::isnan(c.y) ? 0.0 : c.y, std::isnan(c.z) ? 0.0 : c.z};
    }

    // The rows [row_begin, row_end) of camera::render (src/camera.ixx:204-272), each seeded on its own, so any subset
    // of rows gives exactly those rows of the whole image
    std::uint64_t render(int row_begin, int row_end, int threads, double* out) const
    {
        std::atomic<int> next_row = row_begin;
        std::atomic<std::uint64_t> rays = 0;
        auto rows = [&] {
            rays_traced = 0;
            for (int j = next_row++; j < row_end; j = next_row++)
            {
                seed_random((std::uint64_t{seed} << 32) | static_cast<std::uint64_t>(j));
                for (int i = 0; i < frame.image_width; ++i)
                {
                    color pixel{0, 0, 0};
                    for (int s_j = 0; s_j < frame.sqrt_spp; ++s_j)
                    {
                        for (int s_i = 0; s_i < frame.sqrt_spp; ++s_i)
                        {
                            pixel += without_nan(ray_color(get_ray(i, j, s_i, s_j), frame.max_depth));
                        }
                    }
                    const color value = pixel_samples_scale * pixel;
                    double* o = out + 3 * (static_cast<std::size_t>(j - row_begin) * frame.image_width + i);
                    o[0] = value.x;
                    o[1] = value.y;
                    o[2] = value.z;
                }
            }
            rays += rays_traced;
        };
        std::vector<std::thread> workers;
        for (int t = 0; t < threads; ++t)
        {
            workers.emplace_back(rows);
        }
        for (auto& w : workers)
        {
            w.join();
        }
        return rays.load();
    }
};

thread_local std::string last_error;

void put3(std::vector<double>& v, const vec3& a)
{
    v.push_back(a.x);
    v.push_back(a.y);
    v.push_back(a.z);
}
void put_sphere(std::vector<double>& v, const flat_sphere& s)
{
    put3(v, s.center);
    v.push_back(s.radius);
    put3(v, s.displacement);
    v.push_back(s.uv_cos);
    v.push_back(s.uv_sin);
    v.push_back(s.material);
}
void put_quad(std::vector<double>& v, const flat_quad& q)
{
    put3(v, q.q);
    put3(v, q.u);
    put3(v, q.v);
    v.push_back(q.material);
}
}
}

using namespace oracle;

extern "C" {

const char* oracle_last_error() { return last_error.c_str(); }

// Renders rows [row_begin, row_end) of a scene (row_end <= 0: to the last row) on `threads` threads (0: all cores).
// dims receives width, height, sqrt_spp. Returns 0, or -1 with oracle_last_error().
int oracle_render(const char* name, int width, int height, int spp, int sampling, int max_depth, std::uint32_t seed,
                  int threads,
                  int earth_width, std::uint32_t earth_seed, int row_begin, int row_end, double* out,
                  std::int64_t capacity, std::int32_t* dims, std::uint64_t* rays)
{
    try
    {
        scene_options opts;
        opts.earth_width = earth_width > 0 ? earth_width : 2048;
        opts.earth_seed = earth_seed;
        seed_random(std::mt19937_64::default_seed);
        scene s = build_scene(name, opts);
        if (width > 0)
        {
            s.camera.image_width = width;
        }
        if (height > 0)
        {
            // the reference's --height: the aspect ratio that gives the width that many rows (view_options.ixx:101)
            s.camera.aspect_ratio = static_cast<double>(s.camera.image_width) / height;
        }
        if (spp > 0)
        {
            s.camera.samples_per_pixel = spp;
        }
        if (sampling >= 0 && sampling <= 3)
        {
            s.camera.sampling = static_cast<sampling_strategy>(sampling);
        }
        if (max_depth > 0)
        {
            s.camera.max_depth = max_depth;
        }
        renderer r;
        r.frame = frame_of(s.camera);
        r.recip_sqrt_spp = 1.0 / r.frame.sqrt_spp;
        r.pixel_samples_scale = 1.0 / (static_cast<double>(r.frame.sqrt_spp) * r.frame.sqrt_spp);
        r.strategy = effective_sampling(s.camera.sampling, s.lights);
        r.world = &s.world;
        r.lights = &s.lights;
        r.seed = seed;
        if (dims)
        {
            dims[0] = r.frame.image_width;
            dims[1] = r.frame.image_height;
            dims[2] = r.frame.sqrt_spp;
        }
        if (!out)
        {
            return 0;
        }
        if (row_end <= 0 || row_end > r.frame.image_height)
        {
            row_end = r.frame.image_height;
        }
        row_begin = std::max(0, row_begin);
        if (row_begin >= row_end ||
            capacity < 3LL * static_cast<std::int64_t>(row_end - row_begin) * r.frame.image_width)
        {
            throw std::invalid_argument("the output buffer is too small for the rows asked for");
        }
        if (threads <= 0)
        {
            threads = static_cast<int>(std::max(1u, std::thread::hardware_concurrency()));
        }
        const std::uint64_t n = r.render(row_begin, row_end, threads, out);
        if (rays)
        {
            *rays = n;
        }
        return 0;
    }
    catch (const std::exception& e)
    {
        last_error = e.what();
        return -1;
    }
}

// The scene flattened as build_gpu_scene flattens it, as numbers: see tests/flatten.py for the layout. Returns the
// number of doubles (written if it fits in capacity), or -1.
std::int64_t oracle_flatten(const char* name, int width, int spp, int sampling, int max_depth, int earth_width,
                            std::uint32_t earth_seed, double* out, std::int64_t capacity)
{
    try
    {
        scene_options opts;
        opts.earth_width = earth_width > 0 ? earth_width : 2048;
        opts.earth_seed = earth_seed;
        const built_scene b = build_flat(name, width, spp, sampling, max_depth, opts);
        const flat_scene& f = b.flat;
        std::vector<double> v;
        for (std::size_t n : {f.spheres.size(), f.quads.size(), f.materials.size(), f.textures.size(), f.media.size(),
                              f.lights.size(), f.images.size(), f.perlins.size()})
        {
            v.push_back(static_cast<double>(n));
        }
        const camera_frame& c = f.frame;
        v.push_back(c.image_width);
        v.push_back(c.image_height);
        v.push_back(c.sqrt_spp);
        v.push_back(c.max_depth);
        put3(v, c.center);
        put3(v, c.pixel00_loc);
        put3(v, c.pixel_delta_u);
        put3(v, c.pixel_delta_v);
        v.push_back(c.defocus_angle > 0 ? 1 : 0);
        v.push_back(static_cast<double>(f.sampling));
        put3(v, c.defocus_disk_u);
        put3(v, c.defocus_disk_v);
        put3(v, c.background.bottom);
        put3(v, c.background.top);
        for (const auto& s : f.spheres)
        {
            put_sphere(v, s);
        }
        for (const auto& q : f.quads)
        {
            put_quad(v, q);
        }
        for (const auto& m : f.materials)
        {
            v.push_back(m.kind);
            v.push_back(m.texture);
            put3(v, m.albedo);
            v.push_back(m.fuzz);
            v.push_back(m.refraction_index);
        }
        for (const auto& t : f.textures)
        {
            v.push_back(t.kind);
            v.push_back(t.even);
            v.push_back(t.odd);
            v.push_back(t.image);
            v.push_back(t.perlin);
            put3(v, t.rgb);
            v.push_back(t.inv_scale);
            v.push_back(t.scale);
        }
        for (const auto& m : f.media)
        {
            v.push_back(m.boundary);
            v.push_back(m.material);
            v.push_back(m.neg_inv_density);
            put_sphere(v, m.sphere);
            v.push_back(static_cast<double>(m.quads.size()));
            for (const auto& q : m.quads)
            {
                put_quad(v, q);
            }
        }
        for (const auto& l : f.lights)
        {
            v.push_back(l.kind);
            put_quad(v, l.quad);
            put3(v, l.center);
            v.push_back(l.radius);
            v.push_back(l.weight);
        }
        for (const image* img : f.images)
        {
            // the size, then two sums over the texels in order: plain and weighted by position
            v.push_back(img->width);
            v.push_back(img->height);
            double sum = 0, weighted = 0;
            std::size_t k = 0;
            for (const color& p : img->pixels)
            {
                for (const double x : {p.x, p.y, p.z})
                {
                    sum += x;
                    weighted += static_cast<double>(++k % 9973) * x;
                }
            }
            v.push_back(sum);
            v.push_back(weighted);
        }
        for (const perlin* p : f.perlins)
        {
            for (const vec3& g : p->random_vectors)
            {
                put3(v, g);
            }
            for (const auto* perm : {&p->perm_x, &p->perm_y, &p->perm_z})
            {
                for (const int e : *perm)
                {
                    v.push_back(e);
                }
            }
        }
        const auto n = static_cast<std::int64_t>(v.size());
        if (out && capacity >= n)
        {
            std::memcpy(out, v.data(), v.size() * sizeof(double));
        }
        return n;
    }
    catch (const std::exception& e)
    {
        last_error = e.what();
        return -1;
    }
}

// The first outputs of mt19937_64 after seed_random(seed), for known-answer tests
void oracle_mt19937_64(std::uint64_t seed, std::int64_t n, std::uint64_t* out)
{
    std::mt19937_64 g(seed);
    for (std::int64_t i = 0; i < n; ++i)
    {
        out[i] = g();
    }
}

// random_double() after seed_random(seed)
void oracle_random_doubles(std::uint64_t seed, std::int64_t n, double* out)
{
    seed_random(seed);
    for (std::int64_t i = 0; i < n; ++i)
    {
        out[i] = random_double();
    }
}

// The noise image standing in for earthmap.jpg, as linear floats (width x width/2 x 3)
void oracle_noise_image(int width, std::uint32_t seed, float* out)
{
    const image img = noise_image(width, width / 2, seed);
    std::size_t k = 0;
    for (const color& p : img.pixels)
    {
        out[k++] = static_cast<float>(p.x);
        out[k++] = static_cast<float>(p.y);
        out[k++] = static_cast<float>(p.z);
    }
}
}

#ifdef ORACLE_MAIN
// rt_cpu SCENE [-w N] [-s N] [-p uniform|material|lights|mixture] [--seed N] [-t THREADS] [-o FILE.pfm]
int main(int argc, char** argv)
{
    std::string name = "weekend";
    int width = 0, spp = 0, sampling = -1, threads = 0;
    std::uint32_t seed = 0;
    std::string output;
    for (int i = 1; i < argc; ++i)
    {
        const std::string a = argv[i];
        auto next = [&]() { return std::string(i + 1 < argc ? argv[++i] : ""); };
        if (a == "-w")
            width = std::stoi(next());
        else if (a == "-s")
            spp = std::stoi(next());
        else if (a == "--seed")
            seed = static_cast<std::uint32_t>(std::stoul(next()));
        else if (a == "-t")
            threads = std::stoi(next());
        else if (a == "-o")
            output = next();
        else if (a == "-p")
        {
            const std::string p = next();
            sampling = p == "uniform" ? 0 : p == "material" ? 1 : p == "lights" ? 2 : 3;
        }
        else
            name = a;
    }
    std::int32_t dims[3];
    if (oracle_render(name.c_str(), width, 0, spp, sampling, 0, seed, threads, 0, 0, 0, 0, nullptr, 0, dims, nullptr))
    {
        std::fprintf(stderr, "Error: %s\n", oracle_last_error());
        return 2;
    }
    std::vector<double> pixels(3ull * dims[0] * dims[1]);
    std::uint64_t rays = 0;
    const auto start = std::chrono::steady_clock::now();
    oracle_render(name.c_str(), width, 0, spp, sampling, 0, seed, threads, 0, 0, 0, 0, pixels.data(),
                  static_cast<std::int64_t>(pixels.size()), dims, &rays);
    const double seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
    double mean = 0;
    for (const double p : pixels)
    {
        mean += p;
    }
    mean /= static_cast<double>(pixels.size());
    std::printf("Rendering %dx%d, rays per pixel: %d\nRender time: %.3f s\nRays traced: %llu\nMean radiance: %.6f\n",
                dims[0], dims[1], dims[2] * dims[2], seconds, static_cast<unsigned long long>(rays), mean);
    if (!output.empty())
    {
        std::ofstream f(output, std::ios::binary);
        f << "PF\n" << dims[0] << " " << dims[1] << "\n-1.0\n";
        for (int y = dims[1] - 1; y >= 0; --y)
        {
            for (int x = 0; x < dims[0]; ++x)
            {
                for (int c = 0; c < 3; ++c)
                {
                    const float v = static_cast<float>(pixels[3 * (static_cast<std::size_t>(y) * dims[0] + x) + c]);
                    f.write(reinterpret_cast<const char*>(&v), 4);
                }
            }
        }
    }
    return 0;
}
#endif
