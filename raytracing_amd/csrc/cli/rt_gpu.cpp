// rt_gpu: the RayTracingGPU command line (src/main_gpu.cpp:32-294,362-640) over the C ABI: a built-in scene or a glTF
// 2.0 model's file (FILE.glb or FILE.gltf, rendered in a studio), its overrides, a render, the numbers the reference
// prints, and the image as PFM (linear floats) or PNG (the books' bytes).
//
//   rt_gpu [scene] [-w N] [--height N] [-s N] [--strata K] [-p uniform|material|lights|mixture|nee|mis]
//          [--seed N] [--fp64] [--integrator auto|megakernel|wavefront|plain] [--launch-ms N] [-o FILE.pfm|FILE.png]...
//          [--tonemap books|none|agx|neutral] [--exposure EV] [--all-books-kernel | --quality-kernel]
// Exit code: 0 rendered and saved, 2 an error (an option out of range or malformed, a scene the library refuses, a
// file that couldn't be written).
#include <cerrno>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "../../../include/raytracing_amd.h"
#include "pfm.hpp"

namespace
{
uint32_t crc32(const uint8_t* data, std::size_t n, uint32_t crc = 0)
{
    crc = ~crc;
    for (std::size_t i = 0; i < n; ++i)
    {
        crc ^= data[i];
        for (int k = 0; k < 8; ++k)
        {
            crc = (crc >> 1) ^ (0xEDB88320u & (0u - (crc & 1u)));
        }
    }
    return ~crc;
}

void put32(std::vector<uint8_t>& v, uint32_t x)
{
    for (int s = 24; s >= 0; s -= 8)
    {
        v.push_back(static_cast<uint8_t>(x >> s));
    }
}

void chunk(std::ofstream& f, const char* type, const std::vector<uint8_t>& data)
{
    std::vector<uint8_t> out;
    put32(out, static_cast<uint32_t>(data.size()));
    std::vector<uint8_t> body(type, type + 4);
    body.insert(body.end(), data.begin(), data.end());
    out.insert(out.end(), body.begin(), body.end());
    put32(out, crc32(body.data(), body.size()));
    f.write(reinterpret_cast<const char*>(out.data()), static_cast<std::streamsize>(out.size()));
}

// An RGB PNG with uncompressed (stored) deflate blocks
bool save_png(const std::string& path, int w, int h, const std::vector<uint8_t>& rgba)
{
    std::ofstream f(path, std::ios::binary);
    if (!f)
    {
        return false;
    }
    const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    f.write(reinterpret_cast<const char*>(sig), 8);
    std::vector<uint8_t> ihdr;
    put32(ihdr, static_cast<uint32_t>(w));
    put32(ihdr, static_cast<uint32_t>(h));
    ihdr.insert(ihdr.end(), {8, 2, 0, 0, 0});
    chunk(f, "IHDR", ihdr);
    std::vector<uint8_t> raw;
    for (int y = 0; y < h; ++y)
    {
        raw.push_back(0);
        for (int x = 0; x < w; ++x)
        {
            const uint8_t* p = &rgba[4 * (static_cast<std::size_t>(y) * w + x)];
            raw.insert(raw.end(), p, p + 3);
        }
    }
    std::vector<uint8_t> z{0x78, 0x01};
    uint32_t a = 1, b = 0;
    for (const uint8_t c : raw)
    {
        a = (a + c) % 65521;
        b = (b + a) % 65521;
    }
    for (std::size_t off = 0; off < raw.size() || off == 0; off += 65535)
    {
        const std::size_t len = std::min<std::size_t>(65535, raw.size() - off);
        const bool last = off + len >= raw.size();
        z.push_back(last ? 1 : 0);
        z.push_back(static_cast<uint8_t>(len & 255));
        z.push_back(static_cast<uint8_t>(len >> 8));
        z.push_back(static_cast<uint8_t>(~len & 255));
        z.push_back(static_cast<uint8_t>((~len >> 8) & 255));
        z.insert(z.end(), raw.begin() + static_cast<std::ptrdiff_t>(off),
                 raw.begin() + static_cast<std::ptrdiff_t>(off + len));
        if (last)
        {
            break;
        }
    }
    put32(z, (b << 16) | a);
    chunk(f, "IDAT", z);
    chunk(f, "IEND", {});
    return static_cast<bool>(f);
}

int fail(const char* what)
{
    std::printf("Error: %s: %s\n", what, rt_last_error());
    return 2;
}

// A whole number of an option, or the exit with a message (src/main_gpu.cpp refuses malformed numbers alike)
long long whole_number(const std::string& option, const std::string& text, long long lo, long long hi)
{
    char* end = nullptr;
    errno = 0;
    const long long v = std::strtoll(text.c_str(), &end, 10);
    if (text.empty() || *end != '\0' || errno != 0 || v < lo || v > hi)
    {
        std::printf("Error: %s: expected an integer from %lld to %lld, not '%s'\n", option.c_str(), lo, hi,
                    text.c_str());
        std::exit(2);
    }
    return v;
}
double real_number(const std::string& option, const std::string& text)
{
    char* end = nullptr;
    const double v = std::strtod(text.c_str(), &end);
    if (text.empty() || *end != '\0' || !std::isfinite(v))
    {
        std::printf("Error: %s: expected a number, not '%s'\n", option.c_str(), text.c_str());
        std::exit(2);
    }
    return v;
}
}

int main(int argc, char** argv)
{
    std::string scene = "weekend";
    rt_scene_overrides ov{};
    ov.sampling = -1;
    rt_session_options so{};
    std::vector<std::string> outputs;
    int32_t tone_map = RT_TONE_BOOKS;
    double exposure = 0;
    bool all_books_kernel = false;
    bool quality_kernel = false;
    for (int i = 1; i < argc; ++i)
    {
        const std::string a = argv[i];
        const auto value = [&]() -> std::string {
            if (i + 1 >= argc)
            {
                std::printf("Error: %s needs a value\n", a.c_str());
                std::exit(2);
            }
            return argv[++i];
        };
        if (a == "-h" || a == "--help")
        {
            std::printf("Usage: rt_gpu [scene] [-w N] [--height N] [-s N] [--strata K] [-p STRATEGY] [--seed N]\n"
                        "              [--fp64] [--integrator auto|megakernel|wavefront|plain] [--launch-ms N] [-o FILE]\n"
                        "              [--all-books-kernel | --quality-kernel]\n"
                        "              [--tonemap books|none|agx|neutral] [--exposure EV]\n"
                        "Scenes: %s\n"
                        "Or a glTF 2.0 model, FILE.glb or FILE.gltf, rendered in a studio\n",
                        rt_scene_names());
            return 0;
        }
        if (a == "-w" || a == "--width")
            ov.image_width = static_cast<int32_t>(whole_number(a, value(), 1, 65536));
        else if (a == "--height")
            ov.image_height = static_cast<int32_t>(whole_number(a, value(), 1, 65536));
        else if (a == "-s" || a == "--samples")
            ov.samples_per_pixel = static_cast<int32_t>(whole_number(a, value(), 1, 1 << 30));
        else if (a == "--strata")
            ov.strata = static_cast<int32_t>(whole_number(a, value(), 1, 46340));
        else if (a == "--depth")
            ov.max_depth = static_cast<int32_t>(whole_number(a, value(), 1, 1 << 20));
        else if (a == "--seed")
            so.seed = static_cast<uint64_t>(whole_number(a, value(), 0, INT64_MAX));
        else if (a == "--fp64")
            so.precision = RT_FLOAT64;
        else if (a == "--all-books-kernel")
            all_books_kernel = true;
        else if (a == "--quality-kernel")
            quality_kernel = true;
        else if (a == "--launch-ms")
            so.launch_seconds = static_cast<double>(whole_number(a, value(), 1, 500)) / 1000.0;
        else if (a == "-o" || a == "--output")
            outputs.push_back(value());
        else if (a == "--exposure")
            exposure = real_number(a, value());
        else if (a == "--tonemap")
        {
            const std::string v = value();
            const char* names[] = {"books", "none", "agx", "neutral"};
            tone_map = -1;
            for (int k = 0; k < 4; ++k)
            {
                if (v == names[k])
                {
                    tone_map = k;
                }
            }
            if (tone_map < 0)
            {
                std::printf("Error: unknown tone map '%s'\n", v.c_str());
                return 2;
            }
        }
        else if (a == "--integrator")
        {
            const std::string v = value();
            const char* names[] = {"auto", "megakernel", "wavefront", "plain"};
            so.integrator = -1;
            for (int k = 0; k < 4; ++k)
            {
                if (v == names[k])
                {
                    so.integrator = k;
                }
            }
            if (so.integrator < 0)
            {
                std::printf("Error: unknown integrator '%s'\n", v.c_str());
                return 2;
            }
        }
        else if (a == "-p" || a == "--sampling")
        {
            const std::string v = value();
            const char* names[] = {"uniform", "material", "lights", "mixture", "nee", "mis"};
            ov.sampling = -2;
            for (int k = 0; k < 6; ++k)
            {
                if (v == names[k])
                {
                    ov.sampling = k;
                }
            }
            if (ov.sampling == -2)
            {
                std::printf("Error: unknown strategy '%s'\n", v.c_str());
                return 2;
            }
        }
        else if (!a.empty() && a[0] != '-')
            scene = a;
        else
        {
            std::printf("Error: unknown option '%s'\n", a.c_str());
            return 2;
        }
    }
    // the kernel asked for (src/main_gpu.cpp:395-400)
    if (all_books_kernel && quality_kernel)
    {
        std::printf("Error: --all-books-kernel picks the books' kernel for all three books; --quality-kernel needs the "
                    "quality kernel\n");
        return 2;
    }
    so.kernel = all_books_kernel ? RT_KERNEL_ALL_BOOKS : quality_kernel ? RT_KERNEL_QUALITY : RT_KERNEL_AUTO;
    if (outputs.empty())
    {
        outputs.push_back(scene + "-gpu.png");
    }
    rt_scene_handle* h = nullptr;
    if (rt_scene_build(scene.c_str(), &ov, &h) != RT_OK)
    {
        return fail("building the scene");
    }
    const rt_scene* s = rt_scene_get(h);
    // what a model holds that isn't rendered, one line each, before it renders (src/main_gpu.cpp:489)
    if (const char* warnings = rt_scene_warnings(h); warnings[0] != '\0')
    {
        std::printf("The model's parts that aren't rendered:\n%s", warnings);
    }
    rt_device_info dev{};
    if (rt_query_device(&dev) != RT_OK)
    {
        return fail("starting the GPU");
    }
    std::printf("Scene: %s\nGPU: %s (%s), %d compute units, %.0f GiB\n", scene.c_str(), dev.name, dev.arch,
                dev.compute_units, dev.memory_gib);
    const auto start = std::chrono::steady_clock::now();
    rt_session* session = nullptr;
    if (rt_session_create(s, &so, &session) != RT_OK)
    {
        return fail("creating the session");
    }
    const int w = s->camera.image_width, hgt = s->camera.image_height;
    const int64_t passes = rt_scene_passes(h);
    std::printf("Rendering %dx%d, rays per pixel: %lld\n", w, hgt, static_cast<long long>(passes));
    // the passes left, counted down as launches complete them, each count as wide as the first so that a shorter one
    // overwrites it whole; the line is erased when the render ends (src/main_gpu.cpp:516-550)
    int digits = static_cast<int>(std::to_string(passes).size());
    const rt_progress_fn show_progress = [](int64_t done, int64_t total, void* user) {
        if (done < total)
        {
            std::printf("\rPasses remaining: %*lld", *static_cast<int*>(user), static_cast<long long>(total - done));
            std::fflush(stdout);
        }
    };
    show_progress(0, passes, &digits);
    rt_render_report report{};
    const int rendered = rt_session_render_to_progress(session, passes, 0, &report, show_progress, &digits);
    std::printf("\r%30s\r", "");
    if (rendered != RT_OK)
    {
        return fail("rendering");
    }
    std::vector<double> pixels(3ull * w * hgt);
    if (rt_session_pixels(session, pixels.data(), pixels.size()) != RT_OK)
    {
        return fail("reading the image");
    }
    const double seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
    uint64_t rays = 0, shadow = 0;
    double kernel_seconds = 0;
    rt_session_counters(session, &rays, &shadow, &kernel_seconds);
    rt_kernel_info k{};
    rt_session_kernel_info(session, &k);
    double mean = 0;
    for (const double p : pixels)
    {
        mean += p;
    }
    mean /= static_cast<double>(pixels.size());
    const double paths = static_cast<double>(w) * hgt * static_cast<double>(passes);
    std::printf("Render time: %.3f s (kernels %.3f s, %d launches)\n", seconds, kernel_seconds, report.launches);
    std::printf("Kernel: %s, %d registers, %d bytes of scratch, %d bytes of LDS, %d blocks of %d per CU\n", k.name,
                k.registers, k.local_bytes, k.lds_bytes, k.blocks_per_cu, k.threads_per_block);
    std::printf("Rays traced: %llu, %.2f per sample, %.3f G per second\n", static_cast<unsigned long long>(rays),
                static_cast<double>(rays) / paths, static_cast<double>(rays) / kernel_seconds / 1e9);
    std::printf("Paths: %.1f M per second\nMean radiance: %.6f\n", paths / kernel_seconds / 1e6, mean);
    std::vector<uint8_t> rgba(4ull * w * hgt);
    bool saved = true;
    for (const std::string& out : outputs)
    {
        bool ok = false;
        if (out.size() > 4 && (out.substr(out.size() - 4) == ".pfm" || out.substr(out.size() - 4) == ".PFM"))
        {
            ok = rt::cli::save_pfm(out, w, hgt, pixels);
        }
        else
        {
            ok = rt_session_display_toned(session, tone_map, exposure, rgba.data(), rgba.size()) == RT_OK &&
                 save_png(out, w, hgt, rgba);
        }
        std::printf("%s: %s\n", ok ? "Saved" : "Error: could not save", out.c_str());
        saved = saved && ok;
    }
    rt_session_destroy(session);
    rt_scene_free(h);
    return saved ? 0 : 2;
}
