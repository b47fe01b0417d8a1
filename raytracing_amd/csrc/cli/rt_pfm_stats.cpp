// rt_pfm_stats: numbers about renders saved as PFM files, for the checks ImageCompare doesn't make: whether an image
// of a white furnace is exactly 1, how noisy a render is, the RMS rule of the two-strategy protocol, and whether two
// renders are the same bytes. The reference's PfmStats (tests/pfm_stats.cpp:1-60, docs/image-compare.md:95-98,132-);
// its statistics written here from their definitions.
//
//   rt_pfm_stats stats A.pfm [--rect X0 Y0 X1 Y1 [--expect V T]]... [--equal V [--share F]] [--max V]
//       the mean, its standard error, the smallest and the largest value of each channel (and of the channel
//       average) of the whole image or of each rectangle (columns X0..X1, rows Y0..Y1, inclusive, row 0 at the top).
//       --expect: the rectangle's mean of the channel average must lie within 4 standard errors + T of V. --equal:
//       at least the share F (default 0.9999) of the values of every region must be exactly V as a float. --max: no
//       value may exceed V.
//   rt_pfm_stats noise A1.pfm A2.pfm [--trim F] [--tiles N]
//       the noise of one render from two that differ only in their seed: d = A2 - A1 of the channel average has twice
//       one render's variance, so sigma = RMS(d) / sqrt(2); printed untrimmed and with the largest share F (default
//       0.01) of the d^2 left out (fireflies and lamp edges make the plain RMS vary between seeds). --tiles N: sigma
//       per tile of N x N pixels too.
//   rt_pfm_stats rule T1a T1b T2a T2b A1 A2 B1 B2 [--margin M] [--trim F]
//       the RMS rule for two strategies of unequal noise: the controls A1-A2 and B1-B2 measure R_A and R_B (trimmed
//       RMS of their channel-average d), and each test pair (T1a, T1b), (T2a, T2b), A against B with independent
//       seeds, must have a trimmed RMS of d of at most M sqrt((R_A^2 + R_B^2) / 2), M = 1.10 by default.
//   rt_pfm_stats same A.pfm B.pfm [--rect X0 Y0 X1 Y1]... [--ulps N --max-share F]
//       the same floats in every pixel (of the rectangles); with --ulps, at most the share F of the pixels apart, by
//       at most N units in the last place in any channel.
//   rt_pfm_stats self-test
//       checks this program's own arithmetic on small images it writes.
// Exit code: 0 the check passes, 1 it fails, 2 an error (a file that can't be read, sizes that differ, a NaN or an
// infinity in a render, a malformed option).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <optional>
#include <stdexcept>
#include <string>
#include <vector>

#include "pfm.hpp"

namespace
{
using rt::cli::pfm_image;

struct usage_error : std::runtime_error
{
    using std::runtime_error::runtime_error;
};

pfm_image read(const std::string& path)
{
    pfm_image im;
    std::string error;
    if (!rt::cli::load_pfm(path, im, error))
    {
        throw usage_error(error);
    }
    for (std::size_t k = 0; k < im.rgb.size(); ++k)
    {
        if (!std::isfinite(im.rgb[k]))
        {
            const std::size_t p = k / 3;
            throw usage_error(path + ": pixel (" + std::to_string(p % im.width) + ", " + std::to_string(p / im.width) +
                              ") is not finite: a broken render");
        }
    }
    return im;
}

void same_size(const pfm_image& a, const pfm_image& b)
{
    if (a.width != b.width || a.height != b.height)
    {
        throw usage_error("the images have different sizes");
    }
}

double number(const char* text, const char* option)
{
    char* end = nullptr;
    const double v = std::strtod(text, &end);
    if (end == text || *end != '\0' || !std::isfinite(v))
    {
        throw usage_error(std::string(option) + ": expected a number, not '" + text + "'");
    }
    return v;
}
int integer(const char* text, const char* option)
{
    const double v = number(text, option);
    if (v != std::floor(v) || v < 0 || v > 1 << 20)
    {
        throw usage_error(std::string(option) + ": expected a whole number from 0, not '" + text + "'");
    }
    return static_cast<int>(v);
}

// A rectangle of pixels, inclusive bounds, and what its mean is expected to be
struct region
{
    int x0, y0, x1, y1;
    std::optional<double> expect;
    double tolerance = 0;
};
region whole(const pfm_image& im) { return {0, 0, im.width - 1, im.height - 1, std::nullopt, 0}; }
void check_inside(const region& r, const pfm_image& im)
{
    if (r.x0 > r.x1 || r.y0 > r.y1 || r.x0 < 0 || r.y0 < 0 || r.x1 >= im.width || r.y1 >= im.height)
    {
        throw usage_error("a rectangle outside the image, or empty");
    }
}

// Sums for a mean and its standard error, Welford's way (no cancellation when the spread is tiny)
struct moments
{
    int64_t n = 0;
    double mean = 0, m2 = 0, lo = INFINITY, hi = -INFINITY;
    void add(double x)
    {
        ++n;
        const double delta = x - mean;
        mean += delta / static_cast<double>(n);
        m2 += delta * (x - mean);
        lo = std::min(lo, x);
        hi = std::max(hi, x);
    }
    double se() const { return n > 1 ? std::sqrt(m2 / static_cast<double>(n - 1) / static_cast<double>(n)) : 0.0; }
};

double value(const pfm_image& im, int x, int y, int c)
{
    return im.rgb[3 * (static_cast<std::size_t>(y) * im.width + x) + c];
}
double average(const pfm_image& im, int x, int y)
{
    return (value(im, x, y, 0) + value(im, x, y, 1) + value(im, x, y, 2)) / 3.0;
}

// The channel-average differences B - A over the whole image
std::vector<double> differences(const pfm_image& a, const pfm_image& b)
{
    same_size(a, b);
    std::vector<double> d(static_cast<std::size_t>(a.width) * a.height);
    for (int y = 0; y < a.height; ++y)
    {
        for (int x = 0; x < a.width; ++x)
        {
            d[static_cast<std::size_t>(y) * a.width + x] = average(b, x, y) - average(a, x, y);
        }
    }
    return d;
}
// sqrt(mean d^2), leaving out the largest share `trim` of the d^2 (the count left out rounded down)
double trimmed_rms(std::vector<double> d, double trim)
{
    for (double& x : d)
    {
        x *= x;
    }
    std::sort(d.begin(), d.end());
    const std::size_t drop = static_cast<std::size_t>(std::floor(trim * static_cast<double>(d.size())));
    const std::size_t keep = d.size() - std::min(drop, d.size() - 1);
    double sum = 0;
    for (std::size_t k = 0; k < keep; ++k)
    {
        sum += d[k];
    }
    return std::sqrt(sum / static_cast<double>(keep));
}

// --- stats --------------------------------------------------------------------------------------------------------
int stats(const std::vector<std::string>& args)
{
    if (args.empty())
    {
        throw usage_error("stats needs a file");
    }
    const pfm_image im = read(args[0]);
    std::vector<region> regions;
    std::optional<double> equal, maximum;
    double share = 0.9999;
    for (std::size_t k = 1; k < args.size(); ++k)
    {
        const auto need = [&](std::size_t count) {
            if (k + count >= args.size())
            {
                throw usage_error(args[k] + " needs " + std::to_string(count) + " values");
            }
        };
        if (args[k] == "--rect")
        {
            need(4);
            region r{integer(args[k + 1].c_str(), "--rect"), integer(args[k + 2].c_str(), "--rect"),
                     integer(args[k + 3].c_str(), "--rect"), integer(args[k + 4].c_str(), "--rect"), std::nullopt, 0};
            check_inside(r, im);
            regions.push_back(r);
            k += 4;
        }
        else if (args[k] == "--expect")
        {
            need(2);
            if (regions.empty())
            {
                throw usage_error("--expect follows the --rect it is about");
            }
            regions.back().expect = number(args[k + 1].c_str(), "--expect");
            regions.back().tolerance = number(args[k + 2].c_str(), "--expect");
            k += 2;
        }
        else if (args[k] == "--equal")
        {
            need(1);
            equal = number(args[++k].c_str(), "--equal");
        }
        else if (args[k] == "--share")
        {
            need(1);
            share = number(args[++k].c_str(), "--share");
            if (!(share >= 0 && share <= 1))
            {
                throw usage_error("--share: a share from 0 to 1");
            }
        }
        else if (args[k] == "--max")
        {
            need(1);
            maximum = number(args[++k].c_str(), "--max");
        }
        else
        {
            throw usage_error("stats: unknown option '" + args[k] + "'");
        }
    }
    if (regions.empty())
    {
        regions.push_back(whole(im));
    }
    bool pass = true;
    const char* names[4] = {"red", "green", "blue", "average"};
    for (const region& r : regions)
    {
        moments m[4];
        int64_t equal_count = 0, values = 0;
        bool over = false;
        for (int y = r.y0; y <= r.y1; ++y)
        {
            for (int x = r.x0; x <= r.x1; ++x)
            {
                for (int c = 0; c < 3; ++c)
                {
                    const double v = value(im, x, y, c);
                    m[c].add(v);
                    ++values;
                    equal_count += equal && static_cast<float>(v) == static_cast<float>(*equal) ? 1 : 0;
                    over = over || (maximum && v > *maximum);
                }
                m[3].add(average(im, x, y));
            }
        }
        std::printf("%s: %dx%d, columns %d-%d, rows %d-%d\n", args[0].c_str(), im.width, im.height, r.x0, r.x1, r.y0,
                    r.y1);
        std::printf("              mean           SE          min          max\n");
        for (int c = 0; c < 4; ++c)
        {
            std::printf("  %-8s %12.8f %12.8f %12.6g %12.6g\n", names[c], m[c].mean, m[c].se(), m[c].lo, m[c].hi);
        }
        if (r.expect)
        {
            const double limit = 4 * m[3].se() + r.tolerance;
            const bool ok = std::fabs(m[3].mean - *r.expect) <= limit;
            std::printf("  expect %.8f: |mean - expected| %.3g, limit %.3g (4 SE + %g): %s\n", *r.expect,
                        std::fabs(m[3].mean - *r.expect), limit, r.tolerance, ok ? "passed" : "FAILED");
            pass = pass && ok;
        }
        if (equal)
        {
            const double got = static_cast<double>(equal_count) / static_cast<double>(values);
            const bool ok = got >= share;
            std::printf("  equal to %.9g: %lld of %lld values (%.6f, at least %.6f): %s\n", *equal,
                        static_cast<long long>(equal_count), static_cast<long long>(values), got, share,
                        ok ? "passed" : "FAILED");
            pass = pass && ok;
        }
        if (maximum)
        {
            std::printf("  at most %.9g: %s\n", *maximum, over ? "FAILED" : "passed");
            pass = pass && !over;
        }
    }
    return pass ? 0 : 1;
}

// --- noise -------------------------------------------------------------------------------------------------------
int noise(const std::vector<std::string>& args)
{
    if (args.size() < 2)
    {
        throw usage_error("noise needs two files");
    }
    const pfm_image a = read(args[0]), b = read(args[1]);
    double trim = 0.01;
    int tiles = 0;
    for (std::size_t k = 2; k < args.size(); ++k)
    {
        if (args[k] == "--trim" && k + 1 < args.size())
        {
            trim = number(args[++k].c_str(), "--trim");
            if (!(trim >= 0 && trim < 1))
            {
                throw usage_error("--trim: a share from 0 to below 1");
            }
        }
        else if (args[k] == "--tiles" && k + 1 < args.size())
        {
            tiles = integer(args[++k].c_str(), "--tiles");
            if (tiles < 1)
            {
                throw usage_error("--tiles: at least 1 pixel");
            }
        }
        else
        {
            throw usage_error("noise: unknown option '" + args[k] + "'");
        }
    }
    const std::vector<double> d = differences(a, b);
    const double rms = trimmed_rms(d, 0.0), trimmed = trimmed_rms(d, trim);
    std::printf("noise of one render (sigma = RMS(d) / sqrt 2, d of the channel average, %zu pixels)\n", d.size());
    std::printf("  untrimmed  RMS(d) %.8f  sigma %.8f\n", rms, rms / std::sqrt(2.0));
    std::printf("  trimmed    RMS(d) %.8f  sigma %.8f  (the largest %.4g of d^2 left out)\n", trimmed,
                trimmed / std::sqrt(2.0), trim);
    if (tiles > 0)
    {
        std::printf("sigma per tile of %dx%d (untrimmed), rows of tiles from the top:\n", tiles, tiles);
        for (int ty = 0; ty < a.height; ty += tiles)
        {
            for (int tx = 0; tx < a.width; tx += tiles)
            {
                double sum = 0;
                int64_t n = 0;
                for (int y = ty; y < std::min(ty + tiles, a.height); ++y)
                {
                    for (int x = tx; x < std::min(tx + tiles, a.width); ++x)
                    {
                        const double v = d[static_cast<std::size_t>(y) * a.width + x];
                        sum += v * v;
                        ++n;
                    }
                }
                std::printf(" %.5f", std::sqrt(sum / static_cast<double>(n)) / std::sqrt(2.0));
            }
            std::printf("\n");
        }
    }
    return 0;
}

// --- rule --------------------------------------------------------------------------------------------------------
int rule(const std::vector<std::string>& args)
{
    if (args.size() < 8)
    {
        throw usage_error("rule needs eight files: T1a T1b T2a T2b A1 A2 B1 B2");
    }
    double margin = 1.10, trim = 0.01;
    for (std::size_t k = 8; k < args.size(); ++k)
    {
        if (args[k] == "--margin" && k + 1 < args.size())
        {
            margin = number(args[++k].c_str(), "--margin");
        }
        else if (args[k] == "--trim" && k + 1 < args.size())
        {
            trim = number(args[++k].c_str(), "--trim");
            if (!(trim >= 0 && trim < 1))
            {
                throw usage_error("--trim: a share from 0 to below 1");
            }
        }
        else
        {
            throw usage_error("rule: unknown option '" + args[k] + "'");
        }
    }
    std::vector<pfm_image> im;
    for (int k = 0; k < 8; ++k)
    {
        im.push_back(read(args[k]));
    }
    const double ra = trimmed_rms(differences(im[4], im[5]), trim);
    const double rb = trimmed_rms(differences(im[6], im[7]), trim);
    const double limit = margin * std::sqrt((ra * ra + rb * rb) / 2);
    std::printf("controls: R_A %.8f (A1-A2), R_B %.8f (B1-B2); limit %.4g x sqrt((R_A^2 + R_B^2) / 2) = %.8f\n", ra,
                rb, margin, limit);
    bool pass = true;
    for (int t = 0; t < 2; ++t)
    {
        const double r = trimmed_rms(differences(im[2 * t], im[2 * t + 1]), trim);
        const bool ok = r <= limit;
        std::printf("test %d: trimmed RMS of d %.8f (%.3f of the limit): %s\n", t + 1, r, r / limit,
                    ok ? "passed" : "FAILED");
        pass = pass && ok;
    }
    return pass ? 0 : 1;
}

// --- same --------------------------------------------------------------------------------------------------------
// The distance in units in the last place between two floats (their bit patterns in one ordered integer line)
int64_t ulps_apart(float a, float b)
{
    const auto ordered = [](float f) {
        int32_t i;
        std::memcpy(&i, &f, 4);
        return i < 0 ? static_cast<int64_t>(INT32_MIN) - static_cast<int64_t>(i) : static_cast<int64_t>(i);
    };
    const int64_t d = ordered(a) - ordered(b);
    return d < 0 ? -d : d;
}

int same(const std::vector<std::string>& args)
{
    if (args.size() < 2)
    {
        throw usage_error("same needs two files");
    }
    const pfm_image a = read(args[0]), b = read(args[1]);
    same_size(a, b);
    std::vector<region> regions;
    std::optional<int> ulps;
    double max_share = 0;
    for (std::size_t k = 2; k < args.size(); ++k)
    {
        if (args[k] == "--rect" && k + 4 < args.size())
        {
            region r{integer(args[k + 1].c_str(), "--rect"), integer(args[k + 2].c_str(), "--rect"),
                     integer(args[k + 3].c_str(), "--rect"), integer(args[k + 4].c_str(), "--rect"), std::nullopt, 0};
            check_inside(r, a);
            regions.push_back(r);
            k += 4;
        }
        else if (args[k] == "--ulps" && k + 1 < args.size())
        {
            ulps = integer(args[++k].c_str(), "--ulps");
        }
        else if (args[k] == "--max-share" && k + 1 < args.size())
        {
            max_share = number(args[++k].c_str(), "--max-share");
        }
        else
        {
            throw usage_error("same: unknown option '" + args[k] + "'");
        }
    }
    if (regions.empty())
    {
        regions.push_back(whole(a));
    }
    int64_t pixels = 0, apart = 0, worst = 0;
    for (const region& r : regions)
    {
        for (int y = r.y0; y <= r.y1; ++y)
        {
            for (int x = r.x0; x <= r.x1; ++x)
            {
                ++pixels;
                int64_t here = 0;
                for (int c = 0; c < 3; ++c)
                {
                    here = std::max(here, ulps_apart(static_cast<float>(value(a, x, y, c)),
                                                     static_cast<float>(value(b, x, y, c))));
                }
                apart += here > 0 ? 1 : 0;
                worst = std::max(worst, here);
            }
        }
    }
    const double share = static_cast<double>(apart) / static_cast<double>(pixels);
    std::printf("%lld of %lld pixels differ (%.6f), by at most %lld ulps\n", static_cast<long long>(apart),
                static_cast<long long>(pixels), share, static_cast<long long>(worst));
    const bool pass = ulps ? (share <= max_share && worst <= *ulps) : apart == 0;
    std::printf("%s\n", pass ? "passed" : "FAILED");
    return pass ? 0 : 1;
}

// --- self-test ---------------------------------------------------------------------------------------------------
int run(const std::vector<std::string>& args);

int self_test()
{
    const std::filesystem::path dir = std::filesystem::temp_directory_path() /
                                      ("rt_pfm_stats_" + std::to_string(static_cast<long long>(std::rand())));
    std::filesystem::create_directories(dir);
    const auto file = [&](const char* name) { return (dir / name).string(); };
    const auto write = [&](const char* name, int w, int h, auto f) {
        std::vector<double> rgb(3ull * w * h);
        for (int y = 0; y < h; ++y)
        {
            for (int x = 0; x < w; ++x)
            {
                for (int c = 0; c < 3; ++c)
                {
                    rgb[3 * (static_cast<std::size_t>(y) * w + x) + c] = f(x, y, c);
                }
            }
        }
        rt::cli::save_pfm(file(name), w, h, rgb);
    };
    int failures = 0;
    const auto expect = [&](const char* what, bool ok) {
        std::printf("self-test %s: %s\n", what, ok ? "ok" : "FAILED");
        failures += ok ? 0 : 1;
    };
    const auto quiet = [&](const std::vector<std::string>& a) {
        std::fflush(stdout);
        return run(a);
    };
    write("ones.pfm", 16, 8, [](int, int, int) { return 1.0; });
    write("one_off.pfm", 16, 8, [](int x, int y, int) { return x == 3 && y == 2 ? 0.5 : 1.0; });
    // two renders whose channel averages differ by +-s pixel by pixel: sigma = s / sqrt 2 exactly
    write("a.pfm", 16, 8, [](int, int, int) { return 0.5; });
    write("b.pfm", 16, 8, [](int x, int y, int) { return (x + y) % 2 ? 0.5 + 0.125 : 0.5 - 0.125; });
    write("c.pfm", 16, 8, [](int x, int y, int) { return (x + y) % 2 ? 0.5 + 0.25 : 0.5 - 0.25; });
    write("next.pfm", 16, 8, [](int, int, int) { return static_cast<double>(std::nextafter(0.5f, 1.0f)); });
    expect("a white furnace is all ones", quiet({"stats", file("ones.pfm"), "--equal", "1", "--share", "1"}) == 0);
    expect("one value off fails share 1", quiet({"stats", file("one_off.pfm"), "--equal", "1", "--share", "1"}) == 1);
    expect("one value off passes share 0.99", quiet({"stats", file("one_off.pfm"), "--equal", "1", "--share", "0.99"}) == 0);
    expect("a rectangle's mean", quiet({"stats", file("b.pfm"), "--rect", "0", "0", "15", "7", "--expect", "0.5", "0"}) == 0);
    expect("a wrong mean fails", quiet({"stats", file("b.pfm"), "--rect", "0", "0", "15", "7", "--expect", "0.6", "0.01"}) == 1);
    expect("--max", quiet({"stats", file("c.pfm"), "--max", "0.7"}) == 1 && quiet({"stats", file("c.pfm"), "--max", "0.75"}) == 0);
    expect("same images", quiet({"same", file("a.pfm"), file("a.pfm")}) == 0);
    expect("one ulp apart", quiet({"same", file("a.pfm"), file("next.pfm")}) == 1 &&
                                quiet({"same", file("a.pfm"), file("next.pfm"), "--ulps", "1", "--max-share", "1"}) == 0);
    {
        const pfm_image a = read(file("a.pfm")), b = read(file("b.pfm"));
        const double sigma = trimmed_rms(differences(a, b), 0.0) / std::sqrt(2.0);
        expect("noise of +-1/8 is 1/(8 sqrt 2)", std::fabs(sigma - 0.125 / std::sqrt(2.0)) < 1e-12);
    }
    {
        std::vector<double> d(100, 1.0);
        d[7] = 100.0;
        expect("trimming 1 % drops the outlier", std::fabs(trimmed_rms(d, 0.01) - 1.0) < 1e-12 &&
                                                     trimmed_rms(d, 0.0) > 10.0);
    }
    // the RMS rule: tests as noisy as the controls pass; a test twice as far apart fails
    expect("rule passes equal noise", quiet({"rule", file("a.pfm"), file("b.pfm"), file("a.pfm"), file("b.pfm"),
                                             file("a.pfm"), file("b.pfm"), file("a.pfm"), file("b.pfm"), "--trim", "0"}) == 0);
    expect("rule fails a pattern lost", quiet({"rule", file("a.pfm"), file("c.pfm"), file("a.pfm"), file("b.pfm"),
                                               file("a.pfm"), file("b.pfm"), file("a.pfm"), file("b.pfm"), "--trim", "0"}) == 1);
    std::filesystem::remove_all(dir);
    std::printf("self-test: %s\n", failures == 0 ? "passed" : "FAILED");
    return failures == 0 ? 0 : 1;
}

void usage()
{
    std::printf("usage: rt_pfm_stats stats A.pfm [--rect X0 Y0 X1 Y1 [--expect V T]]... [--equal V [--share F]] "
                "[--max V]\n"
                "       rt_pfm_stats noise A1.pfm A2.pfm [--trim F] [--tiles N]\n"
                "       rt_pfm_stats rule T1a T1b T2a T2b A1 A2 B1 B2 [--margin M] [--trim F]\n"
                "       rt_pfm_stats same A.pfm B.pfm [--rect X0 Y0 X1 Y1]... [--ulps N --max-share F]\n"
                "       rt_pfm_stats self-test\n"
                "exit code: 0 the check passes, 1 it fails, 2 an error\n");
}

int run(const std::vector<std::string>& args)
{
    try
    {
        if (args.empty())
        {
            usage();
            return 2;
        }
        const std::vector<std::string> rest(args.begin() + 1, args.end());
        if (args[0] == "stats")
        {
            return stats(rest);
        }
        if (args[0] == "noise")
        {
            return noise(rest);
        }
        if (args[0] == "rule")
        {
            return rule(rest);
        }
        if (args[0] == "same")
        {
            return same(rest);
        }
        if (args[0] == "self-test")
        {
            return self_test();
        }
        if (args[0] == "-h" || args[0] == "--help")
        {
            usage();
            return 0;
        }
        throw usage_error("unknown command '" + args[0] + "'");
    }
    catch (const usage_error& e)
    {
        std::printf("Error: %s\n", e.what());
        return 2;
    }
}
}

int main(int argc, char** argv) { return run(std::vector<std::string>(argv + 1, argv + argc)); }
