// rt_image_compare: are two renders the same picture, up to Monte Carlo noise? The reference's ImageCompare command
// (src/image_compare.cpp), for renders saved as PFM files by rt_gpu (-o FILE.pfm) or any other renderer.
//
//   rt_image_compare A.pfm B.pfm [--tile N]
//
// d = B - A per pixel and channel. If A and B are independent renders of one scene (different seeds, or different
// renderers), every d has expected value 0, which is tested four ways, each with a false-alarm rate of 1/400
// (src/image_compare.cpp:403-458):
//   1. whole image: the mean of d per channel (red, green, blue and their average), as a z-score;
//   2. all tiles: the tiles' z-scores (N x N pixels, 32 by default, of the channel average, each clamped to +-40)
//      summed and divided by the square root of their number (Stouffer);
//   3. largest tile: the largest |z| of one tile, against the threshold for that many tiles (Bonferroni);
//   4. tile count: the number of tiles with |z| > 3, against the binomial count as rare as 1/400.
// A mean's z-score is its t statistic (the mean, less a rounding margin of 1e-6 of the pixels' level, over its
// standard error) turned into the normal z exactly as rare (src/image_compare.cpp:158-170,241-253). A sign test over
// the pixels that differ by more than 0.1 % of their level is printed beside (src/image_compare.cpp:263-297).
// Exit code: 0 consistent, 1 different, 2 an error (a file that can't be read, sizes that differ, identical images,
// NaN or infinite values).
//
// The statistics are this file's own: the regularized incomplete beta function by its continued fraction evaluated with
// the modified Lentz method (Lentz 1976; Press et al., Numerical Recipes, section 6.4), the normal tail by std::erfc and
// its inverse by bisection, the binomial tail through the incomplete beta function.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <limits>
#include <string>
#include <vector>

#include "pfm.hpp"

namespace
{
constexpr double rounding_margin = 1e-6;                     // of a pixel's level: float rounding, not noise
constexpr double sign_threshold = 1000 * rounding_margin;    // the sign test counts pixels apart by more than 0.1 %
constexpr double false_alarm = 1.0 / 400;                    // each check's false-alarm rate
constexpr double infinity = std::numeric_limits<double>::infinity();

using image = rt::cli::pfm_image;
using rt::cli::load_pfm;

// Welford's running mean and sum of squared deviations (one pass, no cancellation when the spread is tiny)
struct running
{
    int64_t n = 0;
    double mean = 0, m2 = 0, squares = 0, level = 0;
    void add(double x, double lvl)
    {
        ++n;
        const double delta = x - mean;
        mean += delta / static_cast<double>(n);
        m2 += delta * (x - mean);
        squares += x * x;
        level += lvl;
    }
    double variance() const { return n > 1 ? m2 / static_cast<double>(n - 1) : 0.0; }
    double rms() const { return n > 0 ? std::sqrt(squares / static_cast<double>(n)) : 0.0; }
    double mean_level() const { return n > 0 ? level / static_cast<double>(n) : 0.0; }
};

// The continued fraction of the regularized incomplete beta function, by the modified Lentz method
double beta_fraction(double a, double b, double x)
{
    const double tiny = 1e-300;
    double c = 1, d = 1 - (a + b) * x / (a + 1);
    d = 1 / (std::fabs(d) < tiny ? tiny : d);
    double h = d;
    for (int m = 1; m <= 10000; ++m)
    {
        const double m2 = 2.0 * m;
        double num = m * (b - m) * x / ((a + m2 - 1) * (a + m2));
        d = 1 + num * d;
        c = 1 + num / c;
        d = 1 / (std::fabs(d) < tiny ? tiny : d);
        c = std::fabs(c) < tiny ? tiny : c;
        h *= d * c;
        num = -(a + m) * (a + b + m) * x / ((a + m2) * (a + m2 + 1));
        d = 1 + num * d;
        c = 1 + num / c;
        d = 1 / (std::fabs(d) < tiny ? tiny : d);
        c = std::fabs(c) < tiny ? tiny : c;
        const double step = d * c;
        h *= step;
        if (std::fabs(step - 1) < 1e-15)
        {
            break;
        }
    }
    return h;
}

// I_x(a, b): the chance that a Beta(a, b) variable is at most x
double incomplete_beta(double a, double b, double x)
{
    if (x <= 0)
    {
        return 0;
    }
    if (x >= 1)
    {
        return 1;
    }
    const double log_front = std::lgamma(a + b) - std::lgamma(a) - std::lgamma(b) + a * std::log(x) + b * std::log1p(-x);
    if (x < (a + 1) / (a + b + 2))
    {
        return std::exp(log_front) * beta_fraction(a, b, x) / a;
    }
    return 1 - std::exp(log_front) * beta_fraction(b, a, 1 - x) / b;
}

// P(|Z| > z) for a standard normal Z, and its inverse: the z with P(|Z| > z) = p
double normal_tail(double z) { return std::erfc(z / std::sqrt(2.0)); }
double normal_quantile(double p)
{
    if (!(p > 0))
    {
        return infinity;
    }
    if (p >= 1)
    {
        return 0;
    }
    double lo = 0, hi = 40;
    while (normal_tail(hi) > p && hi < 1e4)
    {
        hi *= 2;
    }
    for (int k = 0; k < 200 && hi - lo > 1e-15 * hi; ++k)
    {
        const double mid = 0.5 * (lo + hi);
        (normal_tail(mid) > p ? lo : hi) = mid;
    }
    return 0.5 * (lo + hi);
}

// The normal z exactly as rare (two-sided) as Student's t with dof degrees of freedom
double normal_equivalent(double t, double dof)
{
    if (std::isinf(t))
    {
        return std::copysign(infinity, t);
    }
    const double p = incomplete_beta(dof / 2, 0.5, dof / (dof + t * t)); // P(|T| > |t|)
    const double z = p > 0 ? normal_quantile(p) : std::sqrt(dof * std::log1p(t * t / dof));
    return std::copysign(z, t);
}

// A mean of d beyond the rounding margin, in units of its standard error, as a normal z
double mean_z(const running& r)
{
    const double se = std::sqrt(r.variance() / static_cast<double>(r.n));
    const double beyond = std::fabs(r.mean) - rounding_margin * r.mean_level();
    if (!(beyond > 0))
    {
        return 0;
    }
    const double t = se > 0 ? beyond / se : infinity;
    return std::copysign(normal_equivalent(t, static_cast<double>(r.n - 1)), r.mean);
}

// P(X >= k) for X binomial(n, q)
double binomial_tail(int64_t n, double q, int64_t k)
{
    if (k <= 0)
    {
        return 1;
    }
    if (k > n)
    {
        return 0;
    }
    return incomplete_beta(static_cast<double>(k), static_cast<double>(n - k + 1), q);
}

void usage()
{
    std::printf("usage: rt_image_compare A.pfm B.pfm [--tile N]\n"
                "  A, B: independent renders of one scene at one size (different seeds, or different renderers)\n"
                "  --tile N: the side of the tiles tested one by one, in pixels (default 32)\n"
                "exit code: 0 consistent, 1 different, 2 an error\n");
}
}

int main(int argc, char** argv)
{
    std::vector<std::string> files;
    int tile = 32;
    for (int k = 1; k < argc; ++k)
    {
        const std::string arg = argv[k];
        if (arg == "-h" || arg == "--help")
        {
            usage();
            return 0;
        }
        if (arg == "--tile")
        {
            if (k + 1 >= argc || std::atoi(argv[k + 1]) < 2)
            {
                std::printf("Error: --tile needs an integer of at least 2\n");
                return 2;
            }
            tile = std::atoi(argv[++k]);
            continue;
        }
        files.push_back(arg);
    }
    if (files.size() != 2)
    {
        usage();
        return 2;
    }
    image a, b;
    std::string error;
    if (!load_pfm(files[0], a, error) || !load_pfm(files[1], b, error))
    {
        std::printf("Error: %s\n", error.c_str());
        return 2;
    }
    auto mean_radiance = [](const image& im) {
        double s = 0;
        for (const double v : im.rgb)
        {
            s += v;
        }
        return s / static_cast<double>(im.rgb.size());
    };
    std::printf("A: %s: %dx%d, mean radiance %.6f\n", files[0].c_str(), a.width, a.height, mean_radiance(a));
    std::printf("B: %s: %dx%d, mean radiance %.6f\n", files[1].c_str(), b.width, b.height, mean_radiance(b));
    if (a.width != b.width || a.height != b.height)
    {
        std::printf("Error: the images have different sizes\n");
        return 2;
    }
    if (static_cast<int64_t>(a.width) * a.height < 2)
    {
        std::printf("Error: an image of one pixel has no spread to compare with\n");
        return 2;
    }
    for (std::size_t k = 0; k < a.rgb.size(); ++k)
    {
        if (!std::isfinite(a.rgb[k]) || !std::isfinite(b.rgb[k]))
        {
            std::printf("Error: values that are NaN or infinite\n");
            return 2;
        }
    }

    // the whole image per channel (red, green, blue, average), and the sign test on the average
    const int w = a.width, h = a.height;
    running whole[4];
    int64_t b_brighter = 0, a_brighter = 0;
    bool identical = true;
    std::vector<double> avg(static_cast<std::size_t>(w) * h), lvl(avg.size());
    for (std::size_t p = 0; p < avg.size(); ++p)
    {
        double d4[4], l4[4];
        for (int c = 0; c < 3; ++c)
        {
            const double x = a.rgb[3 * p + c], y = b.rgb[3 * p + c];
            d4[c] = y - x;
            l4[c] = (std::fabs(x) + std::fabs(y)) / 2;
            identical = identical && d4[c] == 0;
        }
        d4[3] = (d4[0] + d4[1] + d4[2]) / 3;
        l4[3] = (l4[0] + l4[1] + l4[2]) / 3;
        for (int c = 0; c < 4; ++c)
        {
            whole[c].add(d4[c], l4[c]);
        }
        avg[p] = d4[3];
        lvl[p] = l4[3];
        if (d4[3] > sign_threshold * l4[3])
        {
            ++b_brighter;
        }
        else if (d4[3] < -sign_threshold * l4[3])
        {
            ++a_brighter;
        }
    }
    if (identical)
    {
        std::printf("Error: A and B are identical: the comparison needs two independent renders (different seeds)\n");
        return 2;
    }
    const char* names[4] = {"red", "green", "blue", "average"};
    double whole_largest = 0;
    std::printf("\nWhole image, mean of B - A:\n");
    for (int c = 0; c < 4; ++c)
    {
        const double z = mean_z(whole[c]);
        whole_largest = std::max(whole_largest, std::fabs(z));
        std::printf("  %-8s %+.6e  z %+.9f\n", names[c], whole[c].mean, z);
    }
    const int64_t differing = a_brighter + b_brighter;
    const double sign_z =
        differing ? static_cast<double>(b_brighter - a_brighter) / std::sqrt(static_cast<double>(differing)) : 0.0;
    std::printf("  sign test: B brighter in %lld pixels, A in %lld: z %+.6f\n", static_cast<long long>(b_brighter),
                static_cast<long long>(a_brighter), sign_z);

    // the tiles of the channel average, row by row; partial tiles at the right and bottom edges count
    std::vector<double> zs;
    for (int y0 = 0; y0 < h; y0 += tile)
    {
        for (int x0 = 0; x0 < w; x0 += tile)
        {
            running t;
            for (int y = y0; y < std::min(h, y0 + tile); ++y)
            {
                for (int x = x0; x < std::min(w, x0 + tile); ++x)
                {
                    const std::size_t p = static_cast<std::size_t>(y) * w + x;
                    t.add(avg[p], lvl[p]);
                }
            }
            if (t.n < 2 || t.rms() == 0)
            {
                continue; // too small, or the same in A and B: nothing to test
            }
            zs.push_back(mean_z(t));
        }
    }
    const int64_t tested = static_cast<int64_t>(zs.size());
    double stouffer = 0, largest = 0;
    int64_t over3 = 0;
    for (const double z : zs)
    {
        stouffer += std::max(-40.0, std::min(40.0, z));
        largest = std::max(largest, std::fabs(z));
        over3 += std::fabs(z) > 3;
    }
    if (tested)
    {
        stouffer /= std::sqrt(static_cast<double>(tested));
    }
    std::printf("\nTiles of %dx%d pixels: %lld tested, all tiles z %+.9f, largest |z| %.9f, %lld with |z| > 3\n", tile,
                tile, static_cast<long long>(tested), stouffer, largest, static_cast<long long>(over3));

    struct check
    {
        const char* name;
        double value, limit;
        bool failed;
    };
    std::vector<check> checks;
    const double whole_limit = normal_quantile(false_alarm / 4);
    checks.push_back({"whole image", whole_largest, whole_limit, whole_largest > whole_limit});
    if (tested)
    {
        const double all_limit = normal_quantile(false_alarm);
        checks.push_back({"all tiles", std::fabs(stouffer), all_limit, std::fabs(stouffer) > all_limit});
        const double largest_limit = normal_quantile(false_alarm / static_cast<double>(tested));
        checks.push_back({"largest tile", largest, largest_limit, largest > largest_limit});
        const double q = normal_tail(3.0);
        int64_t k = std::max<int64_t>(1, static_cast<int64_t>(static_cast<double>(tested) * q));
        while (binomial_tail(tested, q, k) > false_alarm)
        {
            ++k;
        }
        checks.push_back({"tile count", static_cast<double>(over3), static_cast<double>(k), over3 >= k});
    }
    std::printf("\nChecks (false-alarm rate 1/400 each):\n");
    bool different = false;
    std::string failed;
    for (const check& c : checks)
    {
        std::printf("  %-12s %.9f (limit %.9f): %s\n", c.name, c.value, c.limit, c.failed ? "FAILED" : "passed");
        if (c.failed)
        {
            failed += (failed.empty() ? "" : ", ") + std::string(c.name);
            different = true;
        }
    }
    if (different)
    {
        std::printf("\nVerdict: DIFFERENT (%s)\n", failed.c_str());
        return 1;
    }
    std::printf("\nVerdict: consistent (A and B differ by noise alone)\n");
    return 0;
}
