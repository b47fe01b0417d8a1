// Generates ggx_energy_table.hpp: E_ss(mu, r), the light a white GGX conductor keeps after one bounce off its facets
// (F = 1), for Turquin's multiple-scattering compensation. Nodes at sqrt(mu) = i/31 and r = j/31 (alpha = r^2); row 0
// (the mirror) and column 0 (a grazing view) are 1 by convention.
//
// The integral is taken in the visible-normal sampler's own sample space (u1 the azimuth, u2 the height in the
// spherical cap, Dupuy and Benyoub 2023), where E = integral over [0,1)^2 of the sample's weight G2/G1(v) wherever the
// reflected direction lies above the horizon: a bounded integrand, smooth but for the horizon's edge. For each u1
// (Gauss-Legendre, symmetric about the plane of the view), the edge's u2 values are found by bisection and each piece
// above the horizon is integrated by Gauss-Legendre. In double throughout.
//
// Build and run: g++ -O2 -std=c++17 ggx_energy_gen.cpp -o gen && ./gen > ggx_energy_table.hpp
#include <cmath>
#include <cstdio>
#include <vector>

namespace
{
struct v3
{
    double x, y, z;
};
v3 operator+(v3 a, v3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
v3 operator-(v3 a, v3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
v3 operator*(double s, v3 a) { return {s * a.x, s * a.y, s * a.z}; }
double dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
v3 unit(v3 a) { return (1 / std::sqrt(dot(a, a))) * a; }

// Gauss-Legendre nodes and weights on [0, 1] (Newton on the Legendre polynomial)
void gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w)
{
    x.resize(n);
    w.resize(n);
    for (int i = 0; i < n; ++i)
    {
        double z = std::cos(M_PI * (i + 0.75) / (n + 0.5));
        double dp = 0;
        for (int it = 0; it < 100; ++it)
        {
            double p0 = 1, p1 = 0;
            for (int k = 1; k <= n; ++k)
            {
                const double p2 = p1;
                p1 = p0;
                p0 = ((2 * k - 1) * z * p1 - (k - 1) * p2) / k;
            }
            dp = n * (z * p0 - p1) / (z * z - 1);
            const double dz = p0 / dp;
            z -= dz;
            if (std::fabs(dz) < 1e-16)
            {
                break;
            }
        }
        x[i] = 0.5 * (1 - z);
        w[i] = 1.0 / ((1 - z * z) * dp * dp);
    }
}

// The sample at (u1, u2) for the view v = (sin, 0, cos) about n = z: the reflected direction's cosine with n, and the
// weight G2/G1(v)
struct sample
{
    double cl;
    double weight;
};
sample draw(double cv, double alpha, double u1, double u2)
{
    const v3 n{0, 0, 1};
    const v3 v{std::sqrt(std::fmax(0.0, 1 - cv * cv)), 0, cv};
    const v3 vs = unit(alpha * (v - cv * n) + cv * n);
    const double zs = vs.z;
    const double z = (1 - u2) * (1 + zs) - zs;
    const double st = std::sqrt(std::fmax(0.0, 1 - z * z));
    const v3 c{st * std::cos(2 * M_PI * u1), st * std::sin(2 * M_PI * u1), z};
    const v3 hs = c + vs;
    const v3 m = alpha * (hs - hs.z * n) + std::fmax(0.0, hs.z) * n;
    const v3 h = unit(m);
    const double vh = dot(v, h);
    const v3 l = (2 * vh) * h - v;
    const double cl = l.z;
    if (!(cl > 0) || !(vh > 0))
    {
        return {cl, 0};
    }
    const double a2 = alpha * alpha;
    const double av = std::sqrt(cv * cv * (1 - a2) + a2);
    const double al = std::sqrt(cl * cl * (1 - a2) + a2);
    return {cl, cl * (cv + av) / (cl * av + cv * al)};
}

double energy(double cv, double alpha)
{
    static std::vector<double> xo, wo, xi, wi;
    if (xo.empty())
    {
        gauss_legendre(192, xo, wo);
        gauss_legendre(48, xi, wi);
    }
    double total = 0;
    const int scan = 256;
    for (std::size_t a = 0; a < xo.size(); ++a)
    {
        const double u1 = 0.5 * xo[a]; // [0, 1/2): the other half mirrors it
        // the pieces of u2 in [0, 1] where the direction lies above the horizon
        std::vector<double> edges{0.0};
        double prev = draw(cv, alpha, u1, 0.0).cl;
        for (int k = 1; k <= scan; ++k)
        {
            const double u2 = static_cast<double>(k) / scan;
            const double now = draw(cv, alpha, u1, u2).cl;
            if ((prev > 0) != (now > 0))
            {
                double lo = static_cast<double>(k - 1) / scan, hi = u2;
                for (int it = 0; it < 60; ++it)
                {
                    const double mid = 0.5 * (lo + hi);
                    ((draw(cv, alpha, u1, mid).cl > 0) == (prev > 0) ? lo : hi) = mid;
                }
                edges.push_back(0.5 * (lo + hi));
            }
            prev = now;
        }
        edges.push_back(1.0);
        double row = 0;
        for (std::size_t e = 0; e + 1 < edges.size(); ++e)
        {
            const double lo = edges[e], hi = edges[e + 1];
            double piece = 0;
            for (std::size_t b = 0; b < xi.size(); ++b)
            {
                piece += wi[b] * draw(cv, alpha, u1, lo + (hi - lo) * xi[b]).weight;
            }
            row += (hi - lo) * piece;
        }
        total += wo[a] * row;
    }
    return total; // the two halves of u1, each of width 1/2
}
}

int main()
{
    const int size = 32;
    std::printf("// Generated by ggx_energy_gen.cpp: E_ss(mu, r) of the GGX conductor with F = 1 (single scattering)\n");
    std::printf("#pragma once\n\n#include <hip/hip_runtime.h>\n\nnamespace rt::dev\n{\n");
    std::printf("constexpr int ggx_energy_size = %d;\n", size);
    std::printf("// [j * %d + i]: roughness r = j/%d (alpha = r^2), sqrt(cos theta_v) = i/%d\n", size, size - 1, size - 1);
    std::printf("__device__ const float ggx_energy_table[ggx_energy_size * ggx_energy_size] = {\n");
    for (int j = 0; j < size; ++j)
    {
        std::printf("    /* r = %d/%d */\n   ", j, size - 1);
        for (int i = 0; i < size; ++i)
        {
            double e = 1.0;
            if (i > 0 && j > 0)
            {
                const double s = static_cast<double>(i) / (size - 1);
                const double r = static_cast<double>(j) / (size - 1);
                e = energy(s * s, r * r);
            }
            std::printf(" %.9ef,", static_cast<double>(static_cast<float>(e)));
            if (i % 6 == 5 && i + 1 < size)
            {
                std::printf("\n   ");
            }
        }
        std::printf("\n");
    }
    std::printf("};\n}\n");
    return 0;
}
