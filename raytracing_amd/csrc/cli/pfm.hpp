// PFM images (the renders' linear floats) for the command-line tools: "PF", the width and height, a scale whose sign
// is the byte order (negative: little-endian), then rows of RGB floats from the bottom (src/image.ixx:150-187). In
// memory, rows from the top, as doubles (a float converts exactly, so the file's values come back bit for bit).
#pragma once

#include <cstdint>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

namespace rt::cli
{
struct pfm_image
{
    int width = 0;
    int height = 0;
    std::vector<double> rgb; // rows from the top, 3 values a pixel
};

inline bool load_pfm(const std::string& path, pfm_image& out, std::string& error)
{
    std::ifstream f(path, std::ios::binary);
    if (!f)
    {
        error = "can't open " + path;
        return false;
    }
    std::string magic;
    double scale = 0;
    f >> magic >> out.width >> out.height >> scale;
    if (!f || magic != "PF" || out.width <= 0 || out.height <= 0 || out.width > 1 << 16 || out.height > 1 << 16 ||
        scale == 0)
    {
        error = path + ": not a color PFM file";
        return false;
    }
    f.get(); // the single white-space character after the scale
    const std::size_t n = static_cast<std::size_t>(out.width) * static_cast<std::size_t>(out.height) * 3;
    std::vector<uint32_t> words(n);
    f.read(reinterpret_cast<char*>(words.data()), static_cast<std::streamsize>(n * 4));
    if (static_cast<std::size_t>(f.gcount()) != n * 4)
    {
        error = path + ": the file is shorter than its size says";
        return false;
    }
    const bool little = scale < 0;
    out.rgb.resize(n);
    const std::size_t row = static_cast<std::size_t>(out.width) * 3;
    for (int y = 0; y < out.height; ++y)
    {
        const std::size_t from = static_cast<std::size_t>(out.height - 1 - y) * row;
        for (std::size_t k = 0; k < row; ++k)
        {
            uint32_t w = words[from + k];
            if (!little)
            {
                w = (w >> 24) | ((w >> 8) & 0xff00u) | ((w << 8) & 0xff0000u) | (w << 24);
            }
            float v;
            std::memcpy(&v, &w, 4);
            out.rgb[static_cast<std::size_t>(y) * row + k] = v;
        }
    }
    return true;
}

// Little-endian, scale -1.0
inline bool save_pfm(const std::string& path, int w, int h, const std::vector<double>& rgb)
{
    std::ofstream f(path, std::ios::binary);
    if (!f)
    {
        return false;
    }
    f << "PF\n" << w << " " << h << "\n-1.0\n";
    for (int y = h - 1; y >= 0; --y)
    {
        for (int x = 0; x < w; ++x)
        {
            for (int c = 0; c < 3; ++c)
            {
                const float v = static_cast<float>(rgb[3 * (static_cast<std::size_t>(y) * w + x) + c]);
                f.write(reinterpret_cast<const char*>(&v), 4);
            }
        }
    }
    return static_cast<bool>(f);
}
}
