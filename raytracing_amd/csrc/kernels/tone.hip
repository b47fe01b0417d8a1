// Tone maps beyond the books' bytes (the reference's gpu::tone_map, src/gpu/renderer.h:282-303, and its display
// kernel, src/gpu/tonemap.cu): a pixel's light (its sums times 1/passes, rounded to float as a PFM holds it, times
// 2^exposure) through a tone curve, display-encoded in [0, 1], dithered and quantized to a byte. One thread per pixel,
// float arithmetic, no fused multiply-adds (-ffp-contract=off): the bytes follow from the order the code is written in.
//
// The curves, each from its published source:
//   - none: the light clipped to [0, 1] and encoded with the sRGB curve (IEC 61966-2-1);
//   - agx: Troy Sobotka's AgX (github.com/sobotka/AgX): an inset matrix that pulls the primaries toward grey, the light
//     on a log2 scale of 16.5 stops around middle grey, an S-curve per channel, and the inset's inverse (the "outset"),
//     as Benjamin Wrensch's Minimal AgX arranges them; the S-curve in linlin's closed form (github.com/bWFuanVzYWth/AgX);
//     the output is encoded for a display already;
//   - neutral: Khronos PBR Neutral (github.com/KhronosGroup/ToneMapping, PBR_Neutral), then the sRGB curve.
// The dither is Jorge Jimenez's interleaved gradient noise ("Next Generation Post Processing in Call of Duty: Advanced
// Warfare", SIGGRAPH 2014), one value per pixel for its three channels.
//
// AgX's matrices and arrangement follow iolite's Minimal AgX, Copyright (c) 2024 Missing Deadlines (Benjamin Wrensch),
// and the closed form of its curve github.com/bWFuanVzYWth/AgX, Copyright (c) 2024 linlin, both under the MIT License:
//
// Permission is hereby granted, free of charge, to any person obtaining a copy of this software and associated
// documentation files (the "Software"), to deal in the Software without restriction, including without limitation the
// rights to use, copy, modify, merge, publish, distribute, sublicense, and/or sell copies of the Software, and to
// permit persons to whom the Software is furnished to do so, subject to the following conditions:
//
// The above copyright notice and this permission notice shall be included in all copies or substantial portions of the
// Software.
//
// THE SOFTWARE IS PROVIDED "AS IS", WITHOUT WARRANTY OF ANY KIND, EXPRESS OR IMPLIED, INCLUDING BUT NOT LIMITED TO THE
// WARRANTIES OF MERCHANTABILITY, FITNESS FOR A PARTICULAR PURPOSE AND NONINFRINGEMENT. IN NO EVENT SHALL THE AUTHORS OR
// COPYRIGHT HOLDERS BE LIABLE FOR ANY CLAIM, DAMAGES OR OTHER LIABILITY, WHETHER IN AN ACTION OF CONTRACT, TORT OR
// OTHERWISE, ARISING FROM, OUT OF OR IN CONNECTION WITH THE SOFTWARE OR THE USE OR OTHER DEALINGS IN THE SOFTWARE.
#include <cstdint>

#include <hip/hip_runtime.h>

#include "../../../include/raytracing_amd.h"
#include "kernel_api.hpp"

namespace rt::dev
{
namespace
{
struct rgb
{
    float r, g, b;
};

__device__ inline float clamp01(float v) { return fminf(fmaxf(v, 0.0f), 1.0f); }

// NaN, negatives and -inf to 0; +inf to a large finite value (so that the matrices' sums stay finite)
__device__ inline float sanitize(float v) { return v > 0.0f ? fminf(v, 3.0e38f) : 0.0f; }

// The sRGB encoding of a linear value in [0, 1]: linear below 0.0031308, a power of 1/2.4 above; 1 stays exactly 1
__device__ inline float srgb_encode(float linear)
{
    if (linear >= 1.0f)
    {
        return 1.0f;
    }
    return linear <= 0.0031308f ? 12.92f * linear : 1.055f * powf(linear, 1.0f / 2.4f) - 0.055f;
}

// AgX's log2 scale: -12.47393 to 4.026069 stops (middle grey 0.18 ten stops above the bottom), as a position in [0, 1]
__device__ inline float agx_position(float linear)
{
    const float min_ev = -12.47393f;
    const float max_ev = 4.026069f;
    return clamp01((log2f(linear) - min_ev) / (max_ev - min_ev));
}

// AgX's S-curve in closed form: 0.5 + 2d (1 + a |d|^b)^c with d = x - 10/16.5, (a, b, c) = (59.507875, 3, -1/3) below
// middle grey and (69.86278913545539, 13/4, -4/13) above
__device__ inline float agx_curve(float x)
{
    const float t = 0.6060606060606061f;
    const float d = x - t;
    const float m = fabsf(d);
    const float cube = m * m * m;
    if (x <= t)
    {
        return 0.5f + 2.0f * d * rcbrtf(1.0f + 59.507875f * cube);
    }
    return 0.5f + 2.0f * d * powf(1.0f + 69.86278913545539f * (cube * sqrtf(sqrtf(m))), -4.0f / 13.0f);
}

// The inset and outset matrices (rows act on a column color): AgX's published "AgX Log (Kraken)" inset and its inverse
__device__ inline rgb agx(rgb c)
{
    const float r = 0.842479062253094f * c.r + 0.0784335999999992f * c.g + 0.0792237451477643f * c.b;
    const float g = 0.0423282422610123f * c.r + 0.878468636469772f * c.g + 0.0791661274605434f * c.b;
    const float b = 0.0423756549057051f * c.r + 0.0784336f * c.g + 0.879142973793104f * c.b;
    const float cr = agx_curve(agx_position(r));
    const float cg = agx_curve(agx_position(g));
    const float cb = agx_curve(agx_position(b));
    return {clamp01(1.19687900512017f * cr - 0.0980208811401368f * cg - 0.0990297440797205f * cb),
            clamp01(-0.0528968517574562f * cr + 1.15190312990417f * cg - 0.0989611768448433f * cb),
            clamp01(-0.0529716355144438f * cr - 0.0980434501171241f * cg + 1.15107367264116f * cb)};
}

// Khronos PBR Neutral (F90 = 0.04, Ks = 0.76, Kd = 0.15): the toe takes a dielectric's 0.04 off every color, colors up
// to 0.76 pass unchanged, and above that the brightest channel is compressed with all three by one factor (the hue
// kept) and faded toward white
__device__ inline rgb pbr_neutral(rgb c)
{
    const float f90 = 0.04f;
    const float ks = 0.8f - f90;
    const float kd = 0.15f;
    const float x = fminf(c.r, fminf(c.g, c.b));
    const float f = x <= 2.0f * f90 ? x - 6.25f * x * x : f90;
    const rgb s{c.r - f, c.g - f, c.b - f};
    const float p = fmaxf(s.r, fmaxf(s.g, s.b));
    if (p <= ks)
    {
        return s;
    }
    const float pn = 1.0f - (1.0f - ks) * (1.0f - ks) / (p + 1.0f - 2.0f * ks);
    const float g = 1.0f / (kd * (p - pn) + 1.0f);
    const float scale = pn / p * g;
    const float white = pn * (1.0f - g);
    return {s.r * scale + white, s.g * scale + white, s.b * scale + white};
}

__device__ inline rgb encode(rgb light, int op)
{
    const rgb c{sanitize(light.r), sanitize(light.g), sanitize(light.b)};
    if (op == RT_TONE_AGX)
    {
        return agx(c);
    }
    if (op == RT_TONE_NEUTRAL)
    {
        const rgb n = pbr_neutral(c);
        return {srgb_encode(clamp01(n.r)), srgb_encode(clamp01(n.g)), srgb_encode(clamp01(n.b))};
    }
    return {srgb_encode(clamp01(c.r)), srgb_encode(clamp01(c.g)), srgb_encode(clamp01(c.b))};
}

__device__ inline float fract(float v) { return v - floorf(v); }

// Interleaved gradient noise at pixel (x, y), in [0, 1)
__device__ inline float ign(int x, int y)
{
    return fract(52.9829189f * fract(0.06711056f * static_cast<float>(x) + 0.00583715f * static_cast<float>(y)));
}

// round(255 v + dither), dither in [-0.5, 0.5); 0 and 1 (and NaN, as 0) never dithered
__device__ inline uint8_t quantize(float v, float dither)
{
    if (!(v > 0.0f))
    {
        return 0;
    }
    if (v >= 1.0f)
    {
        return 255;
    }
    return static_cast<uint8_t>(fminf(fmaxf(floorf(255.0f * v + 0.5f + dither), 0.0f), 255.0f));
}
}

__global__ void __launch_bounds__(256) tone_map_kernel(const double* __restrict__ sums, int width, int64_t pixels,
                                                       double inverse_passes, int op, float exposure_scale,
                                                       uint8_t* __restrict__ rgba)
{
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= pixels)
    {
        return;
    }
    const double* sum = sums + 3 * k;
    const rgb light{static_cast<float>(sum[0] * inverse_passes) * exposure_scale,
                    static_cast<float>(sum[1] * inverse_passes) * exposure_scale,
                    static_cast<float>(sum[2] * inverse_passes) * exposure_scale};
    const rgb v = encode(light, op);
    const float dither = ign(static_cast<int>(k % width), static_cast<int>(k / width)) - 0.5f;
    uint8_t* out = rgba + 4 * k;
    out[0] = quantize(v.r, dither);
    out[1] = quantize(v.g, dither);
    out[2] = quantize(v.b, dither);
    out[3] = 255;
}

// The parity probe: colour k (linear, times the exposure's scale) through the tone map `op` and quantized with
// dither[k] (0 where dither is null), as the display kernel does for a pixel
__global__ void __launch_bounds__(256) tone_probe_kernel(const float* __restrict__ colours,
                                                         const float* __restrict__ dither, int64_t n, int op,
                                                         float exposure_scale, uint8_t* __restrict__ out)
{
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n)
    {
        return;
    }
    const rgb light{colours[3 * k] * exposure_scale, colours[3 * k + 1] * exposure_scale,
                    colours[3 * k + 2] * exposure_scale};
    const rgb v = encode(light, op);
    const float d = dither ? dither[k] : 0.0f;
    out[3 * k] = quantize(v.r, d);
    out[3 * k + 1] = quantize(v.g, d);
    out[3 * k + 2] = quantize(v.b, d);
}
}

namespace rt
{
hipError_t launch_tone_probe(const float* colours, const float* dither, int64_t n, int op, float exposure_scale,
                             uint8_t* out, hipStream_t stream)
{
    if (n <= 0)
    {
        return hipSuccess;
    }
    const int threads = 256;
    const dim3 grid(static_cast<unsigned>((n + threads - 1) / threads));
    hipLaunchKernelGGL(dev::tone_probe_kernel, grid, dim3(threads), 0, stream, colours, dither, n, op, exposure_scale,
                       out);
    return hipGetLastError();
}

hipError_t launch_tone_map(const double* sums, int width, int64_t pixels, double inverse_passes, int op,
                           float exposure_scale, uint8_t* rgba, hipStream_t stream)
{
    const int threads = 256;
    const dim3 grid(static_cast<unsigned>((pixels + threads - 1) / threads));
    hipLaunchKernelGGL(dev::tone_map_kernel, grid, dim3(threads), 0, stream, sums, width, pixels, inverse_passes, op,
                       exposure_scale, rgba);
    return hipGetLastError();
}
}
