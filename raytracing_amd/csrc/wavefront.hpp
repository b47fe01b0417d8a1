// The wavefront integrator's host driver (see wavefront.cpp)
#pragma once

#include <cstdint>
#include <stdexcept>

#include <hip/hip_runtime.h>

#include "device_types.hpp"
#include "kernels/kernel_api.hpp"

namespace rt
{
template <typename Real>
constexpr bool wavefront_supported()
{
    return false;
}

template <typename Real>
class wavefront
{
public:
    wavefront(const dev::scene_view<Real>&, const dev::camera_rec<Real>&, bool, int, hipStream_t)
    {
        throw std::invalid_argument("the wavefront integrator is not built in this library");
    }
    void render(uint64_t, const dev::launch_work&, double*, unsigned long long*) {}
    kernel_attributes attributes() const { return {}; }
    std::size_t bytes() const { return 0; }
};
}
