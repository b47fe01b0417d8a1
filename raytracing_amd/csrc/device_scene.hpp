// The scene in the GPU's memory (host code): the device arrays the kernels read, owned here, and the scene_view that
// tells a kernel where they are (src/gpu/renderer.cu:1051-1152)
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include <hip/hip_runtime.h>

#include "bvh_build.hpp"
#include "device_types.hpp"
#include "scene_tables.hpp"

namespace rt
{
// A HIP error as an exception with HIP's message, the error cleared (src/gpu/device_memory.h:19-24)
inline void check(hipError_t status, const char* what)
{
    if (status != hipSuccess)
    {
        (void)hipGetLastError();
        throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(status));
    }
}

// Device memory owned like a std::vector
template <typename T>
class device_array
{
public:
    device_array() = default;
    explicit device_array(std::size_t n) : size_(n)
    {
        if (n > 0)
        {
            check(hipMalloc(reinterpret_cast<void**>(&data_), n * sizeof(T)), "allocating GPU memory");
        }
    }
    explicit device_array(const std::vector<T>& host) : device_array(host.size())
    {
        if (size_ > 0)
        {
            check(hipMemcpy(data_, host.data(), size_ * sizeof(T), hipMemcpyHostToDevice), "copying to the GPU");
        }
    }
    ~device_array() { release(); }
    device_array(const device_array&) = delete;
    device_array& operator=(const device_array&) = delete;
    device_array(device_array&& o) noexcept : data_(std::exchange(o.data_, nullptr)), size_(std::exchange(o.size_, 0)) {}
    device_array& operator=(device_array&& o) noexcept
    {
        if (this != &o)
        {
            release();
            data_ = std::exchange(o.data_, nullptr);
            size_ = std::exchange(o.size_, 0);
        }
        return *this;
    }
    T* data() const { return data_; }
    std::size_t size() const { return size_; }
    std::size_t bytes() const { return size_ * sizeof(T); }

private:
    void release()
    {
        if (data_)
        {
            (void)hipFree(data_);
            data_ = nullptr;
        }
    }
    T* data_ = nullptr;
    std::size_t size_ = 0;
};

// Rejects what the kernels can't render, before anything is copied to the GPU (src/gpu/renderer.cu:663-900)
void validate_camera(const rt_camera_desc& camera);
void validate_scene(const rt_scene& scene);

// Which kernel a scene needs: the first book's (spheres that stand still, lambertian of one color, metal, glass, the
// material strategy) or the one for all three books (src/gpu/renderer.cu:1011-1023)
bool first_book_scene(const rt_scene& scene);

template <typename Real>
class device_scene
{
public:
    // tree: an rt_tree; nodes: an rt_nodes (where the quality integrator's kernels walk packed nodes)
    // quality_kernel: build the quality integrator's tables even where the scene doesn't need them (RT_KERNEL_QUALITY)
    device_scene(const rt_scene& scene, int32_t tree, int32_t nodes = RT_NODES_AUTO, bool quality_kernel = false);
    const dev::scene_view<Real>& view() const { return view_; }
    const bvh_arrays<Real>& tree() const { return tree_; }
    std::size_t bytes() const { return bytes_; }
    // whether the scene needs the quality integrator (light samples, GGX, an environment map or triangles)
    bool quality() const { return quality_; }

private:
    bvh_arrays<Real> tree_;
    device_array<dev::wide_node<Real>> nodes_;
    device_array<dev::packed_node> packed_; // the same tree packed, for the quality integrator's walks (or empty)
    device_array<dev::sphere_rec<Real>> spheres_;
    device_array<dev::moving_sphere_rec<Real>> moving_;
    device_array<dev::quad_rec<Real>> quads_;
    device_array<dev::instance_rec<Real>> instances_;
    device_array<int32_t> prim_materials_;
    device_array<int32_t> prim_ranks_;
    device_array<dev::sphere_frame<Real>> frames_;
    device_array<dev::material_rec<Real>> materials_;
    device_array<dev::texture_rec<Real>> textures_;
    device_array<dev::image_info> images_;
    device_array<dev::texel> texels_;
    device_array<dev::noise_rec<Real>> noises_;
    device_array<dev::medium_rec<Real>> media_;
    device_array<dev::quad_rec<Real>> boundary_quads_;
    device_array<dev::light_rec<Real>> lights_;
    device_array<dev::triangle_rec<Real>> triangles_;
    // the quality integrator's tables (float scenes only)
    device_array<dev::emitter_rec> lamps_;
    device_array<int32_t> lamp_of_;
    device_array<dev::microfacet_rec> microfacets_;
    device_array<dev::env_texel> env_texels_;
    device_array<dev::alias_bin> env_bins_;
    device_array<float> env_row_cos_;
    bool quality_ = false;
    dev::scene_view<Real> view_{};
    std::size_t bytes_ = 0;
};

template <typename Real>
dev::camera_rec<Real> to_camera(const rt_camera_desc& desc);
}
