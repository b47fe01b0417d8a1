// The scene as it lies in the GPU's memory: plain records shared by the host code that builds them (bvh_build.cpp,
// device_scene.cpp) and the kernels that read them. Every record a hot loop reads is a whole number of 16-byte words,
// read with dwordx4 loads.
//
// Layout choices (the reference's are in src/gpu/bvh.cuh:29-38, sphere.cuh:11-46, quad.cuh:19-30):
//   - BVH: four-wide. The binary tree (the reference's object-median one, or the surface-area one) is collapsed so
//     that a node holds the boxes of up to FOUR children (128 bytes in float: one L2 line per visited node), where the
//     reference's 32-byte node holds its own box and a visit reads two children's records from two places. A child is
//     an inner node's index, or a packed leaf reference (leaf_ref below). The boxes are the binary tree's, the slab
//     test is the reference's, and the closest hit is the same primitive at the same distance (of two primitives hit
//     at exactly the same distance, the one of lower index in the scene, whatever the walk's order).
//   - primitives keep the reference's global ids: spheres that stand still [0, moving_base), moving spheres
//     [moving_base, quad_base), quads [quad_base, triangle_base), triangles in world coordinates [triangle_base,
//     instanced_base), the triangles of meshes placed by several instances [instanced_base, media_base) (each
//     instance its own range), and the media after the surfaces.
#pragma once

#include <cstddef>
#include <cstdint>

#include <hip/hip_runtime.h>

namespace rt::dev
{
// The kinds of primitives in a leaf. An instance leaf holds one instance of a mesh: the walk goes on into the mesh's
// own tree (its triangles in the mesh's coordinates) with the ray moved there.
enum prim_kind : int32_t
{
    prim_sphere = 0,
    prim_moving_sphere = 1,
    prim_quad = 2,
    prim_triangle = 3,
    prim_instance = 4,
};

// A leaf reference: bit 31 set, the kind in bits 28-30, the count minus 1 in bits 24-27 (at most 16 primitives), and
// the first primitive's index in its kind's array in bits 0-23
constexpr int32_t leaf_first_bits = 24;
constexpr int32_t max_leaf_first = (1 << leaf_first_bits) - 1;
__host__ __device__ constexpr int32_t make_leaf_ref(int32_t kind, int32_t count, int32_t first)
{
    return static_cast<int32_t>(0x80000000u | (static_cast<uint32_t>(kind) << 28) |
                                (static_cast<uint32_t>(count - 1) << 24) | static_cast<uint32_t>(first));
}
__host__ __device__ constexpr int32_t leaf_kind(int32_t ref) { return (ref >> 28) & 7; }
__host__ __device__ constexpr int32_t leaf_count(int32_t ref) { return ((ref >> 24) & 15) + 1; }
__host__ __device__ constexpr int32_t leaf_first(int32_t ref) { return ref & max_leaf_first; }

// An inner node of the four-wide tree: the boxes of up to four children, one array per bound and axis (lo[axis][child],
// so the four slab tests are the same instructions on different registers), and the children's references in the
// binary tree's depth-first order. A missing child has a box of +infinity corners, which box_entry never enters (each
// slab gives +inf or -inf), so the walk needs no test for it. 128 bytes in float, 256 in double.
constexpr int wide_children = 4;
constexpr int32_t no_child = 0x7fffffff;
template <typename Real>
struct alignas(16) wide_node
{
    Real lo[3][wide_children];
    Real hi[3][wide_children];
    int32_t ref[wide_children];
    int32_t unused[sizeof(Real) == 4 ? 4 : 12];
};
static_assert(sizeof(wide_node<float>) == 128 && sizeof(wide_node<double>) == 256);
// the kernels read the planes and references at these offsets (child_entries, child_refs)
static_assert(offsetof(wide_node<float>, hi) == 12 * sizeof(float) && offsetof(wide_node<float>, ref) == 24 * sizeof(float));
static_assert(offsetof(wide_node<double>, hi) == 12 * sizeof(double) &&
              offsetof(wide_node<double>, ref) == 24 * sizeof(double));

// The same inner node in 64 bytes, for trees larger than the caches (a packed tree): each child's box on a grid of
// 8-bit steps over the node's own box, one power-of-two step per axis, the grid's corner a multiple of the step, so
// that every plane origin + q x step is a float. The host rounds every lower plane down and every upper one up: the
// packed box holds the float one, and since the slab test is the same multiply-add on either box's planes and
// monotonic in them, the walk enters every child it entered before, at no farther a distance (the closest hit is the
// same primitive at the same distance: the primitives' tests don't change). A missing child has no_child as its
// reference and lo 255, hi 0. Half the bytes of wide_node<float>, read as four dwordx4 loads from one 64-byte
// half-line; a plane is unpacked by a v_cvt_f32_ubyte and an exact multiply-add.
struct alignas(64) packed_node
{
    float origin[3];            // bytes 0-11
    float step[3];              // bytes 12-23: powers of two
    uint8_t lo[3][wide_children]; // bytes 24-35: lo[axis][child]
    uint8_t hi[3][wide_children]; // bytes 36-47
    int32_t ref[wide_children]; // bytes 48-63
};
static_assert(sizeof(packed_node) == 64 && offsetof(packed_node, lo) == 24 && offsetof(packed_node, hi) == 36 &&
              offsetof(packed_node, ref) == 48);

template <typename Real>
struct alignas(16) sphere_rec
{
    Real center[3];
    Real radius;
};

template <typename Real>
struct alignas(16) moving_sphere_rec
{
    Real center[3];
    Real radius;
    Real displacement[3];
    Real unused;
};

// A quad as its hit test reads it: the corner, the unit normal and the dual axes of (u, v) (src/gpu/quad.cuh:19-43)
template <typename Real>
struct alignas(16) quad_rec
{
    Real q[3];
    Real normal[3];
    Real alpha_axis[3];
    Real beta_axis[3];
};
static_assert(sizeof(quad_rec<float>) == 48);

// A triangle for the Moller-Trumbore test: a vertex, the two edges from it, and the unit normal e1 x e2 (its front)
template <typename Real>
struct alignas(16) triangle_rec
{
    Real v0[3];
    Real e1[3];
    Real e2[3];
    Real normal[3];
};
static_assert(sizeof(triangle_rec<float>) == 48);

// An instance of a mesh (two-level BVH, the reference's mesh_set instances, src/gpu/renderer.h:305-376): the world to
// object transform (row-major 3x4: x' = A x + b), the root of the mesh's own tree (in the same node array), where the
// mesh's triangles start in the triangle array, the global id of its first triangle (an instance's triangle k has the
// id id_base + k: every instance has ids of its own, for the self-intersection rule and the ties), and its material.
// A ray keeps its parameter t in the mesh's coordinates (its direction is moved by A, not normalized), so distances
// found there compare with the world's directly.
template <typename Real>
struct alignas(16) instance_rec
{
    Real to_object[12];
    int32_t root_ref;
    int32_t first_triangle;
    int32_t id_base;
    int32_t material;
};
static_assert(sizeof(instance_rec<float>) == 64);

// What only a sphere hit reads besides its material (src/gpu/sphere.cuh:36-46)
template <typename Real>
struct alignas(16) sphere_frame
{
    Real cos_theta;
    Real sin_theta;
    Real k;
    Real limit;
};

enum material_flags : uint32_t
{
    needs_uv = 1,
    needs_exact_point = 2,
};

// A material and how its texture is read (src/gpu/material.cuh:18-40)
template <typename Real>
struct alignas(16) material_rec
{
    int32_t kind;
    int32_t texture; // -1: one solid color, in albedo
    uint32_t flags;
    int32_t unused;
    Real albedo[3];
    Real fuzz;
    Real refraction_index;
    Real unused2[3];
};

template <typename Real>
struct alignas(16) texture_rec
{
    int32_t kind;
    int32_t even;
    int32_t odd;
    int32_t table; // image: index into the image infos; noise: index into the noise tables
    Real rgb[3];
    Real inv_scale;
};

struct alignas(16) image_info
{
    int32_t offset;
    int32_t width;
    int32_t height;
    int32_t unused;
};

struct alignas(16) texel
{
    float r, g, b, unused;
};

template <typename Real>
struct alignas(16) noise_rec
{
    Real gradients[256][4];
    uint8_t perm_x[256];
    uint8_t perm_y[256];
    uint8_t perm_z[256];
    Real scale;
    Real unused[3];
};

// A volume of smoke (src/gpu/constant_medium.cuh:14-26)
template <typename Real>
struct alignas(16) medium_rec
{
    Real box_min[3];
    Real neg_inv_density;
    Real box_max[3];
    Real unused;
    Real center[3];
    Real radius;
    Real displacement[3];
    Real unused2;
    int32_t material;
    int32_t first_quad;
    int32_t quad_count;
    int32_t unused3;
};

// A shape of the lights list (src/gpu/lights.cuh:18-39)
template <typename Real>
struct alignas(16) light_rec
{
    quad_rec<Real> quad;
    Real u[3];
    Real area;
    Real v[3];
    Real cdf;
    Real center[3];
    Real radius;
    Real weight;
    int32_t kind;
    int32_t twin;
    int32_t unused;
};

// The camera as the kernels read it (src/gpu/camera.cuh:33-51)
template <typename Real>
struct camera_rec
{
    int32_t image_width;
    int32_t image_height;
    int32_t sqrt_spp;
    int32_t max_depth;
    Real recip_sqrt_spp;
    Real center[3];
    Real center_to_pixel00[3];
    Real pixel_delta_u[3];
    Real pixel_delta_v[3];
    int32_t defocus;
    Real defocus_disk_u[3];
    Real defocus_disk_v[3];
    Real sky_bottom[3];
    Real sky_top[3];
};

// A lamp of the quality integrator's light set: a glowing quad or sphere of the world (src/gpu/lamps.cuh:41-54)
struct alignas(16) emitter_rec
{
    float corner[3]; // quad: q; sphere: the center
    float size;      // quad: the area; sphere: the radius
    float u[3];
    int32_t kind;    // RT_LIGHT_QUAD or RT_LIGHT_SPHERE
    float v[3];
    int32_t primitive; // its global id: its own shadow ray doesn't count it as a blocker
    float normal[3];
    float unused0;
    float radiance[3];
    float unused1;
};
static_assert(sizeof(emitter_rec) == 80);

// A texel of the environment map with its cell's density per steradian, and a bin of the map's alias table
// (src/gpu/environment_data.h:24-57)
struct alignas(16) env_texel
{
    float r, g, b, pdf;
};
struct alignas(8) alias_bin
{
    float q;
    int32_t alias;
};

// A material's GGX side entry as the kernels read it
struct alignas(16) microfacet_rec
{
    int32_t kind;    // RT_MICROFACET_*
    float alpha;     // r^2, rounded once
    float roughness; // r
    int32_t unused;
};

// What only the quality integrator reads: the light combination, the lamps, the environment, the GGX side entries
struct quality_view
{
    int32_t combination;  // RT_COMBINE_*
    int32_t lamp_count;
    const emitter_rec* lamps;
    const int32_t* lamp_of; // per surface primitive: its lamp, or -1
    float light_pmf;        // 1 / (lamps + the map when it is sampled)
    int32_t env_present;
    const microfacet_rec* microfacets; // one per material, or nullptr
    const env_texel* env_texels;
    const alias_bin* env_bins;          // nullptr: the map isn't sampled (a black map)
    const float* env_row_cos;
    int32_t env_width;
    int32_t env_height;
    int32_t env_log2_bins;
    float env_cos_yaw;
    float env_sin_yaw;
    float env_scale;
};

// Where the scene's arrays are: the kernels take it by value
template <typename Real>
struct scene_view
{
    quality_view quality;
    const wide_node<Real>* nodes;
    const packed_node* packed_nodes; // the same tree packed (null unless the scene's kernels walk it: walk_packed)
    int32_t root_ref;  // the root: an inner node's index (0), a leaf reference, or -1 with no primitive at all
    int32_t stack_depth; // the most entries the traversal stack can hold (at least 1), or 0 with no primitive at all
    const sphere_rec<Real>* spheres;
    const moving_sphere_rec<Real>* moving_spheres;
    const quad_rec<Real>* quads;
    const triangle_rec<Real>* triangles;
    const instance_rec<Real>* instances; // by id_base, increasing
    int32_t instance_count;
    int32_t moving_base;
    int32_t quad_base;
    int32_t triangle_base;
    int32_t instanced_base; // the first instanced triangle's id: prim_materials and prim_ranks end here
    int32_t media_base;
    const int32_t* prim_materials;
    const int32_t* prim_ranks; // by global id: the primitive's index in the scene, which breaks exact ties
    const sphere_frame<Real>* sphere_frames;
    const material_rec<Real>* materials;
    const texture_rec<Real>* textures;
    const image_info* images;
    const texel* texels;
    const noise_rec<Real>* noises;
    const medium_rec<Real>* media;
    const quad_rec<Real>* boundary_quads;
    int32_t medium_count;
    int32_t light_count;
    const light_rec<Real>* lights;
    int32_t sampling; // rt_sampling_strategy
    int32_t motion;   // some sphere or medium boundary moves: camera rays draw a time
};

// What a walk may meet: the first book's trees hold spheres that stand still only; the books' add moving spheres and
// quads; the quality integrator's add triangles, and in a two-level tree instances of meshes (a walk that can enter
// an instance carries the world ray beside the moved one: a kernel of its own, so that the others don't hold it); a
// packed tree (packed_node) is walked by kernels of its own too
enum walk_kinds : int
{
    walk_spheres = 0,
    walk_books = 1,
    walk_all = 2,
    walk_instances = 3,
    walk_packed = 4, // walk_all's kinds in a packed tree (float only)
};
// The walk a scene's quality kernels (and the parity probe) take: into instances, over packed nodes, or plain
template <typename Real>
__host__ __device__ inline int tree_walk(const scene_view<Real>& sc)
{
    return sc.instance_count > 0 ? walk_instances : sc.packed_nodes != nullptr ? walk_packed : walk_all;
}

// Which integrator a scene's kernels run (src/gpu/renderer.cu:1011-1023, src/gpu/quality.cu:53-109): the first book's
// (spheres that stand still, no texture, the material strategy), the three books', or the quality integrator's (light
// samples weighed by MIS, GGX, the environment map, triangles; float only)
enum kernel_mode : int32_t
{
    mode_first_book = 0,
    mode_all_books = 1,
    mode_quality = 2,
};

// One path's sample, staged by the megakernel until its pixel's samples are added in pass order (16 bytes in float)
template <typename Real>
struct alignas(4 * sizeof(Real)) staged_sample
{
    Real x, y, z, unused;
};

// The work of one launch: passes [first_pass, first_pass + pass_count) of rows [first_row, first_row + row_count),
// within one round of strata
struct launch_work
{
    int32_t first_pass;
    int32_t pass_count;
    int32_t first_row;
    int32_t row_count;
};
}
