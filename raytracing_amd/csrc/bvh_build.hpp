// The BVH builder (host code): a binary tree over every sphere, quad and triangle of a scene (the reference's
// object-median splits, src/gpu/bvh.cuh:100-444, or binned surface-area-heuristic ones), then collapsed into the
// four-wide layout the kernels walk (device_types.hpp).
#pragma once

#include <cstdint>
#include <vector>

#include "device_types.hpp"
#include "scene_tables.hpp"

namespace rt
{
// A node in the reference's layout: its own box, and either its second child's index (an inner node; the first child
// follows it) or its primitives (a leaf: count = n | kind << 8)
template <typename Real>
struct ref_node
{
    Real box_min[3];
    int32_t index;
    Real box_max[3];
    int32_t count;
};

template <typename Real>
struct bvh_arrays
{
    std::vector<ref_node<Real>> nodes; // the binary tree, depth first, the reference's order
    std::vector<dev::wide_node<Real>> wide; // the binary tree collapsed four-wide: what the kernels walk
    int32_t root_ref = -1;
    int32_t depth = 0;         // the binary tree's levels
    int32_t stack_entries = 0; // the most entries the four-wide walk's stack can hold
    std::vector<dev::sphere_rec<Real>> spheres;
    std::vector<dev::moving_sphere_rec<Real>> moving_spheres;
    std::vector<dev::quad_rec<Real>> quads;
    std::vector<dev::triangle_rec<Real>> triangles;
    std::vector<int32_t> prim_materials;
    std::vector<int32_t> prim_ranks; // by global id: the primitive's index in the scene (spheres, quads, triangles)
    std::vector<dev::sphere_frame<Real>> sphere_frames;
    int32_t moving_base = 0;
    int32_t quad_base = 0;
    int32_t triangle_base = 0;
    std::vector<int32_t> sphere_ids; // the global id of each sphere of the scene, by its index there
    std::vector<int32_t> quad_ids;
};

// The leaf size the reference uses: one primitive per leaf, or one leaf of up to 8 for a scene of at most 8 shapes
// (src/gpu/renderer.cu:75-82)
int leaf_size_for(std::size_t shapes);

// The tree over every sphere, quad and triangle of a scene (tree: an rt_tree). RT_TREE_MEDIAN on a scene without
// triangles gives the reference's tree (object-median splits, leaves of max_leaf_size); otherwise, and for every scene
// with triangles (meshes baked into world coordinates through their instances), a surface-area-heuristic tree of
// small leaves: the reference's object-median split on millions of triangles would make a tree several times slower
// to walk.
template <typename Real>
bvh_arrays<Real> build_bvh(const rt_scene& scene, int max_leaf_size, int32_t tree);

// The world triangles of a scene's meshes, through their instances (or each mesh once, as it is, when the scene has
// no instances), with their materials: what build_bvh puts in the tree, in that order before it sorts them
struct world_triangle
{
    double p[3][3];
    int32_t material;
};
std::vector<world_triangle> world_triangles(const rt_scene& scene);

// The quad's record from its description, in double, rounded once (src/gpu/quad.cuh:33-43)
template <typename Real>
dev::quad_rec<Real> make_quad_rec(const rt_quad_desc& d);

// A quad's bounding box in double (src/gpu/bvh.cuh:180-214), and a box rounded outward to Real
struct box_d
{
    double min[3] = {INFINITY, INFINITY, INFINITY};
    double max[3] = {-INFINITY, -INFINITY, -INFINITY};
    void enclose(const box_d& o);
};
box_d quad_box(const rt_quad_desc& d);
template <typename Real>
Real round_down(double x);
template <typename Real>
Real round_up(double x);
}
