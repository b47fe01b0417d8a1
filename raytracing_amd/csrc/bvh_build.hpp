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
    std::vector<dev::triangle_rec<Real>> triangles;   // world triangles first, then each instanced mesh's own
    std::vector<dev::instance_rec<Real>> instances;   // the instances of meshes placed more than once, by id_base
    std::vector<int32_t> prim_materials; // by global id below instanced_base
    std::vector<int32_t> prim_ranks; // by global id below instanced_base: the primitive's index in the scene (spheres,
                                     // quads, triangles); an instanced triangle's rank is its id
    std::vector<dev::sphere_frame<Real>> sphere_frames;
    int32_t moving_base = 0;
    int32_t quad_base = 0;
    int32_t triangle_base = 0;
    int32_t instanced_base = 0; // the first id of an instanced mesh's triangle
    int32_t media_base = 0;     // past the last surface id
    std::vector<int32_t> sphere_ids; // the global id of each sphere of the scene, by its index there
    std::vector<int32_t> quad_ids;
};

// The four-wide tree's inner nodes packed (dev::packed_node), node for node: each child's planes on its node's grid, the
// lower planes rounded down and the upper ones up with a whole step to spare, so that every packed box holds the float
// box it replaces. Throws if a node's coordinates leave no such grid (never for finite boxes).
std::vector<dev::packed_node> pack_nodes(const std::vector<dev::wide_node<float>>& wide);

// The leaf size the reference uses: one primitive per leaf, or one leaf of up to 8 for a scene of at most 8 shapes
// (src/gpu/renderer.cu:75-82)
int leaf_size_for(std::size_t shapes);

// The tree over every sphere, quad and triangle of a scene (tree: an rt_tree). RT_TREE_MEDIAN on a scene without
// triangles gives the reference's tree (object-median splits, leaves of max_leaf_size); otherwise, and for every scene
// with triangles, a surface-area-heuristic tree of small leaves: the reference's object-median split on millions of
// triangles would make a tree several times slower to walk. RT_TREE_TWO_LEVEL: two levels where meshes are placed more
// than once: each such mesh gets a tree of its own over its triangles in its own coordinates, once, and each of its
// instances is a leaf of the scene's tree (a mesh placed once, or a scene without instances, is baked into world
// coordinates); otherwise every instance is baked.
template <typename Real>
bvh_arrays<Real> build_bvh(const rt_scene& scene, int max_leaf_size, int32_t tree);

// Which meshes are placed by more than one instance: those are walked through their instances, the others baked
std::vector<bool> instanced_meshes(const rt_scene& scene, bool two_level = true);

// The world triangles of a scene's meshes that aren't instanced (through their instance, or each mesh once, as it
// is, when the scene has no instances), with their materials: what build_bvh puts in the tree in world coordinates, in
// that order before it sorts them (two_level false: every mesh through every instance)
struct world_triangle
{
    double p[3][3];
    int32_t material;
};
std::vector<world_triangle> world_triangles(const rt_scene& scene, bool two_level = true);

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
