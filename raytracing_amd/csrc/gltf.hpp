// glTF 2.0 models (host code): the triangle meshes of a .gltf file (JSON, with base64 data: buffers or buffer files
// beside it) or a .glb file (the binary container), placed by the scene's node hierarchy, with their metallic-roughness
// materials' factors. It takes the place of the reference's load_gltf (src/gltf.ixx:34), which feeds its OptiX tracer;
// here the models render through the HIP kernels' triangle path. Textures, skins, morph targets, cameras and lights are
// not read; what a model holds that isn't rendered is reported in `warnings`, one line each.
#pragma once

#include <array>
#include <cstdint>
#include <string>
#include <vector>

namespace rt
{
struct gltf_material
{
    std::array<double, 3> base_color{1, 1, 1}; // pbrMetallicRoughness.baseColorFactor (linear)
    double metallic = 1;                       // glTF's default material: base 1, metallic 1, roughness 1
    double roughness = 1;
};

struct gltf_mesh
{
    std::vector<float> positions;  // x, y, z per vertex, in the mesh's own coordinates
    std::vector<uint32_t> indices; // three per triangle
    int material = -1;             // into gltf_model::materials; -1 for glTF's default material
};

struct gltf_placement
{
    int mesh = 0;                // into gltf_model::meshes
    double object_to_world[12]; // rows of a 3x4 matrix: the node's world transform
};

struct gltf_model
{
    std::vector<gltf_material> materials;
    std::vector<gltf_mesh> meshes;          // one per (glTF mesh, primitive) with triangles
    std::vector<gltf_placement> placements; // one per node (of the scene) that holds a mesh, per primitive
    std::vector<std::string> warnings;
};

// Loads a model. Throws std::runtime_error with a message that says why for a file it refuses: one it can't read or
// that is cut short, JSON that doesn't parse, a version other than 2, an extension it requires that isn't supported
// (compressed meshes among them), an accessor past its buffer, an index past its vertices, a node hierarchy with a
// cycle, or positions that aren't three floats.
gltf_model load_gltf(const std::string& path);

// Whether a scene name is a model's file: it ends in .gltf or .glb (no built-in scene's name does)
bool is_model_path(const std::string& name);
}
