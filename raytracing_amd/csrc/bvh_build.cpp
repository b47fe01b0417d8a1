// The BVH builder (see bvh_build.hpp)
#include "bvh_build.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <limits>
#include <numeric>
#include <span>
#include <stdexcept>

namespace rt
{
namespace
{
struct v3d
{
    double x, y, z;
};
v3d cross(v3d u, v3d v) { return {u.y * v.z - u.z * v.y, u.z * v.x - u.x * v.z, u.x * v.y - u.y * v.x}; }
double dot(v3d u, v3d v) { return u.x * v.x + u.y * v.y + u.z * v.z; }
v3d scale(double s, v3d v) { return {s * v.x, s * v.y, s * v.z}; }
v3d from(const rt_vec3& v) { return {v.x, v.y, v.z}; }
template <typename Real>
void store(Real* out, v3d v)
{
    out[0] = static_cast<Real>(v.x);
    out[1] = static_cast<Real>(v.y);
    out[2] = static_cast<Real>(v.z);
}

int longest_axis(const box_d& b)
{
    const double x = b.max[0] - b.min[0], y = b.max[1] - b.min[1], z = b.max[2] - b.min[2];
    if (x > y)
    {
        return x > z ? 0 : 2;
    }
    return y > z ? 1 : 2;
}

// The box of a sphere in Real swept from one center to another, each bound moved out by one step of double
// (src/gpu/bvh.cuh:147-176)
template <typename Real>
box_d sphere_box(const Real* a, const Real* b, Real radius)
{
    box_d box;
    for (int k = 0; k < 3; ++k)
    {
        const double lo = std::min(static_cast<double>(a[k]), static_cast<double>(b[k]));
        const double hi = std::max(static_cast<double>(a[k]), static_cast<double>(b[k]));
        box.min[k] = std::nextafter(lo - static_cast<double>(radius), -INFINITY);
        box.max[k] = std::nextafter(hi + static_cast<double>(radius), INFINITY);
    }
    return box;
}

bool stationary(const rt_sphere_desc& s)
{
    return s.displacement.x == 0 && s.displacement.y == 0 && s.displacement.z == 0;
}

template <typename Real>
class builder
{
public:
    builder(const rt_scene& scene, int max_leaf, bool sah, bool two_level) : max_leaf_(max_leaf), sah_(sah)
    {
        const auto n_spheres = static_cast<std::size_t>(scene.sphere_count);
        for (std::size_t i = 0; i < n_spheres; ++i)
        {
            const rt_sphere_desc& s = scene.spheres[i];
            if (!stationary(s))
            {
                continue;
            }
            dev::sphere_rec<Real> rec{};
            store(rec.center, from(s.center));
            rec.radius = static_cast<Real>(s.radius);
            // onto_sphere's k and limit, in double from the sphere in Real (src/gpu/bvh.cuh:290-303)
            const double x = rec.center[0], y = rec.center[1], z = rec.center[2], r = rec.radius;
            const double c2 = x * x + y * y + z * z;
            dev::sphere_frame<Real> frame{static_cast<Real>(s.uv_cos), static_cast<Real>(s.uv_sin),
                                          static_cast<Real>(c2 - r * r), static_cast<Real>(2 * r * std::sqrt(c2))};
            inputs_.push_back({dev::prim_sphere, static_cast<int>(spheres_.size()), static_cast<int>(i), s.material,
                               frame, sphere_box<Real>(rec.center, rec.center, rec.radius)});
            spheres_.push_back(rec);
        }
        out_.moving_base = static_cast<int32_t>(inputs_.size());
        for (std::size_t i = 0; i < n_spheres; ++i)
        {
            const rt_sphere_desc& s = scene.spheres[i];
            if (stationary(s))
            {
                continue;
            }
            dev::moving_sphere_rec<Real> rec{};
            store(rec.center, from(s.center));
            rec.radius = static_cast<Real>(s.radius);
            store(rec.displacement, from(s.displacement));
            Real end[3];
            for (int k = 0; k < 3; ++k)
            {
                end[k] = rec.center[k] + rec.displacement[k]; // in Real, as the GPU computes the center at time 1
            }
            dev::sphere_frame<Real> frame{static_cast<Real>(s.uv_cos), static_cast<Real>(s.uv_sin), 0, 0};
            inputs_.push_back({dev::prim_moving_sphere, static_cast<int>(moving_.size()), static_cast<int>(i),
                               s.material, frame, sphere_box<Real>(rec.center, end, rec.radius)});
            moving_.push_back(rec);
        }
        out_.quad_base = static_cast<int32_t>(inputs_.size());
        for (int64_t i = 0; i < scene.quad_count; ++i)
        {
            inputs_.push_back({dev::prim_quad, static_cast<int>(quads_.size()), static_cast<int>(i),
                               scene.quads[i].material, {}, quad_box(scene.quads[i])});
            quads_.push_back(make_quad_rec<Real>(scene.quads[i]));
        }
        out_.triangle_base = static_cast<int32_t>(inputs_.size());
        for (const world_triangle& w : world_triangles(scene, two_level))
        {
            inputs_.push_back({dev::prim_triangle, static_cast<int>(triangles_.size()), -1, w.material, {},
                               triangle_box(w)});
            triangles_.push_back(make_triangle_rec(w));
        }
        out_.instanced_base = static_cast<int32_t>(inputs_.size());
        // meshes placed more than once: a tree of their own, once, and one leaf per instance in the scene's tree
        const std::vector<bool> instanced = instanced_meshes(scene, two_level);
        meshes_.resize(static_cast<std::size_t>(scene.mesh_count));
        for (int64_t m = 0; m < scene.mesh_count; ++m)
        {
            if (instanced[static_cast<std::size_t>(m)])
            {
                meshes_[static_cast<std::size_t>(m)].tree = builder(scene.meshes[m]).build();
            }
        }
        for (int64_t k = 0; k < scene.instance_count; ++k)
        {
            const rt_mesh_instance_desc& d = scene.instances[k];
            if (!instanced[static_cast<std::size_t>(d.mesh)])
            {
                continue;
            }
            const bvh_arrays<Real>& tree = meshes_[static_cast<std::size_t>(d.mesh)].tree;
            if (tree.nodes.empty())
            {
                continue; // a mesh without triangles
            }
            inputs_.push_back({dev::prim_instance, static_cast<int>(instance_descs_.size()), static_cast<int>(k),
                               scene.meshes[d.mesh].material, {}, instance_box(tree, d.object_to_world)});
            instance_descs_.push_back(d);
        }
        out_.sphere_ids.resize(n_spheres);
        out_.quad_ids.resize(static_cast<std::size_t>(scene.quad_count));
    }

    // A mesh's own tree: its triangles in its own coordinates, in a surface-area tree
    explicit builder(const rt_mesh_desc& mesh) : max_leaf_(sah_leaf), sah_(true)
    {
        for (int64_t k = 0; k < mesh.triangle_count; ++k)
        {
            world_triangle w{};
            w.material = mesh.material;
            for (int c = 0; c < 3; ++c)
            {
                const float* p = mesh.positions + 3 * static_cast<std::size_t>(mesh.indices[3 * k + c]);
                for (int a = 0; a < 3; ++a)
                {
                    w.p[c][a] = p[a];
                }
            }
            inputs_.push_back({dev::prim_triangle, static_cast<int>(triangles_.size()), -1, w.material, {},
                               triangle_box(w)});
            triangles_.push_back(make_triangle_rec(w));
        }
        out_.instanced_base = static_cast<int32_t>(inputs_.size());
    }

    bvh_arrays<Real> build() &&
    {
        out_.prim_materials.resize(static_cast<std::size_t>(out_.instanced_base));
        out_.prim_ranks.resize(static_cast<std::size_t>(out_.instanced_base));
        out_.sphere_frames.resize(static_cast<std::size_t>(out_.quad_base));
        std::vector<int> order(inputs_.size());
        std::iota(order.begin(), order.end(), 0);
        const bool meshes = !triangles_.empty() || !instance_descs_.empty();
        if (!order.empty())
        {
            if (!meshes && !sah_)
            {
                add_subtree(order, 1);
            }
            else
            {
                centroids_.resize(inputs_.size());
                for (std::size_t i = 0; i < inputs_.size(); ++i)
                {
                    for (int a = 0; a < 3; ++a)
                    {
                        centroids_[i][a] = 0.5 * (inputs_[i].box.min[a] + inputs_[i].box.max[a]);
                    }
                }
                out_.triangles.reserve(triangles_.size());
                add_subtree_sah(order, 1);
            }
            pack_wide();
            append_meshes();
        }
        out_.media_base = out_.instanced_base + next_instanced_id_;
        return std::move(out_);
    }

private:
    struct input
    {
        int kind;
        int index;
        int source;
        int material;
        dev::sphere_frame<Real> frame;
        box_d box;
    };
    int max_leaf_;
    bool sah_;
    std::vector<input> inputs_;
    struct mesh_tree
    {
        bvh_arrays<Real> tree; // empty unless the mesh is instanced
        int32_t root_ref = -1; // in the scene's node array, once appended
        int32_t first_triangle = 0;
    };
    std::vector<mesh_tree> meshes_;
    std::vector<rt_mesh_instance_desc> instance_descs_; // of the instanced meshes, by scene order
    std::vector<int32_t> instance_mesh_;                // of out_.instances, by leaf order
    int32_t next_instanced_id_ = 0;
    std::vector<dev::sphere_rec<Real>> spheres_;
    std::vector<dev::moving_sphere_rec<Real>> moving_;
    std::vector<dev::quad_rec<Real>> quads_;
    std::vector<dev::triangle_rec<Real>> triangles_;
    std::vector<std::array<double, 3>> centroids_;
    bvh_arrays<Real> out_;

    int array_size(int kind) const
    {
        return kind == dev::prim_sphere          ? static_cast<int>(out_.spheres.size())
               : kind == dev::prim_moving_sphere ? static_cast<int>(out_.moving_spheres.size())
               : kind == dev::prim_quad          ? static_cast<int>(out_.quads.size())
               : kind == dev::prim_instance      ? static_cast<int>(out_.instances.size())
                                                 : static_cast<int>(out_.triangles.size());
    }

    // An instance's box in the world: the mesh tree's root box (its coordinates) through the instance's transform,
    // every corner, padded as a triangle's
    static box_d instance_box(const bvh_arrays<Real>& tree, const double* m)
    {
        box_d box;
        const ref_node<Real>& root = tree.nodes.front();
        double largest = 0;
        for (int corner = 0; corner < 8; ++corner)
        {
            const double x = corner & 1 ? root.box_max[0] : root.box_min[0];
            const double y = corner & 2 ? root.box_max[1] : root.box_min[1];
            const double z = corner & 4 ? root.box_max[2] : root.box_min[2];
            for (int a = 0; a < 3; ++a)
            {
                const double w = m[4 * a] * x + m[4 * a + 1] * y + m[4 * a + 2] * z + m[4 * a + 3];
                box.min[a] = std::min(box.min[a], w);
                box.max[a] = std::max(box.max[a], w);
                largest = std::max(largest, std::fabs(w));
            }
        }
        const double margin = 1e-6 * (1 + largest);
        for (int a = 0; a < 3; ++a)
        {
            box.min[a] -= margin;
            box.max[a] += margin;
        }
        return box;
    }

    // The world-to-object transform of an object-to-world one (row-major 3x4), inverted in double by cofactors
    static void invert_affine(const double* m, Real* out)
    {
        const double a = m[0], b = m[1], c = m[2], d = m[4], e = m[5], f = m[6], g = m[8], h = m[9], i = m[10];
        const double co[9] = {e * i - f * h, c * h - b * i, b * f - c * e, f * g - d * i, a * i - c * g,
                              c * d - a * f, d * h - e * g, b * g - a * h, a * e - b * d};
        const double det = a * co[0] + b * co[3] + c * co[6];
        double scale = 0;
        for (int k = 0; k < 9; ++k)
        {
            scale = std::max(scale, std::fabs(m[k + k / 3]));
        }
        if (!(std::fabs(det) > 1e-12 * scale * scale * scale))
        {
            throw std::invalid_argument("an instance's transform is singular");
        }
        const double t[3] = {m[3], m[7], m[11]};
        for (int r = 0; r < 3; ++r)
        {
            double shift = 0;
            for (int k = 0; k < 3; ++k)
            {
                const double v = co[3 * r + k] / det;
                out[4 * r + k] = static_cast<Real>(v);
                shift -= v * t[k];
            }
            out[4 * r + 3] = static_cast<Real>(shift);
        }
    }

    // A triangle's box in double, padded as a quad's (a flat box would let rounding in the slab test lose its hits)
    static box_d triangle_box(const world_triangle& w)
    {
        box_d box;
        double largest = 0;
        for (const auto& p : w.p)
        {
            for (int a = 0; a < 3; ++a)
            {
                box.min[a] = std::min(box.min[a], p[a]);
                box.max[a] = std::max(box.max[a], p[a]);
                largest = std::max(largest, std::fabs(p[a]));
            }
        }
        const double margin = 1e-6 * (1 + largest);
        for (int a = 0; a < 3; ++a)
        {
            box.min[a] -= margin;
            box.max[a] += margin;
            if (box.max[a] - box.min[a] < 0.0001)
            {
                box.min[a] -= 0.0001 / 2;
                box.max[a] += 0.0001 / 2;
            }
        }
        return box;
    }

    // v0, the edges from it and the unit normal, in double from the world vertices, each rounded once
    static dev::triangle_rec<Real> make_triangle_rec(const world_triangle& w)
    {
        const v3d p0{w.p[0][0], w.p[0][1], w.p[0][2]};
        const v3d e1{w.p[1][0] - p0.x, w.p[1][1] - p0.y, w.p[1][2] - p0.z};
        const v3d e2{w.p[2][0] - p0.x, w.p[2][1] - p0.y, w.p[2][2] - p0.z};
        const v3d n = cross(e1, e2);
        const double length = std::sqrt(dot(n, n));
        dev::triangle_rec<Real> t{};
        store(t.v0, p0);
        store(t.e1, e1);
        store(t.e2, e2);
        store(t.normal, length > 0 ? scale(1 / length, n) : v3d{0, 0, 0});
        return t;
    }

    void add_to_leaf(const input& in)
    {
        int id = 0;
        if (in.kind == dev::prim_instance)
        {
            const rt_mesh_instance_desc* d = &instance_descs_[static_cast<std::size_t>(in.index)];
            dev::instance_rec<Real> r{};
            invert_affine(d->object_to_world, r.to_object);
            r.id_base = out_.instanced_base + next_instanced_id_;
            r.material = in.material;
            const mesh_tree& mt = meshes_[static_cast<std::size_t>(d->mesh)];
            if (static_cast<int64_t>(next_instanced_id_) + static_cast<int64_t>(mt.tree.triangles.size()) +
                    out_.instanced_base > INT32_MAX / 2)
            {
                throw std::runtime_error("too many instanced triangles for 32-bit primitive ids");
            }
            next_instanced_id_ += static_cast<int32_t>(mt.tree.triangles.size());
            out_.instances.push_back(r);
            instance_mesh_.push_back(d->mesh);
            return;
        }
        switch (in.kind)
        {
        case dev::prim_sphere:
            id = static_cast<int>(out_.spheres.size());
            out_.spheres.push_back(spheres_[in.index]);
            break;
        case dev::prim_moving_sphere:
            id = out_.moving_base + static_cast<int>(out_.moving_spheres.size());
            out_.moving_spheres.push_back(moving_[in.index]);
            break;
        case dev::prim_triangle:
            id = out_.triangle_base + static_cast<int>(out_.triangles.size());
            out_.triangles.push_back(triangles_[in.index]);
            out_.prim_materials[id] = in.material;
            out_.prim_ranks[id] = static_cast<int32_t>(out_.sphere_ids.size() + out_.quad_ids.size()) + in.index;
            return;
        default:
            id = out_.quad_base + static_cast<int>(out_.quads.size());
            out_.quads.push_back(quads_[in.index]);
            out_.prim_materials[id] = in.material;
            out_.prim_ranks[id] = static_cast<int32_t>(out_.sphere_ids.size()) + in.source;
            out_.quad_ids[in.source] = id;
            return;
        }
        out_.prim_materials[id] = in.material;
        out_.prim_ranks[id] = in.source;
        out_.sphere_frames[id] = in.frame;
        out_.sphere_ids[in.source] = id;
    }

    // src/gpu/bvh.cuh:382-435, with a stable sort
    int add_subtree(std::span<int> indices, int level)
    {
        if (level > 32)
        {
            throw std::runtime_error("the BVH is too deep for the traversal stack");
        }
        out_.depth = std::max(out_.depth, level);
        box_d box;
        for (const int i : indices)
        {
            box.enclose(inputs_[i].box);
        }
        const auto index = static_cast<int>(out_.nodes.size());
        ref_node<Real> node{};
        for (int k = 0; k < 3; ++k)
        {
            node.box_min[k] = round_down<Real>(box.min[k]);
            node.box_max[k] = round_up<Real>(box.max[k]);
        }
        out_.nodes.push_back(node);
        std::size_t middle = indices.size() / 2;
        if (static_cast<int>(indices.size()) <= max_leaf_)
        {
            int kind = 3;
            for (const int i : indices)
            {
                kind = std::min(kind, inputs_[i].kind);
            }
            const auto split = std::stable_partition(indices.begin(), indices.end(),
                                                     [&](int i) { return inputs_[i].kind == kind; });
            if (split == indices.end())
            {
                out_.nodes[index].index = array_size(kind);
                out_.nodes[index].count = static_cast<int32_t>(indices.size()) | kind << 8;
                for (const int i : indices)
                {
                    add_to_leaf(inputs_[i]);
                }
                return index;
            }
            middle = static_cast<std::size_t>(split - indices.begin());
        }
        else
        {
            const int axis = longest_axis(box);
            std::stable_sort(indices.begin(), indices.end(),
                             [&](int a, int b) { return inputs_[a].box.min[axis] < inputs_[b].box.min[axis]; });
        }
        add_subtree(indices.first(middle), level + 1);
        out_.nodes[index].index = add_subtree(indices.subspan(middle), level + 1);
        return index;
    }

    static double half_area(const box_d& b)
    {
        const double x = b.max[0] - b.min[0], y = b.max[1] - b.min[1], z = b.max[2] - b.min[2];
        return x * y + y * z + z * x;
    }

    // The tree of a scene with triangles: binned surface-area-heuristic splits (16 bins on each axis of the
    // centroids' box), leaves of one kind and at most sah_leaf primitives where a split wouldn't pay. Below
    // sah_levels levels, object-median splits bound the depth (and the traversal stack).
    static constexpr int sah_bins = 16;
    static constexpr int sah_leaf = 8;
    static constexpr int sah_levels = 40;
    static constexpr int max_levels = 64;

    int add_subtree_sah(std::span<int> indices, int level)
    {
        if (level > max_levels)
        {
            throw std::runtime_error("the BVH is too deep for the traversal stack");
        }
        out_.depth = std::max(out_.depth, level);
        box_d box, centers;
        bool one_kind = true;
        for (const int i : indices)
        {
            box.enclose(inputs_[i].box);
            one_kind = one_kind && inputs_[i].kind == inputs_[indices[0]].kind;
            for (int a = 0; a < 3; ++a)
            {
                centers.min[a] = std::min(centers.min[a], centroids_[i][a]);
                centers.max[a] = std::max(centers.max[a], centroids_[i][a]);
            }
        }
        const auto index = static_cast<int>(out_.nodes.size());
        ref_node<Real> node{};
        for (int k = 0; k < 3; ++k)
        {
            node.box_min[k] = round_down<Real>(box.min[k]);
            node.box_max[k] = round_up<Real>(box.max[k]);
        }
        out_.nodes.push_back(node);
        const std::size_t n = indices.size();
        // an instance is a leaf of its own (the walk enters one mesh tree at a time)
        const bool instances = one_kind && inputs_[indices[0]].kind == dev::prim_instance;
        if (n == 1 || (one_kind && !instances && n <= 2))
        {
            return leaf(index, indices);
        }

        // the best plane over the three axes: cost = A_left N_left + A_right N_right, relative to the node's area
        int best_axis = -1, best_bin = 0;
        double best_cost = std::numeric_limits<double>::infinity();
        if (level < sah_levels)
        {
            for (int a = 0; a < 3; ++a)
            {
                const double extent = centers.max[a] - centers.min[a];
                if (!(extent > 0))
                {
                    continue;
                }
                const double to_bin = sah_bins / extent;
                box_d bin_box[sah_bins];
                std::size_t bin_count[sah_bins] = {};
                for (const int i : indices)
                {
                    const int b = std::min(sah_bins - 1, static_cast<int>((centroids_[i][a] - centers.min[a]) * to_bin));
                    bin_box[b].enclose(inputs_[i].box);
                    ++bin_count[b];
                }
                double right_area[sah_bins];
                std::size_t right_count[sah_bins];
                box_d acc;
                std::size_t count = 0;
                for (int b = sah_bins - 1; b > 0; --b)
                {
                    acc.enclose(bin_box[b]);
                    count += bin_count[b];
                    right_area[b] = count ? half_area(acc) : 0;
                    right_count[b] = count;
                }
                acc = box_d{};
                count = 0;
                for (int b = 0; b + 1 < sah_bins; ++b)
                {
                    acc.enclose(bin_box[b]);
                    count += bin_count[b];
                    if (count == 0 || right_count[b + 1] == 0)
                    {
                        continue;
                    }
                    const double cost = half_area(acc) * static_cast<double>(count) +
                                        right_area[b + 1] * static_cast<double>(right_count[b + 1]);
                    if (cost < best_cost)
                    {
                        best_cost = cost;
                        best_axis = a;
                        best_bin = b;
                    }
                }
            }
        }
        const double parent = half_area(box);
        const double leaf_cost = static_cast<double>(n);
        const double split_cost = parent > 0 ? 1.0 + best_cost / parent : std::numeric_limits<double>::infinity();
        if (n <= static_cast<std::size_t>(sah_leaf) && one_kind && !instances && (best_axis < 0 || leaf_cost <= split_cost))
        {
            return leaf(index, indices);
        }
        std::size_t middle = 0;
        if (n <= static_cast<std::size_t>(sah_leaf) && !one_kind && (best_axis < 0 || leaf_cost <= split_cost))
        {
            // few primitives of several kinds: one subtree per kind, as the reference's builder splits a small leaf
            int kind = dev::prim_instance;
            for (const int i : indices)
            {
                kind = std::min(kind, inputs_[i].kind);
            }
            middle = static_cast<std::size_t>(
                std::stable_partition(indices.begin(), indices.end(), [&](int i) { return inputs_[i].kind == kind; }) -
                indices.begin());
        }
        else if (best_axis >= 0)
        {
            const double extent = centers.max[best_axis] - centers.min[best_axis];
            const double to_bin = sah_bins / extent;
            const auto split = std::partition(indices.begin(), indices.end(), [&](int i) {
                const int b = std::min(sah_bins - 1,
                                       static_cast<int>((centroids_[i][best_axis] - centers.min[best_axis]) * to_bin));
                return b <= best_bin;
            });
            middle = static_cast<std::size_t>(split - indices.begin());
        }
        if (middle == 0 || middle == n)
        {
            // no plane separates the centroids (or below the SAH levels): halves by the longest axis
            const int axis = longest_axis(centers.max[0] > centers.min[0] || centers.max[1] > centers.min[1] ||
                                                  centers.max[2] > centers.min[2]
                                              ? centers
                                              : box);
            middle = n / 2;
            std::nth_element(indices.begin(), indices.begin() + static_cast<std::ptrdiff_t>(middle), indices.end(),
                             [&](int a, int b) {
                                 return centroids_[a][axis] < centroids_[b][axis] ||
                                        (centroids_[a][axis] == centroids_[b][axis] && a < b);
                             });
        }
        add_subtree_sah(indices.first(middle), level + 1);
        out_.nodes[index].index = add_subtree_sah(indices.subspan(middle), level + 1);
        return index;
    }

    int leaf(int index, std::span<int> indices)
    {
        const int kind = inputs_[indices[0]].kind;
        out_.nodes[index].index = array_size(kind);
        out_.nodes[index].count = static_cast<int32_t>(indices.size()) | kind << 8;
        for (const int i : indices)
        {
            add_to_leaf(inputs_[i]);
        }
        return index;
    }

    int32_t leaf_ref_of(int node) const
    {
        const ref_node<Real>& n = out_.nodes[node];
        const int count = n.count & 255;
        if (n.index > dev::max_leaf_first || count > 16)
        {
            throw std::runtime_error("a BVH leaf beyond the packed reference's range");
        }
        return dev::make_leaf_ref(n.count >> 8, count, n.index);
    }

    double node_area(int node) const
    {
        const ref_node<Real>& n = out_.nodes[node];
        const double x = static_cast<double>(n.box_max[0]) - n.box_min[0];
        const double y = static_cast<double>(n.box_max[1]) - n.box_min[1];
        const double z = static_cast<double>(n.box_max[2]) - n.box_min[2];
        return x * y + y * z + z * x;
    }

    // The binary tree collapsed four-wide. A wide node starts from its binary node's two children and, while it has
    // fewer than four, replaces its inner child of largest surface area by that child's own two children (in place,
    // so the children keep the binary tree's depth-first order). Wide nodes are numbered in depth-first order from the
    // root. `pushed` is what the walk can have on its stack on arrival; the stack's exact worst case is the largest
    // pushed + (children - 1) over the wide nodes.
    int32_t collapse(int node, int pushed)
    {
        std::array<int, dev::wide_children> kids{node + 1, out_.nodes[node].index, -1, -1};
        int count = 2;
        while (count < dev::wide_children)
        {
            int best = -1;
            double best_area = -1;
            for (int c = 0; c < count; ++c)
            {
                if (out_.nodes[kids[c]].count == 0 && node_area(kids[c]) > best_area)
                {
                    best = c;
                    best_area = node_area(kids[c]);
                }
            }
            if (best < 0)
            {
                break;
            }
            const int inner = kids[best];
            for (int c = count; c > best + 1; --c)
            {
                kids[c] = kids[c - 1];
            }
            kids[best] = inner + 1;
            kids[best + 1] = out_.nodes[inner].index;
            ++count;
        }
        const auto me = static_cast<int32_t>(out_.wide.size());
        out_.wide.emplace_back();
        out_.stack_entries = std::max(out_.stack_entries, pushed + count - 1);
        for (int c = 0; c < count; ++c)
        {
            entering_instance(kids[c], pushed + count - 1);
        }
        int32_t refs[dev::wide_children];
        for (int c = 0; c < dev::wide_children; ++c)
        {
            refs[c] = c >= count                       ? dev::no_child
                      : out_.nodes[kids[c]].count != 0 ? leaf_ref_of(kids[c])
                                                       : collapse(kids[c], pushed + count - 1);
        }
        dev::wide_node<Real>& w = out_.wide[static_cast<std::size_t>(me)];
        for (int c = 0; c < dev::wide_children; ++c)
        {
            w.ref[c] = refs[c];
            for (int a = 0; a < 3; ++a)
            {
                w.lo[a][c] = c < count ? out_.nodes[kids[c]].box_min[a] : std::numeric_limits<Real>::infinity();
                w.hi[a][c] = c < count ? out_.nodes[kids[c]].box_max[a] : std::numeric_limits<Real>::infinity();
            }
        }
        return me;
    }

    void pack_wide()
    {
        out_.wide.clear();
        out_.stack_entries = 1;
        entering_instance(0, 0);
        out_.root_ref = out_.nodes[0].count != 0 ? leaf_ref_of(0) : collapse(0, 0);
    }

    // A walk that reaches an instance's leaf with `pushed` entries on its stack walks the mesh's tree on top of them
    int32_t instance_stack_ = 0;
    void entering_instance(int node, int pushed)
    {
        const ref_node<Real>& n = out_.nodes[static_cast<std::size_t>(node)];
        if (n.count != 0 && (n.count >> 8) == dev::prim_instance)
        {
            const int32_t mesh = instance_mesh_[static_cast<std::size_t>(n.index)];
            instance_stack_ =
                std::max(instance_stack_, pushed + meshes_[static_cast<std::size_t>(mesh)].tree.stack_entries);
        }
    }

    // The instanced meshes' trees after the scene's: their wide nodes and triangles appended to the scene's arrays
    // (inner references moved by the node offset, leaves by the triangle offset), each instance pointed at its mesh's
    // root. The stack's worst case: the scene tree's, or inside an instance what was on the stack at its leaf plus its
    // mesh tree's worst (entering_instance).
    static int32_t moved(int32_t ref, int32_t node_offset, int32_t triangle_offset)
    {
        if (ref == dev::no_child)
        {
            return ref;
        }
        if (ref >= 0)
        {
            return ref + node_offset;
        }
        const int32_t first = dev::leaf_first(ref) + triangle_offset;
        if (first + dev::leaf_count(ref) - 1 > dev::max_leaf_first)
        {
            throw std::runtime_error("an instanced mesh's leaf beyond the packed reference's range");
        }
        return dev::make_leaf_ref(dev::leaf_kind(ref), dev::leaf_count(ref), first);
    }
    void append_meshes()
    {
        for (mesh_tree& mt : meshes_)
        {
            if (mt.tree.triangles.empty())
            {
                continue;
            }
            const auto node_offset = static_cast<int32_t>(out_.wide.size());
            const auto triangle_offset = static_cast<int32_t>(out_.triangles.size());
            for (dev::wide_node<Real> w : mt.tree.wide)
            {
                for (int32_t& ref : w.ref)
                {
                    ref = moved(ref, node_offset, triangle_offset);
                }
                out_.wide.push_back(w);
            }
            mt.root_ref = moved(mt.tree.root_ref, node_offset, triangle_offset);
            mt.first_triangle = triangle_offset;
            out_.triangles.insert(out_.triangles.end(), mt.tree.triangles.begin(), mt.tree.triangles.end());
        }
        for (std::size_t k = 0; k < out_.instances.size(); ++k)
        {
            const mesh_tree& mt = meshes_[static_cast<std::size_t>(instance_mesh_[k])];
            out_.instances[k].root_ref = mt.root_ref;
            out_.instances[k].first_triangle = mt.first_triangle;
        }
        out_.stack_entries = std::max(out_.stack_entries, instance_stack_);
    }
};
}

void box_d::enclose(const box_d& o)
{
    for (int a = 0; a < 3; ++a)
    {
        min[a] = std::min(min[a], o.min[a]);
        max[a] = std::max(max[a], o.max[a]);
    }
}

template <typename Real>
Real round_down(double x)
{
    const auto r = static_cast<Real>(x);
    return r <= x ? r : std::nextafter(r, -std::numeric_limits<Real>::infinity());
}
template <typename Real>
Real round_up(double x)
{
    const auto r = static_cast<Real>(x);
    return r >= x ? r : std::nextafter(r, std::numeric_limits<Real>::infinity());
}

box_d quad_box(const rt_quad_desc& d)
{
    const v3d q = from(d.q), u = from(d.u), v = from(d.v);
    const v3d corners[4] = {q, {q.x + u.x, q.y + u.y, q.z + u.z}, {q.x + v.x, q.y + v.y, q.z + v.z},
                            {q.x + u.x + v.x, q.y + u.y + v.y, q.z + u.z + v.z}};
    box_d box;
    double largest = 0;
    for (const v3d& c : corners)
    {
        const double xyz[3] = {c.x, c.y, c.z};
        for (int a = 0; a < 3; ++a)
        {
            box.min[a] = std::min(box.min[a], xyz[a]);
            box.max[a] = std::max(box.max[a], xyz[a]);
            largest = std::max(largest, std::fabs(xyz[a]));
        }
    }
    const double margin = 1e-6 * (1 + largest);
    for (int a = 0; a < 3; ++a)
    {
        box.min[a] -= margin;
        box.max[a] += margin;
        if (box.max[a] - box.min[a] < 0.0001)
        {
            box.min[a] -= 0.0001 / 2;
            box.max[a] += 0.0001 / 2;
        }
    }
    return box;
}

template <typename Real>
dev::quad_rec<Real> make_quad_rec(const rt_quad_desc& d)
{
    const v3d n = cross(from(d.u), from(d.v));
    const double nn = dot(n, n);
    dev::quad_rec<Real> q{};
    store(q.q, from(d.q));
    store(q.normal, scale(1 / std::sqrt(nn), n));
    const v3d w = scale(1 / nn, n);
    store(q.alpha_axis, cross(from(d.v), w));
    store(q.beta_axis, cross(w, from(d.u)));
    return q;
}

int leaf_size_for(std::size_t shapes) { return shapes <= 8 ? 8 : 1; }

std::vector<dev::packed_node> pack_nodes(const std::vector<dev::wide_node<float>>& wide)
{
    std::vector<dev::packed_node> out(wide.size());
    for (std::size_t n = 0; n < wide.size(); ++n)
    {
        const dev::wide_node<float>& w = wide[n];
        dev::packed_node& p = out[n];
        for (int a = 0; a < 3; ++a)
        {
            // the node's box on this axis: the union of its children's
            double lo = std::numeric_limits<double>::infinity();
            double hi = -lo;
            for (int c = 0; c < dev::wide_children; ++c)
            {
                if (w.ref[c] != dev::no_child)
                {
                    lo = std::min(lo, static_cast<double>(w.lo[a][c]));
                    hi = std::max(hi, static_cast<double>(w.hi[a][c]));
                }
            }
            if (!(lo <= hi) || !std::isfinite(lo) || !std::isfinite(hi))
            {
                throw std::runtime_error("packing the BVH: a node without a finite box");
            }
            // the step: the smallest power of two that spans the box in at most 254 steps (one more for the corner's
            // rounding down), and no finer than the last place of the box's largest coordinate over 2^22: then every
            // grid point corner + q step (q <= 255) is a whole multiple of the step below 2^24 of them, a float, and
            // the kernels' fma(q, step, corner) gives it exactly
            const double magnitude = std::max(std::fabs(lo), std::fabs(hi));
            int k = magnitude > 0 ? std::ilogb(magnitude) - 22 : -126;
            if (hi > lo)
            {
                k = std::max(k, std::ilogb((hi - lo) / 254.0));
                while (std::ldexp(254.0, k) < hi - lo)
                {
                    ++k;
                }
            }
            k = std::max(k, -126);
            const double step = std::ldexp(1.0, k);
            // the corner: the multiple of the step at or below the box
            const double corner = std::floor(lo / step) * step;
            if (static_cast<double>(static_cast<float>(corner)) != corner ||
                std::fabs(corner / step) + 255 >= 16777216.0)
            {
                throw std::runtime_error("packing the BVH: a node's grid isn't made of floats");
            }
            p.origin[a] = static_cast<float>(corner);
            p.step[a] = static_cast<float>(step);
            for (int c = 0; c < dev::wide_children; ++c)
            {
                if (w.ref[c] == dev::no_child)
                {
                    p.lo[a][c] = 255;
                    p.hi[a][c] = 0;
                    continue;
                }
                // each plane rounded outward to the grid: the nearest grid points at or beyond it (the difference in
                // double may round when a plane is far finer than the step, as a box's rounded-out zero is: the
                // grid point is then moved on until it holds the plane, compared exactly)
                const double lo_c = w.lo[a][c];
                const double hi_c = w.hi[a][c];
                double q_lo = std::floor((lo_c - corner) / step);
                double q_hi = std::ceil((hi_c - corner) / step);
                while (corner + q_lo * step > lo_c)
                {
                    q_lo -= 1;
                }
                while (corner + q_hi * step < hi_c)
                {
                    q_hi += 1;
                }
                if (q_lo < 0 || q_hi > 255 || q_lo > q_hi)
                {
                    throw std::runtime_error("packing the BVH: a child's box leaves its node's grid");
                }
                p.lo[a][c] = static_cast<uint8_t>(q_lo);
                p.hi[a][c] = static_cast<uint8_t>(q_hi);
            }
        }
        for (int c = 0; c < dev::wide_children; ++c)
        {
            p.ref[c] = w.ref[c];
        }
    }
    return out;
}

std::vector<bool> instanced_meshes(const rt_scene& scene, bool two_level)
{
    std::vector<int> uses(static_cast<std::size_t>(std::max<int64_t>(scene.mesh_count, 0)), 0);
    for (int64_t k = 0; k < scene.instance_count; ++k)
    {
        const int32_t m = scene.instances[k].mesh;
        if (m >= 0 && m < scene.mesh_count)
        {
            ++uses[static_cast<std::size_t>(m)];
        }
    }
    std::vector<bool> out(uses.size());
    for (std::size_t m = 0; m < uses.size(); ++m)
    {
        out[m] = two_level && uses[m] > 1;
    }
    return out;
}

std::vector<world_triangle> world_triangles(const rt_scene& scene, bool two_level)
{
    const std::vector<bool> instanced = instanced_meshes(scene, two_level);
    std::vector<world_triangle> out;
    const auto place = [&](const rt_mesh_desc& m, const double* transform) {
        for (int64_t k = 0; k < m.triangle_count; ++k)
        {
            world_triangle w{};
            w.material = m.material;
            for (int c = 0; c < 3; ++c)
            {
                const float* p = m.positions + 3 * static_cast<std::size_t>(m.indices[3 * k + c]);
                const double x = p[0], y = p[1], z = p[2];
                for (int a = 0; a < 3; ++a)
                {
                    w.p[c][a] = transform ? transform[4 * a] * x + transform[4 * a + 1] * y + transform[4 * a + 2] * z +
                                                transform[4 * a + 3]
                                          : (a == 0 ? x : (a == 1 ? y : z));
                }
            }
            out.push_back(w);
        }
    };
    if (scene.instance_count == 0)
    {
        for (int64_t i = 0; i < scene.mesh_count; ++i)
        {
            place(scene.meshes[i], nullptr);
        }
    }
    for (int64_t i = 0; i < scene.instance_count; ++i)
    {
        if (!instanced[static_cast<std::size_t>(scene.instances[i].mesh)])
        {
            place(scene.meshes[scene.instances[i].mesh], scene.instances[i].object_to_world);
        }
    }
    return out;
}

template <typename Real>
bvh_arrays<Real> build_bvh(const rt_scene& scene, int max_leaf_size, int32_t tree)
{
    if (tree < RT_TREE_DEFAULT || tree > RT_TREE_TWO_LEVEL)
    {
        throw std::invalid_argument("unknown tree");
    }
    return builder<Real>(scene, max_leaf_size, tree != RT_TREE_MEDIAN, tree == RT_TREE_TWO_LEVEL).build();
}

template bvh_arrays<float> build_bvh<float>(const rt_scene&, int, int32_t);
template bvh_arrays<double> build_bvh<double>(const rt_scene&, int, int32_t);
template dev::quad_rec<float> make_quad_rec<float>(const rt_quad_desc&);
template dev::quad_rec<double> make_quad_rec<double>(const rt_quad_desc&);
template float round_down<float>(double);
template double round_down<double>(double);
template float round_up<float>(double);
template double round_up<double>(double);
}
