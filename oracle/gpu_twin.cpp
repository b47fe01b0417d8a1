// ORACLE — test infrastructure only (see cpu_model.hpp). The GPU-semantics twin: a plain C++ restatement, in float,
// of the reference GPU renderer's random numbers (src/gpu/random.cuh:25-90), its BVH builder (src/gpu/bvh.cuh:100-444)
// and its closest-hit traversal (src/gpu/bvh.cuh:483-706, src/gpu/sphere.cuh:70-117, src/gpu/quad.cuh:33-84). It is the
// bit-exact oracle for the GPU's closest-hit ids and random numbers.
//
// Three choices are fixed here where the reference leaves them to its tools or its walk, and the HIP kernels make the
// same ones:
//   - the BVH builder sorts with std::stable_sort (the reference's std::ranges::sort leaves the order of equal keys to
//     the standard library);
//   - ties, two rules (tie_rule below). The product's closest-hit walk: of two primitives hit at exactly the same
//     distance the one of lower index in the scene is the closest hit (`better`), so the closest hit doesn't depend on
//     the tree or the walk, and the brute-force search finds the same one. The reference's own (walk_order): its walk
//     keeps the first sphere it meets at a distance (an open interval, src/gpu/sphere.cuh:116) and the last quad (a
//     closed one, src/gpu/quad.cuh:70), in the order of its binary near-first walk (src/gpu/bvh.cuh:576-706), so a quad
//     always wins a tie against a sphere;
//   - multiply-adds: the reference's float code is compiled by nvcc, which fuses a·b + c into one fma wherever it
//     chooses. Here every fusion is written out and nothing else is fused (built with -ffp-contract=off):
//       dot(u, v)        = fma(u.x, v.x, fma(u.y, v.y, u.z·v.z))
//       slab distance    = fma(bound, 1/d, -(origin·(1/d)))              (box_entry, "one multiply-add" per slab)
//       l = oc - (h/a)·d = fma(-(h/a), d, oc)                            (sphere_roots)
//       discriminant     = fma(-l.x, l.x, fma(-l.y, l.y, fma(-l.z, l.z, r·r)))   (as the reference writes it)
//       moving center    = fma(time, displacement, center)
//       quad point       = fma(t, d, -(q - origin))
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <span>
#include <stdexcept>
#include <string>
#include <vector>

#include "cpu_model.hpp"

namespace twin
{
// --- random numbers (src/gpu/random.cuh:25-90) ----------------------------------------------------------------------
struct pcg32
{
    std::uint64_t state = 0, inc = 0;
    pcg32(std::uint64_t initstate, std::uint64_t initseq) : state(0), inc((initseq << 1u) | 1u)
    {
        next();
        state += initstate;
        next();
    }
    std::uint32_t next()
    {
        const std::uint64_t old = state;
        state = old * 6364136223846793005ull + inc;
        const auto xorshifted = static_cast<std::uint32_t>(((old >> 18u) ^ old) >> 27u);
        const auto rot = static_cast<std::uint32_t>(old >> 59u);
        return (xorshifted >> rot) | (xorshifted << ((32u - rot) & 31u));
    }
};
inline std::uint64_t mix64(std::uint64_t z)
{
    z = (z ^ (z >> 30u)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27u)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31u);
}
inline pcg32 sample_generator(std::uint64_t seed, std::uint64_t pixel, std::uint64_t pass)
{
    const std::uint64_t initstate = mix64(mix64(mix64(seed) ^ pixel) ^ pass);
    const std::uint64_t initseq = mix64(mix64(mix64(seed ^ 0x9e3779b97f4a7c15ull) ^ pass) ^ pixel);
    return pcg32{initstate, initseq};
}
inline float random_float(pcg32& g) { return static_cast<float>(g.next() >> 8u) * 0x1p-24f; }

// --- float vectors with the fixed multiply-adds ----------------------------------------------------------------------
struct f3
{
    float x = 0, y = 0, z = 0;
};
inline f3 sub(const f3& a, const f3& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline float dot(const f3& u, const f3& v) { return std::fmaf(u.x, v.x, std::fmaf(u.y, v.y, u.z * v.z)); }
inline f3 to_f3(const oracle::vec3& v)
{
    return {static_cast<float>(v.x), static_cast<float>(v.y), static_cast<float>(v.z)};
}

struct sphere
{
    f3 center;
    float radius;
};
struct moving_sphere
{
    f3 center;
    float radius;
    f3 displacement;
    sphere at(float time) const
    {
        return {{std::fmaf(time, displacement.x, center.x), std::fmaf(time, displacement.y, center.y),
                 std::fmaf(time, displacement.z, center.z)},
                radius};
    }
};
struct quad
{
    f3 q, normal, alpha_axis, beta_axis;
};
struct node
{
    f3 box_min;
    int index = 0;
    f3 box_max;
    int count = 0;
};
struct ray
{
    f3 origin, direction;
    float time = 0;
};

// make_quad (src/gpu/quad.cuh:33-43): the normal and the dual axes in double, rounded at the end
inline quad make_quad(const oracle::flat_quad& d)
{
    using oracle::vec3;
    const vec3 n = oracle::cross(d.u, d.v);
    const double nn = oracle::dot(n, n);
    const vec3 normal = (1 / std::sqrt(nn)) * n;
    const vec3 w = (1 / nn) * n;
    return {to_f3(d.q), to_f3(normal), to_f3(oracle::cross(d.v, w)), to_f3(oracle::cross(w, d.u))};
}

// --- the BVH builder (src/gpu/bvh.cuh:100-444) -----------------------------------------------------------------------
struct build_box
{
    double min[3] = {INFINITY, INFINITY, INFINITY};
    double max[3] = {-INFINITY, -INFINITY, -INFINITY};
    void enclose(const build_box& o)
    {
        for (int a = 0; a < 3; ++a)
        {
            min[a] = std::min(min[a], o.min[a]);
            max[a] = std::max(max[a], o.max[a]);
        }
    }
    int longest_axis() const
    {
        const double x = max[0] - min[0], y = max[1] - min[1], z = max[2] - min[2];
        if (x > y)
        {
            return x > z ? 0 : 2;
        }
        return y > z ? 1 : 2;
    }
};
inline float round_down(double x)
{
    const auto r = static_cast<float>(x);
    return r <= x ? r : std::nextafter(r, -INFINITY);
}
inline float round_up(double x)
{
    const auto r = static_cast<float>(x);
    return r >= x ? r : std::nextafter(r, INFINITY);
}
inline build_box sphere_box(const f3& a, const f3& b, float radius)
{
    const double av[3] = {a.x, a.y, a.z}, bv[3] = {b.x, b.y, b.z};
    build_box box;
    for (int k = 0; k < 3; ++k)
    {
        box.min[k] = std::nextafter(std::min(av[k], bv[k]) - radius, -INFINITY);
        box.max[k] = std::nextafter(std::max(av[k], bv[k]) + radius, INFINITY);
    }
    return box;
}
inline build_box quad_box(const oracle::flat_quad& d)
{
    const oracle::vec3 corners[4] = {d.q, d.q + d.u, d.q + d.v, d.q + d.u + d.v};
    build_box box;
    double largest = 0;
    for (const auto& c : corners)
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

struct tree
{
    std::vector<node> nodes;
    std::vector<sphere> spheres;
    std::vector<moving_sphere> moving;
    std::vector<quad> quads;
    std::vector<int> materials;
    int moving_base = 0, quad_base = 0;
    std::vector<int> sphere_ids, quad_ids;
    std::vector<int> ranks; // by global id: the primitive's index in the scene (spheres, then quads)
    int depth = 0;
};

class builder
{
public:
    builder(const std::vector<oracle::flat_sphere>& spheres, const std::vector<oracle::flat_quad>& quads, int max_leaf)
        : max_leaf_(max_leaf)
    {
        auto stationary = [](const oracle::flat_sphere& s) {
            return s.displacement.x == 0 && s.displacement.y == 0 && s.displacement.z == 0;
        };
        for (std::size_t i = 0; i < spheres.size(); ++i)
        {
            if (stationary(spheres[i]))
            {
                const sphere c{to_f3(spheres[i].center), static_cast<float>(spheres[i].radius)};
                inputs_.push_back({0, static_cast<int>(sph_.size()), static_cast<int>(i), spheres[i].material,
                                   sphere_box(c.center, c.center, c.radius)});
                sph_.push_back(c);
            }
        }
        t_.moving_base = static_cast<int>(inputs_.size());
        for (std::size_t i = 0; i < spheres.size(); ++i)
        {
            if (!stationary(spheres[i]))
            {
                const moving_sphere c{to_f3(spheres[i].center), static_cast<float>(spheres[i].radius),
                                      to_f3(spheres[i].displacement)};
                const f3 end{c.center.x + c.displacement.x, c.center.y + c.displacement.y,
                             c.center.z + c.displacement.z};
                inputs_.push_back({1, static_cast<int>(mov_.size()), static_cast<int>(i), spheres[i].material,
                                   sphere_box(c.center, end, c.radius)});
                mov_.push_back(c);
            }
        }
        t_.quad_base = static_cast<int>(inputs_.size());
        for (std::size_t i = 0; i < quads.size(); ++i)
        {
            inputs_.push_back({2, static_cast<int>(qd_.size()), static_cast<int>(i), quads[i].material,
                               quad_box(quads[i])});
            qd_.push_back(make_quad(quads[i]));
        }
        t_.sphere_ids.resize(spheres.size());
        t_.quad_ids.resize(quads.size());
    }
    tree build() &&
    {
        t_.materials.resize(inputs_.size());
        std::vector<int> order(inputs_.size());
        std::iota(order.begin(), order.end(), 0);
        if (!order.empty())
        {
            add_subtree(order, 1);
        }
        t_.ranks.resize(inputs_.size());
        for (std::size_t k = 0; k < t_.sphere_ids.size(); ++k)
        {
            t_.ranks[static_cast<std::size_t>(t_.sphere_ids[k])] = static_cast<int>(k);
        }
        for (std::size_t k = 0; k < t_.quad_ids.size(); ++k)
        {
            t_.ranks[static_cast<std::size_t>(t_.quad_ids[k])] = static_cast<int>(t_.sphere_ids.size() + k);
        }
        return std::move(t_);
    }

private:
    struct input
    {
        int kind, index, source, material;
        build_box box;
    };
    int max_leaf_;
    std::vector<input> inputs_;
    std::vector<sphere> sph_;
    std::vector<moving_sphere> mov_;
    std::vector<quad> qd_;
    tree t_;

    int tree_size(int kind) const
    {
        return kind == 0 ? static_cast<int>(t_.spheres.size())
                         : kind == 1 ? static_cast<int>(t_.moving.size()) : static_cast<int>(t_.quads.size());
    }
    void add_to_leaf(const input& in)
    {
        int id = 0;
        if (in.kind == 0)
        {
            id = static_cast<int>(t_.spheres.size());
            t_.spheres.push_back(sph_[in.index]);
            t_.sphere_ids[in.source] = id;
        }
        else if (in.kind == 1)
        {
            id = t_.moving_base + static_cast<int>(t_.moving.size());
            t_.moving.push_back(mov_[in.index]);
            t_.sphere_ids[in.source] = id;
        }
        else
        {
            id = t_.quad_base + static_cast<int>(t_.quads.size());
            t_.quads.push_back(qd_[in.index]);
            t_.quad_ids[in.source] = id;
        }
        t_.materials[id] = in.material;
    }
    int add_subtree(std::span<int> indices, int level)
    {
        if (level > 32)
        {
            throw std::runtime_error("the BVH is too deep");
        }
        t_.depth = std::max(t_.depth, level);
        build_box box;
        for (const int i : indices)
        {
            box.enclose(inputs_[i].box);
        }
        const auto index = static_cast<int>(t_.nodes.size());
        t_.nodes.push_back({{round_down(box.min[0]), round_down(box.min[1]), round_down(box.min[2])},
                            0,
                            {round_up(box.max[0]), round_up(box.max[1]), round_up(box.max[2])},
                            0});
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
                t_.nodes[index].index = tree_size(kind);
                t_.nodes[index].count = static_cast<int>(indices.size()) | kind << 8;
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
            const int axis = box.longest_axis();
            std::stable_sort(indices.begin(), indices.end(),
                             [&](int a, int b) { return inputs_[a].box.min[axis] < inputs_[b].box.min[axis]; });
        }
        add_subtree(indices.first(middle), level + 1);
        t_.nodes[index].index = add_subtree(indices.subspan(middle), level + 1);
        return index;
    }
};

// The reference renderer's leaf sizes (src/gpu/renderer.cu:75-82): one primitive per leaf, or one leaf of up to 8 for a
// scene of at most 8 shapes
inline int leaf_size(std::size_t shapes) { return shapes <= 8 ? 8 : 1; }

// --- closest_hit (src/gpu/bvh.cuh:483-706) ---------------------------------------------------------------------------
inline float safe_inverse(float x)
{
    const float tiny = 0x1p-80f;
    return 1.0f / (std::fabs(x) > tiny ? x : std::copysign(tiny, x));
}
inline float box_entry(const node& n, const f3& inv, const f3& os, float t_min, float t_max)
{
    const float x0 = std::fmaf(n.box_min.x, inv.x, -os.x);
    const float x1 = std::fmaf(n.box_max.x, inv.x, -os.x);
    const float y0 = std::fmaf(n.box_min.y, inv.y, -os.y);
    const float y1 = std::fmaf(n.box_max.y, inv.y, -os.y);
    const float z0 = std::fmaf(n.box_min.z, inv.z, -os.z);
    const float z1 = std::fmaf(n.box_max.z, inv.z, -os.z);
    const float enter = std::fmax(std::fmax(std::fmin(x0, x1), std::fmin(y0, y1)), std::fmax(std::fmin(z0, z1), t_min));
    const float exit = std::fmin(std::fmin(std::fmax(x0, x1), std::fmax(y0, y1)), std::fmin(std::fmax(z0, z1), t_max));
    return enter <= exit ? enter : INFINITY;
}
inline bool sphere_roots(const sphere& s, const ray& r, float inverse_a, float& near, float& far)
{
    const f3 oc = sub(s.center, r.origin);
    const float h = dot(r.direction, oc) * inverse_a;
    const f3 l{std::fmaf(-h, r.direction.x, oc.x), std::fmaf(-h, r.direction.y, oc.y),
               std::fmaf(-h, r.direction.z, oc.z)};
    const float disc = std::fmaf(-l.x, l.x, std::fmaf(-l.y, l.y, std::fmaf(-l.z, l.z, s.radius * s.radius)));
    if (disc < 0)
    {
        return false;
    }
    const float half_chord = std::sqrt(disc * inverse_a);
    near = h - half_chord;
    far = h + half_chord;
    return true;
}
// closed: accept t = t_max too (the closest-hit walk's tie rule below); the reference's test is open
inline bool hit_sphere(const sphere& s, const ray& r, float inverse_a, float t_min, float t_max, bool far_only,
                       float& t, bool& entering, bool closed = false)
{
    float near, far;
    if (!sphere_roots(s, r, inverse_a, near, far))
    {
        return false;
    }
    t = near;
    entering = !far_only && t > t_min;
    if (!entering)
    {
        t = far;
    }
    return t_min < t && (t < t_max || (closed && t == t_max));
}
inline bool hit_quad(const quad& s, const ray& r, float t_min, float t_max, float& t, bool& front)
{
    const float denom = dot(s.normal, r.direction);
    if (!(std::fabs(denom) >= 1e-8f))
    {
        return false;
    }
    const f3 otq = sub(s.q, r.origin);
    t = dot(s.normal, otq) / denom;
    if (!(t_min <= t && t <= t_max))
    {
        return false;
    }
    const f3 p{std::fmaf(t, r.direction.x, -otq.x), std::fmaf(t, r.direction.y, -otq.y),
               std::fmaf(t, r.direction.z, -otq.z)};
    const float alpha = dot(p, s.alpha_axis);
    const float beta = dot(p, s.beta_axis);
    if (!(0.0f <= alpha && alpha <= 1.0f && 0.0f <= beta && beta <= 1.0f))
    {
        return false;
    }
    front = denom < 0.0f;
    return true;
}

struct hit
{
    int index = -1;
    float t = INFINITY;
    bool entering = false;
};

// Ties: a primitive hit at exactly the closest distance so far replaces the closest one when its index in the scene is
// lower. The reference keeps whichever its walk meets first (a sphere) or last (a quad), which depends on the walk's
// order; this rule makes the closest hit a function of the ray and the scene alone, the same for any tree and any
// walk, and equal to the brute-force search's. Away from exact ties the two rules agree.
inline bool better(const tree& tr, int id, float t, const hit& closest)
{
    return t < closest.t || (t == closest.t && closest.index >= 0 &&
                             tr.ranks[static_cast<std::size_t>(id)] < tr.ranks[static_cast<std::size_t>(closest.index)]);
}

// Which primitive a tie goes to (see the header)
enum class tie_rule
{
    scene_order, // the lower index in the scene; every test closed
    walk_order,  // the reference's: spheres open (the first met stays), quads closed (the last met wins)
};

struct counters
{
    std::uint64_t node_visits = 0; // inner nodes whose two children's boxes were tested
    std::uint64_t sphere_tests = 0, moving_tests = 0, quad_tests = 0;
};

inline hit closest_hit(const tree& tr, const ray& r, int from, bool inward, counters* count,
                       tie_rule rule = tie_rule::scene_order)
{
    const bool by_walk = rule == tie_rule::walk_order;
    const float t_min = 0.001f;
    hit closest;
    if (tr.nodes.empty())
    {
        return closest;
    }
    const float inverse_a = 1.0f / dot(r.direction, r.direction);
    const f3 inv{safe_inverse(r.direction.x), safe_inverse(r.direction.y), safe_inverse(r.direction.z)};
    const f3 os{r.origin.x * inv.x, r.origin.y * inv.y, r.origin.z * inv.z};
    struct entry
    {
        int node;
        float t;
    };
    entry stack[32];
    int size = 0;
    int n = 0;
    while (n >= 0)
    {
        const node& nd = tr.nodes[n];
        if (nd.count > 0)
        {
            const int kind = nd.count >> 8;
            const int end = nd.index + (nd.count & 255);
            for (int i = nd.index; i < end; ++i)
            {
                if (kind == 2)
                {
                    const int id = tr.quad_base + i;
                    if (id == from)
                    {
                        continue;
                    }
                    if (count)
                    {
                        ++count->quad_tests;
                    }
                    float t;
                    bool front;
                    if (hit_quad(tr.quads[i], r, t_min, closest.t, t, front) && (by_walk || better(tr, id, t, closest)))
                    {
                        closest = {id, t, front};
                    }
                }
                else
                {
                    const bool moving = kind == 1;
                    const int id = moving ? tr.moving_base + i : i;
                    const bool own = id == from;
                    if (own && !inward)
                    {
                        continue;
                    }
                    if (count)
                    {
                        ++(moving ? count->moving_tests : count->sphere_tests);
                    }
                    const sphere s = moving ? tr.moving[i].at(r.time) : tr.spheres[i];
                    float t;
                    bool entering;
                    if (hit_sphere(s, r, inverse_a, t_min, closest.t, own, t, entering, !by_walk) &&
                        (by_walk || better(tr, id, t, closest)))
                    {
                        closest = {id, t, entering};
                    }
                }
            }
            n = -1;
        }
        else
        {
            if (count)
            {
                ++count->node_visits;
            }
            const int first = n + 1;
            const int second = nd.index;
            const float t_first = box_entry(tr.nodes[first], inv, os, t_min, closest.t);
            const float t_second = box_entry(tr.nodes[second], inv, os, t_min, closest.t);
            const bool second_nearer = t_second < t_first;
            const float t_far = std::fmax(t_first, t_second);
            if (t_far < INFINITY)
            {
                stack[size++] = {second_nearer ? first : second, t_far};
            }
            n = std::fmin(t_first, t_second) < INFINITY ? (second_nearer ? second : first) : -1;
        }
        while (n < 0 && size > 0)
        {
            const entry e = stack[--size];
            // scene order: a box entered at exactly the closest distance may hold a tie; the reference's walk pops
            // only boxes entered before it
            if (by_walk ? e.t < closest.t : e.t <= closest.t)
            {
                n = e.node;
            }
        }
    }
    return closest;
}

thread_local std::string last_error;

struct scene_tree
{
    oracle::built_scene scene;
    tree t;
};
inline scene_tree make(const char* name, int width, int spp, int earth_width)
{
    oracle::scene_options opts;
    opts.earth_width = earth_width > 0 ? earth_width : 2048;
    scene_tree st{oracle::build_flat(name, width, spp, -1, 0, opts), {}};
    const auto& f = st.scene.flat;
    st.t = builder(f.spheres, f.quads, leaf_size(f.spheres.size() + f.quads.size())).build();
    return st;
}

// --- shadow rays by brute force (the reference's any-hit check, tests/lamp_tests.cu:1618-1900) ---------------------
// The question of src/gpu/shadow.cuh:26-120: anything but `skip` in (0.001, t_max)? Each primitive is tested over
// (0.001, inf) with closest_hit's rules for the primitive the ray leaves; its distance is then compared with t_max,
// open for a sphere and closed for a quad: neither test's choice of root or distance depends on t_max, so this is the
// same answer as a test over (0.001, t_max) itself.
struct shadow_answer
{
    bool blocked = false;
    bool band = false; // a hit within 1e-4 t_max of t_max: the walk's box tests may round either way there
    bool near = false; // a decision within a few float roundings of going the other way (near_decision)
};

inline float norm(const f3& v) { return std::sqrt(dot(v, v)); }

// Whether primitive g's test for this ray sits within rounding of another outcome: a line that grazes a sphere, a
// root or a quad distance at t_min, a quad's plane nearly parallel to the ray, a hit on a quad's edge. The zones are a
// few float steps (2^-24 ~ 6e-8) of the magnitudes each test subtracts, with a margin: 1e-6 relative.
inline bool near_decision(const sphere* s, const quad* q, const ray& r, float inverse_a, bool own, float t_max)
{
    constexpr float rel = 1e-6f;
    constexpr float t_min = 0.001f;
    if (q)
    {
        const float denom = dot(q->normal, r.direction);
        if (std::fabs(denom) < 1e-6f)
        {
            return true;
        }
        const f3 otq = sub(q->q, r.origin);
        const float t = dot(q->normal, otq) / denom;
        const float size = norm(otq) + std::fabs(t);
        if (std::fabs(t - t_min) <= rel * (1.0f + size) / std::fabs(denom))
        {
            return true;
        }
        if (!(t > 0.5f * t_min && t <= 1.0002f * t_max))
        {
            return false;
        }
        const f3 p{std::fmaf(t, r.direction.x, -otq.x), std::fmaf(t, r.direction.y, -otq.y),
                   std::fmaf(t, r.direction.z, -otq.z)};
        const float alpha = dot(p, q->alpha_axis);
        const float beta = dot(p, q->beta_axis);
        const float ea = rel * (1.0f + size) * norm(q->alpha_axis);
        const float eb = rel * (1.0f + size) * norm(q->beta_axis);
        const bool inside = alpha >= -ea && alpha <= 1.0f + ea && beta >= -eb && beta <= 1.0f + eb;
        const bool edge = std::fabs(alpha) <= ea || std::fabs(alpha - 1.0f) <= ea || std::fabs(beta) <= eb ||
                          std::fabs(beta - 1.0f) <= eb;
        return inside && edge;
    }
    const f3 oc = sub(s->center, r.origin);
    const float h = dot(r.direction, oc) * inverse_a;
    if (!(h > -s->radius - 1.0f && h < t_max + s->radius + 1.0f))
    {
        return false; // the sphere lies well before the origin or beyond t_max
    }
    const f3 l{std::fmaf(-h, r.direction.x, oc.x), std::fmaf(-h, r.direction.y, oc.y),
               std::fmaf(-h, r.direction.z, oc.z)};
    const float r2 = s->radius * s->radius;
    const float disc = r2 - dot(l, l);
    const float size = norm(oc) + std::fabs(h);
    if (std::fabs(disc) <= rel * (1.0f + size) * (1.0f + s->radius) + 10.0f * rel * r2)
    {
        return true; // grazing
    }
    if (disc < 0.0f)
    {
        return false;
    }
    const float half = std::sqrt(disc * inverse_a);
    const float tol = rel * (1.0f + size);
    return (!own && std::fabs(h - half - t_min) <= tol) || std::fabs(h + half - t_min) <= tol;
}

inline shadow_answer shadow_by_brute_force(const tree& tr, const ray& r, float t_max, int from, bool inward, int skip)
{
    shadow_answer a;
    const float inverse_a = 1.0f / dot(r.direction, r.direction);
    const auto blocks = [&](float t, bool closed) {
        a.blocked = a.blocked || t < t_max || (closed && t == t_max);
        a.band = a.band || (std::isfinite(t_max) && std::fabs(t - t_max) <= 1e-4f * t_max);
    };
    for (std::size_t i = 0; i < tr.spheres.size() + tr.moving.size(); ++i)
    {
        const int g = static_cast<int>(i);
        const bool own = g == from;
        if (g == skip || (own && !inward))
        {
            continue;
        }
        const sphere s = g < tr.moving_base ? tr.spheres[i] : tr.moving[i - tr.moving_base].at(r.time);
        float t;
        bool e;
        if (hit_sphere(s, r, inverse_a, 0.001f, INFINITY, own, t, e))
        {
            blocks(t, false);
        }
        a.near = a.near || near_decision(&s, nullptr, r, inverse_a, own, t_max);
    }
    for (std::size_t i = 0; i < tr.quads.size(); ++i)
    {
        const int g = tr.quad_base + static_cast<int>(i);
        if (g == skip || g == from)
        {
            continue;
        }
        float t;
        bool f;
        if (hit_quad(tr.quads[i], r, 0.001f, INFINITY, t, f))
        {
            blocks(t, true);
        }
        a.near = a.near || near_decision(nullptr, &tr.quads[i], r, inverse_a, false, t_max);
    }
    return a;
}
}

using namespace twin;

extern "C" {

const char* twin_last_error() { return last_error.c_str(); }

void twin_pcg32(std::uint64_t initstate, std::uint64_t initseq, std::int64_t n, std::uint32_t* out)
{
    pcg32 g{initstate, initseq};
    for (std::int64_t i = 0; i < n; ++i)
    {
        out[i] = g.next();
    }
}

std::uint64_t twin_mix64(std::uint64_t z) { return mix64(z); }

// The first n random floats of sample_generator(seed, pixel, pass)
void twin_sample_floats(std::uint64_t seed, std::uint64_t pixel, std::uint64_t pass, std::int64_t n, float* out)
{
    pcg32 g = sample_generator(seed, pixel, pass);
    for (std::int64_t i = 0; i < n; ++i)
    {
        out[i] = random_float(g);
    }
}

// The twin's BVH of a built-in scene: per node 6 floats (box_min, box_max) and 2 ints (index, count), and the global
// ids of the scene's spheres and quads. Call with NULL buffers to learn the counts (node_count, and sphere/quad counts
// through counts[0], counts[1]). Returns 0 or -1.
int twin_bvh(const char* name, int width, int spp, int earth_width, float* boxes, std::int32_t* ints,
             std::int64_t* node_count, std::int32_t* sphere_ids, std::int32_t* quad_ids, std::int64_t* counts,
             std::int32_t* depth)
{
    try
    {
        const scene_tree st = make(name, width, spp, earth_width);
        *node_count = static_cast<std::int64_t>(st.t.nodes.size());
        counts[0] = static_cast<std::int64_t>(st.t.sphere_ids.size());
        counts[1] = static_cast<std::int64_t>(st.t.quad_ids.size());
        *depth = st.t.depth;
        if (boxes)
        {
            for (std::size_t i = 0; i < st.t.nodes.size(); ++i)
            {
                const node& n = st.t.nodes[i];
                const float b[6] = {n.box_min.x, n.box_min.y, n.box_min.z, n.box_max.x, n.box_max.y, n.box_max.z};
                std::memcpy(boxes + 6 * i, b, sizeof b);
                ints[2 * i] = n.index;
                ints[2 * i + 1] = n.count;
            }
            std::copy(st.t.sphere_ids.begin(), st.t.sphere_ids.end(), sphere_ids);
            std::copy(st.t.quad_ids.begin(), st.t.quad_ids.end(), quad_ids);
        }
        return 0;
    }
    catch (const std::exception& e)
    {
        last_error = e.what();
        return -1;
    }
}

// closest_hit for n rays in a built-in scene: origins, directions (3 floats each), times, from, inward. Writes the
// global id, t and the entering flag, and (if stats is not NULL) the traversal counts: inner nodes visited, then
// sphere, moving-sphere and quad tests.
int twin_closest_hit(const char* name, int width, int spp, int earth_width, const float* origins,
                     const float* directions, const float* times, const std::int32_t* from, const std::uint8_t* inward,
                     std::int64_t n, std::int32_t* index, float* t, std::uint8_t* entering, std::uint64_t* stats)
{
    try
    {
        const scene_tree st = make(name, width, spp, earth_width);
        counters c;
        for (std::int64_t i = 0; i < n; ++i)
        {
            const ray r{{origins[3 * i], origins[3 * i + 1], origins[3 * i + 2]},
                        {directions[3 * i], directions[3 * i + 1], directions[3 * i + 2]},
                        times[i]};
            const hit h = closest_hit(st.t, r, from[i], inward[i] != 0, stats ? &c : nullptr);
            index[i] = h.index;
            t[i] = h.t;
            entering[i] = h.entering ? 1 : 0;
        }
        if (stats)
        {
            stats[0] = c.node_visits;
            stats[1] = c.sphere_tests;
            stats[2] = c.moving_tests;
            stats[3] = c.quad_tests;
        }
        return 0;
    }
    catch (const std::exception& e)
    {
        last_error = e.what();
        return -1;
    }
}

// twin_closest_hit under the reference's tie rule (tie_rule::walk_order)
int twin_closest_hit_walk_order(const char* name, int width, int spp, int earth_width, const float* origins,
                                const float* directions, const float* times, const std::int32_t* from,
                                const std::uint8_t* inward, std::int64_t n, std::int32_t* index, float* t,
                                std::uint8_t* entering)
{
    try
    {
        const scene_tree st = make(name, width, spp, earth_width);
        for (std::int64_t i = 0; i < n; ++i)
        {
            const ray r{{origins[3 * i], origins[3 * i + 1], origins[3 * i + 2]},
                        {directions[3 * i], directions[3 * i + 1], directions[3 * i + 2]},
                        times[i]};
            const hit h = closest_hit(st.t, r, from[i], inward[i] != 0, nullptr, tie_rule::walk_order);
            index[i] = h.index;
            t[i] = h.t;
            entering[i] = h.entering ? 1 : 0;
        }
        return 0;
    }
    catch (const std::exception& e)
    {
        last_error = e.what();
        return -1;
    }
}

// closest_hit in a scene given as plain primitives (spheres standing still: center x, y, z, radius; quads: q, u, v,
// 9 doubles each), its tree built with the reference's leaf sizes; rule 0 scene order, 1 the reference's walk order.
// Ids are the tree's global ids; sphere_ids and quad_ids (may be null) receive each input primitive's id.
int twin_custom_closest_hit(const double* spheres, std::int64_t sphere_count, const double* quads,
                            std::int64_t quad_count, int rule, const float* origins, const float* directions,
                            std::int64_t n, std::int32_t* index, float* t, std::uint8_t* entering,
                            std::int32_t* sphere_ids, std::int32_t* quad_ids)
{
    try
    {
        std::vector<oracle::flat_sphere> fs;
        for (std::int64_t k = 0; k < sphere_count; ++k)
        {
            const double* v = spheres + 4 * k;
            fs.push_back({{v[0], v[1], v[2]}, v[3], 0, {0, 0, 0}});
        }
        std::vector<oracle::flat_quad> fq;
        for (std::int64_t k = 0; k < quad_count; ++k)
        {
            const double* v = quads + 9 * k;
            fq.push_back({{v[0], v[1], v[2]}, {v[3], v[4], v[5]}, {v[6], v[7], v[8]}, 0});
        }
        const tree tr = builder(fs, fq, leaf_size(fs.size() + fq.size())).build();
        for (std::int64_t i = 0; i < n; ++i)
        {
            const ray r{{origins[3 * i], origins[3 * i + 1], origins[3 * i + 2]},
                        {directions[3 * i], directions[3 * i + 1], directions[3 * i + 2]},
                        0.0f};
            const hit h = closest_hit(tr, r, -1, false, nullptr, rule == 1 ? tie_rule::walk_order : tie_rule::scene_order);
            index[i] = h.index;
            t[i] = h.t;
            entering[i] = h.entering ? 1 : 0;
        }
        if (sphere_ids)
        {
            std::copy(tr.sphere_ids.begin(), tr.sphere_ids.end(), sphere_ids);
        }
        if (quad_ids)
        {
            std::copy(tr.quad_ids.begin(), tr.quad_ids.end(), quad_ids);
        }
        return 0;
    }
    catch (const std::exception& e)
    {
        last_error = e.what();
        return -1;
    }
}

// The first n random doubles of sample_generator(seed, pixel, pass): 53 bits from two outputs, the high one first
// (src/gpu/random.cuh:76-90)
void twin_sample_doubles(std::uint64_t seed, std::uint64_t pixel, std::uint64_t pass, std::int64_t n, double* out)
{
    pcg32 g = sample_generator(seed, pixel, pass);
    for (std::int64_t i = 0; i < n; ++i)
    {
        const std::uint64_t high = g.next();
        const std::uint64_t low = g.next();
        out[i] = static_cast<double>((high << 21u) | (low >> 11u)) * 0x1p-53;
    }
}

// The first n raw 32-bit outputs of sample_generator(seed, pixel, pass)
void twin_sample_words(std::uint64_t seed, std::uint64_t pixel, std::uint64_t pass, std::int64_t n, std::uint32_t* out)
{
    pcg32 g = sample_generator(seed, pixel, pass);
    for (std::int64_t i = 0; i < n; ++i)
    {
        out[i] = g.next();
    }
}

// Brute force over every primitive of the twin's tree: the closest hit as a loop over all of them (the reference's
// T7 check, tests/lamp_tests.cu:1618-1660), to pin the traversal itself. Same arguments as twin_closest_hit, and
// `tied` (may be null): per ray, how many primitives are hit at exactly the closest distance.
int twin_brute_force(const char* name, int width, int spp, int earth_width, const float* origins,
                     const float* directions, const float* times, const std::int32_t* from, const std::uint8_t* inward,
                     std::int64_t n, std::int32_t* index, float* t, std::int32_t* tied)
{
    try
    {
        const scene_tree st = make(name, width, spp, earth_width);
        const tree& tr = st.t;
        for (std::int64_t k = 0; k < n; ++k)
        {
            const ray r{{origins[3 * k], origins[3 * k + 1], origins[3 * k + 2]},
                        {directions[3 * k], directions[3 * k + 1], directions[3 * k + 2]},
                        times[k]};
            const float inverse_a = 1.0f / dot(r.direction, r.direction);
            hit best;
            for (std::size_t i = 0; i < tr.spheres.size() + tr.moving.size(); ++i)
            {
                const int id = static_cast<int>(i);
                const bool own = id == from[k];
                if (own && !inward[k])
                {
                    continue;
                }
                const sphere s = id < tr.moving_base ? tr.spheres[i] : tr.moving[i - tr.moving_base].at(r.time);
                float tt;
                bool e;
                if (hit_sphere(s, r, inverse_a, 0.001f, best.t, own, tt, e, true) && better(tr, id, tt, best))
                {
                    best = {id, tt, e};
                }
            }
            for (std::size_t i = 0; i < tr.quads.size(); ++i)
            {
                const int id = tr.quad_base + static_cast<int>(i);
                if (id == from[k])
                {
                    continue;
                }
                float tt;
                bool f;
                if (hit_quad(tr.quads[i], r, 0.001f, best.t, tt, f) && better(tr, id, tt, best))
                {
                    best = {id, tt, f};
                }
            }
            index[k] = best.index;
            t[k] = best.t;
            if (tied && best.index >= 0)
            {
                // a second sweep with the closest distance known: every primitive hit exactly there
                int count = 0;
                for (std::size_t i = 0; i < tr.spheres.size() + tr.moving.size(); ++i)
                {
                    const int id = static_cast<int>(i);
                    const bool own = id == from[k];
                    const sphere s = id < tr.moving_base ? tr.spheres[i] : tr.moving[i - tr.moving_base].at(r.time);
                    float tt;
                    bool e;
                    count += (!own || inward[k]) && hit_sphere(s, r, inverse_a, 0.001f, best.t, own, tt, e, true) &&
                             tt == best.t;
                }
                for (std::size_t i = 0; i < tr.quads.size(); ++i)
                {
                    float tt;
                    bool f;
                    count += tr.quad_base + static_cast<int>(i) != from[k] &&
                             hit_quad(tr.quads[i], r, 0.001f, best.t, tt, f) && tt == best.t;
                }
                tied[k] = count;
            }
            else if (tied)
            {
                tied[k] = 0;
            }
        }
        return 0;
    }
    catch (const std::exception& e)
    {
        last_error = e.what();
        return -1;
    }
}

// Shadow rays in a scene of plain primitives by brute force: spheres as center x, y, z, radius and displacement
// x, y, z (7 doubles each; a zero displacement stands still), quads as q, u, v (9 doubles each). Per ray: `from` and
// `skip` as an input primitive's index (spheres, then quads) or -1, the ray's time, inward and t_max; out: blocked,
// band and near (shadow_answer), 1 or 0 each. sphere_ids and quad_ids (may be null) receive each input primitive's
// global id in the tree built for the scene with the reference's leaf sizes.
int twin_custom_occluded(const double* spheres, std::int64_t sphere_count, const double* quads,
                         std::int64_t quad_count, const float* origins, const float* directions, const float* times,
                         const std::int32_t* from, const std::uint8_t* inward, const float* t_max,
                         const std::int32_t* skip, std::int64_t n, std::uint8_t* blocked, std::uint8_t* band,
                         std::uint8_t* near, std::int32_t* sphere_ids, std::int32_t* quad_ids)
{
    try
    {
        std::vector<oracle::flat_sphere> fs;
        for (std::int64_t k = 0; k < sphere_count; ++k)
        {
            const double* v = spheres + 7 * k;
            fs.push_back({{v[0], v[1], v[2]}, v[3], 0, {v[4], v[5], v[6]}});
        }
        std::vector<oracle::flat_quad> fq;
        for (std::int64_t k = 0; k < quad_count; ++k)
        {
            const double* v = quads + 9 * k;
            fq.push_back({{v[0], v[1], v[2]}, {v[3], v[4], v[5]}, {v[6], v[7], v[8]}, 0});
        }
        const tree tr = builder(fs, fq, leaf_size(fs.size() + fq.size())).build();
        const auto id_of = [&](std::int32_t input) {
            if (input < 0)
            {
                return -1;
            }
            if (input < sphere_count)
            {
                return tr.sphere_ids[static_cast<std::size_t>(input)];
            }
            if (input < sphere_count + quad_count)
            {
                return tr.quad_ids[static_cast<std::size_t>(input - sphere_count)];
            }
            throw std::out_of_range("a primitive index beyond the scene");
        };
        for (std::int64_t k = 0; k < n; ++k)
        {
            const ray r{{origins[3 * k], origins[3 * k + 1], origins[3 * k + 2]},
                        {directions[3 * k], directions[3 * k + 1], directions[3 * k + 2]},
                        times[k]};
            const shadow_answer a = shadow_by_brute_force(tr, r, t_max[k], id_of(from[k]), inward[k] != 0,
                                                          id_of(skip[k]));
            blocked[k] = a.blocked ? 1 : 0;
            band[k] = a.band ? 1 : 0;
            near[k] = a.near ? 1 : 0;
        }
        if (sphere_ids)
        {
            std::copy(tr.sphere_ids.begin(), tr.sphere_ids.end(), sphere_ids);
        }
        if (quad_ids)
        {
            std::copy(tr.quad_ids.begin(), tr.quad_ids.end(), quad_ids);
        }
        return 0;
    }
    catch (const std::exception& e)
    {
        last_error = e.what();
        return -1;
    }
}
}
