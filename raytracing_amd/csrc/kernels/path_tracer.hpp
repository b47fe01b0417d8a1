// The path tracer's device functions that closest-hit parity rests on, shared by every kernel (the megakernel, the
// wavefront stages, the parity probes): vectors and rays, the per-sample random numbers, the intersectors and the BVH
// walk. Each follows the reference GPU renderer's semantics (cited per function); the kernels differ only in how they
// schedule this work, so they compute the same numbers. The shading is in books.hpp and quality.hpp.
//
// Arithmetic: float (or double for the --fp64 path) as written, compiled with -ffp-contract=off. Multiply-adds are
// fused only where the source writes fma(), in a fixed pattern the CPU twin (oracle/gpu_twin.cpp) restates:
// dot products, the slab test, the sphere discriminant, a moving sphere's center, a quad's hit point and a ray's at().
#pragma once

#include <cstdint>

#include <hip/hip_runtime.h>

#include "../../../include/raytracing_amd.h"
#include "../device_types.hpp"

namespace rt::dev
{
// --- math in both precisions ------------------------------------------------------------------------------------------
__device__ inline float m_sqrt(float x) { return sqrtf(x); }
__device__ inline double m_sqrt(double x) { return sqrt(x); }
__device__ inline float m_rsqrt(float x) { return rsqrtf(x); }
__device__ inline double m_rsqrt(double x) { return rsqrt(x); }
__device__ inline float m_fabs(float x) { return fabsf(x); }
__device__ inline double m_fabs(double x) { return fabs(x); }
__device__ inline float m_min(float x, float y) { return fminf(x, y); }
__device__ inline double m_min(double x, double y) { return fmin(x, y); }
__device__ inline float m_max(float x, float y) { return fmaxf(x, y); }
__device__ inline double m_max(double x, double y) { return fmax(x, y); }
__device__ inline float m_fma(float x, float y, float z) { return fmaf(x, y, z); }
__device__ inline double m_fma(double x, double y, double z) { return fma(x, y, z); }
__device__ inline float m_copysign(float x, float y) { return copysignf(x, y); }
__device__ inline double m_copysign(double x, double y) { return copysign(x, y); }
__device__ inline float m_floor(float x) { return floorf(x); }
__device__ inline double m_floor(double x) { return floor(x); }
__device__ inline float m_atan2(float y, float x) { return atan2f(y, x); }
__device__ inline double m_atan2(double y, double x) { return atan2(y, x); }
__device__ inline float m_acos(float x) { return acosf(x); }
__device__ inline double m_acos(double x) { return acos(x); }
__device__ inline float m_sin(float x) { return sinf(x); }
__device__ inline double m_sin(double x) { return sin(x); }
__device__ inline float m_log(float x) { return logf(x); }
__device__ inline double m_log(double x) { return log(x); }
__device__ inline float m_exp(float x) { return expf(x); }
__device__ inline double m_exp(double x) { return exp(x); }
__device__ inline void m_sincospi(float x, float* s, float* c) { sincospif(x, s, c); }
__device__ inline void m_sincospi(double x, double* s, double* c) { sincospi(x, s, c); }

template <typename Real>
inline constexpr Real inf = static_cast<Real>(__builtin_huge_val());
template <typename Real>
inline constexpr Real pi_v = static_cast<Real>(3.14159265358979323846);

template <typename Real>
struct vec3
{
    Real x, y, z;
};
template <typename Real>
__device__ inline vec3<Real> operator+(vec3<Real> a, vec3<Real> b)
{
    return {a.x + b.x, a.y + b.y, a.z + b.z};
}
template <typename Real>
__device__ inline vec3<Real> operator-(vec3<Real> a, vec3<Real> b)
{
    return {a.x - b.x, a.y - b.y, a.z - b.z};
}
template <typename Real>
__device__ inline vec3<Real> operator-(vec3<Real> a)
{
    return {-a.x, -a.y, -a.z};
}
template <typename Real>
__device__ inline vec3<Real> operator*(vec3<Real> a, vec3<Real> b)
{
    return {a.x * b.x, a.y * b.y, a.z * b.z};
}
template <typename Real>
__device__ inline vec3<Real> operator*(Real t, vec3<Real> v)
{
    return {t * v.x, t * v.y, t * v.z};
}
// dot(u, v) = fma(u.x, v.x, fma(u.y, v.y, u.z * v.z))
template <typename Real>
__device__ inline Real dot(vec3<Real> u, vec3<Real> v)
{
    return m_fma(u.x, v.x, m_fma(u.y, v.y, u.z * v.z));
}
template <typename Real>
__device__ inline Real length_squared(vec3<Real> v)
{
    return dot(v, v);
}
template <typename Real>
__device__ inline vec3<Real> unit_vector(vec3<Real> v)
{
    return m_rsqrt(length_squared(v)) * v;
}
template <typename Real>
__device__ inline vec3<Real> load3(const Real* p)
{
    return {p[0], p[1], p[2]};
}
template <typename Real>
struct ray
{
    vec3<Real> origin;
    vec3<Real> direction;
    Real time;
    // origin + t·direction, one rounding per coordinate
    __device__ vec3<Real> at(Real t) const
    {
        return {m_fma(t, direction.x, origin.x), m_fma(t, direction.y, origin.y), m_fma(t, direction.z, origin.z)};
    }
    // the product rounded, then the sum: the CPU's point (src/gpu/vec3.cuh:201-211)
    __device__ vec3<Real> at_unfused(Real t) const
    {
        return {origin.x + t * direction.x, origin.y + t * direction.y, origin.z + t * direction.z};
    }
};

// --- random numbers (src/gpu/random.cuh:25-139) ------------------------------------------------------------------------
// PCG32 (O'Neill's pcg32_srandom_r / pcg32_random_r), one generator per sample, started from SplitMix64 hashes of
// (seed, pixel, pass): a sample's numbers never depend on the order the GPU runs it in.
struct pcg32
{
    uint64_t state;
    uint64_t inc;
    __device__ uint32_t next()
    {
        const uint64_t old = state;
        state = old * 6364136223846793005ull + inc;
        const auto xorshifted = static_cast<uint32_t>(((old >> 18u) ^ old) >> 27u);
        const auto rot = static_cast<uint32_t>(old >> 59u);
        return (xorshifted >> rot) | (xorshifted << ((32u - rot) & 31u));
    }
};
__device__ inline uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30u)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27u)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31u);
}
__device__ inline pcg32 sample_generator(uint64_t seed, uint64_t pixel, uint64_t pass)
{
    const uint64_t initstate = mix64(mix64(mix64(seed) ^ pixel) ^ pass);
    const uint64_t initseq = mix64(mix64(mix64(seed ^ 0x9e3779b97f4a7c15ull) ^ pass) ^ pixel);
    pcg32 g{0, (initseq << 1u) | 1u};
    g.next();
    g.state += initstate;
    g.next();
    return g;
}
template <typename Real>
__device__ inline Real random_real(pcg32& g);
template <>
__device__ inline float random_real<float>(pcg32& g)
{
    return static_cast<float>(g.next() >> 8u) * 0x1p-24f;
}
template <>
__device__ inline double random_real<double>(pcg32& g)
{
    const uint64_t high = g.next();
    const uint64_t low = g.next();
    return static_cast<double>((high << 21u) | (low >> 11u)) * 0x1p-53;
}
template <typename Real>
__device__ inline vec3<Real> random_unit_vector(pcg32& g)
{
    const Real u = random_real<Real>(g);
    const Real phi_over_pi = Real(2) * random_real<Real>(g);
    Real s, c;
    m_sincospi(phi_over_pi, &s, &c);
    const Real r = Real(2) * m_sqrt(u * (Real(1) - u));
    return {r * c, r * s, Real(1) - Real(2) * u};
}
// --- the BVH walk (src/gpu/bvh.cuh:483-706) ------------------------------------------------------------------------------
template <typename Real>
__device__ inline Real safe_inverse(Real x)
{
    const Real tiny = Real(0x1p-80);
    return Real(1) / (m_fabs(x) > tiny ? x : m_copysign(tiny, x));
}

// The slab test with one multiply-add per slab: where the ray enters the box, or infinity
template <typename Real>
__device__ inline Real box_entry(const Real* lo, const Real* hi, vec3<Real> inv, vec3<Real> os, Real t_min, Real t_max)
{
    const Real x0 = m_fma(lo[0], inv.x, -os.x);
    const Real x1 = m_fma(hi[0], inv.x, -os.x);
    const Real y0 = m_fma(lo[1], inv.y, -os.y);
    const Real y1 = m_fma(hi[1], inv.y, -os.y);
    const Real z0 = m_fma(lo[2], inv.z, -os.z);
    const Real z1 = m_fma(hi[2], inv.z, -os.z);
    const Real enter = m_max(m_max(m_min(x0, x1), m_min(y0, y1)), m_max(m_min(z0, z1), t_min));
    const Real exit = m_min(m_min(m_max(x0, x1), m_max(y0, y1)), m_min(m_max(z0, z1), t_max));
    return enter <= exit ? enter : inf<Real>;
}

template <typename Real>
struct sphere
{
    vec3<Real> center;
    Real radius;
};

// Haines et al.'s discriminant radius² - |l|² with three nested fmas (src/gpu/sphere.cuh:70-87)
template <typename Real>
__device__ inline bool sphere_roots(const sphere<Real>& s, const ray<Real>& r, Real inverse_a, Real& near, Real& far)
{
    const vec3<Real> oc = s.center - r.origin;
    const Real h = dot(r.direction, oc) * inverse_a;
    const vec3<Real> l{m_fma(-h, r.direction.x, oc.x), m_fma(-h, r.direction.y, oc.y), m_fma(-h, r.direction.z, oc.z)};
    const Real disc = m_fma(-l.x, l.x, m_fma(-l.y, l.y, m_fma(-l.z, l.z, s.radius * s.radius)));
    // (the roots computed on every lane, a miss's from 0: no branch, the answer taken by the caller)
    const Real half_chord = m_sqrt(m_max(disc, Real(0)) * inverse_a);
    near = h - half_chord;
    far = h + half_chord;
    return disc >= Real(0);
}

// The nearest root in (t_min, t_max), the far one only for a ray leaving this sphere inward (src/gpu/sphere.cuh:99-117)
template <typename Real>
// closed: accept t = t_max too (the closest-hit walk's tie rule, `better`); the reference's test is open, as the shadow
// rays' stays
__device__ inline bool hit_sphere(const sphere<Real>& s, const ray<Real>& r, Real inverse_a, Real t_min, Real t_max,
                                  bool far_only, Real& t, bool& entering, bool closed = false)
{
    Real near, far;
    const bool real = sphere_roots(s, r, inverse_a, near, far);
    const bool from_outside = !far_only && near > t_min;
    const Real at = from_outside ? near : far;
    const bool met = real & (t_min < at) & ((at < t_max) | (closed & (at == t_max)));
    if (met)
    {
        t = at;
        entering = from_outside;
    }
    return met;
}

template <typename Real>
__device__ inline sphere<Real> load_sphere(const sphere_rec<Real>* spheres, int i)
{
    const sphere_rec<Real> s = spheres[i];
    return {{s.center[0], s.center[1], s.center[2]}, s.radius};
}
// A moving sphere where it is at the ray's time: center + time·displacement, one rounding
template <typename Real>
__device__ inline sphere<Real> moving_at(const moving_sphere_rec<Real>& m, Real time)
{
    return {{m_fma(time, m.displacement[0], m.center[0]), m_fma(time, m.displacement[1], m.center[1]),
             m_fma(time, m.displacement[2], m.center[2])},
            m.radius};
}

// The ray against a quad, closed [t_min, t_max] (src/gpu/quad.cuh:56-84). Every test is evaluated and the answer
// taken at the end (as in hit_triangle): the lanes of a wave meet different primitives, so early exits saved little
// arithmetic and cost a mask save and restore each. t and front_face are written on a hit only.
template <typename Real>
__device__ inline bool hit_quad(const quad_rec<Real>& s, const ray<Real>& r, Real t_min, Real t_max, Real& t,
                                bool& front_face)
{
    const vec3<Real> normal = load3(s.normal);
    const Real denominator = dot(normal, r.direction);
    const vec3<Real> otq = load3(s.q) - r.origin;
    const Real at = dot(normal, otq) / denominator;
    const vec3<Real> p{m_fma(at, r.direction.x, -otq.x), m_fma(at, r.direction.y, -otq.y),
                       m_fma(at, r.direction.z, -otq.z)};
    const Real alpha = dot(p, load3(s.alpha_axis));
    const Real beta = dot(p, load3(s.beta_axis));
    const bool met = (m_fabs(denominator) >= Real(1e-8)) & (t_min <= at) & (at <= t_max) & (Real(0) <= alpha) &
                     (alpha <= Real(1)) & (Real(0) <= beta) & (beta <= Real(1));
    if (met)
    {
        t = at;
        front_face = denominator < Real(0);
    }
    return met;
}
template <typename Real>
__device__ inline void quad_uv(const quad_rec<Real>& s, const ray<Real>& r, Real t, Real& u, Real& v)
{
    const vec3<Real> otq = load3(s.q) - r.origin;
    const vec3<Real> p{m_fma(t, r.direction.x, -otq.x), m_fma(t, r.direction.y, -otq.y), m_fma(t, r.direction.z, -otq.z)};
    u = dot(p, load3(s.alpha_axis));
    v = dot(p, load3(s.beta_axis));
}

// The ray against a triangle (Moller and Trumbore 1997), closed [t_min, t_max] as a quad's; front: the ray meets the
// side its normal e1 x e2 points to. Barycentric (u, v) of the hit for its texture coordinates.
template <typename Real>
__device__ inline bool hit_triangle(const triangle_rec<Real>& tri, const ray<Real>& r, Real t_min, Real t_max, Real& t,
                                    bool& front_face, Real* bary = nullptr)
{
    const vec3<Real> e1 = load3(tri.e1);
    const vec3<Real> e2 = load3(tri.e2);
    const vec3<Real> d = r.direction;
    const vec3<Real> pv{d.y * e2.z - d.z * e2.y, d.z * e2.x - d.x * e2.z, d.x * e2.y - d.y * e2.x};
    const Real det = dot(e1, pv);
    const Real inv = Real(1) / det;
    const vec3<Real> s = r.origin - load3(tri.v0);
    const Real u = dot(s, pv) * inv;
    const vec3<Real> qv{s.y * e1.z - s.z * e1.y, s.z * e1.x - s.x * e1.z, s.x * e1.y - s.y * e1.x};
    const Real v = dot(d, qv) * inv;
    const Real at = dot(e2, qv) * inv;
    // every test evaluated, the answer taken at the end (no early exits: see hit_quad)
    const bool met = (m_fabs(det) > Real(1e-30)) & (u >= Real(0)) & (u <= Real(1)) & (v >= Real(0)) &
                     (u + v <= Real(1)) & (t_min <= at) & (at <= t_max);
    if (met)
    {
        t = at;
        front_face = det > Real(0);
        if (bary)
        {
            bary[0] = u;
            bary[1] = v;
        }
    }
    return met;
}

// (entering as a 32-bit value: a bool carried through the walk's loops is a lane mask, merged with three scalar
// instructions at every join of the leaf loops' branches; a 32-bit value is a vector register and takes a select)
template <typename Real>
struct hit
{
    int index;
    Real t;
    uint32_t entering;
};

// What a path reads from the scene, counted per event for the algorithmic-bytes measurement (an instantiation of the
// kernels with tally<true>); tally<false> compiles to nothing
enum tally_event
{
    tally_node = 0,       // a wide node: its four children's boxes
    tally_sphere,         // a sphere record tested
    tally_moving,         // a moving sphere record tested
    tally_quad,           // a quad record tested
    tally_medium,         // a medium record tested
    tally_boundary_quad,  // a medium's boundary quad tested
    tally_hit,            // a surface hit: primitive material id, material, the hit primitive's record again
    tally_frame,          // a sphere frame
    tally_texture,        // a texture record
    tally_texel,          // an image texel
    tally_perlin,         // a Perlin gradient (and its three permutation bytes)
    tally_light,          // a light record
    tally_triangle,       // a triangle record tested
    tally_lamp,           // a lamp of the quality integrator's light set
    tally_env_texel,      // an environment map texel (radiance and density)
    tally_alias_bin,      // a bin of the environment map's alias table
    tally_instance,       // an instance record (the ray moved into its mesh)
    tally_events
};
// The phases of a path's iteration, timed by the counting instance (lap): the queue (claiming an item, starting its
// path, staging its sample, and a lane's wait while others trace), the walk (closest hits and shadow rays' any-hit
// walks), the media (free flights, transmittance), the shading (vertices, materials, scattering, light samples)
enum tally_phase
{
    phase_queue = 0,
    phase_walk,
    phase_media,
    phase_shade,
    tally_phases
};
template <bool On>
struct tally
{
    __device__ void add(tally_event, uint32_t = 1) {}
    __device__ void lap(tally_phase) {}
};
template <>
struct tally<true>
{
    uint32_t n[tally_events] = {};
    uint64_t cycles[tally_phases] = {};
    uint64_t last = __builtin_readcyclecounter();
    __device__ void add(tally_event e, uint32_t k = 1) { n[e] += k; }
    // the shader clock's cycles since the last lap, to `phase` (per lane: a lane that waits while others trace counts
    // the wait in the phase of its next lap)
    __device__ void lap(tally_phase phase)
    {
        const uint64_t now = __builtin_readcyclecounter();
        cycles[phase] += now - last;
        last = now;
    }
};

// A distance on the traversal stack in 16 bits: the upper half of a float no larger than t (t > 0, so dropping the low
// bits rounds down). The pop's test `entry < closest` can then keep an entry the exact distance would have dropped
// (a wasted visit: every box and primitive below it is still tested against the closest distance), never the reverse.
__device__ inline uint16_t stack_distance(float t) { return static_cast<uint16_t>(__float_as_uint(t) >> 16); }
__device__ inline uint16_t stack_distance(double t) { return stack_distance(__double2float_rd(t)); }
__device__ inline float stack_distance_value(uint16_t bits) { return __uint_as_float(static_cast<uint32_t>(bits) << 16); }

// The traversal stack of one thread in LDS: entries of a node or leaf reference (4 bytes) and the distance the ray
// enters it (2 bytes), interleaved across the block's threads so that a wave's 64 accesses to one level fall into
// distinct banks. 6 bytes an entry keep the four-wide walk's deeper stack within the LDS the binary walk's 8-byte
// entries took, and so the occupancy.
struct lds_stack
{
    int32_t* refs;
    uint16_t* ts;
    int stride; // threads per block
    // a level's place: a 24-bit multiply (full rate; a 32-bit one issues at a quarter of the rate), level < 2^24
    __device__ int at(int level) const
    {
        return static_cast<int>(__umul24(static_cast<unsigned>(level), static_cast<unsigned>(stride)));
    }
    template <typename Real>
    __device__ void store(int level, int32_t ref, Real t) const
    {
        refs[at(level)] = ref;
        ts[at(level)] = stack_distance(t);
    }
    __device__ int32_t ref_at(int level) const { return refs[at(level)]; }
    __device__ float distance_at(int level) const { return stack_distance_value(ts[at(level)]); }
};

// A stack whose first `near` levels are in LDS and whose deeper ones are in HBM, one column per thread of the grid
// (interleaved like the LDS part): a deep tree's worst case then no longer sets the LDS a wave holds, and so the
// occupancy; rays rarely go that deep.
struct split_stack
{
    lds_stack top;
    int near;         // levels in LDS
    int32_t* refs;    // the deeper levels, this thread's column
    uint16_t* ts;
    uint32_t stride;  // threads in the grid
    template <typename Real>
    __device__ void store(int level, int32_t ref, Real t) const
    {
        if (level < near)
        {
            top.store(level, ref, t);
        }
        else
        {
            const size_t at = static_cast<size_t>(level - near) * stride;
            refs[at] = ref;
            ts[at] = stack_distance(t);
        }
    }
    __device__ int32_t ref_at(int level) const
    {
        return level < near ? top.ref_at(level) : refs[static_cast<size_t>(level - near) * stride];
    }
    __device__ float distance_at(int level) const
    {
        return level < near ? top.distance_at(level)
                            : stack_distance_value(ts[static_cast<size_t>(level - near) * stride]);
    }
};

// LDS bytes of a block's stacks, and one thread's stack within them
__host__ __device__ constexpr size_t stack_ref_bytes(int entries, int threads)
{
    return (static_cast<size_t>(entries) * static_cast<size_t>(threads) * 4 + 15) & ~size_t(15);
}
__host__ __device__ constexpr size_t stack_bytes(int entries, int threads)
{
    return stack_ref_bytes(entries, threads) + static_cast<size_t>(entries) * static_cast<size_t>(threads) * 2;
}
__device__ inline lds_stack stack_in(unsigned char* lds, int entries, int threads, int tid)
{
    return {reinterpret_cast<int32_t*>(lds) + tid,
            reinterpret_cast<uint16_t*>(lds + stack_ref_bytes(entries, threads)) + tid, threads};
}
// This thread's stack in the block's dynamic LDS (a launch gives the block stack_bytes(entries, threads))
__device__ inline lds_stack block_stack(int entries)
{
    extern __shared__ __align__(16) unsigned char lds[];
    return stack_in(lds, entries, blockDim.x * blockDim.y, threadIdx.y * blockDim.x + threadIdx.x);
}

// Ties: a primitive hit at exactly the closest distance so far (every test accepts t = t_max in the closest-hit walk)
// replaces the closest one when its index in the scene is lower. The reference keeps whichever its walk meets first
// (a sphere) or last (a quad), which depends on the walk's order; with this rule the closest hit is a function of the
// ray and the scene alone, the same for any tree and for the four-wide walk as for the twin's binary one. The ranks
// are read only on an exact tie.
// (A first candidate at an infinite t, which an overflowing quad or triangle distance could give while nothing is hit
// yet, is no hit: the ranks are read only once a closest hit exists.)
// (An instanced mesh's triangles rank by their ids, which follow every other surface's rank: only a walk that can
// enter an instance meets them.)
template <int Kinds, typename Real>
__device__ inline int32_t rank_of(const scene_view<Real>& sc, int id)
{
    if constexpr (Kinds == walk_instances)
    {
        return id < sc.instanced_base ? sc.prim_ranks[id] : id;
    }
    return sc.prim_ranks[id];
}
template <int Kinds, typename Real>
__device__ inline bool better(const scene_view<Real>& sc, int id, Real t, const hit<Real>& closest)
{
    bool take = t < closest.t;
    if (!take && closest.index >= 0) // (one branch for the rare tie)
    {
        take = rank_of<Kinds>(sc, id) < rank_of<Kinds>(sc, closest.index);
    }
    return take;
}

// The primitives of a leaf against the ray, the closest hit so far updated (src/gpu/bvh.cuh:612-680). A triangle's id
// is its index in the triangle array plus triangle_shift (the world's triangles: triangle_base; an instance's: its
// id_base less where its mesh's triangles start).
// Shadows: the walk may also answer a shadow ray's question (any_hit, a per-lane flag): is any primitive but `skip` hit
// in (t_min, t_max)? Then the first hit found is the answer, and spheres are tested on the open interval as the
// reference's shadow rays test them (src/gpu/shadow.cuh:26-120); quads and triangles keep their closed test.
template <typename Real, int Kinds, bool Shadows = false, typename Tally>
__device__ inline void closest_in_leaf(const scene_view<Real>& sc, const ray<Real>& r, int32_t ref, int from,
                                       bool inward, Real inverse_a, Real t_min, int triangle_shift, hit<Real>& closest,
                                       Tally& tl, int skip = -1, bool any_hit = false)
{
    // (a shadow ray takes any hit; a closest-hit walk's ties go by the scene's order)
    const auto take = [&](int id, Real t) { return (Shadows && any_hit) || better<Kinds>(sc, id, t, closest); };
    const int kind = leaf_kind(ref);
    const int first = leaf_first(ref);
    const int end = first + leaf_count(ref);
    if (Kinds >= walk_all && kind == prim_triangle)
    {
        for (int i = first; i < end; ++i)
        {
            const int id = triangle_shift + i;
            if (id == from || (Shadows && id == skip))
            {
                continue; // a ray never meets the flat face it leaves
            }
            const triangle_rec<Real> tri = sc.triangles[i];
            tl.add(tally_triangle);
            Real t;
            bool front;
            if (hit_triangle(tri, r, t_min, closest.t, t, front) && take(id, t))
            {
                closest = {id, t, front};
            }
        }
    }
    else if (Kinds >= walk_books && kind == prim_quad)
    {
        for (int i = first; i < end; ++i)
        {
            const int id = sc.quad_base + i;
            if (id == from || (Shadows && id == skip))
            {
                continue;
            }
            const quad_rec<Real> face = sc.quads[i];
            tl.add(tally_quad);
            Real t;
            bool front;
            if (hit_quad(face, r, t_min, closest.t, t, front) && take(id, t))
            {
                closest = {id, t, front};
            }
        }
    }
    else
    {
        const bool moving = Kinds >= walk_books && kind == prim_moving_sphere;
        for (int i = first; i < end; ++i)
        {
            const int id = moving ? sc.moving_base + i : i;
            const bool own = id == from;
            if ((own && !inward) || (Shadows && id == skip))
            {
                continue;
            }
            const sphere<Real> s = moving ? moving_at(sc.moving_spheres[i], r.time) : load_sphere(sc.spheres, i);
            tl.add(moving ? tally_moving : tally_sphere);
            Real t;
            bool entering;
            if (hit_sphere(s, r, inverse_a, t_min, closest.t, own, t, entering, !(Shadows && any_hit)) && take(id, t))
            {
                closest = {id, t, entering};
            }
        }
    }
}

// Which plane of each axis a ray meets first in every box: the lo plane where its inverse direction is positive, the
// hi plane where it is negative (safe_inverse never returns zero). Kept per ray as byte offsets of the four children's
// near-plane vectors in a wide_node; the far plane's vector is the other one of the axis.
struct slab_order
{
    uint32_t near[3];
};
template <typename Real>
__device__ inline slab_order slab_order_of(vec3<Real> inv)
{
    const uint32_t r = sizeof(Real);
    return {{(signbit(inv.x) ? 12u : 0u) * r, (signbit(inv.y) ? 16u : 4u) * r, (signbit(inv.z) ? 20u : 8u) * r}};
}

// Where the ray enters each of a wide node's four children, or infinity (a missing child's +infinity box included).
// The near and far planes are loaded by the ray's slab order, so each slab is one multiply-add per plane and no min or
// max: a multiply-add is monotonic in its first operand, so the near plane's distance is the smaller of the two, and
// the entries are box_entry's bit for bit.
template <typename Real>
__device__ inline void child_entries(const scene_view<Real>& sc, int32_t ref, const slab_order& so, vec3<Real> inv,
                                     vec3<Real> os, Real t_min, Real t_max, Real (&t)[wide_children])
{
    typedef Real real4 __attribute__((ext_vector_type(4)));
    const char* const nodes = reinterpret_cast<const char*>(sc.nodes);
    const uint32_t at = static_cast<uint32_t>(ref) * static_cast<uint32_t>(sizeof(wide_node<Real>));
    const uint32_t r = sizeof(Real);
    const real4 nx = *reinterpret_cast<const real4*>(nodes + (at + so.near[0]));
    const real4 ny = *reinterpret_cast<const real4*>(nodes + (at + so.near[1]));
    const real4 nz = *reinterpret_cast<const real4*>(nodes + (at + so.near[2]));
    const real4 fx = *reinterpret_cast<const real4*>(nodes + (at + (12u * r - so.near[0])));
    const real4 fy = *reinterpret_cast<const real4*>(nodes + (at + (20u * r - so.near[1])));
    const real4 fz = *reinterpret_cast<const real4*>(nodes + (at + (28u * r - so.near[2])));
#pragma unroll
    for (int c = 0; c < wide_children; ++c)
    {
        const Real enter = m_max(m_max(m_fma(nx[c], inv.x, -os.x), m_fma(ny[c], inv.y, -os.y)),
                                 m_max(m_fma(nz[c], inv.z, -os.z), t_min));
        const Real exit = m_min(m_min(m_fma(fx[c], inv.x, -os.x), m_fma(fy[c], inv.y, -os.y)),
                                m_min(m_fma(fz[c], inv.z, -os.z), t_max));
        t[c] = enter <= exit ? enter : inf<Real>;
    }
}

// Where the ray enters each child of a packed node (packed_node), or infinity, and the children's references: the four
// 16-byte words of the node's 64 bytes loaded at once, each plane unpacked from its byte to its float, exactly (corner
// + q x step is a float: a multiple of the power-of-two step), then the slab test of child_entries on those planes: the
// same multiply-add, monotonic in the plane, so a packed box's lower planes give distances no larger than its float
// box's and its upper planes no smaller, and the walk enters every child it enters on float nodes, at no farther a
// distance. The ray's slab order picks the near planes (the lower ones where the inverse direction is positive). A
// missing child (no_child) is never entered.
__device__ inline void packed_entries(const scene_view<float>& sc, int32_t ref, vec3<float> inv, vec3<float> os,
                                      float t_min, float t_max, float (&t)[wide_children], int4& refs)
{
    const uint4* const node = reinterpret_cast<const uint4*>(reinterpret_cast<const char*>(sc.packed_nodes) +
                                                             static_cast<uint32_t>(ref) * uint32_t{sizeof(packed_node)});
    const uint4 w0 = node[0]; // corner x, y, z; step x
    const uint4 w1 = node[1]; // step y, z; lo x; lo y
    const uint4 w2 = node[2]; // lo z; hi x, y, z
    const uint4 w3 = node[3]; // the references
    refs = make_int4(static_cast<int>(w3.x), static_cast<int>(w3.y), static_cast<int>(w3.z), static_cast<int>(w3.w));
    const float cx = __uint_as_float(w0.x), cy = __uint_as_float(w0.y), cz = __uint_as_float(w0.z);
    const float sx = __uint_as_float(w0.w), sy = __uint_as_float(w1.x), sz = __uint_as_float(w1.y);
    const bool rx = signbit(inv.x), ry = signbit(inv.y), rz = signbit(inv.z);
    const uint32_t nx = rx ? w2.y : w1.z, fx = rx ? w1.z : w2.y;
    const uint32_t ny = ry ? w2.z : w1.w, fy = ry ? w1.w : w2.z;
    const uint32_t nz = rz ? w2.w : w2.x, fz = rz ? w2.x : w2.w;
    // a plane's distance: its float from its byte (v_cvt_f32_ubyte, then an exact fma), and the slab's multiply-add
    const auto at = [](uint32_t word, int c, float step, float corner, float inv_a, float os_a) {
        const float plane = fmaf(static_cast<float>((word >> (8 * c)) & 0xffu), step, corner);
        return fmaf(plane, inv_a, -os_a);
    };
    const int32_t child[wide_children] = {refs.x, refs.y, refs.z, refs.w};
#pragma unroll
    for (int c = 0; c < wide_children; ++c)
    {
        const float enter = fmaxf(fmaxf(at(nx, c, sx, cx, inv.x, os.x), at(ny, c, sy, cy, inv.y, os.y)),
                                  fmaxf(at(nz, c, sz, cz, inv.z, os.z), t_min));
        const float exit = fminf(fminf(at(fx, c, sx, cx, inv.x, os.x), at(fy, c, sy, cy, inv.y, os.y)),
                                 fminf(at(fz, c, sz, cz, inv.z, os.z), t_max));
        t[c] = ((enter <= exit) & (child[c] != no_child)) ? enter : inf<float>;
    }
}

// A wide node's four child references (read as the planes are: the node array's base plus a 32-bit offset)
template <typename Real>
__device__ inline int4 child_refs(const scene_view<Real>& sc, int32_t ref)
{
    const char* const nodes = reinterpret_cast<const char*>(sc.nodes);
    const uint32_t at = static_cast<uint32_t>(ref) * static_cast<uint32_t>(sizeof(wide_node<Real>));
    return *reinterpret_cast<const int4*>(nodes + (at + static_cast<uint32_t>(24 * sizeof(Real))));
}

// One compare-exchange of the sorting network: the nearer entry first
template <typename Real>
__device__ inline void order(Real& ta, int32_t& ra, Real& tb, int32_t& rb)
{
    const bool swap = tb < ta;
    const Real t = swap ? tb : ta;
    const int32_t r = swap ? rb : ra;
    tb = swap ? ta : tb;
    rb = swap ? ra : rb;
    ta = t;
    ra = r;
}

// An inner node's four children: where the ray enters each (or infinity) and their references, from the tree's own
// node records (Kinds: walk_packed reads packed_node, the others wide_node)
template <int Kinds, typename Real>
__device__ inline void node_children(const scene_view<Real>& sc, int32_t ref, const slab_order& so, vec3<Real> inv,
                                     vec3<Real> os, Real t_min, Real t_max, Real (&t)[wide_children], int4& refs)
{
    if constexpr (Kinds == walk_packed && sizeof(Real) == sizeof(float))
    {
        packed_entries(sc, ref, inv, os, t_min, t_max, t, refs);
    }
    else
    {
        child_entries(sc, ref, so, inv, os, t_min, t_max, t);
        refs = child_refs(sc, ref);
    }
}

// An inner node: its four children's boxes tested and sorted by entry distance (a five-exchange network); the entered
// ones but the nearest go on the stack, farthest first, and the nearest is the next reference (true), or none is
// entered (false)
template <int Kinds, typename Real, typename Tally, typename Stack>
__device__ inline bool descend(const scene_view<Real>& sc, int32_t& ref, const slab_order& so, vec3<Real> inv,
                               vec3<Real> os, Real t_min, Real t_max, const Stack& stack, int& size, Tally& tl)
{
    tl.add(tally_node);
    Real t[wide_children];
    int4 refs;
    node_children<Kinds>(sc, ref, so, inv, os, t_min, t_max, t, refs);
    Real t0 = t[0], t1 = t[1], t2 = t[2], t3 = t[3];
    int32_t r0 = refs.x, r1 = refs.y, r2 = refs.z, r3 = refs.w;
    order(t0, r0, t1, r1);
    order(t2, r2, t3, r3);
    order(t0, r0, t2, r2);
    order(t1, r1, t3, r3);
    order(t1, r1, t2, r2);
    if (t3 < inf<Real>)
    {
        stack.store(size++, r3, t3);
    }
    if (t2 < inf<Real>)
    {
        stack.store(size++, r2, t2);
    }
    if (t1 < inf<Real>)
    {
        stack.store(size++, r1, t1);
    }
    if (t0 < inf<Real>)
    {
        ref = r0;
        return true;
    }
    return false;
}

// The next stack entry above `floor` the ray may enter no farther than the closest hit so far (at exactly that distance
// it may hold a tie), or false when there is none
template <typename Real, typename Stack>
__device__ inline bool pop(const Stack& stack, int& size, Real t_max, int32_t& ref, int floor = 0)
{
    while (size > floor)
    {
        --size;
        if (stack.distance_at(size) <= t_max)
        {
            ref = stack.ref_at(size);
            return true;
        }
    }
    return false;
}

// The ray as the walk sees it, in the world's coordinates or moved into an instance's mesh, with what the box and
// sphere tests take from it
template <typename Real>
struct walk_ray
{
    ray<Real> r;
    Real inverse_a;
    vec3<Real> inv, os;
    slab_order so;
    __device__ void set(const ray<Real>& to)
    {
        r = to;
        inverse_a = Real(1) / length_squared(r.direction);
        inv = {safe_inverse(r.direction.x), safe_inverse(r.direction.y), safe_inverse(r.direction.z)};
        os = {r.origin.x * inv.x, r.origin.y * inv.y, r.origin.z * inv.z};
        so = slab_order_of(inv);
    }
};

// A world ray in an instance's mesh coordinates: x' = A x + b, the direction by A alone, so that a point at t along
// the world ray is the point at t along the moved one
template <typename Real>
__device__ inline ray<Real> into_instance(const Real* m, const ray<Real>& r)
{
    const auto row = [&](int k, vec3<Real> v) { return m_fma(m[4 * k], v.x, m_fma(m[4 * k + 1], v.y, m[4 * k + 2] * v.z)); };
    return {{row(0, r.origin) + m[3], row(1, r.origin) + m[7], row(2, r.origin) + m[11]},
            {row(0, r.direction), row(1, r.direction), row(2, r.direction)},
            r.time};
}

// The closest primitive a ray hits at t in (0.001, ∞), with the self-intersection rule for the primitive it leaves
// (src/gpu/bvh.cuh:576-706). Kinds: what the scene's tree may hold (walk_kinds). Nodes are visited near first and
// leaves tested in the reference's order; the loop takes one node or one leaf per iteration (on MI355X this measured
// faster than Aila and Laine's while-while form, which walks inner nodes in an inner loop until every lane of the wave
// has a leaf). An instance leaf moves the ray into its mesh's coordinates and goes on into its mesh's tree; the stack's
// entries below `floor` are the scene tree's, and when the mesh tree's are used up the ray is moved back.
//
// Shadows: the same walk answers a shadow ray's question for the lanes whose any_hit is set (closest_in_leaf): the
// closest distance starts at t_max, primitive `skip` is passed over and the walk ends at the first hit. Its answer is
// the any-hit walk's (quality.hpp, occluded): the same boxes are entered (children beyond t_max never are, and t_max
// doesn't shrink), only in another order. So the lanes of a wave walk their shadow rays and their paths' next rays in
// one loop.
template <typename Real, int Kinds, typename Tally, typename Stack, bool Shadows = false>
__device__ inline hit<Real> closest_hit(const scene_view<Real>& sc, const ray<Real>& r, int from, bool inward,
                                        const Stack& stack, Tally& tl, Real t_max = inf<Real>, int skip = -1,
                                        bool any_hit = false)
{
    const Real t_min = Real(0.001);
    hit<Real> closest{-1, Shadows ? t_max : inf<Real>, false};
    if (sc.stack_depth == 0)
    {
        return closest;
    }
    walk_ray<Real> w;
    w.set(r);
    int size = 0;
    int floor = 0;                       // in an instance: the scene tree's part of the stack
    int triangle_shift = sc.triangle_base; // in an instance: its triangles' id_base less its first triangle
    int32_t ref = sc.root_ref;
    while (true)
    {
        if (ref < 0)
        {
            if constexpr (Kinds == walk_instances)
            {
                if (leaf_kind(ref) == prim_instance)
                {
                    const instance_rec<Real>& in = sc.instances[leaf_first(ref)];
                    tl.add(tally_instance);
                    w.set(into_instance(in.to_object, r));
                    floor = size;
                    triangle_shift = in.id_base - in.first_triangle;
                    ref = in.root_ref;
                    continue;
                }
            }
            closest_in_leaf<Real, Kinds, Shadows>(sc, w.r, ref, from, inward, w.inverse_a, t_min, triangle_shift, closest,
                                                  tl, skip, any_hit);
            if (Shadows && any_hit && closest.index >= 0)
            {
                return closest;
            }
        }
        else if (descend<Kinds>(sc, ref, w.so, w.inv, w.os, t_min, closest.t, stack, size, tl))
        {
            continue;
        }
        if (pop(stack, size, closest.t, ref, floor))
        {
            continue;
        }
        if constexpr (Kinds == walk_instances)
        {
            if (floor == 0)
            {
                return closest; // (an instance entered on an empty stack leaves nothing of the scene tree to walk)
            }
            // the mesh tree is done: back to the world's coordinates and the scene tree's entries
            w.set(r);
            floor = 0;
            triangle_shift = sc.triangle_base;
            if (!pop(stack, size, closest.t, ref))
            {
                return closest;
            }
        }
        else
        {
            return closest;
        }
    }
}
}
