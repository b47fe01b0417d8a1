/*
 * raytracing_amd.h: the C ABI of the MI355X path tracer.
 *
 * The path tracer is reached through plain C: POD structs in double precision, pointers and sizes, integer status
 * codes and a thread-local error message. No C++ or torch type crosses this boundary, so any host (C, C++, Python
 * ctypes, a Go/Java/Node binding) can call it.
 *
 * Each entry point names the reference interface it replaces (paths relative to DaniilK19/RayTracing):
 *   - the scene tables mirror gpu::scene_desc and its parts      src/gpu/renderer.h:33-395
 *   - rt_scene_build mirrors build_gpu_scene                      src/gpu_scene.ixx:1333-1445
 *   - rt_render mirrors gpu::render                               src/gpu/renderer.h:526
 *   - rt_session_* mirror class gpu::session                      src/gpu/renderer.h:585-673
 *
 * Errors. The reference throws std::runtime_error (GPU and scene errors), std::invalid_argument (options out of range)
 * and std::logic_error (state misuse), renderer.h:513-526 and :588-601. Here each maps to a status code, and
 * rt_last_error() returns the message of the last failure on the calling thread:
 *   RT_OK 0, RT_ERROR_RUNTIME 1, RT_ERROR_INVALID_ARGUMENT 2, RT_ERROR_LOGIC 3.
 * Every check that can fail without the GPU is made before any device memory is touched.
 *
 * Threading. A session belongs to one host thread at a time (renderer.h, session: no reentrancy). Different sessions
 * may be used from different threads; each has its own HIP stream.
 */
#ifndef RAYTRACING_AMD_H
#define RAYTRACING_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RT_ABI_VERSION 4

enum rt_status
{
    RT_OK = 0,
    RT_ERROR_RUNTIME = 1,
    RT_ERROR_INVALID_ARGUMENT = 2,
    RT_ERROR_LOGIC = 3,
};

/* A point, direction or color in double precision: gpu::vec3d, src/gpu/renderer.h:33-41 */
typedef struct rt_vec3
{
    double x, y, z;
} rt_vec3;

/* --- Textures: gpu::texture_desc, image_desc, perlin_desc (src/gpu/renderer.h:45-90) ----------------------------- */

enum rt_texture_kind
{
    RT_TEXTURE_SOLID = 0,
    RT_TEXTURE_CHECKER = 1,
    RT_TEXTURE_IMAGE = 2,
    RT_TEXTURE_NOISE = 3,
};

typedef struct rt_texture_desc
{
    int32_t kind;      /* rt_texture_kind */
    int32_t even;      /* checker: index of the texture of even cubes (smaller than this texture's own index) */
    int32_t odd;       /* checker: index of the texture of odd cubes */
    int32_t image;     /* image: index into rt_scene.images */
    int32_t perlin;    /* noise: index into rt_scene.perlins */
    int32_t reserved;
    rt_vec3 rgb;       /* solid: the color */
    double inv_scale;  /* checker: 1 / the side of its cubes */
    double scale;      /* noise: the marble's stripe frequency along z */
} rt_texture_desc;

/* Linear colors, rows from the top: pixel (x, y) starts at rgb[3 * (y * width + x)] */
typedef struct rt_image_desc
{
    int32_t width;
    int32_t height;
    const float* rgb;
} rt_image_desc;

#define RT_PERLIN_POINT_COUNT 256

typedef struct rt_perlin_desc
{
    rt_vec3 random_vectors[RT_PERLIN_POINT_COUNT];
    int32_t perm_x[RT_PERLIN_POINT_COUNT];
    int32_t perm_y[RT_PERLIN_POINT_COUNT];
    int32_t perm_z[RT_PERLIN_POINT_COUNT];
} rt_perlin_desc;

/* --- Materials: gpu::material_desc (src/gpu/renderer.h:94-111) --------------------------------------------------- */

enum rt_material_kind
{
    RT_MATERIAL_LAMBERTIAN = 0,
    RT_MATERIAL_METAL = 1,
    RT_MATERIAL_DIELECTRIC = 2,
    RT_MATERIAL_DIFFUSE_LIGHT = 3,
    RT_MATERIAL_ISOTROPIC = 4,
};

typedef struct rt_material_desc
{
    int32_t kind;            /* rt_material_kind */
    int32_t texture;         /* lambertian, diffuse_light, isotropic: index into rt_scene.textures */
    rt_vec3 albedo;          /* metal */
    double fuzz;             /* metal, at most 1 */
    double refraction_index; /* dielectric */
} rt_material_desc;

/* --- Shapes in world coordinates: gpu::sphere_desc, quad_desc (src/gpu/renderer.h:115-139) ----------------------- */

typedef struct rt_sphere_desc
{
    rt_vec3 center;       /* at time 0 */
    double radius;
    rt_vec3 displacement; /* the center moves by this much from time 0 to 1; exactly 0 for a sphere standing still */
    double uv_cos;        /* the turn of rotate_y instances baked into it, for its texture coordinates */
    double uv_sin;
    int32_t material;
    int32_t reserved;
} rt_sphere_desc;

typedef struct rt_quad_desc
{
    rt_vec3 q;
    rt_vec3 u;
    rt_vec3 v;
    int32_t material;
    int32_t reserved;
} rt_quad_desc;

/* --- Smoke and fog: gpu::medium_desc (src/gpu/renderer.h:146-161) ------------------------------------------------- */

enum rt_boundary_shape
{
    RT_BOUNDARY_SPHERE = 0,
    RT_BOUNDARY_CONVEX_QUADS = 1,
};

typedef struct rt_medium_desc
{
    int32_t boundary;       /* rt_boundary_shape */
    int32_t material;       /* an isotropic material */
    double neg_inv_density; /* -1 / density */
    rt_sphere_desc sphere;  /* RT_BOUNDARY_SPHERE */
    const rt_quad_desc* quads; /* RT_BOUNDARY_CONVEX_QUADS */
    int64_t quad_count;
} rt_medium_desc;

/* --- The lights list the lights/mixture strategies aim at: gpu::light_desc (src/gpu/renderer.h:172-191) --------- */

enum rt_light_shape
{
    RT_LIGHT_QUAD = 0,
    RT_LIGHT_SPHERE = 1,
};

typedef struct rt_light_desc
{
    int32_t kind; /* rt_light_shape */
    int32_t reserved;
    rt_quad_desc quad;
    rt_vec3 center;
    double radius;
    double weight; /* the chance of aiming at it, in (0, 1] */
} rt_light_desc;

/* --- The camera: gpu::camera_desc (src/gpu/renderer.h:195-232) ---------------------------------------------------- */

enum rt_sampling_strategy
{
    RT_SAMPLING_UNIFORM = 0,
    RT_SAMPLING_MATERIAL = 1,
    RT_SAMPLING_LIGHTS = 2,
    RT_SAMPLING_MIXTURE = 3,
};

typedef struct rt_camera_desc
{
    int32_t image_width;
    int32_t image_height;
    int32_t sqrt_spp;  /* the side of a pixel's grid of strata: a round is sqrt_spp^2 passes */
    int32_t max_depth; /* at most this many rays per path */
    rt_vec3 center;
    rt_vec3 pixel00_loc;
    rt_vec3 pixel_delta_u;
    rt_vec3 pixel_delta_v;
    int32_t defocus;   /* rays start on the lens disk */
    int32_t sampling;  /* rt_sampling_strategy */
    rt_vec3 defocus_disk_u;
    rt_vec3 defocus_disk_v;
    rt_vec3 sky_bottom;
    rt_vec3 sky_top;
} rt_camera_desc;

/* --- Beyond the books: light combination, GGX side entries, environment (src/gpu/renderer.h:240-276) ------------- */

enum rt_light_combination
{
    RT_COMBINE_BSDF_ONLY = 0,       /* -p material */
    RT_COMBINE_LIGHTS_ONLY = 1,     /* -p nee */
    RT_COMBINE_POWER_HEURISTIC = 2, /* -p mis */
};

enum rt_microfacet_kind
{
    RT_MICROFACET_NONE = 0,
    RT_MICROFACET_CONDUCTOR = 1,
    RT_MICROFACET_GLOSS = 2,
};

typedef struct rt_microfacet_desc
{
    int32_t kind; /* rt_microfacet_kind */
    int32_t reserved;
    double roughness; /* r in [0, 1]; the GGX width is r^2 */
} rt_microfacet_desc;

typedef struct rt_environment_desc
{
    rt_image_desc image; /* equirectangular, +y up, row 0 is +y; linear radiance */
    double yaw_degrees;
    double scale;
} rt_environment_desc;

/* --- Triangle meshes (beyond the reference, which renders triangles only through OptiX) -------------------------- */

typedef struct rt_mesh_desc
{
    const float* positions;   /* 3 per vertex */
    int64_t vertex_count;
    const uint32_t* indices;  /* 3 per triangle, counter-clockwise seen from the front */
    int64_t triangle_count;
    int32_t material;         /* index into rt_scene.materials */
    int32_t reserved;
} rt_mesh_desc;

typedef struct rt_mesh_instance_desc
{
    int32_t mesh;
    int32_t reserved;
    double object_to_world[12]; /* row-major 3x4 */
} rt_mesh_instance_desc;

/* --- The scene: gpu::scene_desc (src/gpu/renderer.h:381-395) ----------------------------------------------------- */

typedef struct rt_scene
{
    const rt_sphere_desc* spheres;
    int64_t sphere_count;
    const rt_material_desc* materials;
    int64_t material_count;
    rt_camera_desc camera;
    const rt_texture_desc* textures;
    int64_t texture_count;
    const rt_image_desc* images;
    int64_t image_count;
    const rt_perlin_desc* perlins;
    int64_t perlin_count;
    const rt_quad_desc* quads;
    int64_t quad_count;
    const rt_medium_desc* media;
    int64_t medium_count;
    const rt_light_desc* lights;
    int64_t light_count;
    /* quality_desc */
    int32_t combination; /* rt_light_combination */
    int32_t reserved;
    const rt_microfacet_desc* microfacets; /* empty, or one per material */
    int64_t microfacet_count;
    const rt_environment_desc* environment; /* NULL: the camera's sky */
    /* triangles */
    const rt_mesh_desc* meshes;
    int64_t mesh_count;
    const rt_mesh_instance_desc* instances;
    int64_t instance_count;
} rt_scene;

/* --- Built-in scenes: build_gpu_scene (src/gpu_scene.ixx:1303-1445) ----------------------------------------------
 * The scenes of the books (materials, weekend, bouncing, checker, earth, perlin, quads, light, cornell, smoke, final,
 * aluminum, glass) and the test scenes (furnace), built by this library's own C++ with the reference's mt19937_64 stream
 * from its default seed, then flattened as gpu_scene.ixx does. The earth's image is a seeded noise image (the reference
 * loads assets/earthmap.jpg): see rt_scene_overrides.earth_width. */

typedef struct rt_scene_overrides
{
    int32_t image_width;       /* 0: the scene's own */
    int32_t image_height;      /* 0: from the scene's aspect ratio */
    int32_t samples_per_pixel; /* 0: the scene's own */
    int32_t strata;            /* 0: samples rounded to one k x k grid; > 0: rounds of strata x strata */
    int32_t sampling;          /* -1: the scene's own; else rt_sampling_strategy, or 4 nee, 5 mis */
    int32_t max_depth;         /* 0: the scene's own */
    int32_t earth_width;       /* the noise image standing in for earthmap.jpg: 0 means 2048 (x 1024) */
    uint32_t earth_seed;       /* its generator's seed */
} rt_scene_overrides;

typedef struct rt_scene_handle rt_scene_handle;

/* name: a built-in scene (rt_scene_names), or a glTF 2.0 model's file, FILE.gltf or FILE.glb, which renders in a
 * studio: its triangle meshes placed by its nodes, with their metallic-roughness factors, on a ground under a seeded
 * sky with a sun, framed by the camera (as RayTracingGPU takes a model's file for its scene, src/main_gpu.cpp:141-143,
 * whose models the reference renders through OptiX, src/gltf.ixx:34). A model the loader refuses (a file cut short, a
 * required extension, an index past its vertices, a cycle of nodes...) is RT_ERROR_RUNTIME, one without a triangle
 * RT_ERROR_INVALID_ARGUMENT, with the reason in rt_last_error(). */
int rt_scene_build(const char* name, const rt_scene_overrides* overrides, rt_scene_handle** out);
const rt_scene* rt_scene_get(const rt_scene_handle* handle);
int64_t rt_scene_passes(const rt_scene_handle* handle); /* sqrt_spp^2, or the samples asked for with strata */
/* A model's warnings, one line each ("" for a built-in scene): what the file holds that isn't rendered (textures,
 * points and lines, skins). Valid while the handle is. */
const char* rt_scene_warnings(const rt_scene_handle* handle);
void rt_scene_free(rt_scene_handle* handle);
/* The names of the built-in scenes, separated by commas */
const char* rt_scene_names(void);

/* --- Rendering: gpu::render_options, session_options (src/gpu/renderer.h:405-454, 530-543) --------------------- */

enum rt_precision
{
    RT_FLOAT32 = 0,
    RT_FLOAT64 = 1,     /* the books' scenes (the quality integrator is float32 only, as the reference's; a scene that
                           needs it is RT_ERROR_INVALID_ARGUMENT) */
    RT_FLOAT64_ALL = 2, /* float64 for every scene, the quality integrator's too, on the megakernels: no counterpart in
                           the reference, for the FP64 oracle that renders' RMSE is measured against */
};

enum rt_integrator
{
    RT_INTEGRATOR_AUTO = 0,       /* the wavefront integrator where it is available */
    RT_INTEGRATOR_MEGAKERNEL = 1, /* one thread per path, the whole path in one kernel */
    RT_INTEGRATOR_WAVEFRONT = 2,  /* per-bounce stages over queues of paths */
    RT_INTEGRATOR_PLAIN = 3,      /* the reference's mapping (src/gpu/renderer.cu:340-397,421-451): one thread per pixel
                                     of a 16x8 block, all of a launch's passes in its loop; the baseline the others are
                                     measured against. Same image bytes as the others. */
};

typedef struct rt_session_options
{
    uint64_t seed;
    int32_t precision;      /* rt_precision */
    int32_t integrator;     /* rt_integrator */
    double launch_seconds;  /* how long one launch is planned to take: (0, 0.5]; 0 means 0.05 */
    int32_t block_size;     /* megakernel threads per block: 0 for the tuned default */
    int32_t wavefront_slots;/* wavefront path slots: 0 for the tuned default */
    int32_t wavefront_graph;/* wavefront bounces per captured hipGraph: 0 for the tuned default, -1 for none (every
                               kernel launched on its own, as profilers that don't follow graphs need) */
    int32_t wavefront_drain;/* wavefront: when fewer slots than this are active, each finishes its pixel's passes on
                               its own (0: the tuned default, a sixteenth of the pool; -1: never) */
    int32_t megakernel_waves;/* megakernel: the waves per SIMD its registers are budgeted for: 0 (the compiler's
                               choice), 5, 6 or 7 (float only) */
    int32_t tree;           /* rt_tree: how the session's BVH is split */
    int32_t nodes;          /* rt_nodes: how its inner nodes lie in memory */
    int32_t kernel;         /* rt_kernel: which shading renders the scene */
    int32_t wait;           /* rt_wait: how the host waits for the session's launches */
} rt_session_options;

/* How the host waits for the GPU (gpu::wait_mode, src/gpu/renderer.h:505-516,536-540) */
enum rt_wait
{
    RT_WAIT_SPIN = 0,  /* asks the GPU again and again: reacts at once, keeps a core busy */
    RT_WAIT_SLEEP = 1, /* sleeps until the GPU signals; a session asks for it only after rt_query_device_wait set the
                          device's waits to sleep (RT_ERROR_LOGIC otherwise) */
};

/* Which shading renders a scene (gpu::session_options all_books_kernel and quality_kernel, src/gpu/renderer.h:418-454,
 * src/gpu/renderer.cu:1180-1199): by default the quality integrator for scenes that need it (light samples, GGX, an
 * environment map, triangles), the first book's for scenes of still spheres and the books' basic materials, and the
 * three books' for the rest. The others are checks, and measures of what generality costs. */
enum rt_kernel
{
    RT_KERNEL_AUTO = 0,
    RT_KERNEL_ALL_BOOKS = 1, /* the three books' shading also for the first book's scenes (the same image, bit for
                                bit); refused for scenes that need the quality integrator */
    RT_KERNEL_QUALITY = 2,   /* the quality integrator also for the books' scenes under -p material (the same image
                                within its noise); float32 (or RT_FLOAT64_ALL) only */
};

/* How the BVH's inner nodes lie in memory. Both hold the same tree and give the same closest hits (bit for bit): a
 * packed box holds its float box, so the walk enters every child it entered before. */
enum rt_nodes
{
    RT_NODES_AUTO = 0,   /* packed for the quality integrator's one-level trees whose float nodes pass 16 MiB (half
                            the L2 caches' 32 MiB), float otherwise and for the plain megakernel (RT_INTEGRATOR_PLAIN
                            renders faster on float nodes) */
    RT_NODES_FLOAT = 1,  /* 128 bytes a node: four children's boxes as floats */
    RT_NODES_PACKED = 2, /* 64 bytes a node: each child's planes as bytes on a grid of power-of-two steps over the
                            node's box, rounded outward (no counterpart in the reference, whose nodes are the binary
                            tree's 32-byte records, src/gpu/bvh.cuh:29-38). Float32 scenes that render on the quality
                            integrator with a one-level tree; otherwise RT_ERROR_INVALID_ARGUMENT */
};

/* The BVH's splits. Every tree holds every primitive and finds the same closest hit on scenes without instanced
 * meshes (of two primitives hit at exactly the same distance, the one of lower index in the scene is kept with
 * either); an instanced mesh's hits are found in its own coordinates, so their distances may differ from the baked
 * tree's in the last bits and their ids are the instance's own (each instance a range of ids). */
enum rt_tree
{
    RT_TREE_DEFAULT = 0,   /* binned surface-area-heuristic splits, leaves of one kind and at most 8 primitives, in
                              one level: every instance's triangles baked into world coordinates */
    RT_TREE_MEDIAN = 1,    /* the reference's object-median splits on the longest axis, with its leaf sizes
                              (src/gpu/bvh.cuh:382-444); scenes with triangles take the default tree */
    RT_TREE_SAH = 2,       /* the default tree */
    RT_TREE_TWO_LEVEL = 3, /* the default's splits in two levels: a mesh placed by several instances gets a tree of
                              its own, once, which each instance's leaf enters with the ray moved into the mesh's
                              coordinates (the reference's mesh_set instances, src/gpu/renderer.h:305-376); it holds
                              each such mesh once, and the kernels that can enter an instance are kernels of their own */
};
/* RT_INTEGRATOR_AUTO (float32): the session's autotuner renders its first full passes with each candidate (the
 * megakernel at each register budget, one wave a block, and the wavefront), a few passes each; the three fastest render
 * launches of several passes each, in a forward and a backward sweep, and the fastest of those renders the rest.
 * Every candidate traces the same paths with the same random numbers and sums them in the same order, so the image
 * doesn't depend on which one renders which pass. */

typedef struct rt_render_report
{
    int64_t passes;              /* state after the call */
    int32_t rows;
    int32_t launches;
    double gpu_seconds;
    double longest_launch_seconds;
    int64_t bounces;             /* wavefront: bounce steps launched by the call (extend + shade over the pool) */
} rt_render_report;

typedef struct rt_kernel_info
{
    char name[32];
    int32_t registers;       /* VGPRs per thread (architected + accumulation) */
    int32_t local_bytes;     /* private (scratch) bytes per thread */
    int32_t lds_bytes;       /* LDS per block */
    int32_t threads_per_block;
    int32_t blocks_per_cu;   /* as many as fit at once */
    int32_t compute_units;
} rt_kernel_info;

typedef struct rt_session rt_session;

/* gpu::session::session: validates, builds the BVH, uploads; the view is scene->camera */
int rt_session_create(const rt_scene* scene, const rt_session_options* options, rt_session** out);
void rt_session_destroy(rt_session* session);
/* gpu::session::set_camera: a new view, starting from no samples */
int rt_session_set_camera(rt_session* session, const rt_camera_desc* camera);
/* gpu::session::camera and has_view (src/gpu/renderer.h:595-604): the last view set successfully; whether the
 * session has a view (a set_camera whose image the GPU couldn't hold leaves none, and the renders and their reports
 * refuse with RT_ERROR_LOGIC until another succeeds) */
int rt_session_camera(const rt_session* session, rt_camera_desc* camera);
int rt_session_has_view(const rt_session* session, int32_t* has_view);
/* gpu::session::render_to: render until the image holds `passes` complete passes and `rows` rows of the next */
int rt_session_render_to(rt_session* session, int64_t passes, int32_t rows, rt_render_report* report);
/* Progress (gpu::render_options::progress, src/gpu/renderer.h:423-427; the on_pass of session::render_to): called on
 * the calling thread after each launch that completes passes (one launch may complete several), with the passes done
 * so far and the render's target. It must not call into the library for the same session. */
typedef void (*rt_progress_fn)(int64_t done, int64_t total, void* user);
/* rt_session_render_to with `progress` (may be NULL) called as passes complete, `user` passed through */
int rt_session_render_to_progress(rt_session* session, int64_t passes, int32_t rows, rt_render_report* report,
                                  rt_progress_fn progress, void* user);
/* rt_session_render_to without device allocation. No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops that must not call hipMalloc). The session's
 * whole passes up to `passes` render on the megakernel (at the session's register budget, or the compiler's when the
 * session renders on the wavefront or is still tuning), with the chunk queue's head and the staged samples in a static
 * device array of 64 MiB that the library holds on each GPU it runs on until the process ends; the image's sums, the
 * counters and the scene are the session's, allocated when it was created. Same image bytes as rt_session_render_to.
 * Limit: one pass of the image must fit the array, W x H x 16 bytes (W x H x 32 in float64): about 4.19 M pixels in
 * float32; a larger image is RT_ERROR_INVALID_ARGUMENT. Every call on a GPU shares that GPU's array: calls that
 * overlap on one GPU (sessions on other threads) race, and must be serialized by the caller. */
int rt_session_render_to_static(rt_session* session, int64_t passes, rt_render_report* report);
/* gpu::session::render_for: about `seconds` of GPU time, never past pass_limit */
int rt_session_render_for(rt_session* session, double seconds, int64_t pass_limit, rt_render_report* report);
int rt_session_state(const rt_session* session, int64_t* passes, int32_t* rows);
/* gpu::session::pixels: sums / passes, W*H*3 doubles, rows from the top */
int rt_session_pixels(rt_session* session, double* out, size_t count);
/* The raw per-pixel double sums, W*H*3 (for exact checks of accumulation order) */
int rt_session_sums(rt_session* session, double* out, size_t count);
/* gpu::session::display: the books' bytes (sqrt, clamp 0.999, x256), RGBA, rows from the top */
int rt_session_display_rgba8(rt_session* session, uint8_t* out, size_t count);

/* Tone maps beyond the books' bytes (gpu::tone_map, src/gpu/renderer.h:282-303) */
enum rt_tone_map
{
    RT_TONE_BOOKS = 0,   /* the books' bytes, as rt_session_display_rgba8 (no exposure) */
    RT_TONE_NONE = 1,    /* the light times 2^exposure, clipped to [0, 1], the sRGB curve */
    RT_TONE_AGX = 2,     /* AgX (Troy Sobotka's; Minimal AgX's matrices, a closed-form curve) */
    RT_TONE_NEUTRAL = 3, /* Khronos PBR Neutral, then the sRGB curve */
};
#define RT_MAX_EXPOSURE 24.0 /* stops, either way */
/* display(display_options) (src/gpu/renderer.h:624-635): each pixel's light (its sums / passes rounded to float, times
 * 2^exposure) through the tone map, dithered with interleaved gradient noise and quantized, RGBA, rows from the top.
 * RT_ERROR_INVALID_ARGUMENT for an exposure that isn't finite and within RT_MAX_EXPOSURE, for RT_TONE_BOOKS with an
 * exposure other than 0, and for a tone map other than RT_TONE_BOOKS in a float64 session (the curves compute in
 * float only), as the reference refuses them. */
int rt_session_display_toned(rt_session* session, int32_t tone_map, double exposure, uint8_t* out, size_t count);
/* gpu::session::rays, shadow_rays, kernel_seconds */
int rt_session_counters(const rt_session* session, uint64_t* rays, uint64_t* shadow_rays, double* kernel_seconds);
/* gpu::session::kernel */
int rt_session_kernel_info(const rt_session* session, rt_kernel_info* info);

/* gpu::render: a session rendered to `passes` (0: sqrt_spp^2), its pixels copied to `pixels` (W*H*3 doubles) */
typedef struct rt_render_options
{
    rt_session_options session;
    int64_t passes;
    rt_progress_fn progress; /* may be NULL: called as passes complete (rt_progress_fn) */
    void* progress_user;
    /* render_options::display: a tone map other than RT_TONE_BOOKS fills rgba (W*H*4 bytes) with the display's bytes
     * after the last pass, as rt_session_display_toned; RT_TONE_BOOKS (0) leaves rgba alone (may be NULL) */
    int32_t tone_map;
    double exposure;
    uint8_t* rgba;
} rt_render_options;

typedef struct rt_render_stats
{
    double seconds;
    double kernel_seconds;
    uint64_t rays;
    uint64_t shadow_rays;
    rt_kernel_info kernel; /* the kernel that rendered (render_result::kernel) */
} rt_render_stats;

int rt_render(const rt_scene* scene, const rt_render_options* options, double* pixels, size_t count,
              rt_render_stats* stats);

/* --- Parity probes (the reference's GpuTests drive closest_hit through test kernels, tests/test_scene.h) -------- */

/* The session's BVH traversal on a batch of rays given on the host: origins and directions (3 floats each) and times,
 * with each ray's `from` primitive and `inward` flag (src/gpu/bvh.cuh:576-706). Writes the closest primitive's global
 * id (-1 for none), t, and whether it hit the front. Float32 sessions only. */
int rt_session_closest_hit(rt_session* session, const float* origins, const float* directions, const float* times,
                           const int32_t* from, const uint8_t* inward, int64_t count, int32_t* index, float* t,
                           uint8_t* entering);

/* Shadow rays (src/gpu/shadow.cuh:26-120, the reference's `occluded`): for each ray, whether any primitive but
 * skip[k] (a global id, or -1) lies in (0.001, t_max[k]) along it, with the leaving primitive's rules (from, inward) of
 * rt_session_closest_hit; spheres on the open interval, quads and triangles on the closed one. Both of the product's
 * shadow walks answer: `blocked` the wavefront's (occluded), `walked` the quality megakernel's (its closest-hit walk
 * in any-hit mode), 1 or 0 per ray; either may be NULL. Float32 sessions only. */
int rt_session_occluded(rt_session* session, const float* origins, const float* directions, const float* times,
                        const int32_t* from, const uint8_t* inward, const float* t_max, const int32_t* skip,
                        int64_t count, uint8_t* blocked, uint8_t* walked);

/* The media along given rays (src/gpu/camera.cuh:259-284, src/gpu/constant_medium.cuh:14-125, src/gpu/shadow.cuh:
 * 122-165): for each ray (origins and directions, 3 floats each, times) whose surfaces end at t_surface, the media's
 * event as the paths find it -- each medium's chord within (0.001, t_surface) and a free flight through it with the
 * per-sample generator of (seed, pixel = the ray's index, pass = 0), the nearest collision kept -- as the medium's
 * index among the scene's media (-1: none) and t (infinity: none); and the transmittance of (0.001, t_surface), the
 * weight of a shadow ray through the same media. Float32 sessions only. */
int rt_session_media_events(rt_session* session, const float* origins, const float* directions, const float* times,
                            const float* t_surface, int64_t count, uint64_t seed, int32_t* medium, float* t,
                            float* transmitted);

/* Light samples (src/gpu/light_sampler.cuh:40-102, src/gpu/lamps.cuh:41-146, src/gpu/environment.cuh:59-135): from
 * each of `count` points (3 floats each), one draw of the session's light set with the per-sample generator of
 * (seed, pixel = the point's index, pass = 0): the direction wi (3 floats), the distance to the lamp's point
 * (infinity for the map), its radiance (3), its density per solid angle with the light's choice included, and the lamp
 * primitive aimed at (skip: a global id, -1 for the map, -2 where the draw failed and the other outputs are 0);
 * hit_density: the density the light set gives the same direction when a scattered ray finds that light (what MIS
 * weighs a BSDF sample by). Quality scenes (light samples, a map or GGX) in float32 sessions only. */
int rt_session_light_samples(rt_session* session, const float* points, int64_t count, uint64_t seed, float* wi,
                             float* distance, float* radiance, float* density, int32_t* skip, float* hit_density);

/* The reference's own traversal, for reference-exact closest hits (src/gpu/bvh.cuh:576-706): the session's binary tree
 * (with RT_TREE_MEDIAN, the reference's tree) walked near first in the reference's order (the first child when both
 * are entered at the same distance; the farther child pushed; a popped node visited only when entered before the
 * closest hit), spheres tested on the open interval (0.001, closest) and quads on the closed one (src/gpu/sphere.cuh:116,
 * src/gpu/quad.cuh:70). A tie then goes to the first sphere and the last quad the walk meets, as in the reference,
 * where rt_session_closest_hit gives it to the primitive of lower scene index. Same arguments as
 * rt_session_closest_hit; float32 sessions only. */
int rt_session_reference_hits(rt_session* session, const float* origins, const float* directions, const float* times,
                              const int32_t* from, const uint8_t* inward, int64_t count, int32_t* index, float* t,
                              uint8_t* entering);

/* The per-sample random numbers the kernels draw (src/gpu/random.cuh:25-139), on device 0: for each of `count` items,
 * the generator of (seed, pixels[k], passes[k]) (SplitMix64 hashes into PCG32) and, each from a fresh generator of the
 * item, `draws` raw 32-bit words, `draws` float draws (the top 24 bits of a word times 2^-24), `draws` double draws
 * (53 bits of two words, the first one high), and nine floats: a uniform unit vector, a uniform point of the unit disk
 * (z = 0) and a cosine-weighted direction about +z, as the paths draw them. Any output may be NULL; draws <= 4096. */
int rt_sample_numbers(uint64_t seed, const uint64_t* pixels, const uint64_t* passes, int64_t count, int32_t draws,
                      uint32_t* words, float* floats, double* doubles, float* directions);

/* The display's tone maps on their own (src/gpu/tonemap.cu; the reference's tone tests drive encode and quantize
 * through test kernels, tests/tone_tests.cu:387-400): `count` linear colours (3 floats each) times 2^exposure through
 * RT_TONE_NONE, _AGX or _NEUTRAL, each channel quantized to a byte with dither[k] in [-0.5, 0.5) (NULL: none), into
 * out (3 bytes per colour). On device 0. */
int rt_tone_probe(int32_t tone_map, double exposure, const float* colours, const float* dither, int64_t count,
                  uint8_t* out);

/* The quality path's scattering lobes (src/gpu/bsdf.cuh:58-201, src/gpu/ggx.cuh:59-324), which the reference's tests
 * drive through test kernels (tests/bsdf_tests.cu, tests/ggx_tests.cu, tests/gloss_tests.cu) */
enum rt_lobe
{
    RT_LOBE_LAMBERT = 0,   /* Lambert's cosine lobe of albedo `color` */
    RT_LOBE_PHASE = 1,     /* the isotropic phase function */
    RT_LOBE_CONDUCTOR = 2, /* GGX conductor of F0 `color` and alpha (= roughness^2; below 1e-3 a Fresnel mirror),
                              with the multiple-scattering compensation of the energy table at `roughness` */
    RT_LOBE_GLOSS = 3,     /* glossy coat (GGX dielectric of F0 0.04, alpha) over a matte base of color `color` */
    RT_LOBE_METAL = 4,     /* the books' fuzzed metal: albedo `color`, fuzz `roughness` (a delta lobe: density 0) */
    RT_LOBE_DIELECTRIC_ENTERING = 5, /* the books' glass, from outside: refraction index `alpha` */
    RT_LOBE_DIELECTRIC_LEAVING = 6,  /* the same, from inside (v on the inner side) */
};
/* On device 0, for `count` cases in the tangent frame (the normal is +z; v toward the viewer, l toward the light, 3
 * floats each): value (3 per case) = f(v, l) cos_l and density = the density the lobe draws l with from v (per solid
 * angle; 0 for a mirror); and a direction drawn from v by the per-sample generator of (seed, pixel = case, pass = 0):
 * drawn (3 per case), its weight f cos / density (3) and drawn_density, all zero where the draw ends the path. */
int rt_scatter_probe(int32_t lobe, float alpha, float roughness, const float color[3], int64_t count, const float* v,
                     const float* l, uint64_t seed, float* value, float* density, float* drawn, float* weight,
                     float* drawn_density);

/* The flat BVH this library builds for a scene in float32 with the splits of `tree` (an rt_tree), on the host (no GPU
 * needed), in the reference's depth-first node layout (src/gpu/bvh.cuh:29-38,382-435): per node box_min[3], box_max[3]
 * as floats, and index, count as ints; and the global ids of the scene's spheres and quads. Call with NULL buffers to
 * learn the counts. */
int rt_scene_bvh(const rt_scene* scene, int32_t tree, float* node_boxes, int32_t* node_ints, int64_t* node_count,
                 int32_t* sphere_ids, int32_t* quad_ids, int32_t* depth);

/* The four-wide tree the kernels walk for a scene in float32 with the splits of `tree` (an rt_tree), on the host (no
 * GPU needed): per inner node its 128-byte record of float planes (lo[3][4] then hi[3][4] as floats, then the four
 * children's references as int32: a node's index, a leaf reference with bit 31 set, or 0x7fffffff for no child) and
 * the same node packed into 64 bytes as RT_NODES_PACKED lays it out (origin[3], step[3] as floats, lo[3][4] and hi[3][4]
 * as bytes, the four references). Call with NULL buffers to learn the count. No counterpart in the reference (its
 * nodes are the binary tree's, rt_scene_bvh); for tests of the packing. */
int rt_scene_wide_nodes(const rt_scene* scene, int32_t tree, void* float_nodes, void* packed_nodes, int64_t* count);

/* The quality integrator's light set as the kernels read it, on the host (build_lamps, src/gpu/lamps.cuh:159 and
 * src/gpu/lamp_table.cu:60-119): the scene's glowing quads and still spheres, 80 bytes each (20 floats: corner[3],
 * size, u[3], kind, v[3], primitive, normal[3], 0, radiance[3], 0; kind and primitive are int32 bit patterns). Call
 * with lamps NULL to learn the count. */
int rt_scene_lamps(const rt_scene* scene, float* lamps, int64_t* count);

/* An environment map's tables, on the host (build_environment, src/gpu/environment_data.h:58 and
 * src/gpu/environment_table.cu:139-225): per texel r, g, b and the density per steradian of its cell's directions;
 * the row borders' cosines (height + 1); the alias table's coins and aliases (2^log2_bins bins, none for a black map);
 * and S, the map's luminance over the sphere. Call with texels NULL to learn log2_bins alone. */
int rt_environment_tables(const rt_environment_desc* environment, float* texels, float* row_cos, float* bin_q,
                          int32_t* bin_alias, int32_t* log2_bins, double* total);

/* The scene data the view's paths read, counted by an instrumented instance of the megakernel over passes
 * [0, passes) of the view's first round, into scratch sums (the view's image and counters don't change). events[k] counts
 * reads of one kind of record (in order: wide BVH nodes, spheres, moving spheres, quads, media, media boundary quads,
 * hits, sphere frames, textures, texels, Perlin gradients, the books' lights, triangles, lamps of the light set,
 * environment texels, alias bins, mesh instances) and event_bytes[k] is that record's size; bytes_per_path is their
 * weighted sum per path, plus the double sums' read and write of each pixel once per launch. This is the algorithmic
 * traffic the roofline is measured with. No counterpart in the reference. */
#define RT_TRAFFIC_EVENTS 17
#define RT_TRAFFIC_PHASES 4
typedef struct rt_traffic
{
    uint64_t paths;
    uint64_t rays;
    uint64_t shadow_rays;
    uint64_t events[17];
    double event_bytes[17];
    double bytes_per_path;
    /* the same kernel's lanes' shader-clock cycles by phase of their paths' iterations (in order: the queue, with a
     * lane's wait for an item; the BVH walks; the media; the shading), summed over lanes: where the issue time goes,
     * a lane that idles while others trace counting the wait in the phase it goes on with */
    uint64_t phase_lane_cycles[4];
} rt_traffic;

int rt_session_traffic(rt_session* session, int64_t passes, rt_traffic* traffic);

/* --- Serving: progressive renders of mixed requests on one device (config 4) ------------------------------------- */

typedef struct rt_server rt_server;

typedef struct rt_server_options
{
    uint64_t memory_budget_bytes; /* device memory the server may hold for requests (0: 16 GiB) */
    int32_t max_active;           /* requests rendering at once (0: 8) */
    int32_t passes_per_step;      /* passes a request gets per scheduling step (0: 4) */
    int64_t paths_per_step;       /* a request's step is raised to at least this many paths: a view of P pixels gets
                                     max(passes_per_step, ceil(paths_per_step / P)) passes a step, at most 1024, with
                                     staging for that many (0: passes_per_step alone). Small views then make few,
                                     large launches instead of many launch-bound ones. */
} rt_server_options;

typedef struct rt_request_status
{
    int64_t id;
    int32_t state;         /* 0 queued, 1 active, 2 done, 3 failed */
    int32_t reserved;
    int64_t passes_done;
    int64_t passes_total;
    int64_t milestone;     /* the largest power-of-two pass count reached so far */
    double submit_time;    /* seconds on the server's clock */
    double start_time;
    double milestone_time; /* when `milestone` was reached */
    double finish_time;
} rt_request_status;

int rt_server_create(const rt_server_options* options, rt_server** out);
void rt_server_destroy(rt_server* server);
/* Makes a scene resident: copied, its tree built and uploaded once, shared by every request that views it. Its
 * device memory counts against the budget until it is removed and its last request released. */
int rt_server_add_scene(rt_server* server, const rt_scene* scene, int64_t* scene_id);
int rt_server_remove_scene(rt_server* server, int64_t scene_id);
/* Queues a request: a view of a resident scene (a camera of any resolution, with the scene's sampling strategy), its
 * seed and its passes. Returns its id in *id. */
int rt_server_submit_view(rt_server* server, int64_t scene_id, const rt_camera_desc* camera, uint64_t seed,
                          int64_t passes, int64_t* id);
/* Queues a request with a scene of its own (copied and uploaded for it, freed when it is released) */
int rt_server_submit(rt_server* server, const rt_scene* scene, uint64_t seed, int64_t passes, int64_t* id);
/* One scheduling step (continuous batching): admits queued requests as memory and slots free up, enqueues the next step
 * of every active request that has none in flight (up to its next power-of-two milestone, at most its step's passes:
 * passes_per_step, raised for small views by paths_per_step; on one of the server's few streams, so the steps of
 * different requests run side by side on the device), then returns as soon as one step in flight has ended, with the number of requests still active in *active */
int rt_server_step(rt_server* server, int32_t* active);

/* The server's device memory: its budget; what it has reserved (resident scenes, and for each active request what its
 * session may hold: image, staging for one step, stacks, counters); and what is held now (the scenes and the sessions'
 * actual allocations, never more than reserved). No counterpart in the reference. */
typedef struct rt_server_memory
{
    uint64_t budget_bytes;
    uint64_t reserved_bytes;
    uint64_t held_bytes;
} rt_server_memory;
int rt_server_memory_use(const rt_server* server, rt_server_memory* memory);
int rt_server_status(const rt_server* server, int64_t id, rt_request_status* status);
/* The image of a request's complete passes so far (W*H*3 doubles) */
int rt_server_pixels(rt_server* server, int64_t id, double* out, size_t count);
int rt_server_release(rt_server* server, int64_t id);

/* --- Errors and the device ---------------------------------------------------------------------------------------- */

const char* rt_last_error(void);
int rt_abi_version(void);

typedef struct rt_device_info
{
    char name[128];
    char arch[64];
    int32_t compute_units;
    int32_t reserved;
    double memory_gib;
    int32_t runtime_version;
    int32_t driver_version;
} rt_device_info;

/* gpu::query_device (src/gpu/renderer.h:516): describes device 0 and starts the runtime on it */
int rt_query_device(rt_device_info* info);
/* gpu::query_device(wait): the same, and sets how the device's own waits (synchronous copies, whole-device waits)
 * wait for the rest of the process (an rt_wait); call it before anything else starts the runtime */
int rt_query_device_wait(rt_device_info* info, int32_t wait);

/* A device-to-device copy of `bytes` bytes with a plain dwordx4 copy kernel, `repeats` times, and its average kernel
 * seconds: the calibration the HBM counters (FETCH_SIZE, WRITE_SIZE) are scaled by. No counterpart in the reference. */
int rt_copy_probe(uint64_t bytes, int32_t repeats, double* seconds);

#ifdef __cplusplus
}
#endif

#endif
