/*
 * eos_c.h -- C ABI of the MI355X-native Eos render path (libeos_amd.so).
 *
 * Plain C: opaque handles, plain pointers and sizes, status codes.  Every
 * entry point names the reference interface it replaces (paths relative to
 * the reference root).  The C++ API the reference exposes (Scene, Camera,
 * DeviceRenderer, MultiDeviceRenderer, DeviceCameraHandler, createCudaScene)
 * is kept by the shim in include/eos/ on top of these calls.
 *
 * Errors: every call returns an eos_status and, on failure, sets a
 * thread-local message readable with eos_last_error().  Nothing calls exit().
 * Threading: each scene handle owns one HIP stream; calls on the same handle
 * must not overlap; different handles are independent.
 * Memory: input/output pointers are host memory unless a *_memory field says
 * EOS_MEM_DEVICE (then they are device pointers on the current HIP device).
 */
#ifndef EOS_C_H
#define EOS_C_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EOS_ABI_VERSION 7

typedef enum eos_status {
    EOS_OK = 0,
    EOS_ERR_INVALID = 1,   /* bad argument (null pointer, size, index out of range) */
    EOS_ERR_HIP = 2,       /* a HIP runtime call failed */
    EOS_ERR_NOMEM = 3,     /* device allocation failed */
    EOS_ERR_STATE = 4,     /* call order (e.g. render before eos_build_lbvh) */
    EOS_ERR_LIMIT = 5,     /* a fixed capacity was exceeded */
    EOS_ERR_NODEVICE = 6   /* no HIP device visible */
} eos_status;

enum { EOS_MEM_HOST = 0, EOS_MEM_DEVICE = 1 };

/* Camera: same fields and layout as the reference Camera (src/Camera.h:55-63):
 * position, viewing direction, right and up vectors (their lengths set the
 * field of view), width and height in pixels.  Row 0 of the image is the
 * bottom row (src/RayTracer.cpp:135). */
typedef struct eos_camera {
    float position[3];
    float view[3];
    float right[3];
    float up[3];
    int32_t width;
    int32_t height;
} eos_camera;

/* Material (src/Material.h:68-77): diffuse and specular colours (RGBA, A
 * unused), Phong exponent (the reference stores an int; integer values give
 * the reference's result, other values are used as is), reflection weight and
 * flag, transparency flag and refractive index (the Fresnel-weighted second
 * ray of src/RayTracer.cpp:361-408). */
typedef struct eos_material {
    float kd[4];
    float ks[4];
    float shininess;
    float reflectivity;
    int32_t reflective;
    int32_t transparent;
    float refractive_index;
    int32_t pad[3];
} eos_material;

/* Point light (src/LightSource.h:35-36): position (w must be 1) and colour. */
typedef struct eos_light {
    float position[4];
    float color[4];
} eos_light;

/* Scene description (replaces the host Scene handed to createCudaScene,
 * src/cudaStructures.h:98).  Triangles are in world space unless
 * eos_set_transforms gives them local-to-world transformations. */
typedef struct eos_scene_desc {
    int32_t num_triangles;
    const float* vertices;          /* [N][3][3]: v1, v2, v3 */
    const float* normals;           /* [N][3][3]: n1, n2, n3 */
    const int32_t* material_index;  /* [N] */
    int32_t triangle_memory;        /* EOS_MEM_HOST / EOS_MEM_DEVICE for the three arrays */
    int32_t num_materials;
    const eos_material* materials;  /* host */
    int32_t num_lights;             /* at most EOS_MAX_LIGHTS */
    const eos_light* lights;        /* host */
} eos_scene_desc;

#define EOS_MAX_LIGHTS 32
#define EOS_MAX_DEPTH 8
#define EOS_MAX_AA 16

typedef struct eos_scene eos_scene;

typedef struct eos_build_stats {
    double build_ms;          /* GPU time of the LBVH build (Morton, sort, Karras, refit, layout) */
    int32_t num_internal;     /* internal nodes of the binary tree (N - 1) */
    int32_t root_ref;         /* root reference of the binary tree */
    int32_t leaf_max;         /* triangles per leaf after collapsing */
    int32_t width;            /* tree the renderer walks: 2 (binary) or 4 (4-wide) */
    int32_t num_nodes4;       /* 4-wide nodes */
    int32_t root4;            /* root reference of the 4-wide tree */
} eos_build_stats;

typedef struct eos_build_opts {
    int32_t leaf_max;         /* 1..8 triangles per leaf */
    int32_t width;            /* 2: walk the binary LBVH; 4: collapse it two levels per node */
} eos_build_opts;

/* Render options.  Output pointers are optional (NULL = not wanted). */
typedef struct eos_render_opts {
    int32_t depth;            /* reflection bounces (RayTracer::setTracedDepth, src/RayTracer.cpp:47-50) */
    int32_t output_memory;    /* EOS_MEM_HOST / EOS_MEM_DEVICE for every output below */
    float* rgb_out;           /* [H][W][3] colour before quantisation */
    int32_t* hit_out;         /* [H][W] primary hit: original triangle index or -1 */
    int32_t keep_records;     /* keep hit records for eos_render_adjoint */
    int32_t count_visits;     /* run the counting traversal (node visits, triangle tests) */
    int32_t aa_samples;       /* RayTracer::setAASamples (src/RayTracer.cpp:176-266): n x n samples per
                                 pixel on a regular grid, averaged; 0 or 1: one ray through the pixel
                                 corner.  hit_out then holds sample (0, 0); keep_records needs 0 or 1 */
    int32_t launch_times;     /* 1: time the launches of the frame (trace_closest_ms, trace_shadow_ms,
                                 shade_ms, trace_first_ms; eos_render_frames: trace_first_ms); each
                                 timing event between dependent launches costs a few microseconds of
                                 GPU time.  0: render_ms only */
} eos_render_opts;

typedef struct eos_render_stats {
    int64_t primary_rays;
    int64_t shadow_rays;
    int64_t secondary_rays;
    int64_t hits;
    int64_t node_visits;      /* count_visits only */
    int64_t tri_tests;        /* count_visits only */
    double render_ms;         /* GPU time, first to last kernel of the frame */
    double trace_closest_ms;  /* launch_times: summed over the frame's closest-hit launches */
    double trace_shadow_ms;   /* launch_times: summed over the frame's any-hit launches */
    double shade_ms;          /* launch_times: hit setup + Phong + resolve */
    int32_t trace_closest_launches;
    int32_t trace_shadow_launches;
    int64_t closest_visits_first;  /* count_visits: node visits of the primary-ray launch */
    int64_t closest_tests_first;   /* count_visits: triangle tests of the primary-ray launch */
    double trace_first_ms;         /* launch_times: duration of the primary-ray closest-hit launch (with the
                                      tile test before it) */
    int64_t primary_culled;        /* primary rays the tile test resolved as misses (eos_launch_config.cull) */
} eos_render_stats;

/* ---- scene lifetime ----------------------------------------------------- */

/* Replaces createCudaScene (src/cudaStructures.h:98, .cu:483-543): copies the
 * scene to the current device in the build's own layout.  No BVH yet. */
eos_status eos_create_scene(const eos_scene_desc* desc, eos_scene** out);

/* The reference never frees its device scene; this does. */
eos_status eos_destroy_scene(eos_scene* scene);

/* Replaces Scene::flush's BVH build (src/Scene.cpp:204-209, src/BVH.cpp:97-111)
 * with a GPU Morton-code LBVH: radix sort + Karras hierarchy + bottom-up refit.
 * leaf_max (1..8) collapses subtrees of at most that many triangles into one
 * leaf; the renderer walks the 4-wide collapse of the tree.  Synchronous. */
eos_status eos_build_lbvh(eos_scene* scene, int32_t leaf_max, eos_build_stats* stats);
/* Same, choosing the tree width (2 = binary nodes, 4 = 4-wide nodes). */
eos_status eos_build_lbvh_ex(eos_scene* scene, const eos_build_opts* opts, eos_build_stats* stats);

/* The LBVHs of n scenes built together: every stage of eos_build_lbvh_ex is one
 * launch over the concatenated triangles of all scenes (one sort of (scene,
 * Morton code, index), one Karras pass, one refit sequence, one 4-wide
 * collapse), and each scene gets exactly the tree eos_build_lbvh_ex would build
 * for it alone.  For many-scene workloads (eos_render_batch); the reference
 * builds each scene's BVH on the CPU in Scene::flush (src/Scene.cpp:204-209).
 * stats: n entries (build_ms = the whole batch's GPU
 * time) or NULL.  After it, eos_debug_read gives the scene's Morton codes,
 * permutation, nodes and triangle records, not the Karras arrays or boxes.
 * Synchronous; calls must not overlap. */
eos_status eos_build_lbvh_batch(eos_scene* const* scenes, int32_t n, const eos_build_opts* opts,
                                eos_build_stats* stats);

/* Per-triangle local-to-world transformations.  Replaces Scene::addTransformation
 * (src/Scene.cpp:96-105) with Triangle::setTransformationIndex (src/Triangle.h:59),
 * which createCudaScene copies to the device (copyTransformations and the
 * triangles' transformationIndex, src/cudaStructures.cu:220-262,285-401).
 * transforms: [T][16] float, column-major (glm layout, m[4*c + r]), host memory;
 * transform_index: [N] int32 (-1 = world space), memory per index_memory.
 * Each matrix must be invertible; the inverse is computed once, in glm's
 * operation order, and the transpose of the inverse transforms normals, as the
 * reference's TriangleMesh::setTransformation does (src/triangleMesh.cpp:246-250).
 * A ray meets a triangle in the triangle's local space (src/BVH.cpp:359-375);
 * the LBVH is built over world-space vertices.  num_transforms = 0 removes all
 * transformations.  Takes effect at the next eos_build_lbvh, which rendering
 * requires after this call. */
eos_status eos_set_transforms(eos_scene* scene, int32_t num_transforms, const float* transforms,
                              const int32_t* transform_index, int32_t index_memory);

/* Scene::setAmbientRefractiveIndex (src/Scene.cpp:134-136): the refractive
 * index outside transparent materials, n1 or n2 of their Fresnel term
 * (src/RayTracer.cpp:372-384).  Default 1; the reference driver uses 1.0002
 * (REFRACTIVE_INDEX_AIR, src/Material.cpp:25). */
eos_status eos_set_ambient_refractive_index(eos_scene* scene, float index);

/* Parameter updates without a rebuild (materials / lights only). */
eos_status eos_update_materials(eos_scene* scene, const eos_material* materials, int32_t n);
eos_status eos_update_lights(eos_scene* scene, const eos_light* lights, int32_t n);

/* ---- rendering ---------------------------------------------------------- */

/* Replaces the extern "C" launcher rayTrace_MegaKernel (src/kernelWrapper.h:31)
 * and DeviceRenderer::renderCudaSceneToHostBufferMegaKernel
 * (src/DeviceRenderer.cu:61-109), with the CPU RayTracer's semantics
 * (src/RayTracer.cpp:99-174,268-458): primary rays, nearest hit, one shadow ray
 * per light, Phong, reflection up to opts->depth bounces, 8-bit quantisation.
 * rgba_out: [H][W][4] (may be NULL when only rgb_out/hit_out are wanted).
 * Synchronous: returns after the outputs are written. */
eos_status eos_render(eos_scene* scene, const eos_camera* camera, const eos_render_opts* opts,
                      uint8_t* rgba_out, eos_render_stats* stats);

/* Rows [row_begin, row_end) of the camera's frame (row 0 at the bottom), the
 * pixels eos_render gives there, into outputs sized for the band
 * ([row_end - row_begin][W]...).  One GPU's share of a frame split by rows:
 * replaces the per-GPU band of rayTrace_MultiMegaKernel (src/kernelWrapper.h:33,
 * src/kernelInvocations.cu:34-60,697-711; MultiDeviceRenderer gives GPU k the
 * rows from k * (H / numGPUs)).  keep_records must be 0. */
eos_status eos_render_rows(eos_scene* scene, const eos_camera* camera, int32_t row_begin, int32_t row_end,
                           const eos_render_opts* opts, uint8_t* rgba_out, eos_render_stats* stats);

/* n frames of one scene, frame i from cameras[i] into rgba_out[i] (device
 * memory, [H][W][4]), with two frames in flight: frame i + 1's primary rays
 * trace while frame i's shadow and reflection launches finish.  Each image is
 * the one eos_render gives for that camera.  opts: depth, count_visits;
 * output_memory must be EOS_MEM_DEVICE, no rgb_out / hit_out / keep_records /
 * anti-aliasing.  stats: n entries (render_ms = the sequence's GPU time / n;
 * trace_first_ms = that frame's primary-ray launch, while it overlaps the other frame's launches)
 * or NULL.  No counterpart in the reference, whose driver renders one frame
 * per call of its viewer loop (src/main.cpp:514-606).  Synchronous. */
eos_status eos_render_frames(eos_scene* scene, const eos_camera* cameras, int32_t n, const eos_render_opts* opts,
                             uint8_t* const* rgba_out, eos_render_stats* stats);

/* Adjoint of eos_render's pre-quantisation image with hit sets held fixed
 * (no counterpart in the reference).  dl_drgb: [H][W][3] upstream gradient.
 * Gradients are ACCUMULATED into grad_materials [M][7] (kd rgb, ks rgb,
 * shininess) and grad_lights [L][6] (position xyz, colour rgb), float64.
 * Replays the records of the last eos_render with keep_records on this scene
 * and camera, or re-traces them when recompute != 0 or none are kept. */
typedef struct eos_adjoint_opts {
    int32_t depth;
    int32_t memory;           /* EOS_MEM_HOST / EOS_MEM_DEVICE for dl_drgb and gradients */
    int32_t recompute;        /* 1: re-trace instead of replaying stored records */
    int32_t pad;
} eos_adjoint_opts;

typedef struct eos_adjoint_stats {
    double adjoint_ms;        /* GPU time of the adjoint kernels */
    double retrace_ms;        /* GPU time of the re-trace (recompute mode) */
    int64_t records;          /* hit records replayed */
} eos_adjoint_stats;

eos_status eos_render_adjoint(eos_scene* scene, const eos_camera* camera, const eos_adjoint_opts* opts,
                              const float* dl_drgb, double* grad_materials, double* grad_lights,
                              eos_adjoint_stats* stats);

/* Many scenes, one persistent launch per pipeline stage (no counterpart in the
 * reference; MultiDeviceRenderer's per-GPU loop, src/MultiDeviceRenderer.cu:
 * 112-165, is the nearest relative).  Every scene must have an LBVH.
 * rgba_out[i]: [H_i][W_i][4] for scene i, memory per opts->output_memory. */
typedef struct eos_batch_opts {
    int32_t depth;
    int32_t output_memory;
    int32_t use_graph;        /* capture the frame in a hipGraph and replay it */
    int32_t pad;
} eos_batch_opts;

typedef struct eos_batch_stats {
    double render_ms;
    int64_t rays;             /* primary + shadow + secondary over all scenes */
} eos_batch_stats;

eos_status eos_render_batch(eos_scene* const* scenes, const eos_camera* cameras, int32_t n,
                            const eos_batch_opts* opts, uint8_t* const* rgba_out, eos_batch_stats* stats);

/* ---- ray queries ----------------------------------------------------------- */

/* Which query eos_trace_rays answers (src/Scene.h:69-71). */
enum {
    /* Scene::findMinDistanceIntersectionBVH (src/Scene.cpp:145-153, BVH::intersectStackNearest
     * src/BVH.cpp:609-659): the nearest hit at t >= 0 */
    EOS_QUERY_NEAREST = 0,
    /* Scene::shadowRayVisibilityBVH (src/Scene.cpp:155-158, BVH::intersectStackVisibility
     * src/BVH.cpp:661-716): 1 if any triangle is hit at t >= 0, whatever its distance */
    EOS_QUERY_OCCLUDED = 1,
    /* Scene::findMinDistanceIntersectionLinear (src/Scene.cpp:159-201): every triangle in index
     * order, the first with the least t in (0, 999999); its normal is not normalised again after
     * the transformation */
    EOS_QUERY_NEAREST_LINEAR = 2
};

/* One ray's nearest hit: the fields of RayIntersection (src/RayIntersection.h:29-56) in world
 * space, plus what the reference keeps to itself.  48 bytes. */
typedef struct eos_ray_hit {
    float point[3];           /* RayIntersection::getPoint */
    int32_t hit;              /* the query's return value: 1 found, 0 not */
    float normal[3];          /* RayIntersection::getNormal */
    int32_t material;         /* RayIntersection::getMaterialIndex (-1: no hit) */
    float t, u, v;            /* ray parameter and barycentric coordinates of the hit */
    int32_t triangle;         /* original index of the hit triangle (-1: none) */
} eos_ray_hit;

typedef struct eos_trace_opts {
    int32_t mode;             /* EOS_QUERY_* */
    int32_t memory;           /* EOS_MEM_HOST / EOS_MEM_DEVICE for origins, directions and outputs */
} eos_trace_opts;

/* n arbitrary rays against the scene: origins and directions [n][3] (any length; t is in units of
 * the direction).  Nearest modes write hits[n], EOS_QUERY_OCCLUDED writes occluded[n] (1 =
 * occluded); the other output may be NULL.  Replaces the per-ray Scene queries above
 * (src/Scene.h:69-71) with one GPU launch for all rays.  Needs eos_build_lbvh (all modes).
 * Synchronous. */
eos_status eos_trace_rays(eos_scene* scene, int32_t n, const float* origins, const float* directions,
                          const eos_trace_opts* opts, eos_ray_hit* hits, uint8_t* occluded);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): eos_trace_rays with no allocation.  Its scratch (fetch
 * cursors, error flag and, for EOS_QUERY_NEAREST, 16 bytes of hit record per
 * ray) is one fixed device array of the library, 4 KiB + 16 B x
 * EOS_STATIC_MAX_RAYS (16 MiB), held by every GPU the library runs on until the
 * process ends.  n above EOS_STATIC_MAX_RAYS fails with EOS_ERR_LIMIT; rays and
 * outputs must be device memory (opts->memory = EOS_MEM_DEVICE).  Calls on one
 * GPU share the array: calls that overlap on one GPU race.  The scene must have
 * been built (eos_build_lbvh), which is where its own memory is allocated.
 * Results equal eos_trace_rays's.  Synchronous. */
#define EOS_STATIC_MAX_RAYS (1 << 20)
eos_status eos_trace_rays_static(eos_scene* scene, int32_t n, const float* origins, const float* directions,
                                 const eos_trace_opts* opts, eos_ray_hit* hits, uint8_t* occluded);

/* ---- mesh files ------------------------------------------------------------ */

/* Triangles read from a Wavefront OBJ file, in the layout of eos_scene_desc
 * (replaces TriangleMesh::loadFromFile, src/triangleMesh.cpp:79-160, which
 * reads through assimp): positions and per-corner normals, polygons as fans.
 * Corners without a `vn` get the normalized sum of the unit normals of the
 * faces sharing their position.  Arrays are malloc'ed; free with eos_mesh_free. */
typedef struct eos_mesh {
    int32_t num_triangles;
    float* vertices;          /* [N][3][3] */
    float* normals;           /* [N][3][3] */
} eos_mesh;

eos_status eos_mesh_load_obj(const char* path, eos_mesh* out);
eos_status eos_mesh_free(eos_mesh* mesh);

/* ---- tuning, diagnostics ------------------------------------------------ */

/* Traversal launch configuration (the autotuner's knobs). */
typedef struct eos_launch_config {
    int32_t workgroup;        /* 64, 128 or 256 threads */
    int32_t stack_lds;        /* per-lane LDS stack entries: 8, 16 or 32 */
    int32_t waves_per_cu;     /* persistent waves per CU launched */
    int32_t refill;           /* idle lanes that trigger a refill, 1..64 */
    int32_t batch;            /* ray batch: most items a wave claims per fetch, 1..4096 */
    int32_t primary_packets;  /* 1: primary rays walk the 4-wide tree as 8x8-pixel packets, one per wave
                                 (pays off when triangles cover many pixels; 0, the default, walks per ray) */
    int32_t coop;             /* launches walked by lane quads (four lanes per ray, 4-wide tree): 0 never;
                                 1, 2, 3: launches with at most lanes/4, lanes/2, lanes items; 4 always */
    /* the same three knobs for frames that run two in flight (eos_render_frames,
       anti-aliasing samples): each frame's launches share the GPU with the other's */
    int32_t frames_waves_per_cu;
    int32_t frames_refill;
    int32_t frames_batch;
    int32_t group;            /* launches walked by groups of 16 lanes per ray (4-wide tree): 0 never; g > 0:
                                 launches with at most g x lanes/16 items (the short launches after the primary
                                 rays, whose time is their longest ray's chain of steps) */
    int32_t steal;            /* 1: once a traversal wave has claimed its last rays, its idle lanes walk
                                 pending subtrees of the rays still walking (launches that count visits
                                 never do); 0: each ray is walked by one lane to the end */
    int32_t scatter;          /* launches with at most scatter x lanes items claim them in a scattered
                                 order (item i * p mod n, p prime), so neighbouring rays -- often long
                                 together -- go to different waves; 0: never.  Results do not change */
    int32_t group_lanes;      /* widest group walk: 64 (launches with at most group x lanes/64 items walk
                                 each ray with a whole wave, 16 nodes per step) or 16 (never wider) */
    int32_t cull;             /* 1: before a single scene's primary rays (4-wide tree) are traced, a tile
                                 test resolves the 8x8-pixel tiles whose rays all miss the root node's child
                                 boxes as misses, and only the other tiles are traced (frames that count
                                 visits never do); 0: every tile is traced.  Results do not change */
} eos_launch_config;

eos_status eos_set_launch_config(eos_scene* scene, const eos_launch_config* cfg);
eos_status eos_get_launch_config(const eos_scene* scene, eos_launch_config* cfg);

/* Copies internal arrays to host memory for parity checks. */
enum {
    EOS_DBG_MORTON_SORTED = 1,   /* uint32 [N] */
    EOS_DBG_PERM = 2,            /* int32  [N] sorted slot -> original index */
    EOS_DBG_KCHILD = 3,          /* int32  [N-1][2] Karras children (<0: ~leaf) */
    EOS_DBG_KPARENT = 4,         /* int32  [2N-1] */
    EOS_DBG_IBOX = 5,            /* float  [N-1][6] exact internal boxes */
    EOS_DBG_NODES = 6,           /* float  [N-1][16] 64-byte traversal nodes */
    EOS_DBG_TRIS = 7,            /* float  [N][12] sorted triangle records */
    EOS_DBG_MORTON = 8,          /* uint32 [N] unsorted codes */
    EOS_DBG_NODES4 = 9,          /* float  [num_nodes4][32] 128-byte 4-wide nodes */
    EOS_DBG_XFORMS = 10          /* float  [T][36] rows 0-2 of inverse, M and transpose(inverse) */
};
eos_status eos_debug_read(const eos_scene* scene, int32_t what, void* host_dst, int64_t bytes);

/* Thread-local text of the last failure ("" if none). */
const char* eos_last_error(void);

/* ABI version (EOS_ABI_VERSION) and device name of the current device. */
int32_t eos_abi_version(void);
eos_status eos_device_name(char* buf, int32_t len);

#ifdef __cplusplus
}
#endif
#endif
