// eos_internal.h -- device layouts and host bookkeeping shared by the HIP sources.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/eos_c.h"

namespace eos {

// ---------------------------------------------------------------- errors
void set_error(const std::string& msg);
eos_status fail(eos_status st, const std::string& msg);
#define EOS_HIP(call)                                                                       \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess)                                                               \
            return ::eos::fail(e_ == hipErrorOutOfMemory ? EOS_ERR_NOMEM : EOS_ERR_HIP,     \
                               std::string(#call) + ": " + hipGetErrorString(e_));           \
    } while (0)
#define EOS_TRY(expr)                          \
    do {                                       \
        eos_status s_ = (expr);                \
        if (s_ != EOS_OK) return s_;           \
    } while (0)

// ---------------------------------------------------------------- layouts
// 64-byte BVH node: the two child boxes and the two child references.
//   b0 = c0.lo.x c0.hi.x c0.lo.y c0.hi.y
//   b1 = c1.lo.x c1.hi.x c1.lo.y c1.hi.y
//   b2 = c0.lo.z c0.hi.z c1.lo.z c1.hi.z
//   c  = ref0, ref1, 0, 0
// ref >= 0: internal node index; ref < 0: leaf, ~ref = first | (count-1) << 28.
struct alignas(64) Node {
    float4 b0, b1, b2;
    int4 c;
};
static_assert(sizeof(Node) == 64, "node is one 64-byte line");

// 128-byte 4-wide node (two 64-byte lines), children in structure-of-arrays
// form so one 16-byte load brings one bound of all four children:
//   lox, hix, loy, hiy, loz, hiz (float4 each), ref (int4), pad.
// Empty slots: ref kEmpty, bounds +inf (the slab test rejects them).
struct alignas(64) Node4 {
    float4 lox, hix, loy, hiy, loz, hiz;
    int4 ref;
    int4 pad;
};
static_assert(sizeof(Node4) == 128, "4-wide node is two 64-byte lines");
constexpr int32_t kEmpty = 0x7FFFFFFE;

// Triangle records in leaf (Morton) order, 48 bytes each, read as 3 x 16 B:
//   geo:   v1.xyz e1.x | e1.yz e2.xy | e2.z orig(int) 0 0       (traversal)
//   shade: n1.xyz mat  | n2.xyz 0    | n3.xyz 0                 (shading)
constexpr int kLeafShift = 28;
constexpr int32_t kDone = 0x7FFFFFFF;  // traversal sentinel
constexpr int kMaxStack = 64;           // binary: LBVH depth <= 64 (prefix length <= 63)
constexpr int kMaxStack4 = 96;          // 4-wide: <= 32 levels, <= 3 pushes each

struct DevMaterial {
    float kd[4];
    float ks[4];
    float shininess, reflectivity;
    int32_t reflective, transparent;
    float refractive_index;
    int32_t pad[3];
};
struct DevLight {
    float position[4];
    float color[4];
};
struct DevCamera {
    float position[3], view[3], right[3], up[3];
    int32_t width, height;
};
static_assert(sizeof(DevMaterial) == sizeof(eos_material), "material layout");
static_assert(sizeof(DevLight) == sizeof(eos_light), "light layout");
static_assert(sizeof(DevCamera) == sizeof(eos_camera), "camera layout");

// Per-triangle transformation on the device: rows 0..2 of the inverse, of M
// and of transpose(inverse), each row (m[r], m[4+r], m[8+r], m[12+r]) of the
// column-major matrix, so row r of M * (x, y, z, w) is
// (row.x*x + row.y*y) + (row.z*z + row.w*w), glm's order.  Row 3 is never
// needed: only the xyz of a transformed point or vector is used.
constexpr int kXfRows = 9;
constexpr int kXfInv = 0, kXfM = 3, kXfInvT = 6;

// Device view of one scene, passed by value to kernels.
struct SceneView {
    const Node* nodes;
    const Node4* nodes4;
    const float4* geo;
    const float4* shade;
    const DevMaterial* mats;
    const DevLight* lights;
    const float4* xf;         // [n_xf][kXfRows] transformations (null: none)
    int32_t root;
    int32_t root4;
    int32_t width;            // 2 or 4: which tree the traversal walks
    int32_t n_tri;
    int32_t n_light;
    int32_t n_mat;
    int32_t n_xf;
    float ambient;            // refractive index outside transparent materials
};

// ---------------------------------------------------------------- buffers
struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
    eos_status ensure(size_t want);
    void release();
    template <class T> T* as() const { return static_cast<T*>(ptr); }
};

struct LaunchCfg {
    int workgroup = 256;
    int stack_lds = 16;
    int waves_per_cu = 20;
    int refill = 16;
    int batch = 64;        // most items a traversal wave claims per fetch
    int packets = 0;       // primary rays as one 8x8 packet per wave (4-wide tree)
    int coop = 0;          // quad-cooperative walk (0: never; work stealing serves the short launches)
    int group = 1;         // group walk (16 lanes per ray) for launches with at most one item per 16 lanes
    int group_lanes = 64;  // widest group: a whole wave per ray for launches of at most one item per wave
    int steal = 1;         // per-lane walk: idle lanes of a drained wave help its remaining rays
    int scatter = 4;       // launches of at most four items per lane claim them in scattered order
    int cull = 1;          // single-scene frames: primary tiles whose rays all miss the root's child boxes are not traced
    // frames two in flight share the GPU: fewer waves per launch, larger claims
    int frames_waves_per_cu = 16;
    int frames_refill = 16;
    int frames_batch = 64;
};

}  // namespace eos

// Per-frame device buffers of the wavefront pipeline (sized for the largest
// frame rendered so far on this scene).
struct eos_frame_buffers {
    eos::DevBuf hit0;       // float4 [n_ids]      level-0 hit records (t, slot, u, v)
    eos::DevBuf q[2];       // float4 [cap][2]     ray queues (o.xyz pixel | d.xyz 0)
    eos::DevBuf hitq;       // float4 [cap]        hit records of queued rays
    // Per level l, at fixed strides so one launch can address several levels:
    eos::DevBuf surf;       // float4 [D+1][cap][2] hit surfaces (P.xyz pixel | N.xyz mat)
    eos::DevBuf vis;        // uint8  [D+1][cap][L] light visible from the hit surface
    size_t cap = 0;         // surfaces per level (pixels of the frame)
    int lslots = 1;         // light slots per surface
    size_t surf_stride() const { return 2 * cap; }                // float4 per level
    size_t vis_stride() const { return cap * (size_t)lslots; }    // bytes per level
    float4* surf_at(int l) { return surf.as<float4>() + (size_t)l * surf_stride(); }
    uint8_t* vis_at(int l) { return vis.as<uint8_t>() + (size_t)l * vis_stride(); }
    eos::DevBuf levcol;     // float4 [(D+1) * npix]  level colour rgb + reflection weight
    eos::DevBuf levels;     // uint8  [npix]
    eos::DevBuf counters;   // int64  [64]
    eos::DevBuf gstack;     // int32  overflow stacks of persistent lanes
    eos::DevBuf aa_acc;     // float4 [npix]       anti-aliasing sample sums
    // primary tile test, per frame of n_ids primary items (n_ids / 64 tiles):
    // int32 [tiles] tiles left to trace; per workgroup of the test (256 tiles)
    // int32 tiles left / list offset and int32 rays resolved; uint8 [tiles] flags
    eos::DevBuf tiles;
    static constexpr int kTileBlock = 256;   // tiles per workgroup of the tile test
    static size_t tile_blocks(size_t n_ids) { return (n_ids / 64 + kTileBlock - 1) / kTileBlock; }
    static size_t tiles_bytes(size_t n_ids) { return 4 * (n_ids / 64 + 2 * tile_blocks(n_ids)) + n_ids / 64; }
    uint8_t* tile_flags(size_t n_ids) { return tiles.as<uint8_t>() + 4 * (n_ids / 64 + 2 * tile_blocks(n_ids)); }
    // Ray trees (scenes with transparent materials): level l holds at most
    // tcap[l] rays / hits at offset toff[l] of rayhit (hit index per ray; level
    // 0 per pixel), aux (Fresnel weight, children) and hitcol (colour per hit);
    // surf / vis / queues then hold one level at a time.
    bool tree = false;
    size_t tcap[EOS_MAX_DEPTH + 1] = {}, toff[EOS_MAX_DEPTH + 2] = {};
    eos::DevBuf rayhit, aux, hitcol;
    size_t npix = 0, n_ids = 0;
    int depth = -1, n_light = -1;
    int rec_valid = 0;      // records of the last frame kept (adjoint replay)
    eos_camera rec_cam{};
};

struct eos_scene {
    int device = 0;
    hipStream_t stream = nullptr;
    // Second stream of a frame: level l's shadow rays and shading run on it,
    // overlapping level l+1's closest-hit trace on `stream`.  Both streams
    // keep the default priority.
    hipStream_t stream2 = nullptr;
    hipEvent_t sync[2] = {};  // fork to and join from stream2 (no timing)
    int n_tri = 0, n_mat = 0, n_light = 0;
    // original-order inputs (device)
    eos::DevBuf verts, normals, matidx;
    eos::DevBuf mats, lights;
    // LBVH and sorted layouts
    eos::DevBuf nodes, geo, shade;
    eos::DevBuf codes, codes_sorted, perm, kchild, kparent, first_last, ibox, lbox, flags, cbounds;
    eos::DevBuf sort_tmp_k, sort_tmp_v, sort_hist, bounds_part;
    eos::DevBuf nodes4, flags4, scan_tmp;
    // per-triangle transformations: xf [n_xf][kXfRows] float4, xf_index [N]
    // (original order, -1 = none), world-space vertices for the build
    eos::DevBuf xf, xf_index, wverts;
    int n_xf = 0;
    float ambient_index = 1.0f;   // Scene::setAmbientRefractiveIndex
    bool any_transparent = false; // some material is transparent: frames trace ray trees
    uint64_t mat_generation = 0;  // bumped by every eos_update_materials
    int root = 0;
    int root4 = 0;
    int n_nodes4 = 0;
    int width = 4;
    int leaf_max = 1;
    bool built = false;
    bool batch_built = false;     // built by eos_build_lbvh_batch: the build's intermediate arrays are not kept
    eos::LaunchCfg cfg;
    int num_cu = 256;
    eos_frame_buffers fb;
    hipEvent_t ev[96] = {};  // frame timing events
    // Further frame slots of eos_render_frames (kFrameSlots frames in flight;
    // slot 0 is fb / stream / stream2 / sync): frame buffers, two streams at
    // default priority and fork/join events each; the sequence's ordering and
    // timing events; pinned counter copies.
    static constexpr int kFrameSlots = 2;
    eos_frame_buffers fbx[kFrameSlots - 1];
    hipStream_t xstream[kFrameSlots - 1][2] = {};
    hipEvent_t xsync[kFrameSlots - 1][2] = {};
    hipEvent_t resolved[kFrameSlots] = {}, seq_start = nullptr, seq_ev[kFrameSlots] = {};
    void* host_ctr = nullptr;
    size_t host_ctr_bytes = 0;
    std::vector<hipEvent_t> first_ev;   // frames in flight: timing events around each frame's primary launch
    // Ray queries (eos_trace_rays): staged rays and outputs for host memory,
    // hit records, fetch cursors and error flag, overflow stacks.
    eos::DevBuf q_rays, q_out, q_rec, q_ctr, q_gstack;
};

namespace eos {
// Implemented in lbvh.hip
eos_status build_lbvh(eos_scene* s, int leaf_max, int width, eos_build_stats* st);
eos_status build_lbvh_batch(eos_scene* const* scenes, int n, int leaf_max, int width, eos_build_stats* st);
// Implemented in render.hip
// Rows [row0, row1) of the camera's frame (row1 < 0: to the last row).
eos_status render(eos_scene* s, const eos_camera* cam, const eos_render_opts* o, uint8_t* rgba,
                  eos_render_stats* st, int row0 = 0, int row1 = -1);
eos_status render_adjoint(eos_scene* s, const eos_camera* cam, const eos_adjoint_opts* o, const float* dl,
                          double* gm, double* gl, eos_adjoint_stats* st);
eos_status render_batch(eos_scene* const* scenes, const eos_camera* cams, int n, const eos_batch_opts* o,
                        uint8_t* const* out, eos_batch_stats* st);
eos_status render_frames(eos_scene* s, const eos_camera* cams, int n, const eos_render_opts* o, uint8_t* const* rgba,
                         eos_render_stats* st);
// Implemented in query.hip
eos_status trace_rays(eos_scene* s, int n, const float* o, const float* d, const eos_trace_opts* opts,
                      eos_ray_hit* hits, uint8_t* occ);
eos_status trace_rays_static(eos_scene* s, int n, const float* o, const float* d, const eos_trace_opts* opts,
                             eos_ray_hit* hits, uint8_t* occ);
SceneView view_of(const eos_scene* s);
// One anti-aliasing sample frame: offset inside the pixel and what the resolve
// does with the sample's colour (src/RayTracer.cpp:214-242).
enum { AA_NONE = 0, AA_FIRST = 1, AA_NEXT = 2, AA_LAST = 3 };
struct AaSample {
    float ox = 0.0f, oy = 0.0f;   // p * (1/n), q * (1/n)
    int mode = AA_NONE;
    float inv_n2 = 1.0f;          // (1/n)^2
    float4* acc = nullptr;        // [npix] running sums
};
// One frame of the wavefront pipeline (render.hip); outputs may be null.
// launch_times: timing events around every launch (else only around the frame
// and its primary-ray launch).
eos_status run_frame(eos_scene* s, const eos_camera* cam, int depth, bool count, uchar4* rgba_dev, float* rgb_dev,
                     int32_t* hits_dev, eos_render_stats* st, const AaSample& aa = AaSample(), int row0 = 0,
                     int row1 = -1, bool launch_times = true);
}  // namespace eos
