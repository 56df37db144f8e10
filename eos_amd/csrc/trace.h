// trace.h -- device building blocks of the render path: reference arithmetic
// (camera rays, Moller-Trumbore, shading) and the persistent traversal kernel.
//
// Float expressions follow the reference's glm operation order exactly and the
// sources are compiled with -ffp-contract=off, so (apart from powf) every value
// is bit-identical to the CPU RayTracer's:
//   dot(a,b) = (a.x*b.x + a.y*b.y) + a.z*b.z   glm/detail/func_geometric.inl:52-58
//   cross, normalize = x * (1/sqrt(dot(x,x)))   func_geometric.inl:134-158
//   reflect(I,N) = I - N*dot(N,I)*2             func_geometric.inl:170-173
#pragma once
#include <hip/hip_runtime.h>

#include "eos_internal.h"

namespace eos {

struct F3 {
    float x, y, z;
};
__device__ __forceinline__ F3 f3(float x, float y, float z) { return {x, y, z}; }
__device__ __forceinline__ F3 sub(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ F3 add(F3 a, F3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ F3 scale(float s, F3 a) { return {s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ float dot3(F3 a, F3 b) {
    float tx = a.x * b.x, ty = a.y * b.y, tz = a.z * b.z;
    return (tx + ty) + tz;
}
__device__ __forceinline__ F3 cross3(F3 x, F3 y) {
    return {x.y * y.z - y.y * x.z, x.z * y.x - y.z * x.x, x.x * y.y - y.x * x.y};
}
__device__ __forceinline__ F3 normalize3(F3 x) {
    float s = 1.0f / sqrtf(dot3(x, x));
    return {x.x * s, x.y * s, x.z * s};
}
__device__ __forceinline__ F3 reflect3(F3 I, F3 N) {
    float d = dot3(N, I);
    F3 t = {(N.x * d) * 2.0f, (N.y * d) * 2.0f, (N.z * d) * 2.0f};
    return sub(I, t);
}

// src/RayTracer.cpp:119-142: aa = (x - W/2) * (1/(W/2)), bb likewise,
// dir = normalize(aa*right + bb*up + view).  Row 0 is the bottom row.  With
// anti-aliasing (src/RayTracer.cpp:222-230) the sample's offset inside the
// pixel, ox = p * (1/n) and oy = q * (1/n), is added to x and y first; a zero
// offset leaves x and y exact, so the plain frame is the ox = oy = 0 case.
__device__ __forceinline__ F3 camera_dir(const DevCamera& c, int x, int y, float ox, float oy) {
    const float nw = (float)c.width / 2.0f, nh = (float)c.height / 2.0f;
    const float inw = 1.0f / nw, inh = 1.0f / nh;
    const float bb = (((float)y + oy) - nh) * inh;
    const float aa = (((float)x + ox) - nw) * inw;
    const F3 r = {c.right[0], c.right[1], c.right[2]};
    const F3 u = {c.up[0], c.up[1], c.up[2]};
    const F3 v = {c.view[0], c.view[1], c.view[2]};
    return normalize3(add(add(scale(aa, r), scale(bb, u)), v));
}

// Primary rays are numbered in 8x8 pixel tiles (one wave per tile).
__device__ __forceinline__ bool tile_pixel(int r, int tiles_x, int W, int H, int& x, int& y) {
    const int t = r >> 6, i = r & 63;
    const int ty = t / tiles_x, tx = t - ty * tiles_x;
    x = tx * 8 + (i & 7);
    y = ty * 8 + (i >> 3);
    return x < W && y < H;
}

// src/Triangle.cpp:57-91 with e1 = v2 - v1, e2 = v3 - v1 precomputed (the same
// float subtractions, done once at build time).  Branch-free: every quantity is
// computed with the reference's operations and the reference's early-outs
// become one predicate with the same comparisons (so NaN behaves the same).
// Keeping the arithmetic unconditional stops the compiler from sinking the
// load of v1 behind the determinant test, which would cost a second memory
// round trip per triangle.
__device__ __forceinline__ bool moller_trumbore(F3 o, F3 d, F3 v1, F3 e1, F3 e2, float& t, float& u, float& v) {
    const F3 P = cross3(d, e2);
    const float det = dot3(e1, P);
    const float inv = 1.0f / det;
    const F3 T = sub(o, v1);
    const F3 Q = cross3(T, e1);
    t = dot3(e2, Q) * inv;
    u = dot3(T, P) * inv;
    v = dot3(d, Q) * inv;
    const bool rej = (det > -0.00001f && det < 0.00001f) | (t < 0.0f) | (u < 0.0f || u > 1.0f) |
                     (v < 0.0f || u + v > 1.0f);
    return !rej;
}

// Row r of a column-major matrix applied to (p, w): (row.x*x + row.y*y) +
// (row.z*z + row.w*w), glm's mat4 * vec4 (glm/detail/type_mat4x4.inl:571-616).
__device__ __forceinline__ float xrow(float4 r, F3 p, float w) {
    return (r.x * p.x + r.y * p.y) + (r.z * p.z + r.w * w);
}
__device__ __forceinline__ F3 xform3(const float4* X, F3 p, float w) {
    return f3(xrow(X[0], p, w), xrow(X[1], p, w), xrow(X[2], p, w));
}

// A world-space ray in the local space of transformation xi (-1: none): origin
// with w = 1 and direction with w = 0 through the inverse (src/BVH.cpp:362-364).
template <bool XF>
__device__ __forceinline__ void to_local(const float4* xf, int xi, F3& o, F3& d) {
    if (XF && xi >= 0) {
        const float4* X = xf + kXfRows * (size_t)xi + kXfInv;
        o = xform3(X, o, 1.0f);
        d = xform3(X, d, 0.0f);
    }
}

// src/Triangle.cpp:41-47 and src/BVH.cpp:362-381: point o + t*d and normal
// normalize(n1*(1-u-v) + n2*u + n3*v) in the triangle's local space (the ray
// through the inverse), then the point through M and the normal through
// transpose(inverse), normalised again (renorm) as the BVH query does;
// Scene::findMinDistanceIntersectionLinear (src/Scene.cpp:159-201) leaves the
// transformed normal as it is (renorm false).  Without a transformation the
// point is the world one and the second normalisation repeats the first.
// s0..s2: the triangle's shading record (n1.xyz mat | n2.xyz xf | n3.xyz).
__device__ __forceinline__ void hit_surface(float4 s0, float4 s1, float4 s2, const float4* xf, F3 o, F3 d, float t,
                                            float u, float v, bool renorm, F3& P, F3& N) {
    const int xi = __float_as_int(s1.w);
    const float4* X = xi >= 0 ? xf + kXfRows * (size_t)xi : nullptr;
    const F3 lo = X ? xform3(X + kXfInv, o, 1.0f) : o;
    const F3 ld = X ? xform3(X + kXfInv, d, 0.0f) : d;
    P = add(lo, scale(t, ld));
    const float w = (1.0f - u) - v;
    const F3 na = {s0.x * w, s0.y * w, s0.z * w};
    const F3 nb = {s1.x * u, s1.y * u, s1.z * u};
    const F3 nc = {s2.x * v, s2.y * v, s2.z * v};
    N = normalize3(add(add(na, nb), nc));
    if (X) {
        P = xform3(X + kXfM, P, 1.0f);
        N = xform3(X + kXfInvT, N, 0.0f);
    }
    if (renorm) N = normalize3(N);
}

__device__ __forceinline__ float safe_inv(float d) {
    float x = fabsf(d) < 1e-20f ? copysignf(1e-20f, d) : d;
    return 1.0f / x;
}

// Ray/box slab test against [0, tmax], lo*invd - o*invd in one fma per plane.
// Boxes are inflated at build time, so the test is conservative with respect to
// the float Moller-Trumbore test.
__device__ __forceinline__ bool slab(float lox, float hix, float loy, float hiy, float loz, float hiz, float ix,
                                     float iy, float iz, float px, float py, float pz, float tmax, float& tmin) {
    const float x0 = __builtin_fmaf(lox, ix, -px), x1 = __builtin_fmaf(hix, ix, -px);
    const float y0 = __builtin_fmaf(loy, iy, -py), y1 = __builtin_fmaf(hiy, iy, -py);
    const float z0 = __builtin_fmaf(loz, iz, -pz), z1 = __builtin_fmaf(hiz, iz, -pz);
    const float a = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fmaxf(fminf(z0, z1), 0.0f));
    const float b = fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fminf(fmaxf(z0, z1), tmax));
    tmin = a;
    return a <= b;
}

// The same test with each axis's near and far planes chosen by the sign of the
// inverse direction beforehand (lo first where the component is positive):
// fma(x, inv, -p) is monotone in x for a fixed inverse, so the near plane's
// value is min(lo, hi)'s and the far plane's max(lo, hi)'s, and tmin, tmax and
// the answer equal slab()'s bit for bit with six fewer min/max per box.
__device__ __forceinline__ bool slab_nf(float nx, float fx, float ny, float fy, float nz, float fz, float ix,
                                        float iy, float iz, float px, float py, float pz, float tmax, float& tmin) {
    const float x0 = __builtin_fmaf(nx, ix, -px), x1 = __builtin_fmaf(fx, ix, -px);
    const float y0 = __builtin_fmaf(ny, iy, -py), y1 = __builtin_fmaf(fy, iy, -py);
    const float z0 = __builtin_fmaf(nz, iz, -pz), z1 = __builtin_fmaf(fz, iz, -pz);
    const float a = fmaxf(fmaxf(x0, y0), fmaxf(z0, 0.0f));
    const float b = fminf(fminf(x1, y1), fminf(z1, tmax));
    tmin = a;
    return a <= b;
}

__device__ __forceinline__ uint64_t lanemask_lt() {
    const int lane = __lane_id();
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// Position of the r-th set bit (from 0) of m, for r < popcount(m).
__device__ __forceinline__ int nth_set_bit(uint64_t m, int r) {
    int base = 0;
    for (int w = 32; w >= 1; w >>= 1) {
        const uint64_t lo = m & ((1ull << w) - 1);
        const int c = __popcll(lo);
        if (r >= c) {
            r -= c;
            m >>= w;
            base += w;
        } else {
            m = lo;
        }
    }
    return base;
}

// Wave-aggregated append: one atomic per wave, lanes get consecutive slots in
// lane order.  Every lane of the wave must call it.
__device__ __forceinline__ int wave_append(int* ctr, bool pred) {
    const uint64_t m = __ballot(pred);
    if (m == 0) return -1;
    const int leader = __ffsll((unsigned long long)m) - 1;
    int base = 0;
    if ((int)__lane_id() == leader) base = atomicAdd(ctr, __popcll(m));
    base = __shfl(base, leader);
    return pred ? base + __popcll(m & lanemask_lt()) : -1;
}

// Adds the wave's sum of v to *ctr with one atomic.  Every lane must call it.
__device__ __forceinline__ void wave_sum_add(int* ctr, int v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if (__lane_id() == 0 && v != 0) atomicAdd(ctr, v);
}

// Counter slots (int32) of one frame.
enum : int {
    CTR_Q = 0,         // + level: rays queued for level (level >= 1)
    CTR_HITS = 16,     // + level: hits at level
    CTR_ERR = 64,      // stack overflow / invalid state
    CTR_SHADOW = 65,   // + level: shadow rays of level (sum over its hits of the hit scene's light count);
                       // one level's count stays below 2^31, the frame's total is summed on the host in 64 bits
    CTR_LIVE = 120,    // primary tiles left to trace by the tile test (k_cull_tiles)
    CTR_CULLED = 121,  // primary rays the tile test resolved as misses
    CTR_COUNT = 128
};
static_assert(CTR_Q + EOS_MAX_DEPTH + 1 < CTR_HITS && CTR_HITS + EOS_MAX_DEPTH < CTR_ERR &&
                  CTR_SHADOW + EOS_MAX_DEPTH < CTR_LIVE && CTR_CULLED < CTR_COUNT,
              "per-level counters overlap");
// int64 statistics (count_visits): [2*launch] visits, [2*launch+1] tests;
// launch = level (closest) or 16 + level (any hit).
constexpr int kStats = 64;
// Fetch cursors, after the statistics: per launch (same numbering) one cursor
// per item range, each on its own 128-byte line.  A launch's items are split
// into kRanges contiguous ranges, range r served first to the waves of
// workgroups with blockIdx % kRanges == r: workgroups are dispatched to the
// eight XCDs round robin, so each XCD walks one band of the item space (its own
// L2 holds that band's nodes) and the chunk claims spread over eight counters.
constexpr int kRanges = 8;
constexpr int kCursorStride = 32;                         // ints per cursor (128 B)
constexpr int kCursorBase = CTR_COUNT + 2 * kStats;       // int offset of the cursors
constexpr int kCursorInts = 32 * kRanges * kCursorStride; // 32 launches
constexpr int kCounterInts = kCursorBase + kCursorInts;

// MODE_RAYS / MODE_RAYS_ANY: caller-given rays (origins and directions, [n][3]
// floats each), nearest hit or occlusion, for the ray queries.
enum TraceMode : int { MODE_PRIMARY = 0, MODE_QUEUE = 1, MODE_SHADOW = 2, MODE_RAYS = 3, MODE_RAYS_ANY = 4 };

// Any-hit modes: a ray ends at its first hit, whatever its distance.
__host__ __device__ constexpr bool any_hit_mode(int mode) { return mode == MODE_SHADOW || mode == MODE_RAYS_ANY; }

// One scene of a batched frame (device table, one entry per scene).
struct BatchScene {
    SceneView sc;
    DevCamera cam;
    int tiles_x;
    int id_off;      // first primary ray id of this scene
    int pix_off;     // first pixel of this scene in the concatenated frame
    int n_ids;
};

// Hit surfaces carry the material index and, in a batch, the scene index.
constexpr int kSceneShift = 20;
__device__ __forceinline__ int pack_mat(int mat, int scene) { return mat | (scene << kSceneShift); }
__device__ __forceinline__ int unpack_mat(int v) { return v & ((1 << kSceneShift) - 1); }
__device__ __forceinline__ int unpack_scene(int v) { return (int)((unsigned)v >> kSceneShift); }

// Scene owning primary ray id r: the last entry with id_off <= r.
__device__ __forceinline__ int scene_of_id(const BatchScene* t, int n, int r) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (t[mid].id_off <= r) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

struct TraceArgs {
    SceneView sc;
    DevCamera cam;
    float aa_ox, aa_oy;      // anti-aliasing sample offset inside the pixel (0: pixel corner)
    int row0, band_h;        // primary rays cover camera rows [row0, row0 + band_h)
    int tiles_x;
    const int32_t* live;     // MODE_PRIMARY, single scene: tiles left by the tile test (item i is pixel
                             // i % 64 of tile live[i / 64]); null: item i is pixel i % 64 of tile i / 64
    const BatchScene* table;  // batched frames
    int n_scenes;
    int n_static;            // item count when count_ptr is null
    const int* count_ptr;    // device item count (queued rays / hits)
    int count_mul;           // items per counted entry (light slots for shadow rays)
    int* cursor;             // fetch cursors of the launch: range r at cursor[r * kCursorStride]
    const float4* q;         // MODE_QUEUE: ray queue
    const float* ray_o;      // MODE_RAYS(_ANY): origins [n][3]
    const float* ray_d;      // MODE_RAYS(_ANY): directions [n][3]
    float4* hit_out;         // closest: (t, slot, u, v)
    const float4* surf;      // MODE_SHADOW: hit surfaces (P.xyz pixel | N.xyz mat) of the first level
    uint8_t* vis_out;        // MODE_SHADOW: 1 = light visible, first level; MODE_RAYS_ANY: 1 = occluded
    int nlev;                // MODE_SHADOW: consecutive levels in this launch (count_ptr[0..nlev))
    long long surf_stride;   // MODE_SHADOW: float4 between levels' surfaces
    long long vis_stride;    // MODE_SHADOW: bytes between levels' visibility
    int32_t* gstack;         // overflow stacks
    int* err;
    unsigned long long* stats;
    int refill;
    int batch;               // most items a wave claims per fetch
    int coop;                // quad-cooperative walk: 0 never, 1..3 at most lanes/4, /2, /1 items, 4 always
    int group;               // group walk (16 lanes per ray) for launches of at most group x lanes/16 items (0: never)
    int group_lanes;         // 64: launches of at most group x lanes/64 items walk each ray with a whole wave
    int steal;               // per-lane walk: idle lanes of a drained wave take pending subtrees of its rays
    int scatter;             // launches of at most scatter x lanes items claim them in scattered order (0: never)
};

// Compare-exchange on (entry distance, slot): a total order, so the result does
// not depend on the network's shape.
__device__ __forceinline__ void cmpswap(float& ta, int& ka, int& ra, float& tb, int& kb, int& rb) {
    const bool sw = (tb < ta) | ((tb == ta) & (kb < ka));  // selects, no branch
    const float t0 = sw ? tb : ta, t1 = sw ? ta : tb;
    const int k0 = sw ? kb : ka, k1 = sw ? ka : kb;
    const int r0 = sw ? rb : ra, r1 = sw ? ra : rb;
    ta = t0; tb = t1;
    ka = k0; kb = k1;
    ra = r0; rb = r1;
}

template <int WIDTH>
constexpr int max_stack() { return WIDTH == 4 ? kMaxStack4 : kMaxStack; }

// Wave-uniform state of the two-level work fetch: a wave claims a chunk of
// items from a range cursor with one atomic and hands them out to idle lanes
// (or lane groups) locally.  It starts on its XCD's range and moves on to the
// next range when that one is used up; after every range it is drained.
struct Fetch {
    bool drained = false;   // every range is used up
    int batch = 64;         // most items per claim
    int next = 0, end = 0;  // unissued items of this wave's chunk
    int seen = 0;           // cursor value of the current range after this wave's last claim
    int range = 0;          // current range
    int tried = 0;          // ranges found used up
};

// A primary launch left with under half of its frame's items by the tile test
// traces the tiles around the objects' outlines, where neighbouring rays'
// costs differ most: it claims at most kCulledBatch items at a time (config C,
// 35% of its tiles left: single frame 2.12 -> 1.96 ms, frames in flight
// 1.56 -> 1.50; the same cap on the glass frame, 71% left, cost it 3%).
constexpr int kCulledBatch = 32;

__device__ __forceinline__ Fetch fetch_start(const TraceArgs& a, int n) {
    Fetch f;
    f.range = (int)(blockIdx.x % kRanges);
    f.batch = a.batch;
    if (a.live != nullptr && 2LL * n < (long long)a.n_static && f.batch > kCulledBatch) f.batch = kCulledBatch;
    return f;
}

// Hands out items to `nidle` idle units (lanes or lane groups) of the wave;
// returns the item of the unit ranked `rank` among them (-1: none).  Every lane
// of the wave calls it with the same fetch state, idle mask and count.  Chunks
// follow guided self-scheduling within the range: remaining / (2 * waves per
// range), between the idle count and the launch's ray batch.
__device__ __forceinline__ int fetch_items(Fetch& f, const TraceArgs& a, int n, int waves_total, uint64_t idle,
                                           int nidle, int rank) {
    const int avail = f.end - f.next;
    const int old_next = f.next;
    int base = 0, take = 0;
    if (avail < nidle && !f.drained) {
        const int want = nidle - avail;
        const int leader = __ffsll((unsigned long long)idle) - 1;
        const int wpr = waves_total / kRanges > 0 ? waves_total / kRanges : 1;
        for (int tries = 0; tries < kRanges; ++tries) {  // wave-uniform
            const int r = f.range;
            const int lo = (int)((long long)n * r / kRanges), hi = (int)((long long)n * (r + 1) / kRanges);
            const int len = hi - lo;
            int size = (len - f.seen) / (2 * wpr);
            size = size > f.batch ? f.batch : size;
            size = size < want ? want : size;
            int got = 0;
            if ((int)__lane_id() == leader) got = atomicAdd(a.cursor + r * kCursorStride, size);
            got = __shfl(got, leader);
            if (got < len) {
                f.seen = got + size;
                base = lo + got;
                take = (got + size < len ? size : len - got);
                break;
            }
            f.range = (r + 1) % kRanges;
            f.seen = 0;
            if (++f.tried == kRanges) {
                f.drained = true;
                break;
            }
        }
    }
    int my = -1;
    if (rank < avail) my = old_next + rank;
    else if (rank - avail < take) my = base + (rank - avail);
    if (avail >= nidle) {
        f.next = old_next + nidle;
    } else {
        const int used = nidle - avail < take ? nidle - avail : take;
        f.next = take > used ? base + used : 0;
        f.end = take > used ? base + take : 0;
    }
    return my;
}

// Scattered claim order of a short launch.  Items come in blocks of b (a
// shadow launch: b = its light slots, so a block is b neighbouring surfaces'
// rays toward one light; blocks of 16 or 64 made config C's single frame 6% and
// 12% slower); claim i of block k is item (k * p mod m) * b + i mod b for the m
// whole blocks (the items after them keep their place), with p the first
// integer from floor(m / golden ratio) up that is coprime to m: a bijection on [0, n) whose
// neighbouring claims land far apart and evenly spread (Fibonacci hashing).
// Items next to each other -- often all long, as where shadow rays graze a
// surface -- so go to different waves, each wave holds few long rays and its
// idle lanes can share them out.  The item number keeps naming the ray and its
// outputs; only the order of claims changes.
__device__ __forceinline__ unsigned gcd_u32(unsigned a, unsigned b) {  // binary (Stein) gcd
    if (a == 0) return b;
    if (b == 0) return a;
    const int sh = __builtin_ctz(a | b);
    a >>= __builtin_ctz(a);
    do {
        b >>= __builtin_ctz(b);
        if (a > b) {
            const unsigned t = a;
            a = b;
            b = t;
        }
        b -= a;
    } while (b != 0);
    return a << sh;
}
struct Scatter {
    bool on = false;
    int b = 1, m = 0;
    long long p = 1;
    __device__ __forceinline__ void init(int rounds, int n, long long lanes, int block) {
        on = rounds > 0 && n > 1 && (long long)n <= lanes * (long long)rounds;
        b = block > 1 ? block : 1;
        m = n / b;
        unsigned q = (unsigned)((double)m * 0.6180339887498949);
        q = q < 1 ? 1 : q;
        while (m > 1 && gcd_u32(q, (unsigned)m) != 1) ++q;   // ends at m - 1 at the latest
        p = q;
    }
    __device__ __forceinline__ int map(int i) const {
        if (!on || i < 0 || i >= m * b) return i;
        const int k = i / b, w = i - k * b;
        return (int)(((long long)k * p) % m) * b + w;
    }
};

// Any-hit rays sent into their own surface walk the farthest child first (see
// setup_item); a switch for measuring it.
#ifndef EOS_FAR_FIRST
#define EOS_FAR_FIRST 1
#endif
constexpr bool kFarFirst = EOS_FAR_FIRST != 0;

// Order of a shadow launch's items within a level: tiles of kLightTile hit
// surfaces, light by light inside a tile (item = tile, light slot, surface in
// the tile), so that lanes refilled together trace neighbouring surfaces toward
// one light.  Measured against surface by surface (the light slots of a surface
// next to each other): hit-dense 4K frame 16.5 -> 14.65 ms, glass frames in
// flight 3.66 -> 3.51 ms, config D's batched frame 33.9 -> 33.1 ms; whole
// levels light by light (no tiles) cost config E's frame 5%.  Launches whose
// claims are scattered (Scatter) keep surface by surface: a block then holds
// one surface's light slots, whose walks share their first nodes (config C's
// level-0 shadow launch alone: 0.38 ms, against 0.44 ms for blocks of four
// surfaces toward one light).  0: surface by surface always (a switch for
// measuring it).
#ifndef EOS_LIGHT_TILE
#define EOS_LIGHT_TILE 32
#endif
constexpr int kLightTile = EOS_LIGHT_TILE;

// One traversal item, set up: the ray and the scene it runs in.
struct RayItem {
    F3 o, d;
    bool valid;              // false: no ray (pixel outside the frame, light slot beyond the scene's lights)
    int root, max_visits;
    const Node* nodes;
    const Node4* nodes4;
    const float4* geo;
    const float4* xf;        // transformations of the scene (null: none)
    size_t out_idx;          // any hit: result byte of the item
    bool far_first;          // shadow ray into its own surface: walk the farthest child first
};

// Item of a primary launch whose tiles went through the tile test: the launch
// numbers only the tiles left (TraceArgs::live); the ray keeps its pixel's
// number, which names its hit record.
template <int MODE, bool BATCH>
__device__ __forceinline__ int primary_tile_item(const TraceArgs& a, int my) {
    if (MODE != MODE_PRIMARY || BATCH || a.live == nullptr || my < 0) return my;
    return (a.live[my >> 6] << 6) | (my & 63);
}

// light_tiles: a shadow launch's items come in light tiles (kLightTile), else
// surface by surface
template <int MODE, bool BATCH, int WIDTH>
__device__ __forceinline__ void setup_item(const TraceArgs& a, int my, RayItem& r, bool light_tiles) {
    r.valid = true;
    r.nodes = a.sc.nodes;
    r.nodes4 = a.sc.nodes4;
    r.geo = a.sc.geo;
    r.xf = a.sc.xf;
    r.root = WIDTH == 4 ? a.sc.root4 : a.sc.root;
    r.max_visits = a.sc.n_tri + 64;
    r.out_idx = 0;
    r.far_first = false;
    auto use_scene = [&](const BatchScene& B) {
        r.nodes = B.sc.nodes;
        r.nodes4 = B.sc.nodes4;
        r.geo = B.sc.geo;
        r.xf = B.sc.xf;
        r.root = WIDTH == 4 ? B.sc.root4 : B.sc.root;
        r.max_visits = B.sc.n_tri + 64;
    };
    if (MODE == MODE_PRIMARY) {
        int x, y;
        if (BATCH) {
            const BatchScene& B = a.table[scene_of_id(a.table, a.n_scenes, my)];
            use_scene(B);
            r.valid = tile_pixel(my - B.id_off, B.tiles_x, B.cam.width, B.cam.height, x, y);
            r.o = f3(B.cam.position[0], B.cam.position[1], B.cam.position[2]);
            r.d = camera_dir(B.cam, x, y, 0.0f, 0.0f);
        } else {
            r.valid = tile_pixel(my, a.tiles_x, a.cam.width, a.band_h, x, y);
            r.o = f3(a.cam.position[0], a.cam.position[1], a.cam.position[2]);
            r.d = camera_dir(a.cam, x, y + a.row0, a.aa_ox, a.aa_oy);
        }
    } else if (MODE == MODE_QUEUE) {
        const float4 q0 = a.q[2 * (size_t)my], q1 = a.q[2 * (size_t)my + 1];
        r.o = f3(q0.x, q0.y, q0.z);
        r.d = f3(q1.x, q1.y, q1.z);
        if (BATCH) use_scene(a.table[__float_as_int(q1.w)]);
    } else if (MODE == MODE_RAYS || MODE == MODE_RAYS_ANY) {
        const float* po = a.ray_o + 3 * (size_t)my;
        const float* pd = a.ray_d + 3 * (size_t)my;
        r.o = f3(po[0], po[1], po[2]);
        r.d = f3(pd[0], pd[1], pd[2]);
        r.out_idx = (size_t)my;
    } else {
        // item -> (level, surface j, light slot li)
        int rem = my, lev = 0;
        for (int k = 0; k + 1 < a.nlev; ++k) {
            const int c = a.count_ptr[k] * a.count_mul;
            if (rem < c) break;
            rem -= c;
            lev = k + 1;
        }
        const float4* surf = a.surf + (size_t)lev * (size_t)a.surf_stride;
        int j, li;
        if (kLightTile > 0 && light_tiles) {
            // tile t holds surfaces [t * kLightTile, +hb) (hb < kLightTile in the
            // level's last tile), its items light slot by light slot
            const int nh = a.count_ptr[lev], L = a.count_mul;
            const int t = rem / (kLightTile * L), w = rem - t * (kLightTile * L);
            const int h0 = t * kLightTile;
            const int hb = nh - h0 < kLightTile ? nh - h0 : kLightTile;
            li = w / hb;
            j = h0 + (w - li * hb);
        } else {
            j = rem / a.count_mul;
            li = rem - j * a.count_mul;
        }
        // the result byte keeps the surface-by-surface layout the shading reads
        r.out_idx = (size_t)lev * (size_t)a.vis_stride + (size_t)j * (size_t)a.count_mul + (size_t)li;
        const float4 s0 = surf[2 * (size_t)j];
        const F3 P = f3(s0.x, s0.y, s0.z);
        const DevLight* lights = a.sc.lights;
        int nl = a.sc.n_light;
        if (BATCH) {
            const BatchScene& B = a.table[unpack_scene(__float_as_int(surf[2 * (size_t)j + 1].w))];
            use_scene(B);
            lights = B.sc.lights;
            nl = B.sc.n_light;
        }
        if (li < nl) {
            const DevLight& L = lights[li];
            r.d = normalize3(sub(f3(L.position[0], L.position[1], L.position[2]), P));
            r.o = add(P, scale(1e-3f, r.d));
            // a light behind the surface sends the ray into the object: its
            // occluder, if any, is most likely far away (the far wall), so that
            // ray walks the farthest child first (any order gives the same answer)
            const float4 s1 = surf[2 * (size_t)j + 1];
            r.far_first = kFarFirst && dot3(f3(s1.x, s1.y, s1.z), r.d) < 0.0f;
        } else {
            r.valid = false;
            r.d = f3(0.0f, 0.0f, 1.0f);
            r.o = P;
        }
    }
}

// Items per launch at or below which the cooperative walk is used, as a
// multiple of the launch's lanes / 4: a launch with at most one ray per lane
// is one round, and its time is the longest ray's chain of steps.
__device__ __forceinline__ bool use_coop(int mode, long long n, long long lanes) {
    switch (mode) {
        case 1: return 4 * n <= lanes;
        case 2: return 2 * n <= lanes;
        case 3: return n <= lanes;
        case 4: return true;
        default: return false;
    }
}

// Cooperative walk of the 4-wide tree: four lanes (a quad) share one ray.  At a
// node lane k tests child k; the quad ranks the hit children on (entry
// distance, slot) with quad shuffles, pushes ranks 1.. farthest first onto the
// quad's stack in LDS and continues with rank 0.  At a leaf lane k tests
// triangles k and k + 4 of the leaf, and the quad takes the least (t, original
// index) hit (closest) or the first hit in slot order (any hit).  Every ray
// visits the same nodes in the same order as the per-lane walk, ends with the
// same hit and is counted the same way; what changes is the length of the chain
// of dependent instructions per step, which is what bounds a launch that is one
// round of rays.  XF: triangles may carry transformations (tested in local space).
template <int MODE, int WG, bool COUNT, bool BATCH, bool XF>
__device__ __forceinline__ void trace_coop(const TraceArgs& a, int n, int32_t* wave_stack, unsigned long long& visits,
                                           unsigned long long& tests) {
    constexpr bool ANY = any_hit_mode(MODE);
    const int lane = __lane_id();
    const int sub = lane & 3, grp = lane >> 2;
    const int gbit = 4 * grp;
    int32_t* stk = wave_stack + grp * kMaxStack4;
    const int waves_total = (int)gridDim.x * (WG / 64);
    Scatter scat;
    scat.init(a.scatter, n, (long long)gridDim.x * WG, MODE == MODE_SHADOW ? a.count_mul : 1);
    int gthresh = (a.refill + 3) / 4;
    gthresh = gthresh < 1 ? 1 : (gthresh > 16 ? 16 : gthresh);
    Fetch f = fetch_start(a, n);
    int item = -1;
    size_t out_idx = 0;
    F3 o = {0, 0, 0}, d = {0, 0, 0};
    float ix = 0, iy = 0, iz = 0, px = 0, py = 0, pz = 0;
    float best_t = INFINITY, bu = 0.0f, bv = 0.0f;
    int best_slot = -1, best_orig = 0x7FFFFFFF;
    int cur = kDone, sp = 0, budget = 0;
    const Node4* nodes4 = a.sc.nodes4;
    const float4* geo = a.sc.geo;
    const float4* xf = a.sc.xf;
    for (;;) {
        // one bit per idle quad, at its lane 0
        const uint64_t idle = __ballot(sub == 0 && item < 0);
        const int nidle = __popcll(idle);
        if (nidle >= gthresh && (!f.drained || f.next < f.end)) {
            const int rank = __popcll(idle & ((1ull << gbit) - 1));
            int my = fetch_items(f, a, n, waves_total, idle, nidle, rank);
            my = primary_tile_item<MODE, BATCH>(a, scat.map(my));
            if (item < 0 && my >= 0) {
                RayItem r;
                setup_item<MODE, BATCH, 4>(a, my, r, !scat.on);
                item = my;
                out_idx = r.out_idx;
                o = r.o;
                d = r.d;
                nodes4 = r.nodes4;
                geo = r.geo;
                xf = r.xf;
                ix = safe_inv(d.x);
                iy = safe_inv(d.y);
                iz = safe_inv(d.z);
                px = o.x * ix;
                py = o.y * iy;
                pz = o.z * iz;
                best_t = INFINITY;
                best_slot = -1;
                best_orig = 0x7FFFFFFF;
                bu = bv = 0.0f;
                sp = 0;
                budget = r.max_visits;
                cur = r.valid ? r.root : kDone;
            }
        }
        if (__ballot(item >= 0) == 0) {
            if (f.drained && f.next >= f.end) break;
            continue;
        }
        const int need = (f.drained && f.next >= f.end) ? 16 : gthresh;
        for (;;) {
            if (item >= 0) {  // uniform within a quad from here on
                if (cur >= 0 && cur != kDone && --budget < 0) {
                    if (sub == 0) atomicOr(a.err, 2);
                    cur = kDone;
                } else if (cur >= 0 && cur != kDone) {
                    if (COUNT && sub == 0) ++visits;
                    const float* nd = reinterpret_cast<const float*>(nodes4 + cur);
                    const float lx = nd[sub], hx = nd[4 + sub], ly = nd[8 + sub], hy = nd[12 + sub],
                                lz = nd[16 + sub], hz = nd[20 + sub];
                    const int rf = __float_as_int(nd[24 + sub]);
                    const float tmax = ANY ? INFINITY : best_t;
                    float t;
                    const bool h = slab(lx, hx, ly, hy, lz, hz, ix, iy, iz, px, py, pz, tmax, t) & (rf != kEmpty);
                    const float key = h ? t : INFINITY;
                    const float k1 = __shfl_xor(key, 1), k2 = __shfl_xor(key, 2), k3 = __shfl_xor(key, 3);
                    // rank on (entry distance, slot); misses rank after every hit
                    const int rank = (k1 < key || (k1 == key && (sub ^ 1) < sub)) +
                                     (k2 < key || (k2 == key && (sub ^ 2) < sub)) +
                                     (k3 < key || (k3 == key && (sub ^ 3) < sub));
                    const int hm = (int)((__ballot(h) >> gbit) & 0xFull);
                    const int nh = __popc(hm);
                    if (nh == 0) {
                        cur = sp == 0 ? kDone : stk[--sp];
                    } else if (sp + nh - 1 > kMaxStack4) {
                        if (sub == 0) atomicOr(a.err, 1);
                        cur = kDone;
                    } else {
                        if (h && rank >= 1) stk[sp + (nh - 1 - rank)] = rf;
                        const int src = __ffs((int)((__ballot(h && rank == 0) >> gbit) & 0xFull)) - 1;
                        cur = __shfl(rf, (grp << 2) + src);
                        sp += nh - 1;
                    }
                } else if (cur != kDone) {
                    const int e = ~cur;
                    const int first = e & ((1 << kLeafShift) - 1);
                    const int cnt = ((e >> kLeafShift) & 7) + 1;
                    float lt = INFINITY, lu = 0.0f, lv = 0.0f;
                    int lo_ = 0x7FFFFFFF, ls = -1, lfirst = 8;
                    for (int b = 0; b < cnt; b += 4) {
                        const int k = b + sub;
                        const bool act = k < cnt;
                        const int sl = first + (act ? k : 0);
                        const float4 g0 = geo[3 * (size_t)sl], g1 = geo[3 * (size_t)sl + 1], g2 = geo[3 * (size_t)sl + 2];
                        float tt, uu, vv;
                        F3 lo = o, ld = d;
                        to_local<XF>(xf, __float_as_int(g2.z), lo, ld);
                        const bool hit = moller_trumbore(lo, ld, f3(g0.x, g0.y, g0.z), f3(g0.w, g1.x, g1.y),
                                                         f3(g1.z, g1.w, g2.x), tt, uu, vv) & act;
                        if (ANY) {
                            if (hit && k < lfirst) lfirst = k;
                        } else {
                            const int og = __float_as_int(g2.y);
                            if (hit && (tt < lt || (tt == lt && og < lo_))) {
                                lt = tt;
                                lo_ = og;
                                ls = sl;
                                lu = uu;
                                lv = vv;
                            }
                        }
                    }
                    if (ANY) {
                        int fh = lfirst;
                        fh = min(fh, __shfl_xor(fh, 1));
                        fh = min(fh, __shfl_xor(fh, 2));
                        // counted as the one-at-a-time walk would: it stops at its first hit
                        if (COUNT && sub == 0) tests += fh < 8 ? fh + 1 : cnt;
                        if (fh < 8) {
                            best_slot = 0;  // marks "hit" for the any-hit result
                            cur = kDone;
                        } else {
                            cur = sp == 0 ? kDone : stk[--sp];
                        }
                    } else {
                        if (COUNT && sub == 0) tests += cnt;
                        // least (t, original index) over the quad, then against the ray's best
                        for (int off = 1; off <= 2; off <<= 1) {
                            const float ot = __shfl_xor(lt, off), ou = __shfl_xor(lu, off), ov = __shfl_xor(lv, off);
                            const int oo = __shfl_xor(lo_, off), os = __shfl_xor(ls, off);
                            if (ot < lt || (ot == lt && oo < lo_)) {
                                lt = ot;
                                lo_ = oo;
                                ls = os;
                                lu = ou;
                                lv = ov;
                            }
                        }
                        if (ls >= 0 && (lt < best_t || (lt == best_t && lo_ < best_orig))) {
                            best_t = lt;
                            best_orig = lo_;
                            best_slot = ls;
                            bu = lu;
                            bv = lv;
                        }
                        cur = sp == 0 ? kDone : stk[--sp];
                    }
                }
                if (cur == kDone) {
                    if (sub == 0) {
                        if (ANY) a.vis_out[out_idx] = (best_slot >= 0) == (MODE == MODE_RAYS_ANY) ? 1 : 0;
                        else a.hit_out[item] = make_float4(best_t, __int_as_float(best_slot), bu, bv);
                    }
                    item = -1;
                }
            }
            if ((int)__popcll(__ballot(sub == 0 && item < 0)) >= need) break;
        }
    }
}

// Group walk of the 4-wide tree: kGroupLanes (16) lanes share one ray and its
// stack of pending nodes in LDS.  Every step pops up to four entries from the
// top and quad q takes entry q: at an internal node lane k tests child k and
// the quad pushes its hits farthest first, at a leaf lane k tests triangles k
// and k + 4.  The group's least (t, original index) hit, or its any-hit flag,
// is reduced over the 16 lanes after every step, so four nodes cost one
// dependent step.  It visits nodes in another order than the per-lane walk (so
// launches that count visits never take it), but every node whose box the ray
// meets before its nearest hit so far is visited, so the result is the same
// exact nearest hit, ties to the least original index, or the same occlusion
// flag, bit for bit.  For launches with few rays, whose time is the longest
// ray's chain of dependent steps.
constexpr int kGroupLanes = 16;
constexpr int kWaveGroupStack = 2048;   // stack entries per wave, shared by its groups: 512 per group of
                                        // 16 lanes (at most 16 pushes per step), 2048 for a wave-wide
                                        // group (at most 64); overflow is an error

// Items per launch at or below which the group walk is used, as a multiple of
// the launch's lanes / g (1: one group per ray in one round; 4: four rounds...).
__device__ __forceinline__ bool use_group(int rounds, long long n, long long lanes, int g) {
    return rounds > 0 && n * g <= lanes * (long long)rounds;
}

template <int MODE, int WG, bool BATCH, bool XF, int G>
__device__ __forceinline__ void trace_group(const TraceArgs& a, int n, int32_t* wave_stack) {
    constexpr bool ANY = any_hit_mode(MODE);
    static_assert(G == 16 || G == 64, "groups of 16 lanes or whole waves");
    constexpr int kGroups = 64 / G;
    constexpr int kQuads = G / 4;
    constexpr int kStack = kWaveGroupStack / kGroups;   // entries per group
    const int lane = __lane_id();
    const int gsub = lane & (G - 1), q = gsub >> 2, k = lane & 3;
    const int grp = lane / G;
    const int gbit = G * grp;
    const uint64_t gmask = G == 64 ? ~0ull : (0xFFFFull << gbit);
    const int qbit = lane & ~3;                     // first lane of this lane's quad
    int32_t* stk = wave_stack + grp * kStack;
    const int waves_total = (int)gridDim.x * (WG / 64);
    Scatter scat;
    scat.init(a.scatter, n, (long long)gridDim.x * WG, MODE == MODE_SHADOW ? a.count_mul : 1);
    Fetch f = fetch_start(a, n);
    int item = -1;
    size_t out_idx = 0;
    F3 o = {0, 0, 0}, d = {0, 0, 0};
    float ix = 0, iy = 0, iz = 0, px = 0, py = 0, pz = 0;
    float best_t = INFINITY, bu = 0.0f, bv = 0.0f;
    int best_slot = -1, best_orig = 0x7FFFFFFF;
    int sp = 0, budget = 0;
    const Node4* nodes4 = a.sc.nodes4;
    const float4* geo = a.sc.geo;
    const float4* xf = a.sc.xf;
    for (;;) {
        // one bit per idle group, at its first lane
        const uint64_t idle = __ballot(gsub == 0 && item < 0);
        const int nidle = __popcll(idle);
        if (nidle >= 1 && (!f.drained || f.next < f.end)) {
            const int rank = __popcll(idle & ((1ull << gbit) - 1));
            int my = fetch_items(f, a, n, waves_total, idle, nidle, rank);
            my = primary_tile_item<MODE, BATCH>(a, scat.map(my));
            if (item < 0 && my >= 0) {
                RayItem r;
                setup_item<MODE, BATCH, 4>(a, my, r, !scat.on);
                item = my;
                out_idx = r.out_idx;
                o = r.o;
                d = r.d;
                nodes4 = r.nodes4;
                geo = r.geo;
                xf = r.xf;
                ix = safe_inv(d.x);
                iy = safe_inv(d.y);
                iz = safe_inv(d.z);
                px = o.x * ix;
                py = o.y * iy;
                pz = o.z * iz;
                best_t = INFINITY;
                best_slot = -1;
                best_orig = 0x7FFFFFFF;
                bu = bv = 0.0f;
                budget = r.max_visits;
                sp = 0;
                if (r.valid) {
                    if (gsub == 0) stk[0] = r.root;
                    sp = 1;
                }
            }
        }
        if (__ballot(item >= 0) == 0) {
            if (f.drained && f.next >= f.end) break;
            continue;
        }
        const int need = (f.drained && f.next >= f.end) ? kGroups : 1;
        for (;;) {
            if (item >= 0) {  // uniform within a group from here on
                const int npop = sp < kQuads ? sp : kQuads;
                const int ref = q < npop ? stk[sp - 1 - q] : kDone;
                sp -= npop;
                const bool inner = ref >= 0 && ref != kDone;
                budget -= __popcll(__ballot(inner && k == 0) & gmask);
                bool h = false, anyhit = false;
                int rf = kEmpty;
                float key = INFINITY;
                float lt = INFINITY, lu = 0.0f, lv = 0.0f;
                int lo_ = 0x7FFFFFFF, ls = -1;
                if (inner) {
                    const float* nd = reinterpret_cast<const float*>(nodes4 + ref);
                    const float lx = nd[k], hx = nd[4 + k], ly = nd[8 + k], hy = nd[12 + k], lz = nd[16 + k],
                                hz = nd[20 + k];
                    rf = __float_as_int(nd[24 + k]);
                    float t;
                    h = slab(lx, hx, ly, hy, lz, hz, ix, iy, iz, px, py, pz, ANY ? INFINITY : best_t, t) &
                        (rf != kEmpty);
                    key = h ? t : INFINITY;
                } else if (ref != kDone) {
                    const int e = ~ref;
                    const int first = e & ((1 << kLeafShift) - 1);
                    const int cnt = ((e >> kLeafShift) & 7) + 1;
                    for (int b = 0; b < cnt; b += 4) {
                        const int kk = b + k;
                        const bool act = kk < cnt;
                        const int sl = first + (act ? kk : 0);
                        const float4 g0 = geo[3 * (size_t)sl], g1 = geo[3 * (size_t)sl + 1], g2 = geo[3 * (size_t)sl + 2];
                        float tt, uu, vv;
                        F3 lo = o, ld = d;
                        to_local<XF>(xf, __float_as_int(g2.z), lo, ld);
                        const bool hit = moller_trumbore(lo, ld, f3(g0.x, g0.y, g0.z), f3(g0.w, g1.x, g1.y),
                                                         f3(g1.z, g1.w, g2.x), tt, uu, vv) & act;
                        if (ANY) {
                            anyhit |= hit;
                        } else {
                            const int og = __float_as_int(g2.y);
                            if (hit && (tt < lt || (tt == lt && og < lo_))) {
                                lt = tt;
                                lo_ = og;
                                ls = sl;
                                lu = uu;
                                lv = vv;
                            }
                        }
                    }
                }
                // pushes: each quad its hits farthest first, quads in order, above the popped entries
                const float k1 = __shfl_xor(key, 1), k2 = __shfl_xor(key, 2), k3 = __shfl_xor(key, 3);
                const int rank = (k1 < key || (k1 == key && (k ^ 1) < k)) + (k2 < key || (k2 == key && (k ^ 2) < k)) +
                                 (k3 < key || (k3 == key && (k ^ 3) < k));
                const uint64_t hm = __ballot(h);
                const int nh_q = __popcll((hm >> qbit) & 0xFull);
                const int before = __popcll(hm & gmask & ((1ull << qbit) - 1));
                const int total = __popcll(hm & gmask);
                bool done = false;
                if (budget < 0) {
                    if (gsub == 0) atomicOr(a.err, 2);
                    done = true;
                } else if (sp + total > kStack) {
                    if (gsub == 0) atomicOr(a.err, 1);
                    done = true;
                } else {
                    if (h) stk[sp + before + (nh_q - 1 - rank)] = rf;
                    sp += total;
                }
                if (ANY) {
                    if (__ballot(anyhit) & gmask) {
                        best_slot = 0;  // marks "hit" for the any-hit result
                        done = true;
                    }
                } else {
                    // least (t, original index) over the group, then against the ray's best
                    for (int off = 1; off < G; off <<= 1) {
                        const float ot = __shfl_xor(lt, off), ou = __shfl_xor(lu, off), ov = __shfl_xor(lv, off);
                        const int oo = __shfl_xor(lo_, off), os = __shfl_xor(ls, off);
                        if (ot < lt || (ot == lt && oo < lo_)) {
                            lt = ot;
                            lo_ = oo;
                            ls = os;
                            lu = ou;
                            lv = ov;
                        }
                    }
                    if (ls >= 0 && (lt < best_t || (lt == best_t && lo_ < best_orig))) {
                        best_t = lt;
                        best_orig = lo_;
                        best_slot = ls;
                        bu = lu;
                        bv = lv;
                    }
                }
                if (done || sp == 0) {
                    if (gsub == 0) {
                        if (ANY) a.vis_out[out_idx] = (best_slot >= 0) == (MODE == MODE_RAYS_ANY) ? 1 : 0;
                        else a.hit_out[item] = make_float4(best_t, __int_as_float(best_slot), bu, bv);
                    }
                    item = -1;
                }
            }
            if ((int)__popcll(__ballot(gsub == 0 && item < 0)) >= need) break;
        }
    }
}

// Entries of a lane's traversal stack kept in registers, above the LDS and
// global parts (0: the whole stack is in LDS / global memory).  Measured on the
// 4K frame of the 4M-triangle sphere: 0 entries 2.18 ms for the primary launch,
// 2 entries 2.23 ms, 4 entries 2.28 ms (VGPRs 86 -> 91, occupancy unchanged at
// 5 waves per SIMD): the shifts of a register stack cost more issue slots under
// divergence than the LDS read latency they save, which the other waves of the
// SIMD already hide.
#ifndef EOS_REG_STACK
#define EOS_REG_STACK 0
#endif
constexpr int kRegStack = EOS_REG_STACK;

// Node-step pushes while the lane's LDS stack has room: all slots written
// without branches (measured on the 4K sphere frame: primary launch 1.61 ->
// 1.57 ms).  Any-hit launches push hit children in slot order instead of
// sorting them by entry distance: every box the ray meets is walked whatever
// the order, so only the (rare) occluded rays can end sooner or later (hit-dense
// 4K frame 19.4 -> 18.65 ms).  Both leave every result unchanged; launches that
// count visits keep the sorted order.
#ifndef EOS_FASTPUSH
#define EOS_FASTPUSH 1
#endif
#ifndef EOS_ANY_NOSORT
#define EOS_ANY_NOSORT 1
#endif
constexpr bool kFastPush = EOS_FASTPUSH != 0;
// Idle lanes a wave needs before its walking lanes lend subtrees (measured,
// frames identical: at 1, 16, 32, 48 idle lanes config C's single frame took
// 2.21, 2.11, 2.11, 2.28 ms and config D's 64 sequential frames 102, 93, 84, 87 ms).
#ifndef EOS_STEAL_IDLE
#define EOS_STEAL_IDLE 32
#endif
constexpr int kStealIdle = EOS_STEAL_IDLE;
constexpr bool kAnyNoSort = EOS_ANY_NOSORT != 0;

// A lane's traversal stack: the top kRegStack entries in registers (r[0] the
// top), the rest in LDS (STACK_LDS entries, column per lane) and then in the
// lane's global overflow area.  The memory part is non-empty only while the
// register part is full, so a pop takes its result from a register and refills
// the register bottom from memory off the dependent chain (the value is needed
// kRegStack pops later); a push spills the register bottom.  LIFO order is
// the same as a memory-only stack, so every ray's walk is unchanged.
template <int R, int WG, int STACK_LDS>
struct LaneStack {
    int r[R > 0 ? R : 1];
    int nreg = 0, sp = 0;   // entries in registers, entries in memory
    __device__ __forceinline__ void clear() { nreg = sp = 0; }
    __device__ __forceinline__ int depth() const { return nreg + sp; }
    __device__ __forceinline__ void mem_push(int32_t* lds, int32_t* gst, int tid, int v) {
        if (sp < STACK_LDS) lds[sp * WG + tid] = v;
        else gst[sp - STACK_LDS] = v;
        ++sp;
    }
    __device__ __forceinline__ int mem_pop(const int32_t* lds, const int32_t* gst, int tid) {
        --sp;
        return sp < STACK_LDS ? lds[sp * WG + tid] : gst[sp - STACK_LDS];
    }
    __device__ __forceinline__ void push(int32_t* lds, int32_t* gst, int tid, int v) {
        if (R == 0) {
            mem_push(lds, gst, tid, v);
            return;
        }
        if (nreg == R) mem_push(lds, gst, tid, r[R - 1]);
        else ++nreg;
#pragma unroll
        for (int k = R - 1; k > 0; --k) r[k] = r[k - 1];
        r[0] = v;
    }
    // the next node to visit, kDone when the stack is empty
    __device__ __forceinline__ int pop(const int32_t* lds, const int32_t* gst, int tid) {
        if (R == 0) return sp == 0 ? kDone : mem_pop(lds, gst, tid);
        if (nreg == 0) return kDone;
        const int top = r[0];
#pragma unroll
        for (int k = 0; k + 1 < R; ++k) r[k] = r[k + 1];
        if (sp > 0) r[R - 1] = mem_pop(lds, gst, tid);
        else --nreg;
        return top;
    }
    // Memory-only stacks (R == 0): room for k more entries in LDS; the slot j
    // above the top; k more entries on the stack (written with put first).
    __device__ __forceinline__ bool lds_room(int k) const { return R == 0 && sp + k <= STACK_LDS; }
    __device__ __forceinline__ void put(int32_t* lds, int tid, int j, int v) { lds[(sp + j) * WG + tid] = v; }
    __device__ __forceinline__ void grow(int k) { sp += k; }
    // Removes and returns the bottom entry (the oldest: the farthest sibling
    // pushed nearest the root, usually the largest pending subtree); the
    // entries above move down one place, keeping their order.  depth() > 0,
    // memory-only stacks (R == 0).
    __device__ __forceinline__ int take_bottom(int32_t* lds, int32_t* gst, int tid) {
        const int v = STACK_LDS > 0 ? lds[tid] : gst[0];
        for (int j = 1; j < sp; ++j) {
            const int x = j < STACK_LDS ? lds[j * WG + tid] : gst[j - STACK_LDS];
            if (j - 1 < STACK_LDS) lds[(j - 1) * WG + tid] = x;
            else gst[j - 1 - STACK_LDS] = x;
        }
        --sp;
        return v;
    }
};

// Persistent traversal.  Each wave keeps up to 64 independent rays in flight
// (one per lane) and refills idle lanes from its chunk of items whenever at
// least `refill` lanes are idle (ballot + mbcnt ranks them), so lanes never
// wait for the slowest ray of a fixed batch; the traversal loop returns to the
// refill as soon as that many lanes are idle.  Per lane: ordered near-first
// descent, far children pushed onto a short stack in LDS (STACK_LDS entries,
// column per lane) that overflows into a private global area, pruning against
// the best hit so far.  Per-ray visit order is deterministic and equal to the
// CPU restatement in the oracle.  A launch with few items for its lanes (4-wide
// tree, a.coop) takes the quad-cooperative walk instead.  BATCH: every ray looks
// up its own scene (BVH, triangles, lights, camera) in the batch table.  XF:
// triangles may carry transformations; a ray meets such a triangle in the
// triangle's local space.  LEAN: no group walk, and so no LDS for its stacks
// (24 instead of 32 KiB per 256-thread workgroup), for a launch that runs
// beside another stream's launches.
template <int MODE, int WG, int STACK_LDS, bool COUNT, bool BATCH, int WIDTH, bool XF, bool LEAN = false>
__global__ __launch_bounds__(WG) void k_trace(TraceArgs a) {
    constexpr bool ANY = any_hit_mode(MODE);
    constexpr int MAXS = max_stack<WIDTH>();
    constexpr int LDS_LANE = STACK_LDS * WG, LDS_COOP = WIDTH == 4 ? (WG / 4) * kMaxStack4 : 0;
    // the group walk serves the short launches after the primary rays, never the
    // primary launch, and not batched frames, whose launches hold every scene's
    // rays (its LDS stacks, 32 KiB per workgroup, cost a batched frame 4%)
    constexpr bool GROUP = WIDTH == 4 && !COUNT && MODE != MODE_PRIMARY && !BATCH && !LEAN;
    // work stealing visits nodes in another order (never in launches that count visits)
    constexpr bool STEAL = !COUNT && kRegStack == 0;
    constexpr int LDS_GROUP = GROUP ? (WG / 64) * kWaveGroupStack : 0;
    constexpr int LDS_A = LDS_LANE > LDS_COOP ? LDS_LANE : LDS_COOP;
    __shared__ int32_t lds[LDS_A > LDS_GROUP ? LDS_A : LDS_GROUP];
    const int tid = threadIdx.x;
    const int lane = __lane_id();
    int n = a.n_static;
    if (a.count_ptr) {
        n = 0;
        const int nlev = MODE == MODE_SHADOW ? a.nlev : 1;
        for (int k = 0; k < nlev; ++k) n += a.count_ptr[k] * a.count_mul;
    }
    unsigned long long visits = 0, tests = 0;
    const long long lanes = (long long)gridDim.x * WG;
    // a whole wave per ray for launches of a few thousand rays, else 16 lanes
    const bool wide = GROUP && a.group_lanes == 64 && use_group(a.group, n, lanes, 64);
    const bool grouped = GROUP && (wide || use_group(a.group, n, lanes, kGroupLanes));
    const bool coop = !grouped && WIDTH == 4 && use_coop(a.coop, n, lanes);
    // Waves beyond what the launch's items can occupy at once leave now: a
    // launch with a few items on the full persistent grid would otherwise have
    // every wave claim from every range cursor before it finds nothing left.
    {
        const long long per_wave = wide ? 1 : (grouped ? 64 / kGroupLanes : (coop ? 16 : 64));
        const long long need = (n + per_wave - 1) / per_wave;
        const int wave_id = (int)blockIdx.x * (WG / 64) + (tid >> 6);
        if (wave_id >= (need > 0 ? need : 1)) return;
    }
    if (wide) {
        trace_group<MODE, WG, BATCH, XF, 64>(a, n, lds + (tid >> 6) * kWaveGroupStack);
    } else if (grouped) {
        trace_group<MODE, WG, BATCH, XF, kGroupLanes>(a, n, lds + (tid >> 6) * kWaveGroupStack);
    } else if (coop) {
        trace_coop<MODE, WG, COUNT, BATCH, XF>(a, n, lds + (tid >> 6) * 16 * kMaxStack4, visits, tests);
    } else {
        int32_t* gst = a.gstack + (size_t)(blockIdx.x * WG + tid) * (MAXS - STACK_LDS);
        const Node* __restrict__ nodes = a.sc.nodes;
        const Node4* __restrict__ nodes4 = a.sc.nodes4;
        const float4* __restrict__ geo = a.sc.geo;
        const float4* __restrict__ xf = a.sc.xf;
        int item = -1;
        size_t out_idx = 0;
        Fetch f = fetch_start(a, n);
        const int waves_total = (int)gridDim.x * (WG / 64);
        Scatter scat;
        scat.init(a.scatter, n, (long long)gridDim.x * WG, MODE == MODE_SHADOW ? a.count_mul : 1);
        F3 o = {0, 0, 0}, d = {0, 0, 0};
        float ix = 0, iy = 0, iz = 0, px = 0, py = 0, pz = 0;
        float best_t = INFINITY, bu = 0.0f, bv = 0.0f;
        int best_slot = -1, best_orig = 0x7FFFFFFF;
        int cur = kDone;
        LaneStack<kRegStack, WG, STACK_LDS> stk;
        int lpos = -1;   // next triangle of the current leaf (-1: not in a leaf)
        bool farf = false;   // any hit: walk the farthest hit child first (RayItem::far_first)
        int sgx = 0, sgy = 0, sgz = 0;   // 1 where the inverse direction is negative (hi plane nearer)
        int budget = 0;  // node visits left for this ray: a tree visits each internal node at most once
        // XF: the ray in the local space of transformation lxi, kept while the
        // walk stays among that transformation's triangles (-1: none yet)
        int lxi = -1;
        F3 lo_o = {0, 0, 0}, lo_d = {0, 0, 0};
        const int thresh = a.refill < 1 ? 1 : (a.refill > 64 ? 64 : a.refill);
        // Work stealing once the wave's items are all claimed: a lane with a ray
        // (its home lane) lends the bottom entry of its stack to an idle lane,
        // which walks that subtree for the same ray as a helper and hands its
        // result back.  home: the home lane of the ray a helper works on (-1: this
        // lane is the ray's home, or idle); pending: helpers of a home's ray not yet
        // merged; fin: a helper whose part is walked.
        const bool steal = STEAL && a.steal != 0;
        int home = -1, pending = 0;
        bool fin = false;
        // the ray's answer, written by its home lane
        auto emit = [&]() {
            if (ANY) {
                a.vis_out[out_idx] = (best_slot >= 0) == (MODE == MODE_RAYS_ANY) ? 1 : 0;
            } else {
                a.hit_out[item] = make_float4(best_t, __int_as_float(best_slot), bu, bv);
            }
            item = -1;
            pending = 0;
        };

        for (;;) {
            // ---- refill idle lanes
            const uint64_t idle = __ballot(item < 0);
            const int nidle = __popcll(idle);
            if (nidle >= thresh && (!f.drained || f.next < f.end)) {
                int my = fetch_items(f, a, n, waves_total, idle, nidle, __popcll(idle & lanemask_lt()));
                my = primary_tile_item<MODE, BATCH>(a, scat.map(my));
                if (item < 0 && my >= 0) {
                    RayItem r;
                    setup_item<MODE, BATCH, WIDTH>(a, my, r, !scat.on);
                    item = my;
                    out_idx = r.out_idx;
                    o = r.o;
                    d = r.d;
                    nodes = r.nodes;
                    nodes4 = r.nodes4;
                    geo = r.geo;
                    xf = r.xf;
                    ix = safe_inv(d.x);
                    iy = safe_inv(d.y);
                    iz = safe_inv(d.z);
                    px = o.x * ix;
                    py = o.y * iy;
                    pz = o.z * iz;
                    sgx = ix < 0.0f ? 1 : 0;
                    sgy = iy < 0.0f ? 1 : 0;
                    sgz = iz < 0.0f ? 1 : 0;
                    best_t = INFINITY;
                    best_slot = -1;
                    best_orig = 0x7FFFFFFF;
                    bu = bv = 0.0f;
                    stk.clear();
                    lpos = -1;
                    lxi = -1;
                    budget = r.max_visits;
                    cur = r.valid ? r.root : kDone;
                    farf = ANY && !COUNT && r.far_first;
                    home = -1;
                    pending = 0;
                }
            }
            if (__ballot(item >= 0) == 0) {
                if (f.drained && f.next >= f.end) break;
                continue;
            }
            // ---- traverse until `thresh` lanes are idle (all of them once no work is
            // left to claim), so finished lanes get new items without waiting for the
            // longest ray in the wave
            const int need = (f.drained && f.next >= f.end) ? 64 : thresh;
            const bool stealing = steal && need == 64;
            for (;;) {
                if (ANY && stealing && __ballot(home >= 0) != 0) {
                    // a helper stops once its ray is answered: the home lane left
                    // it (an occluder was found) or took another role
                    const int hl = home >= 0 ? home : lane;
                    const int hi = __shfl(item, hl), hh = __shfl(home, hl);
                    if (home >= 0 && item >= 0 && (hi != item || hh >= 0)) {
                        cur = kDone;
                        stk.clear();
                        lpos = -1;
                    }
                }
                if (item >= 0) {
                    if (cur >= 0 && cur != kDone && --budget < 0) {
                        atomicOr(a.err, 2);
                        cur = kDone;
                    } else if (cur >= 0 && cur != kDone) {
                        if (COUNT) ++visits;
                        const float tmax = ANY ? INFINITY : best_t;
                        if (WIDTH == 4) {
                            // the near and far planes of each axis fetched by the ray's
                            // direction signs (float4 0/1: x lo/hi, 2/3: y, 4/5: z);
                            // batched frames keep the plain test (measured faster there)
                            const float4* nd = reinterpret_cast<const float4*>(nodes4 + cur);
                            const int4 rf = reinterpret_cast<const int4*>(nd)[6];
                            float t0, t1, t2, t3;
                            bool h0, h1, h2, h3;
                            // all four tests unconditionally (no load sinking), empty slots masked after
                            if (BATCH) {
                                const float4 lx = nd[0], hx = nd[1], ly = nd[2], hy = nd[3], lz = nd[4], hz = nd[5];
                                h0 = slab(lx.x, hx.x, ly.x, hy.x, lz.x, hz.x, ix, iy, iz, px, py, pz, tmax, t0) & (rf.x != kEmpty);
                                h1 = slab(lx.y, hx.y, ly.y, hy.y, lz.y, hz.y, ix, iy, iz, px, py, pz, tmax, t1) & (rf.y != kEmpty);
                                h2 = slab(lx.z, hx.z, ly.z, hy.z, lz.z, hz.z, ix, iy, iz, px, py, pz, tmax, t2) & (rf.z != kEmpty);
                                h3 = slab(lx.w, hx.w, ly.w, hy.w, lz.w, hz.w, ix, iy, iz, px, py, pz, tmax, t3) & (rf.w != kEmpty);
                            } else {
                                const float4 nx = nd[sgx], fx = nd[1 - sgx], ny = nd[2 + sgy], fy = nd[3 - sgy],
                                             nz = nd[4 + sgz], fz = nd[5 - sgz];
                                h0 = slab_nf(nx.x, fx.x, ny.x, fy.x, nz.x, fz.x, ix, iy, iz, px, py, pz, tmax, t0) & (rf.x != kEmpty);
                                h1 = slab_nf(nx.y, fx.y, ny.y, fy.y, nz.y, fz.y, ix, iy, iz, px, py, pz, tmax, t1) & (rf.y != kEmpty);
                                h2 = slab_nf(nx.z, fx.z, ny.z, fy.z, nz.z, fz.z, ix, iy, iz, px, py, pz, tmax, t2) & (rf.z != kEmpty);
                                h3 = slab_nf(nx.w, fx.w, ny.w, fy.w, nz.w, fz.w, ix, iy, iz, px, py, pz, tmax, t3) & (rf.w != kEmpty);
                            }
                            const int nh = (int)h0 + (int)h1 + (int)h2 + (int)h3;
                            if (nh == 0) {
                                cur = stk.pop(lds, gst, tid);
                            } else if (stk.depth() + nh - 1 > MAXS) {
                                atomicOr(a.err, 1);
                                cur = kDone;
                            } else if (ANY && !COUNT && kAnyNoSort && !farf && stk.lds_room(4)) {
                                // any hit: the order of the children does not change the
                                // answer; the first hit child next, the others onto the stack
                                // in slot order (non-hits to a slot above the new top)
                                const int c1 = (int)h0, c2 = c1 + (int)h1, c3 = c2 + (int)h2;
                                cur = h0 ? rf.x : (h1 ? rf.y : (h2 ? rf.z : rf.w));
                                stk.put(lds, tid, h1 && c1 >= 1 ? c1 - 1 : 3, rf.y);
                                stk.put(lds, tid, h2 && c2 >= 1 ? c2 - 1 : 3, rf.z);
                                stk.put(lds, tid, h3 && c3 >= 1 ? c3 - 1 : 3, rf.w);
                                stk.grow(nh - 1);
                            } else {
                                float e0 = h0 ? t0 : INFINITY, e1 = h1 ? t1 : INFINITY, e2 = h2 ? t2 : INFINITY,
                                      e3 = h3 ? t3 : INFINITY;
                                int k0 = 0, k1 = 1, k2 = 2, k3 = 3;
                                int r0 = rf.x, r1 = rf.y, r2 = rf.z, r3 = rf.w;
                                cmpswap(e0, k0, r0, e1, k1, r1);
                                cmpswap(e2, k2, r2, e3, k3, r3);
                                cmpswap(e0, k0, r0, e2, k2, r2);
                                cmpswap(e1, k1, r1, e3, k3, r3);
                                cmpswap(e1, k1, r1, e2, k2, r2);
                                if (ANY && !COUNT && farf) {
                                    // farthest hit child next; the nearer ones pushed nearest
                                    // first, so the next farthest is popped next
                                    const int far = nh > 3 ? r3 : (nh > 2 ? r2 : (nh > 1 ? r1 : r0));
                                    if (kFastPush && stk.lds_room(3)) {
                                        stk.put(lds, tid, 0, r0);
                                        stk.put(lds, tid, 1, r1);
                                        stk.put(lds, tid, 2, r2);
                                        stk.grow(nh - 1);
                                    } else {
                                        if (nh > 1) stk.push(lds, gst, tid, r0);
                                        if (nh > 2) stk.push(lds, gst, tid, r1);
                                        if (nh > 3) stk.push(lds, gst, tid, r2);
                                    }
                                    r0 = far;
                                } else if (kFastPush && stk.lds_room(3)) {
                                    // farthest first, all three slots written (those above
                                    // the new top are overwritten later)
                                    stk.put(lds, tid, 0, nh > 3 ? r3 : (nh > 2 ? r2 : r1));
                                    stk.put(lds, tid, 1, nh > 3 ? r2 : r1);
                                    stk.put(lds, tid, 2, r1);
                                    stk.grow(nh - 1);
                                } else {
                                    if (nh > 3) stk.push(lds, gst, tid, r3);
                                    if (nh > 2) stk.push(lds, gst, tid, r2);
                                    if (nh > 1) stk.push(lds, gst, tid, r1);
                                }
                                cur = r0;
                            }
                        } else {
                            const Node* nd = nodes + cur;
                            const float4 b0 = nd->b0, b1 = nd->b1, b2 = nd->b2;
                            const int4 c = nd->c;
                            float t0, t1;
                            const bool h0 = slab(b0.x, b0.y, b0.z, b0.w, b2.x, b2.y, ix, iy, iz, px, py, pz, tmax, t0);
                            const bool h1 = slab(b1.x, b1.y, b1.z, b1.w, b2.z, b2.w, ix, iy, iz, px, py, pz, tmax, t1);
                            if (h0 && h1) {
                                const bool swap = t1 < t0;
                                const int nearc = swap ? c.y : c.x, farc = swap ? c.x : c.y;
                                if (stk.depth() >= MAXS) {
                                    atomicOr(a.err, 1);
                                    cur = kDone;
                                } else {
                                    stk.push(lds, gst, tid, farc);
                                    cur = nearc;
                                }
                            } else if (h0) {
                                cur = c.x;
                            } else if (h1) {
                                cur = c.y;
                            } else {
                                cur = stk.pop(lds, gst, tid);
                            }
                        }
                    } else if (cur != kDone) {
                        // One pair of triangles per step (a longer leaf takes several
                        // steps, so leaf steps cost about what node steps cost and the
                        // lanes of a wave diverge less); both triangles are fetched
                        // before either is tested; an odd leaf tests its last triangle
                        // twice (same result, counted once).
                        const int e = ~cur;
                        const int first = e & ((1 << kLeafShift) - 1);
                        const int cnt = ((e >> kLeafShift) & 7) + 1;
                        bool occluded = false;
                        const int last = first + cnt - 1;
                        {
                            const int s0 = lpos < 0 ? first : lpos;
                            const int s1 = s0 + 1 <= last ? s0 + 1 : s0;
                            lpos = s0 + 2 <= last ? s0 + 2 : -1;
                            const float4 a0 = geo[3 * (size_t)s0], a1 = geo[3 * (size_t)s0 + 1], a2 = geo[3 * (size_t)s0 + 2];
                            const float4 b0 = geo[3 * (size_t)s1], b1 = geo[3 * (size_t)s1 + 1], b2 = geo[3 * (size_t)s1 + 2];
                            float ta, ua, va, tb, ub, vb;
                            F3 oa = o, da = d, ob = o, db = d;
                            if (XF) {
                                const int xa = __float_as_int(a2.z), xb = __float_as_int(b2.z);
                                if (xa >= 0 && xa != lxi) {
                                    lo_o = o;
                                    lo_d = d;
                                    to_local<XF>(xf, xa, lo_o, lo_d);
                                    lxi = xa;
                                }
                                if (xa >= 0) {
                                    oa = lo_o;
                                    da = lo_d;
                                }
                                if (xb >= 0 && xb != lxi) {
                                    lo_o = o;
                                    lo_d = d;
                                    to_local<XF>(xf, xb, lo_o, lo_d);
                                    lxi = xb;
                                }
                                if (xb >= 0) {
                                    ob = lo_o;
                                    db = lo_d;
                                }
                            }
                            const bool ha = moller_trumbore(oa, da, f3(a0.x, a0.y, a0.z), f3(a0.w, a1.x, a1.y),
                                                            f3(a1.z, a1.w, a2.x), ta, ua, va);
                            const bool hb = moller_trumbore(ob, db, f3(b0.x, b0.y, b0.z), f3(b0.w, b1.x, b1.y),
                                                            f3(b1.z, b1.w, b2.x), tb, ub, vb);
                            // counted as the one-at-a-time walk would: an any-hit ray stops at its first hit
                            if (COUNT) tests += (ANY && ha) ? 1 : (s1 != s0 ? 2 : 1);
                            if (ANY) {
                                if (ha | hb) occluded = true;
                            } else {
                                const int oa = __float_as_int(a2.y), ob = __float_as_int(b2.y);
                                if (ha && (ta < best_t || (ta == best_t && oa < best_orig))) {
                                    best_t = ta;
                                    best_orig = oa;
                                    best_slot = s0;
                                    bu = ua;
                                    bv = va;
                                }
                                if (hb && (tb < best_t || (tb == best_t && ob < best_orig))) {
                                    best_t = tb;
                                    best_orig = ob;
                                    best_slot = s1;
                                    bu = ub;
                                    bv = vb;
                                }
                            }
                        }
                        if (ANY && occluded) {
                            best_slot = 0;  // marks "hit" for the any-hit result
                            lpos = -1;
                            cur = kDone;
                        } else if (lpos < 0) {
                            cur = stk.pop(lds, gst, tid);
                        }
                    }
                    if (cur == kDone) {
                        if (STEAL && home >= 0) fin = true;   // merged below
                        else if (!STEAL || pending == 0 || (ANY && best_slot >= 0)) emit();
                        // else: a home whose helpers still walk parts of its ray waits for them
                    }
                }
                if (STEAL && stealing) {
                    // ---- merge the helpers that finished into their home lanes, one at a time
                    uint64_t fm = __ballot(fin);
                    while (fm) {
                        const int src = __ffsll((unsigned long long)fm) - 1;
                        fm &= fm - 1;
                        const int h = __builtin_amdgcn_readlane(home, src);
                        const int it = __builtin_amdgcn_readlane(item, src);
                        const int fs = __builtin_amdgcn_readlane(best_slot, src);
                        const int fo = __builtin_amdgcn_readlane(best_orig, src);
                        const float ft = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(best_t), src));
                        const float fu = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(bu), src));
                        const float fv = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(bv), src));
                        if (lane == h && item == it && home < 0) {
                            --pending;
                            if (ANY) {
                                if (fs >= 0) {
                                    best_slot = 0;
                                    cur = kDone;
                                    stk.clear();
                                    lpos = -1;
                                }
                            } else if (fs >= 0 && (ft < best_t || (ft == best_t && fo < best_orig))) {
                                best_t = ft;
                                best_orig = fo;
                                best_slot = fs;
                                bu = fu;
                                bv = fv;
                            }
                            if (cur == kDone && (pending == 0 || (ANY && best_slot >= 0))) emit();
                        }
                    }
                    if (fin) {
                        fin = false;
                        home = -1;
                        item = -1;
                    }
                    // ---- idle lanes take the bottom stack entry of walking lanes
                    // (once at least kStealIdle lanes are idle: while most lanes still
                    // walk, lending costs the walkers more than the idle lanes gain)
                    const uint64_t idle = __ballot(item < 0);
                    const uint64_t dm = __ballot(item >= 0 && cur != kDone && stk.depth() > 0);
                    if (__popcll(idle) >= kStealIdle && dm != 0) {
                        const int ni = __popcll(idle), nd = __popcll(dm);
                        const int np = ni < nd ? ni : nd;
                        const int irank = __popcll(idle & lanemask_lt());
                        const bool thief = ((idle >> lane) & 1) && irank < np;
                        const bool donor = ((dm >> lane) & 1) && __popcll(dm & lanemask_lt()) < np;
                        const int src = thief ? nth_set_bit(dm, irank) : lane;
                        int give = kDone;
                        if (donor) give = stk.take_bottom(lds, gst, tid);
                        const int g_ref = __shfl(give, src), g_item = __shfl(item, src);
                        const int g_home = __shfl(home >= 0 ? home : lane, src);
                        const float g_ox = __shfl(o.x, src), g_oy = __shfl(o.y, src), g_oz = __shfl(o.z, src);
                        const float g_dx = __shfl(d.x, src), g_dy = __shfl(d.y, src), g_dz = __shfl(d.z, src);
                        const float g_t = __shfl(best_t, src);
                        const int g_orig = __shfl(best_orig, src), g_budget = __shfl(budget, src);
                        const size_t g_out = (size_t)__shfl((long long)out_idx, src);
                        if (BATCH) {
                            const Node* g_nodes = (const Node*)__shfl((long long)nodes, src);
                            const Node4* g_nodes4 = (const Node4*)__shfl((long long)nodes4, src);
                            const float4* g_geo = (const float4*)__shfl((long long)geo, src);
                            const float4* g_xf = (const float4*)__shfl((long long)xf, src);
                            if (thief) {
                                nodes = g_nodes;
                                nodes4 = g_nodes4;
                                geo = g_geo;
                                xf = g_xf;
                            }
                        }
                        if (thief) {
                            item = g_item;
                            home = g_home;
                            out_idx = g_out;
                            cur = g_ref;
                            o = f3(g_ox, g_oy, g_oz);
                            d = f3(g_dx, g_dy, g_dz);
                            ix = safe_inv(d.x);
                            iy = safe_inv(d.y);
                            iz = safe_inv(d.z);
                            px = o.x * ix;
                            py = o.y * iy;
                            pz = o.z * iz;
                            sgx = ix < 0.0f ? 1 : 0;
                            sgy = iy < 0.0f ? 1 : 0;
                            sgz = iz < 0.0f ? 1 : 0;
                            // only hits better than the lender's best so far count
                            best_t = g_t;
                            best_orig = g_orig;
                            best_slot = -1;
                            bu = bv = 0.0f;
                            budget = g_budget;
                            stk.clear();
                            lpos = -1;
                            farf = false;
                            lxi = -1;
                            pending = 0;
                        }
                        uint64_t tm = __ballot(thief);
                        while (tm) {
                            const int s2 = __ffsll((unsigned long long)tm) - 1;
                            tm &= tm - 1;
                            if (lane == __builtin_amdgcn_readlane(home, s2)) ++pending;
                        }
                    }
                }
                if ((int)__popcll(__ballot(item < 0)) >= need) break;
            }
        }
    }
    if (COUNT) {
        // wave-reduce the counters, one atomic per wave
        for (int off = 32; off > 0; off >>= 1) {
            visits += __shfl_xor(visits, off);
            tests += __shfl_xor(tests, off);
        }
        if (lane == 0) {
            atomicAdd(&a.stats[0], visits);
            atomicAdd(&a.stats[1], tests);
        }
    }
}

// Packet traversal of primary rays (4-wide tree).  One wave owns one 8x8 pixel
// tile at a time: the node or leaf being visited is the same for the whole
// wave (a wave-uniform reference, so node and triangle records arrive through
// scalar loads, once per wave), every lane tests its own ray against the four
// child boxes with its own best t, and the wave descends into every child that
// some lane hits.  Children are ordered by the entry distance of the first lane
// (lowest lane index) that hits them, then by slot; the nearest is taken and
// the others are pushed farthest first onto a per-wave stack in LDS.  A lane's
// nearest triangle lies in boxes that lane hits with entry distance <= its
// best t, so the wave visits them: each lane's result is the exact nearest hit
// under the (t, original index) order, as in the per-ray walk.  Visit and test
// counts are per wave (COUNT).
// Read-only views in the constant address space: a wave-uniform load through
// them is a scalar load (scalar cache, SGPRs).  Nodes and triangle records are
// never written while a frame traces.  (The host pass only type-checks kernel
// bodies and has no such address space.)
#if defined(__HIP_DEVICE_COMPILE__)
typedef const __attribute__((address_space(4))) Node4* ConstNode4Ptr;
typedef const __attribute__((address_space(4))) float4* ConstF4Ptr;
#else
typedef const Node4* ConstNode4Ptr;
typedef const float4* ConstF4Ptr;
#endif

template <int WG, bool COUNT, bool BATCH>
__global__ __launch_bounds__(WG) void k_trace_packet(TraceArgs a) {
    constexpr int MAXS = kMaxStack4;
    __shared__ int32_t stk[WG / 64][MAXS];
    const int lane = __lane_id();
    int32_t* st = stk[threadIdx.x >> 6];
    const int n_tiles = a.n_static >> 6;
    unsigned long long visits = 0, tests = 0;
    for (;;) {
        int tile = 0;
        if (lane == 0) tile = atomicAdd(a.cursor, 1);
        tile = __builtin_amdgcn_readfirstlane(__shfl(tile, 0));
        if (tile >= n_tiles) break;
        const int my = tile * 64 + lane;
        ConstNode4Ptr nodes4 = (ConstNode4Ptr)a.sc.nodes4;
        ConstF4Ptr geo = (ConstF4Ptr)a.sc.geo;
        int root = a.sc.root4, max_visits = a.sc.n_tri + 64;
        int x, y;
        bool valid;
        F3 o, d;
        if (BATCH) {
            const BatchScene& B = a.table[scene_of_id(a.table, a.n_scenes, tile * 64)];
            nodes4 = (ConstNode4Ptr)B.sc.nodes4;
            geo = (ConstF4Ptr)B.sc.geo;
            root = B.sc.root4;
            max_visits = B.sc.n_tri + 64;
            valid = tile_pixel(my - B.id_off, B.tiles_x, B.cam.width, B.cam.height, x, y);
            o = f3(B.cam.position[0], B.cam.position[1], B.cam.position[2]);
            d = camera_dir(B.cam, x, y, 0.0f, 0.0f);
        } else {
            valid = tile_pixel(my, a.tiles_x, a.cam.width, a.band_h, x, y);
            o = f3(a.cam.position[0], a.cam.position[1], a.cam.position[2]);
            d = camera_dir(a.cam, x, y + a.row0, a.aa_ox, a.aa_oy);
        }
        const float ix = safe_inv(d.x), iy = safe_inv(d.y), iz = safe_inv(d.z);
        const float px = o.x * ix, py = o.y * iy, pz = o.z * iz;
        float best_t = INFINITY, bu = 0.0f, bv = 0.0f;
        int best_slot = -1, best_orig = 0x7FFFFFFF;
        int cur = __builtin_amdgcn_readfirstlane(root);
        int sp = 0, budget = max_visits;
        while (cur != kDone) {
            if (cur >= 0) {
                if (--budget < 0) {
                    if (lane == 0) atomicOr(a.err, 2);
                    break;
                }
                if (COUNT) ++visits;
                ConstNode4Ptr nd = nodes4 + cur;
                const float4 lx = nd->lox, hx = nd->hix, ly = nd->loy, hy = nd->hiy, lz = nd->loz, hz = nd->hiz;
                const int4 rf = nd->ref;
                float t0, t1, t2, t3;
                const bool h0 = slab(lx.x, hx.x, ly.x, hy.x, lz.x, hz.x, ix, iy, iz, px, py, pz, best_t, t0) & valid;
                const bool h1 = slab(lx.y, hx.y, ly.y, hy.y, lz.y, hz.y, ix, iy, iz, px, py, pz, best_t, t1) & valid;
                const bool h2 = slab(lx.z, hx.z, ly.z, hy.z, lz.z, hz.z, ix, iy, iz, px, py, pz, best_t, t2) & valid;
                const bool h3 = slab(lx.w, hx.w, ly.w, hy.w, lz.w, hz.w, ix, iy, iz, px, py, pz, best_t, t3) & valid;
                const uint64_t m0 = rf.x != kEmpty ? __ballot(h0) : 0, m1 = rf.y != kEmpty ? __ballot(h1) : 0,
                               m2 = rf.z != kEmpty ? __ballot(h2) : 0, m3 = rf.w != kEmpty ? __ballot(h3) : 0;
                // entry distance of the first lane hitting each child (wave-uniform)
                float e0, e1, e2, e3;
                e0 = m0 ? __int_as_float(__builtin_amdgcn_readlane(__float_as_int(t0), __ffsll((unsigned long long)m0) - 1)) : INFINITY;
                e1 = m1 ? __int_as_float(__builtin_amdgcn_readlane(__float_as_int(t1), __ffsll((unsigned long long)m1) - 1)) : INFINITY;
                e2 = m2 ? __int_as_float(__builtin_amdgcn_readlane(__float_as_int(t2), __ffsll((unsigned long long)m2) - 1)) : INFINITY;
                e3 = m3 ? __int_as_float(__builtin_amdgcn_readlane(__float_as_int(t3), __ffsll((unsigned long long)m3) - 1)) : INFINITY;
                const int nh = (m0 != 0) + (m1 != 0) + (m2 != 0) + (m3 != 0);
                if (nh == 0) {
                    cur = sp == 0 ? kDone : __builtin_amdgcn_readfirstlane(st[--sp]);
                } else if (sp + nh - 1 > MAXS) {
                    if (lane == 0) atomicOr(a.err, 1);
                    break;
                } else {
                    int k0 = 0, k1 = 1, k2 = 2, k3 = 3;
                    int r0 = rf.x, r1 = rf.y, r2 = rf.z, r3 = rf.w;
                    cmpswap(e0, k0, r0, e1, k1, r1);
                    cmpswap(e2, k2, r2, e3, k3, r3);
                    cmpswap(e0, k0, r0, e2, k2, r2);
                    cmpswap(e1, k1, r1, e3, k3, r3);
                    cmpswap(e1, k1, r1, e2, k2, r2);
                    if (nh > 3) {
                        if (lane == 0) st[sp] = r3;
                        ++sp;
                    }
                    if (nh > 2) {
                        if (lane == 0) st[sp] = r2;
                        ++sp;
                    }
                    if (nh > 1) {
                        if (lane == 0) st[sp] = r1;
                        ++sp;
                    }
                    cur = __builtin_amdgcn_readfirstlane(r0);
                }
            } else {
                const int e = ~cur;
                const int first = e & ((1 << kLeafShift) - 1);
                const int last = first + ((e >> kLeafShift) & 7);
                for (int s0 = first; s0 <= last; s0 += 2) {
                    const int s1 = s0 + 1 <= last ? s0 + 1 : s0;
                    const float4 a0 = geo[3 * (size_t)s0], a1 = geo[3 * (size_t)s0 + 1], a2 = geo[3 * (size_t)s0 + 2];
                    const float4 b0 = geo[3 * (size_t)s1], b1 = geo[3 * (size_t)s1 + 1], b2 = geo[3 * (size_t)s1 + 2];
                    if (COUNT) tests += s1 != s0 ? 2 : 1;
                    float ta, ua, va, tb, ub, vb;
                    const bool ha = moller_trumbore(o, d, f3(a0.x, a0.y, a0.z), f3(a0.w, a1.x, a1.y),
                                                    f3(a1.z, a1.w, a2.x), ta, ua, va);
                    const bool hb = moller_trumbore(o, d, f3(b0.x, b0.y, b0.z), f3(b0.w, b1.x, b1.y),
                                                    f3(b1.z, b1.w, b2.x), tb, ub, vb);
                    const int oa = __float_as_int(a2.y), ob = __float_as_int(b2.y);
                    if (ha && (ta < best_t || (ta == best_t && oa < best_orig))) {
                        best_t = ta;
                        best_orig = oa;
                        best_slot = s0;
                        bu = ua;
                        bv = va;
                    }
                    if (hb && (tb < best_t || (tb == best_t && ob < best_orig))) {
                        best_t = tb;
                        best_orig = ob;
                        best_slot = s1;
                        bu = ub;
                        bv = vb;
                    }
                }
                cur = sp == 0 ? kDone : __builtin_amdgcn_readfirstlane(st[--sp]);
            }
        }
        a.hit_out[my] = valid ? make_float4(best_t, __int_as_float(best_slot), bu, bv)
                              : make_float4(INFINITY, __int_as_float(-1), 0.0f, 0.0f);
    }
    if (COUNT && lane == 0) {
        atomicAdd(&a.stats[0], visits);
        atomicAdd(&a.stats[1], tests);
    }
}

}  // namespace eos
