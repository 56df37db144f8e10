// render.hip -- the Whitted render path as a wavefront pipeline on gfx950.
//
// Reference semantics: src/RayTracer.cpp:99-174 (frame loop, quantisation),
// :268-300 (rayTrace), :302-416 (shadow rays, Phong sum, reflection),
// :418-458 (calcPhong, findDiffuseColor).  The reference recursion
//     c_l = sum_{unshadowed lights} phong_l  +  r_l * c_{l+1}
// is split into stages per bounce level l = 0..depth:
//   trace closest  (persistent, rays of level l)          -> hit records
//   setup          (hit point, shading normal, material)  -> hit surfaces,
//                  reflection rays of level l+1 (compacted queue)
//   trace any-hit  (persistent, one ray per (hit, light)) -> visibility bytes
//   phong          (sum over visible lights, light order) -> colour of level l
// and a final resolve evaluates c_0 = P_0 + r_0 (P_1 + r_1 (...)) innermost
// first, exactly the reference's order, then quantises.  Every queue length
// lives in device counters, so a frame is one stream of launches with no host
// round trip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstring>
#include <string>
#include <vector>

#include "eos_internal.h"
#include "trace.h"

namespace eos {
namespace {

struct SetupArgs {
    SceneView sc;
    DevCamera cam;
    float aa_ox, aa_oy;       // anti-aliasing sample offset inside the pixel
    int row0, band_h;         // level 0 covers camera rows [row0, row0 + band_h)
    int tiles_x;
    const BatchScene* table;
    int n_scenes;
    int n_ids, depth, level;
    const int* count_ptr;     // level >= 1: queued rays
    const float4* q_in;       // level >= 1
    const float4* hit_in;
    float4* surf_out;
    int* hits_ctr;
    float4* q_out;
    int* q_out_ctr;
    int q_cap;                // rays the next level's queue holds (appends beyond it are dropped and counted)
    int* shadow_ctr;
    uint8_t* levels;
    int32_t* hit_ids;         // level 0, optional
    const uint8_t* tile_live; // level 0, single scene, optional: 0 = the tile test resolved the tile (misses)
    // ray trees (scenes with transparent materials): the hit index of every
    // ray of this level (level 0: per pixel; -1: no hit), and per hit the
    // Fresnel weight and the queue slots of its children
    int32_t* rayhit;
    float4* aux;              // (K, reflection child, Fresnel child, transparent)
};

// src/RayTracer.cpp:52-65 (Bram de Greve's Fresnel reflectance): 1 on total
// internal reflection, else the mean of the squared s and p amplitudes.
__device__ __forceinline__ float fresnel(F3 incident, F3 normal, float n1, float n2) {
    const float n = n1 / n2;
    const float cosI = -dot3(normal, incident);
    const float sinT2 = n * n * (1.0f - cosI * cosI);
    if (sinT2 > 1.0f) return 1.0f;
    const float cosT = sqrtf(1.0f - sinT2);
    const float rOrth = (n1 * cosI - n2 * cosT) / (n1 * cosI + n2 * cosT);
    const float rPar = (n2 * cosI - n1 * cosT) / (n2 * cosI + n1 * cosT);
    return (rOrth * rOrth + rPar * rPar) / 2.0f;
}

// Hit set-up of one level: the hit surface (hit_surface, the reference's
// point and normal), the pixel's level count, the reflection ray and the
// slots of the next stages.
// 256 threads: the slot claims are one atomic per counter per workgroup
// round, and 512 or 1024 threads (2 or 4 times fewer claims) made single
// frames up to 3% faster but frames in flight 5-10% slower (larger workgroups
// share the CUs worse with the other frame's launches)
constexpr int kSetupWG = 256;
template <bool L0, bool BATCH>
__global__ __launch_bounds__(kSetupWG) void k_setup(SetupArgs a) {
    const int n = L0 ? a.n_ids : *a.count_ptr;
    const int stride = gridDim.x * blockDim.x;
    // every wave of the workgroup runs every round (the round's slots are
    // claimed once per workgroup, below), so the trip count is the workgroup's
    __shared__ int s_cnt[3][kSetupWG / 64], s_base[3][kSetupWG / 64];
    // the next round's hit record is loaded while this round is set up (a
    // level-0 pass is many rounds of mostly misses, bound by that load)
    const float4 kMiss = make_float4(INFINITY, __int_as_float(-1), 0.f, 0.f);
    // tiles the tile test resolved were not traced: their rays are misses
    const uint8_t* const tile_live = L0 && !BATCH ? a.tile_live : nullptr;
    auto traced = [tile_live](int j) { return tile_live == nullptr || tile_live[j >> 6] != 0; };
    const int i0 = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    float4 h_next = kMiss;
    if (i0 < n && traced(i0)) h_next = a.hit_in[i0];
    for (int bbase = blockIdx.x * blockDim.x; bbase < n; bbase += stride) {
        const int i = bbase + (int)threadIdx.x;
        const bool valid = i < n;
        const float4 h = h_next;
        h_next = kMiss;
        if ((long long)i + stride < n && traced(i + stride)) h_next = a.hit_in[i + stride];
        const int slot = __float_as_int(h.y);
        int pixel = 0, scene = 0;
        F3 o = {0, 0, 0}, d = {0, 0, 0};
        bool pix_ok = valid;
        const float4* shade = a.sc.shade;
        const float4* geo = a.sc.geo;
        const float4* xf = a.sc.xf;
        const DevMaterial* mats = a.sc.mats;
        if (L0) {
            int x = 0, y = 0;
            if (BATCH) {
                if (valid) {
                    scene = scene_of_id(a.table, a.n_scenes, i);
                    const BatchScene& B = a.table[scene];
                    pix_ok = tile_pixel(i - B.id_off, B.tiles_x, B.cam.width, B.cam.height, x, y);
                    pixel = B.pix_off + y * B.cam.width + x;
                    o = f3(B.cam.position[0], B.cam.position[1], B.cam.position[2]);
                    if (pix_ok && slot >= 0) d = camera_dir(B.cam, x, y, 0.0f, 0.0f);   // hits only use d
                    shade = B.sc.shade;
                    geo = B.sc.geo;
                    xf = B.sc.xf;
                    mats = B.sc.mats;
                }
            } else {
                pix_ok = valid && tile_pixel(i, a.tiles_x, a.cam.width, a.band_h, x, y);
                pixel = y * a.cam.width + x;
                o = f3(a.cam.position[0], a.cam.position[1], a.cam.position[2]);
                if (pix_ok && slot >= 0) d = camera_dir(a.cam, x, y + a.row0, a.aa_ox, a.aa_oy);
            }
        } else if (valid) {
            const float4 q0 = a.q_in[2 * (size_t)i], q1 = a.q_in[2 * (size_t)i + 1];
            o = f3(q0.x, q0.y, q0.z);
            d = f3(q1.x, q1.y, q1.z);
            pixel = __float_as_int(q0.w);
            if (BATCH) {
                scene = __float_as_int(q1.w);
                const BatchScene& B = a.table[scene];
                shade = B.sc.shade;
                geo = B.sc.geo;
                xf = B.sc.xf;
                mats = B.sc.mats;
            }
        }
        const bool hit = pix_ok && slot >= 0;
        F3 P = {0, 0, 0}, N = {0, 0, 0};
        int mat = 0;
        bool refl = false;
        if (hit) {
            const float4 s0 = shade[3 * (size_t)slot], s1 = shade[3 * (size_t)slot + 1], s2 = shade[3 * (size_t)slot + 2];
            hit_surface(s0, s1, s2, xf, o, d, h.x, h.z, h.w, true, P, N);
            mat = __float_as_int(s0.w);
            refl = mats[mat].reflective != 0 && a.level < a.depth;
        }
        if (L0 && pix_ok) {
            a.levels[pixel] = hit ? 1 : 0;
            if (a.hit_ids) a.hit_ids[pixel] = hit ? __float_as_int(geo[3 * (size_t)slot + 2].y) : -1;
        } else if (!L0 && hit) {
            a.levels[pixel] = (uint8_t)(a.level + 1);
        }
        // slots of this round, claimed once per workgroup: hit surfaces, queued
        // rays (reflection, then the Fresnel ray of a transparent material) and
        // the shadow-ray count
        const bool transp = a.rayhit && hit && mats[mat].transparent != 0;
        const bool child = transp && a.level < a.depth;
        const uint64_t lt = lanemask_lt();
        const uint64_t mh = __ballot(hit), mr = __ballot(refl), mc = __ballot(child);
        int j = -1, k = -1, kt = -1;
        // a round in which no ray of the workgroup hit claims nothing (most
        // rounds of a level-0 pass over a frame of misses)
        if (__syncthreads_or(hit ? 1 : 0)) {
            int nl = hit ? (BATCH ? a.table[scene].sc.n_light : a.sc.n_light) : 0;
            for (int off = 32; off > 0; off >>= 1) nl += __shfl_xor(nl, off);
            const int w = (int)threadIdx.x >> 6;
            if (__lane_id() == 0) {
                s_cnt[0][w] = __popcll(mh);
                s_cnt[1][w] = __popcll(mr) + __popcll(mc);
                s_cnt[2][w] = nl;
            }
            __syncthreads();
            if (threadIdx.x == 0) {
                int* const ctrs[3] = {a.hits_ctr, a.q_out_ctr, a.shadow_ctr};
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    int tot = 0;
                    for (int kk = 0; kk < (int)(blockDim.x >> 6); ++kk) {
                        s_base[c][kk] = tot;
                        tot += s_cnt[c][kk];
                    }
                    const int b0 = tot > 0 ? atomicAdd(ctrs[c], tot) : 0;
                    for (int kk = 0; kk < (int)(blockDim.x >> 6); ++kk) s_base[c][kk] += b0;
                }
            }
            __syncthreads();
            j = hit ? s_base[0][w] + __popcll(mh & lt) : -1;
            const int qb = s_base[1][w];
            k = refl ? qb + __popcll(mr & lt) : -1;
            kt = child ? qb + __popcll(mr) + __popcll(mc & lt) : -1;
        }
        if (hit) {
            a.surf_out[2 * (size_t)j] = make_float4(P.x, P.y, P.z, __int_as_float(pixel));
            a.surf_out[2 * (size_t)j + 1] = make_float4(N.x, N.y, N.z, __int_as_float(pack_mat(mat, scene)));
        }
        if (refl && k < a.q_cap) {
            // src/RayTracer.cpp:350-355
            const F3 R = normalize3(reflect3(d, N));
            const F3 ro = add(P, scale(1e-3f, R));
            a.q_out[2 * (size_t)k] = make_float4(ro.x, ro.y, ro.z, __int_as_float(pixel));
            a.q_out[2 * (size_t)k + 1] = make_float4(R.x, R.y, R.z, __int_as_float(scene));
        }
        if (a.rayhit) {
            // Ray trees.  src/RayTracer.cpp:361-408: a transparent material's
            // second ray reflects d about the normal facing the ray (0 - N when
            // the ray leaves the material) from P + 1e-3 * that normal, with
            // K = fresnel(normalize(d), normal, n1, n2); the reference traces
            // this ray twice (:401-403), its colour t enters as K*t + (1-K)*t.
            if (L0) {
                if (pix_ok) a.rayhit[pixel] = hit ? j : -1;
            } else if (valid) {
                a.rayhit[i] = hit ? j : -1;
            }
            float K = 0.0f;
            F3 fo = {0, 0, 0}, fd = {0, 0, 0};
            if (transp) {
                const float amb = BATCH ? a.table[scene].sc.ambient : a.sc.ambient;
                const float eta = mats[mat].refractive_index;
                const F3 inc = normalize3(d);
                F3 nn = N;
                float n1 = eta, n2 = amb;
                if (dot3(inc, N) < 0.0f) {
                    n1 = amb;
                    n2 = eta;
                } else {
                    nn = f3(0.0f - N.x, 0.0f - N.y, 0.0f - N.z);
                }
                K = fresnel(inc, nn, n1, n2);
                fd = normalize3(reflect3(d, nn));
                fo = add(P, scale(1e-3f, nn));
            }
            if (child && kt < a.q_cap) {
                a.q_out[2 * (size_t)kt] = make_float4(fo.x, fo.y, fo.z, __int_as_float(pixel));
                a.q_out[2 * (size_t)kt + 1] = make_float4(fd.x, fd.y, fd.z, __int_as_float(scene));
            }
            if (hit)
                a.aux[j] = make_float4(K, __int_as_float(refl && k < a.q_cap ? k : -1),
                                       __int_as_float(child && kt < a.q_cap ? kt : -1), __int_as_float(transp ? 1 : 0));
        }
    }
}

// src/RayTracer.cpp:418-458 for one light (w = 1 makes the vec4 forms equal to
// the vec3 ones).
__device__ __forceinline__ void phong(const DevMaterial& m, const DevLight& l, F3 cam, F3 P, F3 N, float out[3]) {
    const F3 L = normalize3(sub(f3(l.position[0], l.position[1], l.position[2]), P));
    const F3 V = normalize3(sub(cam, P));
    const float dNL = dot3(N, L);
    const float s2 = 2.0f * dNL;
    const F3 R = normalize3(sub(f3(s2 * N.x, s2 * N.y, s2 * N.z), L));
    float dLN = dot3(L, N);
    dLN = dLN > 0.0f ? dLN : 0.0f;
    const float dVR = dot3(V, R);
    const bool sp = dVR > 0.0f;
    const float st = sp ? powf(dVR, m.shininess) : 0.0f;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float diff = (dLN * m.kd[k]) * l.color[k];
        out[k] = sp ? diff + (st * l.color[k]) * m.ks[k] : diff + 0.0f;
    }
}

struct PhongArgs {
    SceneView sc;
    DevCamera cam;
    const BatchScene* table;
    int level, npix, lstride;
    int nlev;                 // consecutive levels level .. level + nlev - 1 (chains; ray trees: 1)
    const int* count_ptr;     // [nlev] hits per level
    const float4* surf;       // first level's surfaces; the next levels' at surf_stride float4s apart
    const uint8_t* vis;       // likewise, vis_stride bytes apart
    long long surf_stride, vis_stride;
    float4* levcol;
    float4* hitcol;           // ray trees: colour per hit of this level instead of per pixel
};

// Phong of the hits of one or more levels (one launch, the levels one after
// another in the same threads: independent records, so no ordering needed).
template <bool BATCH>
__global__ __launch_bounds__(256) void k_phong(PhongArgs a) {
    for (int li = 0; li < a.nlev; ++li) {
        const int n = a.count_ptr[li];
        const int level = a.level + li;
        const float4* surf = a.surf + li * a.surf_stride;
        const uint8_t* vis = a.vis + li * a.vis_stride;
        for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
            const float4 s0 = surf[2 * (size_t)j], s1 = surf[2 * (size_t)j + 1];
            const F3 P = f3(s0.x, s0.y, s0.z), N = f3(s1.x, s1.y, s1.z);
            const int pixel = __float_as_int(s0.w), packed = __float_as_int(s1.w);
            const SceneView& sc = BATCH ? a.table[unpack_scene(packed)].sc : a.sc;
            const DevCamera& cm = BATCH ? a.table[unpack_scene(packed)].cam : a.cam;
            const F3 cam = f3(cm.position[0], cm.position[1], cm.position[2]);
            const DevMaterial& m = sc.mats[unpack_mat(packed)];
            const int L = sc.n_light;
            float acc[3] = {0.0f, 0.0f, 0.0f};
            for (int i = 0; i < L; ++i) {
                if (!vis[(size_t)j * a.lstride + i]) continue;
                float c[3];
                phong(m, sc.lights[i], cam, P, N, c);
#pragma unroll
                for (int k = 0; k < 3; ++k) acc[k] = acc[k] + c[k];
            }
            const float w = m.reflective ? m.reflectivity : 0.0f;
            if (a.hitcol) a.hitcol[j] = make_float4(acc[0], acc[1], acc[2], w);
            else a.levcol[(size_t)level * a.npix + pixel] = make_float4(acc[0], acc[1], acc[2], w);
        }
    }
}

// src/RayTracer.cpp:150-153: floor(c == 1 ? 255 : min(c*256, 255)); c*256 is
// exact in float, so the float form equals the reference's double form.
__device__ __forceinline__ uint8_t quantize(float c) {
    if (c == 1.0f) return 255;
    const float q = fminf(floorf(c * 256.0f), 255.0f);
    return q > 0.0f ? (uint8_t)q : 0;
}

struct ResolveArgs {
    int npix;
    const uint8_t* levels;
    const float4* levcol;
    uchar4* rgba;
    float* rgb;
    float4* acc;              // anti-aliasing: running sum of the pixel's samples
    int aa_mode;              // AA_NONE, AA_FIRST, AA_NEXT, AA_LAST
    float inv_n2;             // (1/n)^2, the final scale (AA_LAST)
};

// A pixel's final colour: quantised and written, or, with anti-aliasing
// (src/RayTracer.cpp:214-242), one sample of n x n added in (q, p) order to a
// zero vec4 whose sum is scaled by (1/n)^2 before it is quantised.  One sample
// per frame: the first stores 0 + c, the next ones add c, the last adds c,
// scales and writes the pixel.
__device__ __forceinline__ void pixel_out(const ResolveArgs& a, int p, float c0, float c1, float c2) {
    if (a.aa_mode != AA_NONE) {
        const float4 s = a.aa_mode == AA_FIRST ? make_float4(0.0f, 0.0f, 0.0f, 0.0f) : a.acc[p];
        c0 = s.x + c0;
        c1 = s.y + c1;
        c2 = s.z + c2;
        if (a.aa_mode != AA_LAST) {
            a.acc[p] = make_float4(c0, c1, c2, 0.0f);
            return;
        }
        c0 = c0 * a.inv_n2;
        c1 = c1 * a.inv_n2;
        c2 = c2 * a.inv_n2;
    }
    if (a.rgba) a.rgba[p] = make_uchar4(quantize(c0), quantize(c1), quantize(c2), 255);
    if (a.rgb) {
        a.rgb[3 * (size_t)p] = c0;
        a.rgb[3 * (size_t)p + 1] = c1;
        a.rgb[3 * (size_t)p + 2] = c2;
    }
}

// Chains of reflections: c_l = P_l + r_l * c_{l+1}, innermost first.
__global__ __launch_bounds__(256) void k_resolve(ResolveArgs a) {
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < a.npix; p += gridDim.x * blockDim.x) {
        const int nl = a.levels[p];
        float c0 = 0.0f, c1 = 0.0f, c2 = 0.0f;
        for (int l = nl - 1; l >= 0; --l) {
            const float4 col = a.levcol[(size_t)l * a.npix + p];
            c0 = col.x + col.w * c0;
            c1 = col.y + col.w * c1;
            c2 = col.z + col.w * c2;
        }
        pixel_out(a, p, c0, c1, c2);
    }
}

// Ray trees, one launch per level from the deepest up: a hit's colour is
// 0 + ((phong + r * c_reflect) + (0 + (K * t + (1 - K) * t))) with c_reflect
// and t the resolved colours of its children (0 for none), in place.
__global__ __launch_bounds__(256) void k_resolve_tree(const int* count_ptr, float4* hitcol, const float4* aux,
                                                      const int32_t* rayhit_next, const float4* hitcol_next) {
    const int n = *count_ptr;
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        const float4 h = hitcol[j], x = aux[j];
        const int kr = __float_as_int(x.y), kt = __float_as_int(x.z);
        float r0 = 0.0f, r1 = 0.0f, r2 = 0.0f, t0 = 0.0f, t1 = 0.0f, t2 = 0.0f;
        if (kr >= 0) {
            const int jj = rayhit_next[kr];
            if (jj >= 0) {
                const float4 c = hitcol_next[jj];
                r0 = c.x;
                r1 = c.y;
                r2 = c.z;
            }
        }
        if (kt >= 0) {
            const int jj = rayhit_next[kt];
            if (jj >= 0) {
                const float4 c = hitcol_next[jj];
                t0 = c.x;
                t1 = c.y;
                t2 = c.z;
            }
        }
        float c0 = h.x + h.w * r0, c1 = h.y + h.w * r1, c2 = h.z + h.w * r2;
        if (__float_as_int(x.w)) {
            const float K = x.x, L = 1.0f - K;
            c0 = c0 + (0.0f + (K * t0 + L * t0));
            c1 = c1 + (0.0f + (K * t1 + L * t1));
            c2 = c2 + (0.0f + (K * t2 + L * t2));
        }
        hitcol[j] = make_float4(c0, c1, c2, h.w);
    }
}

// Ray trees: the pixel's colour is its primary hit's resolved colour.
__global__ __launch_bounds__(256) void k_resolve_pixels(ResolveArgs a, const int32_t* pixhit, const float4* hitcol) {
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < a.npix; p += gridDim.x * blockDim.x) {
        const int j = pixhit[p];
        float c0 = 0.0f, c1 = 0.0f, c2 = 0.0f;
        if (j >= 0) {
            const float4 c = hitcol[j];
            c0 = c.x;
            c1 = c.y;
            c2 = c.z;
        }
        pixel_out(a, p, c0, c1, c2);
    }
}


// A queue counter past its capacity: clamp it (the appends beyond were
// dropped) and flag the frame.
__global__ void k_clamp_count(int* ctr, int cap, int* err) {
    if (threadIdx.x == 0 && *ctr > cap) {
        *ctr = cap;
        atomicOr(err, 4);
    }
}

__global__ void k_zero_i32(int* p, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

// ---- tile test of the primary rays --------------------------------------------
struct CullArgs {
    const Node4* nodes4;
    int root;                 // internal root node of the 4-wide tree
    double org[3];            // camera position
    double dx[3], dy[3], d0[3];   // unnormalised direction through pixel position (px, py): px dx + py dy + d0
    int width;
    int tiles_x, n_tiles;
    int row0, band_h;         // the frame covers camera rows [row0, row0 + band_h)
    int32_t* live;            // out: tiles left to trace, in tile order, [CTR_LIVE] of them
    int32_t* block_live;      // per 256 tiles: tiles left (k_cull_tiles), then their list offset (k_cull_scan)
    int32_t* block_culled;    // per 256 tiles: rays resolved (summed by the last workgroup of k_cull_compact)
    uint8_t* flags;           // out: per tile 1 = traced, 0 = resolved here
    int* live_ctr;
    int* culled_ctr;          // rays resolved here (pixels of the frame in resolved tiles)
    int scanned;              // k_cull_compact: block_live holds the scanned offsets (k_cull_scan ran)
};
constexpr int kCullBlock = eos_frame_buffers::kTileBlock;   // tiles per workgroup of the tile test
static_assert(kCullBlock == 256, "the tile test's kernels are written for four waves per workgroup");
constexpr int kCullSumBlocks = 4096;   // up to this many workgroups each sums the counts before its own

// Primary rays of an 8x8 tile leave the camera position through pixel
// positions (x + ox, y + oy), ox and oy in [0, 1) (camera_dir); their
// directions are positive multiples of a(px) right + b(py) up + view, which is
// affine in the pixel position, so every one lies in the pyramid spanned by
// the directions through the corners of [x0 - 1, x0 + 9] x [y0 - 1, y0 + 9]
// (the tile's pixels with a pixel to spare on each side).  A box lying wholly
// on the outer side of one of the pyramid's four side planes, or behind the
// camera's plane normal to the tile's centre direction, meets no ray of the
// tile at any t >= 0.  The test asks for a clearance of 1e-3 of the
// scene's scale (camera distance plus coordinate magnitudes) -- far beyond the
// rounding of a ray's direction and of the slab test's fma terms -- and is
// computed in double precision.  A tile whose rays miss all child boxes of the
// root node is resolved as misses: the per-ray walk would visit the root, find
// no child hit and end there with no hit.  Everything else is traced.
#ifndef EOS_CULL_REAL
#define EOS_CULL_REAL double
#endif
__global__ __launch_bounds__(256) void k_cull_tiles(CullArgs a) {
    using creal = EOS_CULL_REAL;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    bool traced = true;
    int culled = 0;
    if (t < a.n_tiles) {
        const int ty = t / a.tiles_x, tx = t - ty * a.tiles_x;
        const creal x0 = creal(tx * 8 - 1), x1 = creal(tx * 8 + 9);
        const creal y0 = creal(ty * 8 + a.row0 - 1), y1 = creal(ty * 8 + a.row0 + 9);
        creal D[4][3], C[3];
        auto dir = [&](creal px, creal py, creal* out) {
#pragma unroll
            for (int k = 0; k < 3; ++k) out[k] = fma(px, creal(a.dx[k]), fma(py, creal(a.dy[k]), creal(a.d0[k])));
        };
        dir(x0, y0, D[0]);
        dir(x1, y0, D[1]);
        dir(x1, y1, D[2]);
        dir(x0, y1, D[3]);
        dir(creal(0.5) * (x0 + x1), creal(0.5) * (y0 + y1), C);
        const creal O[3] = {creal(a.org[0]), creal(a.org[1]), creal(a.org[2])};
        // side planes through the camera, normals towards the pyramid's inside;
        // fifth, the plane through the camera normal to the centre direction when
        // every corner direction lies in front of it (then so does every ray)
        creal N[5][3];
        bool front = true;
#pragma unroll
        for (int k = 0; k < 4; ++k) front = front && D[k][0] * C[0] + D[k][1] * C[1] + D[k][2] * C[2] > creal(0);
#pragma unroll
        for (int c = 0; c < 3; ++c) N[4][c] = front ? C[c] : creal(0);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const creal* p = D[k];
            const creal* q = D[(k + 1) & 3];
            N[k][0] = p[1] * q[2] - p[2] * q[1];
            N[k][1] = p[2] * q[0] - p[0] * q[2];
            N[k][2] = p[0] * q[1] - p[1] * q[0];
            const creal sg = N[k][0] * C[0] + N[k][1] * C[1] + N[k][2] * C[2] < creal(0) ? creal(-1) : creal(1);
#pragma unroll
            for (int c = 0; c < 3; ++c) N[k][c] *= sg;
        }
        creal nn[5];
#pragma unroll
        for (int k = 0; k < 5; ++k) nn[k] = N[k][0] * N[k][0] + N[k][1] * N[k][1] + N[k][2] * N[k][2];
        // box [L, H] wholly outside one of the planes: its farthest corner lies
        // below the plane by more than 1e-3 of the scale (|n| far > 1e-3 scale
        // |n|, compared squared); the scale is the camera distance plus the
        // coordinate magnitudes of the box
        auto outside = [&](const creal* L, const creal* H) {
            creal scale = fabs(O[0]) + fabs(O[1]) + fabs(O[2]);
#pragma unroll
            for (int c = 0; c < 3; ++c)
                scale += fmax(fabs(L[c] - O[c]), fabs(H[c] - O[c])) + fmax(fabs(L[c]), fabs(H[c]));
            bool out = false;
#pragma unroll
            for (int k = 0; k < 5; ++k) {
                creal far = 0;
#pragma unroll
                for (int c = 0; c < 3; ++c) far = fma(N[k][c], (N[k][c] >= creal(0) ? H[c] : L[c]) - O[c], far);
                out = out || (far < creal(0) && far * far > creal(1e-6) * nn[k] * scale * scale);
            }
            return out;
        };
        const Node4& nd = a.nodes4[a.root];
        const float4 lo4[3] = {nd.lox, nd.loy, nd.loz}, hi4[3] = {nd.hix, nd.hiy, nd.hiz};
        const int4 ref = nd.ref;
        const bool used[4] = {ref.x != kEmpty, ref.y != kEmpty, ref.z != kEmpty, ref.w != kEmpty};
        creal CL[4][3], CH[4][3], UL[3], UH[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float l[4] = {lo4[c].x, lo4[c].y, lo4[c].z, lo4[c].w};
            const float h[4] = {hi4[c].x, hi4[c].y, hi4[c].z, hi4[c].w};
            UL[c] = creal(INFINITY);
            UH[c] = -creal(INFINITY);
#pragma unroll
            for (int ch = 0; ch < 4; ++ch) {
                CL[ch][c] = l[ch];
                CH[ch][c] = h[ch];
                if (used[ch]) {
                    UL[c] = fmin(UL[c], creal(l[ch]));
                    UH[c] = fmax(UH[c], creal(h[ch]));
                }
            }
        }
        // the union of the child boxes first (outside it with its own, larger
        // clearance means outside each child with its own), then each child
        bool all_missed = !(used[0] || used[1] || used[2] || used[3]) || outside(UL, UH);
        if (!all_missed) {
            all_missed = true;
#pragma unroll
            for (int ch = 0; ch < 4; ++ch) all_missed = all_missed && (!used[ch] || outside(CL[ch], CH[ch]));
        }
        traced = !all_missed;
        if (!traced) {
            const int w = min(8, a.width - tx * 8), h = min(8, a.band_h - ty * 8);
            culled = max(w, 0) * max(h, 0);
        }
        a.flags[t] = traced ? 1 : 0;
    }
    // tiles left and rays resolved in this workgroup's 256 tiles (the list of
    // tiles left keeps tile order: the primary launch's claim ranges stay image
    // bands, one per XCD); no atomics on one global address (one per wave
    // serialised to about 20 us over a 4K frame)
    __shared__ int s_n[4], s_c[4];
    const uint64_t m = __ballot(t < a.n_tiles && traced);
    for (int off = 32; off > 0; off >>= 1) culled += __shfl_xor(culled, off);
    if (__lane_id() == 0) {
        s_n[threadIdx.x >> 6] = __popcll(m);
        s_c[threadIdx.x >> 6] = culled;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        a.block_live[blockIdx.x] = s_n[0] + s_n[1] + s_n[2] + s_n[3];
        a.block_culled[blockIdx.x] = s_c[0] + s_c[1] + s_c[2] + s_c[3];
    }
}

// Exclusive scan of the workgroups' counts (one workgroup of 1024 threads,
// 1024 counts per step) and the list length.
__global__ __launch_bounds__(1024) void k_cull_scan(int32_t* block_live, int nblocks, int* live_ctr) {
    __shared__ int s_wave[16];
    __shared__ int s_total;
    const int tid = threadIdx.x, lane = __lane_id(), w = tid >> 6;
    int running = 0;
    for (int base = 0; base < nblocks; base += 1024) {
        const int i = base + tid;
        const int v = i < nblocks ? block_live[i] : 0;
        int x = v;   // inclusive scan in the wave
        for (int off = 1; off < 64; off <<= 1) {
            const int y = __shfl_up(x, off);
            if (lane >= off) x += y;
        }
        if (lane == 63) s_wave[w] = x;
        __syncthreads();
        if (tid == 0) {
            int acc = 0;
            for (int k = 0; k < 16; ++k) {
                const int c = s_wave[k];
                s_wave[k] = acc;
                acc += c;
            }
            s_total = acc;
        }
        __syncthreads();
        if (i < nblocks) block_live[i] = running + s_wave[w] + x - v;
        running += s_total;
        __syncthreads();
    }
    if (tid == 0) *live_ctr = running;
}

// The list of tiles left, in tile order: workgroup b writes its tiles from
// its scanned offset on.
__global__ __launch_bounds__(256) void k_cull_compact(CullArgs a) {
    __shared__ int s_wave[4];
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const bool traced = t < a.n_tiles && a.flags[t] != 0;
    const uint64_t m = __ballot(traced);
    const int w = threadIdx.x >> 6;
    if (__lane_id() == 0) s_wave[w] = __popcll(m);
    __syncthreads();
    // this workgroup's offset: scanned, or the sum of the counts before it
    __shared__ int s_part[4];
    int base;
    if (a.scanned) {
        base = a.block_live[blockIdx.x];
    } else {
        int v = 0;
        for (int b = threadIdx.x; b < (int)blockIdx.x; b += blockDim.x) v += a.block_live[b];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        if (__lane_id() == 0) s_part[w] = v;
        __syncthreads();
        base = s_part[0] + s_part[1] + s_part[2] + s_part[3];
        if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0)
            *a.live_ctr = base + s_wave[0] + s_wave[1] + s_wave[2] + s_wave[3];
    }
    if (blockIdx.x == gridDim.x - 1) {
        // the rays resolved over the frame
        int c = 0;
        for (int b = threadIdx.x; b < (int)gridDim.x; b += blockDim.x) c += a.block_culled[b];
        for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
        if (__lane_id() == 0 && c != 0) atomicAdd(a.culled_ctr, c);
    }
    int off = base;
    for (int k = 0; k < w; ++k) off += s_wave[k];
    if (traced) a.live[off + __popcll(m & lanemask_lt())] = t;
}

// ---- traversal launch table ------------------------------------------------
// 4-wide trees: every workgroup / LDS-stack configuration; binary trees: the
// default configuration only (they exist for comparison and parity checks).
// Scenes with transformations: the default configuration (256 threads, 16 LDS
// stack entries), no packets; the frame sets its configuration to match.
// lean (shadow rays, single scenes): the default configuration's kernel without
// the group walk (less LDS per workgroup, for a launch that shares the GPU with
// another stream's).
template <int MODE, bool COUNT, bool BATCH>
eos_status launch_trace(const LaunchCfg& cfg, int width, bool xf, int grid, hipStream_t s, const TraceArgs& a,
                        bool lean) {
#define EOS_TRACE_CASE(WG, SL)                                                                             \
    case WG * 100 + SL:                                                                                     \
        hipLaunchKernelGGL((k_trace<MODE, WG, SL, COUNT, BATCH, 4, false>), dim3(grid), dim3(WG), 0, s, a); \
        break;
    if (xf) {
        if (cfg.workgroup != 256 || cfg.stack_lds != 16 || cfg.packets)
            return fail(EOS_ERR_INVALID, "transformed scenes trace with the default launch configuration");
        if (width == 2) hipLaunchKernelGGL((k_trace<MODE, 256, 16, COUNT, BATCH, 2, true>), dim3(grid), dim3(256), 0, s, a);
        else hipLaunchKernelGGL((k_trace<MODE, 256, 16, COUNT, BATCH, 4, true>), dim3(grid), dim3(256), 0, s, a);
    } else if (MODE == MODE_PRIMARY && width == 4 && cfg.packets) {
        hipLaunchKernelGGL((k_trace_packet<256, COUNT, BATCH>), dim3(std::max(1, grid * cfg.workgroup / 256)),
                           dim3(256), 0, s, a);
    } else if (width == 2) {
        hipLaunchKernelGGL((k_trace<MODE, 256, 16, COUNT, BATCH, 2, false>), dim3(grid), dim3(256), 0, s, a);
    } else if (COUNT) {
        hipLaunchKernelGGL((k_trace<MODE, 256, 16, COUNT, BATCH, 4, false>), dim3(grid), dim3(256), 0, s, a);
    } else if (MODE == MODE_SHADOW && !BATCH && lean && cfg.workgroup == 256 && cfg.stack_lds == 16) {
        if constexpr (MODE == MODE_SHADOW && !BATCH)
            hipLaunchKernelGGL((k_trace<MODE, 256, 16, false, false, 4, false, true>), dim3(grid), dim3(256), 0, s, a);
    } else {
        switch (cfg.workgroup * 100 + cfg.stack_lds) {
            EOS_TRACE_CASE(64, 8)
            EOS_TRACE_CASE(64, 16)
            EOS_TRACE_CASE(64, 32)
            EOS_TRACE_CASE(128, 8)
            EOS_TRACE_CASE(128, 16)
            EOS_TRACE_CASE(128, 32)
            EOS_TRACE_CASE(256, 8)
            EOS_TRACE_CASE(256, 16)
            EOS_TRACE_CASE(256, 32)
            default:
                return fail(EOS_ERR_INVALID, "unsupported launch configuration");
        }
    }
#undef EOS_TRACE_CASE
    return EOS_OK;
}

template <int MODE>
eos_status launch_trace_any(bool count, bool batch, int width, bool xf, const LaunchCfg& cfg, int grid, hipStream_t s,
                            const TraceArgs& a, bool lean = false) {
    if (batch) return launch_trace<MODE, false, true>(cfg, width, xf, grid, s, a, false);
    return count ? launch_trace<MODE, true, false>(cfg, width, xf, grid, s, a, false)
                 : launch_trace<MODE, false, false>(cfg, width, xf, grid, s, a, lean);
}

inline unsigned blocks_for(size_t n, unsigned t, unsigned cap) {
    size_t b = (n + t - 1) / t;
    if (b < 1) b = 1;
    return (unsigned)std::min<size_t>(b, cap);
}

// One frame: a single scene, or a batch of scenes described by a device table.
struct Frame {
    bool batch = false;
    SceneView sc{};
    DevCamera cam{};
    int tiles_x = 0;
    const BatchScene* table = nullptr;
    int n_scenes = 1;
    size_t npix = 0, n_ids = 0;
    int depth = 0;
    int lstride = 0;       // light slots per hit (max lights over the batch)
    int width = 4;         // tree width walked by the traversal
    bool xf = false;       // some triangle has a transformation
    bool tree = false;     // some material is transparent: rays may have two children
    AaSample aa;           // anti-aliasing sample of this frame (mode AA_NONE: plain frame)
    int row0 = 0, band_h = 0;  // single scene: camera rows [row0, row0 + band_h) of the frame
    hipEvent_t resolve_after = nullptr;   // frames in flight: resolve after this event (the previous frame's)
    hipEvent_t resolve_done = nullptr;    // and record this one after it
    hipEvent_t first_ev[2] = {nullptr, nullptr};   // frames in flight: around the primary-ray launch
    bool count = false;
    LaunchCfg cfg;
    int num_cu = 256;
    uchar4* rgba = nullptr;
    float* rgb = nullptr;
    int32_t* hits = nullptr;
};

// Overflow stack entries of one persistent traversal grid.
size_t gstack_ints(const Frame& f, int trace_grid) {
    const int maxs = f.width == 4 ? kMaxStack4 : kMaxStack;
    return (size_t)trace_grid * f.cfg.workgroup * (maxs - f.cfg.stack_lds);
}

struct FrameEvents {
    int start, end, resolve0;
    // per level l: [2l] before, [2l+1] after the launch
    int c[2 * EOS_MAX_DEPTH + 2];  // closest-hit trace (main stream)
    int u[2 * EOS_MAX_DEPTH + 2];  // hit setup (main stream)
    int s[2 * EOS_MAX_DEPTH + 2];  // shadow trace: chains level 0, levels 1..depth; trees per level
    int h[2 * EOS_MAX_DEPTH + 2];  // Phong (second stream)
};

// The streams a frame runs on.  Level 0's shadow rays (most of a frame's
// any-hit work) and its Phong run on S2 after a fork event recorded behind
// level 0's setup, overlapping the reflection levels' closest-hit traces on S;
// the shadow rays of levels 1..depth follow on S in one launch after the last
// setup; S joins S2 before the resolve.  The two traversals that can be in
// flight together use disjoint halves of the overflow stack area.
struct FrameStreams {
    hipStream_t S = nullptr;
    hipStream_t S2 = nullptr;       // null: everything in order on S
    hipEvent_t* sync = nullptr;     // [2]: fork, join
};

// Largest queue of a ray tree level (1 << 26 rays, or a frame's pixels if more).
constexpr size_t kTreeCap = (size_t)1 << 26;

// Traversal launches number their items with 32-bit ints: primary ray ids, and
// a level's shadow rays (its surfaces times the light slots).
constexpr size_t kMaxItems = (size_t)INT_MAX;

// Levels 1.. of a chain frame trace their shadow rays in launches of this many
// consecutive levels (all of them unless their items could pass kMaxItems).
int chain_shadow_span(const Frame& f) {
    const size_t per_level = f.npix * (size_t)std::max(f.lstride, 1);
    return (int)std::min<size_t>((size_t)std::max(f.depth, 1), std::max<size_t>(1, kMaxItems / per_level));
}

eos_status ensure_buffers(eos_frame_buffers& fb, const Frame& f, int trace_grid) {
    if (f.n_ids > kMaxItems || f.npix * (size_t)std::max(f.lstride, 1) > kMaxItems)
        return fail(EOS_ERR_LIMIT, "frame too large: pixels times light slots must stay below 2^31 (" +
                                       std::to_string(f.npix) + " pixels, " + std::to_string(f.lstride) + " lights)");
    size_t cap = f.npix;
    fb.tree = f.tree;
    if (f.tree) {
        // level l holds at most twice the rays of level l - 1 (two children per hit),
        // and its shadow rays (queue times light slots) stay within kMaxItems
        const size_t lim = std::min(std::max(f.npix, kTreeCap), kMaxItems / (size_t)std::max(f.lstride, 1));
        fb.toff[0] = 0;
        for (int l = 0; l <= f.depth; ++l) {
            fb.tcap[l] = l == 0 ? f.npix : std::min(2 * fb.tcap[l - 1], lim);
            fb.toff[l + 1] = fb.toff[l] + fb.tcap[l];
            cap = std::max(cap, fb.tcap[l]);
        }
        EOS_TRY(fb.rayhit.ensure(4 * fb.toff[f.depth + 1]));
        EOS_TRY(fb.aux.ensure(16 * fb.toff[f.depth + 1]));
        EOS_TRY(fb.hitcol.ensure(16 * fb.toff[f.depth + 1]));
    }
    const int surf_levels = f.tree ? 1 : f.depth + 1;   // trees shade one level at a time
    EOS_TRY(fb.hit0.ensure(16 * f.n_ids));
    EOS_TRY(fb.q[0].ensure(32 * cap));
    EOS_TRY(fb.q[1].ensure(32 * cap));
    EOS_TRY(fb.hitq.ensure(16 * cap));
    fb.cap = cap;
    fb.lslots = std::max(f.lstride, 1);
    EOS_TRY(fb.surf.ensure(16 * fb.surf_stride() * (size_t)surf_levels));
    EOS_TRY(fb.vis.ensure(fb.vis_stride() * (size_t)surf_levels));
    EOS_TRY(fb.levcol.ensure(16 * f.npix * (size_t)(f.depth + 1)));
    EOS_TRY(fb.levels.ensure(f.npix));
    EOS_TRY(fb.counters.ensure(4 * (size_t)kCounterInts));
    EOS_TRY(fb.gstack.ensure(2 * gstack_ints(f, trace_grid) * 4));
    EOS_TRY(fb.tiles.ensure(eos_frame_buffers::tiles_bytes(f.n_ids)));
    fb.npix = f.npix;
    fb.n_ids = f.n_ids;
    fb.depth = f.depth;
    fb.n_light = f.lstride;
    return EOS_OK;
}

// Enqueues every launch of one frame.  Timing events are recorded when E is
// non-null (not while a graph is being captured).
eos_status enqueue_frame(eos_frame_buffers& fb, const Frame& f, const FrameStreams& fs, hipEvent_t* E,
                         FrameEvents* fe) {
    const int trace_grid = std::max(1, f.num_cu * f.cfg.waves_per_cu / (f.cfg.workgroup / 64));
    int* ctr = fb.counters.as<int>();
    unsigned long long* stats = reinterpret_cast<unsigned long long*>(fb.counters.as<char>() + 4 * CTR_COUNT);
    const unsigned shade_grid = (unsigned)f.num_cu * 8;
    const hipStream_t S = fs.S;
    const bool two = fs.S2 != nullptr;
    const hipStream_t S2 = two ? fs.S2 : fs.S;
    int ev = 0;
    auto rec = [&](hipStream_t st) -> int {
        if (E) (void)hipEventRecord(E[ev], st);
        return ev++;
    };
    hipLaunchKernelGGL(k_zero_i32, dim3(16), dim3(256), 0, S, ctr, kCounterInts);
    // fetch cursors of launch slot k (closest level l: k = l; any hit from level l: k = 16 + l)
    auto cursors = [&](int k) { return ctr + kCursorBase + k * kRanges * kCursorStride; };
    FrameEvents local;
    FrameEvents& F = fe ? *fe : local;
    // primary tile test: single scenes, 4-wide trees with an internal root node,
    // launches that do not count visits (their counts restate every ray's walk)
    const bool cull = f.cfg.cull != 0 && !f.batch && !f.count && f.width == 4 && !f.cfg.packets &&
                      f.sc.root4 >= 0 && f.sc.root4 != kDone && f.n_ids >= 64;
    F.start = rec(S);
    // Any-hit trace of levels [l0, l1] in one launch, then Phong of each level.
    auto shadow_and_phong = [&](int l0, int l1, hipStream_t st, size_t gstack_off, int slot) -> eos_status {
        F.s[2 * slot] = rec(st);
        if (f.lstride > 0) {
            TraceArgs sh{};
            sh.sc = f.sc;
            sh.cam = f.cam;
            sh.tiles_x = f.tiles_x;
            sh.table = f.table;
            sh.n_scenes = f.n_scenes;
            sh.gstack = fb.gstack.as<int32_t>() + gstack_off;
            sh.err = ctr + CTR_ERR;
            sh.refill = f.cfg.refill;
            sh.batch = f.cfg.batch;
            sh.coop = f.cfg.coop;
            sh.group = f.cfg.group;
            sh.group_lanes = f.cfg.group_lanes;
            sh.steal = f.count ? 0 : f.cfg.steal;
            sh.scatter = f.cfg.scatter;
            sh.cursor = cursors(16 + l0);
            sh.stats = stats + 2 * (16 + l0);
            sh.count_ptr = ctr + CTR_HITS + l0;
            sh.count_mul = f.lstride;
            sh.nlev = l1 - l0 + 1;
            sh.surf = fb.surf_at(f.tree ? 0 : l0);
            sh.vis_out = fb.vis_at(f.tree ? 0 : l0);
            sh.surf_stride = (long long)fb.surf_stride();
            sh.vis_stride = (long long)fb.vis_stride();
            // level 0's shadow launch on the second stream shares the CUs with the
            // reflection chain: it takes the kernel without the group walk, whose
            // 24 KiB of LDS per workgroup (32 KiB with it) leave the chain's
            // workgroups room (hit-dense 4K frame 17.3 -> 16.5 ms, config E
            // 1.84 -> 1.80 ms); its launches are large, so the group walk would
            // not be taken there anyway
            EOS_TRY(launch_trace_any<MODE_SHADOW>(f.count, f.batch, f.width, f.xf, f.cfg, trace_grid, st, sh,
                                                  st != S));
        }
        F.s[2 * slot + 1] = rec(st);
        // Phong of levels [l0, l1] in one launch (ray trees shade one level at a time)
        PhongArgs pa{};
        pa.sc = f.sc;
        pa.cam = f.cam;
        pa.table = f.table;
        pa.level = l0;
        pa.nlev = l1 - l0 + 1;
        pa.npix = (int)f.npix;
        pa.lstride = f.lstride;
        pa.count_ptr = ctr + CTR_HITS + l0;
        pa.surf = fb.surf_at(f.tree ? 0 : l0);
        pa.vis = fb.vis_at(f.tree ? 0 : l0);
        pa.surf_stride = (long long)fb.surf_stride();
        pa.vis_stride = (long long)fb.vis_stride();
        pa.levcol = fb.levcol.as<float4>();
        pa.hitcol = f.tree ? fb.hitcol.as<float4>() + fb.toff[l0] : nullptr;
        F.h[2 * l0] = rec(st);
        if (f.batch) hipLaunchKernelGGL(k_phong<true>, dim3(shade_grid), dim3(256), 0, st, pa);
        else hipLaunchKernelGGL(k_phong<false>, dim3(shade_grid), dim3(256), 0, st, pa);
        F.h[2 * l0 + 1] = rec(st);
        for (int l = l0 + 1; l <= l1; ++l) F.h[2 * l] = F.h[2 * l + 1] = F.h[2 * l0 + 1];   // timed with level l0
        return EOS_OK;
    };
    for (int l = 0; l <= f.depth; ++l) {
        float4* qin = fb.q[l & 1].as<float4>();
        float4* qout = fb.q[(l + 1) & 1].as<float4>();
        TraceArgs ta{};
        ta.sc = f.sc;
        ta.cam = f.cam;
        ta.aa_ox = f.aa.ox;
        ta.aa_oy = f.aa.oy;
        ta.row0 = f.row0;
        ta.band_h = f.band_h;
        ta.tiles_x = f.tiles_x;
        ta.table = f.table;
        ta.n_scenes = f.n_scenes;
        ta.gstack = fb.gstack.as<int32_t>();
        ta.err = ctr + CTR_ERR;
        ta.refill = f.cfg.refill;
        ta.batch = f.cfg.batch;
        ta.coop = f.cfg.coop;
        ta.group = f.cfg.group;
        ta.group_lanes = f.cfg.group_lanes;
        ta.steal = f.count ? 0 : f.cfg.steal;
        ta.scatter = f.cfg.scatter;
        ta.cursor = cursors(l);
        ta.stats = stats + 2 * l;
        if (l == 0) {
            ta.n_static = (int)f.n_ids;
            ta.count_ptr = nullptr;
            ta.count_mul = 1;
            ta.hit_out = fb.hit0.as<float4>();
        } else {
            ta.count_ptr = ctr + CTR_Q + l;
            ta.count_mul = 1;
            ta.q = qin;
            ta.hit_out = fb.hitq.as<float4>();
        }
        F.c[2 * l] = rec(S);
        if (l == 0 && f.first_ev[0]) EOS_HIP(hipEventRecord(f.first_ev[0], S));
        if (l == 0 && cull) {
            // the tile test (timed with the primary launch): the launch then numbers
            // only the tiles left, counted on the device
            const int n_tiles = (int)(f.n_ids / 64);
            const int nblocks = (n_tiles + kCullBlock - 1) / kCullBlock;
            CullArgs ca{};
            ca.nodes4 = f.sc.nodes4;
            ca.root = f.sc.root4;
            // camera_dir's direction before normalisation, affine in the pixel position:
            // ((px - w/2) / (w/2)) right + ((py - h/2) / (h/2)) up + view
            const double nw = f.cam.width / 2.0, nh = f.cam.height / 2.0;
            for (int c = 0; c < 3; ++c) {
                ca.org[c] = f.cam.position[c];
                ca.dx[c] = (double)f.cam.right[c] / nw;
                ca.dy[c] = (double)f.cam.up[c] / nh;
                ca.d0[c] = (double)f.cam.view[c] - (double)f.cam.right[c] - (double)f.cam.up[c];
            }
            ca.width = f.cam.width;
            ca.tiles_x = f.tiles_x;
            ca.n_tiles = n_tiles;
            ca.row0 = f.row0;
            ca.band_h = f.band_h;
            ca.live = fb.tiles.as<int32_t>();
            ca.block_live = ca.live + n_tiles;
            ca.block_culled = ca.block_live + nblocks;
            ca.flags = fb.tile_flags(f.n_ids);
            ca.live_ctr = ctr + CTR_LIVE;
            ca.culled_ctr = ctr + CTR_CULLED;
            ca.scanned = nblocks > kCullSumBlocks;
            hipLaunchKernelGGL(k_cull_tiles, dim3(nblocks), dim3(kCullBlock), 0, S, ca);
            if (ca.scanned)
                hipLaunchKernelGGL(k_cull_scan, dim3(1), dim3(1024), 0, S, ca.block_live, nblocks, ca.live_ctr);
            hipLaunchKernelGGL(k_cull_compact, dim3(nblocks), dim3(kCullBlock), 0, S, ca);
            ta.count_ptr = ctr + CTR_LIVE;
            ta.count_mul = 64;
            ta.live = ca.live;
        }
        if (l == 0) EOS_TRY(launch_trace_any<MODE_PRIMARY>(f.count, f.batch, f.width, f.xf, f.cfg, trace_grid, S, ta));
        else EOS_TRY(launch_trace_any<MODE_QUEUE>(f.count, f.batch, f.width, f.xf, f.cfg, trace_grid, S, ta));
        if (l == 0 && f.first_ev[1]) EOS_HIP(hipEventRecord(f.first_ev[1], S));
        F.c[2 * l + 1] = rec(S);
        SetupArgs sa{};
        sa.sc = f.sc;
        sa.cam = f.cam;
        sa.aa_ox = f.aa.ox;
        sa.aa_oy = f.aa.oy;
        sa.row0 = f.row0;
        sa.band_h = f.band_h;
        sa.tiles_x = f.tiles_x;
        sa.table = f.table;
        sa.n_scenes = f.n_scenes;
        sa.n_ids = (int)f.n_ids;
        sa.depth = f.depth;
        sa.level = l;
        sa.count_ptr = ctr + CTR_Q + l;
        sa.q_in = qin;
        sa.hit_in = l == 0 ? fb.hit0.as<float4>() : fb.hitq.as<float4>();
        sa.surf_out = fb.surf_at(f.tree ? 0 : l);
        sa.hits_ctr = ctr + CTR_HITS + l;
        sa.q_out = qout;
        sa.q_out_ctr = ctr + CTR_Q + l + 1;
        sa.q_cap = (int)(f.tree ? (l < f.depth ? fb.tcap[l + 1] : 0) : fb.cap);
        if (f.tree) {
            sa.rayhit = fb.rayhit.as<int32_t>() + fb.toff[l];
            sa.aux = fb.aux.as<float4>() + fb.toff[l];
        }
        sa.shadow_ctr = ctr + CTR_SHADOW + l;
        sa.levels = fb.levels.as<uint8_t>();
        sa.hit_ids = l == 0 ? f.hits : nullptr;
        sa.tile_live = l == 0 && cull ? fb.tile_flags(f.n_ids) : nullptr;
        F.u[2 * l] = rec(S);
        if (l == 0) {
            if (f.batch)
                hipLaunchKernelGGL((k_setup<true, true>), dim3(blocks_for(f.n_ids, kSetupWG, shade_grid)), dim3(kSetupWG), 0, S, sa);
            else
                hipLaunchKernelGGL((k_setup<true, false>), dim3(blocks_for(f.n_ids, kSetupWG, shade_grid)), dim3(kSetupWG), 0, S, sa);
        } else {
            if (f.batch) hipLaunchKernelGGL((k_setup<false, true>), dim3(shade_grid), dim3(kSetupWG), 0, S, sa);
            else hipLaunchKernelGGL((k_setup<false, false>), dim3(shade_grid), dim3(kSetupWG), 0, S, sa);
        }
        F.u[2 * l + 1] = rec(S);
        if (f.tree) {
            // ray trees: one level's surfaces at a time, shaded before the next level is set up
            if (l < f.depth)
                hipLaunchKernelGGL(k_clamp_count, dim3(1), dim3(64), 0, S, ctr + CTR_Q + l + 1, (int)fb.tcap[l + 1],
                                   ctr + CTR_ERR);
            EOS_TRY(shadow_and_phong(l, l, S, 0, l));
        } else if (l == 0) {
            // level 0's shadow rays and shading overlap the reflection chain
            if (two) {
                EOS_HIP(hipEventRecord(fs.sync[0], S));
                EOS_HIP(hipStreamWaitEvent(S2, fs.sync[0], 0));
            }
            EOS_TRY(shadow_and_phong(0, 0, S2, two ? gstack_ints(f, trace_grid) : 0, 0));
        }
    }
    // deeper levels' shadow rays in one launch once the last level is set up
    // (several when a frame's levels hold more shadow rays than 32-bit item ids
    // number; launch slot l0 times the one that starts at level l0 >= 2)
    if (!f.tree && f.depth >= 1) {
        const int span = chain_shadow_span(f);
        for (int l0 = 1; l0 <= f.depth; l0 += span)
            EOS_TRY(shadow_and_phong(l0, std::min(f.depth, l0 + span - 1), S, 0, l0));
    }
    if (two && !f.tree) {
        EOS_HIP(hipEventRecord(fs.sync[1], S2));
        EOS_HIP(hipStreamWaitEvent(S, fs.sync[1], 0));
    }
    ResolveArgs ra{(int)f.npix, fb.levels.as<uint8_t>(), fb.levcol.as<float4>(), f.rgba, f.rgb,
                   f.aa.acc, f.aa.mode, f.aa.inv_n2};
    if (f.resolve_after) EOS_HIP(hipStreamWaitEvent(S, f.resolve_after, 0));
    F.resolve0 = rec(S);
    if (f.tree) {
        float4* hc = fb.hitcol.as<float4>();
        const float4* ax = fb.aux.as<const float4>();
        const int32_t* rh = fb.rayhit.as<const int32_t>();
        for (int l = f.depth; l >= 0; --l)
            hipLaunchKernelGGL(k_resolve_tree, dim3(blocks_for(fb.tcap[l], 256, shade_grid)), dim3(256), 0, S,
                               ctr + CTR_HITS + l, hc + fb.toff[l], ax + fb.toff[l], rh + fb.toff[l + 1],
                               hc + fb.toff[l + 1]);
        hipLaunchKernelGGL(k_resolve_pixels, dim3(blocks_for(f.npix, 256, shade_grid)), dim3(256), 0, S, ra, rh, hc);
    } else {
        hipLaunchKernelGGL(k_resolve, dim3(blocks_for(f.npix, 256, shade_grid)), dim3(256), 0, S, ra);
    }
    if (f.resolve_done) EOS_HIP(hipEventRecord(f.resolve_done, S));
    F.end = rec(S);
    EOS_HIP(hipGetLastError());
    return EOS_OK;
}

// Reads the frame's counters (synchronises the stream) into stats.
// Counters of a frame as the host reads them: int counters, then the 64-bit
// visit / test statistics (the device layout of the counter block's start).
constexpr size_t kCounterBytes = 4 * CTR_COUNT + 8 * kStats;

eos_status frame_stats(const Frame& f, const int* hctr, const unsigned long long* hstats, hipEvent_t* E,
                       const FrameEvents* F, eos_render_stats* st);

eos_status finish_frame(eos_frame_buffers& fb, const Frame& f, hipStream_t S, hipEvent_t* E, const FrameEvents* F,
                        eos_render_stats* st) {
    int hctr[CTR_COUNT];
    unsigned long long hstats[kStats];
    int* ctr = fb.counters.as<int>();
    unsigned long long* stats = reinterpret_cast<unsigned long long*>(fb.counters.as<char>() + 4 * CTR_COUNT);
    EOS_HIP(hipMemcpyAsync(hctr, ctr, sizeof hctr, hipMemcpyDeviceToHost, S));
    EOS_HIP(hipMemcpyAsync(hstats, stats, sizeof hstats, hipMemcpyDeviceToHost, S));
    EOS_HIP(hipStreamSynchronize(S));
    return frame_stats(f, hctr, hstats, E, F, st);
}

eos_status frame_stats(const Frame& f, const int* hctr, const unsigned long long* hstats, hipEvent_t* E,
                       const FrameEvents* F, eos_render_stats* st) {
    if (hctr[CTR_ERR] & 1) return fail(EOS_ERR_LIMIT, "traversal stack overflow (tree deeper than 64)");
    if (hctr[CTR_ERR] & 2) return fail(EOS_ERR_STATE, "traversal exceeded its visit bound (corrupt BVH)");
    if (hctr[CTR_ERR] & 4)
        return fail(EOS_ERR_LIMIT, "ray tree larger than the queue capacity (2^26 rays per level): lower the depth");
    if (!st) return EOS_OK;
    std::memset(st, 0, sizeof *st);
    st->primary_rays = (int64_t)f.npix;
    for (int l = 0; l <= f.depth; ++l) {
        st->hits += hctr[CTR_HITS + l];
        if (l >= 1) st->secondary_rays += hctr[CTR_Q + l];
    }
    for (int l = 0; l <= f.depth; ++l) st->shadow_rays += hctr[CTR_SHADOW + l];
    st->primary_culled = hctr[CTR_CULLED];
    if (E && F) {
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, E[F->start], E[F->end]);
        st->render_ms = ms;
        for (int l = 0; l <= f.depth; ++l) {
            float a = 0.f, c = 0.f, d2 = 0.f;
            (void)hipEventElapsedTime(&a, E[F->c[2 * l]], E[F->c[2 * l + 1]]);
            (void)hipEventElapsedTime(&c, E[F->u[2 * l]], E[F->u[2 * l + 1]]);
            (void)hipEventElapsedTime(&d2, E[F->h[2 * l]], E[F->h[2 * l + 1]]);
            st->trace_closest_ms += a;
            st->shade_ms += c + d2;
            if (l == 0) st->trace_first_ms = a;
        }
        const int step = f.tree ? 1 : chain_shadow_span(f);
        for (int k = 0; k <= f.depth; k = k == 0 ? 1 : k + step) {
            float b = 0.f;
            (void)hipEventElapsedTime(&b, E[F->s[2 * k]], E[F->s[2 * k + 1]]);
            st->trace_shadow_ms += b;
        }
        float r = 0.f;
        (void)hipEventElapsedTime(&r, E[F->resolve0], E[F->end]);
        st->shade_ms += r;
    }
    st->trace_closest_launches = f.depth + 1;
    st->trace_shadow_launches =
        f.lstride > 0 ? (f.tree ? f.depth + 1 : 1 + (f.depth + chain_shadow_span(f) - 1) / chain_shadow_span(f)) : 0;
    if (f.count) {
        for (int l = 0; l <= f.depth; ++l) {
            st->node_visits += (int64_t)(hstats[2 * l] + hstats[2 * (16 + l)]);
            st->tri_tests += (int64_t)(hstats[2 * l + 1] + hstats[2 * (16 + l) + 1]);
        }
        st->closest_visits_first = (int64_t)hstats[0];
        st->closest_tests_first = (int64_t)hstats[1];
    }
    return EOS_OK;
}

}  // namespace

// Frame driver shared by eos_render and the adjoint's re-trace.
namespace {
// A single-scene frame description (camera, scene view, launch shape).
Frame make_frame(eos_scene* s, const eos_camera* camp, int depth, bool count, uchar4* rgba_dev, float* rgb_dev,
                 int32_t* hits_dev, const AaSample& aa, int row0 = 0, int row1 = -1) {
    Frame f;
    f.aa = aa;
    f.cam = *reinterpret_cast<const DevCamera*>(camp);
    f.sc = view_of(s);
    f.row0 = row0;
    f.band_h = (row1 < 0 ? f.cam.height : row1) - row0;
    f.tiles_x = (f.cam.width + 7) / 8;
    const int tiles_y = (f.band_h + 7) / 8;
    f.npix = (size_t)f.cam.width * f.band_h;
    f.n_ids = (size_t)f.tiles_x * tiles_y * 64;
    f.depth = depth;
    f.lstride = s->n_light;
    f.count = count;
    f.cfg = s->cfg;
    f.width = s->width;
    f.xf = s->n_xf > 0;
    f.tree = s->any_transparent;
    if (count || f.width == 2 || f.xf) {
        f.cfg.workgroup = 256;
        f.cfg.stack_lds = 16;
    }
    if (f.xf) f.cfg.packets = 0;
    f.num_cu = s->num_cu;
    f.rgba = rgba_dev;
    f.rgb = rgb_dev;
    f.hits = hits_dev;
    return f;
}

int trace_grid_of(const Frame& f) { return std::max(1, f.num_cu * f.cfg.waves_per_cu / (f.cfg.workgroup / 64)); }
}  // namespace

eos_status run_frame(eos_scene* s, const eos_camera* camp, int depth, bool count, uchar4* rgba_dev,
                     float* rgb_dev, int32_t* hits_dev, eos_render_stats* st, const AaSample& aa, int row0, int row1,
                     bool launch_times) {
    Frame f = make_frame(s, camp, depth, count, rgba_dev, rgb_dev, hits_dev, aa, row0, row1);
    EOS_TRY(ensure_buffers(s->fb, f, trace_grid_of(f)));
    FrameEvents fe;
    FrameStreams fs{s->stream, s->stream2, s->sync};
    if (launch_times) {
        EOS_TRY(enqueue_frame(s->fb, f, fs, s->ev, &fe));
        return finish_frame(s->fb, f, s->stream, s->ev, &fe, st);
    }
    // two events, around the frame only (an event between two dependent
    // launches delays the second by a few microseconds)
    EOS_HIP(hipEventRecord(s->ev[0], s->stream));
    EOS_TRY(enqueue_frame(s->fb, f, fs, nullptr, nullptr));
    EOS_HIP(hipEventRecord(s->ev[1], s->stream));
    EOS_TRY(finish_frame(s->fb, f, s->stream, nullptr, nullptr, st));
    float ms = 0.f;
    if (st && hipEventElapsedTime(&ms, s->ev[0], s->ev[1]) == hipSuccess) st->render_ms = ms;
    return EOS_OK;
}

// Frames in flight: frame i runs on slot i % kFrameSlots (its own frame
// buffers and pair of streams), so the next frames' primary rays trace while
// frame i's shadow and reflection launches -- a few rays each, bound by the
// longest ray's chain of steps -- still run.  Resolves are ordered (frame i's
// after frame i - 1's), so outputs that alias end with the last frame.
// Counters are copied per frame to pinned host memory and read after the last
// frame.
namespace {
// Runs prepared frames kFrameSlots in flight; per-frame stats (render_ms: the
// sequence's GPU time / frames) into st when non-null.
eos_status run_pipelined(eos_scene* s, std::vector<Frame>& frames, eos_render_stats* st, bool launch_times) {
    constexpr int K = eos_scene::kFrameSlots;
    const int n = (int)frames.size();
    if (!s->seq_start) {
        // the further slots' streams (default priority) and the sequence's events
        hipError_t e = hipSuccess;
        for (int k = 1; k < K; ++k) {
            for (hipStream_t& x : s->xstream[k - 1])
                if (e == hipSuccess) e = hipStreamCreateWithFlags(&x, hipStreamNonBlocking);
            for (hipEvent_t& x : s->xsync[k - 1])
                if (e == hipSuccess) e = hipEventCreateWithFlags(&x, hipEventDisableTiming);
        }
        for (int k = 0; k < K; ++k) {
            if (e == hipSuccess) e = hipEventCreateWithFlags(&s->resolved[k], hipEventDisableTiming);
            if (e == hipSuccess) e = hipEventCreate(&s->seq_ev[k]);
        }
        if (e == hipSuccess) e = hipEventCreate(&s->seq_start);
        if (e != hipSuccess) return fail(EOS_ERR_HIP, std::string("frames in flight: ") + hipGetErrorString(e));
    }
    const size_t need = kCounterBytes * (size_t)n;
    if (s->host_ctr_bytes < need) {
        if (s->host_ctr) (void)hipHostFree(s->host_ctr);
        s->host_ctr = nullptr;
        s->host_ctr_bytes = 0;
        EOS_HIP(hipHostMalloc(&s->host_ctr, need, hipHostMallocDefault));
        s->host_ctr_bytes = need;
    }
    eos_frame_buffers* fbs[K];
    FrameStreams fs[K];
    for (int k = 0; k < K; ++k) {
        fbs[k] = k == 0 ? &s->fb : &s->fbx[k - 1];
        fs[k] = k == 0 ? FrameStreams{s->stream, s->stream2, s->sync}
                       : FrameStreams{s->xstream[k - 1][0], s->xstream[k - 1][1], s->xsync[k - 1]};
    }
    for (Frame& f : frames) {
        f.cfg.waves_per_cu = f.cfg.frames_waves_per_cu;
        f.cfg.refill = f.cfg.frames_refill;
        f.cfg.batch = f.cfg.frames_batch;
    }
    const int slots = std::min(K, std::max(n, 1));
    // launch_times: a pair of timing events per frame around its primary-ray
    // launch (the dispatch duration of that launch while frames overlap)
    while (launch_times && s->first_ev.size() < 2 * (size_t)n) {
        hipEvent_t e = nullptr;
        EOS_HIP(hipEventCreate(&e));
        s->first_ev.push_back(e);
    }
    for (int i = 0; launch_times && i < n; ++i) {
        frames[(size_t)i].first_ev[0] = s->first_ev[2 * (size_t)i];
        frames[(size_t)i].first_ev[1] = s->first_ev[2 * (size_t)i + 1];
    }
    // size every slot in use for every frame first (buffers only grow): no
    // buffer is reallocated while a frame in flight uses it
    for (const Frame& f : frames)
        for (int k = 0; k < slots; ++k) EOS_TRY(ensure_buffers(*fbs[k], f, trace_grid_of(f)));
    s->fb.rec_valid = 0;
    EOS_HIP(hipEventRecord(s->seq_start, s->stream));
    for (int k = 1; k < slots; ++k) EOS_HIP(hipStreamWaitEvent(fs[k].S, s->seq_start, 0));
    for (int i = 0; i < n; ++i) {
        const int k = i % slots;
        Frame& f = frames[(size_t)i];
        EOS_TRY(ensure_buffers(*fbs[k], f, trace_grid_of(f)));
        f.resolve_after = i > 0 ? s->resolved[(i - 1) % slots] : nullptr;
        f.resolve_done = s->resolved[k];
        EOS_TRY(enqueue_frame(*fbs[k], f, fs[k], nullptr, nullptr));
        EOS_HIP(hipMemcpyAsync(static_cast<char*>(s->host_ctr) + kCounterBytes * (size_t)i, fbs[k]->counters.ptr,
                               kCounterBytes, hipMemcpyDeviceToHost, fs[k].S));
    }
    double total = 0.0;
    for (int k = 0; k < slots; ++k) EOS_HIP(hipEventRecord(s->seq_ev[k], fs[k].S));
    for (int k = 0; k < slots; ++k) {
        EOS_HIP(hipStreamSynchronize(fs[k].S));
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, s->seq_start, s->seq_ev[k]);
        total = std::max(total, (double)ms);
    }
    for (int i = 0; i < n; ++i) {
        const char* base = static_cast<const char*>(s->host_ctr) + kCounterBytes * (size_t)i;
        eos_render_stats one{};
        EOS_TRY(frame_stats(frames[(size_t)i], reinterpret_cast<const int*>(base),
                            reinterpret_cast<const unsigned long long*>(base + 4 * CTR_COUNT), nullptr, nullptr,
                            &one));
        one.render_ms = total / n;
        float fm = 0.f;
        if (launch_times &&
            hipEventElapsedTime(&fm, s->first_ev[2 * (size_t)i], s->first_ev[2 * (size_t)i + 1]) == hipSuccess)
            one.trace_first_ms = fm;
        if (st) st[i] = one;
    }
    return EOS_OK;
}
}  // namespace

eos_status render_frames(eos_scene* s, const eos_camera* cams, int n, const eos_render_opts* o, uint8_t* const* rgba,
                         eos_render_stats* st) {
    std::vector<Frame> frames;
    frames.reserve((size_t)n);
    for (int i = 0; i < n; ++i)
        frames.push_back(make_frame(s, &cams[i], o->depth, o->count_visits != 0, reinterpret_cast<uchar4*>(rgba[i]),
                                    nullptr, nullptr, AaSample()));
    return run_pipelined(s, frames, st, o->launch_times != 0);
}

eos_status render(eos_scene* s, const eos_camera* cam, const eos_render_opts* o, uint8_t* rgba,
                  eos_render_stats* st, int row0, int row1) {
    eos_frame_buffers& fb = s->fb;
    if (row1 < 0) row1 = cam->height;
    const size_t npix = (size_t)cam->width * (size_t)(row1 - row0);
    const bool dev_out = o->output_memory == EOS_MEM_DEVICE;
    static thread_local DevBuf stage_rgba, stage_rgb, stage_hit;
    uchar4* rgba_dev = nullptr;
    float* rgb_dev = nullptr;
    int32_t* hit_dev = nullptr;
    if (rgba) {
        if (dev_out) rgba_dev = reinterpret_cast<uchar4*>(rgba);
        else {
            EOS_TRY(stage_rgba.ensure(4 * npix));
            rgba_dev = stage_rgba.as<uchar4>();
        }
    }
    if (o->rgb_out) {
        if (dev_out) rgb_dev = o->rgb_out;
        else {
            EOS_TRY(stage_rgb.ensure(12 * npix));
            rgb_dev = stage_rgb.as<float>();
        }
    }
    if (o->hit_out) {
        if (dev_out) hit_dev = o->hit_out;
        else {
            EOS_TRY(stage_hit.ensure(4 * npix));
            hit_dev = stage_hit.as<int32_t>();
        }
    }
    fb.rec_valid = 0;
    const int n = o->aa_samples > 1 ? o->aa_samples : 1;
    if (n == 1) {
        EOS_TRY(run_frame(s, cam, o->depth, o->count_visits != 0, rgba_dev, rgb_dev, hit_dev, st, AaSample(), row0,
                          row1, o->launch_times != 0));
    } else {
        // one frame per sample, in the reference's (q, p) order, summed by the
        // resolves (ordered across the two slots), two sample frames in flight
        EOS_TRY(fb.aa_acc.ensure(16 * npix));
        AaSample aa;
        aa.acc = fb.aa_acc.as<float4>();
        const float inv_n = (float)(1.0 / (double)(float)n);   // `float inv = 1.0 / (float) n`
        aa.inv_n2 = inv_n * inv_n;
        std::vector<Frame> frames;
        for (int q = 0; q < n; ++q)
            for (int p = 0; p < n; ++p) {
                const int k = q * n + p;
                aa.ox = ((float)p + 0.0f) * inv_n;
                aa.oy = ((float)q + 0.0f) * inv_n;
                aa.mode = k == 0 ? AA_FIRST : (k == n * n - 1 ? AA_LAST : AA_NEXT);
                frames.push_back(make_frame(s, cam, o->depth, o->count_visits != 0,
                                            aa.mode == AA_LAST ? rgba_dev : nullptr,
                                            aa.mode == AA_LAST ? rgb_dev : nullptr, k == 0 ? hit_dev : nullptr, aa,
                                            row0, row1));
            }
        std::vector<eos_render_stats> per(frames.size());
        EOS_TRY(run_pipelined(s, frames, per.data(), o->launch_times != 0));
        eos_render_stats sum{};
        for (const eos_render_stats& one : per) {
            sum.primary_rays += one.primary_rays;
            sum.shadow_rays += one.shadow_rays;
            sum.secondary_rays += one.secondary_rays;
            sum.hits += one.hits;
            sum.node_visits += one.node_visits;
            sum.tri_tests += one.tri_tests;
            sum.render_ms += one.render_ms;
            sum.trace_closest_launches += one.trace_closest_launches;
            sum.trace_shadow_launches += one.trace_shadow_launches;
            sum.closest_visits_first += one.closest_visits_first;
            sum.closest_tests_first += one.closest_tests_first;
        }
        if (st) *st = sum;   // per-launch GPU times are not split out for overlapped frames
    }
    if (!dev_out) {
        if (rgba) EOS_HIP(hipMemcpyAsync(rgba, rgba_dev, 4 * npix, hipMemcpyDeviceToHost, s->stream));
        if (o->rgb_out) EOS_HIP(hipMemcpyAsync(o->rgb_out, rgb_dev, 12 * npix, hipMemcpyDeviceToHost, s->stream));
        if (o->hit_out) EOS_HIP(hipMemcpyAsync(o->hit_out, hit_dev, 4 * npix, hipMemcpyDeviceToHost, s->stream));
        EOS_HIP(hipStreamSynchronize(s->stream));
    }
    if (o->keep_records) {
        fb.rec_valid = 1;
        fb.rec_cam = *cam;
    }
    return EOS_OK;
}

// ---- batched frames ----------------------------------------------------------
namespace {

// Buffers, scene table and captured graph of the last batch (per process;
// eos_render_batch calls must not overlap).
struct BatchState {
    eos_frame_buffers fb;
    DevBuf table, rgba;
    std::vector<BatchScene> host_table;
    std::vector<eos_scene*> scenes;
    std::vector<uint64_t> mat_gens;   // each scene's material generation when the frame was shaped
    std::vector<eos_camera> cams;
    int depth = -1;
    LaunchCfg cfg;
    hipGraphExec_t exec = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    hipEvent_t sync[2] = {};
    Frame frame;
};
BatchState g_batch;

bool same_batch(eos_scene* const* scenes, const eos_camera* cams, int n, int depth, const LaunchCfg& cfg) {
    const BatchState& b = g_batch;
    if ((int)b.scenes.size() != n || b.depth != depth) return false;
    if (b.cfg.workgroup != cfg.workgroup || b.cfg.stack_lds != cfg.stack_lds || b.cfg.waves_per_cu != cfg.waves_per_cu ||
        b.cfg.refill != cfg.refill || b.cfg.batch != cfg.batch || b.cfg.packets != cfg.packets ||
        b.cfg.coop != cfg.coop || b.cfg.group != cfg.group || b.cfg.steal != cfg.steal || b.cfg.group_lanes != cfg.group_lanes ||
        b.cfg.scatter != cfg.scatter)
        return false;
    for (int i = 0; i < n; ++i) {
        if (b.scenes[i] != scenes[i]) return false;
        // material flags (reflective, transparent) shape the frame: a ray-tree or
        // chain schedule and its buffers, captured into the graph
        if (b.mat_gens[i] != scenes[i]->mat_generation) return false;
        if (std::memcmp(&b.cams[i], &cams[i], sizeof(eos_camera)) != 0) return false;
        const SceneView v = view_of(scenes[i]);   // rebuilt BVH or changed light count
        if (std::memcmp(&v, &b.host_table[i].sc, sizeof(SceneView)) != 0) return false;
    }
    return true;
}

}  // namespace

eos_status render_batch(eos_scene* const* scenes, const eos_camera* cams, int n, const eos_batch_opts* o,
                        uint8_t* const* out, eos_batch_stats* st) {
    BatchState& b = g_batch;
    eos_scene* s0 = scenes[0];
    hipStream_t S = s0->stream;
    const LaunchCfg cfg = s0->cfg;
    if (!b.ev[0]) {
        EOS_HIP(hipEventCreate(&b.ev[0]));
        EOS_HIP(hipEventCreate(&b.ev[1]));
        for (hipEvent_t& e : b.sync) EOS_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    const FrameStreams fs{S, s0->stream2, b.sync};
    const bool same = same_batch(scenes, cams, n, o->depth, cfg);
    if (!same) {
        if (b.exec) {
            (void)hipGraphExecDestroy(b.exec);
            b.exec = nullptr;
        }
        b.scenes.clear();   // a failure below leaves nothing that a later call could take as the same batch
        b.host_table.assign(n, BatchScene{});
        size_t ids = 0, pix = 0;
        int lmax = 0;
        for (int i = 0; i < n; ++i)
            if (scenes[i]->width != s0->width) return fail(EOS_ERR_INVALID, "eos_render_batch: scenes built with different tree widths");
        for (int i = 0; i < n; ++i) {
            BatchScene& B = b.host_table[i];
            B.sc = view_of(scenes[i]);
            B.cam = *reinterpret_cast<const DevCamera*>(&cams[i]);
            B.tiles_x = (B.cam.width + 7) / 8;
            B.id_off = (int)ids;
            B.pix_off = (int)pix;
            B.n_ids = B.tiles_x * ((B.cam.height + 7) / 8) * 64;
            ids += (size_t)B.n_ids;
            pix += (size_t)B.cam.width * B.cam.height;
            lmax = std::max(lmax, scenes[i]->n_light);
        }
        if (ids >= ((size_t)1 << 31) || pix >= ((size_t)1 << 31)) return fail(EOS_ERR_LIMIT, "batch too large");
        EOS_TRY(b.table.ensure(sizeof(BatchScene) * n));
        EOS_HIP(hipMemcpyAsync(b.table.ptr, b.host_table.data(), sizeof(BatchScene) * n, hipMemcpyHostToDevice, S));
        EOS_TRY(b.rgba.ensure(4 * pix));
        Frame& f = b.frame;
        f = Frame();
        f.batch = true;
        f.table = b.table.as<const BatchScene>();
        f.n_scenes = n;
        f.npix = pix;
        f.n_ids = ids;
        f.depth = o->depth;
        f.lstride = lmax;
        f.cfg = cfg;
        f.width = s0->width;
        for (int i = 0; i < n; ++i) {
            f.xf |= scenes[i]->n_xf > 0;
            f.tree |= scenes[i]->any_transparent;
        }
        if (f.width == 2 || f.xf) {
            f.cfg.workgroup = 256;
            f.cfg.stack_lds = 16;
        }
        if (f.xf) f.cfg.packets = 0;
        f.num_cu = s0->num_cu;
        f.rgba = b.rgba.as<uchar4>();
        f.sc = view_of(s0);
        f.cam = b.host_table[0].cam;
        const int trace_grid = std::max(1, f.num_cu * f.cfg.waves_per_cu / (f.cfg.workgroup / 64));
        EOS_TRY(ensure_buffers(b.fb, f, trace_grid));
        b.scenes.assign(scenes, scenes + n);
        b.mat_gens.resize(n);
        for (int i = 0; i < n; ++i) b.mat_gens[i] = scenes[i]->mat_generation;
        b.cams.assign(cams, cams + n);
        b.depth = o->depth;
        b.cfg = cfg;
    }
    const Frame& f = b.frame;
    if (o->use_graph) {
        if (!b.exec) {
            hipGraph_t graph = nullptr;
            EOS_HIP(hipStreamBeginCapture(S, hipStreamCaptureModeThreadLocal));
            eos_status es = enqueue_frame(b.fb, f, fs, nullptr, nullptr);
            hipError_t ce = hipStreamEndCapture(S, &graph);
            if (es != EOS_OK) return es;
            if (ce != hipSuccess) return fail(EOS_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ce));
            hipError_t ie = hipGraphInstantiate(&b.exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (ie != hipSuccess) return fail(EOS_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ie));
        }
        EOS_HIP(hipEventRecord(b.ev[0], S));
        EOS_HIP(hipGraphLaunch(b.exec, S));
        EOS_HIP(hipEventRecord(b.ev[1], S));
    } else {
        EOS_HIP(hipEventRecord(b.ev[0], S));
        EOS_TRY(enqueue_frame(b.fb, f, fs, nullptr, nullptr));
        EOS_HIP(hipEventRecord(b.ev[1], S));
    }
    eos_render_stats rs;
    EOS_TRY(finish_frame(b.fb, f, S, nullptr, nullptr, &rs));
    for (int i = 0; i < n; ++i) {
        const BatchScene& B = b.host_table[i];
        const size_t bytes = 4 * (size_t)B.cam.width * B.cam.height;
        EOS_HIP(hipMemcpyAsync(out[i], b.rgba.as<uchar4>() + B.pix_off, bytes,
                               o->output_memory == EOS_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, S));
    }
    EOS_HIP(hipStreamSynchronize(S));
    if (st) {
        float ms = 0.f;
        EOS_HIP(hipEventElapsedTime(&ms, b.ev[0], b.ev[1]));
        st->render_ms = ms;
        st->rays = rs.primary_rays + rs.secondary_rays + rs.shadow_rays;
    }
    return EOS_OK;
}

}  // namespace eos
