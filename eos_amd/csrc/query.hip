// query.hip -- ray queries: the reference Scene's per-ray queries
// (src/Scene.h:69-71) answered for many caller-given rays at once.
//
//   EOS_QUERY_NEAREST         Scene::findMinDistanceIntersectionBVH
//   EOS_QUERY_OCCLUDED        Scene::shadowRayVisibilityBVH
//   EOS_QUERY_NEAREST_LINEAR  Scene::findMinDistanceIntersectionLinear
//
// The BVH queries run the render path's persistent traversal (k_trace in the
// MODE_RAYS / MODE_RAYS_ANY modes, rays read straight from the caller's
// arrays), then one thread per ray turns the hit record into the reference's
// RayIntersection fields with the same arithmetic the frame's set-up kernel
// uses.  The linear query tests every triangle in index order, workgroup tiles
// of triangles staged in LDS and read by every lane as broadcasts.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <string>

#include "eos_internal.h"
#include "trace.h"

namespace eos {
namespace {

// eos_ray_hit as three float4: (point, hit) (normal, material) (t, u, v, triangle)
static_assert(sizeof(eos_ray_hit) == 48, "ray hit record is three 16-byte words");

__device__ __forceinline__ void write_hit(float4* out, size_t i, bool found, F3 P, F3 N, int mat, float t, float u,
                                          float v, int tri) {
    out[3 * i] = make_float4(P.x, P.y, P.z, __int_as_float(found ? 1 : 0));
    out[3 * i + 1] = make_float4(N.x, N.y, N.z, __int_as_float(mat));
    out[3 * i + 2] = make_float4(t, u, v, __int_as_float(tri));
}

// Nearest-hit records (t, slot, u, v) of the traversal -> RayIntersection fields
// (src/BVH.cpp:342-383: point through M, normal through transpose(inverse) and
// normalised; material index of the triangle).  No hit: zero point and normal
// as RayIntersection's constructor leaves them (src/RayIntersection.h:34).
__global__ __launch_bounds__(256) void k_query_hits(int n, const float* ro, const float* rd, const float4* rec,
                                                     const float4* shade, const float4* geo, const float4* xf,
                                                     float4* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 h = rec[i];
    const int slot = __float_as_int(h.y);
    F3 P = {0, 0, 0}, N = {0, 0, 0};
    if (slot < 0) {
        write_hit(out, (size_t)i, false, P, N, -1, INFINITY, 0.0f, 0.0f, -1);
        return;
    }
    const F3 o = f3(ro[3 * (size_t)i], ro[3 * (size_t)i + 1], ro[3 * (size_t)i + 2]);
    const F3 d = f3(rd[3 * (size_t)i], rd[3 * (size_t)i + 1], rd[3 * (size_t)i + 2]);
    const float4 s0 = shade[3 * (size_t)slot], s1 = shade[3 * (size_t)slot + 1], s2 = shade[3 * (size_t)slot + 2];
    hit_surface(s0, s1, s2, xf, o, d, h.x, h.z, h.w, true, P, N);
    write_hit(out, (size_t)i, true, P, N, __float_as_int(s0.w), h.x, h.z, h.w,
              __float_as_int(geo[3 * (size_t)slot + 2].y));
}

constexpr int kLinTile = 256;

// src/Scene.cpp:159-201.  For each triangle in index order: the ray in the
// triangle's local space, Moller-Trumbore (src/Triangle.cpp:26-91) against the
// running distance the reference passes as Triangle::hit's `distance`
// (minTemp: a hit farther than the last accepted one is rejected; the first
// hit is always accepted, the reading of its uninitialised start value), and
// the hit is kept when t < minDist (999999 at first) and t > 0.  found: some
// triangle accepted a hit (the query's return value), even when none was kept.
// The kept hit's point goes through M and its normal through transpose(inverse)
// without a second normalisation.  verts / normals / matidx / xfi are the
// scene's triangles in their original order.
__global__ __launch_bounds__(kLinTile) void k_query_linear(int n, const float* ro, const float* rd, int n_tri,
                                                            const float* verts, const float* normals,
                                                            const int32_t* matidx, const int32_t* xfi,
                                                            const float4* xf, float4* out) {
    __shared__ float tv[9][kLinTile];   // component-major: lane j of the staging writes word j of each row
    __shared__ int32_t tx[kLinTile];
    const int i = blockIdx.x * kLinTile + (int)threadIdx.x;
    const bool act = i < n;
    F3 o = {0, 0, 0}, d = {0, 0, 0};
    if (act) {
        o = f3(ro[3 * (size_t)i], ro[3 * (size_t)i + 1], ro[3 * (size_t)i + 2]);
        d = f3(rd[3 * (size_t)i], rd[3 * (size_t)i + 1], rd[3 * (size_t)i + 2]);
    }
    float min_temp = INFINITY, min_dist = 999999.0f, bu = 0.0f, bv = 0.0f;
    bool found = false;
    int best = -1;
    int lxi = -1;
    F3 lo = o, ld = d;
    for (int base = 0; base < n_tri; base += kLinTile) {
        __syncthreads();
        const int k = base + (int)threadIdx.x;
        if (k < n_tri) {
            for (int c = 0; c < 9; ++c) tv[c][threadIdx.x] = verts[9 * (size_t)k + c];
            tx[threadIdx.x] = xfi ? xfi[k] : -1;
        }
        __syncthreads();
        const int cnt = min(kLinTile, n_tri - base);
        if (!act) continue;
        for (int j = 0; j < cnt; ++j) {
            const int xi = tx[j];
            F3 oo = o, dd = d;
            if (xi >= 0) {
                if (xi != lxi) {
                    lo = o;
                    ld = d;
                    to_local<true>(xf, xi, lo, ld);
                    lxi = xi;
                }
                oo = lo;
                dd = ld;
            }
            const F3 v1 = f3(tv[0][j], tv[1][j], tv[2][j]);
            const F3 e1 = sub(f3(tv[3][j], tv[4][j], tv[5][j]), v1);
            const F3 e2 = sub(f3(tv[6][j], tv[7][j], tv[8][j]), v1);
            float t, u, v;
            if (moller_trumbore(oo, dd, v1, e1, e2, t, u, v) && !(t > min_temp)) {
                found = true;
                min_temp = t;
                if (t < min_dist && t > 0.0f) {
                    min_dist = t;
                    best = base + j;
                    bu = u;
                    bv = v;
                }
            }
        }
    }
    if (!act) return;
    F3 P = {0, 0, 0}, N = {0, 0, 0};
    if (best < 0) {
        write_hit(out, (size_t)i, found, P, N, -1, INFINITY, 0.0f, 0.0f, -1);
        return;
    }
    const float* nn = normals + 9 * (size_t)best;
    const int xi = xfi ? xfi[best] : -1;
    const float4 s0 = make_float4(nn[0], nn[1], nn[2], 0.0f);
    const float4 s1 = make_float4(nn[3], nn[4], nn[5], __int_as_float(xi));
    const float4 s2 = make_float4(nn[6], nn[7], nn[8], 0.0f);
    hit_surface(s0, s1, s2, xf, o, d, min_dist, bu, bv, false, P, N);
    write_hit(out, (size_t)i, found, P, N, matidx[best], min_dist, bu, bv, best);
}

// The BVH traversal for caller-given rays: the default launch shape (256
// threads, 16 LDS stack entries), the scene's tree width and transformations.
template <int MODE>
void launch_rays(int width, bool xf, int grid, hipStream_t s, const TraceArgs& a) {
    if (width == 2) {
        if (xf) hipLaunchKernelGGL((k_trace<MODE, 256, 16, false, false, 2, true>), dim3(grid), dim3(256), 0, s, a);
        else hipLaunchKernelGGL((k_trace<MODE, 256, 16, false, false, 2, false>), dim3(grid), dim3(256), 0, s, a);
    } else {
        if (xf) hipLaunchKernelGGL((k_trace<MODE, 256, 16, false, false, 4, true>), dim3(grid), dim3(256), 0, s, a);
        else hipLaunchKernelGGL((k_trace<MODE, 256, 16, false, false, 4, false>), dim3(grid), dim3(256), 0, s, a);
    }
}

// Fetch cursors of the query launch, then its error flag.
constexpr int kQueryCtrInts = kRanges * kCursorStride + 32;

__global__ void k_zero_i32_query(int* p, int n) {
    for (int i = (int)threadIdx.x; i < n; i += (int)blockDim.x) p[i] = 0;
}

// ---- the allocation-free variant (eos_trace_rays_static) --------------------
// Its scratch is this fixed array (one copy per GPU, held until the process
// ends): the fetch cursors and the error flag in the first kStaticHeader bytes,
// then the nearest-hit records (16 bytes per ray).  Host code reaches it only
// through the address hipGetSymbolAddress gives on each device.
constexpr size_t kStaticHeader = 4096;
static_assert(4 * (size_t)kQueryCtrInts <= kStaticHeader, "cursors and flag fit the header");
constexpr size_t kStaticBytes = kStaticHeader + 16 * (size_t)EOS_STATIC_MAX_RAYS;
__device__ __align__(256) unsigned char scratch[kStaticBytes];

constexpr int kMaxDevices = 64;

// The device address of `scratch` on the current device, looked up once per
// device (null when the lookup fails).
void* static_scratch(int device) {
    static std::mutex mu;
    static void* ptr[kMaxDevices] = {};
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(mu);
    if (!ptr[device]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) return nullptr;
        ptr[device] = p;
    }
    return ptr[device];
}

}  // namespace

eos_status trace_rays_static(eos_scene* s, int n, const float* o, const float* d, const eos_trace_opts* opts,
                             eos_ray_hit* hits, uint8_t* occ) {
    const hipStream_t S = s->stream;
    void* base = static_scratch(s->device);
    if (!base) return fail(EOS_ERR_HIP, "eos_trace_rays_static: hipGetSymbolAddress of the scratch array failed");
    int* ctr = static_cast<int*>(base);
    float4* rec = reinterpret_cast<float4*>(static_cast<char*>(base) + kStaticHeader);
    float4* out = reinterpret_cast<float4*>(hits);
    const unsigned blocks = (unsigned)(((size_t)n + 255) / 256);
    if (opts->mode == EOS_QUERY_NEAREST_LINEAR) {
        hipLaunchKernelGGL(k_query_linear, dim3(blocks), dim3(kLinTile), 0, S, n, o, d, s->n_tri,
                           s->verts.as<const float>(), s->normals.as<const float>(), s->matidx.as<const int32_t>(),
                           s->n_xf > 0 ? s->xf_index.as<const int32_t>() : nullptr,
                           s->n_xf > 0 ? s->xf.as<const float4>() : nullptr, out);
        EOS_HIP(hipGetLastError());
        EOS_HIP(hipStreamSynchronize(S));
        return EOS_OK;
    }
    // every lane's stack lives in LDS: lane quads walk the 4-wide tree (96
    // entries per quad), single lanes walk the binary tree with all of its 64
    // entries in LDS, so no overflow stack is needed
    const int trace_grid = std::max(1, s->num_cu * std::min(s->cfg.waves_per_cu, 8) / 4);
    const int grid = (int)std::min<size_t>((size_t)trace_grid, std::max<size_t>(1, (4 * (size_t)n + 255) / 256));
    hipLaunchKernelGGL(k_zero_i32_query, dim3(1), dim3(256), 0, S, ctr, kQueryCtrInts);
    TraceArgs a{};
    a.sc = view_of(s);
    a.n_static = n;
    a.count_mul = 1;
    a.cursor = ctr;
    a.err = ctr + kRanges * kCursorStride;
    a.gstack = ctr;          // never dereferenced: no lane's stack leaves LDS
    a.refill = s->cfg.refill;
    a.batch = s->cfg.batch;
    a.coop = 4;
    a.ray_o = o;
    a.ray_d = d;
    const bool xf = s->n_xf > 0;
    if (opts->mode == EOS_QUERY_OCCLUDED) {
        a.vis_out = occ;
        if (s->width == 2) {
            if (xf) hipLaunchKernelGGL((k_trace<MODE_RAYS_ANY, 256, kMaxStack, false, false, 2, true>), dim3(grid), dim3(256), 0, S, a);
            else hipLaunchKernelGGL((k_trace<MODE_RAYS_ANY, 256, kMaxStack, false, false, 2, false>), dim3(grid), dim3(256), 0, S, a);
        } else {
            launch_rays<MODE_RAYS_ANY>(4, xf, grid, S, a);
        }
    } else {
        a.hit_out = rec;
        if (s->width == 2) {
            if (xf) hipLaunchKernelGGL((k_trace<MODE_RAYS, 256, kMaxStack, false, false, 2, true>), dim3(grid), dim3(256), 0, S, a);
            else hipLaunchKernelGGL((k_trace<MODE_RAYS, 256, kMaxStack, false, false, 2, false>), dim3(grid), dim3(256), 0, S, a);
        } else {
            launch_rays<MODE_RAYS>(4, xf, grid, S, a);
        }
        hipLaunchKernelGGL(k_query_hits, dim3(blocks), dim3(256), 0, S, n, o, d, rec, s->shade.as<const float4>(),
                           s->geo.as<const float4>(), xf ? s->xf.as<const float4>() : nullptr, out);
    }
    EOS_HIP(hipGetLastError());
    int err = 0;
    EOS_HIP(hipMemcpyAsync(&err, a.err, sizeof err, hipMemcpyDeviceToHost, S));
    EOS_HIP(hipStreamSynchronize(S));
    if (err & 1) return fail(EOS_ERR_LIMIT, "eos_trace_rays_static: traversal stack overflow");
    if (err & 2) return fail(EOS_ERR_STATE, "eos_trace_rays_static: traversal exceeded its visit bound (corrupt BVH)");
    return EOS_OK;
}

eos_status trace_rays(eos_scene* s, int n, const float* o, const float* d, const eos_trace_opts* opts,
                      eos_ray_hit* hits, uint8_t* occ) {
    const bool dev = opts->memory == EOS_MEM_DEVICE;
    const int mode = opts->mode;
    const hipStream_t S = s->stream;
    const size_t nn = (size_t)n;
    const float* ro = o;
    const float* rd = d;
    if (!dev) {
        EOS_TRY(s->q_rays.ensure(24 * nn));
        EOS_HIP(hipMemcpyAsync(s->q_rays.ptr, o, 12 * nn, hipMemcpyHostToDevice, S));
        EOS_HIP(hipMemcpyAsync(s->q_rays.as<char>() + 12 * nn, d, 12 * nn, hipMemcpyHostToDevice, S));
        ro = s->q_rays.as<const float>();
        rd = reinterpret_cast<const float*>(s->q_rays.as<const char>() + 12 * nn);
    }
    float4* out = nullptr;
    uint8_t* occ_dev = nullptr;
    if (mode == EOS_QUERY_OCCLUDED) {
        if (dev) occ_dev = occ;
        else {
            EOS_TRY(s->q_out.ensure(nn));
            occ_dev = s->q_out.as<uint8_t>();
        }
    } else if (dev) {
        out = reinterpret_cast<float4*>(hits);
    } else {
        EOS_TRY(s->q_out.ensure(48 * nn));
        out = s->q_out.as<float4>();
    }
    const unsigned blocks = (unsigned)((nn + 255) / 256);
    int err = 0;
    if (mode == EOS_QUERY_NEAREST_LINEAR) {
        hipLaunchKernelGGL(k_query_linear, dim3(blocks), dim3(kLinTile), 0, S, n, ro, rd, s->n_tri,
                           s->verts.as<const float>(), s->normals.as<const float>(), s->matidx.as<const int32_t>(),
                           s->n_xf > 0 ? s->xf_index.as<const int32_t>() : nullptr,
                           s->n_xf > 0 ? s->xf.as<const float4>() : nullptr, out);
        EOS_HIP(hipGetLastError());
    } else {
        const LaunchCfg& cfg = s->cfg;
        const int trace_grid = std::max(1, s->num_cu * cfg.waves_per_cu / 4);
        // a quad per ray while there are few rays (the cooperative walk), never
        // more workgroups than the persistent grid
        const int grid = (int)std::min<size_t>((size_t)trace_grid, std::max<size_t>(1, (4 * nn + 255) / 256));
        const int maxs = s->width == 4 ? kMaxStack4 : kMaxStack;
        EOS_TRY(s->q_ctr.ensure(4 * (size_t)kQueryCtrInts));
        EOS_TRY(s->q_gstack.ensure(4 * (size_t)grid * 256 * (maxs - 16)));
        int* ctr = s->q_ctr.as<int>();
        hipLaunchKernelGGL(k_zero_i32_query, dim3(1), dim3(256), 0, S, ctr, kQueryCtrInts);
        TraceArgs a{};
        a.sc = view_of(s);
        a.n_static = n;
        a.count_mul = 1;
        a.cursor = ctr;
        a.err = ctr + kRanges * kCursorStride;
        a.gstack = s->q_gstack.as<int32_t>();
        a.refill = cfg.refill;
        a.batch = cfg.batch;
        a.coop = cfg.coop;
        a.group = cfg.group;
        a.group_lanes = cfg.group_lanes;
        a.steal = cfg.steal;
        a.scatter = cfg.scatter;
        a.ray_o = ro;
        a.ray_d = rd;
        if (mode == EOS_QUERY_OCCLUDED) {
            a.vis_out = occ_dev;
            launch_rays<MODE_RAYS_ANY>(s->width, s->n_xf > 0, grid, S, a);
        } else {
            EOS_TRY(s->q_rec.ensure(16 * nn));
            a.hit_out = s->q_rec.as<float4>();
            launch_rays<MODE_RAYS>(s->width, s->n_xf > 0, grid, S, a);
            hipLaunchKernelGGL(k_query_hits, dim3(blocks), dim3(256), 0, S, n, ro, rd, s->q_rec.as<const float4>(),
                               s->shade.as<const float4>(), s->geo.as<const float4>(),
                               s->n_xf > 0 ? s->xf.as<const float4>() : nullptr, out);
        }
        EOS_HIP(hipGetLastError());
        EOS_HIP(hipMemcpyAsync(&err, a.err, sizeof err, hipMemcpyDeviceToHost, S));
    }
    if (!dev) {
        if (mode == EOS_QUERY_OCCLUDED) EOS_HIP(hipMemcpyAsync(occ, occ_dev, nn, hipMemcpyDeviceToHost, S));
        else EOS_HIP(hipMemcpyAsync(hits, out, 48 * nn, hipMemcpyDeviceToHost, S));
    }
    EOS_HIP(hipStreamSynchronize(S));
    if (err & 1) return fail(EOS_ERR_LIMIT, "eos_trace_rays: traversal stack overflow");
    if (err & 2) return fail(EOS_ERR_STATE, "eos_trace_rays: traversal exceeded its visit bound (corrupt BVH)");
    return EOS_OK;
}

}  // namespace eos
