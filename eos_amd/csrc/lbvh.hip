// lbvh.hip -- GPU Morton-code LBVH build for gfx950.
//
// The reference builds its BVH on the CPU (src/BVH.cpp:97-111, hybrid
// median/SAH) and ships only a Morton bit-interleave it never calls
// (src/morton.h:30-48).  This file builds a Karras (2012) LBVH on the GPU:
//   1. centroid bounds (two-level reduction, deterministic)
//   2. 30-bit Morton codes (EncodeMorton3 bit layout: z above y above x)
//   3. stable LSD radix sort of (code, index), 4 passes of 8 bits
//   4. Karras hierarchy, one thread per internal node
//   5. bottom-up AABB refit with per-node arrival counters
//   6. 64-byte traversal nodes and 48-byte triangle records in leaf order
// Every step is integer or exact float min/max work, so the result is
// bit-identical to the CPU restatement of the same algorithm.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "eos_internal.h"

namespace eos {
namespace {

constexpr int kSortThreads = 256;
constexpr int kSortChunks = 16;                      // keys per block = 4096
constexpr int kSortTile = kSortThreads * kSortChunks;
constexpr int kWaves = kSortThreads / 64;
constexpr int kRefitPasses = 64;                     // >= height of a Karras tree over 30-bit codes + index

__device__ __forceinline__ uint32_t part1by2(uint32_t x) {
    x &= 0x000003ffu;
    x = (x ^ (x << 16)) & 0xff0000ffu;
    x = (x ^ (x << 8)) & 0x0300f00fu;
    x = (x ^ (x << 4)) & 0x030c30c3u;
    x = (x ^ (x << 2)) & 0x09249249u;
    return x;
}

// ((v1 + v2) + v3) / 3 per axis, the reference's centroid (src/Triangle.h:38).
__device__ __forceinline__ void centroid(const float* __restrict__ v, size_t i, float c[3]) {
    const float* t = v + 9 * i;
#pragma unroll
    for (int k = 0; k < 3; ++k) c[k] = ((t[k] + t[3 + k]) + t[6 + k]) / 3.0f;
}

__global__ __launch_bounds__(256) void k_cbounds_partial(const float* __restrict__ verts, int n,
                                                         float* __restrict__ part) {
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        float c[3];
        centroid(verts, (size_t)i, c);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            lo[k] = fminf(lo[k], c[k]);
            hi[k] = fmaxf(hi[k], c[k]);
        }
    }
    __shared__ float s[6][256];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        s[k][threadIdx.x] = lo[k];
        s[3 + k][threadIdx.x] = hi[k];
    }
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                s[k][threadIdx.x] = fminf(s[k][threadIdx.x], s[k][threadIdx.x + w]);
                s[3 + k][threadIdx.x] = fmaxf(s[3 + k][threadIdx.x], s[3 + k][threadIdx.x + w]);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < 6) part[6 * blockIdx.x + threadIdx.x] = s[threadIdx.x][0];
}

__global__ __launch_bounds__(64) void k_cbounds_final(const float* __restrict__ part, int nparts,
                                                      float* __restrict__ cb) {
    // 6 lanes each reduce one component over all partials (min and max are exact).
    int k = threadIdx.x;
    if (k >= 6) return;
    float r = k < 3 ? INFINITY : -INFINITY;
    for (int b = 0; b < nparts; ++b) r = k < 3 ? fminf(r, part[6 * b + k]) : fmaxf(r, part[6 * b + k]);
    cb[k] = r;
}

__device__ __forceinline__ uint32_t quant10(float c, float lo, float hi) {
    float e = hi - lo;
    float s = e > 0.0f ? 1024.0f / e : 0.0f;
    float x = (c - lo) * s;
    x = fminf(fmaxf(x, 0.0f), 1023.0f);
    return (uint32_t)x;
}

__global__ __launch_bounds__(256) void k_morton(const float* __restrict__ verts, int n, const float* __restrict__ cb,
                                                uint32_t* __restrict__ codes, uint32_t* __restrict__ keys,
                                                int32_t* __restrict__ vals) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float c[3];
    centroid(verts, (size_t)i, c);
    uint32_t x = quant10(c[0], cb[0], cb[3]);
    uint32_t y = quant10(c[1], cb[1], cb[4]);
    uint32_t z = quant10(c[2], cb[2], cb[5]);
    uint32_t code = (part1by2(z) << 2) + (part1by2(y) << 1) + part1by2(x);
    codes[i] = code;
    keys[i] = code;
    vals[i] = i;
}

// ---- LSD radix sort, 8-bit digits -------------------------------------------
__global__ __launch_bounds__(kSortThreads) void k_radix_hist(const uint32_t* __restrict__ keys, int n, int shift,
                                                             uint32_t* __restrict__ hist, int nblocks) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int base = blockIdx.x * kSortTile;
    const int end = min(n, base + kSortTile);
    for (int i = base + threadIdx.x; i < end; i += kSortThreads) atomicAdd(&h[(keys[i] >> shift) & 255u], 1u);
    __syncthreads();
    hist[threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

// The digit-major histogram is scanned with the multi-block scan below.

// Stable scatter: keys are taken in order, 256 per step; lanes holding the same
// digit are matched with eight ballots, ranked by lane, and waves are ordered
// by a per-digit prefix over the block's waves.
__global__ __launch_bounds__(kSortThreads) void k_radix_scatter(const uint32_t* __restrict__ kin,
                                                                const int32_t* __restrict__ vin,
                                                                uint32_t* __restrict__ kout, int32_t* __restrict__ vout,
                                                                int n, int shift, const uint32_t* __restrict__ hist,
                                                                int nblocks) {
    __shared__ uint32_t running[256];
    __shared__ uint32_t wcount[kWaves][256];
    __shared__ uint32_t woff[kWaves][256];
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    running[tid] = hist[tid * nblocks + blockIdx.x];
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const int base = blockIdx.x * kSortTile;
    for (int ch = 0; ch < kSortChunks; ++ch) {
        const int i = base + ch * kSortThreads + tid;
        const bool valid = i < n;
        uint32_t key = valid ? kin[i] : 0u;
        int32_t val = valid ? vin[i] : 0;
        uint32_t d = (key >> shift) & 255u;
#pragma unroll
        for (int j = 0; j < 4; ++j) wcount[wave][lane * 4 + j] = 0;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int bit = 0; bit < 8; ++bit) {
            uint64_t b = __ballot((d >> bit) & 1u);
            peers &= ((d >> bit) & 1u) ? b : ~b;
        }
        const uint32_t rank = __popcll(peers & lt);
        __syncthreads();
        if (valid && rank == 0) wcount[wave][d] = __popcll(peers);
        __syncthreads();
        {
            uint32_t acc = running[tid];
#pragma unroll
            for (int w = 0; w < kWaves; ++w) {
                woff[w][tid] = acc;
                acc += wcount[w][tid];
            }
            running[tid] = acc;
        }
        __syncthreads();
        if (valid) {
            uint32_t dst = woff[wave][d] + rank;
            kout[dst] = key;
            vout[dst] = val;
        }
        __syncthreads();
    }
}

// ---- Karras hierarchy --------------------------------------------------------
__device__ __forceinline__ int delta(const uint32_t* __restrict__ k, int n, int i, int j) {
    if (j < 0 || j > n - 1) return -1;
    uint32_t a = k[i], b = k[j];
    if (a == b) return 32 + __clz((int)((uint32_t)i ^ (uint32_t)j));
    return __clz((int)(a ^ b));
}

__global__ __launch_bounds__(256) void k_karras(const uint32_t* __restrict__ k, int n, int2* __restrict__ kchild,
                                                int32_t* __restrict__ kparent, int2* __restrict__ fl) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    const int d = (delta(k, n, i, i + 1) - delta(k, n, i, i - 1)) >= 0 ? 1 : -1;
    const int dmin = delta(k, n, i, i - d);
    int lmax = 2;
    while (delta(k, n, i, i + lmax * d) > dmin) lmax *= 2;
    int l = 0;
    for (int t = lmax / 2; t >= 1; t /= 2)
        if (delta(k, n, i, i + (l + t) * d) > dmin) l += t;
    const int j = i + l * d;
    const int dnode = delta(k, n, i, j);
    int s = 0;
    for (int div = 2;; div *= 2) {
        int t = (l + div - 1) / div;
        if (delta(k, n, i, i + (s + t) * d) > dnode) s += t;
        if (t <= 1) break;
    }
    const int g = i + s * d + min(d, 0);
    const int lo = min(i, j), hi = max(i, j);
    const int left = (lo == g) ? ~g : g;
    const int right = (hi == g + 1) ? ~(g + 1) : g + 1;
    kchild[i] = make_int2(left, right);
    fl[i] = make_int2(lo, hi);
    kparent[left >= 0 ? left : (n - 1 + ~left)] = i;
    kparent[right >= 0 ? right : (n - 1 + ~right)] = i;
}

__global__ void k_fill_i32(int32_t* __restrict__ p, int n, int32_t v) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_copy_i32(const int32_t* __restrict__ src, int32_t* __restrict__ dst, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

// Leaf boxes: exact min/max of the three vertices, one thread per sorted slot.
__global__ __launch_bounds__(256) void k_leaf_boxes(const float* __restrict__ verts, const int32_t* __restrict__ perm,
                                                    int n, float* __restrict__ lbox) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const float* t = verts + 9 * (size_t)perm[s];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        lbox[6 * (size_t)s + k] = fminf(fminf(t[k], t[3 + k]), t[6 + k]);
        lbox[6 * (size_t)s + 3 + k] = fmaxf(fmaxf(t[k], t[3 + k]), t[6 + k]);
    }
}

// One level-synchronous refit pass over the internal nodes.  done[i] is the
// pass that boxed node i (0: not yet).  A node is boxed once both children are
// leaves or were boxed in an EARLIER pass, so every box it reads was written by
// an earlier kernel: no cross-workgroup ordering inside a launch is needed
// (device-scope release/acquire per climbing thread costs L2 writebacks on a
// multi-die part).  Passes = tree height; min/max are exact, so the result
// does not depend on the schedule.  Once the root (node 0) is boxed every node
// is, and the remaining passes return at once.
__global__ __launch_bounds__(256) void k_refit_pass(const int2* __restrict__ kchild, int m, int pass,
                                                    const float* __restrict__ lbox, float* __restrict__ ibox,
                                                    int32_t* __restrict__ done) {
    if (done[0] != 0) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m && done[i] == 0) {
        const int2 c = kchild[i];
        const int d0 = c.x >= 0 ? done[c.x] : -1, d1 = c.y >= 0 ? done[c.y] : -1;
        const bool r0 = c.x < 0 || (d0 > 0 && d0 < pass), r1 = c.y < 0 || (d1 > 0 && d1 < pass);
        if (r0 && r1) {
            const float* b0 = c.x >= 0 ? ibox + 6 * (size_t)c.x : lbox + 6 * (size_t)~c.x;
            const float* b1 = c.y >= 0 ? ibox + 6 * (size_t)c.y : lbox + 6 * (size_t)~c.y;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                ibox[6 * (size_t)i + k] = fminf(b0[k], b1[k]);
                ibox[6 * (size_t)i + 3 + k] = fmaxf(b0[3 + k], b1[3 + k]);
            }
            done[i] = pass;
        }
    }
}

__device__ __forceinline__ float inflate_pad(float lo, float hi, float eps) {
    float a = fabsf(lo), b = fabsf(hi);
    float m = a > b ? a : b;
    return m * 0x1p-16f + eps;
}

__device__ __forceinline__ int32_t leaf_ref(int32_t first, int32_t count) {
    return ~(first | ((count - 1) << kLeafShift));
}

__global__ __launch_bounds__(256) void k_nodes(const int2* __restrict__ kchild, const int2* __restrict__ fl,
                                               const float* __restrict__ ibox, const float* __restrict__ lbox,
                                               const float* __restrict__ cb, int n, int leaf_max, Node* __restrict__ nodes) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    float eps = 0.0f;
#pragma unroll
    for (int k = 0; k < 6; ++k) eps = fmaxf(eps, fabsf(cb[k]));
    eps = eps * 0x1p-22f;
    const int2 c = kchild[i];
    float lo[2][3], hi[2][3];
    int32_t ref[2];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const int ch = j ? c.y : c.x;
        const float* b = ch >= 0 ? ibox + 6 * (size_t)ch : lbox + 6 * (size_t)~ch;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            float pad = inflate_pad(b[k], b[3 + k], eps);
            lo[j][k] = b[k] - pad;
            hi[j][k] = b[3 + k] + pad;
        }
        if (ch < 0) {
            ref[j] = leaf_ref(~ch, 1);
        } else {
            const int2 r = fl[ch];
            ref[j] = (r.y - r.x + 1 <= leaf_max) ? leaf_ref(r.x, r.y - r.x + 1) : ch;
        }
    }
    Node nd;
    nd.b0 = make_float4(lo[0][0], hi[0][0], lo[0][1], hi[0][1]);
    nd.b1 = make_float4(lo[1][0], hi[1][0], lo[1][1], hi[1][1]);
    nd.b2 = make_float4(lo[0][2], hi[0][2], lo[1][2], hi[1][2]);
    nd.c = make_int4(ref[0], ref[1], 0, 0);
    nodes[i] = nd;
}

// World-space vertices of transformed triangles (M * (v, 1), glm's order of
// operations; the rows of M are kXfM..kXfM+2 of the triangle's record), the
// input vertices for the rest.  Keys and boxes of the build come from these;
// the triangle records keep the local vertices, as rays meet triangles in
// local space.
__global__ __launch_bounds__(256) void k_world_verts(const float* __restrict__ verts, const int32_t* __restrict__ xi,
                                                     int n, const float4* __restrict__ xf, int n_xf,
                                                     float* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* t = verts + 9 * (size_t)i;
    float* w = out + 9 * (size_t)i;
    const int x = xi[i];
    if (x < 0 || x >= n_xf) {
#pragma unroll
        for (int k = 0; k < 9; ++k) w[k] = t[k];
        return;
    }
    const float4 r[3] = {xf[kXfRows * (size_t)x + kXfM], xf[kXfRows * (size_t)x + kXfM + 1],
                         xf[kXfRows * (size_t)x + kXfM + 2]};
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int k = 0; k < 3; ++k)
            w[3 * c + k] = (r[k].x * t[3 * c] + r[k].y * t[3 * c + 1]) + (r[k].z * t[3 * c + 2] + r[k].w * 1.0f);
}

// Triangle records in leaf order: geo = (v1, e1, e2, original index,
// transformation index), shade = (n1 material, n2 transformation index, n3).
__global__ __launch_bounds__(256) void k_tris(const float* __restrict__ verts, const float* __restrict__ normals,
                                              const int32_t* __restrict__ mat, const int32_t* __restrict__ perm, int n,
                                              int n_mat, const int32_t* __restrict__ xi, int n_xf,
                                              float4* __restrict__ geo, float4* __restrict__ shade,
                                              int32_t* __restrict__ err) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int o = perm[s];
    int x = xi ? xi[o] : -1;
    if (x < -1 || x >= n_xf) {
        atomicAdd(err + 2, 1);
        x = -1;
    }
    const float* t = verts + 9 * (size_t)o;
    const float* nn = normals + 9 * (size_t)o;
    const float v1x = t[0], v1y = t[1], v1z = t[2];
    const float e1x = t[3] - v1x, e1y = t[4] - v1y, e1z = t[5] - v1z;
    const float e2x = t[6] - v1x, e2y = t[7] - v1y, e2z = t[8] - v1z;
    geo[3 * (size_t)s + 0] = make_float4(v1x, v1y, v1z, e1x);
    geo[3 * (size_t)s + 1] = make_float4(e1y, e1z, e2x, e2y);
    geo[3 * (size_t)s + 2] = make_float4(e2z, __int_as_float(o), __int_as_float(x), 0.0f);
    int m = mat[o];
    if (m < 0 || m >= n_mat) {
        atomicAdd(err, 1);
        m = 0;
    }
    shade[3 * (size_t)s + 0] = make_float4(nn[0], nn[1], nn[2], __int_as_float(m));
    shade[3 * (size_t)s + 1] = make_float4(nn[3], nn[4], nn[5], __int_as_float(x));
    shade[3 * (size_t)s + 2] = make_float4(nn[6], nn[7], nn[8], 0.0f);
}

// ---- exclusive scan of int32 (three phases: blocks of 1024, block sums, add) --
constexpr int kScanBlock = 1024;

__global__ __launch_bounds__(256) void k_scan_blocks(const int32_t* __restrict__ in, int32_t* __restrict__ out, int n,
                                                     int32_t* __restrict__ sums) {
    __shared__ int32_t s[256];
    const int base = blockIdx.x * kScanBlock + threadIdx.x * 4;
    int32_t v[4], t = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        v[k] = base + k < n ? in[base + k] : 0;
        t += v[k];
    }
    s[threadIdx.x] = t;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        int32_t x = threadIdx.x >= (unsigned)off ? s[threadIdx.x - off] : 0;
        __syncthreads();
        s[threadIdx.x] += x;
        __syncthreads();
    }
    int32_t run = s[threadIdx.x] - t;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
    if (threadIdx.x == 255) sums[blockIdx.x] = s[255];
}

__global__ __launch_bounds__(1024) void k_scan_sums(int32_t* __restrict__ sums, int nb) {
    __shared__ int32_t s[1024];
    __shared__ int32_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int b0 = 0; b0 < nb; b0 += 1024) {
        const int i = b0 + threadIdx.x;
        const int32_t v = i < nb ? sums[i] : 0;
        s[threadIdx.x] = v;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            int32_t x = threadIdx.x >= (unsigned)off ? s[threadIdx.x - off] : 0;
            __syncthreads();
            s[threadIdx.x] += x;
            __syncthreads();
        }
        if (i < nb) sums[i] = carry + s[threadIdx.x] - v;
        __syncthreads();
        if (threadIdx.x == 1023) carry += s[1023];
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void k_scan_add(int32_t* __restrict__ out, int n, const int32_t* __restrict__ sums) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) out[i] += sums[i / kScanBlock];
}

// ---- 4-wide collapse -----------------------------------------------------------
// A 4-wide node stands for every binary node that stays internal (range larger
// than leaf_max) at even depth; nodes are numbered in binary index order.
__global__ __launch_bounds__(256) void k_flags4(const int32_t* __restrict__ kparent, const int2* __restrict__ fl, int n,
                                                int leaf_max, int32_t* __restrict__ flags) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    const int2 r = fl[i];
    int f = 0;
    if (r.y - r.x + 1 > leaf_max) {
        int depth = 0;
        for (int p = kparent[i]; p >= 0 && depth < 128; p = kparent[p]) ++depth;
        f = (depth & 1) == 0;
    }
    flags[i] = f;
}

__global__ __launch_bounds__(256) void k_nodes4(const int2* __restrict__ kchild, const int2* __restrict__ fl,
                                                const float* __restrict__ ibox, const float* __restrict__ lbox,
                                                const float* __restrict__ cb, const int32_t* __restrict__ flags,
                                                const int32_t* __restrict__ idx, int n, int leaf_max,
                                                Node4* __restrict__ nodes4) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1 || !flags[i]) return;
    float eps = 0.0f;
#pragma unroll
    for (int k = 0; k < 6; ++k) eps = fmaxf(eps, fabsf(cb[k]));
    eps = eps * 0x1p-22f;
    int slots[4] = {0, 0, 0, 0};
    int ns = 0;
    const int2 c = kchild[i];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const int ch = j ? c.y : c.x;
        bool internal = false;
        if (ch >= 0) {
            const int2 r = fl[ch];
            internal = r.y - r.x + 1 > leaf_max;
        }
        if (internal) {
            const int2 g = kchild[ch];
            slots[ns++] = g.x;
            slots[ns++] = g.y;
        } else {
            slots[ns++] = ch;
        }
    }
    float lo[3][4], hi[3][4];
    int32_t ref[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        ref[k] = kEmpty;
#pragma unroll
        for (int a = 0; a < 3; ++a) lo[a][k] = hi[a][k] = INFINITY;
        if (k < ns) {
            const int g = slots[k];
            const float* b = g >= 0 ? ibox + 6 * (size_t)g : lbox + 6 * (size_t)~g;
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                float pad = inflate_pad(b[a], b[3 + a], eps);
                lo[a][k] = b[a] - pad;
                hi[a][k] = b[3 + a] + pad;
            }
            if (g < 0) {
                ref[k] = leaf_ref(~g, 1);
            } else {
                const int2 r = fl[g];
                ref[k] = (r.y - r.x + 1 <= leaf_max) ? leaf_ref(r.x, r.y - r.x + 1) : idx[g];
            }
        }
    }
    Node4 nd;
    nd.lox = make_float4(lo[0][0], lo[0][1], lo[0][2], lo[0][3]);
    nd.hix = make_float4(hi[0][0], hi[0][1], hi[0][2], hi[0][3]);
    nd.loy = make_float4(lo[1][0], lo[1][1], lo[1][2], lo[1][3]);
    nd.hiy = make_float4(hi[1][0], hi[1][1], hi[1][2], hi[1][3]);
    nd.loz = make_float4(lo[2][0], lo[2][1], lo[2][2], lo[2][3]);
    nd.hiz = make_float4(hi[2][0], hi[2][1], hi[2][2], hi[2][3]);
    nd.ref = make_int4(ref[0], ref[1], ref[2], ref[3]);
    nd.pad = make_int4(0, 0, 0, 0);
    nodes4[idx[i]] = nd;
}

inline unsigned blocks(size_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

}  // namespace

eos_status build_lbvh(eos_scene* s, int leaf_max, int width, eos_build_stats* st) {
    const int n = s->n_tri;
    hipStream_t S = s->stream;
    s->built = false;
    s->fb.rec_valid = 0;
    s->leaf_max = leaf_max;
    s->width = width;
    s->n_nodes4 = 0;
    const size_t un = (size_t)n, m = n > 0 ? un - 1 : 0;
    const int nblocks = n > 0 ? (int)blocks(un, kSortTile) : 1;
    const int nparts = n > 0 ? (int)std::min<size_t>(blocks(un, 256), 1024) : 1;
    EOS_TRY(s->codes.ensure(4 * un));
    EOS_TRY(s->codes_sorted.ensure(4 * un));
    EOS_TRY(s->perm.ensure(4 * un));
    EOS_TRY(s->sort_tmp_k.ensure(4 * un));
    EOS_TRY(s->sort_tmp_v.ensure(4 * un));
    EOS_TRY(s->sort_hist.ensure(2 * 4 * 256 * (size_t)nblocks));
    EOS_TRY(s->scan_tmp.ensure(4 * blocks(256 * (size_t)nblocks, kScanBlock) + 4 * blocks(un, kScanBlock) + 8));
    EOS_TRY(s->kchild.ensure(8 * m));
    EOS_TRY(s->kparent.ensure(4 * (2 * un)));
    EOS_TRY(s->first_last.ensure(8 * m));
    EOS_TRY(s->ibox.ensure(24 * m));
    EOS_TRY(s->lbox.ensure(24 * un));
    EOS_TRY(s->flags.ensure(4 * m + 4));
    EOS_TRY(s->cbounds.ensure(64));
    EOS_TRY(s->bounds_part.ensure(24 * (size_t)nparts));
    EOS_TRY(s->nodes.ensure(64 * m));
    EOS_TRY(s->geo.ensure(48 * un));
    EOS_TRY(s->shade.ensure(48 * un));
    EOS_HIP(hipEventRecord(s->ev[0], S));
    int32_t* err = s->cbounds.as<int32_t>() + 8;
    int32_t* root_pass = err + 1;   // refit pass that boxed the root (0: none); err[2]: bad transformation indices
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, S, err, 3, 0);
    const bool xf = s->n_xf > 0;
    if (n > 0) {
        const float* V = s->verts.as<float>();   // keys and boxes: world space
        if (xf) {
            EOS_TRY(s->wverts.ensure(36 * un));
            hipLaunchKernelGGL(k_world_verts, dim3(blocks(un, 256)), dim3(256), 0, S, V,
                               s->xf_index.as<const int32_t>(), n, s->xf.as<const float4>(), s->n_xf,
                               s->wverts.as<float>());
            V = s->wverts.as<const float>();
        }
        float* cb = s->cbounds.as<float>();
        hipLaunchKernelGGL(k_cbounds_partial, dim3(nparts), dim3(256), 0, S, V, n, s->bounds_part.as<float>());
        hipLaunchKernelGGL(k_cbounds_final, dim3(1), dim3(64), 0, S, s->bounds_part.as<const float>(), nparts, cb);
        uint32_t* kA = s->codes_sorted.as<uint32_t>();
        int32_t* vA = s->perm.as<int32_t>();
        uint32_t* kB = s->sort_tmp_k.as<uint32_t>();
        int32_t* vB = s->sort_tmp_v.as<int32_t>();
        hipLaunchKernelGGL(k_morton, dim3(blocks(un, 256)), dim3(256), 0, S, V, n, cb, s->codes.as<uint32_t>(), kA, vA);
        uint32_t* H = s->sort_hist.as<uint32_t>();
        uint32_t* H2 = H + 256 * (size_t)nblocks;
        const int htot = 256 * nblocks, hb = (int)blocks((size_t)htot, kScanBlock);
        int32_t* hsums = s->scan_tmp.as<int32_t>();
        for (int pass = 0; pass < 4; ++pass) {
            const int shift = 8 * pass;
            const uint32_t* ki = (pass & 1) ? kB : kA;
            const int32_t* vi = (pass & 1) ? vB : vA;
            uint32_t* ko = (pass & 1) ? kA : kB;
            int32_t* vo = (pass & 1) ? vA : vB;
            hipLaunchKernelGGL(k_radix_hist, dim3(nblocks), dim3(kSortThreads), 0, S, ki, n, shift, H, nblocks);
            // exclusive scan of the digit-major histogram into H2 (counts < 2^31)
            hipLaunchKernelGGL(k_scan_blocks, dim3(hb), dim3(256), 0, S, reinterpret_cast<const int32_t*>(H),
                               reinterpret_cast<int32_t*>(H2), htot, hsums);
            hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, S, hsums, hb);
            hipLaunchKernelGGL(k_scan_add, dim3(blocks(htot, 256)), dim3(256), 0, S, reinterpret_cast<int32_t*>(H2),
                               htot, hsums);
            hipLaunchKernelGGL(k_radix_scatter, dim3(nblocks), dim3(kSortThreads), 0, S, ki, vi, ko, vo, n, shift, H2,
                               nblocks);
        }
        hipLaunchKernelGGL(k_fill_i32, dim3(blocks(2 * un, 256)), dim3(256), 0, S, s->kparent.as<int32_t>(),
                           (int)(2 * un - 1), -1);
        if (n > 1) {
            hipLaunchKernelGGL(k_karras, dim3(blocks(m, 256)), dim3(256), 0, S, kA, n, s->kchild.as<int2>(),
                               s->kparent.as<int32_t>(), s->first_last.as<int2>());
            hipLaunchKernelGGL(k_fill_i32, dim3(blocks(m, 256)), dim3(256), 0, S, s->flags.as<int32_t>(), (int)m, 0);
        }
        hipLaunchKernelGGL(k_leaf_boxes, dim3(blocks(un, 256)), dim3(256), 0, S, V, vA, n, s->lbox.as<float>());
        if (n > 1) {
            for (int pass = 1; pass <= kRefitPasses; ++pass)
                hipLaunchKernelGGL(k_refit_pass, dim3(blocks(m, 256)), dim3(256), 0, S, s->kchild.as<const int2>(),
                                   (int)m, pass, s->lbox.as<const float>(), s->ibox.as<float>(),
                                   s->flags.as<int32_t>());
            // the root's pass number, 0 if the passes did not reach it
            hipLaunchKernelGGL(k_copy_i32, dim3(1), dim3(64), 0, S, s->flags.as<const int32_t>(), root_pass, 1);
        }
        if (n > 1)
            hipLaunchKernelGGL(k_nodes, dim3(blocks(m, 256)), dim3(256), 0, S, s->kchild.as<const int2>(),
                               s->first_last.as<const int2>(), s->ibox.as<const float>(), s->lbox.as<const float>(),
                               cb, n, leaf_max, s->nodes.as<Node>());
        hipLaunchKernelGGL(k_tris, dim3(blocks(un, 256)), dim3(256), 0, S, s->verts.as<const float>(),
                           s->normals.as<const float>(), s->matidx.as<const int32_t>(), vA, n, s->n_mat,
                           xf ? s->xf_index.as<const int32_t>() : nullptr, s->n_xf, s->geo.as<float4>(),
                           s->shade.as<float4>(), err);
        if (n > leaf_max) {
            // 4-wide collapse: flags, their exclusive scan (node numbers), node count, nodes
            const int nb = (int)blocks(m, kScanBlock);
            EOS_TRY(s->flags4.ensure(8 * m + 8));
            EOS_TRY(s->scan_tmp.ensure(4 * (size_t)nb + 4));
            int32_t* flags = s->flags4.as<int32_t>();
            int32_t* idx = flags + m;
            hipLaunchKernelGGL(k_flags4, dim3(blocks(m, 256)), dim3(256), 0, S, s->kparent.as<const int32_t>(),
                               s->first_last.as<const int2>(), n, leaf_max, flags);
            hipLaunchKernelGGL(k_scan_blocks, dim3(nb), dim3(256), 0, S, flags, idx, (int)m, s->scan_tmp.as<int32_t>());
            hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, S, s->scan_tmp.as<int32_t>(), nb);
            hipLaunchKernelGGL(k_scan_add, dim3(blocks(m, 256)), dim3(256), 0, S, idx, (int)m,
                               s->scan_tmp.as<const int32_t>());
            int32_t tail[2] = {0, 0};
            EOS_HIP(hipMemcpyAsync(&tail[0], flags + (m - 1), 4, hipMemcpyDeviceToHost, S));
            EOS_HIP(hipMemcpyAsync(&tail[1], idx + (m - 1), 4, hipMemcpyDeviceToHost, S));
            EOS_HIP(hipStreamSynchronize(S));
            s->n_nodes4 = tail[0] + tail[1];
            EOS_TRY(s->nodes4.ensure(128 * (size_t)std::max(s->n_nodes4, 1)));
            hipLaunchKernelGGL(k_nodes4, dim3(blocks(m, 256)), dim3(256), 0, S, s->kchild.as<const int2>(),
                               s->first_last.as<const int2>(), s->ibox.as<const float>(), s->lbox.as<const float>(),
                               cb, flags, idx, n, leaf_max, s->nodes4.as<Node4>());
        }
    }
    EOS_HIP(hipGetLastError());
    EOS_HIP(hipEventRecord(s->ev[1], S));
    int32_t herr[3] = {0, 0, 0};
    EOS_HIP(hipMemcpyAsync(herr, err, 12, hipMemcpyDeviceToHost, S));
    EOS_HIP(hipStreamSynchronize(S));
    if (herr[0] != 0)
        return fail(EOS_ERR_INVALID, "eos_build_lbvh: " + std::to_string(herr[0]) + " material indices out of range");
    if (herr[2] != 0)
        return fail(EOS_ERR_INVALID,
                    "eos_build_lbvh: " + std::to_string(herr[2]) + " transformation indices out of range");
    if (n > 1 && herr[1] == 0) return fail(EOS_ERR_STATE, "eos_build_lbvh: refit did not reach the root");
    float ms = 0.0f;
    EOS_HIP(hipEventElapsedTime(&ms, s->ev[0], s->ev[1]));
    if (n == 0) s->root = kDone;
    else if (n <= leaf_max) s->root = ~(0 | ((n - 1) << kLeafShift));
    else s->root = 0;
    s->root4 = s->root;
    s->built = true;
    if (st) {
        st->build_ms = ms;
        st->num_internal = n > 0 ? n - 1 : 0;
        st->root_ref = s->root;
        st->leaf_max = leaf_max;
        st->width = width;
        st->num_nodes4 = s->n_nodes4;
        st->root4 = s->root4;
    }
    return EOS_OK;
}

}  // namespace eos
