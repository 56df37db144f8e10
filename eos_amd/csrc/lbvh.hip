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
#include <climits>
#include <string>
#include <vector>

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
    s->batch_built = false;
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

// ---- many scenes in one launch sequence ---------------------------------------
// The scenes' triangles are concatenated (scene s at global slots
// [tri_off, tri_off + n)); every stage is one launch over all of them.  Each
// scene keeps its own centroid bounds, Morton codes and Karras tree: the sort
// orders (scene, code, index) -- four passes over the codes, then a stable pass
// over the scene index -- and the Karras kernel builds scene s's tree from the
// keys of its own segment with its own local indices, so every per-scene array
// is the one eos_build_lbvh_ex gives that scene alone, bit for bit.  Internal
// nodes of scene s sit at [node_off, node_off + n - 1) with node_off = tri_off - s,
// its parent links at [2 tri_off - s, 2 (tri_off + n) - s - 1); child references
// are local to the scene.
namespace eos {
namespace {

constexpr int kBoundsChunk = 4096;   // triangles per partial-bounds block

struct BuildSeg {
    int tri_off, n, node_off, blk_off;
    const float* verts;       // the scene's local vertices, original order
    const float* normals;
    const int32_t* mat;
    int n_mat;
    const int32_t* xi;        // transformation index per triangle (null: none)
    const float4* xf;
    int n_xf;
    Node* nodes;              // outputs: the scene's own buffers
    Node4* nodes4;
    float4* geo;
    float4* shade;
    int32_t* perm;
    uint32_t* codes_sorted;
    uint32_t* codes;
    int base4;                // global number of the scene's first 4-wide node
};

// The last segment whose offset is <= i (offsets ascending; a segment with no
// entries shares its offset with the next one, which is the one returned).
__device__ __forceinline__ int seg_of(const int* __restrict__ off, int nseg, int i) {
    int lo = 0, hi = nseg - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= i) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// World-space vertices, k_world_verts's arithmetic, into one concatenated array.
__global__ __launch_bounds__(256) void k_bworld(const BuildSeg* __restrict__ seg, const int* __restrict__ seg_tri,
                                                int nseg, int N, float* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const BuildSeg& S = seg[seg_of(seg_tri, nseg, i)];
    const int li = i - S.tri_off;
    const float* t = S.verts + 9 * (size_t)li;
    float* w = out + 9 * (size_t)i;
    const int x = S.xi ? S.xi[li] : -1;
    if (x < 0 || x >= S.n_xf) {
#pragma unroll
        for (int k = 0; k < 9; ++k) w[k] = t[k];
        return;
    }
    const float4 r[3] = {S.xf[kXfRows * (size_t)x + kXfM], S.xf[kXfRows * (size_t)x + kXfM + 1],
                         S.xf[kXfRows * (size_t)x + kXfM + 2]};
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int k = 0; k < 3; ++k)
            w[3 * c + k] = (r[k].x * t[3 * c] + r[k].y * t[3 * c + 1]) + (r[k].z * t[3 * c + 2] + r[k].w * 1.0f);
}

// Partial centroid bounds: block b covers kBoundsChunk triangles of one scene.
__global__ __launch_bounds__(256) void k_bcb_partial(const float* __restrict__ W, const BuildSeg* __restrict__ seg,
                                                     const int* __restrict__ seg_blk, int nseg,
                                                     float* __restrict__ part) {
    const BuildSeg& S = seg[seg_of(seg_blk, nseg, (int)blockIdx.x)];
    const int lb = (int)blockIdx.x - S.blk_off;
    const int a = S.tri_off + lb * kBoundsChunk, b = min(S.tri_off + S.n, a + kBoundsChunk);
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int i = a + (int)threadIdx.x; i < b; i += blockDim.x) {
        float c[3];
        centroid(W, (size_t)i, c);
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

// One thread per (scene, component): min or max over the scene's partials.
__global__ void k_bcb_final(const float* __restrict__ part, const BuildSeg* __restrict__ seg, int nseg,
                            float* __restrict__ cb) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 6 * nseg) return;
    const int sidx = t / 6, k = t - 6 * sidx;
    const BuildSeg& S = seg[sidx];
    const int nb = (S.n + kBoundsChunk - 1) / kBoundsChunk;
    float r = k < 3 ? INFINITY : -INFINITY;
    for (int b = 0; b < nb; ++b) {
        const float v = part[6 * (size_t)(S.blk_off + b) + k];
        r = k < 3 ? fminf(r, v) : fmaxf(r, v);
    }
    cb[t] = r;
}

__global__ __launch_bounds__(256) void k_bmorton(const float* __restrict__ W, const int* __restrict__ seg_tri, int nseg,
                                                 int N, const float* __restrict__ cb, uint32_t* __restrict__ codes,
                                                 uint32_t* __restrict__ keys, int32_t* __restrict__ vals) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const float* c6 = cb + 6 * seg_of(seg_tri, nseg, i);
    float c[3];
    centroid(W, (size_t)i, c);
    const uint32_t x = quant10(c[0], c6[0], c6[3]);
    const uint32_t y = quant10(c[1], c6[1], c6[4]);
    const uint32_t z = quant10(c[2], c6[2], c6[5]);
    const uint32_t code = (part1by2(z) << 2) + (part1by2(y) << 1) + part1by2(x);
    codes[i] = code;
    keys[i] = code;
    vals[i] = i;
}

// Sort keys of the scene pass: the scene of each element.
__global__ __launch_bounds__(256) void k_bscene_keys(const int32_t* __restrict__ vals, const int* __restrict__ seg_tri,
                                                     int nseg, int N, uint32_t* __restrict__ keys) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) keys[i] = (uint32_t)seg_of(seg_tri, nseg, vals[i]);
}

// Sorted codes, and each scene's permutation and codes in its own buffers.
__global__ __launch_bounds__(256) void k_bgather(const int32_t* __restrict__ vals, const uint32_t* __restrict__ codes,
                                                 const BuildSeg* __restrict__ seg, const int* __restrict__ seg_tri,
                                                 int nseg, int N, uint32_t* __restrict__ sorted) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const BuildSeg& S = seg[seg_of(seg_tri, nseg, i)];
    const int g = vals[i], li = i - S.tri_off, o = g - S.tri_off;
    const uint32_t c = codes[g];
    sorted[i] = c;
    S.perm[li] = o;
    S.codes_sorted[li] = c;
    S.codes[o] = c;
}

// Karras over each scene's own segment: k_karras with the scene's keys, length
// and local indices; outputs at the scene's node and parent offsets.
__global__ __launch_bounds__(256) void k_bkarras(const uint32_t* __restrict__ sorted, const BuildSeg* __restrict__ seg,
                                                 const int* __restrict__ seg_tri, int nseg, int N,
                                                 int2* __restrict__ kchild, int32_t* __restrict__ kparent,
                                                 int2* __restrict__ fl) {
    const int gi = blockIdx.x * blockDim.x + threadIdx.x;
    if (gi >= N) return;
    const int sidx = seg_of(seg_tri, nseg, gi);
    const BuildSeg& S = seg[sidx];
    const int n = S.n, i = gi - S.tri_off;
    if (i >= n - 1) return;
    const uint32_t* k = sorted + S.tri_off;
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
    kchild[S.node_off + i] = make_int2(left, right);
    fl[S.node_off + i] = make_int2(lo, hi);
    int32_t* kp = kparent + (2 * (size_t)S.tri_off - sidx);
    kp[left >= 0 ? left : (n - 1 + ~left)] = i;
    kp[right >= 0 ? right : (n - 1 + ~right)] = i;
}

// k_refit_pass over every scene's internal nodes; a scene's root boxed counts
// in *roots, and once every root is boxed the remaining passes return at once.
__global__ __launch_bounds__(256) void k_brefit_pass(const int2* __restrict__ kchild, const BuildSeg* __restrict__ seg,
                                                     const int* __restrict__ seg_node, int nseg, int M, int pass,
                                                     const float* __restrict__ lbox, float* __restrict__ ibox,
                                                     int32_t* __restrict__ done, int n_roots,
                                                     int32_t* __restrict__ roots) {
    if (*roots >= n_roots) return;
    const int gi = blockIdx.x * blockDim.x + threadIdx.x;
    if (gi >= M || done[gi] != 0) return;
    const BuildSeg& S = seg[seg_of(seg_node, nseg, gi)];
    const int2 c = kchild[gi];
    const int g0 = S.node_off + c.x, g1 = S.node_off + c.y;
    const int d0 = c.x >= 0 ? done[g0] : -1, d1 = c.y >= 0 ? done[g1] : -1;
    const bool r0 = c.x < 0 || (d0 > 0 && d0 < pass), r1 = c.y < 0 || (d1 > 0 && d1 < pass);
    if (r0 && r1) {
        const float* b0 = c.x >= 0 ? ibox + 6 * (size_t)g0 : lbox + 6 * (size_t)(S.tri_off + ~c.x);
        const float* b1 = c.y >= 0 ? ibox + 6 * (size_t)g1 : lbox + 6 * (size_t)(S.tri_off + ~c.y);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            ibox[6 * (size_t)gi + k] = fminf(b0[k], b1[k]);
            ibox[6 * (size_t)gi + 3 + k] = fmaxf(b0[3 + k], b1[3 + k]);
        }
        done[gi] = pass;
        if (gi == S.node_off) atomicAdd(roots, 1);
    }
}

// Leaf boxes of the concatenated sorted slots (k_leaf_boxes over the world vertices).
__global__ __launch_bounds__(256) void k_bleaf_boxes(const float* __restrict__ W, const int32_t* __restrict__ vals,
                                                     int N, float* __restrict__ lbox) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= N) return;
    const float* t = W + 9 * (size_t)vals[s];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        lbox[6 * (size_t)s + k] = fminf(fminf(t[k], t[3 + k]), t[6 + k]);
        lbox[6 * (size_t)s + 3 + k] = fmaxf(fmaxf(t[k], t[3 + k]), t[6 + k]);
    }
}

// k_nodes per scene, into the scene's node array.
__global__ __launch_bounds__(256) void k_bnodes(const int2* __restrict__ kchild, const int2* __restrict__ fl,
                                                const float* __restrict__ ibox, const float* __restrict__ lbox,
                                                const float* __restrict__ cbs, const BuildSeg* __restrict__ seg,
                                                const int* __restrict__ seg_node, int nseg, int M, int leaf_max) {
    const int gi = blockIdx.x * blockDim.x + threadIdx.x;
    if (gi >= M) return;
    const int sidx = seg_of(seg_node, nseg, gi);
    const BuildSeg& S = seg[sidx];
    const float* cb = cbs + 6 * sidx;
    float eps = 0.0f;
#pragma unroll
    for (int k = 0; k < 6; ++k) eps = fmaxf(eps, fabsf(cb[k]));
    eps = eps * 0x1p-22f;
    const int2 c = kchild[gi];
    float lo[2][3], hi[2][3];
    int32_t ref[2];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const int ch = j ? c.y : c.x;
        const float* b = ch >= 0 ? ibox + 6 * (size_t)(S.node_off + ch) : lbox + 6 * (size_t)(S.tri_off + ~ch);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            float pad = inflate_pad(b[k], b[3 + k], eps);
            lo[j][k] = b[k] - pad;
            hi[j][k] = b[3 + k] + pad;
        }
        if (ch < 0) {
            ref[j] = leaf_ref(~ch, 1);
        } else {
            const int2 r = fl[S.node_off + ch];
            ref[j] = (r.y - r.x + 1 <= leaf_max) ? leaf_ref(r.x, r.y - r.x + 1) : ch;
        }
    }
    Node nd;
    nd.b0 = make_float4(lo[0][0], hi[0][0], lo[0][1], hi[0][1]);
    nd.b1 = make_float4(lo[1][0], hi[1][0], lo[1][1], hi[1][1]);
    nd.b2 = make_float4(lo[0][2], hi[0][2], lo[1][2], hi[1][2]);
    nd.c = make_int4(ref[0], ref[1], 0, 0);
    S.nodes[gi - S.node_off] = nd;
}

// k_tris per scene, into the scene's triangle records.
__global__ __launch_bounds__(256) void k_btris(const int32_t* __restrict__ vals, const BuildSeg* __restrict__ seg,
                                               const int* __restrict__ seg_tri, int nseg, int N,
                                               int32_t* __restrict__ err) {
    const int gi = blockIdx.x * blockDim.x + threadIdx.x;
    if (gi >= N) return;
    const BuildSeg& S = seg[seg_of(seg_tri, nseg, gi)];
    const int s = gi - S.tri_off, o = vals[gi] - S.tri_off;
    int x = S.xi ? S.xi[o] : -1;
    if (x < -1 || x >= S.n_xf) {
        atomicAdd(err + 2, 1);
        x = -1;
    }
    const float* t = S.verts + 9 * (size_t)o;
    const float* nn = S.normals + 9 * (size_t)o;
    const float v1x = t[0], v1y = t[1], v1z = t[2];
    const float e1x = t[3] - v1x, e1y = t[4] - v1y, e1z = t[5] - v1z;
    const float e2x = t[6] - v1x, e2y = t[7] - v1y, e2z = t[8] - v1z;
    S.geo[3 * (size_t)s + 0] = make_float4(v1x, v1y, v1z, e1x);
    S.geo[3 * (size_t)s + 1] = make_float4(e1y, e1z, e2x, e2y);
    S.geo[3 * (size_t)s + 2] = make_float4(e2z, __int_as_float(o), __int_as_float(x), 0.0f);
    int m = S.mat[o];
    if (m < 0 || m >= S.n_mat) {
        atomicAdd(err, 1);
        m = 0;
    }
    S.shade[3 * (size_t)s + 0] = make_float4(nn[0], nn[1], nn[2], __int_as_float(m));
    S.shade[3 * (size_t)s + 1] = make_float4(nn[3], nn[4], nn[5], __int_as_float(x));
    S.shade[3 * (size_t)s + 2] = make_float4(nn[6], nn[7], nn[8], 0.0f);
}

// k_flags4 per scene (scenes of at most leaf_max triangles have no 4-wide nodes).
__global__ __launch_bounds__(256) void k_bflags4(const int32_t* __restrict__ kparent, const int2* __restrict__ fl,
                                                 const BuildSeg* __restrict__ seg, const int* __restrict__ seg_node,
                                                 int nseg, int M, int leaf_max, int32_t* __restrict__ flags) {
    const int gi = blockIdx.x * blockDim.x + threadIdx.x;
    if (gi >= M) return;
    const int sidx = seg_of(seg_node, nseg, gi);
    const BuildSeg& S = seg[sidx];
    const int2 r = fl[gi];
    int f = 0;
    if (S.n > leaf_max && r.y - r.x + 1 > leaf_max) {
        const int32_t* kp = kparent + (2 * (size_t)S.tri_off - sidx);
        int depth = 0;
        for (int p = kp[gi - S.node_off]; p >= 0 && depth < 128; p = kp[p]) ++depth;
        f = (depth & 1) == 0;
    }
    flags[gi] = f;
}

// Per scene: 4-wide node count and global number of its first node.
__global__ void k_bcount4(const int32_t* __restrict__ flags, const int32_t* __restrict__ idx,
                          const BuildSeg* __restrict__ seg, int nseg, int32_t* __restrict__ out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nseg) return;
    const BuildSeg& S = seg[t];
    int count = 0, base = 0;
    if (S.n > 1) {
        const int first = S.node_off, last = S.node_off + S.n - 2;
        base = idx[first];
        count = idx[last] + flags[last] - base;
    }
    out[2 * t] = count;
    out[2 * t + 1] = base;
}

// k_nodes4 per scene, numbered within the scene.
__global__ __launch_bounds__(256) void k_bnodes4(const int2* __restrict__ kchild, const int2* __restrict__ fl,
                                                 const float* __restrict__ ibox, const float* __restrict__ lbox,
                                                 const float* __restrict__ cbs, const int32_t* __restrict__ flags,
                                                 const int32_t* __restrict__ idx, const BuildSeg* __restrict__ seg,
                                                 const int* __restrict__ seg_node, int nseg, int M, int leaf_max) {
    const int gi = blockIdx.x * blockDim.x + threadIdx.x;
    if (gi >= M || !flags[gi]) return;
    const int sidx = seg_of(seg_node, nseg, gi);
    const BuildSeg& S = seg[sidx];
    const float* cb = cbs + 6 * sidx;
    float eps = 0.0f;
#pragma unroll
    for (int k = 0; k < 6; ++k) eps = fmaxf(eps, fabsf(cb[k]));
    eps = eps * 0x1p-22f;
    int slots[4] = {0, 0, 0, 0};
    int ns = 0;
    const int2 c = kchild[gi];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const int ch = j ? c.y : c.x;
        bool internal = false;
        if (ch >= 0) {
            const int2 r = fl[S.node_off + ch];
            internal = r.y - r.x + 1 > leaf_max;
        }
        if (internal) {
            const int2 g = kchild[S.node_off + ch];
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
            const float* b = g >= 0 ? ibox + 6 * (size_t)(S.node_off + g) : lbox + 6 * (size_t)(S.tri_off + ~g);
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                float pad = inflate_pad(b[a], b[3 + a], eps);
                lo[a][k] = b[a] - pad;
                hi[a][k] = b[3 + a] + pad;
            }
            if (g < 0) {
                ref[k] = leaf_ref(~g, 1);
            } else {
                const int2 r = fl[S.node_off + g];
                ref[k] = (r.y - r.x + 1 <= leaf_max) ? leaf_ref(r.x, r.y - r.x + 1) : idx[S.node_off + g] - S.base4;
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
    S.nodes4[idx[gi] - S.base4] = nd;
}

// Scratch of the batched build (per process; calls must not overlap).
struct BatchBuild {
    DevBuf table, offs, W, part, cb, codes, kA, vA, kB, vB, sorted, hist, scan, kchild, kparent, fl, lbox, ibox,
        done, flags4, counts, misc;
    hipEvent_t ev[2] = {nullptr, nullptr};
};
BatchBuild g_bbuild;

}  // namespace

eos_status build_lbvh_batch(eos_scene* const* scenes, int n_scenes, int leaf_max, int width, eos_build_stats* st) {
    BatchBuild& B = g_bbuild;
    hipStream_t S = scenes[0]->stream;
    if (!B.ev[0]) {
        EOS_HIP(hipEventCreate(&B.ev[0]));
        EOS_HIP(hipEventCreate(&B.ev[1]));
    }
    // scenes with triangles form the segments; every scene's own buffers are sized here
    std::vector<BuildSeg> segs;
    std::vector<int> seg_scene, seg_tri, seg_node, seg_blk;
    size_t N = 0;
    int nblk = 0, n_roots = 0;
    for (int i = 0; i < n_scenes; ++i) {
        eos_scene* s = scenes[i];
        s->built = false;
        s->fb.rec_valid = 0;
        s->leaf_max = leaf_max;
        s->width = width;
        s->n_nodes4 = 0;
        s->batch_built = true;
        const size_t un = (size_t)s->n_tri;
        if (un == 0) continue;
        if (N + un > (size_t)INT_MAX / 2)
            return fail(EOS_ERR_LIMIT, "eos_build_lbvh_batch: more than 2^30 triangles in one batch");
        EOS_TRY(s->nodes.ensure(64 * (un - 1)));
        EOS_TRY(s->geo.ensure(48 * un));
        EOS_TRY(s->shade.ensure(48 * un));
        EOS_TRY(s->perm.ensure(4 * un));
        EOS_TRY(s->codes_sorted.ensure(4 * un));
        EOS_TRY(s->codes.ensure(4 * un));
        BuildSeg g{};
        g.tri_off = (int)N;
        g.n = (int)un;
        g.node_off = (int)N - (int)segs.size();
        g.blk_off = nblk;
        g.verts = s->verts.as<const float>();
        g.normals = s->normals.as<const float>();
        g.mat = s->matidx.as<const int32_t>();
        g.n_mat = s->n_mat;
        g.xi = s->n_xf > 0 ? s->xf_index.as<const int32_t>() : nullptr;
        g.xf = s->n_xf > 0 ? s->xf.as<const float4>() : nullptr;
        g.n_xf = s->n_xf;
        g.nodes = s->nodes.as<Node>();
        g.geo = s->geo.as<float4>();
        g.shade = s->shade.as<float4>();
        g.perm = s->perm.as<int32_t>();
        g.codes_sorted = s->codes_sorted.as<uint32_t>();
        g.codes = s->codes.as<uint32_t>();
        segs.push_back(g);
        seg_scene.push_back(i);
        seg_tri.push_back(g.tri_off);
        seg_node.push_back(g.node_off);
        seg_blk.push_back(g.blk_off);
        nblk += (int)((un + kBoundsChunk - 1) / kBoundsChunk);
        n_roots += un > 1 ? 1 : 0;
        N += un;
    }
    const int nseg = (int)segs.size();
    // wait for work already queued on the scenes' own streams (uploads, transformations)
    for (int i = 0; i < n_scenes; ++i)
        if (scenes[i]->stream != S) EOS_HIP(hipStreamSynchronize(scenes[i]->stream));
    EOS_HIP(hipEventRecord(B.ev[0], S));
    std::vector<int32_t> cnt4(2 * (size_t)std::max(nseg, 1), 0);
    int32_t herr[4] = {0, 0, 0, 0};
    if (nseg > 0) {
        const size_t M = N - (size_t)nseg;       // internal nodes of every scene
        const int Ni = (int)N, Mi = (int)M;
        const int nblocks = (int)blocks(N, kSortTile);
        EOS_TRY(B.table.ensure(sizeof(BuildSeg) * nseg));
        EOS_TRY(B.offs.ensure(4 * 3 * (size_t)nseg));
        EOS_TRY(B.W.ensure(36 * N));
        EOS_TRY(B.part.ensure(24 * (size_t)nblk));
        EOS_TRY(B.cb.ensure(24 * (size_t)nseg));
        EOS_TRY(B.codes.ensure(4 * N));
        EOS_TRY(B.kA.ensure(4 * N));
        EOS_TRY(B.vA.ensure(4 * N));
        EOS_TRY(B.kB.ensure(4 * N));
        EOS_TRY(B.vB.ensure(4 * N));
        EOS_TRY(B.sorted.ensure(4 * N));
        EOS_TRY(B.hist.ensure(2 * 4 * 256 * (size_t)nblocks));
        EOS_TRY(B.scan.ensure(4 * blocks(256 * (size_t)nblocks, kScanBlock) + 4 * blocks(N, kScanBlock) + 8));
        EOS_TRY(B.kchild.ensure(8 * M + 8));
        EOS_TRY(B.kparent.ensure(4 * (2 * N)));
        EOS_TRY(B.fl.ensure(8 * M + 8));
        EOS_TRY(B.lbox.ensure(24 * N));
        EOS_TRY(B.ibox.ensure(24 * M + 24));
        EOS_TRY(B.done.ensure(4 * M + 4));
        EOS_TRY(B.flags4.ensure(8 * M + 8));
        EOS_TRY(B.counts.ensure(8 * (size_t)nseg));
        EOS_TRY(B.misc.ensure(64));
        std::vector<int> offs;
        offs.insert(offs.end(), seg_tri.begin(), seg_tri.end());
        offs.insert(offs.end(), seg_node.begin(), seg_node.end());
        offs.insert(offs.end(), seg_blk.begin(), seg_blk.end());
        EOS_HIP(hipMemcpyAsync(B.table.ptr, segs.data(), sizeof(BuildSeg) * nseg, hipMemcpyHostToDevice, S));
        EOS_HIP(hipMemcpyAsync(B.offs.ptr, offs.data(), 4 * offs.size(), hipMemcpyHostToDevice, S));
        const BuildSeg* T = B.table.as<const BuildSeg>();
        const int* o_tri = B.offs.as<const int>();
        const int* o_node = o_tri + nseg;
        const int* o_blk = o_node + nseg;
        int32_t* err = B.misc.as<int32_t>();           // [0] materials, [1] roots boxed, [2] transformations
        hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, S, err, 4, 0);
        float* W = B.W.as<float>();
        hipLaunchKernelGGL(k_bworld, dim3(blocks(N, 256)), dim3(256), 0, S, T, o_tri, nseg, Ni, W);
        hipLaunchKernelGGL(k_bcb_partial, dim3(nblk), dim3(256), 0, S, W, T, o_blk, nseg, B.part.as<float>());
        hipLaunchKernelGGL(k_bcb_final, dim3(blocks(6 * (size_t)nseg, 64)), dim3(64), 0, S, B.part.as<const float>(),
                           T, nseg, B.cb.as<float>());
        uint32_t* kA = B.kA.as<uint32_t>();
        int32_t* vA = B.vA.as<int32_t>();
        uint32_t* kB = B.kB.as<uint32_t>();
        int32_t* vB = B.vB.as<int32_t>();
        hipLaunchKernelGGL(k_bmorton, dim3(blocks(N, 256)), dim3(256), 0, S, W, o_tri, nseg, Ni, B.cb.as<const float>(),
                           B.codes.as<uint32_t>(), kA, vA);
        uint32_t* H = B.hist.as<uint32_t>();
        uint32_t* H2 = H + 256 * (size_t)nblocks;
        const int htot = 256 * nblocks, hb = (int)blocks((size_t)htot, kScanBlock);
        int32_t* hsums = B.scan.as<int32_t>();
        // four passes over the codes, then one (two beyond 256 scenes) over the scene index
        const int scene_passes = nseg > 256 ? 2 : 1;
        for (int pass = 0; pass < 4 + scene_passes; ++pass) {
            const int shift = pass < 4 ? 8 * pass : 8 * (pass - 4);
            const uint32_t* ki = (pass & 1) ? kB : kA;
            const int32_t* vi = (pass & 1) ? vB : vA;
            uint32_t* ko = (pass & 1) ? kA : kB;
            int32_t* vo = (pass & 1) ? vA : vB;
            if (pass == 4)   // keys of the scene passes: each element's scene (its values move with it)
                hipLaunchKernelGGL(k_bscene_keys, dim3(blocks(N, 256)), dim3(256), 0, S, vi, o_tri, nseg, Ni,
                                   const_cast<uint32_t*>(ki));
            hipLaunchKernelGGL(k_radix_hist, dim3(nblocks), dim3(kSortThreads), 0, S, ki, Ni, shift, H, nblocks);
            hipLaunchKernelGGL(k_scan_blocks, dim3(hb), dim3(256), 0, S, reinterpret_cast<const int32_t*>(H),
                               reinterpret_cast<int32_t*>(H2), htot, hsums);
            hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, S, hsums, hb);
            hipLaunchKernelGGL(k_scan_add, dim3(blocks(htot, 256)), dim3(256), 0, S, reinterpret_cast<int32_t*>(H2),
                               htot, hsums);
            hipLaunchKernelGGL(k_radix_scatter, dim3(nblocks), dim3(kSortThreads), 0, S, ki, vi, ko, vo, Ni, shift, H2,
                               nblocks);
        }
        const int32_t* vS = ((4 + scene_passes) & 1) ? vB : vA;   // values after the last pass
        uint32_t* sorted = B.sorted.as<uint32_t>();
        hipLaunchKernelGGL(k_bgather, dim3(blocks(N, 256)), dim3(256), 0, S, vS, B.codes.as<const uint32_t>(), T, o_tri,
                           nseg, Ni, sorted);
        int32_t* kparent = B.kparent.as<int32_t>();
        hipLaunchKernelGGL(k_fill_i32, dim3(blocks(2 * N, 256)), dim3(256), 0, S, kparent, (int)(2 * N), -1);
        int2* kchild = B.kchild.as<int2>();
        int2* fl = B.fl.as<int2>();
        hipLaunchKernelGGL(k_bkarras, dim3(blocks(N, 256)), dim3(256), 0, S, sorted, T, o_tri, nseg, Ni, kchild,
                           kparent, fl);
        hipLaunchKernelGGL(k_bleaf_boxes, dim3(blocks(N, 256)), dim3(256), 0, S, W, vS, Ni, B.lbox.as<float>());
        if (M > 0) {
            hipLaunchKernelGGL(k_fill_i32, dim3(blocks(M, 256)), dim3(256), 0, S, B.done.as<int32_t>(), Mi, 0);
            for (int pass = 1; pass <= kRefitPasses; ++pass)
                hipLaunchKernelGGL(k_brefit_pass, dim3(blocks(M, 256)), dim3(256), 0, S, kchild, T, o_node, nseg, Mi,
                                   pass, B.lbox.as<const float>(), B.ibox.as<float>(), B.done.as<int32_t>(), n_roots,
                                   err + 1);
            hipLaunchKernelGGL(k_bnodes, dim3(blocks(M, 256)), dim3(256), 0, S, kchild, fl, B.ibox.as<const float>(),
                               B.lbox.as<const float>(), B.cb.as<const float>(), T, o_node, nseg, Mi, leaf_max);
        }
        hipLaunchKernelGGL(k_btris, dim3(blocks(N, 256)), dim3(256), 0, S, vS, T, o_tri, nseg, Ni, err);
        int32_t* flags = B.flags4.as<int32_t>();
        int32_t* idx = flags + M + 1;
        if (M > 0) {
            // 4-wide collapse: flags, one exclusive scan over every scene's nodes, per-scene counts
            const int nb = (int)blocks(M, kScanBlock);
            EOS_TRY(B.scan.ensure(4 * (size_t)nb + 4));
            hipLaunchKernelGGL(k_bflags4, dim3(blocks(M, 256)), dim3(256), 0, S, kparent, fl, T, o_node, nseg, Mi,
                               leaf_max, flags);
            hipLaunchKernelGGL(k_scan_blocks, dim3(nb), dim3(256), 0, S, flags, idx, Mi, B.scan.as<int32_t>());
            hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, S, B.scan.as<int32_t>(), nb);
            hipLaunchKernelGGL(k_scan_add, dim3(blocks(M, 256)), dim3(256), 0, S, idx, Mi, B.scan.as<const int32_t>());
            hipLaunchKernelGGL(k_bcount4, dim3(blocks((size_t)nseg, 64)), dim3(64), 0, S, flags, idx, T, nseg,
                               B.counts.as<int32_t>());
            EOS_HIP(hipMemcpyAsync(cnt4.data(), B.counts.ptr, 8 * (size_t)nseg, hipMemcpyDeviceToHost, S));
            EOS_HIP(hipStreamSynchronize(S));
            for (int k = 0; k < nseg; ++k) {
                eos_scene* s = scenes[seg_scene[k]];
                s->n_nodes4 = cnt4[2 * k];
                EOS_TRY(s->nodes4.ensure(128 * (size_t)std::max(s->n_nodes4, 1)));
                segs[k].nodes4 = s->nodes4.as<Node4>();
                segs[k].base4 = cnt4[2 * k + 1];
            }
            EOS_HIP(hipMemcpyAsync(B.table.ptr, segs.data(), sizeof(BuildSeg) * nseg, hipMemcpyHostToDevice, S));
            hipLaunchKernelGGL(k_bnodes4, dim3(blocks(M, 256)), dim3(256), 0, S, kchild, fl, B.ibox.as<const float>(),
                               B.lbox.as<const float>(), B.cb.as<const float>(), flags, idx, T, o_node, nseg, Mi,
                               leaf_max);
        }
        EOS_HIP(hipGetLastError());
        EOS_HIP(hipMemcpyAsync(herr, err, 12, hipMemcpyDeviceToHost, S));
    }
    EOS_HIP(hipEventRecord(B.ev[1], S));
    EOS_HIP(hipStreamSynchronize(S));
    if (herr[0] != 0)
        return fail(EOS_ERR_INVALID, "eos_build_lbvh_batch: " + std::to_string(herr[0]) + " material indices out of range");
    if (herr[2] != 0)
        return fail(EOS_ERR_INVALID,
                    "eos_build_lbvh_batch: " + std::to_string(herr[2]) + " transformation indices out of range");
    if (herr[1] != n_roots) return fail(EOS_ERR_STATE, "eos_build_lbvh_batch: refit did not reach every root");
    float ms = 0.0f;
    EOS_HIP(hipEventElapsedTime(&ms, B.ev[0], B.ev[1]));
    for (int i = 0; i < n_scenes; ++i) {
        eos_scene* s = scenes[i];
        const int n = s->n_tri;
        if (n == 0) s->root = kDone;
        else if (n <= leaf_max) s->root = ~(0 | ((n - 1) << kLeafShift));
        else s->root = 0;
        s->root4 = s->root;
        s->built = true;
        if (st) {
            st[i].build_ms = ms;
            st[i].num_internal = n > 0 ? n - 1 : 0;
            st[i].root_ref = s->root;
            st[i].leaf_max = leaf_max;
            st[i].width = width;
            st[i].num_nodes4 = s->n_nodes4;
            st[i].root4 = s->root4;
        }
    }
    return EOS_OK;
}

}  // namespace eos
