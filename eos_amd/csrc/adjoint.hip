// adjoint.hip -- reverse-mode derivative of the pre-quantisation image with
// respect to per-material (kd, ks, shininess) and per-light (position, colour)
// parameters, hit sets held fixed.  No counterpart in the reference; the
// function differentiated is its shading (src/RayTracer.cpp:302-458):
//
//   c_0 = sum_l W_l P_l,  W_0 = 1,  W_{l+1} = W_l r_l
//   P_l = sum_{lights i visible at the level-l hit} phong_i
//   phong_i,k = max(0, L.N) kd_k lc_k + [V.R > 0] (V.R)^s lc_k ks_k
//   L = normalize(lp - P), R = normalize(2 (N.L) N - L), V = normalize(cam - P)
//
// Hit points, normals, materials, shadow masks and reflection weights are the
// records the forward pass stored (surfaces and visibility per level, level
// colours with their weights); they are replayed, or re-traced first when the
// caller asks for it or none are stored.  One thread per hit record, FP32
// arithmetic.  Sums are exact: every FP32 contribution becomes a fixed-point
// integer (resolution 2^-64) added into 64-bit integer limbs, per workgroup in
// LDS, then into global limbs; integer addition does not depend on order, so
// the gradients are the same bits whatever the record order (replay, re-trace,
// workgroup scheduling).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "eos_internal.h"
#include "exact_sum.h"
#include "trace.h"

namespace eos {
namespace {

struct AdjArgs {
    SceneView sc;
    DevCamera cam;
    int level, npix;
    const int* count_ptr;          // hits at this level
    const float4* surf;
    const uint8_t* vis;
    const float4* levcol;          // weights of the levels above
    const float* dl;               // [npix][3]
    unsigned long long* gout;      // [3][M*7 + L*6] limbs of the exact sums
    int* err;                      // bit 8: a contribution out of range (non-finite or >= 2^31)
    int lds_acc;                   // accumulate in LDS (else straight to global)
};

__device__ __forceinline__ void acc_add(unsigned long long* lds, unsigned long long* g, int nacc, int idx, float v,
                                        bool use_lds, int* err) {
    if (v == 0.0f) return;
    long long l0, l1, l2;
    if (!to_limbs(v, l0, l1, l2)) {
        atomicOr(err, 8);
        return;
    }
    unsigned long long* base = use_lds ? lds : g;
    if (l0) atomicAdd(base + idx, (unsigned long long)l0);
    if (l1) atomicAdd(base + nacc + idx, (unsigned long long)l1);
    if (l2) atomicAdd(base + 2 * nacc + idx, (unsigned long long)l2);
}

__global__ __launch_bounds__(256) void k_adjoint(AdjArgs a) {
    extern __shared__ unsigned long long acc[];
    const int M = a.sc.n_mat, L = a.sc.n_light;
    const int nacc = 7 * M + 6 * L;
    const bool use_lds = a.lds_acc != 0;
    if (use_lds) {
        for (int i = threadIdx.x; i < 3 * nacc; i += blockDim.x) acc[i] = 0ull;
        __syncthreads();
    }
    const int n = *a.count_ptr;
    const F3 cam = f3(a.cam.position[0], a.cam.position[1], a.cam.position[2]);
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        const float4 s0 = a.surf[2 * (size_t)j], s1 = a.surf[2 * (size_t)j + 1];
        const F3 P = f3(s0.x, s0.y, s0.z), N = f3(s1.x, s1.y, s1.z);
        const int pixel = __float_as_int(s0.w), mi = unpack_mat(__float_as_int(s1.w));
        float W = 1.0f;
        for (int k = 0; k < a.level; ++k) W *= a.levcol[(size_t)k * a.npix + pixel].w;
        const float g0 = W * a.dl[3 * (size_t)pixel], g1 = W * a.dl[3 * (size_t)pixel + 1],
                    g2 = W * a.dl[3 * (size_t)pixel + 2];
        if (g0 == 0.0f && g1 == 0.0f && g2 == 0.0f) continue;
        const float g[3] = {g0, g1, g2};
        const DevMaterial& m = a.sc.mats[mi];
        const F3 V = normalize3(sub(cam, P));
        float gkd[3] = {0, 0, 0}, gks[3] = {0, 0, 0}, gs = 0.0f;
        bool any = false;
        for (int i = 0; i < L; ++i) {
            if (!a.vis[(size_t)j * L + i]) continue;
            any = true;
            const DevLight& lt = a.sc.lights[i];
            const F3 D = sub(f3(lt.position[0], lt.position[1], lt.position[2]), P);
            const float dlen = sqrtf(dot3(D, D));
            const F3 Ld = normalize3(D);
            const float dNL = dot3(N, Ld);
            const float s2 = 2.0f * dNL;
            const F3 Rv = sub(f3(s2 * N.x, s2 * N.y, s2 * N.z), Ld);
            const float rlen = sqrtf(dot3(Rv, Rv));
            const F3 R = normalize3(Rv);
            const float dLN = dot3(Ld, N);
            const bool dif = dLN > 0.0f;
            const float dVR = dot3(V, R);
            const bool sp = dVR > 0.0f;
            const float st = sp ? powf(dVR, m.shininess) : 0.0f;
            float a_dif = 0.0f, b_spec = 0.0f;
            float glc[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const float lc = lt.color[k];
                gkd[k] += dif ? g[k] * dLN * lc : 0.0f;
                gks[k] += g[k] * st * lc;
                glc[k] = g[k] * ((dif ? dLN : 0.0f) * m.kd[k] + st * m.ks[k]);
                a_dif += g[k] * m.kd[k] * lc;
                b_spec += g[k] * m.ks[k] * lc;
            }
            if (sp) gs += b_spec * st * logf(dVR);
            // d/dL of dLN and of dVR (through R), then d/dlp through L = D/|D|
            F3 gL = {0, 0, 0};
            if (dif) gL = scale(a_dif, N);
            if (sp) {
                const float c = b_spec * m.shininess * (st / dVR) / rlen;
                const F3 u = sub(V, scale(dVR, R));              // (I - R R^T) V
                const float nu = dot3(N, u);
                const F3 w = sub(scale(2.0f * nu, N), u);        // (2 N N^T - I) u
                gL = add(gL, scale(c, w));
            }
            const float gLL = dot3(gL, Ld);
            const F3 gD = scale(1.0f / dlen, sub(gL, scale(gLL, Ld)));
            const int lb = 7 * M + 6 * i;
            acc_add(acc, a.gout, nacc, lb + 0, gD.x, use_lds, a.err);
            acc_add(acc, a.gout, nacc, lb + 1, gD.y, use_lds, a.err);
            acc_add(acc, a.gout, nacc, lb + 2, gD.z, use_lds, a.err);
            acc_add(acc, a.gout, nacc, lb + 3, glc[0], use_lds, a.err);
            acc_add(acc, a.gout, nacc, lb + 4, glc[1], use_lds, a.err);
            acc_add(acc, a.gout, nacc, lb + 5, glc[2], use_lds, a.err);
        }
        if (any) {
            const int mb = 7 * mi;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                acc_add(acc, a.gout, nacc, mb + k, gkd[k], use_lds, a.err);
                acc_add(acc, a.gout, nacc, mb + 3 + k, gks[k], use_lds, a.err);
            }
            acc_add(acc, a.gout, nacc, mb + 6, gs, use_lds, a.err);
        }
    }
    if (use_lds) {
        __syncthreads();
        for (int i = threadIdx.x; i < 3 * nacc; i += blockDim.x) {
            const unsigned long long v = acc[i];
            if (v != 0ull) atomicAdd(&a.gout[i], v);
        }
    }
}

__global__ void k_zero_u64(unsigned long long* p, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0ull;
}

// dst_m[M][7] += value of limbs i < 7M, dst_l[L][6] += the rest
__global__ void k_accumulate_out(const unsigned long long* g, int M, int L, double* dm, double* dlg) {
    const int nacc = 7 * M + 6 * L;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nacc; i += gridDim.x * blockDim.x) {
        const double v = limb_value(g, nacc, i);
        if (i < 7 * M) dm[i] += v;
        else dlg[i - 7 * M] += v;
    }
}

constexpr int kLdsAccParams = 2048;   // 48 KiB of limbs (3 x int64 per parameter) per workgroup

}  // namespace

eos_status render_adjoint(eos_scene* s, const eos_camera* cam, const eos_adjoint_opts* o, const float* dl,
                          double* gm, double* gl, eos_adjoint_stats* st) {
    eos_frame_buffers& fb = s->fb;
    hipStream_t S = s->stream;
    const size_t npix = (size_t)cam->width * cam->height;
    const int M = s->n_mat, L = s->n_light;
    const int nacc = 7 * M + 6 * L;
    const bool replay = !o->recompute && fb.rec_valid && fb.depth == o->depth && fb.n_light == L &&
                        std::memcmp(&fb.rec_cam, cam, sizeof(eos_camera)) == 0;
    double retrace_ms = 0.0;
    if (!replay) {
        eos_render_stats rs;
        EOS_TRY(run_frame(s, cam, o->depth, false, nullptr, nullptr, nullptr, &rs));
        retrace_ms = rs.render_ms;
        fb.rec_valid = 1;
        fb.rec_cam = *cam;
    }
    static thread_local DevBuf dl_dev, g_dev;
    const float* dl_ptr = dl;
    if (o->memory != EOS_MEM_DEVICE) {
        EOS_TRY(dl_dev.ensure(12 * npix));
        EOS_HIP(hipMemcpyAsync(dl_dev.ptr, dl, 12 * npix, hipMemcpyHostToDevice, S));
        dl_ptr = dl_dev.as<float>();
    }
    // limbs [3][nacc], then one word for the range flag
    const int words = 3 * nacc + 1;
    EOS_TRY(g_dev.ensure(8 * (size_t)words));
    hipLaunchKernelGGL(k_zero_u64, dim3(std::max(1, std::min(64, (words + 255) / 256))), dim3(256), 0, S,
                       g_dev.as<unsigned long long>(), words);
    int* err = reinterpret_cast<int*>(g_dev.as<unsigned long long>() + 3 * (size_t)nacc);
    int* ctr = fb.counters.as<int>();
    const SceneView sc = view_of(s);
    const bool lds = nacc <= kLdsAccParams;
    const size_t shmem = lds ? 3 * sizeof(unsigned long long) * (size_t)nacc : 0;
    const unsigned grid = (unsigned)s->num_cu * 8;
    EOS_HIP(hipEventRecord(s->ev[0], S));
    for (int l = 0; l <= o->depth && L > 0; ++l) {
        AdjArgs a{};
        a.sc = sc;
        a.cam = *reinterpret_cast<const DevCamera*>(cam);
        a.level = l;
        a.npix = (int)npix;
        a.count_ptr = ctr + CTR_HITS + l;
        a.surf = fb.surf_at(l);
        a.vis = fb.vis_at(l);
        a.levcol = fb.levcol.as<float4>();
        a.dl = dl_ptr;
        a.gout = g_dev.as<unsigned long long>();
        a.err = err;
        a.lds_acc = lds ? 1 : 0;
        hipLaunchKernelGGL(k_adjoint, dim3(grid), dim3(256), shmem, S, a);
    }
    EOS_HIP(hipGetLastError());
    EOS_HIP(hipEventRecord(s->ev[1], S));
    int herr = 0;
    EOS_HIP(hipMemcpyAsync(&herr, err, sizeof herr, hipMemcpyDeviceToHost, S));
    if (o->memory == EOS_MEM_DEVICE) {
        hipLaunchKernelGGL(k_accumulate_out, dim3(std::max(1, std::min(64, (nacc + 255) / 256))), dim3(256), 0, S,
                           g_dev.as<const unsigned long long>(), M, L, gm, gl);
        EOS_HIP(hipStreamSynchronize(S));
    } else {
        std::vector<unsigned long long> h((size_t)std::max(3 * nacc, 1));
        EOS_HIP(hipMemcpyAsync(h.data(), g_dev.ptr, 8 * 3 * (size_t)nacc, hipMemcpyDeviceToHost, S));
        EOS_HIP(hipStreamSynchronize(S));
        for (int i = 0; i < 7 * M; ++i) gm[i] += limb_value(h.data(), nacc, i);
        for (int i = 0; i < 6 * L; ++i) gl[i] += limb_value(h.data(), nacc, 7 * M + i);
    }
    if (herr & 8)
        return fail(EOS_ERR_LIMIT, "eos_render_adjoint: a gradient contribution is not finite or exceeds 2^31 "
                                   "(a light on a surface?); the gradients are incomplete");
    float ms = 0.0f;
    EOS_HIP(hipEventElapsedTime(&ms, s->ev[0], s->ev[1]));
    int hits[EOS_MAX_DEPTH + 1] = {0};
    EOS_HIP(hipMemcpy(hits, ctr + CTR_HITS, sizeof(int) * (o->depth + 1), hipMemcpyDeviceToHost));
    if (st) {
        st->adjoint_ms = ms;
        st->retrace_ms = retrace_ms;
        st->records = 0;
        for (int l = 0; l <= o->depth; ++l) st->records += hits[l];
    }
    return EOS_OK;
}

}  // namespace eos
