// eos_api.cpp -- C ABI entry points: scene lifetime, parameters, diagnostics.
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "eos_internal.h"

namespace eos {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

eos_status fail(eos_status st, const std::string& msg) {
    g_last_error = msg;
    return st;
}

eos_status DevBuf::ensure(size_t want) {
    if (want <= bytes && ptr) return EOS_OK;
    release();
    if (want == 0) return EOS_OK;
    size_t b = (want + 255) & ~size_t(255);
    hipError_t e = hipMalloc(&ptr, b);
    if (e != hipSuccess) {
        ptr = nullptr;
        bytes = 0;
        return fail(EOS_ERR_NOMEM, "hipMalloc(" + std::to_string(b) + "): " + hipGetErrorString(e));
    }
    bytes = b;
    return EOS_OK;
}

void DevBuf::release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

SceneView view_of(const eos_scene* s) {
    SceneView v;
    v.nodes = s->nodes.as<const Node>();
    v.nodes4 = s->nodes4.as<const Node4>();
    v.root4 = s->root4;
    v.width = s->width;
    v.geo = s->geo.as<const float4>();
    v.shade = s->shade.as<const float4>();
    v.mats = s->mats.as<const DevMaterial>();
    v.lights = s->lights.as<const DevLight>();
    v.xf = s->n_xf > 0 ? s->xf.as<const float4>() : nullptr;
    v.n_xf = s->n_xf;
    v.ambient = s->ambient_index;
    v.root = s->root;
    v.n_tri = s->n_tri;
    v.n_light = s->n_light;
    v.n_mat = s->n_mat;
    return v;
}

static eos_status copy_in(DevBuf& dst, const void* src, size_t bytes, int memory, hipStream_t st) {
    EOS_TRY(dst.ensure(bytes));
    if (bytes == 0) return EOS_OK;
    EOS_HIP(hipMemcpyAsync(dst.ptr, src, bytes, memory == EOS_MEM_DEVICE ? hipMemcpyDeviceToDevice
                                                                         : hipMemcpyHostToDevice, st));
    return EOS_OK;
}

static void free_scene(eos_scene* s) {
    DevBuf* bufs[] = {&s->verts, &s->normals, &s->matidx, &s->mats, &s->lights, &s->nodes, &s->geo,
                      &s->shade, &s->codes, &s->codes_sorted, &s->perm, &s->kchild, &s->kparent,
                      &s->first_last, &s->ibox, &s->lbox, &s->flags, &s->cbounds, &s->sort_tmp_k,
                      &s->sort_tmp_v, &s->sort_hist, &s->bounds_part, &s->nodes4, &s->flags4, &s->scan_tmp,
                      &s->xf, &s->xf_index, &s->wverts, &s->q_rays, &s->q_out, &s->q_rec, &s->q_ctr, &s->q_gstack,
                      &s->fb.hit0, &s->fb.q[0], &s->fb.q[1],
                      &s->fb.hitq, &s->fb.levcol, &s->fb.levels, &s->fb.counters, &s->fb.gstack, &s->fb.aa_acc,
                      &s->fb.tiles};
    for (DevBuf* b : bufs) b->release();
    s->fb.surf.release();
    s->fb.vis.release();
    for (int k = 0; k < eos_scene::kFrameSlots; ++k) {
        eos_frame_buffers* fb = k == 0 ? &s->fb : &s->fbx[k - 1];
        for (DevBuf* b : {&fb->hit0, &fb->q[0], &fb->q[1], &fb->hitq, &fb->levcol, &fb->levels, &fb->counters,
                          &fb->gstack, &fb->aa_acc, &fb->surf, &fb->vis, &fb->rayhit, &fb->aux, &fb->hitcol,
                          &fb->tiles})
            b->release();
        if (k > 0) {
            for (hipEvent_t e : s->xsync[k - 1])
                if (e) (void)hipEventDestroy(e);
            for (hipStream_t st : s->xstream[k - 1])
                if (st) (void)hipStreamDestroy(st);
        }
        if (s->resolved[k]) (void)hipEventDestroy(s->resolved[k]);
        if (s->seq_ev[k]) (void)hipEventDestroy(s->seq_ev[k]);
    }
    if (s->seq_start) (void)hipEventDestroy(s->seq_start);
    for (hipEvent_t e : s->first_ev) (void)hipEventDestroy(e);
    s->first_ev.clear();
    if (s->host_ctr) (void)hipHostFree(s->host_ctr);
    for (hipEvent_t& e : s->ev)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t& e : s->sync)
        if (e) (void)hipEventDestroy(e);
    if (s->stream2) (void)hipStreamDestroy(s->stream2);
    if (s->stream) (void)hipStreamDestroy(s->stream);
}

static eos_status check_materials(const eos_material* m, int n) {
    for (int i = 0; i < n; ++i) {
        if (!(m[i].shininess >= 0.0f)) return fail(EOS_ERR_INVALID, "material shininess must be >= 0");
        if (m[i].transparent && !(m[i].refractive_index > 0.0f && std::isfinite(m[i].refractive_index)))
            return fail(EOS_ERR_INVALID, "a transparent material needs a positive finite refractive index");
    }
    return EOS_OK;
}

static eos_status check_lights(const eos_light* l, int n) {
    for (int i = 0; i < n; ++i)
        if (l[i].position[3] != 1.0f)
            return fail(EOS_ERR_INVALID, "light position w must be 1 (src/RayTracer.cpp:433 subtracts a w=1 point)");
    return EOS_OK;
}

// glm 0.9.6 compute_inverse for a column-major 4x4 (glm/detail/type_mat4x4.inl:
// 37-96): the 2x2 cofactors of the lower rows, combined per column as
// (a - b) + c with alternating signs, the determinant from the first row and
// column, and one reciprocal that scales every entry.  Float, in that order, so
// the inverse is the one the reference's glm::inverse computes.
static void inverse_glm(const float* m, float* out) {
    auto M = [&](int c, int r) { return m[4 * c + r]; };
    const float a[18] = {
        M(2, 2) * M(3, 3) - M(3, 2) * M(2, 3), M(1, 2) * M(3, 3) - M(3, 2) * M(1, 3),
        M(1, 2) * M(2, 3) - M(2, 2) * M(1, 3), M(2, 1) * M(3, 3) - M(3, 1) * M(2, 3),
        M(1, 1) * M(3, 3) - M(3, 1) * M(1, 3), M(1, 1) * M(2, 3) - M(2, 1) * M(1, 3),
        M(2, 1) * M(3, 2) - M(3, 1) * M(2, 2), M(1, 1) * M(3, 2) - M(3, 1) * M(1, 2),
        M(1, 1) * M(2, 2) - M(2, 1) * M(1, 2), M(2, 0) * M(3, 3) - M(3, 0) * M(2, 3),
        M(1, 0) * M(3, 3) - M(3, 0) * M(1, 3), M(1, 0) * M(2, 3) - M(2, 0) * M(1, 3),
        M(2, 0) * M(3, 2) - M(3, 0) * M(2, 2), M(1, 0) * M(3, 2) - M(3, 0) * M(1, 2),
        M(1, 0) * M(2, 2) - M(2, 0) * M(1, 2), M(2, 0) * M(3, 1) - M(3, 0) * M(2, 1),
        M(1, 0) * M(3, 1) - M(3, 0) * M(1, 1), M(1, 0) * M(2, 1) - M(2, 0) * M(1, 1)};
    // factor k of cofactor group g: group g's first entry serves columns 0 and 1
    auto F = [&](int g, int k) { return a[3 * g + (k == 0 ? 0 : k - 1)]; };
    auto V = [&](int r, int k) { return k == 0 ? M(1, r) : M(0, r); };
    float inv[16];
    for (int k = 0; k < 4; ++k) {
        const float sa = (k & 1) ? -1.0f : 1.0f, sb = -sa;
        inv[0 + k] = ((V(1, k) * F(0, k) - V(2, k) * F(1, k)) + V(3, k) * F(2, k)) * sa;
        inv[4 + k] = ((V(0, k) * F(0, k) - V(2, k) * F(3, k)) + V(3, k) * F(4, k)) * sb;
        inv[8 + k] = ((V(0, k) * F(1, k) - V(1, k) * F(3, k)) + V(3, k) * F(5, k)) * sa;
        inv[12 + k] = ((V(0, k) * F(2, k) - V(1, k) * F(4, k)) + V(2, k) * F(5, k)) * sb;
    }
    const float det = (M(0, 0) * inv[0] + M(0, 1) * inv[4]) + (M(0, 2) * inv[8] + M(0, 3) * inv[12]);
    const float one_over = 1.0f / det;
    for (int k = 0; k < 16; ++k) out[k] = inv[k] * one_over;
}

}  // namespace eos

using namespace eos;

extern "C" {

const char* eos_last_error(void) { return g_last_error.c_str(); }

int32_t eos_abi_version(void) { return EOS_ABI_VERSION; }

eos_status eos_device_name(char* buf, int32_t len) {
    if (!buf || len <= 0) return fail(EOS_ERR_INVALID, "eos_device_name: null buffer");
    int dev = 0, count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return fail(EOS_ERR_NODEVICE, "no HIP device");
    EOS_HIP(hipGetDevice(&dev));
    hipDeviceProp_t p;
    EOS_HIP(hipGetDeviceProperties(&p, dev));
    std::snprintf(buf, (size_t)len, "%s (%s, %d CUs)", p.name, p.gcnArchName, p.multiProcessorCount);
    return EOS_OK;
}

eos_status eos_create_scene(const eos_scene_desc* d, eos_scene** out) {
    if (!d || !out) return fail(EOS_ERR_INVALID, "eos_create_scene: null argument");
    *out = nullptr;
    if (d->num_triangles < 0 || d->num_triangles >= (1 << kLeafShift))
        return fail(EOS_ERR_LIMIT, "eos_create_scene: num_triangles must be in [0, 2^28)");
    if (d->num_triangles > 0 && (!d->vertices || !d->normals || !d->material_index))
        return fail(EOS_ERR_INVALID, "eos_create_scene: null triangle arrays");
    if (d->num_materials <= 0 || !d->materials)
        return fail(EOS_ERR_INVALID, "eos_create_scene: at least one material is required");
    if (d->num_lights < 0 || d->num_lights > EOS_MAX_LIGHTS || (d->num_lights > 0 && !d->lights))
        return fail(EOS_ERR_LIMIT, "eos_create_scene: num_lights must be in [0, 32]");
    EOS_TRY(check_materials(d->materials, d->num_materials));
    EOS_TRY(check_lights(d->lights, d->num_lights));
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
        return fail(EOS_ERR_NODEVICE, "eos_create_scene: no HIP device");
    // Material indices are validated on the host when they are host memory.
    if (d->triangle_memory == EOS_MEM_HOST)
        for (int i = 0; i < d->num_triangles; ++i)
            if (d->material_index[i] < 0 || d->material_index[i] >= d->num_materials)
                return fail(EOS_ERR_INVALID, "eos_create_scene: material index out of range at triangle " +
                                                 std::to_string(i));
    int device = 0;
    EOS_HIP(hipGetDevice(&device));
    hipDeviceProp_t p;
    EOS_HIP(hipGetDeviceProperties(&p, device));
    eos_scene* s = new eos_scene();
    s->device = device;
    s->num_cu = p.multiProcessorCount;
    hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->stream2, hipStreamNonBlocking);
    if (e != hipSuccess) {
        free_scene(s);
        delete s;
        return fail(EOS_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    for (hipEvent_t& ev : s->ev) {
        if (hipEventCreate(&ev) != hipSuccess) {
            free_scene(s);
            delete s;
            return fail(EOS_ERR_HIP, "hipEventCreate failed");
        }
    }
    for (hipEvent_t& ev : s->sync) {
        if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
            free_scene(s);
            delete s;
            return fail(EOS_ERR_HIP, "hipEventCreate failed");
        }
    }
    s->n_tri = d->num_triangles;
    s->n_mat = d->num_materials;
    for (int i = 0; i < d->num_materials; ++i) s->any_transparent |= d->materials[i].transparent != 0;
    s->n_light = d->num_lights;
    const size_t n = (size_t)d->num_triangles;
    eos_status st = EOS_OK;
    if (st == EOS_OK) st = copy_in(s->verts, d->vertices, n * 9 * sizeof(float), d->triangle_memory, s->stream);
    if (st == EOS_OK) st = copy_in(s->normals, d->normals, n * 9 * sizeof(float), d->triangle_memory, s->stream);
    if (st == EOS_OK) st = copy_in(s->matidx, d->material_index, n * sizeof(int32_t), d->triangle_memory, s->stream);
    if (st == EOS_OK) st = copy_in(s->mats, d->materials, sizeof(eos_material) * d->num_materials, EOS_MEM_HOST, s->stream);
    if (st == EOS_OK) st = s->lights.ensure(sizeof(eos_light) * EOS_MAX_LIGHTS);
    if (st == EOS_OK && d->num_lights > 0)
        st = copy_in(s->lights, d->lights, sizeof(eos_light) * d->num_lights, EOS_MEM_HOST, s->stream);
    if (st == EOS_OK && hipStreamSynchronize(s->stream) != hipSuccess)
        st = fail(EOS_ERR_HIP, "eos_create_scene: upload failed");
    if (st != EOS_OK) {
        free_scene(s);
        delete s;
        return st;
    }
    *out = s;
    return EOS_OK;
}

eos_status eos_destroy_scene(eos_scene* s) {
    if (!s) return EOS_OK;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    free_scene(s);
    delete s;
    return EOS_OK;
}

eos_status eos_set_transforms(eos_scene* s, int32_t nt, const float* mats, const int32_t* index, int32_t memory) {
    if (!s) return fail(EOS_ERR_INVALID, "eos_set_transforms: null scene");
    if (nt < 0) return fail(EOS_ERR_INVALID, "eos_set_transforms: negative transformation count");
    if (nt > 0 && (!mats || (s->n_tri > 0 && !index)))
        return fail(EOS_ERR_INVALID, "eos_set_transforms: null transformation arrays");
    if (memory != EOS_MEM_HOST && memory != EOS_MEM_DEVICE)
        return fail(EOS_ERR_INVALID, "eos_set_transforms: index_memory must be EOS_MEM_HOST or EOS_MEM_DEVICE");
    EOS_HIP(hipSetDevice(s->device));
    s->built = false;  // the LBVH is built over world-space vertices
    if (nt == 0) {
        s->n_xf = 0;
        return EOS_OK;
    }
    std::vector<float> rows((size_t)nt * kXfRows * 4);
    for (int t = 0; t < nt; ++t) {
        const float* m = mats + 16 * (size_t)t;
        float inv[16];
        inverse_glm(m, inv);
        bool finite = true;
        for (float x : inv) finite &= std::isfinite(x);
        if (!finite)
            return fail(EOS_ERR_INVALID, "eos_set_transforms: transformation " + std::to_string(t) +
                                             " is not invertible");
        float* r = rows.data() + (size_t)t * kXfRows * 4;
        for (int k = 0; k < 3; ++k)
            for (int c = 0; c < 4; ++c) {
                r[4 * (kXfInv + k) + c] = inv[4 * c + k];
                r[4 * (kXfM + k) + c] = m[4 * c + k];
                r[4 * (kXfInvT + k) + c] = inv[4 * k + c];  // transpose(inverse)[c][k] = inverse[k][c]
            }
    }
    if (memory == EOS_MEM_HOST)
        for (int i = 0; i < s->n_tri; ++i)
            if (index[i] < -1 || index[i] >= nt)
                return fail(EOS_ERR_INVALID, "eos_set_transforms: transformation index out of range at triangle " +
                                                 std::to_string(i));
    s->n_xf = 0;
    EOS_TRY(copy_in(s->xf, rows.data(), rows.size() * sizeof(float), EOS_MEM_HOST, s->stream));
    EOS_TRY(copy_in(s->xf_index, index, (size_t)s->n_tri * sizeof(int32_t), memory, s->stream));
    EOS_HIP(hipStreamSynchronize(s->stream));
    s->n_xf = nt;
    return EOS_OK;
}

eos_status eos_update_materials(eos_scene* s, const eos_material* m, int32_t n) {
    if (!s || !m) return fail(EOS_ERR_INVALID, "eos_update_materials: null argument");
    if (n != s->n_mat) return fail(EOS_ERR_INVALID, "eos_update_materials: material count differs from the scene's");
    EOS_TRY(check_materials(m, n));
    EOS_HIP(hipMemcpyAsync(s->mats.ptr, m, sizeof(eos_material) * n, hipMemcpyHostToDevice, s->stream));
    EOS_HIP(hipStreamSynchronize(s->stream));
    s->fb.rec_valid = 0;  // reflection flags and weights shape the stored records
    ++s->mat_generation;  // cached batched frames were shaped by the old flags
    s->any_transparent = false;
    for (int i = 0; i < n; ++i) s->any_transparent |= m[i].transparent != 0;
    return EOS_OK;
}

eos_status eos_set_ambient_refractive_index(eos_scene* s, float n) {
    if (!s) return fail(EOS_ERR_INVALID, "eos_set_ambient_refractive_index: null scene");
    if (!(n > 0.0f && std::isfinite(n)))
        return fail(EOS_ERR_INVALID, "eos_set_ambient_refractive_index: the index must be positive and finite");
    s->ambient_index = n;
    return EOS_OK;
}

eos_status eos_update_lights(eos_scene* s, const eos_light* l, int32_t n) {
    if (!s || (!l && n > 0)) return fail(EOS_ERR_INVALID, "eos_update_lights: null argument");
    if (n < 0 || n > EOS_MAX_LIGHTS) return fail(EOS_ERR_LIMIT, "eos_update_lights: at most 32 lights");
    EOS_TRY(check_lights(l, n));
    if (n > 0) EOS_HIP(hipMemcpyAsync(s->lights.ptr, l, sizeof(eos_light) * n, hipMemcpyHostToDevice, s->stream));
    EOS_HIP(hipStreamSynchronize(s->stream));
    s->fb.rec_valid = 0;  // light positions decide the stored shadow masks
    s->n_light = n;
    return EOS_OK;
}

eos_status eos_build_lbvh_ex(eos_scene* s, const eos_build_opts* o, eos_build_stats* st) {
    if (!s || !o) return fail(EOS_ERR_INVALID, "eos_build_lbvh_ex: null argument");
    if (o->leaf_max < 1 || o->leaf_max > 8) return fail(EOS_ERR_INVALID, "eos_build_lbvh: leaf_max must be in [1, 8]");
    if (o->width != 2 && o->width != 4) return fail(EOS_ERR_INVALID, "eos_build_lbvh: width must be 2 or 4");
    EOS_HIP(hipSetDevice(s->device));
    return build_lbvh(s, o->leaf_max, o->width, st);
}

eos_status eos_build_lbvh_batch(eos_scene* const* scenes, int32_t n, const eos_build_opts* o, eos_build_stats* st) {
    if (!scenes || !o || n <= 0) return fail(EOS_ERR_INVALID, "eos_build_lbvh_batch: bad argument");
    if (o->leaf_max < 1 || o->leaf_max > 8)
        return fail(EOS_ERR_INVALID, "eos_build_lbvh_batch: leaf_max must be in [1, 8]");
    if (o->width != 2 && o->width != 4) return fail(EOS_ERR_INVALID, "eos_build_lbvh_batch: width must be 2 or 4");
    if (n > 65536) return fail(EOS_ERR_LIMIT, "eos_build_lbvh_batch: at most 65536 scenes");
    for (int i = 0; i < n; ++i) {
        if (!scenes[i]) return fail(EOS_ERR_INVALID, "eos_build_lbvh_batch: null scene");
        if (scenes[i]->device != scenes[0]->device)
            return fail(EOS_ERR_INVALID, "eos_build_lbvh_batch: scenes on different devices");
        for (int j = 0; j < i; ++j)
            if (scenes[j] == scenes[i]) return fail(EOS_ERR_INVALID, "eos_build_lbvh_batch: a scene appears twice");
    }
    EOS_HIP(hipSetDevice(scenes[0]->device));
    return build_lbvh_batch(scenes, n, o->leaf_max, o->width, st);
}

eos_status eos_build_lbvh(eos_scene* s, int32_t leaf_max, eos_build_stats* st) {
    eos_build_opts o{leaf_max, 4};
    return eos_build_lbvh_ex(s, &o, st);
}

eos_status eos_render(eos_scene* s, const eos_camera* cam, const eos_render_opts* o, uint8_t* rgba,
                      eos_render_stats* st) {
    if (!s || !cam || !o) return fail(EOS_ERR_INVALID, "eos_render: null argument");
    if (!s->built) return fail(EOS_ERR_STATE, "eos_render: call eos_build_lbvh first");
    if (cam->width <= 0 || cam->height <= 0 || (int64_t)cam->width * cam->height > (int64_t)1 << 30)
        return fail(EOS_ERR_INVALID, "eos_render: bad image size");
    if (o->depth < 0 || o->depth > EOS_MAX_DEPTH) return fail(EOS_ERR_INVALID, "eos_render: depth must be in [0, 8]");
    if (o->aa_samples < 0 || o->aa_samples > EOS_MAX_AA)
        return fail(EOS_ERR_INVALID, "eos_render: aa_samples must be in [0, 16]");
    if (o->aa_samples > 1 && o->keep_records)
        return fail(EOS_ERR_INVALID, "eos_render: keep_records needs one sample per pixel (aa_samples 0 or 1)");
    if (o->keep_records && s->any_transparent)
        return fail(EOS_ERR_INVALID, "eos_render: keep_records needs chains of reflections (no transparent materials)");
    EOS_HIP(hipSetDevice(s->device));
    return render(s, cam, o, rgba, st);
}

eos_status eos_render_rows(eos_scene* s, const eos_camera* cam, int32_t row_begin, int32_t row_end,
                           const eos_render_opts* o, uint8_t* rgba, eos_render_stats* st) {
    if (!s || !cam || !o) return fail(EOS_ERR_INVALID, "eos_render_rows: null argument");
    if (!s->built) return fail(EOS_ERR_STATE, "eos_render_rows: call eos_build_lbvh first");
    if (cam->width <= 0 || cam->height <= 0 || (int64_t)cam->width * cam->height > (int64_t)1 << 30)
        return fail(EOS_ERR_INVALID, "eos_render_rows: bad image size");
    if (row_begin < 0 || row_end > cam->height || row_begin >= row_end)
        return fail(EOS_ERR_INVALID, "eos_render_rows: rows must satisfy 0 <= row_begin < row_end <= height");
    if (o->depth < 0 || o->depth > EOS_MAX_DEPTH) return fail(EOS_ERR_INVALID, "eos_render_rows: depth must be in [0, 8]");
    if (o->aa_samples < 0 || o->aa_samples > EOS_MAX_AA)
        return fail(EOS_ERR_INVALID, "eos_render_rows: aa_samples must be in [0, 16]");
    if (o->keep_records)
        return fail(EOS_ERR_INVALID, "eos_render_rows: keep_records needs the whole frame (eos_render)");
    EOS_HIP(hipSetDevice(s->device));
    return render(s, cam, o, rgba, st, row_begin, row_end);
}

eos_status eos_render_frames(eos_scene* s, const eos_camera* cams, int32_t n, const eos_render_opts* o,
                             uint8_t* const* rgba, eos_render_stats* st) {
    if (!s || !cams || !o || !rgba || n < 0) return fail(EOS_ERR_INVALID, "eos_render_frames: null argument");
    if (!s->built) return fail(EOS_ERR_STATE, "eos_render_frames: call eos_build_lbvh first");
    if (o->depth < 0 || o->depth > EOS_MAX_DEPTH)
        return fail(EOS_ERR_INVALID, "eos_render_frames: depth must be in [0, 8]");
    if (o->output_memory != EOS_MEM_DEVICE || o->rgb_out || o->hit_out || o->keep_records || o->aa_samples > 1)
        return fail(EOS_ERR_INVALID, "eos_render_frames: RGBA8 device outputs only (no rgb_out, hit_out, "
                                     "keep_records or anti-aliasing)");
    for (int i = 0; i < n; ++i) {
        if (!rgba[i]) return fail(EOS_ERR_INVALID, "eos_render_frames: null output");
        if (cams[i].width <= 0 || cams[i].height <= 0 || (int64_t)cams[i].width * cams[i].height > (int64_t)1 << 30)
            return fail(EOS_ERR_INVALID, "eos_render_frames: bad image size");
    }
    if (n == 0) return EOS_OK;
    EOS_HIP(hipSetDevice(s->device));
    return render_frames(s, cams, n, o, rgba, st);
}

eos_status eos_render_adjoint(eos_scene* s, const eos_camera* cam, const eos_adjoint_opts* o, const float* dl,
                              double* gm, double* gl, eos_adjoint_stats* st) {
    if (!s || !cam || !o || !dl || !gm || !gl) return fail(EOS_ERR_INVALID, "eos_render_adjoint: null argument");
    if (!s->built) return fail(EOS_ERR_STATE, "eos_render_adjoint: call eos_build_lbvh first");
    if (s->any_transparent)
        return fail(EOS_ERR_INVALID, "eos_render_adjoint: scenes with transparent materials have no adjoint");
    if (o->depth < 0 || o->depth > EOS_MAX_DEPTH) return fail(EOS_ERR_INVALID, "eos_render_adjoint: bad depth");
    EOS_HIP(hipSetDevice(s->device));
    return render_adjoint(s, cam, o, dl, gm, gl, st);
}

eos_status eos_trace_rays(eos_scene* s, int32_t n, const float* o, const float* d, const eos_trace_opts* opts,
                          eos_ray_hit* hits, uint8_t* occ) {
    if (!s || !opts) return fail(EOS_ERR_INVALID, "eos_trace_rays: null argument");
    if (n < 0) return fail(EOS_ERR_INVALID, "eos_trace_rays: negative ray count");
    if (opts->mode != EOS_QUERY_NEAREST && opts->mode != EOS_QUERY_OCCLUDED && opts->mode != EOS_QUERY_NEAREST_LINEAR)
        return fail(EOS_ERR_INVALID, "eos_trace_rays: unknown query mode");
    if (opts->memory != EOS_MEM_HOST && opts->memory != EOS_MEM_DEVICE)
        return fail(EOS_ERR_INVALID, "eos_trace_rays: memory must be EOS_MEM_HOST or EOS_MEM_DEVICE");
    if (n > 0 && (!o || !d)) return fail(EOS_ERR_INVALID, "eos_trace_rays: null ray arrays");
    if (n > 0 && opts->mode == EOS_QUERY_OCCLUDED && !occ)
        return fail(EOS_ERR_INVALID, "eos_trace_rays: EOS_QUERY_OCCLUDED writes `occluded`, which is null");
    if (n > 0 && opts->mode != EOS_QUERY_OCCLUDED && !hits)
        return fail(EOS_ERR_INVALID, "eos_trace_rays: nearest-hit queries write `hits`, which is null");
    if (!s->built) return fail(EOS_ERR_STATE, "eos_trace_rays: call eos_build_lbvh first");
    if (n == 0) return EOS_OK;
    EOS_HIP(hipSetDevice(s->device));
    return trace_rays(s, n, o, d, opts, hits, occ);
}

eos_status eos_trace_rays_static(eos_scene* s, int32_t n, const float* o, const float* d, const eos_trace_opts* opts,
                                 eos_ray_hit* hits, uint8_t* occ) {
    if (!s || !opts) return fail(EOS_ERR_INVALID, "eos_trace_rays_static: null argument");
    if (n < 0) return fail(EOS_ERR_INVALID, "eos_trace_rays_static: negative ray count");
    if (n > EOS_STATIC_MAX_RAYS)
        return fail(EOS_ERR_LIMIT, "eos_trace_rays_static: " + std::to_string(n) + " rays; at most " +
                                       std::to_string(EOS_STATIC_MAX_RAYS) + " per call (the fixed scratch array)");
    if (opts->mode != EOS_QUERY_NEAREST && opts->mode != EOS_QUERY_OCCLUDED && opts->mode != EOS_QUERY_NEAREST_LINEAR)
        return fail(EOS_ERR_INVALID, "eos_trace_rays_static: unknown query mode");
    if (opts->memory != EOS_MEM_DEVICE)
        return fail(EOS_ERR_INVALID, "eos_trace_rays_static: rays and outputs must be device memory (EOS_MEM_DEVICE)");
    if (n > 0 && (!o || !d)) return fail(EOS_ERR_INVALID, "eos_trace_rays_static: null ray arrays");
    if (n > 0 && opts->mode == EOS_QUERY_OCCLUDED && !occ)
        return fail(EOS_ERR_INVALID, "eos_trace_rays_static: EOS_QUERY_OCCLUDED writes `occluded`, which is null");
    if (n > 0 && opts->mode != EOS_QUERY_OCCLUDED && !hits)
        return fail(EOS_ERR_INVALID, "eos_trace_rays_static: nearest-hit queries write `hits`, which is null");
    if (!s->built) return fail(EOS_ERR_STATE, "eos_trace_rays_static: call eos_build_lbvh first");
    if (n == 0) return EOS_OK;
    EOS_HIP(hipSetDevice(s->device));
    return trace_rays_static(s, n, o, d, opts, hits, occ);
}

eos_status eos_render_batch(eos_scene* const* scenes, const eos_camera* cams, int32_t n, const eos_batch_opts* o,
                            uint8_t* const* out, eos_batch_stats* st) {
    if (!scenes || !cams || !o || !out || n <= 0) return fail(EOS_ERR_INVALID, "eos_render_batch: bad argument");
    for (int i = 0; i < n; ++i) {
        if (!scenes[i] || !scenes[i]->built) return fail(EOS_ERR_STATE, "eos_render_batch: scene without LBVH");
        if (cams[i].width <= 0 || cams[i].height <= 0) return fail(EOS_ERR_INVALID, "eos_render_batch: bad image size");
        if (scenes[i]->device != scenes[0]->device) return fail(EOS_ERR_INVALID, "eos_render_batch: scenes on different devices");
    }
    if (o->depth < 0 || o->depth > EOS_MAX_DEPTH) return fail(EOS_ERR_INVALID, "eos_render_batch: bad depth");
    EOS_HIP(hipSetDevice(scenes[0]->device));
    return render_batch(scenes, cams, n, o, out, st);
}

eos_status eos_set_launch_config(eos_scene* s, const eos_launch_config* c) {
    if (!s || !c) return fail(EOS_ERR_INVALID, "eos_set_launch_config: null argument");
    if (c->workgroup != 64 && c->workgroup != 128 && c->workgroup != 256)
        return fail(EOS_ERR_INVALID, "workgroup must be 64, 128 or 256");
    if (c->stack_lds != 8 && c->stack_lds != 16 && c->stack_lds != 32)
        return fail(EOS_ERR_INVALID, "stack_lds must be 8, 16 or 32");
    if (c->waves_per_cu < 1 || c->waves_per_cu > 64) return fail(EOS_ERR_INVALID, "waves_per_cu must be in [1, 64]");
    if (c->refill < 1 || c->refill > 64) return fail(EOS_ERR_INVALID, "refill must be in [1, 64]");
    if (c->batch < 1 || c->batch > 4096) return fail(EOS_ERR_INVALID, "batch must be in [1, 4096]");
    if (c->primary_packets != 0 && c->primary_packets != 1)
        return fail(EOS_ERR_INVALID, "primary_packets must be 0 or 1");
    if (c->coop < 0 || c->coop > 4) return fail(EOS_ERR_INVALID, "coop must be in [0, 4]");
    if (c->frames_waves_per_cu < 1 || c->frames_waves_per_cu > 64)
        return fail(EOS_ERR_INVALID, "frames_waves_per_cu must be in [1, 64]");
    if (c->frames_refill < 1 || c->frames_refill > 64) return fail(EOS_ERR_INVALID, "frames_refill must be in [1, 64]");
    if (c->frames_batch < 1 || c->frames_batch > 4096)
        return fail(EOS_ERR_INVALID, "frames_batch must be in [1, 4096]");
    if (c->group < 0 || c->group > 64) return fail(EOS_ERR_INVALID, "group must be in [0, 64]");
    if (c->steal != 0 && c->steal != 1) return fail(EOS_ERR_INVALID, "steal must be 0 or 1");
    if (c->scatter < 0 || c->scatter > 64) return fail(EOS_ERR_INVALID, "scatter must be in [0, 64]");
    if (c->group_lanes != 16 && c->group_lanes != 64) return fail(EOS_ERR_INVALID, "group_lanes must be 16 or 64");
    if (c->cull != 0 && c->cull != 1) return fail(EOS_ERR_INVALID, "cull must be 0 or 1");
    s->cfg.workgroup = c->workgroup;
    s->cfg.stack_lds = c->stack_lds;
    s->cfg.waves_per_cu = c->waves_per_cu;
    s->cfg.refill = c->refill;
    s->cfg.batch = c->batch;
    s->cfg.packets = c->primary_packets;
    s->cfg.coop = c->coop;
    s->cfg.group = c->group;
    s->cfg.steal = c->steal;
    s->cfg.scatter = c->scatter;
    s->cfg.group_lanes = c->group_lanes;
    s->cfg.cull = c->cull;
    s->cfg.frames_waves_per_cu = c->frames_waves_per_cu;
    s->cfg.frames_refill = c->frames_refill;
    s->cfg.frames_batch = c->frames_batch;
    return EOS_OK;
}

eos_status eos_get_launch_config(const eos_scene* s, eos_launch_config* c) {
    if (!s || !c) return fail(EOS_ERR_INVALID, "eos_get_launch_config: null argument");
    c->workgroup = s->cfg.workgroup;
    c->stack_lds = s->cfg.stack_lds;
    c->waves_per_cu = s->cfg.waves_per_cu;
    c->refill = s->cfg.refill;
    c->batch = s->cfg.batch;
    c->primary_packets = s->cfg.packets;
    c->coop = s->cfg.coop;
    c->group = s->cfg.group;
    c->steal = s->cfg.steal;
    c->scatter = s->cfg.scatter;
    c->group_lanes = s->cfg.group_lanes;
    c->cull = s->cfg.cull;
    c->frames_waves_per_cu = s->cfg.frames_waves_per_cu;
    c->frames_refill = s->cfg.frames_refill;
    c->frames_batch = s->cfg.frames_batch;
    return EOS_OK;
}

eos_status eos_debug_read(const eos_scene* s, int32_t what, void* dst, int64_t bytes) {
    if (!s || !dst || bytes < 0) return fail(EOS_ERR_INVALID, "eos_debug_read: bad argument");
    const DevBuf* b = nullptr;
    size_t need = 0;
    const size_t n = (size_t)s->n_tri, m = n > 0 ? n - 1 : 0;
    switch (what) {
        case EOS_DBG_MORTON_SORTED: b = &s->codes_sorted; need = 4 * n; break;
        case EOS_DBG_PERM: b = &s->perm; need = 4 * n; break;
        case EOS_DBG_KCHILD: b = &s->kchild; need = 8 * m; break;
        case EOS_DBG_KPARENT: b = &s->kparent; need = 4 * (2 * n - (n ? 1 : 0)); break;
        case EOS_DBG_IBOX: b = &s->ibox; need = 24 * m; break;
        case EOS_DBG_NODES: b = &s->nodes; need = 64 * m; break;
        case EOS_DBG_TRIS: b = &s->geo; need = 48 * n; break;
        case EOS_DBG_MORTON: b = &s->codes; need = 4 * n; break;
        case EOS_DBG_NODES4: b = &s->nodes4; need = 128 * (size_t)s->n_nodes4; break;
        case EOS_DBG_XFORMS: b = &s->xf; need = 16 * kXfRows * (size_t)s->n_xf; break;
        default: return fail(EOS_ERR_INVALID, "eos_debug_read: unknown array");
    }
    if (!s->built) return fail(EOS_ERR_STATE, "eos_debug_read: no LBVH built");
    if (s->batch_built && (what == EOS_DBG_KCHILD || what == EOS_DBG_KPARENT || what == EOS_DBG_IBOX))
        return fail(EOS_ERR_STATE, "eos_debug_read: a batched build keeps no per-scene Karras arrays or boxes");
    if ((size_t)bytes < need) return fail(EOS_ERR_INVALID, "eos_debug_read: destination too small");
    if (need == 0) return EOS_OK;
    EOS_HIP(hipSetDevice(s->device));
    EOS_HIP(hipMemcpyAsync(dst, b->ptr, need, hipMemcpyDeviceToHost, s->stream));
    EOS_HIP(hipStreamSynchronize(s->stream));
    return EOS_OK;
}

}  // extern "C"
