// ref_harness.cpp -- TEST INFRASTRUCTURE ONLY.
//
// Drives the reference's own CPU renderer through its public C++ API
// (Scene::addSurface/addLightSource/addMaterial/flush, Camera setters,
// RayTracer::setTracedDepth/render; src/Scene.h:46-86, src/Camera.h:40-45,
// src/RayTracer.h:43-51) so the oracle's restatement can be checked against
// the real thing.  Compiled together with the reference's translation units,
// read in place from the reference tree by oracle/ref.mk; the result lives in
// oracle/_ref/ and is never part of the product.
#include <omp.h>

#include <cstdint>
#include <cstring>
#include <vector>

#include "Camera.h"
#include "LightSource.h"
#include "Material.h"
#include "RayIntersection.h"
#include "RayTracer.h"
#include "Scene.h"
#include "Triangle.h"

#define REF_API extern "C" __attribute__((visibility("default")))

namespace {
struct RefScene {
    Scene scene;
    std::vector<Material> mats;       // stable storage Triangle::setMaterial points into
    glm::mat4 ident{1.0f}, inv{1.0f}, invT{1.0f};
    // per-transformation storage the triangles point into (sized once, never moved)
    std::vector<glm::mat4> xf, xf_inv, xf_invT;
    std::vector<Triangle*> tris;
    std::vector<LightSource*> lights;
};
}  // namespace

static glm::mat4 load_mat4(const float* m) {
    glm::mat4 g;
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r) g[c][r] = m[4 * c + r];
    return g;
}

// glm::inverse of a column-major 4x4 (the reference's own glm).
REF_API void ref_mat4_inverse(const float* m, float* out) {
    const glm::mat4 inv = glm::inverse(load_mat4(m));
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r) out[4 * c + r] = inv[c][r];
}

// Build a reference Scene.  Materials: kd[4] ks[4] shininess reflectivity
// reflective transparent refractive_index pad[3] (64 bytes each); lights:
// position[4] color[4]; the ambient refractive index is 1.  Optional
// per-triangle transformations: xforms [n_xform][16] column-major,
// xform_index [n_tri] (-1 = identity), each given to its triangles the way
// TriangleMesh::setTransformation does (src/triangleMesh.cpp:246-250: the
// inverse, and the transpose of the inverse).
// use_bvh = 0: Scene::useBvh(false) and no flush, so RayTracer and the
// queries take the linear search (src/RayTracer.cpp:281-285,334-337).
REF_API void* ref_scene_create_opts(int32_t n_tri, const float* verts, const float* normals,
                                    const int32_t* mat, int32_t n_mat, const void* mats_raw,
                                    int32_t n_light, const void* lights_raw, int32_t build_threads,
                                    int32_t n_xform, const float* xforms, const int32_t* xform_index,
                                    int32_t use_bvh) {
    RefScene* R = new RefScene();
    R->xf.resize(n_xform > 0 ? n_xform : 0);
    R->xf_inv.resize(R->xf.size());
    R->xf_invT.resize(R->xf.size());
    for (int i = 0; i < n_xform; ++i) {
        R->xf[i] = load_mat4(xforms + 16 * (size_t)i);
        R->xf_inv[i] = glm::inverse(R->xf[i]);
        R->xf_invT[i] = glm::transpose(R->xf_inv[i]);
    }
    R->scene.setMaxTracedDepth(5);
    R->scene.setAmbientRefractiveIndex(1.0f);
    R->scene.useBvh(use_bvh != 0);
    R->scene.setAASamples(1);
    const char* mb = static_cast<const char*>(mats_raw);
    for (int i = 0; i < n_mat; ++i) {
        const float* f = reinterpret_cast<const float*>(mb + 64 * i);
        int32_t refl, transp;
        float eta;
        std::memcpy(&refl, mb + 64 * i + 40, 4);
        std::memcpy(&transp, mb + 64 * i + 44, 4);
        std::memcpy(&eta, mb + 64 * i + 48, 4);
        Material m(0.0f, glm::vec4(f[0], f[1], f[2], f[3]), glm::vec4(f[4], f[5], f[6], f[7]), (int)f[8]);
        m.setReflective(refl != 0);
        m.setReflectionIntensity(f[9]);
        m.setTransparent(transp != 0);
        m.setRefractiveIndex(eta);
        R->mats.push_back(m);
        R->scene.addMaterial(m);
    }
    for (int i = 0; i < n_tri; ++i) {
        const float* v = verts + 9 * (size_t)i;
        const float* n = normals + 9 * (size_t)i;
        Triangle* t = new Triangle(glm::vec3(v[0], v[1], v[2]), glm::vec3(v[3], v[4], v[5]),
                                   glm::vec3(v[6], v[7], v[8]), glm::vec3(n[0], n[1], n[2]),
                                   glm::vec3(n[3], n[4], n[5]), glm::vec3(n[6], n[7], n[8]));
        const int xi = xform_index ? xform_index[i] : -1;
        if (xi >= 0) {
            t->setTransformation(R->xf[xi]);
            t->setInverseTransformation(R->xf_inv[xi]);
            t->setInverseTransposeTransformation(R->xf_invT[xi]);
        } else {
            t->setTransformation(R->ident);
            t->setInverseTransformation(R->inv);
            t->setInverseTransposeTransformation(R->invT);
        }
        t->setMaterial(R->mats[mat[i]]);
        t->setMaterialIndex(mat[i]);
        t->setTransformationIndex(xi >= 0 ? xi : 0);
        R->tris.push_back(t);
        R->scene.addSurface(t);
    }
    const char* lb = static_cast<const char*>(lights_raw);
    for (int i = 0; i < n_light; ++i) {
        const float* f = reinterpret_cast<const float*>(lb + 32 * i);
        LightSource* l = new LightSource(glm::vec4(f[0], f[1], f[2], f[3]), glm::vec4(f[4], f[5], f[6], f[7]));
        R->lights.push_back(l);
        R->scene.addLightSource(l);
    }
    if (use_bvh) {
        int saved = omp_get_max_threads();
        if (build_threads > 0) omp_set_num_threads(build_threads);  // the SAH sweep races (no determinism otherwise)
        R->scene.flush();
        omp_set_num_threads(saved);
    }
    return R;
}

REF_API void* ref_scene_create_xf(int32_t n_tri, const float* verts, const float* normals,
                                  const int32_t* mat, int32_t n_mat, const void* mats_raw,
                                  int32_t n_light, const void* lights_raw, int32_t build_threads,
                                  int32_t n_xform, const float* xforms, const int32_t* xform_index) {
    return ref_scene_create_opts(n_tri, verts, normals, mat, n_mat, mats_raw, n_light, lights_raw, build_threads,
                                 n_xform, xforms, xform_index, 1);
}

REF_API void* ref_scene_create(int32_t n_tri, const float* verts, const float* normals, const int32_t* mat,
                               int32_t n_mat, const void* mats_raw, int32_t n_light, const void* lights_raw,
                               int32_t build_threads) {
    return ref_scene_create_xf(n_tri, verts, normals, mat, n_mat, mats_raw, n_light, lights_raw, build_threads, 0,
                               nullptr, nullptr);
}

// RayTracer::render into rgba[H*W*4].  Camera: position, view, right, up, W, H.
REF_API void ref_render(void* h, const float* cam, int32_t width, int32_t height, int32_t depth,
                        uint8_t* rgba) {
    RefScene* R = static_cast<RefScene*>(h);
    Camera c;
    c.setPosition(glm::vec3(cam[0], cam[1], cam[2]));
    c.setViewingDirection(glm::vec3(cam[3], cam[4], cam[5]));
    c.setRightVector(glm::vec3(cam[6], cam[7], cam[8]));
    c.setUpVector(glm::vec3(cam[9], cam[10], cam[11]));
    c.setWidth(width);
    c.setHeight(height);
    RayTracer rt;
    rt.setAASamples(RAYTRACER_NO_AA);
    rt.setTracedDepth(depth);
    rt.render(R->scene, c, rgba);
}

// Scene::setAmbientRefractiveIndex.
REF_API void ref_scene_set_ambient(void* h, float n) { static_cast<RefScene*>(h)->scene.setAmbientRefractiveIndex(n); }

// RayTracer::render with n x n anti-aliasing samples (renderWithAA).
REF_API void ref_render_aa(void* h, const float* cam, int32_t width, int32_t height, int32_t depth, int32_t n,
                           uint8_t* rgba) {
    RefScene* R = static_cast<RefScene*>(h);
    Camera c;
    c.setPosition(glm::vec3(cam[0], cam[1], cam[2]));
    c.setViewingDirection(glm::vec3(cam[3], cam[4], cam[5]));
    c.setRightVector(glm::vec3(cam[6], cam[7], cam[8]));
    c.setUpVector(glm::vec3(cam[9], cam[10], cam[11]));
    c.setWidth(width);
    c.setHeight(height);
    RayTracer rt;
    rt.setAASamples(n);
    rt.setTracedDepth(depth);
    rt.render(R->scene, c, rgba);
}

// Scene::findMinDistanceIntersectionBVH for n rays: hit flag, point, normal, material.
REF_API void ref_nearest(void* h, int32_t n, const float* orig, const float* dirs, int32_t* hit,
                         float* point, float* normal, int32_t* mat) {
    RefScene* R = static_cast<RefScene*>(h);
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < n; ++i) {
        Ray r;
        r.setOrigin(glm::vec3(orig[3 * i], orig[3 * i + 1], orig[3 * i + 2]));
        r.setDirection(glm::vec3(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]));
        RayIntersection x;
        bool got = R->scene.findMinDistanceIntersectionBVH(r, x);
        hit[i] = got ? 1 : 0;
        if (got) {
            for (int k = 0; k < 3; ++k) { point[3 * i + k] = x.getPoint()[k]; normal[3 * i + k] = x.getNormal()[k]; }
            mat[i] = x.getMaterialIndex();
        }
    }
}

// Scene::shadowRayVisibilityBVH for n rays.
REF_API void ref_occluded(void* h, int32_t n, const float* orig, const float* dirs, int32_t* occ) {
    RefScene* R = static_cast<RefScene*>(h);
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < n; ++i) {
        Ray r;
        r.setOrigin(glm::vec3(orig[3 * i], orig[3 * i + 1], orig[3 * i + 2]));
        r.setDirection(glm::vec3(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]));
        occ[i] = R->scene.shadowRayVisibilityBVH(r) ? 1 : 0;
    }
}

// Scene::findMinDistanceIntersectionLinear for n rays: found flag, point,
// normal and the diffuse colour of the material it returns (the linear search
// sets RayIntersection's material, not its material index).
REF_API void ref_nearest_linear(void* h, int32_t n, const float* orig, const float* dirs, int32_t* hit,
                                float* point, float* normal, float* kd) {
    RefScene* R = static_cast<RefScene*>(h);
#pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < n; ++i) {
        Ray r;
        r.setOrigin(glm::vec3(orig[3 * i], orig[3 * i + 1], orig[3 * i + 2]));
        r.setDirection(glm::vec3(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]));
        RayIntersection x;
        bool got = R->scene.findMinDistanceIntersectionLinear(r, x);
        hit[i] = got ? 1 : 0;
        if (got) {
            for (int k = 0; k < 3; ++k) { point[3 * i + k] = x.getPoint()[k]; normal[3 * i + k] = x.getNormal()[k]; }
            for (int k = 0; k < 4; ++k) kd[4 * i + k] = x.getMaterial().getDiffuseColor()[k];
        }
    }
}

REF_API void ref_scene_free(void* h) {
    RefScene* R = static_cast<RefScene*>(h);
    for (Triangle* t : R->tris) delete t;
    for (LightSource* l : R->lights) delete l;
    delete R;
}
