/*
 * oracle.h -- CPU restatement of the Eos rendering hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP path is
 * compared against; it is loaded by tests/, by __graft_entry__.smoke() and by
 * bench.py's cpu_baseline leg, and never by the product library.
 *
 * What it restates (reference files, relative to the reference root):
 *   camera rays ............ src/RayTracer.cpp:116-146
 *   Moller-Trumbore ........ src/Triangle.cpp:57-91, hit rule :26-55
 *   nearest hit / any hit .. src/BVH.cpp:609-716 (semantics; own acceleration)
 *   shading ................ src/RayTracer.cpp:302-458 (shadow rays, Phong,
 *                            reflection recursion, no ambient)
 *   quantisation ........... src/RayTracer.cpp:150-153
 *   glm op order ........... glm/detail/func_geometric.inl:40-75,134-173
 *   Morton interleave ...... src/morton.h:30-48 (Part1By2 / EncodeMorton3)
 * and, for the parts the reference does not have:
 *   Karras 2012 LBVH, the 64-byte node format and the GPU traversal order
 *   (restated so node-visit counts can be recomputed on the CPU).
 */
#ifndef EOS_ORACLE_H
#define EOS_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { float position[3], view[3], right[3], up[3]; int32_t width, height; } orc_camera;
typedef struct {
    float kd[4], ks[4];
    float shininess, reflectivity;
    int32_t reflective, transparent;
    float refractive_index;
    int32_t pad[3];
} orc_material;
typedef struct { float position[4], color[4]; } orc_light;

typedef struct {
    int32_t n_tri;
    const float* verts;      /* [n_tri][3][3] */
    const float* normals;    /* [n_tri][3][3] */
    const int32_t* mat;      /* [n_tri] */
    int32_t n_mat;
    const orc_material* mats;
    int32_t n_light;
    const orc_light* lights;
    /* Per-triangle local-to-world transforms (the reference's Surface
     * transformation(); src/BVH.cpp:342-383 tests the ray in the triangle's
     * local space).  NULL / -1: none.  The inverse is glm's, the normal matrix
     * the transpose of the inverse (src/triangleMesh.cpp:246-250). */
    int32_t n_xform;
    const float* xforms;        /* [n_xform][16] column-major (glm layout) */
    const int32_t* xform_index; /* [n_tri], -1 = none; may be NULL */
    /* Scene::setAmbientRefractiveIndex (src/Scene.cpp:134-136): n1 or n2 of
     * the Fresnel term at transparent materials (src/RayTracer.cpp:372-384). */
    float ambient_index;
} orc_scene;

typedef struct {
    int64_t primary, shadow, secondary, hits;
    int64_t node_visits, tri_tests;          /* filled by the LBVH-counting render only */
} orc_counts;

/* One shading record per (level, pixel): what the forward pass saw at a hit. */
typedef struct {
    int32_t tri;          /* original triangle index, -1 = no hit at this level */
    float t, u, v;
    float p[3];           /* hit point */
    float n[3];           /* shading normal */
    int32_t mat;
    uint32_t vis;         /* bit i set = light i unshadowed */
    float refl;           /* reflectivity applied to the next level (0 = none) */
    float pad[3];
} orc_rec;                /* 64 bytes */

/* ---- exact nearest-hit acceleration (own median-split BVH or brute force) ---- */
void* orc_accel_build(const orc_scene* s, int brute_force);
void orc_accel_free(void* accel);
/* Nearest hit (any_hit = 0: triangle index or -1, t, u, v, hit point and
 * shading normal) or occlusion flag (any_hit = 1) for n arbitrary rays. */
void orc_nearest(const orc_scene* s, const void* accel, int32_t n, const float* orig, const float* dirs,
                 int32_t any_hit, int32_t* hit_tri, float* hit_t, float* hit_u, float* hit_v, float* point,
                 float* normal);

/* Scene::findMinDistanceIntersectionLinear (src/Scene.cpp:159-201) for n rays:
 * found flag (its return value), kept triangle (original index or -1), t, the
 * world-space point and the transformed (not renormalised) normal.  Outputs
 * other than found / hit_tri may be NULL. */
void orc_nearest_linear(const orc_scene* s, int32_t n, const float* orig, const float* dirs, int32_t* found,
                        int32_t* hit_tri, float* hit_t, float* point, float* normal);

/* Reference-semantics render of rows row_start, row_start+row_step, ...
 * Any output pointer may be NULL.  rgba: [H*W*4]; rgb: [H*W*3] pre-quantisation;
 * hit_ids: [H*W] primary hit (original index or -1); recs: [(depth+1)*H*W]. */
int orc_render(const orc_scene* s, const void* accel, const orc_camera* cam, int depth,
               int row_start, int row_step, uint8_t* rgba, float* rgb, int32_t* hit_ids,
               orc_rec* recs, orc_counts* counts);
/* src/RayTracer.cpp:176-266: n x n samples per pixel (zero jitter), summed in
 * (q, p) order, scaled by (1/n)^2, quantised.  rgb optional. */
int orc_render_aa(const orc_scene* s, const void* accel, const orc_camera* cam, int depth, int n, uint8_t* rgba,
                  float* rgb, orc_counts* counts);

/* glm 0.9.6 4x4 inverse (glm/detail/type_mat4x4.inl:37-96), column-major. */
void orc_mat4_inverse(const float* m, float* out);

/* Primitive-level entry points for the closed-form tests. */
int orc_mt(const float* o, const float* d, const float* tri9, float* tuv);
void orc_camera_ray(const orc_camera* cam, int x, int y, float* dir3);
/* Sample (p, q) of n x n in pixel (x, y), renderWithAA's formula (src/RayTracer.cpp:222-230). */
void orc_camera_ray_aa(const orc_camera* cam, int x, int y, int p, int q, int n, float* dir3);
uint8_t orc_quantize(float c);
float orc_shade_phong(const orc_scene* s, const float* cam_pos, int light, int mat,
                      const float* p, const float* n, float* out3);
uint32_t orc_morton3(uint32_t x, uint32_t y, uint32_t z);

/* ---- LBVH restatement (Karras 2012) ---- */
/* codes[n] (unsorted Morton codes); scene centroid bounds -> cbounds[6]. */
void orc_morton_codes(int32_t n, const float* verts, uint32_t* codes, float* cbounds);
/* Full build.  Outputs (any may be NULL):
 *   sorted_codes[n], perm[n] (sorted slot -> original index),
 *   kchild[2*(n-1)]  Karras children: >=0 internal index, <0 = ~leaf index,
 *   kparent[2n-1]    parent of internal i at [i], of leaf s at [n-1+s] (-1 root),
 *   ibox[(n-1)*6]    exact internal boxes lo.xyz, hi.xyz,
 *   nodes[(n-1)*16]  64-byte traversal nodes (float view; child refs as int bits),
 *   tris[n*12]       sorted triangle records v1.xyz e1.x | e1.yz e2.xy | e2.z orig xform pad
 *                    (local-space vertices; xform = transformation index or -1),
 *   root_ref[1].
 * xforms / xform_index (may be NULL): per-triangle transformations; Morton keys
 * and boxes then come from the world-space vertices. */
void orc_lbvh_build(int32_t n, const float* verts, int32_t leaf_max,
                    uint32_t* sorted_codes, int32_t* perm, int32_t* kchild, int32_t* kparent,
                    float* ibox, float* nodes, float* tris, int32_t* root_ref,
                    float* nodes4, int32_t* n_nodes4, int32_t* root4,
                    const float* xforms, const int32_t* xform_index);
/* nodes4: 128-byte 4-wide nodes [k][32]: lo.x[4] hi.x[4] lo.y[4] hi.y[4]
 * lo.z[4] hi.z[4] ref[4] pad[4]; empty slots have ref 0x7FFFFFFE and boxes at
 * +inf.  Pass nodes4 = NULL with n_nodes4 to get the count first. */
/* Karras hierarchy on an already sorted code array. */
void orc_karras(int32_t n, const uint32_t* sorted_codes, int32_t* kchild, int32_t* kparent,
                int32_t* first, int32_t* last);
/* The GPU traversal, restated: closest (any_hit = 0) or any hit (1). tmax = +inf.
 * hit_orig[n_rays] (-1 miss), t/u/v may be NULL; visits/tests accumulate totals. */
void orc_lbvh_trace(const float* nodes, int32_t root_ref, const float* tris, int32_t n_rays,
                    const float* orig, const float* dirs, int32_t any_hit, int32_t* hit_orig,
                    float* hit_t, float* hit_u, float* hit_v, int64_t* visits, int64_t* tests,
                    int32_t width, int32_t n_xform, const float* xforms);
/* Primary rays of a frame walked as 8x8-pixel packets over the 4-wide tree, as
 * the GPU's packet kernel does: one shared walk per tile, every child some
 * valid lane hits is visited, children ordered by (entry distance of the
 * lowest-index lane hitting them, slot).  Per pixel [H][W]: hit_orig (-1 miss),
 * t/u/v (may be NULL); visits/tests accumulate per-tile totals. */
void orc_lbvh4_packet_primary(const float* nodes4, int32_t root4, const float* tris, const orc_camera* cam,
                              int32_t* hit_orig, float* hit_t, float* hit_u, float* hit_v, int64_t* visits,
                              int64_t* tests, int32_t n_xform, const float* xforms);
/* Whitted render driven by the restated GPU traversal; fills rgba/rgb/hit_ids like
 * orc_render and counts node visits / triangle tests in counts. */
int orc_render_lbvh(const orc_scene* s, const float* nodes, int32_t root_ref, const float* tris,
                    const orc_camera* cam, int depth, int row_start, int row_step, uint8_t* rgba,
                    float* rgb, int32_t* hit_ids, orc_counts* counts, int32_t width);

/* ---- FP64 shading on fixed hit sets (gradient checks) ---- */
/* Parameter vector layout: per material [kd r,g,b, ks r,g,b, shininess] (7*M), then
 * per light [pos x,y,z, color r,g,b] (6*L).  recs/cam_pos as produced by orc_render. */
double orc_loss64(const orc_rec* recs, int32_t levels, int64_t n_pix, const float* cam_pos,
                  int32_t n_mat, int32_t n_light, const double* params, const float* target_rgb,
                  double* image_out);
void orc_grad_fd64(const orc_rec* recs, int32_t levels, int64_t n_pix, const float* cam_pos,
                   int32_t n_mat, int32_t n_light, const double* params, const float* target_rgb,
                   double h, const int32_t* which, int32_t n_which, double* grad_out);
void orc_grad_fwd64(const orc_rec* recs, int32_t levels, int64_t n_pix, const float* cam_pos,
                    int32_t n_mat, int32_t n_light, const double* params, const float* target_rgb,
                    double* grad_out);

#ifdef __cplusplus
}
#endif
#endif
