/*
 * refport.h -- TEST INFRASTRUCTURE ONLY: a timing-faithful C++ port of the
 * reference's CPU render path, used as the CPU baseline (bench.py's
 * cpu_baseline leg) and checked against the reference's own output.
 *
 * It keeps the reference's algorithmic costs:
 *   BVH: hybrid build, spatial-median split above 3000 surfaces, SAH below,
 *        depth cap 18, leaves of <= 6 (src/BVH.cpp:549-606, :161-219, :413-498),
 *        boxes that contain the local origin (src/Box.h:36, src/Triangle.h:35-41),
 *        flattened in preorder (src/BVH.cpp:742-776);
 *   traversal: stack walk with no closest-t pruning, ordered by the ray's sign
 *        on the split axis, slab test over [0.001, 999] (src/BVH.cpp:609-716,
 *        src/Box.cu:228-256);
 *   per triangle: ray moved to local space by the inverse mat4, hit moved back
 *        by the mat4 and the inverse-transpose (src/BVH.cpp:342-383), with the
 *        scene's transformations (identity where a triangle has none); world
 *        centroids and two-corner world boxes for the build (src/BVH.cpp:36-71,
 *        src/Box.cu:84-137);
 *   shading and quantisation as src/RayTracer.cpp:99-174,268-458.
 * Deviations: the SAH sweep uses prefix/suffix box unions (same costs and the
 * same first-minimum split as the reference run with one OpenMP thread, in
 * O(n) instead of O(n^2) per node); leaves over 6 surfaces (depth cap) store
 * a range instead of overflowing a 6-slot array; the any-hit test reads the
 * reference's uninitialised distance as +infinity.
 */
#ifndef EOS_REFPORT_H
#define EOS_REFPORT_H
#include <stdint.h>

#include "oracle.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int64_t primary, shadow, secondary, hits;
    double seconds;       /* CLOCK_MONOTONIC_RAW around the render loop only */
    int32_t threads;
    int32_t pad;
} rp_counts;

/* Builds the reference BVH (single-threaded, deterministic).  build_seconds
 * receives the build time. */
void* rp_build(const orc_scene* s, double* build_seconds);
void rp_free(void* h);
int32_t rp_num_nodes(const void* h);

/* Renders the listed pixels (y*W + x), writing RGBA8 per listed pixel (may be
 * NULL), with `threads` OpenMP threads (0 = all), schedule(dynamic, 2) over
 * chunks of `chunk` pixels (a row's worth). */
int rp_render(const void* h, const orc_camera* cam, int depth, const int32_t* pixels, int64_t n_pixels,
              int32_t chunk, uint8_t* rgba, int32_t threads, rp_counts* counts);

#ifdef __cplusplus
}
#endif
#endif
