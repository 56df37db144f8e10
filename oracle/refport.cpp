// refport.cpp -- TEST INFRASTRUCTURE ONLY: timing-faithful port of the
// reference's CPU render path (see refport.h for what is kept and what is not).
// Built twice by oracle/Makefile: librefport.so with the reference's flags
// (-O3 -ffast-math -fopenmp, src/CMakeLists.txt:1-5,17) for timing, and
// librefport_parity.so (IEEE, no contraction) for the parity test against the
// reference's own images.
#include "refport.h"
#include "glm_mat4.h"

#include <time.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr int kSahCeil = 3000;     // SAH_SURFACE_CEIL, src/BVH.cpp:31
constexpr int kMaxDepth = 18;      // MAX_BVH_DEPTH, src/BVH.cpp:32
constexpr int kLeafSize = 6;       // SURFACES_PER_LEAF, src/BVH.h:32
constexpr float kCostTrav = 0.5f;  // src/BVH.cpp:33-34
constexpr float kCostIsect = 1.0f;

struct V3 { float x, y, z; };
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 mul(float s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline float dot(V3 a, V3 b) { float tx = a.x * b.x, ty = a.y * b.y, tz = a.z * b.z; return (tx + ty) + tz; }
inline V3 cross(V3 x, V3 y) { return {x.y * y.z - y.y * x.z, x.z * y.x - y.z * x.x, x.x * y.y - y.x * x.y}; }
inline V3 normalize(V3 x) { float s = 1.0f / std::sqrt(dot(x, x)); return {x.x * s, x.y * s, x.z * s}; }
inline V3 ld3(const float* p) { return {p[0], p[1], p[2]}; }

// glm mat4 (column-major) times vec4: (m[0]*v.x + m[1]*v.y) + (m[2]*v.z + m[3]*v.w)
struct M4 { float c[4][4]; };
inline void mulv(const M4& m, const float v[4], float out[4]) {
    for (int r = 0; r < 4; ++r)
        out[r] = (m.c[0][r] * v[0] + m.c[1][r] * v[1]) + (m.c[2][r] * v[2] + m.c[3][r] * v[3]);
}

struct RBox { float b[2][3]; };   // [0] = min, [1] = max

struct RRay {
    V3 o, d;
    float inv[3];
    int sign[3];
    void set(V3 origin, V3 dir) {   // src/Ray.h:45-52 (z inverse in double)
        o = origin;
        d = dir;
        inv[0] = 1.0f / dir.x;
        inv[1] = 1.0f / dir.y;
        inv[2] = (float)(1.0 / (double)dir.z);
        for (int k = 0; k < 3; ++k) sign[k] = inv[k] < 0.0f;
    }
};

// src/Box.cu:228-256
inline bool box_opt(const RBox& bx, const RRay& r, float t0, float t1) {
    float tmin = (bx.b[r.sign[0]][0] - r.o.x) * r.inv[0];
    float tmax = (bx.b[1 - r.sign[0]][0] - r.o.x) * r.inv[0];
    float tymin = (bx.b[r.sign[1]][1] - r.o.y) * r.inv[1];
    float tymax = (bx.b[1 - r.sign[1]][1] - r.o.y) * r.inv[1];
    if ((tmin > tymax) || (tymin > tmax)) return false;
    if (tymin > tmin) tmin = tymin;
    if (tymax < tmax) tmax = tymax;
    float tzmin = (bx.b[r.sign[2]][2] - r.o.z) * r.inv[2];
    float tzmax = (bx.b[1 - r.sign[2]][2] - r.o.z) * r.inv[2];
    if ((tmin > tzmax) || (tzmin > tmax)) return false;
    if (tzmin > tmin) tmin = tzmin;
    if (tzmax < tmax) tmax = tzmax;
    return (tmin < t1) && (tmax > t0);
}

struct RTri {
    V3 v1, v2, v3, n1, n2, n3;
    V3 centroid;
    RBox box;
    const M4* M;
    const M4* Minv;
    const M4* MinvT;
    int mat;
};

struct RNode {
    RBox box;
    int leaf;
    int left, right;
    int first, count;
    int split_axis;
};

struct Port {
    std::vector<RTri> tris;
    std::vector<int> surf;       // the reference's surface array (sorted in place by the build)
    std::vector<RNode> nodes;    // preorder
    M4 ident;
    // per-transformation M, glm::inverse(M) and its transpose, as the reference's
    // triangles point at them (src/triangleMesh.cpp:246-250); sized once
    std::vector<M4> xm, xinv, xinvT;
    const orc_scene* s;
};

inline float surface_area(const RBox& b) {   // src/Box.cu:151-158
    float w = b.b[1][0] - b.b[0][0], h = b.b[1][1] - b.b[0][1], l = b.b[1][2] - b.b[0][2];
    return 2.0f * (w * l + h * l + h * w);
}
inline void expand(RBox& a, const RBox& b) {   // src/Box.cu:47-62
    for (int k = 0; k < 3; ++k) {
        a.b[0][k] = std::min(a.b[0][k], b.b[0][k]);
        a.b[1][k] = std::max(a.b[1][k], b.b[1][k]);
    }
}
RBox box_of(const Port& P, const int* s, int a, int b) {
    RBox r = P.tris[s[a]].box;
    for (int i = a + 1; i < b; ++i) expand(r, P.tris[s[i]].box);
    return r;
}
inline float cen(const Port& P, int i, int axis) {
    const V3& c = P.tris[i].centroid;
    return axis == 0 ? c.x : (axis == 1 ? c.y : c.z);
}
void sort_axis(Port& P, int* first, int* last, int axis) {
    std::sort(first, last, [&](int a, int b) { return cen(P, a, axis) < cen(P, b, axis); });
}

// src/BVH.cpp:161-219
int split_median(Port& P, int start, int end, int& axis) {
    RBox cb = box_of(P, P.surf.data(), start, end);
    float dx = cb.b[1][0] - cb.b[0][0], dy = cb.b[1][1] - cb.b[0][1], dz = cb.b[1][2] - cb.b[0][2];
    int big = 0;
    if (dy > dx) big = 1;
    if (dz > dy && dz > dx) big = 2;
    float mid = (cb.b[1][big] + cb.b[0][big]) * 0.5f;
    sort_axis(P, P.surf.data() + start, P.surf.data() + end, big);
    axis = big;
    int i;
    for (i = start; i < end; ++i)
        if (cen(P, P.surf[i], big) > mid) break;
    int split = i;
    if (split == start || split == end) split = start + (end - start) / 2;
    return split;
}

// src/BVH.cpp:413-498 with the cost of every split computed from prefix and
// suffix box unions; first strict minimum in (axis, index) order.
int split_sah(Port& P, const RBox& parent, int start, int end, int& axis_out) {
    const int n = end - start;
    const float parent_area = surface_area(parent);
    float min_cost = FLT_MAX;
    int min_split = start + 1;
    int axis = -1;
    std::vector<int> tmp(n);
    std::vector<RBox> suffix(n);
    for (int dim = 0; dim < 3; ++dim) {
        std::memcpy(tmp.data(), P.surf.data() + start, sizeof(int) * n);
        sort_axis(P, tmp.data(), tmp.data() + n, dim);
        suffix[n - 1] = P.tris[tmp[n - 1]].box;
        for (int k = n - 2; k >= 0; --k) {
            suffix[k] = P.tris[tmp[k]].box;
            expand(suffix[k], suffix[k + 1]);
        }
        RBox left = P.tris[tmp[0]].box;
        for (int i = start; i < end - 1; ++i) {
            int nl = i + 1;
            if (nl - start > 1) expand(left, P.tris[tmp[nl - start - 1]].box);
            const RBox& right = suffix[nl - start];
            float la = surface_area(left), ra = surface_area(right);
            float cost = kCostTrav + (float)(nl - start) * kCostIsect * la / parent_area +
                         (float)(end - nl) * kCostIsect * ra / parent_area;
            if (cost < min_cost) {
                min_cost = cost;
                min_split = nl;
                axis = dim;
            }
        }
    }
    // the reference sorts by an unset axis when no split was evaluated (n == 1): Z
    int sa = axis < 0 ? 2 : axis;
    sort_axis(P, P.surf.data() + start, P.surf.data() + end, sa);
    axis_out = sa;
    return min_split;
}

// src/BVH.cpp:549-606 (hybrid, depth-capped), emitting nodes in preorder.
int build(Port& P, int start, int end, int depth) {
    int me = (int)P.nodes.size();
    P.nodes.push_back(RNode());
    RBox box = box_of(P, P.surf.data(), start, end);
    P.nodes[me].box = box;
    P.nodes[me].first = start;
    P.nodes[me].count = end - start;
    P.nodes[me].leaf = 0;
    P.nodes[me].split_axis = 0;
    auto make_leaf = [&]() { P.nodes[me].leaf = 1; P.nodes[me].left = P.nodes[me].right = -1; };
    if (depth >= kMaxDepth) {
        make_leaf();
        return me;
    }
    int axis = 0, split;
    if (end - start > kSahCeil) {
        split = split_median(P, start, end, axis);
    } else {
        split = split_sah(P, box, start, end, axis);
        if (end - start <= kLeafSize) {
            P.nodes[me].split_axis = axis;
            make_leaf();
            return me;
        }
    }
    P.nodes[me].split_axis = axis;
    int l = build(P, start, split, depth + 1);
    int r = build(P, split, end, depth + 1);
    P.nodes[me].left = l;
    P.nodes[me].right = r;
    return me;
}

struct Isect { V3 p, n; int mat; };

// src/Triangle.cpp:26-91
inline bool tri_hit(const RTri& T, const RRay& r, Isect& x, float& distance) {
    V3 e1 = T.v2 - T.v1, e2 = T.v3 - T.v1;
    V3 Pv = cross(r.d, e2);
    float det = dot(e1, Pv);
    if (det > -0.00001f && det < 0.00001f) return false;
    float inv = 1.0f / det;
    V3 Tv = r.o - T.v1;
    V3 Q = cross(Tv, e1);
    float t = dot(e2, Q) * inv;
    if (t < 0.0f) return false;
    float u = dot(Tv, Pv) * inv;
    if (u < 0.0f || u > 1.0f) return false;
    float v = dot(r.d, Q) * inv;
    if (v < 0.0f || u + v > 1.0f) return false;
    if (t > distance) return false;
    distance = t;
    x.p = r.o + mul(distance, r.d);
    float w = (1.0f - u) - v;
    V3 a = {T.n1.x * w, T.n1.y * w, T.n1.z * w};
    V3 b = {T.n2.x * u, T.n2.y * u, T.n2.z * u};
    V3 c = {T.n3.x * v, T.n3.y * v, T.n3.z * v};
    x.n = normalize((a + b) + c);
    x.mat = T.mat;
    return true;
}

// src/BVH.cpp:342-383
bool leaf_hit(const Port& P, const RNode& nd, const RRay& ray, Isect& x, float& distance) {
    bool found = false;
    float wp[4] = {0, 0, 0, 0}, wn[4] = {0, 0, 0, 0};
    RRay local;
    for (int i = 0; i < nd.count; ++i) {
        const RTri& T = P.tris[P.surf[nd.first + i]];
        float o4[4] = {ray.o.x, ray.o.y, ray.o.z, 1.0f}, d4[4] = {ray.d.x, ray.d.y, ray.d.z, 0.0f}, lo[4], ld[4];
        mulv(*T.Minv, o4, lo);
        mulv(*T.Minv, d4, ld);
        local.set(V3{lo[0], lo[1], lo[2]}, V3{ld[0], ld[1], ld[2]});
        if (tri_hit(T, local, x, distance)) {
            found = true;
            float p4[4] = {x.p.x, x.p.y, x.p.z, 1.0f}, n4[4] = {x.n.x, x.n.y, x.n.z, 0.0f};
            mulv(*T.M, p4, wp);
            mulv(*T.MinvT, n4, wn);
        }
    }
    if (found) {
        x.p = V3{wp[0], wp[1], wp[2]};
        x.n = normalize(V3{wn[0], wn[1], wn[2]});
    }
    return found;
}

// src/BVH.cpp:609-659
bool nearest(const Port& P, const RRay& ray, Isect& x) {
    int stack[1024];
    int sp = 0;
    float dist = FLT_MAX;
    bool found = false;
    stack[sp++] = -1;
    int cur = 0;
    while (cur != -1) {
        const RNode& nd = P.nodes[cur];
        if (!nd.leaf) {
            if (box_opt(nd.box, ray, 0.001f, 999.0f)) {
                if (ray.sign[nd.split_axis]) { stack[sp++] = nd.left; cur = nd.right; }
                else { stack[sp++] = nd.right; cur = nd.left; }
            } else {
                cur = stack[--sp];
            }
        } else {
            if (leaf_hit(P, nd, ray, x, dist)) found = true;
            cur = stack[--sp];
        }
    }
    return found;
}

// src/BVH.cpp:661-716; the leaf's uninitialised distance read as +infinity.
bool occluded(const Port& P, const RRay& ray) {
    if (!box_opt(P.nodes[0].box, ray, 0.001f, 999.0f)) return false;
    int stack[1024];
    int sp = 0;
    stack[sp++] = -1;
    int cur = 0;
    while (cur != -1) {
        const RNode& nd = P.nodes[cur];
        if (!nd.leaf) {
            const RNode& l = P.nodes[nd.left];
            const RNode& r = P.nodes[nd.right];
            bool lh = box_opt(l.box, ray, 0.001f, 9999.0f);
            bool rh = box_opt(r.box, ray, 0.001f, 999.0f);
            if (lh) {
                cur = nd.left;
                if (rh) stack[sp++] = nd.right;
            } else if (rh) {
                cur = nd.right;
            } else {
                cur = stack[--sp];
            }
        } else {
            Isect dummy;
            float distance = INFINITY;
            if (leaf_hit(P, nd, ray, dummy, distance)) return true;
            cur = stack[--sp];
        }
    }
    return false;
}

struct Ctx {
    const Port* P;
    const orc_scene* s;
    V3 cam;
    int depth;
    int64_t shadow, secondary, hits;
};

// src/RayTracer.cpp:52-65
inline float fresnel(V3 incident, V3 normal, float n1, float n2) {
    const float n = n1 / n2;
    const float cosI = -dot(normal, incident);
    const float sinT2 = n * n * (1.0f - cosI * cosI);
    if (sinT2 > 1.0f) return 1.0f;
    const float cosT = std::sqrt(1.0f - sinT2);
    const float rOrth = (n1 * cosI - n2 * cosT) / (n1 * cosI + n2 * cosT);
    const float rPar = (n2 * cosI - n1 * cosT) / (n2 * cosI + n1 * cosT);
    return (rOrth * rOrth + rPar * rPar) / 2.0f;
}

// src/RayTracer.cpp:268-458
void trace(Ctx& c, const RRay& ray, int depth, float out[3]) {
    out[0] = out[1] = out[2] = 0.0f;
    if (depth > c.depth) return;
    Isect x;
    if (!nearest(*c.P, ray, x)) return;
    c.hits++;
    const orc_material& m = c.s->mats[x.mat];
    float phong[3] = {0, 0, 0};
    RRay tmp;
    for (int i = 0; i < c.s->n_light; ++i) {
        const orc_light& l = c.s->lights[i];
        V3 Ld = normalize(ld3(l.position) - x.p);
        tmp.set(x.p + mul(1e-3f, Ld), Ld);
        c.shadow++;
        if (occluded(*c.P, tmp)) continue;
        V3 V = normalize(c.cam - x.p);
        float dNL = dot(x.n, Ld);
        float s2 = 2.0f * dNL;
        V3 R = normalize(V3{s2 * x.n.x, s2 * x.n.y, s2 * x.n.z} - Ld);
        float dLN = std::max(0.0f, dot(Ld, x.n));
        float dVR = dot(V, R);
        for (int k = 0; k < 3; ++k) {
            float diff = (dLN * m.kd[k]) * l.color[k];
            float spec = 0.0f;
            if (dVR > 0.0f) spec = (std::pow(dVR, (float)(int)m.shininess) * l.color[k]) * m.ks[k];
            phong[k] = phong[k] + (diff + spec);
        }
    }
    float refl[3] = {0, 0, 0};
    if (m.reflective) {
        float dNI = dot(x.n, ray.d);
        V3 t = {(x.n.x * dNI) * 2.0f, (x.n.y * dNI) * 2.0f, (x.n.z * dNI) * 2.0f};
        V3 Rd = normalize(ray.d - t);
        tmp.set(x.p + mul(1e-3f, Rd), Rd);
        if (depth + 1 <= c.depth) c.secondary++;
        float sub[3];
        trace(c, tmp, depth + 1, sub);
        for (int k = 0; k < 3; ++k) refl[k] = m.reflectivity * sub[k];
    }
    float refr[3] = {0, 0, 0};
    if (m.transparent) {
        // src/RayTracer.cpp:361-408, tracing the reflected ray twice as :401-403 do
        V3 inc = normalize(ray.d);
        float n1 = m.refractive_index, n2 = c.s->ambient_index;
        V3 nn = x.n;
        if (dot(inc, x.n) < 0.0f) {
            n1 = c.s->ambient_index;
            n2 = m.refractive_index;
        } else {
            nn = V3{0.0f - x.n.x, 0.0f - x.n.y, 0.0f - x.n.z};
        }
        const float K = fresnel(inc, nn, n1, n2);
        float dNI = dot(nn, ray.d);
        V3 t = {(nn.x * dNI) * 2.0f, (nn.y * dNI) * 2.0f, (nn.z * dNI) * 2.0f};
        V3 Rd = normalize(ray.d - t);
        RRay fr;
        fr.set(x.p + mul(1e-3f, nn), Rd);
        if (depth + 1 <= c.depth) c.secondary += 2;
        float a[3], b[3];
        trace(c, fr, depth + 1, a);
        trace(c, fr, depth + 1, b);
        for (int k = 0; k < 3; ++k) refr[k] = 0.0f + (K * a[k] + (1.0f - K) * b[k]);
    }
    for (int k = 0; k < 3; ++k) out[k] = (phong[k] + refl[k]) + refr[k];
}

inline uint8_t quantize(float c) {
    double q = (c == 1.0f) ? 255.0 : std::min((double)c * 256.0, 255.0);
    return (uint8_t)std::floor(q);
}

double now() {
    timespec ts;
    clock_gettime(CLOCK_MONOTONIC_RAW, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

}  // namespace

extern "C" {

void* rp_build(const orc_scene* s, double* build_seconds) {
    double t0 = now();
    Port* P = new Port();
    P->s = s;
    std::memset(&P->ident, 0, sizeof P->ident);
    for (int k = 0; k < 4; ++k) P->ident.c[k][k] = 1.0f;
    const int nx = s->xform_index ? s->n_xform : 0;
    P->xm.resize(nx);
    P->xinv.resize(nx);
    P->xinvT.resize(nx);
    for (int x = 0; x < nx; ++x) {
        float inv[16];
        inverse_glm(s->xforms + 16 * (size_t)x, inv);
        for (int c = 0; c < 4; ++c)
            for (int r = 0; r < 4; ++r) {
                P->xm[x].c[c][r] = s->xforms[16 * (size_t)x + 4 * c + r];
                P->xinv[x].c[c][r] = inv[4 * c + r];
                P->xinvT[x].c[c][r] = inv[4 * r + c];
            }
    }
    P->tris.resize(s->n_tri);
    for (int i = 0; i < s->n_tri; ++i) {
        const float* v = s->verts + 9 * (size_t)i;
        const float* n = s->normals + 9 * (size_t)i;
        RTri& T = P->tris[i];
        T.v1 = ld3(v); T.v2 = ld3(v + 3); T.v3 = ld3(v + 6);
        T.n1 = ld3(n); T.n2 = ld3(n + 3); T.n3 = ld3(n + 6);
        V3 sum = (T.v1 + T.v2) + T.v3;
        T.centroid = {sum.x / 3.0f, sum.y / 3.0f, sum.z / 3.0f};
        for (int k = 0; k < 3; ++k) { T.box.b[0][k] = 0.0f; T.box.b[1][k] = 0.0f; }   // Box() at the origin
        for (const V3* p : {&T.v1, &T.v2, &T.v3}) {
            float c[3] = {p->x, p->y, p->z};
            for (int k = 0; k < 3; ++k) {
                T.box.b[0][k] = std::min(T.box.b[0][k], c[k]);
                T.box.b[1][k] = std::max(T.box.b[1][k], c[k]);
            }
        }
        T.M = T.Minv = T.MinvT = &P->ident;
        const int xi = nx > 0 ? s->xform_index[i] : -1;
        if (xi >= 0) {
            T.M = &P->xm[xi];
            T.Minv = &P->xinv[xi];
            T.MinvT = &P->xinvT[xi];
            // world centroid (the sort comparators' M * centroid, src/BVH.cpp:36-71) and world
            // box from the two corners of the local box (Box::transformBoundingBox, src/Box.cu:84-137)
            const float c4[4] = {T.centroid.x, T.centroid.y, T.centroid.z, 1.0f};
            const float lo4[4] = {T.box.b[0][0], T.box.b[0][1], T.box.b[0][2], 1.0f};
            const float hi4[4] = {T.box.b[1][0], T.box.b[1][1], T.box.b[1][2], 1.0f};
            float wc[4], wlo[4], whi[4];
            mulv(*T.M, c4, wc);
            mulv(*T.M, lo4, wlo);
            mulv(*T.M, hi4, whi);
            T.centroid = {wc[0], wc[1], wc[2]};
            for (int k = 0; k < 3; ++k) {
                T.box.b[0][k] = std::min(wlo[k], whi[k]);
                T.box.b[1][k] = std::max(whi[k], wlo[k]);
            }
        }
        T.mat = s->mat[i];
    }
    P->surf.resize(s->n_tri);
    for (int i = 0; i < s->n_tri; ++i) P->surf[i] = i;
    if (s->n_tri > 0) build(*P, 0, s->n_tri, 0);
    if (build_seconds) *build_seconds = now() - t0;
    return P;
}

void rp_free(void* h) { delete static_cast<Port*>(h); }

int32_t rp_num_nodes(const void* h) { return (int32_t)static_cast<const Port*>(h)->nodes.size(); }

int rp_render(const void* h, const orc_camera* cam, int depth, const int32_t* pixels, int64_t n_pixels,
              int32_t chunk, uint8_t* rgba, int32_t threads, rp_counts* counts) {
    const Port* P = static_cast<const Port*>(h);
    const int W = cam->width, H = cam->height;
    if (chunk < 1) chunk = W;
    const float nw = (float)W / 2.0f, nh = (float)H / 2.0f;
    const float inw = 1.0f / nw, inh = 1.0f / nh;
    const V3 view = ld3(cam->view), right = ld3(cam->right), up = ld3(cam->up), eye = ld3(cam->position);
    const int64_t nchunks = (n_pixels + chunk - 1) / chunk;
    int64_t shadow = 0, secondary = 0, hits = 0;
    int used = 1;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#endif
    const bool empty = P->nodes.empty();
    double t0 = now();
#pragma omp parallel reduction(+ : shadow, secondary, hits)
    {
#ifdef _OPENMP
#pragma omp single
        used = omp_get_num_threads();
#endif
        Ctx c{P, P->s, eye, depth, 0, 0, 0};
#pragma omp for schedule(dynamic, 2)
        for (int64_t ch = 0; ch < nchunks; ++ch) {
            for (int64_t k = ch * chunk; k < std::min<int64_t>(n_pixels, (ch + 1) * chunk); ++k) {
                int p = pixels[k];
                int x = p % W, y = p / W;
                float bb = ((float)y - nh) * inh;
                float aa = ((float)x - nw) * inw;
                V3 d = normalize((mul(aa, right) + mul(bb, up)) + view);
                RRay ray;
                ray.set(eye, d);
                float col[3] = {0, 0, 0};
                if (!empty) trace(c, ray, 0, col);
                if (rgba) {
                    for (int q = 0; q < 3; ++q) rgba[4 * k + q] = quantize(col[q]);
                    rgba[4 * k + 3] = 255;
                }
            }
        }
        shadow += c.shadow;
        secondary += c.secondary;
        hits += c.hits;
    }
    double t1 = now();
    if (counts) {
        counts->primary = n_pixels;
        counts->shadow = shadow;
        counts->secondary = secondary;
        counts->hits = hits;
        counts->seconds = t1 - t0;
        counts->threads = used;
        counts->pad = 0;
    }
    (void)H;
    return 0;
}

}  // extern "C"
