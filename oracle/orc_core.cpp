// orc_core.cpp -- TEST INFRASTRUCTURE ONLY: CPU restatement of the Eos render
// path, used as the checker for the HIP implementation (never linked into it).
//
// Build: see oracle/Makefile.  Must be compiled WITHOUT -ffast-math and with
// -ffp-contract=off so every float operation is the single IEEE operation the
// reference's glm code performs (x86-64 SSE arithmetic, no excess precision).
#include "oracle.h"
#include "glm_mat4.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// ------------------------------------------------------------ glm op order
// glm/detail/func_geometric.inl:52-58 dot(vec3) = (x*x' + y*y') + z*z'
// :134-142 cross; :154-158 normalize(x) = x * (1/sqrt(dot(x,x)));
// :170-173 reflect(I,N) = I - N*dot(N,I)*2 ; func_exponential.inl:148-160.
struct V3 { float x, y, z; };
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 mul(float s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline float dot(V3 a, V3 b) {
    float tx = a.x * b.x, ty = a.y * b.y, tz = a.z * b.z;
    return (tx + ty) + tz;
}
inline V3 cross(V3 x, V3 y) {
    return {x.y * y.z - y.y * x.z, x.z * y.x - y.z * x.x, x.x * y.y - y.x * x.y};
}
inline V3 normalize(V3 x) {
    float s = 1.0f / std::sqrt(dot(x, x));
    return {x.x * s, x.y * s, x.z * s};
}
inline V3 reflect(V3 I, V3 N) {
    float d = dot(N, I);
    V3 t = {(N.x * d) * 2.0f, (N.y * d) * 2.0f, (N.z * d) * 2.0f};
    return I - t;
}
inline V3 ld3(const float* p) { return {p[0], p[1], p[2]}; }

// ------------------------------------------------------------ Moller-Trumbore
// src/Triangle.cpp:57-91: e1=v2-v1, e2=v3-v1, P=cross(d,e2), det=dot(e1,P),
// reject |det|<1e-5 (absolute), inv=1/det, T=o-v1, Q=cross(T,e1),
// t=dot(e2,Q)*inv (t>=0), u=dot(T,P)*inv in [0,1], v=dot(d,Q)*inv, v>=0, u+v<=1.
inline bool mt_e(V3 o, V3 d, V3 v1, V3 e1, V3 e2, float& t, float& u, float& v) {
    V3 P = cross(d, e2);
    float det = dot(e1, P);
    if (det > -0.00001f && det < 0.00001f) return false;
    float inv = 1.0f / det;
    V3 T = o - v1;
    V3 Q = cross(T, e1);
    t = dot(e2, Q) * inv;
    if (t < 0.0f) return false;
    u = dot(T, P) * inv;
    if (u < 0.0f || u > 1.0f) return false;
    v = dot(d, Q) * inv;
    if (v < 0.0f || u + v > 1.0f) return false;
    return true;
}
// glm mat4 * vec4 (glm/detail/type_mat4x4.inl:571-616): per row r,
// (m[0][r]*x + m[1][r]*y) + (m[2][r]*z + m[3][r]*w); column-major m[4c + r].
inline V3 xform(const float* m, V3 p, float w) {
    float r[3];
    for (int k = 0; k < 3; ++k) r[k] = (m[k] * p.x + m[4 + k] * p.y) + (m[8 + k] * p.z + m[12 + k] * w);
    return {r[0], r[1], r[2]};
}

inline void transpose4(const float* a, float* out) {
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r) out[4 * c + r] = a[4 * r + c];
}

inline bool mt(V3 o, V3 d, const float* tri9, float& t, float& u, float& v) {
    V3 v1 = ld3(tri9), v2 = ld3(tri9 + 3), v3 = ld3(tri9 + 6);
    return mt_e(o, d, v1, v2 - v1, v3 - v1, t, u, v);
}

// src/RayTracer.cpp:150-153: floor(c==1 ? 255 : min(c*256.0, 255.0)) in double.
// One triangle against a world-space ray: in the triangle's local space when it
// has a transformation (src/BVH.cpp:359-368: origin and direction through the
// inverse, w = 1 and 0).
inline int xform_of(const orc_scene& s, int tri) { return s.xform_index ? s.xform_index[tri] : -1; }
inline bool tri_hit(const orc_scene& s, int tri, V3 o, V3 d, float& t, float& u, float& v) {
    const int xi = xform_of(s, tri);
    if (xi < 0) return mt(o, d, s.verts + 9 * (size_t)tri, t, u, v);
    float inv[16];
    inverse_glm(s.xforms + 16 * (size_t)xi, inv);
    return mt(xform(inv, o, 1.0f), xform(inv, d, 0.0f), s.verts + 9 * (size_t)tri, t, u, v);
}

inline uint8_t quantize(float c) {
    double q = (c == 1.0f) ? 255.0 : std::min((double)c * 256.0, 255.0);
    return (uint8_t)std::floor(q);
}

// src/RayTracer.cpp:119-142 (viewWidth/2.0f, 1/that, aa, bb, normalize).
inline V3 camera_dir(const orc_camera& c, int x, int y) {
    float nw = (float)c.width / 2.0f, nh = (float)c.height / 2.0f;
    float inw = 1.0f / nw, inh = 1.0f / nh;
    float bb = ((float)y - nh) * inh;
    float aa = ((float)x - nw) * inw;
    V3 r = ld3(c.right), u = ld3(c.up), v = ld3(c.view);
    V3 d = (mul(aa, r) + mul(bb, u)) + v;
    return normalize(d);
}

// src/RayTracer.cpp:196-230: sample (p, q) of an n x n grid in pixel (x, y),
// aa = (x + (p + ksi) * (1/n) - W/2) * (2/W) and the same for y with q.  The
// jitter ksi = rand() / RAND_MAX is an integer division (:222), 0 unless
// rand() returns RAND_MAX; it is taken as 0.  1/n is a double division
// rounded to float (`float inv = 1.0 / (float) n`).
inline V3 camera_dir_aa(const orc_camera& c, int x, int y, int p, int q, float inv_n) {
    float nw = (float)c.width / 2.0f, nh = (float)c.height / 2.0f;
    float inw = 1.0f / nw, inh = 1.0f / nh;
    const float ksi = 0.0f;
    float bb = (((float)y + (((float)q + ksi) * inv_n)) - nh) * inh;
    float aa = (((float)x + (((float)p + ksi) * inv_n)) - nw) * inw;
    V3 r = ld3(c.right), u = ld3(c.up), v = ld3(c.view);
    V3 d = (mul(aa, r) + mul(bb, u)) + v;
    return normalize(d);
}

// ------------------------------------------------------------ exact accel
// A median-split BVH over double-precision boxes inflated far beyond any
// rounding of the float Moller-Trumbore test; traversal only prunes with
// t-intervals widened the same way, so the nearest hit it returns is the
// exact arg-min over all triangles (ties -> smallest original index).
struct ANode { double lo[3], hi[3]; int left, right, first, count; };
struct Accel {
    bool brute = false;
    std::vector<ANode> nodes;
    std::vector<int> idx;
};

void abuild(Accel& A, const float* verts, std::vector<double>& cen, int node, int first, int count) {
    ANode nd;
    double lo[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    double clo[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, chi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    for (int i = first; i < first + count; ++i) {
        int t = A.idx[i];
        for (int k = 0; k < 9; ++k) {
            double x = verts[9 * (size_t)t + k];
            lo[k % 3] = std::min(lo[k % 3], x);
            hi[k % 3] = std::max(hi[k % 3], x);
        }
        for (int k = 0; k < 3; ++k) {
            clo[k] = std::min(clo[k], cen[3 * (size_t)t + k]);
            chi[k] = std::max(chi[k], cen[3 * (size_t)t + k]);
        }
    }
    for (int k = 0; k < 3; ++k) {
        double pad = 1e-4 * (std::fabs(lo[k]) + std::fabs(hi[k]) + (hi[k] - lo[k])) + 1e-9;
        nd.lo[k] = lo[k] - pad;
        nd.hi[k] = hi[k] + pad;
    }
    nd.first = first;
    nd.count = count;
    nd.left = nd.right = -1;
    A.nodes[node] = nd;
    if (count <= 4) return;
    int ax = 0;
    double ext = chi[0] - clo[0];
    for (int k = 1; k < 3; ++k)
        if (chi[k] - clo[k] > ext) { ext = chi[k] - clo[k]; ax = k; }
    int mid = first + count / 2;
    std::nth_element(A.idx.begin() + first, A.idx.begin() + mid, A.idx.begin() + first + count,
                     [&](int a, int b) {
                         double ca = cen[3 * (size_t)a + ax], cb = cen[3 * (size_t)b + ax];
                         return ca < cb || (ca == cb && a < b);
                     });
    int l = (int)A.nodes.size();
    A.nodes.push_back(ANode());
    A.nodes.push_back(ANode());
    A.nodes[node].left = l;
    A.nodes[node].right = l + 1;
    abuild(A, verts, cen, l, first, mid - first);
    abuild(A, verts, cen, l + 1, mid, first + count - mid);
}

inline bool abox(const ANode& n, const double o[3], const double inv[3], double tmax) {
    double t0 = 0.0, t1 = tmax;
    for (int k = 0; k < 3; ++k) {
        double a = (n.lo[k] - o[k]) * inv[k], b = (n.hi[k] - o[k]) * inv[k];
        if (a > b) std::swap(a, b);
        t0 = std::max(t0, a);
        t1 = std::min(t1, b);
    }
    return t0 <= t1;
}

struct Hit { int tri = -1; float t = INFINITY, u = 0, v = 0; };

// Nearest hit: arg-min t over all MT hits, ties to the smaller triangle index.
void nearest(const Accel& A, const orc_scene& s, V3 o, V3 d, Hit& h) {
    auto test = [&](int tri) {
        float t, u, v;
        if (tri_hit(s, tri, o, d, t, u, v)) {
            if (t < h.t || (t == h.t && tri < h.tri)) { h.tri = tri; h.t = t; h.u = u; h.v = v; }
        }
    };
    if (A.brute) {
        for (int i = 0; i < s.n_tri; ++i) test(i);
        return;
    }
    if (A.nodes.empty()) return;
    double od[3] = {o.x, o.y, o.z}, inv[3];
    double dd[3] = {d.x, d.y, d.z};
    for (int k = 0; k < 3; ++k) inv[k] = 1.0 / (dd[k] != 0.0 ? dd[k] : 1e-300);
    int stack[256];
    int sp = 0;
    stack[sp++] = 0;
    while (sp) {
        const ANode& n = A.nodes[stack[--sp]];
        double tm = std::isinf(h.t) ? 1e300 : (double)h.t * (1.0 + 1e-6) + 1e-9;
        if (!abox(n, od, inv, tm)) continue;
        if (n.left < 0) {
            for (int i = n.first; i < n.first + n.count; ++i) test(A.idx[i]);
        } else {
            stack[sp++] = n.right;
            stack[sp++] = n.left;
        }
    }
}

// Any hit with t >= 0, no upper bound (src/BVH.cpp:661-716 with the shadow
// ray's uninitialised distance read as +infinity, src/RayTracer.cpp:326-336).
bool anyhit(const Accel& A, const orc_scene& s, V3 o, V3 d) {
    float t, u, v;
    if (A.brute) {
        for (int i = 0; i < s.n_tri; ++i)
            if (tri_hit(s, i, o, d, t, u, v)) return true;
        return false;
    }
    if (A.nodes.empty()) return false;
    double od[3] = {o.x, o.y, o.z}, inv[3];
    double dd[3] = {d.x, d.y, d.z};
    for (int k = 0; k < 3; ++k) inv[k] = 1.0 / (dd[k] != 0.0 ? dd[k] : 1e-300);
    int stack[256];
    int sp = 0;
    stack[sp++] = 0;
    while (sp) {
        const ANode& n = A.nodes[stack[--sp]];
        if (!abox(n, od, inv, 1e300)) continue;
        if (n.left < 0) {
            for (int i = n.first; i < n.first + n.count; ++i)
                if (tri_hit(s, A.idx[i], o, d, t, u, v)) return true;
        } else {
            stack[sp++] = n.right;
            stack[sp++] = n.left;
        }
    }
    return false;
}

// ------------------------------------------------------------ shading
// src/Triangle.cpp:41-47 + src/BVH.cpp:371-381: point o + t*d and normal
// normalize(n1*(1-u-v) + n2*u + n3*v) in the triangle's space, then the point
// through M (w = 1) and the normal through the transpose of the inverse (w = 0),
// normalised again.  Without a transformation these are the identity (the
// point and the first normalisation pass through unchanged).
inline void hit_geometry(const orc_scene& s, int tri, V3 o, V3 d, float t, float u, float v,
                         V3& P, V3& N, bool renorm = true) {
    const int xi = xform_of(s, tri);
    float inv[16] = {}, invT[16] = {};
    if (xi >= 0) {
        inverse_glm(s.xforms + 16 * (size_t)xi, inv);
        transpose4(inv, invT);
        o = xform(inv, o, 1.0f);
        d = xform(inv, d, 0.0f);
    }
    P = o + mul(t, d);
    const float* n = s.normals + 9 * (size_t)tri;
    float w = (1.0f - u) - v;
    V3 n1 = ld3(n), n2 = ld3(n + 3), n3 = ld3(n + 6);
    V3 a = {n1.x * w, n1.y * w, n1.z * w};
    V3 b = {n2.x * u, n2.y * u, n2.z * u};
    V3 c = {n3.x * v, n3.y * v, n3.z * v};
    N = normalize((a + b) + c);
    if (xi >= 0) {
        P = xform(s.xforms + 16 * (size_t)xi, P, 1.0f);
        N = xform(invT, N, 0.0f);
    }
    if (renorm) N = normalize(N);
}

// World-space vertices of the triangles (through M, w = 1) for bounding
// boxes and build keys; the input array when there are no transformations.
std::vector<float> world_verts(const orc_scene& s) {
    std::vector<float> w(s.verts, s.verts + 9 * (size_t)s.n_tri);
    if (!s.xform_index) return w;
    for (int i = 0; i < s.n_tri; ++i) {
        const int xi = s.xform_index[i];
        if (xi < 0) continue;
        for (int k = 0; k < 3; ++k) {
            V3 p = xform(s.xforms + 16 * (size_t)xi, ld3(s.verts + 9 * (size_t)i + 3 * k), 1.0f);
            w[9 * (size_t)i + 3 * k] = p.x;
            w[9 * (size_t)i + 3 * k + 1] = p.y;
            w[9 * (size_t)i + 3 * k + 2] = p.z;
        }
    }
    return w;
}

// src/RayTracer.cpp:418-458 for one light (the light's w is 1, so the vec4
// normalisations reduce to the vec3 ones bit for bit).
inline void phong(const orc_material& m, const orc_light& l, V3 cam, V3 P, V3 N, float out[3]) {
    V3 L = normalize(ld3(l.position) - P);
    V3 V = normalize(cam - P);
    float dNL = dot(N, L);
    float s2 = 2.0f * dNL;
    V3 R = normalize(V3{s2 * N.x, s2 * N.y, s2 * N.z} - L);
    float dLN = dot(L, N);
    dLN = std::max(0.0f, dLN);
    float dVR = dot(V, R);
    for (int k = 0; k < 3; ++k) {
        float diff = (dLN * m.kd[k]) * l.color[k];
        float spec = 0.0f;
        if (dVR > 0.0f) {
            float st = std::pow(dVR, m.shininess);
            spec = (st * l.color[k]) * m.ks[k];
        }
        out[k] = diff + spec;
    }
}

template <class NearestF, class AnyF>
void shade_path(const orc_scene& s, const orc_camera& cam, int depth, V3 o, V3 d, NearestF&& nearest_f,
                AnyF&& any_f, float rgb[3], int32_t* prim_hit, orc_rec* recs, int64_t stride,
                orc_counts& cnt) {
    const int L = s.n_light;
    float Pd[16][3];
    float Rd[16];
    int levels = 0;
    V3 camp = ld3(cam.position);
    for (int lev = 0; lev <= depth && lev < 16; ++lev) {
        Hit h;
        nearest_f(o, d, h);
        if (lev == 0 && prim_hit) *prim_hit = h.tri;
        if (h.tri < 0) break;
        cnt.hits++;
        V3 P, N;
        hit_geometry(s, h.tri, o, d, h.t, h.u, h.v, P, N);
        int mi = s.mat[h.tri];
        const orc_material& m = s.mats[mi];
        float acc[3] = {0.0f, 0.0f, 0.0f};
        uint32_t vis = 0;
        for (int i = 0; i < L; ++i) {
            const orc_light& l = s.lights[i];
            V3 Ls = normalize(ld3(l.position) - P);
            V3 so = P + mul(1e-3f, Ls);
            cnt.shadow++;
            if (!any_f(so, Ls)) {
                vis |= 1u << (i & 31);
                float c[3];
                phong(m, l, camp, P, N, c);
                for (int k = 0; k < 3; ++k) acc[k] = acc[k] + c[k];
            }
        }
        for (int k = 0; k < 3; ++k) Pd[lev][k] = acc[k];
        Rd[lev] = m.reflective ? m.reflectivity : 0.0f;
        levels = lev + 1;
        if (recs) {
            orc_rec& r = recs[lev * stride];
            r.tri = h.tri; r.t = h.t; r.u = h.u; r.v = h.v;
            r.p[0] = P.x; r.p[1] = P.y; r.p[2] = P.z;
            r.n[0] = N.x; r.n[1] = N.y; r.n[2] = N.z;
            r.mat = mi; r.vis = vis;
            r.refl = (m.reflective && lev < depth) ? m.reflectivity : 0.0f;
        }
        if (!m.reflective || lev == depth) break;
        // src/RayTracer.cpp:348-357
        V3 Rdir = normalize(reflect(d, N));
        o = P + mul(1e-3f, Rdir);
        d = Rdir;
        cnt.secondary++;
    }
    // c_l = P_l + r_l * c_{l+1}: the reference recursion's evaluation order.
    float c[3] = {0.0f, 0.0f, 0.0f};
    for (int lev = levels - 1; lev >= 0; --lev)
        for (int k = 0; k < 3; ++k) c[k] = Pd[lev][k] + Rd[lev] * c[k];
    for (int k = 0; k < 3; ++k) rgb[k] = c[k];
}

// src/RayTracer.cpp:52-65 (Bram de Greve's Fresnel reflectance): 1 on total
// internal reflection, else the mean of the squared s and p amplitudes.
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

// The second ray of a transparent material (src/RayTracer.cpp:361-408): the
// normal faces the incoming ray (negated, as 0 - N, when the ray leaves the
// material), K = fresnel(normalize(d), normal, n1, n2), and the ray is the
// reflection of d about that normal from P + 1e-3 * normal.  The reference
// traces this same ray a second time where it means to trace the refracted
// one (:403), so its colour t enters as K*t + (1-K)*t.
struct FresnelRay {
    V3 o, d;
    float K;
};
inline FresnelRay fresnel_ray(const orc_scene& s, const orc_material& m, V3 d, V3 P, V3 N) {
    const V3 inc = normalize(d);
    const float idn = dot(inc, N);
    float n1, n2;
    V3 nn = N;
    if (idn < 0.0f) {
        n1 = s.ambient_index;
        n2 = m.refractive_index;
    } else {
        n1 = m.refractive_index;
        n2 = s.ambient_index;
        nn = V3{0.0f - N.x, 0.0f - N.y, 0.0f - N.z};
    }
    FresnelRay f;
    f.K = fresnel(inc, nn, n1, n2);
    f.d = normalize(reflect(d, nn));
    f.o = P + mul(1e-3f, nn);
    return f;
}

// RayTracer::rayTrace / shadeIntersection as the reference's recursion, for
// scenes with transparent materials (a ray may have two children): colour =
// 0 + ((phong + 0 + r * c_reflect) + (0 + (K * t + (1 - K) * t))).  Without
// transparent materials it equals shade_path's chain bit for bit.
template <class NearestF, class AnyF>
void shade_tree(const orc_scene& s, const orc_camera& cam, int depth, int level, V3 o, V3 d, NearestF& nearest_f,
                AnyF& any_f, float rgb[3], int32_t* prim_hit, orc_counts& cnt) {
    rgb[0] = rgb[1] = rgb[2] = 0.0f;
    Hit h;
    nearest_f(o, d, h);
    if (level == 0 && prim_hit) *prim_hit = h.tri;
    if (h.tri < 0) return;
    cnt.hits++;
    V3 P, N;
    hit_geometry(s, h.tri, o, d, h.t, h.u, h.v, P, N);
    const orc_material& m = s.mats[s.mat[h.tri]];
    float acc[3] = {0.0f, 0.0f, 0.0f};
    for (int i = 0; i < s.n_light; ++i) {
        const orc_light& l = s.lights[i];
        V3 Ls = normalize(ld3(l.position) - P);
        V3 so = P + mul(1e-3f, Ls);
        cnt.shadow++;
        if (!any_f(so, Ls)) {
            float c[3];
            phong(m, l, ld3(cam.position), P, N, c);
            for (int k = 0; k < 3; ++k) acc[k] = acc[k] + c[k];
        }
    }
    float refl[3] = {0.0f, 0.0f, 0.0f}, refr[3] = {0.0f, 0.0f, 0.0f};
    if (m.reflective) {
        float c[3] = {0.0f, 0.0f, 0.0f};
        if (level < depth) {
            V3 R = normalize(reflect(d, N));
            cnt.secondary++;
            shade_tree(s, cam, depth, level + 1, P + mul(1e-3f, R), R, nearest_f, any_f, c, nullptr, cnt);
        }
        for (int k = 0; k < 3; ++k) refl[k] = 0.0f + m.reflectivity * c[k];
    }
    if (m.transparent) {
        const FresnelRay f = fresnel_ray(s, m, d, P, N);
        float t[3] = {0.0f, 0.0f, 0.0f};
        if (level < depth) {
            cnt.secondary++;
            shade_tree(s, cam, depth, level + 1, f.o, f.d, nearest_f, any_f, t, nullptr, cnt);
        }
        for (int k = 0; k < 3; ++k) refr[k] = 0.0f + (f.K * t[k] + (1.0f - f.K) * t[k]);
    }
    for (int k = 0; k < 3; ++k) rgb[k] = 0.0f + ((acc[k] + refl[k]) + refr[k]);
}

inline bool any_transparent(const orc_scene& s) {
    for (int i = 0; i < s.n_mat; ++i)
        if (s.mats[i].transparent) return true;
    return false;
}

template <class NearestF, class AnyF>
void shade_ray(const orc_scene& s, const orc_camera& cam, int depth, V3 o, V3 d, NearestF&& nearest_f, AnyF&& any_f,
               float rgb[3], int32_t* prim_hit, orc_rec* recs, int64_t stride, orc_counts& cnt, bool tree) {
    if (tree) shade_tree(s, cam, depth, 0, o, d, nearest_f, any_f, rgb, prim_hit, cnt);
    else shade_path(s, cam, depth, o, d, nearest_f, any_f, rgb, prim_hit, recs, stride, cnt);
}


// ------------------------------------------------------------ Morton / LBVH
// src/morton.h:30-48
inline uint32_t part1by2(uint32_t x) {
    x &= 0x000003ff;
    x = (x ^ (x << 16)) & 0xff0000ff;
    x = (x ^ (x << 8)) & 0x0300f00f;
    x = (x ^ (x << 4)) & 0x030c30c3;
    x = (x ^ (x << 2)) & 0x09249249;
    return x;
}
inline uint32_t morton3(uint32_t x, uint32_t y, uint32_t z) {
    return (part1by2(z) << 2) + (part1by2(y) << 1) + part1by2(x);
}

// Centroid as src/Triangle.h:38 ((v1+v2)+v3)/3.
inline void centroid(const float* t, float c[3]) {
    for (int k = 0; k < 3; ++k) c[k] = ((t[k] + t[3 + k]) + t[6 + k]) / 3.0f;
}
// Quantise to 10 bits per axis against the scene's centroid bounds.
inline uint32_t quant10(float c, float lo, float hi) {
    float e = hi - lo;
    float s = e > 0.0f ? 1024.0f / e : 0.0f;
    float x = (c - lo) * s;
    x = std::min(std::max(x, 0.0f), 1023.0f);
    return (uint32_t)x;
}

inline int clz32(uint32_t x) { return x ? __builtin_clz(x) : 32; }
inline int delta(const uint32_t* k, int n, int i, int j) {
    if (j < 0 || j > n - 1) return -1;
    uint32_t a = k[i], b = k[j];
    if (a == b) return 32 + clz32((uint32_t)i ^ (uint32_t)j);
    return clz32(a ^ b);
}

inline float inflate_pad(float lo, float hi, float eps_scene) {
    float a = std::fabs(lo), b = std::fabs(hi);
    float m = a > b ? a : b;
    return m * 0x1p-16f + eps_scene;
}

inline int32_t leaf_ref(int32_t first, int32_t count) { return ~(first | ((count - 1) << 28)); }

}  // namespace

// ====================================================================== C API
extern "C" {

int orc_mt(const float* o, const float* d, const float* tri9, float* tuv) {
    float t, u, v;
    bool h = mt(ld3(o), ld3(d), tri9, t, u, v);
    if (h && tuv) { tuv[0] = t; tuv[1] = u; tuv[2] = v; }
    return h ? 1 : 0;
}

void orc_camera_ray(const orc_camera* cam, int x, int y, float* dir3) {
    V3 d = camera_dir(*cam, x, y);
    dir3[0] = d.x; dir3[1] = d.y; dir3[2] = d.z;
}

void orc_camera_ray_aa(const orc_camera* cam, int x, int y, int p, int q, int n, float* dir3) {
    V3 d = camera_dir_aa(*cam, x, y, p, q, (float)(1.0 / (double)(float)n));
    dir3[0] = d.x; dir3[1] = d.y; dir3[2] = d.z;
}

uint8_t orc_quantize(float c) { return quantize(c); }

void orc_mat4_inverse(const float* m, float* out) { inverse_glm(m, out); }

uint32_t orc_morton3(uint32_t x, uint32_t y, uint32_t z) { return morton3(x, y, z); }

float orc_shade_phong(const orc_scene* s, const float* cam_pos, int light, int mat, const float* p,
                      const float* n, float* out3) {
    phong(s->mats[mat], s->lights[light], ld3(cam_pos), ld3(p), ld3(n), out3);
    return out3[0];
}

void* orc_accel_build(const orc_scene* s, int brute_force) {
    Accel* A = new Accel();
    A->brute = brute_force != 0;
    if (A->brute || s->n_tri == 0) return A;
    const std::vector<float> wv = world_verts(*s);   // boxes in world space
    std::vector<double> cen(3 * (size_t)s->n_tri);
    for (int i = 0; i < s->n_tri; ++i)
        for (int k = 0; k < 3; ++k) {
            const float* t = wv.data() + 9 * (size_t)i;
            cen[3 * (size_t)i + k] = ((double)t[k] + t[3 + k] + t[6 + k]) / 3.0;
        }
    A->idx.resize(s->n_tri);
    std::iota(A->idx.begin(), A->idx.end(), 0);
    A->nodes.reserve(2 * (size_t)s->n_tri / 2 + 8);
    A->nodes.push_back(ANode());
    abuild(*A, wv.data(), cen, 0, 0, s->n_tri);
    return A;
}

void orc_accel_free(void* a) { delete static_cast<Accel*>(a); }

void orc_nearest(const orc_scene* s, const void* accel, int32_t n, const float* orig, const float* dirs,
                 int32_t any_hit, int32_t* hit_tri, float* hit_t, float* hit_u, float* hit_v, float* point,
                 float* normal) {
    const Accel& A = *static_cast<const Accel*>(accel);
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < n; ++i) {
        V3 o = ld3(orig + 3 * (size_t)i), d = ld3(dirs + 3 * (size_t)i);
        if (any_hit) {
            hit_tri[i] = anyhit(A, *s, o, d) ? 1 : 0;
            continue;
        }
        Hit h;
        nearest(A, *s, o, d, h);
        hit_tri[i] = h.tri;
        if (hit_t) hit_t[i] = h.t;
        if (hit_u) hit_u[i] = h.u;
        if (hit_v) hit_v[i] = h.v;
        if (h.tri >= 0 && point && normal) {
            V3 P, N;
            hit_geometry(*s, h.tri, o, d, h.t, h.u, h.v, P, N);
            point[3 * (size_t)i] = P.x; point[3 * (size_t)i + 1] = P.y; point[3 * (size_t)i + 2] = P.z;
            normal[3 * (size_t)i] = N.x; normal[3 * (size_t)i + 1] = N.y; normal[3 * (size_t)i + 2] = N.z;
        }
    }
}

// src/Scene.cpp:159-201, Scene::findMinDistanceIntersectionLinear: every
// triangle in index order, tested in its local space through Triangle::hit
// (src/Triangle.cpp:26-55) with `distance` = minTemp, so a hit farther than the
// last accepted one is rejected (the first is accepted: its start value is
// uninitialised, read as no bound); an accepted hit is kept when
// 0 < t < minDist (999999 at first).  found = some hit was accepted.  The kept
// hit's point goes through M and its normal through transpose(inverse) without
// a second normalisation.
void orc_nearest_linear(const orc_scene* s, int32_t n, const float* orig, const float* dirs, int32_t* found,
                        int32_t* hit_tri, float* hit_t, float* point, float* normal) {
    // glm::inverse of each transformation, once (the reference recomputes the same values per test)
    std::vector<float> inv(16 * (size_t)(s->xform_index ? s->n_xform : 0));
    for (size_t x = 0; x < inv.size() / 16; ++x) inverse_glm(s->xforms + 16 * x, &inv[16 * x]);
#pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < n; ++i) {
        const V3 o = ld3(orig + 3 * (size_t)i), d = ld3(dirs + 3 * (size_t)i);
        float min_temp = INFINITY, min_dist = 999999.0f, bu = 0.0f, bv = 0.0f;
        int best = -1;
        bool any = false;
        for (int k = 0; k < s->n_tri; ++k) {
            float t, u, v;
            const int xi = xform_of(*s, k);
            const bool hit = xi < 0 ? mt(o, d, s->verts + 9 * (size_t)k, t, u, v)
                                    : mt(xform(&inv[16 * (size_t)xi], o, 1.0f), xform(&inv[16 * (size_t)xi], d, 0.0f),
                                         s->verts + 9 * (size_t)k, t, u, v);
            if (!hit || t > min_temp) continue;
            any = true;
            min_temp = t;
            if (t < min_dist && t > 0.0f) {
                min_dist = t;
                best = k;
                bu = u;
                bv = v;
            }
        }
        found[i] = any ? 1 : 0;
        hit_tri[i] = best;
        if (hit_t) hit_t[i] = best >= 0 ? min_dist : INFINITY;
        V3 P = {0, 0, 0}, N = {0, 0, 0};
        if (best >= 0) hit_geometry(*s, best, o, d, min_dist, bu, bv, P, N, false);
        if (point) { point[3 * (size_t)i] = P.x; point[3 * (size_t)i + 1] = P.y; point[3 * (size_t)i + 2] = P.z; }
        if (normal) { normal[3 * (size_t)i] = N.x; normal[3 * (size_t)i + 1] = N.y; normal[3 * (size_t)i + 2] = N.z; }
    }
}

int orc_render(const orc_scene* s, const void* accel, const orc_camera* cam, int depth, int row_start,
               int row_step, uint8_t* rgba, float* rgb, int32_t* hit_ids, orc_rec* recs,
               orc_counts* counts) {
    const Accel& A = *static_cast<const Accel*>(accel);
    const int W = cam->width, H = cam->height;
    const bool tree = any_transparent(*s);
    if (tree && recs) return -1;   // hit records describe chains of reflections only
    if (row_step < 1) row_step = 1;
    const int64_t npix = (int64_t)W * H;
    if (recs)
        for (int64_t i = 0; i < (int64_t)(depth + 1) * npix; ++i) {
            std::memset(&recs[i], 0, sizeof(orc_rec));
            recs[i].tri = -1;
        }
    int64_t prim = 0, shad = 0, sec = 0, hits = 0;
    std::vector<int> rows;
    for (int y = row_start; y < H; y += row_step) rows.push_back(y);
    const int nrows = (int)rows.size();
#pragma omp parallel for schedule(dynamic, 2) reduction(+ : prim, shad, sec, hits)
    for (int ri = 0; ri < nrows; ++ri) {
        int y = rows[ri];
        orc_counts c{};
        for (int x = 0; x < W; ++x) {
            int64_t p = (int64_t)y * W + x;
            V3 d = camera_dir(*cam, x, y);
            float col[3];
            c.primary++;
            shade_ray(
                *s, *cam, depth, ld3(cam->position), d,
                [&](V3 o, V3 dd, Hit& h) { nearest(A, *s, o, dd, h); },
                [&](V3 o, V3 dd) { return anyhit(A, *s, o, dd); }, col, hit_ids ? hit_ids + p : nullptr,
                recs ? recs + p : nullptr, npix, c, tree);
            if (rgb)
                for (int k = 0; k < 3; ++k) rgb[3 * p + k] = col[k];
            if (rgba) {
                for (int k = 0; k < 3; ++k) rgba[4 * p + k] = quantize(col[k]);
                rgba[4 * p + 3] = 255;
            }
        }
        prim += c.primary; shad += c.shadow; sec += c.secondary; hits += c.hits;
    }
    if (counts) {
        counts->primary = prim; counts->shadow = shad; counts->secondary = sec; counts->hits = hits;
        counts->node_visits = 0; counts->tri_tests = 0;
    }
    return 0;
}

// src/RayTracer.cpp:176-266 (renderWithAA): n x n samples per pixel, colours
// summed in (q, p) order into a zero vec4, scaled by (1/n)^2, quantised.
int orc_render_aa(const orc_scene* s, const void* accel, const orc_camera* cam, int depth, int n, uint8_t* rgba,
                  float* rgb, orc_counts* counts) {
    const Accel& A = *static_cast<const Accel*>(accel);
    const int W = cam->width, H = cam->height;
    const bool tree = any_transparent(*s);
    if (n < 1) n = 1;
    const float inv_n = (float)(1.0 / (double)(float)n);
    const float inv_n2 = inv_n * inv_n;
    int64_t prim = 0, shad = 0, sec = 0, hits = 0;
#pragma omp parallel for schedule(dynamic, 2) reduction(+ : prim, shad, sec, hits)
    for (int y = 0; y < H; ++y) {
        orc_counts c{};
        for (int x = 0; x < W; ++x) {
            float acc[3] = {0.0f, 0.0f, 0.0f};
            for (int q = 0; q < n; ++q)
                for (int pp = 0; pp < n; ++pp) {
                    V3 d = camera_dir_aa(*cam, x, y, pp, q, inv_n);
                    float col[3];
                    c.primary++;
                    shade_ray(
                        *s, *cam, depth, ld3(cam->position), d,
                        [&](V3 o, V3 dd, Hit& h) { nearest(A, *s, o, dd, h); },
                        [&](V3 o, V3 dd) { return anyhit(A, *s, o, dd); }, col, nullptr, nullptr, 0, c, tree);
                    for (int k = 0; k < 3; ++k) acc[k] = acc[k] + col[k];
                }
            const int64_t p = (int64_t)y * W + x;
            for (int k = 0; k < 3; ++k) {
                const float v = acc[k] * inv_n2;
                if (rgb) rgb[3 * p + k] = v;
                if (rgba) rgba[4 * p + k] = quantize(v);
            }
            if (rgba) rgba[4 * p + 3] = 255;
        }
        prim += c.primary; shad += c.shadow; sec += c.secondary; hits += c.hits;
    }
    if (counts) {
        counts->primary = prim; counts->shadow = shad; counts->secondary = sec; counts->hits = hits;
        counts->node_visits = 0; counts->tri_tests = 0;
    }
    return 0;
}

// ------------------------------------------------------------------ LBVH
void orc_morton_codes(int32_t n, const float* verts, uint32_t* codes, float* cb) {
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    std::vector<float> cen(3 * (size_t)n);
    for (int i = 0; i < n; ++i) {
        centroid(verts + 9 * (size_t)i, &cen[3 * (size_t)i]);
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], cen[3 * (size_t)i + k]);
            hi[k] = std::max(hi[k], cen[3 * (size_t)i + k]);
        }
    }
    for (int i = 0; i < n; ++i) {
        const float* c = &cen[3 * (size_t)i];
        codes[i] = morton3(quant10(c[0], lo[0], hi[0]), quant10(c[1], lo[1], hi[1]),
                           quant10(c[2], lo[2], hi[2]));
    }
    if (cb)
        for (int k = 0; k < 3; ++k) { cb[k] = lo[k]; cb[3 + k] = hi[k]; }
}

// Karras 2012, "Maximizing Parallelism in the Construction of BVHs, Octrees,
// and k-d Trees", Figure 4: per internal node i, direction from the longer
// common prefix, range end by exponential + binary search, split by binary
// search on delta; duplicate keys break ties by index (delta = 32 + clz(i^j)).
void orc_karras(int32_t n, const uint32_t* k, int32_t* kchild, int32_t* kparent, int32_t* first,
                int32_t* last) {
    if (kparent)
        for (int i = 0; i < 2 * n - 1; ++i) kparent[i] = -1;
    for (int i = 0; i < n - 1; ++i) {
        int d = (delta(k, n, i, i + 1) - delta(k, n, i, i - 1)) >= 0 ? 1 : -1;
        int dmin = delta(k, n, i, i - d);
        int lmax = 2;
        while (delta(k, n, i, i + lmax * d) > dmin) lmax *= 2;
        int l = 0;
        for (int t = lmax / 2; t >= 1; t /= 2)
            if (delta(k, n, i, i + (l + t) * d) > dmin) l += t;
        int j = i + l * d;
        int dnode = delta(k, n, i, j);
        int s = 0;
        for (int div = 2;; div *= 2) {
            int t = (l + div - 1) / div;
            if (delta(k, n, i, i + (s + t) * d) > dnode) s += t;
            if (t <= 1) break;
        }
        int g = i + s * d + std::min(d, 0);
        int lo = std::min(i, j), hi = std::max(i, j);
        int left = (lo == g) ? ~g : g;
        int right = (hi == g + 1) ? ~(g + 1) : g + 1;
        kchild[2 * i] = left;
        kchild[2 * i + 1] = right;
        if (first) first[i] = lo;
        if (last) last[i] = hi;
        if (kparent) {
            kparent[left >= 0 ? left : (n - 1 + ~left)] = i;
            kparent[right >= 0 ? right : (n - 1 + ~right)] = i;
        }
    }
}

void orc_lbvh_build(int32_t n, const float* verts, int32_t leaf_max, uint32_t* sorted_codes,
                    int32_t* perm_out, int32_t* kchild_out, int32_t* kparent_out, float* ibox_out,
                    float* nodes, float* tris, int32_t* root_ref, float* nodes4, int32_t* n_nodes4,
                    int32_t* root4, const float* xforms, const int32_t* xform_index) {
    if (n_nodes4) *n_nodes4 = 0;
    if (root4) *root4 = n <= 0 ? 0x7FFFFFFF : (n <= std::max(1, std::min(leaf_max, 8)) ? leaf_ref(0, n) : 0);
    if (n <= 0) {
        if (root_ref) *root_ref = 0;
        return;
    }
    if (leaf_max < 1) leaf_max = 1;
    if (leaf_max > 8) leaf_max = 8;
    // keys and boxes from world-space vertices; records keep the local ones
    const float* local = verts;
    std::vector<float> wv;
    if (xform_index) {
        orc_scene ws{};
        ws.n_tri = n;
        ws.verts = verts;
        ws.xforms = xforms;
        ws.xform_index = xform_index;
        wv = world_verts(ws);
        verts = wv.data();
    }
    std::vector<uint32_t> codes(n);
    float cb[6];
    orc_morton_codes(n, verts, codes.data(), cb);
    std::vector<int32_t> perm(n);
    std::iota(perm.begin(), perm.end(), 0);
    std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) { return codes[a] < codes[b]; });
    std::vector<uint32_t> sk(n);
    for (int i = 0; i < n; ++i) sk[i] = codes[perm[i]];
    if (sorted_codes) std::memcpy(sorted_codes, sk.data(), sizeof(uint32_t) * n);
    if (perm_out) std::memcpy(perm_out, perm.data(), sizeof(int32_t) * n);
    // triangle records in sorted order
    if (tris)
        for (int s = 0; s < n; ++s) {
            const float* t = local + 9 * (size_t)perm[s];
            float* r = tris + 12 * (size_t)s;
            V3 v1 = ld3(t), v2 = ld3(t + 3), v3 = ld3(t + 6);
            V3 e1 = v2 - v1, e2 = v3 - v1;
            r[0] = v1.x; r[1] = v1.y; r[2] = v1.z; r[3] = e1.x;
            r[4] = e1.y; r[5] = e1.z; r[6] = e2.x; r[7] = e2.y;
            r[8] = e2.z;
            int32_t o = perm[s];
            std::memcpy(&r[9], &o, 4);
            const int32_t xi = xform_index ? xform_index[perm[s]] : -1;
            std::memcpy(&r[10], &xi, 4);   // transformation index, -1 = none
            r[11] = 0.0f;
        }
    float eps_scene = 0.0f;
    for (int k = 0; k < 6; ++k) eps_scene = std::max(eps_scene, std::fabs(cb[k]));
    eps_scene *= 0x1p-22f;
    if (n == 1) {
        if (root_ref) *root_ref = leaf_ref(0, 1);
        if (root4) *root4 = leaf_ref(0, 1);
        return;
    }
    std::vector<int32_t> kc(2 * (size_t)(n - 1)), kp(2 * (size_t)n - 1), fi(n - 1), la(n - 1);
    orc_karras(n, sk.data(), kc.data(), kp.data(), fi.data(), la.data());
    if (kchild_out) std::memcpy(kchild_out, kc.data(), sizeof(int32_t) * 2 * (n - 1));
    if (kparent_out) std::memcpy(kparent_out, kp.data(), sizeof(int32_t) * (2 * n - 1));
    // leaf boxes (exact min/max of the three vertices)
    std::vector<float> lbox(6 * (size_t)n), ib(6 * (size_t)(n - 1));
    for (int s = 0; s < n; ++s) {
        const float* t = verts + 9 * (size_t)perm[s];
        for (int k = 0; k < 3; ++k) {
            lbox[6 * (size_t)s + k] = std::min(std::min(t[k], t[3 + k]), t[6 + k]);
            lbox[6 * (size_t)s + 3 + k] = std::max(std::max(t[k], t[3 + k]), t[6 + k]);
        }
    }
    // bottom-up union (post-order); min/max are exact so any order gives the same boxes
    std::vector<int> order;
    order.reserve(n - 1);
    std::vector<int> st{0};
    while (!st.empty()) {
        int i = st.back();
        st.pop_back();
        order.push_back(i);
        for (int c = 0; c < 2; ++c)
            if (kc[2 * i + c] >= 0) st.push_back(kc[2 * i + c]);
    }
    for (int oi = (int)order.size() - 1; oi >= 0; --oi) {
        int i = order[oi];
        const float* b[2];
        for (int c = 0; c < 2; ++c) {
            int ch = kc[2 * i + c];
            b[c] = ch >= 0 ? &ib[6 * (size_t)ch] : &lbox[6 * (size_t)~ch];
        }
        for (int k = 0; k < 3; ++k) {
            ib[6 * (size_t)i + k] = std::min(b[0][k], b[1][k]);
            ib[6 * (size_t)i + 3 + k] = std::max(b[0][3 + k], b[1][3 + k]);
        }
    }
    if (ibox_out) std::memcpy(ibox_out, ib.data(), sizeof(float) * 6 * (n - 1));
    if (nodes)
        for (int i = 0; i < n - 1; ++i) {
            float* nd = nodes + 16 * (size_t)i;
            int32_t ref[2];
            float lo[2][3], hi[2][3];
            for (int c = 0; c < 2; ++c) {
                int ch = kc[2 * i + c];
                const float* b = ch >= 0 ? &ib[6 * (size_t)ch] : &lbox[6 * (size_t)~ch];
                for (int k = 0; k < 3; ++k) {
                    float pad = inflate_pad(b[k], b[3 + k], eps_scene);
                    lo[c][k] = b[k] - pad;
                    hi[c][k] = b[3 + k] + pad;
                }
                if (ch < 0) ref[c] = leaf_ref(~ch, 1);
                else if (la[ch] - fi[ch] + 1 <= leaf_max) ref[c] = leaf_ref(fi[ch], la[ch] - fi[ch] + 1);
                else ref[c] = ch;
            }
            nd[0] = lo[0][0]; nd[1] = hi[0][0]; nd[2] = lo[0][1]; nd[3] = hi[0][1];
            nd[4] = lo[1][0]; nd[5] = hi[1][0]; nd[6] = lo[1][1]; nd[7] = hi[1][1];
            nd[8] = lo[0][2]; nd[9] = hi[0][2]; nd[10] = lo[1][2]; nd[11] = hi[1][2];
            std::memcpy(&nd[12], &ref[0], 4);
            std::memcpy(&nd[13], &ref[1], 4);
            nd[14] = 0.0f; nd[15] = 0.0f;
        }
    if (root_ref) *root_ref = (n <= leaf_max) ? leaf_ref(0, n) : 0;
    if (!nodes4 && !n_nodes4) return;
    // ---- 4-wide collapse: BVH4 nodes are the binary nodes that stay internal
    // (range > leaf_max) at even depth, numbered in binary index order; a node's
    // slots are its children, with an internal child replaced by its two children.
    const int m = n - 1;
    auto size_of = [&](int i) { return la[i] - fi[i] + 1; };
    std::vector<int32_t> idx4(m, -1);
    int cnt4 = 0;
    for (int i = 0; i < m; ++i) {
        if (size_of(i) <= leaf_max) continue;
        int depth = 0;
        for (int p = kp[i]; p >= 0; p = kp[p]) ++depth;
        if ((depth & 1) == 0) idx4[i] = cnt4++;
    }
    if (n_nodes4) *n_nodes4 = cnt4;
    if (!nodes4) return;
    auto slot_ref = [&](int c) -> int32_t {
        if (c < 0) return leaf_ref(~c, 1);
        if (size_of(c) <= leaf_max) return leaf_ref(fi[c], size_of(c));
        return idx4[c];
    };
    for (int i = 0; i < m; ++i) {
        if (idx4[i] < 0) continue;
        float* nd = nodes4 + 32 * (size_t)idx4[i];
        int slots[4];
        int ns = 0;
        for (int c = 0; c < 2; ++c) {
            int ch = kc[2 * i + c];
            if (ch >= 0 && size_of(ch) > leaf_max) {
                slots[ns++] = kc[2 * ch];
                slots[ns++] = kc[2 * ch + 1];
            } else {
                slots[ns++] = ch;
            }
        }
        for (int k = 0; k < 4; ++k) {
            int32_t ref = 0x7FFFFFFE;   // empty slot
            float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {INFINITY, INFINITY, INFINITY};
            if (k < ns) {
                int g = slots[k];
                const float* b = g >= 0 ? &ib[6 * (size_t)g] : &lbox[6 * (size_t)~g];
                for (int a = 0; a < 3; ++a) {
                    float pad = inflate_pad(b[a], b[3 + a], eps_scene);
                    lo[a] = b[a] - pad;
                    hi[a] = b[3 + a] + pad;
                }
                ref = slot_ref(g);
            }
            nd[0 + k] = lo[0]; nd[4 + k] = hi[0];
            nd[8 + k] = lo[1]; nd[12 + k] = hi[1];
            nd[16 + k] = lo[2]; nd[20 + k] = hi[2];
            std::memcpy(&nd[24 + k], &ref, 4);
            nd[28 + k] = 0.0f;
        }
    }
}

}  // extern "C"

// ----------------------------------------------------- GPU traversal, restated
namespace {

struct LRay { V3 o, d, invd, oid; };
inline LRay make_lray(V3 o, V3 d) {
    LRay r;
    r.o = o;
    r.d = d;
    float dd[3] = {d.x, d.y, d.z}, inv[3];
    for (int k = 0; k < 3; ++k) {
        float x = dd[k];
        if (std::fabs(x) < 1e-20f) x = std::copysign(1e-20f, x);
        inv[k] = 1.0f / x;
    }
    r.invd = {inv[0], inv[1], inv[2]};
    r.oid = {o.x * inv[0], o.y * inv[1], o.z * inv[2]};
    return r;
}

// Slab test of one child box against [0, tmax]; fma form lo*invd - o*invd.
inline bool slab(const LRay& r, float lox, float hix, float loy, float hiy, float loz, float hiz,
                 float tmax, float& tmin_out) {
    float x0 = std::fmaf(lox, r.invd.x, -r.oid.x), x1 = std::fmaf(hix, r.invd.x, -r.oid.x);
    float y0 = std::fmaf(loy, r.invd.y, -r.oid.y), y1 = std::fmaf(hiy, r.invd.y, -r.oid.y);
    float z0 = std::fmaf(loz, r.invd.z, -r.oid.z), z1 = std::fmaf(hiz, r.invd.z, -r.oid.z);
    float tmin = std::fmax(std::fmax(std::fmin(x0, x1), std::fmin(y0, y1)), std::fmax(std::fmin(z0, z1), 0.0f));
    float tmx = std::fmin(std::fmin(std::fmax(x0, x1), std::fmax(y0, y1)), std::fmin(std::fmax(z0, z1), tmax));
    tmin_out = tmin;
    return tmin <= tmx;
}

// One triangle record against a world-space ray, in the triangle's local space
// when its record carries a transformation (inverses in xinv, [n][16]).
inline bool rec_hit(const float* tris, const float* xinv, int32_t s, V3 o, V3 d, float& t, float& u, float& v,
                    int32_t& orig) {
    const float* tr = tris + 12 * (size_t)s;
    V3 v1 = {tr[0], tr[1], tr[2]}, e1 = {tr[3], tr[4], tr[5]}, e2 = {tr[6], tr[7], tr[8]};
    std::memcpy(&orig, &tr[9], 4);
    int32_t xi;
    std::memcpy(&xi, &tr[10], 4);
    if (xi >= 0 && xinv) {
        o = xform(xinv + 16 * (size_t)xi, o, 1.0f);
        d = xform(xinv + 16 * (size_t)xi, d, 0.0f);
    }
    return mt_e(o, d, v1, e1, e2, t, u, v);
}

std::vector<float> inverses(int32_t n_xform, const float* xforms) {
    std::vector<float> inv(16 * (size_t)std::max(n_xform, 0));
    for (int i = 0; i < n_xform; ++i) inverse_glm(xforms + 16 * (size_t)i, inv.data() + 16 * (size_t)i);
    return inv;
}

struct LHit { int32_t orig = -1, slot = -1; float t = INFINITY, u = 0, v = 0; };

// The traversal order the HIP kernels implement: ordered (near child first,
// child 0 on equal entry distance), far child pushed, pruning against the
// current best t, leaves tested in slot order.  Returns visits/tests.
template <bool ANY>
bool lbvh_trace_one(const float* nodes, int32_t root_ref, const float* tris, const float* xinv, V3 o, V3 d, LHit& h,
                    int64_t& visits, int64_t& tests) {
    LRay r = make_lray(o, d);
    int32_t stack[128];
    int sp = 0;
    int32_t cur = root_ref;
    if (cur == 0x7FFFFFFF) return false;
    for (;;) {
        if (cur >= 0) {
            visits++;
            const float* nd = nodes + 16 * (size_t)cur;
            int32_t c0, c1;
            std::memcpy(&c0, &nd[12], 4);
            std::memcpy(&c1, &nd[13], 4);
            float tm0, tm1;
            float tmax = ANY ? INFINITY : h.t;
            bool h0 = slab(r, nd[0], nd[1], nd[2], nd[3], nd[8], nd[9], tmax, tm0);
            bool h1 = slab(r, nd[4], nd[5], nd[6], nd[7], nd[10], nd[11], tmax, tm1);
            if (h0 && h1) {
                if (tm1 < tm0) { stack[sp++] = c0; cur = c1; }
                else { stack[sp++] = c1; cur = c0; }
            } else if (h0) {
                cur = c0;
            } else if (h1) {
                cur = c1;
            } else {
                if (sp == 0) break;
                cur = stack[--sp];
            }
        } else {
            int32_t e = ~cur;
            int32_t first = e & 0x0FFFFFFF, cnt = ((e >> 28) & 7) + 1;
            for (int32_t s = first; s < first + cnt; ++s) {
                tests++;
                int32_t orig;
                float t, u, v;
                if (rec_hit(tris, xinv, s, o, d, t, u, v, orig)) {
                    if (ANY) { h.orig = orig; h.slot = s; h.t = t; h.u = u; h.v = v; return true; }
                    if (t < h.t || (t == h.t && orig < h.orig)) {
                        h.orig = orig; h.slot = s; h.t = t; h.u = u; h.v = v;
                    }
                }
            }
            if (sp == 0) break;
            cur = stack[--sp];
        }
    }
    return h.orig >= 0;
}

// 4-wide: children hit sorted by (entry distance, slot); the nearest is taken,
// the others pushed farthest first.
template <bool ANY>
bool lbvh4_trace_one(const float* nodes, int32_t root_ref, const float* tris, const float* xinv, V3 o, V3 d,
                     LHit& h, int64_t& visits, int64_t& tests) {
    LRay r = make_lray(o, d);
    int32_t stack[256];
    int sp = 0;
    int32_t cur = root_ref;
    if (cur == 0x7FFFFFFF) return false;
    for (;;) {
        if (cur >= 0) {
            visits++;
            const float* nd = nodes + 32 * (size_t)cur;
            float tmax = ANY ? INFINITY : h.t;
            float tk[4];
            int32_t rk[4];
            int kk[4];
            int nh = 0;
            for (int k = 0; k < 4; ++k) {
                int32_t ref;
                std::memcpy(&ref, &nd[24 + k], 4);
                float tm;
                if (ref != 0x7FFFFFFE &&
                    slab(r, nd[0 + k], nd[4 + k], nd[8 + k], nd[12 + k], nd[16 + k], nd[20 + k], tmax, tm)) {
                    tk[nh] = tm; rk[nh] = ref; kk[nh] = k; ++nh;
                }
            }
            // order by (t, slot)
            for (int a = 1; a < nh; ++a)
                for (int b = a; b > 0 && (tk[b] < tk[b - 1] || (tk[b] == tk[b - 1] && kk[b] < kk[b - 1])); --b) {
                    std::swap(tk[b], tk[b - 1]); std::swap(rk[b], rk[b - 1]); std::swap(kk[b], kk[b - 1]);
                }
            if (nh == 0) {
                if (sp == 0) break;
                cur = stack[--sp];
            } else {
                for (int a = nh - 1; a >= 1; --a) stack[sp++] = rk[a];
                cur = rk[0];
            }
        } else {
            int32_t e = ~cur;
            int32_t first = e & 0x0FFFFFFF, cnt = ((e >> 28) & 7) + 1;
            bool stop = false;
            for (int32_t s = first; s < first + cnt; ++s) {
                tests++;
                int32_t orig;
                float t, u, v;
                if (rec_hit(tris, xinv, s, o, d, t, u, v, orig)) {
                    if (ANY) { h.orig = orig; h.slot = s; h.t = t; h.u = u; h.v = v; stop = true; break; }
                    if (t < h.t || (t == h.t && orig < h.orig)) {
                        h.orig = orig; h.slot = s; h.t = t; h.u = u; h.v = v;
                    }
                }
            }
            if (stop) return true;
            if (sp == 0) break;
            cur = stack[--sp];
        }
    }
    return h.orig >= 0;
}

template <bool ANY>
inline bool trace_w(int width, const float* nodes, int32_t root, const float* tris, const float* xinv, V3 o, V3 d,
                    LHit& h, int64_t& v, int64_t& t) {
    return width == 4 ? lbvh4_trace_one<ANY>(nodes, root, tris, xinv, o, d, h, v, t)
                      : lbvh_trace_one<ANY>(nodes, root, tris, xinv, o, d, h, v, t);
}

}  // namespace

extern "C" {

void orc_lbvh_trace(const float* nodes, int32_t root_ref, const float* tris, int32_t n_rays,
                    const float* orig, const float* dirs, int32_t any_hit, int32_t* hit_orig, float* hit_t,
                    float* hit_u, float* hit_v, int64_t* visits, int64_t* tests, int32_t width, int32_t n_xform,
                    const float* xforms) {
    const std::vector<float> xv = inverses(n_xform, xforms);
    const float* xinv = n_xform > 0 ? xv.data() : nullptr;
    int64_t vs = 0, ts = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : vs, ts)
    for (int i = 0; i < n_rays; ++i) {
        LHit h;
        V3 o = ld3(orig + 3 * (size_t)i), d = ld3(dirs + 3 * (size_t)i);
        int64_t v = 0, t = 0;
        if (any_hit) trace_w<true>(width, nodes, root_ref, tris, xinv, o, d, h, v, t);
        else trace_w<false>(width, nodes, root_ref, tris, xinv, o, d, h, v, t);
        vs += v; ts += t;
        if (hit_orig) hit_orig[i] = h.orig;
        if (hit_t) hit_t[i] = h.t;
        if (hit_u) hit_u[i] = h.u;
        if (hit_v) hit_v[i] = h.v;
    }
    if (visits) *visits += vs;
    if (tests) *tests += ts;
}

void orc_lbvh4_packet_primary(const float* nodes4, int32_t root4, const float* tris, const orc_camera* cam,
                              int32_t* hit_orig, float* hit_t, float* hit_u, float* hit_v, int64_t* visits,
                              int64_t* tests, int32_t n_xform, const float* xforms) {
    const std::vector<float> xv = inverses(n_xform, xforms);
    const float* xinv = n_xform > 0 ? xv.data() : nullptr;
    const int W = cam->width, H = cam->height;
    const int tiles_x = (W + 7) / 8, tiles_y = (H + 7) / 8;
    const int n_tiles = tiles_x * tiles_y;
    const V3 o = ld3(cam->position);
    int64_t vs = 0, ts = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : vs, ts)
    for (int tile = 0; tile < n_tiles; ++tile) {
        const int ty = tile / tiles_x, tx = tile - ty * tiles_x;
        LRay r[64];
        V3 dir[64];
        bool valid[64];
        LHit h[64];
        for (int i = 0; i < 64; ++i) {
            const int x = tx * 8 + (i & 7), y = ty * 8 + (i >> 3);
            valid[i] = x < W && y < H;
            dir[i] = camera_dir(*cam, x, y);
            r[i] = make_lray(o, dir[i]);
        }
        int32_t stack[96];
        int sp = 0;
        int32_t cur = root4;
        while (cur != 0x7FFFFFFF) {
            if (cur >= 0) {
                vs++;
                const float* nd = nodes4 + 32 * (size_t)cur;
                float e[4];
                int kk[4], rk[4];
                int nh = 0;
                for (int k = 0; k < 4; ++k) {
                    int32_t ref;
                    std::memcpy(&ref, &nd[24 + k], 4);
                    e[k] = INFINITY;
                    kk[k] = k;
                    rk[k] = ref;
                    if (ref == 0x7FFFFFFE) continue;
                    for (int i = 0; i < 64; ++i) {
                        float tm;
                        if (valid[i] && slab(r[i], nd[0 + k], nd[4 + k], nd[8 + k], nd[12 + k], nd[16 + k],
                                             nd[20 + k], h[i].t, tm)) {
                            e[k] = tm;   // the lowest-index lane that hits
                            ++nh;
                            break;
                        }
                    }
                }
                for (int a = 1; a < 4; ++a)
                    for (int b = a; b > 0 && (e[b] < e[b - 1] || (e[b] == e[b - 1] && kk[b] < kk[b - 1])); --b) {
                        std::swap(e[b], e[b - 1]); std::swap(kk[b], kk[b - 1]); std::swap(rk[b], rk[b - 1]);
                    }
                if (nh == 0) {
                    cur = sp == 0 ? 0x7FFFFFFF : stack[--sp];
                } else {
                    for (int a = nh - 1; a >= 1; --a) stack[sp++] = rk[a];
                    cur = rk[0];
                }
            } else {
                const int32_t ee = ~cur;
                const int32_t first = ee & 0x0FFFFFFF, last = first + ((ee >> 28) & 7);
                for (int32_t s0 = first; s0 <= last; s0 += 2) {
                    const int32_t s1 = s0 + 1 <= last ? s0 + 1 : s0;
                    ts += s1 != s0 ? 2 : 1;
                    for (int32_t s : {s0, s1}) {
                        for (int i = 0; i < 64; ++i) {
                            float t, u, v;
                            int32_t orig;
                            if (rec_hit(tris, xinv, s, o, dir[i], t, u, v, orig) &&
                                (t < h[i].t || (t == h[i].t && orig < h[i].orig))) {
                                h[i].orig = orig; h[i].slot = s; h[i].t = t; h[i].u = u; h[i].v = v;
                            }
                        }
                    }
                }
                cur = sp == 0 ? 0x7FFFFFFF : stack[--sp];
            }
        }
        for (int i = 0; i < 64; ++i) {
            if (!valid[i]) continue;
            const int x = tx * 8 + (i & 7), y = ty * 8 + (i >> 3);
            const size_t p = (size_t)y * W + x;
            if (hit_orig) hit_orig[p] = h[i].orig;
            if (hit_t) hit_t[p] = h[i].t;
            if (hit_u) hit_u[p] = h[i].u;
            if (hit_v) hit_v[p] = h[i].v;
        }
    }
    if (visits) *visits += vs;
    if (tests) *tests += ts;
}

int orc_render_lbvh(const orc_scene* s, const float* nodes, int32_t root_ref, const float* tris,
                    const orc_camera* cam, int depth, int row_start, int row_step, uint8_t* rgba, float* rgb,
                    int32_t* hit_ids, orc_counts* counts, int32_t width) {
    const int W = cam->width, H = cam->height;
    const bool tree = any_transparent(*s);
    if (row_step < 1) row_step = 1;
    int64_t prim = 0, shad = 0, sec = 0, hits = 0, vis = 0, tst = 0;
    std::vector<int> rows;
    for (int y = row_start; y < H; y += row_step) rows.push_back(y);
    const int nrows = (int)rows.size();
    const bool empty = s->n_tri == 0;
    const std::vector<float> xv = inverses(s->xform_index ? s->n_xform : 0, s->xforms);
    const float* xinv = s->xform_index && s->n_xform > 0 ? xv.data() : nullptr;
#pragma omp parallel for schedule(dynamic, 2) reduction(+ : prim, shad, sec, hits, vis, tst)
    for (int ri = 0; ri < nrows; ++ri) {
        int y = rows[ri];
        orc_counts c{};
        int64_t v = 0, t = 0;
        for (int x = 0; x < W; ++x) {
            int64_t p = (int64_t)y * W + x;
            V3 d = camera_dir(*cam, x, y);
            float col[3];
            c.primary++;
            shade_ray(
                *s, *cam, depth, ld3(cam->position), d,
                [&](V3 o, V3 dd, Hit& h) {
                    if (empty) return;
                    LHit lh;
                    if (trace_w<false>(width, nodes, root_ref, tris, xinv, o, dd, lh, v, t)) {
                        h.tri = lh.orig; h.t = lh.t; h.u = lh.u; h.v = lh.v;
                    }
                },
                [&](V3 o, V3 dd) {
                    if (empty) return false;
                    LHit lh;
                    return trace_w<true>(width, nodes, root_ref, tris, xinv, o, dd, lh, v, t);
                },
                col, hit_ids ? hit_ids + p : nullptr, nullptr, 0, c, tree);
            if (rgb)
                for (int k = 0; k < 3; ++k) rgb[3 * p + k] = col[k];
            if (rgba) {
                for (int k = 0; k < 3; ++k) rgba[4 * p + k] = quantize(col[k]);
                rgba[4 * p + 3] = 255;
            }
        }
        prim += c.primary; shad += c.shadow; sec += c.secondary; hits += c.hits;
        vis += v; tst += t;
    }
    if (counts) {
        counts->primary = prim; counts->shadow = shad; counts->secondary = sec; counts->hits = hits;
        counts->node_visits = vis; counts->tri_tests = tst;
    }
    return 0;
}

}  // extern "C"
