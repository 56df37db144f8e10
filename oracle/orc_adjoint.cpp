// orc_adjoint.cpp -- TEST INFRASTRUCTURE ONLY: FP64 shading on fixed hit sets,
// the reference for the HIP adjoint kernel's gradients.
//
// The reference has no gradients.  The forward being differentiated is its
// shading (src/RayTracer.cpp:302-458): per level, the sum over unshadowed lights
// of diffuse max(0, L.N) kd lc plus specular (V.R > 0 ? (V.R)^s lc ks : 0), and
// c_l = P_l + r_l c_{l+1} across reflection levels.  Hit points, normals,
// materials, shadow masks and reflection weights come from the recorded hit
// sets (orc_rec) and are held fixed.  Two independent derivations:
//   orc_grad_fd64  -- central differences of the FP64 loss,
//   orc_grad_fwd64 -- forward mode (dual numbers seeded per material/light).
#include "oracle.h"

#include <cmath>
#include <cstring>
#include <vector>

namespace {

constexpr int K = 13;  // 7 material tangents + 6 light tangents

struct Dual {
    double v;
    double g[K];
    Dual() : v(0.0) { std::memset(g, 0, sizeof g); }
    Dual(double x) : v(x) { std::memset(g, 0, sizeof g); }  // NOLINT: constants promote
};
inline Dual cst(double v) { Dual d; d.v = v; std::memset(d.g, 0, sizeof d.g); return d; }
inline Dual var(double v, int k) { Dual d = cst(v); d.g[k] = 1.0; return d; }
inline Dual operator+(const Dual& a, const Dual& b) { Dual r; r.v = a.v + b.v; for (int k = 0; k < K; ++k) r.g[k] = a.g[k] + b.g[k]; return r; }
inline Dual operator-(const Dual& a, const Dual& b) { Dual r; r.v = a.v - b.v; for (int k = 0; k < K; ++k) r.g[k] = a.g[k] - b.g[k]; return r; }
inline Dual operator*(const Dual& a, const Dual& b) { Dual r; r.v = a.v * b.v; for (int k = 0; k < K; ++k) r.g[k] = a.g[k] * b.v + a.v * b.g[k]; return r; }
inline Dual operator*(double s, const Dual& a) { Dual r; r.v = s * a.v; for (int k = 0; k < K; ++k) r.g[k] = s * a.g[k]; return r; }
inline Dual dsqrt(const Dual& a) { Dual r; r.v = std::sqrt(a.v); double f = 0.5 / r.v; for (int k = 0; k < K; ++k) r.g[k] = a.g[k] * f; return r; }
inline Dual drecip(const Dual& a) { Dual r; r.v = 1.0 / a.v; double f = -r.v * r.v; for (int k = 0; k < K; ++k) r.g[k] = a.g[k] * f; return r; }
inline Dual dpow(const Dual& x, const Dual& s) {
    Dual r;
    r.v = std::pow(x.v, s.v);
    double lx = std::log(x.v);
    double dx = s.v * std::pow(x.v, s.v - 1.0);
    for (int k = 0; k < K; ++k) r.g[k] = dx * x.g[k] + r.v * lx * s.g[k];
    return r;
}
inline double val(double x) { return x; }
inline double val(const Dual& x) { return x.v; }
inline double dsqrt(double a) { return std::sqrt(a); }
inline double drecip(double a) { return 1.0 / a; }
inline double dpow(double x, double s) { return std::pow(x, s); }

template <class T> struct V { T x, y, z; };
template <class T> inline T dot(const V<T>& a, const V<T>& b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
template <class T> inline V<T> norm(const V<T>& a) {
    T s = drecip(dsqrt(dot(a, a)));
    return {a.x * s, a.y * s, a.z * s};
}

// One light's Phong term at a recorded hit; params are already scalars of type T.
template <class T>
inline void phong(const T kd[3], const T ks[3], const T& s, const T lp[3], const T lc[3], const double P[3],
                  const double N[3], const double cam[3], T out[3]) {
    V<T> L = norm(V<T>{lp[0] - T(P[0]), lp[1] - T(P[1]), lp[2] - T(P[2])});
    V<double> Vv = norm(V<double>{cam[0] - P[0], cam[1] - P[1], cam[2] - P[2]});
    V<T> Nn{T(N[0]), T(N[1]), T(N[2])};
    T dNL = dot(Nn, L);
    T s2 = 2.0 * dNL;
    V<T> R = norm(V<T>{s2 * Nn.x - L.x, s2 * Nn.y - L.y, s2 * Nn.z - L.z});
    T dLN = val(dNL) > 0.0 ? dNL : T(0.0);
    V<T> Vt{T(Vv.x), T(Vv.y), T(Vv.z)};
    T dVR = dot(Vt, R);
    bool sp = val(dVR) > 0.0;
    T st = sp ? dpow(dVR, s) : T(0.0);
    for (int k = 0; k < 3; ++k) {
        T diff = (dLN * kd[k]) * lc[k];
        out[k] = sp ? diff + (st * lc[k]) * ks[k] : diff;
    }
}


struct Ctx {
    const orc_rec* recs;
    int levels;
    int64_t npix;
    double cam[3];
    int nm, nl;
    const double* prm;
    const float* target;
};

inline const double* matp(const Ctx& c, int m) { return c.prm + 7 * (size_t)m; }
inline const double* litp(const Ctx& c, int l) { return c.prm + 7 * (size_t)c.nm + 6 * (size_t)l; }

// FP64 pixel colour from fixed records, with an optional parameter override.
inline void pixel64(const Ctx& c, int64_t p, const double* prm, double out[3]) {
    double col[3] = {0, 0, 0};
    int last = -1;
    for (int l = 0; l < c.levels; ++l) {
        if (c.recs[l * c.npix + p].tri < 0) break;
        last = l;
    }
    for (int l = last; l >= 0; --l) {
        const orc_rec& r = c.recs[l * c.npix + p];
        double P[3] = {r.p[0], r.p[1], r.p[2]}, N[3] = {r.n[0], r.n[1], r.n[2]};
        const double* m = prm + 7 * (size_t)r.mat;
        double acc[3] = {0, 0, 0};
        for (int i = 0; i < c.nl; ++i) {
            if (!((r.vis >> (i & 31)) & 1u)) continue;
            const double* L = prm + 7 * (size_t)c.nm + 6 * (size_t)i;
            double o[3];
            phong<double>(m, m + 3, m[6], L, L + 3, P, N, c.cam, o);
            for (int k = 0; k < 3; ++k) acc[k] += o[k];
        }
        double rf = (l < last) ? (double)r.refl : 0.0;
        for (int k = 0; k < 3; ++k) col[k] = acc[k] + rf * col[k];
    }
    for (int k = 0; k < 3; ++k) out[k] = col[k];
}

}  // namespace

extern "C" {

double orc_loss64(const orc_rec* recs, int32_t levels, int64_t n_pix, const float* cam_pos, int32_t n_mat,
                  int32_t n_light, const double* params, const float* target_rgb, double* image_out) {
    Ctx c{recs, levels, n_pix, {cam_pos[0], cam_pos[1], cam_pos[2]}, n_mat, n_light, params, target_rgb};
    double loss = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : loss)
    for (int64_t p = 0; p < n_pix; ++p) {
        double col[3];
        pixel64(c, p, params, col);
        for (int k = 0; k < 3; ++k) {
            if (image_out) image_out[3 * p + k] = col[k];
            double e = col[k] - (double)target_rgb[3 * p + k];
            loss += e * e;
        }
    }
    return loss;
}

void orc_grad_fd64(const orc_rec* recs, int32_t levels, int64_t n_pix, const float* cam_pos, int32_t n_mat,
                   int32_t n_light, const double* params, const float* target_rgb, double h,
                   const int32_t* which, int32_t n_which, double* grad_out) {
    Ctx c{recs, levels, n_pix, {cam_pos[0], cam_pos[1], cam_pos[2]}, n_mat, n_light, params, target_rgb};
    const int np = 7 * n_mat + 6 * n_light;
    for (int w = 0; w < n_which; ++w) {
        int j = which[w];
        int mat = j < 7 * n_mat ? j / 7 : -1;
        int lig = j < 7 * n_mat ? -1 : (j - 7 * n_mat) / 6;
        std::vector<double> pp(params, params + np), pm(params, params + np);
        pp[j] += h;
        pm[j] -= h;
        double dl = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : dl)
        for (int64_t p = 0; p < n_pix; ++p) {
            bool aff = false;
            for (int l = 0; l < levels; ++l) {
                const orc_rec& r = recs[l * n_pix + p];
                if (r.tri < 0) break;
                if ((mat >= 0 && r.mat == mat) || (lig >= 0 && ((r.vis >> (lig & 31)) & 1u))) { aff = true; break; }
            }
            if (!aff) continue;
            double a[3], b[3];
            pixel64(c, p, pp.data(), a);
            pixel64(c, p, pm.data(), b);
            for (int k = 0; k < 3; ++k) {
                double t = target_rgb[3 * p + k];
                dl += (a[k] - t) * (a[k] - t) - (b[k] - t) * (b[k] - t);
            }
        }
        grad_out[w] = dl / (2.0 * h);
    }
}

void orc_grad_fwd64(const orc_rec* recs, int32_t levels, int64_t n_pix, const float* cam_pos, int32_t n_mat,
                    int32_t n_light, const double* params, const float* target_rgb, double* grad_out) {
    Ctx c{recs, levels, n_pix, {cam_pos[0], cam_pos[1], cam_pos[2]}, n_mat, n_light, params, target_rgb};
    const int np = 7 * n_mat + 6 * n_light;
    std::memset(grad_out, 0, sizeof(double) * np);
#pragma omp parallel
    {
        std::vector<double> g(np, 0.0);
#pragma omp for schedule(static)
        for (int64_t p = 0; p < n_pix; ++p) {
            double col[3];
            pixel64(c, p, params, col);
            double res[3];
            for (int k = 0; k < 3; ++k) res[k] = 2.0 * (col[k] - (double)target_rgb[3 * p + k]);
            double w = 1.0;  // product of reflection weights above this level
            for (int l = 0; l < levels; ++l) {
                const orc_rec& r = recs[l * n_pix + p];
                if (r.tri < 0) break;
                double P[3] = {r.p[0], r.p[1], r.p[2]}, N[3] = {r.n[0], r.n[1], r.n[2]};
                const double* m = matp(c, r.mat);
                Dual kd[3], ks[3], s;
                for (int k = 0; k < 3; ++k) { kd[k] = var(m[k], k); ks[k] = var(m[3 + k], 3 + k); }
                s = var(m[6], 6);
                for (int i = 0; i < n_light; ++i) {
                    if (!((r.vis >> (i & 31)) & 1u)) continue;
                    const double* L = litp(c, i);
                    Dual lp[3], lc[3];
                    for (int k = 0; k < 3; ++k) { lp[k] = var(L[k], 7 + k); lc[k] = var(L[3 + k], 10 + k); }
                    Dual o[3];
                    phong<Dual>(kd, ks, s, lp, lc, P, N, c.cam, o);
                    for (int ch = 0; ch < 3; ++ch) {
                        double f = w * res[ch];
                        for (int k = 0; k < 7; ++k) g[7 * (size_t)r.mat + k] += f * o[ch].g[k];
                        for (int k = 0; k < 6; ++k) g[7 * (size_t)n_mat + 6 * (size_t)i + k] += f * o[ch].g[7 + k];
                    }
                }
                w *= (double)r.refl;
                if (r.refl == 0.0f) break;
            }
        }
#pragma omp critical
        for (int j = 0; j < np; ++j) grad_out[j] += g[j];
    }
}

}  // extern "C"
