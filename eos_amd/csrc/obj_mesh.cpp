// obj_mesh.cpp -- Wavefront OBJ triangle meshes for the render path, native
// (the reference reads them through assimp, src/triangleMesh.cpp:79-160, which
// is not part of this build).
//
// What is read: `v x y z` positions, `vn x y z` normals and `f` faces whose
// corners are `v`, `v/vt`, `v//vn` or `v/vt/vn` (1-based, negative = relative
// to the end).  Polygons become triangle fans from their first corner (the
// assimp Triangulate step does the same for convex polygons).  Every face of
// the file goes into the one mesh (the reference keeps assimp's first mesh,
// which is the whole file when the file has no `usemtl`/`o` splits).
//
// Normals: a corner with a `vn` index takes that normal.  When the file has no
// normals for a face, each corner takes a smooth vertex normal: the normalized
// sum of the unit face normals of every face that uses the same position
// (assimp's JoinIdenticalVertices + GenSmoothNormals idea; its exact
// arithmetic is not reproduced).
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "eos_internal.h"

namespace {

struct V3d {
    float x, y, z;
};

// Parses a corner "v", "v/vt", "v//vn" or "v/vt/vn"; indices are resolved to
// 0-based (-1: absent).  Returns false on malformed input.
bool parse_corner(const char* tok, int nv, int nn, int& vi, int& ni) {
    char* end = nullptr;
    long a = std::strtol(tok, &end, 10);
    if (end == tok) return false;
    ni = -1;
    if (*end == '/') {
        const char* p = end + 1;
        if (*p != '/') {
            (void)std::strtol(p, &end, 10);  // texture index, unused
            p = end;
        }
        if (*p == '/') {
            ++p;
            long c = std::strtol(p, &end, 10);
            if (end == p) return false;
            // range-checked as long, before the narrowing (1-based, or -1 for the last)
            if (c == 0 || c > (long)nn || c < -(long)nn) return false;
            ni = c > 0 ? (int)(c - 1) : (int)(nn + c);
        }
    }
    if (a == 0 || a > (long)nv || a < -(long)nv) return false;
    vi = a > 0 ? (int)(a - 1) : (int)(nv + a);
    return true;
}

}  // namespace

extern "C" eos_status eos_mesh_load_obj(const char* path, eos_mesh* out) {
    if (!path || !out) return eos::fail(EOS_ERR_INVALID, "eos_mesh_load_obj: null argument");
    std::memset(out, 0, sizeof *out);
    FILE* fp = std::fopen(path, "r");
    if (!fp) return eos::fail(EOS_ERR_INVALID, std::string("eos_mesh_load_obj: cannot open ") + path + ": " +
                                                   std::strerror(errno));
    std::vector<V3d> pos, nrm;
    std::vector<int> fv, fn;  // per triangle corner: position index, normal index (-1: none)
    char line[4096];
    int lineno = 0;
    std::string err;
    while (std::fgets(line, sizeof line, fp)) {
        ++lineno;
        char* s = line;
        while (*s == ' ' || *s == '\t') ++s;
        if (s[0] == 'v' && (s[1] == ' ' || s[1] == '\t')) {
            V3d p{};
            if (std::sscanf(s + 2, "%f %f %f", &p.x, &p.y, &p.z) != 3) { err = "bad vertex"; break; }
            pos.push_back(p);
        } else if (s[0] == 'v' && s[1] == 'n' && (s[2] == ' ' || s[2] == '\t')) {
            V3d p{};
            if (std::sscanf(s + 3, "%f %f %f", &p.x, &p.y, &p.z) != 3) { err = "bad normal"; break; }
            nrm.push_back(p);
        } else if (s[0] == 'f' && (s[1] == ' ' || s[1] == '\t')) {
            std::vector<int> cv, cn;
            char* tok = std::strtok(s + 2, " \t\r\n");
            while (tok) {
                int vi, ni;
                if (!parse_corner(tok, (int)pos.size(), (int)nrm.size(), vi, ni)) { err = "bad face corner"; break; }
                cv.push_back(vi);
                cn.push_back(ni);
                tok = std::strtok(nullptr, " \t\r\n");
            }
            if (!err.empty()) break;
            if (cv.size() < 3) { err = "face with fewer than 3 corners"; break; }
            for (size_t k = 1; k + 1 < cv.size(); ++k) {
                const size_t c[3] = {0, k, k + 1};
                for (size_t j : c) {
                    fv.push_back(cv[j]);
                    fn.push_back(cn[j]);
                }
            }
        }
    }
    std::fclose(fp);
    if (!err.empty())
        return eos::fail(EOS_ERR_INVALID, std::string("eos_mesh_load_obj: ") + path + ":" + std::to_string(lineno) +
                                              ": " + err);
    const size_t nt = fv.size() / 3;
    if (nt >= ((size_t)1 << 28)) return eos::fail(EOS_ERR_LIMIT, "eos_mesh_load_obj: too many triangles");
    // smooth normals for corners without one: sum of unit face normals per position
    std::vector<double> acc;
    bool need_smooth = false;
    for (int n : fn) need_smooth |= n < 0;
    if (need_smooth) {
        acc.assign(3 * pos.size(), 0.0);
        for (size_t t = 0; t < nt; ++t) {
            const V3d& a = pos[fv[3 * t]];
            const V3d& b = pos[fv[3 * t + 1]];
            const V3d& c = pos[fv[3 * t + 2]];
            const double e1[3] = {(double)b.x - a.x, (double)b.y - a.y, (double)b.z - a.z};
            const double e2[3] = {(double)c.x - a.x, (double)c.y - a.y, (double)c.z - a.z};
            double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
            const double len = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
            if (len == 0.0) continue;
            for (int j = 0; j < 3; ++j)
                for (int k = 0; k < 3; ++k) acc[3 * (size_t)fv[3 * t + j] + k] += n[k] / len;
        }
    }
    float* V = static_cast<float*>(std::malloc(sizeof(float) * 9 * (nt ? nt : 1)));
    float* N = static_cast<float*>(std::malloc(sizeof(float) * 9 * (nt ? nt : 1)));
    if (!V || !N) {
        std::free(V);
        std::free(N);
        return eos::fail(EOS_ERR_NOMEM, "eos_mesh_load_obj: out of host memory");
    }
    for (size_t c = 0; c < 3 * nt; ++c) {
        const V3d& p = pos[fv[c]];
        V[3 * c] = p.x;
        V[3 * c + 1] = p.y;
        V[3 * c + 2] = p.z;
        if (fn[c] >= 0) {
            const V3d& q = nrm[fn[c]];
            N[3 * c] = q.x;
            N[3 * c + 1] = q.y;
            N[3 * c + 2] = q.z;
        } else {
            const double* a = &acc[3 * (size_t)fv[c]];
            const double len = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
            const double s = len > 0.0 ? 1.0 / len : 0.0;
            N[3 * c] = (float)(a[0] * s);
            N[3 * c + 1] = (float)(a[1] * s);
            N[3 * c + 2] = (float)(a[2] * s);
        }
    }
    out->num_triangles = (int32_t)nt;
    out->vertices = V;
    out->normals = N;
    return EOS_OK;
}

extern "C" eos_status eos_mesh_free(eos_mesh* m) {
    if (!m) return eos::fail(EOS_ERR_INVALID, "eos_mesh_free: null argument");
    std::free(m->vertices);
    std::free(m->normals);
    std::memset(m, 0, sizeof *m);
    return EOS_OK;
}
