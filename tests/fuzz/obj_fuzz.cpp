// Random and malformed Wavefront OBJ files through eos_mesh_load_obj, built
// with the host address and undefined-behaviour sanitizers by
// tests/test_obj_fuzz.py: every file must be read or rejected with a status,
// with no invalid memory access, leak or undefined operation.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <random>
#include "eos_c.h"

// The library's error recorder, outside the reader under test.
namespace eos {
eos_status fail(eos_status s, const std::string&) { return s; }
}  // namespace eos
static std::mt19937 rng(12345);
static int R(int n) { return (int)(rng() % (unsigned)n); }
static std::string num() {
    switch (R(8)) {
        case 0: return std::to_string(R(10) - 5);
        case 1: return "1e39";
        case 2: return "-";
        case 3: return "nan";
        case 4: return std::to_string(R(1000)) + "." + std::to_string(R(1000));
        case 5: return "99999999999999999999";
        case 6: return "";
        default: return std::to_string((rng() % 2000) / 100.0 - 10.0);
    }
}
static std::string idx() {
    std::string s;
    switch (R(10)) {
        case 0: s = std::to_string(-(R(6) + 1)); break;
        case 1: s = "0"; break;
        case 2: s = "2147483648"; break;
        case 3: s = "-2147483649"; break;
        case 4: s = "x"; break;
        default: s = std::to_string(R(8) + 1);
    }
    switch (R(6)) {
        case 0: s += "/" + std::to_string(R(3) + 1); break;
        case 1: s += "//" + std::to_string(R(6) - 2); break;
        case 2: s += "/" + std::to_string(R(3)) + "/" + std::to_string(R(6) - 2); break;
        case 3: s += "/"; break;
        case 4: s += "//"; break;
        default: break;
    }
    return s;
}
int main(int argc, char** argv) {
    const int iters = argc > 1 ? atoi(argv[1]) : 20000;
    const char* path = argc > 2 ? argv[2] : "fuzz.obj";
    int ok = 0, bad = 0;
    for (int it = 0; it < iters; ++it) {
        std::string t;
        const int nl = R(30);
        for (int l = 0; l < nl; ++l) {
            switch (R(9)) {
                case 0: case 1: t += "v " + num() + " " + num() + " " + num(); break;
                case 2: t += "vn " + num() + " " + num() + " " + num(); break;
                case 3: case 4: case 5: { t += "f"; int k = R(7); for (int c = 0; c < k; ++c) t += " " + idx(); } break;
                case 6: t += "# comment " + num(); break;
                case 7: t += std::string(R(3) ? 5 : 5000, 'a' + R(26)); break;
                default: t += (R(2) ? "vt 0 0" : "o x\r"); break;
            }
            t += R(10) ? "\n" : "";
        }
        FILE* f = fopen(path, "wb");
        fwrite(t.data(), 1, t.size(), f);
        fclose(f);
        eos_mesh m;
        if (eos_mesh_load_obj(path, &m) == EOS_OK) {
            for (int i = 0; i < m.num_triangles * 9; ++i) { volatile float x = m.vertices[i]; (void)x; volatile float y = m.normals[i]; (void)y; }
            eos_mesh_free(&m);
            ++ok;
        } else {
            ++bad;
        }
    }
    printf("ok %d rejected %d\n", ok, bad);
}
