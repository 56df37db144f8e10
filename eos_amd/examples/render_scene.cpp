// render_scene.cpp -- drives the GPU path through the reference-compatible C++
// API exactly as the reference's driver does (src/main.cpp:355-481): build a
// Scene from Triangle surfaces, materials and lights, construct a
// MultiDeviceRenderer and render once to a host buffer; then the single-GPU
// route (createCudaScene + DeviceCameraHandler + DeviceRenderer).
//
// usage: render_scene <scene.bin> <out_multi.rgba> <out_single.rgba> [mesh.obj mesh_material [tx ty tz]]
// (environment: EOS_AA, EOS_BANDS and EOS_QUERIES add the routes described below)
// With a mesh, a TriangleMesh read from the OBJ file is added after the
// scene's own triangles, the way the reference's driver adds its OBJ model
// (src/main.cpp:386-403): a translation added with Scene::addTransformation
// (identity when tx ty tz are not given) and set on the mesh, material and
// transformation index set before loadFromFile, then Scene::addTriangleMesh.
// scene.bin: int32 n_tri, n_mat, n_light, width, height, depth; float32
// verts[n][9], normals[n][9]; int32 mat[n]; 64-byte materials; 32-byte lights;
// float32 camera position, view, right, up (12 floats).
#include <algorithm>
#include <cstdio>
#include <memory>
#include <cstdlib>
#include <string>
#include <vector>

#include <hip/hip_runtime_api.h>

#include "eos/eos.h"

template <class T>
static void rd(FILE* f, T* p, size_t n) {
    if (std::fread(p, sizeof(T), n, f) != n) {
        std::fprintf(stderr, "short read\n");
        std::exit(2);
    }
}

int main(int argc, char** argv) {
    if (argc != 4 && argc != 6 && argc != 9) {
        std::fprintf(stderr, "usage: %s scene.bin out_multi.rgba out_single.rgba [mesh.obj mesh_material [tx ty tz]]\n",
                     argv[0]);
        return 2;
    }
    FILE* f = std::fopen(argv[1], "rb");
    if (!f) return 2;
    int32_t hdr[6];
    rd(f, hdr, 6);
    const int n = hdr[0], nm = hdr[1], nl = hdr[2], W = hdr[3], H = hdr[4], depth = hdr[5];
    std::vector<float> v(9 * (size_t)n), nn(9 * (size_t)n), cam(12);
    std::vector<int32_t> mat((size_t)n);
    std::vector<eos_material> mats((size_t)nm);
    std::vector<eos_light> lights((size_t)nl);
    rd(f, v.data(), v.size());
    rd(f, nn.data(), nn.size());
    rd(f, mat.data(), mat.size());
    rd(f, mats.data(), mats.size());
    rd(f, lights.data(), lights.size());
    rd(f, cam.data(), cam.size());
    std::fclose(f);

    Scene scene;
    scene.setMaxTracedDepth(5);
    scene.useBvh(true);
    scene.setAASamples(1);
    std::vector<int> matIndex;
    for (const eos_material& m : mats) {
        Material mm(0.0f, eos::vec4(m.kd[0], m.kd[1], m.kd[2], m.kd[3]), eos::vec4(m.ks[0], m.ks[1], m.ks[2], m.ks[3]),
                    (int)m.shininess);
        mm.setTransparent(m.transparent != 0);
        mm.setRefractiveIndex(m.refractive_index);
        mm.setReflective(m.reflective != 0);
        mm.setReflectionIntensity(m.reflectivity);
        matIndex.push_back(scene.addMaterial(mm));
    }
    int ident = scene.addTransformation(eos::mat4(1.0f));
    std::vector<Triangle*> tris;
    for (int i = 0; i < n; ++i) {
        const float* p = &v[9 * (size_t)i];
        const float* q = &nn[9 * (size_t)i];
        Triangle* t = new Triangle(eos::vec3(p[0], p[1], p[2]), eos::vec3(p[3], p[4], p[5]), eos::vec3(p[6], p[7], p[8]),
                                   eos::vec3(q[0], q[1], q[2]), eos::vec3(q[3], q[4], q[5]), eos::vec3(q[6], q[7], q[8]));
        t->setMaterialIndex(matIndex[mat[i]]);
        t->setTransformationIndex(ident);
        scene.addSurface(t);
        tris.push_back(t);
    }
    TriangleMesh mesh;
    if (argc >= 6) {
        // glm::translate(glm::mat4(1.0f), t): the identity with t in column 3
        eos::mat4 meshTransformation(1.0f);
        if (argc == 9)
            for (int k = 0; k < 3; ++k) meshTransformation.m[3][k] = (float)std::atof(argv[6 + k]);
        const int meshTransformationIndex = scene.addTransformation(meshTransformation);
        mesh.setTransformation(meshTransformation);
        mesh.setTransformationIndex(meshTransformationIndex);
        mesh.setMaterialIndex(matIndex[std::atoi(argv[5])]);
        if (!mesh.loadFromFile(argv[4])) {
            std::fprintf(stderr, "cannot read %s: %s\n", argv[4], eos_last_error());
            return 2;
        }
        scene.addTriangleMesh(&mesh);
        std::printf("Triangle Mesh (%s) has %u Triangles.\n", argv[4], mesh.getNumTriangles());
    }
    std::vector<LightSource*> ls;
    for (const eos_light& l : lights) {
        ls.push_back(new LightSource(eos::vec4(l.position[0], l.position[1], l.position[2], l.position[3]),
                                     eos::vec4(l.color[0], l.color[1], l.color[2], l.color[3])));
        scene.addLightSource(ls.back());
    }
    scene.flush();

    Camera camera;
    camera.setPosition(eos::vec3(cam[0], cam[1], cam[2]));
    camera.setViewingDirection(eos::vec3(cam[3], cam[4], cam[5]));
    camera.setRightVector(eos::vec3(cam[6], cam[7], cam[8]));
    camera.setUpVector(eos::vec3(cam[9], cam[10], cam[11]));
    camera.setWidth(W);
    camera.setHeight(H);

    std::vector<unsigned char> image(4 * (size_t)W * H);
    {
        MultiDeviceRenderer renderer(&scene, &camera, W, H);
        renderer.setTracedDepth(depth);
        renderer.renderSceneToHostBuffer(image.data(), &camera);
    }
    FILE* o = std::fopen(argv[2], "wb");
    std::fwrite(image.data(), 1, image.size(), o);
    std::fclose(o);

    cudaPreferL1Cache();
    cudaScene_t* dscene = createCudaScene(&scene);
    DeviceCameraHandler handler;
    handler.setDeviceCamera(&camera);
    DeviceRenderer dr(handler.getDeviceCamera(), W, H);
    dr.setTracedDepth(depth);
    std::vector<unsigned char> image2(4 * (size_t)W * H);
    dr.renderCudaSceneToHostBufferMegaKernel(dscene, image2.data());
    o = std::fopen(argv[3], "wb");
    std::fwrite(image2.data(), 1, image2.size(), o);
    std::fclose(o);
    handler.freeDeviceCamera();
    destroyCudaScene(dscene);

    // the reference's CPU renderer API (src/main.cpp:367-371), here on the GPU:
    // EOS_AA sets its anti-aliasing samples, the image goes to <out_single>.rt
    RayTracer rayTracer;
    rayTracer.setTracedDepth(depth);
    if (const char* aa = std::getenv("EOS_AA")) rayTracer.setAASamples(std::atoi(aa));
    std::vector<unsigned char> image3(4 * (size_t)W * H);
    rayTracer.render(scene, camera, image3.data());
    o = std::fopen((std::string(argv[3]) + ".rt").c_str(), "wb");
    std::fwrite(image3.data(), 1, image3.size(), o);
    std::fclose(o);

    // the split-frame launcher (src/kernelWrapper.h:33) called once per band, as
    // MultiDeviceRenderer's per-GPU threads do (src/MultiDeviceRenderer.cu:122-163):
    // EOS_BANDS sets the band count; band k (H / n rows from row k * (H / n)) goes
    // to a device buffer of its own size, the stitched rows to <out_single>.bands
    if (const char* nb = std::getenv("EOS_BANDS")) {
        const int n = std::atoi(nb), rows = H / std::max(n, 1);
        cudaScene_t* bscene = createCudaScene(&scene);
        DeviceCameraHandler bcam;
        bcam.setDeviceCamera(&camera);
        std::vector<unsigned char> image4(4 * (size_t)W * rows * n);
        void* band = nullptr;
        if (hipMalloc(&band, 4 * (size_t)W * rows + 4) != hipSuccess) return 1;
        int blocks[2] = {(W + 15) / 16, (rows + 15) / 16}, threads[2] = {16, 16};
        for (int k = 0; k < n; ++k) {
            rayTrace_MultiMegaKernel(bscene, bcam.getDeviceCamera(), W, H, static_cast<uchar4*>(band), n, k, blocks,
                                     threads);
            gpuSynch();
            if (hipMemcpy(image4.data() + 4 * (size_t)W * rows * k, band, 4 * (size_t)W * rows,
                          hipMemcpyDeviceToHost) != hipSuccess)
                return 1;
        }
        (void)hipFree(band);
        bcam.freeDeviceCamera();
        destroyCudaScene(bscene);
        o = std::fopen((std::string(argv[3]) + ".bands").c_str(), "wb");
        std::fwrite(image4.data(), 1, image4.size(), o);
        std::fclose(o);
    }
    // the Scene's ray queries (src/Scene.h:69-71), one ray per call as the
    // reference's renderer makes them, then the batched forms: EOS_QUERIES names
    // a file of int32 n, float origins[n][3], directions[n][3]; per ray,
    // <file>.out gets int32 found, float point[3], normal[3], int32 material
    // for the BVH query, int32 occluded, then the same four for the linear query,
    // and <file>.batch the same from the batched calls
    if (const char* qf = std::getenv("EOS_QUERIES")) {
        FILE* q = std::fopen(qf, "rb");
        if (!q) return 2;
        int32_t nq = 0;
        rd(q, &nq, 1);
        std::vector<float> qo(3 * (size_t)nq), qd(3 * (size_t)nq);
        rd(q, qo.data(), qo.size());
        rd(q, qd.data(), qd.size());
        std::fclose(q);
        std::vector<Ray> rays;
        for (int i = 0; i < nq; ++i)
            rays.emplace_back(eos::vec3(qo[3 * i], qo[3 * i + 1], qo[3 * i + 2]),
                              eos::vec3(qd[3 * i], qd[3 * i + 1], qd[3 * i + 2]));
        auto put = [](FILE* f, bool found, const RayIntersection& x) {
            const int32_t fl = found ? 1 : 0, m = found ? x.getMaterialIndex() : -1;
            const float pn[6] = {x.getPoint().x, x.getPoint().y, x.getPoint().z,
                                 x.getNormal().x, x.getNormal().y, x.getNormal().z};
            std::fwrite(&fl, 4, 1, f);
            std::fwrite(pn, 4, 6, f);
            std::fwrite(&m, 4, 1, f);
        };
        FILE* out = std::fopen((std::string(qf) + ".out").c_str(), "wb");
        for (const Ray& r : rays) {
            RayIntersection x, y;
            const bool f1 = scene.findMinDistanceIntersectionBVH(r, x);
            put(out, f1, x);
            const int32_t occ = scene.shadowRayVisibilityBVH(r) ? 1 : 0;
            std::fwrite(&occ, 4, 1, out);
            const bool f2 = scene.findMinDistanceIntersectionLinear(r, y);
            put(out, f2, y);
        }
        std::fclose(out);
        std::vector<RayIntersection> xb((size_t)nq), yb((size_t)nq);
        std::unique_ptr<bool[]> fb(new bool[(size_t)nq]), ob(new bool[(size_t)nq]), lb(new bool[(size_t)nq]);
        scene.findMinDistanceIntersectionsBVH(rays.data(), nq, xb.data(), fb.get());
        scene.shadowRayVisibilitiesBVH(rays.data(), nq, ob.get());
        scene.findMinDistanceIntersectionsLinear(rays.data(), nq, yb.data(), lb.get());
        out = std::fopen((std::string(qf) + ".batch").c_str(), "wb");
        for (int i = 0; i < nq; ++i) {
            put(out, fb[i], xb[(size_t)i]);
            const int32_t occ = ob[i] ? 1 : 0;
            std::fwrite(&occ, 4, 1, out);
            put(out, lb[i], yb[(size_t)i]);
        }
        std::fclose(out);
    }
    for (Triangle* t : tris) delete t;
    for (LightSource* l : ls) delete l;
    return 0;
}
