// eos_shim.cpp -- the reference's host C++ API (include/eos/eos.h) over the C ABI.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/eos/eos.h"

static_assert(sizeof(Camera) == sizeof(eos_camera), "Camera must keep the reference layout (56 bytes)");

namespace {

bool quiet() {
    const char* q = std::getenv("EOS_QUIET");
    return q && q[0] && q[0] != '0';
}

// The reference's error convention (src/cudaQualifiers.h:35-45): print, exit.
[[noreturn]] void die(const char* what, int code = 1) {
    std::fprintf(stderr, "Eos Error: %s: %s\n", what, eos_last_error());
    std::exit(code);
}

void check(eos_status st, const char* what) {
    if (st != EOS_OK) die(what, (int)st);
}

void check_hip(hipError_t e, const char* what) {
    if (e != hipSuccess) {
        std::fprintf(stderr, "GPUassert: %s %s\n", hipGetErrorString(e), what);
        std::exit((int)e);
    }
}

eos_camera to_c(const Camera& c, int width, int height) {
    eos_camera o;
    std::memcpy(&o, &c, sizeof o);
    o.width = width;
    o.height = height;
    return o;
}

Camera download_camera(const Camera* d_camera) {
    Camera h;
    check_hip(hipMemcpy(&h, d_camera, sizeof(Camera), hipMemcpyDeviceToHost), "camera download");
    return h;
}

}  // namespace

bool eos::mat4::is_identity() const {
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r)
            if (m[c][r] != (c == r ? 1.0f : 0.0f)) return false;
    return true;
}

struct cudaScene_t {
    eos_scene* scene = nullptr;
};

// ---------------------------------------------------------------- Scene
bool Scene::addSurface(Surface* s) {
    if (!s) return false;
    m_Surfaces.push_back(s);
    return true;
}

bool Scene::addTriangleMesh(TriangleMesh* mesh) {
    if (!mesh) return false;
    const int start = getNumSurfaces();
    for (unsigned int i = 0; i < mesh->getNumTriangles(); ++i) {
        Triangle* t = mesh->getTriangle((int)i);
        if (t) m_Surfaces.push_back(t);
    }
    mesh->setSceneStartIndex(start);
    mesh->setSceneEndIndex(getNumSurfaces() - start);
    return true;
}

// ---------------------------------------------------------------- TriangleMesh
TriangleMesh::~TriangleMesh() {
    for (Triangle* t : m_Triangles) delete t;
}

bool TriangleMesh::loadFromFile(const char* filename) {
    eos_mesh m;
    if (eos_mesh_load_obj(filename, &m) != EOS_OK) return false;
    for (Triangle* t : m_Triangles) delete t;
    m_Triangles.clear();
    m_Triangles.reserve((size_t)m.num_triangles);
    for (int i = 0; i < m.num_triangles; ++i) {
        const float* v = m.vertices + 9 * (size_t)i;
        const float* n = m.normals + 9 * (size_t)i;
        Triangle* t = new Triangle(eos::vec3(v[0], v[1], v[2]), eos::vec3(v[3], v[4], v[5]), eos::vec3(v[6], v[7], v[8]),
                                   eos::vec3(n[0], n[1], n[2]), eos::vec3(n[3], n[4], n[5]), eos::vec3(n[6], n[7], n[8]));
        t->setMaterial(m_Material);
        t->setMaterialIndex(m_MaterialIndex);
        t->setTransformationIndex(m_TransformationIndex);
        m_Triangles.push_back(t);
    }
    eos_mesh_free(&m);
    return true;
}

bool Scene::addLightSource(LightSource* l) {
    if (!l) return false;
    m_Lights.push_back(l);
    return true;
}

int Scene::addMaterial(Material m) {
    m_Materials.push_back(m);
    return (int)m_Materials.size() - 1;
}

int Scene::addTransformation(const eos::mat4& t) {
    m_Transformations.push_back(t);
    return (int)m_Transformations.size() - 1;
}

namespace {
cudaScene_t* make_device_scene(const Scene* h);
}

void Scene::flush() {
    for (Surface* s : m_Surfaces)
        if (!dynamic_cast<Triangle*>(s)) {
            std::fprintf(stderr, "Eos Error: only triangle surfaces are supported on the GPU path\n");
            std::exit(1);
        }
    m_Device.reset();   // the queries copy the flushed scene to the device when first asked
}

cudaScene_t* Scene::device() const {
    if (!m_Device) m_Device.reset(make_device_scene(this), destroyCudaScene);
    return m_Device.get();
}

// One eos_trace_rays call for n rays; results into the reference's types.
void Scene::query(int mode, const Ray* rays, int n, RayIntersection* out, bool* flags) const {
    if (n <= 0) return;
    if (!rays || !flags || (mode != EOS_QUERY_OCCLUDED && !out)) {
        std::fprintf(stderr, "Eos Error: Scene ray query: null argument\n");
        std::exit(1);
    }
    std::vector<float> o(3 * (size_t)n), d(3 * (size_t)n);
    for (int i = 0; i < n; ++i) {
        const eos::vec3 &ro = rays[i].getOrigin(), &rd = rays[i].getDirection();
        o[3 * (size_t)i] = ro.x; o[3 * (size_t)i + 1] = ro.y; o[3 * (size_t)i + 2] = ro.z;
        d[3 * (size_t)i] = rd.x; d[3 * (size_t)i + 1] = rd.y; d[3 * (size_t)i + 2] = rd.z;
    }
    eos_scene* es = eosSceneOf(device());
    const eos_trace_opts opts{mode, EOS_MEM_HOST};
    if (mode == EOS_QUERY_OCCLUDED) {
        std::vector<uint8_t> occ((size_t)n);
        check(eos_trace_rays(es, n, o.data(), d.data(), &opts, nullptr, occ.data()), "Scene::shadowRayVisibilityBVH");
        for (int i = 0; i < n; ++i) flags[i] = occ[(size_t)i] != 0;
        return;
    }
    std::vector<eos_ray_hit> hits((size_t)n);
    check(eos_trace_rays(es, n, o.data(), d.data(), &opts, hits.data(), nullptr), "Scene ray query");
    for (int i = 0; i < n; ++i) {
        const eos_ray_hit& h = hits[(size_t)i];
        flags[i] = h.hit != 0;
        if (h.triangle < 0) continue;   // the reference leaves the intersection as it was
        out[i].setPoint(eos::vec3(h.point[0], h.point[1], h.point[2]));
        out[i].setNormal(eos::vec3(h.normal[0], h.normal[1], h.normal[2]));
        out[i].setMaterial(m_Materials[(size_t)h.material]);
        out[i].setMaterialIndex(h.material);
    }
}

bool Scene::findMinDistanceIntersectionBVH(const Ray& ray, RayIntersection& x) const {
    bool f = false;
    query(EOS_QUERY_NEAREST, &ray, 1, &x, &f);
    return f;
}

bool Scene::shadowRayVisibilityBVH(const Ray& ray) const {
    bool f = false;
    query(EOS_QUERY_OCCLUDED, &ray, 1, nullptr, &f);
    return f;
}

bool Scene::findMinDistanceIntersectionLinear(const Ray& ray, RayIntersection& x) const {
    bool f = false;
    query(EOS_QUERY_NEAREST_LINEAR, &ray, 1, &x, &f);
    return f;
}

void Scene::findMinDistanceIntersectionsBVH(const Ray* rays, int n, RayIntersection* x, bool* found) const {
    query(EOS_QUERY_NEAREST, rays, n, x, found);
}

void Scene::shadowRayVisibilitiesBVH(const Ray* rays, int n, bool* occluded) const {
    query(EOS_QUERY_OCCLUDED, rays, n, nullptr, occluded);
}

void Scene::findMinDistanceIntersectionsLinear(const Ray* rays, int n, RayIntersection* x, bool* found) const {
    query(EOS_QUERY_NEAREST_LINEAR, rays, n, x, found);
}

cudaScene_t* createCudaScene(Scene* h) {
    if (!h) return nullptr;
    if (!h->isUsingBvh()) return nullptr;   // src/cudaStructures.cu:500-505
    return make_device_scene(h);
}

namespace {
// The scene's triangles, materials, lights and transformations copied to the
// device (eos_create_scene, eos_set_transforms) and its LBVH built.
cudaScene_t* make_device_scene(const Scene* h) {
    const int n = h->getNumSurfaces();
    std::vector<float> verts(9 * (size_t)n), normals(9 * (size_t)n);
    std::vector<int32_t> mat((size_t)n), xform((size_t)n, -1);
    bool any_xform = false;
    for (int i = 0; i < n; ++i) {
        Triangle* t = dynamic_cast<Triangle*>(h->getSurface(i));
        if (!t) {   // src/cudaStructures.cu:340-347
            std::fprintf(stderr, "Eos Error: surface %d is not a triangle; only triangles run on the GPU\n", i);
            std::exit(1);
        }
        // the scene's transformation list, as copyTransformations / copyTriangles
        // hand it to the device (src/cudaStructures.cu:220-262,359); identities
        // are dropped so scenes without real transformations keep the plain walk
        const int ti = t->getTransformationIndex();
        if (ti >= h->getNumTransformations()) {
            std::fprintf(stderr, "Eos Error: triangle %d has transformation index %d of %d\n", i, ti,
                         h->getNumTransformations());
            std::exit(1);
        }
        if (ti >= 0 && !h->getTransformationAtIndex(ti).is_identity()) {
            xform[i] = ti;
            any_xform = true;
        }
        const eos::vec3* v[6] = {&t->m_V1, &t->m_V2, &t->m_V3, &t->m_N1, &t->m_N2, &t->m_N3};
        for (int k = 0; k < 3; ++k) {
            verts[9 * (size_t)i + 3 * k + 0] = v[k]->x;
            verts[9 * (size_t)i + 3 * k + 1] = v[k]->y;
            verts[9 * (size_t)i + 3 * k + 2] = v[k]->z;
            normals[9 * (size_t)i + 3 * k + 0] = v[3 + k]->x;
            normals[9 * (size_t)i + 3 * k + 1] = v[3 + k]->y;
            normals[9 * (size_t)i + 3 * k + 2] = v[3 + k]->z;
        }
        mat[i] = t->getMaterialIndex();
    }
    std::vector<eos_material> mats((size_t)h->getNumMaterials());
    for (int i = 0; i < h->getNumMaterials(); ++i) {
        const Material& m = h->getMaterialAtIndex(i);
        const eos::vec4 &kd = m.getDiffuseColor(), &ks = m.getSpecularColor();
        eos_material& o = mats[i];
        o.kd[0] = kd.x; o.kd[1] = kd.y; o.kd[2] = kd.z; o.kd[3] = kd.w;
        o.ks[0] = ks.x; o.ks[1] = ks.y; o.ks[2] = ks.z; o.ks[3] = ks.w;
        o.shininess = (float)m.getShininess();
        o.reflectivity = m.getReflectiveIntensity();
        o.reflective = m.isReflective() ? 1 : 0;
        o.transparent = m.isTransparent() ? 1 : 0;
        o.refractive_index = m.getRefractiveIndex();
        o.pad[0] = o.pad[1] = o.pad[2] = 0;
    }
    std::vector<eos_light> lights((size_t)h->getNumLightSources());
    for (int i = 0; i < h->getNumLightSources(); ++i) {
        const LightSource* l = h->getLightSource(i);
        const eos::vec4 &p = l->getPosition(), &c = l->getLightColor();
        lights[i] = eos_light{{p.x, p.y, p.z, p.w}, {c.x, c.y, c.z, c.w}};
    }
    eos_scene_desc d;
    d.num_triangles = n;
    d.vertices = verts.data();
    d.normals = normals.data();
    d.material_index = mat.data();
    d.triangle_memory = EOS_MEM_HOST;
    d.num_materials = (int32_t)mats.size();
    d.materials = mats.data();
    d.num_lights = (int32_t)lights.size();
    d.lights = lights.data();
    cudaScene_t* out = new cudaScene_t();
    check(eos_create_scene(&d, &out->scene), "createCudaScene");
    check(eos_set_ambient_refractive_index(out->scene, h->getAmbientRefractiveIndex()), "createCudaScene (ambient index)");
    if (any_xform) {
        std::vector<float> m(16 * (size_t)h->getNumTransformations());
        for (int k = 0; k < h->getNumTransformations(); ++k)
            std::memcpy(&m[16 * (size_t)k], &h->getTransformationAtIndex(k).m[0][0], 16 * sizeof(float));
        check(eos_set_transforms(out->scene, h->getNumTransformations(), m.data(), xform.data(), EOS_MEM_HOST),
              "createCudaScene (transformations)");
    }
    check(eos_build_lbvh(out->scene, 4, nullptr), "createCudaScene (LBVH build)");
    return out;
}
}  // namespace

void destroyCudaScene(cudaScene_t* s) {
    if (!s) return;
    eos_destroy_scene(s->scene);
    delete s;
}

eos_scene* eosSceneOf(cudaScene_t* s) { return s ? s->scene : nullptr; }

// ---------------------------------------------------------------- camera handler
DeviceCameraHandler::DeviceCameraHandler(Camera* h) : m_DeviceCamera(nullptr) { setDeviceCamera(h); }

void DeviceCameraHandler::setDeviceCamera(Camera* h) {
    if (!m_DeviceCamera) check_hip(hipMalloc((void**)&m_DeviceCamera, sizeof(Camera)), "camera alloc");
    check_hip(hipMemcpy(m_DeviceCamera, h, sizeof(Camera), hipMemcpyHostToDevice), "camera upload");
}

Camera* DeviceCameraHandler::getDeviceCamera() { return m_DeviceCamera; }

void DeviceCameraHandler::updateDeviceCamera(Camera* h) {
    check_hip(hipMemcpy(m_DeviceCamera, h, sizeof(Camera), hipMemcpyHostToDevice), "camera upload");
}

void DeviceCameraHandler::freeDeviceCamera() {
    if (m_DeviceCamera) check_hip(hipFree(m_DeviceCamera), "camera free");
    m_DeviceCamera = nullptr;
}

// ---------------------------------------------------------------- renderers
void DeviceRenderer::renderCudaSceneToHostBufferMegaKernel(cudaScene_t* scene, void* image) {
    if (!scene || !image || !m_Camera) {
        std::fprintf(stderr, "Eos Error: DeviceRenderer: null scene, image or camera\n");
        std::exit(1);
    }
    const eos_camera cam = to_c(download_camera(m_Camera), m_Width, m_Height);
    eos_render_opts o{};
    o.depth = m_Depth;
    o.output_memory = EOS_MEM_HOST;
    o.aa_samples = m_AASamples;
    eos_render_stats st;
    check(eos_render(scene->scene, &cam, &o, static_cast<uint8_t*>(image), &st), "render");
    m_LastMs = (float)st.render_ms;
    if (!quiet()) std::printf("Cuda Mega Kernel finished in %f milliseconds\n", m_LastMs);   // src/DeviceRenderer.cu:105
}

void DeviceRenderer::renderCudaSceneToHostBufferWarpShuffleMegaKernel(cudaScene_t* scene, void* image) {
    renderCudaSceneToHostBufferMegaKernel(scene, image);
}

void RayTracer::render(const Scene& scene, const Camera& camera, unsigned char* outputImage) {
    if (!outputImage) {
        std::fprintf(stderr, "Eos Error: RayTracer::render: null image\n");
        std::exit(1);
    }
    // the scene is read, not changed (createCudaScene takes the reference's non-const pointer)
    cudaScene_t* d = createCudaScene(const_cast<Scene*>(&scene));
    if (!d) {
        std::fprintf(stderr, "Eos Error: RayTracer::render: the scene does not use a BVH (Scene::useBvh(true))\n");
        std::exit(1);
    }
    const eos_camera cam = to_c(camera, camera.getWidth(), camera.getHeight());
    eos_render_opts o{};
    o.depth = m_TraceDepth;
    o.output_memory = EOS_MEM_HOST;
    o.aa_samples = m_AASamples;
    check(eos_render(d->scene, &cam, &o, outputImage, nullptr), "RayTracer::render");
    destroyCudaScene(d);
}

MultiDeviceRenderer::MultiDeviceRenderer(Scene* hostScene, Camera* /*hostCamera*/, int width, int height)
    : m_Scene(nullptr), m_Width(width), m_Height(height) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1) {   // src/MultiDeviceRenderer.cu:38-42
        std::fprintf(stderr, "Eos Error:No AMD gpus found\n");
        std::exit(1);
    }
    if (!quiet()) {
        std::printf("number of host CPUs:\t%u\n", std::thread::hardware_concurrency());
        std::printf("number of HIP devices:\t%d (rendering the full frame on device 0)\n", n);
        hipDeviceProp_t p;
        if (hipGetDeviceProperties(&p, 0) == hipSuccess) std::printf("   0: %s\n", p.name);
        std::printf("---------------------------\n");
    }
    check_hip(hipSetDevice(0), "hipSetDevice");
    m_Scene = createCudaScene(hostScene);
    if (!m_Scene) {
        std::fprintf(stderr, "Eos Error: the scene does not use a BVH (Scene::useBvh(true) is required)\n");
        std::exit(1);
    }
}

MultiDeviceRenderer::~MultiDeviceRenderer() { destroyCudaScene(m_Scene); }

void MultiDeviceRenderer::renderSceneToHostBuffer(void* image, Camera* hostCamera) {
    const eos_camera cam = to_c(*hostCamera, m_Width, m_Height);
    eos_render_opts o{};
    o.depth = m_Depth;
    o.output_memory = EOS_MEM_HOST;
    check(eos_render(m_Scene->scene, &cam, &o, static_cast<uint8_t*>(image), nullptr), "render");
}

// ---------------------------------------------------------------- C launchers
extern "C" {

void rayTrace_MegaKernel(cudaScene_t* scene, Camera* camera, int width, int height, uchar4* image, int*, int*) {
    const eos_camera cam = to_c(download_camera(camera), width, height);
    eos_render_opts o{};
    o.depth = RAYTRACER_DEFAULT_TRACE_DEPTH;
    o.output_memory = EOS_MEM_DEVICE;
    check(eos_render(scene->scene, &cam, &o, reinterpret_cast<uint8_t*>(image), nullptr), "rayTrace_MegaKernel");
}

void rayTrace_WarpShuffle_MegaKernel(cudaScene_t* scene, Camera* camera, int width, int height, uchar4* image,
                                     int* blockdim, int) {
    rayTrace_MegaKernel(scene, camera, width, height, image, blockdim, nullptr);
}

// Renders the band of rows [gpuID*H/numGPUS, (gpuID+1)*H/numGPUS) into image
// (src/MultiDeviceRenderer.cu:122-163): only that band's primary rays are
// generated, so each device traces its own share of the frame.
void rayTrace_MultiMegaKernel(cudaScene_t* scene, Camera* camera, int width, int height, uchar4* image,
                              int numGPUS, int gpuID, int*, int*) {
    if (numGPUS <= 1) {
        rayTrace_MegaKernel(scene, camera, width, height, image, nullptr, nullptr);
        return;
    }
    // this GPU's band of rows only (src/kernelInvocations.cu:56: row j of the
    // band is row j + (height / numGPUS) * gpuID of the frame)
    const int rows = height / numGPUS;
    if (rows <= 0 || gpuID < 0 || gpuID >= numGPUS) return;  // an empty band, as the reference's launch of 0 rows
    const eos_camera cam = to_c(download_camera(camera), width, height);
    eos_render_opts o{};
    o.depth = RAYTRACER_DEFAULT_TRACE_DEPTH;
    o.output_memory = EOS_MEM_DEVICE;
    check(eos_render_rows(scene->scene, &cam, rows * gpuID, rows * (gpuID + 1), &o, reinterpret_cast<uint8_t*>(image),
                          nullptr),
          "rayTrace_MultiMegaKernel");
}

void resetDevice(void) { check_hip(hipDeviceReset(), "hipDeviceReset"); }
void setGLDevice(int dev_id) { check_hip(hipSetDevice(dev_id), "hipSetDevice"); }
void gpuSynch(void) { check_hip(hipDeviceSynchronize(), "hipDeviceSynchronize"); }
// CDNA has no L1/LDS carve-out preference to set; kept so callers link unchanged.
void cudaPreferL1Cache(void) {}

}  // extern "C"
