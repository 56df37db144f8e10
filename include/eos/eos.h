/*
 * eos/eos.h -- the reference's host C++ API for the GPU render path, kept
 * intact on top of the C ABI (include/eos_c.h), so code written against the
 * reference compiles against this header instead:
 *
 *   Camera               src/Camera.h:30-64
 *   Material             src/Material.h:43-78
 *   LightSource          src/LightSource.h:27-39
 *   Surface, Triangle    src/surface.h:31-47, src/Triangle.h:32-99
 *   TriangleMesh         src/triangleMesh.h:30-80 (native OBJ reader instead of assimp)
 *   Scene                src/Scene.h:42-118 (with its ray queries, src/Scene.h:69-71)
 *   Ray                  src/Ray.h:31-65
 *   RayIntersection      src/RayIntersection.h:29-56
 *   cudaScene_t, createCudaScene            src/cudaStructures.h:87-98
 *   DeviceCameraHandler  src/DeviceCameraHandler.h:29-40
 *   DeviceRenderer       src/DeviceRenderer.h:33-56
 *   MultiDeviceRenderer  src/MultiDeviceRenderer.h:34-48 (one GPU, full frame)
 *   rayTrace_MegaKernel, rayTrace_MultiMegaKernel  src/kernelWrapper.h:30-34
 *   resetDevice, gpuSynch, cudaPreferL1Cache, setGLDevice  src/cudaWrapper.h:37-44
 *
 * Vectors: eos::vec3 / eos::vec4 / eos::mat4 hold plain floats and convert
 * from any type with the same members (glm's included).  Errors follow the
 * reference: a message on stderr and exit() with a non-zero code.  What the
 * render computes is the reference CPU RayTracer's image (shadow rays, Phong,
 * reflections up to the traced depth, default 4 as RAYTRACER_DEFAULT_TRACE_DEPTH,
 * src/RayTracer.h:36), not the reference megakernel's direct-lighting image.
 */
#ifndef EOS_EOS_H
#define EOS_EOS_H

#include <stdint.h>

#include <hip/hip_vector_types.h>   /* uchar4, the image type of the launchers below */

#include <memory>
#include <utility>
#include <vector>

#include "../eos_c.h"

namespace eos {

struct vec3 {
    float x = 0.0f, y = 0.0f, z = 0.0f;
    vec3() = default;
    explicit vec3(float s) : x(s), y(s), z(s) {}
    vec3(float x_, float y_, float z_) : x(x_), y(y_), z(z_) {}
    template <class V, class = decltype(V().x + V().y + V().z)>
    vec3(const V& v) : x((float)v.x), y((float)v.y), z((float)v.z) {}  // NOLINT: implicit from glm::vec3
};

struct vec4 {
    float x = 0.0f, y = 0.0f, z = 0.0f, w = 0.0f;
    vec4() = default;
    explicit vec4(float s) : x(s), y(s), z(s), w(s) {}
    vec4(float x_, float y_, float z_, float w_) : x(x_), y(y_), z(z_), w(w_) {}
    template <class V, class = decltype(V().x + V().y + V().z + V().w)>
    vec4(const V& v) : x((float)v.x), y((float)v.y), z((float)v.z), w((float)v.w) {}  // NOLINT
};

/* Column-major 4x4 (glm layout, m[column][row]): a triangle's local-to-world
 * transformation (Scene::addTransformation + Triangle::setTransformationIndex). */
struct mat4 {
    float m[4][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}};
    mat4() = default;
    explicit mat4(float d) {
        for (int c = 0; c < 4; ++c)
            for (int r = 0; r < 4; ++r) m[c][r] = c == r ? d : 0.0f;
    }
    /* from glm::mat4 (or any type indexed [column][row]) */
    template <class M, class = decltype(std::declval<const M&>()[0][0] + 0.0f)>
    mat4(const M& g) {  // NOLINT
        for (int c = 0; c < 4; ++c)
            for (int r = 0; r < 4; ++r) m[c][r] = (float)g[c][r];
    }
    bool is_identity() const;
};

}  // namespace eos

#define RAYTRACER_DEFAULT_TRACE_DEPTH 4

class Camera {
public:
    Camera() : m_Width(0), m_Height(0) {}
    void setPosition(const eos::vec3& p) { m_Position = p; }
    void setViewingDirection(const eos::vec3& d) { m_ViewingDirection = d; }
    void setRightVector(const eos::vec3& r) { m_RightVector = r; }
    void setUpVector(const eos::vec3& u) { m_UpVector = u; }
    void setWidth(int w) { m_Width = w; }
    void setHeight(int h) { m_Height = h; }
    const eos::vec3& getPosition() const { return m_Position; }
    const eos::vec3& getViewingDirection() const { return m_ViewingDirection; }
    const eos::vec3& getRightVector() const { return m_RightVector; }
    const eos::vec3& getUpVector() const { return m_UpVector; }
    int getWidth() const { return m_Width; }
    int getHeight() const { return m_Height; }

private:
    // Same member order and sizes as the reference (and eos_camera): 56 bytes.
    eos::vec3 m_Position, m_ViewingDirection, m_RightVector, m_UpVector;
    int m_Width, m_Height;
};

class Material {
public:
    Material() : m_AmbientIntensity(0.2f), m_DiffuseColor(1.0f), m_SpecularColor(1.0f), m_Shininess(128) {}
    Material(float ambient, eos::vec4 diffuse, eos::vec4 specular, int shininess)
        : m_AmbientIntensity(ambient), m_DiffuseColor(diffuse), m_SpecularColor(specular), m_Shininess(shininess) {}
    void setAmbientIntensity(float a) { m_AmbientIntensity = a; }
    void setDiffuseColor(eos::vec4 c) { m_DiffuseColor = c; }
    void setSpecularColor(eos::vec4 c) { m_SpecularColor = c; }
    void setReflective(bool r) { m_Reflective = r; }
    void setReflectionIntensity(float i) { m_ReflectionIntensity = i; }
    void setTransparent(bool t) { m_Transparent = t; }
    void setRefractiveIndex(float n) { m_RefractiveIndex = n; }
    float getAmbientIntensity() const { return m_AmbientIntensity; }
    const eos::vec4& getDiffuseColor() const { return m_DiffuseColor; }
    const eos::vec4& getSpecularColor() const { return m_SpecularColor; }
    int getShininess() const { return m_Shininess; }
    bool isReflective() const { return m_Reflective; }
    float getReflectiveIntensity() const { return m_ReflectionIntensity; }
    bool isTransparent() const { return m_Transparent; }
    float getRefractiveIndex() const { return m_RefractiveIndex; }

private:
    float m_AmbientIntensity;
    eos::vec4 m_DiffuseColor, m_SpecularColor;
    int m_Shininess;
    // The reference leaves these uninitialised (src/Material.h:46-49); here they start false / 0.
    bool m_Reflective = false;
    float m_ReflectionIntensity = 0.0f;
    bool m_Transparent = false;
    float m_RefractiveIndex = 1.0f;
};

class LightSource {
public:
    LightSource() : m_Position(0.0f), m_LightColor(1.0f) {}
    LightSource(const eos::vec4& position, const eos::vec4& color) : m_Position(position), m_LightColor(color) {}
    const eos::vec4& getPosition() const { return m_Position; }
    const eos::vec4& getLightColor() const { return m_LightColor; }

private:
    eos::vec4 m_Position, m_LightColor;
};

class Surface {
public:
    virtual ~Surface() = default;
    virtual int getMaterialIndex() = 0;
    virtual void setMaterialIndex(int index) = 0;
};

class Triangle : public Surface {
public:
    Triangle(const eos::vec3& v1, const eos::vec3& v2, const eos::vec3& v3, const eos::vec3& n1,
             const eos::vec3& n2, const eos::vec3& n3)
        : m_V1(v1), m_V2(v2), m_V3(v3), m_N1(n1), m_N2(n2), m_N3(n3) {}
    int getMaterialIndex() override { return m_MaterialIndex; }
    void setMaterialIndex(int index) override { m_MaterialIndex = index; }
    void setMaterial(Material& m) { m_Material = &m; }
    Material& getMaterial() { return *m_Material; }
    void setTransformationIndex(int index) { m_TransformationIndex = index; }
    int getTransformationIndex() const { return m_TransformationIndex; }
    const eos::vec3& getV1() { return m_V1; }
    const eos::vec3& getV2() { return m_V2; }
    const eos::vec3& getV3() { return m_V3; }
    const eos::vec3& getN1() { return m_N1; }
    const eos::vec3& getN2() { return m_N2; }
    const eos::vec3& getN3() { return m_N3; }

    eos::vec3 m_V1, m_V2, m_V3, m_N1, m_N2, m_N3;

private:
    Material* m_Material = nullptr;
    int m_MaterialIndex = 0;
    int m_TransformationIndex = -1;
};

/* A mesh read from a Wavefront OBJ file (eos_mesh_load_obj; the reference
 * reads through assimp, src/triangleMesh.cpp:79-160).  Like the reference, the
 * material index and transformation index set before loadFromFile are given to
 * every triangle it creates; the mesh owns its triangles. */
class TriangleMesh {
public:
    TriangleMesh() = default;
    ~TriangleMesh();
    TriangleMesh(const TriangleMesh&) = delete;
    TriangleMesh& operator=(const TriangleMesh&) = delete;
    bool loadFromFile(const char* filename);
    const eos::mat4& transformation() { return m_LocalToWorldTransformation; }
    void setTransformation(const eos::mat4& transformation) { m_LocalToWorldTransformation = transformation; }
    Material& getMaterial() { return m_Material; }
    void setMaterial(const Material& material) { m_Material = material; }
    unsigned int getNumTriangles() { return (unsigned int)m_Triangles.size(); }
    Triangle* getTriangle(int index) { return index >= 0 && index < (int)m_Triangles.size() ? m_Triangles[index] : nullptr; }
    void setSceneStartIndex(int start) { m_SceneIndexStart = start; }
    void setSceneEndIndex(int end) { m_SceneIndexEnd = end; }
    void setMaterialIndex(int index) { m_MaterialIndex = index; }
    int getMaterialIndex(int) const { return m_MaterialIndex; }
    void setTransformationIndex(int index) { m_TransformationIndex = index; }
    int getTransformationIndex() const { return m_TransformationIndex; }

private:
    std::vector<Triangle*> m_Triangles;
    eos::mat4 m_LocalToWorldTransformation;
    Material m_Material;
    int m_MaterialIndex = 0;
    int m_TransformationIndex = -1;
    int m_SceneIndexStart = 0, m_SceneIndexEnd = 0;
};

/* src/Ray.h:31-65: origin, direction, the inverse direction (double-precision
 * reciprocals in the constructor, float for x and y in setDirection) and the
 * signs of the inverse.  The GPU queries use the origin and the direction. */
class Ray {
public:
    Ray() : m_InvDirection(0.0f) {}
    Ray(const eos::vec3& origin, const eos::vec3& direction) : m_Origin(origin), m_Direction(direction) {
        m_InvDirection = eos::vec3((float)(1.0 / direction.x), (float)(1.0 / direction.y), (float)(1.0 / direction.z));
        set_signs();
    }
    void setOrigin(const eos::vec3& origin) { m_Origin = origin; }
    void setDirection(const eos::vec3& direction) {
        m_Direction = direction;
        m_InvDirection = eos::vec3(1.0f / direction.x, 1.0f / direction.y, (float)(1.0 / direction.z));
        set_signs();
    }
    const eos::vec3& getOrigin() const { return m_Origin; }
    const eos::vec3& getDirection() const { return m_Direction; }
    const eos::vec3& getInvDirection() const { return m_InvDirection; }

    int m_sign[3] = {0, 0, 0};

private:
    void set_signs() {
        m_sign[0] = m_InvDirection.x < 0.0f;
        m_sign[1] = m_InvDirection.y < 0.0f;
        m_sign[2] = m_InvDirection.z < 0.0f;
    }
    eos::vec3 m_Origin, m_Direction, m_InvDirection;
};

/* src/RayIntersection.h:29-56: the nearest hit's world-space point and normal,
 * its material and material index. */
class RayIntersection {
public:
    RayIntersection() : m_Point(0.0f), m_Normal(0.0f) {}
    RayIntersection(const eos::vec3& point, const eos::vec3& normal) : m_Point(point), m_Normal(normal) {}
    RayIntersection(const eos::vec3& point, const eos::vec3& normal, const Material& material)
        : m_Point(point), m_Normal(normal), m_Material(material) {}
    void setPoint(const eos::vec3& point) { m_Point = point; }
    void setNormal(const eos::vec3& normal) { m_Normal = normal; }
    void setMaterial(const Material& material) { m_Material = material; }
    void setMaterialIndex(int idx) { m_MaterialIndex = idx; }
    const eos::vec3& getPoint() const { return m_Point; }
    const eos::vec3& getNormal() const { return m_Normal; }
    const Material& getMaterial() const { return m_Material; }
    int getMaterialIndex() const { return m_MaterialIndex; }

private:
    eos::vec3 m_Point, m_Normal;
    Material m_Material;
    int m_MaterialIndex = -1;
};

struct cudaScene_t;

class Scene {
public:
    bool addSurface(Surface* surface);
    /* src/Scene.cpp:55-74: appends the mesh's triangles (the mesh keeps owning them). */
    bool addTriangleMesh(TriangleMesh* mesh);
    bool addLightSource(LightSource* light);
    int addMaterial(Material material);
    int addTransformation(const eos::mat4& trans);
    const Material& getMaterialAtIndex(int index) const { return m_Materials[index]; }
    int getNumMaterials() const { return (int)m_Materials.size(); }
    int getNumSurfaces() const { return (int)m_Surfaces.size(); }
    int getNumLightSources() const { return (int)m_Lights.size(); }
    int getNumTransformations() const { return (int)m_Transformations.size(); }
    Surface* getSurface(int id) const { return id < (int)m_Surfaces.size() ? m_Surfaces[id] : nullptr; }
    const LightSource* getLightSource(int id) const { return id < (int)m_Lights.size() ? m_Lights[id] : nullptr; }
    const eos::mat4& getTransformationAtIndex(int i) const { return m_Transformations[i]; }
    void setMaxTracedDepth(int depth) { m_MaxTracedDepth = depth; }
    int getMaxTracedDepth() const { return m_MaxTracedDepth; }
    void setAmbientRefractiveIndex(float n) { m_AmbientRefractiveIndex = n; }
    float getAmbientRefractiveIndex() const { return m_AmbientRefractiveIndex; }
    void setAASamples(int samples) { m_AASamples = samples; }
    int getAASamples() { return m_AASamples; }
    void useBvh(bool use) { m_UsingBvh = use; }
    bool isUsingBvh() const { return m_UsingBvh; }
    /* The reference builds its CPU BVH here (src/Scene.cpp:204-209).  Here the
     * surfaces are checked and the ray queries' device copy is dropped;
     * createCudaScene makes its own copy for rendering. */
    void flush();

    /* The reference's per-ray queries (src/Scene.h:69-71, src/Scene.cpp:145-201),
     * answered on the GPU (eos_trace_rays) by a device copy of the scene with its
     * LBVH, made at the first query after flush().  Each call is one launch and
     * one synchronisation; the batched forms below take many rays at once. */
    bool findMinDistanceIntersectionBVH(const Ray& ray, RayIntersection& intersection) const;
    bool shadowRayVisibilityBVH(const Ray& ray) const;
    /* Every triangle in the order they were added (the reference walks its
     * surfaces in the order its BVH build left them, which differs only when
     * two triangles give the same least distance). */
    bool findMinDistanceIntersectionLinear(const Ray& ray, RayIntersection& intersection) const;
    /* Not in the reference: n rays in one launch; found[i] / occluded[i] are the
     * per-ray queries' return values. */
    void findMinDistanceIntersectionsBVH(const Ray* rays, int n, RayIntersection* intersections, bool* found) const;
    void shadowRayVisibilitiesBVH(const Ray* rays, int n, bool* occluded) const;
    void findMinDistanceIntersectionsLinear(const Ray* rays, int n, RayIntersection* intersections,
                                            bool* found) const;

private:
    void query(int mode, const Ray* rays, int n, RayIntersection* intersections, bool* flags) const;
    cudaScene_t* device() const;

    std::vector<Surface*> m_Surfaces;
    std::vector<LightSource*> m_Lights;
    std::vector<Material> m_Materials;
    std::vector<eos::mat4> m_Transformations;
    int m_MaxTracedDepth = 0;
    float m_AmbientRefractiveIndex = 1.0f;
    bool m_UsingBvh = false;
    int m_AASamples = 1;
    mutable std::shared_ptr<cudaScene_t> m_Device;   // the queries' device copy (copies of the Scene share it)
};

/* Opaque device scene (the reference returns a device pointer that must not be
 * dereferenced on the host; this one is a host handle with the same rule). */
cudaScene_t* createCudaScene(Scene* h_scene);
/* Not in the reference (it never frees the device scene). */
void destroyCudaScene(cudaScene_t* scene);
/* The eos_scene behind a cudaScene_t, for mixing the C ABI in. */
eos_scene* eosSceneOf(cudaScene_t* scene);

class DeviceCameraHandler {
public:
    DeviceCameraHandler() : m_DeviceCamera(nullptr) {}
    explicit DeviceCameraHandler(Camera* h_Camera);
    void setDeviceCamera(Camera* h_Camera);
    Camera* getDeviceCamera();          // device memory
    void updateDeviceCamera(Camera* h_Camera);
    void freeDeviceCamera();

private:
    Camera* m_DeviceCamera;
};

class DeviceRenderer {
public:
    DeviceRenderer(Camera* d_camera, int width, int height)
        : m_Camera(d_camera), m_Width(width), m_Height(height) {}
    void renderCudaSceneToHostBufferMegaKernel(cudaScene_t* deviceScene, void* imageBuffer);
    void renderCudaSceneToHostBufferWarpShuffleMegaKernel(cudaScene_t* deviceScene, void* imageBuffer);
    void setCamera(Camera* d_camera) { m_Camera = d_camera; }
    int getWidth() const { return m_Width; }
    int getHeight() const { return m_Height; }
    Camera* getDeviceCamera() const { return m_Camera; }
    /* Reflection bounces (the CPU RayTracer's setTracedDepth); default 4. */
    void setTracedDepth(int depth) { if (depth > 0) m_Depth = depth; }
    int getTracedDepth() const { return m_Depth; }
    /* n x n anti-aliasing samples per pixel (the CPU RayTracer's setAASamples,
     * src/RayTracer.cpp:33-36,176-266); 0, the default, renders one ray per pixel. */
    void setAASamples(int samples) { if (samples > 0) m_AASamples = samples; }
    int getAASamples() const { return m_AASamples; }
    float getLastRenderMilliseconds() const { return m_LastMs; }

private:
    Camera* m_Camera;
    int m_Width, m_Height;
    int m_Depth = RAYTRACER_DEFAULT_TRACE_DEPTH;
    int m_AASamples = 0;
    float m_LastMs = 0.0f;
};

/* src/RayTracer.h:29-67: the reference's CPU renderer API.  render() gives the
 * image the reference's RayTracer::render computes (renderWithoutAA when the
 * sample count is RAYTRACER_NO_AA, else renderWithAA), rendered on the GPU:
 * the scene is uploaded and its LBVH built on each call, then eos_render runs
 * with the traced depth and anti-aliasing samples set here. */
#define RAYTRACER_AA_SAMPLES_8 8
#define RAYTRACER_AA_SAMPLES_4 4
#define RAYTRACER_AA_SAMPLES_2 2
#define RAYTRACER_NO_AA 0
class RayTracer {
public:
    RayTracer() : m_AASamples(RAYTRACER_NO_AA), m_TraceDepth(RAYTRACER_DEFAULT_TRACE_DEPTH) {}
    void render(const Scene& scene, const Camera& camera, unsigned char* outputImage);
    int getAASamples() const { return m_AASamples; }
    void setAASamples(int samples) { if (samples > 0) m_AASamples = samples; }   // src/RayTracer.cpp:33-36
    int getTracedDepth() const { return m_TraceDepth; }
    void setTracedDepth(int depth) { if (depth > 0) m_TraceDepth = depth; }      // src/RayTracer.cpp:47-50

private:
    int m_AASamples;
    int m_TraceDepth;
};

class MultiDeviceRenderer {
public:
    MultiDeviceRenderer(Scene* hostScene, Camera* hostCamera, int width, int height);
    ~MultiDeviceRenderer();
    void renderSceneToHostBuffer(void* imageBuffer, Camera* hostCamera);
    void setTracedDepth(int depth) { if (depth > 0) m_Depth = depth; }

private:
    cudaScene_t* m_Scene;
    int m_Width, m_Height;
    int m_Depth = RAYTRACER_DEFAULT_TRACE_DEPTH;
};

extern "C" {
/* src/kernelWrapper.h:30-34; imageBuffer is device memory; launch shapes are ignored. */
void rayTrace_MegaKernel(cudaScene_t* deviceScene, Camera* camera, int width, int height, uchar4* imageBuffer,
                         int blockdim[], int tpblock[]);
void rayTrace_WarpShuffle_MegaKernel(cudaScene_t* deviceScene, Camera* camera, int width, int height,
                                     uchar4* imageBuffer, int blockdim[], int tpblock);
void rayTrace_MultiMegaKernel(cudaScene_t* deviceScene, Camera* camera, int width, int height, uchar4* imageBuffer,
                              int numGPUS, int gpuID, int blockdimp[], int tpblock[]);
/* src/cudaWrapper.h:37-44 */
void resetDevice(void);
void setGLDevice(int dev_id);
void gpuSynch(void);
void cudaPreferL1Cache(void);
}

#endif
