// prr.hpp — C++ drop-in for the reference renderer's scene / render API.
//
// The classes keep the reference's names, fields and call pattern
// (main.cpp:11-52, scenes/museum_final.cpp) so a scene script compiles
// against this header and renders on one MI355X through libprr.so:
//
//   Scene scene;                       scene.Add(new MeshTriangle(tris, bounds, mat));
//   scene.camera = Camera::create(...); scene.camera.init(W, H);
//   Integrator integrator(scene, maxDepth, rr);
//   scene.buildBVH();                  // host SAH build + upload (Scene.cpp:15-27)
//   Renderer().Render(scene, integrator, settings);   // Renderer.hpp:57-59
//
// Reference interfaces mirrored (file:line):
//   Vector3f / Vector2f Vector.hpp:7-123, Bounds3 Bounds3.hpp:7-126, Ray Ray.hpp:8-28,
//   Intersection Intersection.hpp:12-34, Object Object.hpp:14-35,
//   Triangle / MeshTriangle Triangle.hpp:13-180, Sphere Sphere.hpp:11-66,
//   Material + subclasses Material.hpp:35-183, Camera Camera.hpp:8-100,
//   EnvMap EnvMap.hpp:15-224, PointLight PointLight.hpp:9-16, Scene Scene.hpp:15-41,
//   Integrator Integrator.hpp:15-45, RenderSettings / Renderer Renderer.hpp:42-59,
//   tonemap::ToneMapper ToneMapping.hpp:32-37 (+ ACES).
//
//   BVHAccel BVH.hpp:86-96 (Intersect / IntersectP), Photon / PhotonMap
//   Photon.hpp:11-47, Scene::sampleLight / pdfLight Scene.hpp:39-40.
//
// Differences a caller sees: the RNG is PCG32 per (pixel, sample) instead of
// a thread-local SplitMix64 (Scene::sampleLight keeps a thread-local
// SplitMix64 of its own), and there is no CPU fallback: without a HIP device
// the calls throw std::runtime_error. The reference's headers are forwarded
// under include/prr/reference/ (Scene.hpp, Renderer.hpp, ...), so a scene
// script that includes them compiles with -I include/prr/reference.
#pragma once

#include <cmath>
#include <cstdint>
#include <limits>
#include <string>
#include <vector>

// ---------------------------------------------------------------------------
class Vector3f {
  public:
    float x, y, z;
    Vector3f() : x(0), y(0), z(0) {}
    Vector3f(float xx) : x(xx), y(xx), z(xx) {}
    Vector3f(float xx, float yy, float zz) : x(xx), y(yy), z(zz) {}
    Vector3f operator*(const float &r) const { return Vector3f(x * r, y * r, z * r); }
    Vector3f operator/(const float &r) const { return Vector3f(x / r, y / r, z / r); }
    float norm() const { return std::sqrt(x * x + y * y + z * z); }
    Vector3f normalized() const {
        float n = std::sqrt(x * x + y * y + z * z);
        return Vector3f(x / n, y / n, z / n);
    }
    Vector3f operator*(const Vector3f &v) const { return Vector3f(x * v.x, y * v.y, z * v.z); }
    Vector3f operator-(const Vector3f &v) const { return Vector3f(x - v.x, y - v.y, z - v.z); }
    Vector3f operator+(const Vector3f &v) const { return Vector3f(x + v.x, y + v.y, z + v.z); }
    Vector3f operator-() const { return Vector3f(-x, -y, -z); }
    Vector3f &operator+=(const Vector3f &v) {
        x += v.x, y += v.y, z += v.z;
        return *this;
    }
    Vector3f &operator*=(const Vector3f &v) {
        x *= v.x, y *= v.y, z *= v.z;
        return *this;
    }
    Vector3f &operator*=(const float &r) {
        x *= r, y *= r, z *= r;
        return *this;
    }
    Vector3f operator/(const Vector3f &v) const { return Vector3f(x / v.x, y / v.y, z / v.z); }
    friend Vector3f operator*(const float &r, const Vector3f &v) { return Vector3f(v.x * r, v.y * r, v.z * r); }
    double operator[](int i) const { return (&x)[i]; }
    float &operator[](int i) { return (&x)[i]; }
    static Vector3f Min(const Vector3f &a, const Vector3f &b) {
        return Vector3f(std::fmin(a.x, b.x), std::fmin(a.y, b.y), std::fmin(a.z, b.z));
    }
    static Vector3f Max(const Vector3f &a, const Vector3f &b) {
        return Vector3f(std::fmax(a.x, b.x), std::fmax(a.y, b.y), std::fmax(a.z, b.z));
    }
};

class Vector2f {
  public:
    float x, y;
    Vector2f() : x(0), y(0) {}
    Vector2f(float xx) : x(xx), y(xx) {}
    Vector2f(float xx, float yy) : x(xx), y(yy) {}
    Vector2f operator*(const float &r) const { return Vector2f(x * r, y * r); }
    Vector2f operator+(const Vector2f &v) const { return Vector2f(x + v.x, y + v.y); }
};

inline Vector3f normalize(const Vector3f &v) {
    float m2 = v.x * v.x + v.y * v.y + v.z * v.z;
    if (m2 > 0) {
        float inv = 1 / std::sqrt(m2);
        return Vector3f(v.x * inv, v.y * inv, v.z * inv);
    }
    return v;
}
inline float dotProduct(const Vector3f &a, const Vector3f &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline Vector3f crossProduct(const Vector3f &a, const Vector3f &b) {
    return Vector3f(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
inline Vector3f lerp(const Vector3f &a, const Vector3f &b, const float &t) { return a * (1 - t) + b * t; }

class Bounds3 {
  public:
    Vector3f pMin, pMax;
    Bounds3()
        : pMin(std::numeric_limits<float>::infinity()), pMax(-std::numeric_limits<float>::infinity()) {}
    Bounds3(const Vector3f p) : pMin(p), pMax(p) {}
    Bounds3(const Vector3f p1, const Vector3f p2) : pMin(Vector3f::Min(p1, p2)), pMax(Vector3f::Max(p1, p2)) {}
    Vector3f Diagonal() const { return pMax - pMin; }
    Vector3f Centroid() const { return 0.5f * pMin + 0.5f * pMax; }
};
inline Bounds3 Union(const Bounds3 &a, const Bounds3 &b) {
    Bounds3 r;
    r.pMin = Vector3f::Min(a.pMin, b.pMin);
    r.pMax = Vector3f::Max(a.pMax, b.pMax);
    return r;
}
inline Bounds3 Union(const Bounds3 &a, const Vector3f &p) {
    Bounds3 r;
    r.pMin = Vector3f::Min(a.pMin, p);
    r.pMax = Vector3f::Max(a.pMax, p);
    return r;
}

struct Ray {
    Vector3f origin;
    Vector3f direction, direction_inv;
    double t, t_min, t_max;
    Ray(const Vector3f &ori, const Vector3f &dir, const double _t = 0.0) : origin(ori), direction(dir), t(_t) {
        direction_inv = Vector3f(1.f / direction.x, 1.f / direction.y, 1.f / direction.z);
        t_min = 0.0;
        t_max = std::numeric_limits<double>::max();
    }
};

// ---------------------------------------------------------------------------
enum LobeType { LOBE_DIFFUSE = 0, LOBE_SPECULAR = 1, LOBE_DELTA = 2, LOBE_ALL = 3 };

// Texture (TextureUtils.hpp:9-19): 8-bit texels, rows top to bottom.
struct Texture {
    std::vector<unsigned char> data;
    int width = 0;
    int height = 0;
    int channels = 4;
    bool empty() const { return data.empty(); }
};

class Material {
  public:
    Vector3f m_emission = Vector3f(0.0f);
    Vector3f baseColor = Vector3f(0.8f, 0.8f, 0.8f);
    float roughness = 0.5f;
    float metallic = 0.0f;
    float ior = 1.5f;
    // Material.hpp:48-51: sRGB colour, metallic-roughness (G, B), normal map,
    // emissive (kept for the asset; castRay does not read it)
    Texture emissiveTex;
    Texture baseColorTex;
    Texture normalTex;
    Texture metallicRoughnessTex;
    virtual ~Material() = default;
    virtual int kind() const = 0;  // prr_material_type
    virtual bool isDelta() const { return false; }
    virtual bool isEmissive() const { return false; }
    bool hasEmission() const { return m_emission.norm() > 1e-3f; }
};
class DiffuseMaterial : public Material {
  public:
    int kind() const override { return 0; }
};
class GlassMaterial : public Material {
  public:
    int kind() const override { return 1; }
    bool isDelta() const override { return true; }
};
class MirrorMaterial : public Material {
  public:
    int kind() const override { return 2; }
    bool isDelta() const override { return true; }
};
class EmissiveMaterial : public Material {
  public:
    int kind() const override { return 3; }
    bool isEmissive() const override { return true; }
};

class Object;
struct Intersection {
    bool happened = false;
    Vector3f coords;
    Vector2f tcoords;  // t0 * (1 - u - v) + t1 * u + t2 * v of the hit triangle (0 without uvs)
    Vector3f normal;
    Vector3f tangent;
    float tangentHandedness = 1.f;
    bool hasTangent = false;
    double tnear = std::numeric_limits<double>::max();
    Object *obj = nullptr;
    Material *material = nullptr;
};

class Object {
  public:
    virtual ~Object() {}
    virtual Bounds3 getBounds() = 0;
    virtual float getArea() = 0;
    virtual bool hasEmit() = 0;
};

class Triangle : public Object {
  public:
    Vector3f v0, v1, v2;
    Vector3f e1, e2;
    Vector2f t0, t1, t2;        // per-vertex texture coordinates
    Vector3f tan0, tan1, tan2;  // per-vertex tangents (xyz)
    float tangentW0 = 1.f, tangentW1 = 1.f, tangentW2 = 1.f;  // glTF handedness
    bool hasTangents = false;
    Vector3f n0, n1, n2;  // per-vertex normals
    bool hasSmoothNormals = false;
    Vector3f normal;
    float area;
    Material *m;
    Triangle(Vector3f _v0, Vector3f _v1, Vector3f _v2, Material *_m = nullptr) : v0(_v0), v1(_v1), v2(_v2), m(_m) {
        e1 = v1 - v0;
        e2 = v2 - v0;
        normal = normalize(crossProduct(e1, e2));
        area = crossProduct(e1, e2).norm() * 0.5f;
    }
    Bounds3 getBounds() override { return Union(Bounds3(v0, v1), v2); }
    float getArea() override { return area; }
    bool hasEmit() override { return m && m->hasEmission(); }
};

class MeshTriangle : public Object {
  public:
    // Wavefront OBJ triangle mesh translated by offset (Triangle.cpp:4-32).
    // Polygons are fan-triangulated; the file must hold one mesh. Throws
    // std::runtime_error when the file cannot be read.
    MeshTriangle(const std::string &filename, Vector3f offset, Material *mt);
    MeshTriangle(std::vector<Triangle> tris, Bounds3 bounds, Material *mt) {
        triangles = std::move(tris);
        bounding_box = bounds;
        m = mt;
        area = 0;
        for (auto &t : triangles) area += t.area;
    }
    Bounds3 getBounds() override { return bounding_box; }
    float getArea() override { return area; }
    bool hasEmit() override { return m && m->hasEmission(); }
    Bounds3 bounding_box;
    std::vector<Triangle> triangles;
    float area = 0.f;
    Material *m = nullptr;
};

class Sphere : public Object {
  public:
    Vector3f center;
    float radius, radius2;
    Material *m;
    float area;
    Sphere(const Vector3f &c, const float &r, Material *mt)
        : center(c), radius(r), radius2(r * r), m(mt), area(4 * 3.141592653589793f * r * r) {}
    Bounds3 getBounds() override {
        return Bounds3(Vector3f(center.x - radius, center.y - radius, center.z - radius),
                       Vector3f(center.x + radius, center.y + radius, center.z + radius));
    }
    float getArea() override { return area; }
    bool hasEmit() override { return m && m->hasEmission(); }
};

class PointLight {
  public:
    PointLight(const Vector3f &p, const Vector3f &i) : position(p), intensity(i) {}
    virtual ~PointLight() = default;
    Vector3f position;
    Vector3f intensity;
};

struct Camera {
    Vector3f eye = Vector3f(0, 0, -1);
    Vector3f target = Vector3f(0, 0, 0);
    Vector3f up = Vector3f(0, 1, 0);
    float fovv = 40.0f;
    float fovh = 40.0f;
    float aperture = 0.0f;
    float focusDistance = -1.0f;
    static Camera create(Vector3f eye_, Vector3f target_, Vector3f up_, float fovv_, float fovh_) {
        Camera c;
        c.eye = eye_;
        c.target = target_;
        c.up = up_;
        c.fovv = fovv_;
        c.fovh = fovh_;
        return c;
    }
    static Camera fromBlender(Vector3f eye_, Vector3f target_, Vector3f up_, float fovh_, float fovv_,
                              float focusDist_) {
        Camera c;
        c.eye = Vector3f(eye_.x, eye_.z, -eye_.y);
        c.target = Vector3f(target_.x, target_.z, -target_.y);
        c.up = normalize(Vector3f(up_.x, up_.z, -up_.y));
        c.fovv = fovv_;
        c.fovh = fovh_;
        c.focusDistance = focusDist_;
        return c;
    }
    Vector3f forward = Vector3f(0, 0, 1);
    Vector3f right = Vector3f(1, 0, 0);
    Vector3f camUp = Vector3f(0, 1, 0);
    // Camera::init (Camera.hpp:43-56) basis; the device derives the same
    // frame from eye / target / up / fovh at render time.
    void init(int width, int height) {
        (void) width;
        (void) height;
        forward = normalize(target - eye);
        right = normalize(crossProduct(forward, up));
        camUp = crossProduct(right, forward);
        if (focusDistance < 0.0f) focusDistance = (target - eye).norm();
    }
};

// Medium (Volume.hpp:7-66): carried for source compatibility; the reference's
// castRay does not use it and neither does the device integrator.
// Volume.hpp:7-66. As in the reference, Scene::medium is carried and not used
// by Integrator::castRay; intersect and the phase functions run on the device
// (prr_medium_intersect, prr_phase_hg, prr_phase_hg_sample).
struct Medium {
    float sigma_a = 0.0f;
    float sigma_s = 0.0f;
    float sigma_t = 0.0f;
    float g = 0.0f;
    bool bounded = false;
    Bounds3 bounds;
    bool active() const { return sigma_t > 1e-6f; }
    // Volume.hpp:17-44: entry and exit t of the ray in the medium's box (0 and
    // infinity when unbounded); false, outputs untouched, when it misses.
    bool intersect(const Ray &ray, float &t_enter, float &t_exit) const;
    static Medium dust() {  // Volume.hpp:46-53
        Medium m;
        m.sigma_s = 0.05f;
        m.sigma_a = 0.005f;
        m.sigma_t = m.sigma_s + m.sigma_a;
        m.g = 0.6f;
        return m;
    }
    static Medium fog() {  // Volume.hpp:55-62
        Medium m;
        m.sigma_s = 0.1f;
        m.sigma_a = 0.01f;
        m.sigma_t = m.sigma_s + m.sigma_a;
        m.g = 0.0f;
        return m;
    }
};
// Volume.cpp:3-24: Henyey-Greenstein phase function, and a direction drawn
// from it about wo with two draws from the calling thread's generator.
float phaseHG(float cos_theta, float g);
Vector3f samplePhaseHG(const Vector3f &wo, float g);

struct EnvMap {
    std::vector<Vector3f> pixels;  // linear RGB, row 0 = zenith
    int width = 0;
    int height = 0;
    bool loaded = false;
    // Radiance RGBE (.hdr) file, uncompressed or RLE scanlines.
    bool load(const std::string &path);
    void set(int w, int h, const float *rgb);
    bool empty() const { return !loaded; }
};

namespace prr_detail {
struct SceneDevice;
}
class Scene;

// BVHAccel (BVH.hpp:86-96): the closest / any-hit queries of a device BVH.
// Scene::buildBVH sets Scene::bvh to the scene's own; a BVHAccel built from
// a list of objects uploads a scene of just those objects (the reference's
// SAH rules: one object per scene-level leaf, four triangles per mesh leaf,
// so maxPrimsInNode must be 1 and splitMethod SAH). Queries go one ray at a
// time through the batch entry points (prr_trace_closest / prr_trace_any);
// batch them there for throughput.
class BVHAccel {
  public:
    enum class SplitMethod { NAIVE, SAH };
    BVHAccel(std::vector<Object *> p, int maxPrimsInNode = 1, SplitMethod splitMethod = SplitMethod::SAH);
    ~BVHAccel();
    BVHAccel(const BVHAccel &) = delete;
    BVHAccel &operator=(const BVHAccel &) = delete;
    Intersection Intersect(const Ray &ray) const;
    bool IntersectP(const Ray &ray, float tMax = std::numeric_limits<float>::infinity()) const;

  private:
    friend class Scene;
    explicit BVHAccel(const Scene *scene) : scene_(scene) {}
    const Scene *scene_ = nullptr;  // the scene whose device tree answers
    Scene *owned_ = nullptr;        // built from an object list
};

// Photon (Photon.hpp:11-16)
struct Photon {
    Vector3f position;
    Vector3f direction;
    Vector3f power;
    bool is_caustic = false;
};

// PhotonMap (Photon.hpp:32-47) on the device: emit() traces num_photons from
// the scene's emitters with the reference's rules (Photon.cpp:96-188;
// prr_photon_emit) and keeps the grid resident on the device; build() has
// nothing left to do (the device grid is built by emit); caustic_map holds a
// host copy of the stored photons; estimateIrradiance (Photon.cpp:82-94)
// queries the device grid (prr_photon_irradiance). Scenes with only a point
// light get the build's caustic map instead (prr_photon_build).
class PhotonMap {
  public:
    std::vector<Photon> caustic_map;
    float photon_radius = 0.01f;
    int maxDepth = 30;
    void emit(int num_photons, const Scene &scene);
    void build() {}
    Vector3f estimateIrradiance(Vector3f pos, Vector3f normal) const;

  private:
    friend class Scene;
    const Scene *scene_ = nullptr;
};

class Scene {
  public:
    Vector3f backgroundColor = Vector3f(0.235294f, 0.67451f, 0.843137f);
    std::vector<Object *> objects;
    std::vector<Object *> emissives;      // objects that hasEmit(), set by buildBVH (Scene.cpp:19-25)
    BVHAccel *bvh = nullptr;              // set by buildBVH; owned by the scene
    std::vector<PointLight> pointLights;  // extension: NEE + photon emission (at most one)
    EnvMap envMap;
    Camera camera;
    float totalEmitArea = 0.f;
    PhotonMap photon_map;
    Medium medium;

    Scene();
    ~Scene();
    Scene(const Scene &) = delete;
    Scene &operator=(const Scene &) = delete;
    Scene(Scene &&other) noexcept;  // `Scene scene = buildScene();` (main.cpp:23)
    Scene &operator=(Scene &&other) noexcept;

    void Add(Object *object) { objects.push_back(object); }
    const std::vector<Object *> &get_objects() const { return objects; }
    // Scene::buildBVH (Scene.cpp:15-27): host SAH build, device upload.
    void buildBVH();
    // Scene::buildPhotonMaps (Scene.cpp:8-13): photon_map.emit + build. With
    // emitters, the reference's own rules; with only a point light, the
    // build's caustic map that Renderer::Render then gathers.
    void buildPhotonMaps(int num_photons);
    Intersection intersect(const Ray &ray) const;
    bool intersectP(const Ray &ray, float tMax = std::numeric_limits<float>::infinity()) const;
    // Scene::sampleLight / pdfLight (Scene.cpp:37-62): a point on an emitter,
    // picked by area (4 draws of a thread-local SplitMix64, as
    // get_random_float), through prr_sample_light. pdf = 1 / totalEmitArea;
    // nothing is written when the scene has no emitter.
    void sampleLight(Intersection &pos, float &pdf) const;
    float pdfLight(const Intersection &lightInter) const;

    prr_detail::SceneDevice *device() const { return dev_; }

  private:
    prr_detail::SceneDevice *dev_;
};

class Integrator {
  public:
    float directClamp = 0.f;
    float indirectClamp = 10.0f;
    Integrator(const Scene &scene_, int maxDepth_, float rrThreshold_)
        : scene(scene_), maxDepth(maxDepth_), rrThreshold(rrThreshold_) {}
    // One path on the device (prr_cast_rays). Use Renderer::Render or
    // prr_cast_rays batches for throughput.
    Vector3f castRay(const Ray &ray) const;
    int depth() const { return maxDepth; }

  private:
    const Scene &scene;
    int maxDepth;
    float rrThreshold;
};

// GLTFLoader (GLTFLoader.hpp:12-37): one MeshTriangle per triangle
// primitive, materials and textures as the reference loads them (PNG images;
// JPEG throws). Returns heap objects the caller owns; throws
// std::runtime_error on an unreadable or malformed file.
enum class MaterialType { Diffuse, Glass, Mirror, Emit };
class GLTFLoader {
  public:
    static std::vector<MeshTriangle *> load(const std::string &filename, MaterialType matType,
                                            Material *overrideMat = nullptr);
};

namespace tonemap {
enum class ToneMapper { Raw, Standard, AgX, PBRNeutral, ACES };
ToneMapper fromString(const char *name);
const char *toString(ToneMapper mapper);
}  // namespace tonemap

struct RenderSettings {
    int width = 1280;
    int height = 1720;
    int minSPP = 64;
    int maxSPP = 64;
    float russianRoulette = 0.95f;
    int maxDepth = 30;
    float varianceThreshold = 0.05f;
    float exposure = 0.18f;
    tonemap::ToneMapper toneMapper = tonemap::ToneMapper::AgX;
    // extensions: PCG32 seed and output file names
    uint64_t seed = 0;
    std::string outputBase = "output";
    std::string sampleCountBase = "sample_count";
};

class Renderer {
  public:
    // Renders maxSPP samples per pixel on the device, then writes
    // <outputBase>.ppm/.png with the selected tone mapper, <outputBase>_raw
    // (Raw) unless Raw is selected, and sample_count.ppm/.png
    // (Renderer.cpp:149-211). The linear mean image is kept in lastImage.
    void Render(const Scene &scene, const Integrator &integrator, const RenderSettings &settings);
    std::vector<Vector3f> lastImage;
};
