// prr_api.cpp — the C++ drop-in classes (prr/prr.hpp) over the C ABI.
//
// Scene::buildBVH flattens the Object / Material graph into the POD
// descriptors of prr_scene_create (host SAH build + upload), Renderer::Render
// runs prr_render (adaptive sampling on the device) and writes the reference's
// output files, Integrator::castRay / Scene::intersect(P) forward single
// queries to the batch entry points. All device work goes through one
// process-wide context on device PRR_DEVICE (default 0).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <random>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "prr/prr.hpp"
#include "prr_abi.h"

namespace prr_detail {

static std::mutex g_mu;
static prr_context *g_ctx = nullptr;

static void check(int rc, const char *what) {
    if (rc != PRR_OK)
        throw std::runtime_error(std::string(what) + ": " + prr_last_error(g_ctx));
}

static prr_context *context() {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ctx) {
        const char *d = std::getenv("PRR_DEVICE");
        prr_context *c = nullptr;
        if (prr_create(d ? std::atoi(d) : 0, &c) != PRR_OK)
            throw std::runtime_error(std::string("prr_create: ") + prr_last_error(nullptr));
        g_ctx = c;
    }
    return g_ctx;
}

// Device buffer for the single-query forwards.
struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
    void *get(size_t bytes) {
        if (bytes > cap) {
            if (p) (void) hipFree(p);
            p = nullptr;
            cap = 0;
            if (hipMalloc(&p, bytes) != hipSuccess) throw std::runtime_error("hipMalloc failed");
            cap = bytes;
        }
        return p;
    }
    ~DevBuf() {
        if (p) (void) hipFree(p);
    }
};

struct SceneDevice {
    prr_scene *scene = nullptr;
    // global primitive id -> hit object and material (Intersection::obj / material)
    std::vector<Object *> primObj;
    std::vector<Material *> primMat;
    std::vector<Vector3f> primNormal;  // triangle geometric normal
    std::vector<Sphere *> primSphere;  // sphere of a sphere id, else null
    int64_t photons = 0;
    DevBuf buf;
    std::mutex mu;
    std::atomic<uint64_t> castCalls{0};
};

}  // namespace prr_detail

using prr_detail::check;
using prr_detail::context;

// ---------------------------------------------------------------------------
// EnvMap: Radiance RGBE (.hdr) reader. Format: text header ending in an empty
// line, resolution line "-Y H +X W", then H scanlines of W RGBE pixels, flat
// or new-style run-length encoded (per channel, runs > 128 repeat a byte).
// RGBE -> float as m * 2^(e - 136) per channel, 0 when e == 0 (Ward; Walter's
// rgbe.c).
// ---------------------------------------------------------------------------
static bool readScanline(FILE *f, int W, std::vector<uint8_t> &rgbe) {
    rgbe.assign(4 * (size_t) W, 0);
    int c0 = fgetc(f), c1 = fgetc(f), c2 = fgetc(f), c3 = fgetc(f);
    if (c3 == EOF) return false;
    const bool rle = W >= 8 && W < 32768 && c0 == 2 && c1 == 2 && !(c2 & 0x80) && ((c2 << 8) | c3) == W;
    if (!rle) {  // flat scanline (first pixel already read)
        rgbe[0] = (uint8_t) c0, rgbe[1] = (uint8_t) c1, rgbe[2] = (uint8_t) c2, rgbe[3] = (uint8_t) c3;
        if (W > 1 && fread(rgbe.data() + 4, 4, (size_t) W - 1, f) != (size_t) W - 1) return false;
        return true;
    }
    for (int ch = 0; ch < 4; ch++) {
        int x = 0;
        while (x < W) {
            int n = fgetc(f);
            if (n == EOF) return false;
            if (n > 128) {
                n -= 128;
                const int v = fgetc(f);
                if (v == EOF || x + n > W) return false;
                for (int k = 0; k < n; k++) rgbe[4 * (size_t) (x + k) + ch] = (uint8_t) v;
            } else {
                if (n == 0 || x + n > W) return false;
                for (int k = 0; k < n; k++) {
                    const int v = fgetc(f);
                    if (v == EOF) return false;
                    rgbe[4 * (size_t) (x + k) + ch] = (uint8_t) v;
                }
            }
            x += n;
        }
    }
    return true;
}

bool EnvMap::load(const std::string &path) {
    FILE *f = fopen(path.c_str(), "rb");
    if (!f) {
        fprintf(stderr, "[EnvMap] Failed to load: %s\n", path.c_str());
        return false;
    }
    std::unique_ptr<FILE, int (*)(FILE *)> guard(f, fclose);
    char line[512];
    if (!fgets(line, sizeof line, f) || (strncmp(line, "#?RADIANCE", 10) && strncmp(line, "#?RGBE", 6))) {
        fprintf(stderr, "[EnvMap] Failed to load: %s (not a Radiance HDR file)\n", path.c_str());
        return false;
    }
    bool rgbeFormat = false;
    for (;;) {
        if (!fgets(line, sizeof line, f)) return false;
        if (line[0] == '\n' || line[0] == '\r') break;
        if (!strncmp(line, "FORMAT=32-bit_rle_rgbe", 22)) rgbeFormat = true;
    }
    (void) rgbeFormat;  // an absent FORMAT line defaults to RGBE
    int H = 0, W = 0;
    if (!fgets(line, sizeof line, f) || sscanf(line, "-Y %d +X %d", &H, &W) != 2 || W <= 0 || H <= 0) {
        fprintf(stderr, "[EnvMap] Failed to load: %s (unsupported resolution line)\n", path.c_str());
        return false;
    }
    // a header is not trusted with the allocation: at most 2^15 on a side and
    // 2^28 pixels (a 32768 x 8192 sky), checked before any pixel is read
    if (W > (1 << 15) || H > (1 << 15) || (int64_t) W * H > ((int64_t) 1 << 28)) {
        fprintf(stderr, "[EnvMap] Failed to load: %s (%d x %d pixels is too large)\n", path.c_str(), W, H);
        return false;
    }
    std::vector<Vector3f> px((size_t) W * H);
    std::vector<uint8_t> sl;
    for (int y = 0; y < H; y++) {
        if (!readScanline(f, W, sl)) {
            fprintf(stderr, "[EnvMap] Failed to load: %s (truncated)\n", path.c_str());
            return false;
        }
        for (int x = 0; x < W; x++) {
            const uint8_t *q = &sl[4 * (size_t) x];
            Vector3f v(0.f);
            if (q[3]) {
                const float s = std::ldexp(1.0f, (int) q[3] - 136);
                v = Vector3f(q[0] * s, q[1] * s, q[2] * s);
            }
            px[(size_t) y * W + x] = v;
        }
    }
    pixels = std::move(px);
    width = W;
    height = H;
    loaded = true;
    return true;
}

void EnvMap::set(int w, int h, const float *rgb) {
    width = w;
    height = h;
    pixels.resize((size_t) w * h);
    for (size_t i = 0; i < pixels.size(); i++) pixels[i] = Vector3f(rgb[3 * i], rgb[3 * i + 1], rgb[3 * i + 2]);
    loaded = w > 0 && h > 0;
}

// ---------------------------------------------------------------------------
// Scene
// ---------------------------------------------------------------------------
Scene::Scene() : dev_(new prr_detail::SceneDevice) {}

Scene::~Scene() {
    delete bvh;
    if (dev_ && dev_->scene && prr_detail::g_ctx) prr_scene_free(prr_detail::g_ctx, dev_->scene);
    delete dev_;
}

Scene::Scene(Scene &&o) noexcept
    : backgroundColor(o.backgroundColor), objects(std::move(o.objects)), emissives(std::move(o.emissives)),
      bvh(o.bvh), pointLights(std::move(o.pointLights)), envMap(std::move(o.envMap)), camera(o.camera),
      totalEmitArea(o.totalEmitArea), photon_map(std::move(o.photon_map)), medium(o.medium), dev_(o.dev_) {
    o.dev_ = new prr_detail::SceneDevice;
    o.bvh = nullptr;
    if (bvh) bvh->scene_ = this;
    if (photon_map.scene_) photon_map.scene_ = this;
}

Scene &Scene::operator=(Scene &&o) noexcept {
    if (this != &o) {
        std::swap(dev_, o.dev_);
        std::swap(bvh, o.bvh);
        if (bvh) bvh->scene_ = this;
        if (o.bvh) o.bvh->scene_ = &o;
        backgroundColor = o.backgroundColor;
        objects = std::move(o.objects);
        emissives = std::move(o.emissives);
        pointLights = std::move(o.pointLights);
        envMap = std::move(o.envMap);
        camera = o.camera;
        totalEmitArea = o.totalEmitArea;
        photon_map = std::move(o.photon_map);
        if (photon_map.scene_) photon_map.scene_ = this;
        medium = o.medium;
    }
    return *this;
}

// ---------------------------------------------------------------------------
// GLTFLoader::load (GLTFLoader.cpp:140-281) over prr_gltf_load: one
// MeshTriangle per mesh of the result, each with a material of the requested
// class holding the loaded factors and textures, or `overrideMat`.
std::vector<MeshTriangle *> GLTFLoader::load(const std::string &filename, MaterialType matType,
                                             Material *overrideMat) {
    const int type = matType == MaterialType::Glass    ? PRR_MAT_GLASS
                     : matType == MaterialType::Mirror ? PRR_MAT_MIRROR
                     : matType == MaterialType::Emit   ? PRR_MAT_EMISSIVE
                                                       : PRR_MAT_DIFFUSE;
    prr_gltf_data *g = nullptr;
    if (prr_gltf_load(filename.c_str(), type, &g) != PRR_OK)
        throw std::runtime_error(std::string("GLTFLoader::load: ") + prr_last_error(nullptr));
    std::unique_ptr<prr_gltf_data, void (*)(prr_gltf_data *)> own(g, prr_gltf_free);
    auto tex = [&](int row, Texture &out) {
        if (row < 0) return;
        const prr_texture_desc &t = g->textures[row];
        out.width = t.width;
        out.height = t.height;
        out.channels = t.channels;
        out.data.assign(t.data, t.data + (size_t) t.width * t.height * t.channels);
    };
    std::vector<MeshTriangle *> out;
    for (int k = 0; k < g->num_meshes; k++) {
        Material *mat = overrideMat;
        if (!mat) {
            const prr_material_desc &md = g->materials[g->mesh_material[k]];
            const int32_t *tx = g->material_textures + 4 * g->mesh_material[k];
            Material *m = md.type == PRR_MAT_GLASS    ? (Material *) new GlassMaterial()
                          : md.type == PRR_MAT_MIRROR ? (Material *) new MirrorMaterial()
                          : md.type == PRR_MAT_EMISSIVE ? (Material *) new EmissiveMaterial()
                                                        : (Material *) new DiffuseMaterial();
            m->baseColor = Vector3f(md.base_color[0], md.base_color[1], md.base_color[2]);
            m->roughness = md.roughness;
            m->metallic = md.metallic;
            m->ior = md.ior;
            m->m_emission = Vector3f(md.emission[0], md.emission[1], md.emission[2]);
            tex(tx[0], m->baseColorTex);
            tex(tx[1], m->metallicRoughnessTex);
            tex(tx[2], m->normalTex);
            tex(tx[3], m->emissiveTex);
            mat = m;
        }
        std::vector<Triangle> tris;
        Vector3f mn(std::numeric_limits<float>::infinity()), mx(-std::numeric_limits<float>::infinity());
        for (int i = g->mesh_first[k]; i < g->mesh_first[k] + g->mesh_count[k]; i++) {
            const float *p = g->positions + 9 * (size_t) i;
            Triangle t(Vector3f(p[0], p[1], p[2]), Vector3f(p[3], p[4], p[5]), Vector3f(p[6], p[7], p[8]), mat);
            for (const Vector3f &v : {t.v0, t.v1, t.v2}) {
                mn = Vector3f(std::min(mn.x, v.x), std::min(mn.y, v.y), std::min(mn.z, v.z));
                mx = Vector3f(std::max(mx.x, v.x), std::max(mx.y, v.y), std::max(mx.z, v.z));
            }
            const uint8_t fl = g->tri_flags[i];
            if (fl & PRR_TRI_SMOOTH_NORMALS) {
                const float *n = g->normals + 9 * (size_t) i;
                t.n0 = Vector3f(n[0], n[1], n[2]);
                t.n1 = Vector3f(n[3], n[4], n[5]);
                t.n2 = Vector3f(n[6], n[7], n[8]);
                t.hasSmoothNormals = true;
                t.normal = normalize(t.n0 + t.n1 + t.n2);
            }
            const float *q = g->uvs + 6 * (size_t) i;
            t.t0 = Vector2f(q[0], q[1]);
            t.t1 = Vector2f(q[2], q[3]);
            t.t2 = Vector2f(q[4], q[5]);
            if (fl & PRR_TRI_TANGENTS) {
                const float *a = g->tangents + 12 * (size_t) i;
                t.tan0 = Vector3f(a[0], a[1], a[2]);
                t.tangentW0 = a[3];
                t.tan1 = Vector3f(a[4], a[5], a[6]);
                t.tangentW1 = a[7];
                t.tan2 = Vector3f(a[8], a[9], a[10]);
                t.tangentW2 = a[11];
                t.hasTangents = true;
            }
            tris.push_back(t);
        }
        out.push_back(new MeshTriangle(std::move(tris), Bounds3(mn, mx), mat));
    }
    return out;
}

// MeshTriangle from a Wavefront OBJ file: "v x y z" and "f a b c ..." records
// (vertex index first in each a/b/c group, 1-based, negative = relative to the
// end); faces become triangle fans in file order.
// ---------------------------------------------------------------------------
MeshTriangle::MeshTriangle(const std::string &filename, Vector3f offset, Material *mt) {
    FILE *f = fopen(filename.c_str(), "r");
    if (!f) throw std::runtime_error("MeshTriangle: cannot open " + filename);
    std::unique_ptr<FILE, int (*)(FILE *)> guard(f, fclose);
    std::vector<Vector3f> verts;
    char line[4096];
    Vector3f lo(std::numeric_limits<float>::infinity()), hi(-std::numeric_limits<float>::infinity());
    while (fgets(line, sizeof line, f)) {
        if (line[0] == 'v' && line[1] == ' ') {
            float x, y, z;
            if (sscanf(line + 2, "%f %f %f", &x, &y, &z) == 3) verts.emplace_back(x, y, z);
        } else if (line[0] == 'f' && line[1] == ' ') {
            std::vector<int> idx;
            char *q = line + 2;
            while (*q) {
                while (*q == ' ' || *q == '\t') q++;
                if (!*q || *q == '\n' || *q == '\r') break;
                errno = 0;
                long v = std::strtol(q, &q, 10);
                const long n = (long) verts.size();
                if (errno == ERANGE || v == 0 || v > n || v < -n)  // 1-based, or -1 = the last vertex
                    throw std::runtime_error("MeshTriangle: bad face index in " + filename);
                if (v < 0) v = n + v + 1;
                idx.push_back((int) (v - 1));
                while (*q && *q != ' ' && *q != '\t' && *q != '\n' && *q != '\r') q++;
            }
            for (size_t k = 2; k < idx.size(); k++) {
                Vector3f a = verts[idx[0]] + offset, b = verts[idx[k - 1]] + offset, c = verts[idx[k]] + offset;
                for (const Vector3f *p : {&a, &b, &c}) {
                    lo = Vector3f::Min(lo, *p);
                    hi = Vector3f::Max(hi, *p);
                }
                triangles.emplace_back(a, b, c, mt);
            }
        }
    }
    if (triangles.empty()) throw std::runtime_error("MeshTriangle: no faces in " + filename);
    bounding_box = Bounds3(lo, hi);
    m = mt;
    area = 0;
    for (auto &t : triangles) area += t.area;
}

static prr_camera_desc cameraDesc(const Camera &c) {
    prr_camera_desc d{};
    const Vector3f *v[3] = {&c.eye, &c.target, &c.up};
    float *o[3] = {d.eye, d.target, d.up};
    for (int k = 0; k < 3; k++) {
        o[k][0] = v[k]->x;
        o[k][1] = v[k]->y;
        o[k][2] = v[k]->z;
    }
    d.fovh = c.fovh;
    d.aperture = c.aperture;
    d.focus_distance = c.focusDistance;
    return d;
}

// Scene::buildBVH (Scene.cpp:15-27): objects in Add order, materials by
// pointer identity, then the host SAH build and upload of prr_scene_create.
void Scene::buildBVH() {
    prr_context *ctx = context();
    auto &D = *dev_;
    std::vector<prr_object_desc> objs;
    std::vector<float> tris, sph;
    std::vector<prr_material_desc> mats;
    std::unordered_map<const Material *, int> matId;
    std::vector<std::pair<Object *, int>> sphereObjs;  // (object, material)
    std::vector<uint8_t> triFlags;
    std::vector<float> triNormals, triUvs, triTan;
    bool anyAttr = false;
    D.primObj.clear();
    D.primMat.clear();
    D.primNormal.clear();
    D.primSphere.clear();
    std::vector<int32_t> matTex;
    std::vector<prr_texture_desc> texs;
    std::unordered_map<const Texture *, int> texId;
    bool anyTex = false;
    auto texture = [&](const Texture &t) -> int {
        if (t.empty()) return -1;
        auto it = texId.find(&t);
        if (it != texId.end()) return it->second;
        if ((size_t) t.width * t.height * t.channels != t.data.size())
            throw std::runtime_error("Scene::buildBVH: texture size does not match its data");
        texs.push_back(prr_texture_desc{t.width, t.height, t.channels, t.data.data()});
        anyTex = true;
        return texId[&t] = (int) texs.size() - 1;
    };
    auto material = [&](Material *m) -> int {
        if (!m) throw std::runtime_error("Scene::buildBVH: object without material");
        auto it = matId.find(m);
        if (it != matId.end()) return it->second;
        matTex.insert(matTex.end(), {texture(m->baseColorTex), texture(m->metallicRoughnessTex),
                                     texture(m->normalTex), texture(m->emissiveTex)});
        prr_material_desc d{};
        d.type = m->kind();
        d.base_color[0] = m->baseColor.x;
        d.base_color[1] = m->baseColor.y;
        d.base_color[2] = m->baseColor.z;
        d.roughness = m->roughness;
        d.metallic = m->metallic;
        d.ior = m->ior;
        d.emission[0] = m->m_emission.x;
        d.emission[1] = m->m_emission.y;
        d.emission[2] = m->m_emission.z;
        mats.push_back(d);
        matId.emplace(m, (int) mats.size() - 1);
        return (int) mats.size() - 1;
    };
    for (Object *o : objects) {
        if (auto *mesh = dynamic_cast<MeshTriangle *>(o)) {
            if (mesh->triangles.empty()) throw std::runtime_error("Scene::buildBVH: empty MeshTriangle");
            prr_object_desc od{PRR_OBJ_MESH, material(mesh->m), (int32_t) (tris.size() / 9),
                               (int32_t) mesh->triangles.size()};
            for (auto &t : mesh->triangles) {
                const Vector3f *v[3] = {&t.v0, &t.v1, &t.v2};
                for (int k = 0; k < 3; k++) {
                    tris.push_back(v[k]->x);
                    tris.push_back(v[k]->y);
                    tris.push_back(v[k]->z);
                }
                // per-vertex attributes (Triangle.hpp:16-22)
                uint8_t fl = PRR_TRI_UVS;
                if (t.hasSmoothNormals) fl |= PRR_TRI_SMOOTH_NORMALS;
                if (t.hasTangents) fl |= PRR_TRI_TANGENTS;
                anyAttr |= t.hasSmoothNormals || t.hasTangents || t.t0.x != 0.f || t.t0.y != 0.f ||
                           t.t1.x != 0.f || t.t1.y != 0.f || t.t2.x != 0.f || t.t2.y != 0.f;
                triFlags.push_back(fl);
                triNormals.insert(triNormals.end(), {t.n0.x, t.n0.y, t.n0.z, t.n1.x, t.n1.y, t.n1.z, t.n2.x, t.n2.y, t.n2.z});
                triUvs.insert(triUvs.end(), {t.t0.x, t.t0.y, t.t1.x, t.t1.y, t.t2.x, t.t2.y});
                triTan.insert(triTan.end(), {t.tan0.x, t.tan0.y, t.tan0.z, t.tangentW0, t.tan1.x, t.tan1.y, t.tan1.z,
                                             t.tangentW1, t.tan2.x, t.tan2.y, t.tan2.z, t.tangentW2});
                D.primObj.push_back(&t);
                D.primMat.push_back(mesh->m);
                D.primNormal.push_back(t.normal);
            }
            objs.push_back(od);
        } else if (auto *s = dynamic_cast<Sphere *>(o)) {
            prr_object_desc od{PRR_OBJ_SPHERE, material(s->m), (int32_t) (sph.size() / 4), 1};
            sph.insert(sph.end(), {s->center.x, s->center.y, s->center.z, s->radius});
            sphereObjs.emplace_back(s, od.material);
            objs.push_back(od);
        } else {
            throw std::runtime_error("Scene::buildBVH: unsupported Object type (MeshTriangle and Sphere only)");
        }
    }
    // sphere ids follow all triangles
    for (auto &so : sphereObjs) {
        auto *s = static_cast<Sphere *>(so.first);
        D.primObj.push_back(s);
        D.primMat.push_back(s->m);
        D.primNormal.push_back(Vector3f(0.f));
    }
    D.primSphere.assign(D.primObj.size(), nullptr);
    for (size_t k = 0; k < sphereObjs.size(); k++)
        D.primSphere[tris.size() / 9 + k] = static_cast<Sphere *>(sphereObjs[k].first);

    std::vector<prr_point_light_desc> pls;
    for (const auto &pl : pointLights)
        pls.push_back({{pl.position.x, pl.position.y, pl.position.z}, {pl.intensity.x, pl.intensity.y, pl.intensity.z}});
    std::vector<float> env;
    prr_scene_desc d{};
    d.num_objects = (int32_t) objs.size();
    d.objects = objs.data();
    d.num_triangles = (int32_t) (tris.size() / 9);
    d.triangles = tris.data();
    d.num_spheres = (int32_t) (sph.size() / 4);
    d.spheres = sph.data();
    d.num_materials = (int32_t) mats.size();
    d.materials = mats.data();
    if (anyAttr) {
        d.tri_flags = triFlags.data();
        d.tri_normals = triNormals.data();
        d.tri_uvs = triUvs.data();
        d.tri_tangents = triTan.data();
    }
    if (anyTex) {
        d.material_textures = matTex.data();
        d.num_textures = (int32_t) texs.size();
        d.textures = texs.data();
    }
    if (envMap.loaded) {
        env.resize(3 * envMap.pixels.size());
        for (size_t i = 0; i < envMap.pixels.size(); i++) {
            env[3 * i] = envMap.pixels[i].x;
            env[3 * i + 1] = envMap.pixels[i].y;
            env[3 * i + 2] = envMap.pixels[i].z;
        }
        d.env.width = envMap.width;
        d.env.height = envMap.height;
        d.env.pixels = env.data();
    }
    d.background[0] = backgroundColor.x;
    d.background[1] = backgroundColor.y;
    d.background[2] = backgroundColor.z;
    d.num_point_lights = (int32_t) pls.size();
    d.point_lights = pls.data();
    d.camera = cameraDesc(camera);
    if (D.scene) {
        check(prr_scene_free(ctx, D.scene), "prr_scene_free");
        D.scene = nullptr;
    }
    check(prr_scene_create(ctx, &d, &D.scene), "prr_scene_create");
    D.photons = 0;
    prr_scene_info info{};
    check(prr_scene_get_info(D.scene, &info), "prr_scene_get_info");
    totalEmitArea = info.total_emit_area;
    // Scene.cpp:19-25: the emitters, in object order
    emissives.clear();
    for (Object *o : objects)
        if (o->hasEmit()) emissives.push_back(o);
    if (!bvh) bvh = new BVHAccel(this);
}

// Scene::buildPhotonMaps (Scene.cpp:8-13) with PhotonMap's radius and depth
// (0.01, 30: Photon.hpp:36-37).
void Scene::buildPhotonMaps(int num_photons) {
    photon_map.emit(num_photons, *this);
    photon_map.build();
}

// PhotonMap::emit: the reference's rules from the emitters (prr_photon_emit),
// or, in a scene with only a point light, the build's caustic map
// (prr_photon_build), which Renderer::Render gathers. The stored photons are
// copied to caustic_map.
void PhotonMap::emit(int num_photons, const Scene &scene) {
    prr_detail::SceneDevice *D = scene.device();
    if (!D->scene) throw std::runtime_error("PhotonMap::emit before Scene::buildBVH");
    prr_context *ctx = context();
    std::lock_guard<std::mutex> lk(D->mu);
    int64_t stored = 0;
    const uint64_t seed = 0x9E3779B97F4A7C15ull;
    if (scene.totalEmitArea > 0.f) {
        check(prr_photon_emit(ctx, D->scene, num_photons, photon_radius, maxDepth, seed, &stored), "prr_photon_emit");
        D->photons = 0;
    } else {
        check(prr_photon_build(ctx, D->scene, num_photons, photon_radius, maxDepth, seed, &stored),
              "prr_photon_build");
        D->photons = num_photons;
    }
    scene_ = &scene;
    caustic_map.assign((size_t) stored, Photon());
    if (stored <= 0) return;
    float *dev = (float *) D->buf.get((size_t) stored * 9 * sizeof(float));
    int64_t cnt = 0;
    check(prr_photon_export(ctx, D->scene, stored, dev, &cnt), "prr_photon_export");
    check(prr_synchronize(ctx), "prr_synchronize");
    std::vector<float> h((size_t) stored * 9);
    if (hipMemcpy(h.data(), dev, h.size() * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
        throw std::runtime_error("hipMemcpy");
    for (int64_t i = 0; i < stored; i++) {
        const float *q = h.data() + 9 * i;
        caustic_map[i] = Photon{Vector3f(q[0], q[1], q[2]), Vector3f(q[3], q[4], q[5]), Vector3f(q[6], q[7], q[8]),
                                true};
    }
}

// PhotonMap::estimateIrradiance (Photon.cpp:82-94) on the device grid.
Vector3f PhotonMap::estimateIrradiance(Vector3f pos, Vector3f normal) const {
    (void) normal;
    if (!scene_) return Vector3f(0.f);
    prr_detail::SceneDevice *D = scene_->device();
    prr_context *ctx = context();
    std::lock_guard<std::mutex> lk(D->mu);
    float *b = (float *) D->buf.get(256);
    const float p[3] = {pos.x, pos.y, pos.z};
    if (hipMemcpy(b, p, sizeof p, hipMemcpyHostToDevice) != hipSuccess) throw std::runtime_error("hipMemcpy");
    check(prr_photon_irradiance(ctx, D->scene, 1, b, nullptr, b + 4), "prr_photon_irradiance");
    check(prr_synchronize(ctx), "prr_synchronize");
    float e[3];
    if (hipMemcpy(e, b + 4, sizeof e, hipMemcpyDeviceToHost) != hipSuccess) throw std::runtime_error("hipMemcpy");
    return Vector3f(e[0], e[1], e[2]);
}

// get_random_float (global.hpp:35-50): a thread-local SplitMix64 seeded from
// std::random_device, top 24 bits.
static float hostRandomFloat() {
    thread_local uint64_t state = [] {
        std::random_device rd;
        return (uint64_t) rd() << 32 | rd();
    }();
    uint64_t z = (state += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z ^= (z >> 31);
    return (float) (z >> 40) * 0x1p-24f;
}

// Device scratch of the media functions (no scene to hold it), taken under
// its lock; kept until the process ends (never freed during static
// destruction, when the HIP runtime may already be gone).
static prr_detail::DevBuf &mediaBuf(std::unique_lock<std::mutex> &lk) {
    static std::mutex mu;
    static prr_detail::DevBuf *buf = new prr_detail::DevBuf;
    lk = std::unique_lock<std::mutex>(mu);
    return *buf;
}

// Medium::intersect (Volume.hpp:17-44) through prr_medium_intersect.
bool Medium::intersect(const Ray &ray, float &t_enter, float &t_exit) const {
    prr_context *ctx = context();
    std::unique_lock<std::mutex> lk;
    prr_detail::DevBuf &buf = mediaBuf(lk);
    float h[16] = {ray.origin.x,    ray.origin.y,    ray.origin.z,    ray.direction.x, ray.direction.y,
                   ray.direction.z, bounds.pMin.x,   bounds.pMin.y,   bounds.pMin.z,   bounds.pMax.x,
                   bounds.pMax.y,   bounds.pMax.z,   0.f,             t_enter,         t_exit,  0.f};
    const int32_t bd = bounded ? 1 : 0;
    std::memcpy(&h[12], &bd, 4);
    float *d = (float *) buf.get(sizeof h);
    if (hipMemcpy(d, h, sizeof h, hipMemcpyHostToDevice) != hipSuccess) throw std::runtime_error("hipMemcpy");
    check(prr_medium_intersect(ctx, 1, d, d + 3, d + 6, (const int32_t *) (d + 12), d + 13, d + 14,
                               (uint8_t *) (d + 15)),
          "prr_medium_intersect");
    check(prr_synchronize(ctx), "prr_synchronize");
    if (hipMemcpy(h, d, sizeof h, hipMemcpyDeviceToHost) != hipSuccess) throw std::runtime_error("hipMemcpy");
    uint8_t hit;
    std::memcpy(&hit, &h[15], 1);
    if (!hit) return false;
    t_enter = h[13];
    t_exit = h[14];
    return true;
}

// phaseHG (Volume.cpp:3-6) through prr_phase_hg.
float phaseHG(float cos_theta, float g) {
    prr_context *ctx = context();
    std::unique_lock<std::mutex> lk;
    prr_detail::DevBuf &buf = mediaBuf(lk);
    float h[3] = {cos_theta, g, 0.f};
    float *d = (float *) buf.get(sizeof h);
    if (hipMemcpy(d, h, sizeof h, hipMemcpyHostToDevice) != hipSuccess) throw std::runtime_error("hipMemcpy");
    check(prr_phase_hg(ctx, 1, d, d + 1, d + 2), "prr_phase_hg");
    check(prr_synchronize(ctx), "prr_synchronize");
    if (hipMemcpy(h, d, sizeof h, hipMemcpyDeviceToHost) != hipSuccess) throw std::runtime_error("hipMemcpy");
    return h[2];
}

// samplePhaseHG (Volume.cpp:8-24): its two draws from the calling thread's
// generator, then prr_phase_hg_sample.
Vector3f samplePhaseHG(const Vector3f &wo, float g) {
    prr_context *ctx = context();
    const float u1 = hostRandomFloat(), u2 = hostRandomFloat();
    std::unique_lock<std::mutex> lk;
    prr_detail::DevBuf &buf = mediaBuf(lk);
    float h[9] = {wo.x, wo.y, wo.z, g, u1, u2, 0.f, 0.f, 0.f};
    float *d = (float *) buf.get(sizeof h);
    if (hipMemcpy(d, h, sizeof h, hipMemcpyHostToDevice) != hipSuccess) throw std::runtime_error("hipMemcpy");
    check(prr_phase_hg_sample(ctx, 1, d, d + 3, d + 4, d + 6), "prr_phase_hg_sample");
    check(prr_synchronize(ctx), "prr_synchronize");
    if (hipMemcpy(h, d, sizeof h, hipMemcpyDeviceToHost) != hipSuccess) throw std::runtime_error("hipMemcpy");
    return Vector3f(h[6], h[7], h[8]);
}

void Scene::sampleLight(Intersection &pos, float &pdf) const {
    if (totalEmitArea <= 0.f) return;  // Scene.cpp:38-39
    if (!dev_->scene) throw std::runtime_error("Scene::sampleLight before buildBVH");
    prr_context *ctx = context();
    std::lock_guard<std::mutex> lk(dev_->mu);
    float u[4];
    for (float &x : u) x = hostRandomFloat();
    float *b = (float *) dev_->buf.get(256);
    if (hipMemcpy(b, u, sizeof u, hipMemcpyHostToDevice) != hipSuccess) throw std::runtime_error("hipMemcpy");
    int32_t *em = (int32_t *) (b + 12);
    check(prr_sample_light(ctx, dev_->scene, 1, b, b + 4, b + 8, em, b + 11), "prr_sample_light");
    check(prr_synchronize(ctx), "prr_synchronize");
    float h[13];
    if (hipMemcpy(h, b, sizeof h, hipMemcpyDeviceToHost) != hipSuccess) throw std::runtime_error("hipMemcpy");
    int32_t k;
    std::memcpy(&k, &h[12], 4);
    pos.coords = Vector3f(h[4], h[5], h[6]);
    pos.normal = Vector3f(h[8], h[9], h[10]);
    pos.happened = true;
    if (k >= 0 && k < (int32_t) emissives.size()) {
        Object *o = emissives[k];
        pos.obj = o;
        if (auto *mesh = dynamic_cast<MeshTriangle *>(o))
            pos.material = mesh->m;
        else if (auto *sph = dynamic_cast<Sphere *>(o))
            pos.material = sph->m;
    }
    pdf = 1.0f / totalEmitArea;  // Scene.cpp:46
}

float Scene::pdfLight(const Intersection &lightInter) const {  // Scene.cpp:60-62
    (void) lightInter;
    return totalEmitArea > 0.f ? 1.f / totalEmitArea : 0.f;
}

// BVHAccel over a list of objects: a scene of just those objects.
BVHAccel::BVHAccel(std::vector<Object *> p, int maxPrimsInNode, SplitMethod splitMethod) {
    if (maxPrimsInNode != 1 || splitMethod != SplitMethod::SAH)
        throw std::invalid_argument("BVHAccel: the device tree follows Scene::buildBVH (maxPrimsInNode 1, SAH)");
    owned_ = new Scene();
    owned_->objects = std::move(p);
    owned_->buildBVH();
    scene_ = owned_;
}

BVHAccel::~BVHAccel() { delete owned_; }

Intersection BVHAccel::Intersect(const Ray &ray) const { return scene_->intersect(ray); }

bool BVHAccel::IntersectP(const Ray &ray, float tMax) const { return scene_->intersectP(ray, tMax); }

static void rayArrays(const Ray &ray, float *od) {
    od[0] = ray.origin.x, od[1] = ray.origin.y, od[2] = ray.origin.z;
    od[3] = ray.direction.x, od[4] = ray.direction.y, od[5] = ray.direction.z;
}

Intersection Scene::intersect(const Ray &ray) const {
    if (!dev_->scene) throw std::runtime_error("Scene::intersect before buildBVH");
    prr_context *ctx = context();
    std::lock_guard<std::mutex> lk(dev_->mu);
    float od[6];
    rayArrays(ray, od);
    char *b = (char *) dev_->buf.get(256);
    float *dO = (float *) b, *dD = dO + 3, *dT = dO + 6, *dU = dO + 7, *dV = dO + 8;
    int32_t *dP = (int32_t *) (dO + 9);
    if (hipMemcpy(dO, od, sizeof od, hipMemcpyHostToDevice) != hipSuccess) throw std::runtime_error("hipMemcpy");
    check(prr_trace_closest(ctx, dev_->scene, 1, dO, dD, dP, dT, dU, dV), "prr_trace_closest");
    check(prr_synchronize(ctx), "prr_synchronize");
    float h[10];
    if (hipMemcpy(h, dO, sizeof h, hipMemcpyDeviceToHost) != hipSuccess) throw std::runtime_error("hipMemcpy");
    int32_t prim;
    std::memcpy(&prim, &h[9], 4);
    Intersection it;
    if (prim < 0 || prim >= (int32_t) dev_->primObj.size()) return it;
    const float t = h[6];
    it.happened = true;
    it.tnear = t;
    it.coords = ray.origin + ray.direction * t;
    it.obj = dev_->primObj[prim];
    it.material = dev_->primMat[prim];
    if (Sphere *s = dev_->primSphere[prim]) {
        it.normal = normalize(it.coords - s->center);
    } else {  // Triangle::finalize (Triangle.cpp:42-56)
        const Triangle &tr = *static_cast<const Triangle *>(dev_->primObj[prim]);
        const float u = h[7], v = h[8], w0 = 1.f - u - v;
        it.tcoords = tr.t0 * w0 + tr.t1 * u + tr.t2 * v;
        it.normal = tr.hasSmoothNormals ? normalize(tr.n0 * w0 + tr.n1 * u + tr.n2 * v) : dev_->primNormal[prim];
        if (tr.hasTangents) {
            it.tangent = normalize(tr.tan0 * w0 + tr.tan1 * u + tr.tan2 * v);
            const float w = tr.tangentW0 * w0 + tr.tangentW1 * u + tr.tangentW2 * v;
            it.tangentHandedness = w >= 0.f ? 1.f : -1.f;
            it.hasTangent = true;
        }
    }
    return it;
}

bool Scene::intersectP(const Ray &ray, float tMax) const {
    if (!dev_->scene) throw std::runtime_error("Scene::intersectP before buildBVH");
    prr_context *ctx = context();
    std::lock_guard<std::mutex> lk(dev_->mu);
    float od[7];
    rayArrays(ray, od);
    od[6] = tMax;
    char *b = (char *) dev_->buf.get(256);
    float *dO = (float *) b;
    uint8_t *dH = (uint8_t *) (dO + 8);
    if (hipMemcpy(dO, od, sizeof od, hipMemcpyHostToDevice) != hipSuccess) throw std::runtime_error("hipMemcpy");
    check(prr_trace_any(ctx, dev_->scene, 1, dO, dO + 3, dO + 6, dH), "prr_trace_any");
    check(prr_synchronize(ctx), "prr_synchronize");
    uint8_t hit = 0;
    if (hipMemcpy(&hit, dH, 1, hipMemcpyDeviceToHost) != hipSuccess) throw std::runtime_error("hipMemcpy");
    return hit != 0;
}

// ---------------------------------------------------------------------------
// Integrator::castRay: one path through prr_cast_rays. Each call takes the
// next PCG32 stream, as successive calls of the reference draw fresh numbers.
// ---------------------------------------------------------------------------
Vector3f Integrator::castRay(const Ray &ray) const {
    prr_detail::SceneDevice *D = scene.device();
    if (!D->scene) throw std::runtime_error("Integrator::castRay before Scene::buildBVH");
    prr_context *ctx = context();
    std::lock_guard<std::mutex> lk(D->mu);
    float od[6];
    rayArrays(ray, od);
    float *b = (float *) D->buf.get(256);
    if (hipMemcpy(b, od, sizeof od, hipMemcpyHostToDevice) != hipSuccess) throw std::runtime_error("hipMemcpy");
    const uint64_t seed = 0x5EEDull + D->castCalls.fetch_add(1);
    check(prr_cast_rays(ctx, D->scene, 1, b, b + 3, seed, maxDepth, indirectClamp, b + 8, nullptr), "prr_cast_rays");
    float L[3];
    if (hipMemcpy(L, b + 8, sizeof L, hipMemcpyDeviceToHost) != hipSuccess) throw std::runtime_error("hipMemcpy");
    return Vector3f(L[0], L[1], L[2]);
}

// ---------------------------------------------------------------------------
// tone-mapper names (ToneMapping.hpp:282-296 plus ACES)
// ---------------------------------------------------------------------------
namespace tonemap {
ToneMapper fromString(const char *name) {
    if (name) {
        const std::string n(name);
        if (n == "agx" || n == "AgX" || n == "AGX") return ToneMapper::AgX;
        if (n == "pbrneutral" || n == "PBRNeutral" || n == "pbr_neutral" || n == "pbr" || n == "neutral")
            return ToneMapper::PBRNeutral;
        if (n == "standard" || n == "Standard" || n == "STANDARD") return ToneMapper::Standard;
        if (n == "raw" || n == "Raw" || n == "RAW") return ToneMapper::Raw;
        if (n == "aces" || n == "ACES" || n == "Aces") return ToneMapper::ACES;
    }
    return ToneMapper::AgX;
}
const char *toString(ToneMapper m) {
    switch (m) {
    case ToneMapper::Raw: return "Raw";
    case ToneMapper::Standard: return "Standard";
    case ToneMapper::AgX: return "AgX";
    case ToneMapper::PBRNeutral: return "PBRNeutral";
    case ToneMapper::ACES: return "ACES";
    }
    return "AgX";
}
static int op(ToneMapper m) {
    switch (m) {
    case ToneMapper::Raw: return PRR_TONEMAP_RAW;
    case ToneMapper::Standard: return PRR_TONEMAP_STANDARD;
    case ToneMapper::AgX: return PRR_TONEMAP_AGX;
    case ToneMapper::PBRNeutral: return PRR_TONEMAP_PBR_NEUTRAL;
    case ToneMapper::ACES: return PRR_TONEMAP_ACES;
    }
    return PRR_TONEMAP_AGX;
}
}  // namespace tonemap

// ---------------------------------------------------------------------------
// image files: binary PPM and PNG (zlib stream of stored deflate blocks)
// ---------------------------------------------------------------------------
static uint32_t crc32Update(uint32_t c, const uint8_t *p, size_t n) {
    static uint32_t table[256];
    static std::once_flag once;
    std::call_once(once, [] {
        for (uint32_t i = 0; i < 256; i++) {
            uint32_t v = i;
            for (int k = 0; k < 8; k++) v = (v & 1) ? 0xEDB88320u ^ (v >> 1) : v >> 1;
            table[i] = v;
        }
    });
    c = ~c;
    for (size_t i = 0; i < n; i++) c = table[(c ^ p[i]) & 0xFF] ^ (c >> 8);
    return ~c;
}

static void put32be(std::vector<uint8_t> &o, uint32_t v) {
    o.push_back((uint8_t) (v >> 24));
    o.push_back((uint8_t) (v >> 16));
    o.push_back((uint8_t) (v >> 8));
    o.push_back((uint8_t) v);
}

static void pngChunk(FILE *f, const char *type, const std::vector<uint8_t> &data) {
    std::vector<uint8_t> b;
    put32be(b, (uint32_t) data.size());
    b.insert(b.end(), type, type + 4);
    b.insert(b.end(), data.begin(), data.end());
    put32be(b, crc32Update(0, b.data() + 4, b.size() - 4));
    fwrite(b.data(), 1, b.size(), f);
}

static bool writePng(const std::string &path, int W, int H, const uint8_t *rgb) {
    FILE *f = fopen(path.c_str(), "wb");
    if (!f) return false;
    static const uint8_t sig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
    fwrite(sig, 1, 8, f);
    std::vector<uint8_t> ihdr;
    put32be(ihdr, (uint32_t) W);
    put32be(ihdr, (uint32_t) H);
    ihdr.insert(ihdr.end(), {8, 2, 0, 0, 0});  // 8-bit RGB, no interlace
    pngChunk(f, "IHDR", ihdr);
    // raw scanlines with filter byte 0
    std::vector<uint8_t> raw;
    raw.reserve((size_t) H * (3 * (size_t) W + 1));
    for (int y = 0; y < H; y++) {
        raw.push_back(0);
        raw.insert(raw.end(), rgb + 3 * (size_t) y * W, rgb + 3 * (size_t) (y + 1) * W);
    }
    std::vector<uint8_t> z = {0x78, 0x01};
    uint32_t a = 1, b2 = 0;  // Adler-32
    for (uint8_t v : raw) {
        a = (a + v) % 65521u;
        b2 = (b2 + a) % 65521u;
    }
    for (size_t off = 0; off < raw.size() || off == 0; off += 65535) {
        const size_t n = std::min<size_t>(65535, raw.size() - off);
        const bool last = off + n >= raw.size();
        z.push_back(last ? 1 : 0);
        z.push_back((uint8_t) n);
        z.push_back((uint8_t) (n >> 8));
        z.push_back((uint8_t) ~n);
        z.push_back((uint8_t) (~n >> 8));
        z.insert(z.end(), raw.begin() + off, raw.begin() + off + n);
        if (last) break;
    }
    put32be(z, (b2 << 16) | a);
    pngChunk(f, "IDAT", z);
    pngChunk(f, "IEND", {});
    return fclose(f) == 0;
}

static bool writePpm(const std::string &path, int W, int H, const uint8_t *rgb) {
    FILE *f = fopen(path.c_str(), "wb");
    if (!f) return false;
    fprintf(f, "P6\n%d %d\n255\n", W, H);
    fwrite(rgb, 1, 3 * (size_t) W * H, f);
    return fclose(f) == 0;
}

// ---------------------------------------------------------------------------
// Renderer::Render (Renderer.cpp:24-212)
// ---------------------------------------------------------------------------
void Renderer::Render(const Scene &scene, const Integrator &integrator, const RenderSettings &settings) {
    prr_detail::SceneDevice *D = scene.device();
    if (!D->scene) throw std::runtime_error("Renderer::Render before Scene::buildBVH");
    prr_context *ctx = context();
    const int W = settings.width, H = settings.height;
    const size_t P = (size_t) W * H;
    printf("SPP: min=%d max=%d threshold=%g\n", settings.minSPP, settings.maxSPP, settings.varianceThreshold);
    printf("Tone mapper: %s\n", tonemap::toString(settings.toneMapper));

    std::lock_guard<std::mutex> lk(D->mu);
    // device: rgb (3P floats), counts (P ints), display bytes (3P)
    char *b = (char *) D->buf.get(P * (12 + 4 + 3) + 256);
    float *dRgb = (float *) b;
    int32_t *dCnt = (int32_t *) (dRgb + 3 * P);
    uint8_t *dByte = (uint8_t *) (dCnt + P);

    prr_render_params p{};
    p.width = W;
    p.height = H;
    p.spp = settings.maxSPP;
    p.min_spp = settings.minSPP;
    p.variance_threshold = settings.varianceThreshold;
    p.max_depth = integrator.depth();
    p.indirect_clamp = integrator.indirectClamp;
    p.seed = settings.seed;
    p.num_photons = D->photons;
    p.photon_radius = 0.01f;
    p.photon_max_depth = 30;
    prr_stats st{};
    check(prr_render(ctx, D->scene, &p, dRgb, dCnt, &st), "prr_render");

    std::vector<float> rgb(3 * P);
    std::vector<int32_t> cnt(P);
    if (hipMemcpy(rgb.data(), dRgb, rgb.size() * 4, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(cnt.data(), dCnt, cnt.size() * 4, hipMemcpyDeviceToHost) != hipSuccess)
        throw std::runtime_error("hipMemcpy");
    lastImage.resize(P);
    for (size_t i = 0; i < P; i++) lastImage[i] = Vector3f(rgb[3 * i], rgb[3 * i + 1], rgb[3 * i + 2]);

    long long total = 0;
    int minS = settings.maxSPP, maxS = 0;
    for (int s : cnt) {
        total += s;
        minS = std::min(minS, s);
        maxS = std::max(maxS, s);
    }
    printf("Adaptive sampling: min=%d max=%d avg=%.1f (budget=%d)\n", minS, maxS, (double) total / (double) P,
           settings.maxSPP);
    double avg[3] = {0, 0, 0};
    float mx[3] = {0, 0, 0};
    for (size_t i = 0; i < P; i++)
        for (int k = 0; k < 3; k++) {
            avg[k] += rgb[3 * i + k];
            mx[k] = std::max(mx[k], rgb[3 * i + k]);
        }
    printf("Avg radiance: (%.4f, %.4f, %.4f)\n", avg[0] / P, avg[1] / P, avg[2] / P);
    printf("Max radiance: (%.4f, %.4f, %.4f)\n", mx[0], mx[1], mx[2]);
    printf("Device: %.1f ms, %llu rays, %llu samples\n", st.ms_total,
           (unsigned long long) (st.rays_closest + st.rays_any), (unsigned long long) st.samples);

    std::vector<uint8_t> bytes(3 * P);
    auto renderAndWrite = [&](tonemap::ToneMapper m, const std::string &base) {
        check(prr_tonemap(ctx, tonemap::op(m), (int64_t) P, settings.exposure, dRgb, nullptr, dByte), "prr_tonemap");
        check(prr_synchronize(ctx), "prr_synchronize");
        if (hipMemcpy(bytes.data(), dByte, bytes.size(), hipMemcpyDeviceToHost) != hipSuccess)
            throw std::runtime_error("hipMemcpy");
        writePpm(base + ".ppm", W, H, bytes.data());
        writePng(base + ".png", W, H, bytes.data());
    };
    renderAndWrite(settings.toneMapper, settings.outputBase);
    if (settings.toneMapper != tonemap::ToneMapper::Raw) renderAndWrite(tonemap::ToneMapper::Raw, settings.outputBase + "_raw");

    // Renderer.cpp:197-211: samples taken as grey levels
    for (size_t i = 0; i < P; i++) {
        const uint8_t v = (uint8_t) (255 * (float) cnt[i] / (float) settings.maxSPP);
        bytes[3 * i] = bytes[3 * i + 1] = bytes[3 * i + 2] = v;
    }
    writePpm(settings.sampleCountBase + ".ppm", W, H, bytes.data());
    writePng(settings.sampleCountBase + ".png", W, H, bytes.data());
}
