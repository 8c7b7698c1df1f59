// ref_harness.cpp — golden-vector generator linked against the reference's
// own sources (built by oracle/Makefile into oracle/_ref/, never shipped).
// TEST INFRASTRUCTURE ONLY.
//
// Usage: ref_harness <mode> <in.bundle> <out.bundle>
//   modes: trace, bsdf, env, camera, tonemap, castray, image, photon, welford, hdr, gltf,
//          bench (only in the ref_bench build: PRR_REF_BENCH)
// A bundle is a sequence of records {char name[32]; int32 dtype; int64 count;
// payload}, dtype 0=f32 1=i32 2=u64 3=u8 4=u32 (tests/golden/bundle.py).
//
// Reference private members (BVHAccel::nodes/orderedPrims, EnvMap::buildCDF)
// are reached by compiling the reference headers with `private` redefined as
// `public` in this translation unit only.
#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <functional>
#include <iostream>
#include <limits>
#include <map>
#include <numeric>
#include <queue>
#include <random>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#define private public
#include "BVH.hpp"
#include "EnvMap.hpp"
#include "GLTFLoader.hpp"
#include "Integrator.hpp"
#include "Material.hpp"
#include "Renderer.hpp"
#include "Scene.hpp"
#include "Sphere.hpp"
#include "ToneMapping.hpp"
#include "Triangle.hpp"
#undef private

// ---------------------------------------------------------------------------
// bundle I/O
// ---------------------------------------------------------------------------
struct Arr {
    int dtype = 0;
    std::vector<uint8_t> bytes;
    template <class T>
    const T *as() const { return reinterpret_cast<const T *>(bytes.data()); }
    template <class T>
    size_t n() const { return bytes.size() / sizeof(T); }
};
using Bundle = std::map<std::string, Arr>;

static Bundle readBundle(const char *path) {
    Bundle b;
    FILE *f = fopen(path, "rb");
    if (!f) {
        fprintf(stderr, "cannot open %s\n", path);
        exit(2);
    }
    while (true) {
        char name[32];
        if (fread(name, 1, 32, f) != 32) break;
        int32_t dt;
        int64_t cnt;
        if (fread(&dt, 4, 1, f) != 1 || fread(&cnt, 8, 1, f) != 1) break;
        static const int sz[] = {4, 4, 8, 1, 4};
        Arr a;
        a.dtype = dt;
        a.bytes.resize((size_t) cnt * sz[dt]);
        if (cnt && fread(a.bytes.data(), 1, a.bytes.size(), f) != a.bytes.size()) break;
        name[31] = 0;
        b[name] = std::move(a);
    }
    fclose(f);
    return b;
}

struct Writer {
    FILE *f;
    explicit Writer(const char *path) : f(fopen(path, "wb")) {
        if (!f) {
            fprintf(stderr, "cannot write %s\n", path);
            exit(2);
        }
    }
    ~Writer() { fclose(f); }
    void put(const std::string &name, int dt, const void *data, int64_t cnt) {
        char nm[32] = {0};
        strncpy(nm, name.c_str(), 31);
        static const int sz[] = {4, 4, 8, 1, 4};
        fwrite(nm, 1, 32, f);
        int32_t d = dt;
        fwrite(&d, 4, 1, f);
        fwrite(&cnt, 8, 1, f);
        if (cnt) fwrite(data, sz[dt], (size_t) cnt, f);
    }
    void f32(const std::string &n, const std::vector<float> &v) { put(n, 0, v.data(), (int64_t) v.size()); }
    void i32(const std::string &n, const std::vector<int32_t> &v) { put(n, 1, v.data(), (int64_t) v.size()); }
    void u32(const std::string &n, const std::vector<uint32_t> &v) { put(n, 4, v.data(), (int64_t) v.size()); }
    void u8(const std::string &n, const std::vector<uint8_t> &v) { put(n, 3, v.data(), (int64_t) v.size()); }
};

static void seedThread(uint32_t v) { prr_refseed::value().store(v); }

// Runs fn on a fresh thread so that the reference's thread_local RNG is
// initialised from seed v ((v<<32)|v).
template <class F>
static void onFreshThread(uint32_t v, F fn) {
    seedThread(v);
    std::thread th(fn);
    th.join();
}

// ---------------------------------------------------------------------------
// scene construction from a bundle
// objects i32 n*4 (kind, material, first, count); tris f32 n*9; spheres f32
// n*4; materials f32 n*12 (type, base3, rough, metal, ior, emission3, pad2)
// ---------------------------------------------------------------------------
struct RefScene {
    Scene scene;
    std::vector<Object *> objs;
    std::vector<MeshTriangle *> meshOf;  // per object (null for spheres)
    std::vector<int> firstOf;
    int numTris = 0;
    std::unordered_map<const Object *, int> primId;  // Triangle*/Sphere* -> global id
};

static Material *makeMaterial(const float *m) {
    Material *mat = nullptr;
    switch ((int) m[0]) {
    case 0: mat = new DiffuseMaterial(); break;
    case 1: mat = new GlassMaterial(); break;
    case 2: mat = new MirrorMaterial(); break;
    default: mat = new EmissiveMaterial(); break;
    }
    mat->baseColor = Vector3f(m[1], m[2], m[3]);
    mat->roughness = m[4];
    mat->metallic = m[5];
    mat->ior = m[6];
    mat->m_emission = Vector3f(m[7], m[8], m[9]);
    return mat;
}

static void buildScene(const Bundle &b, RefScene &rs) {
    const auto &O = b.at("objects");
    const auto &T = b.at("tris");
    const auto &M = b.at("materials");
    const int nObj = (int) O.n<int32_t>() / 4;
    const int nMat = (int) M.n<float>() / 12;
    std::vector<Material *> mats;
    for (int i = 0; i < nMat; i++) mats.push_back(makeMaterial(M.as<float>() + 12 * i));
    // textures: mat_tex (4 per material: base colour, metallic-roughness, normal, emissive)
    if (b.count("mat_tex")) {
        const int32_t *mt = b.at("mat_tex").as<int32_t>();
        const int32_t *whc = b.at("tex_whc").as<int32_t>();
        const uint8_t *px = b.at("texels").as<uint8_t>();
        std::vector<size_t> off(1, 0);
        for (size_t t = 0; t < b.at("tex_whc").n<int32_t>() / 3; t++)
            off.push_back(off.back() + (size_t) whc[3 * t] * whc[3 * t + 1] * whc[3 * t + 2]);
        for (int i = 0; i < nMat; i++) {
            Texture *slots[4] = {&mats[i]->baseColorTex, &mats[i]->metallicRoughnessTex, &mats[i]->normalTex,
                                 &mats[i]->emissiveTex};
            for (int j = 0; j < 4; j++) {
                const int t = mt[4 * i + j];
                if (t < 0) continue;
                slots[j]->width = whc[3 * t];
                slots[j]->height = whc[3 * t + 1];
                slots[j]->channels = whc[3 * t + 2];
                slots[j]->data.assign(px + off[t], px + off[t + 1]);
            }
        }
    }
    const int32_t *tfl = b.count("tri_flags") ? b.at("tri_flags").as<int32_t>() : nullptr;
    const float *tat = b.count("tri_attr") ? b.at("tri_attr").as<float>() : nullptr;
    rs.numTris = (int) T.n<float>() / 9;
    const float *sp = b.count("spheres") ? b.at("spheres").as<float>() : nullptr;
    for (int k = 0; k < nObj; k++) {
        const int32_t *o = O.as<int32_t>() + 4 * k;
        Material *mt = mats[o[1]];
        if (o[0] == 0) {
            std::vector<Triangle> tris;
            Vector3f mn(std::numeric_limits<float>::infinity()), mx(-std::numeric_limits<float>::infinity());
            for (int i = 0; i < o[3]; i++) {
                const float *t = T.as<float>() + 9 * (size_t) (o[2] + i);
                Vector3f v[3] = {Vector3f(t[0], t[1], t[2]), Vector3f(t[3], t[4], t[5]),
                                 Vector3f(t[6], t[7], t[8])};
                for (auto &p : v) {
                    mn = Vector3f(std::min(mn.x, p.x), std::min(mn.y, p.y), std::min(mn.z, p.z));
                    mx = Vector3f(std::max(mx.x, p.x), std::max(mx.y, p.y), std::max(mx.z, p.z));
                }
                tris.emplace_back(v[0], v[1], v[2], mt);
                Triangle &tr = tris.back();
                if (!tat) {
                    tr.t1 = Vector2f(1.f, 0.f);  // tcoords = (u, v)
                    tr.t2 = Vector2f(0.f, 1.f);
                } else {  // GLTFLoader.cpp:244-268: n0..n2 | t0..t2 | tangent xyz w x3
                    const size_t g = (size_t) (o[2] + i);
                    const float *a = tat + 27 * g;
                    if (tfl[g] & 1) {
                        tr.n0 = Vector3f(a[0], a[1], a[2]);
                        tr.n1 = Vector3f(a[3], a[4], a[5]);
                        tr.n2 = Vector3f(a[6], a[7], a[8]);
                        tr.hasSmoothNormals = true;
                        tr.normal = normalize(tr.n0 + tr.n1 + tr.n2);
                    }
                    tr.t0 = Vector2f(a[9], a[10]);
                    tr.t1 = Vector2f(a[11], a[12]);
                    tr.t2 = Vector2f(a[13], a[14]);
                    if (tfl[g] & 2) {
                        tr.tan0 = Vector3f(a[15], a[16], a[17]);
                        tr.tangentW0 = a[18];
                        tr.tan1 = Vector3f(a[19], a[20], a[21]);
                        tr.tangentW1 = a[22];
                        tr.tan2 = Vector3f(a[23], a[24], a[25]);
                        tr.tangentW2 = a[26];
                        tr.hasTangents = true;
                    }
                }
            }
            MeshTriangle *m = new MeshTriangle(tris, Bounds3(mn, mx), mt);
            for (int i = 0; i < (int) m->triangles.size(); i++) rs.primId[&m->triangles[i]] = o[2] + i;
            rs.objs.push_back(m);
            rs.meshOf.push_back(m);
        } else {
            const float *s = sp + 4 * o[2];
            Sphere *sph = new Sphere(Vector3f(s[0], s[1], s[2]), s[3], mt);
            rs.primId[sph] = rs.numTris + o[2];
            rs.objs.push_back(sph);
            rs.meshOf.push_back(nullptr);
        }
        rs.firstOf.push_back(o[2]);
        rs.scene.Add(rs.objs.back());
    }
    if (b.count("background")) {
        const float *bg = b.at("background").as<float>();
        rs.scene.backgroundColor = Vector3f(bg[0], bg[1], bg[2]);
    }
    if (b.count("env_pixels")) {
        const int32_t *wh = b.at("env_wh").as<int32_t>();
        const float *px = b.at("env_pixels").as<float>();
        EnvMap &e = rs.scene.envMap;
        e.width = wh[0];
        e.height = wh[1];
        e.pixels.resize((size_t) wh[0] * wh[1]);
        for (size_t i = 0; i < e.pixels.size(); i++) e.pixels[i] = Vector3f(px[3 * i], px[3 * i + 1], px[3 * i + 2]);
        e.buildCDF();
        e.loaded = true;
    }
    if (b.count("camera")) {
        const float *c = b.at("camera").as<float>();
        Camera &cam = rs.scene.camera;
        cam.eye = Vector3f(c[0], c[1], c[2]);
        cam.target = Vector3f(c[3], c[4], c[5]);
        cam.up = Vector3f(c[6], c[7], c[8]);
        cam.fovh = c[9];
        cam.fovv = c[9];
        cam.aperture = c[10];
        cam.focusDistance = c[11];
    }
    rs.scene.buildBVH();
}

static void exportTree(Writer &w, const std::string &pfx, const BVHAccel *bvh,
                       const std::function<int(const Object *)> &idOf) {
    std::vector<float> bnd;
    std::vector<uint32_t> fc, meta;
    std::vector<int32_t> order;
    std::vector<float> areas(bvh->nodeAreas.begin(), bvh->nodeAreas.end());
    for (const BVHNode &n : bvh->nodes) {
        for (int k = 0; k < 6; k++) bnd.push_back(n.bounds[k]);
        fc.push_back(n.first_child);
        meta.push_back(n.meta);
    }
    for (const Object *o : bvh->orderedPrims) order.push_back(idOf(o));
    w.f32(pfx + "_bounds", bnd);
    w.u32(pfx + "_first", fc);
    w.u32(pfx + "_meta", meta);
    w.i32(pfx + "_order", order);
    w.f32(pfx + "_areas", areas);
}

// ---------------------------------------------------------------------------
static int modeTrace(const Bundle &b, const char *out) {
    RefScene rs;
    buildScene(b, rs);
    const auto &org = b.at("org"), &dir = b.at("dir"), &tmax = b.at("tmax");
    const size_t n = org.n<float>() / 3;
    std::vector<int32_t> prim(n);
    std::vector<float> t(n), u(n), v(n);
    std::vector<uint8_t> any(n);
    for (size_t i = 0; i < n; i++) {
        const float *o = org.as<float>() + 3 * i, *d = dir.as<float>() + 3 * i;
        Ray r(Vector3f(o[0], o[1], o[2]), Vector3f(d[0], d[1], d[2]));
        Intersection h = rs.scene.intersect(r);
        if (h.happened) {
            prim[i] = rs.primId.at(h.obj);
            t[i] = (float) h.tnear;
            u[i] = h.tcoords.x;
            v[i] = h.tcoords.y;
        } else {
            prim[i] = -1;
            t[i] = std::numeric_limits<float>::infinity();
            u[i] = v[i] = 0.f;
        }
        any[i] = rs.scene.intersectP(r, tmax.as<float>()[i]) ? 1 : 0;
    }
    Writer w(out);
    w.i32("prim", prim);
    w.f32("t", t);
    w.f32("u", u);
    w.f32("v", v);
    w.u8("any", any);
    w.f32("total_emit_area", {rs.scene.totalEmitArea});
    std::unordered_map<const Object *, int> objIdx;
    for (size_t k = 0; k < rs.objs.size(); k++) objIdx[rs.objs[k]] = (int) k;
    exportTree(w, "top", rs.scene.bvh, [&](const Object *o) { return objIdx.at(o); });
    for (size_t k = 0; k < rs.objs.size(); k++) {
        MeshTriangle *m = rs.meshOf[k];
        if (!m) continue;
        const int first = rs.firstOf[k];
        exportTree(w, "mesh" + std::to_string(k), m->bvh,
                   [&](const Object *o) { return rs.primId.at(o) - first; });
    }
    // light sampling (Scene::sampleLight, Q4 leaf walk) with seeded streams
    if (b.count("light_seeds")) {
        const auto &ls = b.at("light_seeds");
        const size_t m = ls.n<uint32_t>();
        std::vector<float> pos(3 * m), nrm(3 * m), pdf(m);
        for (size_t i = 0; i < m; i++) {
            onFreshThread(ls.as<uint32_t>()[i], [&]() {
                Intersection li;
                float p = 0.f;
                rs.scene.sampleLight(li, p);
                pos[3 * i] = li.coords.x;
                pos[3 * i + 1] = li.coords.y;
                pos[3 * i + 2] = li.coords.z;
                nrm[3 * i] = li.normal.x;
                nrm[3 * i + 1] = li.normal.y;
                nrm[3 * i + 2] = li.normal.z;
                pdf[i] = p;
            });
        }
        w.f32("light_pos", pos);
        w.f32("light_nrm", nrm);
        w.f32("light_pdf", pdf);
    }
    return 0;
}

static int modeBsdf(const Bundle &b, const char *out) {
    const auto &wi = b.at("wi"), &wo = b.at("wo"), &N = b.at("n"), &P = b.at("params"), &S = b.at("seeds");
    const size_t n = S.n<uint32_t>();
    std::vector<float> f(3 * n), pdf(n), swi(3 * n), sf(3 * n), spdf(n), draws(3 * n);
    std::vector<int32_t> lobe(n);
    for (size_t i = 0; i < n; i++) {
        const float *p = P.as<float>() + 8 * i;
        float m12[12] = {p[0], p[1], p[2], p[3], p[4], p[5], p[6], 0, 0, 0, 0, 0};
        Material *mat = makeMaterial(m12);
        Vector3f vi(wi.as<float>()[3 * i], wi.as<float>()[3 * i + 1], wi.as<float>()[3 * i + 2]);
        Vector3f vo(wo.as<float>()[3 * i], wo.as<float>()[3 * i + 1], wo.as<float>()[3 * i + 2]);
        Vector3f vn(N.as<float>()[3 * i], N.as<float>()[3 * i + 1], N.as<float>()[3 * i + 2]);
        ShadingData sd = mat->buildShadingData(Vector2f(0, 0), vn);
        Vector3f e = mat->eval(vi, vo, sd, LOBE_ALL);
        f[3 * i] = e.x;
        f[3 * i + 1] = e.y;
        f[3 * i + 2] = e.z;
        pdf[i] = mat->pdf(vi, vo, sd, LOBE_ALL);
        const uint32_t seed = S.as<uint32_t>()[i];
        onFreshThread(seed, [&]() {
            BSDFSample s = mat->sample(vo, sd);
            swi[3 * i] = s.wi.x;
            swi[3 * i + 1] = s.wi.y;
            swi[3 * i + 2] = s.wi.z;
            sf[3 * i] = s.f.x;
            sf[3 * i + 1] = s.f.y;
            sf[3 * i + 2] = s.f.z;
            spdf[i] = s.pdf;
            lobe[i] = (s.pdf > 0.f || s.lobe != 0) ? (int) s.lobe : -1;
        });
        onFreshThread(seed, [&]() {
            draws[3 * i] = get_random_float();
            draws[3 * i + 1] = get_random_float();
            draws[3 * i + 2] = get_random_float();
        });
        delete mat;
    }
    Writer w(out);
    w.f32("eval_f", f);
    w.f32("eval_pdf", pdf);
    w.f32("sample_wi", swi);
    w.f32("sample_f", sf);
    w.f32("sample_pdf", spdf);
    w.i32("sample_lobe", lobe);
    w.f32("draws", draws);
    return 0;
}

// Medium::intersect (Volume.hpp:17-44), phaseHG and samplePhaseHG
// (Volume.cpp:3-24): per entry a ray against a box (bounded or not), the HG
// phase function at (cos, g), and a phase sample about wo with its two draws.
static int modeVolume(const Bundle &b, const char *out) {
    const auto &O = b.at("org"), &D = b.at("dir"), &B = b.at("bounds"), &BD = b.at("bounded");
    const auto &C = b.at("cos"), &G = b.at("g"), &WO = b.at("wo"), &S = b.at("seeds");
    const size_t n = S.n<uint32_t>();
    std::vector<float> te(n), tx(n), phase(n), swi(3 * n), draws(2 * n);
    std::vector<int32_t> hit(n);
    for (size_t i = 0; i < n; i++) {
        const float *o = O.as<float>() + 3 * i, *d = D.as<float>() + 3 * i, *bb = B.as<float>() + 6 * i;
        Medium m;
        m.bounded = BD.as<int32_t>()[i] != 0;
        m.bounds = Bounds3(Vector3f(bb[0], bb[1], bb[2]), Vector3f(bb[3], bb[4], bb[5]));
        Ray r(Vector3f(o[0], o[1], o[2]), Vector3f(d[0], d[1], d[2]));
        float a = -7.f, e = -7.f;  // untouched on a miss
        hit[i] = m.intersect(r, a, e) ? 1 : 0;
        te[i] = a;
        tx[i] = e;
        phase[i] = phaseHG(C.as<float>()[i], G.as<float>()[i]);
        const Vector3f wo(WO.as<float>()[3 * i], WO.as<float>()[3 * i + 1], WO.as<float>()[3 * i + 2]);
        const uint32_t seed = S.as<uint32_t>()[i];
        onFreshThread(seed, [&]() {
            const Vector3f w = samplePhaseHG(wo, G.as<float>()[i]);
            swi[3 * i] = w.x;
            swi[3 * i + 1] = w.y;
            swi[3 * i + 2] = w.z;
        });
        onFreshThread(seed, [&]() {
            draws[2 * i] = get_random_float();
            draws[2 * i + 1] = get_random_float();
        });
    }
    Writer w(out);
    w.f32("t_enter", te);
    w.f32("t_exit", tx);
    w.i32("hit", hit);
    w.f32("phase", phase);
    w.f32("sample_wi", swi);
    w.f32("draws", draws);
    return 0;
}

static int modeEnv(const Bundle &b, const char *out) {
    const int32_t *wh = b.at("env_wh").as<int32_t>();
    const float *px = b.at("env_pixels").as<float>();
    EnvMap e;
    e.width = wh[0];
    e.height = wh[1];
    e.pixels.resize((size_t) wh[0] * wh[1]);
    for (size_t i = 0; i < e.pixels.size(); i++) e.pixels[i] = Vector3f(px[3 * i], px[3 * i + 1], px[3 * i + 2]);
    e.buildCDF();
    e.loaded = true;
    const auto &S = b.at("seeds");
    const size_t n = S.n<uint32_t>();
    std::vector<float> idir(3 * n), ipdf(n), draws(2 * n);
    for (size_t i = 0; i < n; i++) {
        const uint32_t seed = S.as<uint32_t>()[i];
        onFreshThread(seed, [&]() {
            float p;
            Vector3f d = e.importanceSample(p);
            idir[3 * i] = d.x;
            idir[3 * i + 1] = d.y;
            idir[3 * i + 2] = d.z;
            ipdf[i] = p;
        });
        onFreshThread(seed, [&]() {
            draws[2 * i] = get_random_float();
            draws[2 * i + 1] = get_random_float();
        });
    }
    const auto &D = b.at("dirs");
    const size_t m = D.n<float>() / 3;
    std::vector<float> rad(3 * m), pdf(m);
    for (size_t i = 0; i < m; i++) {
        Vector3f d(D.as<float>()[3 * i], D.as<float>()[3 * i + 1], D.as<float>()[3 * i + 2]);
        Vector3f r = e.sample(d);
        rad[3 * i] = r.x;
        rad[3 * i + 1] = r.y;
        rad[3 * i + 2] = r.z;
        pdf[i] = e.importanceSamplePdf(d);
    }
    Writer w(out);
    w.f32("cdf", e.cdf);
    w.f32("marg", e.marginalCdf);
    w.f32("total", {e.totalLuminance});
    w.f32("is_dir", idir);
    w.f32("is_pdf", ipdf);
    w.f32("draws", draws);
    w.f32("rad", rad);
    w.f32("pdf", pdf);
    return 0;
}

static int modeCamera(const Bundle &b, const char *out) {
    const float *c = b.at("camera").as<float>();
    const int32_t *wh = b.at("wh").as<int32_t>();
    Camera cam;
    cam.eye = Vector3f(c[0], c[1], c[2]);
    cam.target = Vector3f(c[3], c[4], c[5]);
    cam.up = Vector3f(c[6], c[7], c[8]);
    cam.fovh = c[9];
    cam.fovv = c[9];
    cam.aperture = c[10];
    cam.focusDistance = c[11];
    cam.init(wh[0], wh[1]);
    const auto &pix = b.at("pix"), &offs = b.at("offs");
    const size_t n = pix.n<int32_t>() / 2;
    std::vector<float> o(3 * n), d(3 * n);
    // with "seeds": each ray on a fresh thread whose generator starts from
    // seeds[i] (the thin lens draws its two lens uniforms, Camera.hpp:87-88)
    const Arr *S = b.count("seeds") ? &b.at("seeds") : nullptr;
    for (size_t i = 0; i < n; i++) {
        Ray r(Vector3f(0.f), Vector3f(0.f, 0.f, 1.f));
        auto gen = [&]() {
            r = cam.generateRay(pix.as<int32_t>()[2 * i], pix.as<int32_t>()[2 * i + 1], wh[0], wh[1],
                                offs.as<float>()[2 * i], offs.as<float>()[2 * i + 1]);
        };
        if (S)
            onFreshThread(S->as<uint32_t>()[i], gen);
        else
            gen();
        o[3 * i] = r.origin.x;
        o[3 * i + 1] = r.origin.y;
        o[3 * i + 2] = r.origin.z;
        d[3 * i] = r.direction.x;
        d[3 * i + 1] = r.direction.y;
        d[3 * i + 2] = r.direction.z;
    }
    Writer w(out);
    w.f32("org", o);
    w.f32("dir", d);
    return 0;
}

static int modeTonemap(const Bundle &b, const char *out) {
    const auto &L = b.at("linear");
    const size_t n = L.n<float>() / 3;
    Writer w(out);
    const tonemap::ToneMapper ops[4] = {tonemap::ToneMapper::Raw, tonemap::ToneMapper::Standard,
                                        tonemap::ToneMapper::AgX, tonemap::ToneMapper::PBRNeutral};
    const char *names[4] = {"raw", "standard", "agx", "pbrneutral"};
    for (int k = 0; k < 4; k++) {
        std::vector<float> o(3 * n);
        for (size_t i = 0; i < n; i++) {
            Vector3f c = tonemap::apply(ops[k], Vector3f(L.as<float>()[3 * i], L.as<float>()[3 * i + 1],
                                                         L.as<float>()[3 * i + 2]));
            o[3 * i] = c.x;
            o[3 * i + 1] = c.y;
            o[3 * i + 2] = c.z;
        }
        w.f32(names[k], o);
    }
    return 0;
}

// Renderer.cpp:74-85 per sample: 2 offset draws, generateRay, castRay, NaN->0
static Vector3f oneSample(const RefScene &rs, const Integrator &integ, int i, int j, int W, int H) {
    float xo = get_random_float();
    float yo = get_random_float();
    Ray ray = rs.scene.camera.generateRay(i, j, W, H, xo, yo);
    Vector3f s = integ.castRay(ray);
    if (std::isnan(s.x) || std::isnan(s.y) || std::isnan(s.z)) s = Vector3f(0.f);
    return s;
}

static int modeCastray(const Bundle &b, const char *out) {
    RefScene rs;
    buildScene(b, rs);
    const int32_t *cfg = b.at("cfg").as<int32_t>();  // W, H, maxDepth
    const int W = cfg[0], H = cfg[1];
    rs.scene.camera.init(W, H);
    Integrator integ(rs.scene, cfg[2], 0.95f);
    const auto &pix = b.at("pix"), &S = b.at("seeds");
    const size_t n = S.n<uint32_t>();
    std::vector<float> L(3 * n);
    for (size_t k = 0; k < n; k++) {
        onFreshThread(S.as<uint32_t>()[k], [&]() {
            Vector3f s = oneSample(rs, integ, pix.as<int32_t>()[2 * k], pix.as<int32_t>()[2 * k + 1], W, H);
            L[3 * k] = s.x;
            L[3 * k + 1] = s.y;
            L[3 * k + 2] = s.z;
        });
    }
    Writer w(out);
    w.f32("L", L);
    return 0;
}

// PhotonMap::emit / trace / build / estimateIrradiance (Photon.cpp:9-190)
// as the reference runs them. Photon k is emitted by the reference's own
// PhotonMap::emit(1, scene) on a fresh thread whose generator starts from
// seeds[k] (one OpenMP thread: emit's team is then the calling thread), so
// its draws are known; its stored photons are appended to the map's
// caustic_map and recorded with owner k. num_photons = 1 makes the emitted
// power Le * A (Photon.cpp:168-170), so the diffuse continuation with its
// Russian roulette (Photon.cpp:126-150) is exercised too. Then the grid is
// built and estimateIrradiance evaluated at the query points.
static int modePhoton(const Bundle &b, const char *out) {
    RefScene rs;
    buildScene(b, rs);
    const int32_t *cfg = b.at("cfg").as<int32_t>();  // maxDepth
    PhotonMap pm;
    pm.maxDepth = cfg[0];
    pm.photon_radius = b.at("radius").as<float>()[0];
    const auto &S = b.at("seeds");
    std::vector<float> ph;
    std::vector<int32_t> owner;
    for (size_t k = 0; k < S.n<uint32_t>(); k++) {
        const size_t before = pm.caustic_map.size();
        onFreshThread(S.as<uint32_t>()[k], [&]() {
#ifdef _OPENMP
            omp_set_num_threads(2);  // emit uses omp_get_max_threads() - 1 threads
#endif
            pm.emit(1, rs.scene);
        });
        for (size_t i = before; i < pm.caustic_map.size(); i++) {
            const Photon &q = pm.caustic_map[i];
            for (const Vector3f &v : {q.position, q.direction, q.power}) {
                ph.push_back(v.x);
                ph.push_back(v.y);
                ph.push_back(v.z);
            }
            owner.push_back((int32_t) k);
        }
    }
    pm.build();
    const auto &Q = b.at("queries");
    const size_t nq = Q.n<float>() / 3;
    std::vector<float> irr(3 * nq);
    for (size_t i = 0; i < nq; i++) {
        const float *q = Q.as<float>() + 3 * i;
        Vector3f e = pm.estimateIrradiance(Vector3f(q[0], q[1], q[2]), Vector3f(0.f, 1.f, 0.f));
        irr[3 * i] = e.x;
        irr[3 * i + 1] = e.y;
        irr[3 * i + 2] = e.z;
    }
    Writer w(out);
    w.f32("photons", ph);
    w.i32("owner", owner);
    w.f32("irradiance", irr);
    return 0;
}

static int modeImage(const Bundle &b, const char *out) {
    RefScene rs;
    buildScene(b, rs);
    const int32_t *cfg = b.at("cfg").as<int32_t>();  // W, H, maxDepth, spp
    const int W = cfg[0], H = cfg[1], spp = cfg[3];
    rs.scene.camera.init(W, H);
    Integrator integ(rs.scene, cfg[2], 0.95f);
    std::vector<float> mean(3 * (size_t) W * H), var(3 * (size_t) W * H);
    prr_refseed::counter_mode().store(1);
#pragma omp parallel for schedule(dynamic, 1)
    for (int j = 0; j < H; j++) {
        for (int i = 0; i < W; i++) {
            Vector3f sum(0.f);
            double m1[3] = {0, 0, 0}, m2[3] = {0, 0, 0};
            for (int s = 0; s < spp; s++) {
                Vector3f c = oneSample(rs, integ, i, j, W, H);
                sum += c;
                double cv[3] = {c.x, c.y, c.z};
                for (int k = 0; k < 3; k++) {
                    double d = cv[k] - m1[k];
                    m1[k] += d / (s + 1);
                    m2[k] += d * (cv[k] - m1[k]);
                }
            }
            size_t p = (size_t) j * W + i;
            mean[3 * p] = sum.x / (float) spp;
            mean[3 * p + 1] = sum.y / (float) spp;
            mean[3 * p + 2] = sum.z / (float) spp;
            for (int k = 0; k < 3; k++) var[3 * p + k] = spp > 1 ? (float) (m2[k] / (spp - 1)) : 0.f;
        }
    }
    Writer w(out);
    w.f32("mean", mean);
    w.f32("var", var);
    return 0;
}

// Renderer.cpp:71-99 per pixel on given sample sequences, with the
// reference's own Welford (Renderer.hpp:14-40): samples taken until n reaches
// len or, once n >= minSPP, all three CIs are below the threshold.
static int modeWelford(const Bundle &b, const char *out) {
    const int32_t *cfg = b.at("cfg").as<int32_t>();  // nseq, len
    const int nseq = cfg[0], len = cfg[1];
    const float *seq = b.at("seq").as<float>();
    const int32_t *minSpp = b.at("minspp").as<int32_t>();
    const float *thr = b.at("thr").as<float>();
    std::vector<int32_t> count(nseq);
    std::vector<float> mean(3 * (size_t) nseq);
    for (int q = 0; q < nseq; q++) {
        Welford wR, wG, wB;
        Vector3f fb(0.f);
        int n = 0;
        while (n < len) {
            const float *x = seq + 3 * ((size_t) q * len + n);
            Vector3f s(x[0], x[1], x[2]);
            fb += s;
            n++;
            wR.update(s.x);
            wG.update(s.y);
            wB.update(s.z);
            if (n >= minSpp[q])
                if (wR.ci() < thr[q] && wG.ci() < thr[q] && wB.ci() < thr[q]) break;
        }
        fb = fb / (float) n;
        count[q] = n;
        mean[3 * q] = fb.x;
        mean[3 * q + 1] = fb.y;
        mean[3 * q + 2] = fb.z;
    }
    Writer w(out);
    w.i32("n", count);
    w.f32("mean", mean);
    return 0;
}

// EnvMap::load (EnvMap.hpp:35-55) of the .hdr file named in the bundle:
// decoded linear pixels.
static int modeHdr(const Bundle &b, const char *out) {
    const auto &pa = b.at("path");
    std::string path(reinterpret_cast<const char *>(pa.bytes.data()), pa.bytes.size());
    EnvMap env;
    if (!env.load(path)) return 1;
    std::vector<float> px(3 * env.pixels.size());
    for (size_t i = 0; i < env.pixels.size(); i++) {
        px[3 * i] = env.pixels[i].x;
        px[3 * i + 1] = env.pixels[i].y;
        px[3 * i + 2] = env.pixels[i].z;
    }
    Writer w(out);
    w.i32("wh", {env.width, env.height});
    w.f32("pixels", px);
    return 0;
}

// GLTFLoader::load of one file: per mesh its triangles (vertices, flat and
// per-vertex normals, uvs, tangents, flags) and its material (class, factors,
// the four textures' texels).
static int modeGltf(const Bundle &b, const char *out) {
    const auto &pa = b.at("path");
    std::string path(reinterpret_cast<const char *>(pa.bytes.data()), pa.bytes.size());
    const int type = b.at("type").as<int32_t>()[0];
    const MaterialType mt[4] = {MaterialType::Diffuse, MaterialType::Glass, MaterialType::Mirror, MaterialType::Emit};
    std::vector<MeshTriangle *> meshes = GLTFLoader::load(path, mt[type]);
    std::vector<int32_t> counts, flags, mtype, tw;
    std::vector<float> tri, mat;
    std::vector<uint8_t> texels;
    auto put3 = [&](const Vector3f &v) { tri.insert(tri.end(), {v.x, v.y, v.z}); };
    auto put2 = [&](const Vector2f &v) { tri.insert(tri.end(), {v.x, v.y}); };
    for (MeshTriangle *m : meshes) {
        counts.push_back((int32_t) m->triangles.size());
        for (const Triangle &t : m->triangles) {
            put3(t.v0), put3(t.v1), put3(t.v2), put3(t.normal);
            put3(t.n0), put3(t.n1), put3(t.n2);
            put2(t.t0), put2(t.t1), put2(t.t2);
            put3(t.tan0), put3(t.tan1), put3(t.tan2);
            tri.insert(tri.end(), {t.tangentW0, t.tangentW1, t.tangentW2});
            flags.push_back((t.hasSmoothNormals ? 1 : 0) | (t.hasTangents ? 2 : 0));
        }
        Material *M = m->m;
        mtype.push_back(dynamic_cast<GlassMaterial *>(M) ? 1 : dynamic_cast<MirrorMaterial *>(M) ? 2
                        : dynamic_cast<EmissiveMaterial *>(M) ? 3 : 0);
        mat.insert(mat.end(), {M->baseColor.x, M->baseColor.y, M->baseColor.z, M->roughness, M->metallic, M->ior,
                               M->m_emission.x, M->m_emission.y, M->m_emission.z});
        for (const Texture *t : {&M->baseColorTex, &M->metallicRoughnessTex, &M->normalTex, &M->emissiveTex}) {
            tw.insert(tw.end(), {t->width, t->height, t->channels, (int32_t) t->data.size()});
            texels.insert(texels.end(), t->data.begin(), t->data.end());
        }
    }
    Writer w(out);
    w.i32("counts", counts);
    w.i32("flags", flags);
    w.f32("tri", tri);
    w.i32("mtype", mtype);
    w.f32("mat", mat);
    w.i32("tex", tw);
    w.u8("texels", texels);
    return 0;
}

#ifdef PRR_REF_BENCH
// Scene-level query counts without editing the reference: the link wraps
// Scene::intersect / Scene::intersectP (-Wl,--wrap, oracle/Makefile), so every
// call from castRay's translation unit passes through these counters. The
// link-time-optimised build (PRR_REF_BENCH_NOCOUNT) has no wrap and counts
// nothing: its samples per second are converted with the counting build's
// queries per sample.
static thread_local uint64_t t_closest = 0, t_any = 0;
#ifndef PRR_REF_BENCH_NOCOUNT
Intersection realIntersect(const Scene *, const Ray &) __asm__("__real__ZNK5Scene9intersectERK3Ray");
bool realIntersectP(const Scene *, const Ray &, float) __asm__("__real__ZNK5Scene10intersectPERK3Rayf");
Intersection wrapIntersect(const Scene *s, const Ray &r) __asm__("__wrap__ZNK5Scene9intersectERK3Ray");
bool wrapIntersectP(const Scene *s, const Ray &r, float t) __asm__("__wrap__ZNK5Scene10intersectPERK3Rayf");
Intersection wrapIntersect(const Scene *s, const Ray &r) {
    t_closest++;
    return realIntersect(s, r);
}
bool wrapIntersectP(const Scene *s, const Ray &r, float t) {
    t_any++;
    return realIntersectP(s, r, t);
}
#endif

// The reference's own Render sample loop (Renderer.cpp:63-99, fixed spp,
// rows handed out one at a time) on rows [row0, row1), timed with the wall
// clock; threads stop taking rows once budget_ms have passed (0: no budget).
// The scene build is not timed.
static int modeBench(const Bundle &b, const char *out) {
    RefScene rs;
    const auto tb0 = std::chrono::steady_clock::now();
    buildScene(b, rs);
    const double buildS = std::chrono::duration<double>(std::chrono::steady_clock::now() - tb0).count();
    const int32_t *cfg = b.at("cfg").as<int32_t>();  // W, H, maxDepth, spp, row0, row1, threads, budget_ms
    const int W = cfg[0], H = cfg[1], spp = cfg[3], row0 = cfg[4], row1 = cfg[5], threads = cfg[6];
    const double budget = cfg[7] / 1000.0;
    rs.scene.camera.init(W, H);
    Integrator integ(rs.scene, cfg[2], 0.95f);
    prr_refseed::counter_mode().store(1);
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#endif
    uint64_t closest = 0, any = 0, samples = 0;
    int used = 1;
    std::atomic<int> nextRow(row0);
    const auto t0 = std::chrono::steady_clock::now();
#pragma omp parallel reduction(+ : closest, any, samples)
    {
        t_closest = t_any = 0;
        while (true) {
            if (budget > 0 &&
                std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() > budget)
                break;
            const int j = nextRow.fetch_add(1);
            if (j >= row1) break;
            for (int i = 0; i < W; i++)
                for (int k = 0; k < spp; k++) {
                    oneSample(rs, integ, i, j, W, H);
                    samples++;
                }
        }
        closest += t_closest;
        any += t_any;
#ifdef _OPENMP
#pragma omp single
        used = omp_get_num_threads();
#endif
    }
    const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    Writer w(out);
    w.f32("seconds", {(float) dt, (float) buildS});
    w.u32("counts", {(uint32_t) (closest >> 32), (uint32_t) closest, (uint32_t) (any >> 32), (uint32_t) any,
                     (uint32_t) (samples >> 32), (uint32_t) samples, (uint32_t) used});
    return 0;
}
#endif

int main(int argc, char **argv) {
    if (argc != 4) {
        fprintf(stderr, "usage: %s <mode> <in> <out>\n", argv[0]);
        return 2;
    }
    Bundle b = readBundle(argv[2]);
    std::string m = argv[1];
    if (m == "trace") return modeTrace(b, argv[3]);
    if (m == "bsdf") return modeBsdf(b, argv[3]);
    if (m == "env") return modeEnv(b, argv[3]);
    if (m == "camera") return modeCamera(b, argv[3]);
    if (m == "tonemap") return modeTonemap(b, argv[3]);
    if (m == "castray") return modeCastray(b, argv[3]);
    if (m == "image") return modeImage(b, argv[3]);
    if (m == "photon") return modePhoton(b, argv[3]);
    if (m == "welford") return modeWelford(b, argv[3]);
    if (m == "volume") return modeVolume(b, argv[3]);
    if (m == "hdr") return modeHdr(b, argv[3]);
    if (m == "gltf") return modeGltf(b, argv[3]);
#ifdef PRR_REF_BENCH
    if (m == "bench") return modeBench(b, argv[3]);
#endif
    fprintf(stderr, "unknown mode %s\n", argv[1]);
    return 2;
}
