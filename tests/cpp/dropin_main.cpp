// dropin_main.cpp — a scene script written against the reference's API
// (main.cpp:11-52 pattern) and linked to libprr.so instead of the CPU
// renderer. Used by tests/test_dropin_cpp.py.
//
//   dropin_main render <outdir> <W> <H> <maxSPP> <minSPP> <threshold> <seed>
//       box room (Cornell-style, 5 walls, quad light, metal + glass spheres)
//       through Scene::buildBVH + Renderer::Render; writes the reference's
//       output files into <outdir>, the linear image to <outdir>/linear.f32
//       and single-query results (Scene::intersect / intersectP,
//       Integrator::castRay) to <outdir>/queries.txt
//   dropin_main api <outdir>   BVHAccel, Scene::sampleLight / pdfLight and the
//       photon map on the box room (<outdir>/api.txt, photons.f32)
//   dropin_main hdr <file.hdr> <out.f32>     EnvMap::load; w, h, pixels
//   dropin_main obj <file.obj> <dx> <dy> <dz> MeshTriangle from OBJ; prints
//       triangle count, area and bounds
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "prr/prr.hpp"

static std::vector<Triangle> quad(Vector3f p0, Vector3f p1, Vector3f p2, Vector3f p3, Material *m) {
    return {Triangle(p0, p1, p2, m), Triangle(p0, p2, p3, m)};
}

static MeshTriangle *mesh(std::vector<Triangle> tris, Material *m) {
    Bounds3 b;
    for (auto &t : tris) b = Union(b, t.getBounds());
    return new MeshTriangle(std::move(tris), b, m);
}

static Scene buildScene() {
    Scene scene;
    scene.backgroundColor = Vector3f(0.f);
    auto *grey = new DiffuseMaterial();
    grey->baseColor = Vector3f(0.73f, 0.73f, 0.73f);
    grey->roughness = 1.0f;
    auto *red = new DiffuseMaterial();
    red->baseColor = Vector3f(0.65f, 0.05f, 0.05f);
    red->roughness = 1.0f;
    auto *green = new DiffuseMaterial();
    green->baseColor = Vector3f(0.12f, 0.45f, 0.15f);
    green->roughness = 1.0f;
    auto *light = new EmissiveMaterial();
    light->m_emission = Vector3f(17.f, 12.f, 4.f);
    auto *metal = new DiffuseMaterial();
    metal->baseColor = Vector3f(0.95f, 0.64f, 0.54f);
    metal->roughness = 0.3f;
    metal->metallic = 1.0f;
    auto *glass = new GlassMaterial();
    glass->baseColor = Vector3f(1.f);
    scene.Add(mesh(quad({-1, 0, -1}, {-1, 0, 1}, {1, 0, 1}, {1, 0, -1}, grey), grey));
    scene.Add(mesh(quad({-1, 2, -1}, {1, 2, -1}, {1, 2, 1}, {-1, 2, 1}, grey), grey));
    scene.Add(mesh(quad({-1, 0, -1}, {1, 0, -1}, {1, 2, -1}, {-1, 2, -1}, grey), grey));
    scene.Add(mesh(quad({-1, 0, -1}, {-1, 2, -1}, {-1, 2, 1}, {-1, 0, 1}, red), red));
    scene.Add(mesh(quad({1, 0, -1}, {1, 0, 1}, {1, 2, 1}, {1, 2, -1}, green), green));
    scene.Add(mesh(quad({-0.25f, 1.98f, -0.25f}, {0.25f, 1.98f, -0.25f}, {0.25f, 1.98f, 0.25f},
                        {-0.25f, 1.98f, 0.25f}, light),
                   light));
    scene.Add(new Sphere(Vector3f(-0.45f, 0.4f, -0.3f), 0.4f, metal));
    scene.Add(new Sphere(Vector3f(0.45f, 0.4f, 0.25f), 0.4f, glass));
    scene.camera = Camera::create(Vector3f(0, 1, 3.9f), Vector3f(0, 1, 0), Vector3f(0, 1, 0), 40.f, 40.f);
    return scene;
}

static int render(int argc, char **argv) {
    if (argc < 9) return 2;
    const std::string out = argv[2];
    RenderSettings settings;
    settings.width = std::atoi(argv[3]);
    settings.height = std::atoi(argv[4]);
    settings.maxSPP = std::atoi(argv[5]);
    settings.minSPP = std::atoi(argv[6]);
    settings.varianceThreshold = (float) std::atof(argv[7]);
    settings.seed = std::strtoull(argv[8], nullptr, 10);
    settings.maxDepth = 8;
    settings.exposure = 1.0f;
    settings.toneMapper = tonemap::fromString("pbrneutral");
    settings.outputBase = out + "/output";
    settings.sampleCountBase = out + "/sample_count";

    Scene scene = buildScene();
    scene.camera.init(settings.width, settings.height);
    Integrator integrator(scene, settings.maxDepth, settings.russianRoulette);
    scene.buildBVH();
    Renderer r;
    r.Render(scene, integrator, settings);

    FILE *f = fopen((out + "/linear.f32").c_str(), "wb");
    if (!f) return 1;
    fwrite(r.lastImage.data(), sizeof(Vector3f), r.lastImage.size(), f);
    fclose(f);

    // single queries: rays from the camera through a few points
    FILE *q = fopen((out + "/queries.txt").c_str(), "w");
    if (!q) return 1;
    fprintf(q, "emit_area %.9g\n", scene.totalEmitArea);
    const Vector3f targets[4] = {Vector3f(-0.45f, 0.4f, -0.3f), Vector3f(0.45f, 0.4f, 0.25f),
                                 Vector3f(0.f, 2.f, 0.f), Vector3f(0.f, 1.f, -1.f)};
    for (const Vector3f &tg : targets) {
        Ray ray(scene.camera.eye, normalize(tg - scene.camera.eye));
        Intersection it = scene.intersect(ray);
        const bool occl = scene.intersectP(ray, 100.f);
        const bool occlShort = scene.intersectP(ray, 0.5f);
        Vector3f L = integrator.castRay(ray);
        fprintf(q, "hit %d t %.9g n %.9g %.9g %.9g occluded %d %d L %.9g %.9g %.9g\n", it.happened ? 1 : 0,
                it.tnear, it.normal.x, it.normal.y, it.normal.z, occl ? 1 : 0, occlShort ? 1 : 0, L.x, L.y, L.z);
    }
    fclose(q);
    return 0;
}

// The per-query API of the reference on the box room: BVHAccel (the
// scene's own and one built from the object list), Scene::sampleLight /
// pdfLight, the photon map (Scene::buildPhotonMaps, caustic_map,
// estimateIrradiance) and the media functions (Medium::intersect, phaseHG,
// samplePhaseHG). Writes one line per query to <outdir>/api.txt.
static int api(int argc, char **argv) {
    if (argc < 3) return 2;
    const std::string out = argv[2];
    Scene scene = buildScene();
    scene.camera.init(64, 48);
    scene.buildBVH();
    FILE *q = fopen((out + "/api.txt").c_str(), "w");
    if (!q) return 1;
    fprintf(q, "emissives %d pdf %.9g\n", (int) scene.emissives.size(), scene.pdfLight(Intersection()));
    BVHAccel own(scene.objects, 1, BVHAccel::SplitMethod::SAH);
    const Vector3f targets[4] = {Vector3f(-0.45f, 0.4f, -0.3f), Vector3f(0.45f, 0.4f, 0.25f),
                                 Vector3f(0.f, 2.f, 0.f), Vector3f(0.f, 1.f, -1.f)};
    for (const Vector3f &tg : targets) {
        Ray ray(scene.camera.eye, normalize(tg - scene.camera.eye));
        Intersection a = scene.intersect(ray), b = scene.bvh->Intersect(ray), c = own.Intersect(ray);
        fprintf(q, "bvh %.9g %.9g %.9g %d %d %d\n", a.tnear, b.tnear, c.tnear, scene.bvh->IntersectP(ray, 100.f) ? 1 : 0,
                own.IntersectP(ray, 100.f) ? 1 : 0, own.IntersectP(ray, 0.5f) ? 1 : 0);
    }
    for (int k = 0; k < 64; k++) {
        Intersection it;
        float pdf = 0.f;
        scene.sampleLight(it, pdf);
        fprintf(q, "light %.9g %.9g %.9g %.9g %.9g %.9g %.9g %.9g %d\n", it.coords.x, it.coords.y, it.coords.z,
                it.normal.x, it.normal.y, it.normal.z, pdf, it.material ? it.material->m_emission.x : -1.f,
                it.obj == scene.emissives[0] ? 1 : 0);
    }
    scene.photon_map.photon_radius = 0.05f;
    scene.buildPhotonMaps(200000);
    const auto &cm = scene.photon_map.caustic_map;
    fprintf(q, "photons %d\n", (int) cm.size());
    for (size_t k = 0; k < cm.size() && k < 40; k += 4) {
        const Vector3f p = cm[k].position + Vector3f(0.01f, 0.f, 0.01f);
        const Vector3f e = scene.photon_map.estimateIrradiance(p, Vector3f(0.f, 1.f, 0.f));
        fprintf(q, "irr %.9g %.9g %.9g %.9g %.9g %.9g\n", p.x, p.y, p.z, e.x, e.y, e.z);
    }
    // Medium (Volume.hpp): the dust preset bounded by a box, the fog preset
    // unbounded, and the phase functions
    Medium dust = Medium::dust();
    dust.bounded = true;
    dust.bounds = Bounds3(Vector3f(-1.f, 0.f, -1.f), Vector3f(1.f, 2.f, 1.f));
    const Medium fog = Medium::fog();
    for (const Vector3f &tg : targets) {
        Ray ray(scene.camera.eye, normalize(tg - scene.camera.eye));
        float a = -7.f, b = -7.f, c = -7.f, d = -7.f;
        const int h0 = dust.intersect(ray, a, b) ? 1 : 0, h1 = fog.intersect(ray, c, d) ? 1 : 0;
        fprintf(q, "medium %d %.9g %.9g %d %.9g %.9g\n", h0, a, b, h1, c, d);
    }
    fprintf(q, "phase %.9g %.9g %.9g %d %d\n", phaseHG(0.3f, dust.g), phaseHG(-1.f, 0.f), phaseHG(1.f, 0.6f),
            dust.active() ? 1 : 0, Medium().active() ? 1 : 0);
    for (int k = 0; k < 32; k++) {
        const Vector3f w = samplePhaseHG(Vector3f(0.f, 0.f, 1.f), dust.g);
        fprintf(q, "hg %.9g %.9g %.9g\n", w.x, w.y, w.z);
    }
    fclose(q);
    FILE *f = fopen((out + "/photons.f32").c_str(), "wb");
    if (!f) return 1;
    for (const Photon &ph : cm) {
        const float row[9] = {ph.position.x, ph.position.y, ph.position.z, ph.direction.x, ph.direction.y,
                              ph.direction.z, ph.power.x, ph.power.y, ph.power.z};
        fwrite(row, sizeof(float), 9, f);
    }
    fclose(f);
    return 0;
}

static int hdr(int argc, char **argv) {
    if (argc < 4) return 2;
    EnvMap env;
    if (!env.load(argv[2])) return 1;
    FILE *f = fopen(argv[3], "wb");
    if (!f) return 1;
    const int wh[2] = {env.width, env.height};
    fwrite(wh, sizeof(int), 2, f);
    fwrite(env.pixels.data(), sizeof(Vector3f), env.pixels.size(), f);
    fclose(f);
    return 0;
}

static int obj(int argc, char **argv) {
    if (argc < 6) return 2;
    DiffuseMaterial m;
    MeshTriangle mesh(argv[2], Vector3f((float) std::atof(argv[3]), (float) std::atof(argv[4]),
                                        (float) std::atof(argv[5])),
                      &m);
    Bounds3 b = mesh.getBounds();
    printf("tris %zu area %.9g bounds %.9g %.9g %.9g %.9g %.9g %.9g\n", mesh.triangles.size(), mesh.getArea(),
           b.pMin.x, b.pMin.y, b.pMin.z, b.pMax.x, b.pMax.y, b.pMax.z);
    return 0;
}

// GLTFLoader::load (scenes/museum_final.cpp's call): per mesh its triangle
// count, material class and factors, texture sizes, and a checksum of every
// loaded float and texel, for comparison with the reference loader's output.
static int gltf(int argc, char **argv) {
    if (argc < 4) return 2;
    const MaterialType mt[4] = {MaterialType::Diffuse, MaterialType::Glass, MaterialType::Mirror, MaterialType::Emit};
    std::vector<MeshTriangle *> meshes = GLTFLoader::load(argv[2], mt[std::atoi(argv[3])]);
    for (MeshTriangle *m : meshes) {
        double sum = 0.0;
        for (const Triangle &t : m->triangles) {
            for (const Vector3f &v : {t.v0, t.v1, t.v2, t.normal, t.n0, t.n1, t.n2, t.tan0, t.tan1, t.tan2})
                sum += (double) v.x + 2.0 * v.y + 3.0 * v.z;
            for (const Vector2f &v : {t.t0, t.t1, t.t2}) sum += (double) v.x - 5.0 * v.y;
            sum += t.tangentW0 + t.tangentW1 + t.tangentW2 + (t.hasSmoothNormals ? 7 : 0) + (t.hasTangents ? 11 : 0);
        }
        Material *M = m->m;
        printf("mesh %zu kind %d mat %.9g %.9g %.9g %.9g %.9g %.9g %.9g %.9g %.9g sum %.17g", m->triangles.size(),
               M->kind(), M->baseColor.x, M->baseColor.y, M->baseColor.z, M->roughness, M->metallic, M->ior,
               M->m_emission.x, M->m_emission.y, M->m_emission.z, sum);
        for (const Texture *t : {&M->baseColorTex, &M->metallicRoughnessTex, &M->normalTex, &M->emissiveTex}) {
            uint64_t h = 0;
            for (unsigned char c : t->data) h = h * 131 + c;
            printf(" tex %d %d %d %llu", t->width, t->height, t->channels, (unsigned long long) h);
        }
        printf("\n");
        delete m;
    }
    return 0;
}

int main(int argc, char **argv) {
    if (argc < 2) {
        fprintf(stderr, "usage: %s render|hdr|obj ...\n", argv[0]);
        return 2;
    }
    try {
        if (!std::strcmp(argv[1], "render")) return render(argc, argv);
        if (!std::strcmp(argv[1], "api")) return api(argc, argv);
        if (!std::strcmp(argv[1], "hdr")) return hdr(argc, argv);
        if (!std::strcmp(argv[1], "obj")) return obj(argc, argv);
        if (!std::strcmp(argv[1], "gltf")) return gltf(argc, argv);
    } catch (const std::exception &e) {
        fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 2;
}
