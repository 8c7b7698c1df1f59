// prr_oracle.cpp — TEST INFRASTRUCTURE ONLY (see prr_oracle.h).
//
// C entry points over the CPU restatement in oracle_core.hpp /
// oracle_scene.hpp. Two builds are made from this file (oracle/Makefile):
//   liboracle.so       IEEE-strict (-O2 -ffp-contract=off): the parity checker
//   liboracle_fast.so  the reference's release flags (-O3 -march=native
//                      -ffast-math, OpenMP; CMakeLists.txt:15,41-52): the CPU
//                      baseline timed beside the GPU.
#include "prr_oracle.h"

#include "oracle_cache.hpp"
#include "oracle_scene.hpp"

#include <atomic>
#include <cstdlib>
#include <unordered_map>

#ifdef _OPENMP
#include <omp.h>
#endif

using namespace orc;

struct orc_scene {
    Scene sc;
};

static V3 v3(const float *p) { return V3(p[0], p[1], p[2]); }
static void put3(float *p, const V3 &v) {
    p[0] = v.x;
    p[1] = v.y;
    p[2] = v.z;
}

extern "C" int orc_version(void) { return 3; }

// pcg32_srandom_r(initstate, initseq) followed by n pcg32_random_r outputs
// (O'Neill 2014) and the float each maps to (top 24 bits * 2^-24), through
// the Pcg32 the per-pixel streams use.
extern "C" void orc_pcg32_sequence(uint64_t initstate, uint64_t initseq, int n, uint32_t *out, float *fout) {
    Pcg32 r;
    r.state = 0;
    r.inc = (initseq << 1) | 1u;
    r.next_u32();
    r.state += initstate;
    r.next_u32();
    Pcg32 f = r;
    for (int i = 0; i < n; i++) {
        out[i] = r.next_u32();
        fout[i] = f.next();
    }
}

// The first n floats of the per-pixel stream Pcg32::make(seed, pixel, sample).
extern "C" void orc_pcg32_stream(uint64_t seed, uint32_t pixel, uint32_t sample, int n, float *fout) {
    Pcg32 r = Pcg32::make(seed, pixel, sample);
    for (int i = 0; i < n; i++) fout[i] = r.next();
}

extern "C" orc_scene *orc_scene_create(const prr_scene_desc *d) {
    orc_scene *s = new orc_scene();
    Scene &sc = s->sc;
    sc.tris.resize(d->num_triangles);
    for (int i = 0; i < d->num_triangles; i++) {
        const float *t = d->triangles + 9 * (size_t) i;
        Tri &tr = sc.tris[i] = Tri(v3(t), v3(t + 3), v3(t + 6));
        const int fl = d->tri_flags ? d->tri_flags[i] : 0;
        if (fl & PRR_TRI_SMOOTH_NORMALS) {  // GLTFLoader.cpp:244-251
            const float *n = d->tri_normals + 9 * (size_t) i;
            tr.n0 = v3(n), tr.n1 = v3(n + 3), tr.n2 = v3(n + 6);
            tr.smooth = true;
            tr.n = normalize(tr.n0 + tr.n1 + tr.n2);
        }
        if (fl & PRR_TRI_UVS) {
            const float *q = d->tri_uvs + 6 * (size_t) i;
            tr.t0[0] = q[0], tr.t0[1] = q[1], tr.t1[0] = q[2], tr.t1[1] = q[3], tr.t2[0] = q[4], tr.t2[1] = q[5];
        }
        if (fl & PRR_TRI_TANGENTS) {
            const float *q = d->tri_tangents + 12 * (size_t) i;
            tr.tan0 = v3(q), tr.tw0 = q[3], tr.tan1 = v3(q + 4), tr.tw1 = q[7], tr.tan2 = v3(q + 8), tr.tw2 = q[11];
            tr.hasTan = true;
        }
    }
    for (int i = 0; i < d->num_textures; i++) {
        const prr_texture_desc &t = d->textures[i];
        Tex x;
        x.w = t.width, x.h = t.height, x.c = t.channels;
        x.d.assign(t.data, t.data + (size_t) t.width * t.height * t.channels);
        sc.texs.push_back(std::move(x));
    }
    if (d->material_textures)
        for (int i = 0; i < d->num_materials; i++)
            sc.matTex.push_back({d->material_textures[4 * i], d->material_textures[4 * i + 1],
                                 d->material_textures[4 * i + 2], d->material_textures[4 * i + 3]});
    sc.sphs.resize(d->num_spheres);
    for (int i = 0; i < d->num_spheres; i++) {
        const float *p = d->spheres + 4 * (size_t) i;
        Sph &sp = sc.sphs[i];
        sp.c = v3(p);
        sp.r = p[3];
        sp.r2 = sp.r * sp.r;
        sp.area = 4 * kPi * sp.r * sp.r;  // Sphere.hpp:16-17
    }
    sc.mats.resize(d->num_materials);
    for (int i = 0; i < d->num_materials; i++) {
        const prr_material_desc &m = d->materials[i];
        Material &o = sc.mats[i];
        o.type = m.type;
        o.base = v3(m.base_color);
        o.roughness = m.roughness;
        o.metallic = m.metallic;
        o.ior = m.ior;
        o.emission = v3(m.emission);
    }
    sc.objs.resize(d->num_objects);
    for (int k = 0; k < d->num_objects; k++) {
        sc.objs[k].kind = d->objects[k].kind;
        sc.objs[k].mat = d->objects[k].material;
        sc.objs[k].first = d->objects[k].first;
        sc.objs[k].count = d->objects[k].count;
    }
    sc.finalizeBuild();
    if (d->env.width > 0 && d->env.height > 0 && d->env.pixels)
        sc.env.build(d->env.width, d->env.height, d->env.pixels);
    sc.background = v3(d->background);
    for (int i = 0; i < d->num_point_lights; i++)
        sc.points.push_back(PointLight{v3(d->point_lights[i].position), v3(d->point_lights[i].intensity)});
    Camera &c = sc.camera;
    c.eye = v3(d->camera.eye);
    c.target = v3(d->camera.target);
    c.up = v3(d->camera.up);
    c.fovh = d->camera.fovh;
    c.aperture = d->camera.aperture;
    c.focusDistance = d->camera.focus_distance;
    c.init();
    return s;
}

extern "C" void orc_scene_free(orc_scene *s) { delete s; }

extern "C" float orc_total_emit_area(const orc_scene *s) { return s->sc.totalEmitArea; }

extern "C" int orc_export_level(const orc_scene *s, int level, int cap, float *b6, uint32_t *fc,
                                uint32_t *meta) {
    const BVH *bvh = nullptr;
    if (level < 0)
        bvh = &s->sc.top;
    else if (level < (int) s->sc.objs.size() && s->sc.objs[level].kind == OBJ_MESH)
        bvh = &s->sc.objs[level].sub;
    if (!bvh) return -1;
    int n = (int) bvh->nodes.size();
    for (int i = 0; i < n && i < cap; i++) {
        for (int k = 0; k < 6; k++) b6[6 * i + k] = bvh->nodes[i].b[k];
        fc[i] = bvh->nodes[i].first;
        meta[i] = bvh->nodes[i].meta;
    }
    return n;
}

extern "C" int orc_export_order(const orc_scene *s, int level, int cap, int32_t *order, float *areas) {
    const BVH *bvh = nullptr;
    int base = 0;
    if (level < 0)
        bvh = &s->sc.top;
    else if (level < (int) s->sc.objs.size() && s->sc.objs[level].kind == OBJ_MESH) {
        bvh = &s->sc.objs[level].sub;
        base = s->sc.objs[level].first;
    }
    if (!bvh) return -1;
    for (int i = 0; i < (int) bvh->ordered.size() && i < cap; i++) order[i] = bvh->ordered[i] - base;
    for (int i = 0; i < (int) bvh->nodeAreas.size() && i < cap; i++) areas[i] = bvh->nodeAreas[i];
    return (int) bvh->ordered.size();
}

// Scene::sampleLight with the reference's SplitMix64 stream from state seeds[k]
extern "C" void orc_sample_light(const orc_scene *s, int64_t n, const uint64_t *seeds, float *pos,
                                 float *nrm, float *pdf) {
    for (int64_t k = 0; k < n; k++) {
        SplitMix64 rng{seeds[k]};
        V3 p, q;
        int m;
        float pd = 0.f;
        if (!s->sc.sampleLight(rng, p, q, m, pd)) p = q = V3(0.f);
        put3(pos + 3 * k, p);
        put3(nrm + 3 * k, q);
        pdf[k] = pd;
    }
}

extern "C" void orc_trace_closest(const orc_scene *s, int64_t n, const float *org, const float *dir,
                                  int32_t *prim, float *t, float *u, float *v, uint32_t *counters) {
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < n; i++) {
        Counters c;
        Ray r{v3(org + 3 * i), v3(dir + 3 * i)};
        Hit h = s->sc.intersect(r, &c);
        prim[i] = h.happened ? h.prim : -1;
        if (t) t[i] = h.happened ? h.t : kInf;
        if (u) u[i] = h.u;
        if (v) v[i] = h.v;
        if (counters) {
            counters[3 * i + 0] = (uint32_t) c.inner;
            counters[3 * i + 1] = (uint32_t) c.tris;
            counters[3 * i + 2] = (uint32_t) c.spheres;
        }
    }
}

extern "C" void orc_trace_any(const orc_scene *s, int64_t n, const float *org, const float *dir,
                              const float *tmax, uint8_t *hit, uint32_t *counters) {
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < n; i++) {
        Counters c;
        Ray r{v3(org + 3 * i), v3(dir + 3 * i)};
        hit[i] = s->sc.intersectP(r, tmax[i], &c) ? 1 : 0;
        if (counters) {
            counters[3 * i + 0] = (uint32_t) c.innerAny;
            counters[3 * i + 1] = (uint32_t) c.trisAny;
            counters[3 * i + 2] = (uint32_t) c.spheres;
        }
    }
}

// Brute force: minimum t over all primitives (ties: lowest id), independent
// of any tree. Uses the same primitive tests as the traversal.
extern "C" void orc_brute_closest(const orc_scene *s, int64_t n, const float *org, const float *dir,
                                  int32_t *prim, float *t) {
    const Scene &sc = s->sc;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < n; i++) {
        Ray r{v3(org + 3 * i), v3(dir + 3 * i)};
        float best = kInf;
        int bp = -1;
        for (int k = 0; k < (int) sc.tris.size(); k++) {
            float tt, uu, vv;
            if (triHit(sc.tris[k], r, tt, uu, vv) && tt < best) {
                best = tt;
                bp = k;
            }
        }
        for (int k = 0; k < (int) sc.sphs.size(); k++) {
            float tt;
            if (sphHit(sc.sphs[k], r, tt) && tt < best) {
                best = tt;
                bp = sc.numTris() + k;
            }
        }
        prim[i] = bp;
        t[i] = best;
    }
}

extern "C" void orc_brute_any(const orc_scene *s, int64_t n, const float *org, const float *dir,
                              const float *tmax, uint8_t *hit) {
    const Scene &sc = s->sc;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < n; i++) {
        Ray r{v3(org + 3 * i), v3(dir + 3 * i)};
        bool occ = false;
        for (int k = 0; k < (int) sc.tris.size() && !occ; k++) {
            float tt = triAnyT(sc.tris[k], r, tmax[i]);
            occ = tt >= 0.f && tt < tmax[i];
        }
        for (int k = 0; k < (int) sc.sphs.size() && !occ; k++) {
            float tt;
            occ = sphHit(sc.sphs[k], r, tt) && tt < tmax[i];
        }
        hit[i] = occ ? 1 : 0;
    }
}

static Material paramMat(const float *p) {
    Material m;
    m.type = (int) p[0];
    m.base = V3(p[1], p[2], p[3]);
    m.roughness = p[4];
    m.metallic = p[5];
    m.ior = p[6];
    return m;
}

extern "C" void orc_bsdf_eval(int64_t n, const float *wi, const float *wo, const float *nrm,
                              const float *params, float *f, float *pdf) {
    for (int64_t i = 0; i < n; i++) {
        Material m = paramMat(params + 8 * i);
        V3 fv(0.f);
        float p = 0.f;
        if (m.type == MAT_DIFFUSE) {
            Shading sd = makeShading(m, v3(nrm + 3 * i));
            fv = diffuseEval(v3(wi + 3 * i), v3(wo + 3 * i), sd);
            p = diffusePdf(v3(wi + 3 * i), v3(wo + 3 * i), sd);
        }
        put3(f + 3 * i, fv);
        pdf[i] = p;
    }
}

extern "C" void orc_bsdf_sample(int64_t n, const float *wo, const float *nrm, const float *params,
                                const float *u3, float *wi, float *f, float *pdf, int32_t *lobe) {
    for (int64_t i = 0; i < n; i++) {
        Material m = paramMat(params + 8 * i);
        FixedU rng{u3 + 3 * i};
        BSDFSample bs;
        V3 N = v3(nrm + 3 * i), w = v3(wo + 3 * i);
        if (m.type == MAT_DIFFUSE)
            bs = diffuseSample(w, makeShading(m, N), rng);
        else if (m.type == MAT_GLASS)
            bs = glassSample(w, N, m.ior, rng);
        else if (m.type == MAT_MIRROR)
            bs = mirrorSample(w, N, m.base);
        put3(wi + 3 * i, bs.wi);
        put3(f + 3 * i, bs.f);
        pdf[i] = bs.pdf;
        lobe[i] = bs.lobe;
    }
}

extern "C" void orc_medium(int64_t n, const float *org, const float *dir, const float *bounds,
                           const int32_t *bounded, float *t_enter, float *t_exit, uint8_t *hit) {
    for (int64_t i = 0; i < n; i++) {
        float te = t_enter[i], tx = t_exit[i];
        hit[i] = mediumIntersect(bounded[i] != 0, bounds + 6 * i, bounds + 6 * i + 3, v3(org + 3 * i),
                                 v3(dir + 3 * i), te, tx)
                     ? 1
                     : 0;
        t_enter[i] = te;
        t_exit[i] = tx;
    }
}

extern "C" void orc_phase_hg(int64_t n, const float *cos_theta, const float *g, float *out) {
    for (int64_t i = 0; i < n; i++) out[i] = phaseHG(cos_theta[i], g[i]);
}

extern "C" void orc_phase_hg_sample(int64_t n, const float *wo, const float *g, const float *u2, float *wi) {
    for (int64_t i = 0; i < n; i++) put3(wi + 3 * i, samplePhaseHG(v3(wo + 3 * i), g[i], u2[2 * i], u2[2 * i + 1]));
}

extern "C" void orc_ggx_terms(int64_t n, const float *in4, float *out6) {
    for (int64_t i = 0; i < n; i++) {
        float NdotH = in4[4 * i], NdotWo = in4[4 * i + 1], NdotWi = in4[4 * i + 2], alpha = in4[4 * i + 3];
        out6[6 * i + 0] = D_GGX(NdotH, alpha);
        out6[6 * i + 1] = lambdaGGX(NdotWo, alpha);
        out6[6 * i + 2] = G1_GGX(NdotWo, alpha);
        out6[6 * i + 3] = G2_GGX(NdotWo, NdotWi, alpha);
        out6[6 * i + 4] = pdfVNDF(NdotWo, NdotH, NdotWo, alpha);
        out6[6 * i + 5] = F_Schlick(NdotWo, V3(0.04f)).x;
    }
}

extern "C" void orc_env_build(int w, int h, const float *rgb, float *cdf, float *marg, float *total) {
    Env e;
    e.build(w, h, rgb);
    std::memcpy(cdf, e.cdf.data(), sizeof(float) * (size_t) w * h);
    std::memcpy(marg, e.marg.data(), sizeof(float) * (size_t) h);
    *total = e.total;
}

extern "C" void orc_env_sample(int w, int h, const float *rgb, int64_t n, const float *u2, float *dir,
                               float *pdf) {
    Env e;
    e.build(w, h, rgb);
    for (int64_t i = 0; i < n; i++) {
        float p;
        V3 d = e.importanceSample(u2[2 * i], u2[2 * i + 1], p);
        put3(dir + 3 * i, d);
        pdf[i] = p;
    }
}

extern "C" void orc_env_eval(int w, int h, const float *rgb, int64_t n, const float *dir, float *rad,
                             float *pdf) {
    Env e;
    e.build(w, h, rgb);
    for (int64_t i = 0; i < n; i++) {
        V3 d = v3(dir + 3 * i);
        put3(rad + 3 * i, e.sample(d));
        pdf[i] = e.pdf(d);
    }
}

extern "C" void orc_camera_rays(const prr_camera_desc *cd, int width, int height, int64_t n,
                                const int32_t *pix, const float *offs, float *org, float *dir) {
    Camera c;
    c.eye = v3(cd->eye);
    c.target = v3(cd->target);
    c.up = v3(cd->up);
    c.fovh = cd->fovh;
    c.aperture = cd->aperture;
    c.focusDistance = cd->focus_distance;
    c.init();
    const int stride = c.aperture > 0.f ? 4 : 2;
    for (int64_t i = 0; i < n; i++) {
        FixedU lens{offs + stride * i + 2};
        Ray r = c.generateRay(pix[2 * i], pix[2 * i + 1], width, height, offs[stride * i],
                              offs[stride * i + 1], lens);
        put3(org + 3 * i, r.o);
        put3(dir + 3 * i, r.d);
    }
}

extern "C" void orc_tonemap(int op, int64_t n, float exposure, const float *lin, float *disp,
                            uint8_t *rgb8) {
    for (int64_t i = 0; i < n; i++) {
        V3 c = tonemap(op, exposeLinear(v3(lin + 3 * i), exposure));
        if (disp) put3(disp + 3 * i, c);
        if (rgb8) {
            rgb8[3 * i + 0] = quantize(c.x);
            rgb8[3 * i + 1] = quantize(c.y);
            rgb8[3 * i + 2] = quantize(c.z);
        }
    }
}

extern "C" void orc_castray_splitmix(const orc_scene *s, int width, int height, int64_t n,
                                     const int32_t *pix, const uint64_t *seeds, int max_depth,
                                     float indirect_clamp, float *L) {
    IntegratorParams ip;
    ip.maxDepth = max_depth;
    ip.indirectClamp = indirect_clamp;
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t k = 0; k < n; k++) {
        SplitMix64 rng{seeds[k]};
        float ox = rng.next();
        float oy = rng.next();
        Ray r = s->sc.camera.generateRay(pix[2 * k], pix[2 * k + 1], width, height, ox, oy, rng);
        V3 c = castRay(s->sc, r, rng, ip, nullptr);
        if (std::isnan(c.x) || std::isnan(c.y) || std::isnan(c.z)) c = V3(0.f);  // Renderer.cpp:80-83
        put3(L + 3 * k, c);
    }
}

// castRay on caller rays with the GPU's PCG32 stream (seed, k, sample 0)
extern "C" void orc_castray_pcg(const orc_scene *s, int64_t n, const float *org, const float *dir, uint64_t seed,
                                int max_depth, float indirect_clamp, float *L) {
    IntegratorParams ip;
    ip.maxDepth = max_depth;
    ip.indirectClamp = indirect_clamp;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t k = 0; k < n; k++) {
        Pcg32 rng = Pcg32::make(seed, (uint32_t) k, 0u);
        V3 c = castRay(s->sc, Ray{v3(org + 3 * k), v3(dir + 3 * k)}, rng, ip, nullptr);
        put3(L + 3 * k, c);
    }
}

// ---------------------------------------------------------------------------
// Caustic photon map (Photon.cpp:96-188 restated with the build's documented
// changes: point-light emission, glass handled with the outward normal as in
// castRay's delta branch (the reference flips it, Photon.cpp:105-119), a
// photon is stored at the first diffuse hit after >=1 delta bounce and then
// terminated; storage is a uniform hash grid of cell 2r).
// ---------------------------------------------------------------------------
struct Photon {
    V3 pos, dir, power;
};
struct orc_photon_map {
    std::vector<Photon> ph;
    std::vector<int64_t> owner;  // emitting photon of each stored photon (reference rules)
    std::vector<int32_t> depth;  // its trace depth there
    float r = 0.01f, cell = 0.02f;
    std::unordered_map<uint64_t, std::vector<int>> grid;
    static uint64_t key(int x, int y, int z) {
        return ((uint64_t) (uint32_t) x * 73856093ULL) ^ ((uint64_t) (uint32_t) y * 19349663ULL) ^
               ((uint64_t) (uint32_t) z * 83492791ULL) ^ ((uint64_t) (uint32_t) x << 40) ^
               ((uint64_t) (uint32_t) y << 20);
    }
    V3 irradiance(const V3 &x) const {
        V3 sum(0.f);
        int cx = (int) std::floor(x.x / cell), cy = (int) std::floor(x.y / cell),
            cz = (int) std::floor(x.z / cell);
        for (int dz = -1; dz <= 1; dz++)
            for (int dy = -1; dy <= 1; dy++)
                for (int dx = -1; dx <= 1; dx++) {
                    auto it = grid.find(key(cx + dx, cy + dy, cz + dz));
                    if (it == grid.end()) continue;
                    for (int i : it->second) {
                        const Photon &p = ph[i];
                        if (norm(p.pos - x) <= r) sum += p.power;
                    }
                }
        return sum / (kPi * r * r);
    }
    V3 radiance(const V3 &x, const V3 &wo, const Shading &sd) const {
        V3 sum(0.f);
        int cx = (int) std::floor(x.x / cell), cy = (int) std::floor(x.y / cell),
            cz = (int) std::floor(x.z / cell);
        for (int dz = -1; dz <= 1; dz++)
            for (int dy = -1; dy <= 1; dy++)
                for (int dx = -1; dx <= 1; dx++) {
                    auto it = grid.find(key(cx + dx, cy + dy, cz + dz));
                    if (it == grid.end()) continue;
                    for (int i : it->second) {
                        const Photon &p = ph[i];
                        if (norm(p.pos - x) > r) continue;
                        V3 wi = -p.dir;
                        float c = dot(wi, sd.N);
                        if (c <= 1e-4f) continue;
                        sum += diffuseEval(wi, wo, sd) * p.power / c;
                    }
                }
        return sum / (kPi * r * r);
    }
};

// hash-grid cells of the stored photons (cell = 2r)
static void gridInsert(orc_photon_map *pm) {
    for (int i = 0; i < (int) pm->ph.size(); i++) {
        const V3 &x = pm->ph[i].pos;
        int cx = (int) std::floor(x.x / pm->cell), cy = (int) std::floor(x.y / pm->cell),
            cz = (int) std::floor(x.z / pm->cell);
        pm->grid[orc_photon_map::key(cx, cy, cz)].push_back(i);
    }
}

template <class R>
static bool tracePhoton(const Scene &sc, Ray ray, V3 power, int maxDepth, R &rng, Photon &out) {
    bool caustic = false;
    for (int depth = 0; depth <= maxDepth; depth++) {
        Hit h = sc.intersect(ray, nullptr);
        if (!h.happened) return false;
        const Material &m = sc.mats[sc.objs[h.obj].mat];
        if (m.isEmissive()) return false;
        V3 N = h.normal;
        if (m.isDelta()) {
            const Shading sd = sc.shading(sc.objs[h.obj].mat, h, N);
            BSDFSample bs = (m.type == MAT_GLASS) ? glassSample(-ray.d, sd.N, m.ior, rng)
                                                   : mirrorSample(-ray.d, sd.N, m.base);
            power = power * bs.f / bs.pdf;
            bool trans = dot(bs.wi, N) < 0.f;
            if (trans) power = power * sd.base;
            ray = Ray{h.coords + (trans ? -N : N) * kEps, bs.wi};
            caustic = true;
            continue;
        }
        if (caustic && depth > 0) {
            out = Photon{h.coords, ray.d, power};
            return true;
        }
        return false;
    }
    return false;
}

extern "C" orc_photon_map *orc_photon_build(const orc_scene *s, int64_t num, float radius, int maxDepth,
                                            uint64_t seed, int threads) {
    const Scene &sc = s->sc;
    orc_photon_map *pm = new orc_photon_map();
    pm->r = radius;
    pm->cell = 2.f * radius;
    if (sc.points.empty() || num <= 0) return pm;
    std::vector<Photon> tmp(num);
    std::vector<uint8_t> ok(num, 0);
    const int nl = (int) sc.points.size();
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#endif
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t i = 0; i < num; i++) {
        Pcg32 rng = Pcg32::make(seed, (uint32_t) (i & 0xffffffffu), 0x7000000u + (uint32_t) (i >> 32));
        const PointLight &pl = sc.points[(int) (i % nl)];
        float u1 = rng.next(), u2 = rng.next();
        float z = 1.f - 2.f * u1;
        float rr = std::sqrt(std::max(0.f, 1.f - z * z));
        float phi = 2.f * kPi * u2;
        V3 d(rr * std::cos(phi), z, rr * std::sin(phi));
        // power per photon: 4*pi*I / (photons per light)
        float perLight = (float) ((num + nl - 1 - (i % nl)) / nl);
        V3 power = pl.intensity * (4.f * kPi / perLight);
        Photon p;
        if (tracePhoton(sc, Ray{pl.pos, d}, power, maxDepth, rng, p)) {
            tmp[i] = p;
            ok[i] = 1;
        }
    }
    for (int64_t i = 0; i < num; i++)
        if (ok[i]) pm->ph.push_back(tmp[i]);
    gridInsert(pm);
    return pm;
}

// ---------------------------------------------------------------------------
// The reference's own photon map: PhotonMap::emit's per-photon body
// (Photon.cpp:166-180) and PhotonMap::trace (Photon.cpp:96-152) as written,
// including what the build's point-light map changes: emission from
// Scene::emissives through Scene::sampleLight and a cosine hemisphere about
// the light normal with power Le * A / N (A = 1 / pdf); delta vertices shaded
// with the normal flipped to face the photon (so glass is always "entered",
// eta = 1 / ior, and transmission is not tinted); every diffuse hit after a
// delta vertex stores the photon (depth > 0), then a Russian-roulette
// diffuse continuation on max(power) with the weight f * cos / pdf (f already
// holds the cosine). The trace recursion is a loop over depth.
// ---------------------------------------------------------------------------
template <class R>
static void emitReferencePhoton(const Scene &sc, float numPhotons, int maxDepth, R &rng, std::vector<Photon> &out,
                                std::vector<int32_t> *depths) {
    V3 lp, ln;
    int lm;
    float pdf;
    if (!sc.sampleLight(rng, lp, ln, lm, pdf)) return;
    const float area = 1.0f / pdf;
    // sampleCosineHemisphere (BRDFUtils.hpp:66-73)
    const float r1 = rng.next();
    const float r2 = rng.next();
    const float r = std::sqrt(r1);
    const float phi = 2.f * kPi * r2;
    V3 dir = toWorld(V3(r * std::cos(phi), r * std::sin(phi), std::sqrt(std::max(0.f, 1.f - r1))), ln);
    V3 power = sc.mats[lm].emission * area / numPhotons;
    V3 pos = lp + ln * kEps;
    bool caustic = false;
    for (int depth = 0; depth <= maxDepth; depth++) {
        const Hit h = sc.intersect(Ray{pos, dir}, nullptr);
        if (!h.happened) return;
        const int mi = sc.objs[h.obj].mat;
        const Material &m = sc.mats[mi];
        if (m.isEmissive()) return;
        const V3 geoN = dot(dir, h.normal) < 0.0f ? h.normal : -h.normal;
        if (m.isDelta()) {
            const BSDFSample bs = (m.type == MAT_GLASS) ? glassSample(-dir, geoN, m.ior, rng)
                                                        : mirrorSample(-dir, geoN, m.base);
            caustic = true;
            power = power * bs.f;
            dir = bs.wi;
            pos = h.coords + (dot(bs.wi, geoN) > 0.0f ? geoN : -geoN) * kEps;
            continue;
        }
        if (depth > 0 && caustic) {
            out.push_back(Photon{h.coords, dir, power});
            if (depths) depths->push_back(depth);
        }
        float survival = std::min(1.0f, std::max({power.x, power.y, power.z}));
        if (!(rng.next() < survival && depth < maxDepth)) return;
        const Shading sd = sc.shading(mi, h, geoN);
        const BSDFSample bs = diffuseSample(-dir, sd, rng);
        if (bs.pdf < 1e-6f) return;
        const float cosTheta = std::max(0.0f, dot(bs.wi, geoN));
        const V3 weight = bs.f * cosTheta / bs.pdf;
        power = power * weight;
        survival = std::min(1.0f, std::max({power.x, power.y, power.z}));
        if (rng.next() > survival) return;
        power = power / survival;
        pos = h.coords + geoN * kEps;
        dir = bs.wi;
    }
}

// Scene::buildPhotonMaps (Scene.cpp:8-13) with the reference's rules and the
// GPU's PCG32 stream per photon (the point-light map's streams); stored
// photons in (photon, depth) order.
extern "C" orc_photon_map *orc_photon_emit(const orc_scene *s, int64_t num, float radius, int maxDepth, uint64_t seed,
                                           int threads) {
    const Scene &sc = s->sc;
    orc_photon_map *pm = new orc_photon_map();
    pm->r = radius;
    pm->cell = 2.f * radius;
    if (sc.totalEmitArea <= 0.f || num <= 0) return pm;
    std::vector<std::vector<Photon>> per(num);
    std::vector<std::vector<int32_t>> dep(num);
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#endif
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t i = 0; i < num; i++) {
        Pcg32 rng = Pcg32::make(seed, (uint32_t) (i & 0xffffffffu), 0x7000000u + (uint32_t) (i >> 32));
        emitReferencePhoton(sc, (float) num, maxDepth, rng, per[i], &dep[i]);
    }
    for (int64_t i = 0; i < num; i++)
        for (size_t k = 0; k < per[i].size(); k++) {
            pm->ph.push_back(per[i][k]);
            pm->owner.push_back(i);
            pm->depth.push_back(dep[i][k]);
        }
    gridInsert(pm);
    return pm;
}

// The same with the reference's own SplitMix64 stream from states[k] per
// photon and a given N (power Le * A / N): replays the golden harness.
extern "C" int64_t orc_photon_emit_splitmix(const orc_scene *s, int64_t n, const uint64_t *states, float numPhotons,
                                            int maxDepth, int64_t cap, float *out9, int32_t *owner) {
    std::vector<Photon> ph;
    std::vector<int32_t> own;
    for (int64_t k = 0; k < n; k++) {
        SplitMix64 rng{states[k]};
        emitReferencePhoton(s->sc, numPhotons, maxDepth, rng, ph, nullptr);
        own.resize(ph.size(), (int32_t) k);
    }
    for (int64_t i = 0; i < (int64_t) ph.size() && i < cap; i++) {
        put3(out9 + 9 * i, ph[i].pos);
        put3(out9 + 9 * i + 3, ph[i].dir);
        put3(out9 + 9 * i + 6, ph[i].power);
        owner[i] = own[i];
    }
    return (int64_t) ph.size();
}

// A photon map over given photons (pos, dir, power per row): the gather of
// a map the caller supplies, e.g. the reference's own stored photons.
extern "C" orc_photon_map *orc_photon_map_from(int64_t n, const float *ph9, float radius) {
    orc_photon_map *pm = new orc_photon_map();
    pm->r = radius;
    pm->cell = 2.f * radius;
    for (int64_t i = 0; i < n; i++)
        pm->ph.push_back(Photon{v3(ph9 + 9 * i), v3(ph9 + 9 * i + 3), v3(ph9 + 9 * i + 6)});
    gridInsert(pm);
    return pm;
}

// Emitting photon and trace depth of the stored photons (reference rules).
extern "C" int64_t orc_photon_owners(const orc_photon_map *pm, int64_t cap, int64_t *owner, int32_t *depth) {
    const int64_t n = std::min<int64_t>(cap, (int64_t) pm->owner.size());
    for (int64_t i = 0; i < n; i++) {
        owner[i] = pm->owner[i];
        depth[i] = pm->depth[i];
    }
    return (int64_t) pm->owner.size();
}

extern "C" void orc_photon_free(orc_photon_map *pm) { delete pm; }
extern "C" int64_t orc_photon_count(const orc_photon_map *pm) { return (int64_t) pm->ph.size(); }
extern "C" int64_t orc_photon_export(const orc_photon_map *pm, int64_t cap, float *o) {
    int64_t n = std::min<int64_t>(cap, (int64_t) pm->ph.size());
    for (int64_t i = 0; i < n; i++) {
        put3(o + 9 * i, pm->ph[i].pos);
        put3(o + 9 * i + 3, pm->ph[i].dir);
        put3(o + 9 * i + 6, pm->ph[i].power);
    }
    return (int64_t) pm->ph.size();
}
extern "C" void orc_photon_radiance(const orc_photon_map *pm, int64_t n, const float *pos, const float *wo,
                                    const float *nrm, const float *params, float *o) {
    for (int64_t i = 0; i < n; i++) {
        Material m = paramMat(params + 8 * i);
        Shading sd = makeShading(m, v3(nrm + 3 * i));
        put3(o + 3 * i, pm->radiance(v3(pos + 3 * i), v3(wo + 3 * i), sd));
    }
}

extern "C" void orc_photon_irradiance(const orc_photon_map *pm, int64_t n, const float *pos, float *o) {
    for (int64_t i = 0; i < n; i++) put3(o + 3 * i, pm->irradiance(v3(pos + 3 * i)));
}

// ---------------------------------------------------------------------------
// Renderer::Render sample loop (Renderer.cpp:63-99), PCG32 per (pixel,
// sample), accumulated in sample order, with the adaptive stop rule.
// ---------------------------------------------------------------------------
// Welford (Renderer.hpp:14-40): running mean / M2 of one channel in float.
struct WelfordF {
    int n = 0;
    float mean = 0.f, m2 = 0.f;
    void add(float x) {
        n++;
        float d = x - mean;
        mean += d / (float) n;
        m2 += d * (x - mean);
    }
    // 95% CI half-width relative to |mean|; 1 until two samples / nonzero mean
    float ci() const {
        if (n < 2 || std::fabs(mean) < 1e-6f) return 1.f;
        return 1.96f * std::sqrt(m2 / (float) (n - 1) / (float) n) / std::fabs(mean);
    }
};
struct PixelStop {
    WelfordF w[3];
    int minSpp;
    float thr;
    // Renderer.cpp:88-96: update, then stop once n >= minSPP and all CIs < thr
    bool add(const V3 &c) {
        w[0].add(c.x);
        w[1].add(c.y);
        w[2].add(c.z);
        return w[0].n >= minSpp && w[0].ci() < thr && w[1].ci() < thr && w[2].ci() < thr;
    }
};

extern "C" void orc_welford(int nseq, int len, const float *seq, const int32_t *min_spp, const float *thr,
                            int32_t *count, float *mean) {
    for (int q = 0; q < nseq; q++) {
        PixelStop ps;
        ps.minSpp = min_spp[q];
        ps.thr = thr[q];
        V3 sum(0.f);
        int n = 0;
        while (n < len) {
            V3 c = v3(seq + 3 * ((size_t) q * len + n));
            sum += c;
            n++;
            if (ps.add(c)) break;
        }
        count[q] = n;
        put3(mean + 3 * q, sum / (float) n);
    }
}

// Radiance-cache training-path selection: lowbias32 mix of pixel and sample.
static uint32_t trainHash(uint32_t p, uint32_t s) {
    uint32_t h = p * 0x9E3779B1u + s * 0x85EBCA77u;
    h ^= h >> 16;
    h *= 0x7FEB352Du;
    h ^= h >> 15;
    h *= 0x846CA68Bu;
    h ^= h >> 16;
    return h;
}

static void renderImpl(const orc_scene *s, const prr_render_params *p, int row0, int row1, const orc_photon_map *pm,
                       int threads, float *mean, float *var, int32_t *spp_out, uint64_t *counters, int truncateAt,
                       uint32_t trainMask, int truncateTrainAt);

extern "C" void orc_render(const orc_scene *s, const prr_render_params *p, int row0, int row1,
                           const orc_photon_map *pm, int threads, float *mean, float *var, int32_t *spp_out,
                           uint64_t *counters) {
    renderImpl(s, p, row0, row1, pm, threads, mean, var, spp_out, counters, 0, 0u, 0);
}

// The radiance-cache render with a network that outputs 0: paths with
// trainHash & train_mask != 0 end at their first scattering vertex at bounce
// >= query_bounce; the others (training paths) at bounce >= train_stop, or
// run to full depth when train_stop is 0.
extern "C" void orc_render_truncated(const orc_scene *s, const prr_render_params *p, int row0, int row1,
                                     int threads, int query_bounce, uint32_t train_mask, int train_stop,
                                     float *mean, uint64_t *counters) {
    renderImpl(s, p, row0, row1, nullptr, threads, mean, nullptr, nullptr, counters, query_bounce, train_mask,
               train_stop);
}

static void renderImpl(const orc_scene *s, const prr_render_params *p, int row0, int row1, const orc_photon_map *pm,
                       int threads, float *mean, float *var, int32_t *spp_out, uint64_t *counters, int truncateAt,
                       uint32_t trainMask, int truncateTrainAt) {
    const Scene &sc = s->sc;
    const int W = p->width, H = p->height;
    if (row1 <= 0 || row1 > H) row1 = H;
    IntegratorParams ip;
    ip.maxDepth = p->max_depth;
    ip.indirectClamp = p->indirect_clamp;
    PhotonFn pf;
    const PhotonFn *pfp = nullptr;
    if (pm && !pm->ph.empty()) {
        pf = [pm](const V3 &x, const V3 &wo, const Shading &sd) { return pm->radiance(x, wo, sd); };
        pfp = &pf;
    }
    std::atomic<uint64_t> cClosest(0), cAny(0), cSamples(0), cInner(0), cTris(0), cInnerAny(0), cTrisAny(0);
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#endif
#pragma omp parallel
    {
        Counters cnt;
        uint64_t samples = 0;
#pragma omp for schedule(dynamic, 1)
        for (int j = row0; j < row1; j++) {
            for (int i = 0; i < W; i++) {
                const uint32_t pix = (uint32_t) (j * W + i);
                V3 sum(0.f);
                double m1[3] = {0, 0, 0}, m2[3] = {0, 0, 0};
                PixelStop ps;
                ps.minSpp = p->min_spp;
                ps.thr = p->variance_threshold;
                int taken = 0;
                for (int sIdx = 0; sIdx < p->spp; sIdx++) {
                    Pcg32 rng = Pcg32::make(p->seed, pix, (uint32_t) (p->sample_offset + sIdx));
                    float ox = rng.next();
                    float oy = rng.next();
                    Ray r = sc.camera.generateRay(i, j, W, H, ox, oy, rng);
                    IntegratorParams ipp = ip;
                    if (truncateAt > 0)
                        ipp.truncateAt = (trainHash(pix, (uint32_t) (p->sample_offset + sIdx)) & trainMask) != 0u
                                             ? truncateAt
                                             : truncateTrainAt;
                    V3 c = castRay(sc, r, rng, ipp, &cnt, pfp);
                    if (std::isnan(c.x) || std::isnan(c.y) || std::isnan(c.z)) c = V3(0.f);
                    sum += c;
                    samples++;
                    taken++;
                    if (var) {
                        double cv[3] = {c.x, c.y, c.z};
                        for (int k = 0; k < 3; k++) {
                            double d = cv[k] - m1[k];
                            m1[k] += d / (sIdx + 1);
                            m2[k] += d * (cv[k] - m1[k]);
                        }
                    }
                    if (ps.add(c)) break;
                }
                V3 mv = sum / (float) taken;
                if (mean) put3(mean + 3 * (size_t) pix, mv);
                if (spp_out) spp_out[pix] = taken;
                if (var)
                    for (int k = 0; k < 3; k++)
                        var[3 * (size_t) pix + k] = taken > 1 ? (float) (m2[k] / (taken - 1)) : 0.f;
            }
        }
        cClosest += cnt.closest;
        cAny += cnt.any;
        cSamples += samples;
        cInner += cnt.inner;
        cTris += cnt.tris;
        cInnerAny += cnt.innerAny;
        cTrisAny += cnt.trisAny;
    }
    if (counters) {
        counters[0] = cClosest;
        counters[1] = cAny;
        counters[2] = cSamples;
        counters[3] = cInner;
        counters[4] = cTris;
        counters[5] = cInnerAny;
        counters[6] = cTrisAny;
    }
}

// ---------------------------------------------------------------------------
// Radiance-cache network (oracle_cache.hpp)
// ---------------------------------------------------------------------------
extern "C" void orc_to_half(int64_t n, const float *in, float *out) {
    for (int64_t i = 0; i < n; i++) out[i] = orc_cache::toHalf(in[i]);
}
extern "C" void orc_cache_encode(int64_t n, const float *q, const float *bounds, float *x) {
    for (int64_t i = 0; i < n; i++) orc_cache::encode(q + orc_cache::kQ * i, bounds, x + orc_cache::kIn * i);
}
extern "C" void orc_cache_forward(int64_t n, const float *q, const float *bounds, const float *w, int half,
                                  float *y) {
    const orc_cache::Round R{half != 0};
    std::vector<float> acts[orc_cache::kLayers];
    float x0[orc_cache::kIn];
    for (int64_t i = 0; i < n; i++) {
        orc_cache::encode(q + orc_cache::kQ * i, bounds, x0);
        orc_cache::forward(w, x0, R, acts, y + 3 * i);
    }
}
extern "C" double orc_cache_gradient(int64_t n, const float *q, const float *t, const float *bounds, const float *w,
                                     int half, float clip, float *grad) {
    return orc_cache::gradient(w, n, q, t, bounds, orc_cache::Round{half != 0}, clip, grad);
}
extern "C" void orc_cache_adam(float *w, float *m1, float *m2, const float *grad, int64_t step, float lr, float b1,
                               float b2, float eps) {
    orc_cache::adam(w, m1, m2, grad, step, lr, b1, b2, eps);
}
