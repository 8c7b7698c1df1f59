/*
 * prr_oracle.h — TEST INFRASTRUCTURE ONLY.
 *
 * C entry points of the CPU oracle (a restatement of the reference renderer's
 * hot path; see oracle_core.hpp / oracle_scene.hpp). Only tests/,
 * __graft_entry__.smoke() and bench.py's CPU-baseline leg load it, as the
 * checker. All array arguments are HOST pointers.
 *
 * Parity pinning (what each function is checked against):
 *   orc_trace_*        golden vectors of the compiled reference BVH::Intersect /
 *                      IntersectP + brute force over all primitives
 *   orc_export_level   golden trees of the compiled reference SAH build
 *   orc_bsdf_*         golden vectors of the compiled reference DiffuseMaterial /
 *                      GlassMaterial + closed forms + hemisphere integrals
 *   orc_env_*          golden vectors of the compiled reference EnvMap +
 *                      sample/pdf consistency at texel centres
 *   orc_camera_rays    golden vectors of the compiled reference Camera
 *   orc_tonemap        golden vectors of the compiled reference tonemap::apply +
 *                      fixed points; ACES against its closed form
 *   orc_castray_splitmix  per-path golden vectors of the compiled reference
 *                      castRay with seeded SplitMix64 streams
 *   orc_render         built from orc_castray (PCG32 per pixel)
 *   orc_photon_*       brute-force radius queries; emission statistics
 */
#ifndef PRR_ORACLE_H
#define PRR_ORACLE_H

#include <stdint.h>

#include "../include/prr_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_scene orc_scene;

orc_scene* orc_scene_create(const prr_scene_desc* desc);
void orc_scene_free(orc_scene* s);
/* level -1: scene BVH; k>=0: sub-BVH of object k (meshes only). Same layout
 * as prr_scene_export_level. Returns node count (copies at most cap). */
int orc_export_level(const orc_scene* s, int level, int cap, float* bounds6,
                     uint32_t* first_child, uint32_t* meta);
float orc_total_emit_area(const orc_scene* s);
/* leaf-order primitive list (local index within the object for meshes,
 * object index for level -1) and per-node areas (BVH.hpp:135) */
int orc_export_order(const orc_scene* s, int level, int cap, int32_t* order, float* areas);
/* Scene::sampleLight (Scene.cpp:37-58) on SplitMix64 streams from seeds[k] */
void orc_sample_light(const orc_scene* s, int64_t n, const uint64_t* seeds, float* pos, float* nrm,
                      float* pdf);

/* counters: if non-NULL, 3 uint32 per ray: interior expansions, triangle
 * tests, sphere tests (restated reference traversal). */
void orc_trace_closest(const orc_scene* s, int64_t n, const float* org, const float* dir,
                       int32_t* prim, float* t, float* u, float* v, uint32_t* counters);
void orc_trace_any(const orc_scene* s, int64_t n, const float* org, const float* dir,
                   const float* tmax, uint8_t* hit, uint32_t* counters);
void orc_brute_closest(const orc_scene* s, int64_t n, const float* org, const float* dir,
                       int32_t* prim, float* t);
void orc_brute_any(const orc_scene* s, int64_t n, const float* org, const float* dir,
                   const float* tmax, uint8_t* hit);

/* params: 8 floats per entry (type, base rgb, roughness, metallic, ior, pad) */
void orc_bsdf_eval(int64_t n, const float* wi, const float* wo, const float* nrm,
                   const float* params, float* f, float* pdf);
void orc_bsdf_sample(int64_t n, const float* wo, const float* nrm, const float* params,
                     const float* u3, float* wi, float* f, float* pdf, int32_t* lobe);
/* Medium::intersect (bounds: min xyz, max xyz per entry; t_enter / t_exit kept
 * on a miss), phaseHG, samplePhaseHG with two draws per entry */
void orc_medium(int64_t n, const float* org, const float* dir, const float* bounds, const int32_t* bounded,
                float* t_enter, float* t_exit, uint8_t* hit);
void orc_phase_hg(int64_t n, const float* cos_theta, const float* g, float* out);
void orc_phase_hg_sample(int64_t n, const float* wo, const float* g, const float* u2, float* wi);
void orc_ggx_terms(int64_t n, const float* in4, float* out6); /* NdotH,NdotWo,NdotWi,alpha ->
                                                                 D,Lambda(wo),G1(wo),G2,pdfVNDF(WoDotH=NdotWo),Fschlick(F0=0.04) */

void orc_env_build(int w, int h, const float* rgb, float* cdf, float* marg, float* total);
void orc_env_sample(int w, int h, const float* rgb, int64_t n, const float* u2, float* dir,
                    float* pdf);
void orc_env_eval(int w, int h, const float* rgb, int64_t n, const float* dir, float* radiance,
                  float* pdf);

void orc_camera_rays(const prr_camera_desc* cam, int width, int height, int64_t n,
                     const int32_t* pix, const float* offs, float* org, float* dir);
void orc_tonemap(int op, int64_t n, float exposure, const float* linear, float* display,
                 uint8_t* rgb8);

/* castRay per path with the reference's SplitMix64 stream seeded per path
 * (state = seeds[k]); draws: xOff, yOff, then castRay's own. */
void orc_castray_splitmix(const orc_scene* s, int width, int height, int64_t n,
                          const int32_t* pix, const uint64_t* seeds, int max_depth,
                          float indirect_clamp, float* L);

/* castRay on caller rays with the PCG32 stream of (seed, k, sample 0) */
void orc_castray_pcg(const orc_scene* s, int64_t n, const float* org, const float* dir, uint64_t seed,
                     int max_depth, float indirect_clamp, float* L);

/* Render with PCG32 per (pixel, sample) (the GPU renderer's RNG).
 * mean: W*H*3 (sum in sample order / samples taken), var: W*H*3 sample
 * variance of the per-sample radiance (may be NULL), spp_out: samples taken
 * per pixel (may be NULL; < spp only under the adaptive stop rule of
 * prr_render_params). rows [row0,row1) only (row1<=0: all).
 * counters (may be NULL, 7 entries): closest queries, any-hit queries, samples,
 * closest-hit interior expansions, closest-hit triangle tests, any-hit
 * interior expansions, any-hit triangle tests. threads: OpenMP
 * threads (<=0: default). Photon estimate if photon_map != NULL. */
/* PCG32 (O'Neill 2014): pcg32_srandom_r(initstate, initseq) then n outputs
 * of pcg32_random_r and their floats; and the per-pixel stream's floats. */
void orc_pcg32_sequence(uint64_t initstate, uint64_t initseq, int n, uint32_t* out, float* fout);
void orc_pcg32_stream(uint64_t seed, uint32_t pixel, uint32_t sample, int n, float* fout);

typedef struct orc_photon_map orc_photon_map;
void orc_render(const orc_scene* s, const prr_render_params* p, int row0, int row1,
                const orc_photon_map* pm, int threads, float* mean, float* var,
                int32_t* spp_out, uint64_t* counters);

/* orc_render at fixed spp where every path that is not a radiance-cache
 * training path (lowbias32(pixel, sample) & train_mask != 0) ends at its first
 * diffuse vertex at bounce >= query_bounce, and training paths at bounce >=
 * train_stop (0: they run to full depth): the cache render with a network that
 * outputs zero. mean: W*H*3. */
void orc_render_truncated(const orc_scene* s, const prr_render_params* p, int row0, int row1,
                          int threads, int query_bounce, uint32_t train_mask, int train_stop, float* mean,
                          uint64_t* counters);

/* Renderer.cpp:71-99 on given per-sample values: nseq sequences of len RGB
 * samples; stop rule of Welford (Renderer.hpp:14-40) per sequence with its own
 * min_spp / threshold. Writes samples taken and their mean. */
void orc_welford(int nseq, int len, const float* seq, const int32_t* min_spp, const float* thr,
                 int32_t* count, float* mean);

/* Caustic photon map (Photon.cpp, build extensions: point-light emission and
 * hash-grid storage). */
orc_photon_map* orc_photon_build(const orc_scene* s, int64_t num_photons, float radius,
                                 int max_depth, uint64_t seed, int threads);
/* Scene::buildPhotonMaps with the reference's own rules (Photon.cpp:96-188:
 * area emitters, flipped normal at delta vertices, storage at every diffuse
 * hit after a delta vertex, Russian-roulette diffuse continuation), PCG32
 * stream per photon as orc_photon_build. */
orc_photon_map* orc_photon_emit(const orc_scene* s, int64_t num_photons, float radius, int max_depth,
                                uint64_t seed, int threads);
/* The same per photon k with SplitMix64 state states[k] and power divisor
 * num_photons; writes up to cap stored photons (pos, dir, power) and their
 * emitting photon; returns the count. */
int64_t orc_photon_emit_splitmix(const orc_scene* s, int64_t n, const uint64_t* states, float num_photons,
                                 int max_depth, int64_t cap, float* out9, int32_t* owner);
/* a map over n given photons (pos, dir, power per row of ph9) */
orc_photon_map* orc_photon_map_from(int64_t n, const float* ph9, float radius);
/* emitting photon and trace depth per stored photon (orc_photon_emit maps) */
int64_t orc_photon_owners(const orc_photon_map* pm, int64_t cap, int64_t* owner, int32_t* depth);
void orc_photon_free(orc_photon_map* pm);
int64_t orc_photon_count(const orc_photon_map* pm);
/* copies up to cap photons: pos3, dir3, power3 (9 floats each) */
int64_t orc_photon_export(const orc_photon_map* pm, int64_t cap, float* out9);
/* irradiance estimate sum(power)/(pi r^2) over photons within r (Photon.cpp:82-94) */
void orc_photon_irradiance(const orc_photon_map* pm, int64_t n, const float* pos, float* out3);
/* BSDF-weighted estimate sum f(wi_p, wo) Phi_p / (pi r^2), f = eval / cos
 * (the term the renderer adds at camera/delta-reached diffuse vertices) */
void orc_photon_radiance(const orc_photon_map* pm, int64_t n, const float* pos, const float* wo,
                         const float* nrm, const float* params, float* out3);

int orc_version(void);

/* Radiance-cache network (oracle_cache.hpp). half != 0 reproduces the
 * device's binary16 rounding points; half == 0 is exact float arithmetic. */
void orc_to_half(int64_t n, const float* in, float* out);
void orc_cache_encode(int64_t n, const float* query13, const float* bounds6, float* x96);
void orc_cache_forward(int64_t n, const float* query13, const float* bounds6, const float* weights,
                       int half, float* y3);
double orc_cache_gradient(int64_t n, const float* query13, const float* target3, const float* bounds6,
                          const float* weights, int half, float clip, float* grad);
void orc_cache_adam(float* w, float* m1, float* m2, const float* grad, int64_t step, float lr,
                    float beta1, float beta2, float eps);

#ifdef __cplusplus
}
#endif
#endif
