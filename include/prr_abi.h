/*
 * prr_abi.h — C ABI of libprr.so, the MI355X (gfx950) path-tracing hot path.
 *
 * This is the drop-in boundary that replaces the per-pixel CPU loop of the
 * reference renderer:
 *   Renderer::Render(const Scene&, const Integrator&, const RenderSettings&)
 *       Renderer.hpp:57-59, Renderer.cpp:24-212
 *   Integrator::castRay(const Ray&)          Integrator.hpp:23, Integrator.cpp:104-263
 *   Scene::intersect / Scene::intersectP     Scene.cpp:29-35
 *   BVHAccel::Intersect / IntersectP         BVH.hpp:94-95, BVH.cpp:311-473
 *
 * Conventions
 *  - Every entry point returns an int status: PRR_OK (0) or a negative
 *    PRR_ERR_* code; prr_last_error(ctx) returns the message. No C++
 *    exception crosses this boundary.
 *  - Scene descriptors are HOST memory, borrowed for the duration of the call
 *    and copied. Device-side scene buffers are owned by the prr_scene and
 *    released by prr_scene_free.
 *  - Array arguments of the batch entry points (prr_trace_*, prr_bsdf_*,
 *    prr_env_*, prr_camera_rays, prr_tonemap, prr_render outputs) are DEVICE
 *    pointers (HBM), caller-allocated. Work is enqueued on the context's
 *    stream (prr_set_stream) and is asynchronous unless stated otherwise.
 *  - A context is externally synchronized: one host thread at a time.
 *    Distinct contexts may be used concurrently.
 *  - Vectors are packed float triples (x,y,z), row-major.
 */
#ifndef PRR_ABI_H
#define PRR_ABI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PRR_OK 0
#define PRR_ERR_INVALID (-1)     /* bad argument / shape */
#define PRR_ERR_DEVICE (-2)      /* HIP runtime error or no device */
#define PRR_ERR_NOMEM (-3)       /* device allocation failed / over budget */
#define PRR_ERR_UNSUPPORTED (-4) /* feature not available for this scene */

/* Material classes (Material.hpp:100-183). */
enum prr_material_type {
    PRR_MAT_DIFFUSE = 0,  /* DiffuseMaterial: Lambert*(1-F)*(1-m) + GGX (Material.cpp:28-158) */
    PRR_MAT_GLASS = 1,    /* GlassMaterial: Fresnel dielectric, delta (Material.cpp:169-191) */
    PRR_MAT_MIRROR = 2,   /* MirrorMaterial: delta reflection (Material.cpp:160-167) */
    PRR_MAT_EMISSIVE = 3  /* EmissiveMaterial: emits, never scatters (Material.hpp:135-152) */
};

/* Object kinds of Scene::objects (Scene.hpp:18). */
enum prr_object_kind {
    PRR_OBJ_MESH = 0,   /* MeshTriangle (Triangle.hpp:112-180), owns a sub-BVH of <=4 prims/leaf */
    PRR_OBJ_SPHERE = 1  /* Sphere (Sphere.hpp:11-66) */
};

/* Tone-map operators (ToneMapping.hpp:32-37) plus ACES (Narkowicz 2015 fit). */
enum prr_tonemap_op {
    PRR_TONEMAP_RAW = 0,
    PRR_TONEMAP_STANDARD = 1,
    PRR_TONEMAP_AGX = 2,
    PRR_TONEMAP_PBR_NEUTRAL = 3,
    PRR_TONEMAP_ACES = 4
};

/* Per-triangle attribute flags (Triangle::hasSmoothNormals / hasTangents;
 * PRR_TRI_UVS marks rows whose t0..t2 came from the asset). */
enum prr_tri_flag {
    PRR_TRI_SMOOTH_NORMALS = 1,  /* hit normal = normalize(n0*w + n1*u + n2*v); light-sample normal = normalize(n0+n1+n2) */
    PRR_TRI_TANGENTS = 2,        /* per-vertex tangents: normal maps use the asset's tangent frame */
    PRR_TRI_UVS = 4
};

/* Texture (TextureUtils.hpp:9-19): 8-bit texels, rows top to bottom,
 * `channels` bytes per texel (>= 3 for colour lookups). */
typedef struct prr_texture_desc {
    int32_t width;
    int32_t height;
    int32_t channels;
    const uint8_t* data;  /* host */
} prr_texture_desc;

/* Material record. Field meaning follows Material.hpp:42-47. */
typedef struct prr_material_desc {
    int32_t type;          /* prr_material_type */
    float base_color[3];   /* Material::baseColor (default 0.8) */
    float roughness;       /* Material::roughness; clamped to >=0.08 when shading */
    float metallic;        /* Material::metallic */
    float ior;             /* Material::ior (default 1.5) */
    float emission[3];     /* Material::m_emission */
} prr_material_desc;

/* One entry of Scene::objects, in Scene::Add order. The order fixes the
 * scene-level SAH build input, the emitter list order (Scene.cpp:19-25) and
 * the global primitive ids: a triangle's id is its row in
 * prr_scene_desc::triangles; sphere row s has id num_triangles + s. */
typedef struct prr_object_desc {
    int32_t kind;      /* prr_object_kind */
    int32_t material;  /* row in prr_scene_desc::materials */
    int32_t first;     /* mesh: first triangle row; sphere: sphere row */
    int32_t count;     /* mesh: number of triangles (>=1); sphere: 1 */
} prr_object_desc;

/* Pinhole / thin-lens camera (Camera.hpp:8-20, init :43-56). */
typedef struct prr_camera_desc {
    float eye[3];
    float target[3];
    float up[3];
    float fovh;            /* horizontal field of view, degrees (fovv is unused by the reference) */
    float aperture;        /* 0 = pinhole */
    float focus_distance;  /* <0: |target - eye| */
} prr_camera_desc;

/* Equirectangular linear-RGB environment (EnvMap.hpp). width == 0: none. */
typedef struct prr_envmap_desc {
    int32_t width;
    int32_t height;
    const float* pixels;   /* host, width*height*3 floats, row 0 = zenith (y-up) */
} prr_envmap_desc;

/* Point light (PointLight.hpp:9-16). The reference declares but never uses
 * it; here it contributes NEE (intensity / d^2, delta) and photon emission. */
typedef struct prr_point_light_desc {
    float position[3];
    float intensity[3];    /* radiant intensity (W/sr) per channel */
} prr_point_light_desc;

typedef struct prr_scene_desc {
    int32_t num_objects;
    const prr_object_desc* objects;
    int32_t num_triangles;
    const float* triangles;       /* 9 floats per triangle: v0, v1, v2 (counter-clockwise) */
    int32_t num_spheres;
    const float* spheres;         /* 4 floats per sphere: center xyz, radius */
    int32_t num_materials;
    const prr_material_desc* materials;
    prr_envmap_desc env;
    float background[3];          /* Scene::backgroundColor for escaped rays without env */
    int32_t num_point_lights;
    const prr_point_light_desc* point_lights;
    prr_camera_desc camera;
    /* Per-vertex attributes and textures (Triangle.hpp:20-31, Triangle.cpp:34-59,
     * Material.cpp:3-26). All optional: a zero-initialised tail means flat
     * triangles, uv (0,0) and untextured materials. */
    const uint8_t* tri_flags;        /* NULL, or per triangle: prr_tri_flag bits */
    const float* tri_normals;        /* 9 per triangle: n0, n1, n2 (PRR_TRI_SMOOTH_NORMALS) */
    const float* tri_uvs;            /* 6 per triangle: t0, t1, t2 (PRR_TRI_UVS) */
    const float* tri_tangents;       /* 12 per triangle: tangent xyz + handedness w per vertex (PRR_TRI_TANGENTS) */
    const int32_t* material_textures;/* NULL, or 4 per material: base colour, metallic-roughness,
                                        normal map, emissive; row in textures, -1 none */
    int32_t num_textures;
    const prr_texture_desc* textures;
} prr_scene_desc;

/* RenderSettings (Renderer.hpp:42-55) plus the Integrator parameters
 * (Integrator.hpp:18-21). */
typedef struct prr_render_params {
    int32_t width;
    int32_t height;
    int32_t spp;               /* RenderSettings::maxSPP */
    int32_t max_depth;         /* Integrator maxDepth */
    float indirect_clamp;      /* Integrator::indirectClamp (<=0 disables) */
    int32_t sample_offset;     /* index of the first sample (progressive / split renders) */
    uint64_t seed;             /* per-pixel PCG32 seed */
    /* > 0: gather the scene's resident caustic photon map (prr_photon_build)
     * at diffuse vertices reached from the camera or through delta bounces;
     * radius / depth are those of the build (fields kept for reporting) */
    int64_t num_photons;
    float photon_radius;
    int32_t photon_max_depth;
    int32_t flags;             /* PRR_RENDER_* bits */
    /* Adaptive sampling (Renderer.cpp:74-99, Welford Renderer.hpp:14-40):
     * once a pixel has min_spp samples it stops as soon as the 95% CI of all
     * three channel means is below variance_threshold * |mean|. With
     * variance_threshold <= 0 or min_spp >= spp every pixel takes spp samples. */
    int32_t min_spp;           /* RenderSettings::minSPP */
    float variance_threshold;  /* RenderSettings::varianceThreshold */
    int32_t reserved;
} prr_render_params;

/* prr_render_params::flags */
#define PRR_RENDER_KERNEL_TIMING 1 /* time every kernel launch with HIP events (prr_stats::ms_*) */

typedef struct prr_stats {
    uint64_t rays_closest;     /* scene-level closest-hit queries */
    uint64_t rays_any;         /* scene-level any-hit (shadow) queries */
    uint64_t samples;          /* castRay-equivalents */
    uint64_t photons_stored;   /* caustic photons stored */
    double ms_total;           /* device time of the render, ms */
    /* with PRR_RENDER_KERNEL_TIMING: summed launch durations per kernel */
    double ms_extend;          /* closest-hit traversal (extension rays) */
    double ms_shadow;          /* any-hit traversal (shadow rays) */
    double ms_shade;           /* shading / NEE / BSDF sampling */
    int64_t launches_extend;
    int64_t launches_shadow;
    /* path completion: once few paths remain, one launch runs each of them to
       its end (closest hits, shading and shadow rays in one kernel) */
    uint64_t rays_closest_finish; /* closest-hit queries traced there (part of rays_closest) */
    double ms_finish;             /* its duration (with PRR_RENDER_KERNEL_TIMING) */
    int64_t launches_finish;
    /* samples whose radiance was not finite and was replaced by 0
       (Integrator.cpp:259-261, Renderer.cpp:80-83) */
    uint64_t nan_samples;
    uint64_t rays_any_finish;     /* shadow rays traced by path completion (part of rays_any) */
    double ms_sort;               /* shadow-ray binning before traversal (with PRR_RENDER_KERNEL_TIMING) */
    /* NEE queries the reference makes whose term is exactly zero in every
       channel (the BSDF vanishes for the light direction): the sample cannot
       depend on their answer, so their shadow rays are not traced (not part
       of rays_any) */
    uint64_t rays_any_skipped;
} prr_stats;

typedef struct prr_scene_info {
    int32_t num_nodes;         /* interior nodes of the merged device BVH */
    int32_t num_leaf_prims;    /* primitive slots in leaf order */
    int32_t max_depth;         /* deepest root-to-leaf path */
    int32_t num_emitters;      /* Scene::emissives */
    float total_emit_area;     /* Scene::totalEmitArea */
    float bounds[6];           /* scene root box min xyz, max xyz */
    int64_t device_bytes;      /* HBM held by the scene */
    int32_t num_wide_nodes;    /* nodes of the four-wide occlusion tree (shadow rays, prr_trace_any) */
    int32_t wide_levels_kept;  /* binary levels left uncollapsed there (a child box outside its parent's) */
    int32_t two_level;         /* 1: two-level records of the binary tree on the device (a context created
                                  with PRR_TWO_LEVEL=1, and every box the union of its children's), 0: not */
    int32_t reserved;
} prr_scene_info;

typedef struct prr_context prr_context;
typedef struct prr_scene prr_scene;

/* ---- context ------------------------------------------------------------ */
/* No reference counterpart: the reference has no device. */
int prr_create(int device, prr_context** out);
int prr_destroy(prr_context* ctx);
const char* prr_last_error(const prr_context* ctx); /* ctx may be NULL */
int prr_set_stream(prr_context* ctx, void* hip_stream); /* NULL = context's own stream */
int prr_synchronize(prr_context* ctx);

/* ---- scene: Scene::buildBVH (Scene.cpp:15-27) + upload ------------------ */
/* Builds the reference's two-level binned-SAH BVH on the host (scene level
 * maxPrims 1, mesh level maxPrims 4; BVH.cpp:13-280), merges it into one
 * device tree, builds the emitter tables (Scene.cpp:19-25, BVH.cpp:478-514)
 * and the environment CDF (EnvMap.hpp:189-224), and uploads everything. */
int prr_scene_create(prr_context* ctx, const prr_scene_desc* desc, prr_scene** out);
int prr_scene_free(prr_context* ctx, prr_scene* scene);
int prr_scene_get_info(const prr_scene* scene, prr_scene_info* out);
/* Host copy of the flattened reference-layout tree of one level, for tests:
 * level -1 = scene BVH over objects, level k>=0 = sub-BVH of object k.
 * Writes up to cap nodes (6 bounds floats [minX,maxX,minY,maxY,minZ,maxZ],
 * first_child, meta as in BVH.hpp:34-67) and returns the node count. */
int prr_scene_export_level(const prr_scene* scene, int32_t level, int32_t cap,
                           float* bounds6, uint32_t* first_child, uint32_t* meta,
                           int32_t* out_count);

/* ---- traversal (BVHAccel::Intersect / IntersectP) ----------------------- */
/* Closest hit of n rays (BVH.cpp:311-415 + Triangle.hpp:46-64 + Sphere.hpp:22-44).
 * prim = global primitive id or -1; t, u, v = hit distance and barycentrics
 * (0 for spheres). Rays may be unnormalized exactly as in the reference. */
int prr_trace_closest(prr_context* ctx, const prr_scene* scene, int64_t n,
                      const float* org, const float* dir,
                      int32_t* prim, float* t, float* u, float* v);
/* No counterpart in the reference: prr_trace_closest for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). Same kernel and results (bit for bit) as
 * prr_trace_closest, but the traversal's global spill stacks live in a
 * static device array of PRR_NOALLOC_SCRATCH_BYTES (32 MiB: 256 blocks of
 * 256 lanes, 64 entries of 8 bytes per lane) that each GPU the library runs
 * on holds from the first prr_create until the process ends, instead of the
 * context's own allocation. The call allocates nothing (no hipMalloc, no
 * hipFree). The scratch does not grow with n: larger batches loop over the
 * same lanes; n above PRR_NOALLOC_MAX_RAYS is rejected. Every call on a GPU
 * shares that GPU's array: calls that overlap on one GPU (two contexts or
 * streams at once) race and must be serialised by the caller. Nothing in the
 * array needs clearing between calls (each lane writes its stack before
 * reading it). */
#define PRR_NOALLOC_SCRATCH_BYTES (256ll * 256 * 64 * 8)
#define PRR_NOALLOC_MAX_RAYS (((int64_t) 1 << 31) - 1)
int prr_trace_closest_noalloc(prr_context* ctx, const prr_scene* scene, int64_t n,
                              const float* org, const float* dir,
                              int32_t* prim, float* t, float* u, float* v);
/* Any hit in (1e-3, tmax) (BVH.cpp:421-473, Triangle.hpp:75-92). */
int prr_trace_any(prr_context* ctx, const prr_scene* scene, int64_t n,
                  const float* org, const float* dir, const float* tmax, uint8_t* hit);

/* ---- BSDF / environment / camera / tone map batch entry points ---------- */
/* params: 8 floats per entry: type, base rgb, roughness, metallic, ior, pad.
 * Shading frame is built from N alone (Material.hpp:80-84). */
/* DiffuseMaterial::eval/pdf with LOBE_ALL (Material.cpp:80-158). */
int prr_bsdf_eval(prr_context* ctx, int64_t n, const float* wi, const float* wo,
                  const float* normal, const float* params, float* f, float* pdf);
/* Material::sample with explicit uniforms u (3 per entry: u0,u1,u2 for
 * DiffuseMaterial::sample Material.cpp:28-78, u0 for GlassMaterial::sample
 * Material.cpp:169-191). lobe: 0 diffuse, 1 specular, 2 delta, -1 none. */
int prr_bsdf_sample(prr_context* ctx, int64_t n, const float* wo, const float* normal,
                    const float* params, const float* u, float* wi, float* f, float* pdf,
                    int32_t* lobe);
/* Participating media (Volume.hpp:7-66, Volume.cpp:3-24; Scene::medium is
 * not used by the reference's castRay, so these are batch functions only).
 * Medium::intersect (Volume.hpp:17-44) per ray against its own medium:
 * bounds 6 floats per ray (pMin xyz, pMax xyz as Bounds3 holds them),
 * bounded 0: the ray enters at 0 and leaves at infinity. hit[i] = the
 * function's result; t_enter / t_exit are written only where it is 1 (the
 * reference leaves its outputs untouched otherwise). */
int prr_medium_intersect(prr_context* ctx, int64_t n, const float* org, const float* dir,
                         const float* bounds, const int32_t* bounded, float* t_enter, float* t_exit,
                         uint8_t* hit);
/* phaseHG(cos_theta, g) (Volume.cpp:3-6). */
int prr_phase_hg(prr_context* ctx, int64_t n, const float* cos_theta, const float* g, float* out);
/* samplePhaseHG(wo, g) (Volume.cpp:8-24) with explicit uniforms, 2 per entry
 * (cos theta, then phi); wi is a unit direction (3 floats). */
int prr_phase_hg_sample(prr_context* ctx, int64_t n, const float* wo, const float* g, const float* u2,
                        float* wi);
/* Scene::sampleLight (Scene.cpp:37-58: emitter picked by area,
 * MeshTriangle -> BVHAccel::Sample's sqrt(u) leaf walk, BVH.cpp:478-514, ->
 * Triangle::Sample, Triangle.hpp:98-103; Sphere::Sample, Sphere.hpp:54-63)
 * with explicit uniforms, 4 per sample (pick, sqrt(u), x, y; a sphere uses
 * the first 3). Writes position and light normal (3 floats each), the row of
 * Scene::emissives picked (emitter, may be NULL) and the area pdf
 * 1 / totalEmitArea (pdf, may be NULL; Scene::pdfLight, Scene.cpp:60-62). */
int prr_sample_light(prr_context* ctx, const prr_scene* scene, int64_t n, const float* u4,
                     float* pos, float* normal, int32_t* emitter, float* pdf);
/* EnvMap::importanceSample with explicit uniforms (EnvMap.hpp:145-169). */
int prr_env_sample(prr_context* ctx, const prr_scene* scene, int64_t n, const float* u2,
                   float* dir, float* pdf);
/* EnvMap::sample (bilinear radiance, :97-113) and importanceSamplePdf (:124-133). */
int prr_env_eval(prr_context* ctx, const prr_scene* scene, int64_t n, const float* dir,
                 float* radiance, float* pdf);
/* Camera::generateRay (Camera.hpp:75-95) for pixel (i,j) with offsets (ox,oy);
 * pix: 2 int32 per ray, offs: 2 floats per ray (+2 lens uniforms if aperture>0). */
int prr_camera_rays(prr_context* ctx, const prr_scene* scene, int32_t width, int32_t height,
                    int64_t n, const int32_t* pix, const float* offs, float* org, float* dir);
/* exposure + sanitize (Renderer.cpp:149-161), tonemap::apply (ToneMapping.hpp:267-279)
 * and 8-bit quantisation (Renderer.cpp:169-171). display or rgb8 may be NULL. */
int prr_tonemap(prr_context* ctx, int32_t op, int64_t n, float exposure, const float* linear,
                float* display, uint8_t* rgb8);

/* ---- caustic photon map (Scene::buildPhotonMaps, Scene.cpp:8-13) -------- */
/* Traces num photons from the scene's point light (Photon.cpp:96-188 with the
 * build's changes: point-light emission, glass handled with the outward normal
 * as in castRay's delta branch, storage at the first diffuse hit after >= 1
 * delta bounce) and keeps the hash grid (cell = 2 * radius) resident in the
 * scene. Blocks. stored may be NULL. */
int prr_photon_build(prr_context* ctx, prr_scene* scene, int64_t num_photons, float radius,
                     int32_t max_depth, uint64_t seed, int64_t* stored);
/* Scene::buildPhotonMaps (Scene.cpp:8-13) = PhotonMap::emit + PhotonMap::build
 * (Photon.cpp:156-190) with the reference's own rules: num photons emitted
 * from Scene::emissives (Scene::sampleLight, a cosine hemisphere about the
 * light normal, power Le * A / num), traced as PhotonMap::trace does
 * (Photon.cpp:96-152: normal flipped to face the photon at delta vertices,
 * every diffuse hit after a delta vertex stores the photon, then Russian
 * roulette and a diffuse bounce), PCG32 stream per photon index. The map
 * stays resident in the scene (replacing any earlier one) for
 * prr_photon_irradiance; renders do not gather it (castRay never does).
 * Blocks. Fails with PRR_ERR_UNSUPPORTED when the scene has no emitter. */
int prr_photon_emit(prr_context* ctx, prr_scene* scene, int64_t num_photons, float radius,
                    int32_t max_depth, uint64_t seed, int64_t* stored);
/* PhotonMap::estimateIrradiance (Photon.cpp:82-94) at n points (device
 * arrays): the power of the resident map's photons within its radius over
 * pi r^2. normal is unused, as in the reference, and may be NULL. */
int prr_photon_irradiance(prr_context* ctx, const prr_scene* scene, int64_t n, const float* pos,
                          const float* normal, float* out);
/* For a prr_photon_emit map: per stored photon (the order of
 * prr_photon_export) the index of the photon that stored it and its trace
 * depth there, to HOST arrays of cap entries; *count = photons stored. */
int prr_photon_export_owners(prr_context* ctx, const prr_scene* scene, int64_t cap, int64_t* owner,
                             int32_t* depth, int64_t* count);
/* Copies up to cap stored photons (pos3, dir3, power3 = 9 floats each, bucket
 * order) to device memory out9; *count = photons stored. */
int prr_photon_export(prr_context* ctx, const prr_scene* scene, int64_t cap, float* out9,
                      int64_t* count);
/* BSDF-weighted photon density estimate at n points (the term the renderer
 * adds at camera- or delta-reached diffuse vertices); params as prr_bsdf_eval. */
int prr_photon_radiance(prr_context* ctx, const prr_scene* scene, int64_t n, const float* pos,
                        const float* wo, const float* normal, const float* params, float* out);

/* ---- the render: Renderer::Render sample loop + Integrator::castRay ----- */
/* Writes the per-pixel mean linear radiance (width*height*3 floats, device)
 * and, when spp_out is not NULL, the samples each pixel took (width*height
 * int32, device; the reference's sampleCount, Renderer.cpp:98). Sample k of
 * pixel p uses the PCG32 stream of (seed, p, sample_offset + k), so a pixel's
 * result does not depend on the image size, the batching or the other
 * pixels. Blocks until the render is done; stats may be NULL. */
int prr_render(prr_context* ctx, const prr_scene* scene, const prr_render_params* params,
               float* rgb_out, int32_t* spp_out, prr_stats* stats);

/* Integrator::castRay (Integrator.cpp:104-263) for n caller rays (device
 * arrays): ray k runs the full path with the PCG32 stream of (seed, k,
 * sample 0) and writes its radiance (non-finite -> 0) to L[3k..3k+2].
 * Blocks until done; stats may be NULL. */
int prr_cast_rays(prr_context* ctx, const prr_scene* scene, int64_t n, const float* org,
                  const float* dir, uint64_t seed, int32_t max_depth, float indirect_clamp, float* L,
                  prr_stats* stats);

/* ---- serving: continuous batching of render requests ------------------- */
/* No counterpart in the reference (Renderer::Render is single-shot,
 * Renderer.cpp:24). A server keeps scenes resident, admits requests in
 * arrival order while their memory fits the HBM budget, and runs batches of
 * up to path_slots paths through the wavefront kernels; every batch packs the
 * next samples of the admitted requests (one segment per request, shares set
 * by the policy below), so requests that arrive while others render join at
 * the next batch. A request's image depends only on (scene, camera, size, spp, seed):
 * each pixel takes samples in order with the PCG32 streams of prr_render, so
 * it is the same alone and in any mix. The server owns its context while it
 * runs (its worker thread makes every device call). */
typedef struct prr_server prr_server;

/* Batch shares (prr_server_params::policy) */
#define PRR_SERVE_AGED 0     /* shortest remaining work first, after subtracting the samples the
                                server renders in the time the request has waited */
#define PRR_SERVE_OLDEST 1   /* arrival order, each takes all it can */
#define PRR_SERVE_SHORTEST 2 /* shortest remaining work first, each takes all it can */
#define PRR_SERVE_FAIR 3     /* max-min fair: by remaining work, each takes min(need, equal share of the rest) */

typedef struct prr_server_params {
    int64_t path_slots;    /* paths per batch (1 << 22) */
    int64_t hbm_budget;    /* bytes of admitted request memory (16 B per pixel); 0: half the free HBM */
    int32_t policy;        /* PRR_SERVE_* */
    int32_t reserved[3];
} prr_server_params;

typedef struct prr_request_desc {
    int32_t scene;         /* index into the server's scenes */
    int32_t width, height;
    int32_t spp;
    prr_camera_desc camera;
    uint64_t seed;         /* per-pixel PCG32 seed (prr_render_params::seed) */
} prr_request_desc;

typedef struct prr_request_stats {
    double t_submit_ms;    /* server clock (ms since prr_server_create) */
    double t_admit_ms;
    double t_done_ms;
    int64_t samples;
    int32_t batches;       /* batches the request took part in */
    int32_t reserved;
} prr_request_stats;

typedef struct prr_server_stats {
    int64_t requests_done;
    int64_t batches;
    int64_t samples;       /* camera samples rendered */
    int64_t rays;          /* closest + any-hit queries */
    double busy_ms;        /* wall time the worker spent in batches */
    int64_t peak_admitted_bytes;
    int32_t peak_active;   /* most requests admitted at once */
    int32_t reserved;
} prr_server_stats;

/* scenes: num_scenes resident scenes created on ctx; max_depth and
 * indirect_clamp per scene (Integrator parameters). */
int prr_server_create(prr_context* ctx, const prr_scene* const* scenes, const int32_t* max_depth,
                      const float* indirect_clamp, int32_t num_scenes, const prr_server_params* params,
                      prr_server** out);
/* Stops the worker after the admitted requests finish; frees the server. */
int prr_server_destroy(prr_server* server);
/* Thread-safe; returns at once with the request id. A request larger than
 * the whole budget is rejected (PRR_ERR_NOMEM). */
int prr_server_submit(prr_server* server, const prr_request_desc* request, int64_t* id);
/* Blocks until request id is done; copies its mean linear radiance
 * (width*height*3 floats, HOST memory, may be NULL) and stats, then forgets
 * the request. */
int prr_server_wait(prr_server* server, int64_t id, float* rgb, prr_request_stats* stats);
int prr_server_get_stats(prr_server* server, prr_server_stats* stats);
const char* prr_server_last_error(const prr_server* server);

/* ---- learned radiance cache -------------------------------------------- */
/* No counterpart in the reference (its paths always run to full depth). An
 * FP16 MLP (Mueller et al. 2021, "Real-time Neural Radiance Caching"):
 * 96 encoded inputs, 5 hidden layers of 64 (ReLU), 3 outputs, no biases;
 * MFMA with FP32 accumulation; Adam on FP32 master weights. Trained online on
 * radiance the tracer produces; paths that use it stop after a short prefix
 * and add throughput * cache(vertex). */
#define PRR_CACHE_WEIGHTS 22720 /* 96*64 + 4*64*64 + 64*3 */
#define PRR_CACHE_QUERY 13      /* floats per query: position3, wo3 (towards the viewer), normal3,
                                   roughness, albedo3 */

typedef struct prr_cache prr_cache;

typedef struct prr_cache_params {
    float learning_rate;   /* Adam step size (1e-2) */
    float beta1, beta2;    /* Adam moment decays (0.9, 0.99) */
    float epsilon;         /* Adam epsilon (1e-8) */
    float grad_clip;       /* bound on |dLoss/dy| per output before the FP16 backward pass (100) */
    float bounds[6];       /* position-encoding box: min xyz, max xyz */
    uint64_t seed;         /* He-uniform initialisation */
    int32_t train_batch;   /* samples per Adam step (32768) */
    int32_t reserved[3];
} prr_cache_params;

int prr_cache_create(prr_context* ctx, const prr_cache_params* params, prr_cache** out);
int prr_cache_free(prr_context* ctx, prr_cache* cache);
/* FP32 master weights, PRR_CACHE_WEIGHTS floats (host): per layer W^T
 * [out][in] row-major, layers in order. set_weights also resets Adam. */
int prr_cache_get_weights(prr_context* ctx, const prr_cache* cache, float* weights);
int prr_cache_set_weights(prr_context* ctx, prr_cache* cache, const float* weights);
/* Forward pass on n queries (device arrays: query PRR_CACHE_QUERY floats
 * each, out 3 floats each). */
int prr_cache_infer(prr_context* ctx, const prr_cache* cache, int64_t n, const float* query, float* out);
/* Gradient of the mean relative L2 loss sum_c (y_c - t_c)^2 / (y_c^2 + 0.01)
 * / (3 n) over n samples (device arrays), without an update. grad (host,
 * PRR_CACHE_WEIGHTS floats) and loss (host) may be NULL. */
int prr_cache_gradient(prr_context* ctx, prr_cache* cache, int64_t n, const float* query,
                       const float* target, float* grad, float* loss);
/* steps Adam steps; step s uses samples [a, min(a + B, n)) with B =
 * train_batch and a = (s * B) mod n. loss (host, may be NULL): mean loss of
 * the last step, before its update. */
int prr_cache_train(prr_context* ctx, prr_cache* cache, int64_t n, const float* query,
                    const float* target, int32_t steps, float* loss);

/* One frame of a render that uses the cache. */
typedef struct prr_cache_render_params {
    int32_t use_cache;        /* 1: non-training paths stop at their first diffuse vertex at or after
                                 query_bounce and add throughput * cache; 0: plain paths (the fallback,
                                 per request) */
    int32_t train;            /* 1: one path in train_every runs on (to full depth, or see train_suffix),
                                 its diffuse vertices from bounce 1 on become training samples, and Adam
                                 runs after the frame */
    int32_t query_bounce;     /* 2 */
    int32_t train_every;      /* power of two (16) */
    int32_t max_train_steps;  /* Adam steps per frame at most (batches of train_batch) */
    int32_t train_capacity;   /* training samples kept per frame at most (0: 4M) */
    float divergence_loss;    /* a training loss above this, or not finite, marks the network as
                                 diverged: the next frame re-initialises it and renders plain */
    int32_t train_suffix;     /* 0: training paths run to full depth; n > 0 (with use_cache): they stop
                                 at their first diffuse vertex at or after query_bounce + n and take the
                                 cache's value there as their tail (self-training) */
} prr_cache_render_params;

typedef struct prr_cache_stats {
    prr_stats render;         /* rays, samples, device ms of the frame (render + cache) */
    uint64_t queries;         /* paths ended by a cache query */
    uint64_t train_samples;   /* training samples recorded (before the capacity cap) */
    int32_t train_steps;      /* Adam steps run */
    int32_t fell_back;        /* 1: the frame rendered plain because the network had diverged */
    float train_loss;         /* mean loss of the last step, before its update */
    int32_t diverged;         /* 1: this frame's training diverged (next frame falls back) */
    double ms_render;         /* raygen .. accumulate, including cache inference */
    double ms_train;          /* Adam steps (forward + backward + update) */
    double ms_infer;          /* cache inference launches (part of ms_render) */
} prr_cache_stats;

/* Renderer::Render with the cache: fixed spp (params->variance_threshold must
 * be <= 0), the PCG32 streams of prr_render. Writes the mean linear radiance
 * (device, width*height*3) and blocks until the frame and its training are
 * done. stats may be NULL. */
int prr_cache_render(prr_context* ctx, const prr_scene* scene, prr_cache* cache,
                     const prr_render_params* params, const prr_cache_render_params* cache_params,
                     float* rgb_out, prr_cache_stats* stats);

/* ------------------------------------------------------------------------
 * glTF input (host only, no context)
 * ------------------------------------------------------------------------ */

/* GLTFLoader::load (GLTFLoader.cpp:140-281): one mesh per triangle primitive
 * of every glTF mesh, in file order, positions as stored (node transforms are
 * not applied, as in the reference). Each mesh has its own material:
 * DiffuseMaterial defaults, then the primitive's glTF material (base colour,
 * metallic and roughness factors, emissive factor, the four textures;
 * GLTFLoader.cpp:64-138), then `material_type`'s copy rules
 * (GLTFLoader.cpp:191-224: Glass keeps colour, emission and every texture;
 * Mirror colour, emission and the emissive texture; Emissive a fresh
 * material with zero emission). Textures are decoded to 8-bit RGBA (16-bit
 * images reduced by >> 8). Normals are normalised per vertex. PNG images
 * only; a JPEG image is an error. All arrays are owned by the result. */
typedef struct prr_gltf_data {
    int32_t num_meshes;
    const int32_t* mesh_first;      /* first triangle row of each mesh */
    const int32_t* mesh_count;      /* triangles per mesh */
    const int32_t* mesh_material;   /* material row of each mesh */
    int32_t num_triangles;
    const float* positions;         /* 9 per triangle */
    const float* normals;           /* 9 per triangle (zeros without PRR_TRI_SMOOTH_NORMALS) */
    const float* uvs;               /* 6 per triangle */
    const float* tangents;          /* 12 per triangle */
    const uint8_t* tri_flags;       /* prr_tri_flag bits per triangle */
    int32_t num_materials;
    const prr_material_desc* materials;
    const int32_t* material_textures;  /* 4 per material, as prr_scene_desc::material_textures */
    int32_t num_textures;
    const prr_texture_desc* textures;
    void* impl;                     /* owner of the arrays */
} prr_gltf_data;

/* `material_type`: prr_material_type, GLTFLoader's MaterialType. On error
 * *out is NULL and prr_last_error(NULL) names the file and the cause. */
int prr_gltf_load(const char* path, int32_t material_type, prr_gltf_data** out);
void prr_gltf_free(prr_gltf_data* data);

#ifdef __cplusplus
}
#endif
#endif /* PRR_ABI_H */
