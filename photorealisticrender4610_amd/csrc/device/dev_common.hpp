// dev_common.hpp — device-side types and shading math for gfx950.
//
// Arithmetic keeps the reference's operation order (the library is compiled
// with -ffp-contract=off, so no FMA contraction) so that geometry results are
// bit-identical to the reference's CPU code and shading matches to ~1 ulp of
// the transcendental functions. Reference file:line citations are relative to
// the reference repository root.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace prr {

#define PRR_DEV __device__ __forceinline__
#define PRR_HD __host__ __device__ __forceinline__

constexpr float kPi = 3.141592653589793f;  // global.hpp:9-10
constexpr float kEps = 1e-3f;              // global.hpp:12

// std::max / std::min semantics (a < b ? b : a), including NaN operands.
PRR_HD float smax(float a, float b) { return (a < b) ? b : a; }
PRR_HD float smin(float a, float b) { return (b < a) ? b : a; }

struct V3 {
    float x, y, z;
    PRR_HD V3() : x(0.f), y(0.f), z(0.f) {}
    PRR_HD explicit V3(float s) : x(s), y(s), z(s) {}
    PRR_HD V3(float a, float b, float c) : x(a), y(b), z(c) {}
    PRR_HD float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
PRR_HD V3 operator+(const V3 &a, const V3 &b) { return V3(a.x + b.x, a.y + b.y, a.z + b.z); }
PRR_HD V3 operator-(const V3 &a, const V3 &b) { return V3(a.x - b.x, a.y - b.y, a.z - b.z); }
PRR_HD V3 operator*(const V3 &a, const V3 &b) { return V3(a.x * b.x, a.y * b.y, a.z * b.z); }
PRR_HD V3 operator/(const V3 &a, const V3 &b) { return V3(a.x / b.x, a.y / b.y, a.z / b.z); }
PRR_HD V3 operator*(const V3 &a, float r) { return V3(a.x * r, a.y * r, a.z * r); }
PRR_HD V3 operator*(float r, const V3 &a) { return V3(a.x * r, a.y * r, a.z * r); }
PRR_HD V3 operator/(const V3 &a, float r) { return V3(a.x / r, a.y / r, a.z / r); }
PRR_HD V3 operator-(const V3 &a) { return V3(-a.x, -a.y, -a.z); }
PRR_HD V3 &operator+=(V3 &a, const V3 &b) {
    a.x += b.x;
    a.y += b.y;
    a.z += b.z;
    return a;
}
PRR_HD float dot(const V3 &a, const V3 &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
PRR_HD V3 cross(const V3 &a, const V3 &b) {
    return V3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
// Vector.hpp:103-111
PRR_HD V3 normalize(const V3 &v) {
    float m2 = v.x * v.x + v.y * v.y + v.z * v.z;
    if (m2 > 0) {
        float inv = 1 / sqrtf(m2);
        return V3(v.x * inv, v.y * inv, v.z * inv);
    }
    return v;
}
PRR_HD float norm(const V3 &v) { return sqrtf(v.x * v.x + v.y * v.y + v.z * v.z); }
PRR_HD V3 normalized(const V3 &v) {  // Vector.hpp:28-31
    float n = sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
    return V3(v.x / n, v.y / n, v.z / n);
}
PRR_HD V3 lerp(const V3 &a, const V3 &b, float t) { return a * (1 - t) + b * t; }
PRR_HD float max3(const V3 &v) { return smax(smax(v.x, v.y), v.z); }  // std::max({x,y,z})
PRR_HD float min3(const V3 &v) { return smin(smin(v.x, v.y), v.z); }
PRR_DEV bool finite3(const V3 &v) { return isfinite(v.x) && isfinite(v.y) && isfinite(v.z); }
PRR_DEV V3 ld3(const float4 &f) { return V3(f.x, f.y, f.z); }
PRR_DEV float4 st4(const V3 &v, float w) { return make_float4(v.x, v.y, v.z, w); }

// ---------------------------------------------------------------------------
// PCG32 per (pixel, sample) stream (O'Neill 2014, pcg32_random_r); 24-bit
// floats like the reference's generator (global.hpp:49).
// ---------------------------------------------------------------------------
PRR_HD uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
struct Pcg {
    uint64_t state, inc;
    PRR_HD uint32_t nextU() {
        uint64_t old = state;
        state = old * 6364136223846793005ULL + inc;
        uint32_t xs = (uint32_t) (((old >> 18u) ^ old) >> 27u);
        uint32_t rot = (uint32_t) (old >> 59u);
        return (xs >> rot) | (xs << ((32u - rot) & 31u));
    }
    PRR_HD float next() { return (float) (nextU() >> 8) * 0x1p-24f; }
    PRR_HD static Pcg make(uint64_t seed, uint32_t pixel, uint32_t sample) {
        Pcg r;
        r.state = 0;
        r.inc = ((uint64_t) sample << 1) | 1u;
        r.nextU();
        r.state += mix64(mix64(seed * 0x9E3779B97F4A7C15ULL + pixel) ^ ((uint64_t) sample * 0xD1B54A32D192ED03ULL));
        r.nextU();
        return r;
    }
};

// ---------------------------------------------------------------------------
// Scene view: device pointers + constants, passed by value to kernels.
// ---------------------------------------------------------------------------
// Merged BVH node: the two children's boxes and references (64 B).
struct __align__(16) DNode {
    float4 a;  // L.minx L.maxx L.miny L.maxy
    float4 b;  // L.minz L.maxz R.minx R.maxx
    float4 c;  // R.miny R.maxy R.minz R.maxz
    int4 d;    // Lref Rref pad pad
};
// Two levels of the binary tree in one 128-byte line (index = the binary
// node's): the boxes of the left child's two children (a0 b0 c0, laid out as
// a DNode's pair) and of the right child's (a1 b1 c1); a child that is a leaf
// stands as the first of its pair with an empty (inverted) second box. Each
// child's box is the union of its pair's, exactly as in the binary tree.
struct __align__(16) DNode2 {
    float4 a0, b0, c0;
    float4 a1, b1, c1;
    uint4 gref;  // grandchild refs: left pair, right pair
    uint4 cref;  // left child ref, right child ref, 0, 0
};
// Child reference: bit31 = leaf; leaf: bits 28..30 = count-1, bit 27 =
// sphere leaf, bits 0..26 = first leaf-order primitive slot.
constexpr uint32_t kLeafBit = 0x80000000u;
constexpr uint32_t kSphereBit = 0x08000000u;
PRR_HD bool refIsLeaf(uint32_t r) { return (r & kLeafBit) != 0u; }
PRR_HD uint32_t refCount(uint32_t r) { return ((r >> 28) & 7u) + 1u; }
PRR_HD uint32_t refFirst(uint32_t r) { return r & 0x07FFFFFFu; }
PRR_HD bool refSphere(uint32_t r) { return (r & kSphereBit) != 0u; }
PRR_HD uint32_t makeLeafRef(uint32_t first, uint32_t count, bool sphere) {
    return kLeafBit | ((count - 1u) << 28) | (sphere ? kSphereBit : 0u) | first;
}

// Four-wide node of the occlusion tree (128 B): up to four children's boxes
// (structure of arrays, child k in component k) and references. The tree is
// the merged binary tree with levels collapsed, keeping its leaves; empty
// slots hold an inverted box and kNoChild.
struct __align__(16) DNode4 {
    float4 lox, hix, loy, hiy, loz, hiz;
    uint4 ref;
    uint4 pad;
};
constexpr uint32_t kNoChild = 0x7FFFFFFFu;

// Quantised four-wide occlusion node (64 B, one cache line): the same
// children as a DNode4, boxes stored as bytes q on a per-axis grid, bound =
// fma(q, scale, origin) (exact: q * scale is a power-of-two multiple). The
// builder checks that every decoded box contains the child's box, so a ray
// that enters a child's box enters its decoded box (the slab test is monotone
// in the bounds). A leaf child refers to its leaf record (kLeafBit, the
// leaf's primitive count and sphere bit as in the binary tree, the record's
// float4 offset as `first`), whose exact box is tested before the leaf's
// primitives.
//   w[0..2] origin xyz, w[3..5] scale xyz,
//   w[6..11] lo x, hi x, lo y, hi y, lo z, hi z (child k in byte k),
//   w[12..15] child refs (kNoChild: empty; decoded box inverted)
struct __align__(16) DNode4Q {
    uint4 w0, w1, w2, w3;
};
// Leaf record of the quantised tree, in float4s: the exact box [minX, maxX,
// minY, maxY], [minZ, maxZ, binary-tree leaf ref bits, 0], then for a
// triangle leaf a copy of its triangles (three float4 each, as in the
// triangle array), so the box and the triangles are read together from one
// place; a sphere leaf reaches its spheres through the leaf ref.
constexpr int kLeafQHead = 2;

enum MatType { MAT_DIFFUSE = 0, MAT_GLASS = 1, MAT_MIRROR = 2, MAT_EMISSIVE = 3 };

// prim_n.w holds the material row in its low bits and the triangle's
// per-vertex attribute flags (PRR_TRI_*) above kPrimFlagShift.
constexpr int kPrimFlagShift = 28;
constexpr int kPrimMatMask = (1 << kPrimFlagShift) - 1;
constexpr int kPrimSmooth = 1, kPrimTangents = 2;
// Per-triangle attribute record (float4 rows): n0.xyz n1.x | n1.yz n2.xy |
// n2.z t0.xy t1.x | t1.y t2.xy - | tangent0.xyz w0 | tangent1 w1 | tangent2 w2
constexpr int kTriAttrRows = 7;

struct __align__(16) DMaterial {
    float4 base_type;   // base rgb, type (as float)
    float4 params;      // roughness (raw), metallic, ior, hasEmission (0/1)
    float4 emission;    // emission rgb, pad
};

// Emitter record (Scene::emissives, Scene.cpp:19-25)
struct __align__(16) DEmitter {
    float area;        // Object::getArea
    int kind;          // 0 mesh, 1 sphere
    int material;
    int lt_root;       // mesh: first node of its light tree; sphere: sphere row
    float total;       // mesh sub-BVH totalArea (BVH.cpp:189)
    int pad[3];
};
// Light tree node (reference flatten layout of the emitter mesh's sub-BVH):
// first = right child (interior) or first leaf slot; count 0 = interior;
// left_area = nodeAreas[this+1] (BVH.cpp:496).
struct __align__(16) DLightNode {
    int first;
    int count;
    float left_area;
    int pad;
};

// Camera (Camera.hpp:43-56): basis, aperture, focus distance and
// tan(fovh / 2), precomputed on the host.
struct CamView {
    float eye[3], fwd[3], right[3], up[3];
    float aperture, focus, tanhalf;
};

struct SceneView {
    const DNode *nodes;
    const DNode2 *nodes2;     // the same tree two levels per record (null: not built)
    const DNode4 *nodes4;     // occlusion tree (four-wide)
    const DNode4Q *nodes4q;   // the same tree quantised (one line per node)
    const float4 *leavesq;    // its leaf records (exact box + triangles)
    const float4 *tris;       // 3 float4 per leaf slot: v0.xyz e1.x | e1.yz e2.xy | e2.z id pad pad
    const float4 *sph;        // leaf-order spheres: center.xyz radius
    const int *sph_id;        // leaf-order sphere -> global prim id
    const float4 *prim_n;     // per global prim: geometric normal (triangles), w = material (bits)
    const float4 *sph_by_id;  // per sphere row: center, radius
    const DMaterial *mats;
    const DEmitter *emitters;
    const DLightNode *lnodes;
    const float4 *ltri;       // 4 float4 per light-tree leaf slot: v0.xyz area | v1 | v2 | normal
    const float4 *env_px;     // W*H (rgb, lum)
    const float *env_cdf;     // W*H
    const float *env_marg;    // H
    const int *env_gm;        // marginal guide: env_gk + 1 entries
    const int *env_gc;        // conditional guides: env_gk + 1 per row
    int env_gk;               // guide buckets (power of two), 0: full searches
    uint32_t root;            // root child reference
    uint32_t root4;           // occlusion tree root reference
    uint32_t root4q;          // quantised occlusion tree root reference
    int num_tris;
    int num_emitters;
    float total_emit_area;
    float root_box[6];        // [minX,maxX,minY,maxY,minZ,maxZ]
    // ray origins for which shadow rays may use the short form of the
    // quantised node test (the scene box widened by its extent on every
    // side); quant_fast 0 keeps every ray on the full form
    float safe_box[6];
    int quant_fast;
    int env_w, env_h;
    float env_total;
    float background[3];
    // point light (build extension): at most one
    int num_points;
    float point_pos[3];
    float point_int[3];
    // caustic photon map (build extension; count 0 = none): photons sorted by
    // hash bucket of their cell (cell edge = 2 * radius)
    const float4 *ph_pos;       // position, w unused
    const float4 *ph_dir;       // direction of travel
    const float4 *ph_pow;       // power (flux) rgb
    const uint32_t *ph_start;   // per bucket: first photon
    const uint32_t *ph_end;     // per bucket: one past the last photon
    uint32_t ph_count;
    uint32_t ph_mask;           // buckets - 1 (power of two)
    float ph_radius, ph_cell;
    CamView cam;  // the scene's camera
    // per-vertex attributes and textures (null: none)
    const float4 *tattr;        // kTriAttrRows per triangle
    const int4 *mat_tex;        // per material: base colour, metallic-roughness, normal, emissive rows (-1 none)
    const int4 *tex_info;       // per texture: first texel byte, width, height, channels
    const uint8_t *texels;
};

// ---------------------------------------------------------------------------
// Shading (BRDFUtils.hpp, Material.cpp)
// ---------------------------------------------------------------------------
struct Frame {
    V3 T, B, N;
};

// ---------------------------------------------------------------------------
// Participating media (Volume.hpp:7-66, Volume.cpp:3-24). The reference's
// integrator does not call them; they are here as batch entry points.
// ---------------------------------------------------------------------------
// Medium::intersect (Volume.hpp:17-44). Vector3f's const operator[] returns
// double (Vector.hpp:67,78-80), so each plane time is formed in double from
// the float bounds, origin and 1/d and rounded to float once; std::max(a, b)
// and std::min(a, b) return a when the comparison fails (NaN); dirIsNeg
// holds direction > 0.
PRR_DEV bool mediumIntersect(bool bounded, const float *lo, const float *hi, const V3 &o, const V3 &d, float &tEnter,
                             float &tExit) {
    if (!bounded) {
        tEnter = 0.0f;
        tExit = __builtin_inff();
        return true;
    }
    const float ov[3] = {o.x, o.y, o.z}, dv[3] = {d.x, d.y, d.z};
    float te = -3.40282347e38f, tx = 3.40282347e38f;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const float inv = 1.0f / dv[i];
        float t0 = (float) (((double) lo[i] - (double) ov[i]) * (double) inv);
        float t1 = (float) (((double) hi[i] - (double) ov[i]) * (double) inv);
        if (!(dv[i] > 0.f)) {
            const float t = t0;
            t0 = t1;
            t1 = t;
        }
        te = t0 < te ? te : t0;  // std::max(t_min, tEnter)
        tx = tx < t1 ? tx : t1;  // std::min(t_max, tExit)
    }
    if (te > tx || tx < 0.0f) return false;
    tEnter = te;
    tExit = tx;
    return true;
}
// phaseHG (Volume.cpp:3-6), in float (M_PI is a float literal, global.hpp:9-10).
PRR_DEV float phaseHG(float cosTheta, float g) {
    const float denom = 1.0f + g * g - 2.0f * g * cosTheta;
    return (1.0f - g * g) / (4.0f * kPi * powf(denom, 1.5f));
}
// buildTBN (BRDFUtils.hpp:15-30); the reference rebuilds it in every
// toLocal/toWorld call, it is built once here (identical values).
PRR_DEV Frame makeFrame(const V3 &N) {
    Frame f;
    f.N = N;
    if (fabsf(N.x) > fabsf(N.y)) {
        float invLen = 1.0f / sqrtf(N.x * N.x + N.z * N.z);
        f.T = V3(N.z * invLen, 0.0f, -N.x * invLen);
    } else {
        float invLen = 1.0f / sqrtf(N.y * N.y + N.z * N.z);
        f.T = V3(0.0f, N.z * invLen, -N.y * invLen);
    }
    f.B = cross(N, f.T);
    return f;
}
// samplePhaseHG (Volume.cpp:8-24) with its two draws: u1 for cos theta, u2
// for phi; the frame is buildTBN(wo).
PRR_DEV V3 samplePhaseHG(const V3 &wo, float g, float u1, float u2) {
    float cosTheta;
    if (fabsf(g) < 1e-3f) {
        cosTheta = 1.0f - 2.0f * u1;
    } else {
        const float sqr = (1.0f - g * g) / (1.0f - g + 2.0f * g * u1);
        cosTheta = (1.0f + g * g - sqr * sqr) / (2.0f * g);
    }
    const float sinTheta = sqrtf(fmaxf(0.0f, 1.0f - cosTheta * cosTheta));
    const float phi = 2.0f * kPi * u2;
    const Frame f = makeFrame(wo);
    return normalize(sinTheta * cosf(phi) * f.T + sinTheta * sinf(phi) * f.B + cosTheta * wo);
}

// ---------------------------------------------------------------------------
// Per-vertex attributes and textures (Triangle.cpp:34-59, TextureUtils.hpp,
// Material.cpp:3-26)
// ---------------------------------------------------------------------------
// Interpolated shading normal of a triangle with per-vertex normals:
// normalize(n0 * w0 + n1 * u + n2 * v), w0 = 1 - u - v (Triangle.cpp:47-50).
PRR_DEV V3 smoothNormal(const SceneView &s, int prim, float u, float v) {
    const float4 *a = s.tattr + (int64_t) kTriAttrRows * prim;
    const float4 r0 = a[0], r1 = a[1], r2 = a[2];
    const float w0 = 1.f - u - v;
    return normalize(V3(r0.x, r0.y, r0.z) * w0 + V3(r0.w, r1.x, r1.y) * u + V3(r1.z, r1.w, r2.x) * v);
}

// TextureUtils::uvToTexel (TextureUtils.hpp:24-39); a NaN coordinate reads texel 0.
PRR_DEV void uvToTexel(float u, float v, int w, int h, int &x0, int &y0, int &x1, int &y1, float &fx, float &fy) {
    u = u - floorf(u);
    v = v - floorf(v);
    float px = u * (float) w, py = v * (float) h;
    if (!(px >= 0.f)) px = 0.f;
    if (!(py >= 0.f)) py = 0.f;
    x0 = min((int) px, w) % w;
    y0 = min((int) py, h) % h;
    x1 = (x0 + 1) % w;
    y1 = (y0 + 1) % h;
    fx = px - floorf(px);
    fy = py - floorf(py);
}

// Bilinear lookup of channels 0..2, each texel read as b / 255, or as
// b / 255 * 2 - 1 for a normal map (TextureUtils.hpp:41-103).
template <bool SNORM>
PRR_DEV V3 texBilerp(const SceneView &s, int t, float u, float v) {
    const int4 ti = s.tex_info[t];
    int x0, y0, x1, y1;
    float fx, fy;
    uvToTexel(u, v, ti.y, ti.z, x0, y0, x1, y1, fx, fy);
    const uint8_t *px = s.texels + ti.x;
    auto at = [&](int x, int y) {
        const uint8_t *q = px + ((int64_t) y * ti.y + x) * ti.w;
        if (SNORM) return V3(q[0] / 255.f * 2.f - 1.f, q[1] / 255.f * 2.f - 1.f, q[2] / 255.f * 2.f - 1.f);
        return V3(q[0] / 255.f, q[1] / 255.f, q[2] / 255.f);
    };
    const V3 c00 = at(x0, y0), c10 = at(x1, y0), c01 = at(x0, y1), c11 = at(x1, y1);
    const V3 top = c00 + (c10 - c00) * fx;
    const V3 bottom = c01 + (c11 - c01) * fx;
    return top + (bottom - top) * fy;
}

PRR_DEV float srgbToLinear(float x) {  // TextureUtils.hpp:51-54
    x = fmaxf(0.f, x);
    return x <= 0.04045f ? x / 12.92f : powf((x + 0.055f) / 1.055f, 2.4f);
}

// Material::buildShadingData for a textured material (params.w != 0):
// base colour (sRGB -> linear), roughness (G) and metallic (B) replace the
// material's constants; a normal map tilts N in the frame of the asset's
// tangent (Gram-Schmidt against N, bitangent cross(N, T) * handedness;
// Material.hpp:61-74) or, without one, of buildTBN(N). `N` is the normal the
// caller shades with (the hit normal, or its flip towards the viewer).
struct ShadeData {
    V3 base;
    float rough, metal;
    V3 N;
};
PRR_DEV void texturedShading(const SceneView &s, int mat, int prim, int pflags, float u, float v, ShadeData &sd) {
    const int4 tx = s.mat_tex[mat];
    float tu = 0.f, tv = 0.f;  // Intersection::tcoords (spheres and rows without uvs: 0)
    const bool tri = prim < s.num_tris && s.tattr != nullptr;
    const float4 *a = tri ? s.tattr + (int64_t) kTriAttrRows * prim : nullptr;
    const float w0 = 1.f - u - v;
    if (tri) {
        const float4 r2 = a[2], r3 = a[3];
        tu = r2.y * w0 + r2.w * u + r3.y * v;
        tv = r2.z * w0 + r3.x * u + r3.z * v;
    }
    if (tx.x >= 0) {
        const V3 c = texBilerp<false>(s, tx.x, tu, tv);
        sd.base = V3(srgbToLinear(c.x), srgbToLinear(c.y), srgbToLinear(c.z));
    }
    if (tx.y >= 0) {
        const V3 c = texBilerp<false>(s, tx.y, tu, tv);  // G = roughness, B = metallic
        sd.rough = c.y;
        sd.metal = c.z;
    }
    if (tx.z >= 0) {
        const V3 Ng = sd.N;
        V3 T, B;
        bool haveT = false;
        if (tri && (pflags & kPrimTangents)) {
            const float4 t0 = a[4], t1 = a[5], t2 = a[6];
            // Triangle::finalize normalises the interpolated tangent, castRay again
            const V3 tg = normalize(normalize(V3(t0.x, t0.y, t0.z) * w0 + V3(t1.x, t1.y, t1.z) * u +
                                              V3(t2.x, t2.y, t2.z) * v));
            const float hw = t0.w * w0 + t1.w * u + t2.w * v;
            const float hand = hw >= 0.f ? 1.f : -1.f;
            T = tg - dot(Ng, tg) * Ng;
            const float tlen = sqrtf(T.x * T.x + T.y * T.y + T.z * T.z);
            if (tlen >= 1e-6f) {
                T = T / tlen;
                B = cross(Ng, T) * hand;
                haveT = true;
            }
        }
        if (!haveT) {
            const Frame f = makeFrame(Ng);
            T = f.T;
            B = f.B;
        }
        const V3 tn = normalize(texBilerp<true>(s, tx.z, tu, tv));
        sd.N = normalize(tn.x * T + tn.y * B + tn.z * Ng);
    }
}
PRR_DEV V3 toLocal(const Frame &f, const V3 &w) { return V3(dot(w, f.T), dot(w, f.B), dot(w, f.N)); }
PRR_DEV V3 toWorld(const Frame &f, const V3 &l) { return l.x * f.T + l.y * f.B + l.z * f.N; }
PRR_DEV V3 reflect(const V3 &wo, const V3 &N) { return 2.f * dot(wo, N) * N - wo; }

PRR_DEV float D_GGX(float NdotH, float alpha) {  // BRDFUtils.hpp:90-94
    float a2 = alpha * alpha;
    float d = NdotH * NdotH * (a2 - 1.f) + 1.f;
    return a2 / (kPi * d * d);
}
PRR_DEV float lambdaGGX(float NdotV, float alpha) {  // BRDFUtils.hpp:102-106
    float c2 = NdotV * NdotV;
    float t2 = (1.f - c2) / smax(c2, 1e-6f);
    return 0.5f * (-1.f + sqrtf(1.f + alpha * alpha * t2));
}
PRR_DEV V3 F_Schlick(float WoDotH, const V3 &F0) {  // BRDFUtils.hpp:143-146
    // (1 - WoDotH)^5 by multiplication (within 2 ulp of pow(x, 5))
    const float x = smax(0.f, 1.f - WoDotH), x2 = x * x;
    float t = x2 * x2 * x;
    return F0 + (V3(1.f) - F0) * t;
}
PRR_DEV V3 sampleVNDF(const V3 &wo, float alpha, float u1, float u2) {  // BRDFUtils.hpp:163-181
    V3 Vh = normalize(V3(alpha * wo.x, alpha * wo.y, wo.z));
    float lensq = Vh.x * Vh.x + Vh.y * Vh.y;
    V3 T1 = lensq > 0.f ? V3(-Vh.y, Vh.x, 0.f) / sqrtf(lensq) : V3(1.f, 0.f, 0.f);
    V3 T2 = cross(Vh, T1);
    float r = sqrtf(u1);
    float phi = 2.f * kPi * u2;
    float t1 = r * cosf(phi);
    float t2 = r * sinf(phi);
    float s = 0.5f * (1.f + Vh.z);
    t2 = (1.f - s) * sqrtf(1.f - t1 * t1) + s * t2;
    V3 Nh = t1 * T1 + t2 * T2 + sqrtf(smax(0.f, 1.f - t1 * t1 - t2 * t2)) * Vh;
    return normalize(V3(alpha * Nh.x, alpha * Nh.y, smax(0.f, Nh.z)));
}

// Per-hit material constants of DiffuseMaterial (Material.cpp:28-158).
struct Lobes {
    V3 base, F0;
    float alpha, metallic, specWeight;
};
PRR_DEV Lobes makeLobes(const V3 &base, float roughness, float metallic, float ior) {
    Lobes l;
    float r = smax(roughness, 0.08f);  // Material.cpp:15
    l.alpha = r * r;
    l.base = base;
    l.metallic = metallic;
    float f = (ior - 1.0f) / (ior + 1.0f);  // Material::computeF0 (Material.hpp:117-121)
    f *= f;
    l.F0 = lerp(V3(f), base, metallic);
    l.specWeight = metallic + (1.f - metallic) * max3(l.F0);
    return l;
}

// DiffuseMaterial::eval and ::pdf with LOBE_ALL, fused (they share H).
PRR_DEV void evalPdf(const Lobes &m, const Frame &fr, const V3 &wi, const V3 &wo, V3 &f, float &pdf) {
    V3 wiL = toLocal(fr, wi), woL = toLocal(fr, wo);
    float NdotWi = wiL.z, NdotWo = woL.z;
    f = V3(0.f);
    pdf = 0.f;
    if (NdotWi <= 0.f || NdotWo <= 0.f) return;
    V3 H = wiL + woL;
    if (dot(H, H) < 1e-8f) return;
    H = normalize(H);
    if (H.z < 0.f) H = -H;
    float NdotH = smax(0.f, H.z);
    float WoDotH = smax(0.f, dot(woL, H));
    V3 F = F_Schlick(WoDotH, m.F0);
    float D = D_GGX(NdotH, m.alpha);
    float lo = lambdaGGX(NdotWo, m.alpha);
    float li = lambdaGGX(NdotWi, m.alpha);
    float G = 1.f / (1.f + lo + li);
    float denom = 4.f * smax(NdotWo, 1e-4f) * smax(NdotWi, 1e-4f);
    V3 spec = D * G * F / denom;
    V3 kD = (V3(1.f) - F) * (1.f - m.metallic);
    V3 diff = kD * m.base / kPi;
    f = (diff + spec) * NdotWi;
    float G1 = 1.f / (1.f + lo);
    float ps = D * G1 * smax(0.f, WoDotH) / (smax(1e-6f, NdotWo) * 4.f * smax(1e-4f, WoDotH));
    float pd = smax(0.f, NdotWi) / kPi;
    pdf = m.specWeight * ps + (1.f - m.specWeight) * pd;
}

struct BsdfSample {
    V3 wi, f;
    float pdf;
    int lobe;  // 0 diffuse, 1 specular, 2 delta, -1 none
};

// DiffuseMaterial::sample (Material.cpp:28-78); draws u0..u2 only above the
// horizon, exactly like the reference.
template <class R>
PRR_DEV BsdfSample sampleDiffuse(const Lobes &m, const Frame &fr, const V3 &wo, R &rng) {
    BsdfSample s;
    s.wi = V3(0.f);
    s.f = V3(0.f);
    s.pdf = 0.f;
    s.lobe = -1;
    V3 woL = toLocal(fr, wo);
    if (woL.z <= 0.f) return s;
    float u0 = rng.next();
    float u1 = rng.next();
    float u2 = rng.next();
    V3 wiL;
    int lobe;
    if (u0 < m.specWeight) {
        V3 H = sampleVNDF(woL, m.alpha, u1, u2);
        if (dot(H, H) < 1e-8f) return s;
        wiL = reflect(woL, H);
        lobe = 1;
    } else {
        float r = sqrtf(u1);
        float phi = 2.f * kPi * u2;
        wiL = V3(r * cosf(phi), r * sinf(phi), sqrtf(smax(0.f, 1.f - u1)));
        lobe = 0;
    }
    if (wiL.z <= 0.f) return s;
    V3 wi = toWorld(fr, wiL);
    V3 f;
    float pdf;
    evalPdf(m, fr, wi, wo, f, pdf);
    s.wi = wi;
    s.lobe = lobe;
    s.pdf = pdf;
    s.f = (pdf > 1e-6f) ? f : V3(0.f);
    return s;
}

PRR_DEV float clampf(float lo, float hi, float v) { return smax(lo, smin(hi, v)); }
// refract / fresnel (BRDFUtils.hpp:232-260)
PRR_DEV V3 refract(const V3 &I, const V3 &N, float ior) {
    float cosi = clampf(-1.f, 1.f, dot(I, N));
    float etai = 1.f, etat = ior;
    V3 n = N;
    if (cosi < 0.f) {
        cosi = -cosi;
    } else {
        float t = etai;
        etai = etat;
        etat = t;
        n = -N;
    }
    float eta = etai / etat;
    float k = 1.f - eta * eta * (1.f - cosi * cosi);
    return k < 0.f ? V3(0.f) : eta * I + (eta * cosi - sqrtf(k)) * n;
}
PRR_DEV float fresnel(const V3 &I, const V3 &N, float ior) {
    float cosi = clampf(-1.f, 1.f, dot(I, N));
    float etai = 1.f, etat = ior;
    if (cosi > 0.f) {
        float t = etai;
        etai = etat;
        etat = t;
    }
    float sint = etai / etat * sqrtf(smax(0.f, 1.f - cosi * cosi));
    if (sint >= 1.f) return 1.f;
    float cost = sqrtf(smax(0.f, 1.f - sint * sint));
    cosi = fabsf(cosi);
    float Rs = ((etat * cosi) - (etai * cost)) / ((etat * cosi) + (etai * cost));
    float Rp = ((etai * cosi) - (etat * cost)) / ((etai * cosi) + (etat * cost));
    return (Rs * Rs + Rp * Rp) * 0.5f;
}
// GlassMaterial::sample (Material.cpp:169-191): one draw, always.
PRR_DEV BsdfSample sampleGlass(const V3 &wo, const V3 &N, float ior, float u) {
    BsdfSample s;
    s.pdf = 1.0f;
    s.lobe = 2;
    V3 inc = -wo;
    float kr = fresnel(inc, N, ior);
    V3 refr = refract(inc, N, ior);
    bool tir = norm(refr) < 1e-6f;
    bool entering = dot(inc, N) < 0.f;
    float eta = entering ? (1.0f / ior) : ior;
    if (u < kr || tir) {
        s.wi = normalize(reflect(wo, N));
        s.f = V3(1.0f);
    } else {
        s.wi = normalize(refr);
        s.f = V3(eta * eta);
    }
    return s;
}
// MirrorMaterial::sample (Material.cpp:160-167)
PRR_DEV BsdfSample sampleMirror(const V3 &wo, const V3 &N, const V3 &base) {
    BsdfSample s;
    s.wi = normalize(reflect(wo, N));
    s.f = base;
    s.pdf = 1.0f;
    s.lobe = 2;
    return s;
}

// ---------------------------------------------------------------------------
// Environment map (EnvMap.hpp)
// ---------------------------------------------------------------------------
PRR_DEV float luminance(const V3 &c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; }
PRR_DEV void dirToUV(const V3 &d, float &u, float &v) {  // EnvMap.hpp:71-75
    float phi = atan2f(d.z, d.x);
    float theta = acosf(smax(-1.f, smin(1.f, d.y)));
    u = (phi + kPi) / (2.f * kPi);
    v = theta / kPi;
}
PRR_DEV V3 uvToDir(float u, float v) {  // EnvMap.hpp:84-89
    float phi = u * 2.f * kPi - kPi;
    float theta = v * kPi;
    float sinT = sinf(theta);
    return V3(sinT * cosf(phi), cosf(theta), sinT * sinf(phi));
}
PRR_DEV V3 envLookup(const SceneView &s, const V3 &dir) {  // EnvMap.hpp:97-113
    float u, v;
    dirToUV(dir, u, v);
    const int W = s.env_w, H = s.env_h;
    float px = smax(0.f, u * (W - 1));
    float py = smax(0.f, v * (H - 1));
    int x0 = (int) px, y0 = (int) py;
    int x1 = min(x0 + 1, W - 1);
    int y1 = min(y0 + 1, H - 1);
    float fx = px - x0, fy = py - y0;
    V3 a = ld3(s.env_px[y0 * W + x0]), b = ld3(s.env_px[y0 * W + x1]);
    V3 c = ld3(s.env_px[y1 * W + x0]), d = ld3(s.env_px[y1 * W + x1]);
    V3 top = a + (b - a) * fx;
    V3 bot = c + (d - c) * fx;
    return top + (bot - top) * fy;
}
PRR_DEV float envPdf(const SceneView &s, const V3 &dir) {  // EnvMap.hpp:124-133
    float u, v;
    dirToUV(dir, u, v);
    const int W = s.env_w, H = s.env_h;
    int col = min((int) (u * W), W - 1);
    int row = min((int) (v * H), H - 1);
    float theta = v * kPi;
    float sinT = smax(1e-4f, sinf(theta));
    float lum = s.env_px[row * W + col].w;
    return (lum * sinT / (s.env_total + 1e-6f)) * ((float) (W * H) / (2.f * kPi * kPi * sinT));
}
// first index i in [0,n) with a[i] >= x, or n (std::lower_bound)
PRR_DEV int lowerBound(const float *a, int n, float x) {
    int lo = 0, cnt = n;
    while (cnt > 0) {
        int step = cnt >> 1;
        int mid = lo + step;
        if (a[mid] < x) {
            lo = mid + 1;
            cnt -= step + 1;
        } else {
            cnt = step;
        }
    }
    return lo;
}
// lower_bound(a, n, x) for x in [0, 1) restricted to the guide bucket of x
// (same index: see buildEnvGuides on the host).
PRR_DEV int guidedLowerBound(const float *a, float x, const int *g, int K) {
    const int k = min(max((int) (x * (float) K), 0), K - 1);
    int lo = g[k], cnt = g[k + 1] - lo;
    while (cnt > 0) {
        int step = cnt >> 1;
        int mid = lo + step;
        if (a[mid] < x) {
            lo = mid + 1;
            cnt -= step + 1;
        } else {
            cnt = step;
        }
    }
    return lo;
}
PRR_DEV V3 envImportanceSample(const SceneView &s, float r1, float r2, float &pdf) {  // EnvMap.hpp:145-169
    const int W = s.env_w, H = s.env_h;
    int row, col;
    if (s.env_gk > 0 && r1 < 1.f && r2 < 1.f && r1 >= 0.f && r2 >= 0.f) {
        row = min(guidedLowerBound(s.env_marg, r1, s.env_gm, s.env_gk), H - 1);
        col = min(guidedLowerBound(s.env_cdf + (size_t) row * W, r2, s.env_gc + (size_t) row * (s.env_gk + 1),
                                   s.env_gk),
                  W - 1);
    } else {
        row = min(lowerBound(s.env_marg, H, r1), H - 1);
        col = min(lowerBound(s.env_cdf + (size_t) row * W, W, r2), W - 1);
    }
    float u = (col + 0.5f) / (float) W;
    float v = (row + 0.5f) / (float) H;
    V3 dir = uvToDir(u, v);
    float lum = s.env_px[row * W + col].w;
    float p = (lum / (s.env_total + 1e-6f)) * ((float) (W * H) / (2.f * kPi * kPi));
    pdf = smax(1e-6f, p);
    return dir;
}

// ---------------------------------------------------------------------------
// Camera::generateRay (Camera.hpp:58-95)
// ---------------------------------------------------------------------------
PRR_DEV V3 sampleDisc(float u, float v) {
    float a = 2.0f * u - 1.0f, b = 2.0f * v - 1.0f;
    if (a == 0.0f && b == 0.0f) return V3(0.f);
    float r, phi;
    if (fabsf(a) > fabsf(b)) {
        r = a;
        phi = (kPi / 4.0f) * (b / a);
    } else {
        r = b;
        phi = (kPi / 2.0f) - (kPi / 4.0f) * (a / b);
    }
    return V3(r * cosf(phi), r * sinf(phi), 0.0f);
}
template <class R>
PRR_DEV void cameraRay(const CamView &c, int i, int j, int W, int H, float ox, float oy, R &rng, V3 &org,
                       V3 &dir) {
    float fpx = ((float) W * 0.5f) / c.tanhalf;
    float x = ((float) i + ox - (float) W * 0.5f) / fpx;
    float y = ((float) (H - j) + oy - (float) H * 0.5f) / fpx;
    V3 fwd(c.fwd[0], c.fwd[1], c.fwd[2]);
    V3 right(c.right[0], c.right[1], c.right[2]);
    V3 up(c.up[0], c.up[1], c.up[2]);
    V3 eye(c.eye[0], c.eye[1], c.eye[2]);
    V3 d = normalize(fwd + x * right + y * up);
    if (c.aperture <= 0.0f) {
        org = eye;
        dir = d;
        return;
    }
    V3 focus = eye + d * (c.focus / dot(d, fwd));
    float u = rng.next();
    float v = rng.next();
    V3 lo = sampleDisc(u, v);
    V3 lens = eye + c.aperture * (lo.x * right + lo.y * up);
    org = lens;
    dir = normalize(focus - lens);
}

// ---------------------------------------------------------------------------
// Tone mapping (ToneMapping.hpp) + ACES (Narkowicz 2015 fit)
// ---------------------------------------------------------------------------
PRR_DEV float clamp01f(float x) { return smin(smax(x, 0.f), 1.f); }  // std::clamp(x, 0, 1)
PRR_DEV float srgbOETF(float x) {
    x = clamp01f(x);
    return x < 0.0031308f ? 12.92f * x : 1.055f * powf(x, 1.f / 2.4f) - 0.055f;
}
PRR_DEV V3 srgb3(const V3 &c) { return V3(srgbOETF(c.x), srgbOETF(c.y), srgbOETF(c.z)); }
struct AgxConst {
    float toe_a, slope, w, shoulder_max;
};
PRR_DEV float allenwp(float x, const AgxConst &p) {  // ToneMapping.hpp:151-166
    const float crossover = 0.1841865f;
    x = smax(x, 0.f);
    if (x < crossover) {
        float t = powf(x, 1.25f);
        return t / (t + p.toe_a);
    }
    float s = x - crossover;
    float slope_s = p.slope * s;
    s = slope_s * (1.f + s / p.w) / (1.f + (slope_s / p.shoulder_max));
    s += crossover;
    return s;
}
PRR_DEV V3 mat3mul(const float m[9], const V3 &v) {
    return V3(m[0] * v.x + m[1] * v.y + m[2] * v.z, m[3] * v.x + m[4] * v.y + m[5] * v.z,
              m[6] * v.x + m[7] * v.y + m[8] * v.z);
}
PRR_DEV V3 tonemapAgx(V3 c, const AgxConst &p) {  // ToneMapping.hpp:172-214
    const float inset[9] = {0.544814746488245f, 0.373787398372697f, 0.0813978551390581f,
                            0.140416948464053f, 0.754137554567394f, 0.105445496968552f,
                            0.0888104196149096f, 0.178871756420858f, 0.732317823964232f};
    const float outset[9] = {1.96488741169489f, -0.855988495690215f, -0.108898916004672f,
                             -0.299313364904742f, 1.32639796461980f, -0.0270845997150571f,
                             -0.164352742528393f, -0.238183969428088f, 1.40253671195648f};
    c = V3(smax(c.x, 0.f), smax(c.y, 0.f), smax(c.z, 0.f));
    c = mat3mul(inset, c);
    c = V3(allenwp(c.x, p), allenwp(c.y, p), allenwp(c.z, p));
    c = V3(smin(1.0f, c.x), smin(1.0f, c.y), smin(1.0f, c.z));
    c = mat3mul(outset, c);
    c = V3(clamp01f(c.x), clamp01f(c.y), clamp01f(c.z));
    return srgb3(c);
}
PRR_DEV V3 tonemapPbrNeutral(V3 c) {  // ToneMapping.hpp:234-262
    const float startCompression = 0.8f - 0.04f;
    const float desaturation = 0.15f;
    float x = smin(c.x, smin(c.y, c.z));
    float offset = x < 0.08f ? x - 6.25f * x * x : 0.04f;
    c.x -= offset;
    c.y -= offset;
    c.z -= offset;
    float peak = smax(c.x, smax(c.y, c.z));
    if (peak < startCompression) return srgb3(c);
    float d = 1.f - startCompression;
    float newPeak = 1.f - d * d / (peak + d - startCompression);
    float scale = newPeak / peak;
    c.x *= scale;
    c.y *= scale;
    c.z *= scale;
    float g = 1.f - 1.f / (desaturation * (peak - newPeak) + 1.f);
    c.x = c.x * (1.f - g) + newPeak * g;
    c.y = c.y * (1.f - g) + newPeak * g;
    c.z = c.z * (1.f - g) + newPeak * g;
    return srgb3(c);
}
PRR_DEV float acesCurve(float x) {  // Narkowicz 2015
    x = smax(x, 0.f);
    float n = x * (2.51f * x + 0.03f);
    float dd = x * (2.43f * x + 0.59f) + 0.14f;
    return clamp01f(n / dd);
}
PRR_DEV V3 tonemapApply(int op, const V3 &c, const AgxConst &p) {
    switch (op) {
    case 0: return V3(clamp01f(c.x), clamp01f(c.y), clamp01f(c.z));
    case 1: return srgb3(c);
    case 2: return tonemapAgx(c, p);
    case 3: return tonemapPbrNeutral(c);
    default: return srgb3(V3(acesCurve(c.x), acesCurve(c.y), acesCurve(c.z)));
    }
}
PRR_DEV V3 exposeLinear(const V3 &fb, float exposure) {  // Renderer.cpp:149-161
    V3 c = fb * exposure;
    if (!isfinite(c.x)) c.x = 0.f;
    if (!isfinite(c.y)) c.y = 0.f;
    if (!isfinite(c.z)) c.z = 0.f;
    return V3(smax(0.f, c.x), smax(0.f, c.y), smax(0.f, c.z));
}

// Scene::sampleLight (Scene.cpp:37-58) -> MeshTriangle / BVHAccel::Sample
// (BVH.cpp:478-514) -> Triangle::Sample / Sphere::Sample.
// R: any generator with next() (PCG32 in the renderer, explicit uniforms in
// prr_sample_light). *emitter: the row of Scene::emissives picked.
template <class R>
PRR_DEV void sampleLight(const SceneView &s, R &rng, V3 &pos, V3 &nrm, int &mat, int *emitter = nullptr) {
    const float p = rng.next() * s.total_emit_area;
    float acc = 0.f;
    int chosen = s.num_emitters - 1;
    for (int k = 0; k < s.num_emitters; k++) {
        acc += s.emitters[k].area;
        if (p <= acc) {
            chosen = k;
            break;
        }
    }
    const DEmitter e = s.emitters[chosen];
    mat = e.material;
    if (emitter) *emitter = chosen;
    if (e.kind == 0) {
        float q = sqrtf(rng.next()) * e.total;
        int ni = e.lt_root;
        int slot = -1;
        while (slot < 0) {
            const DLightNode n = s.lnodes[ni];
            if (n.count > 0) {
                for (int i = 0; i < n.count; i++) {
                    q -= s.ltri[4 * (n.first + i)].w;
                    if (q <= 0.f || i == n.count - 1) {
                        slot = n.first + i;
                        break;
                    }
                }
            } else if (q < n.left_area) {
                ni = ni + 1;
            } else {
                q -= n.left_area;
                ni = n.first;
            }
        }
        const float x = sqrtf(rng.next()), y = rng.next();
        const V3 v0 = ld3(s.ltri[4 * slot]), v1 = ld3(s.ltri[4 * slot + 1]), v2 = ld3(s.ltri[4 * slot + 2]);
        pos = v0 * (1.0f - x) + v1 * (x * (1.0f - y)) + v2 * (x * y);
        nrm = ld3(s.ltri[4 * slot + 3]);
    } else {
        const float4 sc = s.sph_by_id[e.lt_root];
        const float theta = (float) (2.0 * (double) kPi * (double) rng.next());
        const float phi = kPi * rng.next();
        const V3 dir(cosf(phi), sinf(phi) * cosf(theta), sinf(phi) * sinf(theta));
        pos = V3(sc.x, sc.y, sc.z) + sc.w * dir;
        nrm = dir;
    }
}

// ---------------------------------------------------------------------------
// Caustic photon map gather (build extension of Photon.cpp:82-94): the
// photon density estimate weighted by the BSDF,
//   L = sum_p f(wi_p, wo) * Phi_p / (pi r^2),  f = eval / cos(wi_p, N),
// over photons within r. The cells overlapping [x - r, x + r] (at most 2 per
// axis, cell = 2r) are hashed to buckets; duplicate buckets are visited once,
// so every photon within r is counted exactly once even under collisions.
// ---------------------------------------------------------------------------
PRR_HD uint32_t photonHash(int cx, int cy, int cz, uint32_t mask) {
    return (((uint32_t) cx * 73856093u) ^ ((uint32_t) cy * 19349663u) ^ ((uint32_t) cz * 83492791u)) & mask;
}
PRR_DEV V3 photonRadiance(const SceneView &s, const V3 &x, const V3 &wo, const Frame &fr, const Lobes &lb) {
    V3 sum(0.f);
    if (s.ph_count == 0) return sum;
    const float r = s.ph_radius, cs = s.ph_cell;
    const int x0 = (int) floorf((x.x - r) / cs), x1 = (int) floorf((x.x + r) / cs);
    const int y0 = (int) floorf((x.y - r) / cs), y1 = (int) floorf((x.y + r) / cs);
    const int z0 = (int) floorf((x.z - r) / cs), z1 = (int) floorf((x.z + r) / cs);
    // cells (cx, cy, cz) in [x0, min(x1, x0 + 1)] x ... in z, y, x order; a
    // cell whose hash bucket an earlier cell already visited is skipped (its
    // hash is recomputed rather than kept in a per-lane array, which would
    // live in scratch memory)
    for (int k = 0; k < 8; k++) {
        const int cx = x0 + (k & 1), cy = y0 + ((k >> 1) & 1), cz = z0 + (k >> 2);
        if (cx > x1 || cy > y1 || cz > z1) continue;
        const uint32_t h = photonHash(cx, cy, cz, s.ph_mask);
        bool dup = false;
        for (int j = 0; j < k; j++) {
            const int jx = x0 + (j & 1), jy = y0 + ((j >> 1) & 1), jz = z0 + (j >> 2);
            dup |= jx <= x1 && jy <= y1 && jz <= z1 && photonHash(jx, jy, jz, s.ph_mask) == h;
        }
        if (dup) continue;
        const uint32_t b = s.ph_start[h], e = s.ph_end[h];
        // positions four at a time (most photons of a cell lie outside r, so
        // the scan is a chain of position loads: four are in flight at once);
        // the photons within r are added in index order
        for (uint32_t i0 = b; i0 < e; i0 += 4) {
            V3 p[4];
#pragma unroll
            for (int u = 0; u < 4; u++) p[u] = i0 + u < e ? ld3(s.ph_pos[i0 + u]) : x + V3(4.f * r);
#pragma unroll
            for (int u = 0; u < 4; u++) {
                if (norm(p[u] - x) > r) continue;
                const uint32_t i = i0 + u;
                const V3 wi = -ld3(s.ph_dir[i]);
                const float c = dot(wi, fr.N);
                if (c <= 1e-4f) continue;
                V3 f;
                float pdf;
                evalPdf(lb, fr, wi, wo, f, pdf);
                sum += f * ld3(s.ph_pow[i]) / c;
            }
        }
    }
    return sum / (kPi * r * r);
}

// PhotonMap::estimateIrradiance (Photon.cpp:82-94): the power of the
// photons within r of x over pi r^2 (the normal is unused, as there). The
// same cells and buckets as photonRadiance.
PRR_DEV V3 photonIrradiance(const SceneView &s, const V3 &x) {
    V3 sum(0.f);
    if (s.ph_count == 0) return sum;
    const float r = s.ph_radius, cs = s.ph_cell;
    const int x0 = (int) floorf((x.x - r) / cs), x1 = (int) floorf((x.x + r) / cs);
    const int y0 = (int) floorf((x.y - r) / cs), y1 = (int) floorf((x.y + r) / cs);
    const int z0 = (int) floorf((x.z - r) / cs), z1 = (int) floorf((x.z + r) / cs);
    for (int k = 0; k < 8; k++) {
        const int cx = x0 + (k & 1), cy = y0 + ((k >> 1) & 1), cz = z0 + (k >> 2);
        if (cx > x1 || cy > y1 || cz > z1) continue;
        const uint32_t h = photonHash(cx, cy, cz, s.ph_mask);
        bool dup = false;
        for (int j = 0; j < k; j++) {
            const int jx = x0 + (j & 1), jy = y0 + ((j >> 1) & 1), jz = z0 + (j >> 2);
            dup |= jx <= x1 && jy <= y1 && jz <= z1 && photonHash(jx, jy, jz, s.ph_mask) == h;
        }
        if (dup) continue;
        const uint32_t b = s.ph_start[h], e = s.ph_end[h];
        for (uint32_t i = b; i < e; i++)
            if (norm(ld3(s.ph_pos[i]) - x) <= r) sum += ld3(s.ph_pow[i]);
    }
    return sum / (kPi * r * r);
}

}  // namespace prr
