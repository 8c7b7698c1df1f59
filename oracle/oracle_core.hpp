// oracle_core.hpp — TEST INFRASTRUCTURE ONLY.
//
// CPU restatement of the reference renderer's hot path, used as the checker
// for the HIP library (tests/, __graft_entry__.smoke(), bench.py's CPU
// baseline leg). Nothing in the product links or calls this code.
//
// Every function keeps the reference's operation order so that an IEEE-strict
// build (-O2 -ffp-contract=off, no -ffast-math) reproduces the reference's
// floats; each cites the reference file:line it restates (paths relative to
// the reference repository root).
//
// Parity status per function (what pins it) is listed in the header comment
// of each section; the tests in tests/test_oracle_*.py carry the checks.
// Parity unpinned: the choice of random streams. The device draws from PCG32
// per (pixel, sample) where the reference draws from a SplitMix64 it seeds
// from std::random_device; PCG32 itself is pinned by its published known
// answers and castRay on SplitMix streams by the reference's own per-path
// radiance, while castRay on PCG streams agrees with the reference only as an
// estimator (the image tests).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

namespace orc {

// global.hpp:9-12
constexpr float kPi = 3.141592653589793f;
constexpr float kEps = 1e-3f;
constexpr float kInf = std::numeric_limits<float>::infinity();

// ---------------------------------------------------------------------------
// Vector math (Vector.hpp:7-123). Operator semantics and evaluation order as
// in the reference class: componentwise, left to right.
// ---------------------------------------------------------------------------
struct V3 {
    float x, y, z;
    V3() : x(0), y(0), z(0) {}
    explicit V3(float s) : x(s), y(s), z(s) {}
    V3(float a, float b, float c) : x(a), y(b), z(c) {}
    float operator[](int i) const { return (&x)[i]; }
    float &operator[](int i) { return (&x)[i]; }
};
inline V3 operator+(const V3 &a, const V3 &b) { return V3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V3 operator-(const V3 &a, const V3 &b) { return V3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V3 operator*(const V3 &a, const V3 &b) { return V3(a.x * b.x, a.y * b.y, a.z * b.z); }
inline V3 operator/(const V3 &a, const V3 &b) { return V3(a.x / b.x, a.y / b.y, a.z / b.z); }
inline V3 operator*(const V3 &a, float r) { return V3(a.x * r, a.y * r, a.z * r); }
inline V3 operator*(float r, const V3 &a) { return V3(a.x * r, a.y * r, a.z * r); }
inline V3 operator/(const V3 &a, float r) { return V3(a.x / r, a.y / r, a.z / r); }
inline V3 operator-(const V3 &a) { return V3(-a.x, -a.y, -a.z); }
inline V3 &operator+=(V3 &a, const V3 &b) { a.x += b.x; a.y += b.y; a.z += b.z; return a; }
inline float dot(const V3 &a, const V3 &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(const V3 &a, const V3 &b) {
    return V3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
// Vector.hpp:103-111 (1/sqrtf then multiply)
inline V3 normalize(const V3 &v) {
    float mag2 = v.x * v.x + v.y * v.y + v.z * v.z;
    if (mag2 > 0) {
        float inv = 1 / sqrtf(mag2);
        return V3(v.x * inv, v.y * inv, v.z * inv);
    }
    return v;
}
// Vector.hpp:25-31
inline float norm(const V3 &v) { return std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z); }
inline V3 normalized(const V3 &v) {
    float n = std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
    return V3(v.x / n, v.y / n, v.z / n);
}
inline V3 vmin(const V3 &a, const V3 &b) {
    return V3(std::min(a.x, b.x), std::min(a.y, b.y), std::min(a.z, b.z));
}
inline V3 vmax(const V3 &a, const V3 &b) {
    return V3(std::max(a.x, b.x), std::max(a.y, b.y), std::max(a.z, b.z));
}
inline V3 lerp(const V3 &a, const V3 &b, float t) { return a * (1 - t) + b * t; }
inline float maxc(const V3 &v) { return std::max({v.x, v.y, v.z}); }
inline bool finite3(const V3 &v) {
    return std::isfinite(v.x) && std::isfinite(v.y) && std::isfinite(v.z);
}

struct Ray {
    V3 o, d;
};

// ---------------------------------------------------------------------------
// Bounds3 (Bounds3.hpp:7-126)
// ---------------------------------------------------------------------------
struct Box {
    V3 mn, mx;
    // default: pMin = +max(double)->+inf, pMax = lowest(double)->-inf (Bounds3.hpp:11-16)
    Box() : mn(kInf), mx(-kInf) {}
    Box(const V3 &a, const V3 &b)
        : mn(fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)),
          mx(fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)) {}
    V3 diag() const { return mx - mn; }
    int maxExtent() const {
        V3 d = diag();
        if (d.x > d.y && d.x > d.z) return 0;
        if (d.y > d.z) return 1;
        return 2;
    }
    // SurfaceArea returns double of a float expression (Bounds3.hpp:37-41)
    float area() const {
        V3 d = diag();
        return (float) (double) (2 * (d.x * d.y + d.x * d.z + d.y * d.z));
    }
    V3 centroid() const { return 0.5f * mn + 0.5f * mx; }
};
inline Box unite(const Box &a, const Box &b) {
    Box r;
    r.mn = vmin(a.mn, b.mn);
    r.mx = vmax(a.mx, b.mx);
    return r;
}
inline Box unite(const Box &a, const V3 &p) {
    Box r;
    r.mn = vmin(a.mn, p);
    r.mx = vmax(a.mx, p);
    return r;
}

// ---------------------------------------------------------------------------
// RNGs. SplitMix64 is the reference's generator (global.hpp:35-50); PCG32
// (O'Neill 2014, pcg32_random_r / pcg32_srandom_r) is the per-pixel generator
// of the GPU renderer, restated here so that the oracle can replay the exact
// sample stream of a GPU path.
// ---------------------------------------------------------------------------
struct SplitMix64 {
    uint64_t state;
    float next() {
        uint64_t z = (state += 0x9E3779B97F4A7C15ULL);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        z ^= (z >> 31);
        return (z >> 40) * 0x1p-24f;
    }
};

inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

struct Pcg32 {
    uint64_t state, inc;
    uint32_t next_u32() {
        uint64_t old = state;
        state = old * 6364136223846793005ULL + inc;
        uint32_t xs = (uint32_t) (((old >> 18u) ^ old) >> 27u);
        uint32_t rot = (uint32_t) (old >> 59u);
        return (xs >> rot) | (xs << ((32u - rot) & 31u));
    }
    float next() { return (next_u32() >> 8) * 0x1p-24f; }
    // per-pixel seeding: initstate from (seed, pixel), stream from sample index
    static Pcg32 make(uint64_t seed, uint32_t pixel, uint32_t sample) {
        Pcg32 r;
        r.state = 0;
        r.inc = ((uint64_t) sample << 1) | 1u;
        r.next_u32();
        r.state += mix64(mix64(seed * 0x9E3779B97F4A7C15ULL + pixel) ^ ((uint64_t) sample * 0xD1B54A32D192ED03ULL));
        r.next_u32();
        return r;
    }
};

// Explicit-uniform generator (for the per-call BSDF / env checks).
struct FixedU {
    const float *u;
    int i = 0;
    float next() { return u[i++]; }
};

// ---------------------------------------------------------------------------
// Shading helpers (BRDFUtils.hpp)
// ---------------------------------------------------------------------------
// BRDFUtils.hpp:15-30
inline void buildTBN(const V3 &N, V3 &T, V3 &B) {
    if (std::fabs(N.x) > std::fabs(N.y)) {
        float invLen = 1.0f / std::sqrt(N.x * N.x + N.z * N.z);
        T = V3(N.z * invLen, 0.0f, -N.x * invLen);
    } else {
        float invLen = 1.0f / std::sqrt(N.y * N.y + N.z * N.z);
        T = V3(0.0f, N.z * invLen, -N.y * invLen);
    }
    B = cross(N, T);
}
// ---------------------------------------------------------------------------
// Participating media (Volume.hpp:7-66, Volume.cpp:3-24); the reference's
// integrator does not call them.
// ---------------------------------------------------------------------------
// Medium::intersect (Volume.hpp:17-44): std::max / std::min as written (the
// first argument wins a comparison with NaN), dirIsNeg = direction > 0. The
// bounds and the origin are read through Vector3f's const operator[], which
// returns double (Vector.hpp:67,78-80), so each plane time is formed in double
// and rounded to float once.
inline bool mediumIntersect(bool bounded, const float *lo, const float *hi, const V3 &o, const V3 &d, float &tEnter,
                            float &tExit) {
    if (!bounded) {
        tEnter = 0.0f;
        tExit = std::numeric_limits<float>::infinity();
        return true;
    }
    const V3 inv(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    float te = -std::numeric_limits<float>::max(), tx = std::numeric_limits<float>::max();
    for (int i = 0; i < 3; i++) {
        float tmin = (float) (((double) lo[i] - (double) o[i]) * (double) inv[i]);
        float tmax = (float) (((double) hi[i] - (double) o[i]) * (double) inv[i]);
        if (!(d[i] > 0)) std::swap(tmin, tmax);
        te = std::max(tmin, te);
        tx = std::min(tmax, tx);
    }
    if (te > tx || tx < 0.0f) return false;
    tEnter = te;
    tExit = tx;
    return true;
}
// phaseHG (Volume.cpp:3-6), in float (M_PI is a float literal, global.hpp:9-10).
inline float phaseHG(float cosTheta, float g) {
    const float denom = 1.0f + g * g - 2.0f * g * cosTheta;
    return (1.0f - g * g) / (4.0f * kPi * std::pow(denom, 1.5f));
}
// samplePhaseHG (Volume.cpp:8-24) with its two draws u1 (cos theta), u2 (phi).
inline V3 samplePhaseHG(const V3 &wo, float g, float u1, float u2) {
    float cosTheta;
    if (std::abs(g) < 1e-3f) {
        cosTheta = 1.0f - 2.0f * u1;
    } else {
        const float sqr = (1.0f - g * g) / (1.0f - g + 2.0f * g * u1);
        cosTheta = (1.0f + g * g - sqr * sqr) / (2.0f * g);
    }
    const float sinTheta = std::sqrt(std::max(0.0f, 1.0f - cosTheta * cosTheta));
    const float phi = 2.0f * kPi * u2;
    V3 T, B;
    buildTBN(wo, T, B);
    return normalize(sinTheta * std::cos(phi) * T + sinTheta * std::sin(phi) * B + cosTheta * wo);
}

// BRDFUtils.hpp:35-49
inline V3 toWorld(const V3 &l, const V3 &N) {
    V3 T, B;
    buildTBN(N, T, B);
    return l.x * T + l.y * B + l.z * N;
}
inline V3 toLocal(const V3 &w, const V3 &N) {
    V3 T, B;
    buildTBN(N, T, B);
    return V3(dot(w, T), dot(w, B), dot(w, N));
}
// BRDFUtils.hpp:56-58
inline V3 reflect(const V3 &wo, const V3 &N) { return 2.f * dot(wo, N) * N - wo; }
// BRDFUtils.hpp:79-81
inline float pdfCosine(float NdotWi) { return std::max(0.f, NdotWi) / kPi; }
// BRDFUtils.hpp:90-94
inline float D_GGX(float NdotH, float alpha) {
    float a2 = alpha * alpha;
    float d = NdotH * NdotH * (a2 - 1.f) + 1.f;
    return a2 / (kPi * d * d);
}
// BRDFUtils.hpp:102-106
inline float lambdaGGX(float NdotV, float alpha) {
    float c2 = NdotV * NdotV;
    float t2 = (1.f - c2) / std::max(c2, 1e-6f);
    return 0.5f * (-1.f + std::sqrt(1.f + alpha * alpha * t2));
}
// BRDFUtils.hpp:112-114, 121-123
inline float G1_GGX(float NdotV, float alpha) { return 1.f / (1.f + lambdaGGX(NdotV, alpha)); }
inline float G2_GGX(float NdotWo, float NdotWi, float alpha) {
    return 1.f / (1.f + lambdaGGX(NdotWo, alpha) + lambdaGGX(NdotWi, alpha));
}
// BRDFUtils.hpp:143-146
inline V3 F_Schlick(float WoDotH, const V3 &F0) {
    float t = std::pow(std::max(0.f, 1.f - WoDotH), 5.f);
    return F0 + (V3(1.f) - F0) * t;
}
// BRDFUtils.hpp:163-181 (Heitz 2018 VNDF)
inline V3 sampleVNDF(const V3 &wo, float alpha, float u1, float u2) {
    V3 Vh = normalize(V3(alpha * wo.x, alpha * wo.y, wo.z));
    float lensq = Vh.x * Vh.x + Vh.y * Vh.y;
    V3 T1 = lensq > 0.f ? V3(-Vh.y, Vh.x, 0.f) / std::sqrt(lensq) : V3(1.f, 0.f, 0.f);
    V3 T2 = cross(Vh, T1);
    float r = std::sqrt(u1);
    float phi = 2.f * kPi * u2;
    float t1 = r * std::cos(phi);
    float t2 = r * std::sin(phi);
    float s = 0.5f * (1.f + Vh.z);
    t2 = (1.f - s) * std::sqrt(1.f - t1 * t1) + s * t2;
    V3 Nh = t1 * T1 + t2 * T2 + std::sqrt(std::max(0.f, 1.f - t1 * t1 - t2 * t2)) * Vh;
    return normalize(V3(alpha * Nh.x, alpha * Nh.y, std::max(0.f, Nh.z)));
}
// BRDFUtils.hpp:195-204
inline float pdfVNDF(float NdotWo, float NdotH, float WoDotH, float alpha) {
    float D = D_GGX(NdotH, alpha);
    float G1 = G1_GGX(NdotWo, alpha);
    return D * G1 * std::max(0.f, WoDotH) /
           (std::max(1e-6f, NdotWo) * 4.f * std::max(1e-4f, WoDotH));
}
inline float clampf(float lo, float hi, float v) { return std::max(lo, std::min(hi, v)); }
// BRDFUtils.hpp:232-245
inline V3 refract(const V3 &I, const V3 &N, float ior) {
    float cosi = clampf(-1.f, 1.f, dot(I, N));
    float etai = 1.f, etat = ior;
    V3 n = N;
    if (cosi < 0.f) {
        cosi = -cosi;
    } else {
        std::swap(etai, etat);
        n = -N;
    }
    float eta = etai / etat;
    float k = 1.f - eta * eta * (1.f - cosi * cosi);
    return k < 0.f ? V3(0) : eta * I + (eta * cosi - sqrtf(k)) * n;
}
// BRDFUtils.hpp:247-260
inline float fresnel(const V3 &I, const V3 &N, float ior) {
    float cosi = clampf(-1.f, 1.f, dot(I, N));
    float etai = 1.f, etat = ior;
    if (cosi > 0.f) std::swap(etai, etat);
    float sint = etai / etat * sqrtf(std::max(0.f, 1.f - cosi * cosi));
    if (sint >= 1.f) return 1.f;
    float cost = sqrtf(std::max(0.f, 1.f - sint * sint));
    cosi = std::abs(cosi);
    float Rs = ((etat * cosi) - (etai * cost)) / ((etat * cosi) + (etai * cost));
    float Rp = ((etai * cosi) - (etat * cost)) / ((etai * cosi) + (etat * cost));
    return (Rs * Rs + Rp * Rp) * 0.5f;
}

// ---------------------------------------------------------------------------
// Materials (Material.hpp / Material.cpp)
// ---------------------------------------------------------------------------
enum MatType { MAT_DIFFUSE = 0, MAT_GLASS = 1, MAT_MIRROR = 2, MAT_EMISSIVE = 3 };
enum Lobe { LOBE_DIFFUSE = 0, LOBE_SPECULAR = 1, LOBE_DELTA = 2, LOBE_NONE = -1 };

struct Material {
    int type = MAT_DIFFUSE;
    V3 base = V3(0.8f);
    float roughness = 0.5f, metallic = 0.f, ior = 1.5f;
    V3 emission = V3(0.f);
    bool isDelta() const { return type == MAT_GLASS || type == MAT_MIRROR; }  // Material.hpp:106-108
    bool isEmissive() const { return type == MAT_EMISSIVE; }                   // :136-138
    bool hasEmission() const { return norm(emission) > kEps; }                // :113-115
};

// ShadingData built from N alone (Material.hpp:80-84 -> Material.cpp:3-26),
// untextured: baseColor/roughness/metallic are the constants.
struct Shading {
    V3 N;
    V3 base;
    float roughness, metallic, ior;
};
inline Shading makeShading(const Material &m, const V3 &N) {
    Shading s;
    s.N = N;
    s.base = m.base;
    s.roughness = std::max(m.roughness, 0.08f);  // Material.cpp:15
    s.metallic = m.metallic;
    s.ior = m.ior;
    return s;
}
// Material::computeF0 (Material.hpp:117-121)
inline V3 computeF0(float ior, float metallic, const V3 &color) {
    float f = (ior - 1.0f) / (ior + 1.0f);
    f *= f;
    return lerp(V3(f), color, metallic);
}

struct BSDFSample {
    V3 wi, f;
    float pdf = 0.f;
    int lobe = LOBE_NONE;
};

// DiffuseMaterial::eval, LOBE_ALL (Material.cpp:80-123)
inline V3 diffuseEval(const V3 &wi, const V3 &wo, const Shading &sd) {
    float alpha = sd.roughness * sd.roughness;
    V3 F0 = computeF0(sd.ior, sd.metallic, sd.base);
    V3 wiL = toLocal(wi, sd.N);
    V3 woL = toLocal(wo, sd.N);
    float NdotWi = wiL.z, NdotWo = woL.z;
    if (NdotWi <= 0.f || NdotWo <= 0.f) return V3(0.f);
    V3 H = wiL + woL;
    if (dot(H, H) < 1e-8f) return V3(0.f);
    H = normalize(H);
    if (H.z < 0.f) H = -H;
    float NdotH = std::max(0.f, H.z);
    float WoDotH = std::max(0.f, dot(woL, H));
    V3 F = F_Schlick(WoDotH, F0);
    float D = D_GGX(NdotH, alpha);
    float G = G2_GGX(NdotWo, NdotWi, alpha);
    float denom = 4.f * std::max(NdotWo, 1e-4f) * std::max(NdotWi, 1e-4f);
    V3 spec = D * G * F / denom;
    V3 kD = (V3(1.f) - F) * (1.f - sd.metallic);
    V3 diff = kD * sd.base / kPi;
    return (diff + spec) * NdotWi;
}

// DiffuseMaterial::pdf, LOBE_ALL (Material.cpp:125-158)
inline float diffusePdf(const V3 &wi, const V3 &wo, const Shading &sd) {
    float alpha = sd.roughness * sd.roughness;
    V3 F0 = computeF0(sd.ior, sd.metallic, sd.base);
    float F0max = std::max({F0.x, F0.y, F0.z});
    float specWeight = sd.metallic + (1.f - sd.metallic) * F0max;
    V3 wiL = toLocal(wi, sd.N);
    V3 woL = toLocal(wo, sd.N);
    float NdotWi = wiL.z, NdotWo = woL.z;
    if (NdotWi <= 0.f || NdotWo <= 0.f) return 0.f;
    V3 H = wiL + woL;
    if (dot(H, H) < 1e-8f) return 0.f;
    H = normalize(H);
    if (H.z < 0.f) H = -H;
    float NdotH = std::max(0.f, H.z);
    float WoDotH = std::max(0.f, dot(woL, H));
    float ps = pdfVNDF(NdotWo, NdotH, WoDotH, alpha);
    float pd = pdfCosine(NdotWi);
    return specWeight * ps + (1.f - specWeight) * pd;
}

// DiffuseMaterial::sample (Material.cpp:28-78). Draws u0,u1,u2 only when
// wo is above the surface, exactly like the reference.
template <class R>
inline BSDFSample diffuseSample(const V3 &wo, const Shading &sd, R &rng) {
    float alpha = sd.roughness * sd.roughness;
    V3 F0 = computeF0(sd.ior, sd.metallic, sd.base);
    float F0max = std::max({F0.x, F0.y, F0.z});
    float specWeight = sd.metallic + (1.f - sd.metallic) * F0max;
    V3 woL = toLocal(wo, sd.N);
    if (woL.z <= 0.f) return BSDFSample();
    float u0 = rng.next();
    float u1 = rng.next();
    float u2 = rng.next();
    V3 wiL;
    int lobe;
    if (u0 < specWeight) {
        V3 H = sampleVNDF(woL, alpha, u1, u2);
        if (dot(H, H) < 1e-8f) return BSDFSample();
        wiL = reflect(woL, H);
        lobe = LOBE_SPECULAR;
    } else {
        float r = std::sqrt(u1);
        float phi = 2.f * kPi * u2;
        wiL = V3(r * std::cos(phi), r * std::sin(phi), std::sqrt(std::max(0.f, 1.f - u1)));
        lobe = LOBE_DIFFUSE;
    }
    if (wiL.z <= 0.f) return BSDFSample();
    V3 wi = toWorld(wiL, sd.N);
    BSDFSample s;
    s.wi = wi;
    s.lobe = lobe;
    s.pdf = diffusePdf(wi, wo, sd);
    s.f = (s.pdf > 1e-6f) ? diffuseEval(wi, wo, sd) : V3(0.f);
    return s;
}

// GlassMaterial::sample (Material.cpp:169-191): one draw, always.
template <class R>
inline BSDFSample glassSample(const V3 &wo, const V3 &N, float ior, R &rng) {
    BSDFSample res;
    res.pdf = 1.0f;
    res.lobe = LOBE_DELTA;
    V3 incident = -wo;
    float kr = fresnel(incident, N, ior);
    V3 refr = refract(incident, N, ior);
    bool tir = norm(refr) < 1e-6f;
    bool entering = dot(incident, N) < 0.f;
    float eta = entering ? (1.0f / ior) : ior;
    if (rng.next() < kr || tir) {
        res.wi = normalize(reflect(wo, N));
        res.f = V3(1.0f);
    } else {
        res.wi = normalize(refr);
        res.f = V3(eta * eta);
    }
    return res;
}

// MirrorMaterial::sample (Material.cpp:160-167)
inline BSDFSample mirrorSample(const V3 &wo, const V3 &N, const V3 &base) {
    BSDFSample res;
    res.wi = normalize(reflect(wo, N));
    res.f = base;
    res.pdf = 1.0f;
    res.lobe = LOBE_DELTA;
    return res;
}

// ---------------------------------------------------------------------------
// Environment map (EnvMap.hpp)
// ---------------------------------------------------------------------------
inline float luminance(const V3 &c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; }

struct Env {
    int w = 0, h = 0;
    std::vector<V3> px;
    std::vector<float> cdf, marg;
    float total = 0.f;
    bool loaded = false;

    // EnvMap.hpp:189-224
    void build(int W, int H, const float *rgb) {
        w = W;
        h = H;
        px.resize((size_t) w * h);
        for (size_t i = 0; i < (size_t) w * h; i++) px[i] = V3(rgb[3 * i], rgb[3 * i + 1], rgb[3 * i + 2]);
        cdf.assign((size_t) w * h, 0.f);
        marg.assign(h, 0.f);
        std::vector<float> rowSums(h, 0.f);
        for (int row = 0; row < h; row++) {
            float theta = (row + 0.5f) / (float) h * kPi;
            float sinT = std::sin(theta);
            float rs = 0.f;
            for (int col = 0; col < w; col++) {
                float lum = luminance(px[(size_t) row * w + col]) * sinT;
                rs += lum;
                cdf[(size_t) row * w + col] = rs;
            }
            if (rs > 0.f)
                for (int col = 0; col < w; col++) cdf[(size_t) row * w + col] /= rs;
            rowSums[row] = rs;
        }
        float t = 0.f;
        for (int row = 0; row < h; row++) {
            t += rowSums[row];
            marg[row] = t;
        }
        total = t;
        if (t > 0.f)
            for (float &v : marg) v /= t;
        loaded = true;
    }
    bool empty() const { return !loaded; }
    // EnvMap.hpp:71-89
    static void dirToUV(const V3 &d, float &u, float &v) {
        float phi = std::atan2(d.z, d.x);
        float theta = std::acos(std::max(-1.f, std::min(1.f, d.y)));
        u = (phi + kPi) / (2.f * kPi);
        v = theta / kPi;
    }
    static V3 uvToDir(float u, float v) {
        float phi = u * 2.f * kPi - kPi;
        float theta = v * kPi;
        float sinT = std::sin(theta);
        return V3(sinT * std::cos(phi), std::cos(theta), sinT * std::sin(phi));
    }
    // EnvMap.hpp:97-113
    V3 sample(const V3 &dir) const {
        if (!loaded) return V3(0);
        float u, v;
        dirToUV(dir, u, v);
        float pxf = std::max(0.f, u * (w - 1));
        float pyf = std::max(0.f, v * (h - 1));
        int x0 = (int) pxf, y0 = (int) pyf;
        int x1 = std::min(x0 + 1, w - 1);
        int y1 = std::min(y0 + 1, h - 1);
        float fx = pxf - x0, fy = pyf - y0;
        auto f = [&](int x, int y) { return px[(size_t) y * w + x]; };
        V3 top = f(x0, y0) + (f(x1, y0) - f(x0, y0)) * fx;
        V3 bot = f(x0, y1) + (f(x1, y1) - f(x0, y1)) * fx;
        return top + (bot - top) * fy;
    }
    // EnvMap.hpp:124-133
    float pdf(const V3 &dir) const {
        float u, v;
        dirToUV(dir, u, v);
        int col = std::min((int) (u * w), w - 1);
        int row = std::min((int) (v * h), h - 1);
        float theta = v * kPi;
        float sinT = std::max(1e-4f, std::sin(theta));
        float lum = luminance(px[(size_t) row * w + col]);
        return (lum * sinT / (total + 1e-6f)) * ((float) (w * h) / (2.f * kPi * kPi * sinT));
    }
    // EnvMap.hpp:145-169 (draws r1 then r2)
    V3 importanceSample(float r1, float r2, float &pdfOut) const {
        int row = (int) (std::lower_bound(marg.begin(), marg.end(), r1) - marg.begin());
        row = std::min(row, h - 1);
        auto rb = cdf.begin() + (size_t) row * w;
        int col = (int) (std::lower_bound(rb, rb + w, r2) - rb);
        col = std::min(col, w - 1);
        float u = (col + 0.5f) / (float) w;
        float v = (row + 0.5f) / (float) h;
        V3 dir = uvToDir(u, v);
        float lum = luminance(px[(size_t) row * w + col]);
        float p = (lum / (total + 1e-6f)) * ((float) (w * h) / (2.f * kPi * kPi));
        pdfOut = std::max(1e-6f, p);
        return dir;
    }
};

// ---------------------------------------------------------------------------
// Camera (Camera.hpp:43-95)
// ---------------------------------------------------------------------------
struct Camera {
    V3 eye = V3(0, 0, -1), target = V3(0, 0, 0), up = V3(0, 1, 0);
    float fovh = 40.f, aperture = 0.f, focusDistance = -1.f;
    V3 fwd, right, camUp;
    float halfFovH = 1.f;
    void init() {
        fwd = normalize(target - eye);
        right = normalize(cross(fwd, up));
        camUp = cross(right, fwd);
        halfFovH = fovh * kPi / 180.0f * 0.5f;
        if (focusDistance < 0.0f) {
            V3 d = target - eye;
            focusDistance = std::sqrt(d.x * d.x + d.y * d.y + d.z * d.z);
        }
    }
    static V3 sampleDisc(float u, float v) {
        float a = 2.0f * u - 1.0f, b = 2.0f * v - 1.0f;
        if (a == 0.0f && b == 0.0f) return V3(0, 0, 0);
        float r, phi;
        if (std::abs(a) > std::abs(b)) {
            r = a;
            phi = (kPi / 4.0f) * (b / a);
        } else {
            r = b;
            phi = (kPi / 2.0f) - (kPi / 4.0f) * (a / b);
        }
        return V3(r * std::cos(phi), r * std::sin(phi), 0.0f);
    }
    template <class R>
    Ray generateRay(int i, int j, int W, int H, float ox, float oy, R &rng) const {
        float fpx = ((float) W * 0.5f) / std::tan(halfFovH);
        float x = ((float) i + ox - (float) W * 0.5f) / fpx;
        float y = ((float) (H - j) + oy - (float) H * 0.5f) / fpx;
        V3 dir = normalize(fwd + x * right + y * camUp);
        if (aperture <= 0.0f) return Ray{eye, dir};
        V3 focusPoint = eye + dir * (focusDistance / dot(dir, fwd));
        float u = rng.next();
        float v = rng.next();
        V3 lo = sampleDisc(u, v);
        V3 lensPoint = eye + aperture * (lo.x * right + lo.y * camUp);
        return Ray{lensPoint, normalize(focusPoint - lensPoint)};
    }
};

// ---------------------------------------------------------------------------
// Tone mapping (ToneMapping.hpp) + ACES (Narkowicz 2015 fitted curve).
// ---------------------------------------------------------------------------
inline float srgbOETF(float x) {  // ToneMapping.hpp:45-48
    x = std::clamp(x, 0.f, 1.f);
    return x < 0.0031308f ? 12.92f * x : 1.055f * std::pow(x, 1.f / 2.4f) - 0.055f;
}
inline V3 srgbOETF(const V3 &c) { return V3(srgbOETF(c.x), srgbOETF(c.y), srgbOETF(c.z)); }
inline V3 clamp01(const V3 &c) {
    return V3(std::clamp(c.x, 0.f, 1.f), std::clamp(c.y, 0.f, 1.f), std::clamp(c.z, 0.f, 1.f));
}
inline V3 mat3(const float m[3][3], const V3 &v) {  // ToneMapping.hpp:55-59
    return V3(m[0][0] * v.x + m[0][1] * v.y + m[0][2] * v.z,
              m[1][0] * v.x + m[1][1] * v.y + m[1][2] * v.z,
              m[2][0] * v.x + m[2][1] * v.y + m[2][2] * v.z);
}
struct AllenwpParams {  // ToneMapping.hpp:114-148
    static constexpr float crossover = 0.1841865f;
    float contrast, toe_a, slope, w, shoulder_max, output_max;
    AllenwpParams(float c = 1.25f, float highClip = 16.0f, float outMax = 1.0f) {
        contrast = c;
        output_max = outMax;
        float high_clip = std::max(highClip, outMax);
        toe_a = ((1.f / crossover) - 1.f) * std::pow(crossover, contrast);
        float denom = std::pow(crossover, contrast) + toe_a;
        slope = (contrast * std::pow(crossover, contrast - 1.f) * toe_a) / (denom * denom);
        shoulder_max = output_max - crossover;
        w = high_clip - crossover;
        w = w * w;
        w = w / shoulder_max;
        w = w * slope;
    }
};
inline float allenwp(float x, const AllenwpParams &p) {  // ToneMapping.hpp:151-166
    x = std::max(x, 0.f);
    if (x < AllenwpParams::crossover) {
        float t = std::pow(x, p.contrast);
        return t / (t + p.toe_a);
    }
    float s = x - AllenwpParams::crossover;
    float slope_s = p.slope * s;
    s = slope_s * (1.f + s / p.w) / (1.f + (slope_s / p.shoulder_max));
    s += AllenwpParams::crossover;
    return s;
}
inline V3 agx(V3 c) {  // ToneMapping.hpp:172-214
    static const float inset[3][3] = {{0.544814746488245f, 0.373787398372697f, 0.0813978551390581f},
                                      {0.140416948464053f, 0.754137554567394f, 0.105445496968552f},
                                      {0.0888104196149096f, 0.178871756420858f, 0.732317823964232f}};
    static const float outset[3][3] = {{1.96488741169489f, -0.855988495690215f, -0.108898916004672f},
                                       {-0.299313364904742f, 1.32639796461980f, -0.0270845997150571f},
                                       {-0.164352742528393f, -0.238183969428088f, 1.40253671195648f}};
    static const AllenwpParams params(1.25f, 16.0f, 1.0f);
    c.x = std::max(c.x, 0.f);
    c.y = std::max(c.y, 0.f);
    c.z = std::max(c.z, 0.f);
    c = mat3(inset, c);
    c = V3(allenwp(c.x, params), allenwp(c.y, params), allenwp(c.z, params));
    c.x = std::min(1.0f, c.x);
    c.y = std::min(1.0f, c.y);
    c.z = std::min(1.0f, c.z);
    c = mat3(outset, c);
    c = clamp01(c);
    return srgbOETF(c);
}
inline V3 pbrNeutral(V3 c) {  // ToneMapping.hpp:234-262
    const float startCompression = 0.8f - 0.04f;
    const float desaturation = 0.15f;
    float x = std::min(c.x, std::min(c.y, c.z));
    float offset = x < 0.08f ? x - 6.25f * x * x : 0.04f;
    c.x -= offset;
    c.y -= offset;
    c.z -= offset;
    float peak = std::max(c.x, std::max(c.y, c.z));
    if (peak < startCompression) return srgbOETF(c);
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
    return srgbOETF(c);
}
// ACES filmic fit, K. Narkowicz 2015 ("ACES Filmic Tone Mapping Curve"):
// (x(ax+b))/(x(cx+d)+e), a=2.51 b=0.03 c=2.43 d=0.59 e=0.14, then sRGB OETF.
inline float acesCurve(float x) {
    x = std::max(x, 0.f);
    float n = x * (2.51f * x + 0.03f);
    float dd = x * (2.43f * x + 0.59f) + 0.14f;
    return std::clamp(n / dd, 0.f, 1.f);
}
inline V3 aces(const V3 &c) { return srgbOETF(V3(acesCurve(c.x), acesCurve(c.y), acesCurve(c.z))); }

inline V3 tonemap(int op, const V3 &c) {  // ToneMapping.hpp:267-279
    switch (op) {
    case 0: return clamp01(c);
    case 1: return srgbOETF(c);
    case 2: return agx(c);
    case 3: return pbrNeutral(c);
    case 4: return aces(c);
    }
    return srgbOETF(c);
}
// Renderer.cpp:149-161
inline V3 exposeLinear(const V3 &fb, float exposure) {
    V3 c = fb * exposure;
    if (!std::isfinite(c.x)) c.x = 0.f;
    if (!std::isfinite(c.y)) c.y = 0.f;
    if (!std::isfinite(c.z)) c.z = 0.f;
    c.x = std::max(0.f, c.x);
    c.y = std::max(0.f, c.y);
    c.z = std::max(0.f, c.z);
    return c;
}
// Renderer.cpp:169-171
inline uint8_t quantize(float v) { return (uint8_t) (std::clamp(v, 0.f, 1.f) * 255.f + 0.5f); }

}  // namespace orc
