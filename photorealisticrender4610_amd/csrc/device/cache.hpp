// cache.hpp — neural radiance cache: network layout, query encoding and the
// MFMA fragment helpers shared by the cache kernels and the host code.
//
// Network (Mueller et al. 2021, "Real-time Neural Radiance Caching"): 96
// encoded inputs -> 5 hidden layers of 64 (ReLU) -> 3 outputs (linear), no
// biases (input 95 is a constant 1). FP16 weights and activations on MFMA
// (v_mfma_f32_16x16x16_f16, and gfx950's v_mfma_f32_16x16x32_f16 for the
// forward layers), FP32 accumulation, FP32 master weights and Adam moments.
//
// Fragment layout of v_mfma_f32_16x16x16_f16 for D = A * B (16x16x16), lane
// l = c + 16 g (c = l % 16, g = l / 16):
//   A (M x K): lane holds A[c][4g + 0..3]
//   B (K x N): lane holds B[4g + 0..3][c]
//   D (M x N): lane holds D[4g + 0..3][c]
// Layers compute Z^T = W^T X^T with queries as the N dimension, so a layer's
// D fragments (features 4g+i of query c) are the next layer's B fragments
// without leaving registers.
#pragma once

#include "dev_common.hpp"

namespace prr {

typedef _Float16 h4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int kCacheIn = 96;     // encoded inputs: 95 features + constant 1
constexpr int kCacheH = 64;      // hidden width
constexpr int kCacheLayers = 6;  // weight matrices: 96->64, 4 x 64->64, 64->3
constexpr int kCacheOut = 3;
constexpr int kCacheQ = 13;      // raw query floats: pos3, wo3, normal3, roughness, albedo3
constexpr int kCacheWeights = kCacheIn * kCacheH + 4 * kCacheH * kCacheH + kCacheOut * kCacheH;  // 22720

PRR_HD constexpr int cacheIn(int l) { return l == 0 ? kCacheIn : kCacheH; }
PRR_HD constexpr int cacheOut(int l) { return l == kCacheLayers - 1 ? kCacheOut : kCacheH; }
PRR_HD constexpr int cacheOutPad(int l) { return l == kCacheLayers - 1 ? 16 : kCacheH; }
// FP32 master weights / gradients: W^T [out][in], real rows only.
PRR_HD constexpr int cacheOff(int l) { return l == 0 ? 0 : kCacheIn * kCacheH + (l - 1) * kCacheH * kCacheH; }
// FP16 W^T [outPad][in] with padded rows (LDS bank spread), forward A operands.
PRR_HD constexpr int cacheStrideT(int l) { return cacheIn(l) + 4; }
PRR_HD constexpr int cacheOffT(int l) { return l == 0 ? 0 : kCacheH * (kCacheIn + 4) + (l - 1) * kCacheH * (kCacheH + 4); }
constexpr int kCacheHalfT = kCacheH * (kCacheIn + 4) + 4 * kCacheH * (kCacheH + 4) + 16 * (kCacheH + 4);  // 24896
// FP16 W [in][outPad] with padded rows, layers 1..5, backward A operands.
PRR_HD constexpr int cacheStrideN(int l) { return cacheOutPad(l) + 4; }
PRR_HD constexpr int cacheOffN(int l) { return (l - 1) * kCacheH * (kCacheH + 4); }
constexpr int kCacheHalfN = 4 * kCacheH * (kCacheH + 4) + kCacheH * (16 + 4);  // 18688

struct CacheView {
    const _Float16 *wt;  // kCacheHalfT
    const _Float16 *wn;  // kCacheHalfN
    float bmin[3];       // scene box for the position encoding
    float inv_ext[3];
};

struct CacheOpt {
    float lr, beta1, beta2, eps;
    float bias1, bias2;  // 1 - beta^t
    float gscale;        // loss normalisation 1 / (3 n)
    float clip;          // |dL/dy| bound before the FP16 backward pass
};

// ---------------------------------------------------------------------------
// Encoding: 96 inputs in 24 groups of 4 (feature f = 4 G + i):
//   G in [0, 18)  position in the scene box p in [0,1]^3: axis a = G % 3,
//                 k = 2 (G / 3): sin(2^k pi p_a), cos(2^k pi p_a), then the
//                 next octave by the double-angle formulas 2 s c, c c - s s
//                 (12 frequencies per axis, 72 features)
//   G 18..21      one-blob (4 bins) of wo theta/pi, wo phi/2pi + 1/2, normal
//                 theta/pi, normal phi/2pi + 1/2 (theta from +y, phi =
//                 atan2(z, x))
//   G 22          one-blob (4 bins) of roughness
//   G 23          albedo rgb, 1 (bias)
// One-blob (Mueller et al. 2019): bin j of scalar s in [0,1] holds
// exp(-0.5 (4 s - j - 0.5)^2). An MFMA input fragment (block kb, lane group
// g) is exactly group G = 4 kb + g, so each lane evaluates one group.
// ---------------------------------------------------------------------------
struct CacheQuery {
    float p[3];      // normalised position
    float ang;       // the one angle this lane's groups use (see cacheQuery)
    float rough;
    float alb[3];
};

// Lane group g (lane / 16) evaluates groups G = 4 kb + g, kb = 0..5, so it
// needs one angle only: G 20 (normal theta) for g = 0, 21 (normal phi) for
// g = 1, 18 (wo theta) for g = 2, 19 (wo phi) for g = 3. Each is one atan2:
// theta = acos(y) = atan2(sqrt(1 - y^2), y) with y clamped to [-1, 1], phi =
// atan2(z, x).
PRR_DEV CacheQuery cacheQuery(const CacheView &cv, const float *q, int g) {
    CacheQuery r;
    for (int a = 0; a < 3; a++) r.p[a] = fminf(1.f, fmaxf(0.f, (q[a] - cv.bmin[a]) * cv.inv_ext[a]));
    const bool theta = (g & 1) == 0;
    const int v = g < 2 ? 6 : 3;  // normal or wo
    const float y = fminf(1.f, fmaxf(-1.f, q[v + 1]));
    const float ay = theta ? sqrtf(fmaxf(0.f, 1.f - y * y)) : q[v + 2];
    const float ax = theta ? y : q[v];
    const float t = atan2f(ay, ax);
    r.ang = theta ? t * (1.f / kPi) : t * (0.5f / kPi) + 0.5f;
    r.rough = q[9];
    r.alb[0] = q[10];
    r.alb[1] = q[11];
    r.alb[2] = q[12];
    return r;
}

// One-blob bin j of s: exp(-0.5 (4 s - j - 0.5)^2), with the hardware exp.
PRR_DEV float oneBlob(float s, int j) {
    const float d = 4.f * s - ((float) j + 0.5f);
    return __expf(-0.5f * d * d);
}

// Element k of a small array as a blend of finite values (exact: one term is
// v[k] * 1, the others are 0), so a per-lane k does not move the query to
// scratch memory as an indexed load would.
PRR_DEV float pick3(const float (&v)[3], int k) {
    return v[0] * (float) (k == 0) + v[1] * (float) (k == 1) + v[2] * (float) (k == 2);
}

// sin(pi x) and cos(pi x) for x >= 0: the hardware's sin and cos of 2 pi r
// (v_sin_f32 / v_cos_f32 take revolutions) on r = fract(x / 2), which is
// exact, so the large octaves lose nothing to argument reduction.
PRR_DEV void sinCosPi(float x, float &s, float &c) {
    const float r = __builtin_amdgcn_fractf(0.5f * x);
    s = __builtin_amdgcn_sinf(r);
    c = __builtin_amdgcn_cosf(r);
}

// Input group G (features 4G..4G+3) of the lane group that evaluates it.
PRR_DEV h4 cacheInputGroup(const CacheQuery &q, int G) {
    float v0, v1, v2, v3;
    if (G < 18) {
        const int k = 2 * (G / 3);
        const float x = pick3(q.p, G - 3 * (G / 3)) * (float) (1 << k);
        float s, c;
        sinCosPi(x, s, c);
        v0 = s;
        v1 = c;
        v2 = 2.f * s * c;
        v3 = c * c - s * s;
    } else if (G < 23) {
        const float a = G < 22 ? q.ang : q.rough;
        v0 = oneBlob(a, 0);
        v1 = oneBlob(a, 1);
        v2 = oneBlob(a, 2);
        v3 = oneBlob(a, 3);
    } else {
        v0 = q.alb[0];
        v1 = q.alb[1];
        v2 = q.alb[2];
        v3 = 1.f;
    }
    h4 r;
    r[0] = (_Float16) v0;
    r[1] = (_Float16) v1;
    r[2] = (_Float16) v2;
    r[3] = (_Float16) v3;
    return r;
}

// B fragment of input block kb for this lane (group g): features kb*16 + 4g + i.
PRR_DEV h4 cacheInputFrag(const CacheQuery &q, int kb, int g) { return cacheInputGroup(q, 4 * kb + g); }

PRR_DEV f4 mfma16(h4 a, h4 b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x16f16(a, b, c, 0, 0, 0); }
// gfx950's double-depth form: each lane's 8 K-elements are two 16-deep
// fragments (K-blocks 2j and 2j + 1) side by side, in A and B alike, so one
// instruction does the work of two mfma16 on the same operands.
typedef _Float16 h8 __attribute__((ext_vector_type(8)));
PRR_DEV h8 cat8(h4 a, h4 b) { return __builtin_shufflevector(a, b, 0, 1, 2, 3, 4, 5, 6, 7); }
PRR_DEV f4 mfma32(h8 a, h8 b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0); }

PRR_DEV h4 reluHalf(f4 v) {
    h4 r;
    r[0] = (_Float16) fmaxf(v[0], 0.f);
    r[1] = (_Float16) fmaxf(v[1], 0.f);
    r[2] = (_Float16) fmaxf(v[2], 0.f);
    r[3] = (_Float16) fmaxf(v[3], 0.f);
    return r;
}

}  // namespace prr
