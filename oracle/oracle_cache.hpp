// oracle_cache.hpp — CPU restatement of the radiance-cache network (TEST
// INFRASTRUCTURE ONLY): the query encoding, the forward pass, the gradient of
// the relative L2 loss and Adam, with the device's FP16 rounding points
// (weights, layer inputs, output deltas) reproduced by round-to-nearest-even
// to binary16, or switched off (exact mode) for finite-difference checks.
//
// Network: 96 inputs -> 5 x 64 ReLU -> 3, no biases (input 95 is 1); weights
// W^T [out][in] per layer, layers in order (22720 floats). Loss: mean over
// samples and channels of (y - t)^2 / (y^2 + 0.01) with y^2 held constant
// (Mueller et al. 2021); dL/dy = 2 (y - t) / (y^2 + 0.01) / (3 n), the
// per-output factor clipped to +-clip before the 1/(3n).
#pragma once

#include <cmath>
#include <cstdint>
#include <vector>

namespace orc_cache {

constexpr int kIn = 96, kH = 64, kOut = 3, kLayers = 6, kQ = 13;
constexpr int kWeights = kIn * kH + 4 * kH * kH + kOut * kH;
inline int layerIn(int l) { return l == 0 ? kIn : kH; }
inline int layerOut(int l) { return l == kLayers - 1 ? kOut : kH; }
inline int layerOff(int l) { return l == 0 ? 0 : kIn * kH + (l - 1) * kH * kH; }

// Round to the nearest binary16 value (ties to even), returned as float;
// overflow to infinity, subnormals with quantum 2^-24.
inline float toHalf(float x) {
    if (!std::isfinite(x) || x == 0.f) return x;
    const float a = std::fabs(x);
    if (a >= 65520.f) return std::copysign(INFINITY, x);
    float q;
    if (a < 6.103515625e-05f) {
        q = std::nearbyint(a * 16777216.f) / 16777216.f;
    } else {
        int e;
        std::frexp(a, &e);  // a = m 2^e, m in [0.5, 1)
        const float quantum = std::ldexp(1.f, e - 11);
        q = std::nearbyint(a / quantum) * quantum;
    }
    return std::copysign(q, x);
}

struct Round {
    bool half;
    float operator()(float x) const { return half ? toHalf(x) : x; }
};

inline float oneBlob(float s, int j) {
    const float d = 4.f * s - ((float) j + 0.5f);
    return std::exp(-0.5f * d * d);
}

// Encoding of one query (pos3, wo3, normal3, roughness, albedo3) in the
// scene box [bmin, bmin + ext]: 24 groups of 4 features, group G:
//   G < 18: axis G % 3, octave k = 2 (G / 3): sin, cos of 2^k pi p, then
//           2 s c and c c - s s (the next octave by the double-angle formulas)
//   18..21: one-blob (4 bins) of wo theta/pi, wo phi/2pi + 1/2, normal
//           theta/pi, normal phi/2pi + 1/2; 22: one-blob of roughness;
//   23: albedo rgb, 1.
inline void encode(const float *q, const float *bounds, float *x) {
    float p[3];
    for (int a = 0; a < 3; a++) {
        const float ext = bounds[3 + a] - bounds[a];
        const float inv = ext > 0.f ? 1.f / ext : 0.f;
        p[a] = std::min(1.f, std::max(0.f, (q[a] - bounds[a]) * inv));
    }
    const double pi = 3.14159265358979323846;
    auto theta = [](float y) { return std::acos(std::min(1.f, std::max(-1.f, y))) * (1.f / 3.141592653589793f); };
    auto phi = [](float x, float z) { return std::atan2(z, x) * (0.5f / 3.141592653589793f) + 0.5f; };
    const float ang[4] = {theta(q[4]), phi(q[3], q[5]), theta(q[7]), phi(q[6], q[8])};
    for (int G = 0; G < 24; G++) {
        float *v = x + 4 * G;
        if (G < 18) {
            const int k = 2 * (G / 3);
            const float arg = p[G % 3] * (float) (1 << k);
            // sin(pi x), cos(pi x) with exact reduction of x modulo 2
            const double r = std::fmod((double) arg, 2.0);
            const float s = (float) std::sin(pi * r), c = (float) std::cos(pi * r);
            v[0] = s;
            v[1] = c;
            v[2] = 2.f * s * c;
            v[3] = c * c - s * s;
        } else if (G < 23) {
            const float a = G < 22 ? ang[G - 18] : q[9];
            for (int j = 0; j < 4; j++) v[j] = oneBlob(a, j);
        } else {
            v[0] = q[10];
            v[1] = q[11];
            v[2] = q[12];
            v[3] = 1.f;
        }
    }
}

// Forward pass of one sample; acts[l] = input of layer l (rounded), y = output.
inline void forward(const float *w, const float *x0, const Round &R, std::vector<float> (&acts)[kLayers], float *y) {
    acts[0].assign(x0, x0 + kIn);
    for (float &v : acts[0]) v = R(v);
    for (int l = 0; l < kLayers; l++) {
        const int in = layerIn(l), out = layerOut(l);
        std::vector<float> z(out);
        for (int o = 0; o < out; o++) {
            double s = 0.0;
            for (int i = 0; i < in; i++) s += (double) R(w[layerOff(l) + o * in + i]) * (double) acts[l][i];
            z[o] = (float) s;
        }
        if (l + 1 < kLayers) {
            acts[l + 1].resize(out);
            for (int o = 0; o < out; o++) acts[l + 1][o] = R(std::max(z[o], 0.f));
        } else {
            for (int o = 0; o < out; o++) y[o] = z[o];
        }
    }
}

// Gradient of the mean loss over n samples (grad: kWeights floats, zeroed
// here); returns the mean loss.
inline double gradient(const float *w, int64_t n, const float *q, const float *t, const float *bounds, const Round &R,
                       float clip, float *grad) {
    std::vector<double> g(kWeights, 0.0);
    double loss = 0.0;
    std::vector<float> acts[kLayers];
    float x0[kIn];
    for (int64_t s = 0; s < n; s++) {
        encode(q + kQ * s, bounds, x0);
        float y[kOut];
        forward(w, x0, R, acts, y);
        std::vector<float> d(kOut);
        for (int k = 0; k < kOut; k++) {
            const float yv = y[k], tv = t[3 * s + k];
            const float den = yv * yv + 0.01f;
            loss += (double) (yv - tv) * (yv - tv) / den;
            d[k] = R(std::min(clip, std::max(-clip, 2.f * (yv - tv) / den)));
        }
        for (int l = kLayers - 1; l >= 0; l--) {
            const int in = layerIn(l), out = layerOut(l);
            for (int o = 0; o < out; o++)
                for (int i = 0; i < in; i++) g[layerOff(l) + o * in + i] += (double) d[o] * (double) acts[l][i];
            if (l == 0) break;
            std::vector<float> nd(in);
            for (int i = 0; i < in; i++) {
                double s2 = 0.0;
                for (int o = 0; o < out; o++) s2 += (double) R(w[layerOff(l) + o * in + i]) * (double) d[o];
                nd[i] = R(acts[l][i] > 0.f ? (float) s2 : 0.f);
            }
            d.swap(nd);
        }
    }
    const double scale = 1.0 / (3.0 * (double) n);
    for (int i = 0; i < kWeights; i++) grad[i] = (float) (g[i] * scale);
    return loss * scale;
}

// One Adam step (step = 1 for the first update), in place.
inline void adam(float *w, float *m1, float *m2, const float *grad, int64_t step, float lr, float b1, float b2,
                 float eps) {
    const float bias1 = 1.f - (float) std::pow((double) b1, (double) step);
    const float bias2 = 1.f - (float) std::pow((double) b2, (double) step);
    for (int i = 0; i < kWeights; i++) {
        const float gr = grad[i];
        const float a = b1 * m1[i] + (1.f - b1) * gr;
        const float v = b2 * m2[i] + (1.f - b2) * gr * gr;
        m1[i] = a;
        m2[i] = v;
        w[i] -= lr * (a / bias1) / (std::sqrt(v / bias2) + eps);
    }
}

}  // namespace orc_cache
