// k_misc.hip — batch kernels for the BSDF, environment, camera and tone-map
// entry points. Each runs the same device functions the renderer uses.
#include "dev_common.hpp"
#include "launch.hpp"

namespace prr {

constexpr int kBlk = 256;

PRR_DEV V3 ldv(const float *p, int64_t i) { return V3(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }
PRR_DEV void stv(float *p, int64_t i, const V3 &v) {
    p[3 * i] = v.x;
    p[3 * i + 1] = v.y;
    p[3 * i + 2] = v.z;
}

struct FixedU {
    const float *u;
    int k;
    PRR_DEV float next() { return u[k++]; }
};

// Scene::sampleLight (Scene.cpp:37-58) with explicit uniforms: 4 per sample.
__global__ void kSampleLight(SceneView s, int64_t n, const float *u, float *pos, float *nrm, int32_t *emitter,
                             float *pdf) {
    const int64_t i = (int64_t) blockIdx.x * kBlk + threadIdx.x;
    if (i >= n) return;
    FixedU r{u + 4 * i, 0};
    V3 p, nn;
    int mat = 0, e = 0;
    sampleLight(s, r, p, nn, mat, &e);
    stv(pos, i, p);
    stv(nrm, i, nn);
    if (emitter) emitter[i] = e;
    if (pdf) pdf[i] = 1.0f / s.total_emit_area;  // Scene.cpp:46
}

__global__ void kBsdfEval(int64_t n, const float *wi, const float *wo, const float *nrm, const float *params, float *f,
                          float *pdf) {
    const int64_t i = (int64_t) blockIdx.x * kBlk + threadIdx.x;
    if (i >= n) return;
    const float *p = params + 8 * i;
    V3 fv(0.f);
    float pv = 0.f;
    if ((int) p[0] == MAT_DIFFUSE) {
        const Lobes m = makeLobes(V3(p[1], p[2], p[3]), p[4], p[5], p[6]);
        const Frame fr = makeFrame(ldv(nrm, i));
        evalPdf(m, fr, ldv(wi, i), ldv(wo, i), fv, pv);
    }
    stv(f, i, fv);
    pdf[i] = pv;
}

__global__ void kMedium(int64_t n, const float *org, const float *dir, const float *bounds, const int32_t *bounded,
                        float *tEnter, float *tExit, uint8_t *hit) {
    const int64_t i = (int64_t) blockIdx.x * kBlk + threadIdx.x;
    if (i >= n) return;
    float te = 0.f, tx = 0.f;
    const bool h = mediumIntersect(bounded[i] != 0, bounds + 6 * i, bounds + 6 * i + 3, ldv(org, i), ldv(dir, i), te, tx);
    hit[i] = h ? 1 : 0;
    if (h) {  // the reference leaves its outputs untouched on a miss
        tEnter[i] = te;
        tExit[i] = tx;
    }
}

__global__ void kPhaseHG(int64_t n, const float *cosTheta, const float *g, float *out) {
    const int64_t i = (int64_t) blockIdx.x * kBlk + threadIdx.x;
    if (i < n) out[i] = phaseHG(cosTheta[i], g[i]);
}

__global__ void kPhaseHGSample(int64_t n, const float *wo, const float *g, const float *u, float *wi) {
    const int64_t i = (int64_t) blockIdx.x * kBlk + threadIdx.x;
    if (i < n) stv(wi, i, samplePhaseHG(ldv(wo, i), g[i], u[2 * i], u[2 * i + 1]));
}

__global__ void kBsdfSample(int64_t n, const float *wo, const float *nrm, const float *params, const float *u,
                            float *wi, float *f, float *pdf, int32_t *lobe) {
    const int64_t i = (int64_t) blockIdx.x * kBlk + threadIdx.x;
    if (i >= n) return;
    const float *p = params + 8 * i;
    const V3 N = ldv(nrm, i), w = ldv(wo, i);
    BsdfSample s;
    s.wi = V3(0.f);
    s.f = V3(0.f);
    s.pdf = 0.f;
    s.lobe = -1;
    const int type = (int) p[0];
    if (type == MAT_DIFFUSE) {
        FixedU r{u + 3 * i, 0};
        s = sampleDiffuse(makeLobes(V3(p[1], p[2], p[3]), p[4], p[5], p[6]), makeFrame(N), w, r);
    } else if (type == MAT_GLASS) {
        s = sampleGlass(w, N, p[6], u[3 * i]);
    } else if (type == MAT_MIRROR) {
        s = sampleMirror(w, N, V3(p[1], p[2], p[3]));
    }
    stv(wi, i, s.wi);
    stv(f, i, s.f);
    pdf[i] = s.pdf;
    lobe[i] = s.lobe;
}

__global__ void kEnvSample(SceneView s, int64_t n, const float *u2, float *dir, float *pdf) {
    const int64_t i = (int64_t) blockIdx.x * kBlk + threadIdx.x;
    if (i >= n) return;
    float p;
    const V3 d = envImportanceSample(s, u2[2 * i], u2[2 * i + 1], p);
    stv(dir, i, d);
    pdf[i] = p;
}

__global__ void kEnvEval(SceneView s, int64_t n, const float *dir, float *rad, float *pdf) {
    const int64_t i = (int64_t) blockIdx.x * kBlk + threadIdx.x;
    if (i >= n) return;
    const V3 d = ldv(dir, i);
    stv(rad, i, envLookup(s, d));
    pdf[i] = envPdf(s, d);
}

__global__ void kCameraRays(SceneView s, int W, int H, int64_t n, const int32_t *pix, const float *offs, float *org,
                            float *dir) {
    const int64_t i = (int64_t) blockIdx.x * kBlk + threadIdx.x;
    if (i >= n) return;
    const int stride = s.cam.aperture > 0.f ? 4 : 2;
    FixedU lens{offs + stride * i + 2, 0};
    V3 o, d;
    cameraRay(s.cam, pix[2 * i], pix[2 * i + 1], W, H, offs[stride * i], offs[stride * i + 1], lens, o, d);
    stv(org, i, o);
    stv(dir, i, d);
}

PRR_DEV uint8_t quantize(float v) { return (uint8_t) (clamp01f(v) * 255.f + 0.5f); }  // Renderer.cpp:169-171

__global__ void kTonemap(int op, int64_t n, float exposure, const float *lin, float *disp, uint8_t *rgb8,
                         AgxConst agx) {
    const int64_t i = (int64_t) blockIdx.x * kBlk + threadIdx.x;
    if (i >= n) return;
    const V3 c = tonemapApply(op, exposeLinear(ldv(lin, i), exposure), agx);
    if (disp) stv(disp, i, c);
    if (rgb8) {
        rgb8[3 * i] = quantize(c.x);
        rgb8[3 * i + 1] = quantize(c.y);
        rgb8[3 * i + 2] = quantize(c.z);
    }
}

static dim3 grid(int64_t n) { return dim3((unsigned) ((n + kBlk - 1) / kBlk)); }

hipError_t launchBsdfEval(int64_t n, const float *wi, const float *wo, const float *nrm, const float *params,
                          float *f, float *pdf, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(kBsdfEval, grid(n), dim3(kBlk), 0, st, n, wi, wo, nrm, params, f, pdf);
    return hipGetLastError();
}
hipError_t launchBsdfSample(int64_t n, const float *wo, const float *nrm, const float *params, const float *u,
                            float *wi, float *f, float *pdf, int32_t *lobe, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(kBsdfSample, grid(n), dim3(kBlk), 0, st, n, wo, nrm, params, u, wi, f, pdf, lobe);
    return hipGetLastError();
}
hipError_t launchMedium(int64_t n, const float *org, const float *dir, const float *bounds, const int32_t *bounded,
                        float *tEnter, float *tExit, uint8_t *hit, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(kMedium, grid(n), dim3(kBlk), 0, st, n, org, dir, bounds, bounded, tEnter, tExit, hit);
    return hipGetLastError();
}
hipError_t launchPhaseHG(int64_t n, const float *cosTheta, const float *g, float *out, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(kPhaseHG, grid(n), dim3(kBlk), 0, st, n, cosTheta, g, out);
    return hipGetLastError();
}
hipError_t launchPhaseHGSample(int64_t n, const float *wo, const float *g, const float *u, float *wi, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(kPhaseHGSample, grid(n), dim3(kBlk), 0, st, n, wo, g, u, wi);
    return hipGetLastError();
}
hipError_t launchEnvSample(const SceneView &s, int64_t n, const float *u2, float *dir, float *pdf, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(kEnvSample, grid(n), dim3(kBlk), 0, st, s, n, u2, dir, pdf);
    return hipGetLastError();
}
hipError_t launchEnvEval(const SceneView &s, int64_t n, const float *dir, float *rad, float *pdf, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(kEnvEval, grid(n), dim3(kBlk), 0, st, s, n, dir, rad, pdf);
    return hipGetLastError();
}
hipError_t launchCameraRays(const SceneView &s, int W, int H, int64_t n, const int32_t *pix, const float *offs,
                            float *org, float *dir, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(kCameraRays, grid(n), dim3(kBlk), 0, st, s, W, H, n, pix, offs, org, dir);
    return hipGetLastError();
}
hipError_t launchTonemap(int op, int64_t n, float exposure, const float *lin, float *disp, uint8_t *rgb8,
                         const AgxConst &agx, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(kTonemap, grid(n), dim3(kBlk), 0, st, op, n, exposure, lin, disp, rgb8, agx);
    return hipGetLastError();
}

hipError_t launchSampleLight(const SceneView &s, int64_t n, const float *u, float *pos, float *nrm, int32_t *emitter,
                            float *pdf, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(kSampleLight, dim3((unsigned) ((n + kBlk - 1) / kBlk)), dim3(kBlk), 0, st, s, n, u, pos, nrm,
                       emitter, pdf);
    return hipGetLastError();
}

}  // namespace prr
