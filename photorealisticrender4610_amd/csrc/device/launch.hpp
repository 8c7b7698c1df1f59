// launch.hpp — host-side launchers of the device kernels (one per .hip file).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dev_common.hpp"

namespace prr {

// k_trace.hip
hipError_t launchTraceClosest(const SceneView &s, int64_t n, const float *org, const float *dir, int32_t *prim,
                              float *t, float *u, float *v, uint2 *spill, int spillLanes, int stackDepth,
                              hipStream_t st);
hipError_t launchTraceAny(const SceneView &s, int64_t n, const float *org, const float *dir, const float *tmax,
                          uint8_t *hit, uint2 *spill, int spillLanes, int stackDepth, int quant, hipStream_t st);

// prr_trace_closest_noalloc: kTraceClosest with its spill region in a
// static device array (kNoAllocBlocks * kTraceBlock lanes of kSpill entries)
// instead of the context's allocation. noAllocScratch() gives the current
// device's address of that array (call it once per device, at set-up).
constexpr int kNoAllocBlocks = 256;
hipError_t noAllocScratch(uint2 **out);
hipError_t launchTraceClosestNoAlloc(const SceneView &s, int64_t n, const float *org, const float *dir, int32_t *prim,
                                     float *t, float *u, float *v, uint2 *scratchPtr, int stackDepth, hipStream_t st);

// k_misc.hip
hipError_t launchSampleLight(const SceneView &s, int64_t n, const float *u, float *pos, float *nrm, int32_t *emitter,
                            float *pdf, hipStream_t st);
hipError_t launchMedium(int64_t n, const float *org, const float *dir, const float *bounds, const int32_t *bounded,
                        float *tEnter, float *tExit, uint8_t *hit, hipStream_t st);
hipError_t launchPhaseHG(int64_t n, const float *cosTheta, const float *g, float *out, hipStream_t st);
hipError_t launchPhaseHGSample(int64_t n, const float *wo, const float *g, const float *u, float *wi, hipStream_t st);
hipError_t launchBsdfEval(int64_t n, const float *wi, const float *wo, const float *nrm, const float *params,
                          float *f, float *pdf, hipStream_t st);
hipError_t launchBsdfSample(int64_t n, const float *wo, const float *nrm, const float *params, const float *u,
                            float *wi, float *f, float *pdf, int32_t *lobe, hipStream_t st);
hipError_t launchEnvSample(const SceneView &s, int64_t n, const float *u2, float *dir, float *pdf, hipStream_t st);
hipError_t launchEnvEval(const SceneView &s, int64_t n, const float *dir, float *rad, float *pdf, hipStream_t st);
hipError_t launchCameraRays(const SceneView &s, int W, int H, int64_t n, const int32_t *pix, const float *offs,
                            float *org, float *dir, hipStream_t st);
hipError_t launchTonemap(int op, int64_t n, float exposure, const float *lin, float *disp, uint8_t *rgb8,
                         const AgxConst &agx, hipStream_t st);

}  // namespace prr
