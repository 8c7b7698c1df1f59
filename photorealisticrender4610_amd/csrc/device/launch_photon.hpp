// launch_photon.hpp — host-side launchers of the photon-map kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>

#include "dev_common.hpp"

namespace prr {

hipError_t launchPhotonTrace(const SceneView &s, int64_t n, uint64_t seed, int maxDepth, float4 *pos, float4 *dir,
                             float4 *pw, uint32_t *stored, uint2 *spill, int spillLanes, int stackDepth,
                             hipStream_t st);
hipError_t launchPhotonKeys(int64_t n, const float4 *pos, float cell, uint32_t mask, uint32_t *keys, uint32_t *vals,
                            hipStream_t st);
// hipCUB stable radix sort of (key, value); temp == nullptr queries tempBytes
hipError_t photonSort(void *temp, size_t &tempBytes, const uint32_t *kin, uint32_t *kout, const uint32_t *vin,
                      uint32_t *vout, int n, hipStream_t st);
hipError_t launchPhotonFinish(uint32_t m, const uint32_t *keys, const uint32_t *vals, const float4 *pos,
                              const float4 *dir, const float4 *pw, float4 *spos, float4 *sdir, float4 *spw,
                              uint32_t *start, uint32_t *end, hipStream_t st);

hipError_t launchPhotonRadiance(const SceneView &s, int64_t n, const float *pos, const float *wo, const float *nrm,
                                const float *params, float *out, hipStream_t st);

}  // namespace prr
