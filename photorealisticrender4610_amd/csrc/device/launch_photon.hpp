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

// the reference's rules (PhotonMap::emit / trace): appends up to cap records
// with key photon * 128 + depth; *count = records stored (may exceed cap)
hipError_t launchPhotonTraceRef(const SceneView &s, int64_t n, uint64_t seed, int maxDepth, float4 *pos, float4 *dir,
                                float4 *pw, uint64_t *key, uint32_t cap, uint32_t *count, uint2 *spill, int spillLanes,
                                int stackDepth, hipStream_t st);
// hipCUB radix sort of (64-bit key, value); temp == nullptr queries tempBytes
hipError_t photonSortByKey(void *temp, size_t &tempBytes, const uint64_t *kin, uint64_t *kout, const uint32_t *vin,
                           uint32_t *vout, int n, hipStream_t st);
hipError_t launchPhotonIota(uint32_t m, uint32_t *vals, hipStream_t st);
// out[j] = record vals[j], j < m
hipError_t launchPhotonGather(uint32_t m, const uint32_t *vals, const float4 *pos, const float4 *dir, const float4 *pw,
                              const uint64_t *key, float4 *opos, float4 *odir, float4 *opw, uint64_t *okey,
                              hipStream_t st);
hipError_t launchPhotonRanges(uint32_t m, const uint32_t *keys, uint32_t *start, uint32_t *end, hipStream_t st);
// PhotonMap::estimateIrradiance at n points
hipError_t launchPhotonIrradiance(const SceneView &s, int64_t n, const float *pos, float *out, hipStream_t st);

hipError_t launchPhotonRadiance(const SceneView &s, int64_t n, const float *pos, const float *wo, const float *nrm,
                                const float *params, float *out, hipStream_t st);

}  // namespace prr
