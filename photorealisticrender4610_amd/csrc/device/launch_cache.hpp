// launch_cache.hpp — host-side launchers of the radiance-cache kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "cache.hpp"

namespace prr {

int cacheTrainBlocks(int64_t n);
hipError_t launchCacheInfer(const CacheView &cv, int64_t n, const uint32_t *nDev, const float *q, float *out,
                            const uint32_t *slot, const float4 *beta, float4 *result, float clampInd, int grid,
                            hipStream_t st);
hipError_t launchCacheTrain(const CacheView &cv, int64_t n, const float *q, const float *target, float *partial,
                            float *lossPart, float clip, int64_t start, int64_t total, int64_t mult,
                            hipStream_t st);
hipError_t launchCacheAdam(float *master, float *m1, float *m2, const float *partial, int nblocks, const CacheOpt &opt,
                           _Float16 *wt, _Float16 *wn, float *gradOut, hipStream_t st);
hipError_t launchCacheLossSum(const float *lossPart, int nblocks, float *out, hipStream_t st);

}  // namespace prr
