// launch_render.hpp — host-side launchers of the wavefront render kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "render.hpp"

namespace prr {

hipError_t launchRaygen(const SceneView &s, const PathBufs &pb, const ChunkInfo &ci, int grid, hipStream_t st);
hipError_t launchRaygenRays(const PathBufs &pb, int64_t n, const float *org, const float *dir, int64_t first,
                            uint64_t seed, int grid, hipStream_t st);
hipError_t launchCopyResults(const PathBufs &pb, int64_t n, int64_t first, float *L, hipStream_t st);
// whileWhile: 0 one node per work check; 1 interior nodes then a leaf; n > 1
// the same with at most n interior steps per work check
// tailPaths: the launch does nothing when the bounce's queue holds at most
// that many entries (kFinish takes them); speculate: speculative while-while
// iteration; twoLevel: while-while on the two-level records where the scene
// has them (both need whileWhile > 0)
hipError_t launchExtend(const SceneView &s, const PathBufs &pb, int bounce, uint2 *spill, int traceGrid,
                        int refillAt, int stackDepth, int whileWhile, bool speculate, bool twoLevel, int fetchRun,
                        int64_t tailPaths, hipStream_t st);
// tree: 0 the binary tree, 1 the four-wide occlusion tree, 2 its quantised copy;
// dense: pb's shadow queue arrays hold the queue in order (the binned copy),
// not in segments
hipError_t launchShadow(const SceneView &s, const PathBufs &pb, int bounce, uint2 *spill, int traceGrid,
                        int refillAt, int stackDepth, int whileWhile, int fetchRun, int tree, bool dense,
                        hipStream_t st);
// after shading pass `bounce`: prefix counts and totals of the ray queue of
// bounce + 1 and of the shadow queue of bounce (segmented queues)
hipError_t launchQueueScan(const PathBufs &pb, int bounce, hipStream_t st);
// Shadow-queue binning (k_sort.hip): the entries of bounce's shadow queue
// grouped by (direction octant, Morton code of the origin's cell on a grid of
// 2^bits cells per axis over the scene box) into sb.slot / org / dir, which
// hold the whole queue (sb.cap entries); shadowBinCount(bits) bins, whose
// histograms take bins * sb.blocks words of sb.hist.
int shadowBinCount(int bits);
hipError_t launchShadowSort(const SceneView &s, const PathBufs &pb, int bounce, int bits, const ShadowSortBufs &sb,
                            hipStream_t st);
// waves: minimum waves per SIMD the shading kernel is compiled for (1, 4 or 5)
hipError_t launchShade(const SceneView &s, const PathBufs &pb, const ChunkInfo &ci, int bounce, int grid, int waves,
                       hipStream_t st);
// Shading sorted by material: the bounce's queue split into the entries
// whose hit is a diffuse surface and the rest (sb), each list shaded by its
// own kernel (lightWaves: waves per SIMD the non-diffuse kernel is compiled
// for, 4, 6 or 8). Same results as launchShade.
hipError_t launchShadeSorted(const SceneView &s, const PathBufs &pb, const ChunkInfo &ci, int bounce, int grid,
                             int waves, int lightWaves, const ShadeBins &sb, hipStream_t st);
// path completion: when bounce0 is the first bounce whose queue holds at
// most ci.tailPaths entries, every entry traced and shaded to the end of its
// path in one launch (the shadow rays of bounce0 - 1 resolved); otherwise
// the launch does nothing
hipError_t launchFinish(const SceneView &s, const PathBufs &pb, const ChunkInfo &ci, int bounce0, uint2 *spill,
                        int traceGrid, int stackDepth, int whileWhile, int tree, int numCU, hipStream_t st);
hipError_t launchAccumulate(const PathBufs &pb, const ChunkInfo &ci, float4 *fb, hipStream_t st);
hipError_t launchAccumulateAdaptive(const PathBufs &pb, const ChunkInfo &ci, float4 *fb, const AdaptiveBufs &ab,
                                    int minSpp, float thr, hipStream_t st);
hipError_t launchCompactActive(const AdaptiveBufs &ab, int pixels, int spp, hipStream_t st);
hipError_t launchResolveAdaptive(const float4 *fb, const AdaptiveBufs &ab, int pixels, float *rgb, int32_t *sppOut,
                                 hipStream_t st);
hipError_t launchRaygenSeg(const PathBufs &pb, const Segment *segs, int nseg, int64_t n, uint32_t *sampleOf, int grid,
                           hipStream_t st);
hipError_t launchAccumulateSeg(const PathBufs &pb, const Segment *segs, const int64_t *pixStart, int nseg,
                               int64_t total, hipStream_t st);
hipError_t launchCacheTargets(const PathBufs &pb, const CacheRec &cr, uint32_t r0, uint32_t r1, float *target,
                              hipStream_t st);
hipError_t launchResolve(const float4 *fb, int pixels, int spp, float *rgb, hipStream_t st);

}  // namespace prr
